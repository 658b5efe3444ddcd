/* fvo_mesh.cpp -- oracle mesh topology and geometry (TEST INFRASTRUCTURE ONLY).
 *
 * Restates:
 *  - UMesh::correctBoundaryFaceOrientation      src/mesh/mesh.cpp:54-82
 *  - UMesh::compute_elementsSurroundingElements src/mesh/mesh.cpp:470-560 (edge matching)
 *  - UMesh::compute_faceConnectivity            src/mesh/mesh.cpp:659-762
 *  - UMesh::compute_face_data                   src/mesh/mesh.cpp:346-365
 *  - UMesh::compute_areas                       src/mesh/mesh.cpp:290-313
 *  - UMesh::compute_cell_centres                src/mesh/mesh.cpp:316-328
 *  - Spatial ctor (face centres, ghost centres) src/spatial/aspatial.cpp:37-119
 *  - levelSchedule                              src/mesh/ameshutils.cpp:158-196
 * Parity: pinned by the reference's level-schedule fixture (squarecoarselevels.dat), by the
 * face-count formulas of the generated meshes and by closed-form geometry checks in tests/.
 */
#include "fvo.hpp"
#include <map>
#include <utility>

namespace fvo {

static inline long long edge_key(int a, int b)
{
	const int lo = a < b ? a : b, hi = a < b ? b : a;
	return (static_cast<long long>(lo) << 32) | static_cast<unsigned int>(hi);
}

void mesh_build(Mesh& m)
{
	const int ne = m.nelem;
	// neighbour across each local face: local face j joins nodes j and j+1 (mesh.cpp:714-723)
	m.esuel.assign(ne*4, -1);
	m.elemface.assign(ne*4, -1);
	std::map<long long, std::pair<int,int>> edges;
	for(int ie = 0; ie < ne; ie++) {
		const int nn = m.nnode[ie];
		for(int j = 0; j < nn; j++) {
			const int a = m.inpoel[ie*4+j], b = m.inpoel[ie*4+(j+1)%nn];
			const long long k = edge_key(a,b);
			auto it = edges.find(k);
			if(it == edges.end())
				edges[k] = std::make_pair(ie, j);
			else {
				const int oe = it->second.first, oj = it->second.second;
				m.esuel[ie*4+j] = oe;
				m.esuel[oe*4+oj] = ie;
			}
		}
	}

	// boundary faces: find host element, correct orientation (mesh.cpp:54-82)
	const int nb = m.nbface;
	std::vector<int> hostelem(nb), hostface(nb);
	for(int ib = 0; ib < nb; ib++) {
		const int a = m.bface[ib*2], b = m.bface[ib*2+1];
		auto it = edges.find(edge_key(a,b));
		if(it == edges.end())
			throw std::runtime_error("boundary face without host element");
		const int he = it->second.first, hj = it->second.second;
		if(m.esuel[he*4+hj] != -1)
			throw std::runtime_error("boundary face is an interior edge");
		hostelem[ib] = he; hostface[ib] = hj;
		const int nn = m.nnode[he];
		if(m.inpoel[he*4+hj] != a || m.inpoel[he*4+(hj+1)%nn] != b) {
			m.bface[ib*2] = b; m.bface[ib*2+1] = a;
		}
	}

	// count interior faces (mesh.cpp:662-673)
	m.ninface = 0;
	for(int ie = 0; ie < ne; ie++)
		for(int j = 0; j < m.nnode[ie]; j++) {
			const int je = m.esuel[ie*4+j];
			if(je > ie && je < ne) m.ninface++;
		}
	m.naface = m.ninface + nb;
	m.intfac.assign(m.naface*4, -1);

	// physical boundary faces first (mesh.cpp:683-700)
	for(int ib = 0; ib < nb; ib++) {
		m.intfac[ib*4+0] = hostelem[ib];
		m.intfac[ib*4+1] = ne + ib;
		m.intfac[ib*4+2] = m.bface[ib*2];
		m.intfac[ib*4+3] = m.bface[ib*2+1];
		m.esuel[hostelem[ib]*4+hostface[ib]] = ne + ib;
		m.elemface[hostelem[ib]*4+hostface[ib]] = ib;
	}
	// interior faces in element order, lower element on the left (mesh.cpp:709-733)
	int fi = nb;
	for(int ie = 0; ie < ne; ie++) {
		const int nn = m.nnode[ie];
		for(int j = 0; j < nn; j++) {
			const int je = m.esuel[ie*4+j];
			if(je > ie && je < ne) {
				m.intfac[fi*4+0] = ie;
				m.intfac[fi*4+1] = je;
				m.intfac[fi*4+2] = m.inpoel[ie*4+j];
				m.intfac[fi*4+3] = m.inpoel[ie*4+(j+1)%nn];
				m.elemface[ie*4+j] = fi;
				for(int k = 0; k < m.nnode[je]; k++)
					if(m.esuel[je*4+k] == ie) m.elemface[je*4+k] = fi;
				fi++;
			}
		}
	}

	// face normals and lengths (mesh.cpp:346-365)
	m.facemetric.assign(m.naface*3, 0.0);
	for(int f = 0; f < m.naface; f++) {
		const int a = m.intfac[f*4+2], b = m.intfac[f*4+3];
		double nx = m.coords[b*2+1] - m.coords[a*2+1];
		double ny = -1.0*(m.coords[b*2+0] - m.coords[a*2+0]);
		const double len = std::sqrt(std::pow(nx,2) + std::pow(ny,2));
		m.facemetric[f*3+0] = nx/len;
		m.facemetric[f*3+1] = ny/len;
		m.facemetric[f*3+2] = len;
	}

	// areas: shoelace, quads as two triangles (mesh.cpp:290-313)
	m.area.assign(ne, 0.0);
	auto tri = [&](int p0, int p1, int p2) {
		const double* c = m.coords.data();
		return 0.5*(c[p0*2]*(c[p1*2+1] - c[p2*2+1]) - c[p0*2+1]*(c[p1*2] - c[p2*2])
		            + c[p1*2]*c[p2*2+1] - c[p2*2]*c[p1*2+1]);
	};
	for(int ie = 0; ie < ne; ie++) {
		const int* in = &m.inpoel[ie*4];
		if(m.nnode[ie] == 3)
			m.area[ie] = tri(in[0], in[1], in[2]);
		else {
			m.area[ie] = tri(in[0], in[1], in[2]);
			m.area[ie] += tri(in[0], in[2], in[3]);
		}
	}

	// cell centres: vertex average (mesh.cpp:316-328)
	m.rc.assign(ne*2, 0.0);
	for(int ie = 0; ie < ne; ie++)
		for(int d = 0; d < 2; d++) {
			double s = 0;
			for(int j = 0; j < m.nnode[ie]; j++)
				s += m.coords[m.inpoel[ie*4+j]*2+d];
			m.rc[ie*2+d] = s / (double)m.nnode[ie];
		}

	// face centres (aspatial.cpp:50-61) and ghost centres about the midpoint (:97-119)
	m.gr.assign(m.naface*2, 0.0);
	for(int f = 0; f < m.naface; f++)
		for(int d = 0; d < 2; d++) {
			double s = 0;
			s += m.coords[m.intfac[f*4+2]*2+d];
			s += m.coords[m.intfac[f*4+3]*2+d];
			m.gr[f*2+d] = s / 2;
		}
	m.rcbp.assign(nb*2, 0.0);
	for(int ib = 0; ib < nb; ib++) {
		const int ie = m.intfac[ib*4];
		for(int d = 0; d < 2; d++) {
			double mid = 0;
			mid += m.coords[m.intfac[ib*4+2]*2+d];
			mid += m.coords[m.intfac[ib*4+3]*2+d];
			mid /= 2;
			m.rcbp[ib*2+d] = 2.0*mid - m.rc[ie*2+d];
		}
	}
}

/* Contiguous levels in which no two cells are neighbours (ameshutils.cpp:158-196). */
std::vector<int> level_schedule(const Mesh& m)
{
	std::vector<int> levels;
	levels.push_back(0);
	int icell = 0;
	std::vector<char> marked(m.nelem, 0);
	while(icell < m.nelem-1) {
		marked[icell] = 1;
		for(int j = 0; j < m.nnode[icell]; j++) {
			const int o = m.esuel[icell*4+j];
			if(o >= 0 && o < m.nelem) marked[o] = 1;
		}
		if(marked[icell+1]) {
			levels.push_back(icell+1);
			std::fill(marked.begin(), marked.end(), 0);
		}
		icell++;
	}
	levels.push_back(m.nelem);
	return levels;
}

}
