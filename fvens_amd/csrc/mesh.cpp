/* mesh.cpp -- host-side mesh preprocessing for libfvx: topology, geometry, RCM and line
 * orderings, and the per-cell stencil coefficients consumed by the device kernels.
 *
 * Topology and numbering follow UMesh (src/mesh/mesh.cpp): faces are numbered physical
 * boundary first (input order, orientation corrected to point out of the domain, :54-82 and
 * :683-700), then interior faces in element order with the lower cell on the left
 * (:709-733). Geometry: normals n = (dy, -dx)/len (:346-365), shoelace areas with quads as
 * two triangles (:290-313), vertex-average cell centres (:316-328), face midpoints and ghost
 * centres reflected about the midpoint (src/spatial/aspatial.cpp:50-119).
 */
#include "fvx_internal.hpp"
#include "wave.hpp"
#include <algorithm>
#include <cmath>
#include <numeric>
#include <queue>
#include <tuple>

namespace fvx {

namespace {

struct EdgeRec {
	long long key;
	int elem, local;
	bool operator<(const EdgeRec& o) const {
		return std::tie(key, elem, local) < std::tie(o.key, o.elem, o.local);
	}
};

long long ekey(int a, int b)
{
	const long long lo = std::min(a, b), hi = std::max(a, b);
	return (lo << 32) | hi;
}

void check_desc(const fvx_mesh_desc& d)
{
	if(d.npoin <= 0 || d.nelem <= 0 || d.nbface < 0 || !d.coords || !d.inpoel || !d.nnode
	   || (d.nbface > 0 && (!d.bface || !d.btag)))
		throw Error(FVX_ERR_INVALID, "mesh description has missing arrays or bad sizes");
	for(int e = 0; e < d.nelem; e++) {
		const int nn = d.nnode[e];
		if(nn != 3 && nn != 4)
			throw Error(FVX_ERR_INVALID, "cells must be triangles or quadrilaterals");
		for(int j = 0; j < nn; j++) {
			const int p = d.inpoel[e*4+j];
			if(p < 0 || p >= d.npoin) throw Error(FVX_ERR_INVALID, "cell node out of range");
		}
	}
	for(int b = 0; b < d.nbface; b++)
		for(int j = 0; j < 2; j++) {
			const int p = d.bface[b*2+j];
			if(p < 0 || p >= d.npoin) throw Error(FVX_ERR_INVALID, "boundary node out of range");
		}
}

/// Sorted list of all cell edges; matching keys are the cell pairs sharing an edge.
std::vector<EdgeRec> cell_edges(const fvx_mesh_desc& d)
{
	std::vector<EdgeRec> e;
	e.reserve((size_t)d.nelem*4);
	for(int c = 0; c < d.nelem; c++) {
		const int nn = d.nnode[c];
		for(int j = 0; j < nn; j++)
			e.push_back({ekey(d.inpoel[c*4+j], d.inpoel[c*4+(j+1)%nn]), c, j});
	}
	std::sort(e.begin(), e.end());
	return e;
}

std::vector<int> neighbours(const fvx_mesh_desc& d, const std::vector<EdgeRec>& e)
{
	std::vector<int> nb((size_t)d.nelem*4, -1);
	for(size_t i = 0; i < e.size(); ) {
		size_t k = i + 1;
		while(k < e.size() && e[k].key == e[i].key) k++;
		if(k - i == 2) {
			nb[(size_t)e[i].elem*4 + e[i].local] = e[i+1].elem;
			nb[(size_t)e[i+1].elem*4 + e[i+1].local] = e[i].elem;
		}
		else if(k - i > 2)
			throw Error(FVX_ERR_INVALID, "non-manifold edge shared by more than two cells");
		i = k;
	}
	return nb;
}

}

void build_mesh(fvx_mesh_s& m, const fvx_mesh_desc& d)
{
	check_desc(d);
	m.npoin = d.npoin; m.nelem = d.nelem; m.nbface = d.nbface;
	m.coords.assign(d.coords, d.coords + (size_t)d.npoin*2);
	m.inpoel.assign(d.inpoel, d.inpoel + (size_t)d.nelem*4);
	m.nnode.assign(d.nnode, d.nnode + d.nelem);
	m.bface.assign(d.bface, d.bface + (size_t)d.nbface*2);
	m.btag.assign(d.btag, d.btag + d.nbface);

	const std::vector<EdgeRec> edges = cell_edges(d);
	m.esuel = neighbours(d, edges);
	m.elemface.assign((size_t)m.nelem*4, -1);

	// boundary faces: host cell by edge lookup, orientation as the host cell's edge
	std::vector<int> host(m.nbface), hostj(m.nbface);
	for(int b = 0; b < m.nbface; b++) {
		const long long k = ekey(m.bface[b*2], m.bface[b*2+1]);
		auto it = std::lower_bound(edges.begin(), edges.end(), EdgeRec{k, -1, -1});
		if(it == edges.end() || it->key != k)
			throw Error(FVX_ERR_INVALID, "boundary face is not an edge of any cell");
		if(m.esuel[(size_t)it->elem*4 + it->local] != -1)
			throw Error(FVX_ERR_INVALID, "boundary face lies between two cells");
		host[b] = it->elem; hostj[b] = it->local;
		const int nn = m.nnode[it->elem];
		m.bface[b*2] = m.inpoel[it->elem*4 + it->local];
		m.bface[b*2+1] = m.inpoel[it->elem*4 + (it->local+1)%nn];
	}

	m.ninface = 0;
	for(int c = 0; c < m.nelem; c++)
		for(int j = 0; j < m.nnode[c]; j++) {
			const int o = m.esuel[(size_t)c*4+j];
			if(o > c) m.ninface++;
		}
	m.naface = m.nbface + m.ninface;
	m.intfac.assign((size_t)m.naface*4, -1);
	for(int b = 0; b < m.nbface; b++) {
		m.intfac[b*4+0] = host[b];
		m.intfac[b*4+1] = m.nelem + b;
		m.intfac[b*4+2] = m.bface[b*2];
		m.intfac[b*4+3] = m.bface[b*2+1];
		m.esuel[(size_t)host[b]*4 + hostj[b]] = m.nelem + b;
		m.elemface[(size_t)host[b]*4 + hostj[b]] = b;
	}
	for(int c = 0; c < m.nelem; c++)
		for(int j = 0; j < m.nnode[c]; j++)
			if(m.esuel[(size_t)c*4+j] == -1)
				throw Error(FVX_ERR_INVALID, "open cell edge without a boundary face");
	int f = m.nbface;
	for(int c = 0; c < m.nelem; c++) {
		const int nn = m.nnode[c];
		for(int j = 0; j < nn; j++) {
			const int o = m.esuel[(size_t)c*4+j];
			if(o > c && o < m.nelem) {
				m.intfac[(size_t)f*4+0] = c;
				m.intfac[(size_t)f*4+1] = o;
				m.intfac[(size_t)f*4+2] = m.inpoel[c*4+j];
				m.intfac[(size_t)f*4+3] = m.inpoel[c*4+(j+1)%nn];
				m.elemface[(size_t)c*4+j] = f;
				for(int k = 0; k < m.nnode[o]; k++)
					if(m.esuel[(size_t)o*4+k] == c) m.elemface[(size_t)o*4+k] = f;
				f++;
			}
		}
	}

	const double* X = m.coords.data();
	m.facemetric.assign((size_t)m.naface*3, 0.0);
	m.gr.assign((size_t)m.naface*2, 0.0);
	for(int i = 0; i < m.naface; i++) {
		const int a = m.intfac[(size_t)i*4+2], b = m.intfac[(size_t)i*4+3];
		const double nx = X[b*2+1] - X[a*2+1], ny = -(X[b*2] - X[a*2]);
		const double len = std::sqrt(nx*nx + ny*ny);
		m.facemetric[(size_t)i*3] = nx/len;
		m.facemetric[(size_t)i*3+1] = ny/len;
		m.facemetric[(size_t)i*3+2] = len;
		m.gr[(size_t)i*2] = (X[a*2] + X[b*2])/2;
		m.gr[(size_t)i*2+1] = (X[a*2+1] + X[b*2+1])/2;
	}
	auto tri = [&](int p, int q, int r) {
		return 0.5*(X[p*2]*(X[q*2+1] - X[r*2+1]) - X[p*2+1]*(X[q*2] - X[r*2])
		            + X[q*2]*X[r*2+1] - X[r*2]*X[q*2+1]);
	};
	m.area.assign(m.nelem, 0.0);
	m.rc.assign((size_t)m.nelem*2, 0.0);
	m.hlong.assign(m.nelem, 0.0);
	for(int c = 0; c < m.nelem; c++) {
		const int* in = &m.inpoel[c*4];
		const int nn = m.nnode[c];
		m.area[c] = nn == 3 ? tri(in[0], in[1], in[2]) : tri(in[0], in[1], in[2]) + tri(in[0], in[2], in[3]);
		if(!(m.area[c] > 0)) throw Error(FVX_ERR_INVALID, "cell with non-positive area (not CCW?)");
		double sx = 0, sy = 0, h2 = 0;
		for(int j = 0; j < nn; j++) {
			sx += X[in[j]*2]; sy += X[in[j]*2+1];
			const int q = in[(j+1)%nn];
			h2 = std::max(h2, (X[in[j]*2]-X[q*2])*(X[in[j]*2]-X[q*2]) + (X[in[j]*2+1]-X[q*2+1])*(X[in[j]*2+1]-X[q*2+1]));
		}
		m.rc[c*2] = sx/nn; m.rc[c*2+1] = sy/nn;
		m.hlong[c] = std::sqrt(h2);
	}
	m.rcbp.assign((size_t)m.nbface*2, 0.0);
	for(int b = 0; b < m.nbface; b++)
		for(int k = 0; k < 2; k++)
			m.rcbp[b*2+k] = 2.0*m.gr[b*2+k] - m.rc[host[b]*2+k];

	// device arrays: per-cell slots and face records
	std::vector<int> nbr((size_t)m.nelem*4, SLOT_UNUSED), fidx((size_t)m.nelem*4, 0);
	for(int c = 0; c < m.nelem; c++)
		for(int j = 0; j < m.nnode[c]; j++) {
			const int o = m.esuel[(size_t)c*4+j];
			nbr[(size_t)c*4+j] = o < m.nelem ? o : -1 - (o - m.nelem);
			fidx[(size_t)c*4+j] = m.elemface[(size_t)c*4+j];
			// the kernels find boundary face ib's record at index ib
			if(o >= m.nelem && fidx[(size_t)c*4+j] != o - m.nelem)
				throw Error(FVX_ERR_INVALID, "boundary faces are not numbered first");
		}
	std::vector<FaceGeom> fg(m.naface);
	for(int i = 0; i < m.naface; i++)
		fg[i] = FaceGeom{m.facemetric[(size_t)i*3], m.facemetric[(size_t)i*3+1],
		                 m.facemetric[(size_t)i*3+2], m.gr[(size_t)i*2], m.gr[(size_t)i*2+1],
		                 m.intfac[(size_t)i*4], i < m.nbface ? -1 - i : m.intfac[(size_t)i*4+1]};
	// per slot, the offset of the face midpoint from the cell centre, contiguous per cell (the
	// limiter's extrapolation distances; the same differences it would form from the face
	// records, without their scattered reads)
	std::vector<double> soff((size_t)m.nelem*8, 0.0);
	for(int c = 0; c < m.nelem; c++)
		for(int j = 0; j < m.nnode[c]; j++) {
			const FaceGeom& G = fg[fidx[(size_t)c*4+j]];
			soff[(size_t)c*8+2*j] = G.mx - m.rc[c*2];
			soff[(size_t)c*8+2*j+1] = G.my - m.rc[c*2+1];
		}
	m.d_nbr.upload(nbr);
	m.d_face.upload(fidx);
	m.d_fgeom.upload(fg);
	m.d_soff.upload(soff);
	m.d_centre.upload(m.rc);
	m.d_area.upload(m.area);
	m.d_bhost.upload(host);
	m.d_btag.upload(m.btag);
	build_wave_tasks(m);
}

/* Tables of the wave-task flux kernel (flux_wave.hpp). Cells are taken in order into groups of
 * at most 64 while the faces touching the group make at most 64 tasks: a cell adds one task
 * per face to the boundary, to a later cell or to a cell outside the group, none for a face to
 * an earlier cell of the group (that face's task then serves both cells). A task is its face's
 * index (the face record holds the two cells); a cell's slot word holds, in byte k, 2 task +
 * side for its local face k (side 1: the cell is the face's right cell), 0xff when unused. */
void build_wave_tasks(fvx_mesh_s& m)
{
	const int n = m.nelem;
	std::vector<int> tasks;
	std::vector<int2> groups;
	std::vector<unsigned> slots;
	int c = 0;
	while(c < n) {
		const int c0 = c;
		int ntask = 0;
		while(c < n && c - c0 < 64) {
			int add = 0;
			for(int k = 0; k < m.nnode[c]; k++) {
				const int o = m.esuel[(size_t)c*4 + k];
				if(!(o < n && o >= c0 && o < c)) add++;
			}
			if(ntask + add > 64) break;
			ntask += add;
			c++;
		}
		const int c1 = c, g = (int)groups.size();
		groups.push_back(make_int2(c0, c1 - c0));
		const size_t tb = tasks.size();
		tasks.resize(tb + 64, -1);
		slots.resize(slots.size() + 64, 0xffffffffu);
		unsigned* sl = &slots[(size_t)g*64];
		auto set_code = [&](int cell, int k, int task, int side) {
			unsigned& w = sl[cell - c0];
			w = (w & ~(0xffu << 8*k)) | ((unsigned)(2*task + side) << 8*k);
		};
		int nt = 0;
		for(int cc = c0; cc < c1; cc++)
			for(int k = 0; k < m.nnode[cc]; k++) {
				const int o = m.esuel[(size_t)cc*4 + k];
				const int f = m.elemface[(size_t)cc*4 + k];
				if(o >= c0 && o < c1) {
					if(o < cc) continue;             // the earlier cell made the task
					int ko = -1;
					for(int j = 0; j < m.nnode[o]; j++)
						if(m.elemface[(size_t)o*4 + j] == f) ko = j;
					if(ko < 0) throw Error(FVX_ERR_INVALID, "face missing from its right cell");
					set_code(o, ko, nt, 1);
				}
				// a boundary face's left cell is its interior cell; otherwise left = lower index
				tasks[tb + nt] = f;
				set_code(cc, k, nt, o < n && o < cc ? 1 : 0);
				nt++;
			}
		if(nt != ntask) throw Error(FVX_ERR_INVALID, "wave-task count mismatch");
	}
	m.nwgroup = (int)groups.size();
	m.d_wtask.upload(tasks);
	m.d_wgroup.upload(groups);
	m.d_wslot.upload(slots);
}

/* Lines of strongly coupled cells (findLines, src/mesh/meshordering.cpp:199-262). Each cell's
 * real neighbours j carry the weight 1/|x_j - x_i| (vertex-average centres) divided by the
 * smallest of them, in decreasing order (ties keep local-face order). From the host cell of
 * every boundary face, in boundary-face order, a line grows while the current cell's largest
 * weight exceeds the threshold, stepping to the first free neighbour (in weight order) whose
 * weight exceeds it; lines of one cell are dropped. celline[c] is the line of cell c or -1. */
void find_lines(const fvx_mesh_desc& d, double threshold, std::vector<std::vector<int>>& lines,
                std::vector<int>& celline)
{
	check_desc(d);
	const int n = d.nelem;
	const std::vector<EdgeRec> edges = cell_edges(d);
	const std::vector<int> nb = neighbours(d, edges);
	std::vector<double> cx(n), cy(n);
	for(int c = 0; c < n; c++) {
		double sx = 0, sy = 0;
		for(int j = 0; j < d.nnode[c]; j++) {
			sx += d.coords[2*d.inpoel[c*4+j]];
			sy += d.coords[2*d.inpoel[c*4+j]+1];
		}
		cx[c] = sx/d.nnode[c]; cy[c] = sy/d.nnode[c];
	}
	// per cell: (weight, local face) of real neighbours, decreasing weight
	std::vector<double> wt((size_t)n*4, 0.0);
	std::vector<int> wface((size_t)n*4, -1), nreal(n, 0);
	for(int c = 0; c < n; c++) {
		std::pair<double, int> e[4];
		int k = 0;
		double minw = 1e20;
		for(int j = 0; j < d.nnode[c]; j++) {
			const int o = nb[(size_t)c*4+j];
			if(o < 0) continue;
			const double dx = cx[c] - cx[o], dy = cy[c] - cy[o];
			e[k] = {1.0/std::sqrt(dx*dx + dy*dy), j};
			minw = std::min(minw, e[k].first);
			k++;
		}
		for(int j = 0; j < k; j++) e[j].first /= minw;
		std::stable_sort(e, e + k, [](const std::pair<double, int>& a,
		                              const std::pair<double, int>& b) { return a.first > b.first; });
		nreal[c] = k;
		for(int j = 0; j < k; j++) { wt[(size_t)c*4+j] = e[j].first; wface[(size_t)c*4+j] = e[j].second; }
	}
	lines.clear();
	celline.assign(n, -1);
	for(int b = 0; b < d.nbface; b++) {
		const long long key = ekey(d.bface[b*2], d.bface[b*2+1]);
		auto it = std::lower_bound(edges.begin(), edges.end(), EdgeRec{key, -1, -1});
		if(it == edges.end() || it->key != key)
			throw Error(FVX_ERR_INVALID, "boundary face is not an edge of any cell");
		int cur = it->elem;
		if(celline[cur] >= 0) continue;
		std::vector<int> line;
		for(;;) {
			if(!(nreal[cur] > 0 && wt[(size_t)cur*4] > threshold)) break;
			line.push_back(cur);
			celline[cur] = (int)lines.size();
			int next = -1;
			for(int j = 0; j < nreal[cur]; j++) {
				const int o = nb[(size_t)cur*4 + wface[(size_t)cur*4+j]];
				if(celline[o] == -1 && wt[(size_t)cur*4+j] > threshold) { next = o; break; }
			}
			if(next < 0) break;
			cur = next;
		}
		if(line.size() > 1) lines.push_back(line);
		else if(line.size() == 1) celline[line[0]] = -1;
	}
}

namespace {

/* Reverse Cuthill-McKee of SPARSPAK (GENRCM with FNROOT, ROOTLS, DEGREE and RCM; George and
 * Liu, Computer Solution of Large Sparse Positive Definite Systems, 1981, ch. 4), which is the
 * "rcm" ordering PETSc's MatGetOrdering applies to a matrix graph. adj lists each vertex's
 * neighbours in increasing order (the vertex itself may appear; it changes no decision).
 * Returns perm[new] = old. */
struct Graph {
	std::vector<int> ptr, adj;
};

void rootls(const Graph& g, int root, std::vector<char>& mask, int& nlvl, std::vector<int>& xls,
            int* ls)
{
	mask[root] = 0;
	ls[0] = root;
	nlvl = 0;
	int lvlend = 0, ccsize = 1;
	xls.clear();
	do {
		const int lbegin = lvlend;
		lvlend = ccsize;
		xls.push_back(lbegin);
		nlvl++;
		for(int i = lbegin; i < lvlend; i++) {
			const int node = ls[i];
			for(int k = g.ptr[node]; k < g.ptr[node+1]; k++) {
				const int nbr = g.adj[k];
				if(!mask[nbr]) continue;
				ls[ccsize++] = nbr;
				mask[nbr] = 0;
			}
		}
	} while(ccsize > lvlend);
	xls.push_back(lvlend);
	for(int i = 0; i < ccsize; i++) mask[ls[i]] = 1;
}

void fnroot(const Graph& g, int& root, std::vector<char>& mask, std::vector<int>& xls, int* ls)
{
	int nlvl;
	rootls(g, root, mask, nlvl, xls, ls);
	const int ccsize = xls[nlvl];
	if(nlvl == 1 || nlvl == ccsize) return;
	for(;;) {
		const int jstrt = xls[nlvl-1];
		int mindeg = ccsize;
		root = ls[jstrt];
		if(ccsize - 1 != jstrt) {
			for(int j = jstrt; j < ccsize; j++) {
				const int node = ls[j];
				int ndeg = 0;
				for(int k = g.ptr[node]; k < g.ptr[node+1]; k++) if(mask[g.adj[k]]) ndeg++;
				if(ndeg >= mindeg) continue;
				root = node;
				mindeg = ndeg;
			}
		}
		int nunlvl;
		rootls(g, root, mask, nunlvl, xls, ls);
		if(nunlvl <= nlvl) return;
		nlvl = nunlvl;
		if(nlvl >= ccsize) return;
	}
}

void degree(const Graph& g, int root, const std::vector<char>& mask, std::vector<int>& deg,
            std::vector<char>& queued, int& ccsize, int* ls)
{
	ls[0] = root;
	queued[root] = 1;
	ccsize = 1;
	int lvlend = 0;
	do {
		const int lbegin = lvlend;
		lvlend = ccsize;
		for(int i = lbegin; i < lvlend; i++) {
			const int node = ls[i];
			int ideg = 0;
			for(int k = g.ptr[node]; k < g.ptr[node+1]; k++) {
				const int nbr = g.adj[k];
				if(!mask[nbr]) continue;
				ideg++;
				if(queued[nbr]) continue;
				queued[nbr] = 1;
				ls[ccsize++] = nbr;
			}
			deg[node] = ideg;
		}
	} while(ccsize > lvlend);
	for(int i = 0; i < ccsize; i++) queued[ls[i]] = 0;
}

void rcm_component(const Graph& g, int root, std::vector<char>& mask, int* perm, int& ccsize,
                   std::vector<int>& deg, std::vector<char>& queued)
{
	degree(g, root, mask, deg, queued, ccsize, perm);
	mask[root] = 0;
	if(ccsize <= 1) return;
	int lvlend = 0, lnbr = 1;
	while(lnbr > lvlend) {
		const int lbegin = lvlend;
		lvlend = lnbr;
		for(int i = lbegin; i < lvlend; i++) {
			const int node = perm[i];
			const int fnbr = lnbr;
			for(int k = g.ptr[node]; k < g.ptr[node+1]; k++) {
				const int nbr = g.adj[k];
				if(!mask[nbr]) continue;
				mask[nbr] = 0;
				perm[lnbr++] = nbr;
			}
			// the new neighbours by increasing degree (linear insertion, stable)
			for(int k = fnbr + 1; k < lnbr; k++) {
				const int v = perm[k];
				int l = k - 1;
				while(l >= fnbr && deg[perm[l]] > deg[v]) { perm[l+1] = perm[l]; l--; }
				perm[l+1] = v;
			}
		}
	}
	std::reverse(perm, perm + ccsize);
}

std::vector<int> genrcm(const Graph& g)
{
	const int n = (int)g.ptr.size() - 1;
	std::vector<char> mask(n, 1), queued(n, 0);
	std::vector<int> perm(n), xls, deg(n, 0);
	int num = 0;
	for(int i = 0; i < n && num < n; i++) {
		if(!mask[i]) continue;
		int root = i, ccsize = 0;
		fnroot(g, root, mask, xls, perm.data() + num);
		rcm_component(g, root, mask, perm.data() + num, ccsize, deg, queued);
		num += ccsize;
	}
	return perm;
}

}

/* Reverse Cuthill-McKee of the cell graph (a vertex per cell, edges between face neighbours,
 * the pattern of the one-variable Jacobian that reorderMeshPetsc orders,
 * src/mesh/ameshutils.cpp:245-288), by the same GENRCM as PETSc's "rcm". */
void rcm_order(const fvx_mesh_desc& d, int* perm)
{
	check_desc(d);
	const int n = d.nelem;
	const std::vector<int> nb = neighbours(d, cell_edges(d));
	Graph g;
	g.ptr.assign(1, 0);
	for(int c = 0; c < n; c++) {
		int row[5], k = 0;
		row[k++] = c;
		for(int j = 0; j < 4; j++) if(nb[(size_t)c*4+j] >= 0) row[k++] = nb[(size_t)c*4+j];
		std::sort(row, row + k);
		g.adj.insert(g.adj.end(), row, row + k);
		g.ptr.push_back((int)g.adj.size());
	}
	const std::vector<int> p = genrcm(g);
	std::copy(p.begin(), p.end(), perm);
}

/* Cells sorted along a Hilbert curve through their vertex-average centres (a 2^30 x 2^30
 * grid over the bounding square; ties keep cell order). Consecutive cells are spatially close,
 * so a block of consecutive cells is a compact patch whose faces mostly join two of its own
 * cells. No reference counterpart. */
void hilbert_order(const fvx_mesh_desc& d, int* perm)
{
	check_desc(d);
	const int n = d.nelem;
	std::vector<double> cx(n), cy(n);
	double lo = 1e300, hi = -1e300;
	for(int c = 0; c < n; c++) {
		double sx = 0, sy = 0;
		for(int j = 0; j < d.nnode[c]; j++) {
			sx += d.coords[2*d.inpoel[c*4+j]];
			sy += d.coords[2*d.inpoel[c*4+j]+1];
		}
		cx[c] = sx/d.nnode[c]; cy[c] = sy/d.nnode[c];
		lo = std::min(lo, std::min(cx[c], cy[c]));
		hi = std::max(hi, std::max(cx[c], cy[c]));
	}
	const int order = 30;
	const long long top = (1LL << order) - 1;
	const double span = hi > lo ? hi - lo : 1.0;
	std::vector<std::pair<long long, int>> key(n);
	for(int c = 0; c < n; c++) {
		long long x = std::min(top, std::max(0LL, (long long)((cx[c] - lo)/span*(double)top)));
		long long y = std::min(top, std::max(0LL, (long long)((cy[c] - lo)/span*(double)top)));
		// position along the curve: per level, the quadrant's offset, then the rotation /
		// reflection that maps the sub-square onto the curve's base orientation
		long long h = 0;
		for(long long q = 1LL << (order - 1); q > 0; q >>= 1) {
			const long long rx = (x & q) ? 1 : 0, ry = (y & q) ? 1 : 0;
			h += q*q*((3*rx) ^ ry);
			if(ry == 0) {
				if(rx == 1) { x = q - 1 - (x & (q - 1)); y = q - 1 - (y & (q - 1)); }
				else { x &= q - 1; y &= q - 1; }
				std::swap(x, y);
			} else { x &= q - 1; y &= q - 1; }
		}
		key[c] = {h, c};
	}
	std::stable_sort(key.begin(), key.end());
	for(int i = 0; i < n; i++) perm[i] = key[i].second;
}

/* Cell order of lineReorder / getHybridLineOrdering (src/mesh/meshordering.cpp:33-128,
 * 264-412). graph_ordering 0: the lines in order, then the other cells in index order.
 * 1: the same through the line-point graph in natural order. 2: the line-point graph (a vertex
 * per line, then one per remaining cell; edges between vertices holding adjacent cells)
 * ordered by reverse Cuthill-McKee, each line vertex expanded to its cells. */
void line_order(const fvx_mesh_desc& d, double threshold, int graph_ordering, int* perm)
{
	std::vector<std::vector<int>> lines;
	std::vector<int> celline;
	find_lines(d, threshold, lines, celline);
	const int n = d.nelem, nl = (int)lines.size();
	std::vector<int> points, pidx(n, -1);
	for(int c = 0; c < n; c++) if(celline[c] < 0) { pidx[c] = nl + (int)points.size(); points.push_back(c); }
	const int nv = nl + (int)points.size();
	std::vector<int> vorder(nv);
	std::iota(vorder.begin(), vorder.end(), 0);
	if(graph_ordering == 2) {
		const std::vector<int> nb = neighbours(d, cell_edges(d));
		auto vert = [&](int c) { return celline[c] >= 0 ? celline[c] : pidx[c]; };
		std::vector<std::vector<int>> rows(nv);
		for(int v = 0; v < nv; v++) {
			const bool isline = v < nl;
			const std::vector<int> cells = isline ? lines[v] : std::vector<int>{points[v - nl]};
			rows[v].push_back(v);
			for(int c : cells)
				for(int j = 0; j < d.nnode[c]; j++) {
					const int o = nb[(size_t)c*4+j];
					if(o >= 0) rows[v].push_back(vert(o));
				}
			std::sort(rows[v].begin(), rows[v].end());
			rows[v].erase(std::unique(rows[v].begin(), rows[v].end()), rows[v].end());
		}
		Graph g;
		g.ptr.assign(1, 0);
		for(int v = 0; v < nv; v++) {
			g.adj.insert(g.adj.end(), rows[v].begin(), rows[v].end());
			g.ptr.push_back((int)g.adj.size());
		}
		vorder = genrcm(g);
	}
	else if(graph_ordering != 0 && graph_ordering != 1)
		throw Error(FVX_ERR_INVALID, "unknown line-graph ordering");
	int k = 0;
	for(int v : vorder) {
		if(v < nl) for(int c : lines[v]) perm[k++] = c;
		else perm[k++] = points[v - nl];
	}
}

/* Stencil coefficients: both gradient schemes become grad_i = sum_j c_ij (w_j - w_i) over the
 * cell's face neighbours (ghost centre for boundary faces).
 *  - weighted least squares (src/spatial/agradientschemes.cpp:218-440):
 *        c_ij = w_ij V_i^{-1} (x_j - x_i),  w_ij = |x_j - x_i|^-2,  V_i = sum_j w_ij dx dx^T
 *  - Green-Gauss with inverse-distance face values (:61-214):
 *        c_ij = len_f n_out,f (d_j/(d_i+d_j)) / A_i,  d = 1/|x_f - x|
 *    (the u_i term carries sum_f len_f n_out,f = 0 for a closed cell). */
GasParams make_gas(const fvx_physics_cfg& p)
{
	GasParams g;
	g.g = p.gamma; g.gm1 = p.gamma - 1.0;
	g.Minf = p.Minf; g.Tinf = p.Tinf; g.Reinf = p.Reinf; g.Pr = p.Pr;
	g.gM2 = p.gamma*p.Minf*p.Minf;
	g.sutherland = 110.5/p.Tinf;
	g.pinf = 1.0/(p.gamma*p.Minf*p.Minf);
	g.uinf[0] = 1.0;
	g.uinf[1] = std::cos(p.aoa);
	g.uinf[2] = std::sin(p.aoa);
	g.uinf[3] = g.pinf/(p.gamma - 1.0) + 0.5;
	g.constvisc = p.const_visc;
	return g;
}

void build_disc(fvx_disc_s& d)
{
	const fvx_mesh_s& m = *d.m;
	const fvx_physics_cfg& p = d.pc;
	d.gas = make_gas(p);

	if(p.nbc < 0 || p.nbc > 8) throw Error(FVX_ERR_INVALID, "at most 8 boundary conditions");
	for(int k = 0; k < 8; k++) { d.bct.kind[k] = -1; d.bct.v0[k] = 0; d.bct.v1[k] = 0; }
	for(int k = 0; k < p.nbc; k++) {
		const int kind = p.bcs[k].kind;
		if(kind < 0 || kind > 7 || kind == FVX_BC_PERIODIC)
			throw Error(FVX_ERR_UNSUPPORTED, "boundary kind not supported");
		d.bct.kind[k] = kind; d.bct.v0[k] = p.bcs[k].vals[0]; d.bct.v1[k] = p.bcs[k].vals[1];
	}
	d.bcid.assign(m.nbface, -1);
	for(int b = 0; b < m.nbface; b++) {
		for(int k = 0; k < p.nbc; k++) if(p.bcs[k].tag == m.btag[b]) d.bcid[b] = k;
		if(d.bcid[b] < 0) throw Error(FVX_ERR_INVALID, "boundary marker without a boundary condition");
	}
	d.d_bcid.upload(d.bcid);
	d.d_gas.upload(std::vector<GasParams>(1, d.gas));
	d.d_bct.upload(std::vector<BCTable>(1, d.bct));

	const fvx_numerics_cfg& n = d.nc;
	if(n.conv_numflux < 0 || n.conv_numflux > 6 || n.conv_numflux_jac < 0 || n.conv_numflux_jac > 6)
		throw Error(FVX_ERR_INVALID, "unknown inviscid flux");
	if(n.gradientscheme < 0 || n.gradientscheme > 2 || n.reconstruction < 0 || n.reconstruction > 4)
		throw Error(FVX_ERR_INVALID, "unknown gradient scheme or reconstruction");
	if(n.reconstruction == FVX_RECON_WENO && !(n.limiter_param >= 0.0 && std::isfinite(n.limiter_param)))
		throw Error(FVX_ERR_INVALID, "WENO reconstruction needs a finite, non-negative central "
		            "weight (limiter parameter)");

	const int ne = m.nelem;
	std::vector<double> coef((size_t)ne*8, 0.0), eps2(ne, 0.0);
	for(int c = 0; c < ne; c++) {
		const double xi = m.rc[c*2], yi = m.rc[c*2+1];
		double xj[4], yj[4];
		int ns = m.nnode[c];
		for(int j = 0; j < ns; j++) {
			const int o = m.esuel[(size_t)c*4+j];
			if(o < ne) { xj[j] = m.rc[o*2]; yj[j] = m.rc[o*2+1]; }
			else { xj[j] = m.rcbp[(o-ne)*2]; yj[j] = m.rcbp[(o-ne)*2+1]; }
		}
		if(n.gradientscheme == FVX_GRAD_LEASTSQUARES) {
			double V[4] = {0, 0, 0, 0};
			for(int j = 0; j < ns; j++) {
				const double dx = xj[j]-xi, dy = yj[j]-yi, w = 1.0/(dx*dx + dy*dy);
				V[0] += w*dx*dx; V[1] += w*dx*dy; V[2] += w*dy*dx; V[3] += w*dy*dy;
			}
			const double det = V[0]*V[3] - V[1]*V[2];
			const double I[4] = { V[3]/det, -V[1]/det, -V[2]/det, V[0]/det };
			for(int j = 0; j < ns; j++) {
				const double dx = xj[j]-xi, dy = yj[j]-yi, w = 1.0/(dx*dx + dy*dy);
				coef[(size_t)c*8 + j*2]     = w*(I[0]*dx + I[1]*dy);
				coef[(size_t)c*8 + j*2 + 1] = w*(I[2]*dx + I[3]*dy);
			}
		}
		else if(n.gradientscheme == FVX_GRAD_GREENGAUSS) {
			for(int j = 0; j < ns; j++) {
				const int f = m.elemface[(size_t)c*4+j];
				const double s = m.intfac[(size_t)f*4] == c ? 1.0 : -1.0;
				const double mx = m.gr[f*2], my = m.gr[f*2+1];
				const double di = 1.0/std::sqrt((mx-xi)*(mx-xi) + (my-yi)*(my-yi));
				const double dj = 1.0/std::sqrt((mx-xj[j])*(mx-xj[j]) + (my-yj[j])*(my-yj[j]));
				const double wgt = dj/(di + dj) * m.facemetric[(size_t)f*3+2] / m.area[c] * s;
				coef[(size_t)c*8 + j*2]     = wgt*m.facemetric[(size_t)f*3];
				coef[(size_t)c*8 + j*2 + 1] = wgt*m.facemetric[(size_t)f*3+1];
			}
		}
		eps2[c] = std::pow(n.limiter_param*m.hlong[c], 3);
	}
	d.d_gcoef.upload(coef);
	d.d_eps2.upload(eps2);
	d.d_grad.alloc((size_t)ne*8);
	d.d_rgrad.alloc((size_t)ne*8);
	d.d_ughost.alloc((size_t)std::max(m.nbface, 1)*4);
	d.d_scratch.alloc(4096);
	// the wave-task flux form where it applies and the cell order packs at least 24 cells per
	// wave64 group on average (compact orders: Hilbert, structured natural orders); it spends
	// fewer than 2.7 face evaluations per cell there against the quad form's 4 lanes per cell
	d.tune.flux_wave = flux_wave_apply(d) && (long)m.nelem >= 24L*m.nwgroup;
}

}
