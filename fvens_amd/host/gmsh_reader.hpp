/* gmsh_reader.hpp -- Gmsh 2.2 ASCII reader and writer for linear triangle / quad meshes.
 *
 * Same reading rules as the reference reader (src/mesh/meshreaders.cpp:66-266): node
 * coordinates with z dropped, linear edges (type 1) are boundary faces whose first tag is the
 * physical marker, triangles (type 2) and quads (type 3) are cells, points (type 15) are
 * skipped; node numbers in the file are 1-based and must be contiguous.
 *
 * write_gmsh2 writes the layout of UMesh::writeGmsh2 (src/mesh/mesh.cpp:204-281): nodes with
 * z = 0 at 20 significant digits, then the boundary faces (physical marker, elementary tag
 * 1 + marker) and the cells (tags 1 1), numbered from 1 in that order.
 */
#ifndef FVENS_GPU_GMSH_READER_HPP
#define FVENS_GPU_GMSH_READER_HPP

#include "fvens_gpu.hpp"
#include <fstream>
#include <iomanip>
#include <sstream>

namespace fvens_gpu {

inline MeshData read_gmsh2(const std::string& path)
{
	std::ifstream f(path);
	if(!f) throw std::runtime_error("cannot open mesh file " + path);
	MeshData m;
	std::string line;
	bool have_nodes = false;
	while(std::getline(f, line)) {
		if(line.rfind("$Nodes", 0) == 0) {
			int npoin;
			f >> npoin;
			m.coords.resize(2*(size_t)npoin);
			for(int i = 0; i < npoin; i++) {
				int id;
				double x, y, z;
				f >> id >> x >> y >> z;
				if(id != i + 1) throw std::runtime_error("node numbering must be contiguous from 1");
				m.coords[2*i] = x;
				m.coords[2*i+1] = y;
			}
			have_nodes = true;
		}
		else if(line.rfind("$Elements", 0) == 0) {
			int nelm;
			f >> nelm;
			std::getline(f, line);
			for(int e = 0; e < nelm; e++) {
				std::getline(f, line);
				std::istringstream is(line);
				int id, type, ntags;
				is >> id >> type >> ntags;
				std::vector<int> tags(ntags);
				for(int& t : tags) is >> t;
				std::vector<int> nodes;
				int v;
				while(is >> v) nodes.push_back(v - 1);
				if(type == 1) {
					m.bface.push_back(nodes.at(0));
					m.bface.push_back(nodes.at(1));
					m.btag.push_back(ntags > 0 ? tags[0] : 0);
				} else if(type == 2) {
					for(int k = 0; k < 3; k++) m.inpoel.push_back(nodes.at(k));
					m.inpoel.push_back(-1);
					m.nnode.push_back(3);
				} else if(type == 3) {
					for(int k = 0; k < 4; k++) m.inpoel.push_back(nodes.at(k));
					m.nnode.push_back(4);
				} else if(type != 15) {
					throw std::runtime_error("unsupported Gmsh element type " + std::to_string(type));
				}
			}
		}
	}
	if(!have_nodes) throw std::runtime_error("no $Nodes section in " + path);
	return m;
}

inline void write_gmsh2(const std::string& path, const MeshData& m)
{
	std::ofstream out(path);
	if(!out) throw std::runtime_error("cannot write " + path);
	out << std::setprecision(20);
	out << "$MeshFormat\n2.2 0 8\n$EndMeshFormat\n$Nodes\n" << m.npoin() << '\n';
	for(int p = 0; p < m.npoin(); p++)
		out << p + 1 << ' ' << m.coords[2*(size_t)p] << ' ' << m.coords[2*(size_t)p + 1] << ' '
		    << 0.0 << '\n';
	out << "$EndNodes\n$Elements\n" << m.nbface() + m.nelem() << '\n';
	for(int b = 0; b < m.nbface(); b++)
		out << b + 1 << " 1 2 " << m.btag[b] << ' ' << 1 + m.btag[b] << ' '
		    << m.bface[2*(size_t)b] + 1 << ' ' << m.bface[2*(size_t)b + 1] + 1 << '\n';
	for(int e = 0; e < m.nelem(); e++) {
		out << m.nbface() + e + 1 << ' ' << (m.nnode[e] == 3 ? 2 : 3) << " 2 1 1";
		for(int k = 0; k < m.nnode[e]; k++) out << ' ' << m.inpoel[4*(size_t)e + k] + 1;
		out << '\n';
	}
	out << "$EndElements\n";
}

}
#endif
