/* gmsh_reader.hpp -- Gmsh 2.2 ASCII reader for linear triangle / quad meshes.
 *
 * Same reading rules as the reference reader (src/mesh/meshreaders.cpp:66-266): node
 * coordinates with z dropped, linear edges (type 1) are boundary faces whose first tag is the
 * physical marker, triangles (type 2) and quads (type 3) are cells, points (type 15) are
 * skipped; node numbers in the file are 1-based and must be contiguous.
 */
#ifndef FVENS_GPU_GMSH_READER_HPP
#define FVENS_GPU_GMSH_READER_HPP

#include "fvens_gpu.hpp"
#include <fstream>
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

}
#endif
