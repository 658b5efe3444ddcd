/* su2_reader.hpp -- SU2 ASCII reader for linear triangle / quad meshes.
 *
 * Same reading rules as the reference reader (src/mesh/meshreaders.cpp:269-397): NDIME=2, NELEM=
 * cells as VTK type (5 triangle, 9 quad) with 0-based node numbers and a trailing index, NPOIN=
 * nodes (x, y, index), NMARK= markers, each MARKER_TAG= (an integer physical marker) and
 * MARKER_ELEMS= line elements (type 3). Boundary faces are kept in marker order. read_mesh picks
 * the format by file name like the reference's readMesh (src/mesh/meshreaders.cpp:35-61).
 */
#ifndef FVENS_GPU_SU2_READER_HPP
#define FVENS_GPU_SU2_READER_HPP

#include "fvens_gpu.hpp"
#include "gmsh_reader.hpp"
#include <fstream>
#include <sstream>

namespace fvens_gpu {

namespace detail {
/// Next non-empty, non-comment line of an SU2 file.
inline std::string su2_line(std::istream& f)
{
	std::string line;
	while(std::getline(f, line)) {
		const size_t b = line.find_first_not_of(" \t\r");
		if(b == std::string::npos || line[b] == '%') continue;
		return line;
	}
	throw std::runtime_error("unexpected end of SU2 file");
}

/// Value of a "KEY= value" line.
inline std::string su2_value(const std::string& line, const char* key)
{
	const size_t eq = line.find('=');
	std::string k = line.substr(0, eq);
	k.erase(k.find_last_not_of(" \t") + 1);
	if(eq == std::string::npos || k != key)
		throw std::runtime_error(std::string("SU2: expected ") + key + "= but read " + line);
	return line.substr(eq + 1);
}
}

inline MeshData read_su2(const std::string& path)
{
	std::ifstream f(path);
	if(!f) throw std::runtime_error("cannot open mesh file " + path);
	MeshData m;
	if(std::stoi(detail::su2_value(detail::su2_line(f), "NDIME")) != 2)
		throw std::runtime_error("only two-dimensional SU2 meshes are supported");
	const int nelem = std::stoi(detail::su2_value(detail::su2_line(f), "NELEM"));
	m.inpoel.assign(4*(size_t)nelem, -1);
	m.nnode.assign(nelem, 0);
	for(int i = 0; i < nelem; i++) {
		std::istringstream is(detail::su2_line(f));
		int type;
		is >> type;
		const int nn = type == 5 ? 3 : type == 9 ? 4 : 0;
		if(nn == 0) throw std::runtime_error("unsupported SU2 element type " + std::to_string(type));
		m.nnode[i] = nn;
		for(int k = 0; k < nn; k++) is >> m.inpoel[4*(size_t)i + k];
	}
	const int npoin = std::stoi(detail::su2_value(detail::su2_line(f), "NPOIN"));
	m.coords.resize(2*(size_t)npoin);
	for(int i = 0; i < npoin; i++) {
		std::istringstream is(detail::su2_line(f));
		is >> m.coords[2*i] >> m.coords[2*i + 1];
	}
	const int nmark = std::stoi(detail::su2_value(detail::su2_line(f), "NMARK"));
	for(int b = 0; b < nmark; b++) {
		const int tag = std::stoi(detail::su2_value(detail::su2_line(f), "MARKER_TAG"));
		const int nf = std::stoi(detail::su2_value(detail::su2_line(f), "MARKER_ELEMS"));
		for(int k = 0; k < nf; k++) {
			std::istringstream is(detail::su2_line(f));
			int type, a, c;
			is >> type >> a >> c;
			if(type != 3) throw std::runtime_error("SU2 boundary elements must be lines (type 3)");
			m.bface.push_back(a);
			m.bface.push_back(c);
			m.btag.push_back(tag);
		}
	}
	return m;
}

inline MeshData read_mesh(const std::string& path)
{
	const size_t dot = path.rfind('.');
	if(dot != std::string::npos && path.substr(dot + 1) == "su2") return read_su2(path);
	return read_gmsh2(path);
}

}
#endif
