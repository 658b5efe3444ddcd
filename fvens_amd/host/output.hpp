/* output.hpp -- solution and convergence output of a steady case.
 *
 * - postprocess_point: cell states averaged to the mesh points with cell-area weights, then
 *   density, Mach number, pressure, temperature and velocity per point
 *   (FlowOutput::postprocess_point, src/spatial/aoutput.cpp:97-147). The reference adds each
 *   cell's area once per conserved variable to the weight sum, which divides its point density
 *   and pressure by four; here the weight is added once per cell and point, so the point values
 *   are the area-weighted averages themselves (Mach number, temperature and velocity are the
 *   same either way).
 * - write_vtu_point_data: VTK XML unstructured grid with those point data
 *   (writeScalarsVectorToVtu_PointData, src/spatial/aoutput.cpp:427-551).
 * - convergence history: the column layout of writeConvergenceHistoryHeader /
 *   writeStepToConvergenceHistory (src/spatial/aoutput.cpp:618-638).
 * - surface files of FlowOutput::exportSurfaceData (src/spatial/aoutput.cpp:181-300):
 *   <prefix>-surf_w<marker>.out with x, y, Cp, Cf per wall face and the marker's CL, CDp, CDf,
 *   and <prefix>-surf_o<marker>.out with x, y and the host cell's velocity per face of the
 *   other listed markers. Faces are written in boundary-face order of the mesh.
 */
#ifndef FVENS_GPU_OUTPUT_HPP
#define FVENS_GPU_OUTPUT_HPP

#include "fvens_gpu.hpp"
#include <algorithm>
#include <cmath>
#include <fstream>
#include <iomanip>
#include <ostream>

namespace fvens_gpu {

struct PointData {
	std::vector<double> density, mach, pressure, temperature;   ///< [npoin]
	std::vector<double> velocity;                                ///< [npoin][2]
};

/// u: host conserved states [nelem][4]; area: cell areas [nelem].
inline PointData postprocess_point(const MeshData& m, const std::vector<double>& u,
                                   const std::vector<double>& area, double gamma, double Minf)
{
	const int np = m.npoin(), ne = m.nelem();
	std::vector<double> up(4*(size_t)np, 0.0), wsum(np, 0.0);
	for(int e = 0; e < ne; e++)
		for(int k = 0; k < m.nnode[e]; k++) {
			const int p = m.inpoel[4*(size_t)e + k];
			for(int v = 0; v < 4; v++) up[4*(size_t)p + v] += u[4*(size_t)e + v]*area[e];
			wsum[p] += area[e];
		}
	PointData o;
	o.density.resize(np); o.mach.resize(np); o.pressure.resize(np); o.temperature.resize(np);
	o.velocity.resize(2*(size_t)np);
	for(int p = 0; p < np; p++) {
		double q[4];
		for(int v = 0; v < 4; v++) q[v] = wsum[p] > 0 ? up[4*(size_t)p + v]/wsum[p] : 0.0;
		const double vx = q[1]/q[0], vy = q[2]/q[0];
		const double pr = (gamma - 1.0)*(q[3] - 0.5*q[0]*(vx*vx + vy*vy));
		const double c = std::sqrt(gamma*pr/q[0]);
		o.density[p] = q[0];
		o.velocity[2*(size_t)p] = vx;
		o.velocity[2*(size_t)p + 1] = vy;
		o.pressure[p] = pr;
		o.mach[p] = std::sqrt(vx*vx + vy*vy)/c;
		// non-dimensional temperature T = gamma M_inf^2 p / rho (aphysics_defs.hpp)
		o.temperature[p] = gamma*Minf*Minf*pr/q[0];
	}
	return o;
}

inline void write_vtu_point_data(const std::string& fname, const MeshData& m, const PointData& d)
{
	std::ofstream out(fname);
	if(!out) throw std::runtime_error("cannot write " + fname);
	out << std::setprecision(10);
	const int np = m.npoin(), ne = m.nelem();
	out << "<VTKFile type=\"UnstructuredGrid\" version=\"0.1\" byte_order=\"LittleEndian\">\n"
	    << "<UnstructuredGrid>\n"
	    << "\t<Piece NumberOfPoints=\"" << np << "\" NumberOfCells=\"" << ne << "\">\n"
	    << "\t\t<PointData Scalars=\"density\" Vectors=\"velocity\">\n";
	const std::pair<const char*, const std::vector<double>*> sc[] = {
		{"density", &d.density}, {"mach-number", &d.mach}, {"pressure", &d.pressure},
		{"temperature", &d.temperature}};
	for(const auto& s : sc) {
		out << "\t\t\t<DataArray type=\"Float64\" Name=\"" << s.first << "\" Format=\"ascii\">\n";
		for(int p = 0; p < np; p++) out << "\t\t\t\t" << (*s.second)[p] << '\n';
		out << "\t\t\t</DataArray>\n";
	}
	out << "\t\t\t<DataArray type=\"Float64\" Name=\"velocity\" NumberOfComponents=\"3\" "
	       "Format=\"ascii\">\n";
	for(int p = 0; p < np; p++)
		out << "\t\t\t\t" << d.velocity[2*(size_t)p] << ' ' << d.velocity[2*(size_t)p + 1]
		    << " 0.0\n";
	out << "\t\t\t</DataArray>\n\t\t</PointData>\n"
	    << "\t\t<Points>\n\t\t<DataArray type=\"Float64\" NumberOfComponents=\"3\" "
	       "Format=\"ascii\">\n";
	for(int p = 0; p < np; p++)
		out << "\t\t\t" << m.coords[2*(size_t)p] << ' ' << m.coords[2*(size_t)p + 1] << " 0.0\n";
	out << "\t\t</DataArray>\n\t\t</Points>\n\t\t<Cells>\n"
	    << "\t\t\t<DataArray type=\"UInt32\" Name=\"connectivity\" Format=\"ascii\">\n";
	for(int e = 0; e < ne; e++) {
		out << "\t\t\t\t";
		for(int k = 0; k < m.nnode[e]; k++) out << m.inpoel[4*(size_t)e + k] << ' ';
		out << '\n';
	}
	out << "\t\t\t</DataArray>\n"
	    << "\t\t\t<DataArray type=\"UInt32\" Name=\"offsets\" Format=\"ascii\">\n";
	long off = 0;
	for(int e = 0; e < ne; e++) { off += m.nnode[e]; out << "\t\t\t\t" << off << '\n'; }
	out << "\t\t\t</DataArray>\n"
	    << "\t\t\t<DataArray type=\"Int32\" Name=\"types\" Format=\"ascii\">\n";
	for(int e = 0; e < ne; e++) out << "\t\t\t\t" << (m.nnode[e] == 4 ? 9 : 5) << '\n';
	out << "\t\t\t</DataArray>\n\t\t</Cells>\n\t</Piece>\n</UnstructuredGrid>\n</VTKFile>";
}

/// Wall-marker surface file; rows: computeSurfaceData's [nfaces][4] (x, y, Cp, Cf).
inline void write_wall_surface_file(const std::string& prefix, int marker,
                                    const std::vector<double>& rows,
                                    const FlowSolutionFunctionals& f)
{
	const std::string fname = prefix + "-surf_w" + std::to_string(marker) + ".out";
	std::ofstream out(fname);
	if(!out) throw std::runtime_error("cannot write " + fname);
	out << "#  x \t y \t Cp  \t Cf \n";
	for(size_t i = 0; i < rows.size()/4; i++) {
		for(int j = 0; j < 4; j++) out << "  " << rows[4*i + j];
		out << '\n';
	}
	out << "# Cl      Cdp      Cdf\n";
	out << "# " << f.CL << "  " << f.CDp << "  " << f.CDsf << '\n';
}

/// Host cell of each boundary face: the cell having both of the face's points.
inline std::vector<int> boundary_host_cells(const MeshData& m)
{
	std::vector<std::vector<int>> cells_of(m.npoin());
	for(int e = 0; e < m.nelem(); e++)
		for(int k = 0; k < m.nnode[e]; k++) cells_of[m.inpoel[4*(size_t)e + k]].push_back(e);
	std::vector<int> host(m.nbface(), -1);
	for(int b = 0; b < m.nbface(); b++) {
		const std::vector<int>& a = cells_of[m.bface[2*(size_t)b]];
		const std::vector<int>& c = cells_of[m.bface[2*(size_t)b + 1]];
		for(int e : a)
			if(std::find(c.begin(), c.end(), e) != c.end()) { host[b] = e; break; }
		if(host[b] < 0) throw std::runtime_error("boundary face without a host cell");
	}
	return host;
}

/// Other-marker surface file: face centre and velocity of the host cell. u: host [nelem][4].
inline void write_other_surface_file(const std::string& prefix, int marker, const MeshData& m,
                                     const std::vector<int>& host, const std::vector<double>& u)
{
	const std::string fname = prefix + "-surf_o" + std::to_string(marker) + ".out";
	std::ofstream out(fname);
	if(!out) throw std::runtime_error("cannot write " + fname);
	out << "#   x         y          u           v\n";
	for(int b = 0; b < m.nbface(); b++) {
		if(m.btag[b] != marker) continue;
		const int p0 = m.bface[2*(size_t)b], p1 = m.bface[2*(size_t)b + 1];
		const double* w = &u[4*(size_t)host[b]];
		out << "  " << 0.5*(m.coords[2*(size_t)p0] + m.coords[2*(size_t)p1])
		    << "  " << 0.5*(m.coords[2*(size_t)p0 + 1] + m.coords[2*(size_t)p1 + 1])
		    << "  " << w[1]/w[0] << "  " << w[2]/w[0] << '\n';
	}
}

inline void write_convergence_history_header(std::ostream& o)
{
	using std::setw;
	o << '#' << setw(6) << "NStep" << setw(14) << "Log rel resi" << setw(14) << "Log abs resi"
	  << setw(12) << "Tot.Wtime" << setw(12) << "Lin.Wtime" << setw(12) << "Lin.iters"
	  << setw(10) << "CFL" << '\n';
	o << "#" << std::string(82, '-') << '\n';
}

inline void write_convergence_step(std::ostream& o, const fvx_step_record& r)
{
	using std::setw;
	o << std::setprecision(6) << setw(7) << r.step << setw(14) << std::log10(r.rel) << setw(14)
	  << std::log10(r.abs) << std::setprecision(4) << setw(12) << r.walltime << setw(12)
	  << r.lin_walltime << setw(12) << r.lin_iters << setw(10) << r.cfl << '\n';
}

}
#endif
