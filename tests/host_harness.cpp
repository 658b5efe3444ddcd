// Test driver for the header-only host code (built by tests/test_host_formats.py with g++):
//   host_harness ctrl FILE        parsed control-file options, one "key value" per line
//   host_harness mesh FILE        mesh sizes and checksums
//   host_harness vtu MESH OUT     VTU file of a synthetic state on the mesh
#include "control_file.hpp"
#include "output.hpp"
#include "su2_reader.hpp"
#include <cstdio>
#include <cstdlib>
#include <cstring>

using namespace fvens_gpu;

static double shoelace(const MeshData& m, int e)
{
	double a = 0;
	const int n = m.nnode[e];
	for(int k = 0; k < n; k++) {
		const int p = m.inpoel[4*(size_t)e + k], q = m.inpoel[4*(size_t)e + (k + 1) % n];
		a += m.coords[2*(size_t)p]*m.coords[2*(size_t)q + 1]
		     - m.coords[2*(size_t)q]*m.coords[2*(size_t)p + 1];
	}
	return 0.5*a;
}

int main(int argc, char** argv)
{
	if(argc < 3) return 2;
	try {
		if(!std::strcmp(argv[1], "ctrl")) {
			const FlowParserOptions o = parse_flow_controlfile(read_info_file(argv[2]));
			std::printf("meshfile %s\nvtu %s\nlogfile %s\nlognres %d\nflowtype %s\n",
			            o.meshfile.c_str(), o.vtu_output_file.c_str(), o.logfile.c_str(),
			            (int)o.lognres, o.flowtype.c_str());
			std::printf("gamma %.17g\nalpha %.17g\nMinf %.17g\nviscsim %d\nTinf %.17g\n"
			            "Reinf %.17g\nPr %.17g\nconstvisc %d\n", o.gamma, o.alpha, o.Minf,
			            (int)o.viscsim, o.Tinf, o.Reinf, o.Pr, (int)o.useconstvisc);
			for(size_t i = 0; i < o.bcconf.size(); i++) {
				std::printf("bc%zu %d %d", i, (int)o.bcconf[i].bc_type, o.bcconf[i].bc_tag);
				for(double v : o.bcconf[i].bc_vals) std::printf(" %.17g", v);
				std::printf("\n");
			}
			std::printf("lwalls");
			for(int w : o.lwalls) std::printf(" %d", w);
			std::printf("\nlothers");
			for(int w : o.lothers) std::printf(" %d", w);
			std::printf("\nsurfprefix %s\nsimtype %s\ninvflux %s\ninvfluxjac %s\ngradient %s\n"
			            "order2 %d\nlimiter %s\nlimparam %.17g\npseudotime %s\n",
			            o.surfnameprefix.c_str(), o.sim_type.c_str(), o.invflux.c_str(),
			            o.invfluxjac.c_str(), o.gradientmethod.c_str(), (int)o.order2,
			            o.limiter.c_str(), o.limiter_param, o.pseudotimetype.c_str());
			std::printf("main %.17g %.17g %.17g %d\nstarter %d %.17g %.17g %.17g %d\n"
			            "nlupdate %s %.17g\n", o.initcfl, o.endcfl, o.tolerance, o.maxiter,
			            (int)o.usestarter, o.firstinitcfl, o.firstendcfl, o.firsttolerance,
			            o.firstmaxiter, o.nl_update_scheme.c_str(), o.min_nl_update);
			return 0;
		}
		if(!std::strcmp(argv[1], "mesh")) {
			const MeshData m = read_mesh(argv[2]);
			double cs = 0;
			long is = 0;
			for(double c : m.coords) cs += c;
			for(int v : m.inpoel) is += v;
			for(int v : m.bface) is += 3*v;
			for(int t : m.btag) is += 7*t;
			std::printf("%d %d %d %.17g %ld\n", m.npoin(), m.nelem(), m.nbface(), cs, is);
			return 0;
		}
		if(!std::strcmp(argv[1], "vtu") && argc > 3) {
			const MeshData m = read_mesh(argv[2]);
			const int ne = m.nelem();
			std::vector<double> u(4*(size_t)ne), area(ne);
			for(int e = 0; e < ne; e++) {
				area[e] = shoelace(m, e);
				const double rho = 1.0 + 0.001*(e % 17), vx = 0.5, vy = 0.01*(e % 5);
				const double p = 1.0/(1.4*0.25);
				u[4*(size_t)e] = rho;
				u[4*(size_t)e + 1] = rho*vx;
				u[4*(size_t)e + 2] = rho*vy;
				u[4*(size_t)e + 3] = p/0.4 + 0.5*rho*(vx*vx + vy*vy);
			}
			write_vtu_point_data(argv[3], m, postprocess_point(m, u, area, 1.4, 0.5));
			return 0;
		}
		if(!std::strcmp(argv[1], "msh") && argc > 3) {
			// any readable mesh written back as Gmsh 2.2
			write_gmsh2(argv[3], read_mesh(argv[2]));
			return 0;
		}
		if(!std::strcmp(argv[1], "surf") && argc > 4) {
			// surface file of the "other" marker argv[4] under prefix argv[3], for the state
			// rho = 1 + 0.001 (e mod 17), velocity (0.5, 0.01 (e mod 5))
			const MeshData m = read_mesh(argv[2]);
			const int ne = m.nelem();
			std::vector<double> u(4*(size_t)ne);
			for(int e = 0; e < ne; e++) {
				const double rho = 1.0 + 0.001*(e % 17);
				u[4*(size_t)e] = rho;
				u[4*(size_t)e + 1] = rho*0.5;
				u[4*(size_t)e + 2] = rho*0.01*(e % 5);
				u[4*(size_t)e + 3] = 2.5;
			}
			const std::vector<int> host = boundary_host_cells(m);
			write_other_surface_file(argv[3], std::atoi(argv[4]), m, host, u);
			for(int h : host) std::printf("%d\n", h);
			return 0;
		}
	}
	catch(const std::exception& e) { std::printf("error: %s\n", e.what()); return 1; }
	return 2;
}
