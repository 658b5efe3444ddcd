/* fvx_case.cpp -- steady flow case driver in C++ over the host adapter.
 *
 * Follows SteadyFlowCase (src/utilities/casesolvers.cpp:225-386) for the inviscid airfoil case
 * of testcases/naca0012/transonic-sanity-test-muscl.ctrl: read the Gmsh mesh, solve a
 * first-order starter problem (a Tolerance_error there is reported and ignored, as the
 * reference does), then the second-order main problem, and print the functionals. With
 * --regression FILE it checks CL / CDp against a reference regression file the way
 * tests/flow_solve.cpp:89-123 does and exits non-zero on a mismatch.
 *
 * usage: fvx_case --mesh FILE [--mach 0.8] [--aoa 1.25] [--flux HLLC]
 *                 [--gradient LEASTSQUARES] [--limiter VANALBADA] [--limiter-param 20]
 *                 [--pc sgs] [--sweeps 1] [--precision f64|f32] [--tol 1e-7] [--maxiter 170]
 *                 [--rcm] [--regression FILE] [--regression-tol 1e-8]
 */
#include "fvens_gpu.hpp"
#include "gmsh_reader.hpp"
#include <hip/hip_runtime.h>
#include <cstdio>
#include <cstring>
#include <map>

using namespace fvens_gpu;

int main(int argc, char** argv)
{
	std::map<std::string, std::string> opt = {
		{"mach", "0.8"}, {"aoa", "1.25"}, {"flux", "HLLC"}, {"gradient", "LEASTSQUARES"},
		{"limiter", "VANALBADA"}, {"limiter-param", "20"}, {"pc", "sgs"}, {"sweeps", "1"},
		{"precision", "f64"}, {"tol", "1e-7"}, {"maxiter", "170"}, {"regression-tol", "1e-8"}};
	bool rcm = false;
	for(int i = 1; i < argc; i++) {
		std::string a = argv[i];
		if(a == "--rcm") { rcm = true; continue; }
		if(a.rfind("--", 0) != 0 || i + 1 >= argc) {
			std::fprintf(stderr, "bad argument %s\n", a.c_str());
			return 2;
		}
		opt[a.substr(2)] = argv[++i];
	}
	if(!opt.count("mesh")) { std::fprintf(stderr, "--mesh is required\n"); return 2; }
	try {
		MeshData md = read_gmsh2(opt["mesh"]);
		if(rcm) md.reorder_rcm();
		UMesh m(md);
		std::printf("mesh: %d cells, %d faces\n", m.gnelem(), m.gnaface());

		FlowPhysicsConfig pconf{1.4, std::stod(opt["mach"]), 288.15, 1e6, 0.72,
		                        std::stod(opt["aoa"])*M_PI/180.0, false, false,
		                        {{2, SLIP_WALL_BC, {}, {}}, {4, FARFIELD_BC, {}, {}}}};
		const FlowNumericsConfig start{opt["flux"], opt["flux"], "NONE", "NONE", 1.0, false};
		const FlowNumericsConfig main{opt["flux"], opt["flux"], opt["gradient"], opt["limiter"],
		                              std::stod(opt["limiter-param"]), true};
		FlowFV startprob(&m, pconf, start), prob(&m, pconf, main);

		const int prec = opt["precision"] == "f32" ? FVX_PREC_F32 : FVX_PREC_F64;
		const size_t n = 4*(size_t)m.gnelem();
		std::vector<double> uh(n);
		const auto uinf = prob.compute_freestream_state();
		for(size_t i = 0; i < n; i++) uh[i] = uinf[i % 4];
		double* u = nullptr;
		if(hipMalloc((void**)&u, n*sizeof(double)) != hipSuccess)
			throw std::runtime_error("hipMalloc failed");
		if(hipMemcpy(u, uh.data(), n*sizeof(double), hipMemcpyHostToDevice) != hipSuccess)
			throw std::runtime_error("hipMemcpy failed");

		SteadySolverConfig sc;
		sc.pc_type = opt["pc"];
		sc.pc_sweeps = std::stoi(opt["sweeps"]);
		sc.pc_precision = sc.jac_precision = prec;
		sc.cflinit = 50; sc.cflfin = 1000; sc.tol = 1e-1; sc.maxiter = 20;
		SteadyBackwardEulerSolver starter(&startprob, sc);
		try { starter.solve(u); }
		catch(const Tolerance_error& e) { std::printf("starter: %s\n", e.what()); }
		sc.cflinit = 500; sc.cflfin = 5000;
		sc.tol = std::stod(opt["tol"]);
		sc.maxiter = std::stoi(opt["maxiter"]);
		SteadyBackwardEulerSolver solver(&prob, sc);
		solver.solve(u);
		const TimingData td = solver.getTimingData();
		const FlowSolutionFunctionals f = prob.computeFunctionals(u, 2);
		std::printf("steps %d lin_iters %d walltime %.4f s rel %.3e\n", td.num_timesteps,
		            td.total_lin_iters, td.ode_walltime, td.final_rel);
		std::printf("CL %.15e CDp %.15e CDsf %.15e entropy %.15e\n", f.CL, f.CDp, f.CDsf,
		            f.entropy);
		(void)hipFree(u);

		if(opt.count("regression")) {
			std::FILE* rf = std::fopen(opt["regression"].c_str(), "r");
			if(!rf) throw std::runtime_error("cannot open regression file");
			double CL, CDp;
			const int got = std::fscanf(rf, "%lf %lf", &CL, &CDp);
			std::fclose(rf);
			if(got != 2) throw std::runtime_error("bad regression file");
			const double tol = std::stod(opt["regression-tol"]);
			const double eCL = std::fabs(CL - f.CL)/std::fabs(CL);
			const double eCD = std::fabs(CDp - f.CDp)/std::fabs(CDp);
			std::printf("regression: CL error %.3e (tol %.1e), CDp error %.3e (tol %.1e)\n",
			            eCL, 100*tol, eCD, tol);
			if(eCL > 100*tol || eCD > tol) { std::printf("REGRESSION FAILED\n"); return 1; }
			std::printf("REGRESSION PASSED\n");
		}
	}
	catch(const Numerical_error& e) { std::printf("Numerical_error: %s\n", e.what()); return 3; }
	catch(const Tolerance_error& e) { std::printf("Tolerance_error: %s\n", e.what()); return 4; }
	catch(const std::exception& e) { std::printf("error: %s\n", e.what()); return 1; }
	return 0;
}
