/* fvx_case.cpp -- steady flow case driver in C++ over the host adapter.
 *
 * Control-file mode follows the reference's steady-case executable (tests/flow_solve.cpp with
 * SteadyFlowCase, src/utilities/casesolvers.cpp:57-163, 225-419): parse the control file
 * (control_file.hpp), read the mesh (Gmsh 2.2 or SU2 by extension), start from the freestream,
 * run the first-order starter solve when the file has one (a Tolerance_error there is reported
 * and ignored), run the main solve (implicit or explicit), report a main solve that does not
 * converge and go on, then write the point-data VTU file, the residual history
 * (<log_file_prefix>-residual_history.log) when convergence_history_required is true, the
 * surface files of each listed wall and other marker (<surface_output_file_prefix>-surf_w<m>.out,
 * -surf_o<m>.out; FlowOutput::exportSurfaceData, src/spatial/aoutput.cpp:181-300), and print
 * the functionals on the first output wall marker. --regression_file checks them like
 * tests/flow_solve.cpp:89-123 (CL to 100 x tol, CDp and CDsf to tol, relative).
 *
 *   fvx_case FILE.ctrl [--mesh_file F] [--regression_file F] [--regression_tol 1e-8]
 *                      [--pc sgs] [--sweeps 16] [--precision f64|f32|bf16] [--matrix_free]
 *                      [--matrix_free_difference_step 1e-7]
 *                      [--rcm] [--ksp_max_it 30] [--ksp_rtol 0.1] [--limiter_param 0]
 *                      [--mesh_reorder rcm|line|line_natural|line_rcm|hilbert]
 *                      [--mesh_anisotropy_threshold T]
 *
 * The preconditioner, its sweeps and storage precision, matrix-free operation and the cell
 * ordering stand in for the PETSc options of the reference (-pc_type, -matrix_free_jacobian,
 * -mesh_reorder with -mesh_anisotropy_threshold; --rcm is --mesh_reorder rcm). The limiter parameter (WENO central weight, Venkatakrishnan K) is
 * zero as in the reference's runs unless --limiter_param gives it.
 *
 * Flag mode (no control file) keeps the inviscid airfoil case of
 * testcases/naca0012/transonic-sanity-test-muscl.ctrl:
 *   fvx_case --mesh FILE [--mach 0.8] [--aoa 1.25] [--flux HLLC] [--gradient LEASTSQUARES]
 *            [--limiter VANALBADA] [--limiter-param 20] [--pc sgs] [--sweeps 1]
 *            [--precision f64|f32|bf16] [--tol 1e-7] [--maxiter 170] [--rcm]
 *            [--regression FILE] [--regression-tol 1e-8]
 */
#include "fvens_gpu.hpp"
#include "control_file.hpp"
#include "gmsh_reader.hpp"
#include "output.hpp"
#include "su2_reader.hpp"
#include <hip/hip_runtime.h>
#include <cstdio>
#include <cstring>
#include <fstream>
#include <map>

using namespace fvens_gpu;

namespace {

struct DeviceState {
	double* p = nullptr;
	explicit DeviceState(const std::vector<double>& h)
	{
		if(hipMalloc((void**)&p, h.size()*sizeof(double)) != hipSuccess)
			throw std::runtime_error("hipMalloc failed");
		if(hipMemcpy(p, h.data(), h.size()*sizeof(double), hipMemcpyHostToDevice) != hipSuccess)
			throw std::runtime_error("hipMemcpy failed");
	}
	~DeviceState() { if(p) (void)hipFree(p); }
	std::vector<double> host(size_t n) const
	{
		std::vector<double> h(n);
		if(hipMemcpy(h.data(), p, n*sizeof(double), hipMemcpyDeviceToHost) != hipSuccess)
			throw std::runtime_error("hipMemcpy failed");
		return h;
	}
};

int regression(const std::string& file, double tol, const FlowSolutionFunctionals& f)
{
	std::FILE* rf = std::fopen(file.c_str(), "r");
	if(!rf) throw std::runtime_error("cannot open regression file " + file);
	double v[3];
	const int got = std::fscanf(rf, "%lf %lf %lf", &v[0], &v[1], &v[2]);
	std::fclose(rf);
	if(got < 2) throw std::runtime_error("bad regression file");
	const double eCL = std::fabs(v[0] - f.CL)/std::fabs(v[0]);
	const double eCD = std::fabs(v[1] - f.CDp)/std::fabs(v[1]);
	const double eSF = got > 2 ? std::fabs(v[2] - f.CDsf)/std::fabs(v[2]) : 0.0;
	std::printf("regression: CL error %.3e (tol %.1e), CDp error %.3e, CDsf error %.3e (tol %.1e)\n",
	            eCL, 100*tol, eCD, eSF, tol);
	if(eCL > 100*tol || eCD > tol || eSF > tol) { std::printf("REGRESSION FAILED\n"); return 1; }
	std::printf("REGRESSION PASSED\n");
	return 0;
}

int run_control_file(const std::string& ctrl, std::map<std::string, std::string>& opt, bool rcm,
                     bool mf)
{
	FlowParserOptions o = parse_flow_controlfile(read_info_file(ctrl));
	if(opt.count("limiter_param")) o.limiter_param = std::stod(opt["limiter_param"]);
	const std::string meshfile = opt.count("mesh_file") ? opt["mesh_file"] : o.meshfile;
	MeshData md = read_mesh(meshfile);
	if(rcm) md.reorder_rcm();
	else if(opt.count("mesh_reorder"))
		md.reorder(opt["mesh_reorder"], opt.count("mesh_anisotropy_threshold")
		                                ? std::stod(opt["mesh_anisotropy_threshold"]) : 0.0);
	UMesh m(md);
	std::printf("mesh %s: %d cells, %d faces\n", meshfile.c_str(), m.gnelem(), m.gnaface());

	const FlowPhysicsConfig pconf = extract_spatial_physics_config(o);
	FlowFV prob(&m, pconf, extract_spatial_numerics_config(o));
	// f64 | f32 | bf16 (bfloat16 preconditioner factors over an FP32 Jacobian)
	const int prec = opt["precision"] == "f64" ? FVX_PREC_F64 : FVX_PREC_F32;
	const int pc_prec = opt["precision"] == "bf16" ? FVX_PREC_BF16 : prec;
	const size_t n = 4*(size_t)m.gnelem();
	std::vector<double> uh(n);
	const auto uinf = prob.compute_freestream_state();
	for(size_t i = 0; i < n; i++) uh[i] = uinf[i % 4];
	DeviceState u(uh);

	SteadySolverConfig sc;
	sc.pc_type = opt["pc"];
	sc.pc_sweeps = std::stoi(opt["sweeps"]);
	sc.pc_precision = pc_prec;
	sc.jac_precision = prec;
	sc.nl_update_scheme = o.nl_update_scheme.empty() ? "ROBUST_FLOW" : o.nl_update_scheme;
	sc.min_nl_update = o.min_nl_update;
	sc.ksp_max_it = std::stoi(opt["ksp_max_it"]);
	sc.ksp_rtol = std::stod(opt["ksp_rtol"]);
	const bool implicit = o.pseudotimetype == "IMPLICIT";
	if(o.usestarter) {
		FlowFV startprob(&m, pconf, firstorder_spatial_numerics_config(o));
		SteadySolverConfig s1 = sc;
		s1.cflinit = o.firstinitcfl; s1.cflfin = o.firstendcfl;
		s1.tol = o.firsttolerance; s1.maxiter = o.firstmaxiter;
		SteadySolver starter(&startprob, s1, implicit);
		try { starter.solve(u.p); }
		catch(const Tolerance_error& e) { std::printf("starter: %s\n", e.what()); }
		std::printf("starter: %d steps\n", starter.getTimingData().num_timesteps);
	}
	sc.cflinit = o.initcfl; sc.cflfin = o.endcfl;
	sc.tol = o.tolerance; sc.maxiter = o.maxiter;
	sc.matrix_free = mf;
	if(opt.count("matrix_free_difference_step"))
		sc.matrix_free_difference_step = std::stod(opt["matrix_free_difference_step"]);
	SteadySolver solver(&prob, sc, implicit);
	try { solver.solve(u.p); }
	catch(const Tolerance_error& e) { std::printf("main: %s\n", e.what()); }
	const TimingData td = solver.getTimingData();
	std::printf("main: %d steps, %d linear iterations, %.4f s, rel %.3e\n", td.num_timesteps,
	            td.total_lin_iters, td.ode_walltime, td.final_rel);

	if(o.lognres) {
		const std::vector<fvx_step_record> h = prob.solve_history();
		std::ofstream conv(o.logfile + "-residual_history.log");
		write_convergence_history_header(conv);
		for(const auto& r : h) write_convergence_step(conv, r);
	}
	const std::vector<double> uf = u.host(n);
	if(!o.vtu_output_file.empty()) {
		const PointData pd = postprocess_point(md, uf, m.areas(), pconf.gamma, pconf.Minf);
		write_vtu_point_data(o.vtu_output_file, md, pd);
	}
	for(int w : o.lwalls) {
		const FlowSolutionFunctionals fw = prob.computeFunctionals(u.p, w);
		write_wall_surface_file(o.surfnameprefix, w, prob.computeSurfaceData(u.p, w), fw);
		std::printf("FlowOutput: CL = %g   CDp = %g    CDf = %g\n", fw.CL, fw.CDp, fw.CDsf);
	}
	if(!o.lothers.empty()) {
		const std::vector<int> host = boundary_host_cells(md);
		for(int w : o.lothers) write_other_surface_file(o.surfnameprefix, w, md, host, uf);
	}
	const int wall = o.lwalls.empty() ? 2 : o.lwalls[0];
	const FlowSolutionFunctionals f = prob.computeFunctionals(u.p, wall);
	std::printf("CL %.15e CDp %.15e CDsf %.15e entropy %.15e\n", f.CL, f.CDp, f.CDsf, f.entropy);
	if(opt.count("regression_file"))
		return regression(opt["regression_file"], std::stod(opt["regression_tol"]), f);
	return 0;
}

int run_flags(std::map<std::string, std::string>& opt, bool rcm)
{
	if(!opt.count("mesh")) { std::fprintf(stderr, "--mesh is required\n"); return 2; }
	MeshData md = read_mesh(opt["mesh"]);
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

	// f64 | f32 | bf16 (bfloat16 preconditioner factors over an FP32 Jacobian)
	const int prec = opt["precision"] == "f64" ? FVX_PREC_F64 : FVX_PREC_F32;
	const int pc_prec = opt["precision"] == "bf16" ? FVX_PREC_BF16 : prec;
	const size_t n = 4*(size_t)m.gnelem();
	std::vector<double> uh(n);
	const auto uinf = prob.compute_freestream_state();
	for(size_t i = 0; i < n; i++) uh[i] = uinf[i % 4];
	DeviceState u(uh);

	SteadySolverConfig sc;
	sc.pc_type = opt["pc"];
	sc.pc_sweeps = std::stoi(opt["sweeps"]);
	sc.pc_precision = pc_prec;
	sc.jac_precision = prec;
	sc.cflinit = 50; sc.cflfin = 1000; sc.tol = 1e-1; sc.maxiter = 20;
	SteadyBackwardEulerSolver starter(&startprob, sc);
	try { starter.solve(u.p); }
	catch(const Tolerance_error& e) { std::printf("starter: %s\n", e.what()); }
	sc.cflinit = 500; sc.cflfin = 5000;
	sc.tol = std::stod(opt["tol"]);
	sc.maxiter = std::stoi(opt["maxiter"]);
	SteadyBackwardEulerSolver solver(&prob, sc);
	solver.solve(u.p);
	const TimingData td = solver.getTimingData();
	const FlowSolutionFunctionals f = prob.computeFunctionals(u.p, 2);
	std::printf("steps %d lin_iters %d walltime %.4f s rel %.3e\n", td.num_timesteps,
	            td.total_lin_iters, td.ode_walltime, td.final_rel);
	std::printf("CL %.15e CDp %.15e CDsf %.15e entropy %.15e\n", f.CL, f.CDp, f.CDsf,
	            f.entropy);
	if(opt.count("regression"))
		return regression(opt["regression"], std::stod(opt["regression-tol"]), f);
	return 0;
}

}

int main(int argc, char** argv)
{
	std::map<std::string, std::string> opt = {
		{"mach", "0.8"}, {"aoa", "1.25"}, {"flux", "HLLC"}, {"gradient", "LEASTSQUARES"},
		{"limiter", "VANALBADA"}, {"limiter-param", "20"}, {"pc", "sgs"}, {"sweeps", "1"},
		{"precision", "f64"}, {"tol", "1e-7"}, {"maxiter", "170"}, {"regression-tol", "1e-8"},
		{"regression_tol", "1e-8"}, {"ksp_max_it", "30"}, {"ksp_rtol", "0.1"}};
	bool rcm = false, mf = false;
	std::string ctrl;
	int first = 1;
	if(argc > 1 && std::strncmp(argv[1], "--", 2) != 0) { ctrl = argv[1]; first = 2; }
	if(!ctrl.empty()) opt["sweeps"] = "16";
	for(int i = first; i < argc; i++) {
		std::string a = argv[i];
		if(a == "--rcm") { rcm = true; continue; }
		if(a == "--matrix_free") { mf = true; continue; }
		if(a.rfind("--", 0) != 0 || i + 1 >= argc) {
			std::fprintf(stderr, "bad argument %s\n", a.c_str());
			return 2;
		}
		opt[a.substr(2)] = argv[++i];
	}
	if(opt["precision"] != "f64" && opt["precision"] != "f32" && opt["precision"] != "bf16") {
		std::fprintf(stderr, "--precision takes f64, f32 or bf16\n");
		return 2;
	}
	try {
		return ctrl.empty() ? run_flags(opt, rcm) : run_control_file(ctrl, opt, rcm, mf);
	}
	catch(const Numerical_error& e) { std::printf("Numerical_error: %s\n", e.what()); return 3; }
	catch(const Tolerance_error& e) { std::printf("Tolerance_error: %s\n", e.what()); return 4; }
	catch(const std::exception& e) { std::printf("error: %s\n", e.what()); return 1; }
}
