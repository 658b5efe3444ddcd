/* fvens_gpu.hpp -- C++ host adapter over the fvx C ABI (include/fvx.h).
 *
 * Mirrors the reference's C++ spatial-discretisation and steady-solver API so that host code
 * written against it keeps its shape:
 *   FlowPhysicsConfig / FlowNumericsConfig     src/spatial/flow_spatial.hpp:33-55
 *   FlowBCConfig                               src/spatial/abc.hpp:34-40
 *   FlowFV_base::compute_residual / getGradients / assemble_jacobian
 *                                              src/spatial/flow_spatial.hpp:61-120,
 *                                              src/spatial/aspatial.hpp:62-92
 *   SteadySolverConfig / SteadyBackwardEulerSolver / SteadyForwardEulerSolver / TimingData
 *                                              src/ode/aodesolver.hpp:18-160
 *   FlowSolutionFunctionals                    src/utilities/casesolvers.hpp:34-41
 *   Numerical_error / Tolerance_error          src/utilities/aerrorhandling.hpp:15-85
 * Differences: vectors are device pointers (double[ncell][4], the layout of the reference's
 * Vec) instead of PETSc Vecs, and the system matrix is an fvx block matrix instead of a Mat.
 * Names of fluxes, gradient schemes and limiters are the reference factory strings
 * (src/utilities/afactory.cpp:30-294).
 */
#ifndef FVENS_GPU_HPP
#define FVENS_GPU_HPP

#include "../../include/fvx.h"
#include <array>
#include <cmath>
#include <stdexcept>
#include <string>
#include <vector>

namespace fvens_gpu {

using StatusCode = int;

struct Numerical_error : std::runtime_error {
	explicit Numerical_error(const std::string& m) : std::runtime_error(m) {}
};
struct Tolerance_error : std::runtime_error {
	explicit Tolerance_error(const std::string& m) : std::runtime_error(m) {}
};

/// Converts an fvx status into the reference's exception types.
inline void check(fvx_status s)
{
	if(s == FVX_OK) return;
	const std::string msg = fvx_last_error();
	if(s == FVX_ERR_NONFINITE) throw Numerical_error(msg);
	if(s == FVX_ERR_NOT_CONVERGED) throw Tolerance_error(msg);
	if(s == FVX_ERR_INVALID || s == FVX_ERR_UNSUPPORTED) throw std::invalid_argument(msg);
	throw std::runtime_error(msg);
}

enum BCType { SLIP_WALL_BC = FVX_BC_SLIPWALL, FARFIELD_BC = FVX_BC_FARFIELD,
              INFLOW_OUTFLOW_BC = FVX_BC_INOUTFLOW, SUBSONIC_INFLOW_BC = FVX_BC_SUBSONIC_INFLOW,
              EXTRAPOLATION_BC = FVX_BC_EXTRAPOLATION, PERIODIC_BC = FVX_BC_PERIODIC,
              ISOTHERMAL_WALL_BC = FVX_BC_ISOTHERMALWALL,
              ADIABATIC_WALL_BC = FVX_BC_ADIABATICWALL };

struct FlowBCConfig {
	int bc_tag;
	BCType bc_type;
	std::vector<double> bc_vals;
	std::vector<int> bc_opts;
};

struct FlowPhysicsConfig {
	double gamma, Minf, Tinf, Reinf, Pr;
	double aoa;                 ///< radians, as in the reference
	bool viscous_sim, const_visc;
	std::vector<FlowBCConfig> bcconf;
};

struct FlowNumericsConfig {
	std::string conv_numflux, conv_numflux_jac, gradientscheme, reconstruction;
	double limiter_param;
	bool order2;
};

struct FlowSolutionFunctionals {
	double meshSizeParameter, entropy, CL, CDp, CDsf;
};

namespace detail {
inline int lookup(const std::string& s, std::initializer_list<std::pair<const char*, int>> t,
                  const char* what)
{
	for(const auto& p : t) if(s == p.first) return p.second;
	throw std::invalid_argument(std::string("unknown ") + what + ": " + s);
}
inline int flux_id(const std::string& s)
{
	return lookup(s, {{"ROE", FVX_FLUX_ROE}, {"HLLC", FVX_FLUX_HLLC}, {"HLL", FVX_FLUX_HLL},
	                  {"LLF", FVX_FLUX_LLF}, {"AUSM", FVX_FLUX_AUSM},
	                  {"AUSMPLUS", FVX_FLUX_AUSMPLUS}, {"VANLEER", FVX_FLUX_VANLEER}},
	              "numerical flux");
}
inline int grad_id(const std::string& s)
{
	if(s == "LEASTSQUARES") return FVX_GRAD_LEASTSQUARES;
	if(s == "GREENGAUSS") return FVX_GRAD_GREENGAUSS;
	return FVX_GRAD_NONE;                        // afactory.cpp: anything else is "zero"
}
inline int recon_id(const std::string& s)
{
	return lookup(s, {{"NONE", FVX_RECON_NONE}, {"VANALBADA", FVX_RECON_VANALBADA},
	                  {"VENKATAKRISHNAN", FVX_RECON_VENKATAKRISHNAN},
	                  {"BARTHJESPERSEN", FVX_RECON_BARTHJESPERSEN}, {"WENO", FVX_RECON_WENO}},
	              "reconstruction");
}
}

/// Mesh as read from a file: hybrid tri/quad cells, boundary edges with markers.
struct MeshData {
	std::vector<double> coords;   ///< [npoin][2]
	std::vector<int> inpoel;      ///< [nelem][4], -1 padded
	std::vector<int> nnode;       ///< [nelem]
	std::vector<int> bface;       ///< [nbface][2]
	std::vector<int> btag;        ///< [nbface]
	int npoin() const { return (int)coords.size()/2; }
	int nelem() const { return (int)nnode.size(); }
	int nbface() const { return (int)btag.size(); }
	fvx_mesh_desc desc() const
	{
		return fvx_mesh_desc{ npoin(), nelem(), nbface(), coords.data(), inpoel.data(),
		                      nnode.data(), bface.data(), btag.data() };
	}
	/// Renumbers cells so that new cell i is old cell perm[i].
	void permute_cells(const std::vector<int>& perm)
	{
		std::vector<int> ip(inpoel.size()), nn(nnode.size());
		for(size_t i = 0; i < perm.size(); i++) {
			for(int k = 0; k < 4; k++) ip[4*i+k] = inpoel[4*(size_t)perm[i]+k];
			nn[i] = nnode[perm[i]];
		}
		inpoel.swap(ip);
		nnode.swap(nn);
	}
	/// Reverse Cuthill-McKee reordering ("-mesh_reorder rcm", src/mesh/ameshutils.cpp:245-288).
	void reorder_rcm()
	{
		std::vector<int> perm(nelem());
		const fvx_mesh_desc d = desc();
		check(fvx_mesh_rcm(&d, perm.data()));
		permute_cells(perm);
	}
	/// "-mesh_reorder line | line_natural | line_rcm" with "-mesh_anisotropy_threshold"
	/// (lineReorder / hybridLineReorder, src/mesh/ameshutils.cpp:56-82).
	void reorder_lines(double threshold, const std::string& ordering = "line")
	{
		const int g = ordering == "line" ? FVX_LINES_ONLY
		            : ordering == "line_natural" ? FVX_LINES_NATURAL
		            : ordering == "line_rcm" ? FVX_LINES_RCM : -1;
		if(g < 0) throw std::invalid_argument("unknown line ordering " + ordering);
		std::vector<int> perm(nelem());
		const fvx_mesh_desc d = desc();
		check(fvx_mesh_line_order(&d, threshold, g, perm.data()));
		permute_cells(perm);
	}
	/// preprocessMesh's "-mesh_reorder" (src/mesh/ameshutils.cpp:38-88): "natural" or empty
	/// leaves the order, "rcm" and the line orderings reorder the cells.
	void reorder(const std::string& ordering, double anisotropy_threshold = 0.0)
	{
		if(ordering.empty() || ordering == "natural") return;
		if(ordering == "rcm") { reorder_rcm(); return; }
		if(ordering == "hilbert") {          // no reference counterpart: compact patches
			std::vector<int> perm(nelem());
			const fvx_mesh_desc d = desc();
			check(fvx_mesh_hilbert(&d, perm.data()));
			permute_cells(perm);
			return;
		}
		if(ordering.find("line") != std::string::npos) {
			if(!(anisotropy_threshold > 0))
				throw std::invalid_argument("No anisotropy threshold given!");
			reorder_lines(anisotropy_threshold, ordering);
			return;
		}
		throw std::invalid_argument("unsupported mesh ordering " + ordering);
	}
};

/// Preprocessed device mesh (UMesh after compute_topological / compute_face_data).
class UMesh {
public:
	explicit UMesh(const MeshData& md)
	{
		const fvx_mesh_desc d = md.desc();
		check(fvx_mesh_create(&d, &h_));
		int s[5];
		check(fvx_mesh_sizes(h_, s));
		npoin_ = s[0]; nelem_ = s[1]; nbface_ = s[2]; ninface_ = s[3]; naface_ = s[4];
	}
	~UMesh() { if(h_) fvx_mesh_destroy(h_); }
	UMesh(const UMesh&) = delete;
	UMesh& operator=(const UMesh&) = delete;
	int gnpoin() const { return npoin_; }
	int gnelem() const { return nelem_; }
	int gnbface() const { return nbface_; }
	int gninface() const { return ninface_; }
	int gnaface() const { return naface_; }
	fvx_mesh handle() const { return h_; }
	/// Cell areas (UMesh::garea, src/mesh/mesh.cpp:290-313).
	std::vector<double> areas() const
	{
		std::vector<double> a(nelem_);
		check(fvx_mesh_faces(h_, nullptr, nullptr, a.data(), nullptr));
		return a;
	}
private:
	fvx_mesh h_ = nullptr;
	int npoin_, nelem_, nbface_, ninface_, naface_;
};

/// Block matrix (SeqBAIJ, bs = 4) of a mesh, double or single precision storage.
class BlockMatrix {
public:
	BlockMatrix(const UMesh& m, int precision = FVX_PREC_F64)
	{ check(fvx_matrix_create(m.handle(), precision, &h_)); }
	~BlockMatrix() { if(h_) fvx_matrix_destroy(h_); }
	BlockMatrix(const BlockMatrix&) = delete;
	BlockMatrix& operator=(const BlockMatrix&) = delete;
	/// MatMult
	void apply(const double* x, double* y, void* stream = nullptr) const
	{ check(fvx_matrix_apply(h_, x, y, stream)); }
	fvx_matrix handle() const { return h_; }
private:
	fvx_matrix h_ = nullptr;
};

/// Device FlowFV: the operations of FlowFV_base on device vectors.
class FlowFV {
public:
	FlowFV(const UMesh* mesh, const FlowPhysicsConfig& pconf, const FlowNumericsConfig& nconf)
		: m_(mesh), pconfig(pconf), nconfig(nconf)
	{
		fvx_physics_cfg p{};
		p.gamma = pconf.gamma; p.Minf = pconf.Minf; p.Tinf = pconf.Tinf;
		p.Reinf = pconf.Reinf; p.Pr = pconf.Pr; p.aoa = pconf.aoa;
		p.viscous = pconf.viscous_sim; p.const_visc = pconf.const_visc;
		if(pconf.bcconf.size() > 8) throw std::invalid_argument("at most 8 boundary conditions");
		p.nbc = (int)pconf.bcconf.size();
		for(int i = 0; i < p.nbc; i++) {
			const FlowBCConfig& b = pconf.bcconf[i];
			p.bcs[i].tag = b.bc_tag;
			p.bcs[i].kind = b.bc_type;
			p.bcs[i].vals[0] = b.bc_vals.size() > 0 ? b.bc_vals[0] : 0.0;
			p.bcs[i].vals[1] = b.bc_vals.size() > 1 ? b.bc_vals[1] : 0.0;
		}
		fvx_numerics_cfg n{};
		n.conv_numflux = detail::flux_id(nconf.conv_numflux);
		n.conv_numflux_jac = detail::flux_id(nconf.conv_numflux_jac.empty()
		                                     ? nconf.conv_numflux : nconf.conv_numflux_jac);
		n.gradientscheme = detail::grad_id(nconf.gradientscheme);
		n.reconstruction = detail::recon_id(nconf.reconstruction);
		n.limiter_param = nconf.limiter_param;
		n.order2 = nconf.order2 && n.gradientscheme != FVX_GRAD_NONE;
		check(fvx_disc_create(mesh->handle(), &p, &n, &h_));
	}
	~FlowFV() { if(h_) fvx_disc_destroy(h_); }
	FlowFV(const FlowFV&) = delete;
	FlowFV& operator=(const FlowFV&) = delete;

	/// Adds -r(u) into residual; overwrites timesteps when requested.
	StatusCode compute_residual(const double* u, double* residual, bool gettimesteps,
	                            double* timesteps, void* stream = nullptr) const
	{
		check(fvx_compute_residual(h_, u, residual, gettimesteps, timesteps, stream));
		return 0;
	}
	/// Overwrites A with the first-order approximate Jacobian (MatZeroEntries + assembly).
	StatusCode assemble_jacobian(const double* u, BlockMatrix& A, void* stream = nullptr) const
	{
		check(fvx_assemble_jacobian(h_, u, A.handle(), stream));
		return 0;
	}
	/// Gradients of the conserved variables, device [ncell][4][2].
	void getGradients(const double* u, double* grads, void* stream = nullptr) const
	{
		check(fvx_get_gradients(h_, u, grads, stream));
	}
	std::array<double,4> compute_freestream_state() const
	{
		std::array<double,4> out;
		check(fvx_freestream_state(h_, out.data()));
		return out;
	}
	/// CL, CDp, CDsf on one wall marker and the entropy error (synchronises the stream).
	FlowSolutionFunctionals computeFunctionals(const double* u, int wall_marker,
	                                           void* stream = nullptr) const
	{
		double out[4];
		check(fvx_functionals(h_, u, wall_marker, out, stream));
		return FlowSolutionFunctionals{ 0.0, out[3], out[0], out[1], out[2] };
	}
	/// computeSurfaceData's output matrix on one wall marker: rows (x, y, Cp, Cf) per face.
	std::vector<double> computeSurfaceData(const double* u, int wall_marker,
	                                       void* stream = nullptr) const
	{
		int n = 0;
		check(fvx_surface_data(h_, u, wall_marker, nullptr, 0, &n, stream));
		std::vector<double> out(4*(size_t)n);
		if(n) check(fvx_surface_data(h_, u, wall_marker, out.data(), n, &n, stream));
		return out;
	}
	/// Step records of the last steady solve on this discretisation (TimingData::convhis,
	/// src/ode/aodesolver.hpp:46-67).
	std::vector<fvx_step_record> solve_history() const
	{
		int n = 0;
		check(fvx_solve_history(h_, nullptr, 0, &n));
		std::vector<fvx_step_record> h(n);
		if(n) check(fvx_solve_history(h_, h.data(), n, &n));
		return h;
	}
	const UMesh* mesh() const { return m_; }
	fvx_disc handle() const { return h_; }

	const UMesh* const m_;
	const FlowPhysicsConfig pconfig;
	const FlowNumericsConfig nconfig;
private:
	fvx_disc h_ = nullptr;
};

/// SteadySolverConfig plus the linear-solver options the reference reads from PETSc.
struct SteadySolverConfig {
	double cflinit = 500, cflfin = 5000;
	double tol = 1e-7;
	int maxiter = 170;
	std::string nl_update_scheme = "ROBUST_FLOW";   ///< or "FULL"
	double min_nl_update = 0.2;
	/// none | jacobi | sgs | ilu0 | line | sgs_level | ilu0_level | bjacobi_ilu0
	std::string pc_type = "sgs";
	int pc_sweeps = 1;
	int pc_precision = FVX_PREC_F64, jac_precision = FVX_PREC_F64;
	int ksp_max_it = 30;
	double ksp_rtol = 1e-1;
	bool matrix_free = false;
	double matrix_free_difference_step = 1e-7;
	double fixed_norm = 0.0;
	double line_threshold = 10.0;                    ///< anisotropy threshold of "line"
	int pc_subdomain_cells = 256;                    ///< subdomain size of "bjacobi_ilu0"
};

struct TimingData {
	double ode_walltime = 0;
	int total_lin_iters = 0, avg_lin_iters = 0, num_timesteps = 0;
	bool converged = false;
	double final_rel = 0, final_abs = 0;
};

/// Steady pseudo-time solver (implicit or explicit) running on the device.
class SteadySolver {
public:
	SteadySolver(const FlowFV* space, const SteadySolverConfig& conf, bool implicit)
		: space_(space), conf_(conf), implicit_(implicit) {}
	/// Solves in place; throws Numerical_error / Tolerance_error like the reference.
	StatusCode solve(double* u, void* stream = nullptr)
	{
		fvx_solver_cfg c{};
		c.cflinit = conf_.cflinit; c.cflfin = conf_.cflfin; c.tol = conf_.tol;
		c.maxiter = conf_.maxiter; c.implicit = implicit_;
		c.fullupdate = conf_.nl_update_scheme == "FULL";
		c.min_relax = conf_.min_nl_update;
		c.pc = detail::lookup(conf_.pc_type, {{"none", FVX_PC_NONE}, {"jacobi", FVX_PC_JACOBI},
		                                      {"sgs", FVX_PC_SGS}, {"ilu0", FVX_PC_ILU0},
		                                      {"line", FVX_PC_LINE},
		                                      {"sgs_level", FVX_PC_SGS_LEVEL},
		                                      {"ilu0_level", FVX_PC_ILU0_LEVEL},
		                                      {"bjacobi_ilu0", FVX_PC_BJACOBI_ILU0}},
		                      "preconditioner");
		c.pc_precision = conf_.pc_precision;
		c.pc_sweeps = conf_.pc_sweeps;
		c.jac_precision = conf_.jac_precision;
		c.krylov_restart = conf_.ksp_max_it;
		c.krylov_rtol = conf_.ksp_rtol;
		c.matrix_free = conf_.matrix_free;
		c.mf_eps = conf_.matrix_free_difference_step;
		c.fixed_norm = conf_.fixed_norm;
		c.line_threshold = conf_.line_threshold;
		c.pc_subdomain_cells = conf_.pc_subdomain_cells;
		fvx_solve_stats st{};
		const fvx_status s = fvx_steady_solve(space_->handle(), &c, u, &st, stream);
		tdata_.ode_walltime = st.walltime;
		tdata_.total_lin_iters = st.lin_iters;
		tdata_.num_timesteps = st.steps;
		tdata_.avg_lin_iters = st.steps ? st.lin_iters/st.steps : 0;
		tdata_.converged = st.converged;
		tdata_.final_rel = st.final_rel;
		tdata_.final_abs = st.final_abs;
		check(s);
		return 0;
	}
	TimingData getTimingData() const { return tdata_; }
private:
	const FlowFV* space_;
	SteadySolverConfig conf_;
	bool implicit_;
	TimingData tdata_;
};

struct SteadyBackwardEulerSolver : SteadySolver {
	SteadyBackwardEulerSolver(const FlowFV* s, const SteadySolverConfig& c)
		: SteadySolver(s, c, true) {}
};
struct SteadyForwardEulerSolver : SteadySolver {
	SteadyForwardEulerSolver(const FlowFV* s, const SteadySolverConfig& c)
		: SteadySolver(s, c, false) {}
};

}
#endif
