/* fvx_internal.hpp -- internal types shared by the host code and kernel launchers of libfvx. */
#ifndef FVX_INTERNAL_HPP
#define FVX_INTERNAL_HPP

#include <hip/hip_runtime.h>
#include <string>
#include <vector>
#include <stdexcept>
#include "../../include/fvx.h"
#include "gas.hpp"

namespace fvx {

/// Error carrying an fvx_status across internal C++ code; converted at the C boundary.
struct Error : std::runtime_error {
	int code;
	Error(int c, const std::string& m) : std::runtime_error(m), code(c) {}
};

#define FVX_HIP(call) do { hipError_t e_ = (call); if(e_ != hipSuccess) \
	throw ::fvx::Error(FVX_ERR_HIP, std::string(#call) + ": " + hipGetErrorString(e_)); } while(0)

/// Marker in a cell's neighbour slot that is not used (triangles have 3 faces).
constexpr int SLOT_UNUSED = -0x40000000;

/// bfloat16 storage of a preconditioner factor (the upper 16 bits of an IEEE binary32).
struct bf16 { unsigned short bits; };

/// Face geometry record: unit normal pointing from the left (lower) cell to the right cell,
/// length, midpoint, and the two cells: left, and right or -1-ib for boundary face ib (48 bytes,
/// three 16-byte loads).
struct alignas(16) FaceGeom {
	double nx, ny, len, mx, my;
	int left, right;
};

/// Instance addressing of a launch. A batched launch covers gridDim.y instances whose ids
/// are list[blockIdx.y]; per-instance arrays are instance-major (instance i of an array of
/// stride S starts at i*S) and each instance has its own gas / boundary parameters. A
/// single-instance launch has list == nullptr, instance 0 and the by-value parameters.
struct BatchArg {
	const int* list = nullptr;
	const GasParams* gp = nullptr;         ///< per-instance gas parameters (null: the single one)
	const double* cfl = nullptr;           ///< per-instance CFL numbers (Jacobian kernel)
};

/// Host-side description of a launch's instances: the device list of active instance ids
/// (null for a single instance) and how many there are (gridDim.y).
struct Inst {
	const int* list = nullptr;
	int nact = 1;
};

/// Device buffer owned by a handle.
template <class T> struct DBuf {
	T* p = nullptr;
	size_t n = 0;
	void alloc(size_t count) {
		release();
		if(count == 0) return;
		if(hipMalloc((void**)&p, count*sizeof(T)) != hipSuccess) {
			p = nullptr;
			throw Error(FVX_ERR_ALLOC, "hipMalloc failed");
		}
		n = count;
	}
	void upload(const std::vector<T>& h) {
		alloc(h.size());
		if(n) FVX_HIP(hipMemcpy(p, h.data(), n*sizeof(T), hipMemcpyHostToDevice));
	}
	void release() { if(p) (void)hipFree(p); p = nullptr; n = 0; }
	~DBuf() { release(); }
	DBuf() = default;
	DBuf(const DBuf&) = delete;
	DBuf& operator=(const DBuf&) = delete;
};

}

/// Host + device mesh. Cells are in the caller's order; faces follow the reference numbering
/// (boundary faces in input order, then interior faces in element order with L < R).
struct fvx_mesh_s {
	int npoin = 0, nelem = 0, nbface = 0, ninface = 0, naface = 0;
	// host topology and geometry
	std::vector<double> coords;
	std::vector<int> inpoel, nnode, bface, btag;
	std::vector<int> esuel, elemface;      // [nelem][4]
	std::vector<int> intfac;               // [naface][4]
	std::vector<double> facemetric;        // [naface][3]
	std::vector<double> area, rc;          // [nelem], [nelem][2]
	std::vector<double> gr, rcbp;          // face midpoints [naface][2], ghost centres [nbface][2]
	std::vector<double> hlong;             // longest edge per cell
	// device copies
	fvx::DBuf<int> d_nbr, d_face;          // [nelem][4] slots: neighbour / face index
	fvx::DBuf<fvx::FaceGeom> d_fgeom;      // [naface]
	fvx::DBuf<double> d_soff;              // [nelem][4][2] face midpoint minus cell centre
	fvx::DBuf<double> d_centre;            // [nelem][2]
	fvx::DBuf<double> d_area;              // [nelem]
	fvx::DBuf<int> d_bhost, d_btag;        // [nbface] host cell and marker of boundary faces
	// wave-task form of the flux pass (flux_wave.hpp): per group of consecutive cells whose
	// faces make at most 64 tasks, the tasks, the group's cells and each cell's slot codes
	int nwgroup = 0;
	fvx::DBuf<int> d_wtask;                // [ngroup][64] face of each task, -1 none
	fvx::DBuf<int2> d_wgroup;              // [ngroup] {first cell, cells}
	fvx::DBuf<unsigned> d_wslot;           // [ngroup][64] per cell: byte k = 2 task + side of
	                                       // slot k, 0xff unused
};

namespace fvx {
/// Optional per-kernel timing with HIP events recorded on the launch stream around each
/// kernel; read back (and the stream synchronised) only on request.
struct Profiler {
	static constexpr int NKERNEL = 8;
	bool on = false;
	std::vector<hipEvent_t> pool;          // start/stop pairs
	std::vector<int> kid;                  // kernel id of each pair
	size_t used = 0;
	double total_ms[NKERNEL] = {};
	long count[NKERNEL] = {};
	void start(int k, hipStream_t s);
	void stop(hipStream_t s);
	void collect();
	void reset();
	~Profiler();
};
enum { PROF_GRADIENTS = 0, PROF_FLUX = 1, PROF_JACOBIAN = 2, PROF_PC = 3 };
struct SolverWork;
void destroy_solver_work(SolverWork* w);
struct Krylov;
void destroy_krylov(Krylov* k);
}

namespace fvx {
/// Launch shapes of the cell kernels (threads per workgroup, XCD-aware block order), chosen
/// by the autotuner; results do not depend on them.
struct Tuning {
	int flux_block = 256, grad_block = 256, jac_block = 256;
	int flux_xcd = 1, grad_xcd = 1, jac_xcd = 1;
	int flux_wave = 0;     ///< 1: wave-task form of the flux pass where it applies
};
}

/// Discretisation handle: configuration plus per-mesh coefficient arrays and scratch.
struct fvx_disc_s {
	mutable fvx::Profiler prof;
	fvx::Tuning tune;
	fvx_mesh_s* m = nullptr;
	fvx_physics_cfg pc;
	fvx_numerics_cfg nc;
	fvx::GasParams gas;
	fvx::BCTable bct;
	std::vector<int> bcid;                 // per boundary face: row of bct
	fvx::DBuf<int> d_bcid;                 // [nbface]
	fvx::DBuf<fvx::GasParams> d_gas;       // [1] device copy of gas
	fvx::DBuf<fvx::BCTable> d_bct;         // [1] device copy of bct
	fvx::DBuf<double> d_gcoef;             // [nelem][4][2] gradient stencil coefficients
	fvx::DBuf<double> d_eps2;              // [nelem] Venkatakrishnan (K h)^3
	fvx::DBuf<double> d_grad;              // [nelem][8] unlimited primitive gradients
	fvx::DBuf<double> d_ughost;            // [nbface][4] ghost states of the boundary faces
	fvx::DBuf<double> d_rgrad;             // [nelem][8] reconstruction gradients (limited)
	fvx::DBuf<double> d_scratch;           // reductions for functionals
	fvx::DBuf<double> mf_aux, mf_raux;     // [nelem][4] matrix-free state and residual, first use
	fvx::DBuf<double> mf_part;             // matrix-free norm partial sums and |x|^2
	fvx::SolverWork* work = nullptr;       // implicit-solver workspace, created on first use
	~fvx_disc_s() { fvx::destroy_solver_work(work); }
};

/// Operating points solved together on one discretisation: per-instance Mach number and
/// angle of attack (gas parameters and free stream), everything else shared.
struct fvx_batch_s {
	fvx_disc_s* d = nullptr;
	int ninst = 0;
	std::vector<double> aoa;
	std::vector<fvx::GasParams> gas;
	fvx::DBuf<fvx::GasParams> d_gas;
	fvx::DBuf<int> d_all;                  // 0 .. ninst-1
	fvx::DBuf<double> d_grad, d_rgrad;     // [ninst][nelem][8]
	fvx::DBuf<double> d_ughost;            // [ninst][nbface][4]
	fvx::DBuf<double> d_scratch;           // functional partial sums
	fvx::DBuf<double> n_r, n_part, n_out;  // residual-norm scratch, allocated on first use
	double* n_host = nullptr;              // pinned host copy of the norms
	std::vector<double> norms;             // fixed convergence normalisers (empty: none)
	~fvx_batch_s() { if(n_host) (void)hipHostFree(n_host); }
};

/// Block matrix on the cell graph: vals[cell][slot][4][4], slot 0 the diagonal block and slot
/// 1+k the block of neighbour nbr[cell][k] (zero where there is none).
struct fvx_matrix_s {
	const fvx_mesh_s* m = nullptr;
	int prec = FVX_PREC_F64;
	int ninst = 1;                         // instances stored, 80*ncell values apart
	fvx::DBuf<double> v64;
	fvx::DBuf<float> v32;
	double* f64() const { return v64.p; }
	float* f32() const { return v32.p; }
	void* raw() const { return prec == FVX_PREC_F32 ? (void*)v32.p : (void*)v64.p; }
};

/// Preconditioner: factors in the matrix layout with slot 0 holding the inverted diagonal
/// block; multicolour ordering (colour per cell, cells listed by colour, lower-colour mask).
struct fvx_pc_s {
	const fvx_mesh_s* m = nullptr;
	int type = FVX_PC_NONE, prec = FVX_PREC_F64;
	int ninst = 1;                         // instances stored, 80*ncell values apart
	int sweeps = 1;                        // stationary iterations per application
	fvx::DBuf<double> one;                 // the constant 1.0 (identity preconditioner)
	const fvx_matrix_s* A = nullptr;       // matrix of the last set-up
	fvx::DBuf<double> t1, t2;              // residual and correction for sweeps > 1
	int ncol = 0;
	std::vector<int> colour, order, coff;
	fvx::DBuf<int> d_colour, d_order;
	fvx::DBuf<unsigned char> d_lowmask;
	fvx::DBuf<int4> d_crec;                // SGS: per position {cell, lower mask, slots 0-3}
	// line Gauss-Seidel: vertices (lines, then the other cells) listed by colour in order /
	// coff, their cells, and per cell the slots of its line neighbours
	double line_threshold = 10.0;
	int nvert = 0;
	std::vector<int> vptr, vcells, lcnt;   // lcnt: lines among the vertices of each colour
	fvx::DBuf<int> d_vptr, d_vcells, d_ocell;   // d_ocell: first cell of each vertex in order
	fvx::DBuf<unsigned char> d_lineinfo;
	// block Jacobi with ILU(0) per subdomain: subdomain of every cell; the cells listed by
	// (subdomain, level, index) with subdomain ranges bptr and level boundaries blv (subdomain b
	// owns blv[blvptr[b] .. blvptr[b+1]]); per listed cell the positions of its neighbours in
	// its subdomain's list (-1: not in the subdomain); the neighbour slots restricted to the
	// subdomain (-1 elsewhere)
	int subdomain_cells = 256, nsub = 0;
	std::vector<int> part;
	fvx::DBuf<int> d_bptr, d_blvptr, d_blv, d_bcells, d_bnbr;
	fvx::DBuf<int4> d_ploc;
	fvx::DBuf<double> v64;
	fvx::DBuf<float> v32;
	fvx::DBuf<fvx::bf16> v16;
	fvx::Krylov* kry = nullptr;            // FGMRES workspace of fvx_linear_solve, first use
	double* f64() const { return v64.p; }
	float* f32() const { return v32.p; }
	fvx::bf16* fb16() const { return v16.p; }
	~fvx_pc_s() { fvx::destroy_krylov(kry); }
};

namespace fvx {
/// Symmetric Gauss-Seidel / ILU(0) preconditioners, in multicolour or natural (level) order.
inline bool pc_sgs(int t) { return t == FVX_PC_SGS || t == FVX_PC_SGS_LEVEL; }
inline bool pc_ilu(int t) { return t == FVX_PC_ILU0 || t == FVX_PC_ILU0_LEVEL; }
inline bool pc_level(int t) { return t == FVX_PC_SGS_LEVEL || t == FVX_PC_ILU0_LEVEL; }
/// Largest subdomain of FVX_PC_BJACOBI_ILU0 (one workgroup's LDS holds its vector).
constexpr int BJ_MAX_CELLS = 1024;
void build_mesh(fvx_mesh_s& m, const fvx_mesh_desc& d);
void rcm_order(const fvx_mesh_desc& d, int* perm);
void find_lines(const fvx_mesh_desc& d, double threshold, std::vector<std::vector<int>>& lines,
                std::vector<int>& celline);
void line_order(const fvx_mesh_desc& d, double threshold, int graph_ordering, int* perm);
void hilbert_order(const fvx_mesh_desc& d, int* perm);
void build_wave_tasks(fvx_mesh_s& m);
void build_disc(fvx_disc_s& d);
GasParams make_gas(const fvx_physics_cfg& p);

// kernel launchers (spatial.hip, flux_inst.hip); a batched launch passes the instance list
// and per-instance parameters in BatchArg and the number of instances in nact
/// ughost: the boundary faces' ghost states, [instance][nbface][4] for a batched launch (null:
/// the discretisation's buffer, for a launch of instance 0 only, like grad and rgrad).
void launch_gradients(const fvx_disc_s& d, const double* u, double* grad, double* rgrad,
                      int conserved, hipStream_t s, const BatchArg& ba = BatchArg(),
                      int nact = 1, double* ughost = nullptr);
/// Whether the wave-task form of the flux pass covers this discretisation's options.
bool flux_wave_apply(const fvx_disc_s& d);
void launch_residual(const fvx_disc_s& d, const double* u, double* r, int gettimesteps,
                     double* dt, hipStream_t s, const BatchArg& ba = BatchArg(), int nact = 1,
                     double* grad = nullptr, double* rgrad = nullptr, double* ughost = nullptr);
void launch_functionals(const fvx_disc_s& d, const double* u, const double* cgrad, int walltag,
                        double* dev_out, hipStream_t s, const GasParams* gp = nullptr,
                        double aoa = 0.0);
int functional_blocks();

// Jacobian assembly (jacobian.hip)
bool jacobian_supported(const fvx_disc_s& d);
void launch_jacobian(const fvx_disc_s& d, const double* u, void* vals, int prec,
                     const double* dt, double cfl, double* dtm, hipStream_t s,
                     const BatchArg& ba = BatchArg(), int nact = 1);

// matrix, preconditioner and vector kernels (linalg.hip); s* arguments are the per-instance
// strides of the vectors, in doubles
void launch_spmv(const fvx_matrix_s& A, const double* x, long sx, double* y, long sy,
                 const Inst& L, hipStream_t s);
void launch_pseudotime(fvx_matrix_s& A, const double* dt, double cfl, double* dtm,
                       hipStream_t s);
void launch_pc_setup(fvx_pc_s& p, const fvx_matrix_s& A, const Inst& L, hipStream_t s);
void launch_pc_apply(const fvx_pc_s& p, const double* r, long sr, double* z, long sz,
                     const Inst& L, hipStream_t s);
int reduction_blocks();
int max_dots();
long partial_stride();
void launch_mdot(long N, const double* V, long ldv, long sv, int nv, const double* w, long sw,
                 double* partial, double* out, long so, const Inst& L, hipStream_t s);
void launch_gs_update(long N, const double* V, long ldv, long sv, int nv, const double* h,
                      long sh, double* w, long sw, double* partial, double* nrm2, long sn,
                      const Inst& L, hipStream_t s);
void launch_scale(long N, const double* b, long sb, const double* nrm2, long sn, double* out,
                  long so, const Inst& L, hipStream_t s);
void launch_lincomb(long N, const double* Z, long ldz, long sz, int nv, const double* y,
                    long sy, double* x, long sx, const Inst& L, hipStream_t s,
                    const int* nvi = nullptr);
/// Device least squares of FGMRES (linalg.hip): state doubles per instance and the launchers
/// of the cycle start, the per-step rotation and the cycle end.
long fgmres_state_size();
void launch_fgmres_begin(int nids, const int* ids, const double* scal, long ns, int beta_slot,
                         double* kst, int* list, int* hcount, hipStream_t s);
void launch_fgmres_rotate(int j, double rtol, int nids, const int* ids, const double* scal,
                          long ns, int h_slot, int nrm_slot, double* kst, int* list,
                          int* hcount, hipStream_t s);
void launch_fgmres_end(int nids, const int* ids, const double* kst, double* scal, long ns,
                       int y_slot, int* its, hipStream_t s);
void launch_resnorm(int n, const double* r, const double* area, double* partial, double* out,
                    long so, const Inst& L, hipStream_t s);
void launch_zero(long N, double* x, long sx, const Inst& L, hipStream_t s);
void launch_mf_prep(long N, const double* u, const double* x, long sx, const double* xx,
                    long sn, double eps, double* aux, const Inst& L, hipStream_t s);
void launch_mf_finish(int n, const double* x, long sx, const double* dtm, const double* ru,
                      const double* raux, const double* xx, long sn, double eps, double* y,
                      long sy, const Inst& L, hipStream_t s);
void launch_update_implicit(int n, double* u, const double* du, int full, double min_relax,
                            double gamma, const Inst& L, hipStream_t s);
void launch_update_explicit(int n, double* u, const double* r, const double* dt,
                            const double* area, double cfl, const Inst& L, hipStream_t s);

// host side of the implicit solver (solver.cpp)
void matrix_init(fvx_matrix_s& A, const fvx_mesh_s* m, int prec, int ninst = 1);
void pc_init(fvx_pc_s& p, const fvx_mesh_s* m, int type, int prec, int ninst = 1);
void pc_set_sweeps(fvx_pc_s& p, int sweeps);
void pc_set_subdomain_cells(fvx_pc_s& p, int cells);
void matrix_export_bsr(const fvx_matrix_s& A, int* rowptr, int* colind, double* vals);
int matrix_bsr_nnzb(const fvx_mesh_s& m);
void matfree_apply(const fvx_disc_s& d, const double* u, const double* ru, const double* dtm,
                   const double* x, double* y, double eps, double* aux, double* raux,
                   double* partial, double* xx, hipStream_t s);
int linear_solve(const fvx_matrix_s& A, fvx_pc_s& p, const double* b, double* x,
                 int restart, double rtol, hipStream_t s);
void steady_solve(fvx_disc_s& d, const fvx_solver_cfg& cfg, double* u, fvx_solve_stats& st,
                  hipStream_t s);
/// Steady solve of ninst instances; norms (host, optional) are fixed per-instance normalisers
/// of the convergence test (otherwise cfg.fixed_norm, or each instance's first residual norm).
void steady_core(fvx_disc_s& d, const fvx_solver_cfg& c, double* u, int ninst,
                 const GasParams* gp, std::vector<fvx_solve_stats>& stats, hipStream_t s,
                 const double* norms = nullptr);

void solve_history(const fvx_disc_s& d, fvx_step_record* out, int cap, int* n);
void launch_surface_data(const fvx_disc_s& d, const double* u, const double* cgrad,
                         const int* list, int nw, double* out, hipStream_t s);

// autotuner (tune.cpp)
void autotune_launch(fvx_disc_s& d, const double* u, int reps, std::vector<fvx_tune_row>& rows,
                     hipStream_t s);
void autotune_pc(fvx_disc_s& d, const double* u, double cfl, int restart, double rtol,
                 int precision, std::vector<fvx_tune_row>& rows, int& best_pc, int& best_sweeps,
                 hipStream_t s);

// learned warm start (mpnn.hip)
namespace mpnn {
constexpr int W = 64, NL = 4, FIN = 8, FE = 4, FO = 4, NSTAT = 6;
// flat fp32 parameter layout; matrices are [in][out] row-major
constexpr int OFF_WENC = 0, OFF_BENC = 512, OFF_L0 = 576, LSZ = 16768;
constexpr int L_WA = 0, L_WB = 4096, L_WE = 8192, L_BM = 8448, L_WU = 8512, L_WV = 12608,
              L_BU = 16704;
constexpr int OFF_WDEC = OFF_L0 + NL*LSZ, OFF_BDEC = OFF_WDEC + 256, NPARAM = OFF_BDEC + 4;
static_assert(NPARAM == 67908, "parameter layout");
// packed MFMA B operands per layer: P|Q forward, update forward, update backward, P|Q backward
enum { M_PQ = 0, M_UPD = 1, M_BUPD = 2, M_BPQ = 3 };
constexpr int FRAG_MAT = 16*64*8;      // bf16 values per packed matrix (16 fragments)
}

/// Graph arrays of the learned warm start built on the host from a mesh.
struct MpnnGraphHost {
	int N = 0;
	std::vector<int> nbr, rev;             // [N][4] interior neighbour / reverse slot, -1 unused
	std::vector<float> invdeg, stat, edge; // [N], [N][6], [N][4][4]
};
void mpnn_build_graph(const fvx_mesh_s& m, int wall_tag, MpnnGraphHost& gh);

struct Mpnn {
	MpnnGraphHost host;
	int N = 0, cap = 0;
	long step = 0;
	bool dirty = true;
	double gamma = 1.4;
	DBuf<int> d_nbr, d_rev, ids;
	DBuf<float> d_invdeg, d_stat, d_edge;
	DBuf<float> w, wq, am, av, grad, loss;
	DBuf<unsigned short> frag;
	DBuf<float> h[mpnn::NL + 1], pq[mpnn::NL];
	DBuf<unsigned short> ab[mpnn::NL];
	DBuf<unsigned long long> umask[mpnn::NL];
	DBuf<float> g0, g1, gpart, da, y, part, gfeat;
	DBuf<unsigned short> dpre, dpq;
	DBuf<double> pp;
	std::vector<float> hgf;
	std::vector<double> hpp;
	std::vector<int> hids;

	void init(const fvx_mesh_s& m, int wall_tag, int max_graphs);
	int npoints = 0;                       // graphs described by the last set_points
	void set_points(int G, const double* mach, const double* aoa, double gamma, hipStream_t s);
	void pack(hipStream_t s);
	void forward(int G, bool train, hipStream_t s);
	void predict(int G, float* yout, hipStream_t s);
	void gradient(int G, const float* targets, const int* sample_ids, hipStream_t s);
	void adam_step(double lr, hipStream_t s);
	void targets(int G, const double* u, float* t, hipStream_t s);
	void warm(int G, double floor, double* u, hipStream_t s);
};
}

struct fvx_mpnn_s {
	fvx::Mpnn net;
};

#endif
