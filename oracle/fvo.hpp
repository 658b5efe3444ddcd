/* fvo.hpp -- CPU oracle for the fvx residual / Jacobian / preconditioner engine.
 *
 * TEST INFRASTRUCTURE ONLY. This library is a plain FP64 CPU restatement of the reference
 * FVENS algorithm for the steady-flow hot path. It exists so that the GPU product can be
 * checked against it. Only tests, the smoke check and the CPU-baseline leg of the benchmark
 * may load it; the product never links or calls it.
 *
 * Conventions follow the reference (cited as path:line relative to the FVENS source tree):
 *  - NDIM = 2, NVARS = 4, double precision, int32 indices (src/aconstants.hpp:32,35,60,70)
 *  - state u is AoS [ncell][4] conserved (rho, rho*vx, rho*vy, rho*E)
 *  - the residual routine ADDS -r(u) into its output (src/spatial/flow_spatial.hpp:74-87)
 *  - face Jacobians: dfdl = -dF/du_L, dfdr = +dF/du_R (src/spatial/anumericalflux.hpp:36-45)
 *
 * Parity status of each part is recorded in the header of the file implementing it.
 */
#ifndef FVO_HPP
#define FVO_HPP

#include <cmath>
#include <vector>
#include <array>
#include <string>
#include <stdexcept>

namespace fvo {

constexpr int NV = 4;   ///< conserved variables
constexpr int ND = 2;   ///< space dimensions

/* Numerical-method identifiers. The numeric values are shared with the product's C ABI
 * (include/fvx.h) so that one configuration struct drives both. */
enum FluxId  { FLUX_ROE = 0, FLUX_HLLC = 1, FLUX_HLL = 2, FLUX_LLF = 3, FLUX_AUSM = 4,
               FLUX_AUSMPLUS = 5, FLUX_VANLEER = 6 };
enum GradId  { GRAD_NONE = 0, GRAD_GREENGAUSS = 1, GRAD_LEASTSQUARES = 2 };
enum ReconId { RECON_NONE = 0, RECON_VANALBADA = 1, RECON_VENKATAKRISHNAN = 2,
               RECON_BARTHJESPERSEN = 3, RECON_WENO = 4 };
/// Boundary-condition kinds, same order as the reference enum (src/spatial/abctypes.hpp:13-21)
enum BCKind  { BC_SLIPWALL = 0, BC_FARFIELD = 1, BC_INOUTFLOW = 2, BC_SUBSONIC_INFLOW = 3,
               BC_EXTRAPOLATION = 4, BC_PERIODIC = 5, BC_ISOTHERMALWALL = 6,
               BC_ADIABATICWALL = 7 };

/// One boundary condition (src/spatial/abc.hpp:34-40)
struct BCConf {
	int tag;
	int kind;
	double vals[2];
};

/// Physics + numerics configuration (src/spatial/flow_spatial.hpp:33-55, explicit limiter K)
struct Config {
	double gamma, Minf, Tinf, Reinf, Pr, aoa;
	int viscous, constvisc;
	int flux, jacflux, gradient, recon;
	double limiter_param;
	int order2;
	int nbc;
	BCConf bcs[8];
};

/// Hybrid tri/quad 2-D mesh with the reference's face structure (src/mesh/mesh.hpp).
struct Mesh {
	int npoin = 0, nelem = 0, nbface = 0, ninface = 0, naface = 0;
	std::vector<double> coords;     // npoin x 2
	std::vector<int> inpoel;        // nelem x 4, -1 padded
	std::vector<int> nnode;         // nelem
	std::vector<int> bface;         // nbface x 2 (after orientation correction)
	std::vector<int> btag;          // nbface
	std::vector<int> esuel;         // nelem x 4: neighbour across local face j (ghost: nelem+ib)
	std::vector<int> elemface;      // nelem x 4: face index of local face j
	std::vector<int> intfac;        // naface x 4: L, R, node0, node1
	std::vector<double> facemetric; // naface x 3: nx, ny, len
	std::vector<double> area;       // nelem
	std::vector<double> rc;         // nelem x 2 cell centres (vertex average)
	std::vector<double> rcbp;       // nbface x 2 ghost centres
	std::vector<double> gr;         // naface x 2 face midpoints
};

void mesh_build(Mesh& m);
std::vector<int> level_schedule(const Mesh& m);

/* ---------------------------------------------------------------- dual numbers --------- */

constexpr int NDUAL = 8;   ///< derivative slots: 0-3 w.r.t. left state, 4-7 w.r.t. right state

/// Forward-mode dual number used to differentiate the flux restatements exactly.
struct Dual {
	double v;
	double d[NDUAL];
	Dual() : v(0) { for(int i = 0; i < NDUAL; i++) d[i] = 0; }
	Dual(double x) : v(x) { for(int i = 0; i < NDUAL; i++) d[i] = 0; }
};

inline Dual operator+(const Dual& a, const Dual& b) { Dual r; r.v = a.v+b.v; for(int i=0;i<NDUAL;i++) r.d[i]=a.d[i]+b.d[i]; return r; }
inline Dual operator-(const Dual& a, const Dual& b) { Dual r; r.v = a.v-b.v; for(int i=0;i<NDUAL;i++) r.d[i]=a.d[i]-b.d[i]; return r; }
inline Dual operator-(const Dual& a) { Dual r; r.v = -a.v; for(int i=0;i<NDUAL;i++) r.d[i]=-a.d[i]; return r; }
inline Dual operator*(const Dual& a, const Dual& b) { Dual r; r.v = a.v*b.v; for(int i=0;i<NDUAL;i++) r.d[i]=a.d[i]*b.v+a.v*b.d[i]; return r; }
inline Dual operator/(const Dual& a, const Dual& b) { Dual r; r.v = a.v/b.v; for(int i=0;i<NDUAL;i++) r.d[i]=(a.d[i]*b.v-a.v*b.d[i])/(b.v*b.v); return r; }
inline Dual operator+(const Dual& a, double b) { Dual r = a; r.v += b; return r; }
inline Dual operator+(double b, const Dual& a) { Dual r = a; r.v += b; return r; }
inline Dual operator-(const Dual& a, double b) { Dual r = a; r.v -= b; return r; }
inline Dual operator-(double b, const Dual& a) { Dual r = -a; r.v += b; return r; }
inline Dual operator*(const Dual& a, double b) { Dual r; r.v = a.v*b; for(int i=0;i<NDUAL;i++) r.d[i]=a.d[i]*b; return r; }
inline Dual operator*(double b, const Dual& a) { return a*b; }
inline Dual operator/(const Dual& a, double b) { Dual r; r.v = a.v/b; for(int i=0;i<NDUAL;i++) r.d[i]=a.d[i]/b; return r; }
inline Dual operator/(double a, const Dual& b) { return Dual(a)/b; }
inline Dual& operator+=(Dual& a, const Dual& b) { a = a+b; return a; }
inline Dual& operator-=(Dual& a, const Dual& b) { a = a-b; return a; }
inline Dual& operator*=(Dual& a, const Dual& b) { a = a*b; return a; }
inline Dual& operator/=(Dual& a, const Dual& b) { a = a/b; return a; }
inline bool operator<(const Dual& a, const Dual& b) { return a.v < b.v; }
inline bool operator>(const Dual& a, const Dual& b) { return a.v > b.v; }
inline bool operator<=(const Dual& a, const Dual& b) { return a.v <= b.v; }
inline bool operator>=(const Dual& a, const Dual& b) { return a.v >= b.v; }
inline bool operator<(const Dual& a, double b) { return a.v < b; }
inline bool operator>(const Dual& a, double b) { return a.v > b; }
inline bool operator<=(const Dual& a, double b) { return a.v <= b; }
inline bool operator>=(const Dual& a, double b) { return a.v >= b; }
inline Dual sqrt(const Dual& a) { Dual r; r.v = std::sqrt(a.v); for(int i=0;i<NDUAL;i++) r.d[i]=0.5*a.d[i]/r.v; return r; }
inline Dual fabs(const Dual& a) { return a.v >= 0 ? a : -a; }
inline Dual pow(const Dual& a, double e) { Dual r; r.v = std::pow(a.v,e); const double dr = e*std::pow(a.v,e-1.0); for(int i=0;i<NDUAL;i++) r.d[i]=dr*a.d[i]; return r; }
inline double value(const Dual& a) { return a.v; }
inline double value(double a) { return a; }

using std::sqrt;
using std::fabs;
using std::pow;

/* ---------------------------------------------------------------- spatial ------------- */

/// Discretisation context: mesh, configuration and precomputed per-mesh data.
struct Disc {
	const Mesh* m;
	Config cfg;
	std::array<double,NV> uinf;
	std::vector<int> bckind;            // per boundary face: BC kind
	std::vector<int> bcindex;           // per boundary face: index into cfg.bcs
	std::vector<double> wlsV;           // nelem x 4: inverse least-squares matrices
	std::vector<double> clength;        // nelem: Venkatakrishnan characteristic length
};

void disc_init(Disc& d, const Mesh* m, const Config& cfg);
std::array<double,NV> freestream_state(const Config& c);

/// -r(u) accumulated into r; optionally the local time steps. parallel=1 runs the face loops
/// with OpenMP threads and atomic scatters (the reference's own structure).
void compute_residual(const Disc& d, const double* u, double* r, bool gettimesteps, double* dt,
                      int parallel);
void flux_magnitude(const Disc& d, const double* u, double* s);
/// Cell-centred gradients of PRIMITIVE variables as used inside the residual (nelem x 8,
/// layout [cell][dim][var]); also returns the limited face states if requested.
void compute_primitive_gradients(const Disc& d, const double* u, double* grads, int parallel);
/// Gradients of conserved variables (src/spatial/flow_spatial.cpp:95-112)
void compute_conserved_gradients(const Disc& d, const double* u, double* grads);

/// Approximate first-order Jacobian in BSR form with sorted columns; the pattern is
/// cell + interior neighbours. Returns ADD-style assembly of dr/du (without pseudo-time term).
struct BSR {
	int n = 0;
	std::vector<int> rowptr, colind, diag;
	std::vector<double> vals;           // nnzb x 16 row-major blocks
};
void bsr_pattern(const Mesh& m, BSR& A);
void assemble_jacobian(const Disc& d, const double* u, BSR& A);
void compute_local_jacobian_interior(const Disc& d, int iface, const double* ul, const double* ur,
                                     double* L, double* U);
void compute_local_jacobian_boundary(const Disc& d, int iface, const double* ul, double* L);

/// Numerical flux and its exact (or reference-approximate) Jacobian for one face.
void inviscid_flux(const Config& c, int fluxid, const double* ul, const double* ur,
                   const double* n, double* f);
void inviscid_flux_jacobian(const Config& c, int fluxid, const double* ul, const double* ur,
                            const double* n, double* dfdl, double* dfdr);
void ghost_state(const Disc& d, int ib, const double* uin, const double* n, double* ug);
void ghost_state_jacobian(const Disc& d, int ib, const double* uin, const double* n,
                          double* ug, double* dug);

/// Output functionals (src/spatial/aoutput.cpp:28-62, src/spatial/flow_spatial.cpp:130-310)
double entropy_error(const Disc& d, const double* u);
void surface_functionals(const Disc& d, const double* u, int walltag, double* cl, double* cdp,
                         double* cdsf);

/* ---------------------------------------------------------------- solvers ------------- */

enum PCType { PC_NONE = 0, PC_JACOBI = 1, PC_SGS = 2, PC_ILU0 = 3 };

struct SolverConfig {
	double cflinit, cflfin, tol;
	int maxiter;
	int implicit;            ///< 1 backward Euler, 0 forward Euler
	int fullupdate;          ///< 1 FULL, 0 ROBUST_FLOW
	double min_relax;
	int pc;                  ///< PCType
	int krylov_restart;      ///< GMRES restart length / max iterations per pseudo-step
	double krylov_rtol;
	int matrix_free;         ///< 1: matrix-free operator with FD step eps, M as preconditioner
	double mf_eps;
	double fixed_norm;       ///< >0: use this normaliser instead of the first-step norm
	int pc_sweeps;           ///< preconditioner applications per Krylov step (>= 1)
};

struct SolveStats {
	int steps = 0, lin_iters = 0, converged = 0, status = 0;
	double final_rel = 0, final_abs = 0, walltime = 0;
};

/// Steady pseudo-time solve, restating src/ode/aodesolver.cpp:135-282 (explicit) and
/// :363-638 (implicit); linear solves by the oracle's own GMRES (PETSc stand-in).
SolveStats steady_solve(const Disc& d, const SolverConfig& sc, double* u);

/// Block ILU(0) / SGS / Jacobi preconditioner on a BSR matrix (host reference semantics).
struct BlockPC {
	int type = PC_NONE;
	int sweeps = 1;                 // stationary iterations z += M^-1 (r - A z) from z = 0
	BSR A;                          // the matrix (for sweeps > 1)
	BSR F;                          // factored copy (ILU0) or the matrix (SGS)
	std::vector<double> dinv;       // inverted diagonal blocks
};
void pc_setup(BlockPC& p, const BSR& A, int type);
void pc_apply(const BlockPC& p, const double* r, double* z);
void bsr_matvec(const BSR& A, const double* x, double* y);
void invert4(const double* a, double* ainv);

}
#endif
