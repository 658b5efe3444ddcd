/* fvx.h -- C ABI of the MI355X-native residual / Jacobian / preconditioner engine.
 *
 * The engine replaces the FVENS hot path behind the reference's own C++ API. Each entry point
 * below names the reference interface it stands in for (paths relative to the FVENS source
 * tree). All functions return an fvx_status; 0 is success (the reference's StatusCode
 * convention, src/aconstants.hpp:93) and fvx_last_error() describes the last failure on the
 * calling thread. No C++ exceptions cross this boundary.
 *
 * Memory: state vectors are device pointers to AoS double arrays [ncell][4] of conserved
 * variables (rho, rho*vx, rho*vy, rho*E), the reference's Vec layout. Scalars per cell
 * (time steps) are device double arrays [ncell]. Every launch goes to the caller's HIP stream
 * (NULL = the default stream). A handle is not thread-safe: as in the reference
 * (src/spatial/flow_spatial.hpp:196-197) it must be used by one thread at a time.
 */
#ifndef FVX_H
#define FVX_H

#ifdef __cplusplus
extern "C" {
#endif

typedef int fvx_status;
enum {
	FVX_OK = 0,
	FVX_ERR_INVALID = 1,        /* bad argument or unsupported option */
	FVX_ERR_HIP = 2,            /* a HIP runtime call failed */
	FVX_ERR_NONFINITE = 3,      /* non-finite residual: maps to Numerical_error */
	FVX_ERR_NOT_CONVERGED = 4,  /* iteration limit hit: maps to Tolerance_error */
	FVX_ERR_UNSUPPORTED = 5,    /* the reference has no implementation of this option */
	FVX_ERR_ALLOC = 6           /* device allocation failed */
};

/* Numerical-method identifiers (names of the reference factories, src/utilities/afactory.cpp:30-294). */
enum { FVX_FLUX_ROE = 0, FVX_FLUX_HLLC = 1, FVX_FLUX_HLL = 2, FVX_FLUX_LLF = 3,
       FVX_FLUX_AUSM = 4, FVX_FLUX_AUSMPLUS = 5, FVX_FLUX_VANLEER = 6 };
enum { FVX_GRAD_NONE = 0, FVX_GRAD_GREENGAUSS = 1, FVX_GRAD_LEASTSQUARES = 2 };
enum { FVX_RECON_NONE = 0, FVX_RECON_VANALBADA = 1, FVX_RECON_VENKATAKRISHNAN = 2,
       FVX_RECON_BARTHJESPERSEN = 3, FVX_RECON_WENO = 4 };
/* Boundary kinds in the order of the reference enum (src/spatial/abctypes.hpp:13-21). */
enum { FVX_BC_SLIPWALL = 0, FVX_BC_FARFIELD = 1, FVX_BC_INOUTFLOW = 2,
       FVX_BC_SUBSONIC_INFLOW = 3, FVX_BC_EXTRAPOLATION = 4, FVX_BC_PERIODIC = 5,
       FVX_BC_ISOTHERMALWALL = 6, FVX_BC_ADIABATICWALL = 7 };

typedef struct fvx_mesh_s* fvx_mesh;
typedef struct fvx_disc_s* fvx_disc;

/* Raw hybrid 2-D mesh as produced by a mesh reader (src/mesh/meshreaders.cpp:66-266):
 * cells are CCW, triangles padded with -1 in inpoel[4*i+3]. */
typedef struct {
	int npoin, nelem, nbface;
	const double* coords;   /* [npoin][2] */
	const int* inpoel;      /* [nelem][4] */
	const int* nnode;       /* [nelem] 3 or 4 */
	const int* bface;       /* [nbface][2] boundary edge nodes */
	const int* btag;        /* [nbface] physical boundary marker */
} fvx_mesh_desc;

/* One boundary condition (src/spatial/abc.hpp:34-40). */
typedef struct {
	int tag;
	int kind;
	double vals[2];
} fvx_bc_cfg;

/* FlowPhysicsConfig (src/spatial/flow_spatial.hpp:33-44). aoa in radians. */
typedef struct {
	double gamma, Minf, Tinf, Reinf, Pr, aoa;
	int viscous, const_visc;
	int nbc;
	fvx_bc_cfg bcs[8];
} fvx_physics_cfg;

/* FlowNumericsConfig (src/spatial/flow_spatial.hpp:47-55); the limiter parameter is explicit
 * (Venkatakrishnan K, WENO lambda). */
typedef struct {
	int conv_numflux, conv_numflux_jac;
	int gradientscheme, reconstruction;
	double limiter_param;
	int order2;
} fvx_numerics_cfg;

/* ---- mesh: UMesh construction and preprocessing (src/mesh/ameshutils.cpp:38-153) ---- */

/* Reverse Cuthill-McKee cell ordering of the cell-adjacency graph; perm[new] = old.
 * Replaces reorderMeshPetsc with "-mesh_reorder rcm" (src/mesh/ameshutils.cpp:245-288). */
fvx_status fvx_mesh_rcm(const fvx_mesh_desc* desc, int* perm);

/* Lines of cells whose neighbour-distance anisotropy exceeds threshold, grown from the
 * boundary cells in boundary-face order (findLines, src/mesh/meshordering.cpp:199-262). Line l
 * holds line_cells[line_ptr[l] .. line_ptr[l+1]); line_ptr needs nelem/2+1 entries at most,
 * line_cells nelem. */
fvx_status fvx_mesh_lines(const fvx_mesh_desc* desc, double threshold, int* nlines,
                          int* line_ptr, int* line_cells);

/* Line orderings, perm[new] = old: graph_ordering FVX_LINES_ONLY is "-mesh_reorder line"
 * (lineReorder, meshordering.cpp:33-57); FVX_LINES_NATURAL and FVX_LINES_RCM are
 * "-mesh_reorder line_natural" / "line_rcm" (getHybridLineOrdering, :68-128, with the
 * line-point graph of :264-385 ordered by reverse Cuthill-McKee as PETSc's MatGetOrdering
 * "rcm" does). */
enum { FVX_LINES_ONLY = 0, FVX_LINES_NATURAL = 1, FVX_LINES_RCM = 2 };

/* Cells sorted along a Hilbert curve through their centres, perm[new] = old. No counterpart in
 * the reference (its orderings are RCM and the line orderings): a compact-patch order for
 * callers that want blocks of consecutive cells to share most of their faces. */
fvx_status fvx_mesh_hilbert(const fvx_mesh_desc* desc, int* perm);
fvx_status fvx_mesh_line_order(const fvx_mesh_desc* desc, double threshold, int graph_ordering,
                               int* perm);

/* Builds topology (UMesh::compute_topological, src/mesh/mesh.cpp:330-340; face numbering of
 * compute_faceConnectivity :659-762), geometry (compute_areas, compute_face_data,
 * compute_cell_centres :290-365) and uploads the device mesh. */
fvx_status fvx_mesh_create(const fvx_mesh_desc* desc, fvx_mesh* out);
/* sizes[0..4] = npoin, nelem, nbface, ninface, naface */
fvx_status fvx_mesh_sizes(fvx_mesh m, int* sizes);
/* Host copies of the face structure: intfac [naface][4] (L, R, node0, node1),
 * facemetric [naface][3] (nx, ny, len), area [nelem], centres [nelem][2]. Any may be NULL. */
fvx_status fvx_mesh_faces(fvx_mesh m, int* intfac, double* facemetric, double* area,
                          double* centres);
fvx_status fvx_mesh_destroy(fvx_mesh m);

/* ---- spatial discretisation: FlowFV_base / FlowFV (src/spatial/flow_spatial.hpp:61-320) ---- */

/* create_const_flowSpatialDiscretization (src/utilities/afactory.cpp:251-275). The mesh must
 * outlive the discretisation (src/spatial/aspatial.hpp:102). */
fvx_status fvx_disc_create(fvx_mesh m, const fvx_physics_cfg* p, const fvx_numerics_cfg* n,
                           fvx_disc* out);
fvx_status fvx_disc_destroy(fvx_disc d);

/* IdealGasPhysics::compute_freestream_state (src/physics/aphysics.cpp:44-58), host out[4]. */
fvx_status fvx_freestream_state(fvx_disc d, double* out);

/* FlowFV::compute_residual (src/spatial/flow_spatial.cpp:636-816): ADDS -r(u) into r;
 * if gettimesteps, overwrites dt with the local pseudo-time steps. u, r, dt are device
 * pointers; stream is a hipStream_t. */
fvx_status fvx_compute_residual(fvx_disc d, const double* u, double* r, int gettimesteps,
                                double* dt, void* stream);

/* FlowFV_base::getGradients (src/spatial/flow_spatial.cpp:95-112): gradients of the
 * CONSERVED variables, device array [ncell][4 vars][2 dims] (the reference's GradBlock_t). */
fvx_status fvx_get_gradients(fvx_disc d, const double* u, double* grads, void* stream);

/* Gradients of the primitive variables as used inside compute_residual (same layout). */
fvx_status fvx_get_primitive_gradients(fvx_disc d, const double* u, double* grads,
                                       void* stream);

/* Output functionals: out[0..3] = CL, CDp, CDsf (FlowFV_base::computeSurfaceData,
 * src/spatial/flow_spatial.cpp:130-310, normalised by wall length) and the entropy error
 * (FlowOutput::compute_entropy_cell, src/spatial/aoutput.cpp:28-62). out is host memory;
 * the call synchronises the stream. */
fvx_status fvx_functionals(fvx_disc d, const double* u, int wall_tag, double* out,
                           void* stream);

/* Surface data on one wall marker: per boundary face carrying the marker, in boundary-face
 * order, the face centre (x, y), Cp and Cf from the cell-centred state -- the output matrix of
 * FlowFV_base::computeSurfaceData (src/spatial/flow_spatial.cpp:130-310) that
 * FlowOutput::exportSurfaceData writes (src/spatial/aoutput.cpp:181-240). Host out[cap][4];
 * *nfaces receives the number of faces (out may be NULL to query it). Synchronises the stream. */
fvx_status fvx_surface_data(fvx_disc d, const double* u, int wall_tag, double* out, int cap,
                            int* nfaces, void* stream);

/* ---- implicit solve: system matrix, preconditioners, Krylov solver, pseudo-time loop ---- */

/* Storage precision of matrices and preconditioner factors; vectors are always double.
 * FVX_PREC_BF16 (preconditioner factors only: 8-bit mantissa, FP32's exponent range) halves
 * the bytes a preconditioner application reads against FP32; the factors are computed in FP64
 * and rounded to nearest even when stored. */
enum { FVX_PREC_F64 = 0, FVX_PREC_F32 = 1, FVX_PREC_BF16 = 2 };
/* Preconditioners (PETSc -pc_type none / pbjacobi / sor / bjacobi+ilu on one rank). SGS and
 * ILU0 run in a multicolour ordering of the cells (greedy colouring in cell order; cells of
 * one colour share no face), which makes each sweep of a colour fully parallel. */
/* FVX_PC_LINE: symmetric line Gauss-Seidel. Its blocks are the lines of fvx_mesh_lines at the
 * preconditioner's anisotropy threshold (default 10) and every other cell on its own, the
 * vertices of the reference's line-point graph (src/mesh/meshordering.cpp:264-385); each line's
 * block-tridiagonal system is solved exactly (block Thomas). The reference reaches the same
 * effect through "-mesh_reorder line" followed by PETSc's ILU or SOR in that order. */
/* FVX_PC_SGS_LEVEL and FVX_PC_ILU0_LEVEL: symmetric Gauss-Seidel and block ILU(0) in the natural
 * (cell) order, the reference's PETSc -pc_type sor and bjacobi + ilu on one rank
 * (src/ode/aodesolver.cpp:483, testcases/defaults.solverc, testcases/naca0012/opts.solverc). The
 * sweeps run over wavefront levels (cells whose lower-numbered neighbours all lie in earlier
 * levels, src/mesh/ameshutils.cpp:158-196), one parallel launch per level; the result is the
 * sequential natural-order one. fvx_pc_ordering reports the levels as colours. */
/* FVX_PC_BJACOBI_ILU0: block Jacobi over subdomains of about fvx_pc_set_subdomain_cells cells
 * (default 256; compact, grown breadth-first over the cell graph from the lowest-numbered
 * unassigned cell) with block ILU(0) in the natural order inside each subdomain -- the
 * reference's -pc_type bjacobi -sub_pc_type ilu when each rank owns one subdomain
 * (src/ode/aodesolver.cpp:483, testcases/defaults.solverc; PETSc's -pc_bjacobi_blocks). Couplings
 * between subdomains are dropped from the factors; the sweeps option adds stationary
 * iterations with the full matrix. One workgroup holds one subdomain and runs its forward and
 * backward substitution over the subdomain's wavefront levels in LDS: one launch per
 * application. fvx_pc_ordering reports the level of each cell inside its subdomain. */
enum { FVX_PC_NONE = 0, FVX_PC_JACOBI = 1, FVX_PC_SGS = 2, FVX_PC_ILU0 = 3, FVX_PC_LINE = 4,
       FVX_PC_SGS_LEVEL = 5, FVX_PC_ILU0_LEVEL = 6, FVX_PC_BJACOBI_ILU0 = 7 };

typedef struct fvx_matrix_s* fvx_matrix;
typedef struct fvx_pc_s* fvx_pc;

/* Block-sparse 4x4 matrix on the cell graph of a mesh: the SeqBAIJ (bs = 4) system matrix of
 * setupSystemMatrix (src/linalg/alinalg.cpp:90-119). Internally one diagonal block and one
 * block per cell face. */
fvx_status fvx_matrix_create(fvx_mesh m, int precision, fvx_matrix* out);
fvx_status fvx_matrix_destroy(fvx_matrix A);
/* Spatial::assemble_jacobian (src/spatial/aspatial.cpp:242-340) preceded by the caller's
 * MatZeroEntries (src/ode/aodesolver.cpp:456): overwrites A with the first-order approximate
 * Jacobian dr/du at u (device pointer). */
fvx_status fvx_assemble_jacobian(fvx_disc d, const double* u, fvx_matrix A, void* stream);
/* SteadyBackwardEulerSolver::addPseudoTimeTerm (src/ode/aodesolver.cpp:300-329):
 * dtm[i] = area_i/(cfl dt[i]) and A_ii += dtm[i] I. dt, dtm: device [ncell]. */
fvx_status fvx_add_pseudotime(fvx_disc d, fvx_matrix A, const double* dt, double cfl,
                              double* dtm, void* stream);
/* MatMult: y = A x (device vectors [ncell][4]). */
fvx_status fvx_matrix_apply(fvx_matrix A, const double* x, double* y, void* stream);
/* Number of nonzero blocks of the equivalent BAIJ matrix (diagonal + interior faces). */
fvx_status fvx_matrix_bsr_nnzb(fvx_matrix A, int* nnzb);
/* Host copy as block CSR with sorted columns in cell order: rowptr[ncell+1],
 * colind[nnzb], vals[nnzb][4][4] row-major. Synchronises the device. */
fvx_status fvx_matrix_export_bsr(fvx_matrix A, int* rowptr, int* colind, double* vals);

/* PCCreate/PCSetType for a matrix of this mesh; precision = storage of the factors. */
fvx_status fvx_pc_create(fvx_mesh m, int type, int precision, fvx_pc* out);
/* FVX_PC_LINE with the given anisotropy threshold of the lines. */
fvx_status fvx_pc_create_lines(fvx_mesh m, double threshold, int precision, fvx_pc* out);
fvx_status fvx_pc_destroy(fvx_pc p);
/* Number of stationary iterations z <- z + M^-1 (r - A z) from z = 0 per application
 * (PETSc -pc_sor_its; for SGS, symmetric forward + backward sweeps). Default 1, at most 64. */
fvx_status fvx_pc_set_sweeps(fvx_pc p, int sweeps);
/* PCSetUp: factorises / inverts from the current values of A; later applications use A for
 * the sweeps, so A must stay alive and unchanged until the next set-up. */
fvx_status fvx_pc_setup(fvx_pc p, fvx_matrix A, void* stream);
/* PCApply: z = M^-1 r (device vectors [ncell][4]). */
fvx_status fvx_pc_apply(fvx_pc p, const double* r, double* z, void* stream);
/* Host copy of the colour of every cell and the number of colours of the PC's ordering. */
fvx_status fvx_pc_ordering(fvx_pc p, int* colour, int* ncolours);
/* FVX_PC_BJACOBI_ILU0: cells per subdomain, in [16, 1024] (default 256); re-partitions the
 * mesh, so a set-up must follow (the role of PETSc's -pc_bjacobi_blocks, given as a size). */
fvx_status fvx_pc_set_subdomain_cells(fvx_pc p, int cells);
/* FVX_PC_BJACOBI_ILU0: host copy of the subdomain of every cell and the subdomain count. */
fvx_status fvx_pc_subdomains(fvx_pc p, int* part, int* nparts);

/* KSPSolve with -ksp_type fgmres (src/ode/aodesolver.cpp:483; options
 * testcases/defaults.solverc): one restart cycle of at most `restart` iterations of right-
 * preconditioned flexible GMRES with classical Gram-Schmidt from x = 0, stopping when
 * ||b - A x|| <= rtol ||b||. x is overwritten; iters receives the iteration count. */
fvx_status fvx_linear_solve(fvx_matrix A, fvx_pc p, const double* b, double* x, int restart,
                            double rtol, int* iters, void* stream);

/* MatrixFreeSpatialJacobian::apply (src/linalg/alinalg.cpp:142-233):
 * y = dtm x + (r(u + eps x/||x||) - r(u)) ||x||/eps, where r_u holds the caller's residual
 * buffer at u in the engine's sign convention (-r(u), as fvx_compute_residual adds it). */
fvx_status fvx_matfree_apply(fvx_disc d, const double* u, const double* r_u, const double* dtm,
                             const double* x, double* y, double eps, void* stream);

/* Largest mesh (cells) of fvx_matfree_apply_noalloc. */
#define FVX_NOALLOC_MAX_CELLS 131072
/* fvx_matfree_apply without any device allocation. No counterpart in the reference: it is for
 * callers that cannot allocate device memory (code under HIP graph capture, serving loops,
 * harnesses that forbid hipMalloc). Same arguments and result bits as fvx_matfree_apply; its
 * scratch (perturbed state, its residual, norm partial sums: 8.5 MB) is a static device array
 * that every GPU the library runs on holds until the process ends. Meshes above
 * FVX_NOALLOC_MAX_CELLS cells are rejected with FVX_ERR_INVALID. Every call on one GPU uses the
 * same scratch: calls that overlap on one GPU (different streams, host threads) race. */
fvx_status fvx_matfree_apply_noalloc(fvx_disc d, const double* u, const double* r_u,
                                     const double* dtm, const double* x, double* y, double eps,
                                     void* stream);

/* SteadySolverConfig (src/ode/aodesolver.hpp:16-42) plus the linear-solver options. */
typedef struct {
	double cflinit, cflfin;   /* CFL ramp bounds */
	double tol;               /* stop when ||r||/||r_0|| <= tol */
	int maxiter;              /* pseudo-time step limit */
	int implicit;             /* 1 backward Euler, 0 forward Euler */
	int fullupdate;           /* 1 FULL update, 0 ROBUST_FLOW local relaxation */
	double min_relax;         /* minimum local relaxation factor */
	int pc;                   /* FVX_PC_* */
	int pc_precision;         /* FVX_PREC_* of the preconditioner factors */
	int pc_sweeps;            /* stationary iterations per application (0 or 1: one) */
	int jac_precision;        /* FVX_PREC_* of the assembled Jacobian */
	int krylov_restart;       /* -ksp_max_it (one FGMRES cycle) */
	double krylov_rtol;       /* -ksp_rtol */
	int matrix_free;          /* 1: matrix-free operator, assembled Jacobian only in the PC */
	double mf_eps;            /* matrix-free difference step */
	double fixed_norm;        /* > 0: normalise residuals by this instead of the first norm */
	double line_threshold;    /* FVX_PC_LINE: anisotropy threshold of the lines (<= 0: 10) */
	int pc_subdomain_cells;   /* FVX_PC_BJACOBI_ILU0: cells per subdomain (<= 0: 256) */
} fvx_solver_cfg;

typedef struct {
	int steps, lin_iters, converged, status;
	double final_rel, final_abs, first_norm, walltime;
} fvx_solve_stats;

/* SteadyForwardEulerSolver::solve / SteadyBackwardEulerSolver::solve
 * (src/ode/aodesolver.cpp:135-282, 363-638) run on the device: u (device [ncell][4]) is
 * updated in place. Returns FVX_ERR_NONFINITE on a non-finite residual norm and
 * FVX_ERR_NOT_CONVERGED when maxiter is reached; stats are filled in either case. */
fvx_status fvx_steady_solve(fvx_disc d, const fvx_solver_cfg* cfg, double* u,
                            fvx_solve_stats* stats, void* stream);

/* ---- batched operating points ----
 * No reference counterpart: the reference runs one operating point per process. A batch holds
 * ninst operating points (Mach number, angle of attack in radians) on one discretisation;
 * every other setting is the discretisation's. Batched state vectors are instance-major device
 * arrays [ninst][ncell][4]. Each launch covers all active instances, while every instance keeps
 * its own Krylov iterations, CFL ramp, convergence test and retirement, so instance i evolves
 * exactly as a single solve of the same operating point would. */
typedef struct fvx_batch_s* fvx_batch;
fvx_status fvx_batch_create(fvx_disc d, int ninst, const double* mach, const double* aoa,
                            fvx_batch* out);
fvx_status fvx_batch_destroy(fvx_batch b);
/* Free-stream states, host out[ninst][4]. */
fvx_status fvx_batch_freestream(fvx_batch b, double* out);
/* fvx_compute_residual for every instance (u, r: [ninst][ncell][4]; dt: [ninst][ncell]). */
fvx_status fvx_batch_compute_residual(fvx_batch b, const double* u, double* r, int gettimesteps,
                                      double* dt, void* stream);
/* One pseudo-time step of a solve: the reference's SteadyStepMonitor
 * (src/ode/aodesolver.hpp:35-43, filled at src/ode/aodesolver.cpp:536-549): step number (from
 * 1), relative and absolute residual norm at the start of the step, elapsed wall time and
 * cumulative linear-solver wall time (seconds), cumulative linear iterations, CFL number. */
typedef struct {
	int step;
	double rel, abs, walltime, lin_walltime;
	int lin_iters;
	double cfl;
} fvx_step_record;
/* Step records of the last steady solve on d that had one instance (host out[cap]); *n receives
 * the number of steps (records beyond cap are not written). */
fvx_status fvx_solve_history(fvx_disc d, fvx_step_record* out, int cap, int* n);

/* fvx_steady_solve for every instance; stats[ninst] receives each instance's outcome in its
 * status field (FVX_OK, FVX_ERR_NONFINITE or FVX_ERR_NOT_CONVERGED); a failing instance does
 * not stop the others, and the call itself returns FVX_OK unless an argument or the device
 * fails. */
fvx_status fvx_batch_steady_solve(fvx_batch b, const fvx_solver_cfg* cfg, double* u,
                                  fvx_solve_stats* stats, void* stream);
/* Fixed normalisers of the convergence test, one per instance (host norms[ninst]): instance i
 * has converged when its residual norm falls to cfg.tol x norms[i]. NULL restores the
 * reference's normaliser, the norm of the first step (src/ode/aodesolver.cpp:536-537). No
 * reference counterpart (a fixed normaliser makes warm and cold starts comparable). */
fvx_status fvx_batch_set_norms(fvx_batch b, const double* norms);
/* The convergence norm sqrt(sum_cells r_energy^2 area) of every instance at u (the reference's
 * residual norm, src/ode/aodesolver.cpp:516-530), host out[ninst]. */
fvx_status fvx_batch_residual_norms(fvx_batch b, const double* u, double* out, void* stream);
/* fvx_functionals for every instance, host out[ninst][4]. */
fvx_status fvx_batch_functionals(fvx_batch b, const double* u, int wall_tag, double* out,
                                 void* stream);

/* ---- autotuning (no reference counterpart: OpenMP loops have no launch shape) ---- */

/* One timing: kernel id (FVX_KERNEL_*), threads per workgroup, XCD-aware block order
 * (0/1) and milliseconds per launch. For fvx_autotune_pc: kernel = FVX_PC_* type, block =
 * sweeps, xcd = Krylov iterations, ms = set-up + solve time (negative: did not reach rtol). */
typedef struct {
	int kernel, block, xcd;
	double ms;
} fvx_tune_row;

/* Times every launch shape of the gradient, face-flux and Jacobian kernels at state u and
 * keeps the fastest in d (results are bitwise identical for every shape). rows[cap]. */
fvx_status fvx_autotune(fvx_disc d, const double* u, int reps, fvx_tune_row* rows, int cap,
                        int* nrows, void* stream);
/* Sets / reads the launch shape of one kernel (block 64, 128, 256 or 512; 256 at most for the
 * Jacobian). */
fvx_status fvx_set_launch(fvx_disc d, int kernel, int block, int xcd);
fvx_status fvx_get_launch(fvx_disc d, int kernel, int* block, int* xcd);
/* Times preconditioner types and sweep counts on the pseudo-time linear system at u and the
 * given CFL number (set-up + one FGMRES(restart) solve to rtol) and returns the fastest that
 * reaches rtol. */
fvx_status fvx_autotune_pc(fvx_disc d, const double* u, double cfl, int restart, double rtol,
                           int precision, fvx_tune_row* rows, int cap, int* nrows, int* best_pc,
                           int* best_sweeps, void* stream);

/* ---- learned warm start (no reference counterpart: the reference starts a solve from the
 * freestream or from a first-order starter solve, src/utilities/casesolvers.cpp:225-386) ----
 *
 * A 4-layer message-passing network over the cell graph of a mesh (width 64, bf16 weights,
 * fp32 accumulation) maps (Mach number, angle of attack, cell geometry) to a predicted state;
 * training and inference run on the device. Network outputs per cell are the normalised
 * deviations (rho - 1, vx - cos(aoa), vy - sin(aoa), p/p_inf - 1) from the reference's
 * non-dimensional freestream (src/physics/aphysics.cpp:44-58). Graphs of one call share the mesh;
 * per-graph arrays are graph-major [ngraph][ncell][...]. */
typedef struct fvx_mpnn_s* fvx_mpnn;
/* Number of fp32 parameters (flat layout: Wenc[8][64], benc[64], then per layer Wa, Wb [64][64],
 * We[4][64], bm[64], Wu, Wv [64][64], bu[64], then Wdec[64][4], bdec[4]). */
int fvx_mpnn_nparams(void);
/* Network on mesh m; wall_tag marks the wall faces (wall distance feature); max_graphs bounds
 * the graphs per call. Parameters start at zero: call fvx_mpnn_set_params. */
fvx_status fvx_mpnn_create(fvx_mesh m, int wall_tag, int max_graphs, fvx_mpnn* out);
fvx_status fvx_mpnn_destroy(fvx_mpnn n);
/* Graph arrays (host outs, any may be NULL): static node features [ncell][6], edge features
 * [ncell][4][4], interior neighbour [ncell][4] and reverse slot [ncell][4] (-1 where unused),
 * 1/degree [ncell]. */
fvx_status fvx_mpnn_features(fvx_mpnn n, float* stat, float* edge, int* nbr, int* rev,
                             float* invdeg);
/* Master weights from / to host arrays [nparams]; setting them also resets the optimiser. */
fvx_status fvx_mpnn_set_params(fvx_mpnn n, const float* w);
fvx_status fvx_mpnn_get_params(fvx_mpnn n, float* w);
/* Operating points of the next calls: host mach[ngraph], aoa[ngraph] (radians). */
fvx_status fvx_mpnn_set_points(fvx_mpnn n, int ngraph, const double* mach, const double* aoa,
                               double gamma, void* stream);
/* Network outputs for the first ngraph points, device y[ngraph][ncell][4]. */
fvx_status fvx_mpnn_predict(fvx_mpnn n, int ngraph, float* y, void* stream);
/* Predicted conserved state, device u[ngraph][ncell][4]; density and pressure are floored at
 * `floor` times their freestream values. */
fvx_status fvx_mpnn_warm_state(fvx_mpnn n, int ngraph, double floor, double* u, void* stream);
/* Training targets (the outputs a perfect network would give) of the conserved fields
 * u[ngraph][ncell][4], device t[ngraph][ncell][4]. */
fvx_status fvx_mpnn_targets(fvx_mpnn n, int ngraph, const double* u, float* t, void* stream);
/* Mean-squared loss over the first ngraph points against targets of a data set (device
 * [nsamples][ncell][4]; graph g uses sample ids[g], host array) and its gradient (device
 * grad[nparams], may be NULL); *loss (host, may be NULL) waits for the result. */
fvx_status fvx_mpnn_gradient(fvx_mpnn n, int ngraph, const float* targets, const int* ids,
                             float* grad, float* loss, void* stream);
/* fvx_mpnn_gradient followed by one Adam step (beta 0.9 / 0.999, eps 1e-8) at rate lr. */
fvx_status fvx_mpnn_train_step(fvx_mpnn n, int ngraph, const float* targets, const int* ids,
                               double lr, float* loss, void* stream);

/* ---- timing: the TimingData role of the reference (src/ode/aodesolver.hpp:46-67) ---- */

/* Kernel ids for fvx_profile_read and fvx_set_launch / fvx_get_launch. FVX_KERNEL_FLUX_WAVE
 * is the wave-task form of the face-flux pass (inviscid, first-order or linear reconstruction):
 * each wave64 owns a group of consecutive cells and evaluates each face touching them once;
 * setting a launch for it (block 64) selects that form (FVX_ERR_INVALID for a viscous or MUSCL
 * discretisation), setting one for FVX_KERNEL_FLUX selects the quad-per-cell form. The two
 * forms give bitwise equal residuals and time steps; the wave form's time is read as
 * FVX_KERNEL_FLUX. */
enum { FVX_KERNEL_GRADIENTS = 0, FVX_KERNEL_FLUX = 1, FVX_KERNEL_JACOBIAN = 2,
       FVX_KERNEL_PC = 3, FVX_KERNEL_FLUX_WAVE = 4 };

/* on != 0 resets the counters and starts recording HIP events around every kernel the handle
 * launches (on the launch stream); on == 0 stops recording. */
fvx_status fvx_profile_enable(fvx_disc d, int on);
/* Waits for the recorded events and returns the accumulated device time (ms) and the number
 * of launches of one kernel since profiling was enabled. */
fvx_status fvx_profile_read(fvx_disc d, int kernel, double* total_ms, long* count);

/* Error text of the last failing call on this thread. */
const char* fvx_last_error(void);

#ifdef __cplusplus
}
#endif

#endif
