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

/* ---- timing: the TimingData role of the reference (src/ode/aodesolver.hpp:46-67) ---- */

/* Kernel ids for fvx_profile_read. */
enum { FVX_KERNEL_GRADIENTS = 0, FVX_KERNEL_FLUX = 1, FVX_KERNEL_JACOBIAN = 2,
       FVX_KERNEL_PC = 3 };

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
