/* spatial.hip -- gradient / limiter kernel, residual dispatch and output functionals.
 *
 * Gradients (GradientScheme::compute_gradients, src/spatial/agradientschemes.cpp:61-440) are
 * evaluated as precomputed face stencils grad_i = sum_j c_ij (w_j - w_i) over the cell's
 * neighbours (ghost state of the cell value for a boundary face), two lanes per cell each
 * visiting two of its faces. The Venkatakrishnan and Barth-Jespersen limiters
 * (src/spatial/limitedlinearreconstruction.cpp:116-268) are fused into the same pass: the
 * neighbour differences that form the gradient also give the max/min bounds, and the limited
 * gradient phi*grad is what the flux pass reads. Boundary neighbours contribute the ghost
 * state (the reference reads past its cell array there; this definition is deliberate).
 */
#include "fvx_internal.hpp"
#include "wave.hpp"
#include "face_state.hpp"

namespace fvx {

#define FVX_DECL_FLUX(ID) void launch_flux_##ID(int mode, bool visc, int ncell, const double* u, \
	const double* rgrad, const double* grad, const int* nbr, const int* face, const FaceGeom* fg, \
	const double* centre, const double* area, const int* bcid, const GasParams* gpa, \
	const BCTable* bta, double* res, double* dt, int getdt, const BatchArg& ba, int nact, \
	int block, int xcd, hipStream_t s);
FVX_DECL_FLUX(0) FVX_DECL_FLUX(1) FVX_DECL_FLUX(2) FVX_DECL_FLUX(3) FVX_DECL_FLUX(4)
FVX_DECL_FLUX(5) FVX_DECL_FLUX(6)
#undef FVX_DECL_FLUX
#define FVX_DECL_FLUXW(ID) void launch_flux_wave_##ID(int mode, int ncell, int ngroup, \
	const double* u, const double* rgrad, const int* tasks, const int2* groups, \
	const unsigned* slots, const FaceGeom* fg, const double* centre, const double* area, \
	const int* bcid, const GasParams* gpa, const BCTable* bta, double* res, double* dt, \
	int getdt, const BatchArg& ba, int nact, int xcd, hipStream_t s);
FVX_DECL_FLUXW(0) FVX_DECL_FLUXW(1) FVX_DECL_FLUXW(2) FVX_DECL_FLUXW(3) FVX_DECL_FLUXW(4)
FVX_DECL_FLUXW(5) FVX_DECL_FLUXW(6)
#undef FVX_DECL_FLUXW

/// Ghost states of the boundary faces from their cells' states (the reference's
/// compute_boundary_states, src/spatial/flow_spatial.cpp:74-82, called at :679 before the
/// gradients): one thread per boundary face ib, the state at ughost[ib], instance-major like
/// the state.
__global__ __launch_bounds__(256)
void k_ghost(int ncell, int nbface, const double* __restrict__ u, const int* __restrict__ bhost,
             const FaceGeom* __restrict__ fgeom, const int* __restrict__ bcid,
             const GasParams* __restrict__ gpa, const BCTable* __restrict__ bta,
             const BatchArg ba, double* __restrict__ ughost)
{
	const long I = ba.list ? ba.list[blockIdx.y] : 0;
	if(I < 0) return;                     // a retired slot of a compacted list
	const int ib = blockIdx.x*blockDim.x + threadIdx.x;
	if(ib >= nbface) return;
	u += I*4L*ncell;
	ughost += I*4L*nbface;
	double uc[4], g[4];
	load4(u + 4*bhost[ib], uc);
	// boundary face ib is face ib (boundary faces are numbered first)
	const FaceGeom G = fgeom[ib];
	const double n[2] = { G.nx, G.ny };
	ghost_state(gpa[I], *bta, bcid[ib], uc, n, g);
	store4(ughost + 4*ib, g);
}

/// LIM: 0 none, 1 Venkatakrishnan, 2 Barth-Jespersen. CONS: gradients of the conserved
/// variables (getGradients) instead of the primitive ones. gout receives the unlimited
/// gradient (may be null when LIM > 0 and it is not needed), rout the limited one.
/// Two lanes per cell, each owning two of the cell's local faces; the pair combines its
/// partial sums, bounds and limiter values with one DPP exchange. A workgroup owns a range of
/// consecutive cells and first stages their primitive states in LDS, so a neighbour inside the
/// range is read from there instead of being fetched and converted again. A boundary face's
/// neighbour value is its ghost state, computed beforehand by k_ghost: built inline, the ghost
/// state's registers were live beside both neighbour gathers (118 VGPRs, 4 waves per SIMD,
/// against 79 and 6 without it).
///
/// The limiter factor of a cell and variable is the minimum over its faces of a ratio N/D with
/// D >= 0 (Venkatakrishnan: N = dp^2 + 2 dp dm + eps^2, D = dp^2 + dp dm + 2 dm^2 + eps^2;
/// Barth-Jespersen: N/D = (max or min difference)/dm), starting from 1. The minimum is tracked
/// as the pair (N, D) by cross-multiplied comparisons and divided once at the end: the factor is
/// then the rounded quotient of the minimising face, the value the face-by-face division gives
/// (up to a rounding-level tie between two faces). A face with D = 0 (0/0 in the reference,
/// ignored by its minimum) never compares smaller.
template <int LIM, bool CONS>
__global__ __launch_bounds__(MAX_BLOCK)
void k_gradients(int ncell, const double* __restrict__ u, const int* __restrict__ nbr,
                 const double* __restrict__ soff, const double* __restrict__ gcoef,
                 const double* __restrict__ eps2, const GasParams* __restrict__ gpa,
                 const BatchArg ba, double* __restrict__ gout, double* __restrict__ rout,
                 const int xcd, int nbface, const double* __restrict__ ughost)
{
	// instance: state and gradients are instance-major, 4 and 8 values per cell; gpa holds
	// the gas parameters of each instance (one entry for a single solve)
	const long I = ba.list ? ba.list[blockIdx.y] : 0;
	if(I < 0) return;                     // a retired slot of a compacted list
	const GasParams& gp = gpa[I];
	u += I*4L*ncell;
	ughost += I*4L*nbface;
	if(gout) gout += I*8L*ncell;
	if(rout) rout += I*8L*ncell;
	const int p = threadIdx.x & 1;
	const int cpb = (int)(blockDim.x >> 1);
	const int c0 = xcd_block(xcd)*cpb;
	const int lc = threadIdx.x >> 1;
	const int nloc = min(cpb, ncell - c0);
	const int cell = c0 + lc;
	const bool valid = lc < nloc;
	const int cc = valid ? cell : c0;
	__shared__ double sw[4][MAX_BLOCK/2];

	double ui[4], wi[4];
	load4(u + 4*cc, ui);
	if constexpr(CONS) { for(int k = 0; k < 4; k++) wi[k] = ui[k]; }
	else prim_lean(gp, ui, wi);
	if(valid && p == 0) for(int k = 0; k < 4; k++) sw[k][lc] = wi[k];
	const int2 nb2 = reinterpret_cast<const int2*>(nbr)[2*cc + p];
	const int nbs[2] = { valid ? nb2.x : SLOT_UNUSED, valid ? nb2.y : SLOT_UNUSED };
	const double2* cf = reinterpret_cast<const double2*>(gcoef) + 4*cc + 2*p;
	// neighbours outside the workgroup's range: both gathers in flight together (any other
	// slot loads the cell itself as a placeholder)
	bool outside[2];
	double uj[2][4];
#pragma unroll
	for(int k = 0; k < 2; k++) {
		const int nb = nbs[k];
		outside[k] = nb >= 0 && (unsigned)(nb - c0) >= (unsigned)nloc;
		// boundary face ib's ghost state (k_ghost) is at ughost[ib]
		const double* src = nb >= 0 ? u + 4*(outside[k] ? nb : cc)
		                  : nb != SLOT_UNUSED ? ughost + 4*(-1 - nb) : u + 4*cc;
		load4(src, uj[k]);
	}
	const double2 c0v = cf[0], c1v = cf[1];
	// face midpoint minus cell centre of the two slots (zero for an unused slot)
	double2 off[2] = { make_double2(0, 0), make_double2(0, 0) };
	if constexpr(LIM > 0) {
		const double2* so = reinterpret_cast<const double2*>(soff) + 4*cc + 2*p;
		off[0] = so[0];
		off[1] = so[1];
	}
	__syncthreads();

	double g[8] = {0, 0, 0, 0, 0, 0, 0, 0};
	double dmax[4] = {0, 0, 0, 0}, dmin[4] = {0, 0, 0, 0};
#pragma unroll
	for(int k = 0; k < 2; k++) {
		const int nb = nbs[k];
		if(nb == SLOT_UNUSED) continue;
		double wj[4];
		if(nb >= 0 && !outside[k]) {
			for(int v = 0; v < 4; v++) wj[v] = sw[v][nb - c0];
		} else {
			if constexpr(CONS) { for(int v = 0; v < 4; v++) wj[v] = uj[k][v]; }
			else prim_lean(gp, uj[k], wj);
		}
		const double2 c = k == 0 ? c0v : c1v;
		for(int v = 0; v < 4; v++) {
			const double dw = wj[v] - wi[v];
			g[2*v] += c.x*dw;
			g[2*v+1] += c.y*dw;
			if constexpr(LIM > 0) { dmax[v] = fmax(dmax[v], dw); dmin[v] = fmin(dmin[v], dw); }
		}
	}
	for(int k = 0; k < 8; k++) g[k] = pair_sum(g[k]);

	if constexpr(LIM == 0) {
		if(valid) store4(gout + 8*cc + 4*p, g + 4*p);
	} else {
		const double e2 = LIM == 1 ? eps2[cc] : 0.0;
		double lg[8];
		for(int v = 0; v < 4; v++) {
			const double mx = pair_max(dmax[v]), mn = pair_min(dmin[v]);
			double bn = 1.0, bd = 1.0;          // the smallest ratio so far, as N / D
			for(int k = 0; k < 2; k++) {
				if(nbs[k] == SLOT_UNUSED) continue;
				// face value by linear extrapolation, then its difference from the cell value
				double uf = wi[v];
				uf += g[2*v]*off[k].x;
				uf += g[2*v+1]*off[k].y;
				const double dm = uf - wi[v];
				double pn, pd;
				if constexpr(LIM == 1) {
					const double dp = dm < 0 ? mn : mx;
					pn = dp*dp + 2*dp*dm + e2;
					pd = dp*dp + dp*dm + 2*dm*dm + e2;
				} else {
					if(dm > 0) { pn = mx; pd = dm; }
					else if(dm < 0) { pn = -mn; pd = -dm; }
					else { pn = 1.0; pd = 1.0; }
				}
				if(pn*bd < bn*pd) { bn = pn; bd = pd; }
			}
			// the pair's two candidates, compared the same way on both lanes
			const double on = dpp_d<DPP_XOR1>(bn), od = dpp_d<DPP_XOR1>(bd);
			if(p == 0) { if(on*bd < bn*od) { bn = on; bd = od; } }
			else { if(!(bn*od < on*bd)) { bn = on; bd = od; } }
			const double phi = bn/bd;
			lg[2*v] = phi*g[2*v];
			lg[2*v+1] = phi*g[2*v+1];
		}
		if(valid) {
			if(gout) store4(gout + 8*cc + 4*p, g + 4*p);
			store4(rout + 8*cc + 4*p, lg + 4*p);
		}
	}
}

/// WENO-weighted gradients (WENOReconstruction::compute_face_values,
/// src/spatial/limitedlinearreconstruction.cpp:39-98): per cell and primitive variable v,
/// g' = (lambda g_i / (|g_i|^2 + eps)^4 + sum_j g_j / (|g_j|^2 + eps)^4) / (sum of the weights)
/// over the cell's non-boundary neighbours j, eps = 1e-5. Face values are then the linear
/// extrapolation with g', which is the flux pass's limited-gradient input. Four lanes per cell,
/// one per variable; each lane reads the 16-byte x/y pair of its variable from every gradient.
__global__ __launch_bounds__(MAX_BLOCK)
void k_weno(int ncell, const double* __restrict__ grad, const int* __restrict__ nbr,
            double lambda, const BatchArg ba, double* __restrict__ rout)
{
	const long I = ba.list ? ba.list[blockIdx.y] : 0;
	if(I < 0) return;                     // a retired slot of a compacted list
	grad += I*8L*ncell;
	rout += I*8L*ncell;
	const int v = threadIdx.x & 3;
	const int cell = blockIdx.x*(int)(blockDim.x >> 2) + (threadIdx.x >> 2);
	if(cell >= ncell) return;
	const int4 nb = reinterpret_cast<const int4*>(nbr)[cell];
	const int nbs[4] = { nb.x, nb.y, nb.z, nb.w };
	double2 gj[4];
#pragma unroll
	for(int k = 0; k < 4; k++)
		gj[k] = reinterpret_cast<const double2*>(grad)[4*(nbs[k] >= 0 ? nbs[k] : cell) + v];
	const double2 gi = reinterpret_cast<const double2*>(grad)[4*cell + v];
	const double eps = 1.0e-5;
	double m = gi.x*gi.x + gi.y*gi.y + eps;
	m *= m;
	double w = lambda/(m*m);
	double wsum = w, lx = w*gi.x, ly = w*gi.y;
#pragma unroll
	for(int k = 0; k < 4; k++) {
		if(nbs[k] < 0) continue;            // boundary (ghost) or unused slot
		double mj = gj[k].x*gj[k].x + gj[k].y*gj[k].y + eps;
		mj *= mj;
		const double wj = 1.0/(mj*mj);
		wsum += wj;
		lx += wj*gj[k].x;
		ly += wj*gj[k].y;
	}
	reinterpret_cast<double2*>(rout)[4*cell + v] = make_double2(lx/wsum, ly/wsum);
}

void launch_weno(const fvx_disc_s& d, const double* grad, double* rgrad, hipStream_t s,
                 const BatchArg& ba, int nact)
{
	const int ncell = d.m->nelem, block = 256;
	hipLaunchKernelGGL(k_weno, dim3((ncell + block/4 - 1)/(block/4), nact), dim3(block), 0, s,
	                   ncell, grad, d.m->d_nbr.p, d.nc.limiter_param, ba, rgrad);
	FVX_HIP(hipGetLastError());
}

/* ------------------------------------------------------------------ profiler ----------- */

void Profiler::start(int k, hipStream_t s)
{
	if(!on) return;
	if(2*used + 2 > pool.size()) {
		for(int i = 0; i < 2; i++) {
			hipEvent_t e;
			FVX_HIP(hipEventCreate(&e));
			pool.push_back(e);
		}
		kid.push_back(0);
	}
	kid[used] = k;
	FVX_HIP(hipEventRecord(pool[2*used], s));
}

void Profiler::stop(hipStream_t s)
{
	if(!on) return;
	FVX_HIP(hipEventRecord(pool[2*used+1], s));
	used++;
}

void Profiler::collect()
{
	for(size_t i = 0; i < used; i++) {
		FVX_HIP(hipEventSynchronize(pool[2*i+1]));
		float ms = 0;
		FVX_HIP(hipEventElapsedTime(&ms, pool[2*i], pool[2*i+1]));
		total_ms[kid[i]] += ms;
		count[kid[i]]++;
	}
	used = 0;
}

void Profiler::reset()
{
	used = 0;
	for(int k = 0; k < NKERNEL; k++) { total_ms[k] = 0; count[k] = 0; }
}

Profiler::~Profiler()
{
	for(hipEvent_t e : pool) (void)hipEventDestroy(e);
}

void launch_gradients(const fvx_disc_s& d, const double* u, double* grad, double* rgrad,
                      int conserved, hipStream_t s, const BatchArg& ba, int nact, double* ughost)
{
	const fvx_mesh_s& m = *d.m;
	if(!ughost) {
		// the discretisation's buffer holds one instance (a single-instance launch, whose list,
		// if any, names instance 0)
		if(nact != 1)
			throw Error(FVX_ERR_INVALID, "a batched gradient launch needs its ghost-state buffer");
		ughost = d.d_ughost.p;
	}
	const int block = d.tune.grad_block, cpb = block/2;
	const int nblk = (m.nelem + cpb - 1)/cpb;
	const int lim = conserved ? 0 : (d.nc.reconstruction == FVX_RECON_VENKATAKRISHNAN ? 1 :
	                                 d.nc.reconstruction == FVX_RECON_BARTHJESPERSEN ? 2 : 0);
#define FVX_GL(L, C) hipLaunchKernelGGL((k_gradients<L, C>), dim3(nblk, nact), dim3(block), 0, \
		s, m.nelem, u, m.d_nbr.p, m.d_soff.p, d.d_gcoef.p, \
		d.d_eps2.p, ba.gp ? ba.gp : d.d_gas.p, ba, grad, rgrad, \
		d.tune.grad_xcd, m.nbface, ughost)
	d.prof.start(PROF_GRADIENTS, s);
	if(m.nbface > 0)
		hipLaunchKernelGGL(k_ghost, dim3((m.nbface + 255)/256, nact), dim3(256), 0, s, m.nelem,
		                   m.nbface, u, m.d_bhost.p, m.d_fgeom.p, d.d_bcid.p,
		                   ba.gp ? ba.gp : d.d_gas.p, d.d_bct.p, ba, ughost);
	if(conserved) FVX_GL(0, true);
	else if(lim == 0) FVX_GL(0, false);
	else if(lim == 1) FVX_GL(1, false);
	else FVX_GL(2, false);
#undef FVX_GL
	FVX_HIP(hipGetLastError());
	d.prof.stop(s);
}

bool flux_wave_apply(const fvx_disc_s& d)
{
	// inviscid, first-order or linear reconstruction (unlimited, limited or WENO-weighted)
	const bool order2 = d.nc.order2 && d.nc.gradientscheme != FVX_GRAD_NONE;
	return !d.pc.viscous && !(order2 && d.nc.reconstruction == FVX_RECON_VANALBADA);
}

void launch_residual(const fvx_disc_s& d, const double* u, double* r, int gettimesteps,
                     double* dt, hipStream_t s, const BatchArg& ba, int nact, double* grad,
                     double* rgrad, double* ughost)
{
	const fvx_mesh_s& m = *d.m;
	if(!grad) grad = d.d_grad.p;
	if(!rgrad) rgrad = d.d_rgrad.p;
	const bool order2 = d.nc.order2 && d.nc.gradientscheme != FVX_GRAD_NONE;
	const bool visc = d.pc.viscous != 0;
	const int rec = d.nc.reconstruction;
	int mode = 0;
	const double* rg = nullptr;
	const double* ug = nullptr;
	if(order2) {
		const bool limited = rec == FVX_RECON_VENKATAKRISHNAN || rec == FVX_RECON_BARTHJESPERSEN;
		if(limited) {
			launch_gradients(d, u, visc ? grad : nullptr, rgrad, 0, s, ba, nact, ughost);
			rg = rgrad; ug = grad; mode = 1;
		} else if(rec == FVX_RECON_WENO) {
			launch_gradients(d, u, grad, nullptr, 0, s, ba, nact, ughost);
			launch_weno(d, grad, rgrad, s, ba, nact);
			rg = rgrad; ug = grad; mode = 1;
		} else {
			launch_gradients(d, u, grad, nullptr, 0, s, ba, nact, ughost);
			rg = grad; ug = grad;
			mode = rec == FVX_RECON_VANALBADA ? 2 : 1;
		}
	}
	d.prof.start(PROF_FLUX, s);
	if(d.tune.flux_wave && flux_wave_apply(d)) {
		// wave-task form (flux_wave.hpp): each face evaluated once per wave64 cell group
		switch(d.nc.conv_numflux) {
#define FVX_W(ID) case ID: launch_flux_wave_##ID(mode, m.nelem, m.nwgroup, u, rg, m.d_wtask.p, \
	m.d_wgroup.p, m.d_wslot.p, m.d_fgeom.p, m.d_centre.p, m.d_area.p, d.d_bcid.p, \
	ba.gp ? ba.gp : d.d_gas.p, d.d_bct.p, r, dt, gettimesteps, ba, nact, d.tune.flux_xcd, s); \
	break;
		FVX_W(0) FVX_W(1) FVX_W(2) FVX_W(3) FVX_W(4) FVX_W(5) FVX_W(6)
#undef FVX_W
		default: throw Error(FVX_ERR_INVALID, "unknown flux");
		}
		FVX_HIP(hipGetLastError());
		d.prof.stop(s);
		return;
	}
	switch(d.nc.conv_numflux) {
#define FVX_C(ID) case ID: launch_flux_##ID(mode, visc, m.nelem, u, rg, ug, m.d_nbr.p, \
	m.d_face.p, m.d_fgeom.p, m.d_centre.p, m.d_area.p, d.d_bcid.p, \
	ba.gp ? ba.gp : d.d_gas.p, d.d_bct.p, r, dt, gettimesteps, ba, nact, d.tune.flux_block, \
	d.tune.flux_xcd, s); break;
	FVX_C(0) FVX_C(1) FVX_C(2) FVX_C(3) FVX_C(4) FVX_C(5) FVX_C(6)
#undef FVX_C
	default: throw Error(FVX_ERR_INVALID, "unknown flux");
	}
	FVX_HIP(hipGetLastError());
	d.prof.stop(s);
}

/* ---------------------------------------------------------------- functionals ---------- */

constexpr int FN_BLOCKS = 128;

/// Per-block partial sums: [0] CL, [1] CDp, [2] CDsf, [3] wall length, [4] entropy error^2.
/// Restates computeSurfaceData (src/spatial/flow_spatial.cpp:130-310, cell-centred states,
/// normalised by wall length) and compute_entropy_cell (src/spatial/aoutput.cpp:28-62).
__global__ __launch_bounds__(BLOCK)
void k_functionals(int ncell, int nbface, const double* __restrict__ u,
                   const double* __restrict__ cgrad, const int* __restrict__ bhost,
                   const int* __restrict__ btag, const FaceGeom* __restrict__ fgeom,
                   const double* __restrict__ area, int walltag, GasParams gp, double cosa,
                   double sina, double* __restrict__ part)
{
	__shared__ double sh[BLOCK][5];
	double acc[5] = {0, 0, 0, 0, 0};
	const double sref = pressure(gp, gp.uinf)/pow(gp.uinf[0], gp.g);
	for(int f = blockIdx.x*BLOCK + threadIdx.x; f < nbface; f += gridDim.x*BLOCK) {
		if(btag[f] != walltag) continue;
		const int c = bhost[f];
		const FaceGeom G = fgeom[f];
		double w[4];
		load4(u + 4*c, w);
		const double cp = (pressure(gp, w) - gp.pinf)*2.0;
		// Sutherland's law also with constant viscosity (flow_spatial.cpp:226)
		const double mu = sutherland(gp, pressure(gp, w)/w[0]*gp.gM2);
		double gu[2][2];
		for(int i = 0; i < 2; i++)
			for(int j = 0; j < 2; j++)
				gu[i][j] = (cgrad[8*c + (i+1)*2 + j]*w[0] - w[i+1]*cgrad[8*c + j])/(w[0]*w[0]);
		const double n[2] = { G.nx, G.ny }, t[2] = { G.ny, -G.nx };
		double fo[2];
		for(int i = 0; i < 2; i++) fo[i] = (gu[i][0] + gu[0][i])*n[0] + (gu[i][1] + gu[1][i])*n[1];
		const double csf = 2*mu*(fo[0]*t[0] + fo[1]*t[1]);
		acc[0] += cp*(n[0]*(-sina) + n[1]*cosa)*G.len;
		acc[1] += cp*(n[0]*cosa + n[1]*sina)*G.len;
		acc[2] += csf*(t[0]*cosa + t[1]*sina)*G.len;
		acc[3] += G.len;
	}
	for(int c = blockIdx.x*BLOCK + threadIdx.x; c < ncell; c += gridDim.x*BLOCK) {
		double w[4];
		load4(u + 4*c, w);
		const double s = pressure(gp, w)/pow(w[0], gp.g);
		const double e = (s - sref)/sref;
		acc[4] += e*e*area[c];
	}
	for(int k = 0; k < 5; k++) sh[threadIdx.x][k] = acc[k];
	__syncthreads();
	for(int st = BLOCK/2; st > 0; st >>= 1) {
		if(threadIdx.x < st)
			for(int k = 0; k < 5; k++) sh[threadIdx.x][k] += sh[threadIdx.x + st][k];
		__syncthreads();
	}
	if(threadIdx.x < 5) part[blockIdx.x*5 + threadIdx.x] = sh[0][threadIdx.x];
}

void launch_functionals(const fvx_disc_s& d, const double* u, const double* cgrad, int walltag,
                        double* dev_out, hipStream_t s, const GasParams* gp, double aoa)
{
	const fvx_mesh_s& m = *d.m;
	const double a = gp ? aoa : d.pc.aoa;
	hipLaunchKernelGGL(k_functionals, dim3(FN_BLOCKS), dim3(BLOCK), 0, s, m.nelem, m.nbface, u,
	                   cgrad, m.d_bhost.p, m.d_btag.p, m.d_fgeom.p, m.d_area.p, walltag,
	                   gp ? *gp : d.gas, cos(a), sin(a), dev_out);
	FVX_HIP(hipGetLastError());
}

int functional_blocks() { return FN_BLOCKS; }

/// Per wall face of `list` (boundary-face indices in boundary order): face centre, Cp and Cf from
/// the cell-centred state and conserved gradients of the host cell -- the output matrix of
/// computeSurfaceData (src/spatial/flow_spatial.cpp:130-310), rows [nw][4].
__global__ __launch_bounds__(BLOCK)
void k_surface_data(int nw, const int* __restrict__ list, const double* __restrict__ u,
                    const double* __restrict__ cgrad, const int* __restrict__ bhost,
                    const FaceGeom* __restrict__ fgeom, GasParams gp, double* __restrict__ out)
{
	const int i = blockIdx.x*BLOCK + threadIdx.x;
	if(i >= nw) return;
	const int f = list[i];
	const int c = bhost[f];
	const FaceGeom G = fgeom[f];
	double w[4];
	load4(u + 4*c, w);
	const double cp = (pressure(gp, w) - gp.pinf)*2.0;
	const double mu = sutherland(gp, pressure(gp, w)/w[0]*gp.gM2);
	double gu[2][2];
	for(int a = 0; a < 2; a++)
		for(int b = 0; b < 2; b++)
			gu[a][b] = (cgrad[8*c + (a+1)*2 + b]*w[0] - w[a+1]*cgrad[8*c + b])/(w[0]*w[0]);
	const double n[2] = { G.nx, G.ny }, t[2] = { G.ny, -G.nx };
	double fo[2];
	for(int a = 0; a < 2; a++) fo[a] = (gu[a][0] + gu[0][a])*n[0] + (gu[a][1] + gu[1][a])*n[1];
	out[4*i] = G.mx;
	out[4*i + 1] = G.my;
	out[4*i + 2] = cp;
	out[4*i + 3] = 2*mu*(fo[0]*t[0] + fo[1]*t[1]);
}

void launch_surface_data(const fvx_disc_s& d, const double* u, const double* cgrad,
                         const int* list, int nw, double* out, hipStream_t s)
{
	if(nw == 0) return;
	hipLaunchKernelGGL(k_surface_data, dim3((nw + BLOCK - 1)/BLOCK), dim3(BLOCK), 0, s, nw, list,
	                   u, cgrad, d.m->d_bhost.p, d.m->d_fgeom.p, d.gas, out);
	FVX_HIP(hipGetLastError());
}

}
