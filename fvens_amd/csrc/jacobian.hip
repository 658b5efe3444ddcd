/* jacobian.hip -- first-order approximate Jacobian assembly (Spatial::assemble_jacobian).
 *
 * Reference semantics: Spatial::assemble_jacobian (src/spatial/aspatial.cpp:242-340) with
 * FlowFV::compute_local_jacobian_interior / _boundary (src/spatial/flow_spatial.cpp:818-875):
 * the Jacobian of the FIRST-ORDER residual evaluated at the cell-centred states. For an interior
 * face with left cell L (lower index) and right cell R, F = F(u_L, u_R) times the face length
 * leaves L and enters R, so the row of L gains +len dF/du_L on its diagonal and +len dF/du_R in
 * column R, and the row of R gains -len dF/du_R on its diagonal and -len dF/du_L in column L.
 * A boundary face adds len (dF/du_L + dF/du_g du_g/du_L) to the diagonal, with the ghost state
 * and its Jacobian from FlowBC::computeGhostStateAndJacobian (src/spatial/abc.cpp).
 *
 * Flux Jacobians (InviscidFlux::get_jacobian): Roe (anumericalflux.cpp:736-965), HLLC
 * (:1230-1397) and AUSM (:317-472) are differentiated exactly here with forward-mode duals
 * through the same templated flux code the residual uses; LLF (:65-107) and HLL (:1012-1061)
 * use the reference's frozen-eigenvalue / frozen-speed forms. The viscous part is the
 * thin-layer flux of flow_spatial.cpp:397-446 / aspatial.cpp:207-240, differentiated exactly.
 *
 * Storage ("slot blocks"): vals[cell][slot][4][4] row-major, slot 0 the diagonal block and slot
 * 1+k the block of the neighbour across local face k (nbr[cell][k]); slots without an interior
 * neighbour hold zeros. Each cell's quad of lanes evaluates its own faces (each interior face
 * twice, once from each side, from identical canonical data), so assembly needs no atomics and
 * is bitwise reproducible.
 */
#include "fvx_internal.hpp"
#include "gas.hpp"
#include "dual.hpp"
#include "wave.hpp"

namespace fvx {

using D4 = Dual<4>;

/// Thin-layer viscous flux of two cell states: face gradient (q_R - q_L)/|dr| along the unit
/// centre-to-centre vector, q = (rho, vx, vy, T) (viscousphysics.cpp:124-246).
template <class T>
__device__ __forceinline__ void thinlayer_flux(const GasParams& gp, const double* n,
                                               const double* xl, const double* xr,
                                               const T* ul, const T* ur, T* vf)
{
	T wl[4], wr[4];
	wl[0] = ul[0]; wl[1] = ul[1]/ul[0]; wl[2] = ul[2]/ul[0];
	wl[3] = pressure(gp, ul)/ul[0]*gp.gM2;
	wr[0] = ur[0]; wr[1] = ur[1]/ur[0]; wr[2] = ur[2]/ur[0];
	wr[3] = pressure(gp, ur)/ur[0]*gp.gM2;
	double dr[2] = { xr[0] - xl[0], xr[1] - xl[1] };
	const double dist = sqrt(dr[0]*dr[0] + dr[1]*dr[1]);
	dr[0] /= dist; dr[1] /= dist;
	T grad[2][4];
	for(int i = 0; i < 4; i++) {
		const T corr = (wr[i] - wl[i])/dist;
		grad[0][i] = corr*dr[0];
		grad[1][i] = corr*dr[1];
	}
	T mu;
	if(gp.constvisc) mu = T(1.0/gp.Reinf);
	else mu = 0.5*(sutherland(gp, wl[3]) + sutherland(gp, wr[3]));
	const T kd = mu/(gp.Minf*gp.Minf*gp.gm1*gp.Pr);
	const T ldiv = 2.0/3.0*mu*(grad[0][1] + grad[1][2]);
	T s[2][2];
	for(int i = 0; i < 2; i++) {
		for(int j = 0; j < 2; j++) s[i][j] = mu*(grad[i][j+1] + grad[j][i+1]);
		s[i][i] -= ldiv;
	}
	vf[0] = T(0.0);
	vf[1] = -(s[0][0]*n[0] + s[0][1]*n[1]);
	vf[2] = -(s[1][0]*n[0] + s[1][1]*n[1]);
	const T va0 = 0.5*(wl[1] + wr[1]), va1 = 0.5*(wl[2] + wr[2]);
	T e = T(0.0);
	for(int i = 0; i < 2; i++) {
		const T comp = s[i][0]*va0 + s[i][1]*va1 + kd*grad[i][3];
		e -= comp*n[i];
	}
	vf[3] = e;
}

/// Ghost state and du_g/du_in (row-major). Every kind except the isothermal wall is the exact
/// derivative of ghost_state (abc.cpp:86-140, 201-210, 228-262, 282-305, 425-437), obtained with
/// duals; the isothermal wall uses the reference's own linearisation, including its ghost
/// energy at the wall temperature (abc.cpp:371-409).
__device__ __forceinline__ void ghost_jacobian(const GasParams& gp, const BCTable& bt, int bc,
                                               const double* ui, const double* n, double* ug,
                                               double* dug)
{
	if(bt.kind[bc] == 6) {
		for(int k = 0; k < 16; k++) dug[k] = 0.0;
		const double v0 = bt.v0[bc], v1 = bt.v1[bc];
		const double tm = v0*ui[0];
		ug[0] = ui[0]; dug[0] = 1.0;
		ug[1] = 2.0*tm*n[1] - ui[1]; dug[4] = 2.0*v0*n[1]; dug[5] = -1.0;
		ug[2] = -2.0*tm*n[0] - ui[2]; dug[8] = -2.0*v0*n[0]; dug[10] = -1.0;
		const double m2 = ug[1]*ug[1] + ug[2]*ug[2];
		const double ir2 = 1.0/(ui[0]*ui[0]);
		const double vmag2 = m2*ir2;
		const double coeff = 1.0/(gp.g*gp.gm1*gp.Minf*gp.Minf);
		ug[3] = ui[0]*(v1*coeff + 0.5*vmag2);
		const double dv0 = -2.0*m2*ir2/ui[0];
		dug[12] = coeff*v1 + 0.5*(vmag2 + ui[0]*dv0);
		dug[13] = ui[0]*0.5*(-ug[1]*ir2);
		dug[14] = ui[0]*0.5*(-ug[2]*ir2);
		dug[15] = 0.0;
		return;
	}
	D4 a[4], g[4];
	for(int k = 0; k < 4; k++) { a[k] = D4(ui[k]); a[k].d[k] = 1.0; }
	ghost_state(gp, bt, bc, a, n, g);
	for(int i = 0; i < 4; i++) {
		ug[i] = g[i].v;
		for(int j = 0; j < 4; j++) dug[i*4+j] = g[i].d[j];
	}
}

/// dF/du of the analytic directional flux (aphysics_defs.hpp:13-23), by duals.
__device__ __forceinline__ void phys_flux_jacobian(const GasParams& gp, const double* u,
                                                   const double* n, double* A)
{
	D4 a[4], f[4];
	for(int k = 0; k < 4; k++) { a[k] = D4(u[k]); a[k].d[k] = 1.0; }
	const D4 vn = (a[1]*n[0] + a[2]*n[1])/a[0];
	phys_flux(a, n, vn, pressure(gp, a), f);
	for(int i = 0; i < 4; i++)
		for(int j = 0; j < 4; j++) A[i*4+j] = f[i].d[j];
}

/// Face flux derivatives JL = dF/du_L, JR = dF/du_R of the Jacobian flux (plus the thin-layer
/// viscous flux when VISC) at cell-centred states.
template <int FLUX, bool VISC>
__device__ __forceinline__ void face_jacobians(const GasParams& gp, const double* n,
                                               const double* xl, const double* xr,
                                               const double* ul, const double* ur,
                                               double* JL, double* JR)
{
	if constexpr(FLUX == 2 || FLUX == 3) {
		phys_flux_jacobian(gp, ul, n, JL);
		phys_flux_jacobian(gp, ur, n, JR);
		const SideVars<double> a = side_vars(gp, ul, n), b = side_vars(gp, ur, n);
		const double ca = sqrt(gp.g*a.p/ul[0]), cb = sqrt(gp.g*b.p/ur[0]);
		double wl, wr, dl, dr;     // JL <- wl JL + dl I, JR <- wr JR + dr I
		if constexpr(FLUX == 3) {
			const double ea = fabs(a.vn) + ca, eb = fabs(b.vn) + cb;
			const double eig = ea >= eb ? ea : eb;
			wl = 0.5; dl = 0.5*eig; wr = 0.5; dr = -0.5*eig;
		} else {
			const RoeAvg<double> m = roe_avg(gp, ul, ur, n, a, b);
			double sl, sr;
			einfeldt(a.vn, ca, b.vn, cb, m, sl, sr);
			const double sr0 = sr > 0.0 ? 0.0 : sr, sl0 = sl > 0.0 ? 0.0 : sl;
			const double t1 = (sr0 - sl0)/(sr - sl), t2 = 1.0 - t1;
			const double t3 = 0.5*(sr*fabs(sl) - sl*fabs(sr))/(sr - sl);
			wl = t2; dl = t3; wr = t1; dr = -t3;
		}
		for(int k = 0; k < 16; k++) { JL[k] *= wl; JR[k] *= wr; }
		for(int i = 0; i < 4; i++) { JL[i*5] += dl; JR[i*5] += dr; }
		if constexpr(VISC) {
			for(int side = 0; side < 2; side++) {
				D4 a4[4], b4[4], vf[4];
				for(int k = 0; k < 4; k++) {
					a4[k] = D4(ul[k]); b4[k] = D4(ur[k]);
					if(side == 0) a4[k].d[k] = 1.0; else b4[k].d[k] = 1.0;
				}
				thinlayer_flux(gp, n, xl, xr, a4, b4, vf);
				double* J = side == 0 ? JL : JR;
				for(int i = 0; i < 4; i++)
					for(int j = 0; j < 4; j++) J[i*4+j] += vf[i].d[j];
			}
		}
	} else {
		for(int side = 0; side < 2; side++) {
			D4 a4[4], b4[4], f[4];
			for(int k = 0; k < 4; k++) {
				a4[k] = D4(ul[k]); b4[k] = D4(ur[k]);
				if(side == 0) a4[k].d[k] = 1.0; else b4[k].d[k] = 1.0;
			}
			num_flux<FLUX>(gp, a4, b4, n, f);
			if constexpr(VISC) {
				D4 vf[4];
				thinlayer_flux(gp, n, xl, xr, a4, b4, vf);
				for(int k = 0; k < 4; k++) f[k] += vf[k];
			}
			double* J = side == 0 ? JL : JR;
			for(int i = 0; i < 4; i++)
				for(int j = 0; j < 4; j++) J[i*4+j] = f[i].d[j];
		}
	}
}

/// Total derivative d/du_L of F(u_L, u_g(u_L)) at a boundary face.
template <int FLUX, bool VISC>
__device__ __forceinline__ void boundary_jacobian(const GasParams& gp, const BCTable& bt, int bc,
                                                  const double* n, const double* xl,
                                                  const double* xg, const double* ul, double* J)
{
	double ug[4], dug[16];
	ghost_jacobian(gp, bt, bc, ul, n, ug, dug);
	if constexpr(FLUX == 2 || FLUX == 3) {
		double JL[16], JR[16];
		face_jacobians<FLUX, VISC>(gp, n, xl, xg, ul, ug, JL, JR);
		for(int i = 0; i < 4; i++)
			for(int j = 0; j < 4; j++) {
				double s = 0.0;
				for(int k = 0; k < 4; k++) s += JR[i*4+k]*dug[k*4+j];
				J[i*4+j] = JL[i*4+j] + s;
			}
	} else {
		D4 a4[4], g4[4], f[4];
		for(int i = 0; i < 4; i++) {
			a4[i] = D4(ul[i]); a4[i].d[i] = 1.0;
			g4[i] = D4(ug[i]);
			for(int j = 0; j < 4; j++) g4[i].d[j] = dug[i*4+j];
		}
		num_flux<FLUX>(gp, a4, g4, n, f);
		if constexpr(VISC) {
			D4 vf[4];
			thinlayer_flux(gp, n, xl, xg, a4, g4, vf);
			for(int k = 0; k < 4; k++) f[k] += vf[k];
		}
		for(int i = 0; i < 4; i++)
			for(int j = 0; j < 4; j++) J[i*4+j] = f[i].d[j];
	}
}

template <class VT> __device__ __forceinline__ void store_row(VT* p, const double* v);
template <> __device__ __forceinline__ void store_row<double>(double* p, const double* v)
{
	reinterpret_cast<double2*>(p)[0] = make_double2(v[0], v[1]);
	reinterpret_cast<double2*>(p)[1] = make_double2(v[2], v[3]);
}
template <> __device__ __forceinline__ void store_row<float>(float* p, const double* v)
{
	*reinterpret_cast<float4*>(p) = make_float4((float)v[0], (float)v[1], (float)v[2],
	                                            (float)v[3]);
}

/// One quad of lanes per cell; lane q owns local face q. When dt is given the pseudo-time term
/// area/(cfl dt) is added to the diagonal and stored in dtm (SteadyBackwardEulerSolver::
/// addPseudoTimeTerm, src/ode/aodesolver.cpp:300-329).
template <int FLUX, bool VISC, class VT>
__global__ __launch_bounds__(BLOCK)
void k_jacobian(int ncell, const double* __restrict__ u, const int* __restrict__ nbr,
                const int* __restrict__ face, const FaceGeom* __restrict__ fgeom,
                const double* __restrict__ centre, const int* __restrict__ bcid,
                const GasParams* __restrict__ gpa, const BCTable* __restrict__ bta,
                const BatchArg ba, VT* __restrict__ vals, const double* __restrict__ dt,
                const double* __restrict__ area, double cfl0, double* __restrict__ dtm,
                const int xcd)
{
	// instance: state, matrix, time steps and CFL number are per instance; gpa holds the gas
	// parameters of each instance (one entry for a single solve)
	const long I = ba.list ? ba.list[blockIdx.y] : 0;
	if(I < 0) return;                     // a retired slot of a compacted list
	const GasParams& gp = gpa[I];
	const BCTable& bt = *bta;
	const double cfl = ba.cfl ? ba.cfl[I] : cfl0;
	u += I*4L*ncell;
	vals += I*80L*ncell;
	if(dt) { dt += I*(long)ncell; dtm += I*(long)ncell; }
	const int q = threadIdx.x & 3;
	const int cell = xcd_block(xcd)*(int)(blockDim.x >> 2) + (threadIdx.x >> 2);
	const bool valid = cell < ncell;
	const int cc = valid ? cell : 0;
	const int nb = nbr[cc*4 + q];
	const bool act = valid && nb != SLOT_UNUSED;

	double Jd[16], Jo[16];
	for(int k = 0; k < 16; k++) { Jd[k] = 0.0; Jo[k] = 0.0; }
	if(act) {
		const FaceGeom G = fgeom[face[cc*4 + q]];
		const double n[2] = { G.nx, G.ny };
		double uo[4], xo[2];
		load4(u + 4*cc, uo);
		xo[0] = centre[2*cc]; xo[1] = centre[2*cc+1];
		if(nb < 0) {
			const int bc = bcid[-1 - nb];
			const double xg[2] = { 2.0*G.mx - xo[0], 2.0*G.my - xo[1] };
			boundary_jacobian<FLUX, VISC>(gp, bt, bc, n, xo, xg, uo, Jd);
			for(int k = 0; k < 16; k++) Jd[k] *= G.len;
		} else {
			double un[4], xn[2];
			load4(u + 4*nb, un);
			xn[0] = centre[2*nb]; xn[1] = centre[2*nb+1];
			const bool selfL = cc < nb;
			double JL[16], JR[16];
			if(selfL) face_jacobians<FLUX, VISC>(gp, n, xo, xn, uo, un, JL, JR);
			else      face_jacobians<FLUX, VISC>(gp, n, xn, xo, un, uo, JL, JR);
			const double s = selfL ? G.len : -G.len;
			for(int k = 0; k < 16; k++) {
				Jd[k] = s*(selfL ? JL[k] : JR[k]);
				Jo[k] = s*(selfL ? JR[k] : JL[k]);
			}
		}
	}
	for(int k = 0; k < 16; k++) Jd[k] = quad_sum(Jd[k]);
	if(!valid) return;
	double add = 0.0;
	if(dt) {
		add = area[cc]/(cfl*dt[cc]);
		if(q == 0) dtm[cc] = add;
	}
	double row[4];
	for(int j = 0; j < 4; j++) {
		// row q of the diagonal block, selected without dynamic register indexing
		double v = Jd[j];
		for(int r = 1; r < 4; r++) v = q == r ? Jd[r*4+j] : v;
		row[j] = v + (j == q ? add : 0.0);
	}
	VT* blk = vals + (size_t)cc*80;
	store_row<VT>(blk + 4*q, row);
	VT* off = blk + 16*(1 + q);
	for(int r = 0; r < 4; r++) store_row<VT>(off + 4*r, Jo + 4*r);
}

template <int FLUX, bool VISC>
static void launch_jac_t(const fvx_disc_s& d, const double* u, void* vals, int prec,
                         const double* dt, double cfl, double* dtm, hipStream_t s,
                         const BatchArg& ba, int nact)
{
	const fvx_mesh_s* m = d.m;
	const int n = m->nelem;
	const int block = d.tune.jac_block, cpb = block/4;
	const dim3 g((n + cpb - 1)/cpb, nact);
	if(prec == FVX_PREC_F32)
		hipLaunchKernelGGL((k_jacobian<FLUX, VISC, float>), g, dim3(block), 0, s, n, u,
		                   m->d_nbr.p, m->d_face.p, m->d_fgeom.p, m->d_centre.p, d.d_bcid.p,
		                   ba.gp ? ba.gp : d.d_gas.p, d.d_bct.p, ba, static_cast<float*>(vals), dt,
		                   m->d_area.p, cfl, dtm, d.tune.jac_xcd);
	else
		hipLaunchKernelGGL((k_jacobian<FLUX, VISC, double>), g, dim3(block), 0, s, n, u,
		                   m->d_nbr.p, m->d_face.p, m->d_fgeom.p, m->d_centre.p, d.d_bcid.p,
		                   ba.gp ? ba.gp : d.d_gas.p, d.d_bct.p, ba, static_cast<double*>(vals),
		                   dt, m->d_area.p, cfl, dtm, d.tune.jac_xcd);
	FVX_HIP(hipGetLastError());
}

template <int FLUX>
static void launch_jac_v(const fvx_disc_s& d, const double* u, void* vals, int prec,
                         const double* dt, double cfl, double* dtm, hipStream_t s,
                         const BatchArg& ba, int nact)
{
	if(d.pc.viscous) launch_jac_t<FLUX, true>(d, u, vals, prec, dt, cfl, dtm, s, ba, nact);
	else launch_jac_t<FLUX, false>(d, u, vals, prec, dt, cfl, dtm, s, ba, nact);
}

bool jacobian_supported(const fvx_disc_s& d)
{
	const int f = d.nc.conv_numflux_jac;
	if(f != FVX_FLUX_ROE && f != FVX_FLUX_HLLC && f != FVX_FLUX_HLL && f != FVX_FLUX_LLF &&
	   f != FVX_FLUX_AUSM)
		return false;
	for(int i = 0; i < d.pc.nbc; i++) {
		const int k = d.pc.bcs[i].kind;
		if(k == FVX_BC_SUBSONIC_INFLOW || k == FVX_BC_PERIODIC) return false;
	}
	return true;
}

void launch_jacobian(const fvx_disc_s& d, const double* u, void* vals, int prec,
                     const double* dt, double cfl, double* dtm, hipStream_t s,
                     const BatchArg& ba, int nact)
{
	d.prof.start(PROF_JACOBIAN, s);
	switch(d.nc.conv_numflux_jac) {
	case FVX_FLUX_ROE:  launch_jac_v<0>(d, u, vals, prec, dt, cfl, dtm, s, ba, nact); break;
	case FVX_FLUX_HLLC: launch_jac_v<1>(d, u, vals, prec, dt, cfl, dtm, s, ba, nact); break;
	case FVX_FLUX_HLL:  launch_jac_v<2>(d, u, vals, prec, dt, cfl, dtm, s, ba, nact); break;
	case FVX_FLUX_LLF:  launch_jac_v<3>(d, u, vals, prec, dt, cfl, dtm, s, ba, nact); break;
	case FVX_FLUX_AUSM: launch_jac_v<4>(d, u, vals, prec, dt, cfl, dtm, s, ba, nact); break;
	default: throw Error(FVX_ERR_UNSUPPORTED, "flux Jacobian not available for this flux");
	}
	d.prof.stop(s);
}

}
