/* flux_kernel.hpp -- the face-flux / residual-gather kernel of FlowFV::compute_residual.
 *
 * Reference semantics: FlowFV::compute_residual and compute_fluxes / compute_max_timestep
 * (src/spatial/flow_spatial.cpp:488-816). The reference computes each face flux once and
 * scatters it to both cells with OpenMP atomics. Here each cell gathers its own faces: the
 * quad of lanes owning a cell evaluates the numerical flux of each of its faces from the
 * face's canonical (left = lower index) states, so the two cells sharing a face evaluate the
 * identical instruction sequence on identical data and receive exactly opposite
 * contributions. No atomics; the per-cell sum order is fixed (local face order).
 *
 * Reconstruction modes:
 *   0  first order: face states are the cell states (flow_spatial.cpp:756-774)
 *   1  linear: w_f = w_c + g . (x_f - x_c) in primitive variables with the (possibly limited)
 *      gradient g (src/spatial/reconstruction_utils.hpp:17-32, areconstruction.cpp:51-103,
 *      limitedlinearreconstruction.cpp:116-268)
 *   2  MUSCL kappa = 1/3 with the Van Albada limiter (musclreconstruction.cpp:33-130)
 * Boundary faces: the ghost state is re-derived from the reconstructed left state
 * (flow_spatial.cpp:777-778); reconstruction and viscous terms use the ghost state of the
 * cell-centred value and the ghost centre reflected about the face midpoint.
 */
#ifndef FVX_FLUX_KERNEL_HPP
#define FVX_FLUX_KERNEL_HPP

#include "fvx_internal.hpp"
#include "face_state.hpp"
#include "wave.hpp"

namespace fvx {

/// Viscous flux across a face (compute_viscous_flux, flow_spatial.cpp:348-395;
/// getPrimitive2StatesAndGradients, viscousphysics.cpp:14-68; modified-average face gradient,
/// aspatial.cpp:172-205; computeViscousFlux, viscousphysics.cpp:70-122).
/// wl, wr: primitive cell-centred states with reciprocal densities irl, irr; gl, gr: primitive
/// gradients [var][dim] (null for first order); A, B: face states.
__device__ __forceinline__ void viscous_flux(const GasParams& gp, const double* n,
                                             const double* xl, const double* xr,
                                             const double* wl, double irl,
                                             const double* wr, double irr,
                                             const double* gl, const double* gr,
                                             const FaceState& A, const FaceState& B, double* vf)
{
	// temperature and its gradient from (rho, p) and their gradients
	const double Tl = wl[3]*irl*gp.gM2, Tr = wr[3]*irr*gp.gM2;
	double gtl[2] = {0, 0}, gtr[2] = {0, 0};
	if(gl) {
		for(int j = 0; j < 2; j++) {
			gtl[j] = (gl[3*2+j]*wl[0] - wl[3]*gl[0*2+j])*(irl*irl)*gp.gM2;
			gtr[j] = (gr[3*2+j]*wr[0] - wr[3]*gr[0*2+j])*(irr*irr)*gp.gM2;
		}
	}
	double dr[2] = { xr[0] - xl[0], xr[1] - xl[1] };
	const double dist = sqrt(dr[0]*dr[0] + dr[1]*dr[1]);
	const double idist = 1.0/dist;
	dr[0] *= idist; dr[1] *= idist;
	const double cl[4] = { wl[0], wl[1], wl[2], Tl }, cr[4] = { wr[0], wr[1], wr[2], Tr };
	double grad[2][4];
	for(int i = 0; i < 4; i++) {
		double da[2];
		for(int j = 0; j < 2; j++) {
			const double a = gl ? (i == 3 ? gtl[j] : gl[i*2+j]) : 0.0;
			const double b = gl ? (i == 3 ? gtr[j] : gr[i*2+j]) : 0.0;
			da[j] = 0.5*(a + b);
		}
		const double corr = (cr[i] - cl[i])*idist;
		const double ddr = da[0]*dr[0] + da[1]*dr[1];
		for(int j = 0; j < 2; j++) grad[j][i] = da[j] - ddr*dr[j] + corr*dr[j];
	}
	double mu;
	if(gp.constvisc) mu = 1.0/gp.Reinf;
	else mu = 0.5*(sutherland(gp, A.p*A.ir*gp.gM2) + sutherland(gp, B.p*B.ir*gp.gM2));
	const double kd = mu/(gp.Minf*gp.Minf*gp.gm1*gp.Pr);
	const double ldiv = (grad[0][1] + grad[1][2])*(2.0/3.0*mu);
	double s[2][2];
	for(int i = 0; i < 2; i++) {
		for(int j = 0; j < 2; j++) s[i][j] = mu*(grad[i][j+1] + grad[j][i+1]);
		s[i][i] -= ldiv;
	}
	vf[0] = 0.0;
	vf[1] = -(s[0][0]*n[0] + s[0][1]*n[1]);
	vf[2] = -(s[1][0]*n[0] + s[1][1]*n[1]);
	const double va[2] = { 0.5*(A.vx + B.vx), 0.5*(A.vy + B.vy) };
	double e = 0.0;
	for(int i = 0; i < 2; i++) {
		const double comp = s[i][0]*va[0] + s[i][1]*va[1] + kd*grad[i][3];
		e -= comp*n[i];
	}
	vf[3] = e;
}

__device__ __forceinline__ void load_grad(const double* __restrict__ g, int c, double* v)
{
	load4(g + 8*c, v);
	load4(g + 8*c + 4, v + 4);
}

template <int FLUX, int MODE, bool VISC>
__global__ __launch_bounds__(MAX_BLOCK)
void k_flux(int ncell, const double* __restrict__ u, const double* __restrict__ rgrad,
            const double* __restrict__ grad, const int* __restrict__ nbr,
            const int* __restrict__ face, const FaceGeom* __restrict__ fgeom,
            const double* __restrict__ centre, const double* __restrict__ area,
            const int* __restrict__ bcid, const GasParams* __restrict__ gpa,
            const BCTable* __restrict__ bta, const BatchArg ba, double* __restrict__ res,
            double* __restrict__ dt, const int getdt, const int xcd)
{
	// instance: state, gradients, residual and time steps are instance-major; gpa holds the
	// gas parameters of each instance (one entry for a single solve)
	const long I = ba.list ? ba.list[blockIdx.y] : 0;
	if(I < 0) return;                     // a retired slot of a compacted list
	const GasParams& gp = gpa[I];
	const BCTable& bt = *bta;
	u += I*4L*ncell;
	if(rgrad) rgrad += I*8L*ncell;
	if(grad) grad += I*8L*ncell;
	res += I*4L*ncell;
	if(dt) dt += I*(long)ncell;
	const int q = threadIdx.x & 3;
	const int cell = xcd_block(xcd)*(int)(blockDim.x >> 2) + (threadIdx.x >> 2);
	const bool valid = cell < ncell;
	const int cc = valid ? cell : 0;
	const int nb = nbr[cc*4 + q];
	const bool act = valid && nb != SLOT_UNUSED;
	const double igm1 = 1.0/gp.gm1;

	double F[4] = {0.0, 0.0, 0.0, 0.0};
	double srad = 0.0;
	if(act) {
		const FaceGeom G = fgeom[face[cc*4 + q]];
		const double n[2] = { G.nx, G.ny };
		const bool bnd = nb < 0;
		const bool selfL = bnd || cc < nb;
		const int cL = selfL ? cc : nb;
		const int cR = selfL ? nb : cc;        // negative for a boundary face

		double uL[4], wL[4], xL[2];
		load4(u + 4*cL, uL);
		xL[0] = centre[2*cL]; xL[1] = centre[2*cL+1];
		const double irL = prim_lean(gp, uL, wL);

		// right cell-centred state: neighbour cell, or ghost of the left cell value
		double uR[4], wR[4], xR[2];
		int bc = 0;
		if(!bnd) {
			load4(u + 4*cR, uR);
			xR[0] = centre[2*cR]; xR[1] = centre[2*cR+1];
		} else {
			bc = bcid[-1 - nb];
			ghost_state(gp, bt, bc, uL, n, uR);
			xR[0] = 2.0*G.mx - xL[0]; xR[1] = 2.0*G.my - xL[1];
		}
		const double irR = prim_lean(gp, uR, wR);

		FaceState A, B;
		if constexpr(MODE == 0) {
			A = fs_both(gp, uL, wL, irL, n);
			B = fs_both(gp, uR, wR, irR, n);
		}
		else if constexpr(MODE == 1) {
			double gL[8], r[4];
			load_grad(rgrad, cL, gL);
			const double dx = G.mx - xL[0], dy = G.my - xL[1];
			for(int k = 0; k < 4; k++) r[k] = wL[k] + gL[2*k]*dx + gL[2*k+1]*dy;
			A = fs_prim(gp, r, n, igm1);
			if(!bnd) {
				double gR[8];
				load_grad(rgrad, cR, gR);
				const double ex = G.mx - xR[0], ey = G.my - xR[1];
				for(int k = 0; k < 4; k++) r[k] = wR[k] + gR[2*k]*ex + gR[2*k+1]*ey;
				B = fs_prim(gp, r, n, igm1);
			}
		}
		else {
			// MUSCL-Van Albada, kappa = 1/3, eps = 1e-8
			double gL[8], gR[8];
			load_grad(grad, cL, gL);
			if(!bnd) load_grad(grad, cR, gR);
			const double ex = xR[0] - xL[0], ey = xR[1] - xL[1];
			const double kap = 1.0/3.0, eps = 1e-8;
			double rl[4], rr[4];
			for(int k = 0; k < 4; k++) {
				const double dw = wR[k] - wL[k];
				const double dm = 2.0*(gL[2*k]*ex + gL[2*k+1]*ey) - dw;
				double pl = (2.0*dm*dw + eps)/(dm*dm + dw*dw + eps);
				if(pl < 0.0) pl = 0.0;
				rl[k] = wL[k] + pl*0.25*((1.0 - kap*pl)*dm + (1.0 + kap*pl)*dw);
				if(!bnd) {
					const double dp = 2.0*(gR[2*k]*ex + gR[2*k+1]*ey) - dw;
					double pr = (2.0*dp*dw + eps)/(dp*dp + dw*dw + eps);
					if(pr < 0.0) pr = 0.0;
					rr[k] = wR[k] - pr*0.25*((1.0 - kap*pr)*dp + (1.0 + kap*pr)*dw);
				}
			}
			A = fs_prim(gp, rl, n, igm1);
			if(!bnd) B = fs_prim(gp, rr, n, igm1);
		}
		if constexpr(MODE != 0) {
			if(bnd) {
				const double UL[4] = { A.r, A.mx, A.my, A.E };
				double UG[4];
				ghost_state(gp, bt, bc, UL, n, UG);
				B = fs_cons(gp, UG, n);
			}
		}

		double FL[4];
		fs_num_flux<FLUX>(gp, A, B, n, FL);
		if constexpr(VISC) {
			double vf[4];
			if constexpr(MODE == 0) {
				viscous_flux(gp, n, xL, xR, wL, irL, wR, irR, nullptr, nullptr, A, B, vf);
			} else {
				double vl[8], vr[8];
				load_grad(grad, cL, vl);
				if(!bnd) load_grad(grad, cR, vr);
				else for(int k = 0; k < 8; k++) vr[k] = vl[k];
				viscous_flux(gp, n, xL, xR, wL, irL, wR, irR, vl, vr, A, B, vf);
			}
			for(int k = 0; k < 4; k++) FL[k] += vf[k];
		}
		const double sgn = selfL ? -G.len : G.len;
		for(int k = 0; k < 4; k++) F[k] = sgn*FL[k];

		if(getdt) {
			const FaceState& S = selfL ? A : B;
			srad = (fabs(S.vn) + S.c)*G.len;
			if constexpr(VISC) {
				const double mu = gp.constvisc ? 1.0/gp.Reinf : sutherland(gp, S.p*S.ir*gp.gM2);
				const double co = fmax(4.0/(3*S.r), gp.g/S.r);
				srad += co*mu/gp.Pr*G.len*G.len/area[cc];
			}
		}
	}
	for(int k = 0; k < 4; k++) F[k] = quad_sum(F[k]);
	srad = quad_sum(srad);
	if(valid) {
		if(q == 0) {
			double rv[4];
			load4(res + 4*cc, rv);
			for(int k = 0; k < 4; k++) rv[k] += F[k];
			store4(res + 4*cc, rv);
		} else if(q == 1 && getdt) {
			dt[cc] = area[cc]/srad;
		}
	}
}

}
#endif
