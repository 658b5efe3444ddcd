/* flux_wave.hpp -- wave-task form of the face-flux / residual-gather pass for first-order and
 * linear reconstruction without viscous terms.
 *
 * Reference semantics: FlowFV::compute_residual with compute_fluxes and compute_max_timestep
 * (src/spatial/flow_spatial.cpp:488-816), reconstruction at the face midpoints
 * (src/spatial/reconstruction_utils.hpp:17-32) with the (limited) gradient of the gradient pass.
 *
 * The quad-per-cell form evaluates every interior face twice, once in each cell's quad. Here a
 * workgroup is one wave64 that owns a group of consecutive cells, chosen on the host so that the
 * faces touching them make at most 64 tasks: a face between two cells of the group is one task
 * serving both, a face to a cell outside the group or to the boundary is one task serving the
 * group's side. Lane t evaluates task t from the face's canonical (left = lower index) states,
 * exactly as the quad form does, and leaves len F and both sides' spectral radii in LDS; then
 * lane t < (cells in the group) sums its cell's four slots in the quad form's order
 * (s0 + s1) + (s2 + s3), with -len F on the left side and +len F on the right. No atomics, no
 * workgroup barrier beyond the one wave. On a compact cell order (Hilbert) a group holds ~25
 * cells instead of the quad form's 16 per wave, so the face arithmetic per cell falls by ~1.5x.
 *
 * Host tables (mesh.cpp, build_wave_tasks): tasks[64 g + t] = the face of task t (-1: no task;
 * the face record holds its left cell and its right cell or -1-ib); groups[g] = {first cell,
 * cells}; cslot[64 g + j] = the slot word of the group's cell j, byte k = 2 task + side for its
 * local face k (side 1: the cell is the face's right cell), 0xff for an unused slot.
 */
#ifndef FVX_FLUX_WAVE_HPP
#define FVX_FLUX_WAVE_HPP

#include "fvx_internal.hpp"
#include "face_state.hpp"
#include "wave.hpp"

namespace fvx {

template <int FLUX, int MODE>
__global__ __launch_bounds__(64, 7)
void k_flux_wave(int ncell, const double* __restrict__ u, const double* __restrict__ rgrad,
                 const int* __restrict__ tasks, const int2* __restrict__ groups,
                 const unsigned* __restrict__ cslot, const FaceGeom* __restrict__ fgeom,
                 const double* __restrict__ centre, const double* __restrict__ area,
                 const int* __restrict__ bcid, const GasParams* __restrict__ gpa,
                 const BCTable* __restrict__ bta, const BatchArg ba, double* __restrict__ res,
                 double* __restrict__ dt, const int getdt, const int xcd)
{
	const long I = ba.list ? ba.list[blockIdx.y] : 0;
	if(I < 0) return;                     // a retired slot of a compacted list
	const GasParams& gp = gpa[I];
	const BCTable& bt = *bta;
	u += I*4L*ncell;
	if(rgrad) rgrad += I*8L*ncell;
	res += I*4L*ncell;
	if(dt) dt += I*(long)ncell;
	const int g = xcd_block(xcd);
	const int t = threadIdx.x;
	__shared__ double sp[6][64];          // per task: len F, left and right spectral radii
	const double igm1 = 1.0/gp.gm1;

	const int f = tasks[(size_t)g*64 + t];
	if(f >= 0) {
		const FaceGeom G = fgeom[f];
		const double n[2] = { G.nx, G.ny };
		const bool bnd = G.right < 0;
		const int cL = G.left;
		double uL[4], wL[4], xL[2];
		load4(u + 4*cL, uL);
		xL[0] = centre[2*cL]; xL[1] = centre[2*cL+1];
		const double irL = prim_lean(gp, uL, wL);
		double uR[4], wR[4], xR[2];
		int bc = 0;
		if(!bnd) {
			load4(u + 4*G.right, uR);
			xR[0] = centre[2*G.right]; xR[1] = centre[2*G.right+1];
		} else {
			bc = bcid[-1 - G.right];
			ghost_state(gp, bt, bc, uL, n, uR);
			xR[0] = 2.0*G.mx - xL[0]; xR[1] = 2.0*G.my - xL[1];
		}
		const double irR = prim_lean(gp, uR, wR);
		FaceState A, B;
		if constexpr(MODE == 0) {
			A = fs_both(gp, uL, wL, irL, n);
			B = fs_both(gp, uR, wR, irR, n);
		} else {
			double gL[8], r[4];
			load4(rgrad + 8*cL, gL);
			load4(rgrad + 8*cL + 4, gL + 4);
			const double dx = G.mx - xL[0], dy = G.my - xL[1];
			for(int k = 0; k < 4; k++) r[k] = wL[k] + gL[2*k]*dx + gL[2*k+1]*dy;
			A = fs_prim(gp, r, n, igm1);
			if(!bnd) {
				double gR[8];
				load4(rgrad + 8*G.right, gR);
				load4(rgrad + 8*G.right + 4, gR + 4);
				const double ex = G.mx - xR[0], ey = G.my - xR[1];
				for(int k = 0; k < 4; k++) r[k] = wR[k] + gR[2*k]*ex + gR[2*k+1]*ey;
				B = fs_prim(gp, r, n, igm1);
			} else {
				const double UL[4] = { A.r, A.mx, A.my, A.E };
				double UG[4];
				ghost_state(gp, bt, bc, UL, n, UG);
				B = fs_cons(gp, UG, n);
			}
		}
		double FL[4];
		fs_num_flux<FLUX>(gp, A, B, n, FL);
		for(int v = 0; v < 4; v++) sp[v][t] = G.len*FL[v];
		sp[4][t] = (fabs(A.vn) + A.c)*G.len;
		sp[5][t] = (fabs(B.vn) + B.c)*G.len;
	}
	__syncthreads();

	const int2 gr = groups[g];
	if(t < gr.y) {
		const int c = gr.x + t;
		const unsigned cs = cslot[(size_t)g*64 + t];
		double s[4][5];
		for(int k = 0; k < 4; k++) {
			const unsigned code = (cs >> 8*k) & 0xffu;
			if(code == 0xffu) {
				for(int v = 0; v < 5; v++) s[k][v] = 0.0;
				continue;
			}
			const int task = (int)(code >> 1), side = (int)(code & 1u);
			for(int v = 0; v < 4; v++) s[k][v] = side ? sp[v][task] : -sp[v][task];
			s[k][4] = sp[4 + side][task];
		}
		double rv[4];
		load4(res + 4*c, rv);
		for(int v = 0; v < 4; v++) rv[v] += (s[0][v] + s[1][v]) + (s[2][v] + s[3][v]);
		store4(res + 4*c, rv);
		if(getdt) dt[c] = area[c]/((s[0][4] + s[1][4]) + (s[2][4] + s[3][4]));
	}
}

}
#endif
