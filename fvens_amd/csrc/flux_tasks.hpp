/* flux_tasks.hpp -- face-task form of the face-flux / residual-gather pass for linear
 * reconstruction without viscous terms (the transonic airfoil, bump and batched operators).
 *
 * Reference semantics: FlowFV::compute_residual with compute_fluxes and compute_max_timestep
 * (src/spatial/flow_spatial.cpp:488-816), linear reconstruction at the face midpoints
 * (src/spatial/reconstruction_utils.hpp:17-32) with the limited gradient of the gradient pass.
 *
 * A workgroup owns FB consecutive cells. It stages their primitive states, reconstruction
 * gradients and centres in LDS once (coalesced loads; one primitive conversion per cell), then
 * evaluates each face that touches its cells exactly once: a face between two of its own cells
 * is one task whose flux goes to both cells' slots, a face to a cell of another workgroup is a
 * task here and one there (both evaluate it from the same canonical left = lower index data, so
 * the two results are identical), a boundary face is one task. Every task deposits -F len / +F len
 * and its side's spectral radius into LDS slots [face k of the cell][cell]; each cell then sums
 * its four slots in the fixed order (s0 + s1) + (s2 + s3), the order of the quad-per-cell form,
 * so the two forms give the same bits. No atomics.
 *
 * The task list of every workgroup is built once per mesh on the host (mesh.cpp): int4
 * {face, left cell, right cell or -1-ib, local slots}, the local slot of the left cell in bits
 * 0..15 and of the right cell in bits 16..31 (0xFFFF: that cell belongs to another workgroup).
 */
#ifndef FVX_FLUX_TASKS_HPP
#define FVX_FLUX_TASKS_HPP

#include "fvx_internal.hpp"
#include "face_state.hpp"
#include "wave.hpp"

namespace fvx {

template <int FLUX>
__global__ __launch_bounds__(FT_THREADS, 4)
void k_flux_tasks(int ncell, const double* __restrict__ u, const double* __restrict__ rgrad,
                  const int4* __restrict__ tasks, const int* __restrict__ tptr,
                  const unsigned char* __restrict__ nslot, const FaceGeom* __restrict__ fgeom,
                  const double* __restrict__ centre, const double* __restrict__ area,
                  const int* __restrict__ bcid, const GasParams* __restrict__ gpa,
                  const BCTable* __restrict__ bta, const BatchArg ba, double* __restrict__ res,
                  double* __restrict__ dt, const int getdt, const int xcd)
{
	const long I = ba.list ? ba.list[blockIdx.y] : 0;
	const GasParams& gp = gpa[I];
	const BCTable& bt = *bta;
	u += I*4L*ncell;
	rgrad += I*8L*ncell;
	res += I*4L*ncell;
	if(dt) dt += I*(long)ncell;
	const int blk = xcd_block(xcd);
	const int c0 = blk*FT_CELLS;
	const int nc = min(FT_CELLS, ncell - c0);
	const int t = threadIdx.x;

	__shared__ double sw[4][FT_CELLS], sg[8][FT_CELLS], sx[2][FT_CELLS];
	__shared__ double sf[5][4*FT_CELLS];

	// stage the workgroup's cells: even threads the primitive state and centre, odd threads
	// the reconstruction gradient
	{
		const int lc = t >> 1;
		if(lc < nc) {
			const int c = c0 + lc;
			if((t & 1) == 0) {
				double uc[4], w[4];
				load4(u + 4*c, uc);
				prim_lean(gp, uc, w);
				for(int k = 0; k < 4; k++) sw[k][lc] = w[k];
				const double2 x = reinterpret_cast<const double2*>(centre)[c];
				sx[0][lc] = x.x;
				sx[1][lc] = x.y;
			} else {
				double g[8];
				load4(rgrad + 8*c, g);
				load4(rgrad + 8*c + 4, g + 4);
				for(int k = 0; k < 8; k++) sg[k][lc] = g[k];
			}
		}
	}
	__syncthreads();

	const double igm1 = 1.0/gp.gm1;
	const int k1 = tptr[blk + 1];
	for(int k = tptr[blk] + t; k < k1; k += FT_THREADS) {
		const int4 T = tasks[k];
		const FaceGeom G = fgeom[T.x];
		const double n[2] = { G.nx, G.ny };
		const bool bnd = T.z < 0;
		// the two cells' states, reconstructed at the face midpoint; a cell of this workgroup
		// comes from LDS, any other from memory through the same conversion
		auto side = [&](int c, double* r) {
			const int lc = c - c0;
			double w[4], g[8], xc[2];
			if((unsigned)lc < (unsigned)nc) {
				for(int v = 0; v < 4; v++) w[v] = sw[v][lc];
				for(int v = 0; v < 8; v++) g[v] = sg[v][lc];
				xc[0] = sx[0][lc];
				xc[1] = sx[1][lc];
			} else {
				double uc[4];
				load4(u + 4*c, uc);
				prim_lean(gp, uc, w);
				load4(rgrad + 8*c, g);
				load4(rgrad + 8*c + 4, g + 4);
				const double2 x = reinterpret_cast<const double2*>(centre)[c];
				xc[0] = x.x;
				xc[1] = x.y;
			}
			const double dx = G.mx - xc[0], dy = G.my - xc[1];
			for(int v = 0; v < 4; v++) r[v] = w[v] + g[2*v]*dx + g[2*v+1]*dy;
		};
		double r[4];
		side(T.y, r);
		const FaceState A = fs_prim(gp, r, n, igm1);
		FaceState B;
		if(!bnd) {
			side(T.z, r);
			B = fs_prim(gp, r, n, igm1);
		} else {
			const double UL[4] = { A.r, A.mx, A.my, A.E };
			double UG[4];
			ghost_state(gp, bt, bcid[-1 - T.z], UL, n, UG);
			B = fs_cons(gp, UG, n);
		}
		double FL[4];
		fs_num_flux<FLUX>(gp, A, B, n, FL);
		const unsigned sl = (unsigned)T.w & 0xFFFFu, sr = (unsigned)T.w >> 16;
		if(sl != 0xFFFFu) {
			for(int v = 0; v < 4; v++) sf[v][sl] = -G.len*FL[v];
			sf[4][sl] = (fabs(A.vn) + A.c)*G.len;
		}
		if(sr != 0xFFFFu) {
			for(int v = 0; v < 4; v++) sf[v][sr] = G.len*FL[v];
			sf[4][sr] = (fabs(B.vn) + B.c)*G.len;
		}
	}
	__syncthreads();

	if(t < nc) {
		const int c = c0 + t;
		const int ns = nslot[c];
		double acc[5];
		for(int v = 0; v < 5; v++) {
			const double s0 = sf[v][t], s1 = sf[v][FT_CELLS + t], s2 = sf[v][2*FT_CELLS + t];
			const double s3 = ns > 3 ? sf[v][3*FT_CELLS + t] : 0.0;
			acc[v] = (s0 + s1) + (s2 + s3);
		}
		double rv[4];
		load4(res + 4*c, rv);
		for(int v = 0; v < 4; v++) rv[v] += acc[v];
		store4(res + 4*c, rv);
		if(getdt) dt[c] = area[c]/acc[4];
	}
}

}
#endif
