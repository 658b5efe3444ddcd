/* tune.cpp -- autotuner for the launch shapes of the cell kernels and for the preconditioner
 * schedule (no reference counterpart: the reference's OpenMP loops have no launch shape).
 *
 * Launch shapes: every candidate (threads per workgroup x XCD-aware block order) of the
 * gradient, face-flux and Jacobian kernels is timed with HIP events on the caller's stream at
 * the caller's state, and the fastest is kept in the discretisation; for the face-flux pass the
 * wave-task form (both block orders) competes with the quad-per-cell shapes where it applies.
 * The kernels compute each cell independently of how cells are grouped into workgroups, and
 * the two flux forms evaluate the same face fluxes and sum them in the same order, so results
 * are bitwise the same for every candidate.
 *
 * Preconditioner schedule: for the linear system of one pseudo-time step at the given state
 * and CFL number, each preconditioner type and sweep count is timed over set-up plus one
 * FGMRES solve to the given tolerance; candidates whose solve does not reach it are discarded.
 */
#include "fvx_internal.hpp"
#include "wave.hpp"
#include <algorithm>

namespace fvx {

namespace {

struct EventTimer {
	hipEvent_t a, b;
	EventTimer() { FVX_HIP(hipEventCreate(&a)); FVX_HIP(hipEventCreate(&b)); }
	~EventTimer() { (void)hipEventDestroy(a); (void)hipEventDestroy(b); }
	template <class F> double ms(F&& f, int reps, hipStream_t s)
	{
		f();                                   // warm-up
		FVX_HIP(hipEventRecord(a, s));
		for(int i = 0; i < reps; i++) f();
		FVX_HIP(hipEventRecord(b, s));
		FVX_HIP(hipEventSynchronize(b));
		float t = 0;
		FVX_HIP(hipEventElapsedTime(&t, a, b));
		return (double)t/reps;
	}
};

}

void autotune_launch(fvx_disc_s& d, const double* u, int reps, std::vector<fvx_tune_row>& rows,
                     hipStream_t s)
{
	const fvx_mesh_s& m = *d.m;
	const size_t n = (size_t)m.nelem;
	DBuf<double> r, dt;
	r.alloc(4*n);
	dt.alloc(n);
	fvx_matrix_s A;
	const bool jac = jacobian_supported(d);
	if(jac) matrix_init(A, &m, FVX_PREC_F64);
	EventTimer T;
	const int blocks[4] = { 64, 128, 256, 512 };
	const bool order2 = d.nc.order2 && d.nc.gradientscheme != FVX_GRAD_NONE;
	const int rec = d.nc.reconstruction;
	const bool limited = rec == FVX_RECON_VENKATAKRISHNAN || rec == FVX_RECON_BARTHJESPERSEN;
	const bool visc = d.pc.viscous != 0;
	const bool wave = flux_wave_apply(d);
	for(int kernel = FVX_KERNEL_GRADIENTS; kernel <= FVX_KERNEL_JACOBIAN; kernel++) {
		if(kernel == FVX_KERNEL_GRADIENTS && !order2) continue;
		if(kernel == FVX_KERNEL_JACOBIAN && !jac) continue;
		double best = 1e300;
		int bb = 256, bx = 1, bw = 0;
		// the face-flux pass: the quad-per-cell form's shapes, then the wave-task form
		for(int bi = 0; bi < (kernel == FVX_KERNEL_FLUX && wave ? 5 : 4); bi++) {
			const bool wform = bi == 4;
			const int block = wform ? 64 : blocks[bi];
			if(kernel == FVX_KERNEL_JACOBIAN && block > 256) continue;   // its launch bound
			for(int xcd = 0; xcd <= 1; xcd++) {
				double t;
				if(kernel == FVX_KERNEL_GRADIENTS) {
					d.tune.grad_block = block; d.tune.grad_xcd = xcd;
					t = T.ms([&] {
						launch_gradients(d, u, limited ? (visc ? d.d_grad.p : nullptr) : d.d_grad.p,
						                 limited ? d.d_rgrad.p : nullptr, 0, s); }, reps, s);
				} else if(kernel == FVX_KERNEL_FLUX) {
					if(!wform) d.tune.flux_block = block;
					d.tune.flux_xcd = xcd; d.tune.flux_wave = wform;
					// the whole residual; the gradient pass is the same for every candidate
					t = T.ms([&] { launch_residual(d, u, r.p, 1, dt.p, s); }, reps, s);
				} else {
					d.tune.jac_block = block; d.tune.jac_xcd = xcd;
					t = T.ms([&] { launch_jacobian(d, u, A.raw(), A.prec, nullptr, 0.0, nullptr, s); },
					         reps, s);
				}
				rows.push_back(fvx_tune_row{wform ? (int)FVX_KERNEL_FLUX_WAVE : kernel, block, xcd, t});
				if(t < best) { best = t; bb = block; bx = xcd; bw = wform; }
			}
		}
		if(kernel == FVX_KERNEL_GRADIENTS) { d.tune.grad_block = bb; d.tune.grad_xcd = bx; }
		else if(kernel == FVX_KERNEL_FLUX) {
			// the quad form keeps its fastest shape even when the wave form wins
			double bq = 1e300;
			for(const fvx_tune_row& w : rows)
				if(w.kernel == FVX_KERNEL_FLUX && w.ms < bq) { bq = w.ms; d.tune.flux_block = w.block; }
			d.tune.flux_xcd = bx; d.tune.flux_wave = bw;
		}
		else { d.tune.jac_block = bb; d.tune.jac_xcd = bx; }
	}
	FVX_HIP(hipStreamSynchronize(s));
}

void autotune_pc(fvx_disc_s& d, const double* u, double cfl, int restart, double rtol,
                 int precision, std::vector<fvx_tune_row>& rows, int& best_pc, int& best_sweeps,
                 hipStream_t s)
{
	if(!jacobian_supported(d))
		throw Error(FVX_ERR_UNSUPPORTED, "the reference has no Jacobian for this flux or BC");
	const fvx_mesh_s& m = *d.m;
	const size_t n = (size_t)m.nelem;
	DBuf<double> r, dt, dtm, x;
	r.alloc(4*n); dt.alloc(n); dtm.alloc(n); x.alloc(4*n);
	FVX_HIP(hipMemsetAsync(r.p, 0, 4*n*sizeof(double), s));
	launch_residual(d, u, r.p, 1, dt.p, s);
	fvx_matrix_s A;
	matrix_init(A, &m, precision);
	launch_jacobian(d, u, A.raw(), A.prec, dt.p, cfl, dtm.p, s);
	EventTimer T;
	double best = 1e300;
	best_pc = FVX_PC_SGS;
	best_sweeps = 1;
	const int types[2] = { FVX_PC_SGS, FVX_PC_ILU0 };
	const int sweeps[5] = { 1, 2, 4, 8, 16 };
	for(int type : types)
		for(int sw : sweeps) {
			fvx_pc_s P;
			pc_init(P, &m, type, precision);
			pc_set_sweeps(P, sw);
			int its = 0;
			bool ok = true;
			const double t = T.ms([&] {
				launch_pc_setup(P, A, Inst(), s);
				try { its = linear_solve(A, P, r.p, x.p, restart, rtol, s); }
				catch(const Error&) { ok = false; }
			}, 1, s);
			// a solve that used every iteration has not shown that it reaches the tolerance
			if(its >= restart) ok = false;
			rows.push_back(fvx_tune_row{type, sw, its, ok ? t : -t});
			if(ok && t < best) { best = t; best_pc = type; best_sweeps = sw; }
		}
	FVX_HIP(hipStreamSynchronize(s));
}

}
