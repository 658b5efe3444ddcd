/* solver.cpp -- host side of the implicit solver: matrix and preconditioner set-up, the
 * flexible GMRES driver, the matrix-free operator and the steady pseudo-time loop. Every
 * vector operation runs on the device; the host only orders launches, keeps the small
 * Hessenberg system of GMRES and the CFL schedule, and reads one scalar per Krylov iteration
 * and per pseudo-time step.
 */
#include "fvx_internal.hpp"
#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstring>
#include <functional>

namespace fvx {

/* ---------------------------------------------------------------- matrix --------------- */

void matrix_init(fvx_matrix_s& A, const fvx_mesh_s* m, int prec)
{
	if(prec != FVX_PREC_F64 && prec != FVX_PREC_F32)
		throw Error(FVX_ERR_INVALID, "unknown matrix precision");
	A.m = m;
	A.prec = prec;
	const size_t cnt = (size_t)m->nelem*80;
	if(prec == FVX_PREC_F32) {
		A.v32.alloc(cnt);
		FVX_HIP(hipMemset(A.v32.p, 0, cnt*sizeof(float)));
	} else {
		A.v64.alloc(cnt);
		FVX_HIP(hipMemset(A.v64.p, 0, cnt*sizeof(double)));
	}
}

static inline bool interior_slot(const fvx_mesh_s& m, int c, int k)
{
	if(k >= m.nnode[c]) return false;
	const int o = m.esuel[(size_t)c*4 + k];
	return o >= 0 && o < m.nelem;
}

int matrix_bsr_nnzb(const fvx_mesh_s& m) { return m.nelem + 2*m.ninface; }

void matrix_export_bsr(const fvx_matrix_s& A, int* rowptr, int* colind, double* vals)
{
	const fvx_mesh_s& m = *A.m;
	const int n = m.nelem;
	std::vector<double> h((size_t)n*80);
	FVX_HIP(hipDeviceSynchronize());
	if(A.prec == FVX_PREC_F32) {
		std::vector<float> f((size_t)n*80);
		FVX_HIP(hipMemcpy(f.data(), A.v32.p, f.size()*sizeof(float), hipMemcpyDeviceToHost));
		for(size_t i = 0; i < f.size(); i++) h[i] = f[i];
	} else {
		FVX_HIP(hipMemcpy(h.data(), A.v64.p, h.size()*sizeof(double), hipMemcpyDeviceToHost));
	}
	int pos = 0;
	for(int c = 0; c < n; c++) {
		if(rowptr) rowptr[c] = pos;
		std::pair<int,int> e[5];
		int ne = 0;
		e[ne++] = { c, 0 };
		for(int k = 0; k < 4; k++)
			if(interior_slot(m, c, k)) e[ne++] = { m.esuel[(size_t)c*4 + k], 1 + k };
		std::sort(e, e + ne);
		for(int t = 0; t < ne; t++, pos++) {
			if(colind) colind[pos] = e[t].first;
			if(vals) std::memcpy(vals + (size_t)pos*16, &h[(size_t)c*80 + 16*e[t].second],
			                     16*sizeof(double));
		}
	}
	if(rowptr) rowptr[n] = pos;
}

/* ------------------------------------------------------------ preconditioner ----------- */

void pc_init(fvx_pc_s& p, const fvx_mesh_s* m, int type, int prec)
{
	if(type < FVX_PC_NONE || type > FVX_PC_ILU0)
		throw Error(FVX_ERR_INVALID, "unknown preconditioner type");
	if(prec != FVX_PREC_F64 && prec != FVX_PREC_F32)
		throw Error(FVX_ERR_INVALID, "unknown preconditioner precision");
	p.m = m;
	p.type = type;
	p.prec = prec;
	const int n = m->nelem;
	// greedy colouring in cell order: the smallest colour unused by lower-numbered neighbours
	p.colour.assign(n, 0);
	p.ncol = 1;
	if(type == FVX_PC_SGS || type == FVX_PC_ILU0) {
		for(int c = 0; c < n; c++) {
			unsigned used = 0;
			for(int k = 0; k < 4; k++)
				if(interior_slot(*m, c, k)) {
					const int o = m->esuel[(size_t)c*4 + k];
					if(o < c) used |= 1u << p.colour[o];
				}
			int col = 0;
			while(used & (1u << col)) col++;
			p.colour[c] = col;
			p.ncol = std::max(p.ncol, col + 1);
		}
	}
	p.coff.assign(p.ncol + 1, 0);
	for(int c = 0; c < n; c++) p.coff[p.colour[c] + 1]++;
	for(int k = 0; k < p.ncol; k++) p.coff[k+1] += p.coff[k];
	p.order.assign(n, 0);
	std::vector<int> fill(p.coff.begin(), p.coff.end() - 1);
	for(int c = 0; c < n; c++) p.order[fill[p.colour[c]]++] = c;
	std::vector<unsigned char> low(n, 0);
	for(int c = 0; c < n; c++)
		for(int k = 0; k < 4; k++)
			if(interior_slot(*m, c, k) && p.colour[m->esuel[(size_t)c*4 + k]] < p.colour[c])
				low[c] |= (unsigned char)(1u << k);
	if(n > 0) {
		p.d_colour.upload(p.colour);
		p.d_order.upload(p.order);
		p.d_lowmask.upload(low);
	}
	if(type != FVX_PC_NONE) {
		const size_t cnt = (size_t)n*80;
		if(prec == FVX_PREC_F32) {
			p.v32.alloc(cnt);
			FVX_HIP(hipMemset(p.v32.p, 0, cnt*sizeof(float)));
		} else {
			p.v64.alloc(cnt);
			FVX_HIP(hipMemset(p.v64.p, 0, cnt*sizeof(double)));
		}
	}
}

void pc_set_sweeps(fvx_pc_s& p, int sweeps)
{
	if(sweeps < 1 || sweeps > 64) throw Error(FVX_ERR_INVALID, "sweep count must be in [1, 64]");
	p.sweeps = sweeps;
	if(sweeps > 1 && p.type != FVX_PC_SGS && p.type != FVX_PC_NONE) {
		const size_t N = 4*(size_t)p.m->nelem;
		if(p.t1.n != N) { p.t1.alloc(N); p.t2.alloc(N); }
	}
}

/* ---------------------------------------------------------------- Krylov --------------- */

/// Device storage of one FGMRES cycle and the pinned host slots for its scalars.
struct Krylov {
	long N = 0;
	int m = 0;
	DBuf<double> V, Z, partial, scal;
	double* hpin = nullptr;
	static constexpr int H0 = 0, NRM = 34, BETA = 35, XX = 36, RES = 37, Y0 = 40, NSCAL = 80;

	void ensure(long n_, int m_)
	{
		if(!hpin) FVX_HIP(hipHostMalloc((void**)&hpin, NSCAL*sizeof(double)));
		if(n_ == N && m_ <= m) return;
		N = n_;
		m = m_;
		V.alloc((size_t)(m + 1)*N);
		Z.alloc((size_t)m*N);
		partial.alloc((size_t)reduction_blocks()*max_dots());
		scal.alloc(NSCAL);
	}
	~Krylov() { if(hpin) (void)hipHostFree(hpin); }

	void fetch(int first, int count, hipStream_t s)
	{
		FVX_HIP(hipMemcpyAsync(hpin + first, scal.p + first, count*sizeof(double),
		                       hipMemcpyDeviceToHost, s));
		FVX_HIP(hipStreamSynchronize(s));
	}

	using Op = std::function<void(const double*, double*)>;

	/// One cycle of right-preconditioned FGMRES from x = 0 with classical Gram-Schmidt.
	int solve(const Op& op, const Op& pc, const double* b, double* x, int mm, double rtol,
	          hipStream_t s)
	{
		double* d_h = scal.p + H0;
		launch_mdot(N, b, 0, 1, b, partial.p, scal.p + BETA, s);
		fetch(BETA, 1, s);
		const double beta = std::sqrt(hpin[BETA]);
		if(!std::isfinite(beta)) throw Error(FVX_ERR_NONFINITE, "non-finite right-hand side");
		if(beta == 0.0) {
			FVX_HIP(hipMemsetAsync(x, 0, N*sizeof(double), s));
			return 0;
		}
		launch_scale_copy(N, b, scal.p + BETA, V.p, s);
		std::vector<double> H((size_t)(mm + 1)*mm, 0.0), cs(mm), sn(mm), g(mm + 1, 0.0);
		g[0] = beta;
		int it = 0;
		for(int j = 0; j < mm; j++) {
			double* vj = V.p + (size_t)j*N;
			double* zj = Z.p + (size_t)j*N;
			double* w = V.p + (size_t)(j + 1)*N;
			pc(vj, zj);
			op(zj, w);
			launch_mdot(N, V.p, N, j + 1, w, partial.p, d_h, s);
			launch_gs_update(N, V.p, N, j + 1, d_h, w, partial.p, scal.p + NRM, s);
			fetch(H0, NRM + 1, s);
			for(int i = 0; i <= j; i++) H[(size_t)i*mm + j] = hpin[H0 + i];
			const double hn = std::sqrt(hpin[NRM]);
			if(!std::isfinite(hn)) throw Error(FVX_ERR_NONFINITE, "non-finite Krylov vector");
			H[(size_t)(j + 1)*mm + j] = hn;
			if(hn != 0.0) launch_normalise(N, w, scal.p + NRM, s);
			for(int i = 0; i < j; i++) {
				const double t = cs[i]*H[(size_t)i*mm + j] + sn[i]*H[(size_t)(i + 1)*mm + j];
				H[(size_t)(i + 1)*mm + j] = -sn[i]*H[(size_t)i*mm + j] + cs[i]*H[(size_t)(i + 1)*mm + j];
				H[(size_t)i*mm + j] = t;
			}
			const double a = H[(size_t)j*mm + j], c = H[(size_t)(j + 1)*mm + j];
			const double den = std::sqrt(a*a + c*c);
			cs[j] = a/den;
			sn[j] = c/den;
			H[(size_t)j*mm + j] = den;
			H[(size_t)(j + 1)*mm + j] = 0.0;
			g[j + 1] = -sn[j]*g[j];
			g[j] = cs[j]*g[j];
			it = j + 1;
			if(std::fabs(g[j + 1]) <= rtol*beta || hn == 0.0) break;
		}
		for(int i = it - 1; i >= 0; i--) {
			double t = g[i];
			for(int k = i + 1; k < it; k++) t -= H[(size_t)i*mm + k]*hpin[Y0 + k];
			hpin[Y0 + i] = t/H[(size_t)i*mm + i];
		}
		FVX_HIP(hipMemcpyAsync(scal.p + Y0, hpin + Y0, it*sizeof(double),
		                       hipMemcpyHostToDevice, s));
		launch_lincomb(N, Z.p, N, it, scal.p + Y0, x, s);
		FVX_HIP(hipStreamSynchronize(s));
		return it;
	}
};

static void check_restart(int restart)
{
	if(restart < 1 || restart > max_dots() - 1)
		throw Error(FVX_ERR_INVALID, "Krylov restart length must be in [1, 31]");
}

int linear_solve(const fvx_matrix_s& A, const fvx_pc_s& p, const double* b, double* x,
                 int restart, double rtol, hipStream_t s)
{
	check_restart(restart);
	Krylov k;
	k.ensure(4L*A.m->nelem, restart);
	return k.solve([&](const double* xi, double* yo) { launch_spmv(A, xi, yo, s); },
	               [&](const double* r, double* z) { launch_pc_apply(p, r, z, s); },
	               b, x, restart, rtol, s);
}

void matfree_apply(const fvx_disc_s& d, const double* u, const double* ru, const double* dtm,
                   const double* x, double* y, double eps, double* aux, double* raux,
                   double* partial, double* xx, hipStream_t s)
{
	const int n = d.m->nelem;
	const long N = 4L*n;
	launch_mdot(N, x, 0, 1, x, partial, xx, s);
	launch_mf_prep(N, u, x, xx, eps, aux, s);
	FVX_HIP(hipMemsetAsync(raux, 0, N*sizeof(double), s));
	launch_residual(d, aux, raux, 0, nullptr, s);
	launch_mf_finish(n, x, dtm, ru, raux, xx, eps, y, s);
}

/* ------------------------------------------------------------ steady solver ------------ */

struct SolverWork {
	int n = -1, jprec = -1, pctype = -1, pcprec = -1;
	fvx_matrix_s A;
	fvx_pc_s P;
	DBuf<double> r, dt, dtm, du, aux, raux;
	Krylov kry;
};

void destroy_solver_work(SolverWork* w) { delete w; }

/// SteadySolver::expResidualRamp (src/ode/aodesolver.cpp:110-120)
static double exp_ramp(double cmin, double cmax, double prev, double ratio, double up,
                       double down)
{
	const double nc = ratio > 1.0 ? prev*std::pow(ratio, up) : prev*std::pow(ratio, down);
	if(nc < cmin) return cmin;
	if(nc > cmax) return cmax;
	return nc;
}

void steady_solve(fvx_disc_s& d, const fvx_solver_cfg& c, double* u, fvx_solve_stats& st,
                  hipStream_t s)
{
	std::memset(&st, 0, sizeof(st));
	if(!(c.tol > 0) || c.maxiter < 0 || !(c.cflinit > 0) || !(c.cflfin >= c.cflinit))
		throw Error(FVX_ERR_INVALID, "invalid pseudo-time settings");
	const fvx_mesh_s& m = *d.m;
	const int n = m.nelem;
	const long N = 4L*n;
	if(c.implicit) {
		check_restart(c.krylov_restart);
		if(!jacobian_supported(d))
			throw Error(FVX_ERR_UNSUPPORTED, "the reference has no Jacobian for this flux or BC");
		if(c.matrix_free && !(c.mf_eps > 0))
			throw Error(FVX_ERR_INVALID, "matrix-free step must be positive");
	}
	if(!d.work) d.work = new SolverWork;
	SolverWork& w = *d.work;
	if(w.n != n) {
		w.r.alloc(N); w.dt.alloc(n); w.dtm.alloc(n); w.du.alloc(N);
		w.aux.alloc(N); w.raux.alloc(N);
		w.n = n;
		w.jprec = w.pctype = w.pcprec = -1;
	}
	if(c.implicit) {
		if(w.jprec != c.jac_precision) { matrix_init(w.A, &m, c.jac_precision); w.jprec = c.jac_precision; }
		if(w.pctype != c.pc || w.pcprec != c.pc_precision) {
			pc_init(w.P, &m, c.pc, c.pc_precision);
			w.pctype = c.pc;
			w.pcprec = c.pc_precision;
		}
		pc_set_sweeps(w.P, c.pc_sweeps > 0 ? c.pc_sweeps : 1);
		w.kry.ensure(N, c.krylov_restart);
	}
	Krylov& K = w.kry;
	if(!K.hpin) K.ensure(N, 1);

	double cfl = 0.0, resi = 1.0, resiold = 1.0;
	double initres = c.fixed_norm > 0 ? c.fixed_norm : 1.0;
	int step = 0;
	FVX_HIP(hipStreamSynchronize(s));
	const auto t0 = std::chrono::steady_clock::now();
	while(resi/initres > c.tol && step < c.maxiter) {
		FVX_HIP(hipMemsetAsync(w.r.p, 0, N*sizeof(double), s));
		launch_residual(d, u, w.r.p, 1, w.dt.p, s);
		launch_resnorm(n, w.r.p, m.d_area.p, K.partial.p, K.scal.p + Krylov::RES, s);
		if(c.implicit) {
			cfl = exp_ramp(c.cflinit, c.cflfin, cfl, resiold/resi, 0.25, 0.3);
			launch_jacobian(d, u, w.A.raw(), w.A.prec, w.dt.p, cfl, w.dtm.p, s);
			d.prof.start(PROF_PC, s);
			launch_pc_setup(w.P, w.A, s);
			d.prof.stop(s);
			Krylov::Op op;
			if(c.matrix_free)
				op = [&](const double* x, double* y) {
					matfree_apply(d, u, w.r.p, w.dtm.p, x, y, c.mf_eps, w.aux.p, w.raux.p,
					              K.partial.p, K.scal.p + Krylov::XX, s);
				};
			else
				op = [&](const double* x, double* y) { launch_spmv(w.A, x, y, s); };
			// the residual norm was reduced before the Krylov solve reuses the partials
			FVX_HIP(hipMemcpyAsync(K.hpin + Krylov::RES, K.scal.p + Krylov::RES,
			                       sizeof(double), hipMemcpyDeviceToHost, s));
			FVX_HIP(hipStreamSynchronize(s));
			const double rn2 = K.hpin[Krylov::RES];
			if(!std::isfinite(rn2)) {
				resiold = resi;
				resi = std::sqrt(rn2);
				st.status = FVX_ERR_NONFINITE;
				break;
			}
			const int its = K.solve(op, [&](const double* r, double* z) {
				                        launch_pc_apply(w.P, r, z, s); },
			                        w.r.p, w.du.p, c.krylov_restart, c.krylov_rtol, s);
			st.lin_iters += its;
			launch_update_implicit(n, u, w.du.p, c.fullupdate, c.min_relax, d.gas.g, s);
			resiold = resi;
			resi = std::sqrt(rn2);
		} else {
			cfl = exp_ramp(c.cflinit, c.cflfin, cfl, resiold/resi, 0.3, 0.25);
			launch_update_explicit(n, u, w.r.p, w.dt.p, m.d_area.p, c.cflinit, s);
			FVX_HIP(hipMemcpyAsync(K.hpin + Krylov::RES, K.scal.p + Krylov::RES,
			                       sizeof(double), hipMemcpyDeviceToHost, s));
			FVX_HIP(hipStreamSynchronize(s));
			resiold = resi;
			resi = std::sqrt(K.hpin[Krylov::RES]);
		}
		if(!std::isfinite(resi)) { st.status = FVX_ERR_NONFINITE; break; }
		if(step == 0 && c.fixed_norm <= 0) initres = resi;
		step++;
	}
	FVX_HIP(hipStreamSynchronize(s));
	st.walltime = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
	st.steps = step;
	st.final_abs = resi;
	st.first_norm = initres;
	st.final_rel = resi/initres;
	st.converged = (st.status == 0 && resi/initres <= c.tol) ? 1 : 0;
	if(st.status == 0 && !st.converged) st.status = FVX_ERR_NOT_CONVERGED;
	if(st.status == FVX_ERR_NONFINITE)
		throw Error(FVX_ERR_NONFINITE, "non-finite residual in the pseudo-time loop");
	if(st.status == FVX_ERR_NOT_CONVERGED)
		throw Error(FVX_ERR_NOT_CONVERGED, "pseudo-time loop reached its step limit");
}

}
