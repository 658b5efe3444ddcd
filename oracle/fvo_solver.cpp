/* fvo_solver.cpp -- oracle steady pseudo-time solver and linear algebra
 * (TEST INFRASTRUCTURE ONLY).
 *
 * Restates:
 *  - SteadySolver::expResidualRamp           src/ode/aodesolver.cpp:110-120
 *  - SteadyForwardEulerSolver::solve         src/ode/aodesolver.cpp:135-282
 *  - SteadyBackwardEulerSolver::addPseudoTimeTerm / solve  src/ode/aodesolver.cpp:300-638
 *  - FlowSimpleUpdate::getLocalRelaxationFactor  src/ode/nonlinearrelaxation.cpp:25-38
 *  - getDeltaPressureFromConserved           src/physics/aphysics_defs.hpp:65-81
 *  - MatrixFreeSpatialJacobian::apply        src/linalg/alinalg.cpp:142-233
 * The linear solver and preconditioners stand in for the reference's external PETSc calls
 * (aodesolver.cpp:483; options testcases/defaults.solverc): FGMRES(m) with right
 * preconditioning and zero initial guess, block ILU(0) on the BSR pattern in the given cell
 * order, symmetric block Gauss-Seidel (PETSc SOR defaults), or point-block Jacobi.
 * Parity: PETSc's own iterates are unpinned (no reference fixture exists); ILU(0) is pinned
 * against a dense LU of the same sparsity-restricted factorisation on tiny matrices, GMRES
 * against a dense solve, and the whole loop against the reference's converged NACA0012
 * functionals (testcases/naca0012/regr-MUSCL_LeastSquares_HLLC.txt).
 */
#include "fvo.hpp"
#include "fvo_physics.hpp"
#include <chrono>
#include <cstring>
#include <algorithm>

namespace fvo {

void invert4(const double* a, double* ainv)
{
	// Gauss-Jordan with partial pivoting on a 4x4 block
	double m[4][8];
	for(int i = 0; i < 4; i++)
		for(int j = 0; j < 8; j++)
			m[i][j] = j < 4 ? a[i*4+j] : (j-4 == i ? 1.0 : 0.0);
	for(int c = 0; c < 4; c++) {
		int p = c;
		for(int r = c+1; r < 4; r++) if(std::fabs(m[r][c]) > std::fabs(m[p][c])) p = r;
		if(p != c) for(int j = 0; j < 8; j++) std::swap(m[c][j], m[p][j]);
		const double piv = m[c][c];
		for(int j = 0; j < 8; j++) m[c][j] /= piv;
		for(int r = 0; r < 4; r++) {
			if(r == c) continue;
			const double fct = m[r][c];
			for(int j = 0; j < 8; j++) m[r][j] -= fct*m[c][j];
		}
	}
	for(int i = 0; i < 4; i++)
		for(int j = 0; j < 4; j++) ainv[i*4+j] = m[i][j+4];
}

static inline void gemv4(const double* A, const double* x, double* y)
{
	for(int i = 0; i < 4; i++) {
		double s = 0;
		for(int j = 0; j < 4; j++) s += A[i*4+j]*x[j];
		y[i] = s;
	}
}
static inline void gemm4(const double* A, const double* B, double* C)
{
	for(int i = 0; i < 4; i++)
		for(int j = 0; j < 4; j++) {
			double s = 0;
			for(int k = 0; k < 4; k++) s += A[i*4+k]*B[k*4+j];
			C[i*4+j] = s;
		}
}

void bsr_matvec(const BSR& A, const double* x, double* y)
{
	for(int i = 0; i < A.n; i++) {
		double acc[4] = {0,0,0,0};
		for(int k = A.rowptr[i]; k < A.rowptr[i+1]; k++) {
			double t[4];
			gemv4(&A.vals[(size_t)k*16], &x[A.colind[k]*4], t);
			for(int v = 0; v < 4; v++) acc[v] += t[v];
		}
		for(int v = 0; v < 4; v++) y[i*4+v] = acc[v];
	}
}

void pc_setup(BlockPC& p, const BSR& A, int type)
{
	p.type = type;
	p.F = A;
	if(p.sweeps > 1) p.A = A;
	const int n = A.n;
	p.dinv.assign((size_t)n*16, 0.0);
	if(type == PC_ILU0) {
		BSR& F = p.F;
		for(int i = 0; i < n; i++) {
			for(int kk = F.rowptr[i]; kk < F.rowptr[i+1]; kk++) {
				const int k = F.colind[kk];
				if(k >= i) break;
				// L_ik = A_ik * inv(U_kk)
				double t[16];
				gemm4(&F.vals[(size_t)kk*16], &p.dinv[(size_t)k*16], t);
				std::memcpy(&F.vals[(size_t)kk*16], t, sizeof(t));
				// A_ij -= L_ik U_kj for j > k present in both rows
				for(int jj = kk+1; jj < F.rowptr[i+1]; jj++) {
					const int j = F.colind[jj];
					for(int kj = F.rowptr[k]; kj < F.rowptr[k+1]; kj++)
						if(F.colind[kj] == j) {
							double u[16];
							gemm4(t, &F.vals[(size_t)kj*16], u);
							for(int q = 0; q < 16; q++) F.vals[(size_t)jj*16+q] -= u[q];
						}
				}
			}
			invert4(&F.vals[(size_t)F.diag[i]*16], &p.dinv[(size_t)i*16]);
		}
	}
	else {
		for(int i = 0; i < n; i++)
			invert4(&A.vals[(size_t)A.diag[i]*16], &p.dinv[(size_t)i*16]);
	}
}

static void pc_apply_once(const BlockPC& p, const double* r, double* z);

/* PETSc -pc_sor_its semantics for every type: `sweeps` steps of the stationary iteration
 * z <- z + M^-1 (r - A z) from z = 0. For symmetric Gauss-Seidel this equals `sweeps` in-place
 * forward + backward sweeps. */
void pc_apply(const BlockPC& p, const double* r, double* z)
{
	pc_apply_once(p, r, z);
	if(p.sweeps <= 1 || p.type == PC_NONE) return;
	const size_t N = (size_t)p.F.n*4;
	std::vector<double> t(N), dz(N);
	for(int s = 1; s < p.sweeps; s++) {
		bsr_matvec(p.A, z, t.data());
		for(size_t i = 0; i < N; i++) t[i] = r[i] - t[i];
		pc_apply_once(p, t.data(), dz.data());
		for(size_t i = 0; i < N; i++) z[i] += dz[i];
	}
}

static void pc_apply_once(const BlockPC& p, const double* r, double* z)
{
	const BSR& F = p.F;
	const int n = F.n;
	if(p.type == PC_NONE) { std::memcpy(z, r, sizeof(double)*n*4); return; }
	if(p.type == PC_JACOBI) {
		for(int i = 0; i < n; i++) gemv4(&p.dinv[(size_t)i*16], &r[i*4], &z[i*4]);
		return;
	}
	if(p.type == PC_ILU0) {
		std::vector<double> y((size_t)n*4);
		for(int i = 0; i < n; i++) {
			double acc[4] = { r[i*4], r[i*4+1], r[i*4+2], r[i*4+3] };
			for(int kk = F.rowptr[i]; kk < F.diag[i]; kk++) {
				double t[4];
				gemv4(&F.vals[(size_t)kk*16], &y[F.colind[kk]*4], t);
				for(int v = 0; v < 4; v++) acc[v] -= t[v];
			}
			for(int v = 0; v < 4; v++) y[i*4+v] = acc[v];
		}
		for(int i = n-1; i >= 0; i--) {
			double acc[4] = { y[i*4], y[i*4+1], y[i*4+2], y[i*4+3] };
			for(int kk = F.diag[i]+1; kk < F.rowptr[i+1]; kk++) {
				double t[4];
				gemv4(&F.vals[(size_t)kk*16], &z[F.colind[kk]*4], t);
				for(int v = 0; v < 4; v++) acc[v] -= t[v];
			}
			gemv4(&p.dinv[(size_t)i*16], acc, &z[i*4]);
		}
		return;
	}
	// symmetric block Gauss-Seidel, one forward and one backward sweep from zero
	std::fill(z, z + (size_t)n*4, 0.0);
	for(int i = 0; i < n; i++) {
		double acc[4] = { r[i*4], r[i*4+1], r[i*4+2], r[i*4+3] };
		for(int kk = F.rowptr[i]; kk < F.diag[i]; kk++) {
			double t[4];
			gemv4(&F.vals[(size_t)kk*16], &z[F.colind[kk]*4], t);
			for(int v = 0; v < 4; v++) acc[v] -= t[v];
		}
		gemv4(&p.dinv[(size_t)i*16], acc, &z[i*4]);
	}
	for(int i = n-1; i >= 0; i--) {
		double acc[4] = { r[i*4], r[i*4+1], r[i*4+2], r[i*4+3] };
		for(int kk = F.rowptr[i]; kk < F.rowptr[i+1]; kk++) {
			if(kk == F.diag[i]) continue;
			double t[4];
			gemv4(&F.vals[(size_t)kk*16], &z[F.colind[kk]*4], t);
			for(int v = 0; v < 4; v++) acc[v] -= t[v];
		}
		gemv4(&p.dinv[(size_t)i*16], acc, &z[i*4]);
	}
}

/* Right-preconditioned flexible GMRES, one cycle of at most m iterations, zero initial guess,
 * stopping when ||b - A x|| <= rtol ||b||. op(x,y) computes y = A x. Returns iterations. */
template <class Op>
static int fgmres(const Op& op, const BlockPC& pc, const double* b, double* x, int N, int m,
                  double rtol)
{
	std::vector<std::vector<double>> V(m+1, std::vector<double>(N)), Z(m, std::vector<double>(N));
	std::vector<double> H((size_t)(m+1)*m, 0.0), cs(m), sn(m), gvec(m+1, 0.0);
	auto dot = [&](const double* a, const double* c) { double s = 0; for(int i = 0; i < N; i++) s += a[i]*c[i]; return s; };
	double beta = std::sqrt(dot(b,b));
	std::fill(x, x+N, 0.0);
	if(beta == 0) return 0;
	for(int i = 0; i < N; i++) V[0][i] = b[i]/beta;
	gvec[0] = beta;
	int it = 0;
	for(int j = 0; j < m; j++) {
		pc_apply(pc, V[j].data(), Z[j].data());
		op(Z[j].data(), V[j+1].data());
		for(int i = 0; i <= j; i++) {
			const double h = dot(V[j+1].data(), V[i].data());
			H[i*m+j] = h;
			for(int q = 0; q < N; q++) V[j+1][q] -= h*V[i][q];
		}
		const double hn = std::sqrt(dot(V[j+1].data(), V[j+1].data()));
		H[(j+1)*m+j] = hn;
		if(hn != 0) for(int q = 0; q < N; q++) V[j+1][q] /= hn;
		for(int i = 0; i < j; i++) {
			const double t = cs[i]*H[i*m+j] + sn[i]*H[(i+1)*m+j];
			H[(i+1)*m+j] = -sn[i]*H[i*m+j] + cs[i]*H[(i+1)*m+j];
			H[i*m+j] = t;
		}
		const double a = H[j*m+j], c = H[(j+1)*m+j];
		const double den = std::sqrt(a*a + c*c);
		cs[j] = a/den; sn[j] = c/den;
		H[j*m+j] = den; H[(j+1)*m+j] = 0;
		gvec[j+1] = -sn[j]*gvec[j];
		gvec[j] = cs[j]*gvec[j];
		it = j+1;
		if(std::fabs(gvec[j+1]) <= rtol*beta) break;
	}
	std::vector<double> y(it);
	for(int i = it-1; i >= 0; i--) {
		double s = gvec[i];
		for(int k = i+1; k < it; k++) s -= H[i*m+k]*y[k];
		y[i] = s/H[i*m+i];
	}
	for(int k = 0; k < it; k++)
		for(int q = 0; q < N; q++) x[q] += y[k]*Z[k][q];
	return it;
}

static double exp_ramp(double cmin, double cmax, double prev, double ratio, double up, double down)
{
	const double nc = ratio > 1.0 ? prev*std::pow(ratio, up) : prev*std::pow(ratio, down);
	if(nc < cmin) return cmin;
	if(nc > cmax) return cmax;
	return nc;
}

static double relax_factor(const Gas& gs, const double* du, const double* u, double minf)
{
	double unew[4];
	for(int i = 0; i < 4; i++) unew[i] = u[i] + du[i];
	double dp = 0;
	for(int i = 2; i < 4; i++)
		dp -= ((u[i]+unew[i])*(u[0]+unew[0])/2.0*du[i] - (unew[i]*unew[i]+u[i]*u[i])/2.0*du[0]);
	dp = (gs.g-1.0)*(du[3] - 1.0/(2*u[0]*unew[0])*dp);
	const double p = pressure_from_conserved(gs, u);
	const double dpr = std::fabs(dp)/p;
	const double drho = std::fabs(du[0])/u[0];
	const double danger = std::max(dpr, drho);
	double omega = minf;
	if(danger < 1.0-minf) omega = 1.0-danger;
	return omega;
}

SolveStats steady_solve(const Disc& d, const SolverConfig& sc, double* u)
{
	SolveStats st;
	const Mesh& m = *d.m;
	const Gas gs(d.cfg);
	const int ne = m.nelem, N = ne*4;
	std::vector<double> r(N), du(N), dtm(ne), aux(N), yg(N);
	BSR M;
	BlockPC pc;
	pc.sweeps = sc.pc_sweeps > 1 ? sc.pc_sweeps : 1;
	double curCFL = 0, resi = 1.0, resiold = 1.0, initres = 1.0;
	if(sc.fixed_norm > 0) initres = sc.fixed_norm;
	int step = 0;
	const auto t0 = std::chrono::steady_clock::now();
	while(resi/initres > sc.tol && step < sc.maxiter) {
		std::fill(r.begin(), r.end(), 0.0);
		compute_residual(d, u, r.data(), true, dtm.data(), 0);
		if(sc.implicit) {
			assemble_jacobian(d, u, M);
			curCFL = exp_ramp(sc.cflinit, sc.cflfin, curCFL, resiold/resi, 0.25, 0.3);
			for(int c = 0; c < ne; c++) {
				dtm[c] = m.area[c]/(curCFL*dtm[c]);
				double* D = &M.vals[(size_t)M.diag[c]*16];
				for(int v = 0; v < 4; v++) D[v*4+v] += dtm[c];
			}
			pc_setup(pc, M, sc.pc);
			int its;
			if(sc.matrix_free) {
				auto op = [&](const double* x, double* y) {
					double xn = 0;
					for(int i = 0; i < N; i++) xn += x[i]*x[i];
					xn = std::sqrt(xn);
					const double pert = sc.mf_eps/xn;
					for(int i = 0; i < N; i++) { aux[i] = u[i] + pert*x[i]; yg[i] = 0; }
					compute_residual(d, aux.data(), yg.data(), false, nullptr, 0);
					for(int c = 0; c < ne; c++)
						for(int v = 0; v < 4; v++)
							y[c*4+v] = dtm[c]*x[c*4+v] + (-yg[c*4+v] + r[c*4+v])/pert;
				};
				its = fgmres(op, pc, r.data(), du.data(), N, sc.krylov_restart, sc.krylov_rtol);
			} else {
				auto op = [&](const double* x, double* y) { bsr_matvec(M, x, y); };
				its = fgmres(op, pc, r.data(), du.data(), N, sc.krylov_restart, sc.krylov_rtol);
			}
			st.lin_iters += its;
			for(int c = 0; c < ne; c++) {
				const double om = sc.fullupdate ? 1.0 : relax_factor(gs, &du[c*4], &u[c*4], sc.min_relax);
				for(int v = 0; v < 4; v++) u[c*4+v] += om*du[c*4+v];
			}
		} else {
			curCFL = exp_ramp(sc.cflinit, sc.cflfin, curCFL, resiold/resi, 0.3, 0.25);
			for(int c = 0; c < ne; c++)
				for(int v = 0; v < 4; v++)
					u[c*4+v] += sc.cflinit*dtm[c] * 1.0/m.area[c]*r[c*4+v];
		}
		double rn = 0;
		for(int c = 0; c < ne; c++) rn += r[c*4+3]*r[c*4+3]*m.area[c];
		resiold = resi;
		resi = std::sqrt(rn);
		if(!std::isfinite(resi)) { st.status = 3; break; }
		if(step == 0 && sc.fixed_norm <= 0) initres = resi;
		step++;
	}
	st.steps = step;
	st.final_abs = resi;
	st.final_rel = resi/initres;
	// converged only below the step limit (src/ode/aodesolver.cpp:618-626)
	st.converged = (st.status == 0 && step < sc.maxiter && resi/initres <= sc.tol) ? 1 : 0;
	if(st.status == 0 && !st.converged) st.status = 4;
	st.walltime = std::chrono::duration<double>(std::chrono::steady_clock::now()-t0).count();
	return st;
}

}
