/* api.cpp -- the extern "C" boundary of libfvx (declared in include/fvx.h). Converts internal
 * exceptions to status codes and keeps the last error text per thread. */
#include "fvx_internal.hpp"
#include "wave.hpp"
#include <cmath>
#include <cstring>
#include <memory>
#include <functional>
#include <mutex>
#include <string>

namespace {
thread_local std::string g_last_error;

template <class F> fvx_status guarded(F&& f)
{
	try { f(); return FVX_OK; }
	catch(const fvx::Error& e) { g_last_error = e.what(); return e.code; }
	catch(const std::bad_alloc&) { g_last_error = "host allocation failed"; return FVX_ERR_ALLOC; }
	catch(const std::exception& e) { g_last_error = e.what(); return FVX_ERR_INVALID; }
	catch(...) { g_last_error = "unknown error"; return FVX_ERR_INVALID; }
}

void require(bool c, const char* msg)
{
	if(!c) throw fvx::Error(FVX_ERR_INVALID, msg);
}
}

namespace fvx {
/// The status conversion of this boundary for entry points defined in other source files.
fvx_status guard(const std::function<void()>& f) { return guarded(f); }
}

namespace {
/* Scratch of fvx_matfree_apply_noalloc: perturbed state and its residual (4 values per cell
 * each, FVX_NOALLOC_MAX_CELLS cells) and the norm partial sums (NOALLOC_PARTIAL values). The
 * array's name is a host-side stand-in: host code uses only the device address that
 * hipGetSymbolAddress returns for the current device, kept per device. */
constexpr long NOALLOC_PARTIAL = 512L*32 + 1;
static __device__ __align__(256) double scratch[8L*FVX_NOALLOC_MAX_CELLS + NOALLOC_PARTIAL];
constexpr int NOALLOC_MAX_DEVICES = 64;
std::mutex g_scratch_mutex;
double* g_scratch_ptr[NOALLOC_MAX_DEVICES] = {};

/// The current device's copy of `scratch`, looked up on the first call on that device.
double* noalloc_scratch()
{
	if(fvx::partial_stride() + 1 > NOALLOC_PARTIAL)
		throw fvx::Error(FVX_ERR_INVALID, "no-allocation scratch smaller than the partial sums");
	int dev = 0;
	FVX_HIP(hipGetDevice(&dev));
	if(dev < 0 || dev >= NOALLOC_MAX_DEVICES)
		throw fvx::Error(FVX_ERR_INVALID, "device index beyond the no-allocation table");
	std::lock_guard<std::mutex> lock(g_scratch_mutex);
	if(!g_scratch_ptr[dev]) {
		void* p = nullptr;
		const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
		if(e != hipSuccess || !p)
			throw fvx::Error(FVX_ERR_HIP, std::string("hipGetSymbolAddress(scratch): ")
			                              + hipGetErrorString(e));
		g_scratch_ptr[dev] = static_cast<double*>(p);
	}
	return g_scratch_ptr[dev];
}
}

extern "C" {

const char* fvx_last_error(void) { return g_last_error.c_str(); }

fvx_status fvx_mesh_rcm(const fvx_mesh_desc* desc, int* perm)
{
	return guarded([&] {
		require(desc && perm, "null argument");
		fvx::rcm_order(*desc, perm);
	});
}

fvx_status fvx_mesh_lines(const fvx_mesh_desc* desc, double threshold, int* nlines,
                          int* line_ptr, int* line_cells)
{
	return guarded([&] {
		require(desc && nlines && line_ptr && line_cells, "null argument");
		std::vector<std::vector<int>> lines;
		std::vector<int> celline;
		fvx::find_lines(*desc, threshold, lines, celline);
		*nlines = (int)lines.size();
		line_ptr[0] = 0;
		int k = 0;
		for(size_t l = 0; l < lines.size(); l++) {
			for(int c : lines[l]) line_cells[k++] = c;
			line_ptr[l+1] = k;
		}
	});
}

fvx_status fvx_mesh_line_order(const fvx_mesh_desc* desc, double threshold, int graph_ordering,
                               int* perm)
{
	return guarded([&] {
		require(desc && perm, "null argument");
		fvx::line_order(*desc, threshold, graph_ordering, perm);
	});
}

fvx_status fvx_mesh_hilbert(const fvx_mesh_desc* desc, int* perm)
{
	return guarded([&] {
		require(desc && perm, "null argument");
		fvx::hilbert_order(*desc, perm);
	});
}

fvx_status fvx_mesh_create(const fvx_mesh_desc* desc, fvx_mesh* out)
{
	return guarded([&] {
		require(desc && out, "null argument");
		std::unique_ptr<fvx_mesh_s> m(new fvx_mesh_s);
		fvx::build_mesh(*m, *desc);
		*out = m.release();
	});
}

fvx_status fvx_mesh_sizes(fvx_mesh m, int* s)
{
	return guarded([&] {
		require(m && s, "null argument");
		s[0] = m->npoin; s[1] = m->nelem; s[2] = m->nbface; s[3] = m->ninface; s[4] = m->naface;
	});
}

fvx_status fvx_mesh_faces(fvx_mesh m, int* intfac, double* facemetric, double* area,
                          double* centres)
{
	return guarded([&] {
		require(m != nullptr, "null mesh");
		if(intfac) std::memcpy(intfac, m->intfac.data(), sizeof(int)*m->intfac.size());
		if(facemetric) std::memcpy(facemetric, m->facemetric.data(), sizeof(double)*m->facemetric.size());
		if(area) std::memcpy(area, m->area.data(), sizeof(double)*m->area.size());
		if(centres) std::memcpy(centres, m->rc.data(), sizeof(double)*m->rc.size());
	});
}

fvx_status fvx_mesh_destroy(fvx_mesh m)
{
	return guarded([&] { delete m; });
}

fvx_status fvx_disc_create(fvx_mesh m, const fvx_physics_cfg* p, const fvx_numerics_cfg* n,
                           fvx_disc* out)
{
	return guarded([&] {
		require(m && p && n && out, "null argument");
		std::unique_ptr<fvx_disc_s> d(new fvx_disc_s);
		d->m = m;
		d->pc = *p;
		d->nc = *n;
		fvx::build_disc(*d);
		*out = d.release();
	});
}

fvx_status fvx_disc_destroy(fvx_disc d)
{
	return guarded([&] { delete d; });
}

fvx_status fvx_freestream_state(fvx_disc d, double* out)
{
	return guarded([&] {
		require(d && out, "null argument");
		for(int k = 0; k < 4; k++) out[k] = d->gas.uinf[k];
	});
}

fvx_status fvx_compute_residual(fvx_disc d, const double* u, double* r, int gettimesteps,
                                double* dt, void* stream)
{
	return guarded([&] {
		require(d && u && r, "null argument");
		require(!gettimesteps || dt, "time steps requested without an output array");
		fvx::launch_residual(*d, u, r, gettimesteps, dt, static_cast<hipStream_t>(stream));
	});
}

fvx_status fvx_get_gradients(fvx_disc d, const double* u, double* grads, void* stream)
{
	return guarded([&] {
		require(d && u && grads, "null argument");
		fvx::launch_gradients(*d, u, grads, nullptr, 1, static_cast<hipStream_t>(stream));
	});
}

fvx_status fvx_get_primitive_gradients(fvx_disc d, const double* u, double* grads,
                                       void* stream)
{
	return guarded([&] {
		require(d && u && grads, "null argument");
		const int rec = d->nc.reconstruction;
		const bool limited = rec == FVX_RECON_VENKATAKRISHNAN || rec == FVX_RECON_BARTHJESPERSEN;
		fvx::launch_gradients(*d, u, grads, limited ? d->d_rgrad.p : nullptr, 0,
		                      static_cast<hipStream_t>(stream));
	});
}

/* ------------------------------------------------------------- implicit solve ---------- */

fvx_status fvx_matrix_create(fvx_mesh m, int precision, fvx_matrix* out)
{
	return guarded([&] {
		require(m && out, "null argument");
		std::unique_ptr<fvx_matrix_s> A(new fvx_matrix_s);
		fvx::matrix_init(*A, m, precision);
		*out = A.release();
	});
}

fvx_status fvx_matrix_destroy(fvx_matrix A)
{
	return guarded([&] { delete A; });
}

fvx_status fvx_assemble_jacobian(fvx_disc d, const double* u, fvx_matrix A, void* stream)
{
	return guarded([&] {
		require(d && u && A, "null argument");
		require(A->m == d->m, "matrix and discretisation belong to different meshes");
		if(!fvx::jacobian_supported(*d))
			throw fvx::Error(FVX_ERR_UNSUPPORTED, "the reference has no Jacobian for this flux or BC");
		fvx::launch_jacobian(*d, u, A->raw(), A->prec, nullptr, 0.0, nullptr,
		                     static_cast<hipStream_t>(stream));
	});
}

fvx_status fvx_add_pseudotime(fvx_disc d, fvx_matrix A, const double* dt, double cfl,
                              double* dtm, void* stream)
{
	return guarded([&] {
		require(d && A && dt && dtm, "null argument");
		require(A->m == d->m, "matrix and discretisation belong to different meshes");
		require(cfl > 0, "CFL number must be positive");
		fvx::launch_pseudotime(*A, dt, cfl, dtm, static_cast<hipStream_t>(stream));
	});
}

fvx_status fvx_matrix_apply(fvx_matrix A, const double* x, double* y, void* stream)
{
	return guarded([&] {
		require(A && x && y, "null argument");
		const long N = 4L*A->m->nelem;
		fvx::launch_spmv(*A, x, N, y, N, fvx::Inst(), static_cast<hipStream_t>(stream));
	});
}

fvx_status fvx_matrix_bsr_nnzb(fvx_matrix A, int* nnzb)
{
	return guarded([&] {
		require(A && nnzb, "null argument");
		*nnzb = fvx::matrix_bsr_nnzb(*A->m);
	});
}

fvx_status fvx_matrix_export_bsr(fvx_matrix A, int* rowptr, int* colind, double* vals)
{
	return guarded([&] {
		require(A != nullptr, "null handle");
		fvx::matrix_export_bsr(*A, rowptr, colind, vals);
	});
}

fvx_status fvx_pc_create(fvx_mesh m, int type, int precision, fvx_pc* out)
{
	return guarded([&] {
		require(m && out, "null argument");
		std::unique_ptr<fvx_pc_s> p(new fvx_pc_s);
		fvx::pc_init(*p, m, type, precision);
		*out = p.release();
	});
}

fvx_status fvx_pc_create_lines(fvx_mesh m, double threshold, int precision, fvx_pc* out)
{
	return guarded([&] {
		require(m && out, "null argument");
		require(threshold > 0, "line threshold must be positive");
		std::unique_ptr<fvx_pc_s> p(new fvx_pc_s);
		p->line_threshold = threshold;
		fvx::pc_init(*p, m, FVX_PC_LINE, precision);
		*out = p.release();
	});
}

fvx_status fvx_pc_destroy(fvx_pc p)
{
	return guarded([&] { delete p; });
}

fvx_status fvx_pc_set_sweeps(fvx_pc p, int sweeps)
{
	return guarded([&] {
		require(p != nullptr, "null handle");
		fvx::pc_set_sweeps(*p, sweeps);
	});
}

fvx_status fvx_pc_setup(fvx_pc p, fvx_matrix A, void* stream)
{
	return guarded([&] {
		require(p && A, "null argument");
		require(p->m == A->m, "preconditioner and matrix belong to different meshes");
		require(A->ninst >= p->ninst, "matrix holds fewer instances than the preconditioner");
		fvx::launch_pc_setup(*p, *A, fvx::Inst(), static_cast<hipStream_t>(stream));
	});
}

fvx_status fvx_pc_apply(fvx_pc p, const double* r, double* z, void* stream)
{
	return guarded([&] {
		require(p && r && z, "null argument");
		const long N = 4L*p->m->nelem;
		fvx::launch_pc_apply(*p, r, N, z, N, fvx::Inst(), static_cast<hipStream_t>(stream));
	});
}

fvx_status fvx_pc_ordering(fvx_pc p, int* colour, int* ncolours)
{
	return guarded([&] {
		require(p != nullptr, "null handle");
		if(colour) std::memcpy(colour, p->colour.data(), sizeof(int)*p->colour.size());
		if(ncolours) *ncolours = p->ncol;
	});
}

fvx_status fvx_pc_set_subdomain_cells(fvx_pc p, int cells)
{
	return guarded([&] {
		require(p != nullptr, "null handle");
		fvx::pc_set_subdomain_cells(*p, cells);
	});
}

fvx_status fvx_pc_subdomains(fvx_pc p, int* part, int* nparts)
{
	return guarded([&] {
		require(p != nullptr, "null handle");
		require(p->type == FVX_PC_BJACOBI_ILU0, "the preconditioner has no subdomains");
		if(part) std::memcpy(part, p->part.data(), sizeof(int)*p->part.size());
		if(nparts) *nparts = p->nsub;
	});
}

fvx_status fvx_linear_solve(fvx_matrix A, fvx_pc p, const double* b, double* x, int restart,
                            double rtol, int* iters, void* stream)
{
	return guarded([&] {
		require(A && p && b && x, "null argument");
		require(p->m == A->m, "preconditioner and matrix belong to different meshes");
		const int it = fvx::linear_solve(*A, *p, b, x, restart, rtol,
		                                 static_cast<hipStream_t>(stream));
		if(iters) *iters = it;
	});
}

fvx_status fvx_matfree_apply(fvx_disc d, const double* u, const double* r_u, const double* dtm,
                             const double* x, double* y, double eps, void* stream)
{
	return guarded([&] {
		require(d && u && r_u && dtm && x && y, "null argument");
		require(eps > 0, "difference step must be positive");
		hipStream_t s = static_cast<hipStream_t>(stream);
		const size_t N = 4*(size_t)d->m->nelem;
		// handle-owned scratch, sized once; later calls reuse it in stream order
		if(d->mf_aux.n != N) {
			d->mf_aux.alloc(N);
			d->mf_raux.alloc(N);
			d->mf_part.alloc((size_t)fvx::partial_stride() + 1);
		}
		fvx::matfree_apply(*d, u, r_u, dtm, x, y, eps, d->mf_aux.p, d->mf_raux.p, d->mf_part.p,
		                   d->mf_part.p + (size_t)fvx::partial_stride(), s);
	});
}

fvx_status fvx_matfree_apply_noalloc(fvx_disc d, const double* u, const double* r_u,
                                     const double* dtm, const double* x, double* y, double eps,
                                     void* stream)
{
	return guarded([&] {
		require(d && u && r_u && dtm && x && y, "null argument");
		require(eps > 0, "difference step must be positive");
		require(d->m->nelem <= FVX_NOALLOC_MAX_CELLS,
		        "mesh larger than the no-allocation scratch (FVX_NOALLOC_MAX_CELLS cells)");
		double* ptr = noalloc_scratch();
		double* aux = ptr;
		double* raux = ptr + 4L*FVX_NOALLOC_MAX_CELLS;
		double* part = ptr + 8L*FVX_NOALLOC_MAX_CELLS;
		// matfree_apply writes aux and the partial sums in full and clears raux itself
		fvx::matfree_apply(*d, u, r_u, dtm, x, y, eps, aux, raux, part,
		                   part + fvx::partial_stride(), static_cast<hipStream_t>(stream));
	});
}

fvx_status fvx_steady_solve(fvx_disc d, const fvx_solver_cfg* cfg, double* u,
                            fvx_solve_stats* stats, void* stream)
{
	fvx_solve_stats local;
	fvx_solve_stats* st = stats ? stats : &local;
	return guarded([&] {
		require(d && cfg && u, "null argument");
		fvx::steady_solve(*d, *cfg, u, *st, static_cast<hipStream_t>(stream));
	});
}

/* ------------------------------------------------------------- batched solve ----------- */

fvx_status fvx_batch_create(fvx_disc d, int ninst, const double* mach, const double* aoa,
                            fvx_batch* out)
{
	return guarded([&] {
		require(d && mach && aoa && out, "null argument");
		require(ninst >= 1 && ninst <= 4096, "instance count must be in [1, 4096]");
		std::unique_ptr<fvx_batch_s> b(new fvx_batch_s);
		b->d = d;
		b->ninst = ninst;
		std::vector<int> all(ninst);
		for(int i = 0; i < ninst; i++) {
			require(mach[i] > 0 && std::isfinite(mach[i]) && std::isfinite(aoa[i]),
			        "invalid Mach number or angle of attack");
			fvx_physics_cfg p = d->pc;
			p.Minf = mach[i];
			p.aoa = aoa[i];
			b->gas.push_back(fvx::make_gas(p));
			b->aoa.push_back(aoa[i]);
			all[i] = i;
		}
		b->d_gas.upload(b->gas);
		b->d_all.upload(all);
		const size_t n = (size_t)d->m->nelem;
		b->d_grad.alloc(8*n*ninst);
		b->d_rgrad.alloc(8*n*ninst);
		b->d_ughost.alloc(4*(size_t)std::max(d->m->nbface, 1)*ninst);
		b->d_scratch.alloc((size_t)fvx::functional_blocks()*5);
		*out = b.release();
	});
}

fvx_status fvx_batch_destroy(fvx_batch b)
{
	return guarded([&] { delete b; });
}

fvx_status fvx_batch_freestream(fvx_batch b, double* out)
{
	return guarded([&] {
		require(b && out, "null argument");
		for(int i = 0; i < b->ninst; i++)
			for(int k = 0; k < 4; k++) out[4*i + k] = b->gas[i].uinf[k];
	});
}

fvx_status fvx_batch_compute_residual(fvx_batch b, const double* u, double* r, int gettimesteps,
                                      double* dt, void* stream)
{
	return guarded([&] {
		require(b && u && r, "null argument");
		require(!gettimesteps || dt, "time steps requested without an output array");
		fvx::BatchArg ba;
		ba.list = b->d_all.p;
		ba.gp = b->d_gas.p;
		fvx::launch_residual(*b->d, u, r, gettimesteps, dt, static_cast<hipStream_t>(stream), ba,
		                     b->ninst, b->d_grad.p, b->d_rgrad.p, b->d_ughost.p);
	});
}

fvx_status fvx_batch_steady_solve(fvx_batch b, const fvx_solver_cfg* cfg, double* u,
                                  fvx_solve_stats* stats, void* stream)
{
	return guarded([&] {
		require(b && cfg && u && stats, "null argument");
		std::vector<fvx_solve_stats> v;
		fvx::steady_core(*b->d, *cfg, u, b->ninst, b->d_gas.p, v,
		                 static_cast<hipStream_t>(stream),
		                 b->norms.empty() ? nullptr : b->norms.data());
		for(int i = 0; i < b->ninst; i++) stats[i] = v[i];
	});
}

fvx_status fvx_solve_history(fvx_disc d, fvx_step_record* out, int cap, int* n)
{
	return guarded([&] {
		require(d && n && (out || cap == 0), "null argument");
		require(cap >= 0, "negative capacity");
		fvx::solve_history(*d, out, cap, n);
	});
}

fvx_status fvx_batch_set_norms(fvx_batch b, const double* norms)
{
	return guarded([&] {
		require(b != nullptr, "null handle");
		if(!norms) { b->norms.clear(); return; }
		for(int i = 0; i < b->ninst; i++)
			require(norms[i] > 0 && std::isfinite(norms[i]), "normalisers must be positive");
		b->norms.assign(norms, norms + b->ninst);
	});
}

fvx_status fvx_batch_residual_norms(fvx_batch b, const double* u, double* out, void* stream)
{
	return guarded([&] {
		require(b && u && out, "null argument");
		hipStream_t s = static_cast<hipStream_t>(stream);
		fvx_disc_s& d = *b->d;
		const int n = d.m->nelem;
		// batch-owned scratch, allocated on the first call
		if(!b->n_host) {
			b->n_r.alloc((size_t)4*n*b->ninst);
			b->n_part.alloc((size_t)fvx::partial_stride()*b->ninst);
			b->n_out.alloc(b->ninst);
			FVX_HIP(hipHostMalloc((void**)&b->n_host, b->ninst*sizeof(double)));
		}
		fvx::Inst L;
		L.list = b->d_all.p;
		L.nact = b->ninst;
		fvx::launch_zero(4L*n, b->n_r.p, 4L*n, L, s);
		fvx::BatchArg ba;
		ba.list = b->d_all.p;
		ba.gp = b->d_gas.p;
		fvx::launch_residual(d, u, b->n_r.p, 0, nullptr, s, ba, b->ninst, b->d_grad.p,
		                     b->d_rgrad.p, b->d_ughost.p);
		fvx::launch_resnorm(n, b->n_r.p, d.m->d_area.p, b->n_part.p, b->n_out.p, 1, L, s);
		FVX_HIP(hipMemcpyAsync(b->n_host, b->n_out.p, b->ninst*sizeof(double),
		                       hipMemcpyDeviceToHost, s));
		// the norms are returned in host memory, so this call waits for them
		FVX_HIP(hipStreamSynchronize(s));
		for(int i = 0; i < b->ninst; i++) out[i] = std::sqrt(b->n_host[i]);
	});
}

fvx_status fvx_batch_functionals(fvx_batch b, const double* u, int wall_tag, double* out,
                                 void* stream)
{
	return guarded([&] {
		require(b && u && out, "null argument");
		hipStream_t s = static_cast<hipStream_t>(stream);
		fvx_disc_s& d = *b->d;
		const size_t n = (size_t)d.m->nelem;
		const int nb = fvx::functional_blocks();
		std::vector<double> part((size_t)nb*5);
		for(int i = 0; i < b->ninst; i++) {
			fvx::BatchArg ba;
			ba.list = b->d_all.p + i;
			ba.gp = b->d_gas.p;
			fvx::launch_gradients(d, u, b->d_grad.p, nullptr, 1, s, ba, 1, b->d_ughost.p);
			fvx::launch_functionals(d, u + 4*n*i, b->d_grad.p + 8*n*i, wall_tag, b->d_scratch.p, s,
			                        &b->gas[i], b->aoa[i]);
			FVX_HIP(hipMemcpyAsync(part.data(), b->d_scratch.p, sizeof(double)*part.size(),
			                       hipMemcpyDeviceToHost, s));
			FVX_HIP(hipStreamSynchronize(s));
			double acc[5] = {0, 0, 0, 0, 0};
			for(int k = 0; k < nb; k++)
				for(int j = 0; j < 5; j++) acc[j] += part[(size_t)k*5 + j];
			require(acc[3] > 0, "no boundary face carries the requested wall marker");
			out[4*i + 0] = acc[0]/acc[3];
			out[4*i + 1] = acc[1]/acc[3];
			out[4*i + 2] = acc[2]/acc[3];
			out[4*i + 3] = std::sqrt(acc[4]);
		}
	});
}

/* ------------------------------------------------------------------ autotuning -------- */

fvx_status fvx_autotune(fvx_disc d, const double* u, int reps, fvx_tune_row* rows, int cap,
                        int* nrows, void* stream)
{
	return guarded([&] {
		require(d && u, "null argument");
		require(reps >= 1, "at least one timed repetition");
		std::vector<fvx_tune_row> v;
		fvx::autotune_launch(*d, u, reps, v, static_cast<hipStream_t>(stream));
		if(nrows) *nrows = (int)v.size();
		if(rows) for(int i = 0; i < (int)v.size() && i < cap; i++) rows[i] = v[i];
	});
}

fvx_status fvx_set_launch(fvx_disc d, int kernel, int block, int xcd)
{
	return guarded([&] {
		require(d != nullptr, "null handle");
		require(block == 64 || block == 128 || block == 256 || block == 512,
		        "block must be 64, 128, 256 or 512");
		require(xcd == 0 || xcd == 1, "xcd must be 0 or 1");
		if(kernel == FVX_KERNEL_GRADIENTS) { d->tune.grad_block = block; d->tune.grad_xcd = xcd; }
		else if(kernel == FVX_KERNEL_FLUX) {
			d->tune.flux_block = block; d->tune.flux_xcd = xcd; d->tune.flux_wave = 0;
		}
		else if(kernel == FVX_KERNEL_FLUX_WAVE) {
			require(block == 64, "the wave-task flux kernel runs one wave64 per workgroup");
			require(fvx::flux_wave_apply(*d), "the wave-task flux form applies to inviscid "
			        "first-order or linear reconstruction only");
			d->tune.flux_xcd = xcd; d->tune.flux_wave = 1;
		}
		else if(kernel == FVX_KERNEL_JACOBIAN) {
			require(block <= 256, "the Jacobian kernel takes at most 256 threads per workgroup");
			d->tune.jac_block = block; d->tune.jac_xcd = xcd;
		}
		else require(false, "kernel has no tunable launch shape");
	});
}

fvx_status fvx_get_launch(fvx_disc d, int kernel, int* block, int* xcd)
{
	return guarded([&] {
		require(d && block && xcd, "null argument");
		if(kernel == FVX_KERNEL_GRADIENTS) { *block = d->tune.grad_block; *xcd = d->tune.grad_xcd; }
		else if(kernel == FVX_KERNEL_FLUX) { *block = d->tune.flux_block; *xcd = d->tune.flux_xcd; }
		else if(kernel == FVX_KERNEL_FLUX_WAVE) {
			*block = d->tune.flux_wave ? 64 : 0; *xcd = d->tune.flux_xcd;
		}
		else if(kernel == FVX_KERNEL_JACOBIAN) { *block = d->tune.jac_block; *xcd = d->tune.jac_xcd; }
		else require(false, "kernel has no tunable launch shape");
	});
}

fvx_status fvx_autotune_pc(fvx_disc d, const double* u, double cfl, int restart, double rtol,
                           int precision, fvx_tune_row* rows, int cap, int* nrows, int* best_pc,
                           int* best_sweeps, void* stream)
{
	return guarded([&] {
		require(d && u && best_pc && best_sweeps, "null argument");
		require(cfl > 0 && rtol > 0, "CFL number and tolerance must be positive");
		std::vector<fvx_tune_row> v;
		fvx::autotune_pc(*d, u, cfl, restart, rtol, precision, v, *best_pc, *best_sweeps,
		                 static_cast<hipStream_t>(stream));
		if(nrows) *nrows = (int)v.size();
		if(rows) for(int i = 0; i < (int)v.size() && i < cap; i++) rows[i] = v[i];
	});
}

fvx_status fvx_profile_enable(fvx_disc d, int on)
{
	return guarded([&] {
		require(d != nullptr, "null handle");
		if(on) { d->prof.collect(); d->prof.reset(); }
		d->prof.on = on != 0;
	});
}

fvx_status fvx_profile_read(fvx_disc d, int kernel, double* total_ms, long* count)
{
	return guarded([&] {
		require(d && total_ms && count, "null argument");
		require(kernel >= 0 && kernel < fvx::Profiler::NKERNEL, "kernel id out of range");
		d->prof.collect();
		*total_ms = d->prof.total_ms[kernel];
		*count = d->prof.count[kernel];
	});
}

fvx_status fvx_functionals(fvx_disc d, const double* u, int wall_tag, double* out, void* stream)
{
	return guarded([&] {
		require(d && u && out, "null argument");
		hipStream_t s = static_cast<hipStream_t>(stream);
		fvx::launch_gradients(*d, u, d->d_grad.p, nullptr, 1, s);
		const int nb = fvx::functional_blocks();
		fvx::launch_functionals(*d, u, d->d_grad.p, wall_tag, d->d_scratch.p, s);
		std::vector<double> part((size_t)nb*5);
		FVX_HIP(hipMemcpyAsync(part.data(), d->d_scratch.p, sizeof(double)*part.size(),
		                       hipMemcpyDeviceToHost, s));
		FVX_HIP(hipStreamSynchronize(s));
		double acc[5] = {0, 0, 0, 0, 0};
		for(int b = 0; b < nb; b++)
			for(int k = 0; k < 5; k++) acc[k] += part[(size_t)b*5+k];
		require(acc[3] > 0, "no boundary face carries the requested wall marker");
		out[0] = acc[0]/acc[3];
		out[1] = acc[1]/acc[3];
		out[2] = acc[2]/acc[3];
		out[3] = std::sqrt(acc[4]);
	});
}

fvx_status fvx_surface_data(fvx_disc d, const double* u, int wall_tag, double* out, int cap,
                            int* nfaces, void* stream)
{
	return guarded([&] {
		require(d && u && nfaces && (out || cap == 0), "null argument");
		const fvx_mesh_s& m = *d->m;
		std::vector<int> list;
		for(int b = 0; b < m.nbface; b++) if(m.btag[b] == wall_tag) list.push_back(b);
		*nfaces = (int)list.size();
		if(!out || list.empty()) return;
		require(cap >= (int)list.size(), "output holds fewer rows than the marker has faces");
		hipStream_t s = static_cast<hipStream_t>(stream);
		fvx::DBuf<int> dl;
		fvx::DBuf<double> dout;
		dl.upload(list);
		dout.alloc(4*list.size());
		fvx::launch_gradients(*d, u, d->d_grad.p, nullptr, 1, s);
		fvx::launch_surface_data(*d, u, d->d_grad.p, dl.p, (int)list.size(), dout.p, s);
		FVX_HIP(hipMemcpyAsync(out, dout.p, 4*list.size()*sizeof(double), hipMemcpyDeviceToHost, s));
		FVX_HIP(hipStreamSynchronize(s));
	});
}

// ---- learned warm start ----

int fvx_mpnn_nparams(void) { return fvx::mpnn::NPARAM; }

fvx_status fvx_mpnn_create(fvx_mesh m, int wall_tag, int max_graphs, fvx_mpnn* out)
{
	return guarded([&] {
		require(m && out, "null argument");
		require(max_graphs >= 1 && (long long)max_graphs*m->nelem < (1ll << 31) - 16,
		        "max_graphs must be at least 1 and keep graphs x cells below 2^31");
		std::unique_ptr<fvx_mpnn_s> n(new fvx_mpnn_s);
		n->net.init(*m, wall_tag, max_graphs);
		*out = n.release();
	});
}

fvx_status fvx_mpnn_destroy(fvx_mpnn n)
{
	return guarded([&] { delete n; });
}

fvx_status fvx_mpnn_features(fvx_mpnn n, float* stat, float* edge, int* nbr, int* rev,
                             float* invdeg)
{
	return guarded([&] {
		require(n != nullptr, "null handle");
		const auto& g = n->net.host;
		if(stat) std::memcpy(stat, g.stat.data(), g.stat.size()*sizeof(float));
		if(edge) std::memcpy(edge, g.edge.data(), g.edge.size()*sizeof(float));
		if(nbr) std::memcpy(nbr, g.nbr.data(), g.nbr.size()*sizeof(int));
		if(rev) std::memcpy(rev, g.rev.data(), g.rev.size()*sizeof(int));
		if(invdeg) std::memcpy(invdeg, g.invdeg.data(), g.invdeg.size()*sizeof(float));
	});
}

fvx_status fvx_mpnn_set_params(fvx_mpnn n, const float* w)
{
	return guarded([&] {
		require(n && w, "null argument");
		fvx::Mpnn& net = n->net;
		const size_t bytes = fvx::mpnn::NPARAM*sizeof(float);
		FVX_HIP(hipDeviceSynchronize());
		FVX_HIP(hipMemcpy(net.w.p, w, bytes, hipMemcpyHostToDevice));
		FVX_HIP(hipMemset(net.am.p, 0, bytes));
		FVX_HIP(hipMemset(net.av.p, 0, bytes));
		net.step = 0;
		net.dirty = true;
	});
}

fvx_status fvx_mpnn_get_params(fvx_mpnn n, float* w)
{
	return guarded([&] {
		require(n && w, "null argument");
		FVX_HIP(hipDeviceSynchronize());
		FVX_HIP(hipMemcpy(w, n->net.w.p, fvx::mpnn::NPARAM*sizeof(float), hipMemcpyDeviceToHost));
	});
}

fvx_status fvx_mpnn_set_points(fvx_mpnn n, int ngraph, const double* mach, const double* aoa,
                               double gamma, void* stream)
{
	return guarded([&] {
		require(n && mach && aoa, "null argument");
		require(gamma > 1, "gamma must exceed 1");
		n->net.set_points(ngraph, mach, aoa, gamma, static_cast<hipStream_t>(stream));
	});
}

namespace {
void require_points(const fvx_mpnn n, int ngraph)
{
	require(ngraph >= 1 && ngraph <= n->net.npoints,
	        "ngraph must be between 1 and the number of points set");
}
}

fvx_status fvx_mpnn_predict(fvx_mpnn n, int ngraph, float* y, void* stream)
{
	return guarded([&] {
		require(n && y, "null argument");
		require_points(n, ngraph);
		n->net.predict(ngraph, y, static_cast<hipStream_t>(stream));
	});
}

fvx_status fvx_mpnn_warm_state(fvx_mpnn n, int ngraph, double floor, double* u, void* stream)
{
	return guarded([&] {
		require(n && u, "null argument");
		require_points(n, ngraph);
		require(floor > 0 && floor <= 1, "floor must be in (0, 1]");
		n->net.warm(ngraph, floor, u, static_cast<hipStream_t>(stream));
	});
}

fvx_status fvx_mpnn_targets(fvx_mpnn n, int ngraph, const double* u, float* t, void* stream)
{
	return guarded([&] {
		require(n && u && t, "null argument");
		require_points(n, ngraph);
		n->net.targets(ngraph, u, t, static_cast<hipStream_t>(stream));
	});
}

namespace {
void mpnn_grad(fvx_mpnn n, int ngraph, const float* targets, const int* ids, hipStream_t s)
{
	require(n && targets && ids, "null argument");
	require_points(n, ngraph);
	n->net.gradient(ngraph, targets, ids, s);
}

/// The device holds the summed squared error; the loss is its mean over cells and outputs.
void read_loss(fvx_mpnn n, int ngraph, float* loss, hipStream_t s)
{
	if(!loss) return;
	float sum = 0;
	FVX_HIP(hipMemcpyAsync(&sum, n->net.loss.p, sizeof(float), hipMemcpyDeviceToHost, s));
	FVX_HIP(hipStreamSynchronize(s));
	*loss = (float)(sum/((double)ngraph*n->net.N*fvx::mpnn::FO));
}
}

fvx_status fvx_mpnn_gradient(fvx_mpnn n, int ngraph, const float* targets, const int* ids,
                             float* grad, float* loss, void* stream)
{
	return guarded([&] {
		hipStream_t s = static_cast<hipStream_t>(stream);
		mpnn_grad(n, ngraph, targets, ids, s);
		if(grad)
			FVX_HIP(hipMemcpyAsync(grad, n->net.grad.p, fvx::mpnn::NPARAM*sizeof(float),
			                       hipMemcpyDeviceToDevice, s));
		read_loss(n, ngraph, loss, s);
	});
}

fvx_status fvx_mpnn_train_step(fvx_mpnn n, int ngraph, const float* targets, const int* ids,
                               double lr, float* loss, void* stream)
{
	return guarded([&] {
		require(lr > 0 && std::isfinite(lr), "learning rate must be positive");
		hipStream_t s = static_cast<hipStream_t>(stream);
		mpnn_grad(n, ngraph, targets, ids, s);
		n->net.adam_step(lr, s);
		read_loss(n, ngraph, loss, s);
	});
}

}
