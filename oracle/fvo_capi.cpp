/* fvo_capi.cpp -- C entry points of the CPU oracle, loaded by the tests through ctypes
 * (TEST INFRASTRUCTURE ONLY). Every function returns 0 on success, non-zero on error with a
 * message available from fvo_last_error(). */
#include "fvo.hpp"
#include "fvo_physics.hpp"
#include <cstring>
#include <string>

using namespace fvo;

static thread_local std::string g_err;

#define FVO_TRY(body) try { body; return 0; } \
	catch(const std::exception& e) { g_err = e.what(); return 1; }

extern "C" {

const char* fvo_last_error(void) { return g_err.c_str(); }

int fvo_mesh_new(int npoin, int nelem, int nbface, const double* coords, const int* inpoel,
                 const int* nnode, const int* bface, const int* btag, void** out)
{
	FVO_TRY({
		Mesh* m = new Mesh;
		m->npoin = npoin; m->nelem = nelem; m->nbface = nbface;
		m->coords.assign(coords, coords + (size_t)npoin*2);
		m->inpoel.assign(inpoel, inpoel + (size_t)nelem*4);
		m->nnode.assign(nnode, nnode + nelem);
		m->bface.assign(bface, bface + (size_t)nbface*2);
		m->btag.assign(btag, btag + nbface);
		try { mesh_build(*m); } catch(...) { delete m; throw; }
		*out = m;
	})
}

void fvo_mesh_free(void* m) { delete static_cast<Mesh*>(m); }

int fvo_mesh_sizes(void* mp, int* out)
{
	const Mesh* m = static_cast<Mesh*>(mp);
	out[0] = m->npoin; out[1] = m->nelem; out[2] = m->nbface; out[3] = m->ninface;
	out[4] = m->naface;
	return 0;
}

int fvo_mesh_arrays(void* mp, int* intfac, double* facemetric, double* area, double* rc,
                    double* rcbp, double* gr, int* esuel, int* elemface, int* bface)
{
	const Mesh* m = static_cast<Mesh*>(mp);
	std::memcpy(intfac, m->intfac.data(), sizeof(int)*m->intfac.size());
	std::memcpy(facemetric, m->facemetric.data(), sizeof(double)*m->facemetric.size());
	std::memcpy(area, m->area.data(), sizeof(double)*m->area.size());
	std::memcpy(rc, m->rc.data(), sizeof(double)*m->rc.size());
	std::memcpy(rcbp, m->rcbp.data(), sizeof(double)*m->rcbp.size());
	std::memcpy(gr, m->gr.data(), sizeof(double)*m->gr.size());
	std::memcpy(esuel, m->esuel.data(), sizeof(int)*m->esuel.size());
	std::memcpy(elemface, m->elemface.data(), sizeof(int)*m->elemface.size());
	std::memcpy(bface, m->bface.data(), sizeof(int)*m->bface.size());
	return 0;
}

int fvo_level_schedule(void* mp, int* levels, int maxlen, int* count)
{
	FVO_TRY({
		const std::vector<int> l = level_schedule(*static_cast<Mesh*>(mp));
		*count = (int)l.size();
		for(int i = 0; i < (int)l.size() && i < maxlen; i++) levels[i] = l[i];
	})
}

int fvo_disc_new(void* mp, const Config* cfg, void** out)
{
	FVO_TRY({
		Disc* d = new Disc;
		try { disc_init(*d, static_cast<Mesh*>(mp), *cfg); } catch(...) { delete d; throw; }
		*out = d;
	})
}

void fvo_disc_free(void* d) { delete static_cast<Disc*>(d); }

int fvo_freestream(const Config* cfg, double* out)
{
	const auto u = freestream_state(*cfg);
	for(int i = 0; i < 4; i++) out[i] = u[i];
	return 0;
}

int fvo_residual(void* d, const double* u, double* r, int gettimesteps, double* dt, int parallel)
{
	FVO_TRY(compute_residual(*static_cast<Disc*>(d), u, r, gettimesteps != 0, dt, parallel))
}

int fvo_flux_magnitude(void* d, const double* u, double* s)
{
	FVO_TRY(flux_magnitude(*static_cast<Disc*>(d), u, s))
}

int fvo_prim_gradients(void* d, const double* u, double* grads, int parallel)
{
	FVO_TRY(compute_primitive_gradients(*static_cast<Disc*>(d), u, grads, parallel))
}

int fvo_cons_gradients(void* d, const double* u, double* grads)
{
	FVO_TRY(compute_conserved_gradients(*static_cast<Disc*>(d), u, grads))
}

int fvo_flux(const Config* cfg, int fluxid, const double* ul, const double* ur, const double* n,
             double* f)
{
	FVO_TRY(inviscid_flux(*cfg, fluxid, ul, ur, n, f))
}

int fvo_flux_jacobian(const Config* cfg, int fluxid, const double* ul, const double* ur,
                      const double* n, double* dfdl, double* dfdr)
{
	FVO_TRY(inviscid_flux_jacobian(*cfg, fluxid, ul, ur, n, dfdl, dfdr))
}

int fvo_ghost(void* dp, int ib, const double* uin, double* ug)
{
	FVO_TRY({
		const Disc* d = static_cast<Disc*>(dp);
		ghost_state(*d, ib, uin, &d->m->facemetric[ib*3], ug);
	})
}

int fvo_ghost_jacobian(void* dp, int ib, const double* uin, double* ug, double* dug)
{
	FVO_TRY({
		const Disc* d = static_cast<Disc*>(dp);
		ghost_state_jacobian(*d, ib, uin, &d->m->facemetric[ib*3], ug, dug);
	})
}

int fvo_local_jacobians(void* dp, int iface, const double* ul, const double* ur, double* L,
                        double* U)
{
	FVO_TRY({
		const Disc* d = static_cast<Disc*>(dp);
		if(iface < d->m->nbface) compute_local_jacobian_boundary(*d, iface, ul, L);
		else compute_local_jacobian_interior(*d, iface, ul, ur, L, U);
	})
}

int fvo_jacobian_nnzb(void* dp, int* nnzb)
{
	FVO_TRY({
		BSR A;
		bsr_pattern(*static_cast<Disc*>(dp)->m, A);
		*nnzb = A.rowptr[A.n];
	})
}

int fvo_assemble_jacobian(void* dp, const double* u, int* rowptr, int* colind, double* vals)
{
	FVO_TRY({
		BSR A;
		assemble_jacobian(*static_cast<Disc*>(dp), u, A);
		std::memcpy(rowptr, A.rowptr.data(), sizeof(int)*A.rowptr.size());
		std::memcpy(colind, A.colind.data(), sizeof(int)*A.colind.size());
		std::memcpy(vals, A.vals.data(), sizeof(double)*A.vals.size());
	})
}

int fvo_functionals(void* dp, const double* u, int walltag, double* out)
{
	FVO_TRY({
		const Disc* d = static_cast<Disc*>(dp);
		surface_functionals(*d, u, walltag, &out[0], &out[1], &out[2]);
		out[3] = entropy_error(*d, u);
	})
}

int fvo_steady_solve(void* dp, const SolverConfig* sc, double* u, double* stats)
{
	FVO_TRY({
		const SolveStats s = steady_solve(*static_cast<Disc*>(dp), *sc, u);
		stats[0] = s.steps; stats[1] = s.lin_iters; stats[2] = s.converged; stats[3] = s.status;
		stats[4] = s.final_rel; stats[5] = s.final_abs; stats[6] = s.walltime;
	})
}

int fvo_pc_apply_bsr(int n, const int* rowptr, const int* colind, const double* vals, int pctype,
                     int sweeps, const double* r, double* z)
{
	FVO_TRY({
		BSR A;
		A.n = n;
		A.rowptr.assign(rowptr, rowptr+n+1);
		A.colind.assign(colind, colind+rowptr[n]);
		A.vals.assign(vals, vals+(size_t)rowptr[n]*16);
		A.diag.assign(n, -1);
		for(int i = 0; i < n; i++)
			for(int k = rowptr[i]; k < rowptr[i+1]; k++) if(colind[k] == i) A.diag[i] = k;
		BlockPC p;
		p.sweeps = sweeps > 1 ? sweeps : 1;
		pc_setup(p, A, pctype);
		pc_apply(p, r, z);
	})
}

}

namespace fvo {
void scalar_face_values(const Disc& d, const double* u, const double* ug, double* grads,
                        double* ufl, double* ufr);
}

extern "C" int fvo_scalar_face_values(void* dp, const double* u, const double* ug, double* grads,
                                      double* ufl, double* ufr)
{
	FVO_TRY(fvo::scalar_face_values(*static_cast<fvo::Disc*>(dp), u, ug, grads, ufl, ufr))
}

namespace fvo {
void reconstructed_face_states(const Disc& d, const double* u, double* ufl, double* ufr);
}

extern "C" int fvo_face_states(void* dp, const double* u, double* ufl, double* ufr)
{
	FVO_TRY(fvo::reconstructed_face_states(*static_cast<fvo::Disc*>(dp), u, ufl, ufr))
}
