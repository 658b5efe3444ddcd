/* fvo_spatial.cpp -- oracle spatial discretisation (TEST INFRASTRUCTURE ONLY).
 *
 * Restates, pass by pass, FlowFV<freal,order2,constVisc>::compute_residual
 * (src/spatial/flow_spatial.cpp:636-816) with the same intermediate arrays (left/right face
 * states, primitive cell states, ghost states) and the same face-loop structure with atomic
 * scatters, so that its OpenMP mode is also the CPU baseline.
 *
 *  gradients:       Green-Gauss src/spatial/agradientschemes.cpp:61-214,
 *                   weighted least squares :218-440
 *  reconstruction:  unlimited linear src/spatial/areconstruction.cpp:51-103,
 *                   MUSCL-Van Albada src/spatial/musclreconstruction.cpp:33-130,
 *                   WENO / Barth-Jespersen / Venkatakrishnan
 *                   src/spatial/limitedlinearreconstruction.cpp:27-268
 *  viscous flux:    src/spatial/flow_spatial.cpp:348-395, src/physics/viscousphysics.cpp:14-122,
 *                   src/spatial/aspatial.cpp:172-205
 *  fluxes + scatter src/spatial/flow_spatial.cpp:488-563; time steps :565-634
 *  Jacobian:        src/spatial/aspatial.cpp:242-340, src/spatial/flow_spatial.cpp:397-446,818-875
 *  functionals:     src/spatial/aoutput.cpp:28-62, src/spatial/flow_spatial.cpp:95-310
 *
 * Deliberate definitions where the reference is undefined:
 *  - Venkatakrishnan / Barth-Jespersen read the ghost primitive state for boundary neighbours
 *    (the reference reads past the end of its cell array, limitedlinearreconstruction.cpp:134,229);
 *  - the limiter parameter K / lambda is an explicit configuration value (never parsed by the
 *    reference, controlparser.cpp:69).
 *
 * Parity: residual pieces pinned by freestream preservation, conservation, the reference's
 * 1-exactness test (tests/finite-volume/testgradientschemes.cpp:36-90), the wall-flux test
 * (tests/flow-general/testwallbcs.cpp:14-78) and the NACA0012 regression functionals
 * (testcases/naca0012/regr-MUSCL_LeastSquares_HLLC.txt). Jacobian blocks pinned by central
 * finite differences of the restated fluxes and BCs. Time steps: closed forms on a uniform
 * Cartesian grid (inviscid and viscous) and, with cell-constant face values and a random state
 * per cell, on the reference's hybrid triangle/quad cylinder mesh.
 */
#include "fvo.hpp"
#include "fvo_physics.hpp"
#include <cstring>
#include <algorithm>

namespace fvo {

std::array<double,NV> freestream_state(const Config& c)
{
	// (aphysics.cpp:44-58)
	std::array<double,NV> u;
	const double pinf = 1.0/(c.gamma*c.Minf*c.Minf);
	u[0] = 1.0;
	u[1] = std::cos(c.aoa)*std::cos(0.0);
	u[2] = std::sin(c.aoa)*std::cos(0.0);
	u[3] = pinf/(c.gamma-1.0) + 0.5*1.0*1.0;
	return u;
}

static BCParams bc_params(const Disc& d, int ib)
{
	BCParams p;
	p.kind = d.bckind[ib];
	for(int i = 0; i < 4; i++) p.uinf[i] = d.uinf[i];
	const BCConf& bc = d.cfg.bcs[d.bcindex[ib]];
	p.v0 = bc.vals[0]; p.v1 = bc.vals[1];
	return p;
}

void disc_init(Disc& d, const Mesh* m, const Config& cfg)
{
	d.m = m;
	d.cfg = cfg;
	d.uinf = freestream_state(cfg);
	d.bckind.assign(m->nbface, -1);
	d.bcindex.assign(m->nbface, -1);
	for(int ib = 0; ib < m->nbface; ib++) {
		for(int k = 0; k < cfg.nbc; k++)
			if(cfg.bcs[k].tag == m->btag[ib]) { d.bckind[ib] = cfg.bcs[k].kind; d.bcindex[ib] = k; }
		if(d.bcindex[ib] < 0)
			throw std::runtime_error("no boundary condition for a boundary tag");
	}

	// least-squares LHS, inverted once (agradientschemes.cpp:218-317)
	const int ne = m->nelem;
	std::vector<double> V(ne*4, 0.0);
	for(int f = 0; f < m->nbface; f++) {
		const int ie = m->intfac[f*4];
		double w2 = 0, dr[2];
		for(int k = 0; k < 2; k++) {
			w2 += (m->rc[ie*2+k]-m->rcbp[f*2+k])*(m->rc[ie*2+k]-m->rcbp[f*2+k]);
			dr[k] = m->rc[ie*2+k]-m->rcbp[f*2+k];
		}
		w2 = 1.0/w2;
		for(int i = 0; i < 2; i++) for(int j = 0; j < 2; j++) V[ie*4+i*2+j] += w2*dr[i]*dr[j];
	}
	for(int f = m->nbface; f < m->naface; f++) {
		const int ie = m->intfac[f*4], je = m->intfac[f*4+1];
		double w2 = 0, dr[2];
		for(int k = 0; k < 2; k++) {
			w2 += (m->rc[ie*2+k]-m->rc[je*2+k])*(m->rc[ie*2+k]-m->rc[je*2+k]);
			dr[k] = m->rc[ie*2+k]-m->rc[je*2+k];
		}
		w2 = 1.0/w2;
		for(int i = 0; i < 2; i++) for(int j = 0; j < 2; j++) {
			V[ie*4+i*2+j] += w2*dr[i]*dr[j];
			V[je*4+i*2+j] += w2*dr[i]*dr[j];
		}
	}
	d.wlsV.assign(ne*4, 0.0);
	for(int ie = 0; ie < ne; ie++) {
		const double a = V[ie*4], b = V[ie*4+1], c = V[ie*4+2], e = V[ie*4+3];
		const double det = a*e - b*c;
		d.wlsV[ie*4+0] = e/det;  d.wlsV[ie*4+1] = -b/det;
		d.wlsV[ie*4+2] = -c/det; d.wlsV[ie*4+3] = a/det;
	}

	// Venkatakrishnan characteristic length: longest edge (limitedlinearreconstruction.cpp:189-205)
	d.clength.assign(ne, 0.0);
	for(int ie = 0; ie < ne; ie++) {
		const int nn = m->nnode[ie];
		for(int j = 0; j < nn; j++) {
			const int a = m->inpoel[ie*4+j], b = m->inpoel[ie*4+(j+1)%nn];
			double l = 0;
			for(int k = 0; k < 2; k++) l += std::pow(m->coords[a*2+k]-m->coords[b*2+k], 2);
			if(d.clength[ie] < l) d.clength[ie] = l;
		}
		d.clength[ie] = std::sqrt(d.clength[ie]);
	}
}

void ghost_state(const Disc& d, int ib, const double* uin, const double* n, double* ug)
{
	const Gas gs(d.cfg);
	bc_ghost(gs, bc_params(d, ib), uin, n, ug);
}

/* Ghost-state Jacobians (abc.cpp:86-140,178-185,201-210,228-262,282-305,371-409,425-437). */
void ghost_state_jacobian(const Disc& d, int ib, const double* ins, const double* n,
                          double* gsv, double* dgs)
{
	const Gas gs(d.cfg);
	const BCParams bc = bc_params(d, ib);
	for(int k = 0; k < 16; k++) dgs[k] = 0;
	switch(bc.kind) {
	case BC_INOUTFLOW: {
		const double vni = dot2(&ins[1],n)/ins[0];
		const double ci = sound_speed(gs, ins[0], pressure_from_conserved(gs, ins));
		const double Mni = vni/ci;
		const double pinf = pressure_from_conserved(gs, bc.uinf);
		if(Mni <= 0) { for(int i = 0; i < 4; i++) gsv[i] = bc.uinf[i]; }
		else if(Mni <= 1) {
			gsv[0] = ins[0]; gsv[1] = ins[1]; gsv[2] = ins[2];
			for(int k = 0; k < 3; k++) dgs[k*4+k] = 1.0;
			gsv[3] = pinf/(gs.g-1.0) + 0.5*ins[0]*(dot2(&ins[1],&ins[1])/(ins[0]*ins[0]));
			dgs[12] = -0.5*dot2(&ins[1],&ins[1])/(ins[0]*ins[0]);
			dgs[13] = ins[1]/ins[0];
			dgs[14] = ins[2]/ins[0];
			dgs[15] = 0;
		}
		else { for(int i = 0; i < 4; i++) { gsv[i] = ins[i]; dgs[i*4+i] = 1.0; } }
		break; }
	case BC_FARFIELD:
		for(int i = 0; i < 4; i++) gsv[i] = bc.uinf[i];
		break;
	case BC_SLIPWALL: {
		const double vni = dot2(&ins[1],n)/ins[0];
		double dvni[4] = { -vni/ins[0], n[0]/ins[0], n[1]/ins[0], 0 };
		gsv[0] = ins[0]; dgs[0] = 1.0;
		for(int i = 1; i < 3; i++) {
			gsv[i] = ins[i] - 2.0*n[i-1]*vni*ins[0];
			dgs[i*4] = -2.0*n[i-1]*(dvni[0]*ins[0] + vni);
			for(int j = 1; j < 3; j++)
				dgs[i*4+j] = (i == j ? 1.0 : 0.0) - 2.0*n[i-1]*dvni[j]*ins[0];
		}
		gsv[3] = ins[3]; dgs[15] = 1.0;
		break; }
	case BC_ADIABATICWALL: {
		const double tm = bc.v0 * ins[0];
		gsv[0] = ins[0]; dgs[0] = 1.0;
		gsv[1] = 2.0*tm*n[1] - ins[1]; dgs[4] = 2.0*bc.v0*n[1]; dgs[5] = -1.0;
		gsv[2] = -2.0*tm*n[0] - ins[2]; dgs[8] = -2.0*bc.v0*n[0]; dgs[10] = -1.0;
		gsv[3] = ins[3]; dgs[15] = 1.0;
		break; }
	case BC_ISOTHERMALWALL: {
		// the reference's own (self-declared inexact) linearisation, abc.cpp:380-408
		const double tm = bc.v0 * ins[0];
		gsv[0] = ins[0]; dgs[0] = 1.0;
		gsv[1] = 2.0*tm*n[1] - ins[1]; dgs[4] = 2.0*bc.v0*n[1]; dgs[5] = -1.0;
		gsv[2] = -2.0*tm*n[0] - ins[2]; dgs[8] = -2.0*bc.v0*n[0]; dgs[10] = -1.0;
		const double vmag2 = dot2(&gsv[1],&gsv[1])/(ins[0]*ins[0]);
		const double coeff = 1.0/(gs.g*(gs.g-1.0)*gs.Minf*gs.Minf);
		gsv[3] = ins[0]*(bc.v1*coeff + 0.5*vmag2);
		double dvmag2[4];
		dvmag2[0] = -2.0*dot2(&gsv[1],&gsv[1])/(ins[0]*ins[0]*ins[0]);
		for(int i = 1; i < 3; i++) dvmag2[i] = 1.0/(ins[0]*ins[0]) * gsv[i] * (-1.0);
		dvmag2[3] = 0;
		// energy row with the temperature derivative taken as zero, as the reference does
		dgs[12] += coeff*bc.v1 + 0.5*(vmag2 + ins[0]*dvmag2[0]);
		for(int i = 1; i < 4; i++) dgs[12+i] += ins[0]*(0.5*dvmag2[i]);
		break; }
	case BC_EXTRAPOLATION:
		for(int k = 0; k < 4; k++) { gsv[k] = ins[k]; dgs[k*4+k] = 1.0; }
		break;
	default:
		throw std::runtime_error("BC Jacobian not implemented for this boundary type");
	}
}

void inviscid_flux(const Config& c, int fluxid, const double* ul, const double* ur,
                   const double* n, double* f)
{
	const Gas gs(c);
	flux_dispatch(gs, fluxid, ul, ur, n, f);
}

/* Flux Jacobians. Roe, HLLC and AUSM: exact derivative of the restated flux (the reference's
 * hand-written Roe/HLLC linearisations are exact derivatives, anumericalflux.cpp:736-965 and
 * 1230-1397). LLF and HLL: the reference's frozen-eigenvalue / frozen-speed approximations
 * (anumericalflux.cpp:65-107, 1012-1061). */
void inviscid_flux_jacobian(const Config& c, int fluxid, const double* ul, const double* ur,
                            const double* n, double* dfdl, double* dfdr)
{
	const Gas gs(c);
	if(fluxid == FLUX_LLF) {
		double vi[2], vj[2], vni, vnj, pi, pj, Hi, Hj;
		vars_from_conserved(gs, ul, n, vi, vni, pi, Hi);
		vars_from_conserved(gs, ur, n, vj, vnj, pj, Hj);
		const double ci = sound_speed(gs, ul[0], pi), cj = sound_speed(gs, ur[0], pj);
		const double eig = std::fabs(vni)+ci >= std::fabs(vnj)+cj ? std::fabs(vni)+ci
		                                                          : std::fabs(vnj)+cj;
		directional_flux_jacobian(gs, ul, n, dfdl);
		directional_flux_jacobian(gs, ur, n, dfdr);
		for(int i = 0; i < 4; i++) { dfdl[i*4+i] -= -eig; dfdr[i*4+i] -= eig; }
		for(int k = 0; k < 16; k++) { dfdl[k] = -0.5*dfdl[k]; dfdr[k] = 0.5*dfdr[k]; }
		return;
	}
	if(fluxid == FLUX_HLL) {
		double vi[2], vj[2], vni, vnj, pi, pj, Hi, Hj;
		vars_from_conserved(gs, ul, n, vi, vni, pi, Hi);
		vars_from_conserved(gs, ur, n, vj, vnj, pj, Hj);
		const double ci = sound_speed(gs, ul[0], pi), cj = sound_speed(gs, ur[0], pj);
		double Rij, rhoij, vij[2], vm2ij, vnij, Hij, cij;
		roe_averages(gs, ul, ur, n, vi, Hi, vj, Hj, Rij, rhoij, vij, vm2ij, vnij, Hij, cij);
		double sl = vni - ci;
		if(sl > vnij-cij) sl = vnij-cij;
		double sr = vnj + cj;
		if(sr < vnij+cij) sr = vnij+cij;
		const double sr0 = sr > 0 ? 0 : sr, sl0 = sl > 0 ? 0 : sl;
		const double t1 = (sr0 - sl0)/(sr-sl), t2 = 1.0 - t1;
		const double t3 = 0.5*(sr*std::fabs(sl)-sl*std::fabs(sr))/(sr-sl);
		directional_flux_jacobian(gs, ul, n, dfdl);
		directional_flux_jacobian(gs, ur, n, dfdr);
		for(int k = 0; k < 16; k++) { dfdl[k] = -t2*dfdl[k]; dfdr[k] = t1*dfdr[k]; }
		for(int i = 0; i < 4; i++) { dfdl[i*4+i] -= t3; dfdr[i*4+i] -= t3; }
		return;
	}
	if(fluxid == FLUX_VANLEER || fluxid == FLUX_AUSMPLUS)
		throw std::runtime_error("flux Jacobian not implemented in the reference for this flux");
	Dual dl[4], dr[4], df[4];
	for(int i = 0; i < 4; i++) {
		dl[i] = Dual(ul[i]); dl[i].d[i] = 1.0;
		dr[i] = Dual(ur[i]); dr[i].d[4+i] = 1.0;
	}
	flux_dispatch(gs, fluxid, dl, dr, n, df);
	for(int i = 0; i < 4; i++)
		for(int j = 0; j < 4; j++) {
			dfdl[i*4+j] = -df[i].d[j];
			dfdr[i*4+j] = df[i].d[4+j];
		}
}

/* ------------------------------------------------------------------ gradients --------- */

// layout of all gradient arrays: [cell][var][dim] (the reference's col-major 2x4 GradBlock_t)
static inline double& G(double* g, int c, int v, int dim) { return g[c*8 + v*2 + dim]; }

static void gradients_greengauss(const Disc& d, const double* u, const double* ug, double* grad,
                                 int nv, int par)
{
	const Mesh& m = *d.m;
	const int ne = m.nelem;
	std::fill(grad, grad + (size_t)ne*nv*2, 0.0);
	auto gi = [&](int c, int v, int k) -> double& { return grad[(size_t)c*nv*2 + v*2 + k]; };
#pragma omp parallel for if(par)
	for(int f = 0; f < m.nbface; f++) {
		const int ie = m.intfac[f*4];
		const double* mid = &m.gr[f*2];
		double dL = 0, dR = 0;
		for(int k = 0; k < 2; k++) {
			dL += (mid[k]-m.rc[ie*2+k])*(mid[k]-m.rc[ie*2+k]);
			dR += (mid[k]-m.rcbp[f*2+k])*(mid[k]-m.rcbp[f*2+k]);
		}
		dL = 1.0/std::sqrt(dL); dR = 1.0/std::sqrt(dR);
		const double ai = 1.0/m.area[ie];
		for(int v = 0; v < nv; v++) {
			const double ut = (u[ie*nv+v]*dL + ug[f*nv+v]*dR)/(dL+dR) * m.facemetric[f*3+2];
			for(int k = 0; k < 2; k++) {
#pragma omp atomic update
				gi(ie,v,k) += (ut * m.facemetric[f*3+k])*ai;
			}
		}
	}
#pragma omp parallel for if(par)
	for(int f = m.nbface; f < m.naface; f++) {
		const int ie = m.intfac[f*4], je = m.intfac[f*4+1];
		const double* mid = &m.gr[f*2];
		double dL = 0, dR = 0;
		for(int k = 0; k < 2; k++) {
			dL += (mid[k]-m.rc[ie*2+k])*(mid[k]-m.rc[ie*2+k]);
			dR += (mid[k]-m.rc[je*2+k])*(mid[k]-m.rc[je*2+k]);
		}
		dL = 1.0/std::sqrt(dL); dR = 1.0/std::sqrt(dR);
		const double a1 = 1.0/m.area[ie], a2 = 1.0/m.area[je];
		for(int v = 0; v < nv; v++) {
			const double ut = (u[ie*nv+v]*dL + u[je*nv+v]*dR)/(dL+dR) * m.facemetric[f*3+2];
			for(int k = 0; k < 2; k++) {
#pragma omp atomic update
				gi(ie,v,k) += (ut * m.facemetric[f*3+k])*a1;
#pragma omp atomic update
				gi(je,v,k) -= (ut * m.facemetric[f*3+k])*a2;
			}
		}
	}
}

static void gradients_leastsquares(const Disc& d, const double* u, const double* ug, double* grad,
                                   int nv, int par)
{
	const Mesh& m = *d.m;
	const int ne = m.nelem;
	std::vector<double> fr((size_t)ne*2*nv, 0.0);       // [cell][dim][var]
	auto F = [&](int c, int k, int v) -> double& { return fr[(size_t)c*2*nv + k*nv + v]; };
#pragma omp parallel for if(par)
	for(int f = 0; f < m.nbface; f++) {
		const int ie = m.intfac[f*4];
		double w2 = 0, dr[2];
		for(int k = 0; k < 2; k++) {
			w2 += (m.rc[ie*2+k]-m.rcbp[f*2+k])*(m.rc[ie*2+k]-m.rcbp[f*2+k]);
			dr[k] = m.rc[ie*2+k]-m.rcbp[f*2+k];
		}
		w2 = 1.0/w2;
		for(int v = 0; v < nv; v++) {
			const double du = u[ie*nv+v] - ug[f*nv+v];
			for(int k = 0; k < 2; k++) {
#pragma omp atomic update
				F(ie,k,v) += w2*dr[k]*du;
			}
		}
	}
#pragma omp parallel for if(par)
	for(int f = m.nbface; f < m.naface; f++) {
		const int ie = m.intfac[f*4], je = m.intfac[f*4+1];
		double w2 = 0, dr[2];
		for(int k = 0; k < 2; k++) {
			w2 += (m.rc[ie*2+k]-m.rc[je*2+k])*(m.rc[ie*2+k]-m.rc[je*2+k]);
			dr[k] = m.rc[ie*2+k]-m.rc[je*2+k];
		}
		w2 = 1.0/w2;
		for(int v = 0; v < nv; v++) {
			const double du = u[ie*nv+v] - u[je*nv+v];
			for(int k = 0; k < 2; k++) {
#pragma omp atomic update
				F(ie,k,v) += w2*dr[k]*du;
#pragma omp atomic update
				F(je,k,v) += w2*dr[k]*du;
			}
		}
	}
#pragma omp parallel for if(par)
	for(int ie = 0; ie < ne; ie++) {
		const double* V = &d.wlsV[ie*4];
		for(int v = 0; v < nv; v++)
			for(int k = 0; k < 2; k++)
				grad[(size_t)ie*nv*2 + v*2 + k] = V[k*2+0]*F(ie,0,v) + V[k*2+1]*F(ie,1,v);
	}
}

static void compute_gradients(const Disc& d, const double* u, const double* ug, double* grad,
                              int nv, int par)
{
	const int ne = d.m->nelem;
	switch(d.cfg.gradient) {
	case GRAD_GREENGAUSS: gradients_greengauss(d, u, ug, grad, nv, par); break;
	case GRAD_LEASTSQUARES: gradients_leastsquares(d, u, ug, grad, nv, par); break;
	default: std::fill(grad, grad + (size_t)ne*nv*2, 0.0);
	}
}

/* ------------------------------------------------------------- reconstruction --------- */

// up: primitive cell states, ug: primitive ghost states (nbface x 4), grads [cell][var][dim]
static void reconstruct(const Disc& d, const double* up, const double* ug, const double* grads,
                        double* ufl, double* ufr, int par)
{
	const Mesh& m = *d.m;
	const int ne = m.nelem;
	auto linext = [&](int c, int v, double lim, const double* gp) {
		double uf = up[c*4+v];
		for(int k = 0; k < 2; k++)
			uf += lim*grads[c*8+v*2+k]*(gp[k] - m.rc[c*2+k]);
		return uf;
	};
	// neighbour primitive value; boundary neighbours use the ghost state (definition B2)
	auto nbval = [&](int jel, int v) {
		return jel < ne ? up[jel*4+v] : ug[(jel-ne)*4+v];
	};

	switch(d.cfg.recon) {
	case RECON_NONE: {
#pragma omp parallel for if(par)
		for(int f = m.nbface; f < m.naface; f++) {
			const int ie = m.intfac[f*4], je = m.intfac[f*4+1];
			for(int v = 0; v < 4; v++) {
				ufl[f*4+v] = linext(ie, v, 1.0, &m.gr[f*2]);
				ufr[f*4+v] = linext(je, v, 1.0, &m.gr[f*2]);
			}
		}
#pragma omp parallel for if(par)
		for(int f = 0; f < m.nbface; f++) {
			const int ie = m.intfac[f*4];
			for(int v = 0; v < 4; v++)
				ufl[f*4+v] = linext(ie, v, 1.0, &m.gr[f*2]);
		}
		break; }
	case RECON_VANALBADA: {
		const double eps = 1e-8, kap = 1.0/3.0;
		auto biased = [&](const double* ri, const double* rj, double ui, double uj,
		                  const double* g) {
			double del = 0;
			for(int k = 0; k < 2; k++) del += g[k]*(rj[k]-ri[k]);
			return 2.0*del - (uj-ui);
		};
		auto left = [&](double ui, double uj, double dm, double phi) {
			return ui + phi/4.0*((1.0-kap*phi)*dm + (1.0+kap*phi)*(uj - ui));
		};
		auto right = [&](double ui, double uj, double dp, double phi) {
			return uj - phi/4.0*((1.0-kap*phi)*dp + (1.0+kap*phi)*(uj - ui));
		};
#pragma omp parallel for if(par)
		for(int f = 0; f < m.nbface; f++) {
			const int ie = m.intfac[f*4];
			for(int v = 0; v < 4; v++) {
				const double ui = up[ie*4+v], uj = ug[f*4+v];
				const double dm = biased(&m.rc[ie*2], &m.rcbp[f*2], ui, uj, &grads[ie*8+v*2]);
				double phl = (2.0*dm*(uj - ui) + eps) / (dm*dm + (uj - ui)*(uj - ui) + eps);
				if(phl < 0.0) phl = 0.0;
				ufl[f*4+v] = left(ui, uj, dm, phl);
			}
		}
#pragma omp parallel for if(par)
		for(int f = m.nbface; f < m.naface; f++) {
			const int ie = m.intfac[f*4], je = m.intfac[f*4+1];
			for(int v = 0; v < 4; v++) {
				const double ui = up[ie*4+v], uj = up[je*4+v];
				const double dm = biased(&m.rc[ie*2], &m.rc[je*2], ui, uj, &grads[ie*8+v*2]);
				const double dp = biased(&m.rc[ie*2], &m.rc[je*2], ui, uj, &grads[je*8+v*2]);
				double phl = (2.0*dm*(uj - ui) + eps) / (dm*dm + (uj - ui)*(uj - ui) + eps);
				if(phl < 0.0) phl = 0.0;
				double phr = (2*dp*(uj - ui) + eps) / (dp*dp + (uj - ui)*(uj - ui) + eps);
				if(phr < 0.0) phr = 0.0;
				ufl[f*4+v] = left(ui, uj, dm, phl);
				ufr[f*4+v] = right(ui, uj, dp, phr);
			}
		}
		break; }
	case RECON_VENKATAKRISHNAN:
	case RECON_BARTHJESPERSEN: {
		const bool venk = d.cfg.recon == RECON_VENKATAKRISHNAN;
#pragma omp parallel for if(par)
		for(int ie = 0; ie < ne; ie++) {
			const int nf = m.nnode[ie];
			const double eps2 = venk ? std::pow(d.cfg.limiter_param*d.clength[ie], 3) : 0.0;
			for(int v = 0; v < 4; v++) {
				double dmin = 0, dmax = 0;
				for(int j = 0; j < nf; j++) {
					const double dui = nbval(m.esuel[ie*4+j], v) - up[ie*4+v];
					if(dui > dmax) dmax = dui;
					if(dui < dmin) dmin = dui;
				}
				double lim = 1.0;
				for(int j = 0; j < nf; j++) {
					const int face = m.elemface[ie*4+j];
					const double uface = linext(ie, v, 1.0, &m.gr[face*2]);
					double phi;
					if(venk) {
						const double dm = uface - up[ie*4+v];
						const double dp = dm < 0 ? dmin : dmax;
						phi = (dp*dp + 2*dp*dm + eps2)/(dp*dp + dp*dm + 2*dm*dm + eps2);
					} else {
						const double diff = uface - up[ie*4+v];
						if(diff > 0) phi = 1 < dmax/diff ? 1 : dmax/diff;
						else if(diff < 0) phi = 1 < dmin/diff ? 1 : dmin/diff;
						else phi = 1;
					}
					if(phi < lim) lim = phi;
				}
				for(int j = 0; j < nf; j++) {
					const int face = m.elemface[ie*4+j];
					const int jel = m.esuel[ie*4+j];
					if(ie < jel) ufl[face*4+v] = linext(ie, v, lim, &m.gr[face*2]);
					else         ufr[face*4+v] = linext(ie, v, lim, &m.gr[face*2]);
				}
			}
		}
		break; }
	case RECON_WENO: {
		const double gam = 4.0, lambda = d.cfg.limiter_param, epsw = 1.0e-5;
#pragma omp parallel for if(par)
		for(int ie = 0; ie < ne; ie++) {
			const int nf = m.nnode[ie];
			for(int v = 0; v < 4; v++) {
				double wsum = 0, lg[2] = {0, 0};
				{
					const double gm = grads[ie*8+v*2]*grads[ie*8+v*2]
					                  + grads[ie*8+v*2+1]*grads[ie*8+v*2+1];
					const double w = lambda / std::pow(gm + epsw, gam);
					wsum += w;
					for(int k = 0; k < 2; k++) lg[k] += w*grads[ie*8+v*2+k];
				}
				for(int j = 0; j < nf; j++) {
					const int je = m.esuel[ie*4+j];
					if(je >= ne) continue;
					const double gm = grads[je*8+v*2]*grads[je*8+v*2]
					                  + grads[je*8+v*2+1]*grads[je*8+v*2+1];
					const double w = 1.0 / std::pow(gm + epsw, gam);
					wsum += w;
					for(int k = 0; k < 2; k++) lg[k] += w*grads[je*8+v*2+k];
				}
				for(int k = 0; k < 2; k++) lg[k] /= wsum;
				for(int j = 0; j < nf; j++) {
					const int face = m.elemface[ie*4+j];
					const int je = m.esuel[ie*4+j];
					double* dst = ie < je ? &ufl[face*4+v] : &ufr[face*4+v];
					*dst = up[ie*4+v];
					for(int k = 0; k < 2; k++) *dst += lg[k]*(m.gr[face*2+k] - m.rc[ie*2+k]);
				}
			}
		}
		break; }
	default:
		throw std::runtime_error("unknown reconstruction");
	}
}

/* ---------------------------------------------------------------- viscous flux ---------- */

// compute_viscous_flux (flow_spatial.cpp:348-395) with getPrimitive2StatesAndGradients
// (viscousphysics.cpp:14-68), modified-average face gradient (aspatial.cpp:172-205) and
// computeViscousFlux (viscousphysics.cpp:70-122). gradsl/gradsr: [var][dim] primitive.
static void viscous_flux(const Disc& d, const double* n, const double* rcl, const double* rcr,
                         const double* ucl, const double* ucr, const double* gradsl,
                         const double* gradsr, const double* ul, const double* ur, double* vflux)
{
	const Gas gs(d.cfg);
	const bool o2 = d.cfg.order2;
	double uctl[4], uctr[4], gl[2][4], gr[2][4];
	if(o2) {
		primitive_from_conserved(gs, ucl, uctl);
		primitive_from_conserved(gs, ucr, uctr);
		for(int i = 0; i < 2; i++)
			for(int j = 0; j < 4; j++) { gl[i][j] = gradsl[j*2+i]; gr[i][j] = gradsr[j*2+i]; }
		for(int j = 0; j < 2; j++) {
			gl[j][3] = grad_temperature(gs, uctl[0], gradsl[0*2+j], uctl[3], gradsl[3*2+j]);
			gr[j][3] = grad_temperature(gs, uctr[0], gradsr[0*2+j], uctr[3], gradsr[3*2+j]);
		}
		uctl[3] = temperature(gs, uctl[0], uctl[3]);
		uctr[3] = temperature(gs, uctr[0], uctr[3]);
	} else {
		primitive_from_conserved(gs, ucl, uctl);
		primitive_from_conserved(gs, ucr, uctr);
		uctl[3] = temperature(gs, ucl[0], uctl[3]);
		uctr[3] = temperature(gs, ucr[0], uctr[3]);
		for(int i = 0; i < 2; i++) for(int j = 0; j < 4; j++) { gl[i][j] = 0; gr[i][j] = 0; }
	}
	// modified average gradient
	double dr[2], dist = 0;
	for(int i = 0; i < 2; i++) { dr[i] = rcr[i]-rcl[i]; dist += dr[i]*dr[i]; }
	dist = std::sqrt(dist);
	for(int i = 0; i < 2; i++) dr[i] /= dist;
	double grad[2][4];
	for(int i = 0; i < 4; i++) {
		double davg[2];
		for(int j = 0; j < 2; j++) davg[j] = 0.5*(gl[j][i] + gr[j][i]);
		const double corr = (uctr[i]-uctl[i])/dist;
		const double ddr = dot2(davg,dr);
		for(int j = 0; j < 2; j++) grad[j][i] = davg[j] - ddr*dr[j] + corr*dr[j];
	}
	// viscous flux
	const double muRe = d.cfg.constvisc ? 1.0/gs.Reinf
		: 0.5*(viscosity_from_conserved(gs, ul) + viscosity_from_conserved(gs, ur));
	const double kdiff = thermal_conductivity(gs, muRe);
	double stress[2][2];
	double ldiv = 0;
	for(int j = 0; j < 2; j++) ldiv += grad[j][j+1];
	ldiv *= 2.0/3.0*muRe;
	for(int i = 0; i < 2; i++) {
		for(int j = 0; j < 2; j++) stress[i][j] = muRe*(grad[i][j+1] + grad[j][i+1]);
		stress[i][i] -= ldiv;
	}
	vflux[0] = 0;
	for(int i = 0; i < 2; i++) {
		vflux[i+1] = 0;
		for(int j = 0; j < 2; j++) vflux[i+1] -= stress[i][j]*n[j];
	}
	double vavg[2];
	for(int j = 0; j < 2; j++) vavg[j] = 0.5*(ul[j+1]/ul[0] + ur[j+1]/ur[0]);
	vflux[3] = 0;
	for(int i = 0; i < 2; i++) {
		double comp = 0;
		for(int j = 0; j < 2; j++) comp += stress[i][j]*vavg[j];
		comp += kdiff*grad[i][3];
		vflux[3] -= comp*n[i];
	}
}

/* ------------------------------------------------------------------ residual ----------- */

static void residual_impl(const Disc& d, const double* u, double* r, bool gettimesteps,
                          double* dt, int par, bool absflux);

void compute_residual(const Disc& d, const double* u, double* r, bool gettimesteps, double* dt,
                      int par)
{
	residual_impl(d, u, r, gettimesteps, dt, par, false);
}

/* Sum over each cell's faces of |F_f| len_f per variable: the cancellation-aware scale of the
 * per-cell parity metric (not a reference quantity). */
void flux_magnitude(const Disc& d, const double* u, double* s)
{
	residual_impl(d, u, s, false, nullptr, 0, true);
}

static void residual_impl(const Disc& d, const double* u, double* r, bool gettimesteps,
                          double* dt, int par, bool absflux)
{
	const Mesh& m = *d.m;
	const Gas gs(d.cfg);
	const int ne = m.nelem, nb = m.nbface, nf = m.naface;
	std::vector<double> ufl((size_t)nf*4, 0.0), ufr((size_t)nf*4, 0.0);
	std::vector<double> ubcell((size_t)nb*4, 0.0);
	std::vector<double> grads;

	// cell values as left states of boundary faces (flow_spatial.cpp:656-666)
#pragma omp parallel for if(par)
	for(int f = 0; f < nb; f++)
		for(int v = 0; v < 4; v++) ufl[f*4+v] = u[m.intfac[f*4]*4+v];

	if(d.cfg.order2) {
		// ghost states from cell values, kept conserved in ubcell, primitive in ufr (:679-700)
#pragma omp parallel for if(par)
		for(int f = 0; f < nb; f++) {
			ghost_state(d, f, &ufl[f*4], &m.facemetric[f*3], &ufr[f*4]);
			for(int v = 0; v < 4; v++) ubcell[f*4+v] = ufr[f*4+v];
			double tmp[4];
			primitive_from_conserved(gs, &ufr[f*4], tmp);
			for(int v = 0; v < 4; v++) ufr[f*4+v] = tmp[v];
		}
		std::vector<double> up((size_t)ne*4);
#pragma omp parallel for if(par)
		for(int c = 0; c < ne; c++) primitive_from_conserved(gs, &u[c*4], &up[c*4]);
		std::vector<double> ug(ufr.begin(), ufr.begin() + (size_t)nb*4);
		grads.assign((size_t)ne*8, 0.0);
		compute_gradients(d, up.data(), ug.data(), grads.data(), 4, par);
		reconstruct(d, up.data(), ug.data(), grads.data(), ufl.data(), ufr.data(), par);
		// back to conserved (:731-754)
#pragma omp parallel for if(par)
		for(int f = nb; f < nf; f++) {
			double t[4];
			conserved_from_primitive(gs, &ufl[f*4], t); for(int v = 0; v < 4; v++) ufl[f*4+v] = t[v];
			conserved_from_primitive(gs, &ufr[f*4], t); for(int v = 0; v < 4; v++) ufr[f*4+v] = t[v];
		}
#pragma omp parallel for if(par)
		for(int f = 0; f < nb; f++) {
			double t[4];
			conserved_from_primitive(gs, &ufl[f*4], t); for(int v = 0; v < 4; v++) ufl[f*4+v] = t[v];
		}
	} else {
#pragma omp parallel for if(par)
		for(int f = nb; f < nf; f++)
			for(int v = 0; v < 4; v++) {
				ufl[f*4+v] = u[m.intfac[f*4]*4+v];
				ufr[f*4+v] = u[m.intfac[f*4+1]*4+v];
			}
	}

	// ghost states from the (reconstructed) left states (:777-778)
#pragma omp parallel for if(par)
	for(int f = 0; f < nb; f++)
		ghost_state(d, f, &ufl[f*4], &m.facemetric[f*3], &ufr[f*4]);

	const double* ugpb = d.cfg.order2 ? ubcell.data() : &ufr[0];

	// fluxes and scatter (:488-563)
#pragma omp parallel for if(par)
	for(int f = 0; f < nf; f++) {
		const double n[2] = { m.facemetric[f*3], m.facemetric[f*3+1] };
		const double len = m.facemetric[f*3+2];
		const int le = m.intfac[f*4], re = m.intfac[f*4+1];
		double fl[4];
		flux_dispatch(gs, d.cfg.flux, &ufl[f*4], &ufr[f*4], n, fl);
		for(int v = 0; v < 4; v++) fl[v] *= len;
		if(d.cfg.viscous) {
			const bool isb = f < nb;
			const double* rcr = isb ? &m.rcbp[f*2] : &m.rc[re*2];
			const double* ucr = isb ? &ugpb[f*4] : &u[re*4];
			static const double zero8[8] = {0,0,0,0,0,0,0,0};
			const double* gL = d.cfg.order2 ? &grads[(size_t)le*8] : zero8;
			const double* gR = d.cfg.order2 ? (isb ? gL : &grads[(size_t)re*8]) : zero8;
			double vf[4];
			viscous_flux(d, n, &m.rc[le*2], rcr, &u[le*4], ucr, gL, gR, &ufl[f*4], &ufr[f*4], vf);
			for(int v = 0; v < 4; v++) fl[v] += vf[v]*len;
		}
		if(absflux) for(int v = 0; v < 4; v++) fl[v] = -std::fabs(fl[v]);
		for(int v = 0; v < 4; v++) {
#pragma omp atomic update
			r[le*4+v] -= fl[v];
		}
		if(absflux) for(int v = 0; v < 4; v++) fl[v] = -fl[v];
		if(re < ne)
			for(int v = 0; v < 4; v++) {
#pragma omp atomic update
				r[re*4+v] += fl[v];
			}
	}

	if(gettimesteps) {
		std::vector<double> integ(ne, 0.0);
#pragma omp parallel for if(par)
		for(int f = 0; f < nf; f++) {
			const double n[2] = { m.facemetric[f*3], m.facemetric[f*3+1] };
			const double len = m.facemetric[f*3+2];
			const int le = m.intfac[f*4], re = m.intfac[f*4+1];
			const double* ul = &ufl[f*4];
			const double* ur = &ufr[f*4];
			const double ci = sound_speed(gs, ul[0], pressure_from_conserved(gs, ul));
			const double cj = sound_speed(gs, ur[0], pressure_from_conserved(gs, ur));
			const double vni = dot2(&ul[1],n)/ul[0];
			const double vnj = dot2(&ur[1],n)/ur[0];
			double sri = (std::fabs(vni)+ci)*len;
			double srj = (std::fabs(vnj)+cj)*len;
			if(d.cfg.viscous) {
				double mui, muj;
				if(d.cfg.constvisc) { mui = 1.0/gs.Reinf; muj = 1.0/gs.Reinf; }
				else { mui = viscosity_from_conserved(gs, ul); muj = viscosity_from_conserved(gs, ur); }
				const double coi = std::max(4.0/(3*ul[0]), gs.g/ul[0]);
				const double coj = std::max(4.0/(3*ur[0]), gs.g/ur[0]);
				sri += coi*mui/gs.Pr * len*len/m.area[le];
				if(re < ne) srj += coj*muj/gs.Pr * len*len/m.area[re];
			}
#pragma omp atomic update
			integ[le] += sri;
			if(re < ne) {
#pragma omp atomic update
				integ[re] += srj;
			}
		}
#pragma omp parallel for if(par)
		for(int c = 0; c < ne; c++) dt[c] = m.area[c]/integ[c];
	}
}

/* Scalar-field gradients with the configured scheme and unlimited linear face values, the
 * pair exercised by the reference's 1-exactness test
 * (tests/finite-volume/testgradientschemes.cpp:36-90). */
void scalar_face_values(const Disc& d, const double* u, const double* ug, double* grads,
                        double* ufl, double* ufr)
{
	const Mesh& m = *d.m;
	compute_gradients(d, u, ug, grads, 1, 0);
	for(int f = 0; f < m.naface; f++) {
		const int ie = m.intfac[f*4], je = m.intfac[f*4+1];
		double l = u[ie], r = 0;
		for(int k = 0; k < 2; k++) l += 1.0*grads[ie*2+k]*(m.gr[f*2+k] - m.rc[ie*2+k]);
		if(f >= m.nbface) {
			r = u[je];
			for(int k = 0; k < 2; k++) r += 1.0*grads[je*2+k]*(m.gr[f*2+k] - m.rc[je*2+k]);
		}
		ufl[f] = l; ufr[f] = r;
	}
}

void compute_primitive_gradients(const Disc& d, const double* u, double* grads, int par)
{
	const Mesh& m = *d.m;
	const Gas gs(d.cfg);
	const int ne = m.nelem, nb = m.nbface;
	std::vector<double> ug((size_t)nb*4), up((size_t)ne*4);
	for(int f = 0; f < nb; f++) {
		double g[4];
		ghost_state(d, f, &u[m.intfac[f*4]*4], &m.facemetric[f*3], g);
		primitive_from_conserved(gs, g, &ug[f*4]);
	}
	for(int c = 0; c < ne; c++) primitive_from_conserved(gs, &u[c*4], &up[c*4]);
	compute_gradients(d, up.data(), ug.data(), grads, 4, par);
}

/* Primitive face states of the second-order reconstruction (ufl, ufr [naface][4]; a boundary
 * face has its left state only), the values residual_impl converts back to conserved before the
 * flux. Exposed so that tests can read a limiter's factor off a face value. */
void reconstructed_face_states(const Disc& d, const double* u, double* ufl, double* ufr)
{
	const Mesh& m = *d.m;
	const Gas gs(d.cfg);
	const int ne = m.nelem, nb = m.nbface;
	std::vector<double> ug((size_t)nb*4), up((size_t)ne*4), grads((size_t)ne*8, 0.0);
	for(int f = 0; f < nb; f++) {
		double g[4];
		ghost_state(d, f, &u[m.intfac[f*4]*4], &m.facemetric[f*3], g);
		primitive_from_conserved(gs, g, &ug[f*4]);
	}
	for(int c = 0; c < ne; c++) primitive_from_conserved(gs, &u[c*4], &up[c*4]);
	compute_gradients(d, up.data(), ug.data(), grads.data(), 4, 0);
	std::fill(ufl, ufl + (size_t)m.naface*4, 0.0);
	std::fill(ufr, ufr + (size_t)m.naface*4, 0.0);
	reconstruct(d, up.data(), ug.data(), grads.data(), ufl, ufr, 0);
}

void compute_conserved_gradients(const Disc& d, const double* u, double* grads)
{
	const Mesh& m = *d.m;
	std::vector<double> ug((size_t)m.nbface*4);
	for(int f = 0; f < m.nbface; f++)
		ghost_state(d, f, &u[m.intfac[f*4]*4], &m.facemetric[f*3], &ug[f*4]);
	compute_gradients(d, u, ug.data(), grads, 4, 0);
}

/* ----------------------------------------------------------------- Jacobian ------------ */

// Thin-layer viscous flux as a function of the two cell states; its exact derivative is the
// reference's viscous Jacobian (flow_spatial.cpp:397-446, aspatial.cpp:207-240,
// viscousphysics.cpp:124-246).
template <class T>
static void viscous_flux_thinlayer(const Gas& gs, const Config& c, const double* n,
                                   const double* rcl, const double* rcr, const T* ul,
                                   const T* ur, T* vflux)
{
	T upl[4], upr[4];
	primitive_from_conserved(gs, ul, upl); upl[3] = temperature(gs, ul[0], upl[3]);
	primitive_from_conserved(gs, ur, upr); upr[3] = temperature(gs, ur[0], upr[3]);
	double dr[2], dist = 0;
	for(int i = 0; i < 2; i++) { dr[i] = rcr[i]-rcl[i]; dist += dr[i]*dr[i]; }
	dist = std::sqrt(dist);
	for(int i = 0; i < 2; i++) dr[i] /= dist;
	T grad[2][4];
	for(int i = 0; i < 4; i++) {
		const T corr = (upr[i]-upl[i])/dist;
		for(int j = 0; j < 2; j++) grad[j][i] = corr*dr[j];
	}
	const T muRe = c.constvisc ? T(1.0/gs.Reinf)
		: 0.5*(viscosity_from_conserved(gs, ul) + viscosity_from_conserved(gs, ur));
	const T kdiff = thermal_conductivity(gs, muRe);
	T div = grad[0][1] + grad[1][2];
	const T ldiv = 2.0/3.0*muRe*div;
	T stress[2][2];
	for(int i = 0; i < 2; i++) {
		for(int j = 0; j < 2; j++) stress[i][j] = muRe*(grad[i][j+1] + grad[j][i+1]);
		stress[i][i] -= ldiv;
	}
	vflux[0] = 0;
	for(int i = 0; i < 2; i++) {
		vflux[i+1] = 0;
		for(int j = 0; j < 2; j++) vflux[i+1] -= stress[i][j]*n[j];
	}
	T vavg[2];
	for(int j = 0; j < 2; j++) vavg[j] = 0.5*(ul[j+1]/ul[0] + ur[j+1]/ur[0]);
	vflux[3] = 0;
	for(int i = 0; i < 2; i++) {
		T comp = 0;
		for(int j = 0; j < 2; j++) comp += stress[i][j]*vavg[j];
		comp += kdiff*grad[i][3];
		vflux[3] -= comp*n[i];
	}
}

static void viscous_jacobian(const Disc& d, const double* n, const double* rcl, const double* rcr,
                             const double* ul, const double* ur, double* dvfi, double* dvfj)
{
	const Gas gs(d.cfg);
	Dual dl[4], dr[4], vf[4];
	for(int i = 0; i < 4; i++) {
		dl[i] = Dual(ul[i]); dl[i].d[i] = 1.0;
		dr[i] = Dual(ur[i]); dr[i].d[4+i] = 1.0;
	}
	viscous_flux_thinlayer(gs, d.cfg, n, rcl, rcr, dl, dr, vf);
	for(int i = 0; i < 4; i++)
		for(int j = 0; j < 4; j++) {
			dvfi[i*4+j] += -vf[i].d[j];
			dvfj[i*4+j] += vf[i].d[4+j];
		}
}

void compute_local_jacobian_interior(const Disc& d, int iface, const double* ul, const double* ur,
                                     double* L, double* U)
{
	const Mesh& m = *d.m;
	const double n[2] = { m.facemetric[iface*3], m.facemetric[iface*3+1] };
	const double len = m.facemetric[iface*3+2];
	inviscid_flux_jacobian(d.cfg, d.cfg.jacflux, ul, ur, n, L, U);
	if(d.cfg.viscous)
		viscous_jacobian(d, n, &m.rc[m.intfac[iface*4]*2], &m.rc[m.intfac[iface*4+1]*2], ul, ur,
		                 L, U);
	for(int k = 0; k < 16; k++) { L[k] *= len; U[k] *= len; }
}

void compute_local_jacobian_boundary(const Disc& d, int iface, const double* ul, double* L)
{
	const Mesh& m = *d.m;
	const double n[2] = { m.facemetric[iface*3], m.facemetric[iface*3+1] };
	const double len = m.facemetric[iface*3+2];
	double ug[4], dug[16], left[16], right[16];
	ghost_state_jacobian(d, iface, ul, n, ug, dug);
	inviscid_flux_jacobian(d.cfg, d.cfg.jacflux, ul, ug, n, left, right);
	if(d.cfg.viscous)
		viscous_jacobian(d, n, &m.rc[m.intfac[iface*4]*2], &m.rcbp[iface*2], ul, ug, left, right);
	for(int i = 0; i < 4; i++)
		for(int j = 0; j < 4; j++) {
			double s = 0;
			for(int k = 0; k < 4; k++) s += right[i*4+k]*dug[k*4+j];
			L[i*4+j] = len*(left[i*4+j] - s);
		}
}

void bsr_pattern(const Mesh& m, BSR& A)
{
	const int ne = m.nelem;
	A.n = ne;
	A.rowptr.assign(ne+1, 0);
	std::vector<std::vector<int>> cols(ne);
	for(int c = 0; c < ne; c++) {
		cols[c].push_back(c);
		for(int j = 0; j < m.nnode[c]; j++) {
			const int o = m.esuel[c*4+j];
			if(o >= 0 && o < ne) cols[c].push_back(o);
		}
		std::sort(cols[c].begin(), cols[c].end());
		A.rowptr[c+1] = A.rowptr[c] + (int)cols[c].size();
	}
	A.colind.resize(A.rowptr[ne]);
	A.diag.resize(ne);
	for(int c = 0; c < ne; c++)
		for(size_t k = 0; k < cols[c].size(); k++) {
			A.colind[A.rowptr[c]+k] = cols[c][k];
			if(cols[c][k] == c) A.diag[c] = A.rowptr[c]+(int)k;
		}
	A.vals.assign((size_t)A.rowptr[ne]*16, 0.0);
}

static int bsr_find(const BSR& A, int row, int col)
{
	for(int k = A.rowptr[row]; k < A.rowptr[row+1]; k++)
		if(A.colind[k] == col) return k;
	throw std::runtime_error("BSR entry not in pattern");
}

// assemble_jacobian (aspatial.cpp:242-340): zeroes and then adds all face contributions
void assemble_jacobian(const Disc& d, const double* u, BSR& A)
{
	const Mesh& m = *d.m;
	if(A.n != m.nelem) bsr_pattern(m, A);
	std::fill(A.vals.begin(), A.vals.end(), 0.0);
	for(int f = 0; f < m.nbface; f++) {
		const int le = m.intfac[f*4];
		double L[16];
		compute_local_jacobian_boundary(d, f, &u[le*4], L);
		double* D = &A.vals[(size_t)A.diag[le]*16];
		for(int k = 0; k < 16; k++) D[k] += -L[k];
	}
	for(int f = m.nbface; f < m.naface; f++) {
		const int le = m.intfac[f*4], re = m.intfac[f*4+1];
		double L[16], U[16];
		compute_local_jacobian_interior(d, f, &u[le*4], &u[re*4], L, U);
		double* Arl = &A.vals[(size_t)bsr_find(A, re, le)*16];
		double* Alr = &A.vals[(size_t)bsr_find(A, le, re)*16];
		double* Dl = &A.vals[(size_t)A.diag[le]*16];
		double* Dr = &A.vals[(size_t)A.diag[re]*16];
		for(int k = 0; k < 16; k++) {
			Arl[k] += L[k];
			Alr[k] += U[k];
			Dl[k] += -L[k];
			Dr[k] += -U[k];
		}
	}
}

/* -------------------------------------------------------------- functionals ------------ */

double entropy_error(const Disc& d, const double* u)
{
	const Mesh& m = *d.m;
	const Gas gs(d.cfg);
	const double sinf = entropy_from_conserved(gs, d.uinf.data());
	double err = 0;
	for(int c = 0; c < m.nelem; c++) {
		const double s = (entropy_from_conserved(gs, &u[c*4]) - sinf)/sinf;
		err += s*s*m.area[c];
	}
	return std::sqrt(err);
}

void surface_functionals(const Disc& d, const double* u, int walltag, double* cl, double* cdp,
                         double* cdsf)
{
	const Mesh& m = *d.m;
	const Gas gs(d.cfg);
	std::vector<double> grads((size_t)m.nelem*8);
	compute_conserved_gradients(d, u, grads.data());
	const double wind[2] = { std::cos(d.cfg.aoa)*std::cos(0.0), std::sin(d.cfg.aoa)*std::cos(0.0) };
	const double fnorm[2] = { -wind[1], wind[0] };
	const double pinf = 1.0/(gs.g*gs.Minf*gs.Minf);
	double Cl = 0, Cdp = 0, Cdf = 0, tot = 0;
	for(int f = 0; f < m.nbface; f++) {
		if(m.btag[f] != walltag) continue;
		const int le = m.intfac[f*4];
		const double n[2] = { m.facemetric[f*3], m.facemetric[f*3+1] };
		const double len = m.facemetric[f*3+2];
		const double tangf[2] = { n[1], -n[0] };
		const double* ur = &u[le*4];
		const double cp = (pressure_from_conserved(gs, ur) - pinf)*2.0;
		const double muhat = viscosity_from_conserved(gs, ur);
		double gradu[2][2];
		for(int i = 0; i < 2; i++)
			for(int j = 0; j < 2; j++)
				gradu[i][j] = (grads[le*8+(i+1)*2+j]*ur[0] - ur[i+1]*grads[le*8+0*2+j])/(ur[0]*ur[0]);
		double force[2];
		for(int i = 0; i < 2; i++) {
			force[i] = 0;
			for(int j = 0; j < 2; j++) force[i] += (gradu[i][j] + gradu[j][i])*n[j];
		}
		const double tauw = muhat*dot2(force,tangf);
		const double csf = 2*tauw;
		tot += len;
		Cl += cp*dot2(n,fnorm)*len;
		Cdp += cp*dot2(n,wind)*len;
		Cdf += csf*dot2(tangf,wind)*len;
	}
	*cl = Cl/tot; *cdp = Cdp/tot; *cdsf = Cdf/tot;
}

}
