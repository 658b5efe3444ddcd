/* fvo_physics.hpp -- oracle gas physics, numerical fluxes and boundary states
 * (TEST INFRASTRUCTURE ONLY).
 *
 * Every routine is templated on the scalar type so that the same restatement yields the
 * flux (T = double) and its exact derivative (T = Dual).
 *
 * Restates:
 *  - IdealGasPhysics         src/physics/aphysics_defs.hpp:13-487, src/physics/aphysics.cpp:28-58
 *  - LLF flux / frozen Jac   src/spatial/anumericalflux.cpp:40-107
 *  - Van Leer flux           src/spatial/anumericalflux.cpp:202-250
 *  - AUSM flux               src/spatial/anumericalflux.cpp:264-315
 *  - AUSM+ flux              src/spatial/anumericalflux.cpp:479-553
 *  - Roe averages            src/spatial/anumericalflux.hpp:175-189
 *  - Roe flux (Harten fix)   src/spatial/anumericalflux.cpp:667-732
 *  - HLL flux / frozen Jac   src/spatial/anumericalflux.cpp:973-1061
 *  - HLLC flux               src/spatial/anumericalflux.cpp:1069-1081,1175-1228
 * The Roe and HLLC reference Jacobians (anumericalflux.cpp:736-965, 1230-1397) are the exact
 * derivatives of their fluxes; the oracle obtains them by forward differentiation of the
 * restated flux. Parity: pinned by central finite differences in tests/test_oracle.py, by flux
 * consistency F(u,u,n) = analytic flux, rotation, supersonic one-sidedness, exact stationary
 * contacts (HLLC, AUSM, AUSM+), the published subsonic split functions of Van Leer (1982), AUSM
 * (Liou-Steffen 1993) and AUSM+ (Liou 1996: beta = 1/8, alpha = 3/16), C1 joins at M = +-1,
 * and by the reference wall-flux test (tests/flow-general/testwallbcs.cpp:14-78).
 */
#ifndef FVO_PHYSICS_HPP
#define FVO_PHYSICS_HPP

#include "fvo.hpp"

namespace fvo {

/// Gas constants carried through the physics routines (src/physics/aphysics.cpp:17-20)
struct Gas {
	double g, Minf, Tinf, Reinf, Pr, sC;
	explicit Gas(const Config& c) : g(c.gamma), Minf(c.Minf), Tinf(c.Tinf), Reinf(c.Reinf),
	                                Pr(c.Pr), sC(110.5) {}
};

template <class T> inline T dot2(const T* a, const T* b) { T s = 0; s += a[0]*b[0]; s += a[1]*b[1]; return s; }
inline Dual dot2(const Dual* a, const double* b) { Dual s = 0; s += a[0]*b[0]; s += a[1]*b[1]; return s; }

// (aphysics_defs.hpp:53-63)
template <class T> inline T pressure_from_conserved(const Gas& gs, const T* uc)
{
	return (gs.g-1.0)*(uc[3] - 0.5*dot2(&uc[1],&uc[1])/uc[0]);
}
// (aphysics_defs.hpp:27-38)
template <class T> inline void vars_from_conserved(const Gas& gs, const T* uc, const double* n,
                                                   T* v, T& vn, T& p, T& H)
{
	v[0] = uc[1]/uc[0];
	v[1] = uc[2]/uc[0];
	vn = dot2(v,n);
	const T vmag2 = dot2(v,v);
	p = (gs.g-1.0)*(uc[3] - 0.5*uc[0]*vmag2);
	H = (uc[3]+p)/uc[0];
}
// (aphysics_defs.hpp:138-143)
template <class T> inline T sound_speed(const Gas& gs, const T& rho, const T& p)
{
	return sqrt(gs.g * p/rho);
}
// (aphysics_defs.hpp:13-23)
template <class T> inline void directional_flux(const T* uc, const double* n, const T& vn,
                                                const T& p, T* f)
{
	f[0] = vn*uc[0];
	f[1] = vn*uc[1] + p*n[0];
	f[2] = vn*uc[2] + p*n[1];
	f[3] = vn*(uc[3] + p);
}
// (aphysics.cpp:28-35)
template <class T> inline void directional_flux_from_conserved(const Gas& gs, const T* u,
                                                               const double* n, T* f)
{
	const T vn = dot2(&u[1],n)/u[0];
	const T p = (gs.g-1.0)*(u[3] - 0.5*dot2(&u[1],&u[1])/u[0]);
	directional_flux(u, n, vn, p, f);
}
// (aphysics_defs.hpp:257-267)
template <class T> inline void primitive_from_conserved(const Gas& gs, const T* uc, T* up)
{
	const T rho = uc[0];
	const T p = pressure_from_conserved(gs, uc);
	const T vx = uc[1]/uc[0], vy = uc[2]/uc[0];
	up[0] = rho; up[1] = vx; up[2] = vy; up[3] = p;
}
// (aphysics_defs.hpp:209-215, 238-244, 285-295)
template <class T> inline void conserved_from_primitive(const Gas& gs, const T* up, T* uc)
{
	const T rho = up[0];
	const T vmag2 = dot2(&up[1],&up[1]);
	const T rhoE = up[3]/(gs.g-1.0) + 0.5*rho*vmag2;
	const T mx = up[0]*up[1], my = up[0]*up[2];
	uc[0] = rho; uc[1] = mx; uc[2] = my; uc[3] = rhoE;
}
// (aphysics_defs.hpp:116-122)
template <class T> inline T temperature(const Gas& gs, const T& rho, const T& p)
{
	return p/rho * gs.g*gs.Minf*gs.Minf;
}
// (aphysics_defs.hpp:347-353)
template <class T> inline T grad_temperature(const Gas& gs, const T& rho, const T& gradrho,
                                             const T& p, const T& gradp)
{
	return (gradp*rho - p*gradrho) / (rho*rho) * gs.g*gs.Minf*gs.Minf;
}
// (aphysics_defs.hpp:408-421)
template <class T> inline T sutherland_viscosity(const Gas& gs, const T& Tem)
{
	return (1.0+gs.sC/gs.Tinf)/(Tem+gs.sC/gs.Tinf) * pow(Tem,1.5) / gs.Reinf;
}
template <class T> inline T viscosity_from_conserved(const Gas& gs, const T* uc)
{
	const T Tem = temperature(gs, uc[0], pressure_from_conserved(gs, uc));
	return sutherland_viscosity(gs, Tem);
}
// (aphysics_defs.hpp:447-451)
template <class T> inline T thermal_conductivity(const Gas& gs, const T& muhat)
{
	return muhat / (gs.Minf*gs.Minf*(gs.g-1.0)*gs.Pr);
}
// (aphysics_defs.hpp:202-207)
inline double entropy_from_conserved(const Gas& gs, const double* uc)
{
	return pressure_from_conserved(gs, uc)/std::pow(uc[0],gs.g);
}

/* -------------------------------------------------------------------- fluxes ---------- */

// Roe averages (anumericalflux.hpp:175-189)
template <class T> inline void roe_averages(const Gas& gs, const T* ul, const T* ur,
                                            const double* n, const T* vi, const T& Hi,
                                            const T* vj, const T& Hj, T& Rij, T& rhoij,
                                            T* vij, T& vm2ij, T& vnij, T& Hij, T& cij)
{
	Rij = sqrt(ur[0]/ul[0]);
	rhoij = Rij*ul[0];
	vij[0] = (Rij*vj[0] + vi[0])/(Rij + 1.0);
	vij[1] = (Rij*vj[1] + vi[1])/(Rij + 1.0);
	Hij = (Rij*Hj + Hi)/(Rij + 1.0);
	vm2ij = dot2(vij,vij);
	vnij = dot2(vij,n);
	cij = sqrt((gs.g-1.0)*(Hij - vm2ij*0.5));
}

// Local Lax-Friedrichs (anumericalflux.cpp:40-61)
template <class T> void flux_llf(const Gas& gs, const T* ul, const T* ur, const double* n, T* f)
{
	T vi[2], vj[2], vni, vnj, pi, pj, Hi, Hj;
	vars_from_conserved(gs, ul, n, vi, vni, pi, Hi);
	vars_from_conserved(gs, ur, n, vj, vnj, pj, Hj);
	const T ci = sound_speed(gs, ul[0], pi), cj = sound_speed(gs, ur[0], pj);
	const T eig = fabs(vni)+ci > fabs(vnj)+cj ? fabs(vni)+ci : fabs(vnj)+cj;
	T fr[4];
	directional_flux_from_conserved(gs, ul, n, f);
	directional_flux_from_conserved(gs, ur, n, fr);
	for(int i = 0; i < 4; i++)
		f[i] = 0.5*(f[i] + fr[i] - eig*(ur[i]-ul[i]));
}

// Van Leer flux-vector splitting (anumericalflux.cpp:202-250)
template <class T> void flux_vanleer(const Gas& gs, const T* ul, const T* ur, const double* n,
                                     T* f)
{
	const double g = gs.g;
	T fp[4], fm[4];
	T vi[2], vj[2], vni, vnj, pi, pj, Hi, Hj;
	vars_from_conserved(gs, ul, n, vi, vni, pi, Hi);
	vars_from_conserved(gs, ur, n, vj, vnj, pj, Hj);
	const T ci = sound_speed(gs, ul[0], pi), cj = sound_speed(gs, ur[0], pj);
	const T Mni = vni/ci, Mnj = vnj/cj;
	if(Mni < -1.0) { for(int i = 0; i < 4; i++) fp[i] = 0; }
	else if(Mni > 1.0) directional_flux(ul, n, vni, pi, fp);
	else {
		const T vmags = (ul[1]/ul[0])*(ul[1]/ul[0]) + (ul[2]/ul[0])*(ul[2]/ul[0]);
		fp[0] = ul[0]*ci*((Mni+1.0)*(Mni+1.0))/4.0;
		fp[1] = fp[0] * (ul[1]/ul[0] + n[0]*(2.0*ci - vni)/g);
		fp[2] = fp[0] * (ul[2]/ul[0] + n[1]*(2.0*ci - vni)/g);
		const T t = (g-1.0)*vni + 2.0*ci;
		fp[3] = fp[0] * ((vmags - vni*vni)/2.0 + (t*t)/(2*(g*g-1)));
	}
	if(Mnj > 1.0) { for(int i = 0; i < 4; i++) fm[i] = 0; }
	else if(Mnj < -1.0) directional_flux(ur, n, vnj, pj, fm);
	else {
		const T vmags = (ur[1]/ur[0])*(ur[1]/ur[0]) + (ur[2]/ur[0])*(ur[2]/ur[0]);
		fm[0] = -ur[0]*cj*((Mnj-1.0)*(Mnj-1.0))/4.0;
		fm[1] = fm[0] * (ur[1]/ur[0] + n[0]*(-2.0*cj - vnj)/g);
		fm[2] = fm[0] * (ur[2]/ur[0] + n[1]*(-2.0*cj - vnj)/g);
		const T t = (g-1.0)*vnj - 2.0*cj;
		fm[3] = fm[0] * ((vmags - vnj*vnj)/2.0 + (t*t)/(2*(g*g-1)));
	}
	for(int i = 0; i < 4; i++) f[i] = fp[i] + fm[i];
}

// Liou-Steffen AUSM (anumericalflux.cpp:264-315)
template <class T> void flux_ausm(const Gas& gs, const T* ul, const T* ur, const double* n, T* f)
{
	T ML, MR, pL, pR;
	T vi[2], vj[2], vni, vnj, pi, pj, Hi, Hj;
	vars_from_conserved(gs, ul, n, vi, vni, pi, Hi);
	vars_from_conserved(gs, ur, n, vj, vnj, pj, Hj);
	const T ci = sound_speed(gs, ul[0], pi), cj = sound_speed(gs, ur[0], pj);
	const T Mni = vni/ci, Mnj = vnj/cj;
	if(fabs(Mni) <= 1.0) { ML = 0.25*(Mni+1.0)*(Mni+1.0); pL = ML*pi*(2.0-Mni); }
	else if(Mni < -1.0) { ML = 0; pL = 0; }
	else { ML = Mni; pL = pi; }
	if(fabs(Mnj) <= 1.0) { MR = -0.25*(Mnj-1.0)*(Mnj-1.0); pR = -MR*pj*(2.0+Mnj); }
	else if(Mnj < -1.0) { MR = Mnj; pR = pj; }
	else { MR = 0; pR = 0; }
	const T Mh = ML+MR, ph = pL+pR;
	f[0] = Mh/2.0*(ul[0]*ci+ur[0]*cj) - fabs(Mh)/2.0*(ur[0]*cj-ul[0]*ci);
	for(int j = 1; j < 3; j++)
		f[j] = Mh/2.0*(ul[j]*ci+ur[j]*cj) - fabs(Mh)/2.0*(ur[j]*cj-ul[j]*ci) + ph*n[j-1];
	f[3] = Mh/2.0*(ci*(ul[3]+pi)+cj*(ur[3]+pj)) - fabs(Mh)/2.0*(cj*(ur[3]+pj)-ci*(ul[3]+pi));
}

// AUSM+ (anumericalflux.cpp:479-553)
template <class T> void flux_ausmplus(const Gas& gs, const T* ul, const T* ur, const double* n,
                                      T* f)
{
	const double g = gs.g;
	T ML, MR, pL, pR;
	T vi[2], vj[2], vni, vnj, pi, pj, Hi, Hj;
	vars_from_conserved(gs, ul, n, vi, vni, pi, Hi);
	vars_from_conserved(gs, ur, n, vj, vnj, pj, Hj);
	const T ci = sound_speed(gs, ul[0], pi), cj = sound_speed(gs, ur[0], pj);
	const T vm2i = dot2(vi,vi), vm2j = dot2(vj,vj);
	T csi = sqrt((ci*ci/(g-1.0)+0.5*vm2i)*2.0*(g-1.0)/(g+1.0));
	T csj = sqrt((cj*cj/(g-1.0)+0.5*vm2j)*2.0*(g-1.0)/(g+1.0));
	const T corri = csi > vni ? csi : vni;
	const T corrj = csj > -vnj ? csj : -vnj;
	csi = csi*csi/corri;
	csj = csj*csj/corrj;
	const T ch = csi < csj ? csi : csj;
	const T Mni = vni/ch, Mnj = vnj/ch;
	if(fabs(Mni) <= 1.0) {
		ML = 0.25*(Mni+1.0)*(Mni+1.0) + 1.0/8.0*(Mni*Mni-1.0)*(Mni*Mni-1.0);
		pL = pi*(0.25*(Mni+1.0)*(Mni+1.0)*(2.0-Mni) + 3.0/16*Mni*(Mni*Mni-1.0)*(Mni*Mni-1.0));
	}
	else if(Mni < -1.0) { ML = 0; pL = 0; }
	else { ML = Mni; pL = pi; }
	if(fabs(Mnj) <= 1.0) {
		MR = -0.25*(Mnj-1.0)*(Mnj-1.0) - 1.0/8.0*(Mnj*Mnj-1.0)*(Mnj*Mnj-1.0);
		pR = pj*(0.25*(Mnj-1.0)*(Mnj-1.0)*(2.0+Mnj) - 3.0/16*Mnj*(Mnj*Mnj-1.0)*(Mnj*Mnj-1.0));
	}
	else if(Mnj < -1.0) { MR = Mnj; pR = pj; }
	else { MR = 0; pR = 0; }
	const T Mh = ML+MR, ph = pL+pR;
	f[0] = ch*(Mh/2.0*(ul[0]+ur[0]) - fabs(Mh)/2.0*(ur[0]-ul[0]));
	for(int j = 1; j < 3; j++)
		f[j] = ch*(Mh/2.0*(ul[j]+ur[j]) - fabs(Mh)/2.0*(ur[j]-ul[j])) + ph*n[j-1];
	f[3] = ch*(Mh/2.0*(ul[3]+pi+ur[3]+pj) - fabs(Mh)/2.0*((ur[3]+pj)-(ul[3]+pi)));
}

// Roe-Pike with Harten entropy fix, delta = 1e-4 * c_Roe (anumericalflux.cpp:664,667-732)
template <class T> void flux_roe(const Gas& gs, const T* ul, const T* ur, const double* n, T* f)
{
	const double fixeps = 1.0e-4;
	T vi[2], vj[2], vni, vnj, pi, pj, Hi, Hj;
	vars_from_conserved(gs, ul, n, vi, vni, pi, Hi);
	vars_from_conserved(gs, ur, n, vj, vnj, pj, Hj);
	T Rij, rhoij, vij[2], vm2ij, vnij, Hij, cij;
	roe_averages(gs, ul, ur, n, vi, Hi, vj, Hj, Rij, rhoij, vij, vm2ij, vnij, Hij, cij);

	T l[4];
	l[0] = fabs(vnij-cij);
	l[1] = fabs(vnij);
	l[2] = fabs(vnij);
	l[3] = fabs(vnij+cij);
	const T delta = fixeps*cij;
	for(int k = 0; k < 4; k++)
		if(l[k] < delta)
			l[k] = (l[k]*l[k] + delta*delta)/(2.0*delta);

	const T devn = vnj-vni, dep = pj-pi, derho = ur[0]-ul[0];
	T la[4];
	la[0] = l[0]*(dep-rhoij*cij*devn)/(2.0*cij*cij);
	la[1] = l[1]*(derho - dep/(cij*cij));
	la[2] = l[1]*rhoij;
	la[3] = l[3]*(dep+rhoij*cij*devn)/(2.0*cij*cij);

	T adu[4];
	adu[0] = la[0];
	adu[1] = la[0]*(vij[0]-cij*n[0]);
	adu[2] = la[0]*(vij[1]-cij*n[1]);
	adu[3] = la[0]*(Hij-cij*vnij);

	adu[0] += la[1];
	adu[1] += la[1]*vij[0] + la[2]*(vj[0]-vi[0] - devn*n[0]);
	adu[2] += la[1]*vij[1] + la[2]*(vj[1]-vi[1] - devn*n[1]);
	adu[3] += la[1]*vm2ij/2.0 + la[2]*(vij[0]*(vj[0]-vi[0]) + vij[1]*(vj[1]-vi[1]) - vnij*devn);

	adu[0] += la[3];
	adu[1] += la[3]*(vij[0]+cij*n[0]);
	adu[2] += la[3]*(vij[1]+cij*n[1]);
	adu[3] += la[3]*(Hij+cij*vnij);

	T fi[4], fj[4];
	directional_flux(ul, n, vni, pi, fi);
	directional_flux(ur, n, vnj, pj, fj);
	for(int k = 0; k < 4; k++)
		f[k] = 0.5*(fi[k]+fj[k] - adu[k]);
}

// HLL with Einfeldt speeds (anumericalflux.cpp:973-1007)
template <class T> void flux_hll(const Gas& gs, const T* ul, const T* ur, const double* n, T* f)
{
	T vi[2], vj[2], vni, vnj, pi, pj, Hi, Hj;
	vars_from_conserved(gs, ul, n, vi, vni, pi, Hi);
	vars_from_conserved(gs, ur, n, vj, vnj, pj, Hj);
	const T ci = sound_speed(gs, ul[0], pi), cj = sound_speed(gs, ur[0], pj);
	T Rij, rhoij, vij[2], vm2ij, vnij, Hij, cij;
	roe_averages(gs, ul, ur, n, vi, Hi, vj, Hj, Rij, rhoij, vij, vm2ij, vnij, Hij, cij);
	T sl = vni - ci;
	if(sl > vnij-cij) sl = vnij-cij;
	T sr = vnj + cj;
	if(sr < vnij+cij) sr = vnij+cij;
	const T sr0 = sr > 0 ? T(0) : sr;
	const T sl0 = sl > 0 ? T(0) : sl;
	const T t1 = (sr0 - sl0)/(sr-sl);
	const T t2 = 1.0 - t1;
	const T t3 = 0.5*(sr*fabs(sl)-sl*fabs(sr))/(sr-sl);
	f[0] = t1*vnj*ur[0] + t2*vni*ul[0] - t3*(ur[0]-ul[0]);
	f[1] = t1*(vnj*ur[1]+pj*n[0]) + t2*(vni*ul[1]+pi*n[0]) - t3*(ur[1]-ul[1]);
	f[2] = t1*(vnj*ur[2]+pj*n[1]) + t2*(vni*ul[2]+pi*n[1]) - t3*(ur[2]-ul[2]);
	f[3] = t1*(vnj*ur[0]*Hj) + t2*(vni*ul[0]*Hi) - t3*(ur[3]-ul[3]);
}

// HLLC star state (anumericalflux.cpp:1069-1081)
template <class T> inline void hllc_star(const T* u, const double* n, const T& vn, const T& p,
                                         const T& ss, const T& sm, T* us)
{
	const T pstar = u[0]*(vn-ss)*(vn-sm) + p;
	us[0] = u[0] * (ss - vn)/(ss-sm);
	us[1] = ((ss-vn)*u[1] + (pstar-p)*n[0])/(ss-sm);
	us[2] = ((ss-vn)*u[2] + (pstar-p)*n[1])/(ss-sm);
	us[3] = ((ss-vn)*u[3] - p*vn + pstar*sm)/(ss-sm);
}

// HLLC (Batten et al.) (anumericalflux.cpp:1175-1228)
template <class T> void flux_hllc(const Gas& gs, const T* ul, const T* ur, const double* n, T* f)
{
	T vi[2], vj[2], vni, vnj, pi, pj, Hi, Hj;
	vars_from_conserved(gs, ul, n, vi, vni, pi, Hi);
	vars_from_conserved(gs, ur, n, vj, vnj, pj, Hj);
	const T ci = sound_speed(gs, ul[0], pi), cj = sound_speed(gs, ur[0], pj);
	T Rij, rhoij, vij[2], vm2ij, vnij, Hij, cij;
	roe_averages(gs, ul, ur, n, vi, Hi, vj, Hj, Rij, rhoij, vij, vm2ij, vnij, Hij, cij);
	T sl = vni - ci;
	if(sl > vnij-cij) sl = vnij-cij;
	T sr = vnj + cj;
	if(sr < vnij+cij) sr = vnij+cij;
	const T sm = (ur[0]*vnj*(sr-vnj) - ul[0]*vni*(sl-vni) + pi-pj)
		/ (ur[0]*(sr-vnj) - ul[0]*(sl-vni));
	if(sl > 0)
		directional_flux(ul, n, vni, pi, f);
	else if(sl <= 0 && sm > 0) {
		directional_flux(ul, n, vni, pi, f);
		T us[4];
		hllc_star(ul, n, vni, pi, sl, sm, us);
		for(int k = 0; k < 4; k++) f[k] += sl*(us[k] - ul[k]);
	}
	else if(sm <= 0 && sr >= 0) {
		directional_flux(ur, n, vnj, pj, f);
		T us[4];
		hllc_star(ur, n, vnj, pj, sr, sm, us);
		for(int k = 0; k < 4; k++) f[k] += sr*(us[k] - ur[k]);
	}
	else
		directional_flux(ur, n, vnj, pj, f);
}

template <class T> void flux_dispatch(const Gas& gs, int id, const T* ul, const T* ur,
                                      const double* n, T* f)
{
	switch(id) {
	case FLUX_ROE: flux_roe(gs, ul, ur, n, f); break;
	case FLUX_HLLC: flux_hllc(gs, ul, ur, n, f); break;
	case FLUX_HLL: flux_hll(gs, ul, ur, n, f); break;
	case FLUX_LLF: flux_llf(gs, ul, ur, n, f); break;
	case FLUX_AUSM: flux_ausm(gs, ul, ur, n, f); break;
	case FLUX_AUSMPLUS: flux_ausmplus(gs, ul, ur, n, f); break;
	case FLUX_VANLEER: flux_vanleer(gs, ul, ur, n, f); break;
	default: throw std::runtime_error("unknown flux");
	}
}

/* Analytic Jacobian of the directional flux (src/physics/aphysics.cpp:61-127). */
inline void directional_flux_jacobian(const Gas& gs, const double* u, const double* n,
                                      double* dfdu)
{
	const double g = gs.g;
	dfdu[0] = 0; dfdu[1] = n[0]; dfdu[2] = n[1]; dfdu[3] = 0;
	const double p = pressure_from_conserved(gs, u);
	(void)p;
	double dp[4];
	dp[0] = (g-1.0)*0.5*dot2(&u[1],&u[1])/(u[0]*u[0]);
	dp[1] = -(g-1.0)*u[1]/u[0];
	dp[2] = -(g-1.0)*u[2]/u[0];
	dp[3] = (g-1.0);
	const double vn = dot2(&u[1],n)/u[0];
	double dvn[4];
	dvn[0] = -vn/u[0]; dvn[1] = n[0]/u[0]; dvn[2] = n[1]/u[0]; dvn[3] = 0;
	for(int i = 1; i < 3; i++) {
		dfdu[i*4] = -vn*u[i]/u[0] + dp[0]*n[i-1];
		for(int j = 1; j < 3; j++)
			dfdu[i*4+j] = dvn[j]*u[i] + (i == j ? vn : 0.0) + dp[j]*n[i-1];
		dfdu[i*4+3] = dp[3]*n[i-1];
	}
	dfdu[12] = -vn/u[0]*(u[3]+p) + vn*dp[0];
	for(int j = 1; j < 3; j++)
		dfdu[12+j] = n[j-1]/u[0]*(u[3]+p) + vn*dp[j];
	dfdu[15] = vn*(1.0 + dp[3]);
}

/* -------------------------------------------------------------- boundary states -------- */

/// Per-BC parameters resolved from the configuration (src/spatial/abc.cpp:460-500)
struct BCParams {
	int kind;
	double uinf[4];
	double v0, v1;     ///< bc_vals
};

// Ghost states (src/spatial/abc.cpp:49-84,152-176,194-199,218-226,272-280,354-369,417-423)
template <class T> void bc_ghost(const Gas& gs, const BCParams& bc, const T* ins,
                                 const double* n, T* gsv)
{
	switch(bc.kind) {
	case BC_SLIPWALL: {
		const T vni = dot2(&ins[1],n)/ins[0];
		gsv[0] = ins[0];
		gsv[1] = ins[1] - 2.0*vni*n[0]*ins[0];
		gsv[2] = ins[2] - 2.0*vni*n[1]*ins[0];
		gsv[3] = ins[3];
		break; }
	case BC_FARFIELD:
		for(int i = 0; i < 4; i++) gsv[i] = bc.uinf[i];
		break;
	case BC_INOUTFLOW: {
		const T vni = dot2(&ins[1],n)/ins[0];
		const T ci = sound_speed(gs, ins[0], pressure_from_conserved(gs, ins));
		const T Mni = vni/ci;
		const double pinf = 1.0/(gs.g*gs.Minf*gs.Minf);
		if(Mni <= 0) { for(int i = 0; i < 4; i++) gsv[i] = bc.uinf[i]; }
		else if(Mni < 1) {
			gsv[0] = ins[0]; gsv[1] = ins[1]; gsv[2] = ins[2];
			gsv[3] = pinf/(gs.g-1.0) + 0.5*ins[0]*(dot2(&ins[1],&ins[1])/(ins[0]*ins[0]));
		}
		else { for(int i = 0; i < 4; i++) gsv[i] = ins[i]; }
		break; }
	case BC_SUBSONIC_INFLOW: {
		const double g = gs.g;
		const T ci = sound_speed(gs, ins[0], pressure_from_conserved(gs, ins));
		const T Rminus = dot2(&ins[1],n)/ins[0] - ci/(2*g - 1.0);
		const T co2 = ci*ci + (g-1.0)/2.0 * dot2(&ins[1],&ins[1])/(ins[0]*ins[0]);
		const T q = sqrt((g+1)*co2/((g-1)*Rminus*Rminus) - (g-1)/2.0);
		const T cg = -Rminus*(g-1)/(g+1) * (1.0 + q);
		const T tg = bc.v1*cg*cg/co2;
		const T pg = bc.v0 * pow(tg/bc.v1, g/(g-1.0));
		gsv[0] = g*gs.Minf*gs.Minf*pg/tg;
		const T vgmag = sqrt(2.0/(g-1.0)*(co2 - cg*cg));
		const T vx = vgmag*1.0*n[0], vy = vgmag*1.0*n[1];
		gsv[1] = gsv[0]*vx;
		gsv[2] = gsv[0]*vy;
		gsv[3] = pg/(g-1.0) + 0.5*gsv[0]*(vgmag*vgmag);
		break; }
	case BC_EXTRAPOLATION:
		for(int i = 0; i < 4; i++) gsv[i] = ins[i];
		break;
	case BC_ADIABATICWALL: {
		const T tm = bc.v0 * ins[0];
		gsv[0] = ins[0];
		gsv[1] = 2.0*tm*n[1] - ins[1];
		gsv[2] = -2.0*tm*n[0] - ins[2];
		gsv[3] = ins[3];
		break; }
	case BC_ISOTHERMALWALL: {
		const T p = pressure_from_conserved(gs, ins);
		const T gtemp = 2.0*bc.v1 - temperature(gs, ins[0], p);
		gsv[0] = ins[0];
		gsv[1] = gsv[0]*(2.0*bc.v0*n[1] - ins[1]/ins[0]);
		gsv[2] = gsv[0]*(-2.0*bc.v0*n[0] - ins[2]/ins[0]);
		const T vmag2 = dot2(&gsv[1],&gsv[1])/(gsv[0]*gsv[0]);
		gsv[3] = gsv[0]*(gtemp/(gs.g*(gs.g-1.0)*gs.Minf*gs.Minf) + 0.5*vmag2);
		break; }
	default:
		throw std::runtime_error("BC type not implemented");
	}
}

}
#endif
