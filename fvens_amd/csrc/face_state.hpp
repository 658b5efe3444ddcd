/* face_state.hpp -- division-lean numerical fluxes for the residual kernels.
 *
 * Each side of a face is converted once into a FaceState holding the primitive and conserved
 * forms, 1/rho, the normal velocity, the speed of sound and the total enthalpy; the fluxes
 * below read these instead of re-deriving velocities and pressure from conserved variables.
 * One reciprocal and one square root per side replace the reference's repeated divisions
 * (src/physics/aphysics_defs.hpp:27-38 evaluated per flux); the formulas are those of
 * src/spatial/anumericalflux.cpp (cited per routine) and agree with the conserved-input forms
 * in gas.hpp to rounding.
 */
#ifndef FVX_FACE_STATE_HPP
#define FVX_FACE_STATE_HPP

#include "gas.hpp"

namespace fvx {

struct FaceState {
	double r, ir;            ///< density and its reciprocal
	double vx, vy, p;        ///< velocity, pressure
	double mx, my, E;        ///< momentum and total energy per unit volume
	double vn, c, H;         ///< normal velocity, speed of sound, total enthalpy
};

__device__ __forceinline__ void fs_finish(const GasParams& gp, const double* n, FaceState& s)
{
	s.vn = s.vx*n[0] + s.vy*n[1];
	s.c = sqrt(gp.g*s.p*s.ir);
	s.H = (s.E + s.p)*s.ir;
}

/// From primitive (rho, vx, vy, p).
__device__ __forceinline__ FaceState fs_prim(const GasParams& gp, const double* w,
                                             const double* n, double inv_gm1)
{
	FaceState s;
	s.r = w[0]; s.ir = 1.0/w[0];
	s.vx = w[1]; s.vy = w[2]; s.p = w[3];
	s.mx = w[0]*w[1]; s.my = w[0]*w[2];
	s.E = w[3]*inv_gm1 + 0.5*w[0]*(w[1]*w[1] + w[2]*w[2]);
	fs_finish(gp, n, s);
	return s;
}

/// From conserved (rho, rho vx, rho vy, rho E).
__device__ __forceinline__ FaceState fs_cons(const GasParams& gp, const double* u,
                                             const double* n)
{
	FaceState s;
	s.r = u[0]; s.ir = 1.0/u[0];
	s.mx = u[1]; s.my = u[2]; s.E = u[3];
	s.vx = u[1]*s.ir; s.vy = u[2]*s.ir;
	s.p = gp.gm1*(u[3] - 0.5*(u[1]*s.vx + u[2]*s.vy));
	fs_finish(gp, n, s);
	return s;
}

/// Primitive variables of a conserved cell state with one reciprocal; returns 1/rho.
__device__ __forceinline__ double prim_lean(const GasParams& gp, const double* u, double* w)
{
	const double ir = 1.0/u[0];
	w[0] = u[0];
	w[1] = u[1]*ir;
	w[2] = u[2]*ir;
	w[3] = gp.gm1*(u[3] - 0.5*(u[1]*w[1] + u[2]*w[2]));
	return ir;
}

/// From a conserved state whose primitive form and 1/rho are already known.
__device__ __forceinline__ FaceState fs_both(const GasParams& gp, const double* u,
                                             const double* w, double ir, const double* n)
{
	FaceState s;
	s.r = u[0]; s.ir = ir;
	s.mx = u[1]; s.my = u[2]; s.E = u[3];
	s.vx = w[1]; s.vy = w[2]; s.p = w[3];
	fs_finish(gp, n, s);
	return s;
}

__device__ __forceinline__ void fs_flux(const FaceState& s, const double* n, double* f)
{
	f[0] = s.vn*s.r;
	f[1] = s.vn*s.mx + s.p*n[0];
	f[2] = s.vn*s.my + s.p*n[1];
	f[3] = s.vn*(s.E + s.p);
}

struct RoeMean { double rho, vx, vy, vn, v2, H, c; };

__device__ __forceinline__ RoeMean roe_mean(const GasParams& gp, const FaceState& a,
                                            const FaceState& b, const double* n)
{
	RoeMean m;
	const double R = sqrt(b.r*a.ir);
	const double iR1 = 1.0/(R + 1.0);
	m.rho = R*a.r;
	m.vx = (R*b.vx + a.vx)*iR1;
	m.vy = (R*b.vy + a.vy)*iR1;
	m.H = (R*b.H + a.H)*iR1;
	m.v2 = m.vx*m.vx + m.vy*m.vy;
	m.vn = m.vx*n[0] + m.vy*n[1];
	m.c = sqrt(gp.gm1*(m.H - 0.5*m.v2));
	return m;
}

// Roe with Harten's entropy fix (anumericalflux.cpp:667-732)
__device__ __forceinline__ void fsf_roe(const GasParams& gp, const FaceState& a,
                                        const FaceState& b, const double* n, double* f)
{
	const RoeMean m = roe_mean(gp, a, b, n);
	double l0 = fabs(m.vn - m.c), l1 = fabs(m.vn), l3 = fabs(m.vn + m.c);
	const double delta = 1.0e-4*m.c;
	const double d2 = delta*delta, i2d = 0.5/delta;
	if(l0 < delta) l0 = (l0*l0 + d2)*i2d;
	if(l1 < delta) l1 = (l1*l1 + d2)*i2d;
	if(l3 < delta) l3 = (l3*l3 + d2)*i2d;
	const double dvn = b.vn - a.vn, dp = b.p - a.p, drho = b.r - a.r;
	const double ic2 = 1.0/(m.c*m.c);
	const double rc = m.rho*m.c;
	const double a0 = l0*(dp - rc*dvn)*0.5*ic2;
	const double a1 = l1*(drho - dp*ic2);
	const double a2 = l1*m.rho;
	const double a3 = l3*(dp + rc*dvn)*0.5*ic2;
	const double dvx = b.vx - a.vx, dvy = b.vy - a.vy;
	double fa[4], fb[4];
	fs_flux(a, n, fa);
	fs_flux(b, n, fb);
	f[0] = 0.5*(fa[0] + fb[0] - (a0 + a1 + a3));
	f[1] = 0.5*(fa[1] + fb[1] - (a0*(m.vx - m.c*n[0]) + a1*m.vx + a2*(dvx - dvn*n[0])
	                             + a3*(m.vx + m.c*n[0])));
	f[2] = 0.5*(fa[2] + fb[2] - (a0*(m.vy - m.c*n[1]) + a1*m.vy + a2*(dvy - dvn*n[1])
	                             + a3*(m.vy + m.c*n[1])));
	f[3] = 0.5*(fa[3] + fb[3] - (a0*(m.H - m.c*m.vn) + a1*m.v2*0.5
	                             + a2*(m.vx*dvx + m.vy*dvy - m.vn*dvn) + a3*(m.H + m.c*m.vn)));
}

// HLLC (anumericalflux.cpp:1069-1081, 1175-1228)
__device__ __forceinline__ void fsf_hllc(const GasParams& gp, const FaceState& a,
                                         const FaceState& b, const double* n, double* f)
{
	const RoeMean m = roe_mean(gp, a, b, n);
	double sl = a.vn - a.c;
	if(sl > m.vn - m.c) sl = m.vn - m.c;
	double sr = b.vn + b.c;
	if(sr < m.vn + m.c) sr = m.vn + m.c;
	const double sm = (b.r*b.vn*(sr - b.vn) - a.r*a.vn*(sl - a.vn) + a.p - b.p)
	                  / (b.r*(sr - b.vn) - a.r*(sl - a.vn));
	bool left, star;
	if(sl > 0.0) { left = true; star = false; }
	else if(sl <= 0.0 && sm > 0.0) { left = true; star = true; }
	else if(sm <= 0.0 && sr >= 0.0) { left = false; star = true; }
	else { left = false; star = false; }
	const FaceState& s = left ? a : b;
	const double ss = left ? sl : sr;
	fs_flux(s, n, f);
	if(star) {
		const double pstar = s.r*(s.vn - ss)*(s.vn - sm) + s.p;
		const double idn = 1.0/(ss - sm);
		const double d = ss - s.vn;
		const double us0 = s.r*d*idn;
		const double us1 = (d*s.mx + (pstar - s.p)*n[0])*idn;
		const double us2 = (d*s.my + (pstar - s.p)*n[1])*idn;
		const double us3 = (d*s.E - s.p*s.vn + pstar*sm)*idn;
		f[0] += ss*(us0 - s.r);
		f[1] += ss*(us1 - s.mx);
		f[2] += ss*(us2 - s.my);
		f[3] += ss*(us3 - s.E);
	}
}

// HLL with Einfeldt speeds (anumericalflux.cpp:973-1007)
__device__ __forceinline__ void fsf_hll(const GasParams& gp, const FaceState& a,
                                        const FaceState& b, const double* n, double* f)
{
	const RoeMean m = roe_mean(gp, a, b, n);
	double sl = a.vn - a.c;
	if(sl > m.vn - m.c) sl = m.vn - m.c;
	double sr = b.vn + b.c;
	if(sr < m.vn + m.c) sr = m.vn + m.c;
	const double sr0 = sr > 0.0 ? 0.0 : sr;
	const double sl0 = sl > 0.0 ? 0.0 : sl;
	const double idn = 1.0/(sr - sl);
	const double t1 = (sr0 - sl0)*idn, t2 = 1.0 - t1;
	const double t3 = 0.5*(sr*fabs(sl) - sl*fabs(sr))*idn;
	f[0] = t1*b.vn*b.r + t2*a.vn*a.r - t3*(b.r - a.r);
	f[1] = t1*(b.vn*b.mx + b.p*n[0]) + t2*(a.vn*a.mx + a.p*n[0]) - t3*(b.mx - a.mx);
	f[2] = t1*(b.vn*b.my + b.p*n[1]) + t2*(a.vn*a.my + a.p*n[1]) - t3*(b.my - a.my);
	f[3] = t1*(b.vn*b.r*b.H) + t2*(a.vn*a.r*a.H) - t3*(b.E - a.E);
}

// Local Lax-Friedrichs (anumericalflux.cpp:40-61)
__device__ __forceinline__ void fsf_llf(const GasParams& gp, const FaceState& a,
                                        const FaceState& b, const double* n, double* f)
{
	const double ea = fabs(a.vn) + a.c, eb = fabs(b.vn) + b.c;
	const double eig = ea > eb ? ea : eb;
	double fa[4], fb[4];
	fs_flux(a, n, fa);
	fs_flux(b, n, fb);
	f[0] = 0.5*(fa[0] + fb[0] - eig*(b.r - a.r));
	f[1] = 0.5*(fa[1] + fb[1] - eig*(b.mx - a.mx));
	f[2] = 0.5*(fa[2] + fb[2] - eig*(b.my - a.my));
	f[3] = 0.5*(fa[3] + fb[3] - eig*(b.E - a.E));
}

// Liou-Steffen AUSM (anumericalflux.cpp:264-315)
__device__ __forceinline__ void fsf_ausm(const GasParams& gp, const FaceState& a,
                                         const FaceState& b, const double* n, double* f)
{
	const double Ma = a.vn/a.c, Mb = b.vn/b.c;
	double ML, MR, pL, pR;
	if(fabs(Ma) <= 1.0) { ML = 0.25*(Ma + 1.0)*(Ma + 1.0); pL = ML*a.p*(2.0 - Ma); }
	else if(Ma < -1.0) { ML = 0.0; pL = 0.0; }
	else { ML = Ma; pL = a.p; }
	if(fabs(Mb) <= 1.0) { MR = -0.25*(Mb - 1.0)*(Mb - 1.0); pR = -MR*b.p*(2.0 + Mb); }
	else if(Mb < -1.0) { MR = Mb; pR = b.p; }
	else { MR = 0.0; pR = 0.0; }
	const double Mh = ML + MR, ph = pL + pR;
	const double h = 0.5*Mh, ah = 0.5*fabs(Mh);
	f[0] = h*(a.r*a.c + b.r*b.c) - ah*(b.r*b.c - a.r*a.c);
	f[1] = h*(a.mx*a.c + b.mx*b.c) - ah*(b.mx*b.c - a.mx*a.c) + ph*n[0];
	f[2] = h*(a.my*a.c + b.my*b.c) - ah*(b.my*b.c - a.my*a.c) + ph*n[1];
	f[3] = h*(a.c*(a.E + a.p) + b.c*(b.E + b.p)) - ah*(b.c*(b.E + b.p) - a.c*(a.E + a.p));
}

// AUSM+ (anumericalflux.cpp:479-553)
__device__ __forceinline__ void fsf_ausmplus(const GasParams& gp, const FaceState& a,
                                             const FaceState& b, const double* n, double* f)
{
	const double g = gp.g;
	const double fac = 2.0*(g - 1.0)/(g + 1.0);
	double csa = sqrt((a.c*a.c/(g - 1.0) + 0.5*(a.vx*a.vx + a.vy*a.vy))*fac);
	double csb = sqrt((b.c*b.c/(g - 1.0) + 0.5*(b.vx*b.vx + b.vy*b.vy))*fac);
	const double ka = csa > a.vn ? csa : a.vn;
	const double kb = csb > -b.vn ? csb : -b.vn;
	csa = csa*csa/ka;
	csb = csb*csb/kb;
	const double ch = csa < csb ? csa : csb;
	const double ich = 1.0/ch;
	const double Ma = a.vn*ich, Mb = b.vn*ich;
	double ML, MR, pL, pR;
	if(fabs(Ma) <= 1.0) {
		const double q = (Ma*Ma - 1.0)*(Ma*Ma - 1.0);
		ML = 0.25*(Ma + 1.0)*(Ma + 1.0) + 0.125*q;
		pL = a.p*(0.25*(Ma + 1.0)*(Ma + 1.0)*(2.0 - Ma) + 0.1875*Ma*q);
	}
	else if(Ma < -1.0) { ML = 0.0; pL = 0.0; }
	else { ML = Ma; pL = a.p; }
	if(fabs(Mb) <= 1.0) {
		const double q = (Mb*Mb - 1.0)*(Mb*Mb - 1.0);
		MR = -0.25*(Mb - 1.0)*(Mb - 1.0) - 0.125*q;
		pR = b.p*(0.25*(Mb - 1.0)*(Mb - 1.0)*(2.0 + Mb) - 0.1875*Mb*q);
	}
	else if(Mb < -1.0) { MR = Mb; pR = b.p; }
	else { MR = 0.0; pR = 0.0; }
	const double Mh = ML + MR, ph = pL + pR;
	const double h = 0.5*Mh, ah = 0.5*fabs(Mh);
	f[0] = ch*(h*(a.r + b.r) - ah*(b.r - a.r));
	f[1] = ch*(h*(a.mx + b.mx) - ah*(b.mx - a.mx)) + ph*n[0];
	f[2] = ch*(h*(a.my + b.my) - ah*(b.my - a.my)) + ph*n[1];
	f[3] = ch*(h*(a.E + a.p + b.E + b.p) - ah*((b.E + b.p) - (a.E + a.p)));
}

// Van Leer flux-vector splitting (anumericalflux.cpp:202-250)
__device__ __forceinline__ void fsf_vanleer(const GasParams& gp, const FaceState& a,
                                            const FaceState& b, const double* n, double* f)
{
	const double g = gp.g;
	const double Ma = a.vn/a.c, Mb = b.vn/b.c;
	double fp[4], fm[4];
	if(Ma < -1.0) { for(int k = 0; k < 4; k++) fp[k] = 0.0; }
	else if(Ma > 1.0) fs_flux(a, n, fp);
	else {
		fp[0] = a.r*a.c*(Ma + 1.0)*(Ma + 1.0)*0.25;
		fp[1] = fp[0]*(a.vx + n[0]*(2.0*a.c - a.vn)/g);
		fp[2] = fp[0]*(a.vy + n[1]*(2.0*a.c - a.vn)/g);
		const double t = (g - 1.0)*a.vn + 2.0*a.c;
		fp[3] = fp[0]*((a.vx*a.vx + a.vy*a.vy - a.vn*a.vn)*0.5 + t*t/(2.0*(g*g - 1.0)));
	}
	if(Mb > 1.0) { for(int k = 0; k < 4; k++) fm[k] = 0.0; }
	else if(Mb < -1.0) fs_flux(b, n, fm);
	else {
		fm[0] = -b.r*b.c*(Mb - 1.0)*(Mb - 1.0)*0.25;
		fm[1] = fm[0]*(b.vx + n[0]*(-2.0*b.c - b.vn)/g);
		fm[2] = fm[0]*(b.vy + n[1]*(-2.0*b.c - b.vn)/g);
		const double t = (g - 1.0)*b.vn - 2.0*b.c;
		fm[3] = fm[0]*((b.vx*b.vx + b.vy*b.vy - b.vn*b.vn)*0.5 + t*t/(2.0*(g*g - 1.0)));
	}
	for(int k = 0; k < 4; k++) f[k] = fp[k] + fm[k];
}

template <int FLUX>
__device__ __forceinline__ void fs_num_flux(const GasParams& gp, const FaceState& a,
                                            const FaceState& b, const double* n, double* f)
{
	if constexpr(FLUX == 0) fsf_roe(gp, a, b, n, f);
	else if constexpr(FLUX == 1) fsf_hllc(gp, a, b, n, f);
	else if constexpr(FLUX == 2) fsf_hll(gp, a, b, n, f);
	else if constexpr(FLUX == 3) fsf_llf(gp, a, b, n, f);
	else if constexpr(FLUX == 4) fsf_ausm(gp, a, b, n, f);
	else if constexpr(FLUX == 5) fsf_ausmplus(gp, a, b, n, f);
	else fsf_vanleer(gp, a, b, n, f);
}

}
#endif
