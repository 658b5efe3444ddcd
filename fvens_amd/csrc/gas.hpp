/* gas.hpp -- ideal-gas physics, numerical fluxes and boundary ghost states for the device
 * kernels (also usable on the host). Non-dimensionalisation and formulas follow
 * IdealGasPhysics (src/physics/aphysics_defs.hpp:13-487); the fluxes follow
 * src/spatial/anumericalflux.cpp (cited per routine); ghost states follow
 * src/spatial/abc.cpp. Routines are templated on the scalar so that the same code yields
 * forward-mode derivatives for the Jacobian kernels.
 */
#ifndef FVX_GAS_HPP
#define FVX_GAS_HPP

#include <hip/hip_runtime.h>
#include <cmath>

#define FVX_HD __host__ __device__ __forceinline__

namespace fvx {

/// Gas constants used by the kernels (passed by value as kernel arguments).
struct GasParams {
	double g, gm1;          ///< gamma, gamma-1
	double Minf, Tinf, Reinf, Pr;
	double gM2;             ///< gamma*Minf^2 (temperature factor)
	double sutherland;      ///< S / Tinf with S = 110.5 K (src/physics/aphysics.cpp:19)
	double pinf;            ///< 1/(gamma Minf^2)
	double uinf[4];
	int constvisc;
};

/// Boundary-condition table: one row per configured BC (kind from the reference enum).
struct BCTable {
	int kind[8];
	double v0[8], v1[8];
};

template <class T> FVX_HD T sq(const T& a) { return a*a; }

/// Pressure from conserved variables (aphysics_defs.hpp:53-63).
template <class T> FVX_HD T pressure(const GasParams& gp, const T* u)
{
	return gp.gm1*(u[3] - 0.5*(u[1]*u[1] + u[2]*u[2])/u[0]);
}

/// Conserved -> primitive (rho, vx, vy, p) (aphysics_defs.hpp:257-267).
template <class T> FVX_HD void to_primitive(const GasParams& gp, const T* u, T* w)
{
	const T p = pressure(gp, u);
	w[1] = u[1]/u[0];
	w[2] = u[2]/u[0];
	w[0] = u[0];
	w[3] = p;
}

/// Primitive -> conserved (aphysics_defs.hpp:238-244, 285-295).
template <class T> FVX_HD void to_conserved(const GasParams& gp, const T* w, T* u)
{
	const T v2 = w[1]*w[1] + w[2]*w[2];
	u[0] = w[0];
	u[1] = w[0]*w[1];
	u[2] = w[0]*w[2];
	u[3] = w[3]/gp.gm1 + 0.5*w[0]*v2;
}

/// Sutherland viscosity coefficient divided by Re (aphysics_defs.hpp:408-421).
template <class T> FVX_HD T sutherland(const GasParams& gp, const T& Tem)
{
	using std::pow;
	return (1.0 + gp.sutherland)/(Tem + gp.sutherland) * pow(Tem, 1.5) / gp.Reinf;
}

template <class T> FVX_HD T viscosity(const GasParams& gp, const T* u)
{
	if(gp.constvisc) return T(1.0/gp.Reinf);
	const T Tem = pressure(gp, u)/u[0] * gp.gM2;
	return sutherland(gp, Tem);
}

/// One-sided quantities needed by every flux (getVarsFromConserved, aphysics_defs.hpp:27-38).
template <class T> struct SideVars {
	T v[2], vn, p, H;
};

template <class T> FVX_HD SideVars<T> side_vars(const GasParams& gp, const T* u, const double* n)
{
	SideVars<T> s;
	s.v[0] = u[1]/u[0];
	s.v[1] = u[2]/u[0];
	s.vn = s.v[0]*n[0] + s.v[1]*n[1];
	const T v2 = s.v[0]*s.v[0] + s.v[1]*s.v[1];
	s.p = gp.gm1*(u[3] - 0.5*u[0]*v2);
	s.H = (u[3] + s.p)/u[0];
	return s;
}

template <class T> FVX_HD void phys_flux(const T* u, const double* n, const T& vn, const T& p,
                                         T* f)
{
	f[0] = vn*u[0];
	f[1] = vn*u[1] + p*n[0];
	f[2] = vn*u[2] + p*n[1];
	f[3] = vn*(u[3] + p);
}

template <class T> FVX_HD T tabs(const T& a) { return a < 0.0 ? -a : a; }
template <class T> FVX_HD T tsqrt(const T& a) { using std::sqrt; return sqrt(a); }

/// Roe averages (anumericalflux.hpp:175-189).
template <class T> struct RoeAvg { T rho, v[2], vn, v2, H, c; };

template <class T> FVX_HD RoeAvg<T> roe_avg(const GasParams& gp, const T* ul, const T* ur,
                                            const double* n, const SideVars<T>& a,
                                            const SideVars<T>& b)
{
	RoeAvg<T> r;
	const T R = tsqrt(ur[0]/ul[0]);
	const T inv = 1.0/(R + 1.0);
	r.rho = R*ul[0];
	r.v[0] = (R*b.v[0] + a.v[0])*inv;
	r.v[1] = (R*b.v[1] + a.v[1])*inv;
	r.H = (R*b.H + a.H)*inv;
	r.v2 = r.v[0]*r.v[0] + r.v[1]*r.v[1];
	r.vn = r.v[0]*n[0] + r.v[1]*n[1];
	r.c = tsqrt(gp.gm1*(r.H - 0.5*r.v2));
	return r;
}

/// Roe flux with Harten's entropy fix, delta = 1e-4 c (anumericalflux.cpp:664,667-732).
template <class T> FVX_HD void flux_roe(const GasParams& gp, const T* ul, const T* ur,
                                        const double* n, T* f)
{
	const SideVars<T> a = side_vars(gp, ul, n), b = side_vars(gp, ur, n);
	const RoeAvg<T> m = roe_avg(gp, ul, ur, n, a, b);
	T l0 = tabs(m.vn - m.c), l1 = tabs(m.vn), l3 = tabs(m.vn + m.c);
	const T delta = 1.0e-4*m.c;
	const T d2 = delta*delta, i2d = 0.5/delta;
	if(l0 < delta) l0 = (l0*l0 + d2)*i2d;
	if(l1 < delta) l1 = (l1*l1 + d2)*i2d;
	if(l3 < delta) l3 = (l3*l3 + d2)*i2d;
	const T dvn = b.vn - a.vn, dp = b.p - a.p, drho = ur[0] - ul[0];
	const T ic2 = 1.0/(m.c*m.c);
	const T rc = m.rho*m.c;
	const T a0 = l0*(dp - rc*dvn)*0.5*ic2;
	const T a1 = l1*(drho - dp*ic2);
	const T a2 = l1*m.rho;
	const T a3 = l3*(dp + rc*dvn)*0.5*ic2;
	const T dvx = b.v[0] - a.v[0], dvy = b.v[1] - a.v[1];
	T ad[4];
	ad[0] = a0 + a1 + a3;
	ad[1] = a0*(m.v[0] - m.c*n[0]) + a1*m.v[0] + a2*(dvx - dvn*n[0]) + a3*(m.v[0] + m.c*n[0]);
	ad[2] = a0*(m.v[1] - m.c*n[1]) + a1*m.v[1] + a2*(dvy - dvn*n[1]) + a3*(m.v[1] + m.c*n[1]);
	ad[3] = a0*(m.H - m.c*m.vn) + a1*m.v2*0.5 + a2*(m.v[0]*dvx + m.v[1]*dvy - m.vn*dvn)
	        + a3*(m.H + m.c*m.vn);
	T fa[4], fb[4];
	phys_flux(ul, n, a.vn, a.p, fa);
	phys_flux(ur, n, b.vn, b.p, fb);
	for(int k = 0; k < 4; k++) f[k] = 0.5*(fa[k] + fb[k] - ad[k]);
}

/// Einfeldt signal speeds shared by HLL and HLLC (anumericalflux.cpp:990-996, 1191-1197).
template <class T> FVX_HD void einfeldt(const T& vna, const T& ca, const T& vnb, const T& cb,
                                        const RoeAvg<T>& m, T& sl, T& sr)
{
	sl = vna - ca;
	if(sl > m.vn - m.c) sl = m.vn - m.c;
	sr = vnb + cb;
	if(sr < m.vn + m.c) sr = m.vn + m.c;
}

/// HLLC (Batten et al.) (anumericalflux.cpp:1069-1081, 1175-1228).
template <class T> FVX_HD void flux_hllc(const GasParams& gp, const T* ul, const T* ur,
                                         const double* n, T* f)
{
	const SideVars<T> a = side_vars(gp, ul, n), b = side_vars(gp, ur, n);
	const T ca = tsqrt(gp.g*a.p/ul[0]), cb = tsqrt(gp.g*b.p/ur[0]);
	const RoeAvg<T> m = roe_avg(gp, ul, ur, n, a, b);
	T sl, sr;
	einfeldt(a.vn, ca, b.vn, cb, m, sl, sr);
	const T sm = (ur[0]*b.vn*(sr - b.vn) - ul[0]*a.vn*(sl - a.vn) + a.p - b.p)
	             / (ur[0]*(sr - b.vn) - ul[0]*(sl - a.vn));
	// the reference's four cases: left, left star, right star, right
	bool left, star;
	if(sl > 0.0) { left = true; star = false; }
	else if(sl <= 0.0 && sm > 0.0) { left = true; star = true; }
	else if(sm <= 0.0 && sr >= 0.0) { left = false; star = true; }
	else { left = false; star = false; }
	const T* u = left ? ul : ur;
	const SideVars<T>& s = left ? a : b;
	const T ss = left ? sl : sr;
	phys_flux(u, n, s.vn, s.p, f);
	if(star) {
		const T pstar = u[0]*(s.vn - ss)*(s.vn - sm) + s.p;
		const T idn = 1.0/(ss - sm);
		T us[4];
		us[0] = u[0]*(ss - s.vn)*idn;
		us[1] = ((ss - s.vn)*u[1] + (pstar - s.p)*n[0])*idn;
		us[2] = ((ss - s.vn)*u[2] + (pstar - s.p)*n[1])*idn;
		us[3] = ((ss - s.vn)*u[3] - s.p*s.vn + pstar*sm)*idn;
		for(int k = 0; k < 4; k++) f[k] += ss*(us[k] - u[k]);
	}
}

/// HLL with Einfeldt speeds (anumericalflux.cpp:973-1007).
template <class T> FVX_HD void flux_hll(const GasParams& gp, const T* ul, const T* ur,
                                        const double* n, T* f)
{
	const SideVars<T> a = side_vars(gp, ul, n), b = side_vars(gp, ur, n);
	const T ca = tsqrt(gp.g*a.p/ul[0]), cb = tsqrt(gp.g*b.p/ur[0]);
	const RoeAvg<T> m = roe_avg(gp, ul, ur, n, a, b);
	T sl, sr;
	einfeldt(a.vn, ca, b.vn, cb, m, sl, sr);
	const T sr0 = sr > 0.0 ? T(0.0) : sr;
	const T sl0 = sl > 0.0 ? T(0.0) : sl;
	const T idn = 1.0/(sr - sl);
	const T t1 = (sr0 - sl0)*idn;
	const T t2 = 1.0 - t1;
	const T t3 = 0.5*(sr*tabs(sl) - sl*tabs(sr))*idn;
	f[0] = t1*b.vn*ur[0] + t2*a.vn*ul[0] - t3*(ur[0] - ul[0]);
	f[1] = t1*(b.vn*ur[1] + b.p*n[0]) + t2*(a.vn*ul[1] + a.p*n[0]) - t3*(ur[1] - ul[1]);
	f[2] = t1*(b.vn*ur[2] + b.p*n[1]) + t2*(a.vn*ul[2] + a.p*n[1]) - t3*(ur[2] - ul[2]);
	f[3] = t1*(b.vn*ur[0]*b.H) + t2*(a.vn*ul[0]*a.H) - t3*(ur[3] - ul[3]);
}

/// Local Lax-Friedrichs (anumericalflux.cpp:40-61).
template <class T> FVX_HD void flux_llf(const GasParams& gp, const T* ul, const T* ur,
                                        const double* n, T* f)
{
	const SideVars<T> a = side_vars(gp, ul, n), b = side_vars(gp, ur, n);
	const T ca = tsqrt(gp.g*a.p/ul[0]), cb = tsqrt(gp.g*b.p/ur[0]);
	const T ea = tabs(a.vn) + ca, eb = tabs(b.vn) + cb;
	const T eig = ea > eb ? ea : eb;
	// analytic fluxes use the conserved-form pressure (aphysics.cpp:28-35)
	T fa[4], fb[4];
	const T vna = (ul[1]*n[0] + ul[2]*n[1])/ul[0], vnb = (ur[1]*n[0] + ur[2]*n[1])/ur[0];
	phys_flux(ul, n, vna, pressure(gp, ul), fa);
	phys_flux(ur, n, vnb, pressure(gp, ur), fb);
	for(int k = 0; k < 4; k++) f[k] = 0.5*(fa[k] + fb[k] - eig*(ur[k] - ul[k]));
}

/// Liou-Steffen AUSM (anumericalflux.cpp:264-315).
template <class T> FVX_HD void flux_ausm(const GasParams& gp, const T* ul, const T* ur,
                                         const double* n, T* f)
{
	const SideVars<T> a = side_vars(gp, ul, n), b = side_vars(gp, ur, n);
	const T ca = tsqrt(gp.g*a.p/ul[0]), cb = tsqrt(gp.g*b.p/ur[0]);
	const T Ma = a.vn/ca, Mb = b.vn/cb;
	T ML, MR, pL, pR;
	if(tabs(Ma) <= 1.0) { ML = 0.25*(Ma + 1.0)*(Ma + 1.0); pL = ML*a.p*(2.0 - Ma); }
	else if(Ma < -1.0) { ML = 0.0; pL = 0.0; }
	else { ML = Ma; pL = a.p; }
	if(tabs(Mb) <= 1.0) { MR = -0.25*(Mb - 1.0)*(Mb - 1.0); pR = -MR*b.p*(2.0 + Mb); }
	else if(Mb < -1.0) { MR = Mb; pR = b.p; }
	else { MR = 0.0; pR = 0.0; }
	const T Mh = ML + MR, ph = pL + pR;
	const T h = 0.5*Mh, ah = 0.5*tabs(Mh);
	f[0] = h*(ul[0]*ca + ur[0]*cb) - ah*(ur[0]*cb - ul[0]*ca);
	f[1] = h*(ul[1]*ca + ur[1]*cb) - ah*(ur[1]*cb - ul[1]*ca) + ph*n[0];
	f[2] = h*(ul[2]*ca + ur[2]*cb) - ah*(ur[2]*cb - ul[2]*ca) + ph*n[1];
	f[3] = h*(ca*(ul[3] + a.p) + cb*(ur[3] + b.p)) - ah*(cb*(ur[3] + b.p) - ca*(ul[3] + a.p));
}

/// AUSM+ (anumericalflux.cpp:479-553).
template <class T> FVX_HD void flux_ausmplus(const GasParams& gp, const T* ul, const T* ur,
                                             const double* n, T* f)
{
	const double g = gp.g;
	const SideVars<T> a = side_vars(gp, ul, n), b = side_vars(gp, ur, n);
	const T ca = tsqrt(g*a.p/ul[0]), cb = tsqrt(g*b.p/ur[0]);
	const T v2a = a.v[0]*a.v[0] + a.v[1]*a.v[1], v2b = b.v[0]*b.v[0] + b.v[1]*b.v[1];
	const double fac = 2.0*(g - 1.0)/(g + 1.0);
	T csa = tsqrt((ca*ca/(g - 1.0) + 0.5*v2a)*fac);
	T csb = tsqrt((cb*cb/(g - 1.0) + 0.5*v2b)*fac);
	const T ka = csa > a.vn ? csa : a.vn;
	const T kb = csb > -b.vn ? csb : -b.vn;
	csa = csa*csa/ka;
	csb = csb*csb/kb;
	const T ch = csa < csb ? csa : csb;
	const T Ma = a.vn/ch, Mb = b.vn/ch;
	T ML, MR, pL, pR;
	if(tabs(Ma) <= 1.0) {
		const T q = (Ma*Ma - 1.0)*(Ma*Ma - 1.0);
		ML = 0.25*(Ma + 1.0)*(Ma + 1.0) + 0.125*q;
		pL = a.p*(0.25*(Ma + 1.0)*(Ma + 1.0)*(2.0 - Ma) + 0.1875*Ma*q);
	}
	else if(Ma < -1.0) { ML = 0.0; pL = 0.0; }
	else { ML = Ma; pL = a.p; }
	if(tabs(Mb) <= 1.0) {
		const T q = (Mb*Mb - 1.0)*(Mb*Mb - 1.0);
		MR = -0.25*(Mb - 1.0)*(Mb - 1.0) - 0.125*q;
		pR = b.p*(0.25*(Mb - 1.0)*(Mb - 1.0)*(2.0 + Mb) - 0.1875*Mb*q);
	}
	else if(Mb < -1.0) { MR = Mb; pR = b.p; }
	else { MR = 0.0; pR = 0.0; }
	const T Mh = ML + MR, ph = pL + pR;
	const T h = 0.5*Mh, ah = 0.5*tabs(Mh);
	f[0] = ch*(h*(ul[0] + ur[0]) - ah*(ur[0] - ul[0]));
	f[1] = ch*(h*(ul[1] + ur[1]) - ah*(ur[1] - ul[1])) + ph*n[0];
	f[2] = ch*(h*(ul[2] + ur[2]) - ah*(ur[2] - ul[2])) + ph*n[1];
	f[3] = ch*(h*(ul[3] + a.p + ur[3] + b.p) - ah*((ur[3] + b.p) - (ul[3] + a.p)));
}

/// Van Leer flux-vector splitting (anumericalflux.cpp:202-250).
template <class T> FVX_HD void flux_vanleer(const GasParams& gp, const T* ul, const T* ur,
                                            const double* n, T* f)
{
	const double g = gp.g;
	const SideVars<T> a = side_vars(gp, ul, n), b = side_vars(gp, ur, n);
	const T ca = tsqrt(g*a.p/ul[0]), cb = tsqrt(g*b.p/ur[0]);
	const T Ma = a.vn/ca, Mb = b.vn/cb;
	T fp[4], fm[4];
	if(Ma < -1.0) { for(int k = 0; k < 4; k++) fp[k] = 0.0; }
	else if(Ma > 1.0) phys_flux(ul, n, a.vn, a.p, fp);
	else {
		const T v2 = a.v[0]*a.v[0] + a.v[1]*a.v[1];
		fp[0] = ul[0]*ca*(Ma + 1.0)*(Ma + 1.0)*0.25;
		fp[1] = fp[0]*(a.v[0] + n[0]*(2.0*ca - a.vn)/g);
		fp[2] = fp[0]*(a.v[1] + n[1]*(2.0*ca - a.vn)/g);
		const T t = (g - 1.0)*a.vn + 2.0*ca;
		fp[3] = fp[0]*((v2 - a.vn*a.vn)*0.5 + t*t/(2.0*(g*g - 1.0)));
	}
	if(Mb > 1.0) { for(int k = 0; k < 4; k++) fm[k] = 0.0; }
	else if(Mb < -1.0) phys_flux(ur, n, b.vn, b.p, fm);
	else {
		const T v2 = b.v[0]*b.v[0] + b.v[1]*b.v[1];
		fm[0] = -ur[0]*cb*(Mb - 1.0)*(Mb - 1.0)*0.25;
		fm[1] = fm[0]*(b.v[0] + n[0]*(-2.0*cb - b.vn)/g);
		fm[2] = fm[0]*(b.v[1] + n[1]*(-2.0*cb - b.vn)/g);
		const T t = (g - 1.0)*b.vn - 2.0*cb;
		fm[3] = fm[0]*((v2 - b.vn*b.vn)*0.5 + t*t/(2.0*(g*g - 1.0)));
	}
	for(int k = 0; k < 4; k++) f[k] = fp[k] + fm[k];
}

template <int FLUX, class T> FVX_HD void num_flux(const GasParams& gp, const T* ul, const T* ur,
                                                  const double* n, T* f)
{
	if constexpr(FLUX == 0) flux_roe(gp, ul, ur, n, f);
	else if constexpr(FLUX == 1) flux_hllc(gp, ul, ur, n, f);
	else if constexpr(FLUX == 2) flux_hll(gp, ul, ur, n, f);
	else if constexpr(FLUX == 3) flux_llf(gp, ul, ur, n, f);
	else if constexpr(FLUX == 4) flux_ausm(gp, ul, ur, n, f);
	else if constexpr(FLUX == 5) flux_ausmplus(gp, ul, ur, n, f);
	else flux_vanleer(gp, ul, ur, n, f);
}

/// Ghost state of a boundary face (src/spatial/abc.cpp: InOutFlow :49-84, InFlow :152-176,
/// Farfield :194-199, Slipwall :218-226, Adiabaticwall2D :272-280, Isothermalwall2D :354-369,
/// Extrapolation :417-423).
template <class T> FVX_HD void ghost_state(const GasParams& gp, const BCTable& bt, int bc,
                                           const T* ui, const double* n, T* ug)
{
	using std::pow;
	switch(bt.kind[bc]) {
	case 0: {       // slip wall: reflect the normal velocity
		const T vn = (ui[1]*n[0] + ui[2]*n[1])/ui[0];
		ug[0] = ui[0];
		ug[1] = ui[1] - 2.0*vn*n[0]*ui[0];
		ug[2] = ui[2] - 2.0*vn*n[1]*ui[0];
		ug[3] = ui[3];
		break; }
	case 1:         // far field
		for(int k = 0; k < 4; k++) ug[k] = gp.uinf[k];
		break;
	case 2: {       // inflow-outflow: free stream in, pressure outlet when subsonic
		const T vn = (ui[1]*n[0] + ui[2]*n[1])/ui[0];
		const T c = tsqrt(gp.g*pressure(gp, ui)/ui[0]);
		const T M = vn/c;
		if(M <= 0.0) { for(int k = 0; k < 4; k++) ug[k] = gp.uinf[k]; }
		else if(M < 1.0) {
			ug[0] = ui[0]; ug[1] = ui[1]; ug[2] = ui[2];
			ug[3] = gp.pinf/gp.gm1 + 0.5*ui[0]*((ui[1]*ui[1] + ui[2]*ui[2])/(ui[0]*ui[0]));
		}
		else { for(int k = 0; k < 4; k++) ug[k] = ui[k]; }
		break; }
	case 3: {       // subsonic inflow with total pressure v0 and total temperature v1
		const double g = gp.g;
		const T c = tsqrt(g*pressure(gp, ui)/ui[0]);
		const T Rm = (ui[1]*n[0] + ui[2]*n[1])/ui[0] - c/(2*g - 1.0);
		const T c02 = c*c + (g - 1.0)/2.0*(ui[1]*ui[1] + ui[2]*ui[2])/(ui[0]*ui[0]);
		const T q = tsqrt((g + 1)*c02/((g - 1)*Rm*Rm) - (g - 1)/2.0);
		const T cg = -Rm*(g - 1)/(g + 1)*(1.0 + q);
		const T tg = bt.v1[bc]*cg*cg/c02;
		const T pg = bt.v0[bc]*pow(tg/bt.v1[bc], g/(g - 1.0));
		ug[0] = g*gp.Minf*gp.Minf*pg/tg;
		const T vm = tsqrt(2.0/(g - 1.0)*(c02 - cg*cg));
		ug[1] = ug[0]*(vm*1.0*n[0]);
		ug[2] = ug[0]*(vm*1.0*n[1]);
		ug[3] = pg/(g - 1.0) + 0.5*ug[0]*(vm*vm);
		break; }
	case 4:         // extrapolation
		for(int k = 0; k < 4; k++) ug[k] = ui[k];
		break;
	case 7: {       // adiabatic wall with tangential wall speed v0
		const T tm = bt.v0[bc]*ui[0];
		ug[0] = ui[0];
		ug[1] = 2.0*tm*n[1] - ui[1];
		ug[2] = -2.0*tm*n[0] - ui[2];
		ug[3] = ui[3];
		break; }
	case 6: {       // isothermal wall: tangential speed v0, wall temperature v1
		const T Ti = pressure(gp, ui)/ui[0]*gp.gM2;
		const T tg = 2.0*bt.v1[bc] - Ti;
		ug[0] = ui[0];
		ug[1] = ug[0]*(2.0*bt.v0[bc]*n[1] - ui[1]/ui[0]);
		ug[2] = ug[0]*(-2.0*bt.v0[bc]*n[0] - ui[2]/ui[0]);
		const T v2 = (ug[1]*ug[1] + ug[2]*ug[2])/(ug[0]*ug[0]);
		ug[3] = ug[0]*(tg/(gp.g*gp.gm1*gp.Minf*gp.Minf) + 0.5*v2);
		break; }
	default:
		for(int k = 0; k < 4; k++) ug[k] = ui[k];
	}
}

}
#endif
