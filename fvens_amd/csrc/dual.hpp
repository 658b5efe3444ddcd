/* dual.hpp -- forward-mode dual numbers for the device Jacobian kernels.
 *
 * Dual<N> carries a value and N directional derivatives. Running the templated fluxes and
 * boundary states of gas.hpp on Dual<4> with the four conserved variables of one side seeded
 * yields that side's 4x4 flux Jacobian exactly (the reference's hand-derived Roe and HLLC
 * linearisations, src/spatial/anumericalflux.cpp:736-965 and 1230-1397, are the exact
 * derivatives of those fluxes).
 */
#ifndef FVX_DUAL_HPP
#define FVX_DUAL_HPP

#include <hip/hip_runtime.h>

namespace fvx {

template <int N> struct Dual {
	double v;
	double d[N];
	__host__ __device__ Dual() : v(0) { for(int i = 0; i < N; i++) d[i] = 0; }
	__host__ __device__ Dual(double x) : v(x) { for(int i = 0; i < N; i++) d[i] = 0; }
};

#define FVX_DD template <int N> __host__ __device__ __forceinline__

FVX_DD Dual<N> operator+(const Dual<N>& a, const Dual<N>& b)
{ Dual<N> r; r.v = a.v + b.v; for(int i = 0; i < N; i++) r.d[i] = a.d[i] + b.d[i]; return r; }
FVX_DD Dual<N> operator-(const Dual<N>& a, const Dual<N>& b)
{ Dual<N> r; r.v = a.v - b.v; for(int i = 0; i < N; i++) r.d[i] = a.d[i] - b.d[i]; return r; }
FVX_DD Dual<N> operator-(const Dual<N>& a)
{ Dual<N> r; r.v = -a.v; for(int i = 0; i < N; i++) r.d[i] = -a.d[i]; return r; }
FVX_DD Dual<N> operator*(const Dual<N>& a, const Dual<N>& b)
{ Dual<N> r; r.v = a.v*b.v; for(int i = 0; i < N; i++) r.d[i] = a.d[i]*b.v + a.v*b.d[i]; return r; }
FVX_DD Dual<N> operator/(const Dual<N>& a, const Dual<N>& b)
{
	Dual<N> r;
	const double ib = 1.0/b.v;
	r.v = a.v*ib;
	for(int i = 0; i < N; i++) r.d[i] = (a.d[i] - r.v*b.d[i])*ib;
	return r;
}
FVX_DD Dual<N> operator+(const Dual<N>& a, double b) { Dual<N> r = a; r.v += b; return r; }
FVX_DD Dual<N> operator+(double b, const Dual<N>& a) { Dual<N> r = a; r.v += b; return r; }
FVX_DD Dual<N> operator-(const Dual<N>& a, double b) { Dual<N> r = a; r.v -= b; return r; }
FVX_DD Dual<N> operator-(double b, const Dual<N>& a) { Dual<N> r = -a; r.v += b; return r; }
FVX_DD Dual<N> operator*(const Dual<N>& a, double b)
{ Dual<N> r; r.v = a.v*b; for(int i = 0; i < N; i++) r.d[i] = a.d[i]*b; return r; }
FVX_DD Dual<N> operator*(double b, const Dual<N>& a) { return a*b; }
FVX_DD Dual<N> operator/(const Dual<N>& a, double b) { return a*(1.0/b); }
FVX_DD Dual<N> operator/(double a, const Dual<N>& b)
{
	Dual<N> r;
	const double ib = 1.0/b.v;
	r.v = a*ib;
	for(int i = 0; i < N; i++) r.d[i] = -r.v*b.d[i]*ib;
	return r;
}
FVX_DD Dual<N>& operator+=(Dual<N>& a, const Dual<N>& b) { a = a + b; return a; }
FVX_DD Dual<N>& operator-=(Dual<N>& a, const Dual<N>& b) { a = a - b; return a; }
FVX_DD bool operator<(const Dual<N>& a, const Dual<N>& b) { return a.v < b.v; }
FVX_DD bool operator>(const Dual<N>& a, const Dual<N>& b) { return a.v > b.v; }
FVX_DD bool operator<=(const Dual<N>& a, const Dual<N>& b) { return a.v <= b.v; }
FVX_DD bool operator>=(const Dual<N>& a, const Dual<N>& b) { return a.v >= b.v; }
FVX_DD bool operator<(const Dual<N>& a, double b) { return a.v < b; }
FVX_DD bool operator>(const Dual<N>& a, double b) { return a.v > b; }
FVX_DD bool operator<=(const Dual<N>& a, double b) { return a.v <= b; }
FVX_DD bool operator>=(const Dual<N>& a, double b) { return a.v >= b; }

FVX_DD Dual<N> sqrt(const Dual<N>& a)
{
	Dual<N> r;
	r.v = ::sqrt(a.v);
	const double h = 0.5/r.v;
	for(int i = 0; i < N; i++) r.d[i] = a.d[i]*h;
	return r;
}
FVX_DD Dual<N> pow(const Dual<N>& a, double e)
{
	Dual<N> r;
	r.v = ::pow(a.v, e);
	const double s = e*r.v/a.v;
	for(int i = 0; i < N; i++) r.d[i] = a.d[i]*s;
	return r;
}

#undef FVX_DD

// keep the scalar overloads visible next to the dual ones inside this namespace
using ::sqrt;
using ::pow;

/// Scalar value of a double or a dual.
__host__ __device__ __forceinline__ double value(double a) { return a; }
template <int N> __host__ __device__ __forceinline__ double value(const Dual<N>& a) { return a.v; }

}
#endif
