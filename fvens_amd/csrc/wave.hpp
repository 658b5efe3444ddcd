/* wave.hpp -- wave64 helpers for the cell kernels.
 *
 * The cell kernels map one cell to a quad of four consecutive lanes (lane k of the quad owns
 * the cell's local face k), so a 256-thread workgroup covers 64 cells and a wavefront 16.
 * Per-cell sums, minima and maxima over the faces are quad reductions done with DPP
 * quad_perm lane exchanges; each reduction combines the four values in a fixed pattern, so
 * results are identical on every lane and from run to run.
 */
#ifndef FVX_WAVE_HPP
#define FVX_WAVE_HPP

#include <hip/hip_runtime.h>

namespace fvx {

constexpr int CELLS_PER_BLOCK = 64;
constexpr int BLOCK = 256;
constexpr int MAX_BLOCK = 512;   ///< largest workgroup the tunable cell kernels are built for

// DPP quad_perm selectors: lane^1 = [1,0,3,2], lane^2 = [2,3,0,1]
constexpr int DPP_XOR1 = 0xB1;
constexpr int DPP_XOR2 = 0x4E;

template <int CTRL> __device__ __forceinline__ double dpp_d(double v)
{
	const long long b = __builtin_bit_cast(long long, v);
	int lo = static_cast<int>(b), hi = static_cast<int>(b >> 32);
	lo = __builtin_amdgcn_mov_dpp(lo, CTRL, 0xF, 0xF, false);
	hi = __builtin_amdgcn_mov_dpp(hi, CTRL, 0xF, 0xF, false);
	const long long r = (static_cast<long long>(hi) << 32) | static_cast<unsigned int>(lo);
	return __builtin_bit_cast(double, r);
}

/// Sum over the four lanes of a quad: (v0+v1)+(v2+v3) on every lane.
__device__ __forceinline__ double quad_sum(double v)
{
	v = v + dpp_d<DPP_XOR1>(v);
	return v + dpp_d<DPP_XOR2>(v);
}
__device__ __forceinline__ double quad_max(double v)
{
	v = fmax(v, dpp_d<DPP_XOR1>(v));
	return fmax(v, dpp_d<DPP_XOR2>(v));
}
__device__ __forceinline__ double quad_min(double v)
{
	v = fmin(v, dpp_d<DPP_XOR1>(v));
	return fmin(v, dpp_d<DPP_XOR2>(v));
}

/// Sum, maximum and minimum over the two lanes of a pair (lane^1).
__device__ __forceinline__ double pair_sum(double v) { return v + dpp_d<DPP_XOR1>(v); }
__device__ __forceinline__ double pair_max(double v) { return fmax(v, dpp_d<DPP_XOR1>(v)); }
__device__ __forceinline__ double pair_min(double v) { return fmin(v, dpp_d<DPP_XOR1>(v)); }

/// Logical workgroup index such that each XCD (workgroups are dealt round-robin over the 8
/// XCDs) processes one contiguous range of cells: neighbours gathered by one workgroup are
/// then mostly resident in the same XCD's L2. With on == false, the hardware order.
__device__ __forceinline__ int xcd_block(bool on = true)
{
	const int b = blockIdx.x, nb = gridDim.x;
	if(!on) return b;
	const int q = nb >> 3, r = nb & 7, x = b & 7, l = b >> 3;
	return x*q + (x < r ? x : r) + l;
}

__device__ __forceinline__ void load4(const double* __restrict__ p, double* v)
{
	const double2 a = reinterpret_cast<const double2*>(p)[0];
	const double2 b = reinterpret_cast<const double2*>(p)[1];
	v[0] = a.x; v[1] = a.y; v[2] = b.x; v[3] = b.y;
}
__device__ __forceinline__ void store4(double* __restrict__ p, const double* v)
{
	reinterpret_cast<double2*>(p)[0] = make_double2(v[0], v[1]);
	reinterpret_cast<double2*>(p)[1] = make_double2(v[2], v[3]);
}

}
#endif
