/* mpnn.hip -- learned warm start: a message-passing network over the cell graph that predicts
 * the converged state of an operating point, trained and evaluated by the kernels below.
 *
 * No reference counterpart: the reference starts a solve from the freestream or from a
 * first-order starter solve (src/utilities/casesolvers.cpp:225-386). The implicit solver still
 * decides the result: a predicted state is only an initial guess.
 *
 * Network (width 64, 4 layers; relu = max(0, .); bf16 weights, fp32 accumulation):
 *   h0 = x Wenc + benc                                  x: 8 node features
 *   P = hb Wa, Q = hb Wb                                hb: h rounded to bf16
 *   s_ik = P_i + Q_j + e_ik We + bm,  a_i = (1/deg_i) sum_k relu(s_ik)   (interior slots k)
 *   h' = h + relu(hb Wu + ab Wv + bu)                   ab: a rounded to bf16
 *   y = h4 Wdec + bdec                                  4 outputs per cell
 *
 * Data layout: rows are cells of G graphs (row = g*N + cell); activations are row-major
 * [rows][64] fp32, the P|Q pair [rows][128] fp32, GEMM-only operands (a, dL/dpre, dL/dP|dQ)
 * bf16. Per row and layer the relu pattern of the update is kept as one 64-bit mask.
 *
 * Matrix products run on the MFMA units (v_mfma_f32_16x16x32_bf16): a wavefront keeps all 16
 * B fragments of a 64x128 or 128x64 weight matrix in registers and streams 16-row tiles of A.
 * Weight gradients (A^T B over all rows) stage 32-row chunks transposed in LDS and write one
 * partial per workgroup; partials are summed in a fixed order, so every result is deterministic.
 */
#include "fvx_internal.hpp"
#include <cmath>

namespace fvx {
namespace mpnn {

typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef float f32x4 __attribute__((ext_vector_type(4)));
typedef unsigned int u32x4 __attribute__((ext_vector_type(4)));
typedef unsigned short bf16_t;

constexpr int RB = 256;                 // threads per workgroup of every kernel here
constexpr int GEMM_BLOCKS = 2048;       // workgroup cap of the row-GEMM kernels
constexpr int RED_BLOCKS = 512;         // workgroups that write reduction partials

__device__ __forceinline__ bf16_t f2bf(float f)
{
	unsigned u = __float_as_uint(f);
	u += 0x7FFFu + ((u >> 16) & 1u);
	return (bf16_t)(u >> 16);
}
__device__ __forceinline__ float bf2f(bf16_t b) { return __uint_as_float((unsigned)b << 16); }
__device__ __forceinline__ unsigned pack2(float a, float b)
{
	return (unsigned)f2bf(a) | ((unsigned)f2bf(b) << 16);
}
__device__ __forceinline__ bf16x8 to_frag(const float* v)
{
	const u32x4 u = { pack2(v[0], v[1]), pack2(v[2], v[3]), pack2(v[4], v[5]), pack2(v[6], v[7]) };
	return __builtin_bit_cast(bf16x8, u);
}
__device__ __forceinline__ bf16x8 zero_frag()
{
	const u32x4 u = { 0u, 0u, 0u, 0u };
	return __builtin_bit_cast(bf16x8, u);
}
__device__ __forceinline__ void load8(const float* p, float* v)
{
	const f32x4 a = *reinterpret_cast<const f32x4*>(p);
	const f32x4 b = *reinterpret_cast<const f32x4*>(p + 4);
	v[0] = a[0]; v[1] = a[1]; v[2] = a[2]; v[3] = a[3];
	v[4] = b[0]; v[5] = b[1]; v[6] = b[2]; v[7] = b[3];
}
__device__ __forceinline__ f32x4 mfma(bf16x8 a, bf16x8 b, f32x4 c)
{
	return __builtin_amdgcn_mfma_f32_16x16x32_bf16(a, b, c, 0, 0, 0);
}

/// Arguments shared by the kernels of one pass.
struct Graph {
	int N;                  // cells per graph
	int R;                  // rows = graphs x N
	const int* nbr;         // [N][4] interior neighbour or -1
	const int* rev;         // [N][4] slot of this cell in the neighbour's list
	const float* invdeg;    // [N]
	const float* stat;      // [N][6] static node features
	const float* edge;      // [N][4][4] edge features
	const float* gfeat;     // [G][2] per-graph features
};

// ------------------------------------------------------------------------------------------
// weight packing: fp32 master -> bf16-rounded copy (VALU weights) and MFMA B fragments
// ------------------------------------------------------------------------------------------

/// B-operand element (k, n) of packed matrix `mat` of layer l (see the file header).
__device__ __forceinline__ float bmat(const float* w, int l, int mat, int k, int n)
{
	const float* L = w + OFF_L0 + l*LSZ;
	switch(mat) {
	case M_PQ:   return n < W ? L[L_WA + k*W + n] : L[L_WB + k*W + n - W];
	case M_UPD:  return k < W ? L[L_WU + k*W + n] : L[L_WV + (k - W)*W + n];
	case M_BUPD: return n < W ? L[L_WU + n*W + k] : L[L_WV + (n - W)*W + k];
	default:     return k < W ? L[L_WA + n*W + k] : L[L_WB + n*W + k - W];
	}
}

__global__ void __launch_bounds__(RB) k_pack(const float* __restrict__ w, float* __restrict__ wq,
                                             bf16_t* __restrict__ frag)
{
	const int t = blockIdx.x*RB + threadIdx.x;
	if(t < NPARAM) {
		// matrices enter the arithmetic rounded to bf16, biases unchanged
		const bool bias = (t >= OFF_BENC && t < OFF_L0) || t >= OFF_BDEC
		                  || (t >= OFF_L0 && t < OFF_WDEC
		                      && (((t - OFF_L0) % LSZ >= L_BM && (t - OFF_L0) % LSZ < L_WU)
		                          || (t - OFF_L0) % LSZ >= L_BU));
		wq[t] = bias ? w[t] : bf2f(f2bf(w[t]));
	}
	if(t < NL*4*FRAG_MAT) {
		const int lm = t / FRAG_MAT, r = t % FRAG_MAT;
		const int l = lm / 4, mat = lm % 4;
		const int e = r % 8, lane = (r / 8) % 64, sj = r / 512;
		const int nt = (mat == M_PQ || mat == M_BUPD) ? 8 : 4;
		const int s = sj / nt, j = sj % nt;
		const int k = 32*s + 8*(lane/16) + e, n = 16*j + lane%16;
		frag[t] = f2bf(bmat(w, l, mat, k, n));
	}
}

// ------------------------------------------------------------------------------------------
// row GEMM driver: each wavefront holds the 16 B fragments and walks 16-row tiles
// ------------------------------------------------------------------------------------------

template <int S, int NT, class ALoad, class Epi>
__device__ __forceinline__ void row_gemm(int ntiles, const bf16_t* __restrict__ frag, ALoad&& aload,
                                         Epi&& epi)
{
	static_assert(S*NT == 16, "every packed matrix has 16 fragments");
	const int lane = threadIdx.x & 63;
	const int wave = blockIdx.x*(RB/64) + (threadIdx.x >> 6);
	const int nwave = gridDim.x*(RB/64);
	bf16x8 b[S][NT];
	const bf16x8* F = reinterpret_cast<const bf16x8*>(frag);
#pragma unroll
	for(int s = 0; s < S; s++)
#pragma unroll
		for(int j = 0; j < NT; j++) b[s][j] = F[(s*NT + j)*64 + lane];
	for(int t = wave; t < ntiles; t += nwave) {
		f32x4 acc[NT];
#pragma unroll
		for(int j = 0; j < NT; j++) acc[j] = f32x4{0.f, 0.f, 0.f, 0.f};
#pragma unroll
		for(int s = 0; s < S; s++) {
			const bf16x8 a = aload(t, s, lane);
#pragma unroll
			for(int j = 0; j < NT; j++) acc[j] = mfma(a, b[s][j], acc[j]);
		}
		epi(t, lane, acc);
	}
}

/// P|Q = hb [Wa | Wb]   (K = 64, 128 outputs)
__global__ void __launch_bounds__(RB) k_pq(int R, const float* __restrict__ h,
                                           const bf16_t* __restrict__ frag, float* __restrict__ pq)
{
	row_gemm<2, 8>((R + 15)/16, frag,
		[&](int t, int s, int lane) {
			const int r = 16*t + (lane & 15);
			if(r >= R) return zero_frag();
			float v[8];
			load8(h + (size_t)r*W + 32*s + 8*(lane >> 4), v);
			return to_frag(v);
		},
		[&](int t, int lane, const f32x4* acc) {
#pragma unroll
			for(int j = 0; j < 8; j++)
#pragma unroll
				for(int q = 0; q < 4; q++) {
					const int row = 16*t + 4*(lane >> 4) + q;
					if(row < R) pq[(size_t)row*2*W + 16*j + (lane & 15)] = acc[j][q];
				}
		});
}

/// Aggregation of the messages of cell row r over channels c0..c0+7: a = (1/deg) sum relu(s).
__device__ __forceinline__ void aggregate8(const Graph& g, int r, int c0, const float* __restrict__ pq,
                                           const float* __restrict__ wl, float* a)
{
	const int node = r % g.N;
	const size_t base = (size_t)(r - node);
	float p[8], we[4][8], bm[8];
	load8(pq + (size_t)r*2*W + c0, p);
#pragma unroll
	for(int f = 0; f < 4; f++) load8(wl + L_WE + f*W + c0, we[f]);
	load8(wl + L_BM + c0, bm);
#pragma unroll
	for(int e = 0; e < 8; e++) a[e] = 0.f;
#pragma unroll
	for(int k = 0; k < 4; k++) {
		const int j = g.nbr[node*4 + k];
		if(j < 0) continue;
		float q[8];
		load8(pq + (base + j)*2*W + W + c0, q);
		const f32x4 ef = *reinterpret_cast<const f32x4*>(g.edge + ((size_t)node*4 + k)*4);
#pragma unroll
		for(int e = 0; e < 8; e++) {
			const float s = p[e] + q[e] + (ef[0]*we[0][e] + ef[1]*we[1][e] + ef[2]*we[2][e]
			                               + ef[3]*we[3][e]) + bm[e];
			a[e] += fmaxf(s, 0.f);
		}
	}
	const float id = g.invdeg[node];
#pragma unroll
	for(int e = 0; e < 8; e++) a[e] *= id;
}

/// Message aggregation and residual update of one layer:
///   h_out = h + relu([hb | ab] [Wu; Wv] + bu)   (K = 128, 64 outputs)
/// Training passes also keep ab (bf16) and the relu mask of every row.
__global__ void __launch_bounds__(RB) k_update(Graph g, const float* __restrict__ h,
                                               const float* __restrict__ pq,
                                               const float* __restrict__ wl,
                                               const bf16_t* __restrict__ frag,
                                               float* __restrict__ hout, bf16_t* __restrict__ ab,
                                               unsigned long long* __restrict__ umask)
{
	const int R = g.R;
	row_gemm<4, 4>((R + 15)/16, frag,
		[&](int t, int s, int lane) {
			const int r = 16*t + (lane & 15);
			if(r >= R) return zero_frag();
			float v[8];
			if(s < 2) {
				load8(h + (size_t)r*W + 32*s + 8*(lane >> 4), v);
			} else {
				const int c0 = 32*(s - 2) + 8*(lane >> 4);
				aggregate8(g, r, c0, pq, wl, v);
				const bf16x8 fr = to_frag(v);
				if(ab) *reinterpret_cast<bf16x8*>(ab + (size_t)r*W + c0) = fr;
				return fr;
			}
			return to_frag(v);
		},
		[&](int t, int lane, const f32x4* acc) {
			unsigned long long bits[4][4];
#pragma unroll
			for(int j = 0; j < 4; j++) {
				const int col = 16*j + (lane & 15);
				const float bu = wl[L_BU + col];
#pragma unroll
				for(int q = 0; q < 4; q++) {
					const int row = 16*t + 4*(lane >> 4) + q;
					const float pre = acc[j][q] + bu;
					bits[j][q] = __ballot(pre > 0.f);
					if(row < R) hout[(size_t)row*W + col] = h[(size_t)row*W + col] + fmaxf(pre, 0.f);
				}
			}
			if(umask && lane < 16) {
				const int gq = lane >> 2, qq = lane & 3;
				unsigned long long m = 0;
#pragma unroll
				for(int j = 0; j < 4; j++) {
					const unsigned long long b = qq == 0 ? bits[j][0] : qq == 1 ? bits[j][1]
					                             : qq == 2 ? bits[j][2] : bits[j][3];
					m |= ((b >> (16*gq)) & 0xFFFFull) << (16*j);
				}
				const int row = 16*t + lane;
				if(row < R) umask[row] = m;
			}
		});
}

/// Backward of the update: dpre = g * mask (kept as bf16), then
///   [dh_part | da] = dpre [Wu^T | Wv^T],  dh_part += g   (K = 64, 128 outputs)
__global__ void __launch_bounds__(RB) k_update_bwd(int R, const float* __restrict__ gin,
                                                   const unsigned long long* __restrict__ umask,
                                                   const bf16_t* __restrict__ frag,
                                                   bf16_t* __restrict__ dpre,
                                                   float* __restrict__ gpart, float* __restrict__ da)
{
	row_gemm<2, 8>((R + 15)/16, frag,
		[&](int t, int s, int lane) {
			const int r = 16*t + (lane & 15);
			if(r >= R) return zero_frag();
			const int c0 = 32*s + 8*(lane >> 4);
			float v[8];
			load8(gin + (size_t)r*W + c0, v);
			const unsigned m = (unsigned)(umask[r] >> c0) & 0xFFu;
#pragma unroll
			for(int e = 0; e < 8; e++) v[e] = (m >> e) & 1u ? v[e] : 0.f;
			const bf16x8 fr = to_frag(v);
			*reinterpret_cast<bf16x8*>(dpre + (size_t)r*W + c0) = fr;
			return fr;
		},
		[&](int t, int lane, const f32x4* acc) {
#pragma unroll
			for(int j = 0; j < 8; j++)
#pragma unroll
				for(int q = 0; q < 4; q++) {
					const int row = 16*t + 4*(lane >> 4) + q;
					const int col = 16*j + (lane & 15);
					if(row >= R) continue;
					if(j < 4) gpart[(size_t)row*W + col] = gin[(size_t)row*W + col] + acc[j][q];
					else da[(size_t)row*W + col - W] = acc[j][q];
				}
		});
}

/// dh = dh_part + [dP | dQ] [Wa^T; Wb^T]   (K = 128, 64 outputs)
__global__ void __launch_bounds__(RB) k_pq_bwd(int R, const bf16_t* __restrict__ dpq,
                                               const bf16_t* __restrict__ frag,
                                               const float* __restrict__ gpart,
                                               float* __restrict__ gout)
{
	row_gemm<4, 4>((R + 15)/16, frag,
		[&](int t, int s, int lane) {
			const int r = 16*t + (lane & 15);
			if(r >= R) return zero_frag();
			return *reinterpret_cast<const bf16x8*>(dpq + (size_t)r*2*W + 32*s + 8*(lane >> 4));
		},
		[&](int t, int lane, const f32x4* acc) {
#pragma unroll
			for(int j = 0; j < 4; j++)
#pragma unroll
				for(int q = 0; q < 4; q++) {
					const int row = 16*t + 4*(lane >> 4) + q;
					const int col = 16*j + (lane & 15);
					if(row < R) gout[(size_t)row*W + col] = gpart[(size_t)row*W + col] + acc[j][q];
				}
		});
}

// ------------------------------------------------------------------------------------------
// VALU kernels: encoder, decoder + loss, message backward (with partial weight sums)
// ------------------------------------------------------------------------------------------

__device__ __forceinline__ void node_x(const Graph& g, int r, float* x)
{
	const int node = r % g.N, gi = r / g.N;
#pragma unroll
	for(int f = 0; f < NSTAT; f++) x[f] = g.stat[node*NSTAT + f];
	x[6] = g.gfeat[2*gi];
	x[7] = g.gfeat[2*gi + 1];
}

/// h0 = x Wenc + benc; one thread per (row, channel).
__global__ void __launch_bounds__(RB) k_encode(Graph g, const float* __restrict__ wq,
                                               float* __restrict__ h0)
{
	const size_t t = (size_t)blockIdx.x*RB + threadIdx.x;
	if(t >= (size_t)g.R*W) return;
	const int r = (int)(t / W), c = (int)(t % W);
	float x[FIN];
	node_x(g, r, x);
	float s = wq[OFF_BENC + c];
#pragma unroll
	for(int f = 0; f < FIN; f++) s += x[f]*wq[OFF_WENC + f*W + c];
	h0[t] = s;
}

/// Decoder y = h4 Wdec + bdec (16 lanes per row, 4 channels each). With targets: the loss
/// gradient dy = 2 (y - t) / count, dh4 = dy Wdec^T and workgroup partials of dWdec, dbdec and
/// the squared error. Targets of graph gi are rows ids[gi] of the target set.
__global__ void __launch_bounds__(RB) k_decode(Graph g, const float* __restrict__ h4,
                                               const float* __restrict__ wq,
                                               float* __restrict__ y,
                                               const float* __restrict__ tgt, const int* __restrict__ ids,
                                               double inv_count, float* __restrict__ gout,
                                               float* __restrict__ part)
{
	__shared__ float red[RB*21];
	const int cg = threadIdx.x & 15, c0 = 4*cg;
	float acc[21];
#pragma unroll
	for(int v = 0; v < 21; v++) acc[v] = 0.f;
	float wd[4][4];
#pragma unroll
	for(int cc = 0; cc < 4; cc++)
#pragma unroll
		for(int o = 0; o < 4; o++) wd[cc][o] = wq[OFF_WDEC + (c0 + cc)*FO + o];
	const int stride = gridDim.x*(RB/16);
	// every lane runs the same number of iterations (the shuffles below need all 16 lanes)
	const int niter = (g.R + stride - 1)/stride;
	for(int it = 0; it < niter; it++) {
		const int r = it*stride + blockIdx.x*(RB/16) + (threadIdx.x >> 4);
		const bool ok = r < g.R;
		f32x4 hv = f32x4{0.f, 0.f, 0.f, 0.f};
		if(ok) hv = *reinterpret_cast<const f32x4*>(h4 + (size_t)r*W + c0);
		float yo[4];
#pragma unroll
		for(int o = 0; o < 4; o++) {
			float s = hv[0]*wd[0][o] + hv[1]*wd[1][o] + hv[2]*wd[2][o] + hv[3]*wd[3][o];
			s += __shfl_xor(s, 8, 16);
			s += __shfl_xor(s, 4, 16);
			s += __shfl_xor(s, 2, 16);
			s += __shfl_xor(s, 1, 16);
			yo[o] = s + wq[OFF_BDEC + o];
		}
		if(!ok) continue;
		if(!tgt) {
			if(cg == 0) *reinterpret_cast<f32x4*>(y + (size_t)r*FO) = f32x4{yo[0], yo[1], yo[2], yo[3]};
			continue;
		}
		const int node = r % g.N, gi = r / g.N;
		const f32x4 tv = *reinterpret_cast<const f32x4*>(tgt + ((size_t)ids[gi]*g.N + node)*FO);
		float dy[4];
#pragma unroll
		for(int o = 0; o < 4; o++) {
			const float d = yo[o] - tv[o];
			dy[o] = (float)(2.0*(double)d*inv_count);
			if(cg == 0) { acc[16 + o] += dy[o]; acc[20] += d*d; }
		}
		f32x4 gv;
#pragma unroll
		for(int cc = 0; cc < 4; cc++) {
			gv[cc] = dy[0]*wd[cc][0] + dy[1]*wd[cc][1] + dy[2]*wd[cc][2] + dy[3]*wd[cc][3];
#pragma unroll
			for(int o = 0; o < 4; o++) acc[cc*4 + o] += hv[cc]*dy[o];
		}
		*reinterpret_cast<f32x4*>(gout + (size_t)r*W + c0) = gv;
	}
	if(!tgt) return;
	// reduce over the 16 row slots; partial layout: Wdec[c][o] (256), bdec (4), loss (1)
	float* out = part + (size_t)blockIdx.x*261;
	const int tid = threadIdx.x, slot = tid >> 4;
#pragma unroll
	for(int v = 0; v < 21; v++) red[(slot*16 + cg)*21 + v] = acc[v];
	__syncthreads();
	for(int i = tid; i < 261; i += RB) {
		float s = 0.f;
		if(i < 256) {
			const int c = i / 4, o = i % 4, grp = c / 4, cc = c % 4;
			for(int sl = 0; sl < RB/16; sl++) s += red[(sl*16 + grp)*21 + cc*4 + o];
		} else {
			const int v = i - 256 + 16;          // bdec 0..3 -> 16..19, loss -> 20
			for(int sl = 0; sl < RB/16; sl++) s += red[(sl*16 + 0)*21 + v];
		}
		out[i] = s;
	}
}

/// Backward of the aggregation (16 lanes per cell, 4 channels each):
///   ds_ik = (1/deg_i) da_i [s_ik > 0];  dP_i = sum_k ds_ik;  dQ_j = sum over the slots of j
/// of ds at the neighbour's reverse slot; workgroup partials of dWe = sum e_ik ds_ik and
/// dbm = sum ds_ik. Writes dP|dQ as bf16 (the operand of the following GEMMs).
__global__ void __launch_bounds__(RB) k_message_bwd(Graph g, const float* __restrict__ pq,
                                                    const float* __restrict__ da,
                                                    const float* __restrict__ wl,
                                                    bf16_t* __restrict__ dpq,
                                                    float* __restrict__ part)
{
	__shared__ float red[RB*20];
	const int cg = threadIdx.x & 15, c0 = 4*cg;
	float we[4][4], bm[4];
#pragma unroll
	for(int f = 0; f < 4; f++)
#pragma unroll
		for(int cc = 0; cc < 4; cc++) we[f][cc] = wl[L_WE + f*W + c0 + cc];
#pragma unroll
	for(int cc = 0; cc < 4; cc++) bm[cc] = wl[L_BM + c0 + cc];
	float acc[20];
#pragma unroll
	for(int v = 0; v < 20; v++) acc[v] = 0.f;
	const int stride = gridDim.x*(RB/16);
	for(int r = blockIdx.x*(RB/16) + (threadIdx.x >> 4); r < g.R; r += stride) {
		const int node = r % g.N;
		const size_t base = (size_t)(r - node);
		const f32x4 dj = *reinterpret_cast<const f32x4*>(da + (size_t)r*W + c0);
		const f32x4 pj = *reinterpret_cast<const f32x4*>(pq + (size_t)r*2*W + c0);
		const f32x4 qj = *reinterpret_cast<const f32x4*>(pq + (size_t)r*2*W + W + c0);
		const float idj = g.invdeg[node];
		float dP[4] = {0.f, 0.f, 0.f, 0.f}, dQ[4] = {0.f, 0.f, 0.f, 0.f};
		for(int k = 0; k < 4; k++) {
			const int j = g.nbr[node*4 + k];
			if(j < 0) continue;
			const size_t rj = base + j;
			const f32x4 e = *reinterpret_cast<const f32x4*>(g.edge + ((size_t)node*4 + k)*4);
			const f32x4 qn = *reinterpret_cast<const f32x4*>(pq + rj*2*W + W + c0);
			// own edge i -> j
#pragma unroll
			for(int cc = 0; cc < 4; cc++) {
				const float s = pj[cc] + qn[cc] + (e[0]*we[0][cc] + e[1]*we[1][cc] + e[2]*we[2][cc]
				                                   + e[3]*we[3][cc]) + bm[cc];
				const float ds = s > 0.f ? idj*dj[cc] : 0.f;
				dP[cc] += ds;
#pragma unroll
				for(int f = 0; f < 4; f++) acc[f*4 + cc] += e[f]*ds;
				acc[16 + cc] += ds;
			}
			// reverse edge j -> this cell, seen from the neighbour's slot
			const int kk = g.rev[node*4 + k];
			const f32x4 e2 = *reinterpret_cast<const f32x4*>(g.edge + ((size_t)j*4 + kk)*4);
			const f32x4 pn = *reinterpret_cast<const f32x4*>(pq + rj*2*W + c0);
			const f32x4 dn = *reinterpret_cast<const f32x4*>(da + rj*W + c0);
			const float idn = g.invdeg[j];
#pragma unroll
			for(int cc = 0; cc < 4; cc++) {
				const float s = pn[cc] + qj[cc] + (e2[0]*we[0][cc] + e2[1]*we[1][cc] + e2[2]*we[2][cc]
				                                   + e2[3]*we[3][cc]) + bm[cc];
				dQ[cc] += s > 0.f ? idn*dn[cc] : 0.f;
			}
		}
		const unsigned p0 = pack2(dP[0], dP[1]), p1 = pack2(dP[2], dP[3]);
		const unsigned q0 = pack2(dQ[0], dQ[1]), q1 = pack2(dQ[2], dQ[3]);
		*reinterpret_cast<uint2*>(dpq + (size_t)r*2*W + c0) = make_uint2(p0, p1);
		*reinterpret_cast<uint2*>(dpq + (size_t)r*2*W + W + c0) = make_uint2(q0, q1);
	}
	// partial layout: We[f][c] (256), bm[c] (64)
	float* out = part + (size_t)blockIdx.x*320;
	const int tid = threadIdx.x, slot = tid >> 4;
#pragma unroll
	for(int v = 0; v < 20; v++) red[(slot*16 + cg)*20 + v] = acc[v];
	__syncthreads();
	for(int i = tid; i < 320; i += RB) {
		int grp, v;
		if(i < 256) { const int f = i / W, c = i % W; grp = c / 4; v = f*4 + c % 4; }
		else { const int c = i - 256; grp = c / 4; v = 16 + c % 4; }
		float s = 0.f;
		for(int sl = 0; sl < RB/16; sl++) s += red[(sl*16 + grp)*20 + v];
		out[i] = s;
	}
}

/// Encoder backward partials: dWenc[f][c] = sum_r x_f g0[r][c], dbenc[c] = sum_r g0[r][c];
/// four row slots of 64 channels per workgroup.
__global__ void __launch_bounds__(RB) k_encode_bwd(Graph g, const float* __restrict__ g0,
                                                   float* __restrict__ part)
{
	__shared__ float red[RB*9];
	const int c = threadIdx.x & 63, slot = threadIdx.x >> 6;
	float acc[9];
#pragma unroll
	for(int v = 0; v < 9; v++) acc[v] = 0.f;
	for(int r = blockIdx.x*4 + slot; r < g.R; r += gridDim.x*4) {
		float x[FIN];
		node_x(g, r, x);
		const float gv = g0[(size_t)r*W + c];
#pragma unroll
		for(int f = 0; f < FIN; f++) acc[f] += x[f]*gv;
		acc[8] += gv;
	}
#pragma unroll
	for(int v = 0; v < 9; v++) red[(slot*64 + c)*9 + v] = acc[v];
	__syncthreads();
	float* out = part + (size_t)blockIdx.x*576;
	for(int i = threadIdx.x; i < 576; i += RB) {
		const int f = i < 512 ? i / W : 8, cc = i < 512 ? i % W : i - 512;
		float s = 0.f;
		for(int sl = 0; sl < 4; sl++) s += red[(sl*64 + cc)*9 + f];
		out[i] = s;
	}
}

// ------------------------------------------------------------------------------------------
// weight gradients on the MFMA units: C[K][NO] = sum_r A[r][k] B[r][n] (+ bias row sum_r B)
// ------------------------------------------------------------------------------------------

constexpr int CH = 32;          // rows per LDS chunk (the MFMA reduction depth)
constexpr int LDS_LD = 40;      // bf16 per transposed LDS row (32 + 8 padding, 80 bytes)

/// A = [h (fp32, rounded to bf16) | a2 (bf16, when K = 128)], B = bf16 [R][NO].
template <int K, int NO, bool BIAS>
__global__ void __launch_bounds__(RB) k_wgrad(int R, int chunks_per_block,
                                              const float* __restrict__ h,
                                              const bf16_t* __restrict__ a2,
                                              const bf16_t* __restrict__ bsrc,
                                              float* __restrict__ part)
{
	__shared__ __attribute__((aligned(16))) bf16_t AT[K*LDS_LD];
	__shared__ __attribute__((aligned(16))) bf16_t BT[NO*LDS_LD];
	constexpr int TN = NO/16, NTILE = (K/16)*TN, PER = NTILE/4;
	const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
	f32x4 acc[PER];
#pragma unroll
	for(int m = 0; m < PER; m++) acc[m] = f32x4{0.f, 0.f, 0.f, 0.f};
	float bsum = 0.f;
	const int c_begin = blockIdx.x*chunks_per_block;
	const int nchunk = (R + CH - 1)/CH;
	const int c_end = min(c_begin + chunks_per_block, nchunk);
	for(int ch = c_begin; ch < c_end; ch++) {
		const int r0 = ch*CH;
		__syncthreads();
		{   // h chunk: 32 rows x 64 fp32, 8 values per thread
			const int row = tid >> 3, c8 = (tid & 7)*8;
			float v[8] = {0.f, 0.f, 0.f, 0.f, 0.f, 0.f, 0.f, 0.f};
			if(r0 + row < R) load8(h + (size_t)(r0 + row)*W + c8, v);
#pragma unroll
			for(int e = 0; e < 8; e++) AT[(c8 + e)*LDS_LD + row] = f2bf(v[e]);
		}
		if(K == 128) {
			const int row = tid >> 3, c8 = (tid & 7)*8;
			u32x4 v = {0u, 0u, 0u, 0u};
			if(r0 + row < R) v = *reinterpret_cast<const u32x4*>(a2 + (size_t)(r0 + row)*W + c8);
#pragma unroll
			for(int e = 0; e < 8; e++)
				AT[(W + c8 + e)*LDS_LD + row] = (bf16_t)(v[e/2] >> (16*(e & 1)));
		}
#pragma unroll
		for(int it = 0; it < NO/64; it++) {
			const int idx = tid + RB*it;
			const int row = idx / (NO/8), c8 = (idx % (NO/8))*8;
			u32x4 v = {0u, 0u, 0u, 0u};
			if(r0 + row < R) v = *reinterpret_cast<const u32x4*>(bsrc + (size_t)(r0 + row)*NO + c8);
#pragma unroll
			for(int e = 0; e < 8; e++)
				BT[(c8 + e)*LDS_LD + row] = (bf16_t)(v[e/2] >> (16*(e & 1)));
		}
		__syncthreads();
#pragma unroll
		for(int m = 0; m < PER; m++) {
			const int tt = wave + 4*m, tk = tt / TN, tn = tt % TN;
			const bf16x8 a = *reinterpret_cast<const bf16x8*>(AT + (16*tk + (lane & 15))*LDS_LD
			                                                   + 8*(lane >> 4));
			const bf16x8 b = *reinterpret_cast<const bf16x8*>(BT + (16*tn + (lane & 15))*LDS_LD
			                                                   + 8*(lane >> 4));
			acc[m] = mfma(a, b, acc[m]);
		}
		if(BIAS && tid < NO) {
			float s = 0.f;
			for(int rr = 0; rr < CH; rr++) s += bf2f(BT[tid*LDS_LD + rr]);
			bsum += s;
		}
	}
	float* out = part + (size_t)blockIdx.x*(K*NO + (BIAS ? NO : 0));
#pragma unroll
	for(int m = 0; m < PER; m++) {
		const int tt = wave + 4*m, tk = tt / TN, tn = tt % TN;
#pragma unroll
		for(int q = 0; q < 4; q++)
			out[(16*tk + 4*(lane >> 4) + q)*NO + 16*tn + (lane & 15)] = acc[m][q];
	}
	if(BIAS && tid < NO) out[K*NO + tid] = bsum;
}

/// Sums `nblk` partials of `width` values into the flat gradient. A workgroup owns 16
/// consecutive values; its 16 lane groups each sum every 16th partial (in increasing order), and
/// the 16 group sums are added in group order, so the result is deterministic. Modes:
/// 0 contiguous at off1; 1 columns split (a [rows][width/rows] matrix whose left half goes to
/// off1 and right half to off2); 2 rows split (rows < half -> off1, the rest -> off2).
constexpr int RED_COLS = 16, RED_GROUPS = RB/RED_COLS;
__global__ void __launch_bounds__(RB) k_reduce(int nblk, int width, int ncols, int mode,
                                               int off1, int off2, const float* __restrict__ part,
                                               float* __restrict__ grad, float* __restrict__ extra)
{
	__shared__ double sh[RB];
	const int c = threadIdx.x % RED_COLS, grp = threadIdx.x / RED_COLS;
	const int i = blockIdx.x*RED_COLS + c;
	double acc = 0.0;
	if(i < width)
		for(int b = grp; b < nblk; b += RED_GROUPS) acc += part[(size_t)b*width + i];
	sh[threadIdx.x] = acc;
	__syncthreads();
	if(grp != 0 || i >= width) return;
	double s = 0.0;
	for(int g = 0; g < RED_GROUPS; g++) s += sh[g*RED_COLS + c];
	const int k = i / ncols, n = i % ncols;
	if(mode == 0) {
		if(extra && i == width - 1) { *extra = (float)s; return; }
		grad[off1 + i] = (float)s;
	} else if(mode == 1) {
		const int hc = ncols/2;
		if(n < hc) grad[off1 + k*hc + n] = (float)s;
		else grad[off2 + k*hc + n - hc] = (float)s;
	} else {
		if(k < W) grad[off1 + k*ncols + n] = (float)s;
		else grad[off2 + (k - W)*ncols + n] = (float)s;
	}
}

/// Adam (Kingma and Ba, 2015) on the fp32 master weights.
__global__ void __launch_bounds__(RB) k_adam(int n, float* __restrict__ w, float* __restrict__ m,
                                             float* __restrict__ v, const float* __restrict__ grad,
                                             float lr, float b1, float b2, float omb1, float omb2,
                                             float eps, float bc1, float bc2)
{
	const int i = blockIdx.x*RB + threadIdx.x;
	if(i >= n) return;
	const float gi = grad[i];
	const float mi = b1*m[i] + omb1*gi;
	const float vi = b2*v[i] + omb2*gi*gi;
	m[i] = mi;
	v[i] = vi;
	w[i] -= lr*(mi/bc1)/(sqrtf(vi/bc2) + eps);
}

/// Targets from conserved fields (see mpnn::targets on the host side).
__global__ void __launch_bounds__(RB) k_targets(int R, int N, const double* __restrict__ u,
                                                const double* __restrict__ pp, double gamma,
                                                float* __restrict__ t)
{
	const int r = blockIdx.x*RB + threadIdx.x;
	if(r >= R) return;
	const int gi = r / N;
	const double ca = pp[3*gi], sa = pp[3*gi + 1], pinf = pp[3*gi + 2];
	const double* q = u + (size_t)r*4;
	const double rho = q[0], vx = q[1]/rho, vy = q[2]/rho;
	const double p = (gamma - 1)*(q[3] - 0.5*rho*(vx*vx + vy*vy));
	float* o = t + (size_t)r*4;
	o[0] = (float)(rho - 1); o[1] = (float)(vx - ca); o[2] = (float)(vy - sa);
	o[3] = (float)(p/pinf - 1);
}

/// Conserved state from outputs, density and pressure floored at `floor` x freestream.
__global__ void __launch_bounds__(RB) k_warm(int R, int N, const float* __restrict__ y,
                                             const double* __restrict__ pp, double gamma,
                                             double floor, double* __restrict__ u)
{
	const int r = blockIdx.x*RB + threadIdx.x;
	if(r >= R) return;
	const int gi = r / N;
	const double ca = pp[3*gi], sa = pp[3*gi + 1], pinf = pp[3*gi + 2];
	const float* o = y + (size_t)r*4;
	const double rho = fmax(1.0 + (double)o[0], floor);
	const double vx = ca + (double)o[1], vy = sa + (double)o[2];
	const double p = fmax(1.0 + (double)o[3], floor)*pinf;
	double* q = u + (size_t)r*4;
	q[0] = rho; q[1] = rho*vx; q[2] = rho*vy; q[3] = p/(gamma - 1) + 0.5*rho*(vx*vx + vy*vy);
}

} // namespace mpnn

// ------------------------------------------------------------------------------------------
// host side
// ------------------------------------------------------------------------------------------

using namespace mpnn;

namespace {

double point_segment(double px, double py, double ax, double ay, double bx, double by)
{
	const double dx = bx - ax, dy = by - ay, ll = dx*dx + dy*dy;
	double t = ll > 0 ? ((px - ax)*dx + (py - ay)*dy)/ll : 0.0;
	t = std::min(1.0, std::max(0.0, t));
	const double qx = ax + t*dx - px, qy = ay + t*dy - py;
	return std::sqrt(qx*qx + qy*qy);
}

int blocks_for(size_t n) { return (int)((n + RB - 1)/RB); }
int reduce_grid(int width) { return (width + RED_COLS - 1)/RED_COLS; }

}

void mpnn_build_graph(const fvx_mesh_s& m, int wall_tag, MpnnGraphHost& gh)
{
	const int N = m.nelem;
	gh.N = N;
	gh.nbr.assign((size_t)N*4, -1);
	gh.rev.assign((size_t)N*4, -1);
	gh.invdeg.assign(N, 0.f);
	gh.stat.assign((size_t)N*NSTAT, 0.f);
	gh.edge.assign((size_t)N*16, 0.f);
	for(int c = 0; c < N; c++) {
		int deg = 0;
		for(int k = 0; k < 4; k++) {
			const int o = m.esuel[(size_t)c*4 + k];
			if(o >= 0 && o < N) { gh.nbr[(size_t)c*4 + k] = o; deg++; }
		}
		gh.invdeg[c] = deg > 0 ? (float)(1.0/deg) : 0.f;
	}
	for(int c = 0; c < N; c++)
		for(int k = 0; k < 4; k++) {
			const int j = gh.nbr[(size_t)c*4 + k];
			if(j < 0) continue;
			for(int kk = 0; kk < 4; kk++)
				if(gh.nbr[(size_t)j*4 + kk] == c) { gh.rev[(size_t)c*4 + k] = kk; break; }
			if(gh.rev[(size_t)c*4 + k] < 0)
				throw Error(FVX_ERR_INVALID, "mesh neighbour lists are not symmetric");
		}
	std::vector<int> wall;
	for(int b = 0; b < m.nbface; b++) if(m.btag[b] == wall_tag) wall.push_back(b);
	for(int c = 0; c < N; c++) {
		const double xc = m.rc[2*c], yc = m.rc[2*c + 1];
		const double r = std::hypot(xc - 0.5, yc);
		double dw = wall.empty() ? 1e4 : INFINITY;
		for(int b : wall) {
			const int p0 = m.bface[2*b], p1 = m.bface[2*b + 1];
			dw = std::min(dw, point_segment(xc, yc, m.coords[2*p0], m.coords[2*p0 + 1],
			                                m.coords[2*p1], m.coords[2*p1 + 1]));
		}
		double walladj = 0;
		for(int k = 0; k < 4; k++) {
			const int o = m.esuel[(size_t)c*4 + k];
			if(o >= N && m.btag[o - N] == wall_tag) walladj = 1;
		}
		float* s = &gh.stat[(size_t)c*NSTAT];
		s[0] = (float)((xc - 0.5)/(1 + r));
		s[1] = (float)(yc/(1 + r));
		s[2] = (float)(std::log1p(r)/std::log(51.0));
		s[3] = (float)(std::log10(dw + 1e-4)/2);
		s[4] = (float)(std::log10(std::sqrt(m.area[c]))/2);
		s[5] = (float)walladj;
		for(int k = 0; k < 4; k++) {
			const int j = gh.nbr[(size_t)c*4 + k];
			if(j < 0) continue;
			const double dx = m.rc[2*j] - xc, dy = m.rc[2*j + 1] - yc, L = std::hypot(dx, dy);
			float* e = &gh.edge[((size_t)c*4 + k)*4];
			e[0] = (float)(dx/L);
			e[1] = (float)(dy/L);
			e[2] = (float)std::log10(L/std::sqrt(m.area[c]));
			e[3] = (float)std::log(m.area[j]/m.area[c]);
		}
	}
}

void Mpnn::init(const fvx_mesh_s& m, int wall_tag, int max_graphs)
{
	mpnn_build_graph(m, wall_tag, host);
	N = host.N;
	cap = max_graphs;
	d_nbr.upload(host.nbr);
	d_rev.upload(host.rev);
	d_invdeg.upload(host.invdeg);
	d_stat.upload(host.stat);
	d_edge.upload(host.edge);
	const size_t R16 = ((size_t)cap*N + 15)/16*16;
	w.alloc(NPARAM); wq.alloc(NPARAM); am.alloc(NPARAM); av.alloc(NPARAM); grad.alloc(NPARAM);
	FVX_HIP(hipMemset(w.p, 0, NPARAM*sizeof(float)));
	FVX_HIP(hipMemset(am.p, 0, NPARAM*sizeof(float)));
	FVX_HIP(hipMemset(av.p, 0, NPARAM*sizeof(float)));
	frag.alloc((size_t)NL*4*FRAG_MAT);
	for(int l = 0; l <= NL; l++) h[l].alloc(R16*W);
	for(int l = 0; l < NL; l++) {
		pq[l].alloc(R16*2*W);
		ab[l].alloc(R16*W);
		umask[l].alloc(R16);
	}
	g0.alloc(R16*W); g1.alloc(R16*W); gpart.alloc(R16*W); da.alloc(R16*W);
	dpre.alloc(R16*W); dpq.alloc(R16*2*W);
	y.alloc(R16*FO);
	part.alloc((size_t)RED_BLOCKS*(2*W*W + W));
	gfeat.alloc(2*(size_t)cap);
	pp.alloc(3*(size_t)cap);
	ids.alloc(cap);
	loss.alloc(1);
	dirty = true;
}

static Graph make_graph(const Mpnn& n, int G)
{
	Graph g;
	g.N = n.N; g.R = G*n.N;
	g.nbr = n.d_nbr.p; g.rev = n.d_rev.p; g.invdeg = n.d_invdeg.p; g.stat = n.d_stat.p;
	g.edge = n.d_edge.p; g.gfeat = n.gfeat.p;
	return g;
}

void Mpnn::set_points(int G, const double* mach, const double* aoa, double gamma, hipStream_t s)
{
	if(G < 1 || G > cap) throw Error(FVX_ERR_INVALID, "graph count outside [1, max_graphs]");
	this->gamma = gamma;
	npoints = G;
	hgf.resize(2*(size_t)G);
	hpp.resize(3*(size_t)G);
	for(int i = 0; i < G; i++) {
		if(!(mach[i] > 0) || !std::isfinite(mach[i]) || !std::isfinite(aoa[i]))
			throw Error(FVX_ERR_INVALID, "invalid Mach number or angle of attack");
		const double a = aoa[i];
		hgf[2*i] = (float)(4.0*(mach[i] - 0.5));
		hgf[2*i + 1] = (float)(a*(180.0/M_PI)/5.0);
		hpp[3*i] = std::cos(a); hpp[3*i + 1] = std::sin(a);
		hpp[3*i + 2] = 1.0/(gamma*mach[i]*mach[i]);
	}
	FVX_HIP(hipMemcpyAsync(gfeat.p, hgf.data(), hgf.size()*sizeof(float), hipMemcpyHostToDevice, s));
	FVX_HIP(hipMemcpyAsync(pp.p, hpp.data(), hpp.size()*sizeof(double), hipMemcpyHostToDevice, s));
}

void Mpnn::pack(hipStream_t s)
{
	if(!dirty) return;
	const int n = std::max(NPARAM, NL*4*FRAG_MAT);
	k_pack<<<blocks_for(n), RB, 0, s>>>(w.p, wq.p, frag.p);
	FVX_HIP(hipGetLastError());
	dirty = false;
}

static int gemm_grid(int R) { return std::max(1, std::min(GEMM_BLOCKS, ((R + 15)/16 + 3)/4)); }

void Mpnn::forward(int G, bool train, hipStream_t s)
{
	pack(s);
	const Graph g = make_graph(*this, G);
	const int R = g.R;
	k_encode<<<blocks_for((size_t)R*W), RB, 0, s>>>(g, wq.p, h[0].p);
	for(int l = 0; l < NL; l++) {
		const bf16_t* fl = frag.p + (size_t)l*4*FRAG_MAT;
		k_pq<<<gemm_grid(R), RB, 0, s>>>(R, h[l].p, fl + M_PQ*FRAG_MAT, pq[l].p);
		k_update<<<gemm_grid(R), RB, 0, s>>>(g, h[l].p, pq[l].p, wq.p + OFF_L0 + l*LSZ,
		                                     fl + M_UPD*FRAG_MAT, h[l + 1].p,
		                                     train ? ab[l].p : nullptr,
		                                     train ? umask[l].p : nullptr);
	}
	FVX_HIP(hipGetLastError());
}

void Mpnn::predict(int G, float* yout, hipStream_t s)
{
	forward(G, false, s);
	const Graph g = make_graph(*this, G);
	k_decode<<<RED_BLOCKS, RB, 0, s>>>(g, h[NL].p, wq.p, yout, nullptr, nullptr, 0.0, nullptr,
	                                   nullptr);
	FVX_HIP(hipGetLastError());
}

void Mpnn::gradient(int G, const float* targets, const int* sample_ids, hipStream_t s)
{
	for(int i = 0; i < G; i++)
		if(sample_ids[i] < 0) throw Error(FVX_ERR_INVALID, "negative sample index");
	hids.assign(sample_ids, sample_ids + G);
	FVX_HIP(hipMemcpyAsync(ids.p, hids.data(), G*sizeof(int), hipMemcpyHostToDevice, s));
	forward(G, true, s);
	const Graph g = make_graph(*this, G);
	const int R = g.R;
	float* gc = g0.p;
	float* gn = g1.p;
	k_decode<<<RED_BLOCKS, RB, 0, s>>>(g, h[NL].p, wq.p, nullptr, targets, ids.p,
	                                   1.0/((double)R*FO), gc, part.p);
	k_reduce<<<reduce_grid(261), RB, 0, s>>>(RED_BLOCKS, 261, 261, 0, OFF_WDEC, 0, part.p, grad.p,
	                                        loss.p);
	const int nchunk = (R + CH - 1)/CH;
	const int nwb = std::min(RED_BLOCKS, nchunk);
	const int cpb = (nchunk + nwb - 1)/nwb;
	const int nwb_used = (nchunk + cpb - 1)/cpb;
	for(int l = NL - 1; l >= 0; l--) {
		const bf16_t* fl = frag.p + (size_t)l*4*FRAG_MAT;
		const float* wl = wq.p + OFF_L0 + l*LSZ;
		const int base = OFF_L0 + l*LSZ;
		k_update_bwd<<<gemm_grid(R), RB, 0, s>>>(R, gc, umask[l].p, fl + M_BUPD*FRAG_MAT, dpre.p,
		                                         gpart.p, da.p);
		k_message_bwd<<<RED_BLOCKS, RB, 0, s>>>(g, pq[l].p, da.p, wl, dpq.p, part.p);
		k_reduce<<<reduce_grid(320), RB, 0, s>>>(RED_BLOCKS, 320, 320, 0, base + L_WE, 0, part.p,
		                                        grad.p, nullptr);
		k_pq_bwd<<<gemm_grid(R), RB, 0, s>>>(R, dpq.p, fl + M_BPQ*FRAG_MAT, gpart.p, gn);
		k_wgrad<W, 2*W, false><<<nwb_used, RB, 0, s>>>(R, cpb, h[l].p, nullptr, dpq.p, part.p);
		k_reduce<<<reduce_grid(W*2*W), RB, 0, s>>>(nwb_used, W*2*W, 2*W, 1, base + L_WA, base + L_WB,
		                                          part.p, grad.p, nullptr);
		k_wgrad<2*W, W, true><<<nwb_used, RB, 0, s>>>(R, cpb, h[l].p, ab[l].p, dpre.p, part.p);
		k_reduce<<<reduce_grid(2*W*W + W), RB, 0, s>>>(nwb_used, 2*W*W + W, W, 2, base + L_WU,
		                                              base + L_WV, part.p, grad.p, nullptr);
		std::swap(gc, gn);
	}
	k_encode_bwd<<<RED_BLOCKS, RB, 0, s>>>(g, gc, part.p);
	k_reduce<<<reduce_grid(576), RB, 0, s>>>(RED_BLOCKS, 576, 576, 0, OFF_WENC, 0, part.p, grad.p,
	                                        nullptr);
	FVX_HIP(hipGetLastError());
}

void Mpnn::adam_step(double lr, hipStream_t s)
{
	step++;
	const double b1 = 0.9, b2 = 0.999;
	const float bc1 = (float)(1.0 - std::pow(b1, step)), bc2 = (float)(1.0 - std::pow(b2, step));
	k_adam<<<blocks_for(NPARAM), RB, 0, s>>>(NPARAM, w.p, am.p, av.p, grad.p, (float)lr, (float)b1,
	                                         (float)b2, (float)(1 - b1), (float)(1 - b2), 1e-8f,
	                                         bc1, bc2);
	FVX_HIP(hipGetLastError());
	dirty = true;
}

void Mpnn::targets(int G, const double* u, float* t, hipStream_t s)
{
	k_targets<<<blocks_for((size_t)G*N), RB, 0, s>>>(G*N, N, u, pp.p, gamma, t);
	FVX_HIP(hipGetLastError());
}

void Mpnn::warm(int G, double floor, double* u, hipStream_t s)
{
	predict(G, y.p, s);
	k_warm<<<blocks_for((size_t)G*N), RB, 0, s>>>(G*N, N, y.p, pp.p, gamma, floor, u);
	FVX_HIP(hipGetLastError());
}

} // namespace fvx
