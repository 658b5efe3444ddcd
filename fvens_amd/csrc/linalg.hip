/* linalg.hip -- kernels of the implicit solver: block-sparse matrix, preconditioners and the
 * vector operations of the Krylov solver.
 *
 * Matrix storage ("slot blocks", shared with jacobian.hip): vals[cell][slot][4][4] row-major,
 * slot 0 the diagonal block, slot 1+k the block coupling the cell to nbr[cell][k]; zero where
 * that slot has no interior neighbour. A preconditioner stores its factors in the same layout
 * with slot 0 holding the INVERTED diagonal block.
 *
 * Multicolour sweeps: cells are coloured so that no two cells of one colour share a face, and
 * the preconditioner sweeps colour by colour. Within a colour every row depends only on rows of
 * other colours, so one launch processes a whole colour; the result equals the sequential
 * sweep over the cells sorted by (colour, index). lowmask[cell] bit k marks neighbour slot k as
 * belonging to a lower colour.
 *
 * Reductions (dot products, norms) are deterministic: fixed grid, fixed per-thread order, fixed
 * tree within the block, then one block sums the per-block partials in order.
 */
#include "fvx_internal.hpp"
#include "wave.hpp"

namespace fvx {

/* ----------------------------------------------------------------- block helpers ------- */

template <class VT> __device__ __forceinline__ void load_row(const VT* p, double* v);
template <> __device__ __forceinline__ void load_row<double>(const double* p, double* v)
{
	const double2 a = reinterpret_cast<const double2*>(p)[0];
	const double2 b = reinterpret_cast<const double2*>(p)[1];
	v[0] = a.x; v[1] = a.y; v[2] = b.x; v[3] = b.y;
}
template <> __device__ __forceinline__ void load_row<float>(const float* p, double* v)
{
	const float4 a = *reinterpret_cast<const float4*>(p);
	v[0] = a.x; v[1] = a.y; v[2] = a.z; v[3] = a.w;
}
/// bfloat16 rows: 4 x 2 bytes (8-byte aligned: rows start at multiples of 4 values)
__device__ __forceinline__ double bf16_value(unsigned int h)
{
	return (double)__uint_as_float(h << 16);
}
template <> __device__ __forceinline__ void load_row<bf16>(const bf16* p, double* v)
{
	const uint2 w = *reinterpret_cast<const uint2*>(p);
	v[0] = bf16_value(w.x & 0xffffu); v[1] = bf16_value(w.x >> 16);
	v[2] = bf16_value(w.y & 0xffffu); v[3] = bf16_value(w.y >> 16);
}
/// FP64 -> FP32 (rounded) -> bfloat16, rounded to nearest even; NaN stays NaN
__device__ __forceinline__ unsigned int bf16_bits(double x)
{
	const unsigned int u = __float_as_uint((float)x);
	if((u & 0x7fffffffu) > 0x7f800000u) return (u >> 16) | 0x40u;
	return (u + 0x7fffu + ((u >> 16) & 1u)) >> 16;
}
template <class VT> __device__ __forceinline__ void put_row(VT* p, const double* v);
template <> __device__ __forceinline__ void put_row<bf16>(bf16* p, const double* v)
{
	*reinterpret_cast<uint2*>(p) = make_uint2(bf16_bits(v[0]) | (bf16_bits(v[1]) << 16),
	                                          bf16_bits(v[2]) | (bf16_bits(v[3]) << 16));
}
template <> __device__ __forceinline__ void put_row<double>(double* p, const double* v)
{
	reinterpret_cast<double2*>(p)[0] = make_double2(v[0], v[1]);
	reinterpret_cast<double2*>(p)[1] = make_double2(v[2], v[3]);
}
template <> __device__ __forceinline__ void put_row<float>(float* p, const double* v)
{
	*reinterpret_cast<float4*>(p) = make_float4((float)v[0], (float)v[1], (float)v[2],
	                                            (float)v[3]);
}
template <class VT> __device__ __forceinline__ void load_block(const VT* p, double* b)
{
	for(int r = 0; r < 4; r++) load_row<VT>(p + 4*r, b + 4*r);
}
template <class VT> __device__ __forceinline__ void put_block(VT* p, const double* b)
{
	for(int r = 0; r < 4; r++) put_row<VT>(p + 4*r, b + 4*r);
}

__device__ __forceinline__ double dot4(const double* a, const double* x)
{
	return a[0]*x[0] + a[1]*x[1] + a[2]*x[2] + a[3]*x[3];
}

/// C = A B for 4x4 row-major blocks.
__device__ __forceinline__ void gemm4(const double* A, const double* B, double* C)
{
	for(int i = 0; i < 4; i++)
		for(int j = 0; j < 4; j++)
			C[i*4+j] = A[i*4]*B[j] + A[i*4+1]*B[4+j] + A[i*4+2]*B[8+j] + A[i*4+3]*B[12+j];
}

/// Inverse of a 4x4 block by Gauss-Jordan elimination with partial pivoting; row exchanges
/// are done with selects so that every index is static (no register-array spills).
__device__ __forceinline__ void invert4(const double* a, double* inv)
{
	double m[4][8];
	for(int i = 0; i < 4; i++)
		for(int j = 0; j < 8; j++) m[i][j] = j < 4 ? a[i*4+j] : (j - 4 == i ? 1.0 : 0.0);
	for(int c = 0; c < 4; c++) {
		int p = c;
		double best = fabs(m[c][c]);
		for(int r = c + 1; r < 4; r++)
			if(fabs(m[r][c]) > best) { best = fabs(m[r][c]); p = r; }
		for(int r = c + 1; r < 4; r++) {
			const bool sw = p == r;
			for(int j = 0; j < 8; j++) {
				const double x = m[c][j], y = m[r][j];
				m[c][j] = sw ? y : x;
				m[r][j] = sw ? x : y;
			}
		}
		const double piv = m[c][c];
		for(int j = 0; j < 8; j++) m[c][j] /= piv;
		for(int r = 0; r < 4; r++) {
			if(r == c) continue;
			const double f = m[r][c];
			for(int j = 0; j < 8; j++) m[r][j] -= f*m[c][j];
		}
	}
	for(int i = 0; i < 4; i++)
		for(int j = 0; j < 4; j++) inv[i*4+j] = m[i][j+4];
}

/// Value of lane k of the quad, on every lane of the quad.
template <int K> __device__ __forceinline__ double quad_bcast(double v)
{
	return dpp_d<K*0x55>(v);
}

__device__ __forceinline__ int4 load_nbr(const int* nbr, int c)
{
	return reinterpret_cast<const int4*>(nbr)[c];
}

/* The kernels below take an instance list `ilist` (null: one instance, index 0) and run
 * gridDim.y instances; every per-instance array argument comes with its instance stride. An
 * entry -1 is a retired slot of a list compacted on the device (the FGMRES cycle below); its
 * workgroups return at once. */
__device__ __forceinline__ long inst_of(const int* ilist)
{
	return ilist ? ilist[blockIdx.y] : 0;
}

/* ------------------------------------------------------------------ matrix ------------- */

/// y = A x (or y = b - A x with RES); one quad per cell, lane q computes row q.
template <class VT, bool RES>
__global__ __launch_bounds__(BLOCK)
void k_spmv(int n, const int* __restrict__ ilist, const VT* __restrict__ vals, long sa,
            const int* __restrict__ nbr, const double* __restrict__ x, long sx,
            const double* __restrict__ b, long sb, double* __restrict__ y, long sy)
{
	const long I = inst_of(ilist);
	if(I < 0) return;                     // a retired slot of a compacted list
	vals += I*sa; x += I*sx; y += I*sy;
	if constexpr(RES) b += I*sb;
	const int q = threadIdx.x & 3;
	const int c = xcd_block()*CELLS_PER_BLOCK + (threadIdx.x >> 2);
	if(c >= n) return;
	const int4 nb = load_nbr(nbr, c);
	const int nbs[4] = { nb.x, nb.y, nb.z, nb.w };
	const VT* blk = vals + (size_t)c*80 + 4*q;
	double a[4], xv[4];
	load_row<VT>(blk, a);
	load4(x + 4*c, xv);
	double acc = dot4(a, xv);
	for(int k = 0; k < 4; k++) {
		if(nbs[k] < 0) continue;
		load_row<VT>(blk + 16*(1 + k), a);
		load4(x + 4*nbs[k], xv);
		acc += dot4(a, xv);
	}
	if constexpr(RES) y[4*c + q] = b[4*c + q] - acc;
	else y[4*c + q] = acc;
}

/// y += x
__global__ __launch_bounds__(BLOCK)
void k_axpy1(long N, const int* __restrict__ ilist, const double* __restrict__ x, long sx,
             double* __restrict__ y, long sy)
{
	const long I = inst_of(ilist);
	if(I < 0) return;                     // a retired slot of a compacted list
	x += I*sx; y += I*sy;
	const long stride = (long)gridDim.x*blockDim.x;
	for(long e = (long)blockIdx.x*blockDim.x + threadIdx.x; e < N; e += stride) y[e] += x[e];
}

/// A_ii += dtm_i I with dtm_i = area_i/(cfl dt_i).
template <class VT>
__global__ __launch_bounds__(BLOCK)
void k_pseudotime(int n, VT* __restrict__ vals, const double* __restrict__ dt,
                  const double* __restrict__ area, double cfl, double* __restrict__ dtm)
{
	const int q = threadIdx.x & 3;
	const int c = xcd_block()*CELLS_PER_BLOCK + (threadIdx.x >> 2);
	if(c >= n) return;
	const double add = area[c]/(cfl*dt[c]);
	if(q == 0) dtm[c] = add;
	VT* p = vals + (size_t)c*80 + 5*q;
	*p = (VT)((double)*p + add);
}

/* ---------------------------------------------------------- preconditioners ------------ */

/// Inverts the diagonal block of every cell into fac slot 0; with copy_off, also copies the
/// off-diagonal blocks (symmetric Gauss-Seidel sweeps read them from the factor array).
template <class VTA, class VTF>
__global__ __launch_bounds__(BLOCK)
void k_invert_diag(int n, const int* __restrict__ ilist, const VTA* __restrict__ vals,
                   VTF* __restrict__ fac, long sa, int copy_off)
{
	const long I = inst_of(ilist);
	if(I < 0) return;                     // a retired slot of a compacted list
	vals += I*sa; fac += I*sa;
	const int c = blockIdx.x*blockDim.x + threadIdx.x;
	if(c >= n) return;
	double a[16], inv[16];
	load_block<VTA>(vals + (size_t)c*80, a);
	invert4(a, inv);
	put_block<VTF>(fac + (size_t)c*80, inv);
	if(copy_off)
		for(int s = 1; s < 5; s++) {
			load_block<VTA>(vals + (size_t)c*80 + 16*s, a);
			put_block<VTF>(fac + (size_t)c*80 + 16*s, a);
		}
}

/// z = D^-1 r, one quad per cell.
template <class VTF>
__global__ __launch_bounds__(BLOCK)
void k_jacobi_apply(int n, const int* __restrict__ ilist, const VTF* __restrict__ fac, long sa,
                    const double* __restrict__ r, long sr, double* __restrict__ z, long sz)
{
	const long I = inst_of(ilist);
	if(I < 0) return;                     // a retired slot of a compacted list
	fac += I*sa; r += I*sr; z += I*sz;
	const int q = threadIdx.x & 3;
	const int c = xcd_block()*CELLS_PER_BLOCK + (threadIdx.x >> 2);
	if(c >= n) return;
	double a[4], rv[4];
	load_row<VTF>(fac + (size_t)c*80 + 4*q, a);
	load4(r + 4*c, rv);
	z[4*c + q] = dot4(a, rv);
}

/// One colour of a block Gauss-Seidel pass (PETSc SOR with omega = 1):
/// z_i = D_i^-1 (r_i - sum_{j != i} A_ij z_j) with the current z (ALL), or with the lower
/// colours only (the first forward pass from z = 0, where the other colours hold no value yet).
template <class VTF, bool ALL>
__global__ __launch_bounds__(BLOCK)
void k_sgs_colour(int cnt, const int* __restrict__ list, const int* __restrict__ ilist,
                  const VTF* __restrict__ fac, long sa, const int* __restrict__ nbr,
                  const unsigned char* __restrict__ lowmask, const double* __restrict__ r,
                  long sr, double* __restrict__ z, long sz)
{
	const long I = inst_of(ilist);
	if(I < 0) return;                     // a retired slot of a compacted list
	fac += I*sa; r += I*sr; z += I*sz;
	const int q = threadIdx.x & 3;
	const int idx = xcd_block()*CELLS_PER_BLOCK + (threadIdx.x >> 2);
	const bool valid = idx < cnt;
	const int i = valid ? list[idx] : list[0];
	const int4 nb = load_nbr(nbr, i);
	const int nbs[4] = { nb.x, nb.y, nb.z, nb.w };
	const int lm = lowmask[i];
	const VTF* blk = fac + (size_t)i*80 + 4*q;
	double t = r[4*i + q];
	double a[4], zv[4];
	for(int k = 0; k < 4; k++) {
		if(nbs[k] < 0 || !(ALL || ((lm >> k) & 1))) continue;
		load_row<VTF>(blk + 16*(1 + k), a);
		load4(z + 4*nbs[k], zv);
		t -= dot4(a, zv);
	}
	const double tv[4] = { quad_bcast<0>(t), quad_bcast<1>(t), quad_bcast<2>(t),
	                       quad_bcast<3>(t) };
	load_row<VTF>(blk, a);
	if(valid) z[4*i + q] = dot4(a, tv);
}

/// The same colour pass with the colour's cells given as records {cell, lower-colour mask,
/// neighbour slots 0-3} (two int4 per position): the identical arithmetic, one dependent load
/// less (the cell's neighbour slots come with its index instead of after it).
template <class VTF, bool ALL>
__global__ __launch_bounds__(BLOCK)
void k_sgs_rec(int cnt, const int4* __restrict__ rec, const int* __restrict__ ilist,
               const VTF* __restrict__ fac, long sa, const double* __restrict__ r, long sr,
               double* __restrict__ z, long sz)
{
	const long I = inst_of(ilist);
	if(I < 0) return;                     // a retired slot of a compacted list
	fac += I*sa; r += I*sr; z += I*sz;
	const int q = threadIdx.x & 3;
	const int idx = xcd_block()*CELLS_PER_BLOCK + (threadIdx.x >> 2);
	const bool valid = idx < cnt;
	const int e = valid ? idx : 0;
	const int4 ra = rec[2*e], rb = rec[2*e + 1];
	const int i = ra.x, lm = ra.y;
	const int nbs[4] = { ra.z, ra.w, rb.x, rb.y };
	const VTF* blk = fac + (size_t)i*80 + 4*q;
	double t = r[4*i + q];
	double a[4], zv[4];
	for(int k = 0; k < 4; k++) {
		if(nbs[k] < 0 || !(ALL || ((lm >> k) & 1))) continue;
		load_row<VTF>(blk + 16*(1 + k), a);
		load4(z + 4*nbs[k], zv);
		t -= dot4(a, zv);
	}
	const double tv[4] = { quad_bcast<0>(t), quad_bcast<1>(t), quad_bcast<2>(t),
	                       quad_bcast<3>(t) };
	load_row<VTF>(blk, a);
	if(valid) z[4*i + q] = dot4(a, tv);
}

/// Forward substitution of block ILU(0) for one colour: y_i = r_i - sum_{lower k} L_ik y_k.
template <class VTF>
__global__ __launch_bounds__(BLOCK)
void k_ilu_fwd(int cnt, const int* __restrict__ list, const int* __restrict__ ilist,
               const VTF* __restrict__ fac, long sa, const int* __restrict__ nbr,
               const unsigned char* __restrict__ lowmask, const double* __restrict__ r, long sr,
               double* __restrict__ z, long sz)
{
	const long I = inst_of(ilist);
	if(I < 0) return;                     // a retired slot of a compacted list
	fac += I*sa; r += I*sr; z += I*sz;
	const int q = threadIdx.x & 3;
	const int idx = xcd_block()*CELLS_PER_BLOCK + (threadIdx.x >> 2);
	if(idx >= cnt) return;
	const int i = list[idx];
	const int4 nb = load_nbr(nbr, i);
	const int nbs[4] = { nb.x, nb.y, nb.z, nb.w };
	const int lm = lowmask[i];
	const VTF* blk = fac + (size_t)i*80 + 4*q;
	double t = r[4*i + q];
	double a[4], zv[4];
	for(int k = 0; k < 4; k++) {
		if(!((lm >> k) & 1)) continue;
		load_row<VTF>(blk + 16*(1 + k), a);
		load4(z + 4*nbs[k], zv);
		t -= dot4(a, zv);
	}
	z[4*i + q] = t;
}

/// Backward substitution of block ILU(0) for one colour, in place on z (which holds y):
/// z_i = D_i^-1 (y_i - sum_{upper j} U_ij z_j).
template <class VTF>
__global__ __launch_bounds__(BLOCK)
void k_ilu_bwd(int cnt, const int* __restrict__ list, const int* __restrict__ ilist,
               const VTF* __restrict__ fac, long sa, const int* __restrict__ nbr,
               const unsigned char* __restrict__ lowmask, double* __restrict__ z, long sz)
{
	const long I = inst_of(ilist);
	if(I < 0) return;                     // a retired slot of a compacted list
	fac += I*sa; z += I*sz;
	const int q = threadIdx.x & 3;
	const int idx = xcd_block()*CELLS_PER_BLOCK + (threadIdx.x >> 2);
	const bool valid = idx < cnt;
	const int i = valid ? list[idx] : list[0];
	const int4 nb = load_nbr(nbr, i);
	const int nbs[4] = { nb.x, nb.y, nb.z, nb.w };
	const int lm = lowmask[i];
	const VTF* blk = fac + (size_t)i*80 + 4*q;
	double t = z[4*i + q];
	double a[4], zv[4];
	for(int k = 0; k < 4; k++) {
		if(nbs[k] < 0 || ((lm >> k) & 1)) continue;
		load_row<VTF>(blk + 16*(1 + k), a);
		load4(z + 4*nbs[k], zv);
		t -= dot4(a, zv);
	}
	const double tv[4] = { quad_bcast<0>(t), quad_bcast<1>(t), quad_bcast<2>(t),
	                       quad_bcast<3>(t) };
	load_row<VTF>(blk, a);
	if(valid) z[4*i + q] = dot4(a, tv);
}

/* Line Gauss-Seidel. The vertices are the lines of strongly coupled cells and every other cell
 * on its own; vertex v holds cells vcells[vptr[v] .. vptr[v+1]) in line order, and lineinfo[c]
 * gives the slot of the cell's line predecessor (bits 0-2) and successor (bits 3-5), 7 for none.
 * The tridiagonal block system of a line is factored by the block Thomas algorithm:
 *   D'_0 = D_0,  D'_k = D_k - L_k C_{k-1},  C_k = D'_k^-1 U_k,
 * with L_k the block coupling cell k to k-1 and U_k the block coupling k to k+1. The factor
 * array keeps the matrix layout: slot 0 holds D'_k^-1, the successor slot C_k, every other
 * slot the matrix block. */
__device__ __forceinline__ int line_prev(unsigned char li) { return li & 7; }
__device__ __forceinline__ int line_next(unsigned char li) { return (li >> 3) & 7; }

template <class VTA, class VTF>
__global__ __launch_bounds__(BLOCK)
void k_line_factor(int nv, const int* __restrict__ vptr, const int* __restrict__ vcells,
                   const unsigned char* __restrict__ lineinfo, const int* __restrict__ ilist,
                   const VTA* __restrict__ vals, VTF* __restrict__ fac, long sa)
{
	const long I = inst_of(ilist);
	if(I < 0) return;                     // a retired slot of a compacted list
	vals += I*sa; fac += I*sa;
	const int v = blockIdx.x*blockDim.x + threadIdx.x;
	if(v >= nv) return;
	double Cp[16] = {};                              // C_{k-1}
	for(int k = vptr[v]; k < vptr[v+1]; k++) {
		const int c = vcells[k];
		const unsigned char li = lineinfo[c];
		const int ps = line_prev(li), ns = line_next(li);
		const VTA* a = vals + (size_t)c*80;
		VTF* f = fac + (size_t)c*80;
		double D[16], B[16], T[16], Dinv[16];
		load_block<VTA>(a, D);
		for(int s = 0; s < 4; s++) {
			if(s == ns) continue;
			load_block<VTA>(a + 16*(1 + s), B);
			put_block<VTF>(f + 16*(1 + s), B);
			if(s == ps) {
				gemm4(B, Cp, T);
				for(int i = 0; i < 16; i++) D[i] -= T[i];
			}
		}
		invert4(D, Dinv);
		put_block<VTF>(f, Dinv);
		if(ns < 4) {
			load_block<VTA>(a + 16*(1 + ns), B);
			gemm4(Dinv, B, Cp);
			put_block<VTF>(f + 16*(1 + ns), Cp);
		}
	}
}

/// One colour of a line Gauss-Seidel pass, one thread per vertex: right-hand side
/// r_c - sum over slots outside the line of A z (current z with ALL, lower colours only in the
/// first forward pass from z = 0), then the forward and backward Thomas sweeps along the line.
/// z of the vertex's cells holds the forward result between the two sweeps.
template <class VTF, bool ALL>
__global__ __launch_bounds__(BLOCK)
void k_line_colour(int cnt, const int* __restrict__ list, const int* __restrict__ vptr,
                   const int* __restrict__ vcells, const unsigned char* __restrict__ lineinfo,
                   const int* __restrict__ ilist, const VTF* __restrict__ fac, long sa,
                   const int* __restrict__ nbr, const unsigned char* __restrict__ lowmask,
                   const double* __restrict__ r, long sr, double* __restrict__ z, long sz)
{
	const long I = inst_of(ilist);
	if(I < 0) return;                     // a retired slot of a compacted list
	fac += I*sa; r += I*sr; z += I*sz;
	const int idx = blockIdx.x*blockDim.x + threadIdx.x;
	if(idx >= cnt) return;
	const int v = list[idx];
	const int k0 = vptr[v], k1 = vptr[v+1];
	double g[4] = {0, 0, 0, 0};
	for(int k = k0; k < k1; k++) {
		const int c = vcells[k];
		const unsigned char li = lineinfo[c];
		const int ps = line_prev(li), ns = line_next(li);
		const int4 nb = load_nbr(nbr, c);
		const int nbs[4] = { nb.x, nb.y, nb.z, nb.w };
		const int lm = lowmask[c];
		const VTF* f = fac + (size_t)c*80;
		double t[4], a[4], zv[4];
		load4(r + 4*c, t);
		for(int s = 0; s < 4; s++) {
			if(nbs[s] < 0 || s == ns) continue;
			const bool prev = s == ps;
			if(!prev && !(ALL || ((lm >> s) & 1))) continue;
			if(prev) for(int i = 0; i < 4; i++) zv[i] = g[i];
			else load4(z + 4*nbs[s], zv);
			for(int i = 0; i < 4; i++) {
				load_row<VTF>(f + 16*(1 + s) + 4*i, a);
				t[i] -= dot4(a, zv);
			}
		}
		for(int i = 0; i < 4; i++) {
			load_row<VTF>(f + 4*i, a);
			g[i] = dot4(a, t);
		}
		store4(z + 4*c, g);
	}
	// backward: x_k = g_k - C_k x_{k+1}; g holds x of the last cell
	for(int k = k1 - 2; k >= k0; k--) {
		const int c = vcells[k];
		const int ns = line_next(lineinfo[c]);
		const VTF* f = fac + (size_t)c*80 + 16*(1 + ns);
		double gk[4], a[4], x[4];
		load4(z + 4*c, gk);
		for(int i = 0; i < 4; i++) {
			load_row<VTF>(f + 4*i, a);
			x[i] = gk[i] - dot4(a, g);
		}
		store4(z + 4*c, x);
		for(int i = 0; i < 4; i++) g[i] = x[i];
	}
}

/// Block ILU(0) factorisation of the rows of one colour (IKJ form on the matrix pattern,
/// rows taken in (colour, index) order; one thread per row). For each lower neighbour k in
/// ascending order: L_ik = A_ik U_kk^-1, then A_ij -= L_ik U_kj for every j > k in the
/// pattern of both rows (the diagonal j = i always; other j only where three cells are
/// mutually adjacent). Finally slot 0 receives inv(A_ii).
template <class VTA, class VTF>
__global__ __launch_bounds__(BLOCK)
void k_ilu_factor(int cnt, const int* __restrict__ list, const int* __restrict__ ilist,
                  const VTA* __restrict__ vals, VTF* __restrict__ fac, long sa,
                  const int* __restrict__ nbr, const unsigned char* __restrict__ lowmask,
                  const int* __restrict__ colour, int n)
{
	const long I = inst_of(ilist);
	if(I < 0) return;                     // a retired slot of a compacted list
	vals += I*sa; fac += I*sa;
	const int idx = blockIdx.x*blockDim.x + threadIdx.x;
	if(idx >= cnt) return;
	const int i = list[idx];
	const int4 nb4 = load_nbr(nbr, i);
	const int nb[4] = { nb4.x, nb4.y, nb4.z, nb4.w };
	const int lm = lowmask[i];
	double A[5][16];
	for(int s = 0; s < 5; s++) load_block<VTA>(vals + (size_t)i*80 + 16*s, A[s]);
	// keys (colour, index) of the neighbours; lower slots sorted ascending by key
	long long key[4];
	for(int k = 0; k < 4; k++)
		key[k] = nb[k] >= 0 ? (long long)colour[nb[k]]*n + nb[k] : -1;
	const long long BIG = 0x7fffffffffffffffLL;
	long long sk[4];
	int ss[4];
	for(int k = 0; k < 4; k++) { sk[k] = ((lm >> k) & 1) ? key[k] : BIG; ss[k] = k; }
	// sorting network on 4 (key, slot) pairs
#define FVX_CSWAP(a, b) { const bool sw = sk[b] < sk[a]; \
		const long long tk = sw ? sk[b] : sk[a]; const long long uk = sw ? sk[a] : sk[b]; \
		const int ts = sw ? ss[b] : ss[a]; const int us = sw ? ss[a] : ss[b]; \
		sk[a] = tk; sk[b] = uk; ss[a] = ts; ss[b] = us; }
	FVX_CSWAP(0, 1) FVX_CSWAP(2, 3) FVX_CSWAP(0, 2) FVX_CSWAP(1, 3) FVX_CSWAP(1, 2)
#undef FVX_CSWAP
	for(int t = 0; t < 4; t++) {
		if(sk[t] == BIG) break;
		const int k = ss[t];
		const int kc = nb[k];
		const long long kkey = sk[t];
		double Dk[16], Ak[16], L[16];
		load_block<VTF>(fac + (size_t)kc*80, Dk);
		for(int e = 0; e < 16; e++) {
			double v = A[1][e];
			for(int s = 1; s < 4; s++) v = k == s ? A[1+s][e] : v;
			Ak[e] = v;
		}
		gemm4(Ak, Dk, L);
		for(int s = 0; s < 4; s++)
			if(k == s) for(int e = 0; e < 16; e++) A[1+s][e] = L[e];
		const int4 kn4 = load_nbr(nbr, kc);
		const int kn[4] = { kn4.x, kn4.y, kn4.z, kn4.w };
		// j = i: the diagonal
		{
			int ti = 0;
			for(int s = 1; s < 4; s++) if(kn[s] == i) ti = s;
			double U[16], P[16];
			load_block<VTF>(fac + (size_t)kc*80 + 16*(1 + ti), U);
			gemm4(L, U, P);
			for(int e = 0; e < 16; e++) A[0][e] -= P[e];
		}
		// j = another neighbour of i that follows k and is also a neighbour of k
		for(int s = 0; s < 4; s++) {
			const int j = nb[s];
			if(s == k || j < 0 || key[s] <= kkey) continue;
			int tj = -1;
			for(int r = 0; r < 4; r++) if(kn[r] == j) tj = r;
			if(tj < 0) continue;
			double U[16], P[16];
			load_block<VTF>(fac + (size_t)kc*80 + 16*(1 + tj), U);
			gemm4(L, U, P);
			for(int e = 0; e < 16; e++) A[1+s][e] -= P[e];
		}
	}
	double inv[16];
	invert4(A[0], inv);
	put_block<VTF>(fac + (size_t)i*80, inv);
	for(int s = 1; s < 5; s++) put_block<VTF>(fac + (size_t)i*80 + 16*s, A[s]);
}

/// Block Jacobi with ILU(0) per subdomain, one application z = M^-1 r. Workgroup b holds
/// subdomain b, whose cells bcells[bptr[b] .. bptr[b+1]) are listed by (level, index); its
/// level boundaries are blv[blvptr[b] .. blvptr[b+1]]. ploc gives, per listed cell and
/// neighbour slot, the neighbour's position in the subdomain's list (-1 outside it), so a
/// position below the cell's own is a lower-numbered neighbour (an earlier level) and one
/// above it a higher-numbered one. The vector of the subdomain stays in LDS: the forward
/// substitution y_i = r_i - sum_{lower k} L_ik y_k runs level by level upwards, the backward
/// substitution z_i = U_ii^-1 (y_i - sum_{upper j} U_ij z_j) level by level downwards, one
/// quad per cell (lane q computes row q), a workgroup barrier between levels. The factors are
/// those of k_ilu_factor on the subdomain-restricted pattern: L_ik in the lower slots, U_ij in
/// the upper slots, inv(U_ii) in slot 0.
template <class VTF>
__global__ __launch_bounds__(BLOCK)
void k_bj_ilu_apply(const int* __restrict__ bptr, const int* __restrict__ blvptr,
                    const int* __restrict__ blv, const int* __restrict__ bcells,
                    const int4* __restrict__ ploc, const int* __restrict__ ilist,
                    const VTF* __restrict__ fac, long sa, const double* __restrict__ r, long sr,
                    double* __restrict__ z, long sz)
{
	const long I = inst_of(ilist);
	if(I < 0) return;                     // a retired slot of a compacted list
	fac += I*sa; r += I*sr; z += I*sz;
	const int b = blockIdx.x;
	const int c0 = bptr[b];
	const int l0 = blvptr[b], nl = blvptr[b+1] - l0;
	__shared__ double zs[BJ_MAX_CELLS*4];
	const int q = threadIdx.x & 3, qd = threadIdx.x >> 2;
	constexpr int NQ = BLOCK/4;
	for(int l = 0; l < nl; l++) {
		const int e0 = blv[l0 + l], e1 = blv[l0 + l + 1];
		for(int idx = e0 + qd; idx < e1; idx += NQ) {
			const int i = bcells[idx];
			const int p = idx - c0;
			const int4 lp4 = ploc[idx];
			const int lp[4] = { lp4.x, lp4.y, lp4.z, lp4.w };
			const VTF* blk = fac + (size_t)i*80 + 4*q;
			double t = r[4*i + q];
			for(int k = 0; k < 4; k++) {
				if(lp[k] < 0 || lp[k] >= p) continue;
				double a[4];
				load_row<VTF>(blk + 16*(1 + k), a);
				const double* y = zs + 4*lp[k];
				t -= a[0]*y[0] + a[1]*y[1] + a[2]*y[2] + a[3]*y[3];
			}
			zs[4*p + q] = t;
		}
		__syncthreads();
	}
	for(int l = nl - 1; l >= 0; l--) {
		const int e0 = blv[l0 + l], e1 = blv[l0 + l + 1];
		for(int idx = e0 + qd; idx < e1; idx += NQ) {
			const int i = bcells[idx];
			const int p = idx - c0;
			const int4 lp4 = ploc[idx];
			const int lp[4] = { lp4.x, lp4.y, lp4.z, lp4.w };
			const VTF* blk = fac + (size_t)i*80 + 4*q;
			double t = zs[4*p + q];
			for(int k = 0; k < 4; k++) {
				if(lp[k] <= p) continue;
				double a[4];
				load_row<VTF>(blk + 16*(1 + k), a);
				const double* y = zs + 4*lp[k];
				t -= a[0]*y[0] + a[1]*y[1] + a[2]*y[2] + a[3]*y[3];
			}
			const double tv[4] = { quad_bcast<0>(t), quad_bcast<1>(t), quad_bcast<2>(t),
			                       quad_bcast<3>(t) };
			double a[4];
			load_row<VTF>(blk, a);
			const double v = dot4(a, tv);
			zs[4*p + q] = v;
			z[4*i + q] = v;
		}
		__syncthreads();
	}
}

/* ------------------------------------------------------------- reductions --------------- */

constexpr int RED_BLOCKS = 512;
constexpr int MAX_DOTS = 32;
constexpr long PART = (long)RED_BLOCKS*MAX_DOTS;   // per-instance partial-sum stride

/// Sum over the 256 threads of a block; the result is valid in every thread.
__device__ __forceinline__ double block_sum(double v, double* sh)
{
	for(int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o);
	const int w = threadIdx.x >> 6;
	__syncthreads();
	if((threadIdx.x & 63) == 0) sh[w] = v;
	__syncthreads();
	return (sh[0] + sh[1]) + (sh[2] + sh[3]);
}

/// partial[b][i] = block b's part of dot(V_i, w), i < nv <= NVM (V_i = V + i*ldv). The NVM
/// accumulators are indexed by compile-time constants only, so they stay in registers; each
/// wave reduces all of them, then one barrier, then thread i adds the four waves' parts of dot
/// i in the fixed order (w0 + w1) + (w2 + w3).
template <int NVM>
__global__ __launch_bounds__(BLOCK)
void k_mdot(long N, const int* __restrict__ ilist, const double* __restrict__ V, long ldv,
            long sv, int nv, const double* __restrict__ w, long sw, double* __restrict__ partial)
{
	const long I = inst_of(ilist);
	if(I < 0) return;                     // a retired slot of a compacted list
	V += I*sv; w += I*sw; partial += I*PART;
	static_assert(BLOCK == 256, "the final sum adds four waves' parts");
	__shared__ double sh[BLOCK/64][NVM];
	double acc[NVM];
#pragma unroll
	for(int i = 0; i < NVM; i++) acc[i] = 0.0;
	const long stride = (long)gridDim.x*blockDim.x;
	for(long e = (long)blockIdx.x*blockDim.x + threadIdx.x; e < N; e += stride) {
		const double we = w[e];
#pragma unroll
		for(int i = 0; i < NVM; i++)
			if(i < nv) acc[i] += V[i*ldv + e]*we;
	}
	const int wv = threadIdx.x >> 6;
#pragma unroll
	for(int i = 0; i < NVM; i++) {
		if(i < nv) {
			double v = acc[i];
			for(int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o);
			if((threadIdx.x & 63) == 0) sh[wv][i] = v;
		}
	}
	__syncthreads();
	const int t = threadIdx.x;
	if(t < nv)
		partial[(size_t)blockIdx.x*MAX_DOTS + t] = (sh[0][t] + sh[1][t]) + (sh[2][t] + sh[3][t]);
}

/// out[i] = sum over blocks of partial[b][i], i < nv (one block per instance, fixed order).
__global__ __launch_bounds__(BLOCK)
void k_reduce(int nblk, const int* __restrict__ ilist, const double* __restrict__ partial,
              int nv, double* __restrict__ out, long so)
{
	const long I = inst_of(ilist);
	if(I < 0) return;                     // a retired slot of a compacted list
	partial += I*PART; out += I*so;
	__shared__ double sh[4];
	for(int i = 0; i < nv; i++) {
		double v = 0.0;
		for(int b = threadIdx.x; b < nblk; b += blockDim.x) v += partial[(size_t)b*MAX_DOTS + i];
		const double s = block_sum(v, sh);
		if(threadIdx.x == 0) out[i] = s;
		__syncthreads();
	}
}

/// w -= sum_i h_i V_i (classical Gram-Schmidt update) and partial[b][0] = block part of |w|^2.
__global__ __launch_bounds__(BLOCK)
void k_gs_update(long N, const int* __restrict__ ilist, const double* __restrict__ V, long ldv,
                 long sv, int nv, const double* __restrict__ h, long sh_, double* __restrict__ w,
                 long sw, double* __restrict__ partial)
{
	const long I = inst_of(ilist);
	if(I < 0) return;                     // a retired slot of a compacted list
	V += I*sv; h += I*sh_; w += I*sw; partial += I*PART;
	__shared__ double sh[4];
	__shared__ double hs[MAX_DOTS];
	if(threadIdx.x < MAX_DOTS) hs[threadIdx.x] = threadIdx.x < nv ? h[threadIdx.x] : 0.0;
	__syncthreads();
	double acc = 0.0;
	const long stride = (long)gridDim.x*blockDim.x;
	for(long e = (long)blockIdx.x*blockDim.x + threadIdx.x; e < N; e += stride) {
		double x = w[e];
		for(int i = 0; i < MAX_DOTS; i++)
			if(i < nv) x -= hs[i]*V[i*ldv + e];
		w[e] = x;
		acc += x*x;
	}
	const double s = block_sum(acc, sh);
	if(threadIdx.x == 0) partial[(size_t)blockIdx.x*MAX_DOTS] = s;
}

/// out = b / sqrt(*nrm2) (out may alias b)
__global__ __launch_bounds__(BLOCK)
void k_scale(long N, const int* __restrict__ ilist, const double* b, long sb,
             const double* __restrict__ nrm2, long sn, double* out, long so)
{
	const long I = inst_of(ilist);
	if(I < 0) return;                     // a retired slot of a compacted list
	b += I*sb; nrm2 += I*sn; out += I*so;
	const double h = sqrt(nrm2[0]);
	const long stride = (long)gridDim.x*blockDim.x;
	for(long e = (long)blockIdx.x*blockDim.x + threadIdx.x; e < N; e += stride) out[e] = b[e]/h;
}

/// x = sum_k y_k Z_k, k < nv (with nvi: nvi[I] terms for instance I; none when negative)
__global__ __launch_bounds__(BLOCK)
void k_lincomb(long N, const int* __restrict__ ilist, const double* __restrict__ Z, long ldz,
               long sz, int nv, const int* __restrict__ nvi, const double* __restrict__ y,
               long sy, double* __restrict__ x, long sx)
{
	const long I = inst_of(ilist);
	if(I < 0) return;                     // a retired slot of a compacted list
	if(nvi) nv = nvi[I];
	Z += I*sz; y += I*sy; x += I*sx;
	__shared__ double ys[MAX_DOTS];
	if(threadIdx.x < MAX_DOTS) ys[threadIdx.x] = threadIdx.x < nv ? y[threadIdx.x] : 0.0;
	__syncthreads();
	const long stride = (long)gridDim.x*blockDim.x;
	for(long e = (long)blockIdx.x*blockDim.x + threadIdx.x; e < N; e += stride) {
		double s = 0.0;
		for(int k = 0; k < MAX_DOTS; k++)
			if(k < nv) s += ys[k]*Z[k*ldz + e];
		x[e] = s;
	}
}

/// x = 0 over N values per instance (a kernel rather than hipMemsetAsync, so that callers may
/// capture it in a graph).
__global__ __launch_bounds__(BLOCK)
void k_zero(long N, const int* __restrict__ ilist, double* __restrict__ x, long sx)
{
	const long I = inst_of(ilist);
	if(I < 0) return;                     // a retired slot of a compacted list
	x += I*sx;
	const long stride = (long)gridDim.x*blockDim.x;
	for(long e = (long)blockIdx.x*blockDim.x + threadIdx.x; e < N; e += stride) x[e] = 0.0;
}

/// partial[b][0] = block part of sum_c r[c][3]^2 area[c] (the reference's energy-residual
/// norm, src/ode/aodesolver.cpp:516-530).
__global__ __launch_bounds__(BLOCK)
void k_resnorm(int n, const int* __restrict__ ilist, const double* __restrict__ r, long sr,
               const double* __restrict__ area, double* __restrict__ partial)
{
	const long I = inst_of(ilist);
	if(I < 0) return;                     // a retired slot of a compacted list
	r += I*sr; partial += I*PART;
	__shared__ double sh[4];
	double acc = 0.0;
	for(int c = blockIdx.x*blockDim.x + threadIdx.x; c < n; c += gridDim.x*blockDim.x) {
		const double e = r[4*c + 3];
		acc += e*e*area[c];
	}
	const double s = block_sum(acc, sh);
	if(threadIdx.x == 0) partial[(size_t)blockIdx.x*MAX_DOTS] = s;
}

/* ------------------------------------------------------------ matrix-free / updates ---- */

/// aux = u + (eps/|x|) x with |x|^2 in *xx
__global__ __launch_bounds__(BLOCK)
void k_mf_prep(long N, const int* __restrict__ ilist, const double* __restrict__ u, long su,
               const double* __restrict__ x, long sx, const double* __restrict__ xx, long sn,
               double eps, double* __restrict__ aux, long sa)
{
	const long I = inst_of(ilist);
	if(I < 0) return;                     // a retired slot of a compacted list
	u += I*su; x += I*sx; xx += I*sn; aux += I*sa;
	const double pert = eps/sqrt(xx[0]);
	const long stride = (long)gridDim.x*blockDim.x;
	for(long e = (long)blockIdx.x*blockDim.x + threadIdx.x; e < N; e += stride)
		aux[e] = u[e] + pert*x[e];
}

/// y = dtm x + (r_u - r_aux)/pert, where r_u and r_aux hold -r() at u and at u + pert x.
__global__ __launch_bounds__(BLOCK)
void k_mf_finish(int n, const int* __restrict__ ilist, const double* __restrict__ x, long sx,
                 const double* __restrict__ dtm, const double* __restrict__ ru,
                 const double* __restrict__ raux, const double* __restrict__ xx, long sn,
                 double eps, double* __restrict__ y, long sy)
{
	const long I = inst_of(ilist);
	if(I < 0) return;                     // a retired slot of a compacted list
	x += I*sx; dtm += I*(long)n; ru += I*4L*n; raux += I*4L*n; xx += I*sn; y += I*sy;
	const double pert = eps/sqrt(xx[0]);
	for(int e = blockIdx.x*blockDim.x + threadIdx.x; e < 4*n; e += gridDim.x*blockDim.x)
		y[e] = dtm[e >> 2]*x[e] + (-raux[e] + ru[e])/pert;
}

/// u += omega du with the local relaxation factor of FlowSimpleUpdate
/// (src/ode/nonlinearrelaxation.cpp:25-38), or omega = 1 for the FULL update. The pressure
/// change is IdealGasPhysics::getDeltaPressureFromConserved (aphysics_defs.hpp:65-81), whose
/// sum runs over conserved components 2 and 3 (y-momentum and energy), not over the two
/// momenta; that sum is kept as the reference has it.
__global__ __launch_bounds__(BLOCK)
void k_update_implicit(int n, const int* __restrict__ ilist, double* __restrict__ u,
                       const double* __restrict__ du, int full, double min_relax, double gamma)
{
	const long I = inst_of(ilist);
	if(I < 0) return;                     // a retired slot of a compacted list
	u += I*4L*n; du += I*4L*n;
	const int c = blockIdx.x*blockDim.x + threadIdx.x;
	if(c >= n) return;
	double uc[4], d[4];
	load4(u + 4*c, uc);
	load4(du + 4*c, d);
	double om = 1.0;
	if(!full) {
		double un[4];
		for(int i = 0; i < 4; i++) un[i] = uc[i] + d[i];
		double dp = 0.0;
		for(int i = 2; i < 4; i++)
			dp -= (uc[i] + un[i])*(uc[0] + un[0])/2.0*d[i] - (un[i]*un[i] + uc[i]*uc[i])/2.0*d[0];
		dp = (gamma - 1.0)*(d[3] - 1.0/(2*uc[0]*un[0])*dp);
		const double p = (gamma - 1.0)*(uc[3] - 0.5*(uc[1]*uc[1] + uc[2]*uc[2])/uc[0]);
		const double danger = fmax(fabs(dp)/p, fabs(d[0])/uc[0]);
		om = danger < 1.0 - min_relax ? 1.0 - danger : min_relax;
	}
	for(int i = 0; i < 4; i++) uc[i] += om*d[i];
	store4(u + 4*c, uc);
}

/// u += cfl dt/area r (SteadyForwardEulerSolver::solve, src/ode/aodesolver.cpp:200-212).
__global__ __launch_bounds__(BLOCK)
void k_update_explicit(int n, const int* __restrict__ ilist, double* __restrict__ u,
                       const double* __restrict__ r, const double* __restrict__ dt,
                       const double* __restrict__ area, double cfl)
{
	const long I = inst_of(ilist);
	if(I < 0) return;                     // a retired slot of a compacted list
	u += I*4L*n; r += I*4L*n; dt += I*(long)n;
	const int c = blockIdx.x*blockDim.x + threadIdx.x;
	if(c >= n) return;
	const double f = cfl*dt[c]*1.0/area[c];
	for(int i = 0; i < 4; i++) u[4*c + i] += f*r[4*c + i];
}

/* ------------------------------------------------------------------ launchers ---------- */

static inline int quad_blocks(int n) { return (n + CELLS_PER_BLOCK - 1)/CELLS_PER_BLOCK; }
static inline int thread_blocks(long n) { return (int)((n + BLOCK - 1)/BLOCK); }
static inline int stream_blocks(long N)
{
	const long b = (N + BLOCK - 1)/BLOCK;
	return (int)(b < 4096 ? (b > 0 ? b : 1) : 4096);
}
static inline long mstride(const fvx_mesh_s* m) { return 80L*m->nelem; }

void launch_spmv(const fvx_matrix_s& A, const double* x, long sx, double* y, long sy,
                 const Inst& L, hipStream_t s)
{
	const int n = A.m->nelem;
	const dim3 g(quad_blocks(n), L.nact);
	if(A.prec == FVX_PREC_F32)
		hipLaunchKernelGGL((k_spmv<float, false>), g, dim3(BLOCK), 0, s, n, L.list, A.f32(),
		                   mstride(A.m), A.m->d_nbr.p, x, sx, nullptr, 0L, y, sy);
	else
		hipLaunchKernelGGL((k_spmv<double, false>), g, dim3(BLOCK), 0, s, n, L.list, A.f64(),
		                   mstride(A.m), A.m->d_nbr.p, x, sx, nullptr, 0L, y, sy);
	FVX_HIP(hipGetLastError());
}

void launch_pseudotime(fvx_matrix_s& A, const double* dt, double cfl, double* dtm,
                       hipStream_t s)
{
	const int n = A.m->nelem;
	if(A.prec == FVX_PREC_F32)
		hipLaunchKernelGGL(k_pseudotime<float>, dim3(quad_blocks(n)), dim3(BLOCK), 0, s, n,
		                   A.f32(), dt, A.m->d_area.p, cfl, dtm);
	else
		hipLaunchKernelGGL(k_pseudotime<double>, dim3(quad_blocks(n)), dim3(BLOCK), 0, s, n,
		                   A.f64(), dt, A.m->d_area.p, cfl, dtm);
	FVX_HIP(hipGetLastError());
}

template <class VTA, class VTF>
static void pc_setup_t(fvx_pc_s& p, const VTA* vals, VTF* fac, const Inst& L, hipStream_t s)
{
	const int n = p.m->nelem;
	const long sa = mstride(p.m);
	if(p.type == FVX_PC_JACOBI || pc_sgs(p.type)) {
		hipLaunchKernelGGL((k_invert_diag<VTA, VTF>), dim3(thread_blocks(n), L.nact), dim3(BLOCK),
		                   0, s, n, L.list, vals, fac, sa, pc_sgs(p.type) ? 1 : 0);
		FVX_HIP(hipGetLastError());
	} else if(p.type == FVX_PC_LINE) {
		hipLaunchKernelGGL((k_line_factor<VTA, VTF>), dim3(thread_blocks(p.nvert), L.nact),
		                   dim3(BLOCK), 0, s, p.nvert, p.d_vptr.p, p.d_vcells.p, p.d_lineinfo.p,
		                   L.list, vals, fac, sa);
		FVX_HIP(hipGetLastError());
	} else if(pc_ilu(p.type) || p.type == FVX_PC_BJACOBI_ILU0) {
		// block Jacobi: the same factorisation on the pattern restricted to each subdomain
		const int* nbr = p.type == FVX_PC_BJACOBI_ILU0 ? p.d_bnbr.p : p.m->d_nbr.p;
		for(int c = 0; c < p.ncol; c++) {
			const int cnt = p.coff[c+1] - p.coff[c];
			if(cnt == 0) continue;
			hipLaunchKernelGGL((k_ilu_factor<VTA, VTF>), dim3(thread_blocks(cnt), L.nact),
			                   dim3(BLOCK), 0, s, cnt, p.d_order.p + p.coff[c], L.list, vals, fac,
			                   sa, nbr, p.d_lowmask.p, p.d_colour.p, n);
			FVX_HIP(hipGetLastError());
		}
	}
}

void launch_pc_setup(fvx_pc_s& p, const fvx_matrix_s& A, const Inst& L, hipStream_t s)
{
	p.A = &A;
	if(p.type == FVX_PC_NONE) return;
	if(A.ninst < p.ninst) throw Error(FVX_ERR_INVALID, "matrix holds fewer instances than the preconditioner");
	if(A.prec == FVX_PREC_F32) {
		if(p.prec == FVX_PREC_F32) pc_setup_t(p, A.f32(), p.f32(), L, s);
		else if(p.prec == FVX_PREC_BF16) pc_setup_t(p, A.f32(), p.fb16(), L, s);
		else pc_setup_t(p, A.f32(), p.f64(), L, s);
	} else {
		if(p.prec == FVX_PREC_F32) pc_setup_t(p, A.f64(), p.f32(), L, s);
		else if(p.prec == FVX_PREC_BF16) pc_setup_t(p, A.f64(), p.fb16(), L, s);
		else pc_setup_t(p, A.f64(), p.f64(), L, s);
	}
}

/// `sweeps` forward + backward passes of symmetric Gauss-Seidel over the colours from z = 0;
/// the first forward pass reads lower colours only. A colour pass reads only the other colours,
/// so a pass that directly follows a pass of the same colour (the highest colour where the
/// forward sweep turns back, the lowest where the next sweep starts) would recompute the same
/// values from the same data, bit for bit: it is skipped, which leaves 2 sweeps (colours - 1) + 1
/// launches instead of 2 sweeps colours.
template <class VTF>
static void sgs_passes(const fvx_pc_s& p, const VTF* fac, const double* r, long sr, double* z,
                       long sz, int sweeps, const Inst& L, hipStream_t s)
{
	const long sa = mstride(p.m);
	int last = -1;                          // colour of the previous launched pass
	auto pass = [&](int c, bool all) {
		const int cnt = p.coff[c+1] - p.coff[c];
		if(cnt == 0 || c == last) return;
		const int4* rec = p.d_crec.p + 2L*p.coff[c];
		const dim3 g(quad_blocks(cnt), L.nact);
		if(all)
			hipLaunchKernelGGL((k_sgs_rec<VTF, true>), g, dim3(BLOCK), 0, s, cnt, rec, L.list,
			                   fac, sa, r, sr, z, sz);
		else
			hipLaunchKernelGGL((k_sgs_rec<VTF, false>), g, dim3(BLOCK), 0, s, cnt, rec, L.list,
			                   fac, sa, r, sr, z, sz);
		FVX_HIP(hipGetLastError());
		last = c;
	};
	for(int k = 0; k < sweeps; k++) {
		for(int c = 0; c < p.ncol; c++) pass(c, k > 0);
		for(int c = p.ncol - 1; c >= 0; c--) pass(c, true);
	}
}

/// `sweeps` forward + backward passes of line Gauss-Seidel over the colours of the vertices
/// from z = 0, skipping a colour pass that directly follows a pass of the same colour (it would
/// recompute the same values, as for point Gauss-Seidel).
template <class VTF>
static void line_passes(const fvx_pc_s& p, const VTF* fac, const double* r, long sr, double* z,
                        long sz, int sweeps, const Inst& L, hipStream_t s)
{
	const int* nbr = p.m->d_nbr.p;
	const long sa = mstride(p.m);
	// per colour: the lines (one thread each), then the lone cells (one quad each; their factor
	// blocks are the point Gauss-Seidel ones: inverted diagonal, matrix off-diagonals)
	auto colour = [&](int c, bool all) {
		const int nl = p.lcnt[c], np = p.coff[c+1] - p.coff[c] - nl;
		if(nl > 0) {
			const int* list = p.d_order.p + p.coff[c];
			const dim3 g(thread_blocks(nl), L.nact);
			if(all)
				hipLaunchKernelGGL((k_line_colour<VTF, true>), g, dim3(BLOCK), 0, s, nl, list,
				                   p.d_vptr.p, p.d_vcells.p, p.d_lineinfo.p, L.list, fac, sa, nbr,
				                   p.d_lowmask.p, r, sr, z, sz);
			else
				hipLaunchKernelGGL((k_line_colour<VTF, false>), g, dim3(BLOCK), 0, s, nl, list,
				                   p.d_vptr.p, p.d_vcells.p, p.d_lineinfo.p, L.list, fac, sa, nbr,
				                   p.d_lowmask.p, r, sr, z, sz);
			FVX_HIP(hipGetLastError());
		}
		if(np > 0) {
			const int* cells = p.d_ocell.p + p.coff[c] + nl;
			const dim3 g(quad_blocks(np), L.nact);
			if(all)
				hipLaunchKernelGGL((k_sgs_colour<VTF, true>), g, dim3(BLOCK), 0, s, np, cells,
				                   L.list, fac, sa, nbr, p.d_lowmask.p, r, sr, z, sz);
			else
				hipLaunchKernelGGL((k_sgs_colour<VTF, false>), g, dim3(BLOCK), 0, s, np, cells,
				                   L.list, fac, sa, nbr, p.d_lowmask.p, r, sr, z, sz);
			FVX_HIP(hipGetLastError());
		}
	};
	int last = -1;                          // colour of the previous pass
	auto pass = [&](int c, bool all) {
		if(c == last || p.coff[c+1] == p.coff[c]) return;
		colour(c, all);
		last = c;
	};
	for(int k = 0; k < sweeps; k++) {
		for(int c = 0; c < p.ncol; c++) pass(c, k > 0);
		for(int c = p.ncol - 1; c >= 0; c--) pass(c, true);
	}
}

/// z = M^-1 r for Jacobi or ILU(0) (one application).
template <class VTF>
static void pc_once(const fvx_pc_s& p, const VTF* fac, const double* r, long sr, double* z,
                    long sz, const Inst& L, hipStream_t s)
{
	const int n = p.m->nelem;
	const int* nbr = p.m->d_nbr.p;
	const long sa = mstride(p.m);
	if(p.type == FVX_PC_JACOBI) {
		hipLaunchKernelGGL(k_jacobi_apply<VTF>, dim3(quad_blocks(n), L.nact), dim3(BLOCK), 0, s,
		                   n, L.list, fac, sa, r, sr, z, sz);
		FVX_HIP(hipGetLastError());
		return;
	}
	if(p.type == FVX_PC_BJACOBI_ILU0) {
		if(p.nsub > 0)
			hipLaunchKernelGGL(k_bj_ilu_apply<VTF>, dim3(p.nsub, L.nact), dim3(BLOCK), 0, s,
			                   p.d_bptr.p, p.d_blvptr.p, p.d_blv.p, p.d_bcells.p, p.d_ploc.p,
			                   L.list, fac, sa, r, sr, z, sz);
		FVX_HIP(hipGetLastError());
		return;
	}
	for(int c = 0; c < p.ncol; c++) {
		const int cnt = p.coff[c+1] - p.coff[c];
		if(cnt == 0) continue;
		hipLaunchKernelGGL(k_ilu_fwd<VTF>, dim3(quad_blocks(cnt), L.nact), dim3(BLOCK), 0, s, cnt,
		                   p.d_order.p + p.coff[c], L.list, fac, sa, nbr, p.d_lowmask.p, r, sr,
		                   z, sz);
		FVX_HIP(hipGetLastError());
	}
	for(int c = p.ncol - 1; c >= 0; c--) {
		const int cnt = p.coff[c+1] - p.coff[c];
		if(cnt == 0) continue;
		hipLaunchKernelGGL(k_ilu_bwd<VTF>, dim3(quad_blocks(cnt), L.nact), dim3(BLOCK), 0, s, cnt,
		                   p.d_order.p + p.coff[c], L.list, fac, sa, nbr, p.d_lowmask.p, z, sz);
		FVX_HIP(hipGetLastError());
	}
}

/// `sweeps` steps of z <- z + M^-1 (r - A z) from z = 0 (PETSc -pc_sor_its). Symmetric
/// Gauss-Seidel does them as in-place passes; Jacobi and ILU(0) form the residual explicitly.
template <class VTF>
static void pc_apply_t(const fvx_pc_s& p, const VTF* fac, const double* r, long sr, double* z,
                       long sz, const Inst& L, hipStream_t s)
{
	if(pc_sgs(p.type)) {
		sgs_passes(p, fac, r, sr, z, sz, p.sweeps, L, s);
		return;
	}
	if(p.type == FVX_PC_LINE) {
		line_passes(p, fac, r, sr, z, sz, p.sweeps, L, s);
		return;
	}
	pc_once(p, fac, r, sr, z, sz, L, s);
	if(p.sweeps <= 1) return;
	const fvx_matrix_s& A = *p.A;
	const int n = p.m->nelem;
	const long N = 4L*n;
	const dim3 g(quad_blocks(n), L.nact);
	for(int k = 1; k < p.sweeps; k++) {
		if(A.prec == FVX_PREC_F32)
			hipLaunchKernelGGL((k_spmv<float, true>), g, dim3(BLOCK), 0, s, n, L.list, A.f32(),
			                   mstride(A.m), p.m->d_nbr.p, z, sz, r, sr, p.t1.p, N);
		else
			hipLaunchKernelGGL((k_spmv<double, true>), g, dim3(BLOCK), 0, s, n, L.list, A.f64(),
			                   mstride(A.m), p.m->d_nbr.p, z, sz, r, sr, p.t1.p, N);
		FVX_HIP(hipGetLastError());
		pc_once(p, fac, p.t1.p, N, p.t2.p, N, L, s);
		hipLaunchKernelGGL(k_axpy1, dim3(stream_blocks(N), L.nact), dim3(BLOCK), 0, s, N, L.list,
		                   p.t2.p, N, z, sz);
		FVX_HIP(hipGetLastError());
	}
}

void launch_pc_apply(const fvx_pc_s& p, const double* r, long sr, double* z, long sz,
                     const Inst& L, hipStream_t s)
{
	if(p.type == FVX_PC_NONE) {
		const long N = 4L*p.m->nelem;
		hipLaunchKernelGGL(k_scale, dim3(stream_blocks(N), L.nact), dim3(BLOCK), 0, s, N, L.list,
		                   r, sr, p.one.p, 0L, z, sz);
		FVX_HIP(hipGetLastError());
		return;
	}
	if(!p.A) throw Error(FVX_ERR_INVALID, "preconditioner applied before its set-up");
	if(p.prec == FVX_PREC_F32) pc_apply_t(p, p.f32(), r, sr, z, sz, L, s);
	else if(p.prec == FVX_PREC_BF16) pc_apply_t(p, p.fb16(), r, sr, z, sz, L, s);
	else pc_apply_t(p, p.f64(), r, sr, z, sz, L, s);
}

int reduction_blocks() { return RED_BLOCKS; }
int max_dots() { return MAX_DOTS; }
long partial_stride() { return PART; }

void launch_mdot(long N, const double* V, long ldv, long sv, int nv, const double* w, long sw,
                 double* partial, double* out, long so, const Inst& L, hipStream_t s)
{
	if(nv < 1 || nv > MAX_DOTS) throw Error(FVX_ERR_INVALID, "multi-dot of 1 to 32 vectors");
#define FVX_MD(M) hipLaunchKernelGGL(k_mdot<M>, dim3(RED_BLOCKS, L.nact), dim3(BLOCK), 0, s, N, \
	L.list, V, ldv, sv, nv, w, sw, partial)
	if(nv <= 8) FVX_MD(8);
	else if(nv <= 16) FVX_MD(16);
	else FVX_MD(MAX_DOTS);
#undef FVX_MD
	hipLaunchKernelGGL(k_reduce, dim3(1, L.nact), dim3(BLOCK), 0, s, RED_BLOCKS, L.list, partial,
	                   nv, out, so);
	FVX_HIP(hipGetLastError());
}

void launch_gs_update(long N, const double* V, long ldv, long sv, int nv, const double* h,
                      long sh, double* w, long sw, double* partial, double* nrm2, long sn,
                      const Inst& L, hipStream_t s)
{
	hipLaunchKernelGGL(k_gs_update, dim3(RED_BLOCKS, L.nact), dim3(BLOCK), 0, s, N, L.list, V,
	                   ldv, sv, nv, h, sh, w, sw, partial);
	hipLaunchKernelGGL(k_reduce, dim3(1, L.nact), dim3(BLOCK), 0, s, RED_BLOCKS, L.list, partial,
	                   1, nrm2, sn);
	FVX_HIP(hipGetLastError());
}

void launch_scale(long N, const double* b, long sb, const double* nrm2, long sn, double* out,
                  long so, const Inst& L, hipStream_t s)
{
	hipLaunchKernelGGL(k_scale, dim3(stream_blocks(N), L.nact), dim3(BLOCK), 0, s, N, L.list, b,
	                   sb, nrm2, sn, out, so);
	FVX_HIP(hipGetLastError());
}

void launch_lincomb(long N, const double* Z, long ldz, long sz, int nv, const double* y,
                    long sy, double* x, long sx, const Inst& L, hipStream_t s, const int* nvi)
{
	hipLaunchKernelGGL(k_lincomb, dim3(stream_blocks(N), L.nact), dim3(BLOCK), 0, s, N, L.list, Z,
	                   ldz, sz, nv, nvi, y, sy, x, sx);
	FVX_HIP(hipGetLastError());
}

/* ------------------------------------------------- FGMRES least-squares on the device ----- */

/* State of one FGMRES cycle per instance (KST doubles, instance-major): the Givens-rotated
 * Hessenberg columns R[r][j] (r <= j), the rotations cs, sn, the rotated right-hand side g,
 * beta = |b|, the iteration count and the status (1 running, 0 stopped, 2 zero right-hand
 * side, -1 non-finite). The kernels below run in one workgroup: thread t serves instance
 * ids[t], t < nids; thread 0 then writes the running instances, in the order of ids, to the
 * front of `list` and -1 behind them (the compacted list the next iteration's launches
 * read), and their number to *hcount (host memory, read by the host after the kernel). */
constexpr int KR = 0, KCS = MAX_DOTS*MAX_DOTS, KSN = KCS + MAX_DOTS, KG = KSN + MAX_DOTS;
constexpr int KBETA = KG + MAX_DOTS + 1, KIT = KBETA + 1, KSTAT = KIT + 1;
constexpr long KST = KSTAT + 1;

__device__ __forceinline__ void fgmres_compact(int nids, const int* __restrict__ ids,
                                               const double* __restrict__ kst,
                                               int* __restrict__ list, int* hcount)
{
	__syncthreads();
	if(threadIdx.x != 0) return;
	int k = 0;
	for(int t = 0; t < nids; t++)
		if(kst[ids[t]*KST + KSTAT] == 1.0) list[k++] = ids[t];
	for(int t = k; t < nids; t++) list[t] = -1;
	if(hcount) __hip_atomic_store(hcount, k, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_SYSTEM);
}

/// Cycle start: beta = sqrt(scal[I][beta_slot]), g = (beta, 0, ...), status.
__global__ __launch_bounds__(BLOCK)
void k_fgmres_begin(int nids, const int* __restrict__ ids, const double* __restrict__ scal,
                    long ns, int beta_slot, double* __restrict__ kst, int* __restrict__ list,
                    int* hcount)
{
	for(int t = threadIdx.x; t < nids; t += blockDim.x) {
		const long I = ids[t];
		double* st = kst + I*KST;
		const double beta = sqrt(scal[I*ns + beta_slot]);
		st[KBETA] = beta;
		st[KG] = beta;
		st[KIT] = 0.0;
		st[KSTAT] = !isfinite(beta) ? -1.0 : beta == 0.0 ? 2.0 : 1.0;
	}
	fgmres_compact(nids, ids, kst, list, hcount);
}

/// Iteration j: the new Hessenberg column h_0..h_j (scal[I][h_slot + r]) and |w|^2
/// (scal[I][nrm_slot]) go through the previous rotations and a new one; an instance stops when
/// |g_{j+1}| <= rtol beta or at an exact breakdown, and is marked non-finite on a non-finite
/// norm.
__global__ __launch_bounds__(BLOCK)
void k_fgmres_rotate(int j, double rtol, int nids, const int* __restrict__ ids,
                     const double* __restrict__ scal, long ns, int h_slot, int nrm_slot,
                     double* __restrict__ kst, int* __restrict__ list, int* hcount)
{
	for(int t = threadIdx.x; t < nids; t += blockDim.x) {
		const long I = ids[t];
		double* st = kst + I*KST;
		if(st[KSTAT] != 1.0) continue;
		const double* h = scal + I*ns + h_slot;
		const double hn = sqrt(scal[I*ns + nrm_slot]);
		if(!isfinite(hn)) { st[KSTAT] = -1.0; continue; }
		double* cs = st + KCS;
		double* sn = st + KSN;
		double* g = st + KG;
		double hr = h[0];
		for(int r = 0; r < j; r++) {
			const double hr1 = h[r + 1];
			st[KR + r*MAX_DOTS + j] = cs[r]*hr + sn[r]*hr1;
			hr = -sn[r]*hr + cs[r]*hr1;
		}
		const double den = sqrt(hr*hr + hn*hn);
		cs[j] = hr/den;
		sn[j] = hn/den;
		st[KR + j*MAX_DOTS + j] = den;
		g[j + 1] = -sn[j]*g[j];
		g[j] = cs[j]*g[j];
		st[KIT] = j + 1;
		if(fabs(g[j + 1]) <= rtol*st[KBETA] || hn == 0.0) st[KSTAT] = 0.0;
	}
	fgmres_compact(nids, ids, kst, list, hcount);
}

/// Cycle end: the weights y of the instances' bases (back substitution of R y = g) into
/// scal[I][y_slot ..], and the iteration count into its[I] (-1: non-finite).
__global__ __launch_bounds__(BLOCK)
void k_fgmres_end(int nids, const int* __restrict__ ids, const double* __restrict__ kst,
                  double* __restrict__ scal, long ns, int y_slot, int* __restrict__ its)
{
	for(int t = threadIdx.x; t < nids; t += blockDim.x) {
		const long I = ids[t];
		const double* st = kst + I*KST;
		if(st[KSTAT] == -1.0) { its[I] = -1; continue; }
		const int it = (int)st[KIT];
		double* y = scal + I*ns + y_slot;
		for(int r = it - 1; r >= 0; r--) {
			double a = st[KG + r];
			for(int k = r + 1; k < it; k++) a -= st[KR + r*MAX_DOTS + k]*y[k];
			y[r] = a/st[KR + r*MAX_DOTS + r];
		}
		its[I] = it;
	}
}

long fgmres_state_size() { return KST; }

void launch_fgmres_begin(int nids, const int* ids, const double* scal, long ns, int beta_slot,
                         double* kst, int* list, int* hcount, hipStream_t s)
{
	hipLaunchKernelGGL(k_fgmres_begin, dim3(1), dim3(BLOCK), 0, s, nids, ids, scal, ns,
	                   beta_slot, kst, list, hcount);
	FVX_HIP(hipGetLastError());
}

void launch_fgmres_rotate(int j, double rtol, int nids, const int* ids, const double* scal,
                          long ns, int h_slot, int nrm_slot, double* kst, int* list,
                          int* hcount, hipStream_t s)
{
	hipLaunchKernelGGL(k_fgmres_rotate, dim3(1), dim3(BLOCK), 0, s, j, rtol, nids, ids, scal, ns,
	                   h_slot, nrm_slot, kst, list, hcount);
	FVX_HIP(hipGetLastError());
}

void launch_fgmres_end(int nids, const int* ids, const double* kst, double* scal, long ns,
                       int y_slot, int* its, hipStream_t s)
{
	hipLaunchKernelGGL(k_fgmres_end, dim3(1), dim3(BLOCK), 0, s, nids, ids, kst, scal, ns,
	                   y_slot, its);
	FVX_HIP(hipGetLastError());
}

void launch_zero(long N, double* x, long sx, const Inst& L, hipStream_t s)
{
	hipLaunchKernelGGL(k_zero, dim3(stream_blocks(N), L.nact), dim3(BLOCK), 0, s, N, L.list, x,
	                   sx);
	FVX_HIP(hipGetLastError());
}

void launch_resnorm(int n, const double* r, const double* area, double* partial, double* out,
                    long so, const Inst& L, hipStream_t s)
{
	hipLaunchKernelGGL(k_resnorm, dim3(RED_BLOCKS, L.nact), dim3(BLOCK), 0, s, n, L.list, r,
	                   4L*n, area, partial);
	hipLaunchKernelGGL(k_reduce, dim3(1, L.nact), dim3(BLOCK), 0, s, RED_BLOCKS, L.list, partial,
	                   1, out, so);
	FVX_HIP(hipGetLastError());
}

void launch_mf_prep(long N, const double* u, const double* x, long sx, const double* xx,
                    long sn, double eps, double* aux, const Inst& L, hipStream_t s)
{
	hipLaunchKernelGGL(k_mf_prep, dim3(stream_blocks(N), L.nact), dim3(BLOCK), 0, s, N, L.list, u,
	                   N, x, sx, xx, sn, eps, aux, N);
	FVX_HIP(hipGetLastError());
}

void launch_mf_finish(int n, const double* x, long sx, const double* dtm, const double* ru,
                      const double* raux, const double* xx, long sn, double eps, double* y,
                      long sy, const Inst& L, hipStream_t s)
{
	hipLaunchKernelGGL(k_mf_finish, dim3(stream_blocks(4L*n), L.nact), dim3(BLOCK), 0, s, n,
	                   L.list, x, sx, dtm, ru, raux, xx, sn, eps, y, sy);
	FVX_HIP(hipGetLastError());
}

void launch_update_implicit(int n, double* u, const double* du, int full, double min_relax,
                            double gamma, const Inst& L, hipStream_t s)
{
	hipLaunchKernelGGL(k_update_implicit, dim3(thread_blocks(n), L.nact), dim3(BLOCK), 0, s, n,
	                   L.list, u, du, full, min_relax, gamma);
	FVX_HIP(hipGetLastError());
}

void launch_update_explicit(int n, double* u, const double* r, const double* dt,
                            const double* area, double cfl, const Inst& L, hipStream_t s)
{
	hipLaunchKernelGGL(k_update_explicit, dim3(thread_blocks(n), L.nact), dim3(BLOCK), 0, s, n,
	                   L.list, u, r, dt, area, cfl);
	FVX_HIP(hipGetLastError());
}

}
