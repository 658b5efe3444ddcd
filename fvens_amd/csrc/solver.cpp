/* solver.cpp -- host side of the implicit solver: matrix and preconditioner set-up, the
 * flexible GMRES driver, the matrix-free operator and the steady pseudo-time loop.
 *
 * Everything is written for a set of instances (operating points on one mesh) solved
 * together: each launch covers all active instances, while every instance keeps its own
 * Krylov basis, Hessenberg system, CFL ramp and stopping decisions, so an instance evolves
 * exactly as it would alone. A single solve is the one-instance case (no instance list, the
 * discretisation's own parameters). All vector work runs on the device; the host keeps the
 * small Hessenberg systems and the CFL schedules and reads a few scalars per instance per
 * Krylov iteration and per pseudo-time step.
 */
#include "fvx_internal.hpp"
#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstring>
#include <functional>
#include <numeric>

namespace fvx {

/* ---------------------------------------------------------------- matrix --------------- */

void matrix_init(fvx_matrix_s& A, const fvx_mesh_s* m, int prec, int ninst)
{
	if(prec != FVX_PREC_F64 && prec != FVX_PREC_F32)
		throw Error(FVX_ERR_INVALID, "unknown matrix precision");
	A.m = m;
	A.prec = prec;
	A.ninst = ninst;
	const size_t cnt = (size_t)m->nelem*80*ninst;
	if(prec == FVX_PREC_F32) {
		A.v32.alloc(cnt);
		FVX_HIP(hipMemset(A.v32.p, 0, cnt*sizeof(float)));
	} else {
		A.v64.alloc(cnt);
		FVX_HIP(hipMemset(A.v64.p, 0, cnt*sizeof(double)));
	}
}

static inline bool interior_slot(const fvx_mesh_s& m, int c, int k)
{
	if(k >= m.nnode[c]) return false;
	const int o = m.esuel[(size_t)c*4 + k];
	return o >= 0 && o < m.nelem;
}

int matrix_bsr_nnzb(const fvx_mesh_s& m) { return m.nelem + 2*m.ninface; }

void matrix_export_bsr(const fvx_matrix_s& A, int* rowptr, int* colind, double* vals)
{
	const fvx_mesh_s& m = *A.m;
	const int n = m.nelem;
	std::vector<double> h((size_t)n*80);
	FVX_HIP(hipDeviceSynchronize());
	if(A.prec == FVX_PREC_F32) {
		std::vector<float> f((size_t)n*80);
		FVX_HIP(hipMemcpy(f.data(), A.v32.p, f.size()*sizeof(float), hipMemcpyDeviceToHost));
		for(size_t i = 0; i < f.size(); i++) h[i] = f[i];
	} else {
		FVX_HIP(hipMemcpy(h.data(), A.v64.p, h.size()*sizeof(double), hipMemcpyDeviceToHost));
	}
	int pos = 0;
	for(int c = 0; c < n; c++) {
		if(rowptr) rowptr[c] = pos;
		std::pair<int,int> e[5];
		int ne = 0;
		e[ne++] = { c, 0 };
		for(int k = 0; k < 4; k++)
			if(interior_slot(m, c, k)) e[ne++] = { m.esuel[(size_t)c*4 + k], 1 + k };
		std::sort(e, e + ne);
		for(int t = 0; t < ne; t++, pos++) {
			if(colind) colind[pos] = e[t].first;
			if(vals) std::memcpy(vals + (size_t)pos*16, &h[(size_t)c*80 + 16*e[t].second],
			                     16*sizeof(double));
		}
	}
	if(rowptr) rowptr[n] = pos;
}

/* ------------------------------------------------------------ preconditioner ----------- */

/* Vertices of line Gauss-Seidel: the lines of findLines at p.line_threshold, then every cell
 * outside them on its own (the line-point graph of getHybridLineOrdering,
 * src/mesh/meshordering.cpp:264-385). Vertices are coloured greedily in vertex order so that
 * no two vertices of one colour hold adjacent cells; p.colour gets each cell's vertex colour,
 * p.order the vertices by colour. */
static void line_vertices(fvx_pc_s& p, const fvx_mesh_s& m)
{
	const int n = m.nelem;
	const fvx_mesh_desc d{ m.npoin, m.nelem, m.nbface, m.coords.data(), m.inpoel.data(),
	                       m.nnode.data(), m.bface.data(), m.btag.data() };
	std::vector<std::vector<int>> lines;
	std::vector<int> celline;
	find_lines(d, p.line_threshold, lines, celline);
	p.vptr.assign(1, 0);
	p.vcells.clear();
	std::vector<int> vert(n, -1);
	for(const auto& l : lines) {
		for(int c : l) { vert[c] = (int)p.vptr.size() - 1; p.vcells.push_back(c); }
		p.vptr.push_back((int)p.vcells.size());
	}
	for(int c = 0; c < n; c++)
		if(celline[c] < 0) { vert[c] = (int)p.vptr.size() - 1; p.vcells.push_back(c); p.vptr.push_back((int)p.vcells.size()); }
	p.nvert = (int)p.vptr.size() - 1;
	// slots of each cell's predecessor and successor in its line (7: none)
	std::vector<unsigned char> info(n, 7 | (7 << 3));
	for(int v = 0; v < p.nvert; v++)
		for(int k = p.vptr[v]; k < p.vptr[v+1]; k++) {
			const int c = p.vcells[k];
			int ps = 7, ns = 7;
			for(int s = 0; s < 4; s++) {
				if(!interior_slot(m, c, s)) continue;
				const int o = m.esuel[(size_t)c*4 + s];
				if(k > p.vptr[v] && o == p.vcells[k-1]) ps = s;
				if(k + 1 < p.vptr[v+1] && o == p.vcells[k+1]) ns = s;
			}
			if((k > p.vptr[v] && ps == 7) || (k + 1 < p.vptr[v+1] && ns == 7))
				throw Error(FVX_ERR_INVALID, "consecutive cells of a line are not neighbours");
			info[c] = (unsigned char)(ps | (ns << 3));
		}
	// greedy colouring of the vertices in vertex order
	std::vector<int> vcol(p.nvert, 0);
	p.ncol = 1;
	for(int v = 0; v < p.nvert; v++) {
		unsigned long long used = 0;
		for(int k = p.vptr[v]; k < p.vptr[v+1]; k++) {
			const int c = p.vcells[k];
			for(int s = 0; s < 4; s++) {
				if(!interior_slot(m, c, s)) continue;
				const int w = vert[m.esuel[(size_t)c*4 + s]];
				if(w < v) used |= 1ull << vcol[w];
			}
		}
		int col = 0;
		while(col < 63 && (used & (1ull << col))) col++;
		if(col == 63) throw Error(FVX_ERR_INVALID, "line-point graph needs more than 63 colours");
		vcol[v] = col;
		p.ncol = std::max(p.ncol, col + 1);
	}
	for(int c = 0; c < n; c++) p.colour[c] = vcol[vert[c]];
	p.coff.assign(p.ncol + 1, 0);
	for(int v = 0; v < p.nvert; v++) p.coff[vcol[v] + 1]++;
	for(int k = 0; k < p.ncol; k++) p.coff[k+1] += p.coff[k];
	// within a colour the lines come first (they have the lower vertex numbers); the lone cells
	// after them are swept by the point Gauss-Seidel kernel through their cell numbers
	p.order.assign(p.nvert, 0);
	std::vector<int> fill(p.coff.begin(), p.coff.end() - 1), ocell(p.nvert);
	p.lcnt.assign(p.ncol, 0);
	for(int v = 0; v < p.nvert; v++) {
		p.order[fill[vcol[v]]++] = v;
		if(v < (int)lines.size()) p.lcnt[vcol[v]]++;
	}
	for(int k = 0; k < p.nvert; k++) ocell[k] = p.vcells[p.vptr[p.order[k]]];
	p.d_ocell.upload(ocell);
	p.d_vptr.upload(p.vptr);
	p.d_vcells.upload(p.vcells);
	p.d_lineinfo.upload(info);
}

/* Subdomains of block Jacobi + ILU(0): grown breadth-first over the cell graph from the
 * lowest-numbered unassigned cell until they hold p.subdomain_cells cells or their connected
 * region is used up; a piece smaller than half that size is then merged into the adjacent
 * subdomain found first, if the union holds at most twice that size and at most BJ_MAX_CELLS
 * cells. Inside a subdomain the natural
 * order applies: a cell's level (p.colour) is one more than the highest level of its
 * lower-numbered neighbours in the same subdomain. Fills the device lists of the fused
 * substitution kernel and, in p.order / p.coff, the cells by (level, index) for the
 * factorisation launches. */
static void bj_subdomains(fvx_pc_s& p, const fvx_mesh_s& m)
{
	const int n = m.nelem, P = p.subdomain_cells;
	std::vector<int> part(n, -1), size;
	std::vector<int> queue;
	queue.reserve(P);
	for(int seed = 0; seed < n; seed++) {
		if(part[seed] >= 0) continue;
		const int id = (int)size.size();
		queue.assign(1, seed);
		part[seed] = id;
		for(size_t h = 0; h < queue.size() && (int)queue.size() < P; h++) {
			const int c = queue[h];
			for(int k = 0; k < 4 && (int)queue.size() < P; k++) {
				if(!interior_slot(m, c, k)) continue;
				const int o = m.esuel[(size_t)c*4 + k];
				if(part[o] >= 0) continue;
				part[o] = id;
				queue.push_back(o);
			}
		}
		size.push_back((int)queue.size());
	}
	// merge small pieces into a neighbour (first adjacent subdomain in cell and slot order)
	const int ns0 = (int)size.size();
	std::vector<int> target(ns0);
	std::iota(target.begin(), target.end(), 0);
	std::vector<char> merged(ns0, 0);
	for(int c = 0; c < n; c++) {
		const int a = part[c];
		if(merged[a] || size[a] >= P/2) continue;
		for(int k = 0; k < 4; k++) {
			if(!interior_slot(m, c, k)) continue;
			int b = part[m.esuel[(size_t)c*4 + k]];
			while(target[b] != b) b = target[b];
			if(b == a || size[b] + size[a] > std::min(BJ_MAX_CELLS, 2*P)) continue;
			target[a] = b;
			size[b] += size[a];
			merged[a] = 1;
			break;
		}
	}
	std::vector<int> renum(ns0, -1);
	int ns = 0;
	for(int a = 0; a < ns0; a++) {
		int b = a;
		while(target[b] != b) b = target[b];
		if(renum[b] < 0) renum[b] = ns++;
		renum[a] = renum[b];
	}
	for(int c = 0; c < n; c++) part[c] = renum[part[c]];
	p.part = part;
	p.nsub = ns;
	auto same = [&](int c, int k) {
		return interior_slot(m, c, k) && part[m.esuel[(size_t)c*4 + k]] == part[c];
	};
	p.ncol = 1;
	std::vector<unsigned char> low(n, 0);
	std::vector<int> bnbr((size_t)n*4, -1);
	for(int c = 0; c < n; c++) {
		int lev = 0;
		for(int k = 0; k < 4; k++) {
			if(!same(c, k)) continue;
			const int o = m.esuel[(size_t)c*4 + k];
			bnbr[(size_t)c*4 + k] = o;
			if(o < c) { lev = std::max(lev, p.colour[o] + 1); low[c] |= (unsigned char)(1u << k); }
		}
		p.colour[c] = lev;
		p.ncol = std::max(p.ncol, lev + 1);
	}
	// factorisation order: by level, then index
	p.coff.assign(p.ncol + 1, 0);
	for(int c = 0; c < n; c++) p.coff[p.colour[c] + 1]++;
	for(int k = 0; k < p.ncol; k++) p.coff[k+1] += p.coff[k];
	p.order.assign(n, 0);
	{
		std::vector<int> fill(p.coff.begin(), p.coff.end() - 1);
		for(int c = 0; c < n; c++) p.order[fill[p.colour[c]]++] = c;
	}
	// substitution lists: by (subdomain, level, index)
	std::vector<int> cells(n);
	std::iota(cells.begin(), cells.end(), 0);
	std::stable_sort(cells.begin(), cells.end(), [&](int a, int b) {
		return part[a] != part[b] ? part[a] < part[b] : p.colour[a] < p.colour[b]; });
	std::vector<int> bptr(ns + 1, 0), blvptr(ns + 1, 0), blv, pos(n);
	for(int i = 0; i < n; i++) {
		const int c = cells[i], b = part[c];
		if(i == 0 || part[cells[i-1]] != b) {
			bptr[b] = i;
			blvptr[b] = (int)blv.size();
		}
		if(i == 0 || part[cells[i-1]] != b || p.colour[cells[i-1]] != p.colour[c])
			blv.push_back(i);
		pos[c] = i - bptr[b];
	}
	bptr[ns] = n;
	blv.push_back(n);
	blvptr[ns] = (int)blv.size() - 1;
	std::vector<int4> ploc(n);
	for(int i = 0; i < n; i++) {
		const int c = cells[i];
		int l[4];
		for(int k = 0; k < 4; k++) l[k] = same(c, k) ? pos[m.esuel[(size_t)c*4 + k]] : -1;
		ploc[i] = make_int4(l[0], l[1], l[2], l[3]);
	}
	for(int b = 0; b < ns; b++)
		if(bptr[b+1] - bptr[b] > BJ_MAX_CELLS)
			throw Error(FVX_ERR_INVALID, "subdomain larger than a workgroup can hold");
	if(n > 0) {
		p.d_bptr.upload(bptr);
		p.d_blvptr.upload(blvptr);
		p.d_blv.upload(blv);
		p.d_bcells.upload(cells);
		p.d_ploc.upload(ploc);
		p.d_bnbr.upload(bnbr);
		p.d_colour.upload(std::vector<int>(n, 0));
		p.d_order.upload(p.order);
		p.d_lowmask.upload(low);
	}
}

void pc_set_subdomain_cells(fvx_pc_s& p, int cells)
{
	if(cells < 16 || cells > BJ_MAX_CELLS)
		throw Error(FVX_ERR_INVALID, "subdomain size must be in [16, 1024] cells");
	p.subdomain_cells = cells;
	if(p.type == FVX_PC_BJACOBI_ILU0 && p.m) {
		p.colour.assign(p.m->nelem, 0);
		bj_subdomains(p, *p.m);
		p.A = nullptr;
	}
}

void pc_init(fvx_pc_s& p, const fvx_mesh_s* m, int type, int prec, int ninst)
{
	if(type < FVX_PC_NONE || type > FVX_PC_BJACOBI_ILU0)
		throw Error(FVX_ERR_INVALID, "unknown preconditioner type");
	if(prec != FVX_PREC_F64 && prec != FVX_PREC_F32 && prec != FVX_PREC_BF16)
		throw Error(FVX_ERR_INVALID, "unknown preconditioner precision");
	p.m = m;
	p.type = type;
	p.prec = prec;
	p.ninst = ninst;
	p.A = nullptr;
	const int n = m->nelem;
	p.colour.assign(n, 0);
	p.ncol = 1;
	if(type == FVX_PC_SGS || type == FVX_PC_ILU0) {
		// greedy colouring in cell order: the smallest colour unused by lower-numbered neighbours
		for(int c = 0; c < n; c++) {
			unsigned used = 0;
			for(int k = 0; k < 4; k++)
				if(interior_slot(*m, c, k)) {
					const int o = m->esuel[(size_t)c*4 + k];
					if(o < c) used |= 1u << p.colour[o];
				}
			int col = 0;
			while(used & (1u << col)) col++;
			p.colour[c] = col;
			p.ncol = std::max(p.ncol, col + 1);
		}
	}
	if(pc_level(type)) {
		// wavefronts of the natural order: a cell's level is one more than the highest level
		// of its lower-numbered neighbours, so the cells of one level are independent and every
		// lower-numbered neighbour of a cell lies in an earlier level; sweeping the levels in
		// turn gives the sequential natural-order result (the role of levelSchedule,
		// src/mesh/ameshutils.cpp:158-196, with dependency levels instead of contiguous ranges)
		for(int c = 0; c < n; c++) {
			int lev = 0;
			for(int k = 0; k < 4; k++)
				if(interior_slot(*m, c, k)) {
					const int o = m->esuel[(size_t)c*4 + k];
					if(o < c) lev = std::max(lev, p.colour[o] + 1);
				}
			p.colour[c] = lev;
			p.ncol = std::max(p.ncol, lev + 1);
		}
	}
	if(type == FVX_PC_LINE) {
		line_vertices(p, *m);
	} else if(type == FVX_PC_BJACOBI_ILU0) {
		bj_subdomains(p, *m);
	} else {
		p.coff.assign(p.ncol + 1, 0);
		for(int c = 0; c < n; c++) p.coff[p.colour[c] + 1]++;
		for(int k = 0; k < p.ncol; k++) p.coff[k+1] += p.coff[k];
		p.order.assign(n, 0);
		std::vector<int> fill(p.coff.begin(), p.coff.end() - 1);
		for(int c = 0; c < n; c++) p.order[fill[p.colour[c]]++] = c;
	}
	std::vector<unsigned char> low(n, 0);
	for(int c = 0; c < n; c++)
		for(int k = 0; k < 4; k++)
			if(interior_slot(*m, c, k) && p.colour[m->esuel[(size_t)c*4 + k]] < p.colour[c])
				low[c] |= (unsigned char)(1u << k);
	if(n > 0 && type != FVX_PC_BJACOBI_ILU0) {
		// ILU(0) takes the lower neighbours of a row in (colour, index) order: the multicolour
		// order's, or for the natural order the index alone
		p.d_colour.upload(pc_level(type) ? std::vector<int>(n, 0) : p.colour);
		p.d_order.upload(p.order);
		p.d_lowmask.upload(low);
		if(pc_sgs(type)) {
			// per position of the colour order: {cell, lower-colour mask, neighbour slots 0-3},
			// the slots as the mesh's (neighbour, -1-ib on a boundary, SLOT_UNUSED)
			std::vector<int4> rec((size_t)2*n);
			for(int e = 0; e < n; e++) {
				const int c = p.order[e];
				int nb[4];
				for(int k = 0; k < 4; k++) {
					const int o = m->esuel[(size_t)c*4 + k];
					nb[k] = k >= m->nnode[c] ? SLOT_UNUSED : o < n ? o : -1 - (o - n);
				}
				rec[2*e] = make_int4(c, low[c], nb[0], nb[1]);
				rec[2*e + 1] = make_int4(nb[2], nb[3], 0, 0);
			}
			p.d_crec.upload(rec);
		}
	}
	p.one.upload(std::vector<double>(1, 1.0));
	if(type != FVX_PC_NONE) {
		const size_t cnt = (size_t)n*80*ninst;
		if(prec == FVX_PREC_F32) {
			p.v32.alloc(cnt);
			FVX_HIP(hipMemset(p.v32.p, 0, cnt*sizeof(float)));
		} else if(prec == FVX_PREC_BF16) {
			p.v16.alloc(cnt);
			FVX_HIP(hipMemset(p.v16.p, 0, cnt*sizeof(bf16)));
		} else {
			p.v64.alloc(cnt);
			FVX_HIP(hipMemset(p.v64.p, 0, cnt*sizeof(double)));
		}
	}
	p.t1.release();
	p.t2.release();
	pc_set_sweeps(p, p.sweeps);
}

void pc_set_sweeps(fvx_pc_s& p, int sweeps)
{
	if(sweeps < 1 || sweeps > 64) throw Error(FVX_ERR_INVALID, "sweep count must be in [1, 64]");
	p.sweeps = sweeps;
	if(sweeps > 1 && !pc_sgs(p.type) && p.type != FVX_PC_LINE && p.type != FVX_PC_NONE) {
		const size_t N = 4*(size_t)p.m->nelem*p.ninst;
		if(p.t1.n != N) { p.t1.alloc(N); p.t2.alloc(N); }
	}
}

/* ---------------------------------------------------------------- Krylov --------------- */

using Op = std::function<void(const double* x, long sx, double* y, long sy, const Inst& L)>;

/// Device storage of one FGMRES cycle for up to `cap` instances: bases, per-instance scalar
/// slots (h column, norms, solution weights), the least-squares state of the cycle
/// (k_fgmres_* in linalg.hip), the cycle's instance list and its compacted running list, and
/// pinned host words through which the device reports how many instances still run.
struct Krylov {
	long N = 0;
	int m = 0, cap = 0;
	DBuf<double> V, Z, partial, scal, kst;
	DBuf<int> dlist, dlist_in, dids, dits;  // caller's lists / running list / cycle / counts
	std::vector<int> dlist_ids;             // the list dlist holds on the device
	double* hpin = nullptr;
	int* ipin = nullptr;
	int* hcnt = nullptr;                    // [m + 2] running instances after each step
	int* itspin = nullptr;                  // [cap] iteration counts of the last cycle
	std::vector<hipEvent_t> ev;
	static constexpr int H0 = 0, NRM = 34, BETA = 35, XX = 36, RES = 37, Y0 = 40, NSCAL = 80;

	void ensure(long n_, int m_, int cap_)
	{
		if(n_ == N && m_ <= m && cap_ <= cap) return;
		release_host();
		N = n_;
		m = m_;
		cap = cap_;
		V.alloc((size_t)cap*(m + 1)*N);
		Z.alloc((size_t)cap*m*N);
		partial.alloc((size_t)cap*partial_stride());
		scal.alloc((size_t)cap*NSCAL);
		kst.alloc((size_t)cap*fgmres_state_size());
		dlist.alloc(cap);
		dlist_ids.clear();
		dlist_in.alloc(cap);
		dids.alloc(cap);
		dits.alloc(cap);
		FVX_HIP(hipHostMalloc((void**)&hpin, (size_t)cap*NSCAL*sizeof(double)));
		FVX_HIP(hipHostMalloc((void**)&ipin, (size_t)cap*sizeof(int)));
		FVX_HIP(hipHostMalloc((void**)&itspin, (size_t)cap*sizeof(int)));
		FVX_HIP(hipHostMalloc((void**)&hcnt, (size_t)(m + 2)*sizeof(int), hipHostMallocCoherent));
		std::memset(hpin, 0, (size_t)cap*NSCAL*sizeof(double));
		ev.assign(m + 2, nullptr);
		for(auto& e : ev) FVX_HIP(hipEventCreateWithFlags(&e, hipEventDisableTiming));
	}
	void release_host()
	{
		if(hpin) (void)hipHostFree(hpin);
		if(ipin) (void)hipHostFree(ipin);
		if(itspin) (void)hipHostFree(itspin);
		if(hcnt) (void)hipHostFree(hcnt);
		for(auto& e : ev) if(e) (void)hipEventDestroy(e);
		ev.clear();
		hpin = nullptr;
		ipin = nullptr;
		itspin = nullptr;
		hcnt = nullptr;
	}
	~Krylov() { release_host(); }

	long sV() const { return (long)(m + 1)*N; }
	long sZ() const { return (long)m*N; }
	double& h(int i, int slot) { return hpin[(size_t)i*NSCAL + slot]; }

	/// Launch description for the instances `ids`: no list for a single-instance solve,
	/// otherwise the list on the device. A list that differs from the one the device holds is
	/// uploaded and the upload waited for, so the pinned staging buffer can be reused at once;
	/// an unchanged list (most pseudo-steps of a batch) costs neither a copy nor a wait.
	Inst inst(const std::vector<int>& ids, bool single, hipStream_t s)
	{
		if(single) return Inst{nullptr, 1};
		if(ids != dlist_ids) {
			std::memcpy(ipin, ids.data(), ids.size()*sizeof(int));
			FVX_HIP(hipMemcpyAsync(dlist.p, ipin, ids.size()*sizeof(int), hipMemcpyHostToDevice,
			                       s));
			FVX_HIP(hipStreamSynchronize(s));
			dlist_ids = ids;
		}
		return Inst{dlist.p, (int)ids.size()};
	}

	/// Device scalar slots of every instance to the host (one copy, then waits).
	void fetch(hipStream_t s)
	{
		FVX_HIP(hipMemcpyAsync(hpin, scal.p, (size_t)cap*NSCAL*sizeof(double),
		                       hipMemcpyDeviceToHost, s));
		FVX_HIP(hipStreamSynchronize(s));
	}

	/// One cycle of right-preconditioned FGMRES from x = 0 with classical Gram-Schmidt for the
	/// instances `ids`. b and x are instance-major with stride N. Returns the iteration count
	/// per instance in `its`, or -1 where a non-finite value appeared (that instance's x is then
	/// undefined). The least-squares update runs on the device after every Arnoldi step and
	/// compacts the list of running instances there, so the launches of the next step skip the
	/// stopped ones; the host only looks one step behind (has any instance still been running
	/// after the previous step?) to decide whether to enqueue another step, so it never waits
	/// for the step it has just enqueued. The launches take the device list also for a single
	/// instance (the list {0}).
	void solve(const Op& op, const Op& pc, const double* b, double* x, int mm, double rtol,
	           const std::vector<int>& ids, bool single, std::vector<int>& its, hipStream_t s)
	{
		(void)single;
		its.assign(cap, 0);
		const int nids = (int)ids.size();
		if(nids == 0) return;
		std::memcpy(ipin, ids.data(), nids*sizeof(int));
		FVX_HIP(hipMemcpyAsync(dids.p, ipin, nids*sizeof(int), hipMemcpyHostToDevice, s));
		const Inst A{dids.p, nids};            // every instance of the cycle
		const Inst L{dlist_in.p, nids};        // the running ones, compacted on the device
		launch_mdot(N, b, 0, N, 1, b, N, partial.p, scal.p + BETA, NSCAL, A, s);
		launch_fgmres_begin(nids, dids.p, scal.p, NSCAL, BETA, kst.p, dlist_in.p, hcnt, s);
		FVX_HIP(hipEventRecord(ev[0], s));
		launch_scale(N, b, N, scal.p + BETA, NSCAL, V.p, sV(), L, s);
		for(int j = 0; j < mm; j++) {
			double* vj = V.p + (size_t)j*N;
			double* zj = Z.p + (size_t)j*N;
			double* w = V.p + (size_t)(j + 1)*N;
			pc(vj, sV(), zj, sZ(), L);
			op(zj, sZ(), w, sV(), L);
			launch_mdot(N, V.p, N, sV(), j + 1, w, sV(), partial.p, scal.p + H0, NSCAL, L, s);
			launch_gs_update(N, V.p, N, sV(), j + 1, scal.p + H0, NSCAL, w, sV(), partial.p,
			                 scal.p + NRM, NSCAL, L, s);
			// normalise v_{j+1}; after an exact breakdown (norm 0) the vector is never used
			launch_scale(N, w, sV(), scal.p + NRM, NSCAL, w, sV(), L, s);
			launch_fgmres_rotate(j, rtol, nids, dids.p, scal.p, NSCAL, H0, NRM, kst.p,
			                     dlist_in.p, hcnt + j + 1, s);
			FVX_HIP(hipEventRecord(ev[j + 1], s));
			// step j is queued; stop if nothing ran after step j - 1 (step j then found an
			// empty list and did nothing)
			FVX_HIP(hipEventSynchronize(ev[j]));
			if(__atomic_load_n(hcnt + j, __ATOMIC_ACQUIRE) == 0) break;
		}
		int* dits_p = dits.p;
		launch_fgmres_end(nids, dids.p, kst.p, scal.p, NSCAL, Y0, dits_p, s);
		launch_lincomb(N, Z.p, N, sZ(), mm, scal.p + Y0, NSCAL, x, N, A, s, dits_p);
		FVX_HIP(hipMemcpyAsync(itspin, dits_p, cap*sizeof(int), hipMemcpyDeviceToHost, s));
		FVX_HIP(hipStreamSynchronize(s));
		for(int i : ids) its[i] = itspin[i];
	}
};

static void check_restart(int restart)
{
	if(restart < 1 || restart > max_dots() - 1)
		throw Error(FVX_ERR_INVALID, "Krylov restart length must be in [1, 31]");
}

void destroy_krylov(Krylov* k) { delete k; }

int linear_solve(const fvx_matrix_s& A, fvx_pc_s& p, const double* b, double* x,
                 int restart, double rtol, hipStream_t s)
{
	check_restart(restart);
	if(!p.kry) p.kry = new Krylov;
	Krylov& k = *p.kry;
	k.ensure(4L*A.m->nelem, restart, 1);
	std::vector<int> its;
	k.solve([&](const double* xi, long sx, double* yo, long sy, const Inst& L) {
	            launch_spmv(A, xi, sx, yo, sy, L, s); },
	        [&](const double* r, long sr, double* z, long sz, const Inst& L) {
	            launch_pc_apply(p, r, sr, z, sz, L, s); },
	        b, x, restart, rtol, {0}, true, its, s);
	if(its[0] < 0) throw Error(FVX_ERR_NONFINITE, "non-finite value in the Krylov solve");
	return its[0];
}

/// y = dtm x + (r(u + eps x/|x|) - r(u)) |x|/eps for the instances of L; raux (instance stride
/// 4 ncell) is cleared for those instances before the residual is accumulated.
static void matfree_batch(const fvx_disc_s& d, const double* u, const double* ru,
                          const double* dtm, const double* x, long sx, double* y, long sy,
                          double eps, double* aux, double* raux,
                          double* partial, double* xx, long sxx, const Inst& L,
                          const BatchArg& ba, double* grad, double* rgrad, double* ughost,
                          hipStream_t s)
{
	const int n = d.m->nelem;
	const long N = 4L*n;
	launch_mdot(N, x, 0, sx, 1, x, sx, partial, xx, sxx, L, s);
	launch_mf_prep(N, u, x, sx, xx, sxx, eps, aux, L, s);
	launch_zero(N, raux, N, L, s);
	launch_residual(d, aux, raux, 0, nullptr, s, ba, L.nact, grad, rgrad, ughost);
	launch_mf_finish(n, x, sx, dtm, ru, raux, xx, sxx, eps, y, sy, L, s);
}

void matfree_apply(const fvx_disc_s& d, const double* u, const double* ru, const double* dtm,
                   const double* x, double* y, double eps, double* aux, double* raux,
                   double* partial, double* xx, hipStream_t s)
{
	const long N = 4L*d.m->nelem;
	matfree_batch(d, u, ru, dtm, x, N, y, N, eps, aux, raux, partial, xx, 0, Inst(),
	              BatchArg(), nullptr, nullptr, nullptr, s);
}

/* ------------------------------------------------------------ steady solver ------------ */

struct SolverWork {
	int n = -1, ninst = 0, jprec = -1, pctype = -1, pcprec = -1;
	fvx_matrix_s A;
	fvx_pc_s P;
	DBuf<double> r, dt, dtm, du, aux, raux, grad, rgrad, ughost, cfl;
	double* cfl_pin = nullptr;
	Krylov kry;
	std::vector<fvx_step_record> hist;     // per-step record of the last one-instance solve
	~SolverWork() { if(cfl_pin) (void)hipHostFree(cfl_pin); }
};

void solve_history(const fvx_disc_s& d, fvx_step_record* out, int cap, int* n)
{
	const int k = d.work ? (int)d.work->hist.size() : 0;
	*n = k;
	for(int i = 0; i < k && i < cap; i++) out[i] = d.work->hist[i];
}

void destroy_solver_work(SolverWork* w) { delete w; }

/// SteadySolver::expResidualRamp (src/ode/aodesolver.cpp:110-120)
static double exp_ramp(double cmin, double cmax, double prev, double ratio, double up,
                       double down)
{
	const double nc = ratio > 1.0 ? prev*std::pow(ratio, up) : prev*std::pow(ratio, down);
	if(nc < cmin) return cmin;
	if(nc > cmax) return cmax;
	return nc;
}

static void validate(const fvx_disc_s& d, const fvx_solver_cfg& c)
{
	if(!(c.tol > 0) || c.maxiter < 0 || !(c.cflinit > 0) || !(c.cflfin >= c.cflinit))
		throw Error(FVX_ERR_INVALID, "invalid pseudo-time settings");
	if(c.implicit) {
		check_restart(c.krylov_restart);
		if(!jacobian_supported(d))
			throw Error(FVX_ERR_UNSUPPORTED, "the reference has no Jacobian for this flux or BC");
		if(c.matrix_free && !(c.mf_eps > 0))
			throw Error(FVX_ERR_INVALID, "matrix-free step must be positive");
	}
}

void steady_core(fvx_disc_s& d, const fvx_solver_cfg& c, double* u, int ninst,
                 const GasParams* gp, std::vector<fvx_solve_stats>& stats, hipStream_t s,
                 const double* norms)
{
	validate(d, c);
	const fvx_mesh_s& m = *d.m;
	const int n = m.nelem;
	const long N = 4L*n;
	const bool single = gp == nullptr;
	if(single && ninst != 1) throw Error(FVX_ERR_INVALID, "a single solve has one instance");
	if(!d.work) d.work = new SolverWork;
	SolverWork& w = *d.work;
	if(w.n != n || w.ninst != ninst) {
		w.r.alloc(N*ninst); w.dt.alloc((size_t)n*ninst); w.dtm.alloc((size_t)n*ninst);
		w.du.alloc(N*ninst); w.aux.alloc(N*ninst); w.raux.alloc(N*ninst);
		if(!single) {
			w.grad.alloc(8L*n*ninst); w.rgrad.alloc(8L*n*ninst);
			w.ughost.alloc(4L*std::max(m.nbface, 1)*ninst);
		}
		w.cfl.alloc(ninst);
		if(w.cfl_pin) (void)hipHostFree(w.cfl_pin);
		w.cfl_pin = nullptr;
		FVX_HIP(hipHostMalloc((void**)&w.cfl_pin, ninst*sizeof(double)));
		w.n = n;
		w.ninst = ninst;
		w.jprec = w.pctype = w.pcprec = -1;
	}
	if(c.implicit) {
		if(w.jprec != c.jac_precision) {
			matrix_init(w.A, &m, c.jac_precision, ninst);
			w.jprec = c.jac_precision;
		}
		const double thr = c.line_threshold > 0 ? c.line_threshold : 10.0;
		const int sub = c.pc_subdomain_cells > 0 ? c.pc_subdomain_cells : 256;
		if(w.pctype != c.pc || w.pcprec != c.pc_precision
		   || (c.pc == FVX_PC_LINE && w.P.line_threshold != thr)
		   || (c.pc == FVX_PC_BJACOBI_ILU0 && w.P.subdomain_cells != sub)) {
			w.P.line_threshold = thr;
			if(sub < 16 || sub > BJ_MAX_CELLS)
				throw Error(FVX_ERR_INVALID, "subdomain size must be in [16, 1024] cells");
			w.P.subdomain_cells = sub;
			pc_init(w.P, &m, c.pc, c.pc_precision, ninst);
			w.pctype = c.pc;
			w.pcprec = c.pc_precision;
		}
		pc_set_sweeps(w.P, c.pc_sweeps > 0 ? c.pc_sweeps : 1);
	}
	Krylov& K = w.kry;
	K.ensure(N, c.implicit ? c.krylov_restart : 1, ninst);
	double* grad = single ? nullptr : w.grad.p;
	double* rgrad = single ? nullptr : w.rgrad.p;
	double* ughost = single ? nullptr : w.ughost.p;

	struct Ist { double cfl = 0, resi = 1, resiold = 1, initres = 1; int step = 0; };
	std::vector<Ist> is(ninst);
	stats.assign(ninst, fvx_solve_stats{});
	for(int i = 0; i < ninst; i++)
		is[i].initres = norms ? norms[i] : c.fixed_norm > 0 ? c.fixed_norm : 1.0;
	auto active = [&](int i) {
		return stats[i].status == 0 && is[i].resi/is[i].initres > c.tol && is[i].step < c.maxiter;
	};
	FVX_HIP(hipStreamSynchronize(s));
	const auto t0 = std::chrono::steady_clock::now();
	auto now = [&] { return std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count(); };
	std::vector<double> t_end(ninst, 0.0);
	std::vector<int> its;
	// the reference's SteadyStepMonitor (src/ode/aodesolver.cpp:536-549), one-instance solves
	w.hist.clear();
	double linw = 0.0;
	for(;;) {
		std::vector<int> act;
		for(int i = 0; i < ninst; i++) if(active(i)) act.push_back(i);
		if(act.empty()) break;
		Inst L = K.inst(act, single, s);
		BatchArg ba;
		if(!single) { ba.list = L.list; ba.gp = gp; ba.cfl = w.cfl.p; }
		FVX_HIP(hipMemsetAsync(w.r.p, 0, (size_t)N*ninst*sizeof(double), s));
		launch_residual(d, u, w.r.p, 1, w.dt.p, s, ba, L.nact, grad, rgrad, ughost);
		launch_resnorm(n, w.r.p, m.d_area.p, K.partial.p, K.scal.p + Krylov::RES, Krylov::NSCAL,
		               L, s);
		if(c.implicit) {
			for(int i : act) {
				is[i].cfl = exp_ramp(c.cflinit, c.cflfin, is[i].cfl, is[i].resiold/is[i].resi,
				                     0.25, 0.3);
				w.cfl_pin[i] = is[i].cfl;
			}
			if(!single)
				FVX_HIP(hipMemcpyAsync(w.cfl.p, w.cfl_pin, ninst*sizeof(double),
				                       hipMemcpyHostToDevice, s));
			launch_jacobian(d, u, w.A.raw(), w.A.prec, w.dt.p, is[act[0]].cfl, w.dtm.p, s, ba,
			                L.nact);
			d.prof.start(PROF_PC, s);
			launch_pc_setup(w.P, w.A, L, s);
			d.prof.stop(s);
			K.fetch(s);
			// an instance whose residual norm is not finite stops here; the others go on
			std::vector<int> ok;
			for(int i : act) {
				const double rn2 = K.h(i, Krylov::RES);
				if(!std::isfinite(rn2)) {
					is[i].resiold = is[i].resi;
					is[i].resi = std::sqrt(rn2);
					stats[i].status = FVX_ERR_NONFINITE;
					t_end[i] = now();
				} else ok.push_back(i);
			}
			if(ok.empty()) continue;
			if(ok.size() != act.size()) {
				act = ok;
				L = K.inst(act, single, s);
				ba.list = L.list;
			}
			Op op;
			if(c.matrix_free)
				op = [&](const double* x, long sx, double* y, long sy, const Inst& LL) {
					BatchArg bb = ba;
					bb.list = LL.list;
					matfree_batch(d, u, w.r.p, w.dtm.p, x, sx, y, sy, c.mf_eps, w.aux.p, w.raux.p,
					              K.partial.p, K.scal.p + Krylov::XX,
					              Krylov::NSCAL, LL, bb, grad, rgrad, ughost, s);
				};
			else
				op = [&](const double* x, long sx, double* y, long sy, const Inst& LL) {
					launch_spmv(w.A, x, sx, y, sy, LL, s);
				};
			std::vector<double> rn2(ninst);
			for(int i : act) rn2[i] = K.h(i, Krylov::RES);
			const double tl0 = now();
			K.solve(op, [&](const double* r, long sr, double* z, long sz, const Inst& LL) {
			            launch_pc_apply(w.P, r, sr, z, sz, LL, s); },
			        w.r.p, w.du.p, c.krylov_restart, c.krylov_rtol, act, single, its, s);
			linw += now() - tl0;
			std::vector<int> upd;
			for(int i : act) {
				if(its[i] < 0) { stats[i].status = FVX_ERR_NONFINITE; t_end[i] = now(); continue; }
				stats[i].lin_iters += its[i];
				upd.push_back(i);
			}
			if(!upd.empty()) {
				const Inst U = upd.size() == act.size() ? L : K.inst(upd, single, s);
				launch_update_implicit(n, u, w.du.p, c.fullupdate, c.min_relax, d.gas.g, U, s);
			}
			for(int i : upd) {
				is[i].resiold = is[i].resi;
				is[i].resi = std::sqrt(rn2[i]);
			}
			act = upd;
		} else {
			for(int i : act)
				is[i].cfl = exp_ramp(c.cflinit, c.cflfin, is[i].cfl, is[i].resiold/is[i].resi,
				                     0.3, 0.25);
			launch_update_explicit(n, u, w.r.p, w.dt.p, m.d_area.p, c.cflinit, L, s);
			K.fetch(s);
			for(int i : act) {
				is[i].resiold = is[i].resi;
				is[i].resi = std::sqrt(K.h(i, Krylov::RES));
			}
		}
		for(int i : act) {
			if(!std::isfinite(is[i].resi)) stats[i].status = FVX_ERR_NONFINITE;
			else {
				if(is[i].step == 0 && !norms && c.fixed_norm <= 0) is[i].initres = is[i].resi;
				is[i].step++;
				if(ninst == 1)
					w.hist.push_back(fvx_step_record{is[i].step, is[i].resi/is[i].initres,
					                                 is[i].resi, now(), linw,
					                                 stats[i].lin_iters, is[i].cfl});
			}
			if(!active(i)) t_end[i] = now();
		}
	}
	FVX_HIP(hipStreamSynchronize(s));
	const double wall = now();
	for(int i = 0; i < ninst; i++) {
		fvx_solve_stats& st = stats[i];
		st.walltime = single ? wall : t_end[i];
		st.steps = is[i].step;
		st.final_abs = is[i].resi;
		st.first_norm = is[i].initres;
		st.final_rel = is[i].resi/is[i].initres;
		// converged only below the step limit: a solve that reaches the tolerance exactly on
		// step maxiter counts as not converged (src/ode/aodesolver.cpp:618-626)
		st.converged = (st.status == 0 && st.steps < c.maxiter && st.final_rel <= c.tol) ? 1 : 0;
		if(st.status == 0 && !st.converged) st.status = FVX_ERR_NOT_CONVERGED;
	}
}

void steady_solve(fvx_disc_s& d, const fvx_solver_cfg& c, double* u, fvx_solve_stats& st,
                  hipStream_t s)
{
	std::vector<fvx_solve_stats> v;
	steady_core(d, c, u, 1, nullptr, v, s);
	st = v[0];
	if(st.status == FVX_ERR_NONFINITE)
		throw Error(FVX_ERR_NONFINITE, "non-finite residual in the pseudo-time loop");
	if(st.status == FVX_ERR_NOT_CONVERGED)
		throw Error(FVX_ERR_NOT_CONVERGED, "pseudo-time loop reached its step limit");
}

}
