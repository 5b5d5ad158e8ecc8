// Root sums of an Aberth sweep by a fast multipole method.
//
// Every full Aberth update needs S_k = sum_{j != k} 1/(z_k - z_j) over the R
// roots of its row (ccf_functions.cuh:232-239), O(R^2) per row and sweep.
// For the rows of a sweep in which many roots are active this file forms
// the sums in O(R p^2 / leaf + R leaf) instead:
//
//  * The field is covered by a uniform quadtree: the top level is a grid of
//    nx x ny square cells of side s0 (one square for a disc, a strip of
//    squares for an elongated rectangle) and level l splits every top cell
//    into 2^l x 2^l. Its geometry is fixed per field; the roots of each row
//    are binned into the leaf cells at every sweep (counting sort, then a
//    sort by root index inside each cell so that sums are deterministic).
//  * Expansions are of the Cauchy kernel 1/(z - w), scaled by the half side
//    h of the cell's level (Greengard & Rokhlin 1987, written for the
//    complex kernel rather than for log|z - w|):
//      multipole   A_n = sum_j ((z_j - c)/h)^n,
//                  sum_j 1/(z - z_j) = (1/h) sum_n A_n / w^(n+1),  w = (z - c)/h
//      local       sum_far 1/(z - z_j) = sum_m L_m w^m
//      M2M         A_n(parent) = sum_k C(n,k) 2^-k A_k(child) e^(n-k),  e = (c_child - c_parent)/h_parent
//      M2L         L_m += (-1)^m / h sum_n C(n+m,m) A_n d^-(n+m+1),    d = (c_target - c_source)/h
//      L2L         L_k(child) = 2^-k sum_{m>=k} C(m,k) L_m(parent) e^(m-k)
//    each formed without a table of binomials (whose constants would
//    crowd the scalar registers): M2M as the in-place binomial transform
//    a_n += e a_(n-1), L2L as a Taylor shift by repeated synthetic
//    division, and M2L's sum_n C(n+m,m) b_n as element 0 after m + 1
//    suffix-sum passes over b_n = A_n d^-n.
//    The interaction list of a cell at level >= 1 is the set of children of
//    its parent's neighbours that are not its own neighbours; at level 0 it
//    is every top cell that is not a neighbour. A leaf's near field is its
//    3 x 3 block of leaf cells, summed directly in FP64 (the own root gives
//    a zero difference and, as in the direct kernels, a zero term).
//  * Roots outside the grid are summed directly, as sources and as targets.
//
// The Aberth map z - f0 / (f1 + f0 S) has the roots of f0 as fixed points
// whatever S is; the truncation error of the far field (at most about
// 0.71^(p+1) of the far part, for the closest well-separated cells, far less
// on average) steers the iteration, it does not move the roots it reaches.

constexpr int ROOT_FMM_MIN_ROOTS = 64;     // smaller rows are always summed directly
// sweeps of a phase that may use the FMM: the dense ones (the initial solve
// converges over a few dozen sweeps; a continuation step mostly in one or two)
constexpr int ROOT_FMM_INIT_SWEEPS = 24;
constexpr int ROOT_FMM_STEP_SWEEPS = 2;
constexpr int ROOT_FMM_PROBE = 8;
constexpr int OUT_BLOCKS = 8; // workgroups per row summing the roots outside the grid

// listed roots from which a row's sums are cheaper by the FMM: its pass
// costs about as much as R/32 direct updates on a row of 20 000 roots; on
// larger rows the pass's cost per root falls while every direct update
// sums R terms, and 512 updates outweigh it
inline int root_fmm_min(int R) { return std::max(32, std::min(R / 32, 512)); }

struct RootView
{
	RootGrid g;
	int* slot;
	int* rmode;
	double2* rsum;
	int* ccnt;
	int* cstart;
	int* ccur;
	int* cell;
	int* sidx;
	double2* spos;
	int* ocnt;
	int* oidx;
	double2* mp;
	double2* lc;
};

RootView root_view(const SweepBufs& B)
{
	return RootView{B.grid, B.slot, B.rmode, B.rsum, B.ccnt, B.cstart, B.ccur, B.cell, B.sidx, B.spos, B.ocnt, B.oidx, B.mp, B.lc};
}

__device__ inline cd cell_center(const RootGrid& g, int l, int ix, int iy)
{
	const double s = ldexp(g.s0, -l);
	return cd(g.x0 + (ix + 0.5) * s, g.y0 + (iy + 0.5) * s);
}

// counts and the outlier count of the rows of this sweep
__global__ void k_rzero(RootView V)
{
	const int r = blockIdx.y;
	if (!V.rmode[r]) return;
	const int c = blockIdx.x * blockDim.x + threadIdx.x;
	if (c < V.g.nleaf) V.ccur[(size_t)r * V.g.nleaf + c] = 0;
	if (c == 0) V.ocnt[r] = 0;
}

// leaf cell of every root of the row (-1 outside the grid or not finite;
// k_rscan lists those)
__global__ __launch_bounds__(BLOCK) void k_rbin(RootView V, const double2* __restrict__ roots, RowMap M)
{
	const int r = blockIdx.y;
	if (!V.rmode[r]) return;
	const int R = M.R;
	const int j = blockIdx.x * BLOCK + threadIdx.x;
	int c = -1;
	if (j < R)
	{
		int row;
		double phi;
		M.get(r, row, phi);
		const double2 z = roots[(size_t)row * R + j];
		const double sl = ldexp(V.g.s0, -V.g.leaf);
		const double fx = (z.x - V.g.x0) / sl, fy = (z.y - V.g.y0) / sl;
		const int nxl = V.g.nx << V.g.leaf, nyl = V.g.ny << V.g.leaf;
		if (fx >= 0 && fx < nxl && fy >= 0 && fy < nyl) c = (int)fy * nxl + (int)fx;
		V.cell[(size_t)r * R + j] = c;
	}
	// neighbouring roots mostly share a cell: one atomic per run of them
	(void)run_add(V.ccur + (size_t)r * V.g.nleaf, c);
}

// exclusive prefix of the leaf counts: first sorted position of each cell
__global__ __launch_bounds__(1024) void k_rscan(RootView V, int R, unsigned long long* __restrict__ stats)
{
	const int r = blockIdx.x;
	if (!V.rmode[r]) return;
	if (threadIdx.x == 0) atomicAdd(stat_slot(stats, C_RROWS), 1ull);
	__shared__ int wsum[16];
	__shared__ int carry;
	const int lane = threadIdx.x & 63, w = threadIdx.x >> 6, nw = blockDim.x >> 6;
	const int nl = V.g.nleaf;
	int* cnt = V.ccur + (size_t)r * nl;
	int* st = V.cstart + (size_t)r * (nl + 1);
	if (threadIdx.x == 0) carry = 0;
	__syncthreads();
	for (int base = 0; base < nl; base += blockDim.x)
	{
		const int c = base + threadIdx.x;
		const int v = c < nl ? cnt[c] : 0;
		int x = v;
		for (int d = 1; d < 64; d <<= 1)
		{
			const int y = __shfl_up(x, d, 64);
			if (lane >= d) x += y;
		}
		if (lane == 63) wsum[w] = x;
		__syncthreads();
		int woff = 0, tot = 0;
		for (int k = 0; k < nw; k++)
		{
			if (k < w) woff += wsum[k];
			tot += wsum[k];
		}
		if (c < nl)
		{
			const int e = carry + woff + x - v;
			st[c] = e;
			cnt[c] = e;
		}
		__syncthreads();
		if (threadIdx.x == 0) carry += tot;
		__syncthreads();
	}
	if (threadIdx.x == 0) st[nl] = carry;
	// the row's roots outside the grid (cell -1), listed in index order so
	// that every sum over them runs in one order
	const int* cell = V.cell + (size_t)r * R;
	int* oidx = V.oidx + (size_t)r * R;
	if (threadIdx.x == 0) carry = 0;
	__syncthreads();
	const unsigned long long below = lane == 0 ? 0ull : (~0ull >> (64 - lane));
	for (int base = 0; base < R; base += blockDim.x)
	{
		const int j = base + threadIdx.x;
		const bool o = j < R && cell[j] < 0;
		const unsigned long long m = __ballot(o);
		if (lane == 0) wsum[w] = (int)__popcll(m);
		__syncthreads();
		int woff = 0, tot = 0;
		for (int k = 0; k < nw; k++)
		{
			if (k < w) woff += wsum[k];
			tot += wsum[k];
		}
		if (o) oidx[carry + woff + (int)__popcll(m & below)] = j;
		__syncthreads();
		if (threadIdx.x == 0) carry += tot;
		__syncthreads();
	}
	if (threadIdx.x == 0) V.ocnt[r] = carry;
}

__global__ __launch_bounds__(BLOCK) void k_rscatter(RootView V, int R)
{
	const int r = blockIdx.y;
	if (!V.rmode[r]) return;
	const int j = blockIdx.x * BLOCK + threadIdx.x;
	const int c = j < R ? V.cell[(size_t)r * R + j] : -1;
	const int pos = run_add(V.ccur + (size_t)r * V.g.nleaf, c);
	if (c >= 0) V.sidx[(size_t)r * R + pos] = j;
}

// exclusive prefix of v over the workgroup (blockDim.x a multiple of 64, at
// most 64 waves); tot: the sum of all. wsum: one int per wave, reusable on
// return.
__device__ inline int block_scan_excl(int v, int* wsum, int& tot)
{
	const int lane = threadIdx.x & 63, w = threadIdx.x >> 6, nw = blockDim.x >> 6;
	int x = v;
	for (int d = 1; d < 64; d <<= 1)
	{
		const int y = __shfl_up(x, d, 64);
		if (lane >= d) x += y;
	}
	if (lane == 63) wsum[w] = x;
	__syncthreads();
	int woff = 0;
	tot = 0;
	for (int k = 0; k < nw; k++)
	{
		if (k < w) woff += wsum[k];
		tot += wsum[k];
	}
	__syncthreads();
	return woff + x - v;
}

// The binning of a row in one workgroup when its leaf counts fit in LDS
// (k_rzero, k_rbin, k_rscan and k_rscatter in one launch). The counts and
// then the cells' cursors are LDS atomics; the outliers (cell -1) of every
// 64 consecutive roots are a ballot mask in LDS, whose popcounts' prefix
// places them in index order. Within a cell the order is the atomics'
// (k_rleaf sorts it). LDS: nleaf ints, then per 64 roots a mask and an int.
constexpr int BIN_THREADS = 1024;
constexpr int BIN_BATCH = 4;        // roots per thread whose loads k_rbin_row issues at once
constexpr int BIN_STATIC_LDS = 256; // k_rbin_row's own LDS beside the dynamic part

inline size_t bin_row_lds(int nleaf, int R)
{
	const size_t ng = ((size_t)R + 63) / 64;
	return ((size_t)nleaf * sizeof(int) + 7) / 8 * 8 + ng * (sizeof(unsigned long long) + sizeof(int));
}

__global__ __launch_bounds__(BIN_THREADS) void k_rbin_row(RootView V, const double2* __restrict__ roots, RowMap M,
	unsigned long long* __restrict__ stats)
{
	extern __shared__ unsigned long long bin_lds[];
	__shared__ int wsum[BIN_THREADS / 64];
	const int r = blockIdx.x;
	if (!V.rmode[r]) return;
	if (threadIdx.x == 0) atomicAdd(stat_slot(stats, C_RROWS), 1ull);
	const int R = M.R, nl = V.g.nleaf, ng = (R + 63) / 64;
	int* cnt = (int*)bin_lds;
	unsigned long long* omask = bin_lds + (nl + 1) / 2;
	int* obase = (int*)(omask + ng);
	const int lane = threadIdx.x & 63;
	for (int c = threadIdx.x; c < nl; c += BIN_THREADS) cnt[c] = 0;
	__syncthreads();
	int row;
	double phi;
	M.get(r, row, phi);
	const double2* __restrict__ zr = roots + (size_t)row * R;
	int* __restrict__ cell = V.cell + (size_t)r * R;
	const double sl = ldexp(V.g.s0, -V.g.leaf);
	const int nxl = V.g.nx << V.g.leaf, nyl = V.g.ny << V.g.leaf;
	// every wave covers whole groups of 64: j - lane is a multiple of 64.
	// BIN_BATCH roots per thread and pass, their loads issued together.
	constexpr int step = BIN_THREADS * BIN_BATCH;
	const int nj = (R + step - 1) / step * step;
	for (int j0 = threadIdx.x; j0 < nj; j0 += step)
	{
		double2 z[BIN_BATCH];
#pragma unroll
		for (int k = 0; k < BIN_BATCH; k++)
		{
			const int j = j0 + k * BIN_THREADS;
			z[k] = j < R ? zr[j] : make_double2(0, 0);
		}
#pragma unroll
		for (int k = 0; k < BIN_BATCH; k++)
		{
			const int j = j0 + k * BIN_THREADS;
			int c = -1;
			if (j < R)
			{
				const double fx = (z[k].x - V.g.x0) / sl, fy = (z[k].y - V.g.y0) / sl;
				if (fx >= 0 && fx < nxl && fy >= 0 && fy < nyl) c = (int)fy * nxl + (int)fx;
				cell[j] = c;
				if (c >= 0) atomicAdd(&cnt[c], 1);
			}
			const unsigned long long m = __ballot(j < R && c < 0);
			if (lane == 0 && j < R) omask[j >> 6] = m;
		}
	}
	__syncthreads();
	// the outliers in index order: prefix of the groups' popcounts
	int run = 0;
	for (int base = 0; base < ng; base += BIN_THREADS)
	{
		const int g = base + threadIdx.x;
		const int v = g < ng ? (int)__popcll(omask[g]) : 0;
		int t;
		const int e = run + block_scan_excl(v, wsum, t);
		if (g < ng) obase[g] = e;
		run += t;
	}
	if (threadIdx.x == 0) V.ocnt[r] = run;
	__syncthreads();
	if (run > 0)
	{
		int* oidx = V.oidx + (size_t)r * R;
		const unsigned long long below = lane == 0 ? 0ull : (~0ull >> (64 - lane));
		for (int j = threadIdx.x; j < R; j += BIN_THREADS)
		{
			const unsigned long long m = omask[j >> 6];
			if ((m >> lane) & 1ull) oidx[obase[j >> 6] + (int)__popcll(m & below)] = j;
		}
	}
	// exclusive prefix of the leaf counts: cstart, and the cells' cursors
	int* st = V.cstart + (size_t)r * (nl + 1);
	run = 0;
	for (int base = 0; base < nl; base += BIN_THREADS)
	{
		const int c = base + threadIdx.x;
		const int v = c < nl ? cnt[c] : 0;
		int t;
		const int e = run + block_scan_excl(v, wsum, t);
		if (c < nl)
		{
			st[c] = e;
			cnt[c] = e;
		}
		run += t;
	}
	if (threadIdx.x == 0) st[nl] = run;
	__syncthreads();
	int* __restrict__ sidx = V.sidx + (size_t)r * R;
	for (int j0 = threadIdx.x; j0 < R; j0 += step)
	{
		int c[BIN_BATCH];
#pragma unroll
		for (int k = 0; k < BIN_BATCH; k++)
		{
			const int j = j0 + k * BIN_THREADS;
			c[k] = j < R ? cell[j] : -1;
		}
#pragma unroll
		for (int k = 0; k < BIN_BATCH; k++)
			if (c[k] >= 0) sidx[atomicAdd(&cnt[c[k]], 1)] = j0 + k * BIN_THREADS;
	}
}

constexpr int LEAF_REG = 32; // cells up to this many roots sort in registers

// per leaf cell: the cell's roots in index order, their positions copied
// contiguously, and the multipole expansion (P2M)
template <int P>
__global__ __launch_bounds__(BLOCK) void k_rleaf(RootView V, const double2* __restrict__ roots, RowMap M)
{
	const int r = blockIdx.y;
	if (!V.rmode[r]) return;
	const int c = blockIdx.x * BLOCK + threadIdx.x;
	if (c >= V.g.nleaf) return;
	const int R = M.R;
	int row;
	double phi;
	M.get(r, row, phi);
	const int* st = V.cstart + (size_t)r * (V.g.nleaf + 1);
	const int b = st[c], e = st[c + 1], n = e - b;
	const int L = V.g.leaf;
	const size_t cid = (size_t)r * V.g.ncells + V.g.off[L] + c;
	V.ccnt[cid] = n;
	if (n == 0) return;
	int* ix = V.sidx + (size_t)r * R;
	const int nxl = V.g.nx << L;
	const cd cc = cell_center(V.g, L, c % nxl, c / nxl);
	const double ih = 2.0 / ldexp(V.g.s0, -L);
	cd A[P + 1];
#pragma unroll
	for (int k = 0; k <= P; k++) A[k] = cd(0, 0);
	if (n <= LEAF_REG)
	{
		// a small cell sorts in registers: each index's rank among the
		// cell's (the indices are distinct), then every load of the cell's
		// roots at once instead of one dependent chain per root
		int v[LEAF_REG], srt[LEAF_REG];
#pragma unroll
		for (int i = 0; i < LEAF_REG; i++) v[i] = i < n ? ix[b + i] : INT_MAX;
#pragma unroll
		for (int i = 0; i < LEAF_REG; i++) srt[i] = INT_MAX;
#pragma unroll
		for (int i = 0; i < LEAF_REG; i++)
		{
			int rk = 0;
#pragma unroll
			for (int j = 0; j < LEAF_REG; j++) rk += v[j] < v[i];
#pragma unroll
			for (int q = 0; q < LEAF_REG; q++)
				if (rk == q && i < n) srt[q] = v[i];
		}
		double2 zz[LEAF_REG];
#pragma unroll
		for (int q = 0; q < LEAF_REG; q++)
			if (q < n) zz[q] = roots[(size_t)row * R + srt[q]];
#pragma unroll
		for (int q = 0; q < LEAF_REG; q++)
			if (q < n)
			{
				ix[b + q] = srt[q];
				V.spos[(size_t)r * R + b + q] = zz[q];
				const cd w = (cd(zz[q].x, zz[q].y) - cc) * ih;
				cd t(1, 0);
#pragma unroll
				for (int k = 0; k <= P; k++)
				{
					A[k] += t;
					t = t * w;
				}
			}
		double2* out = V.mp + cid * (P + 1);
#pragma unroll
		for (int k = 0; k <= P; k++) out[k] = make_double2(A[k].re, A[k].im);
		return;
	}
	if (n <= 128)
	{
		for (int q = b + 1; q < e; q++)
		{
			const int v = ix[q];
			int t = q - 1;
			while (t >= b && ix[t] > v)
			{
				ix[t + 1] = ix[t];
				t--;
			}
			ix[t + 1] = v;
		}
	}
	for (int q = b; q < e; q++)
	{
		const double2 z = roots[(size_t)row * R + ix[q]];
		V.spos[(size_t)r * R + q] = z;
		const cd w = (cd(z.x, z.y) - cc) * ih;
		cd t(1, 0);
#pragma unroll
		for (int k = 0; k <= P; k++)
		{
			A[k] += t;
			t = t * w;
		}
	}
	double2* out = V.mp + cid * (P + 1);
#pragma unroll
	for (int k = 0; k <= P; k++) out[k] = make_double2(A[k].re, A[k].im);
}

// M2M: level l from its children at level l + 1
template <int P>
__global__ __launch_bounds__(BLOCK) void k_rm2m(RootView V, int l)
{
	const int r = blockIdx.y;
	if (!V.rmode[r]) return;
	const int nxl = V.g.nx << l, nyl = V.g.ny << l;
	const int c = blockIdx.x * BLOCK + threadIdx.x;
	if (c >= nxl * nyl) return;
	const int ix = c % nxl, iy = c / nxl, nxc = 2 * nxl;
	const size_t base = (size_t)r * V.g.ncells;
	int tot = 0;
	cd A[P + 1];
#pragma unroll
	for (int k = 0; k <= P; k++) A[k] = cd(0, 0);
	for (int q = 0; q < 4; q++)
	{
		const int a = q & 1, b = q >> 1;
		const size_t ch = base + V.g.off[l + 1] + (size_t)(2 * iy + b) * nxc + 2 * ix + a;
		const int n = V.ccnt[ch];
		if (n == 0) continue;
		tot += n;
		const double2* src = V.mp + ch * (P + 1);
		cd ak[P + 1];
		const cd e(a - 0.5, b - 0.5);
		double sc = 1;
#pragma unroll
		for (int k = 0; k <= P; k++)
		{
			ak[k] = cd(src[k].x * sc, src[k].y * sc);
			sc *= 0.5;
		}
		// binomial transform in place: a_n <- sum_k C(n, k) a_k e^(n-k)
#pragma unroll
		for (int i = 1; i <= P; i++)
#pragma unroll
			for (int n = P; n >= i; n--) ak[n] += e * ak[n - 1];
#pragma unroll
		for (int m = 0; m <= P; m++) A[m] += ak[m];
	}
	const size_t cid = base + V.g.off[l] + c;
	V.ccnt[cid] = tot;
	if (tot == 0) return;
	double2* out = V.mp + cid * (P + 1);
#pragma unroll
	for (int k = 0; k <= P; k++) out[k] = make_double2(A[k].re, A[k].im);
}

// L2L and M2L of level l with 32 lanes per target cell: lane k translates
// the k-th source of the target's interaction list into a partial local
// expansion, the partials are added through LDS in lane order, and lane m
// adds coefficient m of the parent's shifted expansion.
template <int P>
constexpr int down_targets() { return (P + 1) <= 14 ? 8 : 4; } // 32 (P+1) double2 of LDS per target

// The interaction list of a cell below level 0: the children of its
// parent's 3 x 3 block of cells (a 6 x 6 block, parent block row-major,
// four children each) that are not its neighbours. Entry k of list o (the
// cell's child position x + 2 y in its parent) is x + 6 y in that block; 27
// entries, so one per lane.
struct FarList
{
	unsigned char v[4][27];
};

constexpr FarList make_far_list()
{
	FarList t{};
	for (int o = 0; o < 4; o++)
	{
		const int tx = 2 + (o & 1), ty = 2 + (o >> 1);
		int n = 0;
		for (int b = 0; b < 9; b++)
			for (int q = 0; q < 4; q++)
			{
				const int jx = 2 * (b % 3) + (q & 1), jy = 2 * (b / 3) + (q >> 1);
				const int dx = jx > tx ? jx - tx : tx - jx, dy = jy > ty ? jy - ty : ty - jy;
				if (dx >= 2 || dy >= 2) t.v[o][n++] = (unsigned char)(jx + 6 * jy);
			}
	}
	return t;
}

__constant__ const FarList kFarList = make_far_list();
static_assert(make_far_list().v[3][26] != 0, "27 far cells per child position");

// M2L: adds to Lc the local expansion about a target cell of the multipole
// am of a source cell at offset (ix - jx, iy - jy) = (dx, dy) cells, on a
// level of cell side 2 / ih
template <int P>
__device__ __forceinline__ void m2l_add(const double2* am, int dx, int dy, double ih, cd (&Lc)[P + 1])
{
	const cd rho = cd(2.0 * dx, 2.0 * dy).inv();
	cd bq[P + 1];
	cd t(1, 0);
#pragma unroll
	for (int n = 0; n <= P; n++)
	{
		const double2 a = am[n];
		bq[n] = cd(a.x, a.y) * t;
		t = t * rho;
	}
	// sum_n C(n+m, m) b_n is element 0 after the (m+1)-th suffix-sum pass
	cd rp = rho * ih;
#pragma unroll
	for (int m = 0; m <= P; m++)
	{
#pragma unroll
		for (int k = P - 1; k >= 0; k--) bq[k] += bq[k + 1];
		const cd v = rp * bq[0];
		if (m & 1) Lc[m] -= v;
		else Lc[m] += v;
		rp = rp * rho;
	}
}

template <int P>
__global__ __launch_bounds__(32 * down_targets<P>()) void k_rdown32(RootView V, int l, unsigned long long* __restrict__ stats)
{
	constexpr int TPB = down_targets<P>();
	__shared__ double2 part[TPB][32][P + 1];
	__shared__ unsigned long long nsum[TPB * 32 / 64];
	const int r = blockIdx.y;
	if (!V.rmode[r]) return;
	const int nxl = V.g.nx << l, nyl = V.g.ny << l;
	const int tl = threadIdx.x >> 5, lane = threadIdx.x & 31;
	const int c = blockIdx.x * TPB + tl;
	const size_t base = (size_t)r * V.g.ncells;
	const size_t cid = base + V.g.off[l] + c;
	const bool live = c < nxl * nyl && V.ccnt[cid] > 0;
	const int ix = live ? c % nxl : 0, iy = live ? c / nxl : 0;
	const double ih = 2.0 / ldexp(V.g.s0, -l);
	cd Lc[P + 1];
#pragma unroll
	for (int k = 0; k <= P; k++) Lc[k] = cd(0, 0);
	unsigned long long nt = 0;
	auto m2l = [&](int jx, int jy) {
		const size_t sid = base + V.g.off[l] + (size_t)jy * nxl + jx;
		if (V.ccnt[sid] == 0) return;
		nt++;
		const double2* src = V.mp + sid * (P + 1);
		double2 am[P + 1];
#pragma unroll
		for (int n = 0; n <= P; n++) am[n] = src[n];
		m2l_add<P>(am, ix - jx, iy - jy, ih, Lc);
	};
	if (live)
	{
		if (l == 0)
		{
			for (int k = lane; k < nxl * nyl; k += 32)
			{
				const int jx = k % nxl, jy = k / nxl;
				if (abs(jx - ix) >= 2 || abs(jy - iy) >= 2) m2l(jx, jy);
			}
		}
		else
		{
			// the 27 far children of the parent's 3 x 3 block, one per lane
			if (lane < 27)
			{
				const int v = kFarList.v[(ix & 1) + 2 * (iy & 1)][lane];
				const int jx = 2 * ((ix >> 1) - 1) + v % 6, jy = 2 * ((iy >> 1) - 1) + v / 6;
				if (jx >= 0 && jy >= 0 && jx < nxl && jy < nyl) m2l(jx, jy);
			}
		}
	}
#pragma unroll
	for (int k = 0; k <= P; k++) part[tl][lane][k] = make_double2(Lc[k].re, Lc[k].im);
	{
		const unsigned long long s = wave_sum(nt);
		if ((threadIdx.x & 63) == 0) nsum[threadIdx.x >> 6] = s;
	}
	__syncthreads();
	if (live && lane <= P)
	{
		const int m = lane;
		double sx = 0, sy = 0;
		for (int k = 0; k < 32; k++)
		{
			const double2 v = part[tl][k][m];
			sx += v.x;
			sy += v.y;
		}
		if (l >= 1)
		{
			// L2L: 2^-m sum_{n >= m} C(n, m) L_n(parent) e^(n - m), the
			// parent's polynomial shifted by e (repeated synthetic division)
			const double2* lp = V.lc + (base + V.g.off[l - 1] + (size_t)(iy >> 1) * (nxl >> 1) + (ix >> 1)) * (P + 1);
			const cd e((ix & 1) - 0.5, (iy & 1) - 0.5);
			cd c[P + 1];
#pragma unroll
			for (int n = 0; n <= P; n++) c[n] = cd(lp[n].x, lp[n].y);
#pragma unroll
			for (int i = 0; i < P; i++)
#pragma unroll
				for (int j = P - 1; j >= i; j--) c[j] += e * c[j + 1];
			cd s(0, 0);
#pragma unroll
			for (int n = 0; n <= P; n++)
				if (n == m) s = c[n];
			const double sc = ldexp(1.0, -m);
			sx += s.re * sc;
			sy += s.im * sc;
		}
		V.lc[cid * (P + 1) + m] = make_double2(sx, sy);
	}
	if (threadIdx.x == 0)
	{
		unsigned long long s = 0;
		for (int w = 0; w < TPB * 32 / 64; w++) s += nsum[w];
		if (s) atomicAdd(stat_slot(stats, C_RM2L), s);
	}
}

// M2L and L2L of a level whose rows and columns of cells are multiples of
// DOWN_TILE (levels 3 and finer): a workgroup takes a DOWN_TILE^2 tile of
// target cells and stages the multipoles of the tile's whole interaction
// region, (DOWN_TILE + 4)^2 cells, in LDS once, where the 32-lane kernel has
// every target read its 27 sources through the cache. Four threads per
// target take every fourth entry of its list; lane shuffles add their
// partials, and each adds every fourth coefficient of the parent's shifted
// expansion.
constexpr int DOWN_TILE = 8;
constexpr int DOWN_SRC = DOWN_TILE + 4;
constexpr int DOWN_TILE_THREADS = DOWN_TILE * DOWN_TILE * 4;
template <int P>
constexpr bool down_tiled() { return (P + 1) <= 13; } // higher orders: the 32-lane kernel

template <int P>
__global__ __launch_bounds__(DOWN_TILE_THREADS) void k_rdown_tile(RootView V, int l, unsigned long long* __restrict__ stats)
{
	__shared__ double2 src[DOWN_SRC * DOWN_SRC][P + 1];
	__shared__ int sok[DOWN_SRC * DOWN_SRC];
	__shared__ unsigned long long nsum[DOWN_TILE_THREADS / 64];
	const int r = blockIdx.y;
	if (!V.rmode[r]) return;
	const int nxl = V.g.nx << l, nyl = V.g.ny << l;
	const int tiles_x = nxl / DOWN_TILE;
	const int x0 = (blockIdx.x % tiles_x) * DOWN_TILE, y0 = (blockIdx.x / tiles_x) * DOWN_TILE;
	// the region's first cell: the far lists reach two cells beyond the tile
	const int sx0 = x0 - 2, sy0 = y0 - 2;
	const size_t rowb = (size_t)r * V.g.ncells, lb = rowb + V.g.off[l];
	for (int i = threadIdx.x; i < DOWN_SRC * DOWN_SRC; i += DOWN_TILE_THREADS)
	{
		const int jx = sx0 + i % DOWN_SRC, jy = sy0 + i / DOWN_SRC;
		sok[i] = jx >= 0 && jy >= 0 && jx < nxl && jy < nyl && V.ccnt[lb + (size_t)jy * nxl + jx] > 0;
	}
	__syncthreads();
	for (int i = threadIdx.x; i < DOWN_SRC * DOWN_SRC * (P + 1); i += DOWN_TILE_THREADS)
	{
		const int c = i / (P + 1), n = i - c * (P + 1);
		if (sok[c])
		{
			const int jx = sx0 + c % DOWN_SRC, jy = sy0 + c / DOWN_SRC;
			src[c][n] = V.mp[(lb + (size_t)jy * nxl + jx) * (P + 1) + n];
		}
	}
	__syncthreads();
	const int t = threadIdx.x >> 2, q = threadIdx.x & 3;
	const int ix = x0 + t % DOWN_TILE, iy = y0 + t / DOWN_TILE;
	const size_t cid = lb + (size_t)iy * nxl + ix;
	const bool live = ix < nxl && iy < nyl && V.ccnt[cid] > 0;
	const double ih = 2.0 / ldexp(V.g.s0, -l);
	cd Lc[P + 1];
#pragma unroll
	for (int k = 0; k <= P; k++) Lc[k] = cd(0, 0);
	unsigned long long nt = 0;
	if (live)
	{
		// entry v of the far list is cell (v % 6, v / 6) of the 6 x 6 block
		// whose first cell is (2 (ix / 2) - 2, 2 (iy / 2) - 2): local
		// coordinates bx + v % 6, by + v / 6 in [0, DOWN_SRC)
		const int o = (ix & 1) + 2 * (iy & 1);
		const int bx = 2 * (ix >> 1) - x0, by = 2 * (iy >> 1) - y0;
		for (int k = q; k < 27; k += 4)
		{
			const int v = kFarList.v[o][k];
			const int lx = bx + v % 6, ly = by + v / 6;
			const int c = ly * DOWN_SRC + lx;
			if (!sok[c]) continue;
			nt++;
			m2l_add<P>(src[c], (ix - sx0) - lx, (iy - sy0) - ly, ih, Lc);
		}
	}
	// the four threads of a target: (q0 + q1) + (q2 + q3) in every one of them
#pragma unroll
	for (int k = 0; k <= P; k++)
		for (int d = 1; d <= 2; d <<= 1)
		{
			Lc[k].re += __shfl_xor(Lc[k].re, d, 64);
			Lc[k].im += __shfl_xor(Lc[k].im, d, 64);
		}
	{
		const unsigned long long sn = wave_sum(nt);
		if ((threadIdx.x & 63) == 0) nsum[threadIdx.x >> 6] = sn;
	}
	if (live)
	{
		// L2L: 2^-m sum_{n >= m} C(n, m) L_n(parent) e^(n - m), the parent's
		// polynomial shifted by e (repeated synthetic division)
		const double2* lp = V.lc + (rowb + V.g.off[l - 1] + (size_t)(iy >> 1) * (nxl >> 1) + (ix >> 1)) * (P + 1);
		const cd e((ix & 1) - 0.5, (iy & 1) - 0.5);
		cd c[P + 1];
#pragma unroll
		for (int n = 0; n <= P; n++) c[n] = cd(lp[n].x, lp[n].y);
#pragma unroll
		for (int i = 0; i < P; i++)
#pragma unroll
			for (int j = P - 1; j >= i; j--) c[j] += e * c[j + 1];
#pragma unroll
		for (int m = 0; m <= P; m++)
			if ((m & 3) == q)
			{
				const double sc = ldexp(1.0, -m);
				V.lc[cid * (P + 1) + m] = make_double2(Lc[m].re + c[m].re * sc, Lc[m].im + c[m].im * sc);
			}
	}
	__syncthreads();
	if (threadIdx.x == 0)
	{
		unsigned long long sn = 0;
		for (int w = 0; w < DOWN_TILE_THREADS / 64; w++) sn += nsum[w];
		if (sn) atomicAdd(stat_slot(stats, C_RM2L), sn);
	}
}

// per listed root of a row in FMM mode, in cell order: far field by L2P,
// near field and outliers directly; rsum = -S (the root part of the
// Aberth correction, the sign that the direct kernels' partials carry)
// one direct Cauchy term (z - u) / |z - u|^2 added to (ax, ay), in FP64.
// At orders up to 10, whose far field carries far larger relative errors,
// the reciprocal is the hardware estimate alone (relative error about
// 2^-23); above, rcp64 refines it to full precision. The own root gives a
// zero difference and, floored, a zero term.
template <int P>
__device__ __forceinline__ void direct_term(cd z, double2 u, double& ax, double& ay)
{
	const double dx = z.re - u.x, dy = z.im - u.y;
	const double d2 = fmax(fma(dx, dx, dy * dy), 1e-300);
	const double iv = P <= 10 ? __builtin_amdgcn_rcp(d2) : rcp64(d2);
	ax = fma(dx, iv, ax);
	ay = fma(dy, iv, ay);
}

template <int P>
__global__ __launch_bounds__(BLOCK) void k_rsum(RootView V, const double2* __restrict__ roots, RowMap M,
	unsigned long long* __restrict__ stats)
{
	const int r = blockIdx.y;
	if (!V.rmode[r]) return;
	const int R = M.R;
	const int k = blockIdx.x * BLOCK + threadIdx.x;
	const int* st = V.cstart + (size_t)r * (V.g.nleaf + 1);
	unsigned long long nn = 0, nt = 0;
	if (k < st[V.g.nleaf])
	{
		const int j = V.sidx[(size_t)r * R + k];
		const int s = V.slot[(size_t)r * R + j];
		if (s >= 0)
		{
			const double2 zz = V.spos[(size_t)r * R + k];
			const cd z(zz.x, zz.y);
			const int L = V.g.leaf, nxl = V.g.nx << L, nyl = V.g.ny << L;
			const int c = V.cell[(size_t)r * R + j];
			const int ix = c % nxl, iy = c / nxl;
			const cd w = (z - cell_center(V.g, L, ix, iy)) * (2.0 / ldexp(V.g.s0, -L));
			const double2* lp = V.lc + ((size_t)r * V.g.ncells + V.g.off[L] + c) * (P + 1);
			cd S(lp[P].x, lp[P].y);
#pragma unroll
			for (int m = P - 1; m >= 0; m--) S = S * w + cd(lp[m].x, lp[m].y);
			double ax = 0, ay = 0;
			const double2* sp = V.spos + (size_t)r * R;
			for (int y = max(iy - 1, 0); y <= min(iy + 1, nyl - 1); y++)
			{
				const int q0 = st[y * nxl + max(ix - 1, 0)], q1 = st[y * nxl + min(ix + 1, nxl - 1) + 1];
				nn += q1 - q0;
#pragma unroll 4
				for (int q = q0; q < q1; q++) direct_term<P>(z, sp[q], ax, ay);
			}
			const int no = V.ocnt[r];
			if (no > 0)
			{
				int row;
				double phi;
				M.get(r, row, phi);
				for (int o = 0; o < no; o++) direct_term<P>(z, roots[(size_t)row * R + V.oidx[(size_t)r * R + o]], ax, ay);
			}
			S += cd(ax, -ay);
			V.rsum[(size_t)r * R + s] = make_double2(-S.re, -S.im);
			nt = 1;
		}
	}
	const int which[2] = {C_RNEAR, C_RTGT};
	unsigned long long v[2] = {nn, nt};
	block_stats(stats, which, v);
}

// listed roots outside the grid: direct sums over the whole row. Block b
// of a row takes its outliers b, b + OUT_BLOCKS, ..., OUT_GROUP at a time,
// so that each pass over the row's roots serves a group of them.
constexpr int OUT_GROUP = 8;
constexpr int OUT_THREADS = 1024; // a pass over the row's roots: R / 1024 loads per lane

template <int P>
__global__ __launch_bounds__(OUT_THREADS) void k_rsum_out(RootView V, const double2* __restrict__ roots, RowMap M,
	unsigned long long* __restrict__ stats)
{
	// grid (rows, OUT_BLOCKS): the outliers of a row split over its blocks
	const int r = blockIdx.x;
	if (!V.rmode[r]) return;
	const int no = V.ocnt[r];
	if (no == 0) return;
	const int R = M.R;
	int row;
	double phi;
	M.get(r, row, phi);
	const double2* zr = roots + (size_t)row * R;
	__shared__ double red[OUT_GROUP][2][OUT_THREADS / 64];
	for (int o0 = blockIdx.y; o0 < no; o0 += gridDim.y * OUT_GROUP)
	{
		// the group: outliers o0 + k gridDim.y that are listed (slot >= 0)
		int j[OUT_GROUP], sl[OUT_GROUP];
		cd z[OUT_GROUP];
		double ax[OUT_GROUP], ay[OUT_GROUP];
#pragma unroll
		for (int k = 0; k < OUT_GROUP; k++)
		{
			const int o = o0 + k * gridDim.y;
			j[k] = o < no ? V.oidx[(size_t)r * R + o] : -1;
			sl[k] = j[k] >= 0 ? V.slot[(size_t)r * R + j[k]] : -1;
			const double2 v = sl[k] >= 0 ? zr[j[k]] : make_double2(0, 0);
			z[k] = cd(v.x, v.y);
			ax[k] = ay[k] = 0;
		}
		for (int q = threadIdx.x; q < R; q += OUT_THREADS)
		{
			const double2 u = zr[q];
#pragma unroll
			for (int k = 0; k < OUT_GROUP; k++)
				if (sl[k] >= 0 && q != j[k]) direct_term<P>(z[k], u, ax[k], ay[k]);
		}
#pragma unroll
		for (int k = 0; k < OUT_GROUP; k++)
		{
			for (int d = 32; d >= 1; d >>= 1)
			{
				ax[k] += __shfl_xor(ax[k], d, 64);
				ay[k] += __shfl_xor(ay[k], d, 64);
			}
			if ((threadIdx.x & 63) == 0)
			{
				red[k][0][threadIdx.x >> 6] = ax[k];
				red[k][1][threadIdx.x >> 6] = ay[k];
			}
		}
		__syncthreads();
		if (threadIdx.x < OUT_GROUP && sl[threadIdx.x] >= 0)
		{
			const int k = threadIdx.x;
			double sx = 0, sy = 0;
			for (int w = 0; w < OUT_THREADS / 64; w++)
			{
				sx += red[k][0][w];
				sy += red[k][1][w];
			}
			V.rsum[(size_t)r * R + sl[k]] = make_double2(-sx, sy);
			atomicAdd(stat_slot(stats, C_RTGT), 1ull);
		}
		__syncthreads();
	}
}

// the FMM pass of one sweep over nrows rows (rows not in FMM mode exit at once)
template <int P>
void root_fmm_pass_p(const SweepBufs& B, const double2* roots, const RowMap& M, int nrows, unsigned long long* stats, hipStream_t s)
{
	const RootView V = root_view(B);
	const RootGrid& g = B.grid;
	const int R = M.R;
	if (g.bin_row)
		k_rbin_row<<<nrows, BIN_THREADS, bin_row_lds(g.nleaf, R), s>>>(V, roots, M, stats);
	else
	{
		k_rzero<<<dim3(grid(g.nleaf), nrows), BLOCK, 0, s>>>(V);
		k_rbin<<<dim3(grid(R), nrows), BLOCK, 0, s>>>(V, roots, M);
		k_rscan<<<nrows, 1024, 0, s>>>(V, R, stats);
		k_rscatter<<<dim3(grid(R), nrows), BLOCK, 0, s>>>(V, R);
	}
	k_rleaf<P><<<dim3(grid(g.nleaf), nrows), BLOCK, 0, s>>>(V, roots, M);
	for (int l = g.leaf - 1; l >= 0; l--)
		k_rm2m<P><<<dim3(grid((g.nx << l) * (g.ny << l)), nrows), BLOCK, 0, s>>>(V, l);
	constexpr int TPB = down_targets<P>();
	for (int l = 0; l <= g.leaf; l++)
	{
		// from 1024 cells (16 tiles) per row, with nx 2^l and ny 2^l multiples
		// of DOWN_TILE (l >= 3); on smaller levels a few workgroups would
		// each carry a whole tile's lists
		if constexpr (down_tiled<P>())
			if (l >= 3 && (g.nx << l) * (g.ny << l) >= 1024)
			{
				k_rdown_tile<P><<<dim3((g.nx << l) / DOWN_TILE * ((g.ny << l) / DOWN_TILE), nrows), DOWN_TILE_THREADS, 0, s>>>(V, l, stats);
				continue;
			}
		k_rdown32<P><<<dim3(grid((g.nx << l) * (g.ny << l), TPB), nrows), 32 * TPB, 0, s>>>(V, l, stats);
	}
	k_rsum<P><<<dim3(grid(R), nrows), BLOCK, 0, s>>>(V, roots, M, stats);
	k_rsum_out<P><<<dim3(nrows, OUT_BLOCKS), OUT_THREADS, 0, s>>>(V, roots, M, stats);
}

// compiled expansion orders; a requested order runs at the next one up
constexpr int ROOT_ORDERS[] = {6, 8, 10, 12, 14, 16, 20, 24};

int root_order_compiled(int p)
{
	for (int q : ROOT_ORDERS)
		if (q >= p) return q;
	return MAX_ROOT_FMM_ORDER;
}

void root_fmm_pass(const SweepBufs& B, const double2* roots, const RowMap& M, int nrows, unsigned long long* stats, hipStream_t s)
{
	switch (B.grid.order)
	{
	case 6: root_fmm_pass_p<6>(B, roots, M, nrows, stats, s); break;
	case 8: root_fmm_pass_p<8>(B, roots, M, nrows, stats, s); break;
	case 10: root_fmm_pass_p<10>(B, roots, M, nrows, stats, s); break;
	case 12: root_fmm_pass_p<12>(B, roots, M, nrows, stats, s); break;
	case 14: root_fmm_pass_p<14>(B, roots, M, nrows, stats, s); break;
	case 16: root_fmm_pass_p<16>(B, roots, M, nrows, stats, s); break;
	case 20: root_fmm_pass_p<20>(B, roots, M, nrows, stats, s); break;
	default: root_fmm_pass_p<24>(B, roots, M, nrows, stats, s); break;
	}
}

// grid geometry for a field: top cells square, the depth chosen so that the
// mean number of roots per leaf cell is closest (in ratio) to `leaf_target`
RootGrid plan_root_grid(double hx, double hy, int R, int order, double leaf_target)
{
	RootGrid g;
	const double s0 = 2.0 * std::min(hx, hy);
	g.nx = std::max(1, (int)std::ceil(2.0 * hx / s0 - 1e-9));
	g.ny = std::max(1, (int)std::ceil(2.0 * hy / s0 - 1e-9));
	g.s0 = s0;
	g.x0 = -0.5 * g.nx * s0;
	g.y0 = -0.5 * g.ny * s0;
	int best = 0;
	double bd = 1e300;
	for (int l = 0; l <= MAX_ROOT_GRID_LEVELS; l++)
	{
		const double cells = (double)g.nx * g.ny * std::ldexp(1.0, 2 * l);
		if (cells > 4.0 * R + 64) break;
		const double d = std::fabs(std::log((double)R / cells / leaf_target));
		if (d < bd)
		{
			bd = d;
			best = l;
		}
	}
	g.leaf = best;
	g.order = order;
	int tot = 0;
	for (int l = 0; l <= g.leaf; l++)
	{
		g.off[l] = tot;
		tot += (g.nx << l) * (g.ny << l);
	}
	g.ncells = tot;
	g.nleaf = (g.nx << g.leaf) * (g.ny << g.leaf);
	return g;
}
