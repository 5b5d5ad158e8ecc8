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
//    sort by root index inside each cell so that sums are deterministic),
//    the leaf cells numbered region by region (leaf_key) so that the roots
//    of a block of up to 8 x 8 leaf cells are contiguous.
//  * Expansions are of the Cauchy kernel 1/(z - w), scaled by the half side
//    h of the cell's level (Greengard & Rokhlin 1987, written for the
//    complex kernel rather than for log|z - w|):
//      multipole   A_n = sum_j ((z_j - c)/h)^n,
//                  sum_j 1/(z - z_j) = (1/h) sum_n A_n / w^(n+1),  w = (z - c)/h
//      local       sum_far 1/(z - z_j) = sum_m L_m w^m
//      M2M         A_n(parent) = sum_k C(n,k) 2^-k A_k(child) e^(n-k),  e = (c_child - c_parent)/h_parent
//      M2L         L_m += (-1)^m / h sum_n C(n+m,m) A_n d^-(n+m+1),    d = (c_target - c_source)/h
//      L2L         L_k(child) = 2^-k sum_{m>=k} C(m,k) L_m(parent) e^(m-k)
//    M2M with a table of binomials in LDS, one coefficient per thread; L2L
//    as a Taylor shift by repeated synthetic division, and M2L's
//    sum_n C(n+m,m) b_n as element 0 after m + 1 suffix-sum passes over
//    b_n = A_n d^-n (no table: its constants would crowd the registers).
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
constexpr int ROOT_FMM_ORDER = 8;          // expansion order unless given
constexpr int ROOT_FMM_LEAF = 6;           // mean roots per leaf cell the grid depth aims at unless given
// sweeps of a phase that may use the FMM: the dense ones (the initial solve
// converges over a few dozen sweeps; a continuation step mostly in one or two)
constexpr int ROOT_FMM_INIT_SWEEPS = 24;
constexpr int ROOT_FMM_STEP_SWEEPS = 2;
constexpr int ROOT_FMM_PROBE = 8;

// regions: blocks of up to REGION_SIDE x REGION_SIDE leaf cells, one cell of
// level rlev = max(0, leaf - REGION_LOG) each, and the workgroup size of the
// kernels over them
constexpr int REGION_LOG = 3;
constexpr int REGION_SIDE = 1 << REGION_LOG;
constexpr int REGION_CELLS = REGION_SIDE * REGION_SIDE;
constexpr int REGION_TREE = REGION_CELLS + REGION_CELLS / 4 + REGION_CELLS / 16 + 1; // its cells on levels leaf .. rlev
constexpr int REGION_THREADS = 256;
constexpr int RUP_STAGE = 1024; // roots of a region that k_rup sorts and expands from LDS

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
	int* bidx;
	int* sidx;
	double2* spos;
	int* ocnt;
	int* oidx;
	double2* mp;
	double2* lcq;
};

RootView root_view(const SweepBufs& B)
{
	return RootView{B.grid, B.slot, B.rmode, B.rsum, B.ccnt, B.cstart, B.ccur, B.cell, B.bidx, B.sidx, B.spos, B.ocnt, B.oidx,
		B.mp, B.lcq};
}

__device__ inline cd cell_center(const RootGrid& g, int l, int ix, int iy)
{
	const double s = ldexp(g.s0, -l);
	return cd(g.x0 + (ix + 0.5) * s, g.y0 + (iy + 0.5) * s);
}

// sort key of leaf cell (ix, iy): its region's row-major index over level
// rlev, then the cell's row-major place in the region's block
__device__ __forceinline__ int leaf_key(const RootGrid& g, int ix, int iy)
{
	const int s = g.leaf - g.rlev, m = (1 << s) - 1;
	const int rg = (iy >> s) * (g.nx << g.rlev) + (ix >> s);
	return (rg << (2 * s)) + ((iy & m) << s) + (ix & m);
}

// cells of levels 0 .. rlev (the coarse cells, whose local expansions
// k_rcoarse forms once per row)
__host__ __device__ inline int coarse_cells(const RootGrid& g) { return g.rlev < g.leaf ? g.off[g.rlev + 1] : g.ncells; }

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
		const double isl = ldexp(1.0 / V.g.s0, V.g.leaf);
		const double fx = (z.x - V.g.x0) * isl, fy = (z.y - V.g.y0) * isl;
		const int nxl = V.g.nx << V.g.leaf, nyl = V.g.ny << V.g.leaf;
		if (fx >= 0 && fx < nxl && fy >= 0 && fy < nyl) c = leaf_key(V.g, (int)fx, (int)fy);
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
	if (c >= 0) V.bidx[(size_t)r * R + pos] = j;
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
// (k_rup sorts it). LDS: nleaf ints, then per 64 roots a mask and an int.
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
	int* __restrict__ cell = V.cell + (size_t)r * R;
	for (int c = threadIdx.x; c < nl; c += BIN_THREADS) cnt[c] = 0;
	__syncthreads();
	int row;
	double phi;
	M.get(r, row, phi);
	const double2* __restrict__ zr = roots + (size_t)row * R;
	const double isl = ldexp(1.0 / V.g.s0, V.g.leaf);
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
				const double fx = (z[k].x - V.g.x0) * isl, fy = (z[k].y - V.g.y0) * isl;
				if (fx >= 0 && fx < nxl && fy >= 0 && fy < nyl) c = leaf_key(V.g, (int)fx, (int)fy);
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
	int* __restrict__ bidx = V.bidx + (size_t)r * R;
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
			if (c[k] >= 0) bidx[atomicAdd(&cnt[c[k]], 1)] = j0 + k * BIN_THREADS;
	}
}

// The interaction list of a cell below level 0: the children of its
// parent's 3 x 3 block of cells (a 6 x 6 block, parent block row-major,
// four children each) that are not its neighbours. Entry k of list o (the
// cell's child position x + 2 y in its parent) is x + 6 y in that block; 27
// entries.
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

// listed roots outside the grid: direct sums over the whole row. Workgroup
// b of nb takes the row's outliers b, b + nb, ..., OUT_GROUP at a time, so
// that each pass over the row's roots serves a group of them.
constexpr int OUT_GROUP = 8;
constexpr int RSUM_OUT_BLOCKS = 16;

template <int P>
__device__ void rsum_outliers(const RootView& V, const double2* __restrict__ zr, int R, int r, int b, int nb,
	unsigned long long* __restrict__ stats)
{
	constexpr int NT = REGION_THREADS;
	__shared__ double red[OUT_GROUP][2][NT / 64];
	const int no = V.ocnt[r];
	for (int o0 = b; o0 < no; o0 += nb * OUT_GROUP)
	{
		// the group: outliers o0 + k nb that are listed (slot >= 0)
		int j[OUT_GROUP], sl[OUT_GROUP];
		cd z[OUT_GROUP];
		double ax[OUT_GROUP], ay[OUT_GROUP];
#pragma unroll
		for (int k = 0; k < OUT_GROUP; k++)
		{
			const int o = o0 + k * nb;
			j[k] = o < no ? V.oidx[(size_t)r * R + o] : -1;
			sl[k] = j[k] >= 0 ? V.slot[(size_t)r * R + j[k]] : -1;
			const double2 v = sl[k] >= 0 ? zr[j[k]] : make_double2(0, 0);
			z[k] = cd(v.x, v.y);
			ax[k] = ay[k] = 0;
		}
		for (int q = threadIdx.x; q < R; q += NT)
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
			for (int w = 0; w < NT / 64; w++)
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

// After the binning a pass is three launches. Regions are the cells of
// level g.rlev, each a block of RS x RS leaf cells (RS = 2^(leaf - rlev) <=
// REGION_SIDE) whose roots leaf_key keeps contiguous in the sorted arrays:
//  k_rup     one workgroup per region: each leaf cell's roots sorted by
//            index, P2M on the leaf cells and M2M up to level rlev in LDS.
//  k_rcoarse one workgroup per row: M2M above rlev, then the local
//            expansions of every cell of levels 0 .. rlev (M2L, L2L).
//  k_rdown   one workgroup per region: from its own cell's local expansion,
//            M2L and L2L over its cells down to the leaf (the sources of a
//            level's targets staged in LDS), then L2P and the direct near
//            field of each listed root of the region.
// k_rup's launch also carries RSUM_OUT_BLOCKS workgroups per row that sum
// the roots outside the grid directly.

// coefficient n of a child's multipole shifted to its parent,
// sum_k C(n, k) 2^-k a_k e^(n - k) (the child's unit is half the parent's)
template <int P>
__device__ __forceinline__ cd m2m_term(const double2* a, int n, cd e, const double (*binom)[P + 1])
{
	cd sum(0, 0), ep(1, 0);
	for (int k = n; k >= 0; k--)
	{
		const double2 v = a[k];
		sum += cd(v.x, v.y) * ep * (binom[n][k] * ldexp(1.0, -k));
		ep = ep * e;
	}
	return sum;
}

// L2L: L_m(child) = 2^-m sum_{n >= m} C(n, m) L_n(parent) e^(n - m), the
// parent's polynomial shifted by e (repeated synthetic division), in place
template <int P>
__device__ __forceinline__ void l2l_shift(cd (&Lc)[P + 1], cd e)
{
#pragma unroll
	for (int i = 0; i < P; i++)
#pragma unroll
		for (int j = P - 1; j >= i; j--) Lc[j] += e * Lc[j + 1];
#pragma unroll
	for (int m = 0; m <= P; m++) Lc[m] *= ldexp(1.0, -m);
}

// the cell of sorted position q among nc <= 128 cells: the last k with
// cst[k] <= q (cst non-decreasing, cst[0] <= q < cst[nc]), by a fixed
// number of halving steps
__device__ __forceinline__ int region_cell(const int* cst, int nc, int q)
{
	int lo = 0;
#pragma unroll
	for (int step = 64; step > 0; step >>= 1)
	{
		const int j = min(lo + step, nc); // cst[nc] > q: never taken
		lo = cst[j] <= q ? j : lo;
	}
	return lo;
}
static_assert(REGION_CELLS <= 128 && (REGION_SIDE + 2) * (REGION_SIDE + 2) <= 128, "region_cell searches at most 128 cells");

template <int P>
__global__ __launch_bounds__(REGION_THREADS) void k_rup(RootView V, const double2* __restrict__ roots, RowMap M,
	unsigned long long* __restrict__ stats)
{
	constexpr int NT = REGION_THREADS, NC = P + 1, NPT = (NC + 3) / 4;
	__shared__ int cst[REGION_CELLS + 1];
	__shared__ int tcnt[REGION_TREE];
	__shared__ double2 am[REGION_TREE][NC];
	__shared__ double binom[NC][NC];
	__shared__ int sbi[RUP_STAGE];     // the region's binned root indices
	__shared__ double2 sz[RUP_STAGE];  // and their positions in sorted order
	const int r = blockIdx.y;
	if (!V.rmode[r]) return;
	const RootGrid& g = V.g;
	const int Q = g.rlev, L = g.leaf, s = L - Q, RS = 1 << s, RC = RS * RS;
	const int nxq = g.nx << Q, nreg = nxq * (g.ny << Q), rg = blockIdx.x, rx = rg % nxq, ry = rg / nxq;
	const int R = M.R;
	int row;
	double phi;
	M.get(r, row, phi);
	if (rg >= nreg)
	{
		rsum_outliers<P>(V, roots + (size_t)row * R, R, r, rg - nreg, gridDim.x - nreg, stats);
		return;
	}
	const size_t rowc = (size_t)r * g.ncells;
	for (int i = threadIdx.x; i <= RC; i += NT) cst[i] = V.cstart[(size_t)r * (g.nleaf + 1) + (size_t)rg * RC + i];
	// C(n, k) for n, k <= P by the product formula: every partial product
	// is an integer below 2^53, so each step is exact
	for (int i = threadIdx.x; i < NC * NC; i += NT)
	{
		const int n = i / NC, k = i - n * NC;
		double c = k <= n ? 1.0 : 0.0;
		for (int t = 1; t <= k && k <= n; t++) c = c * (n - k + t) / t;
		binom[n][k] = c;
	}
	__syncthreads();
	// each cell's roots in index order (the binning leaves them in the order
	// of its atomics): a root's rank among its cell's indices, then its
	// position copied beside it. A region of at most RUP_STAGE roots works
	// from a copy of its indices in LDS and leaves its positions there.
	const int b0 = cst[0], e0 = cst[RC];
	const bool staged = e0 - b0 <= RUP_STAGE;
	const int* __restrict__ bi = V.bidx + (size_t)r * R;
	if (staged)
		for (int i = threadIdx.x; i < e0 - b0; i += NT) sbi[i] = bi[b0 + i];
	__syncthreads();
	{
		int* __restrict__ si = V.sidx + (size_t)r * R;
		double2* __restrict__ sp = V.spos + (size_t)r * R;
		const double2* __restrict__ zr = roots + (size_t)row * R;
		// binned root index at position q of the row's cell order
		auto lb = [&](int q) { return staged ? sbi[q - b0] : bi[q]; };
		for (int q = b0 + threadIdx.x; q < e0; q += NT)
		{
			const int k = region_cell(cst, RC, q);
			const int cb = cst[k], ce = cst[k + 1], v = lb(q);
			int rank = 0;
			for (int t = cb; t < ce; t++) rank += lb(t) < v;
			const double2 z = zr[v];
			si[cb + rank] = v;
			sp[cb + rank] = z;
			if (staged) sz[cb + rank - b0] = z;
		}
	}
	__syncthreads();
	// P2M: NT / RC threads per leaf cell, thread p forming the coefficients
	// n = p, p + tpc, ... as sums over the cell's roots in index order
	{
		const int tpc = NT / RC, k = threadIdx.x / tpc, p = threadIdx.x - k * tpc;
		const int ix = (rx << s) + (k & (RS - 1)), iy = (ry << s) + (k >> s);
		const int cb = cst[k], ce = cst[k + 1];
		const cd cc = cell_center(g, L, ix, iy);
		const double ih = 2.0 / ldexp(g.s0, -L);
		const double2* __restrict__ sp = V.spos + (size_t)r * R;
		cd acc[NPT];
#pragma unroll
		for (int i = 0; i < NPT; i++) acc[i] = cd(0, 0);
		if (p <= P)
			for (int q = cb; q < ce; q++)
			{
				const double2 z = staged ? sz[q - b0] : sp[q];
				const cd w = (cd(z.x, z.y) - cc) * ih;
				cd wp(1, 0), wt(1, 0);
				for (int t = 0; t < p; t++) wp = wp * w;
				if (tpc <= P)
					for (int t = 0; t < tpc; t++) wt = wt * w;
#pragma unroll
				for (int i = 0; i < NPT; i++)
				{
					if (p + i * tpc <= P) acc[i] += wp;
					wp = wp * wt;
				}
			}
		const size_t cid = rowc + g.off[L] + (size_t)iy * (g.nx << L) + ix;
#pragma unroll
		for (int i = 0; i < NPT; i++)
		{
			const int n = p + i * tpc;
			if (n <= P)
			{
				am[k][n] = make_double2(acc[i].re, acc[i].im);
				V.mp[cid * NC + n] = am[k][n];
			}
		}
		if (p == 0)
		{
			tcnt[k] = ce - cb;
			V.ccnt[cid] = ce - cb;
		}
	}
	__syncthreads();
	// M2M inside the region, level by level up to rlev; a level of side sl
	// follows its children's level (side 2 sl) in am and tcnt
	int cofs = 0;
	for (int l = L - 1; l >= Q; l--)
	{
		const int sl = 1 << (l - Q), nc = sl * sl, pofs = cofs + 4 * nc, nxl = g.nx << l;
		for (int i = threadIdx.x; i < nc * NC; i += NT)
		{
			const int c = i / NC, n = i - c * NC, cx = c & (sl - 1), cy = c / sl;
			cd sum(0, 0);
			int tot = 0;
			for (int qd = 0; qd < 4; qd++)
			{
				const int a = qd & 1, b = qd >> 1, ch = cofs + (2 * cy + b) * (2 * sl) + 2 * cx + a;
				tot += tcnt[ch];
				if (tcnt[ch] > 0) sum += m2m_term<P>(am[ch], n, cd(a - 0.5, b - 0.5), binom);
			}
			const size_t cid = rowc + g.off[l] + (size_t)((ry << (l - Q)) + cy) * nxl + (rx << (l - Q)) + cx;
			am[pofs + c][n] = make_double2(sum.re, sum.im);
			V.mp[cid * NC + n] = am[pofs + c][n];
			if (n == 0)
			{
				tcnt[pofs + c] = tot;
				V.ccnt[cid] = tot;
			}
		}
		__syncthreads();
		cofs = pofs;
	}
}

// One workgroup per row between k_rup and k_rdown: the multipoles of the
// cells above the regions (levels rlev - 1 .. 0, M2M from their children),
// then the local expansion of every cell of levels 0 .. rlev: L2L from its
// parent (levels >= 1) and M2L from its interaction list (at level 0 every
// top cell that is not a neighbour), tpt threads per cell. They go to lcq
// (the row's coarse cells in the grid's cell order), where each region of
// k_rdown finds its own cell's; before, every region formed this chain of
// ancestors itself.
// (512 threads while the registers of an order allow two waves a SIMD)
constexpr size_t RCOARSE_LDS = 48 * 1024;

template <int P>
constexpr int coarse_threads()
{
	return P <= 16 ? 512 : 256;
}

template <int P>
__global__ __launch_bounds__(coarse_threads<P>()) void k_rcoarse(RootView V, unsigned long long* __restrict__ stats, int lds)
{
	constexpr int NT = coarse_threads<P>(), NC = P + 1;
	__shared__ double binom[NC][NC];
	// with lds set, the coarse cells' multipoles, local expansions and
	// counts (the first coarse_cells(g) cells of the row, levels 0 .. rlev)
	// live in LDS, launched with their size; level rlev's multipoles and
	// counts come in from k_rup, and only level rlev's local expansions,
	// which k_rdown reads, and the counts go back out
	extern __shared__ double2 dyn[];
	const int r = blockIdx.x;
	if (!V.rmode[r]) return;
	const RootGrid& g = V.g;
	const int Q = g.rlev, CC = coarse_cells(g);
	const size_t rowc = (size_t)r * g.ncells, rowq = (size_t)r * CC;
	double2* const mpb = lds ? dyn : V.mp + rowc * NC;
	double2* const lqb = lds ? dyn + (size_t)CC * NC : V.lcq + rowq * NC;
	int* const ccb = lds ? reinterpret_cast<int*>(dyn + (size_t)2 * CC * NC) : V.ccnt + rowc;
	const int* __restrict__ cc = ccb;
	if (lds)
	{
		const int q0 = g.off[Q];
		for (int i = threadIdx.x; i < (CC - q0) * NC; i += NT) mpb[(size_t)q0 * NC + i] = V.mp[(rowc + q0) * NC + i];
		for (int i = threadIdx.x; i < CC - q0; i += NT) ccb[q0 + i] = V.ccnt[rowc + q0 + i];
	}
	// C(n, k) for n, k <= P as in k_rup
	for (int i = threadIdx.x; i < NC * NC; i += NT)
	{
		const int n = i / NC, k = i - n * NC;
		double c = k <= n ? 1.0 : 0.0;
		for (int t = 1; t <= k && k <= n; t++) c = c * (n - k + t) / t;
		binom[n][k] = c;
	}
	__syncthreads();
	for (int l = Q - 1; l >= 0; l--)
	{
		const int nxl = g.nx << l, nyl = g.ny << l, nxc = 2 * nxl;
		for (int i = threadIdx.x; i < nxl * nyl * NC; i += NT)
		{
			const int c = i / NC, n = i - c * NC, ix = c % nxl, iy = c / nxl;
			cd sum(0, 0);
			int tot = 0;
			for (int qd = 0; qd < 4; qd++)
			{
				const int a = qd & 1, b = qd >> 1;
				const size_t ch = g.off[l + 1] + (size_t)(2 * iy + b) * nxc + 2 * ix + a;
				const int nch = ccb[ch];
				tot += nch;
				if (nch > 0) sum += m2m_term<P>(mpb + ch * NC, n, cd(a - 0.5, b - 0.5), binom);
			}
			const size_t cid = g.off[l] + c;
			mpb[cid * NC + n] = make_double2(sum.re, sum.im);
			if (n == 0) ccb[cid] = tot;
		}
		__syncthreads();
	}
	unsigned long long nm = 0;
	for (int l = 0; l <= Q; l++)
	{
		const int nxl = g.nx << l, nyl = g.ny << l, T = nxl * nyl;
		int tpt = 32;
		while (tpt > 1 && T * tpt > NT) tpt >>= 1;
		const double ih = 2.0 / ldexp(g.s0, -l);
		for (int t0 = 0; t0 < T; t0 += NT / tpt)
		{
			const int t = t0 + (int)threadIdx.x / tpt, q = (int)threadIdx.x % tpt;
			const int ix = t % nxl, iy = t / nxl;
			const bool live = t < T && cc[g.off[l] + t] > 0;
			cd Lc[NC];
#pragma unroll
			for (int k = 0; k < NC; k++) Lc[k] = cd(0, 0);
			if (live)
			{
				if (l >= 1 && q == 0)
				{
					const double2* lp = lqb + (g.off[l - 1] + (size_t)(iy >> 1) * (nxl >> 1) + (ix >> 1)) * NC;
#pragma unroll
					for (int n = 0; n < NC; n++) Lc[n] = cd(lp[n].x, lp[n].y);
					l2l_shift<P>(Lc, cd((ix & 1) - 0.5, (iy & 1) - 0.5));
				}
				if (l == 0)
				{
					for (int k = q; k < T; k += tpt)
					{
						const int jx = k % nxl, jy = k / nxl;
						if ((abs(jx - ix) >= 2 || abs(jy - iy) >= 2) && cc[k] > 0)
						{
							nm++;
							m2l_add<P>(mpb + (size_t)k * NC, ix - jx, iy - jy, ih, Lc);
						}
					}
				}
				else
				{
					const int o = (ix & 1) + 2 * (iy & 1), bx = 2 * (ix >> 1) - 2, by = 2 * (iy >> 1) - 2;
					for (int k = q; k < 27; k += tpt)
					{
						const int v = kFarList.v[o][k];
						const int jx = bx + v % 6, jy = by + v / 6;
						if (jx < 0 || jy < 0 || jx >= nxl || jy >= nyl) continue;
						const int c = g.off[l] + jy * nxl + jx;
						if (cc[c] <= 0) continue;
						nm++;
						m2l_add<P>(mpb + (size_t)c * NC, ix - jx, iy - jy, ih, Lc);
					}
				}
			}
			for (int d = 1; d < tpt; d <<= 1)
#pragma unroll
				for (int k = 0; k < NC; k++)
				{
					Lc[k].re += __shfl_xor(Lc[k].re, d, 64);
					Lc[k].im += __shfl_xor(Lc[k].im, d, 64);
				}
			if (live)
#pragma unroll
				for (int m = 0; m < NC; m++)
					if ((m & (tpt - 1)) == q) lqb[(g.off[l] + t) * NC + m] = make_double2(Lc[m].re, Lc[m].im);
		}
		// the level's expansions before the next level's L2L reads them
		__syncthreads();
	}
	if (lds)
	{
		// (the last level's barrier is behind)
		const int q0 = g.off[Q];
		for (int i = threadIdx.x; i < (CC - q0) * NC; i += NT) V.lcq[(rowq + q0) * NC + i] = lqb[(size_t)q0 * NC + i];
		for (int i = threadIdx.x; i < q0; i += NT) V.ccnt[rowc + i] = ccb[i];
	}
	const int which[1] = {C_RM2L};
	unsigned long long v[1] = {nm};
	block_stats<1, NT>(stats, which, v);
}

constexpr int HALO_SIDE = REGION_SIDE + 2;        // a region's leaf cells and their neighbours
constexpr int HALO_CELLS = HALO_SIDE * HALO_SIDE;
constexpr int STAGE_BATCH = 4;                    // loads per thread issued together when staging
constexpr int ROOT_BATCH = 2;                     // roots per thread whose list loads k_rdown issues together

// rsum = -S for every listed root of a row in FMM mode (the root part of
// the Aberth correction, the sign that the direct kernels' partials carry)
template <int P>
__global__ __launch_bounds__(REGION_THREADS) void k_rdown(RootView V, const double2* __restrict__ roots, RowMap M,
	unsigned long long* __restrict__ stats)
{
	constexpr int NT = REGION_THREADS, NC = P + 1, W = REGION_SIDE + 4, SRC = W * W * NC, SB = STAGE_BATCH;
	// a level's sources (coefficient n of staged cell c at c * NC + n), then
	// the positions of the roots of the region's cells and their neighbours
	__shared__ double2 src[SRC];
	__shared__ int sok[W * W];
	__shared__ double2 lbuf[2][REGION_CELLS][NC]; // local expansions of a level and of its parents
	__shared__ int cst[REGION_CELLS + 1];
	__shared__ int hst[HALO_CELLS + 1], hsrc[HALO_CELLS];
	__shared__ int wsum[NT / 64];
	const int r = blockIdx.y;
	if (!V.rmode[r]) return;
	const RootGrid& g = V.g;
	const int Q = g.rlev, L = g.leaf, s = L - Q, RS = 1 << s, RC = RS * RS;
	const int nxq = g.nx << Q;
	const int R = M.R;
	int row;
	double phi;
	M.get(r, row, phi);
	const double2* __restrict__ zr = roots + (size_t)row * R;
	const int rg = blockIdx.x, rx = rg % nxq, ry = rg / nxq;
	const size_t rowc = (size_t)r * g.ncells;
	const int* __restrict__ cc = V.ccnt + rowc;
	const int* __restrict__ st = V.cstart + (size_t)r * (g.nleaf + 1);
	for (int i = threadIdx.x; i <= RC; i += NT) cst[i] = st[(size_t)rg * RC + i];
	// the region's own cell (level rlev): its local expansion from k_rcoarse
	if ((int)threadIdx.x < NC)
		lbuf[Q & 1][0][threadIdx.x] = V.lcq[((size_t)r * coarse_cells(g) + g.off[Q] + (size_t)ry * nxq + rx) * NC + threadIdx.x];
	unsigned long long nm = 0;
	for (int l = Q + 1; l <= L; l++)
	{
		// the level's targets: the region's side x side cells, tpt threads
		// per target
		const int nxl = g.nx << l, nyl = g.ny << l, lofs = g.off[l];
		const int side = 1 << (l - Q), x0 = rx << (l - Q), y0 = ry << (l - Q);
		const int T = side * side, tpt = min(32, NT / T);
		const int t = threadIdx.x / tpt, q = threadIdx.x - t * tpt;
		const int ix = x0 + t % side, iy = y0 + t / side;
		const bool live = t < T && cc[lofs + iy * nxl + ix] > 0;
		const double ih = 2.0 / ldexp(g.s0, -l);
		double2(*lcur)[NC] = lbuf[l & 1];
		const double2(*lpar)[NC] = lbuf[(l & 1) ^ 1];
		cd Lc[NC];
#pragma unroll
		for (int k = 0; k < NC; k++) Lc[k] = cd(0, 0);
		// the sources, staged in LDS: the block of cells that holds the
		// targets' far lists, cells x0 - 2 .. x0 + side + 1 (x0, y0 even)
		const int bx0 = x0 - 2, by0 = y0 - 2, Wl = side + 4, ne = Wl * Wl * NC;
		for (int i0 = threadIdx.x; i0 < ne; i0 += NT * SB)
		{
			double2 v[SB];
			int n0[SB];
#pragma unroll
			for (int b = 0; b < SB; b++)
			{
				const int i = i0 + b * NT, c = i / NC, n = i - c * NC;
				const int jx = bx0 + c % Wl, jy = by0 + c / Wl;
				const bool in = i < ne && jx >= 0 && jy >= 0 && jx < nxl && jy < nyl;
				const size_t cid = lofs + (size_t)jy * nxl + jx;
				v[b] = in ? V.mp[(rowc + cid) * NC + n] : make_double2(0, 0);
				n0[b] = in && n == 0 ? cc[cid] : 0;
			}
#pragma unroll
			for (int b = 0; b < SB; b++)
			{
				const int i = i0 + b * NT, c = i / NC;
				if (i < ne)
				{
					src[i] = v[b];
					if (i == c * NC) sok[c] = n0[b] > 0;
				}
			}
		}
		__syncthreads();
		if (live)
		{
			// L2L into the group's first lane (the parents' expansions were
			// written before the barrier above), then the M2L partials
			if (q == 0)
			{
				const int ps = side >> 1;
				const int pi = ((iy >> 1) - (y0 >> 1)) * ps + ((ix >> 1) - (x0 >> 1));
#pragma unroll
				for (int n = 0; n < NC; n++) Lc[n] = cd(lpar[pi][n].x, lpar[pi][n].y);
				l2l_shift<P>(Lc, cd((ix & 1) - 0.5, (iy & 1) - 0.5));
			}
			// entry v of the far list is cell (v % 6, v / 6) of the 6 x 6
			// block whose first cell is (2 (ix / 2) - 2, 2 (iy / 2) - 2)
			const int o = (ix & 1) + 2 * (iy & 1), bx = 2 * (ix >> 1) - 2, by = 2 * (iy >> 1) - 2;
			for (int k = q; k < 27; k += tpt)
			{
				const int v = kFarList.v[o][k];
				const int jx = bx + v % 6, jy = by + v / 6;
				const int c = (jy - by0) * Wl + (jx - bx0);
				if (!sok[c]) continue;
				nm++;
				m2l_add<P>(src + c * NC, ix - jx, iy - jy, ih, Lc);
			}
		}
		__syncthreads();
		// the target's partials over its tpt threads (aligned groups of
		// lanes; the butterfly leaves the same sum in each)
		for (int d = 1; d < tpt; d <<= 1)
#pragma unroll
			for (int k = 0; k < NC; k++)
			{
				Lc[k].re += __shfl_xor(Lc[k].re, d, 64);
				Lc[k].im += __shfl_xor(Lc[k].im, d, 64);
			}
		// read by the next level after its staging barrier; the buffer it
		// overwrites held the parents, read before this level's barrier
		if (live)
#pragma unroll
			for (int m = 0; m < NC; m++)
				if ((m & (tpt - 1)) == q) lcur[t][m] = make_double2(Lc[m].re, Lc[m].im);
	}
	// the near-field sources: the roots of the region's leaf cells and of
	// their neighbours, HS x HS cells from (hx0, hy0), copied into LDS when
	// they fit (cell k's at hst[k] ..), else read where they lie
	const int nxl = g.nx << L, nyl = g.ny << L, HS = RS + 2, hx0 = (rx << s) - 1, hy0 = (ry << s) - 1;
	const double2* __restrict__ sp = V.spos + (size_t)r * R;
	int hc = 0;
	if ((int)threadIdx.x < HS * HS)
	{
		const int jx = hx0 + threadIdx.x % HS, jy = hy0 + threadIdx.x / HS;
		if (jx >= 0 && jy >= 0 && jx < nxl && jy < nyl)
		{
			const int key = leaf_key(g, jx, jy), b = st[key];
			hc = st[key + 1] - b;
			hsrc[threadIdx.x] = b;
		}
	}
	int htot;
	const int ho = block_scan_excl(hc, wsum, htot);
	if ((int)threadIdx.x < HS * HS) hst[threadIdx.x] = ho;
	if (threadIdx.x == 0) hst[HS * HS] = htot;
	const bool hlds = htot <= SRC;
	__syncthreads();
	if (hlds)
#pragma unroll 4
		for (int i = threadIdx.x; i < htot; i += NT)
		{
			const int k = region_cell(hst, HS * HS, i);
			src[i] = sp[hsrc[k] + i - hst[k]];
		}
	__syncthreads();
	// L2P, the near field (the 3 x 3 leaf cells around the root's) and the
	// roots outside the grid, per listed root of the region, ROOT_BATCH at a
	// time per thread
	const double2(*lleaf)[NC] = lbuf[L & 1];
	const double ih = 2.0 / ldexp(g.s0, -L);
	const int no = V.ocnt[r];
	const int* __restrict__ si = V.sidx + (size_t)r * R;
	const int* __restrict__ oi = V.oidx + (size_t)r * R;
	unsigned long long nn = 0, nt = 0;
	for (int q0 = cst[0] + threadIdx.x; q0 < cst[RC]; q0 += NT * ROOT_BATCH)
	{
		int sl[ROOT_BATCH];
#pragma unroll
		for (int b = 0; b < ROOT_BATCH; b++)
		{
			const int qq = q0 + b * NT;
			sl[b] = qq < cst[RC] ? si[qq] : -1;
		}
#pragma unroll
		for (int b = 0; b < ROOT_BATCH; b++) sl[b] = sl[b] >= 0 ? V.slot[(size_t)r * R + sl[b]] : -1;
#pragma unroll
		for (int b = 0; b < ROOT_BATCH; b++)
		{
			const int qq = q0 + b * NT;
			if (sl[b] < 0) continue;
			const int k = region_cell(cst, RC, qq);
			const int kx = k & (RS - 1), ky = k >> s, ix = (rx << s) + kx, iy = (ry << s) + ky;
			// the root's own cell in the near-field block: (kx + 1, ky + 1)
			const int hk = (ky + 1) * HS + kx + 1;
			const double2 zz = hlds ? src[hst[hk] + qq - hsrc[hk]] : sp[qq];
			const cd z(zz.x, zz.y);
			const cd w = (z - cell_center(g, L, ix, iy)) * ih;
			cd S(lleaf[k][P].x, lleaf[k][P].y);
#pragma unroll
			for (int m = P - 1; m >= 0; m--) S = S * w + cd(lleaf[k][m].x, lleaf[k][m].y);
			double ax = 0, ay = 0;
			for (int d = 0; d < 9; d++)
			{
				const int h = hk + (d / 3 - 1) * HS + d % 3 - 1;
				const int u0 = hst[h], u1 = hst[h + 1];
				nn += u1 - u0;
				if (hlds)
				{
					for (int u = u0; u < u1; u++) direct_term<P>(z, src[u], ax, ay);
				}
				else
				{
					const int gb = hsrc[h] - u0;
#pragma unroll 4
					for (int u = u0; u < u1; u++) direct_term<P>(z, sp[gb + u], ax, ay);
				}
			}
			for (int o = 0; o < no; o++) direct_term<P>(z, zr[oi[o]], ax, ay);
			S += cd(ax, -ay);
			V.rsum[(size_t)r * R + sl[b]] = make_double2(-S.re, -S.im);
			nt++;
		}
	}
	const int which[3] = {C_RM2L, C_RNEAR, C_RTGT};
	unsigned long long v[3] = {nm, nn, nt};
	block_stats<3, REGION_THREADS>(stats, which, v);
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
	const int nreg = (g.nx << g.rlev) * (g.ny << g.rlev);
	k_rup<P><<<dim3(nreg + RSUM_OUT_BLOCKS, nrows), REGION_THREADS, 0, s>>>(V, roots, M, stats);
	// the coarse levels in LDS when they fit in RCOARSE_LDS bytes
	const size_t cl = (size_t)2 * coarse_cells(g) * (P + 1) * sizeof(double2) + (size_t)coarse_cells(g) * sizeof(int);
	const int lds = cl <= RCOARSE_LDS ? 1 : 0;
	k_rcoarse<P><<<nrows, coarse_threads<P>(), lds ? cl : 0, s>>>(V, stats, lds);
	k_rdown<P><<<dim3(nreg, nrows), REGION_THREADS, 0, s>>>(V, roots, M, stats);
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
	g.rlev = std::max(0, best - REGION_LOG);
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
