// Quadtree and 2-D complex FMM build on gfx950.
//
// Same tree as weisluke/ccf (tree_node.cuh:183-267, 321-519; driver
// ccf.cuh:732-904): split every node whose own plus neighbour stars exceed
// 32, children in quadrant order (+,+), (-,+), (-,-), (+,-), neighbours
// closer than 3 (same level) or 3.5 (coarser leaf) half-lengths, the rest
// in the interaction lists. The construction differs in how, not what:
//  - nodes live in structure-of-arrays with integer ids, children of a node
//    are 4 consecutive ids given by an exclusive prefix sum over the split
//    flags (deterministic, where the reference hands out slots by atomics);
//  - each split node's stars are stably partitioned by one workgroup using
//    wavefront ballots, so the tree order of the stars is reproducible.
// The FMM passes (fmm.cuh) run one thread per (node, coefficient).
#include "internal.hpp"

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <iostream>

namespace ccf
{
namespace
{

constexpr int BLOCK = 256;

__device__ inline int quadrant(double2 z, double2 c)
{
	if (z.y >= c.y) return z.x >= c.x ? 0 : 1;
	return z.x < c.x ? 2 : 3;
}

__device__ inline cd ld(const double2* p, size_t i) { double2 v = p[i]; return cd(v.x, v.y); }
__device__ inline void st(double2* p, size_t i, cd v) { p[i] = make_double2(v.re, v.im); }

// near-field run offset of every node: its prefix rank for a leaf, -1
// otherwise
__global__ void k_leaf_off(int nn, const int* __restrict__ child, const int* __restrict__ rank, int* __restrict__ off)
{
	const int n = blockIdx.x * blockDim.x + threadIdx.x;
	if (n < nn) off[n] = child[n] < 0 ? rank[n] : -1;
}

// level 0 of a one-field tree: the root node (centre 0, half-length hl0,
// stars [s0, s0 + n)) and the level loop's four counters cleared
__global__ void k_root_node(double hl0, int s0, int n, double2* center, double* hl, int* parent, int* child, int* sstart,
	int* scount, int* nnbr, int* nsl, int* ndl, int* field, int* scal)
{
	center[0] = make_double2(0, 0);
	hl[0] = hl0;
	parent[0] = -1;
	child[0] = -1;
	sstart[0] = s0;
	scount[0] = n;
	nnbr[0] = 0;
	nsl[0] = 0;
	ndl[0] = 0;
	field[0] = 0;
	for (int k = 0; k < 4; k++) scal[k] = 0;
}

// scal[0] = max stars in node + neighbours over the level, scal[2] = list overflow flag
__global__ void k_level_count(int off, int cnt, const int* __restrict__ scount, const int* __restrict__ nbr,
	const int* __restrict__ nnbr, int* __restrict__ split, int* __restrict__ scal)
{
	int i = blockIdx.x * blockDim.x + threadIdx.x;
	int tot = 0;
	if (i < cnt)
	{
		int n = off + i;
		tot = scount[n];
		int k = nnbr[n];
		for (int j = 0; j < k; j++) tot += scount[nbr[n * MAX_NEIGHBORS + j]];
		split[i] = tot > MAX_NUM_STARS_DIRECT ? 1 : 0;
	}
	for (int d = 32; d >= 1; d >>= 1) tot = max(tot, __shfl_xor(tot, d, 64));
	if ((threadIdx.x & 63) == 0) atomicMax(&scal[0], tot);
}

// exclusive prefix sum of flags[0..n) into rank, total into *total (one block)
__global__ void k_scan(const int* __restrict__ flags, int n, int* __restrict__ rank, int* __restrict__ total)
{
	__shared__ int wsum[16];
	__shared__ int carry;
	const int lane = threadIdx.x & 63, w = threadIdx.x >> 6, nw = blockDim.x >> 6;
	if (threadIdx.x == 0) carry = 0;
	__syncthreads();
	for (int base = 0; base < n; base += blockDim.x)
	{
		int i = base + threadIdx.x;
		int v = i < n ? flags[i] : 0;
		int x = v;
		for (int d = 1; d < 64; d <<= 1)
		{
			int y = __shfl_up(x, d, 64);
			if (lane >= d) x += y;
		}
		if (lane == 63) wsum[w] = x;
		__syncthreads();
		int woff = 0;
		for (int k = 0; k < w; k++) woff += wsum[k];
		if (i < n) rank[i] = carry + woff + x - v;
		int blocktot = 0;
		for (int k = 0; k < nw; k++) blocktot += wsum[k];
		__syncthreads();
		if (threadIdx.x == 0) carry += blocktot;
		__syncthreads();
	}
	if (threadIdx.x == 0) *total = carry;
}

// (also clears scal[0], the next level's maximum, which the host has read)
__global__ void k_make_children(int off, int cnt, int new_off, const int* __restrict__ split, const int* __restrict__ rank,
	double2* center, double* hl, int* parent, int* child, int* sstart, int* scount, int* nnbr, int* nsl, int* ndl, int* field,
	int* scal)
{
	int i = blockIdx.x * blockDim.x + threadIdx.x;
	if (i == 0) scal[0] = 0;
	if (i >= cnt || !split[i]) return;
	int n = off + i;
	int c0 = new_off + 4 * rank[i];
	child[n] = c0;
	const double h = hl[n] / 2;
	const double2 c = center[n];
	const double ox[4] = {1, -1, -1, 1}, oy[4] = {1, 1, -1, -1};
	for (int q = 0; q < 4; q++)
	{
		int k = c0 + q;
		hl[k] = h;
		center[k] = make_double2(c.x + h * ox[q], c.y + h * oy[q]);
		parent[k] = n;
		child[k] = -1;
		sstart[k] = sstart[n];
		scount[k] = 0;
		nnbr[k] = 0;
		nsl[k] = 0;
		ndl[k] = 0;
		field[k] = field[n];
	}
}

// levels of at most this many nodes are partitioned by 1024-thread workgroups
constexpr int PARTITION_WIDE_NODES = 64;

// one workgroup per node of the level: stable 4-way partition of the stars
// of every split node into its children (tree_node.cuh:392-501); a star's
// place depends only on the order of the stars, not on BS
template <int BS>
__global__ __launch_bounds__(BS) void k_partition(int off, const int* __restrict__ split, const int* __restrict__ child,
	const double2* __restrict__ center, int* sstart, int* scount,
	double2* spos, double* smass, double2* tpos, double* tmass)
{
	const int i = blockIdx.x;
	if (!split[i]) return;
	const int n = off + i;
	const int c0 = child[n];
	const int start = sstart[n], count = scount[n];
	if (count == 0)
	{
		if (threadIdx.x < 4)
		{
			sstart[c0 + threadIdx.x] = start;
			scount[c0 + threadIdx.x] = 0;
		}
		return;
	}
	const double2 c = center[n];
	__shared__ int qcnt[4];
	__shared__ int run[4];
	__shared__ int wtot[BS / 64][4];
	if (threadIdx.x < 4) { qcnt[threadIdx.x] = 0; run[threadIdx.x] = 0; }
	__syncthreads();
	int lc[4] = {0, 0, 0, 0};
	for (int s = threadIdx.x; s < count; s += BS) lc[quadrant(spos[start + s], c)]++;
	for (int q = 0; q < 4; q++) atomicAdd(&qcnt[q], lc[q]);
	__syncthreads();
	const int base[4] = {0, qcnt[0], qcnt[0] + qcnt[1], qcnt[0] + qcnt[1] + qcnt[2]};
	const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
	const unsigned long long below = (lane == 0) ? 0ull : (~0ull >> (64 - lane));
	for (int chunk = 0; chunk < count; chunk += BS)
	{
		const int s = chunk + threadIdx.x;
		const int q = s < count ? quadrant(spos[start + s], c) : -1;
		int myrank = 0;
		for (int k = 0; k < 4; k++)
		{
			unsigned long long m = __ballot(q == k);
			if (q == k) myrank = __popcll(m & below);
			if (lane == 0) wtot[w][k] = __popcll(m);
		}
		__syncthreads();
		if (q >= 0)
		{
			int pos = base[q] + run[q] + myrank;
			for (int k = 0; k < w; k++) pos += wtot[k][q];
			tpos[start + pos] = spos[start + s];
			tmass[start + pos] = smass[start + s];
		}
		__syncthreads();
		if (threadIdx.x < 4)
		{
			int t = 0;
			for (int k = 0; k < BS / 64; k++) t += wtot[k][threadIdx.x];
			run[threadIdx.x] += t;
		}
		__syncthreads();
	}
	for (int s = threadIdx.x; s < count; s += BS)
	{
		spos[start + s] = tpos[start + s];
		smass[start + s] = tmass[start + s];
	}
	if (threadIdx.x < 4)
	{
		sstart[c0 + threadIdx.x] = start + base[threadIdx.x];
		scount[c0 + threadIdx.x] = qcnt[threadIdx.x];
	}
}

// tree_node.cuh:183-267 for every node of a new level
__global__ void k_set_neighbors(int off, int cnt, const double2* __restrict__ center, const double* __restrict__ hl,
	const int* __restrict__ parent, const int* __restrict__ child, const int* __restrict__ scount,
	int* nbr, int* nnbr, int* sl, int* nsl, int* dl, int* ndl, int* scal)
{
	int i = blockIdx.x * blockDim.x + threadIdx.x;
	if (i >= cnt) return;
	const int n = off + i;
	const int p = parent[n];
	const double h = hl[n];
	const double2 c = center[n];
	int a = 0, b = 0, d = 0;
	bool overflow = false;
	for (int q = 0; q < 4; q++)
	{
		int s = child[p] + q;
		if (s == n || scount[s] == 0) continue;
		if (a < MAX_NEIGHBORS) nbr[n * MAX_NEIGHBORS + a++] = s; else overflow = true;
	}
	const int pn = nnbr[p];
	for (int k = 0; k < pn; k++)
	{
		const int m = nbr[p * MAX_NEIGHBORS + k];
		if (scount[m] == 0) continue;
		if (child[m] < 0)
		{
			const double2 cm = center[m];
			if (fabs(cm.x - c.x) < 3.5 * h && fabs(cm.y - c.y) < 3.5 * h)
			{
				if (a < MAX_NEIGHBORS) nbr[n * MAX_NEIGHBORS + a++] = m; else overflow = true;
			}
			else
			{
				if (d < MAX_DIFF_LEVEL) dl[n * MAX_DIFF_LEVEL + d++] = m; else overflow = true;
			}
		}
		else
		{
			for (int q = 0; q < 4; q++)
			{
				const int s = child[m] + q;
				if (scount[s] == 0) continue;
				const double2 cs = center[s];
				if (fabs(cs.x - c.x) < 3 * h && fabs(cs.y - c.y) < 3 * h)
				{
					if (a < MAX_NEIGHBORS) nbr[n * MAX_NEIGHBORS + a++] = s; else overflow = true;
				}
				else
				{
					if (b < MAX_SAME_LEVEL) sl[n * MAX_SAME_LEVEL + b++] = s; else overflow = true;
				}
			}
		}
	}
	nnbr[n] = a;
	nsl[n] = b;
	ndl[n] = d;
	if (overflow) atomicOr(&scal[2], 1);
}

// star ranges summed directly at a node: its own stars then each neighbour's
__global__ void k_near_field(int nnodes, const int* __restrict__ sstart, const int* __restrict__ scount,
	const int* __restrict__ nbr, const int* __restrict__ nnbr, int2* nf, int* nnf)
{
	int n = blockIdx.x * blockDim.x + threadIdx.x;
	if (n >= nnodes) return;
	int k = 0;
	nf[n * MAX_NEAR_RANGES + k++] = make_int2(sstart[n], scount[n]);
	for (int j = 0; j < nnbr[n]; j++)
	{
		int m = nbr[n * MAX_NEIGHBORS + j];
		nf[n * MAX_NEAR_RANGES + k++] = make_int2(sstart[m], scount[m]);
	}
	nnf[n] = k;
}

// packed node box for the descent of tree_node.cuh:529-561 (one load per level)
__global__ void k_node_box(int nnodes, const double2* __restrict__ center, const double* __restrict__ hl,
	const int* __restrict__ child, double4* nbox)
{
	int n = blockIdx.x * blockDim.x + threadIdx.x;
	if (n >= nnodes) return;
	nbox[n] = make_double4(center[n].x, center[n].y, hl[n], (double)child[n]);
}

// point-location table: cell (ix, iy) of the 2^lv x 2^lv grid over the root
// box gets the node the descent (tree_node.cuh:529-561) reaches from the
// cell's centre after at most lv levels
constexpr int MAX_LOOKUP_LEVEL = 10;

__global__ void k_lookup(int lv, int root, double x0, double cell, const double4* __restrict__ nbox, int* lookup)
{
	const int n = 1 << lv;
	const long long c = (long long)blockIdx.x * blockDim.x + threadIdx.x;
	if (c >= (long long)n * n) return;
	const double px = x0 + ((int)(c % n) + 0.5) * cell, py = x0 + ((int)(c / n) + 0.5) * cell;
	int node = root;
	for (int l = 0; l < lv; l++)
	{
		const double4 b = nbox[node];
		if (b.w < 0) break;
		int q;
		if (px >= b.x && py >= b.y) q = 0;
		else if (px < b.x && py >= b.y) q = 1;
		else if (px < b.x && py < b.y) q = 2;
		else q = 3;
		node = (int)b.w + q;
	}
	lookup[c] = node;
}

// number of near-field stars of every leaf (0 for other nodes)
__global__ void k_near_count(int nnodes, const int* __restrict__ child, const int2* __restrict__ nf, const int* __restrict__ nnf,
	int* cnt)
{
	int n = blockIdx.x * blockDim.x + threadIdx.x;
	if (n >= nnodes) return;
	int c = 0;
	if (child[n] < 0)
		for (int k = 0; k < nnf[n]; k++) c += nf[n * MAX_NEAR_RANGES + k].y;
	cnt[n] = c;
}

// contiguous copy {x, y, m} of each leaf's near-field stars, in the order
// of its ranges (own stars, then each neighbour's)
__global__ void k_near_fill(int nnodes, const int* __restrict__ off, const int2* __restrict__ nf, const int* __restrict__ nnf,
	const double2* __restrict__ spos, const double* __restrict__ smass, double4* __restrict__ nfs)
{
	int n = blockIdx.x * blockDim.x + threadIdx.x;
	if (n >= nnodes || off[n] < 0) return;
	int t = off[n];
	for (int k = 0; k < nnf[n]; k++)
	{
		const int2 rg = nf[n * MAX_NEAR_RANGES + k];
#pragma unroll 4
		for (int s = rg.x; s < rg.x + rg.y; s++) nfs[t++] = make_double4(spos[s].x, spos[s].y, smass[s], 0.0);
	}
}

__device__ inline double bin(const double* __restrict__ B, int n, int k) { return n < k ? 0.0 : B[n * (n + 1) / 2 + k]; }
__device__ inline double pow2_divisor(int l) { return (double)(int)(1u << l); } // (1 << l) as the reference's int

// upward pass for one level: P2M at leaves with stars (fmm.cuh:21-49),
// M2M from the four children elsewhere (fmm.cuh:109-139). With NS = 2 the
// same pass also builds a second expansion set from the second masses
// (the unit masses of the pole sum): the geometry is shared, and each set
// takes exactly the operations it would take alone.
template <int NS>
__global__ __launch_bounds__(BLOCK) void k_upward(int off, int cnt, int p, const double2* __restrict__ center, const double* __restrict__ hl,
	const int* __restrict__ child, const int* __restrict__ sstart, const int* __restrict__ scount,
	const double2* __restrict__ spos, const double* __restrict__ smass, const double* __restrict__ smass2,
	const double* __restrict__ B, double2* mp, double2* mp2)
{
	const int idx = blockIdx.x * blockDim.x + threadIdx.x;
	const int i = idx / (p + 1), l = idx % (p + 1);
	if (i >= cnt) return;
	const int n = off + i;
	const int P1 = p + 1;
	const double* ms[2] = {smass, smass2};
	double2* out[2] = {mp, mp2};
	cd r[NS];
	if (child[n] < 0)
	{
		const int s0 = sstart[n], ns = scount[n];
		if (ns == 0) return;
		const cd c = ld(center, n);
		const double h = hl[n];
		if (l == 0)
		{
			for (int s = 0; s < ns; s++)
#pragma unroll
				for (int q = 0; q < NS; q++) r[q] += cd(ms[q][s0 + s]);
		}
		else
		{
			// the stars four at a time: their loads ahead of the terms, which
			// are added in star order
			for (int s4 = 0; s4 < ns; s4 += 4)
			{
				double2 zv[4];
				double mv[NS][4];
#pragma unroll
				for (int u = 0; u < 4; u++)
				{
					const int sv = s0 + min(s4 + u, ns - 1);
					zv[u] = spos[sv];
#pragma unroll
					for (int q = 0; q < NS; q++) mv[q][u] = ms[q][sv];
				}
#pragma unroll
				for (int u = 0; u < 4; u++)
					if (s4 + u < ns)
					{
						const cd t = (cd(zv[u].x, zv[u].y) - c) / h;
						const cd tl = t.ipow(l);
#pragma unroll
						for (int q = 0; q < NS; q++) r[q] -= tl * mv[q][u];
					}
			}
#pragma unroll
			for (int q = 0; q < NS; q++) r[q] /= (double)l;
		}
	}
	else
	{
		for (int c4 = 0; c4 < 4; c4++)
		{
			const int ch = child[n] + c4;
			if (l == 0)
			{
#pragma unroll
				for (int q = 0; q < NS; q++) r[q] += ld(out[q] + (size_t)ch * P1, 0);
				continue;
			}
			const cd dz = (ld(center, ch) - ld(center, n)) / hl[ch];
			const cd idz = dz.inv();
			const cd dzl = dz.ipow(l);
			const double dv = pow2_divisor(l);
			// Horner from k = l down, every mass set's chain at once, the
			// coefficients and binomials of four steps loaded ahead of them
			cd t[NS];
			for (int k0 = l; k0 >= 1; k0 -= 4)
			{
				cd av[NS][4];
				double bv[4];
#pragma unroll
				for (int u = 0; u < 4; u++)
				{
					const int k = max(k0 - u, 1);
#pragma unroll
					for (int q = 0; q < NS; q++) av[q][u] = ld(out[q] + (size_t)ch * P1, k);
					bv[u] = bin(B, l - 1, k - 1);
				}
#pragma unroll
				for (int u = 0; u < 4; u++)
					if (k0 - u >= 1)
#pragma unroll
						for (int q = 0; q < NS; q++)
						{
							t[q] += av[q][u] * bv[u];
							t[q] *= idz;
						}
			}
#pragma unroll
			for (int q = 0; q < NS; q++)
			{
				t[q] -= ld(out[q] + (size_t)ch * P1, 0) / (double)l;
				t[q] *= dzl;
				t[q] /= dv;
				r[q] += t[q];
			}
		}
	}
#pragma unroll
	for (int q = 0; q < NS; q++) st(out[q], (size_t)n * P1 + l, r[q]);
}

// downward pass for one level >= 2: L2L from the parent (fmm.cuh:204-223),
// M2L from the same-level list (fmm.cuh:281-322), P2L from the coarser
// leaves (fmm.cuh:387-413); NS = 2 as in k_upward. A node takes DOWN_LANES
// lanes of a wave, lane l forming coefficient l (l <= p <= 31). The M2L sum
// of a source, sum_{k=1..p} C(l+k-1, k-1) b_k (-dz)^-k, is formed from
// u_k = b_k (-dz)^-k, which lane k computes once for every l and leaves in
// LDS; lane l then adds its binomial-weighted u_k from k = p down to 1.
constexpr int DOWN_LANES = 32;
constexpr int DOWN_NODES = BLOCK / DOWN_LANES;

// a node's coefficients, and its u_k (k <= MAX_EXPANSION_ORDER), fit its lanes
static_assert(MAX_EXPANSION_ORDER < DOWN_LANES, "a node's coefficients must fit its lanes");
static_assert(MAX_SAME_LEVEL <= DOWN_LANES, "a node's same-level list must fit its lanes");

template <int NS>
__global__ __launch_bounds__(BLOCK) void k_downward(int off, int cnt, int p, const double2* __restrict__ center, const double* __restrict__ hl,
	const int* __restrict__ parent, const int* __restrict__ sl, const int* __restrict__ nsl,
	const int* __restrict__ dl, const int* __restrict__ ndl, const int* __restrict__ sstart, const int* __restrict__ scount,
	const double2* __restrict__ spos, const double* __restrict__ smass, const double* __restrict__ smass2,
	const double* __restrict__ B, const double2* __restrict__ mp, const double2* __restrict__ mp2, double2* loc, double2* loc2)
{
	// the binomial rows 0..2p in LDS (dynamic, launched with their size);
	// the Horner loops below take their coefficients and binomials four at a
	// time ahead of the dependent products, in the same order of operations
	extern __shared__ double sB[];
	__shared__ double2 su[DOWN_NODES][NS][DOWN_LANES];
	const int nB = (2 * p + 1) * (2 * p + 2) / 2;
	for (int q = threadIdx.x; q < nB; q += blockDim.x) sB[q] = B[q];
	__syncthreads();
	const int slot = threadIdx.x / DOWN_LANES, l = threadIdx.x % DOWN_LANES;
	const int i = blockIdx.x * DOWN_NODES + slot;
	if (i >= cnt) return;
	const int n = off + i;
	const int P1 = p + 1;
	const double* ms[2] = {smass, smass2};
	const double2* mps[2] = {mp, mp2};
	double2* locs[2] = {loc, loc2};
	const cd cn = ld(center, n);
	const double hn = hl[n];
	cd r[NS];
	{
		const int pa = parent[n];
		const cd dz = (cn - ld(center, pa)) / hl[pa];
		for (int k0 = p; k0 >= 1; k0 -= 4)
		{
			cd av[NS][4];
			double bv[4];
#pragma unroll
			for (int u = 0; u < 4; u++)
			{
				const int k = max(k0 - u, 1);
#pragma unroll
				for (int q = 0; q < NS; q++) av[q][u] = ld(locs[q] + (size_t)pa * P1, k);
				bv[u] = bin(sB, k, l);
			}
#pragma unroll
			for (int u = 0; u < 4; u++)
				if (k0 - u >= 1)
				{
#pragma unroll
					for (int q = 0; q < NS; q++)
					{
						r[q] += av[q][u] * bv[u];
						r[q] *= dz;
					}
				}
		}
		const cd w = dz.inv().ipow(l);
		const double dv = pow2_divisor(l);
#pragma unroll
		for (int q = 0; q < NS; q++)
		{
			r[q] *= w;
			r[q] /= dv;
		}
	}
	// the node's same-level list, entry j in lane j of its DOWN_LANES
	const int nsl_n = nsl[n];
	const int my_src = l < nsl_n ? sl[n * MAX_SAME_LEVEL + l] : 0;
	for (int j = 0; j < nsl_n; j++)
	{
		const int m = __shfl(my_src, j, DOWN_LANES);
		const cd dz = (ld(center, m) - cn) / hl[m];
		const cd imdz = (-dz).inv();
		// lane k (1 <= k <= p): u_k of every mass set
		{
			cd pw;
			if (l >= 1 && l <= p) pw = imdz.ipow(l);
#pragma unroll
			for (int q = 0; q < NS; q++)
			{
				const cd u = l >= 1 && l <= p ? ld(mps[q] + (size_t)m * P1, l) * pw : cd();
				su[slot][q][l] = make_double2(u.re, u.im);
			}
		}
		// the node's lanes are one wave's: its LDS writes above are seen by
		// its reads below once the wave passes this point
		__builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
		__builtin_amdgcn_wave_barrier();
		__builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
		cd t[NS];
		if (l <= p)
		{
			for (int k = p; k >= 1; k--)
			{
				const double c = l == 0 ? 1.0 : bin(sB, l + k - 1, k - 1);
#pragma unroll
				for (int q = 0; q < NS; q++)
				{
					const double2 u = su[slot][q][k];
					t[q] += cd(u.x, u.y) * c;
				}
			}
			if (l == 0)
			{
				const cd lg = (-dz).log();
				const double lh = log(hl[m]);
#pragma unroll
				for (int q = 0; q < NS; q++)
				{
					const cd b0 = ld(mps[q] + (size_t)m * P1, 0);
					t[q] += b0 * lg;
					t[q] += b0 * lh;
				}
			}
			else
			{
				const cd w = dz.inv().ipow(l);
#pragma unroll
				for (int q = 0; q < NS; q++)
				{
					t[q] -= ld(mps[q] + (size_t)m * P1, 0) / (double)l;
					t[q] *= w;
				}
			}
#pragma unroll
			for (int q = 0; q < NS; q++) r[q] += t[q];
		}
		// every lane's reads of this source's u_k are done before the next
		// source's writes
		__builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
		__builtin_amdgcn_wave_barrier();
		__builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
	}
	for (int j = 0; j < ndl[n]; j++)
	{
		const int m = dl[n * MAX_DIFF_LEVEL + j];
		const int s0 = sstart[m], ns = scount[m];
		// the stars four at a time: their loads ahead of the terms, which
		// are added in star order
		for (int s4 = 0; s4 < ns; s4 += 4)
		{
			double2 zv[4];
			double mv[NS][4];
#pragma unroll
			for (int u = 0; u < 4; u++)
			{
				const int sv = s0 + min(s4 + u, ns - 1);
				zv[u] = spos[sv];
#pragma unroll
				for (int q = 0; q < NS; q++) mv[q][u] = ms[q][sv];
			}
#pragma unroll
			for (int u = 0; u < 4; u++)
			{
				if (s4 + u >= ns) break;
				const cd dz = (cd(zv[u].x, zv[u].y) - cn) / hn;
				if (l == 0)
				{
					const cd lg = (-dz).log();
					const double lh = log(hn);
#pragma unroll
					for (int q = 0; q < NS; q++)
					{
						const double mq = mv[q][u];
						r[q] += mq * lg;
						r[q] += cd(mq * lh);
					}
				}
				else
				{
					const cd w = dz.inv().ipow(l);
#pragma unroll
					for (int q = 0; q < NS; q++) r[q] -= (mv[q][u] / (double)l) * w;
				}
			}
		}
	}
	if (l > p) return; // (lanes beyond the order only formed u_k)
#pragma unroll
	for (int q = 0; q < NS; q++) st(locs[q], (size_t)n * P1 + l, r[q]);
}

__global__ void k_fill_ones(double* __restrict__ p, int n)
{
	const int i = blockIdx.x * blockDim.x + threadIdx.x;
	if (i < n) p[i] = 1.0;
}

__global__ void k_to_float(const double2* __restrict__ a, float2* b, size_t n)
{
	size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
	if (i < n) b[i] = make_float2((float)a[i].x, (float)a[i].y);
}

inline int grid(long long n, int b = BLOCK) { return (int)((n + b - 1) / b); }

template <typename V>
bool grow(ccf_ctx* ctx, V*& ptr, int old_n, int new_n, int width)
{
	V* p = nullptr;
	if (!CCF_HIP(ctx, ctx->dmalloc(&p, sizeof(V) * (size_t)new_n * width))) return false;
	if (ptr)
	{
		// the old buffer goes back to the context's pool, whose next user
		// queues on the same stream after this copy
		if (old_n > 0 && !CCF_HIP(ctx, hipMemcpyAsync(p, ptr, sizeof(V) * (size_t)old_n * width, hipMemcpyDeviceToDevice, ctx->stream))) return false;
		ctx->dfree(ptr);
	}
	ptr = p;
	return true;
}

bool ensure_capacity(ccf_ctx* ctx, int need)
{
	DeviceTree& T = ctx->tree;
	if (need <= T.capacity) return true;
	int cap = std::max(need, 2 * T.capacity);
	const int n = T.nnodes;
	bool ok = grow(ctx, T.center, n, cap, 1) && grow(ctx, T.hl, n, cap, 1) && grow(ctx, T.parent, n, cap, 1) &&
		grow(ctx, T.child, n, cap, 1) && grow(ctx, T.sstart, n, cap, 1) && grow(ctx, T.scount, n, cap, 1) &&
		grow(ctx, T.nbr, n, cap, MAX_NEIGHBORS) && grow(ctx, T.nnbr, n, cap, 1) &&
		grow(ctx, T.sl, n, cap, MAX_SAME_LEVEL) && grow(ctx, T.nsl, n, cap, 1) &&
		grow(ctx, T.dl, n, cap, MAX_DIFF_LEVEL) && grow(ctx, T.ndl, n, cap, 1) && grow(ctx, T.field, n, cap, 1) &&
		// split flags and ranks of the level being split are still live
		grow(ctx, T.split, T.capacity, cap, 1) && grow(ctx, T.rank, T.capacity, cap, 1);
	if (ok) T.capacity = cap;
	return ok;
}

// Pascal rows 0..n as the reference tabulates them (binomial_coefficients.cuh:11-44):
// int rows that wrap past row 33, or exact values
// LDS bytes of k_downward: the binomial rows 0..2p
inline size_t binom_lds(int p) { return (size_t)(2 * p + 1) * (2 * p + 2) / 2 * sizeof(double); }

std::vector<double> binomial_table(int n, bool exact)
{
	std::vector<double> t((size_t)(n + 1) * (n + 2) / 2);
	std::vector<uint32_t> w(t.size());
	std::vector<long double> e(t.size());
	for (int r = 0; r <= n; r++)
	{
		const int s = r * (r + 1) / 2, ps = (r - 1) * r / 2;
		for (int k = 0; k <= r; k++)
		{
			if (k == 0 || k == r) { w[s + k] = 1; e[s + k] = 1; }
			else { w[s + k] = w[ps + k - 1] + w[ps + k]; e[s + k] = e[ps + k - 1] + e[ps + k]; }
			t[s + k] = exact ? (double)e[s + k] : (double)(int32_t)w[s + k];
		}
	}
	return t;
}

} // namespace

bool build_tree(ccf_ctx* ctx, int verbose)
{
	ctx->root_half_length = (ctx->rectangular ? std::max(ctx->corner.re, ctx->corner.im) : ctx->corner.abs()) * 1.1;
	if (verbose >= 2) std::cout << "root_half_length set to: " << ctx->root_half_length << "\n\n";
	ForestSpec f;
	f.soff = {0};
	f.nstars = {ctx->num_stars};
	f.root_hl = {ctx->root_half_length};
	f.theta = {ctx->theta};
	f.alpha_error = {ctx->alpha_error};
	if (!build_forest(ctx, f, verbose)) return false;
	ctx->expansion_order = f.order;
	ctx->info.tree_levels = ctx->tree.levels;
	ctx->info.num_nodes = ctx->tree.nnodes;
	ctx->info.expansion_order = ctx->tree.order;
	return true;
}

// The trees of spec.nstars.size() independent fields built together: level 0
// holds one root per field (centre 0, half-length root_hl[f], the field's
// stars [soff[f], soff[f] + nstars[f]) of ctx->spos), every level splits the
// nodes that the reference's rule splits, and since siblings and neighbours
// never leave a parent's subtree the fields stay apart. One expansion order
// serves all fields: fmm_order if set, else the largest the reference's rule
// gives over the fields (the rule of ccf.cuh:840-852 itself for one field).
bool build_forest(ccf_ctx* ctx, ForestSpec& spec, int verbose)
{
	DeviceTree& T = ctx->tree;
	const ccf_params& P = ctx->P;
	const int F = (int)spec.nstars.size();
	int N = 0;
	for (int f = 0; f < F; f++) N = std::max(N, spec.soff[f] + spec.nstars[f]);
	hipStream_t s = ctx->stream;
	auto t0 = std::chrono::steady_clock::now();
	if (verbose >= 1) std::cout << "Creating children and sorting stars...\n";

	T.nnodes = 0;
	T.capacity = 0;
	// the node count of the last tree of the same star and field counts, if
	// any, so that a repeated build does not grow its arrays level by level
	int cap0 = 4096 + N / 2 + 4 * F;
	if (ctx->tree_hint_stars == N && ctx->tree_hint_fields == F) cap0 = std::max(cap0, ctx->tree_hint_nodes);
	if (!ensure_capacity(ctx, cap0)) return false;
	int* scal = nullptr;
	if (!CCF_HIP(ctx, ctx->dmalloc(&scal, 4 * sizeof(int)))) return false;

	// root nodes: one field's by a kernel (no host copies to wait for),
	// several fields' from host arrays
	bool roots_set = false;
	if (F == 1)
	{
		k_root_node<<<1, 1, 0, s>>>(spec.root_hl[0], spec.soff[0], spec.nstars[0], T.center, T.hl, T.parent, T.child, T.sstart,
			T.scount, T.nnbr, T.nsl, T.ndl, T.field, scal);
		if (!CCF_LAUNCH(ctx, "tree root node")) { ctx->dfree(scal); return false; }
		roots_set = true;
	}
	if (!roots_set)
	{
		std::vector<double2> c(F, make_double2(0, 0));
		std::vector<int> m1(F, -1), z(F, 0), fid(F);
		for (int f = 0; f < F; f++) fid[f] = f;
		bool ok = CCF_HIP(ctx, hipMemcpyAsync(T.center, c.data(), F * sizeof(double2), hipMemcpyHostToDevice, s)) &&
			CCF_HIP(ctx, hipMemcpyAsync(T.hl, spec.root_hl.data(), F * sizeof(double), hipMemcpyHostToDevice, s)) &&
			CCF_HIP(ctx, hipMemcpyAsync(T.parent, m1.data(), F * sizeof(int), hipMemcpyHostToDevice, s)) &&
			CCF_HIP(ctx, hipMemcpyAsync(T.child, m1.data(), F * sizeof(int), hipMemcpyHostToDevice, s)) &&
			CCF_HIP(ctx, hipMemcpyAsync(T.sstart, spec.soff.data(), F * sizeof(int), hipMemcpyHostToDevice, s)) &&
			CCF_HIP(ctx, hipMemcpyAsync(T.scount, spec.nstars.data(), F * sizeof(int), hipMemcpyHostToDevice, s)) &&
			CCF_HIP(ctx, hipMemcpyAsync(T.nnbr, z.data(), F * sizeof(int), hipMemcpyHostToDevice, s)) &&
			CCF_HIP(ctx, hipMemcpyAsync(T.nsl, z.data(), F * sizeof(int), hipMemcpyHostToDevice, s)) &&
			CCF_HIP(ctx, hipMemcpyAsync(T.ndl, z.data(), F * sizeof(int), hipMemcpyHostToDevice, s)) &&
			CCF_HIP(ctx, hipMemcpyAsync(T.field, fid.data(), F * sizeof(int), hipMemcpyHostToDevice, s)) &&
			CCF_HIP(ctx, hipStreamSynchronize(s));
		if (!ok) { ctx->dfree(scal); return false; }
	}
	T.nnodes = F;
	T.level_off = {0};
	T.level_cnt = {F};
	T.levels = 0;

	// scal[2], the list overflow flag, stays set once a level's (capped)
	// lists overflow: it is read with the next level's counts (the loop
	// ends at such a read-back), one read-back per level
	bool ok = roots_set || CCF_HIP(ctx, hipMemsetAsync(scal, 0, 4 * sizeof(int), s));
	int h_scal[4] = {0, 0, 0, 0};
	// the level counts come back through the pinned staging buffer
	// (STAGE_DOUBLES doubles, until ccf_destroy)
	if (ok && !ctx->host_stage)
		ok = CCF_HIP(ctx, hipHostMalloc((void**)&ctx->host_stage, sizeof(double) * STAGE_DOUBLES, hipHostMallocDefault));
	int* const hs = ok ? reinterpret_cast<int*>(ctx->host_stage) : nullptr;
	while (ok && !P.direct)
	{
		const int off = T.level_off[T.levels], cnt = T.level_cnt[T.levels];
		// scal[0] was cleared at level 0's set-up or by the previous level's
		// k_make_children; k_scan writes scal[1]
		k_level_count<<<grid(cnt), BLOCK, 0, s>>>(off, cnt, T.scount, T.nbr, T.nnbr, T.split, scal);
		k_scan<<<1, 1024, 0, s>>>(T.split, cnt, T.rank, scal + 1);
		ok = CCF_LAUNCH(ctx, "tree level count") && CCF_HIP(ctx, hipMemcpyAsync(hs, scal, 4 * sizeof(int), hipMemcpyDeviceToHost, s)) &&
			CCF_HIP(ctx, hipStreamSynchronize(s));
		if (!ok) break;
		for (int k = 0; k < 4; k++) h_scal[k] = hs[k];
		if (h_scal[2]) break;
		if (verbose >= 3)
		{
			std::cout << "\nProcessing level " << T.levels << "\n";
			std::cout << "Maximum number of stars in a node and its neighbors is " << h_scal[0] << "\n";
		}
		if (h_scal[0] <= MAX_NUM_STARS_DIRECT) break;
		const int nsplit = h_scal[1];
		const int new_off = off + cnt, new_cnt = 4 * nsplit;
		if (verbose >= 3) std::cout << "Number of non-empty children: " << new_cnt << "\n";
		if (!(ok = ensure_capacity(ctx, new_off + new_cnt))) break;
		k_make_children<<<grid(cnt), BLOCK, 0, s>>>(off, cnt, new_off, T.split, T.rank, T.center, T.hl, T.parent, T.child,
			T.sstart, T.scount, T.nnbr, T.nsl, T.ndl, T.field, scal);
		// (levels of few nodes, each with many stars: wider workgroups)
		if (cnt <= PARTITION_WIDE_NODES)
			k_partition<1024><<<cnt, 1024, 0, s>>>(off, T.split, T.child, T.center, T.sstart, T.scount, ctx->spos, ctx->smass,
				ctx->spos_tmp, ctx->smass_tmp);
		else
			k_partition<BLOCK><<<cnt, BLOCK, 0, s>>>(off, T.split, T.child, T.center, T.sstart, T.scount, ctx->spos, ctx->smass,
				ctx->spos_tmp, ctx->smass_tmp);
		k_set_neighbors<<<grid(new_cnt), BLOCK, 0, s>>>(new_off, new_cnt, T.center, T.hl, T.parent, T.child, T.scount,
			T.nbr, T.nnbr, T.sl, T.nsl, T.dl, T.ndl, scal);
		ok = CCF_LAUNCH(ctx, "tree split level");
		if (!ok) break;
		T.nnodes = new_off + new_cnt;
		T.level_off.push_back(new_off);
		T.level_cnt.push_back(new_cnt);
		T.levels++;
	}
	if (ok && h_scal[2])
	{
		ctx->fail("Error. A node's neighbour or interaction list exceeded its maximum size.\n");
		ok = false;
	}
	ctx->dfree(scal);
	if (!ok) return false;
	if (verbose >= 2) std::cout << "tree_levels set to: " << T.levels << "\n\n";
	ctx->tree_hint_stars = N;
	ctx->tree_hint_fields = F;
	ctx->tree_hint_nodes = T.nnodes;

	auto t1 = std::chrono::steady_clock::now();
	if (verbose >= 1)
		std::cout << "Done creating children and sorting stars. Elapsed time: " << std::chrono::duration<double>(t1 - t0).count() << " seconds.\n\n";

	// near-field ranges of every node
	if (!CCF_HIP(ctx, ctx->dmalloc(&T.nf, sizeof(int2) * (size_t)T.nnodes * MAX_NEAR_RANGES))) return false;
	if (!CCF_HIP(ctx, ctx->dmalloc(&T.nnf, sizeof(int) * (size_t)T.nnodes))) return false;
	k_near_field<<<grid(T.nnodes), BLOCK, 0, s>>>(T.nnodes, T.sstart, T.scount, T.nbr, T.nnbr, T.nf, T.nnf);
	if (!CCF_LAUNCH(ctx, "near field ranges")) return false;
	// packed boxes and contiguous near-field copies of the leaves
	{
		const int nn = T.nnodes;
		if (!CCF_HIP(ctx, ctx->dmalloc(&T.nbox, sizeof(double4) * nn)) || !CCF_HIP(ctx, ctx->dmalloc(&T.nf_off, sizeof(int) * nn)) ||
			!CCF_HIP(ctx, ctx->dmalloc(&T.nf_cnt, sizeof(int) * nn)))
			return false;
		k_node_box<<<grid(nn), BLOCK, 0, s>>>(nn, T.center, T.hl, T.child, T.nbox);
		// point-location table over the root box (centre 0, half-length root_half_length)
		ctx->dfree(T.lookup);
		T.lookup = nullptr;
		T.lookup_level = 0;
		if (F == 1 && T.levels >= 2)
		{
			T.lookup_level = std::min(T.levels, MAX_LOOKUP_LEVEL);
			const int n = 1 << T.lookup_level;
			const double H = spec.root_hl[0];
			T.lookup_x0 = -H;
			T.lookup_inv = n / (2 * H);
			if (!CCF_HIP(ctx, ctx->dmalloc(&T.lookup, sizeof(int) * (size_t)n * n))) return false;
			k_lookup<<<grid((long long)n * n), BLOCK, 0, s>>>(T.lookup_level, 0, -H, 2 * H / n, T.nbox, T.lookup);
		}
		k_near_count<<<grid(nn), BLOCK, 0, s>>>(nn, T.child, T.nf, T.nnf, T.nf_cnt);
		// the leaves' runs in node order: an exclusive prefix of the counts
		// (non-leaves count 0) on the device, the total read back to size
		// the copies (T.rank and T.split are the level loop's scratch, of
		// the tree's capacity)
		k_scan<<<1, 1024, 0, s>>>(T.nf_cnt, nn, T.rank, T.split);
		k_leaf_off<<<grid(nn), BLOCK, 0, s>>>(nn, T.child, T.rank, T.nf_off);
		int* const hst = reinterpret_cast<int*>(ctx->host_stage); // (pinned, allocated before the levels)
		if (!CCF_LAUNCH(ctx, "node boxes") || !CCF_HIP(ctx, hipMemcpyAsync(hst, T.split, sizeof(int), hipMemcpyDeviceToHost, s)) ||
			!CCF_HIP(ctx, hipStreamSynchronize(s)))
			return false;
		const int tot = hst[0];
		if (tot < 0)
		{
			ctx->fail("Error. The near-field copies exceed 2^31 stars.\n");
			return false;
		}
		if (!CCF_HIP(ctx, ctx->dmalloc(&T.nfs, sizeof(double4) * (size_t)std::max(tot, 1)))) return false;
		k_near_fill<<<grid(nn), BLOCK, 0, s>>>(nn, T.nf_off, T.nf, T.nnf, ctx->spos, ctx->smass, T.nfs);
		if (!CCF_LAUNCH(ctx, "near-field copies")) return false;
	}

	// depth of every field's tree
	spec.levels.assign(F, 0);
	if (F == 1)
		spec.levels[0] = T.levels;
	else
	{
		std::vector<int> fid(T.nnodes);
		if (!CCF_HIP(ctx, hipMemcpy(fid.data(), T.field, sizeof(int) * T.nnodes, hipMemcpyDeviceToHost))) return false;
		for (int l = 0; l <= T.levels; l++)
			for (int n = T.level_off[l]; n < T.level_off[l] + T.level_cnt[l]; n++) spec.levels[fid[n]] = l;
	}
	// a forest's point-location tables: one per field over its own root box
	// (node f at level 0, centre 0), at the field's depth, back to back
	spec.lookup_off.assign(F, 0);
	spec.lookup_level.assign(F, 0);
	if (F > 1)
	{
		size_t tot = 0;
		for (int f = 0; f < F; f++)
		{
			const int lv = std::min(spec.levels[f], MAX_LOOKUP_LEVEL);
			if (lv < 2) continue;
			spec.lookup_level[f] = lv;
			spec.lookup_off[f] = tot;
			tot += (size_t)1 << (2 * lv);
		}
		if (tot)
		{
			if (!CCF_HIP(ctx, ctx->dmalloc(&T.lookup, sizeof(int) * tot))) return false;
			for (int f = 0; f < F; f++)
			{
				const int lv = spec.lookup_level[f];
				if (!lv) continue;
				const int n = 1 << lv;
				const double H = spec.root_hl[f];
				k_lookup<<<grid((long long)n * n), BLOCK, 0, s>>>(lv, f, -H, 2 * H / n, T.nbox, T.lookup + spec.lookup_off[f]);
			}
			if (!CCF_LAUNCH(ctx, "forest point-location tables")) return false;
		}
	}
	// expansion order (ccf.cuh:840-852)
	int p = 0;
	if (P.fmm_order > 0) p = P.fmm_order;
	else
		for (int f = 0; f < F; f++)
			p = std::max(p, (int)std::ceil(2 * std::log2(spec.theta[f]) + spec.levels[f] - std::log2(spec.root_hl[f]) -
				std::log2(spec.alpha_error[f])));
	if (verbose >= 2) std::cout << "expansion_order set to: " << p << "\n\n";
	if (!P.direct)
	{
		if (p < 3) { ctx->fail("Error. Expansion order needs to be >= 3\n"); return false; }
		if (p > MAX_EXPANSION_ORDER)
		{
			ctx->fail("Error. Maximum allowed expansion order is " + std::to_string(MAX_EXPANSION_ORDER) + "\n");
			return false;
		}
	}
	spec.order = p;
	T.order = P.direct ? 0 : p;

	// FMM: upward from the deepest level, downward from level 2 (ccf.cuh:866-893)
	if (verbose >= 1) std::cout << "Calculating multipole and local coefficients...\n";
	const int P1 = T.order + 1;
	if (!CCF_HIP(ctx, ctx->dmalloc(&T.mp, sizeof(double2) * (size_t)T.nnodes * P1))) return false;
	if (!CCF_HIP(ctx, ctx->dmalloc(&T.loc, sizeof(double2) * (size_t)T.nnodes * P1))) return false;
	if (!CCF_HIP(ctx, ctx->dmalloc(&T.loc32, sizeof(float2) * (size_t)T.nnodes * P1))) return false;
	if (!CCF_HIP(ctx, hipMemsetAsync(T.mp, 0, sizeof(double2) * (size_t)T.nnodes * P1, s))) return false;
	if (!CCF_HIP(ctx, hipMemsetAsync(T.loc, 0, sizeof(double2) * (size_t)T.nnodes * P1, s))) return false;
	// the mixed form also needs the expansions of the unit masses: local
	// expansions of sum log(z - s_i), whose first derivative is the far part
	// of the Aberth pole sum; both sets come from the same passes
	const bool dual = P.precision == 1 && T.order > 0;
	if (dual)
	{
		bool ok2 = CCF_HIP(ctx, ctx->dmalloc(&ctx->ones, sizeof(double) * N));
		if (ok2)
		{
			k_fill_ones<<<grid(N), BLOCK, 0, s>>>(ctx->ones, N);
			ok2 = CCF_LAUNCH(ctx, "unit masses");
		}
		ok2 = ok2 && CCF_HIP(ctx, ctx->dmalloc(&T.mpu, sizeof(double2) * (size_t)T.nnodes * P1)) &&
			CCF_HIP(ctx, ctx->dmalloc(&T.locu, sizeof(double2) * (size_t)T.nnodes * P1)) &&
			CCF_HIP(ctx, ctx->dmalloc(&T.locu32, sizeof(float2) * (size_t)T.nnodes * P1)) &&
			CCF_HIP(ctx, hipMemsetAsync(T.mpu, 0, sizeof(double2) * (size_t)T.nnodes * P1, s)) &&
			CCF_HIP(ctx, hipMemsetAsync(T.locu, 0, sizeof(double2) * (size_t)T.nnodes * P1, s));
		if (!ok2) return false;
	}
	if (T.order > 0)
	{
		// through pinned memory, so the copy does not hold the host (the
		// table of a previous build was copied before that build's end)
		std::vector<double> B = binomial_table(2 * p, P.binomial_exact != 0);
		std::copy(B.begin(), B.end(), ctx->host_stage);
		if (!CCF_HIP(ctx, ctx->dmalloc(&T.binom, sizeof(double) * B.size()))) return false;
		if (!CCF_HIP(ctx, hipMemcpyAsync(T.binom, ctx->host_stage, sizeof(double) * B.size(), hipMemcpyHostToDevice, s))) return false;
		for (int l = T.levels; l >= 0; l--)
		{
			const long long work = (long long)T.level_cnt[l] * P1;
			if (dual)
				k_upward<2><<<grid(work), BLOCK, 0, s>>>(T.level_off[l], T.level_cnt[l], p, T.center, T.hl, T.child, T.sstart,
					T.scount, ctx->spos, ctx->smass, ctx->ones, T.binom, T.mp, T.mpu);
			else
				k_upward<1><<<grid(work), BLOCK, 0, s>>>(T.level_off[l], T.level_cnt[l], p, T.center, T.hl, T.child, T.sstart,
					T.scount, ctx->spos, ctx->smass, nullptr, T.binom, T.mp, nullptr);
		}
		for (int l = 2; l <= T.levels; l++)
		{
			const int blocks = grid(T.level_cnt[l], DOWN_NODES);
			if (dual)
				k_downward<2><<<blocks, BLOCK, binom_lds(p), s>>>(T.level_off[l], T.level_cnt[l], p, T.center, T.hl, T.parent,
					T.sl, T.nsl, T.dl, T.ndl, T.sstart, T.scount, ctx->spos, ctx->smass, ctx->ones, T.binom, T.mp, T.mpu, T.loc, T.locu);
			else
				k_downward<1><<<blocks, BLOCK, binom_lds(p), s>>>(T.level_off[l], T.level_cnt[l], p, T.center, T.hl, T.parent,
					T.sl, T.nsl, T.dl, T.ndl, T.sstart, T.scount, ctx->spos, ctx->smass, nullptr, T.binom, T.mp, nullptr, T.loc, nullptr);
		}
		if (!CCF_LAUNCH(ctx, "calculate_coeffs_kernels")) return false;
	}
	k_to_float<<<grid((long long)T.nnodes * P1), BLOCK, 0, s>>>(T.loc, T.loc32, (size_t)T.nnodes * P1);
	if (!CCF_LAUNCH(ctx, "local coefficients to float")) return false;
	if (dual)
	{
		k_to_float<<<grid((long long)T.nnodes * P1), BLOCK, 0, s>>>(T.locu, T.locu32, (size_t)T.nnodes * P1);
		if (!CCF_LAUNCH(ctx, "unit-mass expansions")) return false;
	}
	// FP32 star positions for the mixed-precision Cauchy sums
	k_to_float<<<grid(N), BLOCK, 0, s>>>(ctx->spos, ctx->spos32, (size_t)N);
	if (!CCF_LAUNCH(ctx, "star positions to float")) return false;
	if (!CCF_HIP(ctx, hipStreamSynchronize(s))) return false;
	auto t2 = std::chrono::steady_clock::now();
	if (verbose >= 1)
		std::cout << "Done calculating multipole and local coefficients. Elapsed time: " << std::chrono::duration<double>(t2 - t1).count() << " seconds.\n\n";
	ctx->info.t_tree = std::chrono::duration<double>(t2 - t0).count();
	return true;
}

} // namespace ccf
