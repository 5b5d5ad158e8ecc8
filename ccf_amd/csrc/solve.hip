// Aberth-Ehrlich critical-curve solve, phi continuation, 1/mu error bound,
// transpose and caustic map on gfx950.
//
// weisluke/ccf runs a (root, side, branch) grid of single-thread updates that
// write the roots in place (ccf_functions.cuh:321-394) and launches 20 of
// them per phi step whatever the convergence state (ccf.cuh:1013-1018).
// Here a "row" is one phi value of one branch side and a sweep is one Jacobi
// update of every still-active root of every row of a phase, in three
// launches:
//   K1 k_witt      per active root: Witt's function f0, its derivative f1
//                  (near-field P2P + far-field L2P + smooth sheet) and the
//                  1/mu early exit; roots that need the Cauchy sums are
//                  appended to a per-row list (the few roots outside the
//                  tree by their whole wave); its last workgroup runs
//   K2 plan_sums   cuts the (stars + roots) term range of every tile of
//                  RPT*256 listed roots into S slices so that dense and sparse
//                  sweeps alike give enough equal work items to fill the chip;
//   K3 k_sums      one work item = one tile x one slice: the slice's terms are
//                  streamed through LDS and every lane accumulates
//                  2 sum 1/(z - s_i) - sum 1/(z - z_j) for its RPT roots;
//                  partials go to a buffer so the result is deterministic;
//   K4 k_finalize  reduces the partials, forms z - f0 / (f1 + f0 S), tests
//                  |dz| / theta < 1e-9, writes the root and appends the
//                  unconverged ones to the next sweep's active list.
// K3 exists in FP64 (the reference's arithmetic) and in FP32 with a per-chunk
// fallback: a lane whose nearest root in an LDS chunk is closer than
// CLUSTER_REL times the field's half-length (where an FP32 difference of two
// positions keeps fewer than ~12 bits) redoes that chunk in FP64.
#include "lens.hpp"

#include <hiprand/hiprand_kernel.h>

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <functional>
#include <iostream>
#include <mutex>
#include <string>
#include <vector>

namespace ccf
{

// device statistics (ccf_ctx::counters)
enum Counter
{
	C_FULL = 0,    // updates that formed the Cauchy sums
	C_EARLY = 1,   // updates that returned on the 1/mu bound
	C_NEAR = 2,    // near-field stars summed by updates that formed f1 (full and Newton)
	C_CONV = 3,    // roots converged
	C_ERRBITS = 4, // max 1/mu error (bits of a non-negative double)
	C_NAN = 5,     // non-finite error seen
	C_CLUSTER = 6, // FP32 chunks redone in FP64
	C_RNEAR = 7,   // root-sum FMM: near-field terms
	C_RM2L = 8,    // root-sum FMM: M2L translations
	C_RROWS = 9,   // root-sum FMM: (row, sweep) pairs summed by the FMM
	C_RTGT = 10,   // root-sum FMM: updates whose root sums came from the FMM
	C_RESID = 11,  // max |f0| over the solved points (bits of a non-negative double)
	C_NEWTON = 12, // mixed form: roots ended by a Newton step below the convergence threshold
	C_NEAR_EARLY = 13, // near-field stars summed by updates that returned on the 1/mu bound
	C_TAIL_FULL = 14,  // k_tail: Aberth updates (not in C_FULL)
	C_TAIL_EARLY = 15, // k_tail: 1/mu exits (not in C_EARLY)
	C_TAIL_NEWTON = 16, // k_tail: Newton steps (not in C_NEWTON)
	C_TAIL_NEAR = 17,  // k_tail: near-field stars of its evaluations
	C_TAIL_SWEEPS = 18, // k_tail: sweeps, summed over rows
	C_OUTSIDE = 19,     // error pass: points outside the tree (their near field is every star)
	C_COUNT = 24
};

namespace
{

constexpr int BLOCK = 256;
constexpr int CHUNK64 = 1024;         // FP64 complex values per LDS chunk (16 KiB)
constexpr int CHUNK32 = 2048;         // FP32 complex values per LDS chunk (16 KiB)
constexpr int MAX_SLICES = 64;
constexpr int MIN_SLICE = 1024;       // fewest terms per slice

inline int grid(long long n, int b = BLOCK) { return (int)((n + b - 1) / b); }

__device__ inline unsigned long long wave_sum(unsigned long long v)
{
	for (int d = 32; d >= 1; d >>= 1) v += __shfl_xor(v, d, 64);
	return v;
}

// counters[key] += 1 for every lane with key >= 0, one atomic per run of
// equal keys in consecutive lanes (the run's first lane adds the run's
// length); returns the counter's value before the lane's own increment
// (-1 for key < 0). Every lane of the wave calls it.
__device__ inline int run_add(int* counters, int key)
{
	const int lane = threadIdx.x & 63;
	const int prev = __shfl_up(key, 1, 64);
	const unsigned long long starts = __ballot(lane == 0 || prev != key);
	// the lane's run: from the highest start at or below the lane to the
	// next start above it (or the wave's end)
	const unsigned long long upto = lane == 63 ? ~0ull : ((2ull << lane) - 1);
	const int first = 63 - __clzll(starts & upto);
	const unsigned long long above = starts & ~upto;
	const int end = above ? __ffsll(above) - 1 : 64;
	int base = 0;
	if (lane == first && key >= 0) base = atomicAdd(&counters[key], end - first);
	base = __shfl(base, first, 64);
	return key >= 0 ? base + (lane - first) : -1;
}

// Device statistics are spread over STAT_SLOTS copies of each counter, one
// per 128-byte line, and added once per workgroup: atomics from thousands of
// workgroups on one address serialise at the L2. The host folds the copies.
constexpr int STAT_SLOTS = 64;
constexpr int STAT_STRIDE = 16; // unsigned long longs per slot (128 bytes)
constexpr size_t STAT_WORDS = (size_t)C_COUNT * STAT_SLOTS * STAT_STRIDE;

__device__ inline unsigned long long* stat_slot(unsigned long long* stats, int c)
{
	return stats + ((size_t)c * STAT_SLOTS + ((blockIdx.x + 31u * blockIdx.y) & (STAT_SLOTS - 1))) * STAT_STRIDE;
}


// the per-slot copies of every counter folded into out[k]: max for the
// error bits and the residual, or for the NaN flag, sums otherwise
__global__ void k_fold_stats(const unsigned long long* __restrict__ stats, unsigned long long* __restrict__ out)
{
	const int k = threadIdx.x;
	if (k >= C_COUNT) return;
	unsigned long long v = 0;
	for (int q = 0; q < STAT_SLOTS; q++)
	{
		const unsigned long long x = stats[((size_t)k * STAT_SLOTS + q) * STAT_STRIDE];
		if (k == C_ERRBITS || k == C_RESID) v = x > v ? x : v;
		else if (k == C_NAN) v |= x;
		else v += x;
	}
	out[k] = v;
}

// adds v[k] of every thread of the block to counter which[k]; every thread
// of a BS-sized workgroup calls this
template <int NC, int BS = BLOCK>
__device__ __forceinline__ void block_stats(unsigned long long* stats, const int (&which)[NC], unsigned long long (&v)[NC])
{
	__shared__ unsigned long long red[NC][BS / 64];
	const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
#pragma unroll
	for (int k = 0; k < NC; k++)
	{
		const unsigned long long s = wave_sum(v[k]);
		if (lane == 0) red[k][w] = s;
	}
	__syncthreads();
	if (threadIdx.x < NC)
	{
		unsigned long long s = 0;
		for (int q = 0; q < BS / 64; q++) s += red[threadIdx.x][q];
		// the counter by a select chain: a local array indexed by the thread
		// would be placed in scratch memory
		int c = which[0];
#pragma unroll
		for (int k = 1; k < NC; k++)
			if ((int)threadIdx.x == k) c = which[k];
		if (s) atomicAdd(stat_slot(stats, c), s);
	}
	__syncthreads();
}

// Append to a list whose length is *counter with one atomic per workgroup;
// every thread of a BLOCK-sized workgroup calls this. Returns the slot of a
// thread with pred, -1 otherwise.
__device__ inline int block_append(int* counter, bool pred)
{
	__shared__ int wcnt[BLOCK / 64];
	__shared__ int base;
	const unsigned long long m = __ballot(pred);
	const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
	if (lane == 0) wcnt[w] = (int)__popcll(m);
	__syncthreads();
	if (threadIdx.x == 0)
	{
		int t = 0;
		for (int q = 0; q < BLOCK / 64; q++)
		{
			const int c = wcnt[q];
			wcnt[q] = t;
			t += c;
		}
		base = t ? atomicAdd(counter, t) : 0;
	}
	__syncthreads();
	const unsigned long long below = lane == 0 ? 0ull : (~0ull >> (64 - lane));
	const int pos = pred ? base + wcnt[w] + (int)__popcll(m & below) : -1;
	__syncthreads();
	return pos;
}

// 1/x for x > 0: hardware reciprocal estimate refined by two Newton steps
// (each squares the relative error of the ~2^-23 estimate)
__device__ inline double rcp64(double x)
{
	double r = __builtin_amdgcn_rcp(x);
	double e = fma(-x, r, 1.0);
	r = fma(r, e, r);
	e = fma(-x, r, 1.0);
	return fma(r, e, r);
}

// phi-row index and angle of solved row r of a phase (ccf_functions.cuh:338-374).
// rows_per_branch 1: the branch centres (j = 0); 2: both sides of step j;
// 0: row r of a caller-supplied array, all at angle phi_fixed.
// angle of phi step j on side sgn of branch c (ccf_functions.cuh:338-374)
__device__ inline double branch_phi(int c, int sgn, int j, int np, int nb)
{
	const double phi0 = PI / nb + c * 2 * PI / nb;
	const double dphi = 2 * PI / np * j;
	return phi0 + sgn * dphi;
}

struct RowMap
{
	int R, np, nb, j, rows_per_branch;
	double phi_fixed;
	const double2* cs = nullptr; // (cos phi, sin phi) of every phi row, or null
	int c0 = 0;                  // the first branch of the rows (a group of branches solved on its own stream)
	__device__ inline void get(int r, int& row, double& phi) const
	{
		if (rows_per_branch == 0)
		{
			row = r;
			phi = phi_fixed;
			return;
		}
		const int c = c0 + r / rows_per_branch;
		const int sgn = rows_per_branch == 1 ? 0 : ((r % rows_per_branch) == 0 ? -1 : 1);
		const int center = np / (2 * nb) + c * (np / nb + 1);
		row = center + sgn * j;
		phi = branch_phi(c, sgn, j, np, nb);
	}
	__device__ inline void phase(int row, double phi, double& cp, double& sp) const
	{
		if (cs && rows_per_branch != 0)
		{
			const double2 t = cs[row];
			cp = t.x;
			sp = t.y;
		}
		else
		{
			sincos(phi, &sp, &cp);
		}
	}
};

// (cos phi, sin phi) of every phi row of the output, row = c (np/nb + 1) + k
__global__ void k_phase(int np, int nb, double2* cs)
{
	const int row = blockIdx.x * blockDim.x + threadIdx.x;
	if (row >= np + nb) return;
	const int per = np / nb + 1;
	const int c = row / per, k = row % per - np / (2 * nb);
	const int sgn = k < 0 ? -1 : 1;
	double sp, cp;
	sincos(branch_phi(c, sgn, sgn * k, np, nb), &sp, &cp);
	cs[row] = make_double2(cp, sp);
}

// device pointers of a SweepBufs, with the parity of the current sweep applied
struct SweepView
{
	int R, rows, tile;
	const int* act;
	int* act_next;
	int* cnt;      // cleared by the sweep's plan once read
	int* cnt_next;
	int* fl;
	int* fcnt;
	int* fcnt_next;
	double2* f0;
	double2* f1;
	double2* ex;
	int* tile_off;
	int* plan;
	int* ticket;
	double2* part;
	int* slot;     // root-sum FMM only (else null)
	int* rmode;
	double2* rsum;
};

SweepView view(const SweepBufs& B, int parity)
{
	SweepView v;
	v.R = B.R;
	v.rows = B.rows;
	v.tile = B.tile;
	const size_t list = (size_t)B.rows * B.R;
	v.act = B.act + parity * list;
	v.act_next = B.act + (parity ^ 1) * list;
	v.cnt = B.cnt + parity * B.rows;
	v.cnt_next = B.cnt + (parity ^ 1) * B.rows;
	v.fl = B.fl;
	v.fcnt = B.fcnt + parity * B.rows;
	v.fcnt_next = B.fcnt + (parity ^ 1) * B.rows;
	v.f0 = B.f0;
	v.f1 = B.f1;
	v.ex = B.ex;
	v.tile_off = B.tile_off;
	v.plan = B.plan;
	v.ticket = B.ticket;
	v.part = B.part;
	v.slot = B.slot;
	v.rmode = B.rmode;
	v.rsum = B.rsum;
	return v;
}

// what K1 does with a root once its Witt sums are known (ccf_functions.cuh:
// 197-227): the 1/mu early exit, or f1 and the Aberth terms formed per root.
// MIXED: the mixed-precision form, whose early-exit roots also take one FP64
// Newton step z - f0/f1 (a final polish; the reference keeps z as it is) and
// whose pole sum is the near-field sum plus the FP32 far field. With newton
// set (the continuation steps), a root whose Newton step is already below
// the convergence threshold takes that step and ends without the sums; one
// whose step is below NEWTON_STEP_TOL takes it and stays active.
template <bool MIXED>
__device__ inline int witt_root(const LensView& L, double2* p, cd z, const Witt& w, int circ_exact, int fmm_pole, int newton,
	cd& f1, cd& ex)
{
	if (w.early)
	{
		if (MIXED)
		{
			const cd dz = w.f0 / witt_f1(L, z, w);
			if (dz.abs() < 1e-6 * L.theta) *p = make_double2(z.re - dz.re, z.im - dz.im);
		}
		return ROOT_EARLY;
	}
	f1 = witt_f1(L, z, w);
	if (MIXED && newton)
	{
		const cd dz = w.f0 / f1;
		const double adz = dz.abs();
		if (adz < NEWTON_STEP_TOL * L.theta) // false for a non-finite step
		{
			*p = make_double2(z.re - dz.re, z.im - dz.im);
			return adz < CONV_TOL * L.theta ? ROOT_NEWTON : ROOT_STEPPED;
		}
	}
	ex = cd();
	if (circ_exact) ex = 2.0 * z.inv(); // the disc's pole at 0 (ccf_functions.cuh:223-227)
	if (MIXED && fmm_pole) ex += 2.0 * (w.nr.s0 + far_pole32(L, w.node, z));
	return ROOT_FULL;
}

// The near-field sums of z's node (a node that is not a leaf: z lies
// outside the tree, whose root then holds every star) over its star
// ranges, the terms of p2p's range loop split over the workgroup's lanes
// and reduced in a fixed order; the result is valid in thread 0. Every
// thread of the workgroup calls it.
template <int BS = BLOCK>
__device__ inline Near coop_near(const LensView& L, int node, cd z, double (*red)[8])
{
	Near nr;
	const int ranges = L.nnf[node];
	for (int k = 0; k < ranges; k++)
	{
		const int2 rg = L.nf[node * MAX_NEAR_RANGES + k];
		for (int s = rg.x + threadIdx.x; s < rg.x + rg.y; s += BS)
		{
			const double2 q = L.spos[s];
			const cd t = cd(z.re - q.x, z.im - q.y).inv();
			nr.s0 += t;
			const cd t1 = t * L.smass[s];
			nr.s1 += t1;
			const cd t2 = t1 * t;
			nr.s2 += t2;
			nr.s3 += t2 * t;
		}
		nr.count += rg.y;
	}
	double v[8] = {nr.s0.re, nr.s0.im, nr.s1.re, nr.s1.im, nr.s2.re, nr.s2.im, nr.s3.re, nr.s3.im};
#pragma unroll
	for (int q = 0; q < 8; q++)
		for (int o = 32; o >= 1; o >>= 1) v[q] += __shfl_xor(v[q], o, 64);
	if ((threadIdx.x & 63) == 0)
		for (int q = 0; q < 8; q++) red[threadIdx.x >> 6][q] = v[q];
	__syncthreads();
	Near sum;
	sum.count = nr.count;
	if (threadIdx.x == 0)
	{
		double t[8] = {0, 0, 0, 0, 0, 0, 0, 0};
		for (int w = 0; w < BS / 64; w++)
			for (int q = 0; q < 8; q++) t[q] += red[w][q];
		sum.s0 = cd(t[0], t[1]);
		sum.s1 = cd(t[2], t[3]);
		sum.s2 = cd(t[4], t[5]);
		sum.s3 = cd(t[6], t[7]);
	}
	return sum;
}

// K2: tiles per row (exclusive prefix in tile_off) and slices per tile; a
// row with at least fmm_min listed roots (fmm_min > 0) takes its root sums
// from the FMM pass instead and gets no tiles. Formed by the last workgroup
// of k_witt to finish, once every row's listed count is final.

struct PlanArgs
{
	int nrows, terms, target, fmm_min, sweep;
};

__device__ inline void plan_sums(SweepView V, const PlanArgs& P)
{
	const int nrows = P.nrows, terms = P.terms, target = P.target, fmm_min = P.fmm_min, sweep = P.sweep;
	__shared__ int wsum[16];
	__shared__ int carry, nfmm;
	const int lane = threadIdx.x & 63, w = threadIdx.x >> 6, nw = blockDim.x >> 6;
	if (threadIdx.x == 0) carry = nfmm = 0;
	__syncthreads();
	for (int base = 0; base < nrows; base += blockDim.x)
	{
		const int r = base + threadIdx.x;
		const int nf = r < nrows ? __hip_atomic_load(&V.fcnt[r], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) : 0;
		const bool fmm = r < nrows && fmm_min > 0 && nf >= fmm_min;
		if (r < nrows && V.rmode) V.rmode[r] = fmm ? 1 : 0;
		if (fmm) atomicAdd(&nfmm, 1);
		// this sweep's active count is consumed (k_witt read it): cleared
		// here, it is the list the next sweep's k_witt appends stepped roots to
		if (r < nrows) V.cnt[r] = 0;
		const int v = (r < nrows && !fmm) ? (nf + V.tile - 1) / V.tile : 0;
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
		if (r < nrows) V.tile_off[r] = carry + woff + x - v;
		__syncthreads();
		if (threadIdx.x == 0) carry += tot;
		__syncthreads();
	}
	if (threadIdx.x == 0)
	{
		const int tiles = carry;
		int S = 1;
		if (tiles > 0 && tiles < target)
		{
			S = (target + tiles - 1) / tiles;
			S = min(S, MAX_SLICES);
			S = min(S, max(1, terms / MIN_SLICE));
		}
		V.tile_off[nrows] = tiles;
		V.plan[0] = tiles;
		V.plan[1] = S;
		// FMM rows of the phase's first and second sweep (cleared by k_prepare)
		if (sweep < 2 && nfmm) V.plan[2 + sweep] += nfmm;
	}
}

// The near-field sums of a root outside the tree (its node is not a leaf:
// every star of the node's ranges), by the lanes of one wave: lane l sums
// the stars l, l + 64, ... of the ranges (p2p's terms), and a butterfly of
// fixed order gives every lane the total. Every lane of the wave calls it
// with the same z and node.
__device__ inline Near wave_near(const LensView& L, int node, cd z)
{
	Near nr;
	const int lane = threadIdx.x & 63;
	const int ranges = L.nnf[node];
	for (int k = 0; k < ranges; k++)
	{
		const int2 rg = L.nf[node * MAX_NEAR_RANGES + k];
		const int e = rg.x + rg.y;
		// the lane's stars in order, two loads issued before their terms
		for (int s0 = rg.x + lane; s0 < e; s0 += 2 * 64)
		{
			double2 q[2];
			double m[2];
#pragma unroll
			for (int u = 0; u < 2; u++)
			{
				const int s = min(s0 + u * 64, e - 1);
				q[u] = L.spos[s];
				m[u] = L.smass[s];
			}
#pragma unroll
			for (int u = 0; u < 2; u++)
				if (s0 + u * 64 < e)
				{
					const cd t = cd(z.re - q[u].x, z.im - q[u].y).inv();
					nr.s0 += t;
					const cd t1 = t * m[u];
					nr.s1 += t1;
					const cd t2 = t1 * t;
					nr.s2 += t2;
					nr.s3 += t2 * t;
				}
		}
		nr.count += rg.y;
	}
	double v[8] = {nr.s0.re, nr.s0.im, nr.s1.re, nr.s1.im, nr.s2.re, nr.s2.im, nr.s3.re, nr.s3.im};
#pragma unroll
	for (int q = 0; q < 8; q++)
		for (int o = 32; o >= 1; o >>= 1) v[q] += __shfl_xor(v[q], o, 64);
	nr.s0 = cd(v[0], v[1]);
	nr.s1 = cd(v[2], v[3]);
	nr.s2 = cd(v[4], v[5]);
	nr.s3 = cd(v[6], v[7]);
	return nr;
}

// Counts of a workgroup's Witt evaluations (block_stats adds them to the
// device counters)
struct EvalCounts
{
	int full = 0, early = 0, newton = 0, conv = 0, nnear = 0, near_early = 0;
};

// One Witt evaluation of root a of the row xrow and its outcome
// (witt_root), called by every lane of a wave (valid: the lane has a root).
// A root outside the tree (whose near field is every star,
// tree_node.cuh:529-561) takes its near-field sums from its whole wave
// (wave_near), one such root at a time.
template <bool MIXED>
__device__ __forceinline__ int witt_eval(const LensView& L, double2* xrow, int a, bool valid, double cp, double sp, int circ_exact,
	int fmm_pole, int newton, cd& f0, cd& f1, cd& ex, EvalCounts& c)
{
	double2* const p = xrow + a;
	cd z;
	int node = 0;
	if (valid)
	{
		z = cd(p->x, p->y);
		node = nearest_node(L, z);
	}
	const bool outside = valid && L.nf_off[node] < 0;
	Near nr;
	for (unsigned long long m = __ballot(outside); m; m &= m - 1)
	{
		const int src = __ffsll(m) - 1;
		const Near t = wave_near(L, __shfl(node, src, 64), cd(__shfl(z.re, src, 64), __shfl(z.im, src, 64)));
		if ((int)(threadIdx.x & 63) == src) nr = t;
	}
	if (!valid) return -1;
	if (!outside) nr = p2p(L, node, z);
	const Witt w = witt_from(L, z, cp, sp, node, nr);
	const int st = witt_root<MIXED>(L, p, z, w, circ_exact, fmm_pole, newton, f1, ex);
	if (st == ROOT_FULL) f0 = w.f0;
	c.full += st == ROOT_FULL;
	c.early += st == ROOT_EARLY;
	c.newton += st == ROOT_NEWTON || st == ROOT_STEPPED;
	c.conv += st == ROOT_EARLY || st == ROOT_NEWTON;
	c.nnear += st != ROOT_EARLY ? w.nr.count : 0;
	c.near_early += st == ROOT_EARLY ? w.nr.count : 0;
	return st;
}

// appends root a (full) to row r's list for the sweep's Cauchy sums with its
// f0, f1 and extra pole sums; every thread of the block calls it
__device__ __forceinline__ void append_full(const SweepView& V, int r, int R, int a, bool full, cd f0, cd f1, cd ex)
{
	const int pos = block_append(&V.fcnt[r], full);
	if (full)
	{
		const size_t k = (size_t)r * R + pos;
		V.fl[k] = a;
		if (V.slot) V.slot[(size_t)r * R + a] = pos;
		V.f0[k] = make_double2(f0.re, f0.im);
		V.f1[k] = make_double2(f1.re, f1.im);
		V.ex[k] = make_double2(ex.re, ex.im);
	}
}

// K1 (ccf_functions.cuh:197-213 and the early return at :207-211), one
// thread per listed root: the Witt evaluation, then the 1/mu early exit, a
// Newton step or a listing for the sweep's Cauchy sums. In the mixed form a
// root that took a Newton step without converging is evaluated once more in
// the same launch: the workgroup gathers its stepped roots in LDS and its
// first threads evaluate them, so that the roots left for the next sweep
// are those that still need it. Every workgroup then takes a ticket; the
// last to finish sees every row's listed count and forms the plan of the
// sweep's Cauchy sums.
template <bool MIXED>
__global__ __launch_bounds__(BLOCK) void k_witt(LensView L, double2* __restrict__ roots, RowMap M, SweepView V,
	unsigned long long* __restrict__ stats, int circ_exact, int fmm_pole, PlanArgs P)
{
	__shared__ int s_last;
	const int2 sl = xcd_slot();
	const int r = sl.y;
	const int R = M.R;
	// (cnt_next was cleared by the previous sweep's plan or by k_prepare:
	// stepped roots are appended to it below)
	if (xcd_row_leader(sl) && threadIdx.x == 0) V.fcnt_next[r] = 0;
	const int n = V.cnt[r];
	const int i = xcd_x(sl.x, (n + BLOCK - 1) / BLOCK) * BLOCK + threadIdx.x;
	if (i - (int)threadIdx.x < n) // a block with listed roots (uniform)
	{
		EvalCounts c;
		int row;
		double phi, cp, sp;
		M.get(r, row, phi);
		M.phase(row, phi, cp, sp);
		double2* const xrow = roots + (size_t)row * R;
		const bool valid = i < n;
		int a = valid ? V.act[(size_t)r * R + i] : 0;
		cd f0, f1, ex;
		const int st = witt_eval<MIXED>(L, xrow, a, valid, cp, sp, circ_exact, fmm_pole, M.j > 0, f0, f1, ex, c);
		append_full(V, r, R, a, st == ROOT_FULL, f0, f1, ex);
		bool stepped = st == ROOT_STEPPED;
		if (MIXED)
		{
			// the workgroup's stepped roots, in thread order, to its first threads
			__shared__ int s_a2[BLOCK];
			__shared__ int s_w2[BLOCK / 64];
			const unsigned long long m = __ballot(stepped);
			const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
			if (lane == 0) s_w2[wv] = (int)__popcll(m);
			__syncthreads();
			int off = 0, n2 = 0;
			for (int q = 0; q < BLOCK / 64; q++)
			{
				if (q < wv) off += s_w2[q];
				n2 += s_w2[q];
			}
			const unsigned long long below = lane == 0 ? 0ull : (~0ull >> (64 - lane));
			if (stepped) s_a2[off + (int)__popcll(m & below)] = a;
			__syncthreads();
			stepped = false;
			if (n2 > 0) // uniform
			{
				const bool v2 = (int)threadIdx.x < n2;
				a = v2 ? s_a2[threadIdx.x] : 0;
				const int st2 = witt_eval<MIXED>(L, xrow, a, v2, cp, sp, circ_exact, fmm_pole, 1, f0, f1, ex, c);
				append_full(V, r, R, a, st2 == ROOT_FULL, f0, f1, ex);
				stepped = st2 == ROOT_STEPPED;
			}
		}
		const int spos = block_append(&V.cnt_next[r], stepped);
		if (stepped) V.act_next[(size_t)r * R + spos] = a;
		const int which[6] = {C_FULL, C_EARLY, C_CONV, C_NEAR, C_NEWTON, C_NEAR_EARLY};
		unsigned long long v[6] = {(unsigned long long)c.full, (unsigned long long)c.early, (unsigned long long)c.conv,
			(unsigned long long)c.nnear, (unsigned long long)c.newton, (unsigned long long)c.near_early};
		block_stats(stats, which, v);
	}
	// every workgroup takes a ticket once its appends are done; the last one
	// sees every row's listed count and forms the sums' plan, then resets
	// the ticket for the next sweep. The plan reads only the list counts,
	// device-scope atomics whose returned values block_append waited for
	// before the ticket is taken, so no release fence precedes it (one per
	// workgroup would write back the L2 of its XCD).
	if (threadIdx.x == 0) s_last = atomicAdd(V.ticket, 1) == (int)(gridDim.x * gridDim.y) - 1;
	__syncthreads();
	if (s_last)
	{
		__threadfence();
		plan_sums(V, P);
		if (threadIdx.x == 0) *V.ticket = 0;
	}
}

// row of global tile g: the largest r with tile_off[r] <= g (rows without
// tiles have tile_off[r] == tile_off[r + 1] and are skipped)
__device__ inline int row_of_tile(const int* __restrict__ tile_off, int nrows, int g)
{
	int lo = 0, hi = nrows;
	while (hi - lo > 1)
	{
		const int mid = (lo + hi) >> 1;
		if (tile_off[mid] <= g) lo = mid;
		else hi = mid;
	}
	return lo;
}

#include "rootfmm.hpp"

// K3 in FP64: partial 2 sum_stars 1/(z - s) - sum_roots 1/(z - w) over one slice
template <int RPT>
__global__ __launch_bounds__(BLOCK) void k_sums64(const double2* __restrict__ spos, int Np, const double2* __restrict__ roots,
	RowMap M, SweepView V, int nrows)
{
	__shared__ double2 buf[CHUNK64];
	const int tiles = V.plan[0], S = V.plan[1];
	const long long items = (long long)tiles * S;
	const int R = M.R, T = BLOCK * RPT;
	const long long Lt = (long long)Np + R;
	for (long long item = blockIdx.x; item < items; item += gridDim.x)
	{
		const int g = (int)(item / S), s = (int)(item % S);
		const int r = row_of_tile(V.tile_off, nrows, g);
		const int t = g - V.tile_off[r];
		const int nfr = V.fcnt[r];
		int row;
		double phi;
		M.get(r, row, phi);
		const double2* __restrict__ xrow = roots + (size_t)row * R;
		double zx[RPT], zy[RPT], ax[RPT], ay[RPT], bx[RPT], by[RPT];
#pragma unroll
		for (int k = 0; k < RPT; k++)
		{
			const int i = t * T + k * BLOCK + threadIdx.x;
			const int a = V.fl[(size_t)r * R + (i < nfr ? i : t * T)];
			const double2 v = xrow[a];
			zx[k] = v.x;
			zy[k] = v.y;
			ax[k] = ay[k] = bx[k] = by[k] = 0;
		}
		const long long lo = s * Lt / S, hi = (s + 1) * Lt / S;
		// star terms (ccf_functions.cuh:218-227)
		const int s1 = (int)min(hi, (long long)Np);
		for (int c0 = (int)lo; c0 < s1; c0 += CHUNK64)
		{
			const int m = min(CHUNK64, s1 - c0);
			for (int q = threadIdx.x; q < m; q += BLOCK) buf[q] = spos[c0 + q];
			__syncthreads();
#pragma unroll 4
			for (int q = 0; q < m; q++)
			{
				const double2 w = buf[q];
#pragma unroll
				for (int k = 0; k < RPT; k++)
				{
					const double dx = zx[k] - w.x, dy = zy[k] - w.y;
					const double iv = rcp64(fma(dx, dx, dy * dy));
					ax[k] = fma(dx, iv, ax[k]);
					ay[k] = fma(dy, iv, ay[k]);
				}
			}
			__syncthreads();
		}
		// root terms (ccf_functions.cuh:232-239); the own root gives dx = dy = 0
		// exactly and the floor on |d|^2 turns its term into 0
		const int r0 = (int)(max(lo, (long long)Np) - Np), r1 = (int)(hi - Np);
		for (int c0 = r0; c0 < r1; c0 += CHUNK64)
		{
			const int m = min(CHUNK64, r1 - c0);
			for (int q = threadIdx.x; q < m; q += BLOCK) buf[q] = xrow[c0 + q];
			__syncthreads();
#pragma unroll 4
			for (int q = 0; q < m; q++)
			{
				const double2 w = buf[q];
#pragma unroll
				for (int k = 0; k < RPT; k++)
				{
					const double dx = zx[k] - w.x, dy = zy[k] - w.y;
					const double iv = rcp64(fmax(fma(dx, dx, dy * dy), 1e-300));
					bx[k] = fma(dx, iv, bx[k]);
					by[k] = fma(dy, iv, by[k]);
				}
			}
			__syncthreads();
		}
#pragma unroll
		for (int k = 0; k < RPT; k++)
			V.part[(size_t)item * T + k * BLOCK + threadIdx.x] = make_double2(2 * ax[k] - bx[k], -(2 * ay[k] - by[k]));
	}
}

// K3 in FP32 with the FP64 fallback for clustered roots
// The items of tiles with few listed roots (at most SPARSE_TARGETS), in the
// same launch as the dense items (a branch uniform per item): each wave
// takes groups of SPARSE_GROUP targets (wave w the groups w, w + waves, ...),
// its lanes run over the slice's sources, SPARSE_LOADS at a time, and a
// butterfly of fixed order over the wave gives each target's partial sum,
// written where the tile's own lane would have written it (part[i] for
// target i of the tile). A root within the cluster distance has its one term
// formed in FP64; the target's own index is skipped by index, not by
// distance.
constexpr int SPARSE_TARGETS = 32;
constexpr int SPARSE_GROUP = 4;
constexpr int SPARSE_LOADS = 2;

__device__ inline void sums32_sparse(const float2* __restrict__ spos32, int Np, const double2* __restrict__ xrow,
	const int* __restrict__ fl, int nt, long long lo, long long hi, float cluster_d2, double2* __restrict__ part,
	unsigned long long& nclus)
{
	const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
	const int s1 = (int)min(hi, (long long)Np);
	const int r0 = (int)(max(lo, (long long)Np) - Np), r1 = (int)(hi - Np);
#pragma unroll 1
	for (int g0 = wv * SPARSE_GROUP; g0 < nt; g0 += (BLOCK / 64) * SPARSE_GROUP)
	{
		int ia[SPARSE_GROUP];
		double cx[SPARSE_GROUP], cy[SPARSE_GROUP];
		float fx[SPARSE_GROUP], fy[SPARSE_GROUP], ax[SPARSE_GROUP], ay[SPARSE_GROUP], bx[SPARSE_GROUP], by[SPARSE_GROUP];
#pragma unroll
		for (int g = 0; g < SPARSE_GROUP; g++)
		{
			// slots past the last target repeat it; their sums are not written
			ia[g] = fl[min(g0 + g, nt - 1)];
			const double2 v = xrow[ia[g]];
			fx[g] = (float)v.x;
			fy[g] = (float)v.y;
			ax[g] = ay[g] = bx[g] = by[g] = 0;
			cx[g] = cy[g] = 0;
		}
#pragma unroll 1
		for (int q0 = (int)lo + lane; q0 < s1; q0 += 64 * SPARSE_LOADS)
		{
			float2 w[SPARSE_LOADS];
#pragma unroll
			for (int u = 0; u < SPARSE_LOADS; u++) w[u] = spos32[min(q0 + 64 * u, s1 - 1)];
#pragma unroll
			for (int u = 0; u < SPARSE_LOADS; u++)
				if (q0 + 64 * u < s1)
#pragma unroll
					for (int g = 0; g < SPARSE_GROUP; g++)
					{
						const float dx = fx[g] - w[u].x, dy = fy[g] - w[u].y;
						const float iv = __builtin_amdgcn_rcpf(fmaf(dx, dx, dy * dy));
						ax[g] = fmaf(dx, iv, ax[g]);
						ay[g] = fmaf(dy, iv, ay[g]);
					}
		}
#pragma unroll 1
		for (int q0 = r0 + lane; q0 < r1; q0 += 64 * SPARSE_LOADS)
		{
			float2 w[SPARSE_LOADS];
#pragma unroll
			for (int u = 0; u < SPARSE_LOADS; u++)
			{
				const double2 v = xrow[min(q0 + 64 * u, r1 - 1)];
				w[u] = make_float2((float)v.x, (float)v.y);
			}
#pragma unroll
			for (int u = 0; u < SPARSE_LOADS; u++)
			{
				const int q = q0 + 64 * u;
				if (q >= r1) continue;
#pragma unroll
				for (int g = 0; g < SPARSE_GROUP; g++)
				{
					const float dx = fx[g] - w[u].x, dy = fy[g] - w[u].y;
					const float d2 = fmaf(dx, dx, dy * dy);
					if (d2 < cluster_d2 && q != ia[g])
					{
						// rare: the source and the target in FP64 again
						const double2 z = xrow[ia[g]], wd = xrow[q];
						const double ex = z.x - wd.x, ey = z.y - wd.y;
						const double iv = rcp64(fmax(fma(ex, ex, ey * ey), 1e-300));
						cx[g] = fma(ex, iv, cx[g]);
						cy[g] = fma(ey, iv, cy[g]);
						nclus++;
					}
					else
					{
						// the own root gives dx = dy = 0 and, floored, a zero term
						const float iv = __builtin_amdgcn_rcpf(fmaxf(d2, 1e-30f));
						bx[g] = fmaf(dx, iv, bx[g]);
						by[g] = fmaf(dy, iv, by[g]);
					}
				}
			}
		}
#pragma unroll
		for (int g = 0; g < SPARSE_GROUP; g++)
		{
			double vx = 2.0 * (double)ax[g] - ((double)bx[g] + cx[g]);
			double vy = 2.0 * (double)ay[g] - ((double)by[g] + cy[g]);
			for (int o = 32; o >= 1; o >>= 1)
			{
				vx += __shfl_xor(vx, o, 64);
				vy += __shfl_xor(vy, o, 64);
			}
			if (lane == 0 && g0 + g < nt) part[g0 + g] = make_double2(vx, -vy);
		}
	}
}

template <int RPT>
__global__ __launch_bounds__(BLOCK) void k_sums32(const float2* __restrict__ spos32, int Np, const double2* __restrict__ roots,
	RowMap M, SweepView V, int nrows, float cluster_d2, unsigned long long* __restrict__ stats)
{
	__shared__ float2 buf[CHUNK32];
	const int tiles = V.plan[0], S = V.plan[1];
	const long long items = (long long)tiles * S;
	const int R = M.R, T = BLOCK * RPT;
	const long long Lt = (long long)Np + R;
	unsigned long long nclus = 0;
	for (long long item = blockIdx.x; item < items; item += gridDim.x)
	{
		const int g = (int)(item / S), s = (int)(item % S);
		const int r = row_of_tile(V.tile_off, nrows, g);
		const int t = g - V.tile_off[r];
		const int nfr = V.fcnt[r];
		int row;
		double phi;
		M.get(r, row, phi);
		const double2* __restrict__ xrow = roots + (size_t)row * R;
		const long long lo = s * Lt / S, hi = (s + 1) * Lt / S;
		if (nfr - t * T <= SPARSE_TARGETS)
		{
			// a tile of few listed roots: lanes over the sources (uniform per item)
			sums32_sparse(spos32, Np, xrow, V.fl + (size_t)r * R + (size_t)t * T, nfr - t * T, lo, hi, cluster_d2,
				V.part + (size_t)item * T, nclus);
			continue;
		}
		double zx[RPT], zy[RPT], bx[RPT], by[RPT];
		float fx[RPT], fy[RPT], ax[RPT], ay[RPT];
		int ia[RPT];
#pragma unroll
		for (int k = 0; k < RPT; k++)
		{
			const int i = t * T + k * BLOCK + threadIdx.x;
			const int a = V.fl[(size_t)r * R + (i < nfr ? i : t * T)];
			ia[k] = a;
			const double2 v = xrow[a];
			zx[k] = v.x;
			zy[k] = v.y;
			fx[k] = (float)v.x;
			fy[k] = (float)v.y;
			ax[k] = ay[k] = 0;
			bx[k] = by[k] = 0;
		}
		const int s1 = (int)min(hi, (long long)Np);
		for (int c0 = (int)lo; c0 < s1; c0 += CHUNK32)
		{
			const int m = min(CHUNK32, s1 - c0);
			for (int q = threadIdx.x; q < m; q += BLOCK) buf[q] = spos32[c0 + q];
			__syncthreads();
#pragma unroll 4
			for (int q = 0; q < m; q++)
			{
				const float2 w = buf[q];
#pragma unroll
				for (int k = 0; k < RPT; k++)
				{
					const float dx = fx[k] - w.x, dy = fy[k] - w.y;
					const float iv = __builtin_amdgcn_rcpf(fmaf(dx, dx, dy * dy));
					ax[k] = fmaf(dx, iv, ax[k]);
					ay[k] = fmaf(dy, iv, ay[k]);
				}
			}
			__syncthreads();
		}
		const int r0 = (int)(max(lo, (long long)Np) - Np), r1 = (int)(hi - Np);
		for (int c0 = r0; c0 < r1; c0 += CHUNK32)
		{
			const int m = min(CHUNK32, r1 - c0);
			for (int q = threadIdx.x; q < m; q += BLOCK)
			{
				const double2 w = xrow[c0 + q];
				buf[q] = make_float2((float)w.x, (float)w.y);
			}
			__syncthreads();
			float cx[RPT], cy[RPT], dmin[RPT];
#pragma unroll
			for (int k = 0; k < RPT; k++) { cx[k] = cy[k] = 0; dmin[k] = 3.0e38f; }
#pragma unroll 4
			for (int q = 0; q < m; q++)
			{
				const float2 w = buf[q];
#pragma unroll
				for (int k = 0; k < RPT; k++)
				{
					const float dx = fx[k] - w.x, dy = fy[k] - w.y;
					const float d2 = fmaxf(fmaf(dx, dx, dy * dy), 1e-30f);
					const float iv = __builtin_amdgcn_rcpf(d2);
					cx[k] = fmaf(dx, iv, cx[k]);
					cy[k] = fmaf(dy, iv, cy[k]);
					// every root but the own one counts as a neighbour, also
					// one at the same FP32 position
					dmin[k] = c0 + q != ia[k] ? fminf(dmin[k], d2) : dmin[k];
				}
			}
#pragma unroll
			for (int k = 0; k < RPT; k++)
			{
				if (dmin[k] < cluster_d2)
				{
					// a root of this chunk is close: redo the chunk in FP64
					double ex = 0, ey = 0;
					for (int q = 0; q < m; q++)
					{
						const double2 w = xrow[c0 + q];
						const double dx = zx[k] - w.x, dy = zy[k] - w.y;
						const double iv = rcp64(fmax(fma(dx, dx, dy * dy), 1e-300));
						ex = fma(dx, iv, ex);
						ey = fma(dy, iv, ey);
					}
					bx[k] += ex;
					by[k] += ey;
					nclus++;
				}
				else
				{
					bx[k] += cx[k];
					by[k] += cy[k];
				}
			}
			__syncthreads();
		}
#pragma unroll
		for (int k = 0; k < RPT; k++)
			V.part[(size_t)item * T + k * BLOCK + threadIdx.x] =
				make_double2(2.0 * ax[k] - bx[k], -(2.0 * ay[k] - by[k]));
	}
	nclus = wave_sum(nclus);
	if ((threadIdx.x & 63) == 0 && nclus) atomicAdd(stat_slot(stats, C_CLUSTER), nclus);
}

// K4 (ccf_functions.cuh:241-242 and 379-389)
__global__ __launch_bounds__(BLOCK) void k_finalize(double theta, double2* __restrict__ roots, RowMap M, SweepView V,
	unsigned long long* __restrict__ stats)
{
	const int r = blockIdx.y;
	const int R = M.R;
	const int i = blockIdx.x * BLOCK + threadIdx.x;
	const int n = V.fcnt[r];
	if (blockIdx.x * BLOCK >= n) return; // a block without listed roots (uniform)
	unsigned long long nc = 0;
	bool again = false;
	int a = 0;
	if (i < n)
	{
		int row;
		double phi;
		M.get(r, row, phi);
		const size_t k = (size_t)r * R + i;
		a = V.fl[k];
		double2* p = roots + (size_t)row * R + a;
		const cd z(p->x, p->y);
		const int S = V.plan[1];
		const size_t g = (size_t)(V.tile_off[r] + i / V.tile);
		const int loc = i % V.tile;
		double2 e = V.ex[k];
		cd sum(e.x, e.y);
		if (V.rmode && V.rmode[r])
		{
			const double2 q = V.rsum[k];
			sum += cd(q.x, q.y);
		}
		else
		{
			for (int s = 0; s < S; s++)
			{
				const double2 q = V.part[(g * S + s) * V.tile + loc];
				sum += cd(q.x, q.y);
			}
		}
		if (V.slot) V.slot[(size_t)r * R + a] = -1;
		const double2 a0 = V.f0[k], a1 = V.f1[k];
		const cd f0(a0.x, a0.y), f1(a1.x, a1.y);
		const cd zn = z - f0 / (f1 + f0 * sum);
		*p = make_double2(zn.re, zn.im);
		if ((zn - z).abs() / theta < CONV_TOL) nc = 1;
		else again = true;
	}
	const int pos = block_append(&V.cnt_next[r], again);
	if (again) V.act_next[(size_t)r * R + pos] = a;
	const int which[1] = {C_CONV};
	unsigned long long v[1] = {nc};
	block_stats(stats, which, v);
}

// The sweeps after the first few of a continuation step in one launch
// (mixed form): one workgroup per row runs the row's remaining sweeps (at
// most cap) with the sweep of the kernels above: the Witt evaluation with
// its early exit and Newton steps (witt_root, roots outside the tree
// through p2p's star ranges), the Cauchy sums of the listed roots, the
// Aberth update and the |dz| / theta < 1e-9 test, Jacobi within the row.
// The sums take one wave per listed root with its lanes over the row's
// roots (FP32, a close root's term in FP64, the own root skipped by index)
// and, without the FMM pole sum, the stars. The active lists ping-pong
// between two buffers, and the listed roots carry f0, f1 and their Aberth
// terms; entry k of each lives in LDS for k < TAIL_LDS and otherwise in the
// row's slice of act (both halves), fl, f0, f1 and ex (the sweep kernels are
// done with them), so a row of any size fits. The waves of the workgroup
// see each other's global writes across its barriers.
constexpr int TAIL_BLOCK = 512;
constexpr int TAIL_LDS = 128;
constexpr int TAIL_DEFER = 64; // roots outside the tree per sweep taken by the whole workgroup
// sweeps of a continuation step before the tail launch: the first sweep
// evaluates every root of every row, which takes the whole device; after it
// a row has a few per cent of its roots left, mostly ending by Newton steps,
// which one workgroup per row finishes while the other group of branches'
// first sweep fills the device
constexpr int TAIL_AT = 1;

// k_tail: the outcome of root a's evaluation w, as in k_witt: listed for
// the sums (entry k of the listed roots: LDS for k < TAIL_LDS, else the
// global slices), active again next sweep (the next list), or done
__device__ __forceinline__ void tail_settle(const LensView& L, int a, double2* pz, cd z, const Witt& w, int circ_exact, int fmm_pole,
	bool newton, int& s_nf, int& s_nn, int* s_fl, double2* s_f0, double2* s_f1, double2* s_ex, int* s_next, int* gfl, double2* gf0,
	double2* gf1, double2* gex, int* gnxt, unsigned long long& nfull, unsigned long long& nearly, unsigned long long& nnewton,
	unsigned long long& nconv, unsigned long long& nnear)
{
	cd f1, ex;
	const int st = witt_root<true>(L, pz, z, w, circ_exact, fmm_pole, newton, f1, ex);
	nnear += w.nr.count;
	// counted without branches: increments chosen by branch are merged into
	// one through a pointer, which places the counters in scratch memory
	const bool settled = st != ROOT_FULL && st != ROOT_STEPPED;
	nfull += st == ROOT_FULL;
	nnewton += st == ROOT_STEPPED || st == ROOT_NEWTON;
	nconv += settled;
	nearly += settled && st != ROOT_NEWTON;
	if (st == ROOT_FULL)
	{
		const int k = atomicAdd(&s_nf, 1);
		const double2 v0 = make_double2(w.f0.re, w.f0.im), v1 = make_double2(f1.re, f1.im), ve = make_double2(ex.re, ex.im);
		if (k < TAIL_LDS)
		{
			s_fl[k] = a;
			s_f0[k] = v0;
			s_f1[k] = v1;
			s_ex[k] = ve;
		}
		else
		{
			gfl[k] = a;
			gf0[k] = v0;
			gf1[k] = v1;
			gex[k] = ve;
		}
	}
	else if (st == ROOT_STEPPED)
	{
		const int k = atomicAdd(&s_nn, 1);
		if (k < TAIL_LDS) s_next[k] = a;
		else gnxt[k] = a;
	}
}

__global__ __launch_bounds__(TAIL_BLOCK) void k_tail(LensView L, double2* __restrict__ roots, RowMap M, SweepView V,
	unsigned long long* __restrict__ stats, int circ_exact, int fmm_pole, int cap, float cluster_d2,
	const float2* __restrict__ spos32, int Np, double theta)
{
	__shared__ int s_a[2][TAIL_LDS];
	__shared__ int s_fl[TAIL_LDS];
	__shared__ double2 s_f0[TAIL_LDS], s_f1[TAIL_LDS], s_ex[TAIL_LDS];
	__shared__ int s_nf, s_nn; // listed for the sums this sweep, active for the next
	__shared__ int s_def[TAIL_DEFER], s_nd; // roots outside the tree, evaluated by the whole workgroup
	__shared__ double s_red[TAIL_BLOCK / 64][8];
	constexpr int NW = TAIL_BLOCK / 64;
	const int r = blockIdx.x, R = M.R;
	const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
	int row;
	double phi;
	M.get(r, row, phi);
	double cp, sp;
	M.phase(row, phi, cp, sp);
	double2* __restrict__ xrow = roots + (size_t)row * R;
	const size_t base = (size_t)r * R;
	// the lists' global parts: cur this sweep, nxt the next (swapped with p)
	int* gcur = const_cast<int*>(V.act) + base;
	int* gnxt = V.act_next + base;
	int* __restrict__ gfl = V.fl + base;
	double2* __restrict__ gf0 = V.f0 + base;
	double2* __restrict__ gf1 = V.f1 + base;
	double2* __restrict__ gex = V.ex + base;
	unsigned long long nfull = 0, nearly = 0, nnewton = 0, nconv = 0, nnear = 0, nclus = 0, nsweeps = 0;
	// the first list: the roots the previous sweep's Witt kernels stepped
	// (its head into LDS; its other entries stay where they are), then the
	// previous sweep's updates of the row (k_finalize's work): its listed
	// roots (the other parity's count) take their sums and move, and those
	// not converged join the list
	const int n0 = V.cnt[r];
	for (int q = threadIdx.x; q < min(n0, TAIL_LDS); q += TAIL_BLOCK) s_a[0][q] = gcur[q];
	if (threadIdx.x == 0)
	{
		s_nf = s_nd = 0;
		s_nn = n0;
	}
	__syncthreads();
	{
		const int nl = V.fcnt_next[r], S = V.plan[1];
		for (int i = threadIdx.x; i < nl; i += TAIL_BLOCK)
		{
			const int a = gfl[i];
			double2* pz = xrow + a;
			const cd z(pz->x, pz->y);
			const double2 e = gex[i];
			cd sum(e.x, e.y);
			if (V.rmode && V.rmode[r])
			{
				const double2 q = V.rsum[base + i];
				sum += cd(q.x, q.y);
			}
			else
			{
				const size_t g = (size_t)(V.tile_off[r] + i / V.tile);
				const int loc = i % V.tile;
				for (int sl = 0; sl < S; sl++)
				{
					const double2 q = V.part[(g * S + sl) * V.tile + loc];
					sum += cd(q.x, q.y);
				}
			}
			if (V.slot) V.slot[base + a] = -1;
			const double2 a0 = gf0[i], a1 = gf1[i];
			const cd f0(a0.x, a0.y), f1(a1.x, a1.y);
			const cd zn = z - f0 / (f1 + f0 * sum);
			*pz = make_double2(zn.re, zn.im);
			if ((zn - z).abs() / theta < CONV_TOL) nconv++;
			else
			{
				const int k = atomicAdd(&s_nn, 1);
				if (k < TAIL_LDS) s_a[0][k] = a;
				else gcur[k] = a;
			}
		}
	}
	__syncthreads();
	int n = s_nn;
	__syncthreads();
	if (threadIdx.x == 0) s_nn = 0;
	__syncthreads();
	int p = 0;
	for (int it = 0; it < cap && n > 0; it++)
	{
		nsweeps++;
		for (int q = threadIdx.x; q < n; q += TAIL_BLOCK)
		{
			const int a = q < TAIL_LDS ? s_a[p][q] : gcur[q];
			double2* pz = xrow + a;
			const cd z(pz->x, pz->y);
			const int node = nearest_node(L, z);
			// outside the tree (every star in the near field): below, with
			// the whole workgroup (inline only past TAIL_DEFER of them)
			if (L.nf_off[node] < 0)
			{
				const int k = atomicAdd(&s_nd, 1);
				if (k < TAIL_DEFER)
				{
					s_def[k] = a;
					continue;
				}
			}
			const Witt w = witt_from(L, z, cp, sp, node, p2p(L, node, z));
			tail_settle(L, a, pz, z, w, circ_exact, fmm_pole, M.j > 0, s_nf, s_nn, s_fl, s_f0, s_f1, s_ex, s_a[p ^ 1], gfl, gf0, gf1, gex,
				gnxt, nfull, nearly, nnewton, nconv, nnear);
		}
		__syncthreads();
		const int nd = min(s_nd, TAIL_DEFER);
		for (int d = 0; d < nd; d++)
		{
			const int a = s_def[d];
			double2* pz = xrow + a;
			const cd z(pz->x, pz->y);
			const int node = nearest_node(L, z);
			const Near nr = coop_near<TAIL_BLOCK>(L, node, z, s_red);
			if (threadIdx.x == 0)
			{
				const Witt w = witt_from(L, z, cp, sp, node, nr);
				tail_settle(L, a, pz, z, w, circ_exact, fmm_pole, M.j > 0, s_nf, s_nn, s_fl, s_f0, s_f1, s_ex, s_a[p ^ 1], gfl, gf0, gf1,
					gex, gnxt, nfull, nearly, nnewton, nconv, nnear);
			}
			__syncthreads();
		}
		const int nf = s_nf;
		// waves per listed root: all of them share the row when few roots
		// are listed (a power of two, lanes strided over the row)
		const int wpt = nf >= NW ? 1 : (nf >= NW / 2 ? 2 : (nf >= NW / 4 ? 4 : NW));
		const int sub = wv % wpt, lanes = 64 * wpt, ln = sub * 64 + lane;
		for (int g0 = 0; g0 < nf; g0 += NW / wpt)
		{
			const int g = g0 + wv / wpt;
			const bool live = g < nf;
			const int a = live ? (g < TAIL_LDS ? s_fl[g] : gfl[g]) : 0;
			const double2 zg = xrow[a];
			const float fx = (float)zg.x, fy = (float)zg.y;
			float ax = 0, ay = 0, bx = 0, by = 0;
			double cx = 0, cy = 0;
			for (int q = live ? ln : Np; q < Np; q += lanes)
			{
				const float2 w = spos32[q];
				const float dx = fx - w.x, dy = fy - w.y;
				const float iv = __builtin_amdgcn_rcpf(fmaf(dx, dx, dy * dy));
				ax = fmaf(dx, iv, ax);
				ay = fmaf(dy, iv, ay);
			}
			for (int q = live ? ln : R; q < R; q += lanes)
			{
				const double2 wd = xrow[q];
				const float dx = fx - (float)wd.x, dy = fy - (float)wd.y;
				const float d2 = fmaf(dx, dx, dy * dy);
				if (d2 < cluster_d2 && q != a)
				{
					const double ex = zg.x - wd.x, ey = zg.y - wd.y;
					const double iv = rcp64(fmax(fma(ex, ex, ey * ey), 1e-300));
					cx = fma(ex, iv, cx);
					cy = fma(ey, iv, cy);
					nclus++;
				}
				else
				{
					// the own root gives dx = dy = 0 and, floored, a zero term
					const float iv = __builtin_amdgcn_rcpf(fmaxf(d2, 1e-30f));
					bx = fmaf(dx, iv, bx);
					by = fmaf(dy, iv, by);
				}
			}
			double vx = 2.0 * (double)ax - ((double)bx + cx), vy = 2.0 * (double)ay - ((double)by + cy);
			for (int o = 32; o >= 1; o >>= 1)
			{
				vx += __shfl_xor(vx, o, 64);
				vy += __shfl_xor(vy, o, 64);
			}
			if (wpt > 1)
			{
				// the waves of a root add their sums in wave order
				if (lane == 0)
				{
					s_red[wv][0] = vx;
					s_red[wv][1] = vy;
				}
				__syncthreads();
				if (lane == 0 && sub == 0)
					for (int k = 1; k < wpt; k++)
					{
						vx += s_red[wv + k][0];
						vy += s_red[wv + k][1];
					}
				__syncthreads();
			}
			// the whole Aberth sum: the root part (conjugated) and the terms formed per root
			if (lane == 0 && sub == 0 && live)
			{
				double2* e = g < TAIL_LDS ? &s_ex[g] : &gex[g];
				*e = make_double2(vx + e->x, -vy + e->y);
			}
		}
		__syncthreads();
		for (int g = threadIdx.x; g < nf; g += TAIL_BLOCK)
		{
			const bool lds = g < TAIL_LDS;
			const int a = lds ? s_fl[g] : gfl[g];
			const double2 v0 = lds ? s_f0[g] : gf0[g], v1 = lds ? s_f1[g] : gf1[g], ve = lds ? s_ex[g] : gex[g];
			const double2 zz = xrow[a];
			const cd z(zz.x, zz.y), f0(v0.x, v0.y), f1(v1.x, v1.y);
			const cd sum(ve.x, ve.y);
			const cd zn = z - f0 / (f1 + f0 * sum);
			xrow[a] = make_double2(zn.re, zn.im);
			if ((zn - z).abs() / L.theta < CONV_TOL) nconv++;
			else
			{
				const int k = atomicAdd(&s_nn, 1);
				if (k < TAIL_LDS) s_a[p ^ 1][k] = a;
				else gnxt[k] = a;
			}
		}
		__syncthreads();
		n = s_nn;
		p ^= 1;
		int* const t = gcur;
		gcur = gnxt;
		gnxt = t;
		__syncthreads();
		if (threadIdx.x == 0) s_nf = s_nn = s_nd = 0;
		__syncthreads();
	}
	// the row is finished (or out of sweeps): its list is empty for any
	// sweep kernel that follows
	if (threadIdx.x == 0) V.cnt[r] = 0;
	if (threadIdx.x != 0) nsweeps = 0;
	const int which[8] = {C_TAIL_FULL, C_TAIL_EARLY, C_CONV, C_TAIL_NEAR, C_TAIL_NEWTON, C_CLUSTER, C_TAIL_SWEEPS, C_TAIL_SWEEPS};
	unsigned long long v[8] = {nfull, nearly, nconv, nnear, nnewton, nclus, nsweeps, 0};
	block_stats<8, TAIL_BLOCK>(stats, which, v);
}

// initial roots of a phase: row centre+sgn*j from the rows of the previous
// steps (prepare_roots_kernel, ccf_functions.cuh:261-295; order > 0
// extrapolates), every root active, the sweep counters cleared
__global__ void k_prepare(double2* roots, RowMap M, SweepBufs B, int order, const int* __restrict__ perm)
{
	const int r = blockIdx.y;
	const int R = M.R;
	const int a = blockIdx.x * blockDim.x + threadIdx.x;
	int row;
	double phi;
	M.get(r, row, phi);
	if (a < R)
	{
		if (M.j > 0)
		{
			const int sgn = (r % M.rows_per_branch) == 0 ? -1 : 1;
			const int q = min(order, M.j - 1); // previous steps beyond j - 1 that exist
			roots[(size_t)row * R + a] = predict_root(roots + (size_t)(row - sgn) * R + a, (long long)sgn * R, q);
		}
		B.act[(size_t)r * R + a] = perm ? perm[a] : a;
	}
	if (a == 0)
	{
		B.cnt[r] = R;
		B.cnt[B.rows + r] = 0;
		B.fcnt[r] = 0;
		B.fcnt[B.rows + r] = 0;
		if (r == 0)
		{
			B.plan[2] = 0;
			B.plan[3] = 0;
		}
	}
}

// 1/mu error bound at every point of the solved rows, reduced to its max
// with a non-finite flag (find_errors_kernel + has_nan_err_kernel +
// thrust::max_element, ccf_functions.cuh:419-489). With caus / mu set, the
// same evaluation also gives the caustic point and the length scale
// (find_caustics_kernel, find_mu_length_scales_kernel), written in file
// order: root a, phi row r at a * rows + r.
__global__ __launch_bounds__(BLOCK) void k_errors(LensView L, const double2* __restrict__ roots, int R, RowMap M, int rows,
	const int* __restrict__ perm, unsigned long long* stats, double2* __restrict__ caus, double* __restrict__ mu)
{
	// grid: (xcd_round(blocks per row), rows of M); a row's points in the
	// roots' processing order; rows: the phi rows of the output
	const int2 sl = xcd_slot();
	int row;
	double phi;
	M.get(sl.y, row, phi);
	const int k = xcd_x(sl.x, (R + BLOCK - 1) / BLOCK) * BLOCK + threadIdx.x;
	const long long idx = (long long)row * R + (k < R ? perm[k] : 0);
	double e = 0, res = 0;
	unsigned long long bad = 0;
	cd z;
	int node = 0;
	if (k < R)
	{
		const double2 v = roots[idx];
		z = cd(v.x, v.y);
		node = nearest_node(L, z);
	}
	// a point outside the tree sums every star: its whole wave forms the
	// sums (wave_near), one such point at a time
	const bool outside = k < R && L.nf_off[node] < 0;
	const unsigned long long mo = __ballot(outside);
	Near nr;
	for (unsigned long long m = mo; m; m &= m - 1)
	{
		const int src = __ffsll(m) - 1;
		const Near t = wave_near(L, __shfl(node, src, 64), cd(__shfl(z.re, src, 64), __shfl(z.im, src, 64)));
		if ((int)(threadIdx.x & 63) == src) nr = t;
	}
	if (k < R)
	{
		{
			if (!outside) nr = p2p(L, node, z);
			double cp, sp;
			M.phase(row, phi, cp, sp);
			Witt w = witt_from(L, z, cp, sp, node, nr);
			e = err_bound(w);
			res = w.f0.abs();
			if (!isfinite(e)) { bad = 1; e = 0; res = 0; }
			if (caus || mu)
			{
				cd ws;
				double m;
				lens_map_from(L, z, w.nr, w.fr, w.d1, w.d2, ws, m);
				const size_t o = (size_t)(idx - (long long)row * R) * rows + row;
				if (caus) caus[o] = make_double2(ws.re, ws.im);
				if (mu) mu[o] = m;
			}
		}
	}
	for (int d = 32; d >= 1; d >>= 1)
	{
		e = fmax(e, __shfl_xor(e, d, 64));
		res = fmax(res, __shfl_xor(res, d, 64));
	}
	bad = wave_sum(bad);
	if ((threadIdx.x & 63) == 0)
	{
		atomicMax(stat_slot(stats, C_ERRBITS), (unsigned long long)__double_as_longlong(e));
		atomicMax(stat_slot(stats, C_RESID), (unsigned long long)__double_as_longlong(res));
		if (bad) atomicOr(stat_slot(stats, C_NAN), 1ull);
		if (mo) atomicAdd(stat_slot(stats, C_OUTSIDE), (unsigned long long)__popcll(mo));
	}
}

// [rows][R] -> [R][rows] through an LDS tile (transpose_array_kernel)
__global__ void k_transpose(const double2* __restrict__ in, int rows, int R, double2* __restrict__ out)
{
	__shared__ double2 t[32][33];
	const int a0 = blockIdx.x * 32, r0 = blockIdx.y * 32;
	const int tx = threadIdx.x & 31, ty = threadIdx.x >> 5; // 256 threads: 32 x 8
	for (int y = ty; y < 32; y += 8)
	{
		const int r = r0 + y, a = a0 + tx;
		if (r < rows && a < R) t[y][tx] = in[(size_t)r * R + a];
	}
	__syncthreads();
	for (int y = ty; y < 32; y += 8)
	{
		const int a = a0 + y, r = r0 + tx;
		if (r < rows && a < R) out[(size_t)a * rows + r] = t[tx][y];
	}
}

// lens map and length scale at every point (find_caustics_kernel,
// find_mu_length_scales_kernel; ccf_functions.cuh:510-563)
__global__ void k_caustics(LensView L, const double2* __restrict__ z, long long n, double2* w, double* mu)
{
	// file order is root-major and roots are seeded in the stars' tree order:
	// XCD k takes the k-th eighth of the blocks (gridDim.x a multiple of NUM_XCD)
	const long long per = gridDim.x / NUM_XCD;
	const long long blk = (long long)(blockIdx.x % NUM_XCD) * per + blockIdx.x / NUM_XCD;
	const long long i = blk * blockDim.x + threadIdx.x;
	if (i >= n) return;
	cd ws;
	double m;
	lens_map(L, cd(z[i].x, z[i].y), ws, m);
	if (w) w[i] = make_double2(ws.re, ws.im);
	if (mu) mu[i] = m;
}

__global__ void k_eval(LensView L, const double2* __restrict__ z, int n, double phi, double* out)
{
	const int i = blockIdx.x * blockDim.x + threadIdx.x;
	if (i >= n) return;
	const cd zz(z[i].x, z[i].y);
	Witt w = witt_f0(L, zz, phi);
	const cd f1 = witt_f1(L, zz, w);
	cd ws;
	double m;
	lens_map(L, zz, ws, m);
	double* o = out + 8 * (size_t)i;
	o[0] = w.f0.re; o[1] = w.f0.im;
	o[2] = f1.re; o[3] = f1.im;
	o[4] = ws.re; o[5] = ws.im;
	o[6] = m;
	o[7] = err_bound(w);
}

// star positions and masses (generate_star_field_kernel, star.cuh:64-107) from
// a counter-based Philox stream: star i uses subsequence i of the seed
__global__ void k_generate_stars(int n, unsigned long long seed, int rect, double cre, double cim, double cabs, Imf imf,
	double2* pos, double* mass)
{
	const int i = blockIdx.x * blockDim.x + threadIdx.x;
	if (i >= n) return;
	hiprandStatePhilox4_32_10_t s;
	hiprand_init(seed, i, 0, &s);
	double x1, x2;
	if (rect)
	{
		x1 = hiprand_uniform_double(&s) * 2 * cre - cre;
		x2 = hiprand_uniform_double(&s) * 2 * cim - cim;
	}
	else
	{
		const double a = hiprand_uniform_double(&s) * 2 * PI;
		const double r = sqrt(hiprand_uniform_double(&s)) * cabs;
		x1 = r * cos(a);
		x2 = r * sin(a);
	}
	pos[i] = make_double2(x1, x2);
	mass[i] = imf_mass(imf, hiprand_uniform_double(&s));
}

// ---------------------------------------------------------------------------

struct SumsLaunch
{
	int grid = 0;
	int target = 0;
	int rpt = 2;
};

SumsLaunch sums_launch(ccf_ctx* ctx, bool f32)
{
	SumsLaunch l;
	l.rpt = ctx->P.roots_per_thread > 0 ? ctx->P.roots_per_thread : 2;
	int per_cu = 0;
	const void* fn;
	if (f32) fn = l.rpt == 1 ? (const void*)k_sums32<1> : (l.rpt == 4 ? (const void*)k_sums32<4> : (const void*)k_sums32<2>);
	else fn = l.rpt == 1 ? (const void*)k_sums64<1> : (l.rpt == 4 ? (const void*)k_sums64<4> : (const void*)k_sums64<2>);
	if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, fn, BLOCK, 0) != hipSuccess || per_cu < 1) per_cu = 2;
	const int cus = ctx->num_cus > 0 ? ctx->num_cus : 256;
	l.grid = cus * per_cu;
	l.target = 4 * l.grid;
	return l;
}

// `cap` sweeps over the rows of M (ccf.cuh:929-939 and 1013-1018): the
// sweep kernels exit at once when no root of a row is active any more
// whether to read the active counts back after `done` of `cap` sweeps: from
// sweep `first` (the sweeps the previous phi step needed, 3 at first) after
// each of three sweeps, then every 4th. A read-back waits for the device,
// about as long as a few sweeps of stragglers take.
inline bool sweep_check(int done, int cap, int first)
{
	if (done >= cap || done < first) return false;
	return done < first + 3 || done % 4 == 0;
}

// pinned host slot of the root-sum FMM row counts of step j's first two
// sweeps for a group of branches (two slots per group, by the parity of j)
inline int* fmm_lag_slot(ccf_ctx* ctx, int group, int j)
{
	return ctx->host_counts + ctx->max_rows + 2 * CONT_GROUPS + 4 * group + 2 * (j & 1);
}

// pinned host slot of a read-back's two root-sum FMM row counts for a group
inline int* fmm_read_slot(ccf_ctx* ctx, int group) { return ctx->host_counts + ctx->max_rows + 2 * group; }

// One group of rows of a phase, solved on its own stream with its own sweep
// buffers: the rows of M (nrows of them).
struct SweepGroup
{
	SweepBufs* B;
	RowMap M;
	int nrows;
	hipStream_t s;
	unsigned fmm_mask;  // sweeps that may take the root sums of dense rows from the FMM pass
	int id = 0;         // its host slots (FMM row counts) and events, one set per group of branches
	int row0 = 0;       // first slot of its rows' active counts in a read-back
	bool done = false;  // no root active, or handed to the tail launch
	bool fmm_any = false;
};

// `cap` sweeps over the rows of every group (ccf.cuh:929-939 and 1013-1018).
// Each sweep queues the groups' launches on their own streams, so that one
// group's short, latency-bound launches run beside the others'. The first
// sweeps may take the root sums of dense rows from the FMM pass (when the
// buffers carry a root grid). A read-back (after sweep_check's sweeps)
// waits for every group still running and stops the groups whose rows have
// no active root. With tail_at, the sweeps after the first tail_at of each
// group run in one tail launch, with no read-back; before_tail(g), when
// set, runs once group g's first tail_at sweeps are queued.
bool run_sweeps_groups(ccf_ctx* ctx, SweepGroup* gs, int G, double2* roots, int cap, int first_check = 3, int tail_at = 0,
	const std::function<bool(int)>& before_tail = {})
{
	ctx->phase_sweeps = cap;
	const LensView L = ctx->lens();
	const bool f32 = ctx->P.precision == 1;
	const bool fmm_pole = f32 && ctx->tree.order > 0;
	const int circ_exact = (!ctx->rectangular && !ctx->P.approx) ? 1 : 0;
	const int Np = fmm_pole ? 0 : ctx->num_stars;
	const SumsLaunch sl = sums_launch(ctx, f32);
	ctx->fmm_rows_sweep[0] = ctx->fmm_rows_sweep[1] = -1; // unknown until a read-back
	int rows_total = 0, left = 0;
	for (int g = 0; g < G; g++)
	{
		if (gs[g].id < 0 || gs[g].id >= CONT_GROUPS)
		{
			ctx->fail("Error. sweep group out of range\n");
			return false;
		}
		left += !gs[g].done;
		if (gs[g].B->tile != BLOCK * sl.rpt)
		{
			ctx->fail("Error. sweep buffers were sized for another tile\n");
			return false;
		}
		gs[g].row0 = rows_total;
		rows_total += gs[g].nrows;
	}
	const bool readback = ctx->host_counts && rows_total <= ctx->max_rows;
	int swept = 0, pend = -1, ended = -1;
	// the counts read back after sweep it (pend): groups with no active root
	// stop, and the FMM pass stops once no row has as many active roots as it
	// takes
	const auto take_counts = [&](int it) -> bool {
		for (int g = 0; g < G; g++)
			if (!gs[g].done && !CCF_HIP(ctx, hipEventSynchronize(ctx->read_event[gs[g].id]))) return false;
		for (int g = 0; g < G; g++)
		{
			SweepGroup& q = gs[g];
			if (q.done) continue;
			long long active = 0;
			int amax = 0;
			for (int r = 0; r < q.nrows; r++)
			{
				active += ctx->host_counts[q.row0 + r];
				amax = std::max(amax, ctx->host_counts[q.row0 + r]);
			}
			// a row's listed roots are among its active ones
			if (q.B->grid.order > 0 && amax < root_fmm_min(q.M.R)) q.fmm_mask = 0;
			if (q.id == 0)
			{
				ctx->fmm_rows_sweep[0] = fmm_read_slot(ctx, 0)[0];
				ctx->fmm_rows_sweep[1] = fmm_read_slot(ctx, 0)[1];
			}
			if (ctx->sweep_verbose >= 3)
				std::cout << "sweep " << it + 1 << (G > 1 ? ", group " + std::to_string(g) : std::string()) << ": " << active
						  << " roots active, at most " << amax << " in a row\n";
			if (active == 0)
			{
				q.done = true;
				if (--left == 0) ended = it + 1;
			}
		}
		return true;
	};
	for (int it = 0; it < cap && left > 0; it++)
	{
		swept = it + 1;
		for (int g = 0; g < G; g++)
		{
			SweepGroup& q = gs[g];
			if (q.done) continue;
			SweepBufs& B = *q.B;
			const RowMap& M = q.M;
			const int nrows = q.nrows, R = M.R;
			hipStream_t s = q.s;
			const dim3 g_rows(grid(R), nrows);
			const dim3 g_witt(xcd_round(grid(R)), nrows);
			const SweepView V = view(B, it & 1);
			const bool fmm = B.grid.order > 0 && it < 32 && ((q.fmm_mask >> it) & 1u);
			q.fmm_any = q.fmm_any || fmm;
			const PlanArgs plan{nrows, Np + R, sl.target, fmm ? root_fmm_min(R) : 0, it};
			ccf::Timer tw;
			if (ctx->P.timing)
			{
				tw.a = ctx->next_event();
				tw.b = ctx->next_event();
				if (tw.a) (void)hipEventRecord(tw.a, s);
			}
			if (f32) k_witt<true><<<g_witt, BLOCK, 0, s>>>(L, roots, M, V, ctx->counters, circ_exact, fmm_pole ? 1 : 0, plan);
			else k_witt<false><<<g_witt, BLOCK, 0, s>>>(L, roots, M, V, ctx->counters, circ_exact, 0, plan);
			if (ctx->P.timing && tw.a && tw.b)
			{
				(void)hipEventRecord(tw.b, s);
				ctx->witt_timers.push_back(tw);
			}
			ccf::Timer tm;
			if (ctx->P.timing)
			{
				tm.a = ctx->next_event();
				tm.b = ctx->next_event();
				if (tm.a) (void)hipEventRecord(tm.a, s);
			}
			if (fmm) root_fmm_pass(B, roots, M, nrows, ctx->counters, s);
			if (f32)
			{
				const double d = CLUSTER_REL * ctx->root_half_length;
				const float cd2 = (float)(d * d);
				if (sl.rpt == 1) k_sums32<1><<<sl.grid, BLOCK, 0, s>>>(ctx->spos32, Np, roots, M, V, nrows, cd2, ctx->counters);
				else if (sl.rpt == 4) k_sums32<4><<<sl.grid, BLOCK, 0, s>>>(ctx->spos32, Np, roots, M, V, nrows, cd2, ctx->counters);
				else k_sums32<2><<<sl.grid, BLOCK, 0, s>>>(ctx->spos32, Np, roots, M, V, nrows, cd2, ctx->counters);
			}
			else
			{
				if (sl.rpt == 1) k_sums64<1><<<sl.grid, BLOCK, 0, s>>>(ctx->spos, Np, roots, M, V, nrows);
				else if (sl.rpt == 4) k_sums64<4><<<sl.grid, BLOCK, 0, s>>>(ctx->spos, Np, roots, M, V, nrows);
				else k_sums64<2><<<sl.grid, BLOCK, 0, s>>>(ctx->spos, Np, roots, M, V, nrows);
			}
			if (ctx->P.timing && tm.a && tm.b)
			{
				(void)hipEventRecord(tm.b, s);
				ctx->timers.push_back(tm);
			}
			// mixed form, continuation: the sweeps after the first tail_at of
			// the step in one launch, with no read-back, which begins with
			// this sweep's updates. The rows the root-sum FMM took in the
			// step's first two sweeps go to the host for the choice two steps
			// later.
			const bool tail_now = tail_at > 0 && it + 1 == tail_at && it + 1 < cap;
			if (!tail_now) k_finalize<<<g_rows, BLOCK, 0, s>>>(ctx->theta, roots, M, V, ctx->counters);
			if (!CCF_LAUNCH(ctx, "find_critical_curve_roots_kernel")) return false;
			ctx->info.sweeps++;
			if (tail_now)
			{
				if (ctx->host_counts && nrows <= ctx->max_rows)
				{
					if (q.fmm_any)
					{
						hipEvent_t& e = ctx->fmm_rows_event[2 * q.id + (M.j & 1)];
						if (!e && !CCF_HIP(ctx, hipEventCreateWithFlags(&e, hipEventDisableTiming))) return false;
						if (!CCF_HIP(ctx, hipMemcpyAsync(fmm_lag_slot(ctx, q.id, M.j), B.plan + 2, 2 * sizeof(int),
								hipMemcpyDeviceToHost, s)) ||
							!CCF_HIP(ctx, hipEventRecord(e, s)))
							return false;
					}
					else
					{
						// no pass ran: its row counts are 0 (the slot's previous
						// copy, two steps back, was waited for when this step was
						// planned)
						int* slot = fmm_lag_slot(ctx, q.id, M.j);
						slot[0] = slot[1] = 0;
					}
				}
				if (before_tail && !before_tail(q.id)) return false;
				const SweepView Vn = view(B, (it + 1) & 1);
				const double d = CLUSTER_REL * ctx->root_half_length;
				ccf::Timer tt;
				if (ctx->P.timing)
				{
					tt.a = ctx->next_event();
					tt.b = ctx->next_event();
					if (tt.a && tt.b) (void)hipEventRecord(tt.a, s);
				}
				k_tail<<<nrows, TAIL_BLOCK, 0, s>>>(L, roots, M, Vn, ctx->counters, circ_exact, fmm_pole ? 1 : 0, cap - (it + 1),
					(float)(d * d), ctx->spos32, Np, ctx->theta);
				if (ctx->P.timing && tt.a && tt.b)
				{
					(void)hipEventRecord(tt.b, s);
					ctx->tail_timers.push_back(tt);
				}
				if (!CCF_LAUNCH(ctx, "find_critical_curve_roots_kernel")) return false;
				ctx->info.tail_launches++;
				q.done = true;
				left--;
			}
		}
		// stop the groups whose rows have no active root: the remaining
		// sweeps up to the cap would do nothing (checked at a few sweeps
		// only). A check's counts come back while the next sweep runs: they
		// are copied after their sweep, and waited for once the next sweep is
		// queued, so the device does not idle on the host's decision. The
		// sweep queued meanwhile after a group's last active root finds none
		// and changes nothing; a cleared FMM flag takes effect a sweep later,
		// when no row lists enough roots for the pass any more either.
		if (pend >= 0)
		{
			if (!take_counts(pend)) return false;
			pend = -1;
		}
		if (left > 0 && readback && (ctx->sweep_verbose >= 3 || sweep_check(it + 1, cap, first_check)))
		{
			for (int g = 0; g < G; g++)
			{
				SweepGroup& q = gs[g];
				if (q.done) continue;
				const SweepView V = view(*q.B, it & 1);
				hipEvent_t& e = ctx->read_event[q.id];
				if (!e && !CCF_HIP(ctx, hipEventCreateWithFlags(&e, hipEventDisableTiming))) return false;
				if (!CCF_HIP(ctx, hipMemcpyAsync(ctx->host_counts + q.row0, V.cnt_next, q.nrows * sizeof(int), hipMemcpyDeviceToHost, q.s)) ||
					!CCF_HIP(ctx, hipMemcpyAsync(fmm_read_slot(ctx, q.id), q.B->plan + 2, 2 * sizeof(int), hipMemcpyDeviceToHost, q.s)) ||
					!CCF_HIP(ctx, hipEventRecord(e, q.s)))
					return false;
			}
			pend = it;
		}
	}
	if (pend >= 0 && !take_counts(pend)) return false;
	ctx->phase_sweeps = ended > 0 ? ended : swept;
	return true;
}

// run_sweeps_groups for one group on `stream` (the context's by default),
// with the host slots of group `group`
bool run_sweeps(ccf_ctx* ctx, SweepBufs& B, double2* roots, const RowMap& M, int nrows, int cap, unsigned fmm_mask,
	int first_check = 3, int tail_at = 0, const std::function<bool()>& before_tail = {}, hipStream_t stream = nullptr,
	int group = 0)
{
	SweepGroup q{&B, M, nrows, stream ? stream : ctx->stream, fmm_mask, group};
	return run_sweeps_groups(ctx, &q, 1, roots, cap, first_check, tail_at,
		before_tail ? std::function<bool(int)>([&](int) { return before_tail(); }) : std::function<bool(int)>());
}

// The sweep buffers of `rows` rows of R roots (and, when the root-sum FMM
// applies, its grid over the field): take(p, count) provides each array of
// count elements, from the context's device pool or from a slice of the
// fixed scratch array. exact_items sizes the split-K partials for the plan
// plan_sums can make for these terms (at most MAX_SLICES slices of at least
// MIN_SLICE terms per tile, ceil(target / tiles) wanted); otherwise for any
// slice count up to the target.
template <class Take>
bool root_layout(ccf_ctx* ctx, SweepBufs& b, Take&& take);

template <class Take>
bool sweep_layout(ccf_ctx* ctx, SweepBufs& b, int rows, int R, bool exact_items, Take&& take)
{
	const bool f32 = ctx->P.precision == 1;
	const SumsLaunch sl = sums_launch(ctx, f32);
	b.rows = rows;
	b.R = R;
	b.tile = BLOCK * sl.rpt;
	// items per sweep <= max(tiles, target) + tiles
	const long long tiles_max = (long long)rows * ((R + b.tile - 1) / b.tile);
	b.max_items = (int)(tiles_max + sl.target + tiles_max);
	if (exact_items)
	{
		const int Np = (f32 && ctx->tree.order > 0) ? 0 : ctx->num_stars;
		const long long scap = std::min(MAX_SLICES, std::max(1, (Np + R) / MIN_SLICE));
		b.max_items = (int)(std::min(tiles_max * scap, tiles_max + sl.target) + tiles_max);
	}
	const size_t list = (size_t)rows * R;
	if (!(take(b.act, 2 * list) && take(b.cnt, 2 * (size_t)rows) && take(b.fl, list) && take(b.fcnt, 2 * (size_t)rows) &&
			take(b.f0, list) && take(b.f1, list) && take(b.ex, list) &&
			take(b.tile_off, (size_t)rows + 1) && take(b.plan, 4) && take(b.ticket, 1) &&
			take(b.part, (size_t)b.max_items * b.tile)))
		return false;
	if (ctx->P.root_fmm != 1 || R < ROOT_FMM_MIN_ROOTS) return true;
	return root_layout(ctx, b, take);
}

// the root-sum FMM's arrays of b.rows rows of b.R roots and its grid over
// the field (with a tenth of margin, as the star tree's root node,
// ccf.cuh:738-746)
template <class Take>
bool root_layout(ccf_ctx* ctx, SweepBufs& b, Take&& take)
{
	const int rows = b.rows, R = b.R;
	const size_t list = (size_t)rows * R;
	double hx, hy;
	if (ctx->rectangular)
	{
		hx = 1.1 * std::fabs(ctx->corner.re);
		hy = 1.1 * std::fabs(ctx->corner.im);
	}
	else
	{
		hx = hy = 1.1 * ctx->corner.abs();
	}
	const int order = root_order_compiled(ctx->P.root_fmm_order > 0 ? ctx->P.root_fmm_order : ROOT_FMM_ORDER);
	const double leaf = ctx->P.root_fmm_leaf > 0 ? ctx->P.root_fmm_leaf : ROOT_FMM_LEAF;
	b.grid = plan_root_grid(hx, hy, R, order, leaf);
	int lds = 0;
	if (!CCF_HIP(ctx, hipDeviceGetAttribute(&lds, hipDeviceAttributeMaxSharedMemoryPerBlock, 0))) return false;
	b.grid.bin_row = bin_row_lds(b.grid.nleaf, R) + BIN_STATIC_LDS <= (size_t)lds;
	const RootGrid& g = b.grid;
	const size_t ce = (size_t)rows * g.ncells * (g.order + 1);
	return take(b.slot, list) && take(b.rmode, (size_t)rows) && take(b.rsum, list) && take(b.ccnt, (size_t)rows * g.ncells) &&
		take(b.cstart, (size_t)rows * (g.nleaf + 1)) && take(b.ccur, (size_t)rows * g.nleaf) && take(b.cell, list) &&
		take(b.bidx, list) && take(b.sidx, list) && take(b.spos, list) && take(b.ocnt, (size_t)rows) && take(b.oidx, list) &&
		take(b.mp, ce) && take(b.lcq, (size_t)rows * coarse_cells(g) * (g.order + 1));
}

// the root-sum FMM's arrays alone (ccf_root_sums), from the context's pool
bool alloc_root_fmm(ccf_ctx* ctx, SweepBufs& b)
{
	auto take = [ctx](auto*& p, size_t count) { return CCF_HIP(ctx, ctx->dmalloc(&p, count * sizeof(*p))); };
	return root_layout(ctx, b, take) &&
		CCF_HIP(ctx, hipMemsetAsync(b.slot, 0xff, (size_t)b.rows * b.R * sizeof(int), ctx->stream)) &&
		CCF_HIP(ctx, hipMemsetAsync(b.rmode, 0, b.rows * sizeof(int), ctx->stream)) &&
		CCF_HIP(ctx, hipStreamSynchronize(ctx->stream));
}

} // namespace

void free_sweep(ccf_ctx* ctx, SweepBufs& b)
{
	auto F = [ctx](auto*& p) { ctx->dfree(p); p = nullptr; };
	F(b.act); F(b.cnt); F(b.fl); F(b.fcnt); F(b.f0); F(b.f1); F(b.ex); F(b.tile_off); F(b.plan); F(b.ticket);
	F(b.part);
	F(b.slot); F(b.rmode); F(b.rsum); F(b.ccnt); F(b.cstart); F(b.ccur); F(b.cell); F(b.bidx); F(b.sidx); F(b.spos);
	F(b.ocnt); F(b.oidx); F(b.mp); F(b.lcq);
	b = SweepBufs{};
}

bool alloc_sweep(ccf_ctx* ctx, SweepBufs& b, int rows, int R)
{
	free_sweep(ctx, b);
	auto take = [ctx](auto*& p, size_t count) { return CCF_HIP(ctx, ctx->dmalloc(&p, count * sizeof(*p))); };
	if (!sweep_layout(ctx, b, rows, R, false, take)) return false;
	const size_t list = (size_t)rows * R;
	if (!CCF_HIP(ctx, hipMemsetAsync(b.cnt, 0, 2 * rows * sizeof(int), ctx->stream)) ||
		!CCF_HIP(ctx, hipMemsetAsync(b.fcnt, 0, 2 * rows * sizeof(int), ctx->stream)) ||
		!CCF_HIP(ctx, hipMemsetAsync(b.ticket, 0, sizeof(int), ctx->stream)))
		return false;
	if (!b.slot) return true;
	return CCF_HIP(ctx, hipMemsetAsync(b.slot, 0xff, list * sizeof(int), ctx->stream)) &&
		CCF_HIP(ctx, hipMemsetAsync(b.rmode, 0, rows * sizeof(int), ctx->stream)) &&
		CCF_HIP(ctx, hipStreamSynchronize(ctx->stream));
}

// the seeds in the rows of branch centres c0 .. c0 + gridDim.y - 1
// (ccf.cuh:711-726: every centre starts from the same seeds)
__global__ void k_seed_rows(double2* __restrict__ roots, const double2* __restrict__ seeds, int R, int np, int nb, int c0)
{
	const int a = blockIdx.x * blockDim.x + threadIdx.x, c = c0 + (int)blockIdx.y;
	if (a < R) roots[(size_t)(np / (2 * nb) + c * np / nb + c) * R + a] = seeds[a];
}

bool seed_rows(ccf_ctx* ctx, int c0, int n, hipStream_t s)
{
	const int R = ctx->num_roots;
	k_seed_rows<<<dim3(grid(R), n), BLOCK, 0, s>>>(ctx->ccs_init, ctx->dseeds, R, ctx->P.num_phi, ctx->P.num_branches, c0);
	return CCF_LAUNCH(ctx, "seeds");
}

bool launch_phase(ccf_ctx* ctx)
{
	const int rows = ctx->P.num_phi + ctx->P.num_branches;
	k_phase<<<grid(rows), BLOCK, 0, ctx->stream>>>(ctx->P.num_phi, ctx->P.num_branches, ctx->phase);
	return CCF_LAUNCH(ctx, "phase table");
}

// groups of branches solved side by side: the mixed form's (at most
// CONT_GROUPS, at most one per branch); the FP64 form reads its sweeps back
// and runs one
inline int group_count(const ccf_ctx* ctx, bool init)
{
	return ctx->P.precision == 1 ? std::max(1, std::min(init ? CONT_GROUPS : CONT_STEP_GROUPS, ctx->P.num_branches)) : 1;
}

// the groups a solve runs: group_count's, if their buffers were made for
// the parameters the context has now (ccf_load makes them), else one
inline int solve_groups(const ccf_ctx* ctx, bool init)
{
	const int G = group_count(ctx, init);
	for (int g = 1; g < G; g++)
		if (!ctx->swg[g - 1].act) return 1;
	return G;
}

// group g's sweep buffers and stream: the context's own for group 0
inline SweepBufs& group_bufs(ccf_ctx* ctx, int g) { return g == 0 ? ctx->sw : ctx->swg[g - 1]; }
inline hipStream_t group_stream(ccf_ctx* ctx, int g) { return g == 0 ? ctx->stream : ctx->group_streams[g - 1]; }

// the sweep buffers: the first group's (also every phase that runs one
// group) for all rows of a step, each other group's for its 2 (c1 - c0)
// rows, its branches c0 .. c1 - 1
bool alloc_groups(ccf_ctx* ctx)
{
	const int R = ctx->num_roots, nb = ctx->P.num_branches, Gi = group_count(ctx, true), Gc = group_count(ctx, false);
	if (!alloc_sweep(ctx, ctx->sw, ctx->max_rows, R)) return false;
	for (int g = 1; g < std::max(Gi, Gc); g++)
	{
		const int ri = g < Gi ? (g + 1) * nb / Gi - g * nb / Gi : 0, rc = g < Gc ? 2 * ((g + 1) * nb / Gc - g * nb / Gc) : 0;
		if (!alloc_sweep(ctx, ctx->swg[g - 1], std::max(ri, rc), R)) return false;
	}
	return true;
}

bool alloc_solver(ccf_ctx* ctx)
{
	const int R = ctx->num_roots, np = ctx->P.num_phi, nb = ctx->P.num_branches;
	const size_t pts = (size_t)(np + nb) * R;
	ctx->max_rows = 2 * nb;
	if (!CCF_HIP(ctx, ctx->dmalloc(&ctx->dorder, (size_t)R * sizeof(int))) ||
		!CCF_HIP(ctx, hipMemcpyAsync(ctx->dorder, ctx->order.data(), (size_t)R * sizeof(int), hipMemcpyHostToDevice, ctx->stream)))
		return false;
	if (ctx->batch_member)
		return CCF_HIP(ctx, ctx->dmalloc(&ctx->ccs_init, pts * sizeof(double2))) &&
			CCF_HIP(ctx, ctx->dmalloc(&ctx->phase, (size_t)(np + nb) * sizeof(double2))) && launch_phase(ctx) &&
			CCF_HIP(ctx, ctx->dmalloc(&ctx->dseeds, (size_t)R * sizeof(double2))) &&
			CCF_HIP(ctx, hipMemcpyAsync(ctx->dseeds, ctx->seeds.data(), (size_t)R * sizeof(double2), hipMemcpyHostToDevice, ctx->stream)) &&
			CCF_HIP(ctx, hipMemsetAsync(ctx->ccs_init, 0, pts * sizeof(double2), ctx->stream)) &&
			CCF_HIP(ctx, hipStreamSynchronize(ctx->stream));
	return CCF_HIP(ctx, ctx->dmalloc(&ctx->dseeds, (size_t)R * sizeof(double2))) &&
		CCF_HIP(ctx, hipMemcpyAsync(ctx->dseeds, ctx->seeds.data(), (size_t)R * sizeof(double2), hipMemcpyHostToDevice, ctx->stream)) &&
		CCF_HIP(ctx, ctx->dmalloc(&ctx->ccs_init, pts * sizeof(double2))) && alloc_groups(ctx) &&
		CCF_HIP(ctx, hipHostMalloc((void**)&ctx->host_counts, (ctx->max_rows + 6 * CONT_GROUPS) * sizeof(int), hipHostMallocDefault)) &&
		CCF_HIP(ctx, ctx->dmalloc(&ctx->phase, (size_t)(np + nb) * sizeof(double2))) && launch_phase(ctx) &&
		CCF_HIP(ctx, ctx->dmalloc(&ctx->counters, (STAT_WORDS + C_COUNT) * sizeof(unsigned long long))) &&
		CCF_HIP(ctx, hipMemsetAsync(ctx->counters, 0, STAT_WORDS * sizeof(unsigned long long), ctx->stream));
}

// the streams of groups 1 .. G - 1 (created once, default priority) start
// after everything queued on the context's stream so far
bool group_streams_start(ccf_ctx* ctx, int G)
{
	if (!ctx->group_event && !CCF_HIP(ctx, hipEventCreateWithFlags(&ctx->group_event, hipEventDisableTiming))) return false;
	if (!CCF_HIP(ctx, hipEventRecord(ctx->group_event, ctx->stream))) return false;
	for (int g = 1; g < G; g++)
	{
		hipStream_t& q = ctx->group_streams[g - 1];
		if (!q && !CCF_HIP(ctx, hipStreamCreateWithFlags(&q, hipStreamNonBlocking))) return false;
		if (!CCF_HIP(ctx, hipStreamWaitEvent(q, ctx->group_event, 0))) return false;
	}
	return true;
}

// what follows on the context's stream waits for the other groups' work
bool group_streams_join(ccf_ctx* ctx, int G)
{
	for (int g = 1; g < G; g++)
		if (!CCF_HIP(ctx, hipEventRecord(ctx->group_event, ctx->group_streams[g - 1])) ||
			!CCF_HIP(ctx, hipStreamWaitEvent(ctx->stream, ctx->group_event, 0)))
			return false;
	return true;
}

// params.autotune: the context's first solve times, on the initial solve of
// the first group's branch centres (the rows, sweeps and FMM passes of the
// real one, from the seeds), the automatic choices of the sums' roots per
// thread and then, in the mixed form with the root-sum FMM, of its (order,
// mean roots per leaf); the fastest stay for the context's later runs.
// None of them moves a fixed point: the tile only splits the sums, the
// FMM's truncation only steers the iteration. The trial solves leave the
// run's counters and statistics as they found them.
bool autotune_dev(ccf_ctx* ctx, int verbose)
{
	const int R = ctx->num_roots, np = ctx->P.num_phi, nb = ctx->P.num_branches;
	hipStream_t s = ctx->stream;
	const ccf_params& Pin = ctx->P_in;
	const bool tune_rpt = Pin.roots_per_thread == 0;
	const bool tune_fmm = ctx->P.precision == 1 && ctx->P.root_fmm == 1 && Pin.root_fmm_order == 0 && Pin.root_fmm_leaf == 0 &&
		R >= ROOT_FMM_MIN_ROOTS;
	auto t0 = std::chrono::steady_clock::now();
	const int G = solve_groups(ctx, true);
	const int nbg = nb / G;
	const RowMap M{R, np, nb, 0, 1, 0.0, ctx->phase, 0};
	const int cap = ctx->P.init_iters > 0 ? ctx->P.init_iters : 60;
	const ccf_info keep = ctx->info;
	const int timing = ctx->P.timing;
	ctx->P.timing = 0;
	hipEvent_t e0 = nullptr, e1 = nullptr;
	bool ok = CCF_HIP(ctx, hipEventCreate(&e0)) && CCF_HIP(ctx, hipEventCreate(&e1));
	auto trial = [&](int rpt, int order, int leaf, float& ms) -> bool {
		ctx->P.roots_per_thread = rpt;
		ctx->P.root_fmm_order = order;
		ctx->P.root_fmm_leaf = leaf;
		SweepBufs B;
		bool good = alloc_sweep(ctx, B, nbg, R) && seed_rows(ctx, 0, nbg, s);
		if (good)
		{
			k_prepare<<<dim3(grid(R), nbg), BLOCK, 0, s>>>(ctx->ccs_init, M, B, 0, ctx->dorder);
			SweepGroup q{&B, M, nbg, s, (1u << ROOT_FMM_INIT_SWEEPS) - 1u, 0};
			good = CCF_LAUNCH(ctx, "prepare_roots_kernel") && CCF_HIP(ctx, hipEventRecord(e0, s)) &&
				run_sweeps_groups(ctx, &q, 1, ctx->ccs_init, cap) && CCF_HIP(ctx, hipEventRecord(e1, s)) &&
				CCF_HIP(ctx, hipEventSynchronize(e1)) && CCF_HIP(ctx, hipEventElapsedTime(&ms, e0, e1));
		}
		free_sweep(ctx, B);
		if (good && verbose >= 2)
			std::cout << "autotune: roots_per_thread " << rpt << ", root-sum FMM order " << order << " leaf " << leaf << ": " << ms
					  << " ms\n";
		return good;
	};
	int best_rpt = sums_launch(ctx, ctx->P.precision == 1).rpt;
	int best_order = ctx->P.root_fmm_order > 0 ? ctx->P.root_fmm_order : ROOT_FMM_ORDER;
	int best_leaf = ctx->P.root_fmm_leaf > 0 ? ctx->P.root_fmm_leaf : ROOT_FMM_LEAF;
	float best = 1e30f;
	if (ok && tune_rpt)
		for (int rpt : {1, 2, 4})
		{
			float ms = 0;
			if (!(ok = trial(rpt, best_order, best_leaf, ms))) break;
			if (ms < best)
			{
				best = ms;
				best_rpt = rpt;
			}
		}
	if (ok && tune_fmm)
	{
		// the defaults first; lower orders and larger cells trade far-field
		// terms for near-field ones
		const int cand[][2] = {{ROOT_FMM_ORDER, ROOT_FMM_LEAF}, {6, 6}, {10, 6}, {8, 10}, {6, 10}, {8, 4}};
		best = 1e30f;
		for (const auto& c : cand)
		{
			float ms = 0;
			if (!(ok = trial(best_rpt, c[0], c[1], ms))) break;
			if (ms < best)
			{
				best = ms;
				best_order = c[0];
				best_leaf = c[1];
			}
		}
	}
	if (e0) (void)hipEventDestroy(e0);
	if (e1) (void)hipEventDestroy(e1);
	ctx->P.timing = timing;
	if (!ok) return false;
	// the tuned choices, and the solver's buffers laid out for them
	ctx->P.roots_per_thread = best_rpt;
	ctx->P.root_fmm_order = best_order;
	ctx->P.root_fmm_leaf = best_leaf;
	ctx->tuned = true;
	ctx->tuned_rpt = best_rpt;
	ctx->tuned_root_order = tune_fmm ? best_order : 0;
	ctx->tuned_root_leaf = tune_fmm ? best_leaf : 0;
	ctx->info = keep;
	ok = alloc_groups(ctx) &&
		CCF_HIP(ctx, hipMemsetAsync(ctx->counters, 0, STAT_WORDS * sizeof(unsigned long long), s)) &&
		CCF_HIP(ctx, hipStreamSynchronize(s));
	ctx->info.t_autotune = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
	if (verbose >= 1)
		std::cout << "Autotuned: roots_per_thread " << best_rpt << (tune_fmm ? ", root-sum FMM order " + std::to_string(best_order) +
			" leaf " + std::to_string(best_leaf) : std::string()) << " (" << ctx->info.t_autotune << " s)\n";
	return ok;
}

bool find_initial_roots_dev(ccf_ctx* ctx, int verbose)
{
	const int R = ctx->num_roots, np = ctx->P.num_phi, nb = ctx->P.num_branches;
	hipStream_t s = ctx->stream;
	if (ctx->P.autotune && !ctx->tuned && !autotune_dev(ctx, verbose)) return false;
	if (verbose >= 1) std::cout << "Finding initial roots...\n";
	auto t0 = std::chrono::steady_clock::now();
	// seeds at every branch centre (ccf.cuh:711-726)
	if (!seed_rows(ctx, 0, nb, s)) return false;
	ctx->sweep_verbose = verbose;
	const int cap = ctx->P.init_iters > 0 ? ctx->P.init_iters : 60;
	// the mixed form solves its branch centres in groups on their own
	// streams, sweep by sweep side by side (the root-sum FMM pass of a dense
	// sweep is a chain of short launches); each group's rows are independent
	const int G = solve_groups(ctx, true);
	if (G > 1 && !group_streams_start(ctx, G)) return false;
	SweepGroup gs[CONT_GROUPS];
	for (int g = 0; g < G; g++)
	{
		const int c0 = g * nb / G, nbg = (g + 1) * nb / G - c0;
		const RowMap M{R, np, nb, 0, 1, 0.0, ctx->phase, c0};
		gs[g] = SweepGroup{&group_bufs(ctx, g), M, nbg, group_stream(ctx, g), (1u << ROOT_FMM_INIT_SWEEPS) - 1u, g};
	}
	for (int g = 0; g < G; g++)
	{
		k_prepare<<<dim3(grid(R), gs[g].nrows), BLOCK, 0, gs[g].s>>>(ctx->ccs_init, gs[g].M, *gs[g].B, 0, ctx->dorder);
		if (!CCF_LAUNCH(ctx, "prepare_roots_kernel")) return false;
	}
	if (!run_sweeps_groups(ctx, gs, G, ctx->ccs_init, cap)) return false;
	if (G > 1 && !group_streams_join(ctx, G)) return false;
	if (!CCF_HIP(ctx, hipStreamSynchronize(s))) return false;
	auto t1 = std::chrono::steady_clock::now();
	ctx->info.t_init_roots = std::chrono::duration<double>(t1 - t0).count();
	if (verbose >= 1) std::cout << "\nDone finding initial roots. Elapsed time: " << ctx->info.t_init_roots << " seconds.\n";
	return true;
}

bool find_ccs_dev(ccf_ctx* ctx, int verbose)
{
	const int R = ctx->num_roots, np = ctx->P.num_phi, nb = ctx->P.num_branches;
	hipStream_t s = ctx->stream;
	const int steps = np / (2 * nb);
	const int cap = ctx->P.step_iters > 0 ? ctx->P.step_iters : 20;
	if (verbose >= 1) std::cout << "Finding critical curve positions...\n";
	auto t0 = std::chrono::steady_clock::now();
	// rows the root-sum FMM took in the previous step's first and second
	// sweep (-1: not read back); a sweep whose counterpart took none skips
	// the pass, whose launches cost more than its empty rows, and every
	// ROOT_FMM_PROBE-th step tries again
	int prev[2] = {-1, -1};
	// sweeps the previous step needed (FP64 form: its first read-back)
	const int first0 = 3;
	int first = first0;
	// the mixed form runs the sweeps after the first TAIL_AT of a step (the
	// first TAIL_AT + 1 while the predictor still has fewer than two earlier
	// steps) in the tail launch, without read-backs: the FMM choice of its
	// first sweeps uses the row counts of the step before the previous one,
	// copied back while the previous step ran (a fixed lag, so the choice
	// does not depend on timing)
	const bool tailed = ctx->P.precision == 1;
	ctx->sweep_verbose = verbose;

	// the error pass (max 1/mu error and the NaN abort, ccf.cuh:1039-1069)
	// takes each row set once its roots are final: the branch centres now,
	// every phi step's rows after the step, on a second stream beside the
	// later steps (whose tails and launch gaps leave most of the device
	// idle). Its statistics are integer sums and maxima and every point
	// writes its own outputs, so the results do not depend on the overlap.
	const LensView L = ctx->lens();
	// the caustic map and length scales the run will write come from the
	// same evaluation of the field; the caustics stage then has them
	const long long total = (long long)(np + nb) * R;
	const bool wc = ctx->P.write_caustics != 0, wm = ctx->P.write_mu_length_scales != 0;
	ctx->caustics_done = ctx->mu_done = false;
	if (wc && !ctx->caustics && !CCF_HIP(ctx, ctx->dmalloc(&ctx->caustics, (size_t)total * sizeof(double2)))) return false;
	if (wm && !ctx->mu_scales && !CCF_HIP(ctx, ctx->dmalloc(&ctx->mu_scales, (size_t)total * sizeof(double)))) return false;
	if (!ctx->err_stream && !CCF_HIP(ctx, hipStreamCreateWithFlags(&ctx->err_stream, hipStreamNonBlocking))) return false;
	if (!ctx->err_event && !CCF_HIP(ctx, hipEventCreateWithFlags(&ctx->err_event, hipEventDisableTiming))) return false;
	hipStream_t es = ctx->err_stream;
	auto errors = [&](const RowMap& Me, int nrows, hipStream_t after) -> bool {
		if (!CCF_HIP(ctx, hipEventRecord(ctx->err_event, after)) || !CCF_HIP(ctx, hipStreamWaitEvent(es, ctx->err_event, 0)))
			return false;
		k_errors<<<dim3(xcd_round(grid(R)), nrows), BLOCK, 0, es>>>(L, ctx->ccs_init, R, Me, np + nb, ctx->dorder, ctx->counters,
			wc ? ctx->caustics : nullptr, wm ? ctx->mu_scales : nullptr);
		return CCF_LAUNCH(ctx, "find_errors_kernel");
	};
	// The mixed form (no read-backs) runs its branches in CONT_STEP_GROUPS groups,
	// each on its own stream with its own sweep buffers: a group's steps are
	// a chain of short, latency-bound launches (the later sweeps, the tail of
	// one workgroup per row) that leaves most of the device idle, which the
	// other group's first sweep over all of its roots fills. The groups share
	// only read-only inputs, the statistics (integer atomics) and the error
	// pass's stream; each writes its own rows.
	const int G = tailed ? solve_groups(ctx, false) : 1;
	int gb[CONT_GROUPS + 1] = {};
	hipStream_t gs[CONT_GROUPS] = {};
	SweepBufs* gsw[CONT_GROUPS] = {};
	for (int g = 0; g <= G; g++) gb[g] = g * nb / G;
	if (G > 1 && !group_streams_start(ctx, G)) return false;
	for (int g = 0; g < G; g++)
	{
		gs[g] = group_stream(ctx, g);
		gsw[g] = &group_bufs(ctx, g);
	}
	// the rows of a group's previous step (the centres before the first)
	// take their pass while this step's tail runs, a launch of one workgroup
	// per row; without a tail, after the step
	auto prev_rows = [&](int j, int g) { return RowMap{R, np, nb, j - 1, j == 1 ? 1 : 2, 0.0, ctx->phase, gb[g]}; };
	auto prev_count = [&](int j, int g) { return (j == 1 ? 1 : 2) * (gb[g + 1] - gb[g]); };
	for (int j = 1; j <= steps; j++)
	{
		for (int g = 0; g < G; g++)
		{
			const int nbg = gb[g + 1] - gb[g];
			const RowMap M{R, np, nb, j, 2, 0.0, ctx->phase, gb[g]};
			unsigned mask = 0;
			if (tailed)
			{
				int rows2[2] = {-1, -1};
				hipEvent_t e = ctx->fmm_rows_event[2 * g + (j & 1)];
				if (j >= 3 && e)
				{
					if (!CCF_HIP(ctx, hipEventSynchronize(e))) return false;
					rows2[0] = fmm_lag_slot(ctx, g, j)[0];
					rows2[1] = fmm_lag_slot(ctx, g, j)[1];
				}
				// the first steps (a predictor without history) can list
				// thousands of roots a row; later the pass runs while the
				// step before the previous one gave it rows. No periodic
				// probe: its launches over empty rows cost more than the
				// direct FP32 sums of a continuation sweep, O(listed R),
				// that it could replace
				for (int q = 0; q < ROOT_FMM_STEP_SWEEPS; q++)
					if (j < 3 || rows2[q] > 0) mask |= 1u << q;
			}
			else
				for (int q = 0; q < ROOT_FMM_STEP_SWEEPS; q++)
					if (prev[q] != 0 || j % ROOT_FMM_PROBE == 1) mask |= 1u << q;
			// the exact smooth sheets are discontinuous at the field's edge
			// (boxcar / heaviside terms), where extrapolated roots could cross
			// onto another branch: those keep the previous step's roots
			const int order = ctx->P.approx ? ctx->P.predictor : 0;
			k_prepare<<<dim3(grid(R), 2 * nbg), BLOCK, 0, gs[g]>>>(ctx->ccs_init, M, *gsw[g], order, ctx->dorder);
			if (!CCF_LAUNCH(ctx, "prepare_roots_kernel")) return false;
			bool prev_done = false;
			const auto at_tail = [&]() {
				prev_done = true;
				return errors(prev_rows(j, g), prev_count(j, g), gs[g]);
			};
			if (!run_sweeps(ctx, *gsw[g], ctx->ccs_init, M, 2 * nbg, cap, mask, tailed ? cap + 1 : first,
					tailed ? (j < 3 ? TAIL_AT + 1 : TAIL_AT) : 0, at_tail, gs[g], g))
				return false;
			if (!prev_done && !errors(prev_rows(j, g), prev_count(j, g), gs[g])) return false;
		}
		prev[0] = ctx->fmm_rows_sweep[0];
		prev[1] = ctx->fmm_rows_sweep[1];
		// a step checked first at k can end no earlier than k: every
		// ROOT_FMM_PROBE-th step starts from 3 again
		first = (j + 1) % ROOT_FMM_PROBE == 1 ? first0 : std::max(2, std::min(ctx->phase_sweeps, 6));
		if (verbose >= 1 && j * 100 / steps > (j - 1) * 100 / steps)
		{
			if (!CCF_HIP(ctx, hipStreamSynchronize(s))) return false;
			std::cout << "\r[";
			for (int b = 0; b < 50; b++) std::cout << (b <= (long long)j * 50 / steps ? "=" : " ");
			std::cout << "] " << j * 100 / steps << " %" << std::flush;
		}
	}
	for (int g = 0; g < G; g++)
		if (!errors(prev_rows(steps + 1, g), steps == 0 ? gb[g + 1] - gb[g] : 2 * (gb[g + 1] - gb[g]), gs[g])) return false;
	// the stages after the continuation run on the context's stream
	if (G > 1 && !group_streams_join(ctx, G)) return false;
	if (!CCF_HIP(ctx, hipStreamSynchronize(s))) return false;
	auto t1 = std::chrono::steady_clock::now();
	ctx->info.t_ccs = std::chrono::duration<double>(t1 - t0).count();
	if (verbose >= 1) std::cout << "\nDone finding critical curve positions. Elapsed time: " << ctx->info.t_ccs << " seconds.\n";

	// transpose to file order (ccf.cuh:1072-1080), beside the error pass's
	// last launches (a failed error check below still writes no file)
	if (!ctx->ccs && !CCF_HIP(ctx, ctx->dmalloc(&ctx->ccs, (size_t)total * sizeof(double2)))) return false;
	k_transpose<<<dim3(grid(R, 32), grid(np + nb, 32)), BLOCK, 0, s>>>(ctx->ccs_init, np + nb, R, ctx->ccs);
	if (!CCF_LAUNCH(ctx, "transpose_array_kernel")) return false;

	unsigned long long c[C_COUNT] = {};
	{
		// the per-slot copies folded on the device (k_fold_stats, after the
		// error pass) and the C_COUNT totals read back through the pinned
		// staging buffer
		if (!ctx->host_stage &&
			!CCF_HIP(ctx, hipHostMalloc((void**)&ctx->host_stage, sizeof(double) * STAGE_DOUBLES, hipHostMallocDefault)))
			return false;
		unsigned long long* const hs = reinterpret_cast<unsigned long long*>(ctx->host_stage);
		if (!CCF_HIP(ctx, hipEventRecord(ctx->err_event, es)) || !CCF_HIP(ctx, hipStreamWaitEvent(s, ctx->err_event, 0)))
			return false;
		k_fold_stats<<<1, 64, 0, s>>>(ctx->counters, ctx->counters + STAT_WORDS);
		if (!CCF_LAUNCH(ctx, "statistics") ||
			!CCF_HIP(ctx, hipMemcpyAsync(hs, ctx->counters + STAT_WORDS, C_COUNT * sizeof(unsigned long long), hipMemcpyDeviceToHost, s)) ||
			!CCF_HIP(ctx, hipStreamSynchronize(s)))
			return false;
		for (int k = 0; k < C_COUNT; k++) c[k] = hs[k];
	}
	if (c[C_NAN])
	{
		ctx->fail("Error. Errors in 1/mu contain values which are not positive real numbers.\n");
		return false;
	}
	ctx->caustics_done = wc;
	ctx->mu_done = wm;
	double e;
	std::memcpy(&e, &c[C_ERRBITS], sizeof(e));
	ctx->max_error = e;
	std::memcpy(&ctx->info.max_witt_residual, &c[C_RESID], sizeof(double));
	if (verbose >= 1) std::cout << "Maximum error in 1/mu: " << e << "\n\n";
	ctx->info.points_total = total;
	ctx->info.points_converged = (long long)c[C_CONV];
	ctx->info.points_outside_tree = (long long)c[C_OUTSIDE];
	if (verbose >= 3) std::cout << "error pass: " << c[C_OUTSIDE] << " points outside the tree\n";
	ctx->info.updates_full = (long long)(c[C_FULL] + c[C_TAIL_FULL]);
	ctx->info.updates_early = (long long)(c[C_EARLY] + c[C_TAIL_EARLY]);
	ctx->info.updates_newton = (long long)(c[C_NEWTON] + c[C_TAIL_NEWTON]);
	ctx->info.tail_sweeps = (long long)c[C_TAIL_SWEEPS];
	ctx->info.cluster_chunks = (long long)c[C_CLUSTER];
	{
		// algorithmic flops: 10 per Cauchy term of the sums; per Witt
		// evaluation 36 per near-field star and 16 per expansion order when
		// f0 and f1 are formed (full updates, Newton-ended roots and, in the
		// mixed form, the polished early exits), 18 and 8 for f0 alone (the
		// FP64 form's early exits), 8 per rectangular smooth-sheet term; the
		// FMM pole sum of a full update adds 8 per order
		const bool fmm_pole = ctx->P.precision == 1 && ctx->tree.order > 0;
		const double Np = fmm_pole ? 0.0 : ctx->num_stars + ((!ctx->rectangular && !ctx->P.approx) ? 1 : 0);
		const double p = ctx->tree.order;
		const double sm = (ctx->rectangular && ctx->P.approx) ? 8.0 * (ctx->taylor_smooth - 1) : 0.0;
		const double full = (double)c[C_FULL];
		// root-sum FMM: 10 per near-field term, 8 (p+1)^2 per M2L (the
		// textbook complex matrix-vector product; the suffix-sum form takes
		// about as many), 8 p per root for P2M and for L2P, about 20 (p+1)^2
		// per cell for M2M and L2L
		const RootGrid& g = ctx->sw.grid;
		const double pr = g.order, rt = (double)c[C_RTGT], rows = (double)c[C_RROWS];
		const double fmm = 10.0 * (double)c[C_RNEAR] + (double)c[C_RM2L] * 8.0 * (pr + 1) * (pr + 1) +
			rows * R * 8.0 * pr + rt * 8.0 * pr + rows * g.ncells * 20.0 * (pr + 1) * (pr + 1);
		ctx->info.flops_sums = (full - rt) * 10.0 * (Np + R - 1) + (g.order > 0 ? fmm : 0.0);
		ctx->info.root_fmm_updates = (long long)c[C_RTGT];
		ctx->info.root_fmm_rows = (long long)c[C_RROWS];
		ctx->info.root_fmm_order = g.order;
		ctx->info.root_fmm_leaf_level = g.leaf;
		ctx->info.root_fmm_leaf = g.order > 0 ? (ctx->P.root_fmm_leaf > 0 ? ctx->P.root_fmm_leaf : ROOT_FMM_LEAF) : 0;
		ctx->info.roots_per_thread = sums_launch(ctx, ctx->P.precision == 1).rpt;
		// the tails: every evaluation forms f0 and f1 (mixed form), every
		// full update sums the row (and the stars without the FMM pole sum)
		const double tail_evals = (double)(c[C_TAIL_FULL] + c[C_TAIL_EARLY] + c[C_TAIL_NEWTON]);
		ctx->info.flops_tail = 36.0 * (double)c[C_TAIL_NEAR] + tail_evals * (16.0 * p + sm) +
			(double)c[C_TAIL_FULL] * (10.0 * (Np + R - 1) + (fmm_pole ? 8.0 * p : 0.0));
		const double ef = ctx->P.precision == 1 ? 1.0 : 0.5; // share of f0 + f1 in an early exit
		const double early = (double)c[C_EARLY];
		ctx->info.flops_update = ctx->info.flops_sums + 36.0 * (double)c[C_NEAR] + (full + (double)c[C_NEWTON]) * (16.0 * p + sm) +
			ef * (36.0 * (double)c[C_NEAR_EARLY] + early * 16.0 * p) + early * sm + (fmm_pole ? full * 8.0 * p : 0.0);
	}

	// (the transpose to file order was queued after the continuation)
	if (!CCF_HIP(ctx, hipStreamSynchronize(s))) return false;
	if (ctx->P.timing)
	{
		double ms = 0;
		for (auto& t : ctx->timers)
		{
			float x = 0;
			(void)hipEventElapsedTime(&x, t.a, t.b);
			ms += x;
		}
		ctx->info.aberth_kernel_ms = ms;
		ctx->info.aberth_launches = (long long)ctx->timers.size();
		ctx->timers.clear();
		double mw = 0;
		for (auto& t : ctx->witt_timers)
		{
			float x = 0;
			(void)hipEventElapsedTime(&x, t.a, t.b);
			mw += x;
		}
		ctx->info.witt_kernel_ms = mw;
		ctx->info.witt_launches = (long long)ctx->witt_timers.size();
		ctx->witt_timers.clear();
		double mt = 0;
		for (auto& t : ctx->tail_timers)
		{
			float x = 0;
			(void)hipEventElapsedTime(&x, t.a, t.b);
			mt += x;
		}
		ctx->info.tail_kernel_ms = mt;
		ctx->tail_timers.clear();
		ctx->events_used = 0;
	}
	ctx->info.flops_witt = ctx->info.flops_update - ctx->info.flops_sums;
	return true;
}

bool find_caustics_dev(ccf_ctx* ctx, bool caustics, bool mu)
{
	const long long n = (long long)(ctx->P.num_phi + ctx->P.num_branches) * ctx->num_roots;
	hipStream_t s = ctx->stream;
	// what the error pass of find_ccs already formed
	if (ctx->caustics_done) caustics = false;
	if (ctx->mu_done) mu = false;
	if (!caustics && !mu) return true;
	if (caustics && !ctx->caustics && !CCF_HIP(ctx, ctx->dmalloc(&ctx->caustics, (size_t)n * sizeof(double2)))) return false;
	if (mu && !ctx->mu_scales && !CCF_HIP(ctx, ctx->dmalloc(&ctx->mu_scales, (size_t)n * sizeof(double)))) return false;
	k_caustics<<<xcd_round(grid(n)), BLOCK, 0, s>>>(ctx->lens(), ctx->ccs, n, caustics ? ctx->caustics : nullptr, mu ? ctx->mu_scales : nullptr);
	return CCF_LAUNCH(ctx, "find_caustics_kernel") && CCF_HIP(ctx, hipStreamSynchronize(s));
}

bool generate_stars_dev(ccf_ctx* ctx)
{
	const int n = ctx->num_stars;
	k_generate_stars<<<grid(n), BLOCK, 0, ctx->stream>>>(n, (unsigned long long)(unsigned int)ctx->random_seed, ctx->rectangular,
		ctx->corner.re, ctx->corner.im, ctx->corner.abs(), ctx->imf, ctx->spos, ctx->smass);
	return CCF_LAUNCH(ctx, "generate_star_field_kernel") && CCF_HIP(ctx, hipStreamSynchronize(ctx->stream));
}

bool eval_points_dev(ccf_ctx* ctx, const double* z, int n, double phi, double* out)
{
	if (n <= 0) return true;
	double2* dz = nullptr;
	double* dout = nullptr;
	hipStream_t s = ctx->stream;
	bool ok = CCF_HIP(ctx, ctx->dmalloc(&dz, (size_t)n * sizeof(double2))) && CCF_HIP(ctx, ctx->dmalloc(&dout, (size_t)n * 8 * sizeof(double))) &&
		CCF_HIP(ctx, hipMemcpyAsync(dz, z, (size_t)n * sizeof(double2), hipMemcpyHostToDevice, s));
	if (ok)
	{
		k_eval<<<grid(n), BLOCK, 0, s>>>(ctx->lens(), dz, n, phi, dout);
		ok = CCF_LAUNCH(ctx, "eval kernel") && CCF_HIP(ctx, hipMemcpyAsync(out, dout, (size_t)n * 8 * sizeof(double), hipMemcpyDeviceToHost, s)) &&
			CCF_HIP(ctx, hipStreamSynchronize(s));
	}
	ctx->dfree(dz);
	ctx->dfree(dout);
	return ok;
}

bool aberth_sweep_dev(ccf_ctx* ctx, const double* roots, int n, double phi, double* out)
{
	if (n <= 0) return true;
	hipStream_t s = ctx->stream;
	double2* row = nullptr;
	SweepBufs B;
	bool ok = CCF_HIP(ctx, ctx->dmalloc(&row, (size_t)n * sizeof(double2))) && alloc_sweep(ctx, B, 1, n) &&
		CCF_HIP(ctx, hipMemcpyAsync(row, roots, (size_t)n * sizeof(double2), hipMemcpyHostToDevice, s));
	if (ok)
	{
		const RowMap M{n, 2, 1, 0, 0, phi};
		k_prepare<<<dim3(grid(n), 1), BLOCK, 0, s>>>(row, M, B, 0, nullptr);
		ok = CCF_LAUNCH(ctx, "prepare") && run_sweeps(ctx, B, row, M, 1, 1, 1) &&
			CCF_HIP(ctx, hipMemcpyAsync(out, row, (size_t)n * sizeof(double2), hipMemcpyDeviceToHost, s)) &&
			CCF_HIP(ctx, hipStreamSynchronize(s));
	}
	ctx->dfree(row);
	free_sweep(ctx, B);
	return ok;
}

// ---------------------------------------------------------------------------
// One Aberth sweep that allocates nothing (ccf_aberth_sweep_noalloc): the
// row's staging copy and every buffer of the sweep are slices of `scratch`,
// a file-scope device array that each device the library runs on holds
// until the process ends. In host code the array's name is no device
// address: its address on a device comes from hipGetSymbolAddress, taken
// once per device with that device current (scratch_setup, called when a
// context initialises the device), and only that address is handed to
// kernels and copies. Every call on a device shares that device's array and
// finds the last call's contents: what must start at a fixed value is set
// by each call, and calls that overlap on one device race.
constexpr size_t NOALLOC_SCRATCH_BYTES = size_t(32) << 20;
static __device__ __align__(256) unsigned char scratch[NOALLOC_SCRATCH_BYTES];

namespace
{
constexpr int SCRATCH_DEVICES = 64;
std::mutex scratch_mu;
unsigned char* scratch_addr[SCRATCH_DEVICES] = {};
hipError_t scratch_err[SCRATCH_DEVICES] = {}; // a failed look-up of the address, reported by the sweep

// slices of the scratch array, 256-byte aligned; with a null base it only
// adds up the bytes the slices take
struct ScratchCarve
{
	unsigned char* base;
	size_t used = 0;
	template <class T>
	bool operator()(T*& p, size_t count)
	{
		used = (used + 255) & ~size_t(255);
		p = base ? reinterpret_cast<T*>(base + used) : nullptr;
		used += count * sizeof(T);
		return true;
	}
};

__global__ void k_fill_int(int* __restrict__ p, size_t n, int v)
{
	const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
	if (i < n) p[i] = v;
}

bool fill_int(ccf_ctx* ctx, int* p, size_t n, int v)
{
	k_fill_int<<<(unsigned)((n + BLOCK - 1) / BLOCK), BLOCK, 0, ctx->stream>>>(p, n, v);
	return CCF_LAUNCH(ctx, "fill kernel");
}
} // namespace

bool scratch_setup(ccf_ctx* ctx)
{
	int dev = 0;
	if (!CCF_HIP(ctx, hipGetDevice(&dev))) return false;
	if (dev < 0 || dev >= SCRATCH_DEVICES) return true; // the no-allocation sweep refuses to run there
	std::lock_guard<std::mutex> lock(scratch_mu);
	if (scratch_addr[dev]) return true;
	void* p = nullptr;
	const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
	if (e != hipSuccess || !p)
	{
		// only the no-allocation sweep needs the array: it reports this
		scratch_err[dev] = e != hipSuccess ? e : hipErrorInvalidSymbol;
		(void)hipGetLastError();
		return true;
	}
	scratch_addr[dev] = static_cast<unsigned char*>(p);
	scratch_err[dev] = hipSuccess;
	return true;
}

bool aberth_sweep_noalloc_dev(ccf_ctx* ctx, const double* roots, int n, double phi, double* out)
{
	if (n <= 0) return true;
	int dev = 0;
	if (!CCF_HIP(ctx, hipGetDevice(&dev))) return false;
	unsigned char* base = nullptr;
	hipError_t why = hipSuccess;
	{
		std::lock_guard<std::mutex> lock(scratch_mu);
		if (dev >= 0 && dev < SCRATCH_DEVICES)
		{
			base = scratch_addr[dev];
			why = scratch_err[dev];
		}
	}
	if (!base)
	{
		ctx->fail("Error. ccf_aberth_sweep_noalloc: no scratch array on device " + std::to_string(dev) +
			(why != hipSuccess ? std::string(" (hipGetSymbolAddress: ") + hipGetErrorString(why) + ")" : std::string()) + "\n");
		return false;
	}
	// the bytes first: the row, then the sweep buffers (and the root-sum
	// FMM's when the context takes it)
	SweepBufs B;
	double2* row = nullptr;
	ScratchCarve size{nullptr};
	size(row, (size_t)n);
	if (!sweep_layout(ctx, B, 1, n, true, size)) return false;
	if (size.used > NOALLOC_SCRATCH_BYTES)
	{
		ctx->fail("Error. ccf_aberth_sweep_noalloc: a sweep of " + std::to_string(n) + " roots needs " + std::to_string(size.used) +
			" bytes of scratch, more than the " + std::to_string(NOALLOC_SCRATCH_BYTES) + " it has\n");
		return false;
	}
	B = SweepBufs{};
	ScratchCarve carve{base};
	carve(row, (size_t)n);
	if (!sweep_layout(ctx, B, 1, n, true, carve)) return false;
	hipStream_t s = ctx->stream;
	// the Witt kernel's ticket starts at 0, the slots of listed roots at -1,
	// the row's FMM mode at 0; k_prepare sets the active list and clears the
	// counts
	if (!fill_int(ctx, B.ticket, 1, 0)) return false;
	if (B.slot && !(fill_int(ctx, B.slot, (size_t)n, -1) && fill_int(ctx, B.rmode, 1, 0))) return false;
	if (!CCF_HIP(ctx, hipMemcpyAsync(row, roots, (size_t)n * sizeof(double2), hipMemcpyHostToDevice, s))) return false;
	const RowMap M{n, 2, 1, 0, 0, phi};
	k_prepare<<<dim3(grid(n), 1), BLOCK, 0, s>>>(row, M, B, 0, nullptr);
	if (!CCF_LAUNCH(ctx, "prepare")) return false;
	// no timing events (creating them would allocate)
	const int timing = ctx->P.timing;
	ctx->P.timing = 0;
	const bool ok = run_sweeps(ctx, B, row, M, 1, 1, 1);
	ctx->P.timing = timing;
	return ok && CCF_HIP(ctx, hipMemcpyAsync(out, row, (size_t)n * sizeof(double2), hipMemcpyDeviceToHost, s)) &&
		CCF_HIP(ctx, hipStreamSynchronize(s));
}

bool root_sums_dev(ccf_ctx* ctx, const double* roots, int n, double* out)
{
	if (n <= 0) return true;
	hipStream_t s = ctx->stream;
	double2* row = nullptr;
	unsigned long long* stats = nullptr;
	SweepBufs B;
	B.rows = 1;
	B.R = n;
	std::vector<int> iota(n), one{1};
	for (int i = 0; i < n; i++) iota[i] = i;
	std::vector<double2> res(n);
	bool ok = CCF_HIP(ctx, ctx->dmalloc(&row, (size_t)n * sizeof(double2))) && alloc_root_fmm(ctx, B) &&
		CCF_HIP(ctx, ctx->dmalloc(&stats, STAT_WORDS * sizeof(unsigned long long))) &&
		CCF_HIP(ctx, hipMemcpyAsync(row, roots, (size_t)n * sizeof(double2), hipMemcpyHostToDevice, s)) &&
		CCF_HIP(ctx, hipMemcpyAsync(B.slot, iota.data(), (size_t)n * sizeof(int), hipMemcpyHostToDevice, s)) &&
		CCF_HIP(ctx, hipMemcpyAsync(B.rmode, one.data(), sizeof(int), hipMemcpyHostToDevice, s));
	if (ok)
	{
		const RowMap M{n, 2, 1, 0, 0, 0.0};
		root_fmm_pass(B, row, M, 1, stats, s);
		ok = CCF_LAUNCH(ctx, "root-sum FMM") &&
			CCF_HIP(ctx, hipMemcpyAsync(res.data(), B.rsum, (size_t)n * sizeof(double2), hipMemcpyDeviceToHost, s)) &&
			CCF_HIP(ctx, hipStreamSynchronize(s));
	}
	if (ok)
		for (int i = 0; i < n; i++)
		{
			out[2 * i] = -res[i].x;
			out[2 * i + 1] = -res[i].y;
		}
	ctx->dfree(row);
	ctx->dfree(stats);
	free_sweep(ctx, B);
	return ok;
}

} // namespace ccf
