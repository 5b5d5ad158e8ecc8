// Many independent star fields solved together: the batched form.
//
// The reference solves one field per process (main_ccf.cu:514-515). Here a
// batch of fields of uneven size shares every launch:
//  * each field keeps its own context (parameters, stars, seeds, output rows,
//    the files ccf_save writes), loaded as a single field would be;
//  * their trees and FMM coefficients are built as one forest
//    (build_forest), level by level for all fields at once;
//  * the Aberth solves are two persistent launches over a task queue:
//      phase 0, one task per (field, branch): the branch centre row from the
//               seeds (ccf.cuh:906-977);
//      phase 1, one task per (field, branch, side): the chain of phi steps of
//               that side toward 0 or 2 pi (ccf.cuh:979-1083), each step
//               predicted from the previous ones as in the single-field path.
//    One workgroup runs a task. The R roots of its row live in LDS (with the
//    field's stars in the FP64 form, when they fit), or, for a field whose
//    rows exceed the LDS a workgroup can hold (20 000 roots at 10 000 stars),
//    in the output row itself in global memory, which the L2 serves; a
//    Jacobi sweep is three
//    barrier-separated phases over a compacted list of the active roots:
//    the Witt function and the 1/mu early exit (ccf_functions.cuh:197-213),
//    the Cauchy sums over stars and roots (ccf_functions.cuh:214-242), then
//    the Aberth update and the |dz| / theta < 1e-9 test (ccf_functions.cuh:
//    379-389). Workgroups take tasks from the queue, largest rows first,
//    until it is empty.
//  * the 1/mu error pass (which also forms the caustic map from the same
//    field evaluation) and the transpose run once over all fields;
//  * the launches of a solve are captured in a HIP graph and replayed.
// An autotuner picks the workgroup size of each phase by timing the phase's
// launch on the batch itself.
#include "lens.hpp"

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstring>
#include <iostream>
#include <numeric>
#include <vector>

namespace ccf
{
namespace
{

constexpr int BLOCK = 256;
// workgroups per CU the list buffers of the persistent grid are sized for
constexpr int MAX_BATCH_PER_CU = 8;

inline int grid(long long n, int b = BLOCK) { return (int)((n + b - 1) / b); }

__device__ inline double rcp64(double x)
{
	double r = __builtin_amdgcn_rcp(x);
	double e = fma(-x, r, 1.0);
	r = fma(r, e, r);
	e = fma(-x, r, 1.0);
	return fma(r, e, r);
}

__device__ inline unsigned long long wave_sum(unsigned long long v)
{
	for (int d = 32; d >= 1; d >>= 1) v += __shfl_xor(v, d, 64);
	return v;
}

// one field of the batch as the kernels see it
struct BField
{
	LensView L;              // lens model and tree (L.root = the field's root node)
	double2* roots;          // [(np + nb)][R] phi rows, row-major
	const double2* seeds;    // [R] initial roots
	const double2* cs;       // (cos phi, sin phi) per phi row
	const double2* stars;    // the field's star positions, tree order
	const float2* stars32;   // the same in FP32 (mixed form)
	const int* order;        // [R] processing order of the roots
	double2* ccs;            // [R nb][np/nb + 1] output (file order)
	double2* caustics;       // same layout
	double* mu;              // same layout
	unsigned long long* stats; // [BSTAT]
	int N, R, np, nb, circ_exact, predictor, cap_init, cap_step;
	float cluster_d2;        // mixed form: squared root separation below which a Cauchy term is FP64
};

// per-field counters
enum BStat
{
	B_FULL = 0,
	B_EARLY = 1,
	B_CONV = 2,
	B_NEAR = 3,
	B_SWEEPS = 4,
	B_ERRBITS = 5,
	B_NAN = 6,
	B_RESID = 7,
	B_NEWTON = 8,
	BSTAT = 9
};

struct BTask
{
	int field, branch, side, cost;
};

__device__ inline int centre_row(const BField& F, int c) { return F.np / (2 * F.nb) + c * (F.np / F.nb + 1); }

// All sweeps of one row (at most cap): zc holds the row's roots (LDS),
// la/lb are the ping-pong active lists (the workgroup's global slice). The
// threads take the listed roots of a sweep in strides, so the row's length
// costs no registers; a root that needs the Cauchy sums takes a slot k in
// the workgroup's list (fl, f0, f1, ex: global, L2-resident). Returns after the first sweep that leaves no
// root active. newton (MIXED, continuation steps): a root whose FP64 Newton
// step is below NEWTON_STEP_TOL takes it instead of an Aberth update, and
// ends there when the step is below the convergence threshold, as in the
// single-field path.
constexpr int BCNT = 6; // full, early, converged, near-field stars, sweeps, Newton
// a sweep that lists at most COOP_MAX roots for the Cauchy sums sums them
// cooperatively: wave w takes listed roots w, w + waves, ... and
// its lanes split the row (and the stars), so a few listed roots do not
// leave the workgroup waiting on their owners' serial loops; more listed
// roots take a thread each. The two forms add a root's terms in different
// orders, so the limit is fixed: were it tuned by timing, results would
// depend on the timing.
constexpr int COOP_MAX = 256;
// static LDS of k_batch (the cooperative sums, counters, reductions), kept
// out of the dynamic limit
constexpr int BATCH_STATIC_LDS = 8192;

// a workgroup's list of the roots that take the Cauchy sums in a sweep
struct ListSlice
{
	int* la;      // the ping-pong active lists
	int* lb;
	int* fl;      // row index of slot k
	double2* f0;  // Witt function
	double2* f1;  // its derivative
	double2* ex;  // the Aberth terms formed per root, then (serial sums) the whole sum
};

template <int BS, bool MIXED>
__device__ void solve_row(const BField& F, int row, int cap, bool newton, double2* zc, const double2* st, int* la, int* lb,
	const ListSlice& ls, int* s_n, unsigned long long (&cnt)[BCNT])
{
	const LensView& L = F.L;
	const int R = F.R;
	const double2 e = F.cs[row];
	const bool pole_fmm = MIXED && L.order > 0;
	__shared__ int s_nf;                  // roots listed for the sums this sweep
	__shared__ double2 s_fs[COOP_MAX];    // cooperative sums 2 sum 1/(z - s) - sum 1/(z - z_j), (re, -im)
	constexpr int NW = BS / 64;
	const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
	for (int q = threadIdx.x; q < R; q += BS) la[q] = F.order[q];
	if (threadIdx.x == 0)
	{
		s_n[0] = R;
		s_n[1] = 0;
		s_nf = 0;
	}
	__syncthreads();
	int* cur = la;
	int* nxt = lb;
	for (int it = 0; it < cap; it++)
	{
		const int n = s_n[0];
		if (n == 0) break;
		cnt[4] += threadIdx.x == 0 ? 1 : 0;
		// Witt function and 1/mu early exit of the listed roots
		for (int q = threadIdx.x; q < n; q += BS)
		{
			const int i = cur[q];
			const cd z(zc[i].x, zc[i].y);
			const int node = nearest_node(L, z);
			const Witt w = witt_from(L, z, e.x, e.y, node, p2p(L, node, z));
			if (w.early)
			{
				cnt[1]++;
				cnt[2]++;
				if (MIXED)
				{
					const cd dz = w.f0 / witt_f1(L, z, w);
					if (dz.abs() < 1e-6 * L.theta) zc[i] = make_double2(z.re - dz.re, z.im - dz.im);
				}
				continue;
			}
			const cd f1 = witt_f1(L, z, w);
			if (MIXED && newton)
			{
				const cd dz = w.f0 / f1;
				const double adz = dz.abs();
				if (adz < NEWTON_STEP_TOL * L.theta) // false for a non-finite step
				{
					zc[i] = make_double2(z.re - dz.re, z.im - dz.im);
					if (adz < CONV_TOL * L.theta) cnt[2]++;
					else nxt[atomicAdd(&s_n[1], 1)] = i; // stepped: checked again next sweep
					cnt[3] += w.nr.count;
					cnt[5]++;
					continue;
				}
			}
			cd ex;
			if (F.circ_exact) ex = 2.0 * z.inv(); // ccf_functions.cuh:223-227
			if (pole_fmm) ex += 2.0 * (w.nr.s0 + far_pole32(L, node, z));
			const int k = atomicAdd(&s_nf, 1);
			ls.fl[k] = i;
			ls.f0[k] = make_double2(w.f0.re, w.f0.im);
			ls.f1[k] = make_double2(f1.re, f1.im);
			ls.ex[k] = make_double2(ex.re, ex.im);
			cnt[0]++;
			cnt[3] += w.nr.count;
		}
		__syncthreads();
		const int nf = s_nf;
		if (nf <= COOP_MAX)
		{
			// cooperative sums, one wave per listed root, lanes over the sources
			for (int g = wv; g < nf; g += NW)
			{
				const int ig = ls.fl[g];
				const double2 zg = zc[ig];
				double vx, vy;
				if (MIXED)
				{
					// FP32 terms, a close root's term in FP64
					const float fx = (float)zg.x, fy = (float)zg.y;
					float ax = 0, ay = 0, bx = 0, by = 0;
					double cx = 0, cy = 0;
					if (!pole_fmm)
						for (int s = lane; s < F.N; s += 64)
						{
							const float2 w = F.stars32[s];
							const float dx = fx - w.x, dy = fy - w.y;
							const float iv = __builtin_amdgcn_rcpf(fmaf(dx, dx, dy * dy));
							ax = fmaf(dx, iv, ax);
							ay = fmaf(dy, iv, ay);
						}
					for (int j = lane; j < R; j += 64)
					{
						const double2 w = zc[j];
						const float dx = fx - (float)w.x, dy = fy - (float)w.y;
						const float d2 = fmaf(dx, dx, dy * dy);
						if (d2 < F.cluster_d2 && j != ig)
						{
							const double ex = zg.x - w.x, ey = zg.y - w.y;
							const double iv = rcp64(fmax(fma(ex, ex, ey * ey), 1e-300));
							cx = fma(ex, iv, cx);
							cy = fma(ey, iv, cy);
						}
						else
						{
							// the own root gives dx = dy = 0 and, floored, a zero term
							const float iv = __builtin_amdgcn_rcpf(fmaxf(d2, 1e-30f));
							bx = fmaf(dx, iv, bx);
							by = fmaf(dy, iv, by);
						}
					}
					vx = 2.0 * (double)ax - ((double)bx + cx);
					vy = 2.0 * (double)ay - ((double)by + cy);
				}
				else
				{
					double ax = 0, ay = 0, bx = 0, by = 0;
					if (!pole_fmm)
						for (int s = lane; s < F.N; s += 64)
						{
							const double2 w = st[s];
							const double dx = zg.x - w.x, dy = zg.y - w.y;
							const double iv = rcp64(fma(dx, dx, dy * dy));
							ax = fma(dx, iv, ax);
							ay = fma(dy, iv, ay);
						}
					for (int j = lane; j < R; j += 64)
					{
						// the own root gives dx = dy = 0 and, floored, a zero term
						const double2 w = zc[j];
						const double dx = zg.x - w.x, dy = zg.y - w.y;
						const double iv = rcp64(fmax(fma(dx, dx, dy * dy), 1e-300));
						bx = fma(dx, iv, bx);
						by = fma(dy, iv, by);
					}
					vx = 2 * ax - bx;
					vy = 2 * ay - by;
				}
				for (int o = 32; o >= 1; o >>= 1)
				{
					vx += __shfl_xor(vx, o, 64);
					vy += __shfl_xor(vy, o, 64);
				}
				if (lane == 0) s_fs[g] = make_double2(vx, -vy);
			}
		}
		else
		{
			// each thread sums its listed roots over the stars and the row
			for (int g = threadIdx.x; g < nf; g += BS)
			{
				const int ig = ls.fl[g];
				const double2 zg = zc[ig];
				double ax = 0, ay = 0, bx = 0, by = 0;
				if (MIXED)
				{
					// FP32 terms, a close root's term in FP64, as in the cooperative sums
					const float fx = (float)zg.x, fy = (float)zg.y;
					float sax = 0, say = 0, sbx = 0, sby = 0;
					if (!pole_fmm)
						for (int s = 0; s < F.N; s++)
						{
							const float2 w = F.stars32[s];
							const float dx = fx - w.x, dy = fy - w.y;
							const float iv = __builtin_amdgcn_rcpf(fmaf(dx, dx, dy * dy));
							sax = fmaf(dx, iv, sax);
							say = fmaf(dy, iv, say);
						}
					for (int j = 0; j < R; j++)
					{
						const double2 w = zc[j];
						const float dx = fx - (float)w.x, dy = fy - (float)w.y;
						const float d2 = fmaf(dx, dx, dy * dy);
						if (d2 < F.cluster_d2 && j != ig)
						{
							const double ex = zg.x - w.x, ey = zg.y - w.y;
							const double iv = rcp64(fmax(fma(ex, ex, ey * ey), 1e-300));
							bx = fma(ex, iv, bx);
							by = fma(ey, iv, by);
						}
						else
						{
							// the own root gives dx = dy = 0 and, floored, a zero term
							const float iv = __builtin_amdgcn_rcpf(fmaxf(d2, 1e-30f));
							sbx = fmaf(dx, iv, sbx);
							sby = fmaf(dy, iv, sby);
						}
					}
					ax = (double)sax;
					ay = (double)say;
					bx += (double)sbx;
					by += (double)sby;
				}
				else
				{
					if (!pole_fmm)
						for (int s = 0; s < F.N; s++)
						{
							const double2 w = st[s];
							const double dx = zg.x - w.x, dy = zg.y - w.y;
							const double iv = rcp64(fma(dx, dx, dy * dy));
							ax = fma(dx, iv, ax);
							ay = fma(dy, iv, ay);
						}
					for (int j = 0; j < R; j++)
					{
						// the own root gives dx = dy = 0 and, floored, a zero term
						const double2 w = zc[j];
						const double dx = zg.x - w.x, dy = zg.y - w.y;
						const double iv = rcp64(fmax(fma(dx, dx, dy * dy), 1e-300));
						bx = fma(dx, iv, bx);
						by = fma(dy, iv, by);
					}
				}
				const double2 t = ls.ex[g];
				ls.ex[g] = make_double2(2 * ax - bx + t.x, -(2 * ay - by) + t.y);
			}
		}
		__syncthreads();
		// the Aberth update (every sum is formed, so the roots may move), and
		// the roots that moved by 1e-9 theta or more listed for the next sweep
		for (int g = threadIdx.x; g < nf; g += BS)
		{
			const int i = ls.fl[g];
			const cd z(zc[i].x, zc[i].y);
			const double2 a0 = ls.f0[g], a1 = ls.f1[g], ae = ls.ex[g];
			const cd f0(a0.x, a0.y), f1(a1.x, a1.y);
			const cd sum = nf <= COOP_MAX ? cd(s_fs[g].x, s_fs[g].y) + cd(ae.x, ae.y) : cd(ae.x, ae.y);
			const cd zn = z - f0 / (f1 + f0 * sum);
			zc[i] = make_double2(zn.re, zn.im);
			if (!((zn - z).abs() / L.theta < CONV_TOL)) nxt[atomicAdd(&s_n[1], 1)] = i; // a non-finite step stays active
			else cnt[2]++;
		}
		__syncthreads();
		if (threadIdx.x == 0)
		{
			s_n[0] = s_n[1];
			s_n[1] = 0;
			s_nf = 0;
		}
		int* t = cur;
		cur = nxt;
		nxt = t;
		__syncthreads();
	}
}

// The persistent solve of one phase (0: branch centres, 1: phi chains).
// Dynamic LDS: rlds roots and nlds stars (fields with more stars read them
// from global memory; a field with more than rlds roots solves its rows in
// place in global memory). lists: the active lists and the listed roots,
// rmax entries each per workgroup of the grid (L2-resident; in LDS they would
// leave room for fewer workgroups per CU). At most 128 VGPRs: four waves
// per SIMD.
constexpr int BATCH_WAVES_PER_SIMD = 4;
template <int BS, bool MIXED>
__global__ __launch_bounds__(BS, BATCH_WAVES_PER_SIMD) void k_batch(const BField* __restrict__ fields, const BTask* __restrict__ tasks,
	int ntasks, int* __restrict__ next, int phase, int rmax, int rlds, int nlds, ListSlice lists)
{
	extern __shared__ double2 smem[];
	double2* sl = smem + rlds;
	__shared__ int s_task;
	__shared__ int s_n[2];
	__shared__ unsigned long long s_red[BCNT][BS / 64];
	const size_t lo = (size_t)blockIdx.x * rmax;
	const ListSlice ls{lists.la + lo, lists.lb + lo, lists.fl + lo, lists.f0 + lo, lists.f1 + lo, lists.ex + lo};
	int* la = ls.la;
	int* lb = ls.lb;
	for (;;)
	{
		__syncthreads();
		if (threadIdx.x == 0) s_task = atomicAdd(next, 1);
		__syncthreads();
		const int t = __builtin_amdgcn_readfirstlane(s_task);
		if (t >= ntasks) break;
		const BTask tk = tasks[t];
		const BField& F = fields[tk.field];
		const int R = F.R;
		const double2* st = F.stars;
		if (F.N <= nlds)
		{
			for (int s = threadIdx.x; s < F.N; s += BS) sl[s] = F.stars[s];
			st = sl;
		}
		unsigned long long cnt[BCNT] = {0, 0, 0, 0, 0, 0};
		// the row being solved: in LDS, or in place in the output rows
		const bool lds_row = R <= rlds;
		if (phase == 0)
		{
			const int row = centre_row(F, tk.branch);
			double2* zc = lds_row ? smem : F.roots + (size_t)row * R;
			for (int i = threadIdx.x; i < R; i += BS) zc[i] = F.seeds[i];
			solve_row<BS, MIXED>(F, row, F.cap_init, false, zc, st, la, lb, ls, s_n, cnt);
			if (lds_row)
				for (int i = threadIdx.x; i < R; i += BS) F.roots[(size_t)row * R + i] = zc[i];
		}
		else
		{
			const int c0 = centre_row(F, tk.branch);
			const int sgn = tk.side ? 1 : -1;
			const int steps = F.np / (2 * F.nb);
			for (int j = 1; j <= steps; j++)
			{
				const int row = c0 + sgn * j;
				// prepare_roots_kernel (ccf_functions.cuh:261-295); a predictor
				// extrapolates from the rows this thread itself wrote
				const int q = min(F.predictor, j - 1);
				double2* zc = lds_row ? smem : F.roots + (size_t)row * R;
				// (in place: row is written, the rows before it are read)
				for (int i = threadIdx.x; i < R; i += BS)
					zc[i] = predict_root(F.roots + (size_t)(row - sgn) * R + i, (long long)sgn * R, q);
				solve_row<BS, MIXED>(F, row, F.cap_step, true, zc, st, la, lb, ls, s_n, cnt);
				if (lds_row)
					for (int i = threadIdx.x; i < R; i += BS) F.roots[(size_t)row * R + i] = zc[i];
			}
		}
		// per-field counters, one atomic per counter and task
#pragma unroll
		for (int k = 0; k < BCNT; k++)
		{
			const unsigned long long s = wave_sum(cnt[k]);
			if ((threadIdx.x & 63) == 0) s_red[k][threadIdx.x >> 6] = s;
		}
		__syncthreads();
		if (threadIdx.x < BCNT)
		{
			unsigned long long s = 0;
			for (int w = 0; w < BS / 64; w++) s += s_red[threadIdx.x][w];
			const int which[BCNT] = {B_FULL, B_EARLY, B_CONV, B_NEAR, B_SWEEPS, B_NEWTON};
			if (s) atomicAdd(&F.stats[which[threadIdx.x]], s);
		}
	}
}

__global__ void k_zero_int(int* p, int n)
{
	const int i = blockIdx.x * blockDim.x + threadIdx.x;
	if (i < n) p[i] = 0;
}

__global__ void k_zero_u64(unsigned long long* p, int n)
{
	const int i = blockIdx.x * blockDim.x + threadIdx.x;
	if (i < n) p[i] = 0;
}

// 1/mu error bound and non-finite flag over every point of every field
// (find_errors_kernel + has_nan_err_kernel, ccf_functions.cuh:419-489), and
// the caustic map and length scales of the fields that write them
// (find_caustics_kernel, find_mu_length_scales_kernel)
__global__ __launch_bounds__(BLOCK) void k_batch_errors(const BField* __restrict__ fields, const int* __restrict__ boff, int nf)
{
	// the field of this block: blocks boff[f] .. boff[f + 1] - 1 cover field
	// f's points (uniform binary search)
	const int blk = blockIdx.x;
	int lo = 0, hi = nf;
	while (hi - lo > 1)
	{
		const int mid = (lo + hi) >> 1;
		if (boff[mid] <= blk) lo = mid;
		else hi = mid;
	}
	const BField& F = fields[lo];
	const unsigned R = (unsigned)F.R;
	// points in the rows' processing (Morton) order, so that a wave's
	// evaluations share tree nodes (fewer than 2^31 points per field)
	const unsigned total = (unsigned)(F.np + F.nb) * R;
	const unsigned t0 = (unsigned)(blk - boff[lo]) * BLOCK + threadIdx.x;
	double e = 0, res = 0;
	unsigned long long bad = 0;
	if (t0 < total)
	{
		const unsigned row = t0 / R;
		const long long idx = (long long)row * R + F.order[t0 - row * R];
		const double2 t = F.cs[row];
		const double2 v = F.roots[idx];
		const cd z(v.x, v.y);
		const int node = nearest_node(F.L, z);
		const Witt w = witt_from(F.L, z, t.x, t.y, node, p2p(F.L, node, z));
		const double b = err_bound(w);
		if (!isfinite(b)) bad = 1;
		else
		{
			e = b;
			res = w.f0.abs();
		}
		// the caustic point and length scale from the same sums, in file
		// order (root a, phi row r at a * rows + r)
		if (F.caustics || F.mu)
		{
			cd ws;
			double m;
			lens_map_from(F.L, z, w.nr, w.fr, w.d1, w.d2, ws, m);
			const size_t o = (size_t)(idx - (long long)row * F.R) * (F.np + F.nb) + row;
			if (F.caustics) F.caustics[o] = make_double2(ws.re, ws.im);
			if (F.mu) F.mu[o] = m;
		}
	}
	for (int d = 32; d >= 1; d >>= 1)
	{
		e = fmax(e, __shfl_xor(e, d, 64));
		res = fmax(res, __shfl_xor(res, d, 64));
	}
	bad = wave_sum(bad);
	// one set of atomics per block
	__shared__ double s_e[BLOCK / 64], s_r[BLOCK / 64];
	__shared__ unsigned long long s_b[BLOCK / 64];
	const int wv = threadIdx.x >> 6;
	if ((threadIdx.x & 63) == 0)
	{
		s_e[wv] = e;
		s_r[wv] = res;
		s_b[wv] = bad;
	}
	__syncthreads();
	if (threadIdx.x == 0)
	{
		for (int k = 1; k < BLOCK / 64; k++)
		{
			e = fmax(e, s_e[k]);
			res = fmax(res, s_r[k]);
			bad |= s_b[k];
		}
		atomicMax(&F.stats[B_ERRBITS], (unsigned long long)__double_as_longlong(e));
		atomicMax(&F.stats[B_RESID], (unsigned long long)__double_as_longlong(res));
		if (bad) atomicOr(&F.stats[B_NAN], 1ull);
	}
}

// [rows][R] -> [R][rows] per field (transpose_array_kernel), 32 x 32 tiles
__global__ __launch_bounds__(BLOCK) void k_batch_transpose(const BField* __restrict__ fields)
{
	const BField& F = fields[blockIdx.z];
	const int rows = F.np + F.nb, R = F.R;
	__shared__ double2 t[32][33];
	const int tx = threadIdx.x & 31, ty = threadIdx.x >> 5;
	for (int a0 = blockIdx.x * 32; a0 < R; a0 += gridDim.x * 32)
		for (int r0 = blockIdx.y * 32; r0 < rows; r0 += gridDim.y * 32)
		{
			for (int y = ty; y < 32; y += 8)
			{
				const int r = r0 + y, a = a0 + tx;
				if (r < rows && a < R) t[y][tx] = F.roots[(size_t)r * R + a];
			}
			__syncthreads();
			for (int y = ty; y < 32; y += 8)
			{
				const int a = a0 + y, r = r0 + tx;
				if (r < rows && a < R) F.ccs[(size_t)a * rows + r] = t[tx][y];
			}
			__syncthreads();
		}
}

using BatchKernel = void (*)(const BField*, const BTask*, int, int*, int, int, int, int, ListSlice);

template <bool MIXED>
BatchKernel pick_kernel(int bs)
{
	switch (bs)
	{
	case 256: return k_batch<256, MIXED>;
	case 512: return k_batch<512, MIXED>;
	default: return k_batch<1024, MIXED>;
	}
}

} // namespace

// ---------------------------------------------------------------------------

// one field's stars in its own buffers and at off in the batch's
struct StarSlice
{
	double2* pos;
	double* mass;
	int off, n;
};

// the fields' stars into the batch's arrays (to_fields = 0) or back
__global__ __launch_bounds__(BLOCK) void k_batch_stars(const StarSlice* __restrict__ sl, double2* pos, double* mass, int to_fields)
{
	const StarSlice t = sl[blockIdx.y];
	for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < t.n; i += gridDim.x * blockDim.x)
	{
		if (to_fields)
		{
			t.pos[i] = pos[t.off + i];
			t.mass[i] = mass[t.off + i];
		}
		else
		{
			pos[t.off + i] = t.pos[i];
			mass[t.off + i] = t.mass[i];
		}
	}
}

struct BatchDevice
{
	StarSlice* slices = nullptr;
	ListSlice lists{};       // the listed roots of a sweep, rmax entries per workgroup of the grid
	char* list_mem = nullptr;
	int list_wgs = 0;        // workgroups the lists hold
	int* eoff = nullptr;     // [nf] first block of each field in the error pass
	int eblocks = 0;
	BField* fields = nullptr;
	BTask* tasks0 = nullptr;
	BTask* tasks1 = nullptr;
	int* next = nullptr;
	unsigned long long* stats = nullptr;
	int n0 = 0, n1 = 0, nf = 0;
	int rmax = 0, rlds = 0, nlds = 0, maxrows = 0; // roots of the largest row, rows held in LDS, stars in LDS
	long long maxpts = 0;
};

namespace
{

void free_bdev(ccf_ctx* m, BatchDevice& d)
{
	auto F = [m](auto*& p) { m->dfree(p); p = nullptr; };
	F(d.slices); F(d.eoff); F(d.list_mem); F(d.fields); F(d.tasks0); F(d.tasks1); F(d.next); F(d.stats);
	d = BatchDevice{};
}

// LDS bytes of a workgroup: the row (up to rlds roots) and staged stars
size_t batch_lds(int rlds, int nlds) { return (size_t)(rlds + nlds) * sizeof(double2); }

// persistent grid: as many workgroups as fit the device at once, at most
// one per task (queried outside any graph capture)
int batch_grid(ccf_batch* b, int phase, int bs)
{
	ccf_ctx* m = b->master;
	BatchDevice& d = *b->dev;
	const bool mixed = m->P.precision == 1;
	const BatchKernel fn = mixed ? pick_kernel<true>(bs) : pick_kernel<false>(bs);
	int per_cu = 0;
	if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, (const void*)fn, bs, batch_lds(d.rlds, d.nlds)) != hipSuccess || per_cu < 1)
		per_cu = 1;
	const int cus = m->num_cus > 0 ? m->num_cus : 256;
	// the list buffers hold d.list_wgs workgroups
	return std::max(1, std::min(std::min(phase == 0 ? d.n0 : d.n1, cus * per_cu), d.list_wgs));
}

bool launch_batch_phase(ccf_batch* b, int phase, int bs, int g, hipStream_t s)
{
	ccf_ctx* m = b->master;
	BatchDevice& d = *b->dev;
	const bool mixed = m->P.precision == 1;
	const BatchKernel fn = mixed ? pick_kernel<true>(bs) : pick_kernel<false>(bs);
	const size_t lds = batch_lds(d.rlds, d.nlds);
	const int ntasks = phase == 0 ? d.n0 : d.n1;
	if (g > d.list_wgs)
	{
		m->fail("Error. The batched grid exceeds its list buffers\n");
		return false;
	}
	k_zero_int<<<1, 64, 0, s>>>(d.next, 1);
	hipLaunchKernelGGL(fn, dim3(g), dim3(bs), lds, s, (const BField*)d.fields, (const BTask*)(phase == 0 ? d.tasks0 : d.tasks1),
		ntasks, d.next, phase, d.rmax, d.rlds, d.nlds, d.lists);
	return CCF_LAUNCH(m, "batched Aberth solve");
}

bool launch_post(ccf_batch* b, hipStream_t s)
{
	ccf_ctx* m = b->master;
	BatchDevice& d = *b->dev;
	k_batch_errors<<<d.eblocks, BLOCK, 0, s>>>(d.fields, d.eoff, d.nf);
	k_batch_transpose<<<dim3(std::min(grid(d.rmax, 32), 64), std::min(grid(d.maxrows, 32), 64), d.nf), BLOCK, 0, s>>>(d.fields);
	return CCF_LAUNCH(m, "batched errors, transpose, caustics");
}

// the whole solve of a loaded and built batch, on stream s
bool enqueue_solve(ccf_batch* b, hipStream_t s)
{
	BatchDevice& d = *b->dev;
	k_zero_u64<<<grid((long long)d.nf * BSTAT), BLOCK, 0, s>>>(d.stats, d.nf * BSTAT);
	return launch_batch_phase(b, 0, b->block, b->grid0, s) && launch_batch_phase(b, 1, b->block1, b->grid1, s) &&
		launch_post(b, s);
}

} // namespace

void free_batch_device(ccf_batch* b)
{
	if (b->graph_exec) (void)hipGraphExecDestroy(b->graph_exec);
	if (b->graph) (void)hipGraphDestroy(b->graph);
	b->graph_exec = nullptr;
	b->graph = nullptr;
	if (b->dev)
	{
		free_bdev(b->master, *b->dev);
		delete b->dev;
		b->dev = nullptr;
	}
}

int batch_lds_limit(ccf_ctx* m);

// forest of all fields' trees, then the device tables of the batch
bool batch_build(ccf_batch* b, int verbose)
{
	ccf_ctx* m = b->master;
	hipStream_t s = m->stream;
	const int nf = (int)b->fields.size();
	const auto tb0 = std::chrono::steady_clock::now();
	auto lap = [&](const char* what) {
		if (verbose < 2) return;
		(void)hipStreamSynchronize(s);
		std::cout << "batch build: " << what << " at "
				  << std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - tb0).count() << " ms\n";
	};
	free_batch_device(b);
	lap("previous batch freed");
	b->dev = new BatchDevice();
	BatchDevice& d = *b->dev;
	d.nf = nf;
	// the fields' stars back to back
	ForestSpec spec;
	int N = 0;
	for (ccf_ctx* f : b->fields)
	{
		spec.soff.push_back(N);
		spec.nstars.push_back(f->num_stars);
		f->root_half_length = (f->rectangular ? std::max(f->corner.re, f->corner.im) : f->corner.abs()) * 1.1;
		spec.root_hl.push_back(f->root_half_length);
		spec.theta.push_back(f->theta);
		spec.alpha_error.push_back(f->alpha_error);
		N += f->num_stars;
	}
	m->free_device();
	lap("master freed");
	m->num_stars = N;
	std::vector<StarSlice> slices(nf);
	int nmax_stars = 1;
	for (int f = 0; f < nf; f++)
	{
		ccf_ctx* c = b->fields[f];
		slices[f] = StarSlice{c->spos, c->smass, spec.soff[f], c->num_stars};
		nmax_stars = std::max(nmax_stars, c->num_stars);
	}
	const dim3 g_stars(std::min(grid(nmax_stars), 64), nf);
	if (!CCF_HIP(m, m->dmalloc(&m->spos, (size_t)N * sizeof(double2))) || !CCF_HIP(m, m->dmalloc(&m->smass, (size_t)N * sizeof(double))) ||
		!CCF_HIP(m, m->dmalloc(&m->spos_tmp, (size_t)N * sizeof(double2))) || !CCF_HIP(m, m->dmalloc(&m->smass_tmp, (size_t)N * sizeof(double))) ||
		!CCF_HIP(m, m->dmalloc(&m->spos32, (size_t)N * sizeof(float2))) || !CCF_HIP(m, m->dmalloc(&d.slices, sizeof(StarSlice) * nf)) ||
		!CCF_HIP(m, hipMemcpyAsync(d.slices, slices.data(), sizeof(StarSlice) * nf, hipMemcpyHostToDevice, s)))
		return false;
	k_batch_stars<<<g_stars, BLOCK, 0, s>>>(d.slices, m->spos, m->smass, 0);
	if (!CCF_LAUNCH(m, "batched stars")) return false;
	lap("stars gathered");
	if (!build_forest(m, spec, verbose)) return false;
	lap("forest built");
	// tree-order stars back into each field (ccf_stars.bin is written in tree order)
	k_batch_stars<<<g_stars, BLOCK, 0, s>>>(d.slices, m->spos, m->smass, 1);
	if (!CCF_LAUNCH(m, "batched stars")) return false;
	for (int f = 0; f < nf; f++)
	{
		ccf_ctx* c = b->fields[f];
		c->expansion_order = spec.order;
		c->info.tree_levels = spec.levels[f];
		c->info.expansion_order = m->tree.order;
		c->info.num_nodes = m->tree.nnodes;
		c->info.t_tree = m->info.t_tree;
		c->prepared = true;
	}
	lap("stars scattered");
	// the fields' results: slices of one arena, kept while it is large enough
	{
		auto room = [](size_t n) { return (n + 255) / 256 * 256; };
		size_t need = 0;
		for (ccf_ctx* c : b->fields)
		{
			const size_t pts = (size_t)(c->P.num_phi + c->P.num_branches) * c->num_roots;
			need += room(pts * sizeof(double2)) + (c->P.write_caustics ? room(pts * sizeof(double2)) : 0) +
				(c->P.write_mu_length_scales ? room(pts * sizeof(double)) : 0);
		}
		for (ccf_ctx* c : b->fields)
		{
			// a field's own buffers (from before it was borrowing) go first
			if (!c->results_borrowed)
			{
				c->dfree(c->ccs);
				c->dfree(c->caustics);
				c->dfree(c->mu_scales);
			}
			c->ccs = c->caustics = nullptr;
			c->mu_scales = nullptr;
			c->results_borrowed = false;
		}
		if (need > b->results_bytes)
		{
			if (b->results) (void)hipFree(b->results);
			b->results = nullptr;
			b->results_bytes = 0;
			if (!CCF_HIP(m, hipMalloc(&b->results, need))) return false;
			b->results_bytes = need;
		}
		char* q = (char*)b->results;
		for (ccf_ctx* c : b->fields)
		{
			const size_t pts = (size_t)(c->P.num_phi + c->P.num_branches) * c->num_roots;
			c->ccs = (double2*)q;
			q += room(pts * sizeof(double2));
			if (c->P.write_caustics)
			{
				c->caustics = (double2*)q;
				q += room(pts * sizeof(double2));
			}
			if (c->P.write_mu_length_scales)
			{
				c->mu_scales = (double*)q;
				q += room(pts * sizeof(double));
			}
			c->results_borrowed = true;
		}
	}
	lap("result arena");
	// device tables
	std::vector<BField> hf(nf);
	std::vector<BTask> t0, t1;
	for (int f = 0; f < nf; f++)
	{
		ccf_ctx* c = b->fields[f];
		LensView L = m->lens();
		const LensView Lc = c->lens();
		L.kappa = Lc.kappa;
		L.gamma = Lc.gamma;
		L.theta = Lc.theta;
		L.theta2 = Lc.theta2;
		L.kappastar = Lc.kappastar;
		L.rectangular = Lc.rectangular;
		L.approx = Lc.approx;
		L.taylor_smooth = Lc.taylor_smooth;
		L.corner_re = Lc.corner_re;
		L.corner_im = Lc.corner_im;
		L.corner_abs = Lc.corner_abs;
		L.corner_abs2 = Lc.corner_abs2;
		L.corner_arg = Lc.corner_arg;
		for (int k = 0; k < 4; k++) L.clog[k] = Lc.clog[k];
		L.nstars = c->num_stars;
		L.nts = Lc.nts;
		L.ts_a = Lc.ts_a;
		L.ts_b = Lc.ts_b;
		L.root = f;
		L.lookup = nullptr;
		if (spec.lookup_level[f] > 0)
		{
			const int n = 1 << spec.lookup_level[f];
			L.lookup = m->tree.lookup + spec.lookup_off[f];
			L.lookup_n = n;
			L.lookup_x0 = -spec.root_hl[f];
			L.lookup_inv = n / (2 * spec.root_hl[f]);
		}
		const int R = c->num_roots, np = c->P.num_phi, nb = c->P.num_branches;
		const size_t pts = (size_t)(np + nb) * R;
		BField& F = hf[f];
		F.L = L;
		F.roots = c->ccs_init;
		F.seeds = c->dseeds;
		F.cs = c->phase;
		F.stars = m->spos + spec.soff[f];
		F.stars32 = m->spos32 + spec.soff[f];
		{
			const double dc = CLUSTER_REL * c->root_half_length;
			F.cluster_d2 = (float)(dc * dc);
		}
		F.order = c->dorder;
		F.ccs = c->ccs;
		F.caustics = c->P.write_caustics ? c->caustics : nullptr;
		F.mu = c->P.write_mu_length_scales ? c->mu_scales : nullptr;
		F.N = c->num_stars;
		F.R = R;
		F.np = np;
		F.nb = nb;
		F.circ_exact = (!c->rectangular && !c->P.approx) ? 1 : 0;
		F.predictor = c->P.approx ? c->P.predictor : 0;
		F.cap_init = c->P.init_iters > 0 ? c->P.init_iters : 60;
		F.cap_step = c->P.step_iters > 0 ? c->P.step_iters : 20;
		d.rmax = std::max(d.rmax, R);
		d.maxrows = std::max(d.maxrows, np + nb);
		d.maxpts = std::max(d.maxpts, (long long)pts);
		const int cost = (int)std::min<long long>((long long)R * (R + c->num_stars) / 64, 1 << 30);
		for (int br = 0; br < nb; br++)
		{
			t0.push_back(BTask{f, br, 0, cost});
			t1.push_back(BTask{f, br, 0, cost});
			t1.push_back(BTask{f, br, 1, cost});
		}
	}
	// LDS of a workgroup: the largest row that fits (larger rows are solved
	// in place in global memory), then as many stars as fit
	const long long lim = batch_lds_limit(m);
	for (const BField& F : hf)
		if ((long long)batch_lds(F.R, 0) <= lim) d.rlds = std::max(d.rlds, F.R);
	// largest rows first (longest-processing-time order of the queue)
	auto by_cost = [](const BTask& a, const BTask& b2) { return a.cost > b2.cost; };
	std::stable_sort(t0.begin(), t0.end(), by_cost);
	std::stable_sort(t1.begin(), t1.end(), by_cost);
	d.n0 = (int)t0.size();
	d.n1 = (int)t1.size();
	const long long room = lim - (long long)batch_lds(d.rlds, 0);
	int nmax = 0;
	for (ccf_ctx* c : b->fields) nmax = std::max(nmax, c->num_stars);
	d.nlds = (int)std::max(0LL, std::min<long long>(nmax, room / (long long)sizeof(double2)));
	// the mixed form takes the pole sum from the star FMM, or else its FP32
	// terms from the FP32 star copy: no star is read from LDS, which then
	// holds only the row (more workgroups per CU)
	if (m->P.precision == 1) d.nlds = 0;
	// list buffers for the largest persistent grid (no more workgroups than
	// tasks): rmax slots per workgroup
	{
		const int cus = m->num_cus > 0 ? m->num_cus : 256;
		d.list_wgs = std::min(cus * MAX_BATCH_PER_CU, std::max(d.n0, d.n1));
		const size_t slots = (size_t)d.list_wgs * d.rmax;
		if (!CCF_HIP(m, m->dmalloc(&d.list_mem, slots * (3 * sizeof(int) + 3 * sizeof(double2))))) return false;
		d.lists.f0 = (double2*)d.list_mem;
		d.lists.f1 = d.lists.f0 + slots;
		d.lists.ex = d.lists.f1 + slots;
		d.lists.fl = (int*)(d.lists.ex + slots);
		d.lists.la = d.lists.fl + slots;
		d.lists.lb = d.lists.la + slots;
	}
	lap("field tables");
	// error-pass blocks of each field (fewer than 2^31 points per field and blocks in all)
	std::vector<int> eoff(nf);
	long long eb = 0;
	for (int f = 0; f < nf; f++)
	{
		const long long pts = (long long)(hf[f].np + hf[f].nb) * hf[f].R;
		if (pts >= (1LL << 31))
		{
			m->fail("Error. A batched field has 2^31 critical curve points or more\n");
			return false;
		}
		eoff[f] = (int)eb;
		eb += grid(pts);
	}
	if (eb >= (1LL << 31))
	{
		m->fail("Error. The batch has too many critical curve points for one error pass\n");
		return false;
	}
	d.eblocks = (int)std::max(1LL, eb);
	if (!CCF_HIP(m, m->dmalloc(&d.eoff, sizeof(int) * nf)) ||
		!CCF_HIP(m, hipMemcpyAsync(d.eoff, eoff.data(), sizeof(int) * nf, hipMemcpyHostToDevice, s)))
		return false;
	if (!CCF_HIP(m, m->dmalloc(&d.stats, sizeof(unsigned long long) * nf * BSTAT))) return false;
	for (int f = 0; f < nf; f++) hf[f].stats = d.stats + (size_t)f * BSTAT;
	return CCF_HIP(m, m->dmalloc(&d.fields, sizeof(BField) * nf)) && CCF_HIP(m, m->dmalloc(&d.tasks0, sizeof(BTask) * d.n0)) &&
		CCF_HIP(m, m->dmalloc(&d.tasks1, sizeof(BTask) * d.n1)) && CCF_HIP(m, m->dmalloc(&d.next, sizeof(int))) &&
		CCF_HIP(m, hipMemcpyAsync(d.fields, hf.data(), sizeof(BField) * nf, hipMemcpyHostToDevice, s)) &&
		CCF_HIP(m, hipMemcpyAsync(d.tasks0, t0.data(), sizeof(BTask) * d.n0, hipMemcpyHostToDevice, s)) &&
		CCF_HIP(m, hipMemcpyAsync(d.tasks1, t1.data(), sizeof(BTask) * d.n1, hipMemcpyHostToDevice, s)) &&
		CCF_HIP(m, hipStreamSynchronize(s));
}

// workgroup size of each phase by timing its launch on this batch at every
// candidate (idempotent: phase 0 recomputes the centre rows from the seeds,
// phase 1 the chains from those rows). A root's arithmetic does not depend
// on the workgroup size.
bool batch_autotune(ccf_batch* b, int verbose)
{
	ccf_ctx* m = b->master;
	hipStream_t s = m->stream;
	const int cands[3] = {256, 512, 1024};
	double best[2] = {1e300, 1e300};
	int pick[2] = {0, 0};
	hipEvent_t e0, e1, e2;
	if (!CCF_HIP(m, hipEventCreate(&e0)) || !CCF_HIP(m, hipEventCreate(&e1)) || !CCF_HIP(m, hipEventCreate(&e2))) return false;
	b->tuned_ms.clear();
	for (int bs : cands)
	{
		float ms[2] = {0, 0};
		const int g0 = batch_grid(b, 0, bs), g1 = batch_grid(b, 1, bs);
		if (!launch_batch_phase(b, 0, bs, g0, s) || !CCF_HIP(m, hipEventRecord(e0, s)) ||
			!launch_batch_phase(b, 0, bs, g0, s) || !CCF_HIP(m, hipEventRecord(e1, s)) ||
			!launch_batch_phase(b, 1, bs, g1, s) || !CCF_HIP(m, hipEventRecord(e2, s)) ||
			!CCF_HIP(m, hipEventSynchronize(e2)) || !CCF_HIP(m, hipEventElapsedTime(&ms[0], e0, e1)) ||
			!CCF_HIP(m, hipEventElapsedTime(&ms[1], e1, e2)))
			return false;
		if (verbose >= 2)
			std::cout << "batch autotune: workgroup " << bs << ": centres " << ms[0] << " ms, chains " << ms[1] << " ms\n";
		for (int q = 0; q < 2; q++)
		{
			b->tuned_ms.push_back(ms[q]);
			if (ms[q] < best[q])
			{
				best[q] = ms[q];
				pick[q] = bs;
			}
		}
	}
	(void)hipEventDestroy(e0);
	(void)hipEventDestroy(e1);
	(void)hipEventDestroy(e2);
	if (!pick[0] || !pick[1])
	{
		m->fail("Error. no workgroup size fits the batch\n");
		return false;
	}
	b->block = pick[0];
	b->block1 = pick[1];
	return true;
}

int batch_lds_limit(ccf_ctx* m)
{
	int v = 0;
	if (hipDeviceGetAttribute(&v, hipDeviceAttributeMaxSharedMemoryPerBlock, 0) != hipSuccess || v <= 0) v = 65536;
	return v - BATCH_STATIC_LDS;
}

// solve, errors, transpose and caustics of every field: captured once into a
// HIP graph per build, replayed afterwards
bool batch_solve(ccf_batch* b, int verbose)
{
	ccf_ctx* m = b->master;
	hipStream_t s = m->stream;
	BatchDevice& d = *b->dev;
	// tuned once per batch
	if (b->block == 0 && !batch_autotune(b, verbose)) return false;
	b->grid0 = batch_grid(b, 0, b->block);
	b->grid1 = batch_grid(b, 1, b->block1);
	auto t0 = std::chrono::steady_clock::now();
	if (!b->graph_exec)
	{
		if (!CCF_HIP(m, hipStreamBeginCapture(s, hipStreamCaptureModeThreadLocal))) return false;
		const bool ok = enqueue_solve(b, s);
		hipGraph_t g = nullptr;
		const hipError_t e = hipStreamEndCapture(s, &g);
		if (!ok || !CCF_HIP(m, e)) return false;
		b->graph = g;
		if (!CCF_HIP(m, hipGraphInstantiate(&b->graph_exec, g, nullptr, nullptr, 0))) return false;
		if (verbose >= 2)
			std::cout << "batch solve: graph captured and instantiated in "
					  << std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count() << " ms\n";
	}
	if (!CCF_HIP(m, hipGraphLaunch(b->graph_exec, s)) || !CCF_HIP(m, hipStreamSynchronize(s))) return false;
	auto t1 = std::chrono::steady_clock::now();
	const double dt = std::chrono::duration<double>(t1 - t0).count();
	std::vector<unsigned long long> st((size_t)d.nf * BSTAT);
	if (!CCF_HIP(m, hipMemcpy(st.data(), d.stats, st.size() * sizeof(unsigned long long), hipMemcpyDeviceToHost))) return false;
	bool nan = false;
	b->info = ccf_info{};
	for (int f = 0; f < d.nf; f++)
	{
		ccf_ctx* c = b->fields[f];
		const unsigned long long* x = st.data() + (size_t)f * BSTAT;
		const long long pts = (long long)(c->P.num_phi + c->P.num_branches) * c->num_roots;
		double e;
		std::memcpy(&e, &x[B_ERRBITS], sizeof(e));
		c->max_error = e;
		c->info.max_error = e;
		std::memcpy(&c->info.max_witt_residual, &x[B_RESID], sizeof(double));
		b->info.max_witt_residual = std::max(b->info.max_witt_residual, c->info.max_witt_residual);
		c->info.points_total = pts;
		c->info.points_converged = (long long)x[B_CONV];
		c->info.updates_full = (long long)x[B_FULL];
		c->info.updates_early = (long long)x[B_EARLY];
		c->info.updates_newton = (long long)x[B_NEWTON];
		c->info.sweeps = (long long)x[B_SWEEPS];
		c->info.t_init_roots = 0;
		c->info.t_ccs = dt;
		c->have_roots = c->have_ccs = true;
		c->have_caustics = c->P.write_caustics != 0;
		c->have_mu = c->P.write_mu_length_scales != 0;
		if (x[B_NAN]) nan = true;
		b->info.points_total += pts;
		b->info.points_converged += (long long)x[B_CONV];
		b->info.updates_full += (long long)x[B_FULL];
		b->info.updates_early += (long long)x[B_EARLY];
		b->info.updates_newton += (long long)x[B_NEWTON];
		b->info.sweeps += (long long)x[B_SWEEPS];
		b->info.max_error = std::max(b->info.max_error, e);
		const double Np = c->num_stars + ((!c->rectangular && !c->P.approx) ? 1 : 0);
		b->info.flops_sums += (double)x[B_FULL] * 10.0 * (Np + c->num_roots - 1);
	}
	b->info.t_ccs = dt;
	b->info.t_tree = m->info.t_tree;
	b->info.num_stars = m->num_stars;
	if (nan)
	{
		m->fail("Error. Errors in 1/mu contain values which are not positive real numbers.\n");
		return false;
	}
	return true;
}

} // namespace ccf
