// Aberth-Ehrlich critical-curve solve, phi continuation, 1/mu error bound,
// transpose and caustic map on gfx950.
//
// Work decomposition (weisluke/ccf runs a (root, side, branch) grid of
// single-thread updates that write roots in place, ccf_functions.cuh:321-394):
//  - a "row" is one phi value of one branch side; every sweep updates the
//    still-unconverged roots of all rows of a phase in one launch;
//  - sweeps are Jacobi: the update kernel reads the row as left by the
//    previous sweep and writes new values to a compacted side buffer; a
//    commit kernel scatters them back and appends the roots that have not
//    converged to the next sweep's active list, so converged roots cost no
//    lanes (the reference re-tests a fin[] flag in every launch);
//  - one workgroup owns a tile of active roots of one row and streams the
//    row's roots and all stars through LDS in chunks; every lane reads the
//    same LDS word (broadcast) and keeps RPT roots in registers, so each LDS
//    read feeds RPT complex reciprocal-accumulates.
#include "lens.hpp"

#include <hiprand/hiprand_kernel.h>

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <iostream>
#include <vector>

namespace ccf
{

// device statistics (ccf_ctx::counters)
enum Counter
{
	C_FULL = 0,    // updates that formed the Cauchy sums
	C_EARLY = 1,   // updates that returned on the 1/mu bound
	C_NEAR = 2,    // near-field stars summed by full updates
	C_CONV = 3,    // roots converged
	C_ERRBITS = 4, // max 1/mu error (bits of a non-negative double)
	C_NAN = 5,     // non-finite error seen
	C_COUNT = 8
};

namespace
{

constexpr int BLOCK = 256;
constexpr int CHUNK = 1024; // complex values per LDS chunk (16 KiB in FP64)

inline int grid(long long n, int b = BLOCK) { return (int)((n + b - 1) / b); }

__device__ inline unsigned long long wave_sum(unsigned long long v)
{
	for (int d = 32; d >= 1; d >>= 1) v += __shfl_xor(v, d, 64);
	return v;
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
struct RowMap
{
	int R, np, nb, j, rows_per_branch;
	double phi_fixed;
	__device__ inline void get(int r, int& row, double& phi) const
	{
		if (rows_per_branch == 0)
		{
			row = r;
			phi = phi_fixed;
			return;
		}
		const int c = r / rows_per_branch;
		const int sgn = rows_per_branch == 1 ? 0 : ((r % rows_per_branch) == 0 ? -1 : 1);
		const int center = np / (2 * nb) + c * (np / nb + 1);
		row = center + sgn * j;
		const double phi0 = PI / nb + c * 2 * PI / nb;
		const double dphi = 2 * PI / np * j;
		phi = phi0 + sgn * dphi;
	}
};

// One Aberth update of every active root of every row (ccf_functions.cuh:193-243).
// grid: (tiles, rows); block: BLOCK threads, RPT roots per thread.
template <int RPT>
__global__ __launch_bounds__(BLOCK) void k_aberth(LensView L, const double2* __restrict__ roots, RowMap M,
	const int* __restrict__ act, const int* __restrict__ cnt, int* __restrict__ cnt_other,
	double2* __restrict__ nv, unsigned char* __restrict__ conv, unsigned long long* __restrict__ stats, int circ_exact)
{
	__shared__ double2 tile[CHUNK];
	const int r = blockIdx.y;
	const int R = M.R;
	if (blockIdx.x == 0 && threadIdx.x == 0) cnt_other[r] = 0; // next sweep's list is refilled by k_commit
	const int n = cnt[r];
	const int base = blockIdx.x * (BLOCK * RPT);
	if (base >= n) return;
	int row;
	double phi;
	M.get(r, row, phi);
	const double2* __restrict__ xrow = roots + (size_t)row * R;

	cd z[RPT], f0[RPT], f1[RPT];
	bool live[RPT], full[RPT];
	unsigned long long nfull = 0, nearly = 0, nnear = 0;
#pragma unroll
	for (int k = 0; k < RPT; k++)
	{
		const int i = base + k * BLOCK + threadIdx.x;
		live[k] = i < n;
		full[k] = false;
		const int a = live[k] ? act[(size_t)r * R + i] : 0;
		const double2 v = xrow[a];
		z[k] = cd(v.x, v.y);
		if (live[k])
		{
			Witt w = witt_f0(L, z[k], phi);
			f0[k] = w.f0;
			full[k] = !w.early;
			if (full[k])
			{
				f1[k] = witt_f1(L, z[k], w);
				nfull++;
				nnear += w.nr.count;
			}
			else
			{
				nearly++;
			}
		}
	}

	// 2 sum 1/(z - s_i) over all stars (ccf_functions.cuh:218-227)
	double ax[RPT], ay[RPT];
#pragma unroll
	for (int k = 0; k < RPT; k++) { ax[k] = 0; ay[k] = 0; }
	for (int c0 = 0; c0 < L.nstars; c0 += CHUNK)
	{
		const int m = min(CHUNK, L.nstars - c0);
		for (int t = threadIdx.x; t < m; t += BLOCK) tile[t] = L.spos[c0 + t];
		__syncthreads();
#pragma unroll 4
		for (int j = 0; j < m; j++)
		{
			const double2 w = tile[j];
#pragma unroll
			for (int k = 0; k < RPT; k++)
			{
				const double dx = z[k].re - w.x, dy = z[k].im - w.y;
				const double q = rcp64(fma(dx, dx, dy * dy));
				ax[k] = fma(dx, q, ax[k]);
				ay[k] = fma(dy, q, ay[k]);
			}
		}
		__syncthreads();
	}
	cd S[RPT];
#pragma unroll
	for (int k = 0; k < RPT; k++)
	{
		S[k] = cd(2 * ax[k], -2 * ay[k]);
		if (circ_exact) S[k] += z[k].inv();
		ax[k] = 0;
		ay[k] = 0;
	}
	// sum over the other roots of the row (ccf_functions.cuh:232-239); the
	// self term has dx = dy = 0 exactly and the floor on |d|^2 makes it add 0
	for (int c0 = 0; c0 < R; c0 += CHUNK)
	{
		const int m = min(CHUNK, R - c0);
		for (int t = threadIdx.x; t < m; t += BLOCK) tile[t] = xrow[c0 + t];
		__syncthreads();
#pragma unroll 4
		for (int j = 0; j < m; j++)
		{
			const double2 w = tile[j];
#pragma unroll
			for (int k = 0; k < RPT; k++)
			{
				const double dx = z[k].re - w.x, dy = z[k].im - w.y;
				const double q = rcp64(fmax(fma(dx, dx, dy * dy), 1e-300));
				ax[k] = fma(dx, q, ax[k]);
				ay[k] = fma(dy, q, ay[k]);
			}
		}
		__syncthreads();
	}
#pragma unroll
	for (int k = 0; k < RPT; k++)
	{
		if (!live[k]) continue;
		const int i = base + k * BLOCK + threadIdx.x;
		cd zn = z[k];
		if (full[k])
		{
			const cd s = S[k] - cd(ax[k], -ay[k]);
			zn = z[k] - f0[k] / (f1[k] + f0[k] * s);
		}
		nv[(size_t)r * R + i] = make_double2(zn.re, zn.im);
		conv[(size_t)r * R + i] = ((zn - z[k]).abs() / L.theta) < 1e-9 ? 1 : 0;
	}
	nfull = wave_sum(nfull);
	nearly = wave_sum(nearly);
	nnear = wave_sum(nnear);
	if ((threadIdx.x & 63) == 0)
	{
		if (nfull) atomicAdd(&stats[C_FULL], nfull);
		if (nearly) atomicAdd(&stats[C_EARLY], nearly);
		if (nnear) atomicAdd(&stats[C_NEAR], nnear);
	}
}

// scatter the sweep's updates into the rows and build the next active lists
__global__ void k_commit(double2* __restrict__ roots, RowMap M, const int* __restrict__ act, const int* __restrict__ cnt,
	int* __restrict__ act_next, int* __restrict__ cnt_next, const double2* __restrict__ nv,
	const unsigned char* __restrict__ conv, unsigned long long* __restrict__ stats)
{
	const int r = blockIdx.y;
	const int R = M.R;
	const int i = blockIdx.x * blockDim.x + threadIdx.x;
	const int n = cnt[r];
	unsigned long long nc = 0;
	if (i < n)
	{
		int row;
		double phi;
		M.get(r, row, phi);
		const int a = act[(size_t)r * R + i];
		roots[(size_t)row * R + a] = nv[(size_t)r * R + i];
		if (conv[(size_t)r * R + i]) nc = 1;
		else act_next[(size_t)r * R + atomicAdd(&cnt_next[r], 1)] = a;
	}
	nc = wave_sum(nc);
	if ((threadIdx.x & 63) == 0 && nc) atomicAdd(&stats[C_CONV], nc);
}

// copy row centre+sgn*(j-1) to centre+sgn*j and make every root active
// (prepare_roots_kernel, ccf_functions.cuh:261-295)
__global__ void k_prepare(double2* roots, RowMap M, int* act, int* cnt)
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
			roots[(size_t)row * R + a] = roots[(size_t)(row - sgn) * R + a];
		}
		act[(size_t)r * R + a] = a;
	}
	if (a == 0) cnt[r] = R;
}

// 1/mu error bound at every point of the solved rows, reduced to its max
// with a non-finite flag (find_errors_kernel + has_nan_err_kernel +
// thrust::max_element, ccf_functions.cuh:419-489)
__global__ void k_errors(LensView L, const double2* __restrict__ roots, int R, int np, int nb, unsigned long long* stats)
{
	const long long total = (long long)(np + nb) * R;
	const long long idx = (long long)blockIdx.x * blockDim.x + threadIdx.x;
	double e = 0;
	unsigned long long bad = 0;
	if (idx < total)
	{
		const int row = (int)(idx / R);
		const int per = np / nb + 1;
		const int c = row / per;
		const int k = row % per - np / (2 * nb);
		const int sgn = k < 0 ? -1 : 1;
		const double phi = PI / nb + c * 2 * PI / nb + sgn * (2 * PI / np * (sgn * k));
		const double2 v = roots[idx];
		Witt w = witt_f0(L, cd(v.x, v.y), phi);
		e = err_bound(w);
		if (!isfinite(e)) { bad = 1; e = 0; }
	}
	for (int d = 32; d >= 1; d >>= 1) e = fmax(e, __shfl_xor(e, d, 64));
	bad = wave_sum(bad);
	if ((threadIdx.x & 63) == 0)
	{
		atomicMax(&stats[C_ERRBITS], (unsigned long long)__double_as_longlong(e));
		if (bad) atomicOr(&stats[C_NAN], 1ull);
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
	const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
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

bool run_sweeps(ccf_ctx* ctx, const RowMap& M, int nrows, int cap, int buf0)
{
	const LensView L = ctx->lens();
	const int R = ctx->num_roots;
	hipStream_t s = ctx->stream;
	const int circ_exact = (!ctx->rectangular && !ctx->P.approx) ? 1 : 0;
	const int rpt = ctx->P.roots_per_thread > 0 ? ctx->P.roots_per_thread : 2;
	const int per_tile = BLOCK * rpt;
	int cur = buf0;
	int* act[2] = {ctx->act, ctx->act + (size_t)ctx->max_rows * R};
	int* cnt[2] = {ctx->act_cnt, ctx->act_cnt + ctx->max_rows};
	for (int it = 0; it < cap; it++)
	{
		const dim3 g(grid(R, per_tile), nrows);
		ccf::Timer tm;
		if (ctx->P.timing)
		{
			hipEventCreate(&tm.a);
			hipEventCreate(&tm.b);
			hipEventRecord(tm.a, s);
		}
		if (rpt == 1)
			k_aberth<1><<<g, BLOCK, 0, s>>>(L, ctx->ccs_init, M, act[cur], cnt[cur], cnt[cur ^ 1], ctx->nv, ctx->conv, ctx->counters, circ_exact);
		else if (rpt == 4)
			k_aberth<4><<<g, BLOCK, 0, s>>>(L, ctx->ccs_init, M, act[cur], cnt[cur], cnt[cur ^ 1], ctx->nv, ctx->conv, ctx->counters, circ_exact);
		else
			k_aberth<2><<<g, BLOCK, 0, s>>>(L, ctx->ccs_init, M, act[cur], cnt[cur], cnt[cur ^ 1], ctx->nv, ctx->conv, ctx->counters, circ_exact);
		if (ctx->P.timing)
		{
			hipEventRecord(tm.b, s);
			ctx->timers.push_back(tm);
		}
		k_commit<<<dim3(grid(R), nrows), BLOCK, 0, s>>>(ctx->ccs_init, M, act[cur], cnt[cur], act[cur ^ 1], cnt[cur ^ 1], ctx->nv, ctx->conv, ctx->counters);
		if (!CCF_LAUNCH(ctx, "find_critical_curve_roots_kernel")) return false;
		cur ^= 1;
	}
	return true;
}

} // namespace

bool alloc_solver(ccf_ctx* ctx)
{
	const int R = ctx->num_roots, np = ctx->P.num_phi, nb = ctx->P.num_branches;
	const size_t pts = (size_t)(np + nb) * R;
	ctx->max_rows = 2 * nb;
	bool ok = CCF_HIP(ctx, hipMalloc(&ctx->ccs_init, pts * sizeof(double2))) &&
		CCF_HIP(ctx, hipMalloc(&ctx->act, 2 * (size_t)ctx->max_rows * R * sizeof(int))) &&
		CCF_HIP(ctx, hipMalloc(&ctx->act_cnt, 2 * (size_t)ctx->max_rows * sizeof(int))) &&
		CCF_HIP(ctx, hipMalloc(&ctx->nv, (size_t)ctx->max_rows * R * sizeof(double2))) &&
		CCF_HIP(ctx, hipMalloc(&ctx->conv, (size_t)ctx->max_rows * R)) &&
		CCF_HIP(ctx, hipMalloc(&ctx->counters, C_COUNT * sizeof(unsigned long long))) &&
		CCF_HIP(ctx, hipMemsetAsync(ctx->ccs_init, 0, pts * sizeof(double2), ctx->stream)) &&
		CCF_HIP(ctx, hipMemsetAsync(ctx->act_cnt, 0, 2 * (size_t)ctx->max_rows * sizeof(int), ctx->stream)) &&
		CCF_HIP(ctx, hipMemsetAsync(ctx->counters, 0, C_COUNT * sizeof(unsigned long long), ctx->stream));
	return ok;
}

bool find_initial_roots_dev(ccf_ctx* ctx, int verbose)
{
	const int R = ctx->num_roots, np = ctx->P.num_phi, nb = ctx->P.num_branches;
	hipStream_t s = ctx->stream;
	if (verbose >= 1) std::cout << "Finding initial roots...\n";
	auto t0 = std::chrono::steady_clock::now();
	// seeds at every branch centre (ccf.cuh:711-726)
	for (int c = 0; c < nb; c++)
	{
		const int center = np / (2 * nb) + c * np / nb + c;
		if (!CCF_HIP(ctx, hipMemcpyAsync(ctx->ccs_init + (size_t)center * R, ctx->seeds.data(), (size_t)R * sizeof(double2), hipMemcpyHostToDevice, s)))
			return false;
	}
	const RowMap M{R, np, nb, 0, 1, 0.0};
	k_prepare<<<dim3(grid(R), nb), BLOCK, 0, s>>>(ctx->ccs_init, M, ctx->act, ctx->act_cnt);
	if (!CCF_LAUNCH(ctx, "prepare_roots_kernel")) return false;
	const int cap = ctx->P.init_iters > 0 ? ctx->P.init_iters : 60;
	if (!run_sweeps(ctx, M, nb, cap, 0)) return false;
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
	for (int j = 1; j <= steps; j++)
	{
		const RowMap M{R, np, nb, j, 2, 0.0};
		k_prepare<<<dim3(grid(R), 2 * nb), BLOCK, 0, s>>>(ctx->ccs_init, M, ctx->act, ctx->act_cnt);
		if (!CCF_LAUNCH(ctx, "prepare_roots_kernel")) return false;
		if (!run_sweeps(ctx, M, 2 * nb, cap, 0)) return false;
		if (verbose >= 1 && j * 100 / steps > (j - 1) * 100 / steps)
		{
			if (!CCF_HIP(ctx, hipStreamSynchronize(s))) return false;
			std::cout << "\r[";
			for (int b = 0; b < 50; b++) std::cout << (b <= (long long)j * 50 / steps ? "=" : " ");
			std::cout << "] " << j * 100 / steps << " %" << std::flush;
		}
	}
	if (!CCF_HIP(ctx, hipStreamSynchronize(s))) return false;
	auto t1 = std::chrono::steady_clock::now();
	ctx->info.t_ccs = std::chrono::duration<double>(t1 - t0).count();
	if (verbose >= 1) std::cout << "\nDone finding critical curve positions. Elapsed time: " << ctx->info.t_ccs << " seconds.\n";

	// max 1/mu error and the NaN abort (ccf.cuh:1039-1069)
	const long long total = (long long)(np + nb) * R;
	const LensView L = ctx->lens();
	k_errors<<<grid(total), BLOCK, 0, s>>>(L, ctx->ccs_init, R, np, nb, ctx->counters);
	if (!CCF_LAUNCH(ctx, "find_errors_kernel")) return false;
	unsigned long long c[C_COUNT];
	if (!CCF_HIP(ctx, hipMemcpyAsync(c, ctx->counters, sizeof(c), hipMemcpyDeviceToHost, s)) || !CCF_HIP(ctx, hipStreamSynchronize(s)))
		return false;
	if (c[C_NAN])
	{
		ctx->fail("Error. Errors in 1/mu contain values which are not positive real numbers.\n");
		return false;
	}
	double e;
	std::memcpy(&e, &c[C_ERRBITS], sizeof(e));
	ctx->max_error = e;
	if (verbose >= 1) std::cout << "Maximum error in 1/mu: " << e << "\n\n";
	ctx->info.points_total = total;
	ctx->info.points_converged = (long long)c[C_CONV];
	ctx->info.updates_full = (long long)c[C_FULL];
	ctx->info.updates_early = (long long)c[C_EARLY];
	{
		const double Np = ctx->num_stars + ((!ctx->rectangular && !ctx->P.approx) ? 1 : 0);
		const double p = ctx->tree.order;
		const double sm = (ctx->rectangular && ctx->P.approx) ? 8.0 * (ctx->taylor_smooth - 1) : 0.0;
		ctx->info.flops_update = (double)c[C_FULL] * (10.0 * (Np + R - 1) + 16.0 * p + sm) + 36.0 * (double)c[C_NEAR] + (double)c[C_EARLY] * 8.0 * p;
	}

	// transpose to file order (ccf.cuh:1072-1080)
	if (!ctx->ccs && !CCF_HIP(ctx, hipMalloc(&ctx->ccs, (size_t)total * sizeof(double2)))) return false;
	k_transpose<<<dim3(grid(R, 32), grid(np + nb, 32)), BLOCK, 0, s>>>(ctx->ccs_init, np + nb, R, ctx->ccs);
	if (!CCF_LAUNCH(ctx, "transpose_array_kernel")) return false;
	if (!CCF_HIP(ctx, hipStreamSynchronize(s))) return false;
	if (ctx->P.timing)
	{
		double ms = 0;
		for (auto& t : ctx->timers)
		{
			float x = 0;
			hipEventElapsedTime(&x, t.a, t.b);
			ms += x;
			hipEventDestroy(t.a);
			hipEventDestroy(t.b);
		}
		ctx->info.aberth_kernel_ms = ms;
		ctx->info.aberth_launches = (long long)ctx->timers.size();
		ctx->timers.clear();
	}
	return true;
}

bool find_caustics_dev(ccf_ctx* ctx, bool caustics, bool mu)
{
	const long long n = (long long)(ctx->P.num_phi + ctx->P.num_branches) * ctx->num_roots;
	hipStream_t s = ctx->stream;
	if (caustics && !ctx->caustics && !CCF_HIP(ctx, hipMalloc(&ctx->caustics, (size_t)n * sizeof(double2)))) return false;
	if (mu && !ctx->mu_scales && !CCF_HIP(ctx, hipMalloc(&ctx->mu_scales, (size_t)n * sizeof(double)))) return false;
	k_caustics<<<grid(n), BLOCK, 0, s>>>(ctx->lens(), ctx->ccs, n, caustics ? ctx->caustics : nullptr, mu ? ctx->mu_scales : nullptr);
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
	bool ok = CCF_HIP(ctx, hipMalloc(&dz, (size_t)n * sizeof(double2))) && CCF_HIP(ctx, hipMalloc(&dout, (size_t)n * 8 * sizeof(double))) &&
		CCF_HIP(ctx, hipMemcpyAsync(dz, z, (size_t)n * sizeof(double2), hipMemcpyHostToDevice, s));
	if (ok)
	{
		k_eval<<<grid(n), BLOCK, 0, s>>>(ctx->lens(), dz, n, phi, dout);
		ok = CCF_LAUNCH(ctx, "eval kernel") && CCF_HIP(ctx, hipMemcpyAsync(out, dout, (size_t)n * 8 * sizeof(double), hipMemcpyDeviceToHost, s)) &&
			CCF_HIP(ctx, hipStreamSynchronize(s));
	}
	hipFree(dz);
	hipFree(dout);
	return ok;
}

bool aberth_sweep_dev(ccf_ctx* ctx, const double* roots, int n, double phi, double* out)
{
	if (n <= 0) return true;
	hipStream_t s = ctx->stream;
	double2* row = nullptr;
	int* act = nullptr;
	int* cnt = nullptr;
	double2* nv = nullptr;
	unsigned char* conv = nullptr;
	unsigned long long* st = nullptr;
	bool ok = CCF_HIP(ctx, hipMalloc(&row, (size_t)n * sizeof(double2))) && CCF_HIP(ctx, hipMalloc(&act, (size_t)n * sizeof(int))) &&
		CCF_HIP(ctx, hipMalloc(&cnt, 2 * sizeof(int))) && CCF_HIP(ctx, hipMalloc(&nv, (size_t)n * sizeof(double2))) &&
		CCF_HIP(ctx, hipMalloc(&conv, (size_t)n)) && CCF_HIP(ctx, hipMalloc(&st, C_COUNT * sizeof(unsigned long long))) &&
		CCF_HIP(ctx, hipMemsetAsync(st, 0, C_COUNT * sizeof(unsigned long long), s)) &&
		CCF_HIP(ctx, hipMemcpyAsync(row, roots, (size_t)n * sizeof(double2), hipMemcpyHostToDevice, s));
	if (ok)
	{
		const RowMap M{n, 2, 1, 0, 0, phi};
		const int circ_exact = (!ctx->rectangular && !ctx->P.approx) ? 1 : 0;
		k_prepare<<<dim3(grid(n), 1), BLOCK, 0, s>>>(row, M, act, cnt);
		k_aberth<1><<<dim3(grid(n), 1), BLOCK, 0, s>>>(ctx->lens(), row, M, act, cnt, cnt + 1, nv, conv, st, circ_exact);
		ok = CCF_LAUNCH(ctx, "aberth sweep") && CCF_HIP(ctx, hipMemcpyAsync(out, nv, (size_t)n * sizeof(double2), hipMemcpyDeviceToHost, s)) &&
			CCF_HIP(ctx, hipStreamSynchronize(s));
	}
	hipFree(row);
	hipFree(act);
	hipFree(cnt);
	hipFree(nv);
	hipFree(conv);
	hipFree(st);
	return ok;
}

} // namespace ccf
