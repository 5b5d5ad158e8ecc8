// Device-side lens model: star, far-field and smooth-matter deflections and
// Witt's parametric critical-curve function (weisluke/ccf
// include/alpha_star.cuh, alpha_local.cuh, alpha_smooth.cuh and
// ccf_functions.cuh:33-166).
//
// The reference evaluates each derivative in its own loop with a Smith
// division per star; here one pass over the near-field stars forms
// t = 1/(z - s) once and accumulates m t, m t^2 and m t^3 together, and one
// Horner pass over the local coefficients produces the first three
// derivative sums of the far-field potential.
#pragma once

#include "internal.hpp"

namespace ccf
{

// the descent of tree_node.cuh:529-561 from node n (box b)
__device__ inline int descend(const LensView& L, cd z, int n, double4 b)
{
	while (b.w >= 0)
	{
		if (fabs(z.re - b.x) > b.z || fabs(z.im - b.y) > b.z) return n;
		int q;
		if (z.re >= b.x && z.im >= b.y) q = 0;
		else if (z.re < b.x && z.im >= b.y) q = 1;
		else if (z.re < b.x && z.im < b.y) q = 2;
		else q = 3;
		n = (int)b.w + q;
		b = L.nbox[n];
	}
	return n;
}

// tree_node.cuh:529-561. The point-location table gives the node the
// descent reaches at the table's level in one load; it is used only when z
// lies inside that node's box by a margin far above the rounding of the
// stored centres, where the descent from the root takes the same path.
__device__ inline int nearest_node(const LensView& L, cd z)
{
	if (L.lookup)
	{
		const double fx = (z.re - L.lookup_x0) * L.lookup_inv, fy = (z.im - L.lookup_x0) * L.lookup_inv;
		if (fx >= 0 && fx < L.lookup_n && fy >= 0 && fy < L.lookup_n)
		{
			const int n = L.lookup[(int)fy * L.lookup_n + (int)fx];
			const double4 b = L.nbox[n];
			const double m = b.z * (1 - 1e-9);
			if (fabs(z.re - b.x) < m && fabs(z.im - b.y) < m) return descend(L, z, n, b);
		}
	}
	return descend(L, z, L.root, L.nbox[L.root]);
}

// near-field sums over a node and its neighbours:
// s1 = sum m/(z-s), s2 = sum m/(z-s)^2, s3 = sum m/(z-s)^3, s0 = sum 1/(z-s)
struct Near
{
	cd s0, s1, s2, s3;
	int count = 0;
};

__device__ inline void p2p_add(Near& r, cd z, double4 s)
{
	const cd t = cd(z.re - s.x, z.im - s.y).inv();
	r.s0 += t;
	const cd t1 = t * s.z;
	r.s1 += t1;
	const cd t2 = t1 * t;
	r.s2 += t2;
	r.s3 += t2 * t;
}

__device__ inline Near p2p(const LensView& L, int node, cd z)
{
	Near r;
	const int off = L.nf_off[node];
	if (off >= 0)
	{
		// a leaf: one contiguous run of at most ~32 stars, two at a time
		// four stars per batch of loads, summed in order
		const int n = L.nf_cnt[node];
		const double4* __restrict__ p = L.nfs + off;
		for (int t = 0; t < n; t += 4)
		{
			double4 s[4];
#pragma unroll
			for (int k = 0; k < 4; k++) s[k] = p[min(t + k, n - 1)];
#pragma unroll
			for (int k = 0; k < 4; k++)
				if (t + k < n) p2p_add(r, z, s[k]);
		}
		r.count = n;
		return r;
	}
	// the root (positions outside the tree) or a box edge case: star ranges
	const int nr = L.nnf[node];
	for (int k = 0; k < nr; k++)
	{
		const int2 rg = L.nf[node * MAX_NEAR_RANGES + k];
		for (int s = rg.x; s < rg.x + rg.y; s++)
		{
			const double2 p = L.spos[s];
			const cd t = cd(z.re - p.x, z.im - p.y).inv();
			r.s0 += t;
			const cd t1 = t * L.smass[s];
			r.s1 += t1;
			const cd t2 = t1 * t;
			r.s2 += t2;
			r.s3 += t2 * t;
		}
		r.count += rg.y;
	}
	return r;
}

// local-expansion sums with dz = (z - c)/hl (alpha_local.cuh:18-102):
// h1 = sum i c_i dz^(i-1), h2 = sum i(i-1) c_i dz^(i-2), h3 = sum i(i-1)(i-2) c_i dz^(i-3)
struct Far
{
	cd h1, h2, h3;
	double hl = 1;
};

__device__ inline Far l2p(const LensView& L, int node, cd z)
{
	Far f;
	f.hl = L.hl[node];
	if (L.order == 0) return f;
	const double2 c = L.center[node];
	const cd dz = cd(z.re - c.x, z.im - c.y) / f.hl;
	const double2* a = L.loc + (size_t)node * (L.order + 1);
	// Horner from the top, eight coefficients per batch of loads
	for (int i0 = L.order; i0 >= 1; i0 -= 8)
	{
		double2 cb[8];
#pragma unroll
		for (int k = 0; k < 8; k++) cb[k] = a[max(i0 - k, 0)];
#pragma unroll
		for (int k = 0; k < 8; k++)
		{
			const int i = i0 - k;
			if (i < 1) break;
			const cd c1(cb[k].x * i, cb[k].y * i);
			f.h1 = f.h1 * dz + c1;
			if (i >= 2) f.h2 = f.h2 * dz + c1 * (double)(i - 1);
			if (i >= 3) f.h3 = f.h3 * dz + c1 * (double)((i - 1) * (i - 2));
		}
	}
	return f;
}

// far part of sum 1/(z - s_i) over all stars: first derivative of the
// unit-mass local expansion, evaluated in FP32 (its error only changes the
// Aberth correction, not the root it converges to)
__device__ inline cd far_pole32(const LensView& L, int node, cd z)
{
	if (L.order == 0) return cd();
	const double2 c = L.center[node];
	const double hl = L.hl[node];
	const float ih = (float)(1.0 / hl);
	const cf dz((float)(z.re - c.x) * ih, (float)(z.im - c.y) * ih);
	const float2* a = L.locu32 + (size_t)node * (L.order + 1);
	cf h;
	for (int i0 = L.order; i0 >= 1; i0 -= 8)
	{
		float2 cb[8];
#pragma unroll
		for (int k = 0; k < 8; k++) cb[k] = a[max(i0 - k, 0)];
#pragma unroll
		for (int k = 0; k < 8; k++)
		{
			const int i = i0 - k;
			if (i < 1) break;
			h = h * dz + cf(cb[k].x * i, cb[k].y * i);
		}
	}
	return cd(h.re * (double)ih, h.im * (double)ih);
}

__device__ inline double boxcar(cd z, double cre, double cim)
{
	return (-cre <= z.re && z.re <= cre && -cim <= z.im && z.im <= cim) ? 1.0 : 0.0;
}

// alpha_smooth.cuh:141-170
__device__ inline double smooth_dz(const LensView& L, cd z)
{
	if (L.rectangular) return L.approx ? -L.kappastar : -L.kappastar * boxcar(z, L.corner_re, L.corner_im);
	return L.approx ? -L.kappastar : -L.kappastar * (L.corner_abs - z.abs() >= 0 ? 1.0 : 0.0);
}

// d alpha_smooth / d zbar and d^2 alpha_smooth / d zbar^2 (alpha_smooth.cuh:185-301)
__device__ inline void smooth_dzbar(const LensView& L, cd z, cd& d1, cd& d2)
{
	const double ks = L.kappastar;
	const cd zb = z.conj();
	d1 = cd();
	d2 = cd();
	if (L.rectangular)
	{
		const cd c(L.corner_re, L.corner_im);
		if (L.approx)
		{
			// Taylor series in r^2, r = zbar / corner; the z-independent
			// coefficients (1 - e^{2 i k arg c})/k and (1 - e^{2 i k arg c}) are precomputed
			const cd r = zb / c;
			const cd r2 = r * r;
			for (int k = 0; k < L.nts; k++)
			{
				const double2 a = L.ts_a[k], b = L.ts_b[k];
				d1 = (d1 + cd(a.x, a.y)) * r2;
				d2 = (d2 + cd(b.x, b.y)) * r2;
			}
			const cd m(0, -ks / PI);
			d1 = d1 * 2.0 * m + cd(ks - 4 * ks * L.corner_arg / PI);
			d2 = d2 * zb.inv() * 2.0 * m;
		}
		else
		{
			const cd c1 = c.conj() - zb, c2 = c - zb, c3 = -c - zb, c4 = -c.conj() - zb;
			const cd m(0, -ks / PI);
			d1 = (c1.log() - c2.log() - c3.log() + c4.log()) * m - cd(ks * boxcar(z, L.corner_re, L.corner_im));
			d2 = (-(c1.inv()) + c2.inv() + c3.inv() - c4.inv()) * m;
		}
	}
	else if (!L.approx)
	{
		const double a2 = ks * L.corner_abs2;
		const cd izb = zb.inv();
		d1 = a2 * izb * izb;
		d2 = -2 * a2 * izb * izb * izb;
	}
}

// alpha_smooth.cuh:62-127
__device__ inline cd smooth_alpha(const LensView& L, cd z)
{
	const double ks = L.kappastar;
	const cd zb = z.conj();
	if (L.rectangular)
	{
		const cd c(L.corner_re, L.corner_im);
		if (L.approx)
		{
			// s(r) = sum_{i=1..T} r^i / i at r1 = zbar / c, r2 = zbar / conj(c)
			// and at -r1, -r2 (the corners -c, -conj(c)): with u = r^2,
			// s(+-r) = B(u) +- r A(u), A over the odd powers, B over the even
			const cd r1 = zb / c, r2 = zb / c.conj();
			const cd u1 = r1 * r1, u2 = r2 * r2;
			cd a1, b1, a2, b2;
			for (int i = L.taylor_smooth; i >= 1; i--)
			{
				const double q = 1.0 / i;
				if (i & 1)
				{
					a1 = a1 * u1 + q;
					a2 = a2 * u2 + q;
				}
				else
				{
					b1 = (b1 + q) * u1;
					b2 = (b2 + q) * u2;
				}
			}
			const cd o1 = r1 * a1, o2 = r2 * a2;
			const cd s1 = b1 + o1, s3 = b1 - o1, s2 = b2 + o2, s4 = b2 - o2;
			const cd l1(L.clog[0].x, L.clog[0].y), l2(L.clog[1].x, L.clog[1].y), l3(L.clog[2].x, L.clog[2].y),
				l4(L.clog[3].x, L.clog[3].y);
			cd a = (c - zb) * (l1 - s1) - (c.conj() - zb) * (l2 - s2) + (-c - zb) * (l3 - s3) - (-c.conj() - zb) * (l4 - s4);
			return a * cd(0, -ks / PI) - cd(ks * 2 * (c.re + z.re));
		}
		const cd c1 = c - zb, c2 = c.conj() - zb, c3 = -c - zb, c4 = -c.conj() - zb;
		cd a = (c1 * c1.log() - c2 * c2.log() + c3 * c3.log() - c4 * c4.log()) * cd(0, -ks / PI);
		a -= cd(ks * 2 * (c.re + z.re) * boxcar(z, c.re, c.im));
		const double hv = (c.im + z.im >= 0 ? 1.0 : 0.0) * (c.im - z.im >= 0 ? 1.0 : 0.0) * (z.re - c.re >= 0 ? 1.0 : 0.0);
		a -= cd(ks * 4 * c.re * hv);
		return a;
	}
	if (L.approx || z.abs() <= L.corner_abs) return -ks * z;
	return -ks * L.corner_abs2 * zb.inv();
}

// Witt's function f0 and the pieces the Aberth update and the 1/mu bound
// need (ccf_functions.cuh:69-86, 193-211)
struct Witt
{
	cd f0;
	double a = 0;   // 1 - kappa - d alpha_smooth / dz
	bool early = false;
	int node = 0;
	Near nr;
	Far fr;
	cd d1, d2;      // d alpha_smooth / d zbar and its zbar-derivative
};

// the rest of witt_f0 once the near-field sums of z's node are known
__device__ inline Witt witt_from(const LensView& L, cd z, double cp, double sp, int node, const Near& nr)
{
	Witt w;
	w.node = node;
	w.nr = nr;
	w.fr = l2p(L, w.node, z);
	smooth_dzbar(L, z, w.d1, w.d2);
	w.a = 1 - L.kappa - smooth_dz(L, z);
	const double ih2 = 1.0 / (w.fr.hl * w.fr.hl);
	w.f0 = cd(L.gamma) + w.nr.s2 * L.theta2 - w.fr.h2 * (L.theta2 * ih2) - w.d1.conj() - w.a * cd(cp, -sp);
	const double a0 = w.f0.abs();
	w.early = fabs(a0 * (a0 + 2 * w.a)) < 1e-9 && fabs(a0 * (a0 - 2 * w.a)) < 1e-9;
	return w;
}

__device__ inline Witt witt_f0(const LensView& L, cd z, double phi)
{
	const int node = nearest_node(L, z);
	double sp, cp;
	sincos(phi, &sp, &cp);
	return witt_from(L, z, cp, sp, node, p2p(L, node, z));
}

// f1 = d f0 / dz from the sums already formed for f0 (ccf_functions.cuh:108-121)
__device__ inline cd witt_f1(const LensView& L, cd z, const Witt& w)
{
	const double ih3 = 1.0 / (w.fr.hl * w.fr.hl * w.fr.hl);
	return w.nr.s3 * (-2 * L.theta2) - w.fr.h3 * (L.theta2 * ih3) - w.d2.conj();
}

// The device hands a launch's workgroups to its XCDs round-robin by linear
// id (y * gridDim.x + x), and every XCD has an L2 of its own. Kernels that
// walk a row's roots in Morton order along x place their blocks with
// xcd_slot / xcd_x: of the `used` blocks of row y that have work, XCD k
// takes the k-th eighth, so that one L2 holds the tree data of about an
// eighth of the field instead of all of it, and a row with few active roots
// still spreads over every XCD. Only locality depends on the dispatch order;
// each of the used blocks is taken exactly once. gridDim.x must be a
// multiple of NUM_XCD and at least the largest `used`.
constexpr int NUM_XCD = 8;

inline int xcd_round(int nx) { return (nx + NUM_XCD - 1) / NUM_XCD * NUM_XCD; }

// .x: the block's slot among its XCD's blocks of the row, .y: the row
__device__ inline int2 xcd_slot()
{
	const int per = gridDim.x / NUM_XCD;
	const int i = (blockIdx.y * gridDim.x + blockIdx.x) / NUM_XCD;
	return make_int2(i % per, i / per);
}

// whether this block is the row's first on XCD 0 (one block per row)
__device__ inline bool xcd_row_leader(int2 slot) { return slot.x == 0 && (blockIdx.y * gridDim.x + blockIdx.x) % NUM_XCD == 0; }

// the x block of slot j when the row has `used` blocks with work; `used`
// (no work) for the slots beyond them
__device__ inline int xcd_x(int j, int used)
{
	const int pu = (used + NUM_XCD - 1) / NUM_XCD;
	const int x = (int)((blockIdx.y * gridDim.x + blockIdx.x) % NUM_XCD) * pu + j;
	return j < pu && x < used ? x : used;
}

// |dz| / theta below which an Aberth update ends a root (ccf_functions.cuh:384)
constexpr double CONV_TOL = 1e-9;
// mixed form: root separation, relative to the field's half-length, below
// which an FP32 Cauchy term is redone in FP64 (an FP32 difference of two
// positions that close keeps fewer than ~12 bits)
constexpr double CLUSTER_REL = 2e-4;
// mixed form, continuation steps: |f0/f1| / theta below which a root takes
// a plain FP64 Newton step instead of an Aberth update. The root is then
// within a millionth of theta_E of a root of Witt's equation, where Newton
// converges quadratically (the next sweep's step is about 1e-12) and the
// Aberth sums, which only keep far-off iterates apart, change nothing.
constexpr double NEWTON_STEP_TOL = 1e-6;

// what the Witt evaluation of a listed root decided
enum RootState
{
	ROOT_EARLY = 0,   // the 1/mu early exit (ccf_functions.cuh:207-211)
	ROOT_FULL = 1,    // the root needs the Cauchy sums and an Aberth update
	ROOT_NEWTON = 2,  // mixed form: ended by a Newton step below CONV_TOL
	ROOT_STEPPED = 3, // mixed form: took a Newton step below NEWTON_STEP_TOL, still active
};

// Initial root of a phi step (prepare_roots_kernel, ccf_functions.cuh:284-289
// takes q = 0): the polynomial of degree q through the roots of the last
// q + 1 steps, evaluated one step on. p points at the root of the previous
// step and p[-i * back] at the one i steps earlier. The weight of that root
// is (-1)^i C(q+1, i+1) (exact in the recurrence below); the weights sum to
// one, so the extrapolation is formed from differences to the previous
// root, whose rounding scales with the roots' motion rather than with |z|.
// Above degree 2 the quadratic's value is the guard: where the two differ
// by more than PREDICT_GUARD times the root's last step, the higher degree
// is overshooting (a root turning sharply) and the quadratic is taken.
constexpr double PREDICT_GUARD = 0.1;

__device__ inline double2 predict_root(const double2* p, long long back, int q)
{
	const double2 z1 = p[0];
	double dx = 0, dy = 0;
	double w = q + 1; // (-1)^i C(q+1, i+1), from i = 0
	for (int i = 1; i <= q; i++)
	{
		w = -w * (q + 1 - i) / (i + 1);
		const double2 zi = p[-i * back];
		dx += w * (zi.x - z1.x);
		dy += w * (zi.y - z1.y);
	}
	if (q > 2)
	{
		const double2 z2 = p[-back], z3 = p[-2 * back];
		const double qx = -3 * (z2.x - z1.x) + (z3.x - z1.x), qy = -3 * (z2.y - z1.y) + (z3.y - z1.y);
		const double ex = dx - qx, ey = dy - qy, mx = z1.x - z2.x, my = z1.y - z2.y;
		if (ex * ex + ey * ey > PREDICT_GUARD * PREDICT_GUARD * (mx * mx + my * my))
		{
			dx = qx;
			dy = qy;
		}
	}
	return make_double2(z1.x + dx, z1.y + dy);
}

// 1/mu error bound at a root (ccf_functions.cuh:457-463)
__device__ inline double err_bound(const Witt& w)
{
	const double a0 = w.f0.abs();
	return fmax(fabs(a0 * (a0 + 2 * w.a)), fabs(a0 * (a0 - 2 * w.a)));
}

// lens map and magnification length scale (ccf_functions.cuh:33-45, 144-166)
// from the near- and far-field sums of z's node and the smooth-matter
// derivatives d1, d2 (smooth_dzbar) at z
__device__ inline void lens_map_from(const LensView& L, cd z, const Near& nr, const Far& fr, cd d1, cd d2, cd& wsrc,
	double& mu_scale)
{
	const double t2 = L.theta2, h = fr.hl;
	const cd a_star = (nr.s1 * t2).conj();
	const cd da_star = (nr.s2 * (-t2)).conj();
	const cd d2a_star = (nr.s3 * (2 * t2)).conj();
	const cd a_loc = (fr.h1 * (t2 / h)).conj();
	const cd da_loc = (fr.h2 * (t2 / (h * h))).conj();
	const cd d2a_loc = (fr.h3 * (t2 / (h * h * h))).conj();
	wsrc = (1 - L.kappa) * z + L.gamma * z.conj() - a_star - a_loc - smooth_alpha(L, z);
	const double dwdz = 1 - L.kappa - smooth_dz(L, z);
	const cd dwdzb = cd(L.gamma) - da_star - da_loc - d1;
	const cd d2wdzb2 = -d2a_star - d2a_loc - d2;
	const cd cct = cd(0, -2) * dwdzb.conj() * d2wdzb2;
	const cd ct = dwdz * cct + dwdzb * cct.conj();
	mu_scale = 1 / (2 * ct.abs());
}

__device__ inline void lens_map(const LensView& L, cd z, cd& wsrc, double& mu_scale)
{
	const int node = nearest_node(L, z);
	cd d1, d2;
	smooth_dzbar(L, z, d1, d2);
	lens_map_from(L, z, p2p(L, node, z), l2p(L, node, z), d1, d2, wsrc, mu_scale);
}

} // namespace ccf
