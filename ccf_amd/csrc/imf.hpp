// Initial mass functions as piecewise power laws, usable on host and device.
//
// weisluke/ccf implements each mass function as a class with virtual
// mass()/mean_mass()/... (include/mass_functions/*.cuh). Here every
// continuous IMF (uniform = slope 0, Salpeter = slope -2.35, Kroupa =
// slopes -0.3/-1.3/-2.3 broken at 0.08 and 0.5 m_solar) is one table of
// up to three normalised segments built on the host; a kernel only needs
// the table to invert the CDF. equal and optical_depth are special kinds.
#pragma once

#include <hip/hip_runtime.h>

#include <cmath>
#include <cstring>

namespace ccf
{

struct Imf
{
	int kind = 0;  // 0 equal (m = 1), 1 piecewise power law, 2 optical depth, -1 unknown
	int nseg = 0;
	double ml = 0, mu = 0;
	double lo[3] = {0, 0, 0}, hi[3] = {0, 0, 0}, slope[3] = {0, 0, 0};
	double b[3] = {0, 0, 0};     // density b_i m^slope_i on [lo_i, hi_i]
	double cum[3] = {0, 0, 0};   // probability below lo_i
	double prob[3] = {0, 0, 0};  // probability of segment i
};

// integral of b x^a over [x1, x2] (power_law.cuh:30-46)
__host__ __device__ inline double imf_pint(double x1, double x2, double a, double b)
{
	if (a != -1) return b * (pow(x2, a + 1) - pow(x1, a + 1)) / (a + 1);
	return b * (log(x2) - log(x1));
}

// integral of b x^a ln x over [x1, x2] (power_law.cuh:61-77)
__host__ __device__ inline double imf_plogint(double x1, double x2, double a, double b)
{
	if (a != -1) return b * (pow(x2, a + 1) * ((a + 1) * log(x2) - 1) - pow(x1, a + 1) * ((a + 1) * log(x1) - 1)) / ((a + 1) * (a + 1));
	return b * (log(x2) * log(x2) - log(x1) * log(x1)) / 2;
}

// x with integral of b t^a over [x1, x] equal to p (power_law.cuh:90-104)
__host__ __device__ inline double imf_pinv(double p, double x1, double a, double b)
{
	if (a != -1) return pow(p * (a + 1) / b + pow(x1, a + 1), 1 / (a + 1));
	return x1 * exp(p / b);
}

// Build the table. ml and mu are the arbitrary-unit cutoffs the reference
// passes (already multiplied by m_solar, ccf.cuh:359-368).
inline Imf make_imf(const char* name, double ml, double mu, double m_solar)
{
	Imf f;
	f.ml = ml;
	f.mu = mu;
	if (!std::strcmp(name, "equal")) { f.kind = 0; return f; }
	if (!std::strcmp(name, "optical_depth")) { f.kind = 2; return f; }
	double slopes[3];
	double brk[4];
	int n;
	if (!std::strcmp(name, "uniform") || !std::strcmp(name, "salpeter"))
	{
		n = 1;
		slopes[0] = name[0] == 'u' ? 0.0 : -2.35;
		brk[0] = ml;
		brk[1] = mu;
	}
	else if (!std::strcmp(name, "kroupa"))
	{
		// segments of the broken power law that intersect [ml, mu], with the
		// reference's case boundaries (kroupa.cuh:50-125)
		const double s[3] = {-0.3, -1.3, -2.3};
		const double m0 = 0.08 * m_solar, m1 = 0.5 * m_solar;
		const bool use[3] = {ml < m0, mu >= m0 && ml < m1, mu >= m1};
		const double edge[4] = {ml, m0, m1, mu};
		n = 0;
		for (int i = 0; i < 3; i++)
		{
			if (!use[i]) continue;
			slopes[n] = s[i];
			brk[n] = i == 0 ? ml : (n == 0 ? ml : edge[i]);
			brk[n + 1] = (i == 2 || !use[i + 1]) ? mu : edge[i + 1];
			n++;
		}
	}
	else
	{
		f.kind = -1;
		return f;
	}
	f.kind = 1;
	f.nseg = n;
	if (ml == mu) return f;
	// unnormalised densities continuous at the breaks, then normalise
	double rel = 1, tot = 0;
	for (int i = 0; i < n; i++)
	{
		if (i > 0) rel *= pow(brk[i], slopes[i - 1] - slopes[i]);
		f.lo[i] = brk[i];
		f.hi[i] = brk[i + 1];
		f.slope[i] = slopes[i];
		f.b[i] = rel;
		f.prob[i] = imf_pint(brk[i], brk[i + 1], slopes[i], rel);
		tot += f.prob[i];
	}
	double c = 0;
	for (int i = 0; i < n; i++)
	{
		f.b[i] /= tot;
		f.prob[i] /= tot;
		f.cum[i] = c;
		c += f.prob[i];
	}
	return f;
}

// mass for a uniform deviate p in (0, 1] (generate_star_field_kernel, star.cuh:105)
__host__ __device__ inline double imf_mass(const Imf& f, double p)
{
	if (f.kind == 0) return 1;
	if (f.ml == f.mu) return f.ml;
	if (f.kind == 2) return p <= f.mu / (f.ml + f.mu) ? f.ml : f.mu; // optical_depth.cuh:31-42
	int i = 0;
	while (i + 1 < f.nseg && p >= f.cum[i] + f.prob[i]) i++;
	return imf_pinv(p - f.cum[i], f.lo[i], f.slope[i], f.b[i]);
}

// which: 1 <m>, 2 <m^2>, 3 <m^2 ln m>
inline double imf_moment(const Imf& f, int which)
{
	if (f.kind == 0) return which == 3 ? 0.0 : 1.0;
	if (f.ml == f.mu) return which == 1 ? f.ml : (which == 2 ? f.ml * f.ml : f.ml * f.ml * std::log(f.ml));
	if (f.kind == 2)
	{
		const double a = f.ml, b = f.mu;
		if (which == 1) return 2 * a * b / (a + b);
		if (which == 2) return a * b;
		return a * b * (a * std::log(a) + b * std::log(b)) / (a + b);
	}
	double m = 0;
	for (int i = 0; i < f.nseg; i++)
	{
		if (which == 1) m += imf_pint(f.lo[i], f.hi[i], f.slope[i] + 1, f.b[i]);
		else if (which == 2) m += imf_pint(f.lo[i], f.hi[i], f.slope[i] + 2, f.b[i]);
		else m += imf_plogint(f.lo[i], f.hi[i], f.slope[i] + 2, f.b[i]);
	}
	return m;
}

} // namespace ccf
