/*
 * CCF<T>: drop-in C++ mirror of weisluke/ccf's class template (include/ccf.cuh:34-1289).
 *
 * Same public fields with the same defaults (ccf.cuh:42-62), same
 * bool run(int verbose) and bool save(int verbose). The work is done by the
 * MI355X library through the C-ABI in ccf_amd.h. Extension fields follow the
 * reference's; leaving them at 0 keeps the reference's behaviour.
 *
 * Differences by design: the device is never reset (the reference calls
 * cudaDeviceReset at the start of run(), ccf.cuh:147-171), and computation is
 * FP64 for any T (T sets only the type of the public fields).
 */
#ifndef CCF_AMD_CCF_HPP
#define CCF_AMD_CCF_HPP

#include "ccf_amd.h"

#include <cstring>
#include <string>

template <typename T>
class CCF
{
public:
	T kappa_tot = static_cast<T>(0.3);
	T shear = static_cast<T>(0.3);
	T kappa_star = static_cast<T>(0.27);
	T theta_star = static_cast<T>(1);
	std::string mass_function_str = "equal";
	T m_solar = static_cast<T>(1);
	T m_lower = static_cast<T>(0.01);
	T m_upper = static_cast<T>(50);
	int rectangular = 0;
	int approx = 1;
	T safety_scale = static_cast<T>(1.37);
	int num_stars = 137;
	std::string starfile = "";
	int num_phi = 100;
	int num_branches = 1;
	int random_seed = 0;
	int write_stars = 1;
	int write_critical_curves = 1;
	int write_caustics = 1;
	int write_mu_length_scales = 0;
	std::string outfile_prefix = "./";

	/* extensions (ccf_params in ccf_amd.h) */
	int fmm_order = 0;
	int direct = 0;
	int binomial_exact = 0;
	int precision = 0;
	int init_iters = 0;
	int step_iters = 0;
	int roots_per_thread = 0;
	int timing = 0;
	int predictor = 0;
	int root_fmm = 0;
	int root_fmm_order = 0;
	int root_fmm_leaf = 0;
	int autotune = 0;

	CCF() : ctx_(ccf_create()) {}
	~CCF() { ccf_destroy(ctx_); }
	CCF(const CCF&) = delete;
	CCF& operator=(const CCF&) = delete;

	bool run(int verbose)
	{
		if (!ctx_) return false;
		ccf_params p = params();
		if (ccf_set_params(ctx_, &p)) return false;
		const bool ok = ccf_run(ctx_, verbose) == 0;
		pull();
		return ok;
	}

	bool save(int verbose)
	{
		return ctx_ && ccf_save(ctx_, verbose) == 0;
	}

	/* derived values and timings of the last run */
	ccf_info info() const
	{
		ccf_info i{};
		if (ctx_) ccf_get_info(ctx_, &i);
		return i;
	}

	ccf_ctx* handle() { return ctx_; }

	ccf_params params() const
	{
		ccf_params p;
		ccf_default_params(&p);
		p.kappa_tot = kappa_tot;
		p.shear = shear;
		p.kappa_star = kappa_star;
		p.theta_star = theta_star;
		copy(p.mass_function, sizeof(p.mass_function), mass_function_str);
		p.m_solar = m_solar;
		p.m_lower = m_lower;
		p.m_upper = m_upper;
		p.rectangular = rectangular;
		p.approx = approx;
		p.safety_scale = safety_scale;
		p.num_stars = num_stars;
		copy(p.starfile, sizeof(p.starfile), starfile);
		p.num_phi = num_phi;
		p.num_branches = num_branches;
		p.random_seed = random_seed;
		p.write_stars = write_stars;
		p.write_critical_curves = write_critical_curves;
		p.write_caustics = write_caustics;
		p.write_mu_length_scales = write_mu_length_scales;
		copy(p.outfile_prefix, sizeof(p.outfile_prefix), outfile_prefix);
		p.fmm_order = fmm_order;
		p.direct = direct;
		p.binomial_exact = binomial_exact;
		p.precision = precision;
		p.init_iters = init_iters;
		p.step_iters = step_iters;
		p.roots_per_thread = roots_per_thread;
		p.timing = timing;
		p.predictor = predictor;
		p.root_fmm = root_fmm;
		p.root_fmm_order = root_fmm_order;
		p.root_fmm_leaf = root_fmm_leaf;
		p.autotune = autotune;
		return p;
	}

private:
	ccf_ctx* ctx_;

	static void copy(char* dst, size_t n, const std::string& s)
	{
		std::strncpy(dst, s.c_str(), n - 1);
		dst[n - 1] = 0;
	}

	/* fields the reference's run() assigns (set_param on its own members) */
	void pull()
	{
		ccf_params p;
		if (ccf_get_params(ctx_, &p)) return;
		kappa_star = static_cast<T>(p.kappa_star);
		theta_star = static_cast<T>(p.theta_star);
		m_lower = static_cast<T>(p.m_lower);
		m_upper = static_cast<T>(p.m_upper);
		rectangular = p.rectangular;
		num_stars = p.num_stars;
		random_seed = p.random_seed;
	}
};

#endif
