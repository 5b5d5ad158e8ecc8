// Host side of ccf_amd: the CCF pipeline of weisluke/ccf include/ccf.cuh
// (validation, derived parameters, star field, stage order, output files)
// behind the C-ABI declared in include/ccf_amd.h. Device work is in
// tree.hip and solve.hip.
#include "internal.hpp"

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <filesystem>
#include <fstream>
#include <iostream>
#include <limits>
#include <random>
#include <sstream>
#include <string>
#include <vector>

using ccf::cd;

// ---------------------------------------------------------------------------
// context helpers

ccf::LensView ccf_ctx::lens() const
{
	ccf::LensView L{};
	L.kappa = P.kappa_tot;
	L.gamma = P.shear;
	L.theta = theta;
	L.theta2 = theta * theta;
	L.kappastar = kappa_star;
	L.rectangular = rectangular;
	L.approx = P.approx;
	L.taylor_smooth = taylor_smooth;
	L.corner_re = corner.re;
	L.corner_im = corner.im;
	L.corner_abs = corner.abs();
	L.corner_abs2 = corner.abs() * corner.abs();
	L.corner_arg = corner.arg();
	L.order = tree.order;
	L.nstars = num_stars;
	L.nts = nts;
	{
		const ccf::cd c(corner.re, corner.im);
		const ccf::cd lg[4] = {c.log(), c.conj().log(), (-c).log(), (-c).conj().log()};
		for (int k = 0; k < 4; k++) L.clog[k] = make_double2(lg[k].re, lg[k].im);
	}
	L.ts_a = ts_a;
	L.ts_b = ts_b;
	L.spos = spos;
	L.smass = smass;
	L.center = tree.center;
	L.hl = tree.hl;
	L.child = tree.child;
	L.nf = tree.nf;
	L.nnf = tree.nnf;
	L.nbox = tree.nbox;
	L.nf_off = tree.nf_off;
	L.nf_cnt = tree.nf_cnt;
	L.nfs = tree.nfs;
	L.loc = tree.loc;
	L.loc32 = tree.loc32;
	L.locu32 = tree.locu32;
	L.root = 0;
	L.lookup = tree.lookup;
	L.lookup_n = 1 << tree.lookup_level;
	L.lookup_x0 = tree.lookup_x0;
	L.lookup_inv = tree.lookup_inv;
	return L;
}

void ccf_ctx::fail(const std::string& msg)
{
	last_error = msg;
	std::cerr << msg;
}

bool ccf_ctx::hip_ok(hipError_t e, const char* what, const char* file, int line)
{
	if (e == hipSuccess) return true;
	std::ostringstream os;
	os << "HIP error check for " << what << " failed at " << file << ":" << line << "\n";
	os << "Error code: " << (int)e << " (" << hipGetErrorString(e) << ")\n";
	fail(os.str());
	return false;
}

hipError_t ccf_ctx::dalloc(void** p, size_t bytes)
{
	*p = nullptr;
	if (bytes == 0) bytes = 1;
	const auto it = pool.lower_bound(bytes);
	if (it != pool.end() && it->first <= bytes + bytes / 4 + (size_t(1) << 16))
	{
		*p = it->second;
		pool.erase(it);
		return hipSuccess;
	}
	hipError_t e = hipMalloc(p, bytes);
	if (e == hipErrorOutOfMemory && !pool.empty())
	{
		(void)hipGetLastError();
		pool_release();
		e = hipMalloc(p, bytes);
	}
	if (e == hipSuccess) owned[*p] = bytes;
	return e;
}

void ccf_ctx::dfree(void* p)
{
	if (!p) return;
	const auto it = owned.find(p);
	if (it == owned.end())
	{
		(void)hipFree(p);
		return;
	}
	pool.emplace(it->second, p);
}

void ccf_ctx::pool_release()
{
	if (pool.empty()) return;
	// work queued on the streams may still use a pooled buffer
	if (stream) (void)hipStreamSynchronize(stream);
	if (err_stream) (void)hipStreamSynchronize(err_stream);
	for (hipStream_t g : group_streams)
		if (g) (void)hipStreamSynchronize(g);
	for (auto& [n, p] : pool)
	{
		owned.erase(p);
		(void)hipFree(p);
	}
	pool.clear();
}

void ccf_ctx::free_device()
{
	// a failed solve can leave error-pass or continuation work queued on
	// the other streams
	if (err_stream) (void)hipStreamSynchronize(err_stream);
	for (hipStream_t g : group_streams)
		if (g) (void)hipStreamSynchronize(g);
	auto F = [this](auto*& p) { dfree(p); p = nullptr; };
	ccf::DeviceTree& T = tree;
	F(T.center); F(T.hl); F(T.parent); F(T.child); F(T.sstart); F(T.scount);
	F(T.nbr); F(T.nnbr); F(T.sl); F(T.nsl); F(T.dl); F(T.ndl); F(T.split); F(T.rank);
	F(T.field); F(T.nf); F(T.nnf); F(T.nbox); F(T.nf_off); F(T.nf_cnt); F(T.nfs); F(T.mp); F(T.loc); F(T.loc32); F(T.mpu); F(T.locu); F(T.locu32); F(T.binom); F(T.lookup);
	T = ccf::DeviceTree{};
	F(spos); F(smass); F(spos_tmp); F(smass_tmp); F(spos32); F(ts_a); F(ts_b); F(ones);
	F(ccs_init); F(phase); F(dseeds); F(dorder);
	if (results_borrowed)
	{
		ccs = caustics = nullptr;
		mu_scales = nullptr;
		results_borrowed = false;
	}
	else
	{
		F(ccs); F(caustics); F(mu_scales);
	}
	ccf::free_sweep(this, sw);
	for (ccf::SweepBufs& b : swg) ccf::free_sweep(this, b);
	F(counters);
	if (host_counts) (void)hipHostFree(host_counts);
	host_counts = nullptr;
	// the events themselves live until ccf_destroy
	timers.clear();
	witt_timers.clear();
	tail_timers.clear();
	events_used = 0;
	loaded = prepared = have_roots = have_ccs = have_caustics = have_mu = false;
	caustics_done = mu_done = false;
}

namespace
{

template <typename V>
void set_param(const char* name, V& param, V what, int verbose, bool newline = false)
{
	param = what;
	if (verbose >= 2)
	{
		std::cout << name << " set to: " << param << "\n";
		if (newline) std::cout << "\n";
	}
}

std::string dmin_str()
{
	std::ostringstream os;
	os << std::numeric_limits<double>::min();
	return os.str();
}

bool known_mass_function(const std::string& s)
{
	return s == "equal" || s == "uniform" || s == "salpeter" || s == "kroupa" || s == "optical_depth";
}

// ccf.cuh:218-346
bool check_input_params(ccf_ctx* c)
{
	const ccf_params& P = c->P;
	const bool nofile = P.starfile[0] == 0;
	const double dmin = std::numeric_limits<double>::min();
	auto err = [c](const std::string& m) { c->fail(m); return false; };
	if (P.kappa_tot < dmin) return err("Error. kappa_tot must be >= " + dmin_str() + "\n");
	if (P.kappa_star < dmin) return err("Error. kappa_star must be >= " + dmin_str() + "\n");
	if (nofile && P.kappa_star > P.kappa_tot) return err("Error. kappa_star must be <= kappa_tot\n");
	if (nofile && P.theta_star < dmin) return err("Error. theta_star must be >= " + dmin_str() + "\n");
	if (nofile && !known_mass_function(P.mass_function))
		return err("Error. mass_function must be equal, uniform, Salpeter, Kroupa, or optical_depth.\n");
	if (nofile && P.m_solar < dmin) return err("Error. m_solar must be >= " + dmin_str() + "\n");
	if (nofile && P.m_lower < dmin) return err("Error. m_lower must be >= " + dmin_str() + "\n");
	if (nofile && P.m_upper < P.m_lower) return err("Error. m_upper must be >= m_lower.\n");
	if (nofile && P.rectangular != 0 && P.rectangular != 1) return err("Error. rectangular must be 1 (rectangular) or 0 (circular).\n");
	if (P.approx != 0 && P.approx != 1) return err("Error. approx must be 1 (approximate) or 0 (exact).\n");
	if (P.safety_scale < 1.1) return err("Error. safety_scale must be >= 1.1\n");
	if (P.num_stars < 1) return err("Error. num_stars must be an integer > 0\n");
	if (P.num_phi < 1 || P.num_phi % 2 != 0) return err("Error. num_phi must be an even integer > 0\n");
	if (P.num_branches < 1) return err("Error. num_branches must be an integer > 0\n");
	if (P.num_phi % (2 * P.num_branches) != 0) return err("Error. num_phi must be a multiple of 2 * num_branches\n");
	if (P.write_stars != 0 && P.write_stars != 1) return err("Error. write_stars must be 1 (true) or 0 (false).\n");
	if (P.write_critical_curves != 0 && P.write_critical_curves != 1) return err("Error. write_critical_curves must be 1 (true) or 0 (false).\n");
	if (P.write_caustics != 0 && P.write_caustics != 1) return err("Error. write_caustics must be 1 (true) or 0 (false).\n");
	if (P.write_mu_length_scales != 0 && P.write_mu_length_scales != 1) return err("Error. write_mu_length_scales must be 1 (true) or 0 (false).\n");
	if (P.fmm_order < 0 || P.fmm_order > ccf::MAX_EXPANSION_ORDER || (P.fmm_order > 0 && P.fmm_order < 3))
		return err("Error. fmm_order must be 0 (automatic) or in [3, " + std::to_string(ccf::MAX_EXPANSION_ORDER) + "]\n");
	if (P.direct != 0 && P.direct != 1) return err("Error. direct must be 1 (true) or 0 (false).\n");
	if (P.precision != 0 && P.precision != 1) return err("Error. precision must be 0 (fp64) or 1 (mixed).\n");
	if (P.roots_per_thread != 0 && P.roots_per_thread != 1 && P.roots_per_thread != 2 && P.roots_per_thread != 4)
		return err("Error. roots_per_thread must be 0 (automatic), 1, 2 or 4.\n");
	if (P.init_iters < 0 || P.step_iters < 0) return err("Error. iteration caps must be >= 0\n");
	if (P.predictor < 0 || P.predictor > ccf::MAX_PREDICTOR) return err("Error. predictor must be 0, 1, 2, 3 or 4.\n");
	if (P.root_fmm != 0 && P.root_fmm != 1) return err("Error. root_fmm must be 1 (true) or 0 (false).\n");
	if (P.root_fmm_order < 0 || P.root_fmm_order > ccf::MAX_ROOT_FMM_ORDER || (P.root_fmm_order > 0 && P.root_fmm_order < 3))
		return err("Error. root_fmm_order must be 0 (automatic) or in [3, " + std::to_string(ccf::MAX_ROOT_FMM_ORDER) + "]\n");
	if (P.root_fmm_leaf < 0) return err("Error. root_fmm_leaf must be >= 0\n");
	if (P.autotune != 0 && P.autotune != 1) return err("Error. autotune must be 0 or 1\n");
	return true;
}

// star.cuh:124-161
void star_params(int n, int rect, cd corner, double theta, const std::vector<double>& st, double& kappastar,
	double& ml, double& mu, double& m1, double& m2, double& m2l)
{
	ml = std::numeric_limits<double>::max();
	mu = std::numeric_limits<double>::min();
	double t = 0, t2 = 0, t2l = 0;
	for (int i = 0; i < n; i++)
	{
		const double m = st[3 * i + 2];
		t += m;
		t2 += m * m;
		t2l += m * m * std::log(m);
		ml = std::min(ml, m);
		mu = std::max(mu, m);
	}
	m1 = t / n;
	m2 = t2 / n;
	m2l = t2l / n;
	if (rect) kappastar = t * ccf::PI * theta * theta / (4 * corner.re * corner.im);
	else kappastar = t * theta * theta / (corner.abs() * corner.abs());
}

// star.cuh:175-434
bool read_star_file(ccf_ctx* c, const std::string& fname)
{
	std::filesystem::path path = fname;
	std::vector<double>& st = c->stars_in;
	if (path.extension() == ".bin")
	{
		std::error_code ec;
		const auto fsize = std::filesystem::file_size(fname, ec);
		if (ec) { c->fail("Error determining size of star input file " + fname + "\n"); return false; }
		std::ifstream in(fname, std::ios::binary);
		if (!in.is_open()) { c->fail("Error. Failed to open file " + fname + "\n"); return false; }
		int32_t n = 0, rect = 0;
		in.read((char*)&n, 4);
		if (n < 1) { c->fail("Error. Invalid num_stars input. num_stars must be an integer > 0\n"); return false; }
		in.read((char*)&rect, 4);
		if (rect != 0 && rect != 1)
		{
			c->fail("Error. Invalid rectangular input. rectangular must be 1 (rectangular) or 0 (circular).\n");
			return false;
		}
		const bool dbl = fsize == 8 + 16 + 8 + (uint64_t)n * 24, flt = fsize == 8 + 8 + 4 + (uint64_t)n * 12;
		if (!dbl && !flt)
		{
			c->fail("Error. Star input file " + fname + " does not contain validly formatted single or double precision stars and accompanying information.\n");
			return false;
		}
		// sized only once the file holds n stars: a header's n alone allocates nothing
		st.assign(3 * (size_t)n, 0.0);
		if (dbl)
		{
			double cc[2], t;
			in.read((char*)cc, 16);
			in.read((char*)&t, 8);
			in.read((char*)st.data(), (size_t)n * 24);
			c->corner = cd(cc[0], cc[1]);
			c->theta = t;
		}
		else
		{
			float cc[2], t;
			in.read((char*)cc, 8);
			in.read((char*)&t, 4);
			std::vector<float> buf(3 * (size_t)n);
			in.read((char*)buf.data(), buf.size() * 4);
			for (size_t i = 0; i < buf.size(); i++) st[i] = buf[i];
			c->corner = cd(cc[0], cc[1]);
			c->theta = t;
		}
		if (c->corner.re < 0 || c->corner.im < 0)
		{
			c->fail("Error. Real and imaginary parts of the corner of the star field must both be >= 0\n");
			return false;
		}
		if (c->theta < std::numeric_limits<double>::min())
		{
			c->fail("Error. theta_star must be >= " + dmin_str() + "\n");
			return false;
		}
		c->num_stars = n;
		c->rectangular = rect;
		return true;
	}
	if (path.extension() == ".txt")
	{
		std::ifstream in(fname);
		if (!in.is_open()) { c->fail("Error. Failed to open file " + fname + "\n"); return false; }
		st.clear();
		double x1, x2, m, mx1 = 0, mx2 = 0, mr = 0, tot = 0, I = 0;
		while (in >> x1 >> x2 >> m)
		{
			st.push_back(x1);
			st.push_back(x2);
			st.push_back(m);
			tot += m;
			I += m * (x1 * x1 + x2 * x2);
			mx1 = std::max(mx1, std::abs(x1));
			mx2 = std::max(mx2, std::abs(x2));
			mr = std::max(mr, std::hypot(x1, x2));
		}
		const int n = (int)(st.size() / 3);
		if (n < 1) { c->fail("Error. Invalid num_stars input. num_stars must be an integer > 0\n"); return false; }
		// circular or rectangular by moment of inertia (star.cuh:411-429)
		const double Ic = tot * mr * mr / 2, Ir = tot * (mx1 * mx1 + mx2 * mx2) / 3;
		if (I / Ic > I / Ir) { c->rectangular = 0; c->corner = cd(mr, 0); }
		else { c->rectangular = 1; c->corner = cd(mx1, mx2); }
		c->theta = 1;
		c->num_stars = n;
		return true;
	}
	c->fail("Error. Star input file " + fname + " is not a .bin or .txt file.\n");
	return false;
}

// star.cuh:498-534; stars are {x1, x2, m} triples
bool write_star_file(ccf_ctx* c, const std::string& fname, const std::vector<double>& st)
{
	std::filesystem::path path = fname;
	if (path.extension() != ".bin") { c->fail("Error. Star file " + fname + " is not a .bin file.\n"); return false; }
	std::ofstream out(fname, std::ios::binary);
	if (!out.is_open()) { c->fail("Error. Failed to open file " + fname + "\n"); return false; }
	const int32_t n = c->num_stars, r = c->rectangular;
	out.write((const char*)&n, 4);
	out.write((const char*)&r, 4);
	const double cc[2] = {r ? c->corner.re : c->corner.abs(), r ? c->corner.im : 0.0};
	out.write((const char*)cc, 16);
	out.write((const char*)&c->theta, 8);
	out.write((const char*)st.data(), st.size() * 8);
	return true;
}

// array_functions.cuh:280-306
bool write_array(ccf_ctx* c, const std::string& fname, const void* vals, size_t elem, int nrows, int ncols)
{
	std::filesystem::path path = fname;
	if (path.extension() != ".bin") { c->fail("Error. File " + fname + " is not a .bin file.\n"); return false; }
	std::ofstream out(fname, std::ios::binary);
	if (!out.is_open()) { c->fail("Error. Failed to open file " + fname + "\n"); return false; }
	out.write((const char*)&nrows, 4);
	out.write((const char*)&ncols, 4);
	out.write((const char*)vals, elem * (size_t)nrows * ncols);
	return true;
}

// ccf.cuh:348-497
bool calculate_derived_params(ccf_ctx* c, int verbose)
{
	ccf_params& P = c->P;
	const bool nofile = P.starfile[0] == 0;
	if (verbose >= 3) std::cout << "Calculating derived parameters...\n";
	c->from_file = !nofile;
	if (nofile)
	{
		double ml = P.m_lower, mu = P.m_upper;
		if (!std::strcmp(P.mass_function, "equal")) { ml = 1; mu = 1; }
		else { ml *= P.m_solar; mu *= P.m_solar; }
		set_param("m_lower", P.m_lower, ml, verbose);
		set_param("m_upper", P.m_upper, mu, verbose);
		c->imf = ccf::make_imf(P.mass_function, P.m_lower, P.m_upper, P.m_solar);
		set_param("mean_mass", c->mean_mass, ccf::imf_moment(c->imf, 1), verbose);
		set_param("mean_mass2", c->mean_mass2, ccf::imf_moment(c->imf, 2), verbose);
		set_param("mean_mass2_ln_mass", c->mean_mass2_ln_mass, ccf::imf_moment(c->imf, 3), verbose);
		c->num_stars = P.num_stars;
		c->rectangular = P.rectangular;
		c->theta = P.theta_star;
		c->kappa_star = P.kappa_star;
	}
	else
	{
		if (!read_star_file(c, P.starfile))
		{
			c->fail(std::string("Error. Unable to read star field parameters from file ") + P.starfile + "\n");
			return false;
		}
		double ks, ml, mu, m1, m2, m2l;
		star_params(c->num_stars, c->rectangular, c->corner, c->theta, c->stars_in, ks, ml, mu, m1, m2, m2l);
		c->kappa_star = ks;
		c->mean_mass = m1;
		c->mean_mass2 = m2;
		c->mean_mass2_ln_mass = m2l;
		P.num_stars = c->num_stars;
		P.rectangular = c->rectangular;
		P.theta_star = c->theta;
		P.kappa_star = ks;
		P.m_lower = ml;
		P.m_upper = mu;
		if (verbose >= 2)
		{
			std::cout << "num_stars set to: " << c->num_stars << "\nrectangular set to: " << c->rectangular << "\n";
			std::cout << "corner set to: (" << c->corner.re << ", " << c->corner.im << ")\ntheta_star set to: " << c->theta << "\n";
			std::cout << "kappa_star set to: " << ks << "\n";
		}
		if (ks > P.kappa_tot) std::cerr << "Warning. kappa_star > kappa_tot\n";
	}
	set_param("mu_ave", c->mu_ave, 1 / ((1 - P.kappa_tot) * (1 - P.kappa_tot) - P.shear * P.shear), verbose);
	if (nofile)
	{
		cd cr(1 / std::abs(1 - P.kappa_tot + P.shear), 1 / std::abs(1 - P.kappa_tot - P.shear));
		if (c->rectangular)
		{
			cr = cd(std::sqrt(cr.re / cr.im), std::sqrt(cr.im / cr.re));
			cr *= std::sqrt(ccf::PI * c->theta * c->theta * c->num_stars * c->mean_mass / (4 * c->kappa_star));
		}
		else
		{
			cr = cr / cr.abs();
			cr *= std::sqrt(c->theta * c->theta * c->num_stars * c->mean_mass / c->kappa_star);
		}
		c->corner = cr;
		if (verbose >= 2) std::cout << "corner set to: (" << cr.re << ", " << cr.im << ")\n";
	}
	c->alpha_error = c->theta * 0.0000001; // 1e-7 Einstein radii
	// degree of the Taylor series of the rectangular smooth sheet (ccf.cuh:441-464)
	int ts = 1;
	const double cabs = c->corner.abs(), carg = c->corner.arg(), ss = P.safety_scale;
	while (c->kappa_star / ccf::PI * 4 / (ts + 1) * cabs * (ss + 1) / (ss - 1) * std::pow(1 / ss, ts + 1) > c->alpha_error && ts <= ccf::MAX_TAYLOR_SMOOTH)
		ts += 2;
	while ((std::fmod(carg * (ts - 1), ccf::PI) < 0.1 * ccf::PI || std::fmod(carg * (ts - 1), ccf::PI) > 0.9 * ccf::PI) && ts <= ccf::MAX_TAYLOR_SMOOTH)
		ts += 2;
	c->taylor_smooth = ts;
	if (verbose >= 2 && c->rectangular && P.approx) std::cout << "taylor_smooth set to: " << ts << "\n";
	if (c->rectangular && ts > ccf::MAX_TAYLOR_SMOOTH)
	{
		c->fail("Error. taylor_smooth must be <= " + std::to_string(ccf::MAX_TAYLOR_SMOOTH) + "\n");
		return false;
	}
	// number of roots (ccf.cuh:469-491)
	if (c->rectangular) c->num_roots = P.approx ? 2 * c->num_stars + ts - 1 : 2 * c->num_stars;
	else c->num_roots = P.approx ? 2 * c->num_stars : 2 * c->num_stars + 2;
	if (verbose >= 2) std::cout << "num_roots set to: " << c->num_roots << "\n";
	return true;
}

// ccf.cuh:499-730: device buffers, star field, seeds
bool populate(ccf_ctx* c, int verbose)
{
	ccf_params& P = c->P;
	const int N = c->num_stars;
	hipStream_t s = c->stream;
	if (!CCF_HIP(c, c->dmalloc(&c->spos, (size_t)N * sizeof(double2))) || !CCF_HIP(c, c->dmalloc(&c->smass, (size_t)N * sizeof(double))) ||
		!CCF_HIP(c, c->dmalloc(&c->spos_tmp, (size_t)N * sizeof(double2))) || !CCF_HIP(c, c->dmalloc(&c->smass_tmp, (size_t)N * sizeof(double))) ||
		!CCF_HIP(c, c->dmalloc(&c->spos32, (size_t)N * sizeof(float2))))
		return false;
	if (!c->from_file)
	{
		if (verbose >= 1) std::cout << "Generating star field...\n";
		auto t0 = std::chrono::steady_clock::now();
		c->random_seed = P.random_seed;
		while (c->random_seed == 0) // ccf.cuh:619-622
			c->random_seed = (int)std::chrono::system_clock::now().time_since_epoch().count();
		if (verbose >= 2) std::cout << "random_seed set to: " << c->random_seed << "\n";
		P.random_seed = c->random_seed;
		if (!ccf::generate_stars_dev(c)) return false;
		std::vector<double2> pos(N);
		std::vector<double> m(N);
		if (!CCF_HIP(c, hipMemcpy(pos.data(), c->spos, (size_t)N * sizeof(double2), hipMemcpyDeviceToHost)) ||
			!CCF_HIP(c, hipMemcpy(m.data(), c->smass, (size_t)N * sizeof(double), hipMemcpyDeviceToHost)))
			return false;
		c->stars_in.resize(3 * (size_t)N);
		for (int i = 0; i < N; i++)
		{
			c->stars_in[3 * i] = pos[i].x;
			c->stars_in[3 * i + 1] = pos[i].y;
			c->stars_in[3 * i + 2] = m[i];
		}
		auto t1 = std::chrono::steady_clock::now();
		if (verbose >= 1)
			std::cout << "Done generating star field. Elapsed time: " << std::chrono::duration<double>(t1 - t0).count() << " seconds.\n\n";
	}
	else
	{
		c->random_seed = 0;
		P.random_seed = 0;
		std::vector<double2> pos(N);
		std::vector<double> m(N);
		for (int i = 0; i < N; i++)
		{
			pos[i] = make_double2(c->stars_in[3 * i], c->stars_in[3 * i + 1]);
			m[i] = c->stars_in[3 * i + 2];
		}
		if (!CCF_HIP(c, hipMemcpyAsync(c->spos, pos.data(), (size_t)N * sizeof(double2), hipMemcpyHostToDevice, s)) ||
			!CCF_HIP(c, hipMemcpyAsync(c->smass, m.data(), (size_t)N * sizeof(double), hipMemcpyHostToDevice, s)) ||
			!CCF_HIP(c, hipStreamSynchronize(s)))
			return false;
	}
	// actual star-field statistics (ccf.cuh:676-684)
	star_params(N, c->rectangular, c->corner, c->theta, c->stars_in, c->kappa_star_actual, c->m_lower_actual, c->m_upper_actual,
		c->mean_mass_actual, c->mean_mass2_actual, c->mean_mass2_ln_mass_actual);
	if (verbose >= 2) std::cout << "kappa_star_actual set to: " << c->kappa_star_actual << "\n";
	// corner from the actual mean mass (ccf.cuh:686-700)
	if (!c->from_file)
	{
		cd cr = c->corner;
		if (c->rectangular)
		{
			cr = cd(std::sqrt(cr.re / cr.im), std::sqrt(cr.im / cr.re));
			cr *= std::sqrt(ccf::PI * c->theta * c->theta * N * c->mean_mass_actual / (4 * c->kappa_star));
		}
		else
		{
			cr = cr / cr.abs();
			cr *= std::sqrt(c->theta * c->theta * N * c->mean_mass_actual / c->kappa_star);
		}
		c->corner = cr;
	}
	// seeds at z_i + 1, z_i - 1 and on the circle |z| = |corner| (ccf.cuh:711-726)
	const int R = c->num_roots;
	c->seeds.assign(2 * (size_t)R, 0.0);
	for (int i = 0; i < N; i++)
	{
		c->seeds[2 * i] = c->stars_in[3 * i] + 1;
		c->seeds[2 * i + 1] = c->stars_in[3 * i + 1];
		c->seeds[2 * (i + N)] = c->stars_in[3 * i] - 1;
		c->seeds[2 * (i + N) + 1] = c->stars_in[3 * i + 1];
	}
	const int nx = R - 2 * N;
	for (int i = 0; i < nx; i++)
	{
		const double a = 2 * ccf::PI / nx * i;
		c->seeds[2 * (2 * N + i)] = c->corner.abs() * std::cos(a);
		c->seeds[2 * (2 * N + i) + 1] = c->corner.abs() * std::sin(a);
	}
	// order in which the roots are processed (not their indices, which stay
	// the reference's): by the Morton key of their seeds, so that the lanes
	// of a wave evaluate roots in neighbouring tree leaves
	{
		double x0 = 1e300, x1 = -1e300, y0 = 1e300, y1 = -1e300;
		for (int i = 0; i < R; i++)
		{
			x0 = std::min(x0, c->seeds[2 * i]);
			x1 = std::max(x1, c->seeds[2 * i]);
			y0 = std::min(y0, c->seeds[2 * i + 1]);
			y1 = std::max(y1, c->seeds[2 * i + 1]);
		}
		const double span = std::max(std::max(x1 - x0, y1 - y0), 1e-300);
		auto spread = [](uint32_t v) {
			v &= 0xffff;
			v = (v | (v << 8)) & 0x00ff00ff;
			v = (v | (v << 4)) & 0x0f0f0f0f;
			v = (v | (v << 2)) & 0x33333333;
			return (v | (v << 1)) & 0x55555555;
		};
		std::vector<uint32_t> key(R);
		for (int i = 0; i < R; i++)
		{
			const double fx = std::clamp((c->seeds[2 * i] - x0) / span, 0.0, 1.0), fy = std::clamp((c->seeds[2 * i + 1] - y0) / span, 0.0, 1.0);
			key[i] = spread((uint32_t)(fx * 65535.0)) | (spread((uint32_t)(fy * 65535.0)) << 1);
		}
		c->order.resize(R);
		for (int i = 0; i < R; i++) c->order[i] = i;
		std::stable_sort(c->order.begin(), c->order.end(), [&](int a2, int b2) { return key[a2] < key[b2]; });
	}
	// z-independent terms of the rectangular smooth-sheet series (alpha_smooth.cuh:201-204, 267-270)
	if (c->rectangular && P.approx)
	{
		std::vector<double2> a, b;
		const double arg = c->corner.arg();
		for (int i = c->taylor_smooth - 1; i >= 2; i -= 2)
		{
			const cd e = cd(0, 2.0 * arg * i).exp();
			const cd one_m = cd(1) - e;
			b.push_back(make_double2(one_m.re, one_m.im));
			a.push_back(make_double2(one_m.re / i, one_m.im / i));
		}
		c->nts = (int)a.size();
		if (c->nts > 0)
		{
			if (!CCF_HIP(c, c->dmalloc(&c->ts_a, a.size() * sizeof(double2))) || !CCF_HIP(c, c->dmalloc(&c->ts_b, b.size() * sizeof(double2))) ||
				!CCF_HIP(c, hipMemcpy(c->ts_a, a.data(), a.size() * sizeof(double2), hipMemcpyHostToDevice)) ||
				!CCF_HIP(c, hipMemcpy(c->ts_b, b.data(), b.size() * sizeof(double2), hipMemcpyHostToDevice)))
				return false;
		}
	}
	else
	{
		c->nts = 0;
	}
	return ccf::alloc_solver(c);
}

bool device_init(ccf_ctx* c)
{
	int n = 0;
	if (!CCF_HIP(c, hipGetDeviceCount(&n))) return false;
	if (n < 1) { c->fail("Error. No HIP capable devices detected.\n"); return false; }
	if (!CCF_HIP(c, hipSetDevice(0))) return false;
	if (!c->stream && !CCF_HIP(c, hipStreamCreate(&c->stream))) return false;
	if (!CCF_HIP(c, hipDeviceGetAttribute(&c->num_cus, hipDeviceAttributeMultiprocessorCount, 0))) return false;
	// the no-allocation sweep's scratch array on this device (once per device)
	return ccf::scratch_setup(c);
}

bool need(ccf_ctx* c, bool ok, const char* what)
{
	if (!ok) c->fail(std::string("Error. ") + what + " must run first.\n");
	return ok;
}

double ms_seconds(double s) { return std::floor(s * 1000.0) / 1000.0; } // stopwatch.hpp:22-33

// Every extern "C" entry point that can allocate runs inside guarded(): an
// exception (std::bad_alloc, a filesystem error) becomes the reference's
// error convention -- a message and status 1 -- instead of unwinding through
// the C boundary into the caller.
template <class F>
int guarded(ccf_ctx* c, F&& f)
{
	try
	{
		return f();
	}
	catch (const std::bad_alloc&)
	{
		if (c) c->fail("Error. Out of host memory.\n");
		else std::cerr << "Error. Out of host memory.\n";
	}
	catch (const std::exception& e)
	{
		if (c) c->fail(std::string("Error. ") + e.what() + "\n");
		else std::cerr << "Error. " << e.what() << "\n";
	}
	return 1;
}

} // namespace

// ---------------------------------------------------------------------------
// C-ABI

extern "C" {

ccf_ctx* ccf_create(void)
{
	ccf_ctx* c = new (std::nothrow) ccf_ctx();
	if (c)
	{
		ccf_default_params(&c->P_in);
		c->P = c->P_in;
	}
	return c;
}

void ccf_destroy(ccf_ctx* c)
{
	if (!c) return;
	c->free_device();
	c->pool_release();
	for (hipEvent_t e : c->events) (void)hipEventDestroy(e);
	for (hipEvent_t e : c->fmm_rows_event)
		if (e) (void)hipEventDestroy(e);
	for (hipEvent_t e : c->read_event)
		if (e) (void)hipEventDestroy(e);
	if (c->host_stage) (void)hipHostFree(c->host_stage);
	if (c->err_event) (void)hipEventDestroy(c->err_event);
	if (c->err_stream) (void)hipStreamDestroy(c->err_stream);
	if (c->group_event) (void)hipEventDestroy(c->group_event);
	for (hipStream_t g : c->group_streams)
		if (g) (void)hipStreamDestroy(g);
	if (c->stream && c->own_stream) (void)hipStreamDestroy(c->stream);
	delete c;
}

void ccf_default_params(ccf_params* p)
{
	std::memset(p, 0, sizeof(*p));
	p->kappa_tot = 0.3;
	p->shear = 0.3;
	p->kappa_star = 0.27;
	p->theta_star = 1;
	std::strcpy(p->mass_function, "equal");
	p->m_solar = 1;
	p->m_lower = 0.01;
	p->m_upper = 50;
	p->rectangular = 0;
	p->approx = 1;
	p->safety_scale = 1.37;
	p->num_stars = 137;
	p->num_phi = 100;
	p->num_branches = 1;
	p->random_seed = 0;
	p->write_stars = 1;
	p->write_critical_curves = 1;
	p->write_caustics = 1;
	p->write_mu_length_scales = 0;
	std::strcpy(p->outfile_prefix, "./");
}

int ccf_set_params(ccf_ctx* c, const ccf_params* p)
{
	if (!c || !p) return 1;
	const ccf_params before = c->P_in;
	c->P_in = *p;
	c->P_in.mass_function[sizeof(c->P_in.mass_function) - 1] = 0;
	c->P_in.starfile[CCF_PATH_MAX - 1] = 0;
	c->P_in.outfile_prefix[CCF_PATH_MAX - 1] = 0;
	c->P = c->P_in;
	// tuned choices hold for the field and parameters they were timed on
	// (turning the timing events on or off changes neither)
	ccf_params a = before, b = c->P_in;
	a.timing = b.timing = 0;
	if (std::memcmp(&a, &b, sizeof(a)) != 0) c->tuned = false;
	return 0;
}

int ccf_get_params(const ccf_ctx* c, ccf_params* p)
{
	if (!c || !p) return 1;
	*p = c->P;
	return 0;
}

int ccf_check_input_params(ccf_ctx* c, int verbose)
{
	if (!c) return 1;
	c->P = c->P_in;
	if (verbose >= 3) std::cout << "Checking input parameters...\n";
	if (!check_input_params(c)) return 1;
	if (verbose >= 3) std::cout << "Done checking input parameters.\n\n";
	return 0;
}

// validation and derived parameters need no device (the reference
// initialises the device first, ccf.cuh:1267-1270)
static bool derive(ccf_ctx* c, int verbose)
{
	c->last_error.clear();
	c->info = ccf_info{};
	c->stars_in.clear();
	c->P = c->P_in;
	if (!check_input_params(c)) return false;
	// the choices an earlier solve of this context tuned (params.autotune)
	if (c->P.autotune && c->tuned)
	{
		if (c->P.roots_per_thread == 0) c->P.roots_per_thread = c->tuned_rpt;
		if (c->P.root_fmm_order == 0 && c->P.root_fmm_leaf == 0 && c->tuned_root_order > 0)
		{
			c->P.root_fmm_order = c->tuned_root_order;
			c->P.root_fmm_leaf = c->tuned_root_leaf;
		}
	}
	if (!calculate_derived_params(c, verbose)) return false;
	c->info.num_stars = c->num_stars;
	c->info.num_roots = c->num_roots;
	c->info.taylor_smooth = c->taylor_smooth;
	c->info.kappa_star = c->kappa_star;
	c->info.mu_ave = c->mu_ave;
	c->info.corner_re = c->corner.re;
	c->info.corner_im = c->corner.im;
	c->info.alpha_error = c->alpha_error;
	return true;
}

int ccf_calculate_derived_params(ccf_ctx* c, int verbose)
{
	return guarded(c, [&]() -> int {
		if (!c) return 1;
		return derive(c, verbose) ? 0 : 1;
	});
}

int ccf_load(ccf_ctx* c, int verbose)
{
	return guarded(c, [&]() -> int {
		if (!c) return 1;
		c->free_device();
		c->loaded = false;
		if (!derive(c, verbose)) return 1;
		if (!device_init(c)) return 1;
		if (!populate(c, verbose)) return 1;
		c->loaded = true;
		return 0;
	});
}

int ccf_build_tree_fmm(ccf_ctx* c, int verbose)
{
	return guarded(c, [&]() -> int {
		if (!c || !need(c, c->loaded && !c->prepared, "ccf_load")) return 1;
		if (!ccf::build_tree(c, verbose)) return 1;
		c->info.num_stars = c->num_stars;
		c->info.num_roots = c->num_roots;
		c->info.taylor_smooth = c->taylor_smooth;
		c->info.random_seed = c->random_seed;
		c->info.kappa_star = c->kappa_star;
		c->info.kappa_star_actual = c->kappa_star_actual;
		c->info.mu_ave = c->mu_ave;
		c->info.corner_re = c->corner.re;
		c->info.corner_im = c->corner.im;
		c->info.alpha_error = c->alpha_error;
		c->prepared = true;
		return 0;
	});
}

int ccf_prepare(ccf_ctx* c, int verbose)
{
	return guarded(c, [&]() -> int {
		if (ccf_load(c, verbose)) return 1;
		return ccf_build_tree_fmm(c, verbose);
	});
}

int ccf_find_initial_roots(ccf_ctx* c, int verbose)
{
	return guarded(c, [&]() -> int {
		if (!c || !need(c, c->prepared, "ccf_prepare")) return 1;
		if (!ccf::find_initial_roots_dev(c, verbose)) return 1;
		c->have_roots = true;
		return 0;
	});
}

int ccf_find_ccs(ccf_ctx* c, int verbose)
{
	return guarded(c, [&]() -> int {
		if (!c || !need(c, c->have_roots, "ccf_find_initial_roots")) return 1;
		if (!ccf::find_ccs_dev(c, verbose)) return 1;
		c->info.max_error = c->max_error;
		c->have_ccs = true;
		return 0;
	});
}

int ccf_find_caustics(ccf_ctx* c, int verbose)
{
	return guarded(c, [&]() -> int {
		if (!c || !need(c, c->have_ccs, "ccf_find_ccs")) return 1;
		if (!c->P.write_caustics) return 0;
		if (verbose >= 2) std::cout << "Finding caustic positions...\n";
		auto t0 = std::chrono::steady_clock::now();
		if (!ccf::find_caustics_dev(c, true, false)) return 1;
		c->info.t_caustics = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
		if (verbose >= 2) std::cout << "Done finding caustic positions. Elapsed time: " << c->info.t_caustics << " seconds.\n\n";
		c->have_caustics = true;
		return 0;
	});
}

int ccf_find_mu_length_scales(ccf_ctx* c, int verbose)
{
	return guarded(c, [&]() -> int {
		if (!c || !need(c, c->have_ccs, "ccf_find_ccs")) return 1;
		if (!c->P.write_mu_length_scales) return 0;
		if (verbose >= 2) std::cout << "Finding magnification length scales...\n";
		if (!ccf::find_caustics_dev(c, false, true)) return 1;
		if (verbose >= 2) std::cout << "Done finding magnification length scales.\n\n";
		c->have_mu = true;
		return 0;
	});
}

int ccf_run(ccf_ctx* c, int verbose)
{
	return guarded(c, [&]() -> int {
		if (!c) return 1;
		if (ccf_prepare(c, verbose)) return 1;
		if (ccf_find_initial_roots(c, verbose)) return 1;
		if (ccf_find_ccs(c, verbose)) return 1;
		if (ccf_find_caustics(c, verbose)) return 1;
		if (ccf_find_mu_length_scales(c, verbose)) return 1;
		return 0;
	});
}

int ccf_save(ccf_ctx* c, int verbose)
{
	return guarded(c, [&]() -> int {
		if (!c || !need(c, c->have_ccs, "ccf_run")) return 1;
		const ccf_params& P = c->P;
		const std::string pre = P.outfile_prefix;
		const bool nofile = !c->from_file;
		const int R = c->num_roots, np = P.num_phi, nb = P.num_branches;
		// ccf.cuh:1129-1192
		if (verbose >= 2) std::cout << "Writing parameter info...\n";
		{
			const std::string fname = pre + "ccf_parameter_info.txt";
			std::ofstream out(fname);
			if (!out.is_open()) { c->fail("Error. Failed to open file " + fname + "\n"); return 1; }
			out.precision(9);
			out << "kappa_tot " << P.kappa_tot << "\n";
			out << "shear " << P.shear << "\n";
			out << "mu_ave " << c->mu_ave << "\n";
			out << "smooth_fraction " << (1 - c->kappa_star / P.kappa_tot) << "\n";
			out << "kappa_star " << c->kappa_star << "\n";
			if (nofile) out << "kappa_star_actual " << c->kappa_star_actual << "\n";
			out << "theta_star " << c->theta << "\n";
			out << "random_seed " << c->random_seed << "\n";
			if (nofile)
			{
				out << "mass_function " << P.mass_function << "\n";
				if (!std::strcmp(P.mass_function, "salpeter") || !std::strcmp(P.mass_function, "kroupa")) out << "m_solar " << P.m_solar << "\n";
				out << "m_lower " << P.m_lower << "\n";
				out << "m_upper " << P.m_upper << "\n";
				out << "mean_mass " << c->mean_mass << "\n";
				out << "mean_mass2 " << c->mean_mass2 << "\n";
				out << "mean_mass2_ln_mass " << c->mean_mass2_ln_mass << "\n";
			}
			out << "m_lower_actual " << c->m_lower_actual << "\n";
			out << "m_upper_actual " << c->m_upper_actual << "\n";
			out << "mean_mass_actual " << c->mean_mass_actual << "\n";
			out << "mean_mass2_actual " << c->mean_mass2_actual << "\n";
			out << "mean_mass2_ln_mass_actual " << c->mean_mass2_ln_mass_actual << "\n";
			out << "num_stars " << c->num_stars << "\n";
			if (c->rectangular)
			{
				out << "corner_x1 " << c->corner.re << "\n";
				out << "corner_x2 " << c->corner.im << "\n";
				if (P.approx) out << "taylor_smooth " << c->taylor_smooth << "\n";
			}
			else
			{
				out << "rad " << c->corner.abs() << "\n";
			}
			out << "num_roots " << R << "\n";
			out << "num_phi " << np << "\n";
			out << "num_branches " << nb << "\n";
			out << "max_error_1/mu " << c->max_error << "\n";
			out << "t_init_roots " << ms_seconds(c->info.t_init_roots) << "\n";
			out << "t_ccs " << ms_seconds(c->info.t_ccs) << "\n";
			out << "t_caustics " << ms_seconds(c->info.t_caustics) << "\n";
			if (verbose >= 1) std::cout << "Done writing parameter info to file " << fname << "\n";
		}
		if (P.write_stars)
		{
			std::vector<double> st(3 * (size_t)c->num_stars);
			int nr, nc;
			if (ccf_get_array(c, CCF_ARRAY_STARS, st.data(), (long long)st.size() * 8, &nr, &nc)) return 1;
			const std::string fname = pre + "ccf_stars.bin";
			if (!write_star_file(c, fname, st)) { c->fail("Error. Unable to write star info to file " + fname + "\n"); return 1; }
			if (verbose >= 1) std::cout << "Done writing star info to file " << fname << "\n";
		}
		const size_t pts = (size_t)R * (np + nb);
		if (P.write_critical_curves)
		{
			std::vector<double> buf(2 * pts);
			int nr, nc;
			if (ccf_get_array(c, CCF_ARRAY_CCS, buf.data(), (long long)buf.size() * 8, &nr, &nc)) return 1;
			const std::string fname = pre + "ccf_ccs.bin";
			if (!write_array(c, fname, buf.data(), 16, nr, nc)) { c->fail("Error. Unable to write ccs info to file " + fname + "\n"); return 1; }
			if (verbose >= 1) std::cout << "Done writing critical curve positions to file " << fname << "\n";
		}
		if (P.write_caustics)
		{
			std::vector<double> buf(2 * pts);
			int nr, nc;
			if (ccf_get_array(c, CCF_ARRAY_CAUSTICS, buf.data(), (long long)buf.size() * 8, &nr, &nc)) return 1;
			const std::string fname = pre + "ccf_caustics.bin";
			if (!write_array(c, fname, buf.data(), 16, nr, nc)) { c->fail("Error. Unable to write caustic info to file " + fname + "\n"); return 1; }
			if (verbose >= 1) std::cout << "Done writing caustic positions to file " << fname << "\n";
		}
		if (P.write_mu_length_scales)
		{
			std::vector<double> buf(pts);
			int nr, nc;
			if (ccf_get_array(c, CCF_ARRAY_MU_LENGTH_SCALES, buf.data(), (long long)buf.size() * 8, &nr, &nc)) return 1;
			const std::string fname = pre + "ccf_mu_length_scales.bin";
			if (!write_array(c, fname, buf.data(), 8, nr, nc))
			{
				c->fail("Error. Unable to write magnification length scales info to file " + fname + "\n");
				return 1;
			}
			if (verbose >= 1) std::cout << "Done writing magnification length scales to file " << fname << "\n";
		}
		return 0;
	});
}

int ccf_get_info(const ccf_ctx* c, ccf_info* info)
{
	if (!c || !info) return 1;
	*info = c->info;
	return 0;
}

int ccf_get_array(ccf_ctx* c, int which, void* dst, long long capacity, int* nrows, int* ncols)
{
	return guarded(c, [&]() -> int {
		if (!c || !dst) return 1;
		const int R = c->num_roots, np = c->P.num_phi, nb = c->P.num_branches;
		const size_t pts = (size_t)R * (np + nb);
		const void* src = nullptr;
		size_t bytes = 0;
		int nr = R * nb, nc = np / nb + 1;
		if (which == CCF_ARRAY_STARS)
		{
			if (!need(c, c->prepared, "ccf_prepare")) return 1;
			const int N = c->num_stars;
			if (capacity < (long long)N * 24) { c->fail("Error. ccf_get_array: destination too small\n"); return 1; }
			std::vector<double2> pos(N);
			std::vector<double> m(N);
			if (!CCF_HIP(c, hipMemcpy(pos.data(), c->spos, (size_t)N * sizeof(double2), hipMemcpyDeviceToHost)) ||
				!CCF_HIP(c, hipMemcpy(m.data(), c->smass, (size_t)N * sizeof(double), hipMemcpyDeviceToHost)))
				return 1;
			double* d = (double*)dst;
			for (int i = 0; i < N; i++)
			{
				d[3 * i] = pos[i].x;
				d[3 * i + 1] = pos[i].y;
				d[3 * i + 2] = m[i];
			}
			if (nrows) *nrows = N;
			if (ncols) *ncols = 3;
			return 0;
		}
		if (which == CCF_ARRAY_CCS) { if (!need(c, c->have_ccs, "ccf_find_ccs")) return 1; src = c->ccs; bytes = pts * 16; }
		else if (which == CCF_ARRAY_CAUSTICS) { if (!need(c, c->have_caustics, "ccf_find_caustics")) return 1; src = c->caustics; bytes = pts * 16; }
		else if (which == CCF_ARRAY_MU_LENGTH_SCALES) { if (!need(c, c->have_mu, "ccf_find_mu_length_scales")) return 1; src = c->mu_scales; bytes = pts * 8; }
		else { c->fail("Error. ccf_get_array: unknown array\n"); return 1; }
		if (capacity < (long long)bytes) { c->fail("Error. ccf_get_array: destination too small\n"); return 1; }
		if (!CCF_HIP(c, hipMemcpy(dst, src, bytes, hipMemcpyDeviceToHost))) return 1;
		if (nrows) *nrows = nr;
		if (ncols) *ncols = nc;
		return 0;
	});
}

const char* ccf_last_error(const ccf_ctx* c)
{
	return c ? c->last_error.c_str() : "";
}

int ccf_eval_points(ccf_ctx* c, const double* z, int n, double phi, double* out)
{
	return guarded(c, [&]() -> int {
		if (!c || !need(c, c->prepared, "ccf_prepare")) return 1;
		return ccf::eval_points_dev(c, z, n, phi, out) ? 0 : 1;
	});
}

int ccf_aberth_sweep(ccf_ctx* c, const double* roots, int n, double phi, double* out)
{
	return guarded(c, [&]() -> int {
		if (!c || !need(c, c->prepared, "ccf_prepare")) return 1;
		return ccf::aberth_sweep_dev(c, roots, n, phi, out) ? 0 : 1;
	});
}

int ccf_aberth_sweep_noalloc(ccf_ctx* c, const double* roots, int n, double phi, double* out)
{
	return guarded(c, [&]() -> int {
		if (!c || !need(c, c->prepared, "ccf_prepare")) return 1;
		return ccf::aberth_sweep_noalloc_dev(c, roots, n, phi, out) ? 0 : 1;
	});
}

int ccf_root_sums(ccf_ctx* c, const double* roots, int n, double* out)
{
	return guarded(c, [&]() -> int {
		if (!c || !need(c, c->prepared, "ccf_prepare")) return 1;
		return ccf::root_sums_dev(c, roots, n, out) ? 0 : 1;
	});
}

double ccf_mass_function(const char* name, int which, double p, double m_lower, double m_upper, double m_solar)
{
	const ccf::Imf f = ccf::make_imf(name, m_lower, m_upper, m_solar);
	if (f.kind < 0) return std::numeric_limits<double>::quiet_NaN();
	if (which == 0) return ccf::imf_mass(f, p);
	return ccf::imf_moment(f, which);
}

// ---------------------------------------------------------------------------
// batched fields

ccf_batch* ccf_batch_create(void)
{
	ccf_batch* b = new (std::nothrow) ccf_batch();
	if (b && !(b->master = ccf_create()))
	{
		delete b;
		return nullptr;
	}
	return b;
}

void ccf_batch_destroy(ccf_batch* b)
{
	if (!b) return;
	ccf::free_batch_device(b);
	for (ccf_ctx* c : b->fields) ccf_destroy(c);
	if (b->results) (void)hipFree(b->results);
	ccf_destroy(b->master);
	delete b;
}

int ccf_batch_add(ccf_batch* b, const ccf_params* p)
{
	if (!b || !p) return -1;
	ccf_ctx* c = ccf_create();
	if (!c) return -1;
	c->P_in = *p;
	c->P = *p;
	try
	{
		b->fields.push_back(c);
	}
	catch (const std::bad_alloc&)
	{
		ccf_destroy(c);
		return -1;
	}
	b->loaded = b->built = b->solved = false;
	return (int)b->fields.size() - 1;
}

int ccf_batch_size(const ccf_batch* b) { return b ? (int)b->fields.size() : 0; }

int ccf_batch_load(ccf_batch* b, int verbose)
{
	return guarded(b ? b->master : nullptr, [&]() -> int {
		if (!b) return 1;
		ccf_ctx* m = b->master;
		m->last_error.clear();
		b->loaded = b->built = b->solved = false;
		ccf::free_batch_device(b);
		if (b->fields.empty())
		{
			m->fail("Error. The batch has no fields.\n");
			return 1;
		}
		const ccf_params& p0 = b->fields[0]->P_in;
		for (ccf_ctx* c : b->fields)
		{
			const ccf_params& p = c->P_in;
			if (p.fmm_order != p0.fmm_order || p.direct != p0.direct || p.binomial_exact != p0.binomial_exact || p.precision != p0.precision)
			{
				m->fail("Error. The fields of a batch must share fmm_order, direct, binomial_exact and precision.\n");
				return 1;
			}
		}
		m->P_in = p0;
		m->P = p0;
		if (!device_init(m)) return 1;
		for (size_t i = 0; i < b->fields.size(); i++)
		{
			ccf_ctx* c = b->fields[i];
			c->batch_member = true;
			if (c->stream && c->own_stream) (void)hipStreamDestroy(c->stream);
			c->stream = m->stream;
			c->own_stream = false;
			if (ccf_load(c, verbose))
			{
				m->last_error = "Error. Field " + std::to_string(i) + ": " + c->last_error;
				return 1;
			}
		}
		b->loaded = true;
		return 0;
	});
}

int ccf_batch_build(ccf_batch* b, int verbose)
{
	return guarded(b ? b->master : nullptr, [&]() -> int {
		if (!b) return 1;
		if (!b->loaded)
		{
			b->master->fail("Error. ccf_batch_load must run first.\n");
			return 1;
		}
		b->built = b->solved = false;
		if (!ccf::batch_build(b, verbose)) return 1;
		b->built = true;
		return 0;
	});
}

int ccf_batch_solve(ccf_batch* b, int verbose)
{
	return guarded(b ? b->master : nullptr, [&]() -> int {
		if (!b) return 1;
		if (!b->built)
		{
			b->master->fail("Error. ccf_batch_build must run first.\n");
			return 1;
		}
		if (!ccf::batch_solve(b, verbose)) return 1;
		b->solved = true;
		return 0;
	});
}

int ccf_batch_run(ccf_batch* b, int verbose)
{
	return guarded(b ? b->master : nullptr, [&]() -> int {
		if (ccf_batch_load(b, verbose) || ccf_batch_build(b, verbose)) return 1;
		return ccf_batch_solve(b, verbose);
	});
}

int ccf_batch_save(ccf_batch* b, int field, int verbose)
{
	return guarded(b ? b->master : nullptr, [&]() -> int {
		if (!b || field < 0 || field >= (int)b->fields.size()) return 1;
		if (!b->solved)
		{
			b->master->fail("Error. ccf_batch_solve must run first.\n");
			return 1;
		}
		return ccf_save(b->fields[field], verbose);
	});
}

int ccf_batch_get_info(const ccf_batch* b, int field, ccf_info* info)
{
	return guarded(b ? b->master : nullptr, [&]() -> int {
		if (!b || !info || field < -1 || field >= (int)b->fields.size()) return 1;
		*info = field < 0 ? b->info : b->fields[field]->info;
		return 0;
	});
}

int ccf_batch_get_array(ccf_batch* b, int field, int which, void* dst, long long capacity, int* nrows, int* ncols)
{
	return guarded(b ? b->master : nullptr, [&]() -> int {
		if (!b || field < 0 || field >= (int)b->fields.size()) return 1;
		return ccf_get_array(b->fields[field], which, dst, capacity, nrows, ncols);
	});
}

int ccf_batch_workgroup(const ccf_batch* b) { return b ? b->block1 : 0; }

const char* ccf_batch_last_error(const ccf_batch* b) { return b && b->master ? b->master->last_error.c_str() : ""; }

int ccf_write_synthetic_stars(const ccf_params* p, unsigned long long seed, const char* path)
{
	return guarded(nullptr, [&]() -> int {
		if (!p || !path) return 1;
		const int N = p->num_stars;
		if (N < 1) { std::cerr << "Error. num_stars must be an integer > 0\n"; return 1; }
		double ml = p->m_lower, mu = p->m_upper;
		if (!std::strcmp(p->mass_function, "equal")) { ml = 1; mu = 1; }
		else { ml *= p->m_solar; mu *= p->m_solar; }
		const ccf::Imf f = ccf::make_imf(p->mass_function, ml, mu, p->m_solar);
		if (f.kind < 0) { std::cerr << "Error. mass_function must be equal, uniform, Salpeter, Kroupa, or optical_depth.\n"; return 1; }
		const double mean = ccf::imf_moment(f, 1);
		const double th = p->theta_star, ks = p->kappa_star, k = p->kappa_tot, g = p->shear;
		// field the reference derives for these parameters (ccf.cuh:420-436)
		cd cr(1 / std::abs(1 - k + g), 1 / std::abs(1 - k - g));
		if (p->rectangular)
		{
			cr = cd(std::sqrt(cr.re / cr.im), std::sqrt(cr.im / cr.re));
			cr *= std::sqrt(ccf::PI * th * th * N * mean / (4 * ks));
		}
		else
		{
			cr = cr / cr.abs();
			cr *= std::sqrt(th * th * N * mean / ks);
		}
		std::mt19937_64 rng(seed);
		auto u = [&rng]() { return 1.0 - (double)(rng() >> 11) * 0x1.0p-53; }; // (0, 1]
		std::vector<double> st(3 * (size_t)N);
		double mt = 0;
		for (int i = 0; i < N; i++)
		{
			double x1, x2;
			if (p->rectangular)
			{
				x1 = u() * 2 * cr.re - cr.re;
				x2 = u() * 2 * cr.im - cr.im;
			}
			else
			{
				const double a = u() * 2 * ccf::PI;
				const double r = std::sqrt(u()) * cr.abs();
				x1 = r * std::cos(a);
				x2 = r * std::sin(a);
			}
			st[3 * i] = x1;
			st[3 * i + 1] = x2;
			st[3 * i + 2] = ccf::imf_mass(f, u());
			mt += st[3 * i + 2];
		}
		// the file's field size follows the actual mean mass, as the reference
		// rescales its corner after drawing the stars (ccf.cuh:686-700)
		const double mact = mt / N;
		if (p->rectangular) cr *= std::sqrt(mact / mean);
		else cr *= std::sqrt(mact / mean);
		std::ofstream out(path, std::ios::binary);
		if (!out.is_open()) { std::cerr << "Error. Failed to open file " << path << "\n"; return 1; }
		const int32_t n32 = N, r32 = p->rectangular;
		out.write((const char*)&n32, 4);
		out.write((const char*)&r32, 4);
		const double cc[2] = {p->rectangular ? cr.re : cr.abs(), p->rectangular ? cr.im : 0.0};
		out.write((const char*)cc, 16);
		out.write((const char*)&th, 8);
		out.write((const char*)st.data(), st.size() * 8);
		return out.good() ? 0 : 1;
	});
}

} // extern "C"
