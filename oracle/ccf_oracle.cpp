/*
 * FP64 CPU oracle of cc_finder's hot path -- TEST INFRASTRUCTURE ONLY.
 *
 * The checker for the HIP product path and the CPU baseline timed by
 * bench.py. The product (ccf_amd/) never links or calls this file.
 * Each section cites the reference file and lines it restates
 * (paths relative to the weisluke/ccf repository root). The interface and
 * the deliberate differences from the reference are listed in oracle.h.
 */
#include "oracle.h"

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstdint>
#include <cstdio>
#include <cstring>
#include <filesystem>
#include <fstream>
#include <iostream>
#include <limits>
#include <memory>
#include <numbers>
#include <string>
#include <vector>

#ifdef _OPENMP
#include <omp.h>
#endif

namespace orc
{

constexpr double PI = std::numbers::pi;

/* ---------------------------------------------------------------------------
   complex arithmetic with the reference's conventions (complex.cuh):
   abs = hypot, arg = atan2, log = (log|z|, arg z), Smith division
   (complex.cuh:218-280), integer powers by repeated multiplication (:285-306)
   ------------------------------------------------------------------------- */
struct cx
{
	double re = 0, im = 0;
	cx() = default;
	cx(double r) : re(r), im(0) {}
	cx(double r, double i) : re(r), im(i) {}
	cx conj() const { return cx(re, -im); }
	double abs() const { return std::hypot(re, im); }
	double arg() const { return std::atan2(im, re); }
	cx log() const { return cx(std::log(abs()), arg()); }
	cx exp() const { return cx(std::exp(re) * std::cos(im), std::exp(re) * std::sin(im)); }
	cx operator-() const { return cx(-re, -im); }
	cx& operator+=(cx c) { re += c.re; im += c.im; return *this; }
	cx& operator-=(cx c) { re -= c.re; im -= c.im; return *this; }
	cx& operator*=(cx c) { double r = re * c.re - im * c.im; im = re * c.im + im * c.re; re = r; return *this; }
	cx& operator*=(double a) { re *= a; im *= a; return *this; }
	cx& operator/=(double a) { re /= a; im /= a; return *this; }
	cx& operator/=(cx c);
	cx pow(int n) const;
};
inline cx operator+(cx a, cx b) { return cx(a.re + b.re, a.im + b.im); }
inline cx operator-(cx a, cx b) { return cx(a.re - b.re, a.im - b.im); }
inline cx operator*(cx a, cx b) { return cx(a.re * b.re - a.im * b.im, a.re * b.im + a.im * b.re); }
inline cx operator*(cx a, double s) { return cx(a.re * s, a.im * s); }
inline cx operator*(double s, cx a) { return cx(s * a.re, s * a.im); }
inline cx operator/(cx a, double s) { return cx(a.re / s, a.im / s); }
inline cx operator/(cx a, cx b)
{
	if (std::abs(b.im) < std::abs(b.re))
	{
		double f = b.im / b.re;
		return cx((a.re + a.im * f) / (b.re + b.im * f), (a.im - a.re * f) / (b.re + b.im * f));
	}
	double f = b.re / b.im;
	return cx((a.re * f + a.im) / (b.re * f + b.im), (a.im * f - a.re) / (b.re * f + b.im));
}
inline cx operator/(double s, cx b)
{
	if (std::abs(b.im) < std::abs(b.re))
	{
		double f = b.im / b.re;
		return cx(s / (b.re + b.im * f), (-s * f) / (b.re + b.im * f));
	}
	double f = b.re / b.im;
	return cx((s * f) / (b.re * f + b.im), -s / (b.re * f + b.im));
}
cx& cx::operator/=(cx c) { *this = *this / c; return *this; }
cx cx::pow(int n) const
{
	cx r(1, 0);
	if (n > 0) { for (int i = 0; i < n; i++) r *= *this; }
	else if (n < 0) { for (int i = 0; i > n; i--) r *= *this; r = 1.0 / r; }
	return r;
}

/* ---------------------------------------------------------------------------
   binomial coefficients (binomial_coefficients.cuh:11-64). The reference
   fills Pascal's rows 0..2p in `int`; rows past 33 overflow and wrap, and
   M2L (fmm.cuh:314) reads binom(l+k-1, k-1) up to row 2p-1. The wrapped
   table is rebuilt here with uint32 addition (the same residues mod 2^32)
   read back as int32.
   ------------------------------------------------------------------------- */
struct Binomials
{
	int nmax = 0;
	std::vector<double> tab;
	void build(int n, bool wrapped)
	{
		nmax = n;
		tab.assign((n + 1) * (n + 2) / 2, 0.0);
		std::vector<uint32_t> w((n + 1) * (n + 2) / 2, 0);
		std::vector<long double> e((n + 1) * (n + 2) / 2, 0);
		for (int r = 0; r <= n; r++)
		{
			int s = r * (r + 1) / 2, ps = (r - 1) * r / 2;
			for (int i = 0; i <= r; i++)
			{
				if (i == 0 || i == r) { w[s + i] = 1; e[s + i] = 1; }
				else { w[s + i] = w[ps + i - 1] + w[ps + i]; e[s + i] = e[ps + i - 1] + e[ps + i]; }
				tab[s + i] = wrapped ? static_cast<double>(static_cast<int32_t>(w[s + i])) : static_cast<double>(e[s + i]);
			}
		}
	}
	double get(int n, int k) const
	{
		if (n < k) return 0;
		return tab[n * (n + 1) / 2 + k];
	}
};

/* ---------------------------------------------------------------------------
   mass functions (the headers under include/mass_functions)
   ------------------------------------------------------------------------- */
double power_integral(double x1, double x2, double a, double b = 1)
{
	if (a != -1) return b * (std::pow(x2, a + 1) - std::pow(x1, a + 1)) / (a + 1);
	return b * (std::log(x2) - std::log(x1));
}
double power_log_integral(double x1, double x2, double a, double b = 1)
{
	if (a != -1)
		return b * (std::pow(x2, a + 1) * ((a + 1) * std::log(x2) - 1) - std::pow(x1, a + 1) * ((a + 1) * std::log(x1) - 1)) / ((a + 1) * (a + 1));
	return b * (std::log(x2) * std::log(x2) - std::log(x1) * std::log(x1)) / 2;
}
double invert_power_integral(double p, double x1, double a, double b)
{
	if (a != -1) return std::pow(p * (a + 1) / b + std::pow(x1, a + 1), 1 / (a + 1));
	return x1 * std::exp(p / b);
}

/* power law p(m) ~ m^a on [ml, mu] (power_law.cuh:126-228) */
double powerlaw_moment(double a, int which, double p, double ml, double mu)
{
	if (ml == mu)
	{
		if (which == 0 || which == 1) return ml;
		if (which == 2) return ml * ml;
		return ml * ml * std::log(ml);
	}
	double b = 1 / power_integral(ml, mu, a);
	if (which == 0) return invert_power_integral(p, ml, a, b);
	if (which == 1) return power_integral(ml, mu, a + 1, b);
	if (which == 2) return power_integral(ml, mu, a + 2, b);
	return power_log_integral(ml, mu, a + 2, b);
}

/* Kroupa broken power law, slopes -0.3/-1.3/-2.3 with breaks 0.08 and 0.5
   solar masses (kroupa.cuh:22-505). The three segments are normalised
   jointly and matched at the breaks; moments sum the segment integrals. */
double kroupa(int which, double p, double ml, double mu, double ms)
{
	if (ml == mu) return powerlaw_moment(0, which, p, ml, mu);
	const double sl[3] = {-0.3, -1.3, -2.3};
	double m0 = 0.08 * ms, m1 = 0.5 * ms;
	/* segment bounds clipped to [ml, mu] */
	double lo[3] = {ml, std::max(ml, m0), std::max(ml, m1)};
	double hi[3] = {std::min(mu, m0), std::min(mu, m1), mu};
	/* segments present in the reference's case analysis (kroupa.cuh:50-125) */
	bool used[3] = {ml < m0, mu >= m0 && ml < m1, mu >= m1};
	/* continuity factors relative to segment 0: b1 = b0 m0^(s0-s1), b2 = b1 m1^(s1-s2) */
	double rel[3] = {1, std::pow(m0, sl[0] - sl[1]), std::pow(m0, sl[0] - sl[1]) * std::pow(m1, sl[1] - sl[2])};
	/* the reference normalises relative to the first segment that is present */
	int first = used[0] ? 0 : (used[1] ? 1 : 2);
	double probs[3] = {0, 0, 0}, tot = 0;
	for (int i = first; i < 3; i++)
	{
		if (!used[i]) continue;
		probs[i] = power_integral(lo[i], hi[i], sl[i]) * rel[i] / rel[first];
		tot += probs[i];
	}
	double b[3];
	for (int i = 0; i < 3; i++) b[i] = rel[i] / rel[first] / tot;
	for (int i = 0; i < 3; i++) probs[i] /= tot;
	if (which == 0)
	{
		/* invert the CDF segment by segment (kroupa.cuh:128-158) */
		double acc = 0;
		for (int i = first; i < 3; i++)
		{
			if (!used[i]) continue;
			bool last = (i == 2) || !used[i + 1];
			if (p < acc + probs[i] || last) return invert_power_integral(p - acc, lo[i], sl[i], b[i]);
			acc += probs[i];
		}
		return mu;
	}
	double m = 0;
	for (int i = first; i < 3; i++)
	{
		if (!used[i]) continue;
		if (which == 1) m += power_integral(lo[i], hi[i], sl[i] + 1, b[i]);
		else if (which == 2) m += power_integral(lo[i], hi[i], sl[i] + 2, b[i]);
		else m += power_log_integral(lo[i], hi[i], sl[i] + 2, b[i]);
	}
	return m;
}

/* two masses contributing equally to the convergence (optical_depth.cuh:24-93) */
double optical_depth(int which, double p, double ml, double mu)
{
	if (ml == mu) return powerlaw_moment(0, which, p, ml, mu);
	if (which == 0) return p <= mu / (ml + mu) ? ml : mu;
	if (which == 1) return 2 * ml * mu / (ml + mu);
	if (which == 2) return ml * mu;
	return ml * mu * (ml * std::log(ml) + mu * std::log(mu)) / (ml + mu);
}

double mass_function(const std::string& name, int which, double p, double ml, double mu, double ms)
{
	if (name == "equal") return which == 3 ? 0.0 : 1.0; /* equal.cuh:24-72 */
	if (name == "uniform") return powerlaw_moment(0, which, p, ml, mu);
	if (name == "salpeter") return powerlaw_moment(-2.35, which, p, ml, mu);
	if (name == "kroupa") return kroupa(which, p, ml, mu, ms);
	if (name == "optical_depth") return optical_depth(which, p, ml, mu);
	return std::numeric_limits<double>::quiet_NaN();
}

/* ---------------------------------------------------------------------------
   star files (star.cuh:124-534)
   ------------------------------------------------------------------------- */
struct Star { cx pos; double mass; };

struct Field
{
	int nstars = 0;
	int rectangular = 0;
	cx corner;
	double theta = 1;
	std::vector<Star> stars;
	double kappa_star = 0, m_lower = 0, m_upper = 0, mean_mass = 0, mean_mass2 = 0, mean_mass2_ln_mass = 0;
};

bool read_star_file(const std::string& fname, Field& f)
{
	std::filesystem::path path = fname;
	if (path.extension() == ".bin")
	{
		std::error_code err;
		auto fsize = std::filesystem::file_size(fname, err);
		if (err) { std::cerr << "Error determining size of star input file " << fname << "\n"; return false; }
		std::ifstream in(fname, std::ios::binary);
		if (!in) { std::cerr << "Error. Failed to open file " << fname << "\n"; return false; }
		int32_t n = 0, rect = 0;
		in.read((char*)&n, 4);
		if (n < 1) { std::cerr << "Error. Invalid num_stars input. num_stars must be an integer > 0\n"; return false; }
		in.read((char*)&rect, 4);
		if (rect != 0 && rect != 1) { std::cerr << "Error. Invalid rectangular input. rectangular must be 1 (rectangular) or 0 (circular).\n"; return false; }
		f.nstars = n;
		f.rectangular = rect;
		f.stars.resize(n);
		if (fsize == 8 + 16 + 8 + (uint64_t)n * 24)
		{
			double c[2], t;
			in.read((char*)c, 16);
			in.read((char*)&t, 8);
			f.corner = cx(c[0], c[1]);
			f.theta = t;
			std::vector<double> buf(3 * (size_t)n);
			in.read((char*)buf.data(), buf.size() * 8);
			for (int i = 0; i < n; i++) f.stars[i] = {cx(buf[3 * i], buf[3 * i + 1]), buf[3 * i + 2]};
		}
		else if (fsize == 8 + 8 + 4 + (uint64_t)n * 12)
		{
			float c[2], t;
			in.read((char*)c, 8);
			in.read((char*)&t, 4);
			f.corner = cx(c[0], c[1]);
			f.theta = t;
			std::vector<float> buf(3 * (size_t)n);
			in.read((char*)buf.data(), buf.size() * 4);
			for (int i = 0; i < n; i++) f.stars[i] = {cx(buf[3 * i], buf[3 * i + 1]), buf[3 * i + 2]};
		}
		else
		{
			std::cerr << "Error. Star input file " << fname << " does not contain validly formatted single or double precision stars and accompanying information.\n";
			return false;
		}
		if (f.corner.re < 0 || f.corner.im < 0) { std::cerr << "Error. Real and imaginary parts of the corner of the star field must both be >= 0\n"; return false; }
		if (f.theta < std::numeric_limits<double>::min()) { std::cerr << "Error. theta_star must be >= " << std::numeric_limits<double>::min() << "\n"; return false; }
	}
	else if (path.extension() == ".txt")
	{
		std::ifstream in(fname);
		if (!in) { std::cerr << "Error. Failed to open file " << fname << "\n"; return false; }
		double x1, x2, m;
		f.stars.clear();
		double max_x1 = 0, max_x2 = 0, max_rad = 0, total = 0, I = 0;
		while (in >> x1 >> x2 >> m)
		{
			f.stars.push_back({cx(x1, x2), m});
			total += m;
			I += m * (x1 * x1 + x2 * x2);
			max_x1 = std::max(max_x1, std::abs(x1));
			max_x2 = std::max(max_x2, std::abs(x2));
			max_rad = std::max(max_rad, std::hypot(x1, x2));
		}
		f.nstars = (int)f.stars.size();
		if (f.nstars < 1) { std::cerr << "Error. Invalid num_stars input. num_stars must be an integer > 0\n"; return false; }
		/* circular vs rectangular by moment of inertia (star.cuh:411-429) */
		double I_circ = total * max_rad * max_rad / 2, I_rect = total * (max_x1 * max_x1 + max_x2 * max_x2) / 3;
		if (I / I_circ > I / I_rect) { f.rectangular = 0; f.corner = cx(max_rad, 0); }
		else { f.rectangular = 1; f.corner = cx(max_x1, max_x2); }
		f.theta = 1;
	}
	else
	{
		std::cerr << "Error. Star input file " << fname << " is not a .bin or .txt file.\n";
		return false;
	}
	return true;
}

/* star.cuh:124-161 */
void calculate_star_params(Field& f)
{
	f.m_lower = std::numeric_limits<double>::max();
	f.m_upper = std::numeric_limits<double>::min();
	double mt = 0, m2 = 0, m2l = 0;
	for (auto& s : f.stars)
	{
		mt += s.mass;
		m2 += s.mass * s.mass;
		m2l += s.mass * s.mass * std::log(s.mass);
		f.m_lower = std::min(f.m_lower, s.mass);
		f.m_upper = std::max(f.m_upper, s.mass);
	}
	f.mean_mass = mt / f.nstars;
	f.mean_mass2 = m2 / f.nstars;
	f.mean_mass2_ln_mass = m2l / f.nstars;
	if (f.rectangular) f.kappa_star = mt * PI * f.theta * f.theta / (4 * f.corner.re * f.corner.im);
	else f.kappa_star = mt * f.theta * f.theta / (f.corner.abs() * f.corner.abs());
}

bool write_star_file(const std::string& fname, const Field& f, const std::vector<Star>& stars)
{
	std::ofstream out(fname, std::ios::binary);
	if (!out) { std::cerr << "Error. Failed to open file " << fname << "\n"; return false; }
	int32_t n = f.nstars, r = f.rectangular;
	out.write((char*)&n, 4);
	out.write((char*)&r, 4);
	double c[2] = {f.rectangular ? f.corner.re : f.corner.abs(), f.rectangular ? f.corner.im : 0.0};
	out.write((char*)c, 16);
	out.write((char*)&f.theta, 8);
	for (auto& s : stars)
	{
		double v[3] = {s.pos.re, s.pos.im, s.mass};
		out.write((char*)v, 24);
	}
	return true;
}

/* array_functions.cuh:280-306 */
template <typename V>
bool write_array(const std::string& fname, const V* vals, int nrows, int ncols)
{
	std::ofstream out(fname, std::ios::binary);
	if (!out) { std::cerr << "Error. Failed to open file " << fname << "\n"; return false; }
	out.write((char*)&nrows, 4);
	out.write((char*)&ncols, 4);
	out.write((const char*)vals, (size_t)nrows * ncols * sizeof(V));
	return true;
}

/* ---------------------------------------------------------------------------
   quadtree (tree_node.cuh:32-561, driver ccf.cuh:732-835)
   ------------------------------------------------------------------------- */
constexpr int MAX_DIRECT = 32;     /* tree_node.cuh:14 */
constexpr int MAX_ORDER = 31;      /* tree_node.cuh:20 */

struct Node
{
	cx center;
	double hl = 0;
	int level = 0;
	Node* parent = nullptr;
	Node* children[4] = {nullptr, nullptr, nullptr, nullptr};
	int nchildren = 0;
	Node* nbrs[8];
	int nnbrs = 0;
	Node* sl[27];
	int nsl = 0;
	Node* dl[5];
	int ndl = 0;
	int stars = 0, numstars = 0;
	cx mp[MAX_ORDER + 1];
	cx loc[MAX_ORDER + 1];
};

[[noreturn]] void list_overflow(const char* what)
{
	std::fprintf(stderr, "oracle: %s list overflow\n", what);
	std::abort();
}

/* tree_node.cuh:183-267 */
void set_neighbors(Node* n)
{
	if (n->level == 0) return;
	Node* p = n->parent;
	for (int i = 0; i < p->nchildren; i++)
	{
		Node* c = p->children[i];
		if (c->numstars == 0 || c == n) continue;
		if (n->nnbrs >= 8) list_overflow("neighbour");
		n->nbrs[n->nnbrs++] = c;
	}
	for (int i = 0; i < p->nnbrs; i++)
	{
		Node* nb = p->nbrs[i];
		if (nb->numstars == 0) continue;
		if (nb->nchildren == 0)
		{
			if (std::fabs(nb->center.re - n->center.re) < 3.5 * n->hl && std::fabs(nb->center.im - n->center.im) < 3.5 * n->hl)
			{
				if (n->nnbrs >= 8) list_overflow("neighbour");
				n->nbrs[n->nnbrs++] = nb;
			}
			else
			{
				if (n->ndl >= 5) list_overflow("coarser-level interaction");
				n->dl[n->ndl++] = nb;
			}
		}
		else
		{
			for (int j = 0; j < nb->nchildren; j++)
			{
				Node* c = nb->children[j];
				if (c->numstars == 0) continue;
				if (std::fabs(c->center.re - n->center.re) < 3 * n->hl && std::fabs(c->center.im - n->center.im) < 3 * n->hl)
				{
					if (n->nnbrs >= 8) list_overflow("neighbour");
					n->nbrs[n->nnbrs++] = c;
				}
				else
				{
					if (n->nsl >= 27) list_overflow("same-level interaction");
					n->sl[n->nsl++] = c;
				}
			}
		}
	}
}

/* child quadrant of a position (tree_node.cuh:155-160, 430-479, 542-557) */
inline int quadrant(cx z, cx c)
{
	if (z.im >= c.im) return z.re >= c.re ? 0 : 1;
	return z.re < c.re ? 2 : 3;
}

struct Tree
{
	std::vector<std::unique_ptr<std::vector<Node>>> levels;
	int tree_levels = 0;
	int order = 0;
	Node* root() { return &(*levels[0])[0]; }
	int num_nodes() const
	{
		int n = 0;
		for (auto& l : levels) n += (int)l->size();
		return n;
	}
};

/* build the tree over `stars` (sorted in place into node order) */
void build_tree(Tree& t, std::vector<Star>& stars, double root_hl, bool direct)
{
	t.levels.clear();
	t.levels.emplace_back(new std::vector<Node>(1));
	Node& r = (*t.levels[0])[0];
	r.center = cx(0, 0);
	r.hl = root_hl;
	r.level = 0;
	r.numstars = (int)stars.size();
	t.tree_levels = 0;
	if (direct) return;
	std::vector<Star> tmp(stars.size());
	while (true)
	{
		auto& lvl = *t.levels[t.tree_levels];
		int max_n = 0, nsplit = 0;
		std::vector<char> split(lvl.size(), 0);
		for (size_t i = 0; i < lvl.size(); i++)
		{
			Node& n = lvl[i];
			int cnt = n.numstars;
			for (int j = 0; j < n.nnbrs; j++) cnt += n.nbrs[j]->numstars;
			if (cnt > MAX_DIRECT) { split[i] = 1; nsplit++; }
			max_n = std::max(max_n, cnt);
		}
		if (max_n <= MAX_DIRECT) break;
		t.levels.emplace_back(new std::vector<Node>(4 * (size_t)nsplit));
		auto& next = *t.levels[t.tree_levels + 1];
		int k = 0;
		const cx off[4] = {cx(1, 1), cx(-1, 1), cx(-1, -1), cx(1, -1)};
		for (size_t i = 0; i < lvl.size(); i++)
		{
			if (!split[i]) continue;
			Node& n = lvl[i];
			for (int q = 0; q < 4; q++)
			{
				Node& c = next[4 * k + q];
				c.hl = n.hl / 2;
				c.center = n.center + c.hl * off[q];
				c.level = n.level + 1;
				c.parent = &n;
				n.children[q] = &c;
			}
			n.nchildren = 4;
			k++;
			/* stable 4-way partition of the node's stars by quadrant */
			if (n.numstars > 0)
			{
				int cnt[4] = {0, 0, 0, 0};
				for (int s = 0; s < n.numstars; s++) cnt[quadrant(stars[n.stars + s].pos, n.center)]++;
				int base[4] = {0, cnt[0], cnt[0] + cnt[1], cnt[0] + cnt[1] + cnt[2]};
				int fill[4] = {base[0], base[1], base[2], base[3]};
				for (int s = 0; s < n.numstars; s++)
				{
					const Star& st = stars[n.stars + s];
					tmp[n.stars + fill[quadrant(st.pos, n.center)]++] = st;
				}
				for (int s = 0; s < n.numstars; s++) stars[n.stars + s] = tmp[n.stars + s];
				for (int q = 0; q < 4; q++)
				{
					n.children[q]->stars = n.stars + base[q];
					n.children[q]->numstars = cnt[q];
				}
			}
		}
		t.tree_levels++;
		for (auto& c : next) set_neighbors(&c);
	}
}

/* fmm.cuh:21-49 (P2M) */
void p2m(Node& n, int p, const std::vector<Star>& stars)
{
	for (int k = 0; k <= p; k++)
	{
		cx r;
		if (k == 0)
		{
			for (int i = 0; i < n.numstars; i++) r += cx(stars[n.stars + i].mass);
		}
		else
		{
			for (int i = 0; i < n.numstars; i++)
				r -= ((stars[n.stars + i].pos - n.center) / n.hl).pow(k) * stars[n.stars + i].mass;
			r /= (double)k;
		}
		n.mp[k] = r;
	}
}

/* fmm.cuh:109-139 (M2M, child -> parent) */
cx m2m_coeff(const Node& c, int power, const Binomials& B)
{
	if (power == 0) return c.mp[0];
	cx r;
	cx dz = (c.center - c.parent->center) / c.hl;
	for (int i = power; i >= 1; i--)
	{
		r += c.mp[i] * B.get(power - 1, i - 1);
		r /= dz;
	}
	r -= c.mp[0] / (double)power;
	r *= dz.pow(power);
	r /= (double)(int32_t)(1u << power);
	return r;
}

/* fmm.cuh:204-223 (L2L, parent -> child) */
cx l2l_coeff(const Node& n, int power, int maxpower, const Binomials& B)
{
	cx r;
	cx dz = (n.center - n.parent->center) / n.parent->hl;
	for (int i = maxpower; i >= 1; i--)
	{
		r += n.parent->loc[i] * B.get(i, power);
		r *= dz;
	}
	r /= dz.pow(power);
	r /= (double)(int32_t)(1u << power);
	return r;
}

/* fmm.cuh:281-322 (M2L) */
cx m2l_coeff(const Node& from, const Node& to, int power, int maxpower, const Binomials& B)
{
	cx r;
	cx dz = (from.center - to.center) / from.hl;
	if (power == 0)
	{
		for (int i = maxpower; i >= 1; i--)
		{
			r += from.mp[i];
			r /= -dz;
		}
		r += from.mp[0] * (-dz).log();
		r += from.mp[0] * std::log(from.hl);
	}
	else
	{
		for (int i = maxpower; i >= 1; i--)
		{
			r += from.mp[i] * B.get(power + i - 1, i - 1);
			r /= -dz;
		}
		r -= from.mp[0] / (double)power;
		r /= dz.pow(power);
	}
	return r;
}

/* fmm.cuh:387-413 (P2L) */
cx p2l_coeff(const Node& from, const Node& to, int power, const std::vector<Star>& stars)
{
	cx r;
	for (int i = 0; i < from.numstars; i++)
	{
		const Star& s = stars[from.stars + i];
		cx dz = (s.pos - to.center) / to.hl;
		if (power == 0)
		{
			r += s.mass * (-dz).log();
			r += cx(s.mass * std::log(to.hl));
		}
		else
		{
			r -= s.mass / ((double)power * dz.pow(power));
		}
	}
	return r;
}

/* ccf.cuh:866-893 */
void build_fmm(Tree& t, int p, const Binomials& B, const std::vector<Star>& stars)
{
	t.order = p;
	for (int i = t.tree_levels; i >= 0; i--)
	{
		auto& lvl = *t.levels[i];
#pragma omp parallel for schedule(dynamic, 8)
		for (size_t k = 0; k < lvl.size(); k++)
		{
			Node& n = lvl[k];
			if (n.nchildren == 0 && n.numstars > 0) p2m(n, p, stars);
			for (int c = 0; c < n.nchildren; c++)
			{
				cx sh[MAX_ORDER + 1];
				for (int l = 0; l <= p; l++) sh[l] = m2m_coeff(*n.children[c], l, B);
				for (int l = 0; l <= p; l++) n.mp[l] += sh[l];
			}
		}
	}
	for (int i = 2; i <= t.tree_levels; i++)
	{
		auto& lvl = *t.levels[i];
#pragma omp parallel for schedule(dynamic, 8)
		for (size_t k = 0; k < lvl.size(); k++)
		{
			Node& n = lvl[k];
			cx loc[MAX_ORDER + 1];
			for (int l = 0; l <= p; l++) loc[l] = l2l_coeff(n, l, p, B);
			for (int j = 0; j < n.nsl; j++)
				for (int l = 0; l <= p; l++) loc[l] += m2l_coeff(*n.sl[j], n, l, p, B);
			for (int j = 0; j < n.ndl; j++)
				for (int l = 0; l <= p; l++) loc[l] += p2l_coeff(*n.dl[j], n, l, stars);
			for (int l = 0; l <= p; l++) n.loc[l] = loc[l];
		}
	}
}

/* tree_node.cuh:529-561 */
const Node* nearest_node(cx z, const Node* n)
{
	while (n->nchildren > 0)
	{
		if (std::fabs(z.re - n->center.re) > n->hl || std::fabs(z.im - n->center.im) > n->hl) return n;
		if (z.re >= n->center.re && z.im >= n->center.im) n = n->children[0];
		else if (z.re < n->center.re && z.im >= n->center.im) n = n->children[1];
		else if (z.re < n->center.re && z.im < n->center.im) n = n->children[2];
		else n = n->children[3];
	}
	return n;
}

/* ---------------------------------------------------------------------------
   lens model and deflections
   ------------------------------------------------------------------------- */
struct Lens
{
	double kappa = 0, gamma = 0, theta = 1, kappastar = 0;
	int rectangular = 0, approx = 1, taylor_smooth = 1;
	cx corner;
	int p = 0;
	int plant = 0; /* oracle_params::plant: a planted mistake (tests of the tests only) */
	bool use_tree = true;
	const std::vector<Star>* stars = nullptr;
	const std::vector<double>* sx = nullptr; /* star x and y, tree order */
	const std::vector<double>* sy = nullptr;
	const Node* root = nullptr;
};

/* sums over the stars of a node and its neighbours (alpha_star.cuh:18-115):
   s1 = sum m/(z-s), s2 = sum m/(z-s)^2, s3 = sum m/(z-s)^3 */
void p2p(const Lens& L, const Node* node, cx z, cx& s1, cx& s2, cx& s3, int& nnf)
{
	const auto& st = *L.stars;
	s1 = s2 = s3 = cx();
	nnf = 0;
	auto acc = [&](const Node* n) {
		for (int i = 0; i < n->numstars; i++)
		{
			const Star& s = st[n->stars + i];
			cx d = z - s.pos;
			s1 += s.mass / d;
			s2 += s.mass / (d * d);
			s3 += s.mass / (d * d * d);
		}
		nnf += n->numstars;
	};
	for (int j = 0; j < node->nnbrs; j++) acc(node->nbrs[j]);
	acc(node);
}

/* Horner sums of the local expansion (alpha_local.cuh:18-102):
   h[k] = sum_i i(i-1)..(i-k+1) c_i dz^(i-k), k = 1, 2, 3 */
void l2p(const Lens& L, const Node* node, cx z, cx h[4])
{
	h[1] = h[2] = h[3] = cx();
	if (!L.use_tree) return;
	cx dz = (z - node->center) / node->hl;
	for (int i = L.p; i >= 1; i--) { h[1] *= dz; h[1] += node->loc[i] * (double)i; }
	for (int i = L.p; i >= 2; i--) { h[2] *= dz; h[2] += node->loc[i] * (double)(i * (i - 1)); }
	for (int i = L.p; i >= 3; i--) { h[3] *= dz; h[3] += node->loc[i] * (double)(i * (i - 1) * (i - 2)); }
}

double heaviside(double x) { return x >= 0 ? 1 : 0; }
double boxcar(cx z, cx c) { return (-c.re <= z.re && z.re <= c.re && -c.im <= z.im && z.im <= c.im) ? 1 : 0; }

/* alpha_smooth.cuh:62-127 */
cx alpha_smooth(const Lens& L, cx z)
{
	const double ks = L.kappastar;
	const cx c = L.corner;
	cx a;
	if (L.rectangular)
	{
		if (L.approx)
		{
			cx s1, s2, s3, s4;
			for (int i = L.taylor_smooth; i >= 1; i--)
			{
				s1 += cx(1.0 / i); s2 += cx(1.0 / i); s3 += cx(1.0 / i); s4 += cx(1.0 / i);
				s1 *= z.conj() / c;
				s2 *= z.conj() / c.conj();
				s3 *= z.conj() / -c;
				s4 *= z.conj() / -c.conj();
			}
			a = (c - z.conj()) * (c.log() - s1) - (c.conj() - z.conj()) * (c.conj().log() - s2)
				+ (-c - z.conj()) * ((-c).log() - s3) - (-c.conj() - z.conj()) * ((-c).conj().log() - s4);
			a *= cx(0, -ks / PI);
			a -= cx(ks * 2 * (c.re + z.re));
		}
		else
		{
			cx c1 = c - z.conj(), c2 = c.conj() - z.conj(), c3 = -c - z.conj(), c4 = -c.conj() - z.conj();
			a = c1 * c1.log() - c2 * c2.log() + c3 * c3.log() - c4 * c4.log();
			a *= cx(0, -ks / PI);
			a -= cx(ks * 2 * (c.re + z.re) * boxcar(z, c));
			a -= cx(ks * 4 * c.re * heaviside(c.im + z.im) * heaviside(c.im - z.im) * heaviside(z.re - c.re));
		}
	}
	else
	{
		if (L.approx || z.abs() <= c.abs()) a = -ks * z;
		else a = -ks * c.abs() * c.abs() / (L.plant == ORACLE_PLANT_DISC_OUTSIDE_Z ? z : z.conj());
	}
	return a;
}

/* alpha_smooth.cuh:141-170 */
double d_alpha_smooth_d_z(const Lens& L, cx z)
{
	if (L.rectangular) return L.approx ? -L.kappastar : -L.kappastar * boxcar(z, L.corner);
	return L.approx ? -L.kappastar : -L.kappastar * heaviside(L.corner.abs() - z.abs());
}

/* alpha_smooth.cuh:185-236 */
cx d_alpha_smooth_d_zbar(const Lens& L, cx z)
{
	const double ks = L.kappastar;
	const cx c = L.corner;
	cx d;
	if (L.rectangular)
	{
		if (L.approx)
		{
			cx r = z.conj() / c;
			cx phase(0, c.arg());
			for (int i = L.taylor_smooth - 1; i >= 2; i -= 2)
			{
				d += (cx(1) - (2.0 * phase * (double)i).exp()) / (double)i;
				d *= r * r;
			}
			d *= 2.0;
			d *= cx(0, -ks / PI);
			d += cx(ks - 4 * ks * c.arg() / PI);
		}
		else
		{
			cx c1 = c.conj() - z.conj(), c2 = c - z.conj(), c3 = -c - z.conj(), c4 = -c.conj() - z.conj();
			d = c1.log() - c2.log() - c3.log() + c4.log();
			d *= cx(0, -ks / PI);
			d -= cx(ks * boxcar(z, c));
		}
	}
	else
	{
		if (!L.approx && !(L.plant == ORACLE_PLANT_DISC_DZBAR_INSIDE && z.abs() <= c.abs()))
			d = ks * c.abs() * c.abs() / z.conj().pow(2);
	}
	return d;
}

/* alpha_smooth.cuh:251-301 */
cx d2_alpha_smooth_d_zbar2(const Lens& L, cx z)
{
	const double ks = L.kappastar;
	const cx c = L.corner;
	cx d;
	if (L.rectangular)
	{
		if (L.approx)
		{
			cx r = z.conj() / c;
			cx phase(0, c.arg());
			for (int i = L.taylor_smooth - 1; i >= 2; i -= 2)
			{
				d += cx(1) - (2.0 * phase * (double)i).exp();
				d *= r * r;
			}
			d /= z.conj();
			d *= 2.0;
			d *= cx(0, -ks / PI);
		}
		else
		{
			cx c1 = c.conj() - z.conj(), c2 = c - z.conj(), c3 = -c - z.conj(), c4 = -c.conj() - z.conj();
			d = -1.0 / c1 + 1.0 / c2 + 1.0 / c3 - 1.0 / c4;
			d *= cx(0, -ks / PI);
		}
	}
	else
	{
		if (!L.approx) d = -2 * ks * c.abs() * c.abs() / z.conj().pow(3);
	}
	return d;
}

/* everything the solver and the output kernels need at one position */
struct Eval
{
	cx f0, f1, w;
	double mu_scale = 0, err = 0, dsz = 0;
	int nnf = 0;
};

/* ccf_functions.cuh:33-166, 449-463 */
Eval evaluate(const Lens& L, cx z, double phi, bool want_f1 = true)
{
	Eval e;
	const Node* node = nearest_node(z, L.root);
	cx s1, s2, s3;
	p2p(L, node, z, s1, s2, s3, e.nnf);
	cx h[4];
	l2p(L, node, z, h);
	double t2 = L.theta * L.theta;
	double hl = node->hl;
	/* deflections and derivatives with the reference's signs and conjugations */
	cx a_star = (s1 * t2).conj();
	cx da_star = (s2 * (-t2)).conj();
	cx d2a_star = (s3 * (2 * t2)).conj();
	cx a_loc = (h[1] * t2 / hl).conj();
	cx da_loc = (h[2] * t2 / (hl * hl)).conj();
	cx d2a_loc = (h[3] * t2 / (hl * hl * hl)).conj();
	double dsz = d_alpha_smooth_d_z(L, z);
	cx dsb = d_alpha_smooth_d_zbar(L, z);
	e.dsz = dsz;
	e.f0 = cx(L.gamma) - da_star.conj() - da_loc.conj() - dsb.conj() - (1 - L.kappa - dsz) * cx(std::cos(phi), -std::sin(phi));
	double a0 = e.f0.abs();
	double am = 1 - L.kappa - (L.plant == ORACLE_PLANT_BOUND_SMOOTH_SIGN ? -dsz : dsz);
	e.err = std::max(std::fabs(a0 * (a0 + 2 * am)), std::fabs(a0 * (a0 - 2 * am)));
	if (!want_f1) return e;
	cx d2sb = d2_alpha_smooth_d_zbar2(L, z);
	e.f1 = -d2a_star.conj() - d2a_loc.conj() - d2sb.conj();
	e.w = (1 - L.kappa) * z + L.gamma * z.conj() - (L.plant == ORACLE_PLANT_W_STAR_UNCONJUGATED ? a_star.conj() : a_star)
		- a_loc - alpha_smooth(L, z);
	/* ccf_functions.cuh:145-166 */
	double dwdz = 1 - L.kappa - dsz;
	cx dwdzb = cx(L.gamma) - da_star - da_loc - dsb;
	cx d2wdzb2 = -d2a_star - d2a_loc - d2sb;
	cx cct = cx(0, -2) * dwdzb.conj() * d2wdzb2;
	cx ct = dwdz * cct + dwdzb * cct.conj();
	e.mu_scale = 1 / ((L.plant == ORACLE_PLANT_MU_SCALE_FACTOR ? 1 : 2) * ct.abs());
	return e;
}

/* sum over i in [a, b) of 1/(z - (xs[i] + i ys[i])) */
cx cauchy_sum(cx z, const double* xs, const double* ys, int a, int b)
{
	double ax = 0, ay = 0;
#pragma omp simd reduction(+ : ax, ay)
	for (int i = a; i < b; i++)
	{
		const double dx = z.re - xs[i], dy = z.im - ys[i];
		const double q = 1.0 / (dx * dx + dy * dy);
		ax += dx * q;
		ay += dy * q;
	}
	return cx(ax, -ay);
}

struct Counters
{
	long long full = 0, early = 0;
	double flops = 0;
};

/* one Aberth update (ccf_functions.cuh:193-243); roots is the whole current
   row (Jacobi: all reads from the previous sweep) with rx, ry its real and
   imaginary parts, k the updated slot */
cx aberth_update(const Lens& L, int k, const cx* roots, const double* rx, const double* ry, int nroots, double phi, Counters& cnt)
{
	cx z = roots[k];
	Eval e = evaluate(L, z, phi, false);
	double a0 = e.f0.abs();
	double am = 1 - L.kappa - e.dsz;
	const int nsm = (L.rectangular && L.approx) ? 8 * (L.taylor_smooth - 1) : 0;
	if (std::fabs(a0 * (a0 + 2 * am)) < 1e-9 && std::fabs(a0 * (a0 - 2 * am)) < 1e-9)
	{
		cnt.early++;
		cnt.flops += 18.0 * e.nnf + 8.0 * L.p;
		return z;
	}
	const Node* node = nearest_node(z, L.root);
	cx s1, s2, s3;
	int nnf;
	p2p(L, node, z, s1, s2, s3, nnf);
	cx h[4];
	l2p(L, node, z, h);
	double t2 = L.theta * L.theta;
	cx d2a_star = (s3 * (2 * t2)).conj();
	cx d2a_loc = (h[3] * t2 / (node->hl * node->hl * node->hl)).conj();
	cx f1 = -d2a_star.conj() - d2a_loc.conj() - d2_alpha_smooth_d_zbar2(L, z).conj();
	/* the O(N + R) Cauchy sums, as 1/(z - w) = conj(z - w) / |z - w|^2 over
	   structure-of-arrays copies so the compiler vectorises them */
	int np = (int)L.sx->size();
	cx psum = cauchy_sum(z, L.sx->data(), L.sy->data(), 0, np);
	if (!L.rectangular && !L.approx) { psum += 1.0 / z; np++; }
	psum *= 2.0;
	cx rsum = cauchy_sum(z, rx, ry, 0, k) + cauchy_sum(z, rx, ry, k + 1, nroots);
	cnt.full++;
	cnt.flops += 10.0 * (np + nroots - 1) + 36.0 * nnf + 16.0 * L.p + nsm;
	return z - e.f0 / (f1 + e.f0 * (psum - rsum));
}

/* ---------------------------------------------------------------------------
   pipeline (ccf.cuh:218-1260)
   ------------------------------------------------------------------------- */
struct Setup
{
	Field field;
	std::vector<Star> sorted;  /* tree order */
	std::vector<double> sx, sy;
	Tree tree;
	Binomials binom;
	Lens lens;
	int num_roots = 0;
	double mu_ave = 0, alpha_error = 0, root_hl = 0;
};

bool check_params(const oracle_params& p)
{
	auto err = [](const std::string& m) { std::cerr << m; return false; };
	const double dmin = std::numeric_limits<double>::min();
	if (p.kappa_tot < dmin) return err("Error. kappa_tot must be >= " + std::to_string(dmin) + "\n");
	if (p.approx != 0 && p.approx != 1) return err("Error. approx must be 1 (approximate) or 0 (exact).\n");
	if (p.safety_scale < 1.1) return err("Error. safety_scale must be >= 1.1\n");
	if (p.num_phi < 1 || p.num_phi % 2 != 0) return err("Error. num_phi must be an even integer > 0\n");
	if (p.num_branches < 1) return err("Error. num_branches must be an integer > 0\n");
	if (p.num_phi % (2 * p.num_branches) != 0) return err("Error. num_phi must be a multiple of 2 * num_branches\n");
	if (p.starfile[0] == 0) return err("oracle: a star file is required\n");
	return true;
}

/* ccf.cuh:348-497 (star-file branch) and 732-904 */
bool setup(const oracle_params& p, Setup& S, oracle_stats* stats, std::vector<cx>* seeds)
{
	if (!check_params(p)) return false;
	Field& f = S.field;
	if (!read_star_file(p.starfile, f))
	{
		std::cerr << "Error. Unable to read star field parameters from file " << p.starfile << "\n";
		return false;
	}
	calculate_star_params(f);
	S.mu_ave = 1 / ((1 - p.kappa_tot) * (1 - p.kappa_tot) - p.shear * p.shear);
	S.alpha_error = f.theta * 0.0000001;
	/* taylor_smooth (ccf.cuh:441-464) */
	int ts = 1;
	const double kst = f.kappa_star;
	while ((kst / PI * 4 / (ts + 1) * f.corner.abs() * (p.safety_scale + 1) / (p.safety_scale - 1) * std::pow(1 / p.safety_scale, ts + 1) > S.alpha_error) && ts <= 101)
		ts += 2;
	while ((std::fmod(f.corner.arg() * (ts - 1), PI) < 0.1 * PI || std::fmod(f.corner.arg() * (ts - 1), PI) > 0.9 * PI) && ts <= 101)
		ts += 2;
	if (f.rectangular && ts > 101) { std::cerr << "Error. taylor_smooth must be <= 101\n"; return false; }
	/* number of roots (ccf.cuh:469-491) */
	if (f.rectangular) S.num_roots = p.approx ? 2 * f.nstars + ts - 1 : 2 * f.nstars;
	else S.num_roots = p.approx ? 2 * f.nstars : 2 * f.nstars + 2;
	/* root seeds in file order (ccf.cuh:711-726) */
	if (seeds)
	{
		seeds->assign(S.num_roots, cx());
		for (int i = 0; i < f.nstars; i++)
		{
			(*seeds)[i] = f.stars[i].pos + cx(1);
			(*seeds)[i + f.nstars] = f.stars[i].pos - cx(1);
		}
		int nx = S.num_roots - 2 * f.nstars;
		for (int i = 0; i < nx; i++)
			(*seeds)[2 * f.nstars + i] = f.corner.abs() * cx(std::cos(2 * PI / nx * i), std::sin(2 * PI / nx * i));
	}
	/* tree and FMM (ccf.cuh:738-893) */
	auto t0 = std::chrono::steady_clock::now();
	S.root_hl = (f.rectangular ? std::max(f.corner.re, f.corner.im) : f.corner.abs()) * 1.1;
	S.sorted = f.stars;
	build_tree(S.tree, S.sorted, S.root_hl, p.direct != 0);
	int order;
	if (p.fmm_order > 0) order = p.fmm_order;
	else order = (int)std::ceil(2 * std::log2(f.theta) + S.tree.tree_levels - std::log2(S.root_hl) - std::log2(S.alpha_error));
	if (!p.direct)
	{
		if (order < 3) { std::cerr << "Error. Expansion order needs to be >= 3\n"; return false; }
		if (order > MAX_ORDER) { std::cerr << "Error. Maximum allowed expansion order is " << MAX_ORDER << "\n"; return false; }
		S.binom.build(2 * order, p.binomial_exact == 0);
		build_fmm(S.tree, order, S.binom, S.sorted);
	}
	auto t1 = std::chrono::steady_clock::now();
	Lens& L = S.lens;
	L.kappa = p.kappa_tot;
	L.gamma = p.shear;
	L.theta = f.theta;
	L.kappastar = f.kappa_star;
	L.rectangular = f.rectangular;
	L.approx = p.approx;
	L.taylor_smooth = ts;
	L.corner = f.corner;
	L.p = p.direct ? 0 : order;
	L.use_tree = !p.direct;
	L.plant = p.plant;
	L.stars = &S.sorted;
	S.sx.resize(S.sorted.size());
	S.sy.resize(S.sorted.size());
	for (size_t i = 0; i < S.sorted.size(); i++)
	{
		S.sx[i] = S.sorted[i].pos.re;
		S.sy[i] = S.sorted[i].pos.im;
	}
	L.sx = &S.sx;
	L.sy = &S.sy;
	L.root = S.tree.root();
	if (stats)
	{
		stats->num_stars = f.nstars;
		stats->num_roots = S.num_roots;
		stats->tree_levels = S.tree.tree_levels;
		stats->expansion_order = p.direct ? 0 : order;
		stats->taylor_smooth = ts;
		stats->num_nodes = S.tree.num_nodes();
		stats->kappa_star = f.kappa_star;
		stats->corner_re = f.corner.re;
		stats->corner_im = f.corner.im;
		stats->t_tree = std::chrono::duration<double>(t1 - t0).count();
	}
	return true;
}

/* Jacobi sweeps over one row until every root is converged or the cap is
   hit; fin[] holds per-root convergence (ccf_functions.cuh:362-389) */
void solve_row(const Lens& L, cx* row, int R, double phi, int cap, std::vector<char>& fin, Counters& cnt)
{
	std::vector<int> active(R);
	for (int a = 0; a < R; a++) { active[a] = a; fin[a] = 0; }
	std::vector<cx> next(R);
	std::vector<double> rx(R), ry(R);
	for (int it = 0; it < cap && !active.empty(); it++)
	{
		for (int a = 0; a < R; a++)
		{
			rx[a] = row[a].re;
			ry[a] = row[a].im;
		}
		int na = (int)active.size();
		std::vector<char> conv(na, 0);
		long long full = 0, early = 0;
		double flops = 0;
#pragma omp parallel for schedule(dynamic, 4) reduction(+ : full, early, flops)
		for (int i = 0; i < na; i++)
		{
			Counters c;
			int a = active[i];
			cx zn = aberth_update(L, a, row, rx.data(), ry.data(), R, phi, c);
			next[i] = zn;
			conv[i] = ((zn - row[a]).abs() / L.theta) < 1e-9;
			full += c.full;
			early += c.early;
			flops += c.flops;
		}
		cnt.full += full;
		cnt.early += early;
		cnt.flops += flops;
		std::vector<int> still;
		for (int i = 0; i < na; i++)
		{
			int a = active[i];
			row[a] = next[i];
			if (conv[i]) fin[a] = 1;
			else still.push_back(a);
		}
		active.swap(still);
	}
}

int run(const oracle_params& p, oracle_stats* stats)
{
#ifdef _OPENMP
	if (p.num_threads > 0) omp_set_num_threads(p.num_threads);
#endif
	Setup S;
	std::vector<cx> seeds;
	oracle_stats local{};
	if (!stats) stats = &local;
	*stats = oracle_stats{};
	if (!setup(p, S, stats, &seeds)) return 1;
	const Lens& L = S.lens;
	const int R = S.num_roots, np = p.num_phi, nb = p.num_branches;
	const int rows = np + nb, half = np / (2 * nb);
	const int init_cap = p.init_iters > 0 ? p.init_iters : 60;
	const int step_cap = p.step_iters > 0 ? p.step_iters : 20;
	const int nbs = p.branch_limit > 0 ? std::min(p.branch_limit, nb) : nb;
	const int steps = p.step_limit > 0 ? std::min(p.step_limit, half) : half;
	std::vector<cx> ccs_init((size_t)rows * R);
	std::vector<char> solved(rows, 0);
	Counters cnt;
	long long conv = 0;
	std::vector<char> fin(R);

	/* initial roots at every branch centre (ccf.cuh:906-941) */
	auto t0 = std::chrono::steady_clock::now();
	for (int c = 0; c < nbs; c++)
	{
		int center = half + c * np / nb + c;
		cx* row = &ccs_init[(size_t)center * R];
		std::copy(seeds.begin(), seeds.end(), row);
		solve_row(L, row, R, PI / nb + c * 2 * PI / nb, init_cap, fin, cnt);
		solved[center] = 1;
		for (int a = 0; a < R; a++) conv += fin[a];
	}
	auto t1 = std::chrono::steady_clock::now();
	/* phi continuation outward from every centre (ccf.cuh:979-1035) */
	for (int j = 1; j <= steps; j++)
	{
		for (int c = 0; c < nbs; c++)
		{
			int center = half + c * np / nb + c;
			double phi0 = PI / nb + c * 2 * PI / nb;
			for (int b = 0; b < 2; b++)
			{
				int sgn = b == 0 ? -1 : 1;
				cx* row = &ccs_init[(size_t)(center + sgn * j) * R];
				const cx* prev = &ccs_init[(size_t)(center + sgn * (j - 1)) * R];
				std::copy(prev, prev + R, row);
				solve_row(L, row, R, phi0 + sgn * 2 * PI / np * j, step_cap, fin, cnt);
				solved[center + sgn * j] = 1;
				for (int a = 0; a < R; a++) conv += fin[a];
			}
		}
	}
	auto t2 = std::chrono::steady_clock::now();
	/* 1/mu error bound at every solved point (ccf_functions.cuh:419-467) */
	double max_err = 0, max_res = 0;
	bool nan = false;
	for (int c = 0; c < nb; c++)
	{
		int center = half + c * np / nb + c;
		double phi0 = PI / nb + c * 2 * PI / nb;
		for (int k = -half; k <= half; k++)
		{
			int r = center + k;
			if (!solved[r]) continue;
			double phi = phi0 + 2 * PI / np * k;
			double lmax = 0, lres = 0;
			bool lnan = false;
#pragma omp parallel for reduction(max : lmax, lres) reduction(|| : lnan)
			for (int a = 0; a < R; a++)
			{
				const Eval ev = evaluate(L, ccs_init[(size_t)r * R + a], phi, false);
				const double e = ev.err;
				if (!std::isfinite(e)) lnan = true;
				else
				{
					lmax = std::max(lmax, e);
					lres = std::max(lres, ev.f0.abs());
				}
			}
			max_err = std::max(max_err, lmax);
			max_res = std::max(max_res, lres);
			nan = nan || lnan;
		}
	}
	if (nan)
	{
		std::cerr << "Error. Errors in 1/mu contain values which are not positive real numbers.\n";
		return 1;
	}
	/* transpose [row][root] -> [root][row] (ccf.cuh:1072-1080); with
	   compact output only the first branch_limit branches' rows, in the same
	   layout for that many branches */
	const int nb_out = (p.compact_output && p.branch_limit > 0) ? std::min(nb, p.branch_limit) : nb;
	const int cols = np / nb + 1, rows_out = nb_out * cols;
	std::vector<cx> ccs((size_t)rows_out * R);
	for (int c = 0; c < nb_out; c++)
		for (int k = 0; k < cols; k++)
		{
			const int r = c * (np / nb + 1) + k;
#pragma omp parallel for schedule(static, 4096)
			for (int a = 0; a < R; a++) ccs[((size_t)a * nb_out + c) * cols + k] = ccs_init[(size_t)r * R + a];
		}
	/* caustics and length scales (ccf.cuh:1085-1121) */
	auto t3 = std::chrono::steady_clock::now();
	std::vector<cx> caus;
	std::vector<double> mus;
	if (p.write_caustics || p.write_mu_length_scales)
	{
		if (p.write_caustics) caus.resize(ccs.size());
		if (p.write_mu_length_scales) mus.resize(ccs.size());
#pragma omp parallel for schedule(static, 256)
		for (size_t i = 0; i < ccs.size(); i++)
		{
			Eval e = evaluate(L, ccs[i], 0.0, true);
			if (p.write_caustics) caus[i] = e.w;
			if (p.write_mu_length_scales) mus[i] = e.mu_scale;
		}
	}
	auto t4 = std::chrono::steady_clock::now();

	stats->max_error = max_err;
	stats->max_witt_residual = max_res;
	stats->t_init_roots = std::chrono::duration<double>(t1 - t0).count();
	stats->t_ccs = std::chrono::duration<double>(t2 - t1).count();
	stats->t_caustics = std::chrono::duration<double>(t4 - t3).count();
	long long nsolved = 0;
	for (int r = 0; r < rows; r++) nsolved += solved[r];
	stats->points_solved = nsolved * R;
	stats->points_converged = conv;
	stats->updates_full = cnt.full;
	stats->updates_early = cnt.early;
	stats->flops = cnt.flops;

	/* output files (ccf.cuh:1123-1260) */
	const std::string pre = p.outfile_prefix;
	const Field& f = S.field;
	{
		std::ofstream out(pre + "ccf_parameter_info.txt");
		if (!out) { std::cerr << "Error. Failed to open file " << pre << "ccf_parameter_info.txt\n"; return 1; }
		out.precision(9);
		out << "kappa_tot " << p.kappa_tot << "\n";
		out << "shear " << p.shear << "\n";
		out << "mu_ave " << S.mu_ave << "\n";
		out << "smooth_fraction " << (1 - f.kappa_star / p.kappa_tot) << "\n";
		out << "kappa_star " << f.kappa_star << "\n";
		out << "theta_star " << f.theta << "\n";
		out << "random_seed " << 0 << "\n";
		out << "m_lower_actual " << f.m_lower << "\n";
		out << "m_upper_actual " << f.m_upper << "\n";
		out << "mean_mass_actual " << f.mean_mass << "\n";
		out << "mean_mass2_actual " << f.mean_mass2 << "\n";
		out << "mean_mass2_ln_mass_actual " << f.mean_mass2_ln_mass << "\n";
		out << "num_stars " << f.nstars << "\n";
		if (f.rectangular)
		{
			out << "corner_x1 " << f.corner.re << "\n";
			out << "corner_x2 " << f.corner.im << "\n";
			if (p.approx) out << "taylor_smooth " << L.taylor_smooth << "\n";
		}
		else
		{
			out << "rad " << f.corner.abs() << "\n";
		}
		out << "num_roots " << R << "\n";
		out << "num_phi " << np << "\n";
		out << "num_branches " << nb << "\n";
		out << "max_error_1/mu " << max_err << "\n";
		out << "t_init_roots " << stats->t_init_roots << "\n";
		out << "t_ccs " << stats->t_ccs << "\n";
		out << "t_caustics " << stats->t_caustics << "\n";
	}
	if (p.write_stars && !write_star_file(pre + "ccf_stars.bin", f, S.sorted)) return 1;
	if (p.write_critical_curves && !write_array(pre + "ccf_ccs.bin", ccs.data(), R * nb_out, cols)) return 1;
	if (p.write_caustics && !write_array(pre + "ccf_caustics.bin", caus.data(), R * nb_out, cols)) return 1;
	if (p.write_mu_length_scales && !write_array(pre + "ccf_mu_length_scales.bin", mus.data(), R * nb_out, cols)) return 1;
	return 0;
}

} // namespace orc

extern "C" {

void oracle_default_params(oracle_params* p)
{
	std::memset(p, 0, sizeof(*p));
	p->kappa_tot = 0.3;
	p->shear = 0.3;
	p->kappa_star = 0.27;
	p->theta_star = 1;
	p->m_solar = 1;
	p->m_lower = 0.01;
	p->m_upper = 50;
	p->safety_scale = 1.37;
	p->rectangular = 0;
	p->approx = 1;
	p->num_stars = 137;
	p->num_phi = 100;
	p->num_branches = 1;
	p->write_stars = 1;
	p->write_critical_curves = 1;
	p->write_caustics = 1;
	p->write_mu_length_scales = 0;
	p->verbose = 0;
	std::strcpy(p->mass_function, "equal");
	std::strcpy(p->outfile_prefix, "./");
}

int oracle_run(const oracle_params* p, oracle_stats* stats)
{
	try
	{
		return orc::run(*p, stats);
	}
	catch (const std::exception& e)
	{
		std::cerr << "oracle: " << e.what() << "\n";
		return 2;
	}
}

double oracle_binomial(int n, int k, int wrapped)
{
	orc::Binomials b;
	b.build(n, wrapped != 0);
	return b.get(n, k);
}

double oracle_mass_function(const char* name, int which, double p, double m_lower, double m_upper, double m_solar)
{
	return orc::mass_function(name, which, p, m_lower, m_upper, m_solar);
}

int oracle_eval(const oracle_params* p, int use_tree, const double* z, int n, double phi, double* out, oracle_stats* stats)
{
	orc::Setup S;
	oracle_params q = *p;
	if (!use_tree) q.direct = 1;
	if (!orc::setup(q, S, stats, nullptr)) return 1;
#pragma omp parallel for schedule(static, 64)
	for (int i = 0; i < n; i++)
	{
		orc::Eval e = orc::evaluate(S.lens, orc::cx(z[2 * i], z[2 * i + 1]), phi, true);
		double* o = out + 8 * (size_t)i;
		o[0] = e.f0.re; o[1] = e.f0.im;
		o[2] = e.f1.re; o[3] = e.f1.im;
		o[4] = e.w.re; o[5] = e.w.im;
		o[6] = e.mu_scale;
		o[7] = e.err;
	}
	return 0;
}

int oracle_eval_phis(const oracle_params* p, int use_tree, const double* z, const double* phis, int n, double* out)
{
	orc::Setup S;
	oracle_params q = *p;
	if (!use_tree) q.direct = 1;
	if (!orc::setup(q, S, nullptr, nullptr)) return 1;
#pragma omp parallel for schedule(static, 64)
	for (int i = 0; i < n; i++)
	{
		orc::Eval e = orc::evaluate(S.lens, orc::cx(z[2 * i], z[2 * i + 1]), phis[i], true);
		double* o = out + 8 * (size_t)i;
		o[0] = e.f0.re; o[1] = e.f0.im;
		o[2] = e.f1.re; o[3] = e.f1.im;
		o[4] = e.w.re; o[5] = e.w.im;
		o[6] = e.mu_scale;
		o[7] = e.err;
	}
	return 0;
}

int oracle_aberth_sweep(const oracle_params* p, const double* roots, int n, double phi, double* out)
{
	orc::Setup S;
	if (!orc::setup(*p, S, nullptr, nullptr)) return 1;
	std::vector<orc::cx> r(n);
	std::vector<double> rx(n), ry(n);
	for (int i = 0; i < n; i++)
	{
		r[i] = orc::cx(roots[2 * i], roots[2 * i + 1]);
		rx[i] = roots[2 * i];
		ry[i] = roots[2 * i + 1];
	}
#pragma omp parallel for schedule(dynamic, 4)
	for (int i = 0; i < n; i++)
	{
		orc::Counters c;
		orc::cx zn = orc::aberth_update(S.lens, i, r.data(), rx.data(), ry.data(), n, phi, c);
		out[2 * i] = zn.re;
		out[2 * i + 1] = zn.im;
	}
	return 0;
}

} // extern "C"
