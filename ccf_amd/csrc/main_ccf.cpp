// ccf_gpu: command line of cc_finder for MI355X.
//
// Same options, defaults, usage text, parse rules and exit codes as
// weisluke/ccf src/main_ccf.cu:16-522 (every option takes a value; unknown
// options and an odd argument count are errors; with -sf the star-field
// options are ignored; the first occurrence of a repeated option wins;
// -s sets kappa_star = (1 - s) kappa_tot only without -ks and -sf).
// Extension options (long form only) expose the fields added in ccf_amd.h.
#include "../../include/ccf.hpp"

#include <algorithm>
#include <cctype>
#include <iostream>
#include <string>

namespace
{

CCF<double> ccf;

const std::string OPTS[] = {
	"-h", "--help",
	"-v", "--verbose",
	"-k", "--kappa_tot",
	"-y", "--shear",
	"-s", "--smooth_fraction",
	"-ks", "--kappa_star",
	"-t", "--theta_star",
	"-mf", "--mass_function",
	"-ms", "--m_solar",
	"-ml", "--m_lower",
	"-mh", "--m_upper",
	"-r", "--rectangular",
	"-a", "--approx",
	"-ss", "--safety_scale",
	"-ns", "--num_stars",
	"-sf", "--starfile",
	"-np", "--num_phi",
	"-nb", "--num_branches",
	"-rs", "--random_seed",
	"-ws", "--write_stars",
	"-wcc", "--write_critical_curves",
	"-wc", "--write_caustics",
	"-wls", "--write_length_scales",
	"-o", "--outfile_prefix",
	"--fmm_order", "--direct", "--binomial_exact", "--precision",
	"--init_iters", "--step_iters", "--roots_per_thread", "--predictor",
	"--root_fmm", "--root_fmm_order", "--root_fmm_leaf", "--autotune",
};

int verbose = 1;
double smooth_fraction = 1 - ccf.kappa_star / ccf.kappa_tot;

bool exists(char** b, char** e, const std::string& o) { return std::find(b, e, o) != e; }

char* value(char** b, char** e, const std::string& o)
{
	char** it = std::find(b, e, o);
	if (it != e && ++it != e) return *it;
	return nullptr;
}

void usage(const char* name)
{
	std::cout << "Usage: " << (name ? name : "programname") << " opt1 val1 opt2 val2 opt3 val3 ...\n";
	std::cout
		<< "                                                                               \n"
		<< "Options:\n"
		<< "  -h,--help                  Show this help message\n"
		<< "  -v,--verbose               Specify verbosity of output from 0 (none) to 1, 2,\n"
		<< "                             or 3 (low, medium, high). Default value: " << verbose << "\n"
		<< "  -k,--kappa_tot             Specify the total convergence. Default value: " << ccf.kappa_tot << "\n"
		<< "  -y,--shear                 Specify the shear. Default value: " << ccf.shear << "\n"
		<< "  -s,--smooth_fraction       Specify the fraction of convergence due to\n"
		<< "                             smoothly distributed mass. Default value: " << smooth_fraction << "\n"
		<< "  -ks,--kappa_star           Specify the convergence in point mass lenses. If\n"
		<< "                             provided, this overrides any supplied value for\n"
		<< "                             the smooth fraction. Default value: " << ccf.kappa_star << "\n"
		<< "  -t,--theta_star            Specify the size of the Einstein radius of a unit\n"
		<< "                             mass point lens in arbitrary units.\n"
		<< "                             Default value: " << ccf.theta_star << "\n"
		<< "  -mf,--mass_function        Specify the mass function to use for the point\n"
		<< "                             mass lenses. Options are: equal, uniform,\n"
		<< "                             Salpeter, and, Kroupa, or optical_depth.\n"
		<< "                             Default value: " << ccf.mass_function_str << "\n"
		<< "  -ms,--m_solar              Specify the solar mass in arbitrary units.\n"
		<< "                             Default value: " << ccf.m_solar << "\n"
		<< "  -ml,--m_lower              Specify the lower mass cutoff in solar mass units.\n"
		<< "                             Default value: " << ccf.m_lower << "\n"
		<< "  -mh,--m_upper              Specify the upper mass cutoff in solar mass units.\n"
		<< "                             Default value: " << ccf.m_upper << "\n"
		<< "  -r,--rectangular           Specify whether the star field should be\n"
		<< "                             rectangular (1) or circular (0). Default value: " << ccf.rectangular << "\n"
		<< "  -a,--approx                Specify whether terms for alpha_smooth should be\n"
		<< "                             approximated (1) or exact (0). Default value: " << ccf.approx << "\n"
		<< "  -ss,--safety_scale         Specify the ratio of the size of the star field to\n"
		<< "                             the radius of convergence for alpha_smooth.\n"
		<< "                             Default value: " << ccf.safety_scale << "\n"
		<< "  -ns,--num_stars            Specify the number of stars desired.\n"
		<< "                             Default value: " << ccf.num_stars << "\n"
		<< "  -sf,--starfile             Specify the location of a binary file containing\n"
		<< "                             values for num_stars, rectangular, corner,\n"
		<< "                             theta_star, and the star positions and masses, in\n"
		<< "                             an order as defined in this source code.\n"
		<< "                             A whitespace delimited text file where each line\n"
		<< "                             contains the x1 and x2 coordinates and the mass of\n"
		<< "                             a microlens, in units where theta_star = 1, is\n"
		<< "                             also accepted. If provided, this takes precedence\n"
		<< "                             for all star information.\n"
		<< "  -np,--num_phi              Specify the number of steps used to vary phi in\n"
		<< "                             the range [0, 2*pi]. Default value: " << ccf.num_phi << "\n"
		<< "  -nb,--num_branches         Specify the number of branches to use for phi in\n"
		<< "                             the range [0, 2*pi]. Default value: " << ccf.num_branches << "\n"
		<< "  -rs,--random_seed          Specify the random seed for star field generation.\n"
		<< "                             A value of 0 is reserved for star input files.\n"
		<< "  -ws,--write_stars          Specify whether to write stars (1) or not (0).\n"
		<< "                             Default value: " << ccf.write_stars << "\n"
		<< "  -wcc,                      Specify whether to write critical curves (1) or\n"
		<< "   --write_critical_curves   not (0). Default value: " << ccf.write_critical_curves << "\n"
		<< "  -wc,--write_caustics       Specify whether to write caustics (1) or not (0).\n"
		<< "                             Default value: " << ccf.write_caustics << "\n"
		<< "  -wls,                      Specify whether to write magnification length\n"
		<< "   --write_length_scales     scales (1) or not (0). Default value: " << ccf.write_mu_length_scales << "\n"
		<< "  -o,--outfile_prefix        Specify the prefix to be used in output\n"
		<< "                             filenames. Default value: " << ccf.outfile_prefix << "\n"
		<< "Extensions:\n"
		<< "  --fmm_order                FMM expansion order in [3, 31]; 0 derives it from\n"
		<< "                             the tree as the reference does. Default value: 0\n"
		<< "  --direct                   Sum every star directly, no FMM (1) or not (0).\n"
		<< "                             Default value: 0\n"
		<< "  --binomial_exact           Exact binomial coefficients in M2L (1) or the\n"
		<< "                             reference's int32 table (0). Default value: 0\n"
		<< "  --precision                0: FP64 throughout; 1: mixed (FP32 Cauchy sums,\n"
		<< "                             FP64 Witt function and polish). Default value: 0\n"
		<< "  --init_iters               Cap of Aberth sweeps at the branch centres.\n"
		<< "                             Default value: 60\n"
		<< "  --step_iters               Cap of Aberth sweeps per phi step. Default value: 20\n"
		<< "  --roots_per_thread         Aberth tile roots per thread (1, 2, 4; 0 automatic).\n"
		<< "  --predictor                Initial roots of each phi step: 0 the previous\n"
		<< "                             step's roots; 1..4 polynomial extrapolation of\n"
		<< "                             that degree.\n"
		<< "                             Default value: 0\n"
		<< "  --root_fmm                 Root sums of dense sweeps by a fast multipole\n"
		<< "                             method (1) or directly (0). Default value: 0\n"
		<< "  --root_fmm_order           Expansion order of the root-sum FMM (0 automatic).\n"
		<< "  --root_fmm_leaf            Mean roots per leaf cell of its grid (0 automatic).\n"
		<< "  --autotune                 1: time the automatic choices of roots_per_thread,\n"
		<< "                             root_fmm_order and root_fmm_leaf on the first\n"
		<< "                             branch centres and keep the fastest.\n"
		<< "                             Default value: 0\n";
}

template <typename V, typename F>
bool parse(const char* name, char* in, V& dst, F conv)
{
	try
	{
		dst = static_cast<V>(conv(std::string(in ? in : "")));
		if (verbose >= 2) std::cout << name << " set to: " << dst << "\n";
		return true;
	}
	catch (...)
	{
		std::cerr << "Error. Invalid " << name << " input.\n";
		return false;
	}
}

double to_d(const std::string& s) { return std::stod(s); }
int to_i(const std::string& s) { return std::stoi(s); }

} // namespace

int main(int argc, char* argv[])
{
	char** b = argv;
	char** e = argv + argc;
	const std::string* ob = OPTS;
	const std::string* oe = OPTS + sizeof(OPTS) / sizeof(OPTS[0]);

	if (exists(b, e, "-h") || exists(b, e, "--help"))
	{
		usage(argv[0]);
		return -1;
	}
	if ((argc - 1) % 2 != 0)
	{
		std::cerr << "Error. Invalid input syntax.\n";
		usage(argv[0]);
		return -1;
	}
	for (int i = 1; i < argc; i += 2)
	{
		if (std::find(ob, oe, std::string(argv[i])) == oe)
		{
			std::cerr << "Error. Invalid input syntax. Unknown option " << argv[i] << "\n";
			usage(argv[0]);
			return -1;
		}
		if (argv[i] == std::string("-v") || argv[i] == std::string("--verbose"))
		{
			char* in = value(b, e, argv[i]);
			try
			{
				verbose = std::stoi(in ? in : "");
				if (verbose >= 2) std::cout << "verbose set to: " << verbose << "\n";
			}
			catch (...)
			{
				std::cerr << "Error. Invalid verbose input.\n";
				return -1;
			}
		}
	}

	const bool has_sf = exists(b, e, "-sf") || exists(b, e, "--starfile");
	const bool has_ks = exists(b, e, "-ks") || exists(b, e, "--kappa_star");
	for (int i = 1; i < argc; i += 2)
	{
		const std::string o = argv[i];
		char* in = value(b, e, o);
		bool ok = true;
		if (o == "-k" || o == "--kappa_tot") ok = parse("kappa_tot", in, ccf.kappa_tot, to_d);
		else if (o == "-y" || o == "--shear") ok = parse("shear", in, ccf.shear, to_d);
		else if (o == "-s" || o == "--smooth_fraction")
		{
			if (has_sf || has_ks) continue;
			ok = parse("smooth_fraction", in, smooth_fraction, to_d);
			if (ok && smooth_fraction < 0)
			{
				std::cerr << "Error. Invalid smooth_fraction input. smooth_fraction must be >= 0\n";
				return -1;
			}
			if (ok && smooth_fraction >= 1)
			{
				std::cerr << "Error. Invalid smooth_fraction input. smooth_fraction must be < 1\n";
				return -1;
			}
		}
		else if (o == "-ks" || o == "--kappa_star") { if (has_sf) continue; ok = parse("kappa_star", in, ccf.kappa_star, to_d); }
		else if (o == "-t" || o == "--theta_star") { if (has_sf) continue; ok = parse("theta_star", in, ccf.theta_star, to_d); }
		else if (o == "-mf" || o == "--mass_function")
		{
			if (has_sf) continue;
			std::string s = in ? in : "";
			std::transform(s.begin(), s.end(), s.begin(), [](unsigned char ch) { return std::tolower(ch); });
			ccf.mass_function_str = s;
			if (verbose >= 2) std::cout << "mass_function set to: " << s << "\n";
		}
		else if (o == "-ms" || o == "--m_solar") ok = parse("m_solar", in, ccf.m_solar, to_d);
		else if (o == "-ml" || o == "--m_lower") { if (has_sf) continue; ok = parse("m_lower", in, ccf.m_lower, to_d); }
		else if (o == "-mh" || o == "--m_upper") { if (has_sf) continue; ok = parse("m_upper", in, ccf.m_upper, to_d); }
		else if (o == "-r" || o == "--rectangular") { if (has_sf) continue; ok = parse("rectangular", in, ccf.rectangular, to_i); }
		else if (o == "-a" || o == "--approx") ok = parse("approx", in, ccf.approx, to_i);
		else if (o == "-ss" || o == "--safety_scale") ok = parse("safety_scale", in, ccf.safety_scale, to_d);
		else if (o == "-ns" || o == "--num_stars") { if (has_sf) continue; ok = parse("num_stars", in, ccf.num_stars, to_i); }
		else if (o == "-sf" || o == "--starfile")
		{
			ccf.starfile = in ? in : "";
			if (verbose >= 2) std::cout << "starfile set to: " << ccf.starfile << "\n";
		}
		else if (o == "-np" || o == "--num_phi") ok = parse("num_phi", in, ccf.num_phi, to_i);
		else if (o == "-nb" || o == "--num_branches") ok = parse("num_branches", in, ccf.num_branches, to_i);
		else if (o == "-rs" || o == "--random_seed") { if (has_sf) continue; ok = parse("random_seed", in, ccf.random_seed, to_i); }
		else if (o == "-ws" || o == "--write_stars") ok = parse("write_stars", in, ccf.write_stars, to_i);
		else if (o == "-wcc" || o == "--write_critical_curves") ok = parse("write_critical_curves", in, ccf.write_critical_curves, to_i);
		else if (o == "-wc" || o == "--write_caustics") ok = parse("write_caustics", in, ccf.write_caustics, to_i);
		else if (o == "-wls" || o == "--write_length_scales") ok = parse("write_length_scales", in, ccf.write_mu_length_scales, to_i);
		else if (o == "-o" || o == "--outfile_prefix")
		{
			ccf.outfile_prefix = in ? in : "";
			if (verbose >= 2) std::cout << "outfile_prefix set to: " << ccf.outfile_prefix << "\n";
		}
		else if (o == "--fmm_order") ok = parse("fmm_order", in, ccf.fmm_order, to_i);
		else if (o == "--direct") ok = parse("direct", in, ccf.direct, to_i);
		else if (o == "--binomial_exact") ok = parse("binomial_exact", in, ccf.binomial_exact, to_i);
		else if (o == "--precision") ok = parse("precision", in, ccf.precision, to_i);
		else if (o == "--init_iters") ok = parse("init_iters", in, ccf.init_iters, to_i);
		else if (o == "--step_iters") ok = parse("step_iters", in, ccf.step_iters, to_i);
		else if (o == "--roots_per_thread") ok = parse("roots_per_thread", in, ccf.roots_per_thread, to_i);
		else if (o == "--predictor") ok = parse("predictor", in, ccf.predictor, to_i);
		else if (o == "--root_fmm") ok = parse("root_fmm", in, ccf.root_fmm, to_i);
		else if (o == "--root_fmm_order") ok = parse("root_fmm_order", in, ccf.root_fmm_order, to_i);
		else if (o == "--root_fmm_leaf") ok = parse("root_fmm_leaf", in, ccf.root_fmm_leaf, to_i);
		else if (o == "--autotune") ok = parse("autotune", in, ccf.autotune, to_i);
		if (!ok) return -1;
	}
	if (!has_sf && !has_ks && (exists(b, e, "-s") || exists(b, e, "--smooth_fraction")))
	{
		ccf.kappa_star = (1 - smooth_fraction) * ccf.kappa_tot;
		if (verbose >= 2) std::cout << "kappa_star set to: " << ccf.kappa_star << "\n";
	}
	if (verbose >= 2) std::cout << "\n";

	if (!ccf.run(verbose)) return -1;
	if (!ccf.save(verbose)) return -1;
	return 0;
}
