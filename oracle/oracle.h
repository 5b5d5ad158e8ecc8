/*
 * FP64 CPU oracle of cc_finder's hot path (weisluke/ccf).
 *
 * TEST INFRASTRUCTURE ONLY. This library is the checker for the HIP product
 * path and the timed CPU baseline in bench.py. Nothing in the product
 * (ccf_amd/) links, loads or calls it; only tests/, __graft_entry__.smoke()
 * and bench.py's cpu_baseline leg use it.
 *
 * It restates, in plain sequential C++ (OpenMP over independent roots), the
 * reference's algorithm: quadtree (include/tree_node.cuh), 2-D complex FMM
 * (include/fmm.cuh, include/binomial_coefficients.cuh), star/local/smooth
 * deflections (include/alpha_*.cuh), Witt's parametric critical-curve
 * equation and the Aberth update (include/ccf_functions.cuh) and the
 * pipeline of include/ccf.cuh (derived parameters, root seeding, initial
 * solve, phi continuation, 1/mu error bound, caustics, length scales and
 * the output file layout).
 *
 * Deliberate, documented differences from the reference:
 *  - Aberth sweeps are Jacobi (every update of a sweep reads the roots of
 *    the previous sweep), where the reference updates roots in place while
 *    other threads read them. The fixed points are the same.
 *  - The initial solve at a branch centre updates each root once per sweep,
 *    with a cap of 60 sweeps; the reference launches 30 kernels in which
 *    the b=0 and b=1 threads both update the same slot (2 updates/launch).
 *  - Tree construction is deterministic (children in parent order, stars
 *    stably partitioned); the reference orders both by atomics.
 *  - Stars must come from a star file (the reference's cuRAND stream is
 *    not reproducible here).
 */
#pragma once

#ifdef __cplusplus
extern "C" {
#endif

typedef struct oracle_params
{
	double kappa_tot;
	double shear;
	double kappa_star;
	double theta_star;
	double m_solar;
	double m_lower;
	double m_upper;
	double safety_scale;
	int rectangular;
	int approx;
	int num_stars;
	int num_phi;
	int num_branches;
	int random_seed;
	int write_stars;
	int write_critical_curves;
	int write_caustics;
	int write_mu_length_scales;
	int verbose;
	/* extensions (0 selects the reference behaviour) */
	int fmm_order;      /* >0 overrides the reference's order rule */
	int direct;         /* 1: no tree splitting, every star summed directly */
	int binomial_exact; /* 1: exact binomials in M2L; 0: int32 Pascal rows as the reference */
	int init_iters;     /* cap of Jacobi sweeps at the branch centres (0 -> 60) */
	int step_iters;     /* cap of sweeps per phi step (0 -> 20) */
	int step_limit;     /* >0: stop the continuation after this many phi steps (timing samples) */
	int branch_limit;   /* >0: solve only the first branch_limit branches (timing samples) */
	int num_threads;    /* OpenMP threads (0 -> runtime default) */
	int compact_output; /* 1 with branch_limit > 0: the output arrays hold only the solved branches
	                       (the file layout for num_branches = branch_limit) */
	int plant;          /* tests of the tests only: 0 the restated algorithm; k > 0 plants
	                       mistake k (ORACLE_PLANT_*) so a test can show that its check fails */
	char mass_function[32];
	char starfile[512];
	char outfile_prefix[512];
} oracle_params;

typedef struct oracle_stats
{
	int num_stars;
	int num_roots;
	int tree_levels;
	int expansion_order;
	int taylor_smooth;
	int num_nodes;
	double kappa_star;       /* kappa_star used by the solve */
	double corner_re;
	double corner_im;
	double max_error;        /* max error in 1/mu over all solved points */
	double t_tree;           /* seconds */
	double t_init_roots;
	double t_ccs;
	double t_caustics;
	long long points_solved;    /* cc points whose rows were solved */
	long long points_converged; /* of those, converged (|dz|/theta < 1e-9 or 1/mu exit) */
	long long updates_full;     /* executed full Aberth updates */
	long long updates_early;    /* updates that returned on the 1/mu test */
	double flops;               /* algorithmic flops of the executed updates */
	double max_witt_residual;   /* max |f0| over the solved points */
} oracle_stats;

/* mistakes a test can plant in the lens-map family, each one a plausible slip */
enum
{
	ORACLE_PLANT_NONE = 0,
	ORACLE_PLANT_W_STAR_UNCONJUGATED = 1, /* w subtracts theta^2 sum m/(z - s), not its conjugate */
	ORACLE_PLANT_BOUND_SMOOTH_SIGN = 2,   /* 1/mu bound with 1 - kappa + d(alpha_smooth)/dz */
	ORACLE_PLANT_MU_SCALE_FACTOR = 3,     /* mu length scale 1/|caustic tangent| (factor 2 dropped) */
	ORACLE_PLANT_DISC_OUTSIDE_Z = 4,      /* exact disc outside: -kappa_star R^2 / z, not / zbar */
	ORACLE_PLANT_DISC_DZBAR_INSIDE = 5    /* exact disc: d(alpha_smooth)/dzbar = 0 inside (not the reference's) */
};

/* default-initialise a parameter block with the reference's defaults (ccf.cuh:42-62) */
void oracle_default_params(oracle_params* p);

/* full pipeline; writes the reference's output files under outfile_prefix.
   returns 0 on success, non-zero on any failure (message on stderr). */
int oracle_run(const oracle_params* p, oracle_stats* stats);

/* binomial coefficient as the reference tabulates it (wrapped=1: int32 Pascal
   rows, binomial_coefficients.cuh:11-44) or exactly (wrapped=0). */
double oracle_binomial(int n, int k, int wrapped);

/* mass function moments and inverse CDF (the mass_functions headers). which:
   0 mass(p), 1 mean_mass, 2 mean_mass2, 3 mean_mass2_ln_mass. m_lower and
   m_upper are in the arbitrary units the reference passes (already scaled
   by m_solar). returns NaN for an unknown mass function. */
double oracle_mass_function(const char* name, int which, double p, double m_lower, double m_upper, double m_solar);

/* evaluate, for the star field and tree that oracle_run would build for p,
   at n points z (interleaved re,im) and parametric angle phi:
     out[8*i+0..1] f0 (Witt's equation), out[8*i+2..3] f1 = d f0/dz,
     out[8*i+4..5] lens map w(z),       out[8*i+6]    mu length scale,
     out[8*i+7]    1/mu error bound.
   use_tree=0 sums every star directly (no FMM) for the same field. */
int oracle_eval(const oracle_params* p, int use_tree, const double* z, int n, double phi, double* out, oracle_stats* stats);

/* as oracle_eval, with its own angle phis[i] for each point (one tree build
   for points of many phi columns) */
int oracle_eval_phis(const oracle_params* p, int use_tree, const double* z, const double* phis, int n, double* out);

/* one Aberth update of every root in roots[0..n) (Jacobi), writing the
   updated roots to out; uses the same field/tree as oracle_eval. */
int oracle_aberth_sweep(const oracle_params* p, const double* roots, int n, double phi, double* out);

#ifdef __cplusplus
}
#endif
