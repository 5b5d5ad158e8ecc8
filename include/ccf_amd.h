/*
 * ccf_amd -- C-ABI of the MI355X (gfx950) critical-curve / caustic finder.
 *
 * The reference's drop-in boundary is the class template CCF<T>
 * (weisluke/ccf include/ccf.cuh:34-1289): callers set public fields,
 * then call bool run(int verbose) and bool save(int verbose). This header
 * exports that surface as plain C (POD parameter block, opaque context,
 * int status codes) so that the C++ class mirror in ccf.hpp, the ccf_gpu
 * command line, Python (ctypes) and any other FFI bind the same symbols.
 *
 * Conventions
 *  - Every int-returning call returns 0 on success and non-zero on failure.
 *    Failures print the reference's "Error. ..." message on stderr and keep
 *    it for ccf_last_error().
 *  - A context owns all of its device memory until ccf_destroy(). Unlike
 *    the reference (ccf.cuh:147-171) it never resets the device.
 *  - One context per host thread; contexts are independent. Device 0 only.
 */
#ifndef CCF_AMD_H
#define CCF_AMD_H

#ifdef __cplusplus
extern "C" {
#endif

#define CCF_PATH_MAX 1024

/* Public fields of CCF<T> (ccf.cuh:42-62) followed by extensions. */
typedef struct ccf_params
{
	double kappa_tot;            /* ccf.cuh:42 */
	double shear;                /* ccf.cuh:43 */
	double kappa_star;           /* ccf.cuh:44 */
	double theta_star;           /* ccf.cuh:45 */
	char mass_function[32];      /* ccf.cuh:46 (mass_function_str) */
	double m_solar;              /* ccf.cuh:47 */
	double m_lower;              /* ccf.cuh:48 */
	double m_upper;              /* ccf.cuh:49 */
	int rectangular;             /* ccf.cuh:50 */
	int approx;                  /* ccf.cuh:51 */
	double safety_scale;         /* ccf.cuh:52 */
	int num_stars;               /* ccf.cuh:53 */
	char starfile[CCF_PATH_MAX]; /* ccf.cuh:54 */
	int num_phi;                 /* ccf.cuh:55 */
	int num_branches;            /* ccf.cuh:56 */
	int random_seed;             /* ccf.cuh:57 */
	int write_stars;             /* ccf.cuh:58 */
	int write_critical_curves;   /* ccf.cuh:59 */
	int write_caustics;          /* ccf.cuh:60 */
	int write_mu_length_scales;  /* ccf.cuh:61 */
	char outfile_prefix[CCF_PATH_MAX]; /* ccf.cuh:62 */

	/* ---- extensions; 0 keeps the reference's behaviour ---- */
	int fmm_order;       /* >0: FMM expansion order instead of the rule at ccf.cuh:840-852 */
	int direct;          /* 1: no tree splitting; every star summed directly (no FMM) */
	int binomial_exact;  /* 1: exact binomials in M2L; 0: the reference's int32 Pascal rows */
	int precision;       /* 0: FP64 throughout; 1: mixed (FP32 Cauchy sums and far field, FP64 f0/polish) */
	int init_iters;      /* cap of Aberth sweeps at the branch centres (0 -> 60) */
	int step_iters;      /* cap of Aberth sweeps per phi step (0 -> 20) */
	int roots_per_thread;/* Aberth tile: roots per thread (0 -> autotuned default) */
	int timing;          /* 1: time every Aberth launch with HIP events (see ccf_info) */
	int predictor;       /* initial roots of phi step j: 0 the roots of step j-1 (the reference,
	                        ccf_functions.cuh:284-289); q = 1..4 the polynomial of degree q
	                        through steps j-1 .. j-1-q (fewer near the branch centre) */
	int root_fmm;        /* 1: the root sum sum_{j!=k} 1/(z_k - z_j) of sweeps in which many roots
	                        are active by a fast multipole method over a uniform grid of the field
	                        (exact FP64 near field, far field to order root_fmm_order); 0: direct
	                        sums as ccf_functions.cuh:232-239 */
	int root_fmm_order;  /* expansion order of the root-sum FMM (0 -> 8; at most 24) */
	int root_fmm_leaf;   /* mean roots per leaf cell the grid depth aims at (0 -> 6) */
	int autotune;        /* 1: the context's first solve times the automatic (0) choices of
	                        roots_per_thread (1, 2, 4) and, in the mixed form with root_fmm, of
	                        (root_fmm_order, root_fmm_leaf) over a few pairs, on the initial solve
	                        of the first group of branch centres, and keeps the fastest for its
	                        later runs (until ccf_set_params); none of them moves a root */
} ccf_params;

/* Derived values and measurements of the last run. */
typedef struct ccf_info
{
	int num_stars;
	int num_roots;
	int tree_levels;
	int expansion_order;
	int taylor_smooth;
	int num_nodes;
	int random_seed;
	double kappa_star;
	double kappa_star_actual;
	double mu_ave;
	double corner_re;
	double corner_im;
	double alpha_error;
	double max_error;          /* max error in 1/mu (ccf.cuh:972, 1068) */
	double t_tree;             /* seconds: tree + FMM build */
	double t_init_roots;       /* seconds (ccf.cuh:940) */
	double t_ccs;              /* seconds (ccf.cuh:1034) */
	double t_caustics;         /* seconds (ccf.cuh:1097) */
	long long points_total;    /* num_roots * (num_phi + num_branches) */
	long long points_converged;/* |dz|/theta < 1e-9 on the last sweep, or the 1/mu exit */
	long long updates_full;    /* executed Aberth updates that formed the sums */
	long long updates_early;   /* updates that returned on the 1/mu bound */
	double flops_update;       /* algorithmic flops of the executed updates */
	double aberth_kernel_ms;   /* sum of timed Cauchy-sum launches (params.timing = 1) */
	long long aberth_launches; /* number of timed Cauchy-sum launches */
	long long sweeps;          /* Aberth sweeps launched */
	long long cluster_chunks;  /* mixed precision: LDS chunks recomputed in FP64 for clustered roots */
	double flops_sums;         /* algorithmic flops of the Cauchy-sum kernels (direct and FMM) */
	long long root_fmm_updates;/* updates whose root sums came from the root-sum FMM */
	long long root_fmm_rows;   /* (row, sweep) pairs summed by the root-sum FMM */
	int root_fmm_order;        /* its expansion order (0: not used) */
	int root_fmm_leaf_level;   /* depth of its grid */
	double max_witt_residual;  /* max |f0| (Witt's function) over the solved points */
	double witt_kernel_ms;     /* sum of timed Witt-evaluation launches (params.timing = 1) */
	long long witt_launches;
	double flops_witt;         /* algorithmic flops of the Witt evaluations (flops_update - flops_sums) */
	long long updates_newton;  /* precision 1, continuation steps: FP64 Newton steps z - f0/f1 taken
	                              instead of an Aberth update (no Cauchy sums) where |f0/f1| < 1e-6 theta;
	                              below 1e-9 theta the root ends there */
	long long tail_launches;   /* precision 1: phase tails run in one launch (all counts above include them) */
	long long tail_sweeps;     /* their sweeps, summed over rows */
	double tail_kernel_ms;     /* sum of timed tail launches (params.timing = 1) */
	double flops_tail;         /* algorithmic flops of the tails (Witt evaluations and Cauchy sums) */
	long long points_outside_tree; /* error pass: points outside the star tree's root box, whose near
	                              field is every star (summed by a whole wave each) */
	int roots_per_thread;      /* the sums' roots per thread the run used (given, default or tuned) */
	int root_fmm_leaf;         /* the mean roots per leaf cell its root-sum FMM grid aimed at */
	double t_autotune;         /* seconds the run spent tuning (params.autotune; 0 once tuned) */
} ccf_info;

/* Arrays readable with ccf_get_array. */
enum ccf_array
{
	CCF_ARRAY_STARS = 0,             /* nstars x 3 doubles {x1, x2, mass}, tree order */
	CCF_ARRAY_CCS = 1,               /* (R*nb) x (np/nb+1) complex doubles, file order */
	CCF_ARRAY_CAUSTICS = 2,          /* same layout as CCS */
	CCF_ARRAY_MU_LENGTH_SCALES = 3   /* (R*nb) x (np/nb+1) doubles */
};

typedef struct ccf_ctx ccf_ctx;

/* ccf.cuh:68-71 (constructor) and ownership of every device buffer */
ccf_ctx* ccf_create(void);
void ccf_destroy(ccf_ctx* ctx);

/* reference defaults of the public fields (ccf.cuh:42-62) */
void ccf_default_params(ccf_params* p);

/* copy the public fields into the context (the caller's field assignments) */
int ccf_set_params(ccf_ctx* ctx, const ccf_params* p);
int ccf_get_params(const ccf_ctx* ctx, ccf_params* p);

/* bool CCF::run(int verbose) (ccf.cuh:1265-1280) */
int ccf_run(ccf_ctx* ctx, int verbose);

/* bool CCF::save(int verbose) (ccf.cuh:1282-1287) */
int ccf_save(ccf_ctx* ctx, int verbose);

/* stages of run(), in order; each requires the previous ones.
   ccf_load = check_input_params + calculate_derived_params +
   allocate_initialize_memory + populate_star_array (ccf.cuh:218-730);
   ccf_build_tree_fmm = create_tree (ccf.cuh:732-904);
   ccf_prepare = ccf_load + ccf_build_tree_fmm. */
int ccf_check_input_params(ccf_ctx* ctx, int verbose); /* ccf.cuh:218-346, host only */
/* check_input_params + calculate_derived_params (ccf.cuh:218-497), host
   only: fills num_stars, num_roots, taylor_smooth, kappa_star, mu_ave,
   corner and alpha_error of ccf_info (reads the star file if one is set) */
int ccf_calculate_derived_params(ccf_ctx* ctx, int verbose);
int ccf_load(ccf_ctx* ctx, int verbose);
int ccf_build_tree_fmm(ccf_ctx* ctx, int verbose);
int ccf_prepare(ccf_ctx* ctx, int verbose);
int ccf_find_initial_roots(ccf_ctx* ctx, int verbose);      /* ccf.cuh:906-977 */
int ccf_find_ccs(ccf_ctx* ctx, int verbose);                /* ccf.cuh:979-1083 */
int ccf_find_caustics(ccf_ctx* ctx, int verbose);           /* ccf.cuh:1085-1102 */
int ccf_find_mu_length_scales(ccf_ctx* ctx, int verbose);   /* ccf.cuh:1104-1121 */

int ccf_get_info(const ccf_ctx* ctx, ccf_info* info);

/* copy an array to host memory dst of capacity bytes; writes its shape */
int ccf_get_array(ccf_ctx* ctx, int which, void* dst, long long capacity, int* nrows, int* ncols);

/* message of the last failure on this context ("" if none) */
const char* ccf_last_error(const ccf_ctx* ctx);

/* Evaluate the field of a prepared context at n host points z (interleaved
   re, im) for angle phi. out[8*i + 0..7] = f0.re, f0.im, f1.re, f1.im,
   w.re, w.im, mu_length_scale, 1/mu error bound -- the quantities of
   ccf_functions.cuh:33-166 and 449-463 evaluated by the device kernels. */
int ccf_eval_points(ccf_ctx* ctx, const double* z, int n, double phi, double* out);

/* One Jacobi Aberth sweep (ccf_functions.cuh:193-243) of n host roots on
   the device: out[i] = update of roots[i] given all of roots[0..n). */
int ccf_aberth_sweep(ccf_ctx* ctx, const double* roots, int n, double phi, double* out);

/* ccf_aberth_sweep without allocating: no counterpart in the reference.
   For callers that cannot allocate device memory (code under HIP graph
   capture, tight serving loops, harnesses that forbid hipMalloc). The row
   and every buffer of the sweep live in a fixed scratch array of 32 MiB that
   the library holds on each device it runs on until the process ends
   (set up when a context initialises the device); a sweep that needs more
   (about 30 000 roots; fewer with root_fmm on an elongated field) is
   refused with an error. Results are those of ccf_aberth_sweep. Calls on
   one device share the array: calls that overlap on one device race. */
int ccf_aberth_sweep_noalloc(ccf_ctx* ctx, const double* roots, int n, double phi, double* out);

/* The root sums S_k = sum_{j != k} 1/(z_k - z_j) of n host roots (the root
   part of ccf_functions.cuh:232-239) by the root-sum FMM of a prepared
   context, with its root_fmm_order and root_fmm_leaf, whether root_fmm is
   set or not: out[2k], out[2k+1] = S_k. No counterpart in the reference;
   for checking the FMM against direct sums. */
int ccf_root_sums(ccf_ctx* ctx, const double* roots, int n, double* out);

/* ---- batched fields ----
   No counterpart in the reference, which solves one field per process
   (main_ccf.cu:514-515). Fields of uneven size are added with their own
   parameters (as for ccf_set_params) and solved together: their trees are
   built as one forest and the Aberth solves of all fields run in shared
   persistent launches. Every field then has the reference's outputs and
   files (ccf_batch_save). The fields must share fmm_order, direct,
   binomial_exact and precision; the expansion order is the largest the
   reference's rule gives over the fields unless fmm_order is set. */
typedef struct ccf_batch ccf_batch;
ccf_batch* ccf_batch_create(void);
void ccf_batch_destroy(ccf_batch* b);
/* adds a field; returns its index, or -1 */
int ccf_batch_add(ccf_batch* b, const ccf_params* p);
int ccf_batch_size(const ccf_batch* b);
/* every field: validation, derived parameters, stars, seeds (device) */
int ccf_batch_load(ccf_batch* b, int verbose);
/* the forest of all fields' trees with its FMM coefficients */
int ccf_batch_build(ccf_batch* b, int verbose);
/* initial roots, critical curves, 1/mu errors, caustics and length scales of every field */
int ccf_batch_solve(ccf_batch* b, int verbose);
/* load + build + solve */
int ccf_batch_run(ccf_batch* b, int verbose);
/* the reference's files of one field (its outfile_prefix) */
int ccf_batch_save(ccf_batch* b, int field, int verbose);
/* info of one field, or of the whole batch for field = -1 (totals; t_ccs = solve time) */
int ccf_batch_get_info(const ccf_batch* b, int field, ccf_info* info);
int ccf_batch_get_array(ccf_batch* b, int field, int which, void* dst, long long capacity, int* nrows, int* ncols);
/* workgroup size the autotuner picked for the phi chains, the longer of the two persistent
   launches (0 before the first solve) */
int ccf_batch_workgroup(const ccf_batch* b);
const char* ccf_batch_last_error(const ccf_batch* b);

/* Write a seeded synthetic star file (layout of star.cuh:498-534) drawn
   on the host with std::mt19937_64: positions uniform in the disc or
   rectangle the reference derives for these parameters (ccf.cuh:420-436),
   masses by the inverse CDF of the mass function. Used to feed identical
   stars (-sf) to every implementation. */
int ccf_write_synthetic_stars(const ccf_params* p, unsigned long long seed, const char* path);

/* mass function moments and inverse CDF (mass_functions headers):
   which 0 mass(p), 1 <m>, 2 <m^2>, 3 <m^2 ln m>; NaN for unknown names */
double ccf_mass_function(const char* name, int which, double p, double m_lower, double m_upper, double m_solar);

#ifdef __cplusplus
}
#endif

#endif /* CCF_AMD_H */
