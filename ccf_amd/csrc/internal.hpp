// Internal state of a ccf_amd context: host-side derived parameters and the
// device buffers of one star field. Not part of the public interface.
#pragma once

#include "../../include/ccf_amd.h"
#include "cplx.hpp"
#include "imf.hpp"

#include <hip/hip_runtime.h>

#include <map>
#include <sstream>
#include <unordered_map>
#include <string>
#include <vector>

namespace ccf
{

constexpr int MAX_NUM_STARS_DIRECT = 32; // tree_node.cuh:14
constexpr int MAX_EXPANSION_ORDER = 31;  // tree_node.cuh:20
constexpr int MAX_NEIGHBORS = 8;         // tree_node.cuh:23
constexpr int MAX_SAME_LEVEL = 27;       // tree_node.cuh:24
// the pinned staging buffer of a context (ccf_ctx::host_stage), in doubles:
// the binomial table of the largest expansion order, which also holds a
// tree level's counts or the run's folded statistics
constexpr size_t STAGE_DOUBLES = (size_t)(2 * MAX_EXPANSION_ORDER + 1) * (2 * MAX_EXPANSION_ORDER + 2) / 2;
constexpr int MAX_DIFF_LEVEL = 5;        // tree_node.cuh:25
constexpr int MAX_NEAR_RANGES = MAX_NEIGHBORS + 1;
constexpr int MAX_TAYLOR_SMOOTH = 101;   // ccf.cuh:79
constexpr int MAX_ROOT_FMM_ORDER = 24;
// highest degree of the continuation predictor
constexpr int MAX_PREDICTOR = 4;
constexpr int MAX_ROOT_GRID_LEVELS = 10;
// groups of branches whose mixed-form solves run on their own streams side
// by side (one group's latency-bound sweeps beside the others' device-wide
// ones), at most one per branch: CONT_GROUPS in the initial solve (chains of
// short FMM launches), CONT_STEP_GROUPS in the continuation (more were
// slower: each group's chain slows with the contention). The group streams
// and the error pass's stay within the 4 hardware queues a process gets by
// default; streams that share a queue serialise.
constexpr int CONT_GROUPS = 3;
constexpr int CONT_STEP_GROUPS = 2;

// Quadtree over the stars, structure of arrays indexed by node id. Nodes of
// a level are contiguous ([level_off[l], level_off[l] + level_cnt[l])) and
// the four children of a node are consecutive ids starting at child[n].
struct DeviceTree
{
	int nnodes = 0;
	int capacity = 0;
	int levels = 0; // deepest level index (the reference's tree_levels)
	int order = 0;  // expansion order p (0: no far field)
	std::vector<int> level_off, level_cnt;

	double2* center = nullptr;
	double* hl = nullptr;
	int* parent = nullptr;
	int* child = nullptr;  // first child id, -1 for a leaf
	int* sstart = nullptr; // first star (tree order)
	int* scount = nullptr;
	int* nbr = nullptr;    // [capacity][MAX_NEIGHBORS]
	int* nnbr = nullptr;
	int* sl = nullptr;     // same-level interaction list [capacity][27]
	int* nsl = nullptr;
	int* dl = nullptr;     // coarser-level interaction list [capacity][5]
	int* ndl = nullptr;
	int* field = nullptr;  // field of each node (a forest of several fields' trees)
	int* split = nullptr;  // per-level scratch: split flags / ranks
	int* rank = nullptr;
	int2* nf = nullptr;    // near-field star ranges: self then neighbours [capacity][9]
	int* nnf = nullptr;
	double4* nbox = nullptr;  // {centre x, centre y, half-length, first child} per node
	int* nf_off = nullptr;    // leaves: first entry of their near-field copy in nfs, -1 otherwise
	int* nf_cnt = nullptr;    // leaves: number of near-field stars
	double4* nfs = nullptr;   // near-field stars of every leaf, contiguous: {x, y, m, 0}
	double2* mp = nullptr;  // multipole coefficients [nnodes][order + 1]
	double2* loc = nullptr; // local coefficients [nnodes][order + 1]
	float2* loc32 = nullptr;
	// expansions of the same tree with every mass set to 1: their L2P gives
	// the far part of sum 1/(z - s_i), the pole sum of the Aberth correction
	double2* mpu = nullptr;
	double2* locu = nullptr;
	float2* locu32 = nullptr;
	double* binom = nullptr;
	int* lookup = nullptr;  // [4^lookup_level]: node reached at that level by each cell of a uniform grid
	int lookup_level = 0;
	double lookup_x0 = 0, lookup_inv = 0;
};

// Kernel-side view of the lens model and tree (passed by value).
struct LensView
{
	double kappa, gamma, theta, theta2, kappastar;
	int rectangular, approx, taylor_smooth;
	double corner_re, corner_im, corner_abs, corner_abs2, corner_arg;
	int order;
	int nstars;
	int nts;                   // number of precomputed smooth-matter series terms
	const double2* ts_a;       // (1 - e^{2 i k arg c}) / k for k = ts-1, ts-3, ..., 2
	const double2* ts_b;       // (1 - e^{2 i k arg c})
	double2 clog[4];           // log c, log conj(c), log(-c), log(-conj(c)) of the corner c (rectangular fields)
	const double2* spos;       // star positions, tree order
	const double* smass;
	const double2* center;
	const double* hl;
	const int* child;
	const int2* nf;
	const int* nnf;
	const double4* nbox;
	const int* nf_off;
	const int* nf_cnt;
	const double4* nfs;
	const double2* loc;
	const float2* loc32;
	const float2* locu32;
	int root;                // root node of the field (several fields share one forest)
	const int* lookup;       // point-location table (null: descend from the root)
	int lookup_n;            // its cells per side
	double lookup_x0, lookup_inv; // lower-left corner of the root box, cells per unit length
};

struct Timer
{
	hipEvent_t a = nullptr, b = nullptr;
};

// Uniform quadtree over the field for the root-sum FMM. Level l is a grid of
// (nx << l) x (ny << l) square cells of side s0 / 2^l whose lower-left corner
// is (x0, y0); its cells are numbered row-major from off[l].
struct RootGrid
{
	int nx = 0, ny = 0;  // top-level cells
	int leaf = 0;        // leaf level
	int order = 0;       // expansion order (0: no FMM)
	int ncells = 0;      // cells of all levels
	int nleaf = 0;       // leaf cells
	int bin_row = 0;     // a row's leaf counts fit in one workgroup's LDS (one-launch binning)
	int rlev = 0;        // region level: its cells group the leaf cells in blocks of at most 8 x 8
	double x0 = 0, y0 = 0, s0 = 0;
	int off[MAX_ROOT_GRID_LEVELS + 1] = {};
};

// Work buffers of the Aberth sweeps over up to `rows` rows of R roots.
struct SweepBufs
{
	int rows = 0, R = 0, tile = 0, max_items = 0;
	int* act = nullptr;      // [2][rows][R] active root indices (ping-pong)
	int* cnt = nullptr;      // [2][rows]
	int* fl = nullptr;       // [rows][R] roots that need the Cauchy sums this sweep
	int* fcnt = nullptr;     // [2][rows]
	double2* f0 = nullptr;   // [rows][R] Witt function at the roots of fl
	double2* f1 = nullptr;   // [rows][R] its z-derivative
	double2* ex = nullptr;   // [rows][R] terms of the Aberth sum formed per root
	int* tile_off = nullptr; // [rows + 1] first tile of each row
	int* plan = nullptr;     // [4] tiles, slices per tile; rows summed by the root-sum FMM in a phase's first and second sweep
	int* ticket = nullptr;   // [1] workgroups of the sweep's Witt kernel done (the last forms the plan and resets it)
	double2* part = nullptr; // [max_items][tile] partial Cauchy sums
	// root-sum FMM (grid.order > 0 only)
	RootGrid grid;
	int* slot = nullptr;     // [rows][R] position of root a in fl this sweep, -1 otherwise
	int* rmode = nullptr;    // [rows] 1: the row's root sums of this sweep come from the FMM
	double2* rsum = nullptr; // [rows][R] FMM root sums, indexed like fl
	int* ccnt = nullptr;     // [rows][ncells] roots in each cell's subtree
	int* cstart = nullptr;   // [rows][nleaf + 1] first sorted position of each leaf cell
	int* ccur = nullptr;     // [rows][nleaf] counts, then scatter cursors
	int* cell = nullptr;     // [rows][R] leaf cell (sort key) of each root, -1 outside the grid
	int* bidx = nullptr;     // [rows][R] root indices in leaf cell order, each cell's in any order
	int* sidx = nullptr;     // [rows][R] root indices sorted by leaf cell, then by index
	double2* spos = nullptr; // [rows][R] their positions
	int* ocnt = nullptr;     // [rows] roots outside the grid
	int* oidx = nullptr;     // [rows][R] their indices
	double2* mp = nullptr;   // [rows][ncells][order + 1] multipole coefficients
	double2* lcq = nullptr;  // [rows][cells of levels 0 .. rlev][order + 1] their local expansions (k_rcoarse)
};

} // namespace ccf

struct ccf_ctx
{
	ccf_params P_in{}; // the caller's fields; every run starts from them
	ccf_params P{};    // the fields as run() leaves them (derived values assigned)
	ccf_info info{};
	std::string last_error;

	// derived parameters (ccf.cuh:96-126)
	ccf::Imf imf;
	double mean_mass = 0, mean_mass2 = 0, mean_mass2_ln_mass = 0;
	double kappa_star_actual = 0, m_lower_actual = 0, m_upper_actual = 0;
	double mean_mass_actual = 0, mean_mass2_actual = 0, mean_mass2_ln_mass_actual = 0;
	double mu_ave = 0, alpha_error = 0, root_half_length = 0, max_error = 0;
	ccf::cd corner;
	int taylor_smooth = 1, num_roots = 0, num_stars = 0, expansion_order = 0;
	int rectangular = 0;
	double theta = 1, kappa_star = 0;
	int random_seed = 0;
	bool from_file = false;

	// stage completion
	bool loaded = false, prepared = false, have_roots = false, have_ccs = false, have_caustics = false, have_mu = false;

	// host copies
	std::vector<double> stars_in;   // {x1, x2, m} in generation/file order
	std::vector<double> seeds;      // initial roots (re, im) interleaved
	std::vector<int> order;         // processing order of the roots (Morton order of the seeds)

	// device state
	hipStream_t stream = nullptr;
	int num_cus = 0;
	ccf::DeviceTree tree;
	double2* spos = nullptr;      // tree order
	double* smass = nullptr;
	double2* spos_tmp = nullptr;
	double* smass_tmp = nullptr;
	float2* spos32 = nullptr;
	double2* ts_a = nullptr;
	double2* ts_b = nullptr;
	int nts = 0;
	double2* ccs_init = nullptr;  // [(np + nb) rows][R], row-major by phi
	double2* ccs = nullptr;       // transposed, file order
	double2* caustics = nullptr;
	double* mu_scales = nullptr;
	// Aberth work buffers
	int max_rows = 0;
	ccf::SweepBufs sw;
	double* ones = nullptr;       // unit masses for the pole-sum expansions
	unsigned long long* counters = nullptr; // device statistics, see solve.hip
	std::vector<ccf::Timer> timers;
	std::vector<ccf::Timer> witt_timers;
	std::vector<ccf::Timer> tail_timers;
	// the timers' events, created once and reused once the timers are read
	std::vector<hipEvent_t> events;
	size_t events_used = 0;
	hipEvent_t next_event()
	{
		if (events_used == events.size())
		{
			hipEvent_t e = nullptr;
			if (hipEventCreate(&e) != hipSuccess) return nullptr;
			events.push_back(e);
		}
		return events[events_used++];
	}
	int* host_counts = nullptr;   // pinned: active roots per row, the two FMM row counts of a read-back, then per continuation group two slots of two for the mixed continuation
	int tree_hint_stars = -1, tree_hint_fields = -1, tree_hint_nodes = 0; // the last tree build's node count
	double* host_stage = nullptr; // pinned (STAGE_DOUBLES), until ccf_destroy: a tree build's level counts and binomial table, a run's folded statistics
	hipStream_t err_stream = nullptr; // the error pass of find_ccs, beside the continuation (default priority)
	hipEvent_t err_event = nullptr;   // a row set's roots are final (no timing)
	hipEvent_t fmm_rows_event[2 * ccf::CONT_GROUPS] = {}; // mixed continuation: after the copy of a step's FMM row count, per group (no timing)
	hipEvent_t read_event[ccf::CONT_GROUPS] = {}; // after the copy of a sweep's active counts, per group (no timing)
	hipStream_t group_streams[ccf::CONT_GROUPS - 1] = {}; // mixed form: the streams of the groups of branches after the first (default priority)
	hipEvent_t group_event = nullptr;   // orders the group streams around the continuation (no timing)
	ccf::SweepBufs swg[ccf::CONT_GROUPS - 1]; // mixed form: the sweep buffers of the groups after the first
	int fmm_rows_sweep[2] = {-1, -1}; // root-sum FMM rows of the last phase's first and second sweep (-1: unknown)
	int phase_sweeps = 0;         // sweeps the last phase ran (its cap when no read-back saw it end)
	// params.autotune: the choices the context's first solve timed fastest
	bool tuned = false;
	int tuned_rpt = 0, tuned_root_order = 0, tuned_root_leaf = 0;
	int sweep_verbose = 0;        // verbose level of the running solve stage (3: a read-back and a line after every sweep)
	bool caustics_done = false;   // caustics / length scales formed by the error pass of find_ccs
	bool mu_done = false;
	double2* phase = nullptr;     // (cos phi, sin phi) of every phi row
	bool batch_member = false;    // a field of a ccf_batch: no sweep buffers, seeds on the device
	bool results_borrowed = false; // ccs, caustics and mu_scales are slices of the batch's result arena (not freed here)
	bool own_stream = true;       // ccf_destroy destroys the stream
	double2* dseeds = nullptr;    // [R] initial roots (batch members)
	int* dorder = nullptr;        // [R] processing order of the roots

	ccf::LensView lens() const;
	void fail(const std::string& msg);
	bool hip_ok(hipError_t e, const char* what, const char* file, int line);
	void free_device();
	// Device memory of this context. Buffers that free_device or a resize
	// gives up stay in a pool and serve later requests of about the same
	// size (at most a quarter plus 64 KiB larger), so that rebuilding a
	// field or a batch of the same shape allocates nothing; pool_release
	// (also on an allocation that fails) and ccf_destroy return them.
	// Reuse is ordered by the context's stream, which every user of these
	// buffers works on.
	hipError_t dalloc(void** p, size_t bytes);
	template <typename T>
	hipError_t dmalloc(T** p, size_t bytes)
	{
		void* v = nullptr;
		const hipError_t e = dalloc(&v, bytes);
		*p = (T*)v;
		return e;
	}
	void dfree(void* p);
	void pool_release();
	std::multimap<size_t, void*> pool;        // free buffers by capacity
	std::unordered_map<void*, size_t> owned;  // every buffer from dalloc, by capacity
};

namespace ccf
{
struct BatchDevice;
}

// several fields solved together (batch.hip)
struct ccf_batch
{
	std::vector<ccf_ctx*> fields;
	ccf_ctx* master = nullptr;      // stream, forest, concatenated stars, errors
	ccf::BatchDevice* dev = nullptr;
	hipGraph_t graph = nullptr;
	hipGraphExec_t graph_exec = nullptr;
	int block = 0, block1 = 0;      // autotuned workgroup sizes of the branch-centre and phi-chain phases
	int grid0 = 0, grid1 = 0;       // persistent grids of the two phases
	std::vector<float> tuned_ms;
	void* results = nullptr;        // the fields' ccs, caustics and length scales, one allocation kept across builds
	size_t results_bytes = 0;
	ccf_info info{};
	bool loaded = false, built = false, solved = false;
};

#define CCF_HIP(ctx, call) ((ctx)->hip_ok((call), #call, __FILE__, __LINE__))
#define CCF_LAUNCH(ctx, name) ((ctx)->hip_ok(hipGetLastError(), name, __FILE__, __LINE__))

namespace ccf
{
// tree.hip
bool build_tree(ccf_ctx* ctx, int verbose);
// trees of several fields at once (ctx->spos holds their stars back to back)
struct ForestSpec
{
	std::vector<int> soff, nstars;
	std::vector<double> root_hl, theta, alpha_error;
	std::vector<int> levels; // out: depth of each field's tree
	std::vector<size_t> lookup_off; // out (forests): offset of each field's point-location table in tree.lookup
	std::vector<int> lookup_level;  // out (forests): its level (0: none)
	int order = 0;           // out: the expansion order used
};
bool build_forest(ccf_ctx* ctx, ForestSpec& spec, int verbose);
// solve.hip
bool alloc_sweep(ccf_ctx* ctx, SweepBufs& b, int rows, int R);
void free_sweep(ccf_ctx* ctx, SweepBufs& b);
bool alloc_solver(ccf_ctx* ctx);
bool find_initial_roots_dev(ccf_ctx* ctx, int verbose);
bool find_ccs_dev(ccf_ctx* ctx, int verbose);
bool find_caustics_dev(ccf_ctx* ctx, bool caustics, bool mu);
bool generate_stars_dev(ccf_ctx* ctx);
bool eval_points_dev(ccf_ctx* ctx, const double* z, int n, double phi, double* out);
bool aberth_sweep_dev(ccf_ctx* ctx, const double* roots, int n, double phi, double* out);
bool root_sums_dev(ccf_ctx* ctx, const double* roots, int n, double* out);
// the no-allocation sweep: its scratch array's address on the current
// device (once per device), and the sweep over slices of that array
bool scratch_setup(ccf_ctx* ctx);
bool aberth_sweep_noalloc_dev(ccf_ctx* ctx, const double* roots, int n, double phi, double* out);
// batch.hip
bool batch_build(ccf_batch* b, int verbose);
bool batch_solve(ccf_batch* b, int verbose);
void free_batch_device(ccf_batch* b);
} // namespace ccf
