/* gt_hip.h: C ABI of libgt_hip.so, the MI355X (gfx950) gravity backend for
 * galactic-tides' grouped Barnes-Hut path (device = hip).
 *
 * Plain C types only; every entry point returns an int status (GT_HIP_OK or
 * an error code) and never lets a C++ exception cross the boundary.  The
 * message for the last failure on a context is gt_hip_last_error(ctx).
 * A context is not re-entrant: use one per host thread.  All calls are
 * synchronous: when a call returns, its outputs are in the caller's memory.
 *
 * Structs keep the byte layouts of the reference's kernel structs
 * (include/galactic_tides/kernels/gravity_kernels.h:36-38,51-60,87-105):
 * gt_float4 16 B, gt_params 32 B, gt_cell 48 B, gt_group 32 B.
 */
#ifndef GT_HIP_H
#define GT_HIP_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define GT_HIP_ABI_VERSION 1

enum {
  GT_HIP_OK = 0,
  GT_HIP_ERR_ARGUMENT = 1,    /* invalid argument (reference: std::invalid_argument) */
  GT_HIP_ERR_DEVICE = 2,      /* no usable GPU, or a HIP call failed (std::runtime_error) */
  GT_HIP_ERR_MEMORY = 3,      /* device allocation failed */
  GT_HIP_ERR_LIMIT = 4,       /* too many particles or list entries for 32-bit indices */
  GT_HIP_ERR_UNSUPPORTED = 5  /* input outside what this build handles (message says why) */
};

typedef struct gt_hip_context gt_hip_context;

/* kern::float4 (gravity_kernels.h:36-38): x, y, z and mass (or ax, ay, az, pot). */
typedef struct {
  float x, y, z, w;
} gt_float4;

/* kern::GTParams (gravity_kernels.h:51-60). */
typedef struct {
  uint32_t n, n_groups;
  float G, eps2;
  uint32_t quadrupole, pad0, pad1, pad2;
} gt_params;

/* kern::GTCell (gravity_kernels.h:87-91). */
typedef struct {
  gt_float4 com_mass, q0, q1;
} gt_cell;

/* kern::GTGroup (gravity_kernels.h:96-105). */
typedef struct {
  uint32_t begin, end, cell_begin, cell_end, body_begin, body_end, pad0, pad1;
} gt_group;

/* Solver settings: GravityParams (include/galactic_tides/octree.h:13-18) plus
 * the GroupTreeGravity constructor arguments (gravity_backends.h:62-63). */
typedef struct {
  double G;          /* gravitational constant */
  double theta;      /* opening angle */
  double eps;        /* Plummer softening length */
  int quadrupole;    /* 1: cells carry quadrupole moments */
  int group_size;    /* targets per group, 1..64 */
  int leaf_size;     /* bodies per leaf (>= 1) */
  int pack;          /* merge small siblings into groups; < 0: device default (on) */
  int exact;         /* 1: IEEE 1/sqrt, no contraction, reference summation order:
                        forces bit-identical to evaluate_group_lists_cpu */
  int eval_waves;    /* tuning: waves per SIMD the list-evaluation kernel is built
                        for (2 or 3; 0 = library default, 3) */
  int walk_lanes;    /* tuning: 0 or 3 (default): eight lanes per group test a node's
                        children together (one round of loads per child set, FP32 node
                        images with an exact FP64 fallback); 4: the same with four lanes
                        per group, two children each; 2: one walk per lane, one node
                        record per step; 1: the same, fetching both possible next nodes'
                        records ahead; 16, 32 or 64: that many groups share one
                        wave-coherent walk.  Every choice gives the same lists */
  int split_len;     /* tuning: lists longer than this many entries are evaluated in
                        parts, by several waves with the longest-first orders and by
                        one wave in tree order (fast evaluation only; parts added in
                        part order, so the results are the same either way); 0 =
                        library default (8192), < 0 = never split */
  int eval_order;    /* tuning (fast evaluation): the order in which each XCD takes its
                        share of the groups: 1 = tree order, 2 = longest first, 3 / 4 =
                        longest first by cost classes of half / quarter octaves (tree
                        order within a class); 0 = library default (4; 1 with the
                        persistent grid, whose queue balances the load).  Results do
                        not depend on it. */
  int hilo;          /* 1: fast evaluation with every position as an FP32 hi + lo pair
                        taken relative to its group's centre (bounding-box centre in
                        FP64), instead of FP32 relative to the root centre as the
                        reference does; same lists, more precise differences of nearby
                        positions.  Ignored by exact evaluation. */
  int predict;       /* list slots of the one-pass walk: 0 = uniform (the previous
                        evaluation's longest lists + 25%), 1 = per group from the learned
                        predictor (an MLP trained on the first evaluations' lists; uniform
                        until it is trained), 2 = no slots: a count pass and a fill pass
                        every evaluation (the plain two-pass solver, for comparison) */
  int overlap;       /* tuning (fast evaluation): the walk runs in this many ranges of
                        groups and each range is evaluated on a second stream while the
                        next is walked; 1 = no overlap, 0 = library default (1), at most
                        8.  Pays only where the walk has many more groups than the GPU
                        has SIMDs: a range's walk lasts as long as its longest group's */
  int persistent;    /* tuning (fast evaluation, whole evaluations): 0 (default) or 1 = a
                        persistent grid (the GPU's CUs x waves per SIMD) whose waves
                        take the work items of their XCD's share one at a time from a
                        counter, in the eval_order; -1 = one wave per work item, dealt
                        out by the hardware dispatcher.  Results do not depend on it. */
} gt_group_config;

/* GroupLists statistics (group_lists.h:37-55) and tree size of the last call. */
typedef struct {
  uint64_t n, n_nodes, depth, n_groups, n_cell_list, n_body_list;
  uint64_t cell_interactions, body_interactions, max_list, max_group;
} gt_lists_info;

/* GroupTreeGravity::Timings (gravity_backends.h:54-58), in milliseconds,
 * measured with HIP events on the context's stream. */
typedef struct {
  double tree, lists, evaluate, scatter, total;
} gt_timings;

/* ---- context (replaces metal::Context, include/galactic_tides/metal_backend.h:23-54) ---- */

/* metal::compiled_in() (metal_backend.h:23): 1 when built with HIP kernels. */
int gt_hip_compiled_in(void);
/* metal::Context::Context() (metal_backend.h:27): fails with GT_HIP_ERR_DEVICE
 * when no GPU is visible. */
int gt_hip_create(gt_hip_context **ctx, int device);
void gt_hip_destroy(gt_hip_context *ctx);
/* Message for the last failing call on ctx ("" if none). */
const char *gt_hip_last_error(const gt_hip_context *ctx);
/* metal::Context::device_name() (metal_backend.h:32). */
int gt_hip_device_name(const gt_hip_context *ctx, char *buf, size_t len);
/* metal::Context::last_gpu_seconds() (metal_backend.h:47): kernel time of the last call. */
double gt_hip_last_gpu_seconds(const gt_hip_context *ctx);

/* ---- the reference's two GPU entry points, host arrays in and out ---- */

/* metal::Context::direct (metal_backend.h:35-36, kernel gt_direct,
 * src/metal/gravity_kernels.metal:18-43): out[i] = (ax, ay, az, pot) of body i
 * against all others; p->n bodies. */
int gt_hip_direct(gt_hip_context *ctx, const gt_float4 *bodies, const gt_params *p,
                  gt_float4 *out);
/* metal::Context::tree (metal_backend.h:41-44, kernel gt_tree,
 * src/metal/gravity_kernels.metal:53-112): evaluates grouped interaction lists,
 * out[i] in tree order.  p->n must equal the number of bodies, p->n_groups the
 * number of groups; groups may hold at most 64 targets.  exact = 1 gives
 * results bit-identical to evaluate_group_lists_cpu (gravity_backends.cpp:70-84). */
int gt_hip_tree(gt_hip_context *ctx, const gt_float4 *bodies, uint64_t n_cells,
                const gt_cell *cells, uint64_t n_cell_list, const uint32_t *cell_list,
                uint64_t n_body_list, const uint32_t *body_list, const gt_group *groups,
                const gt_params *p, int exact, gt_float4 *out);

/* gt_hip_tree for callers that cannot allocate device memory (no counterpart
 * in the reference; for code that must not call hipMalloc: tight serving
 * loops, harnesses that forbid it).  Same arguments and results as gt_hip_tree,
 * but every array (bodies, cells, the two lists, groups and out) is in device
 * memory, written by the caller before the call (its writes complete); p is a
 * host struct.  The call allocates nothing: the evaluation's scratch (status
 * and count words, the input check's flags: 1 KiB) is a fixed device array,
 * one per GPU, that the library holds until the process ends.  The inputs
 * are checked on the device as gt_hip_tree checks them on the host (same
 * messages, GT_HIP_ERR_ARGUMENT), before anything is evaluated.  Calls on one
 * GPU share that scratch: calls that overlap on one GPU (from several host
 * threads or contexts) race.  Synchronous; limits: at most 2^32 - 1 cells
 * and list entries (32-bit indices), groups of at most 64 targets. */
int gt_hip_tree_noalloc(gt_hip_context *ctx, const gt_float4 *bodies, uint64_t n_cells,
                        const gt_cell *cells, uint64_t n_cell_list, const uint32_t *cell_list,
                        uint64_t n_body_list, const uint32_t *body_list, const gt_group *groups,
                        const gt_params *p, int exact, gt_float4 *out);

/* ---- the whole force evaluation on the GPU ---- */

/* GroupTreeGravity::compute (src/gravity_backends.cpp:109-137) with device = hip:
 * octree build, group selection, grouped walk, list evaluation and the scatter
 * back to particle order all run on the GPU.  pos/acc are n x 3 doubles. */
int gt_hip_group_compute(gt_hip_context *ctx, const gt_group_config *cfg, uint64_t n,
                         const double *pos, const double *mass, double *acc, double *pot);
/* TreeGravity::compute / compute_gravity_tree (include/galactic_tides/integrator.h:20-34,
 * src/gravity.cpp:7-26, walk Octree::accel_pot src/octree.cpp:275-344), which the
 * reference runs on the CPU only ("gravity = tree"): the per-particle FP64 tree
 * walk, here on the GPU over the GPU-built tree, with the reference's arithmetic
 * op for op, so acc/pot equal the reference's bit for bit.  Uses cfg->G, theta,
 * eps, quadrupole and leaf_size.  stats (may be NULL): [0] cell interactions,
 * [1] particle interactions (WalkStats, include/galactic_tides/octree.h:22-30). */
int gt_hip_tree_gravity(gt_hip_context *ctx, const gt_group_config *cfg, uint64_t n,
                        const double *pos, const double *mass, double *acc, double *pot,
                        uint64_t *stats);
/* Octree::accel_pot (src/octree.cpp:275-344) at np points with nobody left out
 * (self = Octree::kNoSelf) on the tree of the n particles, built and walked
 * as gt_hip_tree_gravity does: acc np x 3 and pot np, bit for bit the
 * reference's.  The reference's ICs generator samples the realised potential
 * this way (src/hernquist_ics.cpp:206-330).  stats as for gt_hip_tree_gravity. */
int gt_hip_tree_gravity_at(gt_hip_context *ctx, const gt_group_config *cfg, uint64_t n,
                           const double *pos, const double *mass, uint64_t np,
                           const double *points, double *acc, double *pot, uint64_t *stats);
/* Octree::build (src/octree.cpp:169-273) of the n particles with leaf_size, then
 * Octree::for_each_within (include/galactic_tides/octree.h:73-75,103-127) for nq
 * queries at once, on the GPU.  Query q has centre centres[3q .. 3q+2] and radius
 * radii[q] (n_radii = nq) or radii[0] (n_radii = 1).  Its neighbours are every
 * particle within the radius, inclusive, in the order the reference's callback
 * receives them (its LIFO tree walk; tree order within a leaf), as CSR arrays:
 * entries offsets[q] .. offsets[q + 1] - 1 (nq + 1 offsets, from 0) of index
 * (original particle index) and dist2 (squared distance, the reference's FP64
 * expression, bit for bit).  mass may be NULL (the tree's shape does not depend
 * on it).  offsets and *total (may be NULL) are always written.  index and dist2
 * (each may be NULL; both NULL: the call only counts) are written when the total
 * fits max_out; otherwise the call returns GT_HIP_ERR_LIMIT and can be made again
 * with room for *total entries.  At most 2^32 - 1 neighbours in all per call.
 * gt_hip_last_timings: tree = the build, lists = query ordering, the walk
 * (counting; when the lists are wanted, also writing each query's first
 * neighbours into a slot of its own) and the scan, evaluate = the lists moved
 * to their places and the queries whose lists outgrew their slots walked
 * again (the filling walk when slots did not fit 4 GiB), total = their sum
 * (GPU time; host copies and waits between the phases excluded). */
int gt_hip_neighbours(gt_hip_context *ctx, int leaf_size, uint64_t n, const double *pos,
                      const double *mass, uint64_t nq, const double *centres,
                      const double *radii, uint64_t n_radii, uint64_t max_out,
                      uint64_t *offsets, uint64_t *index, double *dist2, uint64_t *total);
/* Independent systems in one evaluation (no counterpart in the reference,
 * which builds one GroupTreeGravity per system): system s is particles
 * offsets[s] .. offsets[s + 1] (n_systems + 1 offsets from 0 to n, non-decreasing;
 * empty systems allowed).  Each system gets exactly the tree, lists and forces
 * gt_hip_group_compute gives it alone (its own root cube and FP32 origin; systems
 * do not attract each other), but all of them are built, walked and evaluated by
 * the same launches.  n = offsets[n_systems]. */
int gt_hip_group_compute_batch(gt_hip_context *ctx, const gt_group_config *cfg,
                               uint64_t n_systems, const uint64_t *offsets, const double *pos,
                               const double *mass, double *acc, double *pot);
/* Autotuner (no counterpart in the reference): times force evaluations of the
 * given particles on the GPU (HIP events, best of 2 after a warm-up) while it
 * varies one tuning setting at a time (eval_waves {2, 3}, walk_lanes {2, 4,
 * 16, 32, 64}, eval_order {1, 2, 3, 4}, split_len {2048, 8192, never},
 * persistent {-1, 1}, and group_size
 * {32, 64} when search_group_size is 1), keeping each change that is at least 1%
 * faster,
 * and writes the fastest settings into *best (the other fields copied from
 * *cfg) and their time per evaluation in ms into *best_ms (may be NULL).
 * The tuning settings do not change the lists (results are identical or, for
 * split lists, within the fast evaluation's tolerance); group_size is the
 * reference's own parameter and changes the lists, so it is only searched on
 * request.  Exact evaluation (cfg->exact) has nothing to tune but walk_lanes. */
int gt_hip_autotune(gt_hip_context *ctx, const gt_group_config *cfg, uint64_t n,
                    const double *pos, const double *mass, int search_group_size,
                    gt_group_config *best, double *best_ms);
/* The same autotuner (no counterpart in the reference) timing batched
 * evaluations of independent systems (offsets as for gt_hip_group_compute_batch):
 * the settings that evaluate the whole batch fastest. */
int gt_hip_autotune_batch(gt_hip_context *ctx, const gt_group_config *cfg, uint64_t n_systems,
                          const uint64_t *offsets, const double *pos, const double *mass,
                          int search_group_size, gt_group_config *best, double *best_ms);
/* Walk and evaluate on a tree supplied by the caller (the reference's node
 * layout: geom n_nodes x 16 doubles = center xyz, half, com xyz, mass,
 * quad xx yy zz xy xz yz, delta, 0; links n_nodes x 4 int32 = first_child,
 * n_children, begin, end; tree_pos/tree_mass in tree order).  Lists use the
 * caller's node numbering, so they compare raw with build_group_lists
 * (src/group_lists.cpp:18-207) on the same tree.  out (n x 4, tree order) may be NULL. */
int gt_hip_walk_tree(gt_hip_context *ctx, const gt_group_config *cfg, uint64_t n,
                     uint64_t n_nodes, const double *geom, const int32_t *links,
                     const double *tree_pos, const double *tree_mass, gt_float4 *out);
/* Statistics and phase timings of the last gt_hip_group_compute / walk / step. */
int gt_hip_last_info(const gt_hip_context *ctx, gt_lists_info *info);
int gt_hip_last_timings(const gt_hip_context *ctx, gt_timings *t);
/* Octree accessors (Octree::nodes()/order(), octree.h:77-79) for the last GPU
 * build: geom/links as above (n_nodes rows, level order), order n entries. */
int gt_hip_get_tree(gt_hip_context *ctx, double *geom, int32_t *links, uint64_t *order);
/* GroupLists accessor (GroupTreeGravity::lists(), gravity_backends.h:68) for
 * the last walk: n_groups groups, the two lists, bodies (n), cells (n_nodes)
 * and the origin (3 doubles).  Any pointer may be NULL. */
int gt_hip_get_lists(gt_hip_context *ctx, gt_group *groups, uint32_t *cell_list,
                     uint32_t *body_list, gt_float4 *bodies, gt_cell *cells, double *origin);

/* ---- device-resident leapfrog (LeapfrogKDK, include/galactic_tides/integrator.h:39-47) ---- */

/* Upload particles (n x 3 pos/vel, n mass) to device memory kept by ctx. */
int gt_hip_state_set(gt_hip_context *ctx, uint64_t n, const double *pos, const double *vel,
                     const double *mass);
/* The resident particles as a batch of independent systems (offsets as for
 * gt_hip_group_compute_batch, ending at the n of gt_hip_state_set): later
 * prime/steps evaluate each system on its own.  gt_hip_state_set resets to
 * one system.  Forces must be primed again. */
int gt_hip_state_set_systems(gt_hip_context *ctx, uint64_t n_systems, const uint64_t *offsets);
/* Download, in the order the particles were set; any pointer may be NULL. */
int gt_hip_state_get(gt_hip_context *ctx, double *pos, double *vel, double *acc, double *pot);

/* Diagnostics (include/galactic_tides/integrator.h:49-58): total energy is
 * kinetic + potential, the virial ratio -2 kinetic / potential. */
typedef struct {
  double time;
  double kinetic;              /* sum 0.5 m |v|^2 */
  double potential;            /* sum 0.5 m phi */
  double momentum[3];          /* sum m v */
  double angular_momentum[3];  /* sum m (x cross v) */
  double centre_of_mass[3];    /* sum m x / sum m (0 when the mass is 0) */
} gt_diagnostics;
/* compute_diagnostics (src/integrator.cpp:18-33) on the resident state, one
 * entry per system (max_out must be at least the number of systems; the run
 * loop of apps/galactic_tides_run.cpp:65-112 logs these every few steps).
 * Uses the potentials, so forces must be current (after gt_hip_prime).  Sums
 * run in a fixed parallel order: deterministic, equal to the reference's
 * sequential sums to FP64 rounding. */
int gt_hip_state_diagnostics(gt_hip_context *ctx, double time, uint64_t max_out,
                             gt_diagnostics *out);
/* LeapfrogKDK::prime (integrator.h:42): forces for the resident state.  The
 * library then keeps the resident particles in the order of this force
 * evaluation's tree (so each step's gathers into tree order are nearly
 * sequential); results, downloads and gt_hip_get_tree's order are in the
 * caller's order all the same. */
int gt_hip_prime(gt_hip_context *ctx, const gt_group_config *cfg);
/* nsteps x LeapfrogKDK::step (src/integrator.cpp:5-16) on the resident state:
 * FP64 kick and drift kernels around each GPU force evaluation.  An
 * evaluation is enqueued with no host round trip: sizes it depends on (nodes,
 * groups, list slots) are bounded from the previous evaluation, and one that
 * outgrows its bounds is run again with larger ones.  use_graph = 1 captures
 * the step's four phases as HIP graphs (tree, lists, evaluation, kick;
 * recaptured only when the bounds or settings change) and runs the steps, up
 * to 64 at a time, as one graph that chains them, with one host
 * synchronisation at the end; a step that outgrows the bounds stops the rest,
 * and the call finishes them without graphs.  Results are bit-identical to
 * use_graph = 0.  Timings are per step (means over up to 64 steps). */
int gt_hip_steps(gt_hip_context *ctx, const gt_group_config *cfg, int nsteps, double dt,
                 int use_graph);
/* Captures and instantiates the graphs gt_hip_steps(use_graph = 1) with the
 * same arguments would launch first, without running anything: a caller that
 * times the steps calls it beforehand, so the timed call only launches.  No
 * counterpart in the reference (its steps have no graphs).  Does nothing while
 * the sizes or the predictor are still being learned. */
int gt_hip_prepare_steps(gt_hip_context *ctx, const gt_group_config *cfg, int nsteps, double dt);

/* Counters of how evaluations ran on ctx since it was created (no counterpart
 * in the reference; for monitoring and tests of the bound handling). */
typedef struct {
  uint64_t evaluations;    /* force evaluations completed */
  uint64_t reruns;         /* evaluations run again with larger bounds */
  uint64_t list_rewrites;  /* lists rewritten at exact offsets (slots too small) */
  uint64_t graph_captures; /* step graphs captured */
  uint64_t graph_steps;    /* leapfrog steps run from captured graphs */
  uint64_t graph_stops;    /* graph steps that outgrew their bounds and were finished
                              without graphs */
  uint64_t interactions;   /* cell + body interactions of all completed evaluations */
  uint64_t last_exact_tests; /* opening tests of the last evaluation's walk that the FP32
                                filter left to the FP64 test */
  uint64_t last_slot_entries; /* list entries the last evaluation's slots spanned: the
                                 predicted slots' total, groups x the uniform slots, or
                                 (two passes, no slots) the lists' own entries */
  uint64_t sort_recounts;  /* radix-sort digit totals added up again by a workgroup whose
                              bounded wait for another's published total gave up (results
                              are right either way; as of the last evaluation's end) */
} gt_pipeline_stats;
int gt_hip_pipeline_stats(const gt_hip_context *ctx, gt_pipeline_stats *stats);
/* The learned list-length predictor (no counterpart in the reference): with
 * cfg->predict = 1, the first `collect` evaluations (default 50) gather samples
 * (the features of a spread of groups and their true list lengths), then an FP32
 * MLP 8 -> 64 -> 64 -> 2 is trained with Adam on the GPU and sizes every group's
 * slots from then on.  gt_hip_predictor_reset forgets the model and sets
 * `collect` (0: default). */
typedef struct {
  uint64_t samples;          /* training samples gathered */
  uint64_t trained;          /* 1 once the model is in use */
  uint64_t train_iterations; /* Adam iterations run */
  double train_loss;         /* final mean squared error, standardised targets */
  double margin;             /* log-space headroom added to predictions */
  uint64_t predicted;        /* evaluations with predicted slots */
  uint64_t overflows;        /* ... of which some group outgrew its slots (exact rewrite) */
  double mean_abs_log_error; /* last predicted evaluation: mean |log(1+n) - prediction| */
  uint64_t slot_entries;     /* last predicted evaluation: list entries reserved */
  uint64_t list_entries;     /* ... and used */
} gt_predictor_stats;
int gt_hip_predictor_reset(gt_hip_context *ctx, int collect);
/* The predictor's kernels on caller data (no counterpart in the reference; for
 * checking them against a CPU restatement).  Parameter vector (params, m, v:
 * 4866 floats): W1 (64 x 8, row-major), b1 (64), W2 (64 x 64), b2 (64),
 * W3 (2 x 64), b3 (2).  gt_hip_predictor_forward: y (n x 2, standardised
 * outputs) of the network at normalised inputs x (n x 8), as the slot
 * inference computes it.  gt_hip_predictor_train_step: one training iteration
 * as the library runs it: `blocks` workgroups of 64 samples each take
 * samples (iter * blocks * 64 + b * 64 + t) * 2654435761 mod n of
 * samples (n x 10: 8 raw features, log(1 + cells), log(1 + bodies)),
 * standardised with stats (means of the 10 columns, then their deviations);
 * each writes the gradient of its share of the loss (the mean over the
 * minibatch and both outputs of the squared error) to grad[b] (blocks x 4866)
 * and the sum of its squared errors to loss[b]; the gradients are added in
 * workgroup order and one Adam step (beta 0.9 / 0.999, eps 1e-8, bias
 * correction for step iter + 1) updates params, m and v in place. */
int gt_hip_predictor_forward(gt_hip_context *ctx, uint64_t n, const float *x, const float *params,
                             float *y);
int gt_hip_predictor_train_step(gt_hip_context *ctx, uint64_t n, const float *samples,
                                const float *stats, float *params, float *m, float *v,
                                uint32_t blocks, uint32_t iter, float lr, float *grad,
                                float *loss);
int gt_hip_predictor_stats(const gt_hip_context *ctx, gt_predictor_stats *stats);

/* Test hook (no counterpart in the reference): while shrink > 0, evaluations
 * on ctx get list slots `shrink` entries shorter than the previous
 * evaluation's longest lists, so that the slot-overflow rewrite and the
 * graph-stop recovery run on every evaluation; results must not change.
 * 0 (the default) restores the normal bounds. */
int gt_hip_debug_shrink_slots(gt_hip_context *ctx, uint32_t shrink);
/* Test hook (no counterpart in the reference): with polls > 0, every wait of
 * the radix sorts' digit scans for another workgroup's published total gives
 * up after `polls` polls and the total is added up from the histogram instead
 * (counted in gt_pipeline_stats.sort_recounts); results must not change.
 * 0 (the default) restores the normal bounded wait. */
int gt_hip_debug_sort_waits(gt_hip_context *ctx, uint32_t polls);

/* ---- measurement helpers ---- */

/* The GPU's sustained FP32 VALU rate, in lane operations per second (no
 * counterpart in the reference; the practical ceiling of the VALU-bound
 * kernels, which are built without packed FP32): every lane runs 16
 * independent FMA chains at eight waves per SIMD, best of 3 timed launches;
 * rsq = 1 makes every fourth operation a v_rsq_f32. */
int gt_hip_debug_valu_rate(gt_hip_context *ctx, int rsq, double *lane_ops_per_s);

/* FP64 direct summation for a sample of targets (apps/galactic_tides_bench.cpp:80-95):
 * acc is ns x 3: the reference solution force errors are measured against. */
int gt_hip_direct_f64_sample(gt_hip_context *ctx, uint64_t n, const double *pos,
                             const double *mass, double G, double eps, uint64_t ns,
                             const uint64_t *sample, double *acc);

#ifdef __cplusplus
}
#endif
#endif
