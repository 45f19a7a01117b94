// C-ABI entry points of libgt_hip.so (declared in include/gt_hip.h): the
// context, the reference's two GPU calls (direct, tree), the whole force
// evaluation on the GPU, and the device-resident leapfrog.
#include <algorithm>
#include <cstring>
#include <cmath>
#include <new>
#include <random>
#include <string>
#include <utility>
#include <vector>

#include "gt_build.h"
#include "gt_internal.h"
#include "gt_predict.h"

using namespace gth;

// What the previous evaluation on a context needed; bounds for the next one
// are derived from it.  Zero / -1 / false mean "not known".
struct History {
  uint32_t n = 0;  // particles of the evaluation described
  uint32_t systems = 1;
  int words = 1;
  bool full_paths = false;  // a one-word build needed all 21 levels of the path sort
  uint32_t nodes = 0;
  int depth = -1;
  uint32_t groups = 0;
  bool lists_known = false;
  uint64_t most_c = 0, most_b = 0;
  bool items_known = false;
  uint64_t items = 0, split_parts = 0;  // evaluation work items, parts of split groups
  uint64_t all_c = 0, all_b = 0;        // list entries (lists_known)
  bool pred_known = false;
  uint64_t pred_c = 0, pred_b = 0;      // predicted slot totals
  bool levels_known = false;
  uint32_t level_nodes[frame::kLevels] = {};  // nodes per tree level
};

// The learned list-length predictor of a context.
struct Predictor {
  uint32_t collect = 50;     // evaluations to gather samples from
  uint32_t collected = 0;    // evaluations sampled so far
  uint32_t samples = 0;      // samples in the buffer (at most kSampleCap)
  uint32_t next = 0;         // where the next sample goes
  bool trained = false;
  uint64_t iterations = 0;
  double loss = 0.0;
  float margin = 0.0f;        // log-space headroom (its copy in device memory is used)
  float margin_floor = 0.0f;  // the margin right after training
  uint64_t predicted = 0, overflows = 0;
  double err = 0.0;
  uint64_t slot_entries = 0, list_entries = 0;
};
constexpr uint32_t kSampleCap = 1u << 18;      // samples kept
constexpr uint32_t kSamplesPerEval = 4096;     // taken from each evaluation
constexpr uint32_t kTrainIterations = 1500;
constexpr uint32_t kTrainBlocks = 32;          // x 64 samples per minibatch

// Settings a captured step graph was made for.
struct GraphKey {
  Plan plan;
  double G = 0, theta = 0, eps = 0, dt = 0;
  int quadrupole = 0, group_size = 0, leaf_size = 0, pack = 0, exact = 0, eval_waves = 0,
      walk_lanes = 0, hilo = 0, persistent = 0;
  uint64_t generation = 0;
};

bool same_plan(const Plan &a, const Plan &b) {
  return a.n == b.n && a.systems == b.systems && a.words == b.words &&
         a.path_levels == b.path_levels && a.node_cap == b.node_cap &&
         a.depth_cap == b.depth_cap && a.group_cap == b.group_cap && a.cap_c == b.cap_c &&
         a.cap_b == b.cap_b && a.split_len == b.split_len && a.item_cap == b.item_cap &&
         a.split_cap == b.split_cap && a.longest_first == b.longest_first &&
         a.cost_buckets == b.cost_buckets && a.hilo == b.hilo && a.predicted == b.predicted &&
         a.slot_total_c == b.slot_total_c && a.slot_total_b == b.slot_total_b &&
         a.chunks == b.chunks && a.levels_sized == b.levels_sized &&
         std::equal(a.level_blocks, a.level_blocks + frame::kLevels, b.level_blocks);
}

bool same_key(const GraphKey &a, const GraphKey &b) {
  return same_plan(a.plan, b.plan) && a.G == b.G && a.theta == b.theta && a.eps == b.eps &&
         a.dt == b.dt && a.quadrupole == b.quadrupole && a.group_size == b.group_size &&
         a.leaf_size == b.leaf_size && a.pack == b.pack && a.exact == b.exact &&
         a.eval_waves == b.eval_waves && a.walk_lanes == b.walk_lanes && a.hilo == b.hilo &&
         a.persistent == b.persistent &&
         a.generation == b.generation;
}

constexpr int kTimedSteps = 64;  // graph steps whose phases are timed with events
// Leapfrog steps between reorders of the resident state into tree order.
constexpr int kReorderSteps = 64;
constexpr int kPhases = 4;       // tree, lists, evaluation, scatter + kick

constexpr int kMaxChunks = 8;  // overlap: walk chunks whose evaluation runs alongside

struct gt_hip_context {
  int device = 0;
  hipStream_t stream = nullptr;
  hipStream_t stream2 = nullptr;  // evaluation of walked chunks, alongside the walk
  hipEvent_t ev[6] = {};
  hipEvent_t walk_done[kMaxChunks] = {}, eval_done = nullptr;
  hipEvent_t eval_t0[kMaxChunks] = {}, eval_t1[kMaxChunks] = {};
  uint32_t overlapped = 0;  // chunks of the last evaluation run alongside its walk (0: none)
  std::string err;
  double gpu_seconds = 0.0;
  gt_lists_info info{};
  gt_timings tm{};
  gt_pipeline_stats pstats{};
  hipDeviceProp_t prop{};

  TreeBuilder tree;
  // device frame of the current evaluation, and its host copy
  DBuf<unsigned long long> frame, raw_frame, misc;
  unsigned long long fw[frame::kWords] = {};
  History hist;
  Plan plan;  // bounds of the last evaluation
  uint32_t debug_shrink = 0;  // test hook: list slots this much too short
  // caller-supplied tree (gt_hip_walk_tree)
  DBuf<double> up_geom, up_tpos, up_tmass;
  DBuf<int4> up_lnk;
  // lists of the last walk
  DBuf<WalkRec> recs;
  DBuf<WalkRec32> images;
  DBuf<uint2> span;
  DBuf<Cell> cells;
  DBuf<float4> bodies, out;
  DBuf<uint32_t> parent, skip, depth, flags, offs, starts, ccount, bcount, coff, boff, cell_list,
      body_list, group_root, flag_node, group_node;
  // the node depths the last lists were built with (depth above, or the
  // tree builder's own)
  const uint32_t *lists_depth = nullptr;
  // systems of the current evaluation (a batch when more than one): count and
  // particle offsets in device memory
  uint32_t systems = 1;
  const unsigned long long *sys_off = nullptr;
  DBuf<unsigned long long> batch_off, state_off;
  uint32_t state_systems = 1;
  DBuf<gt_group> groups;
  DBuf<unsigned long long> scan_tmp;
  // long lists in parts
  DBuf<uint32_t> parts, split_count, item_off, split_off;
  DBuf<uint2> items;
  DBuf<float4> partial;
  // hi/lo precision mode
  DBuf<float4> bodies_lo, cells_lo, group_hilo;
  // learned list lengths
  Predictor pred;
  DBuf<float> pred_params, pred_m, pred_v, pred_stats, pred_samples, pred_grad, pred_loss, feat,
      pred_out;
  DBuf<uint32_t> slot_c, slot_b, slot_off_c, slot_off_b;
  DBuf<uint64_t> item_keys, item_keys_alt;
  DBuf<uint32_t> item_order, item_order_alt;
  DBuf<unsigned long long> item_sort_tmp;
  uint32_t lists_n = 0, lists_nn = 0;
  const double *lists_geom = nullptr;  // geometry rows the last lists refer to
  // particle state (resident leapfrog) and inputs of single evaluations
  DBuf<double> pos, vel, mass, acc, pot;
  // The resident state is kept in the order of the tree of its last priming
  // (gt_hip_prime), so the gathers into tree order and the force scatter back
  // are nearly sequential; orig[k] is the caller's index of resident particle k.
  DBuf<uint32_t> orig, orig_tmp;
  DBuf<double> state_tmp;
  bool permuted = false;       // orig is not the identity
  int steps_in_order = 0;      // steps since the state was last put in tree order
  bool tree_on_state = false;  // the last tree was built from the resident state
  DBuf<double> diag_partial, diag_total;  // diagnostics sums
  DBuf<double> in_pos, in_mass, in_acc, in_pot;
  uint64_t n_state = 0;
  bool state_forces = false;
  // captured step graphs: the four phases, and k steps of them chained into
  // one graph (run_exec[k], made when a run of k steps first needs it)
  hipGraph_t step_graph[kPhases] = {};
  hipGraphExec_t run_exec[kReorderSteps + 1] = {};
  // overlapped steps: the tree and scatter phases as graphs of their own
  hipGraphExec_t step_exec[kPhases] = {};
  // overlapped steps: the lists phase and the evaluation as graphs per walk
  // range, launched on the two streams with the events between them
  uint32_t ov_chunks = 0;
  hipGraphExec_t ov_prep = nullptr, ov_finish = nullptr;
  hipGraphExec_t ov_walk[kMaxChunks] = {}, ov_items[kMaxChunks] = {}, ov_eval[kMaxChunks] = {},
                 ov_comb[kMaxChunks] = {};
  bool step_valid = false;
  GraphKey step_key;
  std::vector<hipEvent_t> step_ev;  // kTimedSteps x (kPhases + 1)
  // raw-array entry points and sampling
  DBuf<float4> raw_bodies, raw_out;
  DBuf<float> probe_out;
  DBuf<Cell> raw_cells;
  DBuf<uint32_t> raw_cl, raw_bl;
  DBuf<gt_group> raw_groups;
  DBuf<uint64_t> sample_idx;
  DBuf<double> sample_acc;
  // neighbour queries
  DBuf<double> nb_centres, nb_radii, nb_dist2;
  DBuf<uint32_t> nb_counts, nb_offs, nb_perm, nb_perm_alt;
  DBuf<uint64_t> nb_keys, nb_keys_alt;
  DBuf<unsigned long long> nb_index, nb_tmp, nb_sort_tmp;
  DBuf<uint4> nb_nodes;  // 64-byte query-walk records, 4 uint4 each
  DBuf<unsigned long long> nb_stage_index;  // per-query slots of the one-walk fill
  DBuf<double> nb_stage_dist2;
  DBuf<uint32_t> nb_over;  // queries whose lists outgrew their slots
  uint32_t nb_stage = 64;  // slot entries per query (grows when many overflow)
  DBuf<uint32_t> eval_queue;  // per-XCD item counters of the persistent evaluation and its
                              // count of waves out (kQueueStride apart; zero between launches)
  void *noalloc_scratch = nullptr;  // this device's gt_hip_tree_noalloc scratch
  DBuf<double> pt_pos;  // gt_hip_tree_gravity_at's points
  // predictor kernels on caller data (test hooks)
  DBuf<float> tk_in, tk_stats, tk_params, tk_m, tk_v, tk_grad, tk_loss, tk_out;
};

namespace {

thread_local std::string g_create_error;

template <class F>
int guarded(gt_hip_context *ctx, F &&f) {
  try {
    f();
    if (ctx) ctx->err.clear();
    return GT_HIP_OK;
  } catch (const Error &e) {
    if (ctx) ctx->err = e.what();
    return e.code;
  } catch (const std::bad_alloc &) {
    if (ctx) ctx->err = "host allocation failed";
    return GT_HIP_ERR_MEMORY;
  } catch (const std::exception &e) {
    if (ctx) ctx->err = e.what();
    return GT_HIP_ERR_DEVICE;
  } catch (...) {
    if (ctx) ctx->err = "unknown error";
    return GT_HIP_ERR_DEVICE;
  }
}

void require(bool ok, const char *msg) {
  if (!ok) fail(GT_HIP_ERR_ARGUMENT, msg);
}

void check_config(const gt_group_config *cfg) {
  require(cfg != nullptr, "config is NULL");
  require(cfg->group_size >= 1, "group_size must be at least 1");
  if (cfg->group_size > kMaxGroup)
    fail(GT_HIP_ERR_ARGUMENT,
         "group_size can be at most " + std::to_string(kMaxGroup) + " on hip (one wave64)");
  require(cfg->leaf_size >= 1, "leaf_size must be at least 1");
  require(cfg->theta > 0.0, "theta must be positive");
}

template <class T>
void upload(T *dst, const T *src, size_t n, hipStream_t s) {
  if (n) GTH_CHECK(hipMemcpyAsync(dst, src, n * sizeof(T), hipMemcpyHostToDevice, s));
}

template <class T>
void download(T *dst, const T *src, size_t n, hipStream_t s) {
  if (n) GTH_CHECK(hipMemcpyAsync(dst, src, n * sizeof(T), hipMemcpyDeviceToHost, s));
}

float elapsed_ms(hipEvent_t a, hipEvent_t b) {
  float ms = 0.0f;
  GTH_CHECK(hipEventElapsedTime(&ms, a, b));
  return ms;
}

// ---- bounds --------------------------------------------------------------------

constexpr uint32_t kMaxNodes = kRecIndexMask - 1;  // node ids must fit the walk records

uint32_t grown(uint64_t x, uint64_t extra, uint64_t cap) {
  return static_cast<uint32_t>(std::min<uint64_t>(x + x / 4 + extra, cap));
}

// List slot for groups whose longest list so far is `most`: a quarter more, a
// multiple of 4 (slots start 16-byte aligned for the walk's vector stores).
uint32_t slot_for(uint64_t most) {
  return static_cast<uint32_t>(std::min<uint64_t>((most + most / 4 + 32 + 3) & ~3ull, 1u << 30));
}

// Path levels for the sort of a one-word build whose previous tree was depth
// levels deep: at least one level more, and as many as the sort's 8-bit passes
// for that hold (3 bits a level).
int path_levels_for(int depth) {
  const int need = std::min(depth + 1, 21);
  const int passes = (3 * need + 7) / 8;
  return std::min(21, 8 * passes / 3);
}

// Whether the bounds p still hold what the history says with some headroom
// (then they are kept, so buffers and captured graphs stay as they are).
bool plan_holds(const Plan &p, const History &h, uint32_t n, uint32_t systems) {
  if (p.n != n || h.n != n || p.systems != systems || h.systems != systems || p.words < h.words || h.nodes == 0 || h.depth < 0 ||
      h.groups == 0 || !h.lists_known)
    return false;
  if (static_cast<uint64_t>(h.nodes) + h.nodes / 16 > p.node_cap) return false;
  if (h.depth + 1 > p.depth_cap && p.depth_cap < 60) return false;
  if (p.words == 1 && p.path_levels < 21 && (h.full_paths || h.depth + 1 > p.path_levels))
    return false;
  // moment grids: sized once the levels are known, then kept while no level
  // needs more than two passes of its grid
  if (h.levels_known && !p.levels_sized) return false;
  if (p.levels_sized)
    for (int d = 0; d < frame::kLevels; ++d)
      if (h.level_nodes[d] > 2ull * 256 * p.level_blocks[d]) return false;
  if (p.group_cap < n && static_cast<uint64_t>(h.groups) + h.groups / 16 > p.group_cap) return false;
  if (h.most_c + h.most_c / 16 + 8 > p.cap_c || h.most_b + h.most_b / 16 + 8 > p.cap_b)
    return false;
  if (p.longest_first && (!h.items_known || h.items + h.items / 16 > p.item_cap ||
                      h.split_parts + h.split_parts / 16 > p.split_cap))
    return false;
  if (p.predicted && (!h.pred_known || h.pred_c + h.pred_c / 16 > p.slot_total_c ||
                      h.pred_b + h.pred_b / 16 > p.slot_total_b))
    return false;
  return true;
}

// Whether the history describes every size of an evaluation of n particles.
bool sizes_known(const History &h, uint32_t n, uint32_t systems) {
  return h.n == n && h.systems == systems && h.nodes && h.depth >= 0 && h.groups && h.lists_known;
}

// Bounds for an evaluation of n particles: the last ones while they hold,
// else new ones from the history with headroom (kept in c->plan).
bool uses_hilo(const gt_group_config *cfg) { return cfg->hilo != 0 && !cfg->exact; }

constexpr uint32_t kDefaultSplitLen = 8192;
// One wave per work item: quarter-octave cost classes, longest first (at 2M
// the evaluation takes 2.45 ms against 2.79 in tree order and 2.46 fully
// sorted, with less L2 miss traffic than the full sort, since a class keeps
// tree order).  Persistent grid: tree order, with no work items to make or
// sort (2.42 ms either way at 2M, since the waves' queue balances the load,
// and the lists phase is ~70 us shorter).
constexpr int kDefaultEvalOrder = 4;
constexpr int kDefaultPersistentOrder = 1;

// The persistent evaluation grid: the default (persistent 0 or 1); -1 turns it
// off.  It is faster on every workload measured (2M encounter 4.65 -> 4.56
// ms per step, 4M galaxy 9.32 -> 8.97, 128 galaxies 6.49 -> 6.20).
bool persistent_for(const gt_group_config *cfg) { return cfg->persistent >= 0 && !cfg->exact; }

int eval_order_for(const gt_group_config *cfg) {
  if (cfg->eval_order != 0) return cfg->eval_order;
  return persistent_for(cfg) ? kDefaultPersistentOrder : kDefaultEvalOrder;
}

bool longest_first_for(const gt_group_config *cfg) {
  return !cfg->exact && eval_order_for(cfg) >= 2;
}

// eval_order 3, 4: cost classes of half and quarter octaves
int cost_buckets_for(const gt_group_config *cfg) {
  const int o = eval_order_for(cfg);
  return o == 3 ? 2 : o == 4 ? 4 : 0;
}

// List entries per evaluation part for these settings (0: no work items, the
// groups are evaluated as they are; UINT32_MAX: items, but no list is split).
uint32_t split_len_for(const gt_group_config *cfg) {
  if (cfg->exact) return 0;
  const uint32_t never = longest_first_for(cfg) ? 0xffffffffu : 0u;
  if (cfg->split_len < 0) return never;
  return cfg->split_len == 0 ? kDefaultSplitLen : static_cast<uint32_t>(cfg->split_len);
}

Plan bounds_for(gt_hip_context *c, uint32_t n, uint32_t split_len, bool longest_first,
                int buckets, bool predicted) {
  const History &h = c->hist;
  if (c->plan.split_len == split_len && c->plan.longest_first == longest_first &&
      c->plan.cost_buckets == buckets && c->plan.predicted == predicted &&
      plan_holds(c->plan, h, n, c->systems))
    return c->plan;
  Plan p;
  p.n = n;
  p.systems = c->systems;
  const bool same = h.n == n && h.systems == c->systems;
  p.words = same ? h.words : 1;
  p.node_cap = same && h.nodes ? grown(h.nodes, 1024, kMaxNodes)
                               : static_cast<uint32_t>(std::min<uint64_t>(n + 1024ull, kMaxNodes));
  p.depth_cap = same && h.depth >= 0 ? std::min(60, h.depth + 3) : 60;
  p.path_levels = same && h.depth >= 0 && !h.full_paths ? path_levels_for(h.depth) : 21;
  if (same && h.levels_known) {
    p.levels_sized = true;
    for (int d = 0; d < frame::kLevels; ++d) {
      // a quarter of headroom, rounded up to a power of two
      const uint64_t need = (h.level_nodes[d] + h.level_nodes[d] / 4 + 255) / 256;
      uint32_t b = 1;
      while (b < need && b < kMomentBlocks) b <<= 1;
      p.level_blocks[d] = static_cast<uint16_t>(b);
    }
  }
  p.group_cap = same && h.groups ? grown(h.groups, 64, n) : n;
  if (h.lists_known) {
    p.cap_c = slot_for(h.most_c);
    p.cap_b = slot_for(h.most_b);
  }
  if (static_cast<uint64_t>(p.group_cap) * p.cap_c > 0xffffffffull ||
      static_cast<uint64_t>(p.group_cap) * p.cap_b > 0xffffffffull)
    p.cap_c = p.cap_b = 0;
  p.predicted = predicted;
  if (predicted) {  // the list buffers: the predicted totals (or the true ones) with headroom
    const bool known = same && h.pred_known;
    const uint64_t lim = 0xffffffffull;  // list offsets are 32 bit
    p.slot_total_c = std::min(known ? h.pred_c + h.pred_c / 4 + 4096 : 2 * h.all_c + 4096, lim);
    p.slot_total_b = std::min(known ? h.pred_b + h.pred_b / 4 + 4096 : 2 * h.all_b + 4096, lim);
  }
  p.split_len = split_len;
  p.longest_first = longest_first && split_len != 0;
  p.cost_buckets = p.longest_first ? buckets : 0;
  if (p.longest_first) {  // work items
    const bool known = same && h.items_known;
    p.split_cap = known ? grown(h.split_parts, 64, 1u << 26) : p.group_cap / 8 + 64;
    const uint64_t extra = known && h.items > h.groups ? h.items - h.groups : p.split_cap;
    p.item_cap = static_cast<uint32_t>(
        std::min<uint64_t>(p.group_cap + extra + extra / 4 + 64, 0xffffffffull));
  }
  c->plan = p;
  return p;
}

// One range by default: the 2M encounter's 87k group walks fill the SIMDs
// about 1.3 times, so each range's walk lasts about as long as the whole walk
// and ranges add up (6.1 ms per step at 1 range, 6.9 at 2, 8.7 at 4).
constexpr uint32_t kDefaultChunks = 1;

// Walk chunks whose evaluation runs alongside the walk (1: none).  Exact
// evaluation and coherent walks take all groups at once.
uint32_t chunks_for(const gt_group_config *cfg, const Plan &p) {
  if (cfg->exact) return 1;
  const int l = cfg->walk_lanes;
  if ((l == 16 || l == 32 || l == 64) && p.systems <= 1 && !p.predicted) return 1;
  const int k = cfg->overlap == 0 ? static_cast<int>(kDefaultChunks) : cfg->overlap;
  return static_cast<uint32_t>(std::min(std::max(k, 1), kMaxChunks));
}

Plan plan_for(gt_hip_context *c, uint32_t n, const gt_group_config *cfg) {
  const bool predicted = cfg->predict == 1 && c->pred.trained && c->hist.lists_known;
  Plan p = bounds_for(c, n, split_len_for(cfg), longest_first_for(cfg), cost_buckets_for(cfg),
                      predicted);
  p.hilo = uses_hilo(cfg);
  p.chunks = chunks_for(cfg, p);
  if (cfg->predict == 2) p.cap_c = p.cap_b = 0;  // two passes: count, then fill
  const History &h = c->hist;
  if (c->debug_shrink && h.lists_known) {  // test hook: slots too short
    const uint64_t k = c->debug_shrink;
    p.cap_c = h.most_c > k ? static_cast<uint32_t>(h.most_c - k) & ~3u : 0u;
    p.cap_b = h.most_b > k ? static_cast<uint32_t>(h.most_b - k) & ~3u : 0u;
  }
  return p;
}

// What an evaluation that ended with status st says about the sizes needed.
void learn(gt_hip_context *c, const Plan &p, unsigned long long st) {
  const unsigned long long *fw = c->fw;
  History &h = c->hist;
  if (h.n != p.n || h.systems != p.systems) h = History{};
  h.n = p.n;
  h.systems = p.systems;
  h.words = p.words;
  if (st & kAbortPaths) {
    if (p.words == 1 && p.path_levels < 21) {
      h.full_paths = true;
    } else {
      if (p.words >= 3) fail(GT_HIP_ERR_DEVICE, "octree build: leaf deeper than 63 levels");
      h.words = p.words + 1;
    }
  }
  if (fw[frame::kNodes] > kMaxNodes)
    fail(GT_HIP_ERR_LIMIT, "too many tree nodes for the packed walk records (2^26 - 1)");
  h.nodes = static_cast<uint32_t>(fw[frame::kNodes]);
  if (!(st & kAbortNodes)) {
    h.depth = static_cast<int>(fw[frame::kDepth]);
    h.groups = static_cast<uint32_t>(fw[frame::kGroups]);
  }
  if (!(st & (kAbortNodes | kAbortPaths | kAbortDepth))) {
    h.levels_known = true;
    for (int d = 0; d < frame::kLevels; ++d)
      h.level_nodes[d] = static_cast<uint32_t>(fw[frame::kLevelNodes + d]);
  }
  // List counts hold whenever the tree and groups were complete (slots or
  // work items too small do not change them); item counts need whole lists.
  if ((st & ~static_cast<unsigned long long>(kAbortLists | kAbortItems)) == 0) {
    h.lists_known = true;
    h.most_c = fw[frame::kStats + 4];
    h.most_b = fw[frame::kStats + 5];
    h.all_c = fw[frame::kStats + 6];
    h.all_b = fw[frame::kStats + 7];
  }
  if (p.predicted && !(st & (kAbortNodes | kAbortGroups | kAbortPaths | kAbortDepth))) {
    h.pred_known = true;
    h.pred_c = fw[frame::kPredCells];
    h.pred_b = fw[frame::kPredBodies];
  }
  if ((st & ~static_cast<unsigned long long>(kAbortItems)) == 0 && p.longest_first) {
    // (with the walk in chunks the frame holds the last chunk's counts: after
    // an overflow the bounds double instead)
    const bool over = st & kAbortItems;
    h.items_known = true;
    h.items = over ? std::max<uint64_t>(2 * p.item_cap, fw[frame::kItems]) : fw[frame::kItems];
    h.split_parts = over ? std::max<uint64_t>(2 * p.split_cap, fw[frame::kSplitParts])
                         : fw[frame::kSplitParts];
  }
}

void read_frame(gt_hip_context *c) {
  download(c->fw, c->frame.get(), frame::kWords, c->stream);
  GTH_CHECK(hipStreamSynchronize(c->stream));
}

// Walk errors: a malformed tree, or node half sizes that are not the root's
// halved (code: the error for a caller-supplied tree, or an internal one).
void check_walk(const gt_hip_context *c, unsigned long long st, int code) {
  if (c->fw[frame::kWalkError] & 2ull)
    fail(code, "tree: a node's half size is not root half / 2^depth");
  if (st & kAbortWalk) fail(code, "group walk: malformed tree (walk did not terminate)");
}

// ---- the phases of one evaluation, enqueued with no host round trip --------------

// Every device buffer an evaluation with bounds p needs.
void reserve_all(gt_hip_context *c, const Plan &p) {
  const size_t n = std::max<uint32_t>(p.n, 1), nn = std::max<uint32_t>(p.node_cap, 1),
               gc = std::max<uint32_t>(p.group_cap, 1);
  c->tree.reserve(p);
  c->parent.ensure(nn);
  c->skip.ensure(nn);
  c->depth.ensure(nn);
  c->recs.ensure(nn);
  c->images.ensure(nn);
  c->span.ensure(nn);
  c->cells.ensure(nn);
  c->bodies.ensure(n);
  c->flags.ensure(n);
  c->flag_node.ensure(n);
  c->offs.ensure(n);
  c->starts.ensure(n);
  c->scan_tmp.ensure(scan_tmp_words(n) + 8);
  c->groups.ensure(gc);
  c->group_node.ensure(gc);
  if (p.systems > 1) c->group_root.ensure(gc);
  c->ccount.ensure(gc);
  c->bcount.ensure(gc);
  if (p.predicted) {
    c->cell_list.ensure(std::max<uint64_t>(p.slot_total_c, 1));
    c->body_list.ensure(std::max<uint64_t>(p.slot_total_b, 1));
    c->feat.ensure(static_cast<size_t>(kPredFeatures) * gc);
    c->pred_out.ensure(static_cast<size_t>(kPredOutputs) * gc);
    c->slot_c.ensure(gc);
    c->slot_b.ensure(gc);
    c->slot_off_c.ensure(gc);
    c->slot_off_b.ensure(gc);
  } else {
    c->cell_list.ensure(std::max<uint64_t>(static_cast<uint64_t>(gc) * p.cap_c, 1));
    c->body_list.ensure(std::max<uint64_t>(static_cast<uint64_t>(gc) * p.cap_b, 1));
  }
  c->out.ensure(n);
  if (p.longest_first) {
    c->parts.ensure(gc);
    c->split_count.ensure(gc);
    c->item_off.ensure(gc);
    c->split_off.ensure(gc);
    c->items.ensure(std::max<uint32_t>(p.item_cap, 1));
    c->partial.ensure(static_cast<size_t>(std::max<uint32_t>(p.split_cap, 1)) * kMaxGroup);
  }
  if (p.hilo) {
    c->bodies_lo.ensure(n);
    c->cells_lo.ensure(nn);
    c->group_hilo.ensure(2 * gc);
  }
  if (p.longest_first) {
    const size_t ic = std::max<uint32_t>(p.item_cap, 1);
    c->item_keys.ensure(ic);
    c->item_keys_alt.ensure(ic);
    c->item_order.ensure(ic);
    c->item_order_alt.ensure(ic);
    c->item_sort_tmp.ensure(radix_tmp_words(ic));
  }
}

// 3: eight lanes per group over child sets (the default), 4: four lanes per
// group, 2: one walk per lane one record per step, 1: the same fetching the
// next records ahead, 16/32/64: coherent walks.
int walk_lanes(const gt_group_config *cfg, const Plan &p) {
  const int l = cfg->walk_lanes;
  const int per_lane = l == 1 || l == 2 || l == 4 ? l : 3;
  if (p.systems > 1 || p.predicted) return per_lane;  // a coherent walk shares one tree
  return l == 16 || l == 32 || l == 64 ? l : per_lane;
}

// Whether the walk for these settings (a walk over child sets) writes the
// hi/lo group centres itself; the other walks leave them to group_centres.
bool walk_writes_centres(const gt_group_config *cfg, const Plan &p) {
  const int l = walk_lanes(cfg, p);
  return l == 3 || l == 4;
}

// Stack entries per lane of the walk over child sets: one per level below the
// root (the tree is at most depth_cap deep, or the tree phase stops first).
uint32_t levels_for(const Plan &p) {
  return static_cast<uint32_t>(std::min(std::max(p.depth_cap, 1), 63) + 1);
}

// Clears the counts of the frame (not its status or step counter).
void clear_counts(unsigned long long *frame, hipStream_t s) {
  launch_fill_u64(frame + frame::kNodes, frame::kCountsEnd - frame::kNodes, 0ull, s);
}

// Group selection, walk into per-group slots, statistics and the status check,
// for a tree in device memory whose node count is frame[kNodes].
// A tree the lists are built for (tsid/roots/root_node: a batch's systems,
// null for one system).
// parent/skip/depth: the tree's per-node links as TreeBuilder keeps them
// (null for a caller's tree: then they are derived from lnk here).
struct ListsInput {
  const double *geom = nullptr;
  const int4 *lnk = nullptr;
  const double *tpos = nullptr, *tmass = nullptr;
  const uint32_t *tsid = nullptr;
  const double *roots = nullptr;
  const uint32_t *root_node = nullptr;
  const uint32_t *parent = nullptr, *skip = nullptr, *depth = nullptr;
  bool records = false;  // the walk's and evaluation's records are written already
  bool groups = false;   // the group starts are flagged already (by the tree build)
};

// The context's record arrays for evaluations with bounds p (reserved).
ListRecords list_records(gt_hip_context *c, const gt_group_config *cfg, const Plan &p) {
  ListRecords r;
  r.inv_theta = 1.0 / cfg->theta;
  r.recs = c->recs.get();
  r.images = c->images.get();
  r.span = c->span.get();
  r.cells = c->cells.get();
  r.bodies = c->bodies.get();
  r.cells_lo = p.hilo ? c->cells_lo.get() : nullptr;
  r.bodies_lo = p.hilo ? c->bodies_lo.get() : nullptr;
  return r;
}

// Everything of the lists before the walk: links, records, groups, and the
// predicted slots.
// The flags of the group starts, for the tree build to write (node_links;
// the lists then take them instead of flagging them again).
GroupFlags tree_group_flags(gt_hip_context *c, const gt_group_config *cfg) {
  GroupFlags f;
  f.gs = static_cast<uint32_t>(cfg->group_size);
  f.pack = cfg->pack < 0 ? 1 : (cfg->pack != 0);
  f.flags = c->flags.get();
  f.node_of = c->flag_node.get();
  return f;
}

void enqueue_lists_prep(gt_hip_context *c, const gt_group_config *cfg, const Plan &p,
                        const ListsInput &in, hipStream_t s) {
  unsigned long long *frame = c->frame.get();
  reserve_all(c, p);
  const bool hilo = p.hilo;
  const uint32_t n = p.n, cap = p.node_cap;
  const uint32_t *parent = in.parent, *skip = in.skip, *depth = in.depth;
  if (!parent || !skip || !depth) {
    launch_link_up(in.lnk, frame, cap, c->parent.get(), c->skip.get(), c->depth.get(), s);
    parent = c->parent.get();
    skip = c->skip.get();
    depth = c->depth.get();
  }
  c->lists_depth = depth;
  if (!in.records)
    launch_prep_lists(in.geom, in.lnk, skip, depth, in.tpos, in.tmass, n, frame, cap, in.tsid,
                      in.roots, list_records(c, cfg, p), s);
  const int pack = cfg->pack < 0 ? 1 : (cfg->pack != 0);
  if (!in.groups)
    launch_group_starts(in.lnk, parent, frame, cap, n, static_cast<uint32_t>(cfg->group_size),
                        pack, c->flags.get(), s, c->flag_node.get());
  exclusive_scan_u32(c->flags.get(), c->offs.get(), n, frame + frame::kGroups, c->scan_tmp.get(),
                     s, c->starts.get());
  uint32_t *groot = in.tsid ? c->group_root.get() : nullptr;
  launch_make_groups(n, frame, p.group_cap, in.tsid, in.root_node, c->starts.get(),
                     c->groups.get(), groot, s, c->flag_node.get(), c->group_node.get());
  // (the child-set walks write the hi/lo group centres themselves)
  if (hilo && !walk_writes_centres(cfg, p))
    launch_group_centres(c->groups.get(), frame, p.group_cap, in.tpos, in.tsid, in.roots, in.geom,
                         c->group_hilo.get(), s);
  if (p.predicted) {
    // slots from the learned lengths, packed at scanned offsets
    launch_group_features(c->groups.get(), frame, p.group_cap, cap, in.tpos, in.lnk, in.geom,
                          groot, c->group_node.get(), depth, c->feat.get(), s);
    launch_predict_slots(c->feat.get(), frame, p.group_cap, c->pred_params.get(),
                         c->pred_stats.get(), c->slot_c.get(), c->slot_b.get(), c->pred_out.get(),
                         s);
    exclusive_scan_u32(c->slot_c.get(), c->slot_off_c.get(), p.group_cap,
                       frame + frame::kPredCells, c->scan_tmp.get(), s);
    exclusive_scan_u32(c->slot_b.get(), c->slot_off_b.get(), p.group_cap,
                       frame + frame::kPredBodies, c->scan_tmp.get(), s);
    launch_check_slot_totals(frame, p.slot_total_c, p.slot_total_b, s);
  }
}

// The walk of range k of p.chunks ranges of groups.
void enqueue_walk_range(gt_hip_context *c, const gt_group_config *cfg, const Plan &p,
                        const ListsInput &in, uint32_t k, hipStream_t s) {
  unsigned long long *frame = c->frame.get();
  const uint32_t chunks = std::max<uint32_t>(p.chunks, 1);
  const uint32_t span = (p.group_cap + chunks - 1) / chunks;
  uint32_t *groot = in.tsid ? c->group_root.get() : nullptr;
  float4 *ghl = p.hilo && walk_writes_centres(cfg, p) ? c->group_hilo.get() : nullptr;
  if (p.predicted)
    launch_walk(kWalkCapacity, walk_lanes(cfg, p), c->recs.get(), c->images.get(),
                levels_for(p), c->span.get(), in.tpos,
                c->groups.get(), frame, p.group_cap, p.node_cap, in.geom, groot, 0, 0,
                c->slot_off_c.get(), c->slot_off_b.get(), c->ccount.get(), c->bcount.get(),
                c->cell_list.get(), c->body_list.get(), s, c->slot_c.get(), c->slot_b.get(),
                k * span, span, ghl);
  else
    launch_walk(kWalkCapacity, walk_lanes(cfg, p), c->recs.get(), c->images.get(),
                levels_for(p), c->span.get(), in.tpos,
                c->groups.get(), frame, p.group_cap, p.node_cap, in.geom, groot, p.cap_c, p.cap_b,
                nullptr, nullptr, c->ccount.get(), c->bcount.get(), c->cell_list.get(),
                c->body_list.get(), s, nullptr, nullptr, k * span, span, ghl);
}

// After the walk: the predictor's error, the statistics and the status check.
void enqueue_lists_finish(gt_hip_context *c, const Plan &p, hipStream_t s) {
  unsigned long long *frame = c->frame.get();
  if (p.predicted)
    launch_prediction_error(frame, p.group_cap, c->pred_out.get(), c->ccount.get(),
                            c->bcount.get(), frame + frame::kPredError, s);
  launch_list_stats(c->ccount.get(), c->bcount.get(), c->groups.get(), frame, p.group_cap, s,
                    &p);
}

// The whole lists phase on the main stream; with p.chunks > 1 the walk runs
// in ranges and walk_done[k] marks the end of range k.
void enqueue_lists(gt_hip_context *c, const gt_group_config *cfg, const Plan &p,
                   const ListsInput &in) {
  hipStream_t s = c->stream;
  enqueue_lists_prep(c, cfg, p, in, s);
  const uint32_t chunks = std::max<uint32_t>(p.chunks, 1);
  for (uint32_t k = 0; k < chunks; ++k) {
    enqueue_walk_range(c, cfg, p, in, k, s);
    if (chunks > 1) GTH_CHECK(hipEventRecord(c->walk_done[k], s));
  }
  enqueue_lists_finish(c, p, s);
}

// The tree the GPU just built, as lists input.
ListsInput tree_input(const gt_hip_context *c) {
  const TreeBuilder &t = c->tree;
  ListsInput in;
  in.geom = t.geom();
  in.lnk = t.links();
  in.tpos = t.tree_pos();
  in.tmass = t.tree_mass();
  in.tsid = t.tree_sid();
  in.roots = t.root();
  in.root_node = t.root_nodes();
  in.parent = t.parents();
  in.skip = t.skips();
  in.depth = t.depths();
  in.records = t.wrote_records();
  in.groups = t.flagged_groups();
  return in;
}

// Lists for the tree the GPU just built.
void enqueue_tree_lists(gt_hip_context *c, const gt_group_config *cfg, const Plan &p) {
  enqueue_lists(c, cfg, p, tree_input(c));
}

// The lists again at exact (scanned) offsets, for an evaluation whose only
// problem was slots too small.  Needs the counts of the capacity walk.
void lists_exact(gt_hip_context *c, const gt_group_config *cfg, const Plan &p, const double *geom,
                 const double *tpos) {
  hipStream_t s = c->stream;
  unsigned long long *frame = c->frame.get();
  const unsigned long long ng = c->fw[frame::kGroups], all_c = c->fw[frame::kStats + 6],
                           all_b = c->fw[frame::kStats + 7];
  if (all_c > 0xffffffffull || all_b > 0xffffffffull)
    fail(GT_HIP_ERR_LIMIT, "interaction lists too long for 32 bit offsets");
  ++c->pstats.list_rewrites;
  uint32_t *co = c->coff.ensure(std::max<unsigned long long>(ng, 1));
  uint32_t *bo = c->boff.ensure(std::max<unsigned long long>(ng, 1));
  unsigned long long *stmp = c->scan_tmp.ensure(scan_tmp_words(ng) + 8);
  unsigned long long *tot = c->misc.ensure(4);
  exclusive_scan_u32(c->ccount.get(), co, ng, tot, stmp, s);
  exclusive_scan_u32(c->bcount.get(), bo, ng, tot + 1, stmp, s);
  uint32_t *cl = c->cell_list.ensure(std::max<unsigned long long>(all_c, 1));
  uint32_t *bl = c->body_list.ensure(std::max<unsigned long long>(all_b, 1));
  launch_fill_u64(frame + frame::kStatus, 1, 0ull, s);
  uint32_t *groot = p.systems > 1 ? c->group_root.get() : nullptr;
  launch_walk(kWalkExact, walk_lanes(cfg, p), c->recs.get(), c->images.get(), levels_for(p),
              c->span.get(), tpos, c->groups.get(),
              frame, p.group_cap, p.node_cap, geom, groot, 0, 0, co, bo, c->ccount.get(),
              c->bcount.get(), cl, bl, s);
  launch_check_frame(frame, p, false, s);
}

// The fast evaluation's work item arrays for bounds p.
EvalSplit split_of(gt_hip_context *c, const Plan &p) {
  EvalSplit sp;
  sp.split_len = p.split_len;
  sp.item_cap = p.item_cap;
  sp.split_cap = p.split_cap;
  sp.parts = c->parts.get();
  sp.split_count = c->split_count.get();
  sp.item_off = c->item_off.get();
  sp.split_off = c->split_off.get();
  sp.items = c->items.get();
  sp.partial = c->partial.get();
  if (p.longest_first) {
    sp.longest_first = true;
    sp.cost_buckets = p.cost_buckets;
    sp.keys = c->item_keys.get();
    sp.keys_alt = c->item_keys_alt.get();
    sp.order = c->item_order.get();
    sp.order_alt = c->item_order_alt.get();
    sp.sort_tmp = c->item_sort_tmp.get();
  }
  return sp;
}

// Work items (made, and sorted longest first, in the lists phase) for the
// longest-first orders; in tree order the evaluation cuts long lists into
// the same parts itself, with the same results.
bool uses_items(const gt_group_config *cfg, const Plan &p) {
  return p.longest_first && !cfg->exact;
}

// Entries per part of the lists that the evaluation cuts itself (0: none).
uint32_t inline_split_for(const gt_group_config *cfg, const Plan &p) {
  if (cfg->exact || uses_items(cfg, p) || p.split_len == 0xffffffffu) return 0;
  return p.split_len;
}

// Work items of the fast evaluation (long lists in parts, each XCD's share
// longest first): part of the lists phase.
void enqueue_items(gt_hip_context *c, const gt_group_config *cfg, const Plan &p,
                   hipStream_t s = nullptr, uint32_t g0 = 0, uint32_t span = 0) {
  if (!uses_items(cfg, p)) return;
  reserve_all(c, p);
  const EvalSplit sp = split_of(c, p);
  launch_eval_items(c->groups.get(), c->frame.get(), p.group_cap, sp, c->scan_tmp.get(),
                    s ? s : c->stream, g0, span);
}

// The list evaluation kernel alone (groups g0 .. g0 + span; span 0: all).
void enqueue_eval(gt_hip_context *c, const gt_group_config *cfg, const Plan &p,
                  hipStream_t s = nullptr, uint32_t g0 = 0, uint32_t span = 0) {
  const float G = static_cast<float>(cfg->G);
  const float eps2 = static_cast<float>(cfg->eps * cfg->eps);
  const bool items = uses_items(cfg, p);
  const EvalSplit sp = items ? split_of(c, p) : EvalSplit{};
  EvalHiLo hl;
  if (p.hilo) {
    hl.bodies_lo = c->bodies_lo.get();
    hl.cells_lo = c->cells_lo.get();
    hl.group_hilo = c->group_hilo.get();
  }
  const int waves = cfg->eval_waves == 2 ? 2 : 3;
  // persistent grid (whole evaluation only): each CU holds `waves` blocks
  const bool persistent = span == 0 && g0 == 0 && persistent_for(cfg);
  const uint32_t pblocks =
      persistent ? static_cast<uint32_t>(c->prop.multiProcessorCount) * waves : 0u;
  launch_eval(c->bodies.get(), c->cells.get(), c->cell_list.get(), c->body_list.get(),
              c->groups.get(), c->frame.get(), p.group_cap, G, eps2, cfg->quadrupole ? 1u : 0u,
              cfg->exact != 0, waves, c->out.get(), s ? s : c->stream, items ? &sp : nullptr,
              p.hilo ? &hl : nullptr, g0, span, c->eval_queue.get(), pblocks,
              inline_split_for(cfg, p));
}

// Results in tree order: split groups' parts added up.
void enqueue_combine(gt_hip_context *c, const gt_group_config *cfg, const Plan &p,
                     hipStream_t s = nullptr, uint32_t g0 = 0, uint32_t span = 0) {
  if (!uses_items(cfg, p)) return;
  launch_combine_parts(c->groups.get(), c->frame.get(), p.group_cap, split_of(c, p), c->out.get(),
                       s ? s : c->stream, g0, span);
}

// The evaluation of the walk's p.chunks ranges on the second stream, each as
// soon as its range is walked (so it runs alongside the walk of the next
// range); eval_done marks the end.  eval_t0/eval_t1 time each range's
// evaluation kernel.  The lists a range's evaluation reads stay inside their
// slots even when they overflowed (the walk cuts the ranges), and an
// evaluation whose lists turn out incomplete is discarded by the status.
void enqueue_overlapped_eval(gt_hip_context *c, const gt_group_config *cfg, const Plan &p) {
  hipStream_t b = c->stream2;
  const uint32_t chunks = p.chunks, span = (p.group_cap + chunks - 1) / chunks;
  for (uint32_t k = 0; k < chunks; ++k) {
    GTH_CHECK(hipStreamWaitEvent(b, c->walk_done[k], 0));
    enqueue_items(c, cfg, p, b, k * span, span);
    GTH_CHECK(hipEventRecord(c->eval_t0[k], b));
    enqueue_eval(c, cfg, p, b, k * span, span);
    GTH_CHECK(hipEventRecord(c->eval_t1[k], b));
    enqueue_combine(c, cfg, p, b, k * span, span);
  }
  GTH_CHECK(hipEventRecord(c->eval_done, b));
  GTH_CHECK(hipStreamWaitEvent(c->stream, c->eval_done, 0));
}

double overlapped_eval_ms(const gt_hip_context *c) {
  double t = 0.0;
  for (uint32_t k = 0; k < c->overlapped; ++k) t += elapsed_ms(c->eval_t0[k], c->eval_t1[k]);
  return t;
}

// After a complete evaluation: sizes into the history and the statistics
// (counted in the pipeline counters unless `count` is false).
void finish(gt_hip_context *c, const Plan &p, const double *geom, bool count = true) {
  const unsigned long long *fw = c->fw;
  learn(c, p, 0);
  const uint32_t nn = static_cast<uint32_t>(fw[frame::kNodes]);
  c->info = gt_lists_info{};
  c->info.n = p.n;
  c->info.n_nodes = nn;
  c->info.depth = fw[frame::kDepth];
  c->info.n_groups = fw[frame::kGroups];
  c->info.cell_interactions = fw[frame::kStats + 0];
  c->info.body_interactions = fw[frame::kStats + 1];
  c->info.max_list = fw[frame::kStats + 2];
  c->info.max_group = fw[frame::kStats + 3];
  c->info.n_cell_list = fw[frame::kStats + 6];
  c->info.n_body_list = fw[frame::kStats + 7];
  c->pstats.last_exact_tests = fw[frame::kExactTests];
  // entries the slot layout spanned: predicted slots, uniform slots per
  // group, or (two passes) the lists themselves
  if (p.predicted)
    c->pstats.last_slot_entries = fw[frame::kPredCells] + fw[frame::kPredBodies];
  else if (p.cap_c && p.cap_b)
    c->pstats.last_slot_entries = c->info.n_groups * (static_cast<uint64_t>(p.cap_c) + p.cap_b);
  else
    c->pstats.last_slot_entries = c->info.n_cell_list + c->info.n_body_list;
  c->lists_n = p.n;
  c->lists_nn = nn;
  c->lists_geom = geom;
  if (count) {
    ++c->pstats.evaluations;
    c->pstats.interactions += c->info.cell_interactions + c->info.body_interactions;
  }
}

// ---- the learned list lengths --------------------------------------------------------

// The slot headroom, in the predictor's device statistics (read by the
// slot kernel, so captured graphs follow changes).
void set_margin(gt_hip_context *c, float margin) {
  c->pred.margin = margin;  // (the copy reads the member, which outlives the call)
  upload(c->pred_stats.get() + kPredStats, &c->pred.margin, 1, c->stream);
}

// Fits the MLP to the gathered samples (Adam, kTrainIterations minibatches of
// kTrainBlocks x 64 samples, learning rate decaying linearly to a tenth) and
// sets the slot headroom from the final error.
void train_predictor(gt_hip_context *c) {
  Predictor &P = c->pred;
  hipStream_t s = c->stream;
  float *params = c->pred_params.ensure(kPredParams), *m = c->pred_m.ensure(kPredParams),
        *v = c->pred_v.ensure(kPredParams), *stats = c->pred_stats.ensure(kPredStats + 1);
  float *grad = c->pred_grad.ensure(static_cast<size_t>(kTrainBlocks) * kPredParams);
  float *loss = c->pred_loss.ensure(kTrainBlocks);
  // He-uniform initial weights, zero biases, from a fixed seed
  std::vector<float> init(kPredParams, 0.0f);
  std::mt19937 rng(12345u);
  auto layer = [&](int off, int rows, int cols) {
    const float a = std::sqrt(6.0f / static_cast<float>(cols));
    std::uniform_real_distribution<float> u(-a, a);
    for (int k = 0; k < rows * cols; ++k) init[off + k] = u(rng);
  };
  constexpr int F = kPredFeatures, H = kPredHidden, O = kPredOutputs;
  layer(0, H, F);
  layer(H * F + H, H, H);
  layer(H * F + H + H * H + H, O, H);
  upload(params, init.data(), kPredParams, s);
  launch_fill_u32(reinterpret_cast<uint32_t *>(m), kPredParams, 0u, s);
  launch_fill_u32(reinterpret_cast<uint32_t *>(v), kPredParams, 0u, s);
  launch_sample_stats(c->pred_samples.get(), P.samples, stats, s);
  for (uint32_t it = 0; it < kTrainIterations; ++it) {
    const float lr = 3e-3f * (1.0f - 0.9f * static_cast<float>(it) / kTrainIterations);
    launch_train_iteration(c->pred_samples.get(), P.samples, stats, params, m, v, grad, loss,
                           kTrainBlocks, it, lr, s);
  }
  unsigned int *rmax = reinterpret_cast<unsigned int *>(c->misc.ensure(4) + 3);
  launch_sample_residual_max(c->pred_samples.get(), P.samples, params, stats, rmax, s);
  std::vector<float> lh(kTrainBlocks);
  unsigned int rbits = 0;
  download(lh.data(), loss, kTrainBlocks, s);
  download(&rbits, rmax, 1, s);
  GTH_CHECK(hipStreamSynchronize(s));
  double sum = 0.0;
  for (float x : lh) sum += x;
  P.loss = sum / (static_cast<double>(kTrainBlocks) * 64 * O);
  // Headroom: the largest residual over the samples (the samples are a subset
  // of the groups, and lists grow between steps: a little more on top).
  float rmaxf = 0.0f;
  std::memcpy(&rmaxf, &rbits, sizeof rmaxf);
  P.margin_floor = std::min(rmaxf + 0.1f, 3.0f);
  set_margin(c, P.margin_floor);
  P.iterations = kTrainIterations;
  P.trained = true;
}

// After a complete evaluation with cfg->predict = 1: the predictor's
// statistics, and (until it is trained) samples of this evaluation's groups.
void predictor_after_eval(gt_hip_context *c, const Plan &p, bool overflowed, const double *geom,
                          const int4 *lnk, const double *tpos) {
  Predictor &P = c->pred;
  const unsigned long long *fw = c->fw;
  const uint64_t ng = fw[frame::kGroups];
  if (p.predicted) {
    ++P.predicted;
    if (overflowed) ++P.overflows;
    // widen the headroom after an overflow, let it settle back slowly
    const float m = overflowed ? std::min(P.margin + 0.1f, 3.0f)
                               : std::max(P.margin_floor, P.margin - 0.005f);
    if (m != P.margin) set_margin(c, m);
    P.err = ng ? static_cast<double>(fw[frame::kPredError]) * 1e-6 / static_cast<double>(ng) : 0.0;
    P.slot_entries = fw[frame::kPredCells] + fw[frame::kPredBodies];
    P.list_entries = fw[frame::kStats + 6] + fw[frame::kStats + 7];
  }
  if (P.trained || P.collected >= P.collect || ng == 0) return;
  hipStream_t s = c->stream;
  const uint32_t gc = std::max<uint32_t>(p.group_cap, 1);
  float *feat = c->feat.ensure(static_cast<size_t>(kPredFeatures) * gc);
  float *samples = c->pred_samples.ensure(static_cast<size_t>(kPredSampleFloats) * kSampleCap);
  const uint32_t *groot = p.systems > 1 ? c->group_root.get() : nullptr;
  launch_group_features(c->groups.get(), c->frame.get(), p.group_cap, p.node_cap, tpos, lnk, geom,
                        groot, c->group_node.get(), c->lists_depth, feat, s);
  const uint32_t stride = static_cast<uint32_t>(std::max<uint64_t>(1, ng / kSamplesPerEval));
  const uint32_t count = static_cast<uint32_t>((ng + stride - 1) / stride);
  launch_collect_samples(c->frame.get(), p.group_cap, stride, feat, c->ccount.get(),
                         c->bcount.get(), P.next, kSampleCap, count, samples, s);
  P.next = (P.next + count) % kSampleCap;
  P.samples = std::min<uint32_t>(P.samples + count, kSampleCap);
  ++P.collected;
  if (P.collected >= P.collect) train_predictor(c);
}

// One force evaluation of particles in device memory: build, lists, evaluate,
// scatter, enqueued with no host round trip, then one synchronisation to read
// the status.  Slots too small: the lists are rewritten at exact offsets and
// evaluated.  Any other bound too small: everything again with the sizes the
// status reported.  Phase times go to c->tm.
void force_eval(gt_hip_context *c, const gt_group_config *cfg, const double *pos,
                const double *mass, uint32_t n, double *acc, double *pot,
                const uint32_t *orig = nullptr) {
  hipStream_t s = c->stream;
  c->tree_on_state = orig != nullptr;
  unsigned long long *frame = c->frame.get();
  for (int attempt = 0;; ++attempt) {
    if (attempt == 8) fail(GT_HIP_ERR_DEVICE, "force evaluation: bounds did not settle");
    const Plan p = plan_for(c, n, cfg);
    const double *geom = c->tree.geom();
    launch_fill_u64(frame + frame::kStatus, 1, 0ull, s);
    clear_counts(frame, s);
    GTH_CHECK(hipEventRecord(c->ev[0], s));
    reserve_all(c, p);
    const ListRecords rec = list_records(c, cfg, p);
    const GroupFlags gf = tree_group_flags(c, cfg);
    c->tree.enqueue(pos, mass, p, cfg->leaf_size, c->sys_off, frame, s, orig, &rec, &gf);
    geom = c->tree.geom();
    GTH_CHECK(hipEventRecord(c->ev[1], s));
    enqueue_tree_lists(c, cfg, p);
    // Without list slots yet (a first evaluation) the lists go to exact
    // offsets below; evaluating now would only meet the overflow status.
    const bool slots = p.predicted || (p.cap_c != 0 && p.cap_b != 0);
    c->overlapped = 0;
    if (slots && p.chunks > 1) {  // evaluation alongside the walk
      GTH_CHECK(hipEventRecord(c->ev[2], s));
      enqueue_overlapped_eval(c, cfg, p);
      c->overlapped = p.chunks;
      GTH_CHECK(hipEventRecord(c->ev[3], s));
      launch_scatter(c->out.get(), c->tree.inverse_order(), n, acc, pot, frame + frame::kStatus,
                     s);
      GTH_CHECK(hipEventRecord(c->ev[4], s));
    } else if (slots) {
      enqueue_items(c, cfg, p);
      GTH_CHECK(hipEventRecord(c->ev[2], s));
      enqueue_eval(c, cfg, p);
      GTH_CHECK(hipEventRecord(c->ev[3], s));
      enqueue_combine(c, cfg, p);
      launch_scatter(c->out.get(), c->tree.inverse_order(), n, acc, pot, frame + frame::kStatus,
                     s);
      GTH_CHECK(hipEventRecord(c->ev[4], s));
    }
    read_frame(c);
    unsigned long long st = c->fw[frame::kStatus];
    check_walk(c, st, GT_HIP_ERR_DEVICE);
    if (!slots && st == 0) st = kAbortLists;  // lists complete but not evaluated: do both below
    const bool overflowed = (st & ~static_cast<unsigned long long>(kAbortItems)) == kAbortLists;
    if (overflowed) {
      // slots too small (the item counts came from incomplete lists): rewrite
      learn(c, p, st);
      lists_exact(c, cfg, p, geom, c->tree.tree_pos());
      c->overlapped = 0;
      enqueue_items(c, cfg, p);
      GTH_CHECK(hipEventRecord(c->ev[2], s));
      enqueue_eval(c, cfg, p);
      GTH_CHECK(hipEventRecord(c->ev[3], s));
      enqueue_combine(c, cfg, p);
      launch_scatter(c->out.get(), c->tree.inverse_order(), n, acc, pot, frame + frame::kStatus,
                     s);
      GTH_CHECK(hipEventRecord(c->ev[4], s));
      read_frame(c);
      st = c->fw[frame::kStatus];
      check_walk(c, st, GT_HIP_ERR_DEVICE);
      if (st & ~static_cast<unsigned long long>(kAbortItems))
        fail(GT_HIP_ERR_DEVICE, "group walk: exact lists incomplete");
    }
    if (st == 0) {
      finish(c, p, geom);
      c->tree.set_counts(static_cast<uint32_t>(c->info.n_nodes), static_cast<int>(c->info.depth));
      if (cfg->predict == 1)
        predictor_after_eval(c, p, overflowed && slots, geom, c->tree.links(),
                             c->tree.tree_pos());
      break;
    }
    learn(c, p, st);
    ++c->pstats.reruns;
  }
  c->tm.tree = elapsed_ms(c->ev[0], c->ev[1]);
  c->tm.lists = elapsed_ms(c->ev[1], c->ev[2]);
  // overlapped: the evaluation kernels' own times (they ran alongside the
  // walk, which `lists` covers), and the wait for the last of them
  c->tm.evaluate = c->overlapped ? overlapped_eval_ms(c) : elapsed_ms(c->ev[2], c->ev[3]);
  c->tm.scatter = elapsed_ms(c->ev[3], c->ev[4]);
  c->tm.total = elapsed_ms(c->ev[0], c->ev[4]);
  c->gpu_seconds = c->tm.total * 1e-3;
}

// ---- leapfrog steps ---------------------------------------------------------------

// The resident state (and its forces) into the order of the tree just built
// from it; orig follows, and the tree's order becomes the identity.
void permute_state(gt_hip_context *c) {
  const uint64_t n = c->n_state;
  if (!n || !c->tree_on_state || c->tree.n() != n) return;
  hipStream_t s = c->stream;
  const uint32_t *order = c->tree.order();
  double *tmp = c->state_tmp.ensure(3 * n);
  const std::pair<double *, int> rows[] = {{c->pos.get(), 3}, {c->vel.get(), 3},
                                           {c->acc.get(), 3}, {c->mass.get(), 1},
                                           {c->pot.get(), 1}};
  for (const auto &r : rows) {
    launch_gather_rows(r.first, order, n, r.second, tmp, s);
    GTH_CHECK(hipMemcpyAsync(r.first, tmp, sizeof(double) * r.second * n,
                             hipMemcpyDeviceToDevice, s));
  }
  uint32_t *ot = c->orig_tmp.ensure(n);
  launch_compose_u32(c->orig.get(), order, n, ot, s);
  GTH_CHECK(hipMemcpyAsync(c->orig.get(), ot, sizeof(uint32_t) * n, hipMemcpyDeviceToDevice, s));
  launch_compose_u32(nullptr, nullptr, n, c->tree.order_mutable(), s);
  launch_compose_u32(nullptr, nullptr, n, c->tree.inverse_order_mutable(), s);
  GTH_CHECK(hipStreamSynchronize(s));
  c->permuted = true;
  c->steps_in_order = 0;
}

// The resident state back into the caller's order (orig the identity again).
void unpermute_state(gt_hip_context *c) {
  const uint64_t n = c->n_state;
  if (!n || !c->permuted) return;
  hipStream_t s = c->stream;
  double *tmp = c->state_tmp.ensure(3 * n);
  const std::pair<double *, int> rows[] = {{c->pos.get(), 3}, {c->vel.get(), 3},
                                           {c->acc.get(), 3}, {c->mass.get(), 1},
                                           {c->pot.get(), 1}};
  for (const auto &r : rows) {
    launch_scatter_rows(r.first, c->orig.get(), n, r.second, tmp, s);
    GTH_CHECK(hipMemcpyAsync(r.first, tmp, sizeof(double) * r.second * n,
                             hipMemcpyDeviceToDevice, s));
  }
  launch_compose_u32(nullptr, nullptr, n, c->orig.get(), s);
  GTH_CHECK(hipStreamSynchronize(s));
  c->permuted = false;
  c->tree_on_state = false;  // its order refers to the permuted rows
}

void eager_step(gt_hip_context *c, const gt_group_config *cfg, double dt) {
  const uint64_t n = c->n_state;
  const double h = 0.5 * dt;
  unsigned long long *status = c->frame.get() + frame::kStatus;
  launch_kick_drift(c->pos.get(), c->vel.get(), c->acc.get(), n, h, dt, status, c->stream);
  force_eval(c, cfg, c->pos.get(), c->mass.get(), static_cast<uint32_t>(n), c->acc.get(),
             c->pot.get(), c->orig.get());
  launch_kick(c->vel.get(), c->acc.get(), n, h, status, c->stream);
}

void destroy_graph(hipGraphExec_t &e) {
  if (e) (void)hipGraphExecDestroy(e);
  e = nullptr;
}

void destroy_step_graphs(gt_hip_context *c) {
  for (auto &e : c->step_exec) destroy_graph(e);
  for (auto &e : c->run_exec) destroy_graph(e);
  for (auto &g : c->step_graph) {
    if (g) (void)hipGraphDestroy(g);
    g = nullptr;
  }
  destroy_graph(c->ov_prep);
  destroy_graph(c->ov_finish);
  for (hipGraphExec_t *arr : {c->ov_walk, c->ov_items, c->ov_eval, c->ov_comb})
    for (uint32_t k = 0; k < kMaxChunks; ++k) destroy_graph(arr[k]);
  c->ov_chunks = 0;
  c->step_valid = false;
}

// What f enqueues on stream st, as a graph.
template <class F>
hipGraph_t capture_only(hipStream_t st, F &&f) {
  GTH_CHECK(hipStreamBeginCapture(st, hipStreamCaptureModeThreadLocal));
  try {
    f();
  } catch (...) {
    hipGraph_t g = nullptr;
    (void)hipStreamEndCapture(st, &g);
    if (g) (void)hipGraphDestroy(g);
    throw;
  }
  hipGraph_t g = nullptr;
  GTH_CHECK(hipStreamEndCapture(st, &g));
  return g;
}

// What f enqueues on stream st, as an instantiated graph.
template <class F>
hipGraphExec_t capture_graph(hipStream_t st, F &&f) {
  hipGraph_t g = capture_only(st, std::forward<F>(f));
  hipGraphExec_t e = nullptr;
  const hipError_t r = hipGraphInstantiate(&e, g, nullptr, nullptr, 0);
  (void)hipGraphDestroy(g);
  GTH_CHECK(r);
  return e;
}

// The kernels of one leapfrog step in phase `phase` (all of them read the
// frame's status, so a step after one that outgrew its bounds does nothing
// to the state).
void enqueue_step_phase(gt_hip_context *c, const gt_group_config *cfg, const Plan &p, double dt,
                        int phase) {
  hipStream_t s = c->stream;
  unsigned long long *frame = c->frame.get();
  const uint64_t n = c->n_state;
  const double h = 0.5 * dt;
  switch (phase) {
    case 0:
      launch_kick_drift(c->pos.get(), c->vel.get(), c->acc.get(), n, h, dt,
                        frame + frame::kStatus, s);
      clear_counts(frame, s);
      {
        reserve_all(c, p);
        const ListRecords rec = list_records(c, cfg, p);
        const GroupFlags gf = tree_group_flags(c, cfg);
        c->tree.enqueue(c->pos.get(), c->mass.get(), p, cfg->leaf_size, c->sys_off, frame, s,
                        c->orig.get(), &rec, &gf);
      }
      break;
    case 1:
      // (overlapped steps launch this phase and the next from graphs per range)
      enqueue_tree_lists(c, cfg, p);
      enqueue_items(c, cfg, p);
      break;
    case 2:
      enqueue_eval(c, cfg, p);
      break;
    default:
      if (p.chunks <= 1) enqueue_combine(c, cfg, p);
      // the scatter and the closing kick in one pass over the rows, which
      // also counts the step
      launch_scatter_kick(c->out.get(), c->tree.inverse_order(), static_cast<uint32_t>(n),
                          c->acc.get(), c->pot.get(), c->vel.get(), h, frame + frame::kStatus, s,
                          frame);
      break;
  }
}

GraphKey key_for(const gt_group_config *cfg, const Plan &p, double dt) {
  GraphKey k;
  k.plan = p;
  k.G = cfg->G;
  k.theta = cfg->theta;
  k.eps = cfg->eps;
  k.dt = dt;
  k.quadrupole = cfg->quadrupole != 0;
  k.group_size = cfg->group_size;
  k.leaf_size = cfg->leaf_size;
  k.pack = cfg->pack < 0 ? 1 : (cfg->pack != 0);
  k.exact = cfg->exact != 0;
  k.eval_waves = cfg->eval_waves == 2 ? 2 : 3;
  k.walk_lanes = walk_lanes(cfg, p);
  k.hilo = uses_hilo(cfg);
  k.persistent = persistent_for(cfg);
  k.generation = alloc_generation;
  return k;
}

// Captures the step's four phases as graphs for bounds p (unless the ones held
// were made for the same bounds, settings and buffers).
void prepare_step_graphs(gt_hip_context *c, const gt_group_config *cfg, const Plan &p, double dt) {
  reserve_all(c, p);
  const GraphKey key = key_for(cfg, p, dt);
  if (c->step_valid && same_key(c->step_key, key)) return;
  destroy_step_graphs(c);
  hipStream_t s = c->stream, b = c->stream2;
  const bool overlap = p.chunks > 1;
  for (int phase = 0; phase < kPhases; ++phase) {
    if (!overlap)
      c->step_graph[phase] =
          capture_only(s, [&] { enqueue_step_phase(c, cfg, p, dt, phase); });
    else if (phase == 0 || phase == 3)
      c->step_exec[phase] = capture_graph(s, [&] { enqueue_step_phase(c, cfg, p, dt, phase); });
  }
  if (overlap) {
    const ListsInput in = tree_input(c);
    const uint32_t chunks = p.chunks, span = (p.group_cap + chunks - 1) / chunks;
    const bool items = uses_items(cfg, p);
    c->ov_prep = capture_graph(s, [&] { enqueue_lists_prep(c, cfg, p, in, s); });
    c->ov_finish = capture_graph(s, [&] { enqueue_lists_finish(c, p, s); });
    for (uint32_t k = 0; k < chunks; ++k) {
      c->ov_walk[k] = capture_graph(s, [&] { enqueue_walk_range(c, cfg, p, in, k, s); });
      if (items) {
        c->ov_items[k] = capture_graph(b, [&] { enqueue_items(c, cfg, p, b, k * span, span); });
        c->ov_comb[k] = capture_graph(b, [&] { enqueue_combine(c, cfg, p, b, k * span, span); });
      }
      c->ov_eval[k] = capture_graph(b, [&] { enqueue_eval(c, cfg, p, b, k * span, span); });
    }
    c->ov_chunks = chunks;
  }
  if (alloc_generation != key.generation)
    fail(GT_HIP_ERR_DEVICE, "step graph: a buffer was allocated during capture");
  c->step_key = key;
  c->step_valid = true;
  ++c->pstats.graph_captures;
  if (c->step_ev.empty()) {
    c->step_ev.assign(static_cast<size_t>(kTimedSteps) * (kPhases + 1), nullptr);
    for (auto &e : c->step_ev) GTH_CHECK(hipEventCreate(&e));
  }
}

void launch_graph(hipGraphExec_t e, hipStream_t st) {
  if (e) GTH_CHECK(hipGraphLaunch(e, st));
}

// k steps (k <= kTimedSteps) as one graph: each step's four phase graphs as
// child graphs in a chain, with the step's timing events recorded between
// them as event record nodes (events recorded on a capturing stream are not
// captured, so they are added as nodes).  Made once per k and kept until the
// phase graphs change.
hipGraphExec_t run_graph(gt_hip_context *c, int k) {
  if (k < 1 || k > kReorderSteps || k > kTimedSteps)
    fail(GT_HIP_ERR_DEVICE, "step graph: run length out of range");
  if (c->run_exec[k]) return c->run_exec[k];
  for (hipGraph_t g : c->step_graph)
    if (!g) fail(GT_HIP_ERR_DEVICE, "step graph: phases not captured");
  hipGraph_t g = nullptr;
  GTH_CHECK(hipGraphCreate(&g, 0));
  hipGraphExec_t e = nullptr;
  try {
    hipGraphNode_t prev = nullptr;
    for (int i = 0; i < k; ++i) {
      hipEvent_t *ev = &c->step_ev[static_cast<size_t>(i) * (kPhases + 1)];
      for (int phase = 0;; ++phase) {
        hipGraphNode_t node = nullptr;
        GTH_CHECK(hipGraphAddEventRecordNode(&node, g, prev ? &prev : nullptr, prev ? 1 : 0,
                                             ev[phase]));
        prev = node;
        if (phase == kPhases) break;
        GTH_CHECK(hipGraphAddChildGraphNode(&node, g, &prev, 1, c->step_graph[phase]));
        prev = node;
      }
    }
    GTH_CHECK(hipGraphInstantiate(&e, g, nullptr, nullptr, 0));
  } catch (...) {
    (void)hipGraphDestroy(g);
    throw;
  }
  (void)hipGraphDestroy(g);
  c->run_exec[k] = e;
  return e;
}

// The lists phase and the evaluation of an overlapped step from their graphs:
// each walk range's evaluation on the second stream as soon as the range is
// walked, with the same events as the launches without graphs.
void launch_overlapped_lists_eval(gt_hip_context *c) {
  hipStream_t s = c->stream, b = c->stream2;
  launch_graph(c->ov_prep, s);
  for (uint32_t k = 0; k < c->ov_chunks; ++k) {
    launch_graph(c->ov_walk[k], s);
    GTH_CHECK(hipEventRecord(c->walk_done[k], s));
    GTH_CHECK(hipStreamWaitEvent(b, c->walk_done[k], 0));
    launch_graph(c->ov_items[k], b);
    GTH_CHECK(hipEventRecord(c->eval_t0[k], b));
    launch_graph(c->ov_eval[k], b);
    GTH_CHECK(hipEventRecord(c->eval_t1[k], b));
    launch_graph(c->ov_comb[k], b);
  }
  launch_graph(c->ov_finish, s);
  GTH_CHECK(hipEventRecord(c->eval_done, b));
  GTH_CHECK(hipStreamWaitEvent(s, c->eval_done, 0));
}

// The graphs for the next run of graph steps (at most nsteps of them): the
// phases for bounds p and, without overlap, the run's steps as one graph.
// Returns the run's length: at most kReorderSteps steps between reorders of
// the state (the graphs stay valid: the state keeps its buffers).
int prepare_run(gt_hip_context *c, const gt_group_config *cfg, const Plan &p, int nsteps,
                double dt) {
  prepare_step_graphs(c, cfg, p, dt);
  const int k = std::min(nsteps, std::max(1, kReorderSteps - c->steps_in_order));
  if (!c->ov_chunks) run_graph(c, k);
  return k;
}

// nsteps leapfrog steps from captured graphs, launched back to back; a step
// that outgrows the bounds stops the rest (their kernels see the status), is
// finished here without graphs, and the remaining steps go on from new graphs.
void graph_steps(gt_hip_context *c, const gt_group_config *cfg, int nsteps, double dt) {
  hipStream_t s = c->stream;
  unsigned long long *frame = c->frame.get();
  const uint32_t n = static_cast<uint32_t>(c->n_state);
  int done_total = 0;
  while (done_total < nsteps) {
    const bool collecting = cfg->predict == 1 && !c->pred.trained;
    if (collecting || !sizes_known(c->hist, n, c->systems)) {
      // sizes not known yet, or the predictor still learning: a step without graphs
      eager_step(c, cfg, dt);
      ++done_total;
      if (++c->steps_in_order >= kReorderSteps) permute_state(c);
      continue;
    }
    const Plan p = plan_for(c, n, cfg);
    const int k = prepare_run(c, cfg, p, nsteps - done_total, dt);
    launch_fill_u64(frame + frame::kStatus, 1, 0ull, s);
    launch_fill_u64(frame + frame::kStepsDone, 2, 0ull, s);  // and kStepInteractions
    if (!c->ov_chunks) {
      GTH_CHECK(hipGraphLaunch(run_graph(c, k), s));
    } else {
      for (int i = 0; i < k; ++i) {
        hipEvent_t *ev = i < kTimedSteps ? &c->step_ev[static_cast<size_t>(i) * (kPhases + 1)]
                                         : nullptr;
        if (ev) GTH_CHECK(hipEventRecord(ev[0], s));
        for (int phase = 0; phase < kPhases; ++phase) {
          if (phase == 1) launch_overlapped_lists_eval(c);
          if (c->step_exec[phase]) GTH_CHECK(hipGraphLaunch(c->step_exec[phase], s));
          if (ev) GTH_CHECK(hipEventRecord(ev[phase + 1], s));
        }
      }
    }
    read_frame(c);
    const int done = static_cast<int>(c->fw[frame::kStepsDone]);
    const unsigned long long st = c->fw[frame::kStatus];
    c->pstats.graph_steps += done;
    c->pstats.evaluations += done;
    c->pstats.interactions += c->fw[frame::kStepInteractions];
    if (done > 0) {
      double t[kPhases] = {}, total = 0.0;
      const int timed = std::min(done, kTimedSteps);
      for (int i = 0; i < timed; ++i) {
        hipEvent_t *ev = &c->step_ev[static_cast<size_t>(i) * (kPhases + 1)];
        for (int phase = 0; phase < kPhases; ++phase) t[phase] += elapsed_ms(ev[phase], ev[phase + 1]);
        total += elapsed_ms(ev[0], ev[kPhases]);
      }
      c->tm.tree = t[0] / timed;
      c->tm.lists = t[1] / timed;
      // overlapped: the last step's evaluation kernels
      c->overlapped = p.chunks > 1 ? p.chunks : 0;
      c->tm.evaluate = c->overlapped ? overlapped_eval_ms(c) : t[2] / timed;
      c->tm.scatter = t[3] / timed;
      c->tm.total = total / timed;
      c->gpu_seconds = c->tm.total * 1e-3;
    }
    done_total += done;
    if (st == 0) {
      // counts of the last step
      finish(c, p, c->tree.geom(), false);
      c->steps_in_order += done;
      if (c->steps_in_order >= kReorderSteps) permute_state(c);
      if (p.predicted) {
        Predictor &P = c->pred;
        P.predicted += done;
        const uint64_t ng = c->fw[frame::kGroups];
        P.err = ng ? static_cast<double>(c->fw[frame::kPredError]) * 1e-6 / ng : 0.0;
        P.slot_entries = c->fw[frame::kPredCells] + c->fw[frame::kPredBodies];
        P.list_entries = c->fw[frame::kStats + 6] + c->fw[frame::kStats + 7];
      }
      c->tree.set_counts(static_cast<uint32_t>(c->info.n_nodes), static_cast<int>(c->info.depth));
      continue;
    }
    check_walk(c, st, GT_HIP_ERR_DEVICE);
    // Step done_total ran its kick and drift only: its forces and second kick here.
    ++c->pstats.graph_stops;
    if (p.predicted && (st & kAbortLists)) {
      ++c->pred.overflows;
      set_margin(c, std::min(c->pred.margin + 0.1f, 3.0f));
    }
    learn(c, p, st);
    force_eval(c, cfg, c->pos.get(), c->mass.get(), n, c->acc.get(), c->pot.get(),
               c->orig.get());
    launch_kick(c->vel.get(), c->acc.get(), n, 0.5 * dt, frame + frame::kStatus, s);
    ++done_total;
    // the graph steps before the stop and the one finished here
    c->steps_in_order += done + 1;
    if (c->steps_in_order >= kReorderSteps) permute_state(c);
  }
}

// The tree phase's sizes into the history (not the lists': a tree solve has
// none, and the lists history of the same particles stays valid).
void learn_tree(gt_hip_context *c, const Plan &p) {
  History &h = c->hist;
  if (h.n != p.n || h.systems != p.systems) h = History{};
  h.n = p.n;
  h.systems = p.systems;
  h.words = p.words;
  h.nodes = static_cast<uint32_t>(c->fw[frame::kNodes]);
  h.depth = static_cast<int>(c->fw[frame::kDepth]);
  h.levels_known = true;
  for (int d = 0; d < frame::kLevels; ++d)
    h.level_nodes[d] = static_cast<uint32_t>(c->fw[frame::kLevelNodes + d]);
}

// The octree of n particles (one system) from device arrays, with its parent
// and skip links: the tree alone, sized again until it fits its bounds.  The
// last lists no longer refer to it.
Plan build_single_tree(gt_hip_context *c, const gt_group_config *cfg, uint32_t n,
                       const double *dp, const double *dm) {
  hipStream_t s = c->stream;
  c->systems = 1;
  c->sys_off = nullptr;
  unsigned long long *frame = c->frame.get();
  Plan p;
  for (int attempt = 0;; ++attempt) {
    if (attempt == 8) fail(GT_HIP_ERR_DEVICE, "tree build: bounds did not settle");
    p = plan_for(c, n, cfg);
    launch_fill_u64(frame + frame::kStatus, 1, 0ull, s);
    clear_counts(frame, s);
    GTH_CHECK(hipEventRecord(c->ev[0], s));
    c->tree_on_state = false;
    c->tree.enqueue(dp, dm, p, cfg->leaf_size, nullptr, frame, s);
    GTH_CHECK(hipEventRecord(c->ev[1], s));
    read_frame(c);
    const unsigned long long *fw = c->fw;
    unsigned long long st = 0;
    if (fw[frame::kNodes] > p.node_cap) st |= kAbortNodes;
    if (fw[frame::kUnresolved] & 1ull) st |= kAbortPaths;
    if (fw[frame::kDepth] > static_cast<unsigned long long>(p.depth_cap)) st |= kAbortDepth;
    if (st == 0) break;
    learn(c, p, st);
    ++c->pstats.reruns;
  }
  learn_tree(c, p);
  const uint32_t cap = p.node_cap;
  c->tree.set_counts(static_cast<uint32_t>(c->fw[frame::kNodes]),
                     static_cast<int>(c->fw[frame::kDepth]));
  c->lists_n = 0;  // (the last lists referred to the tree just replaced)
  c->info = gt_lists_info{};
  c->info.n = n;
  c->info.n_nodes = c->fw[frame::kNodes];
  c->info.depth = c->fw[frame::kDepth];
  return p;
}

}  // namespace

extern "C" {

int gt_hip_compiled_in(void) { return 1; }

int gt_hip_create(gt_hip_context **out, int device) {
  if (!out) return GT_HIP_ERR_ARGUMENT;
  *out = nullptr;
  gt_hip_context *c = nullptr;
  int rc = GT_HIP_OK;
  try {
    int count = 0;
    if (hipGetDeviceCount(&count) != hipSuccess || count == 0)
      fail(GT_HIP_ERR_DEVICE, "hip: no GPU device available");
    if (device < 0 || device >= count)
      fail(GT_HIP_ERR_ARGUMENT, "hip: device " + std::to_string(device) + " does not exist");
    GTH_CHECK(hipSetDevice(device));
    c = new gt_hip_context();
    c->device = device;
    GTH_CHECK(hipGetDeviceProperties(&c->prop, device));
    GTH_CHECK(hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking));
    GTH_CHECK(hipStreamCreateWithFlags(&c->stream2, hipStreamNonBlocking));
    for (auto &e : c->ev) GTH_CHECK(hipEventCreate(&e));
    for (auto &e : c->eval_t0) GTH_CHECK(hipEventCreate(&e));
    for (auto &e : c->eval_t1) GTH_CHECK(hipEventCreate(&e));
    for (auto &e : c->walk_done) GTH_CHECK(hipEventCreateWithFlags(&e, hipEventDisableTiming));
    GTH_CHECK(hipEventCreateWithFlags(&c->eval_done, hipEventDisableTiming));
    launch_fill_u64(c->frame.ensure(frame::kWords), frame::kWords, 0ull, c->stream);
    launch_fill_u32(c->eval_queue.ensure((kXcds + 1) * kQueueStride), (kXcds + 1) * kQueueStride,
                    0u, c->stream);
    launch_fill_u64(c->raw_frame.ensure(frame::kWords), frame::kWords, 0ull, c->stream);
    // the no-allocation entry point's device scratch on this device (current)
    c->noalloc_scratch = noalloc_scratch(device);
    GTH_CHECK(hipStreamSynchronize(c->stream));
    g_create_error.clear();
  } catch (const Error &e) {
    g_create_error = e.what();
    rc = e.code;
  } catch (const std::exception &e) {
    g_create_error = e.what();
    rc = GT_HIP_ERR_DEVICE;
  }
  if (rc != GT_HIP_OK) {
    delete c;
    return rc;
  }
  *out = c;
  return GT_HIP_OK;
}

void gt_hip_destroy(gt_hip_context *c) {
  if (!c) return;
  (void)hipSetDevice(c->device);
  if (c->stream) (void)hipStreamSynchronize(c->stream);
  destroy_step_graphs(c);
  for (auto &e : c->step_ev)
    if (e) (void)hipEventDestroy(e);
  if (c->stream2) (void)hipStreamSynchronize(c->stream2);
  for (auto &e : c->ev)
    if (e) (void)hipEventDestroy(e);
  for (auto *arr : {c->eval_t0, c->eval_t1, c->walk_done})
    for (int k = 0; k < kMaxChunks; ++k)
      if (arr[k]) (void)hipEventDestroy(arr[k]);
  if (c->eval_done) (void)hipEventDestroy(c->eval_done);
  if (c->stream2) (void)hipStreamDestroy(c->stream2);
  if (c->stream) (void)hipStreamDestroy(c->stream);
  delete c;
}

const char *gt_hip_last_error(const gt_hip_context *c) {
  return c ? c->err.c_str() : g_create_error.c_str();
}

int gt_hip_device_name(const gt_hip_context *c, char *buf, size_t len) {
  if (!c || !buf || !len) return GT_HIP_ERR_ARGUMENT;
  std::snprintf(buf, len, "%s (%s, %d CUs)", c->prop.name, c->prop.gcnArchName,
                c->prop.multiProcessorCount);
  return GT_HIP_OK;
}

double gt_hip_last_gpu_seconds(const gt_hip_context *c) { return c ? c->gpu_seconds : 0.0; }

int gt_hip_direct(gt_hip_context *c, const gt_float4 *bodies, const gt_params *p,
                  gt_float4 *out) {
  if (!c) return GT_HIP_ERR_ARGUMENT;
  return guarded(c, [&] {
    require(p != nullptr, "params is NULL");
    const uint32_t n = p->n;
    c->gpu_seconds = 0.0;
    if (n == 0) return;
    require(bodies && out, "bodies/out is NULL");
    GTH_CHECK(hipSetDevice(c->device));
    float4 *db = c->raw_bodies.ensure(n), *dout = c->raw_out.ensure(n);
    upload(db, reinterpret_cast<const float4 *>(bodies), n, c->stream);
    GTH_CHECK(hipEventRecord(c->ev[0], c->stream));
    launch_direct(db, n, p->G, p->eps2, dout, c->stream);
    GTH_CHECK(hipEventRecord(c->ev[1], c->stream));
    download(reinterpret_cast<float4 *>(out), dout, n, c->stream);
    GTH_CHECK(hipStreamSynchronize(c->stream));
    c->gpu_seconds = elapsed_ms(c->ev[0], c->ev[1]) * 1e-3;
  });
}

int gt_hip_tree(gt_hip_context *c, const gt_float4 *bodies, uint64_t n_cells,
                const gt_cell *cells, uint64_t n_cell_list, const uint32_t *cell_list,
                uint64_t n_body_list, const uint32_t *body_list, const gt_group *groups,
                const gt_params *p, int exact, gt_float4 *out) {
  if (!c) return GT_HIP_ERR_ARGUMENT;
  return guarded(c, [&] {
    require(p != nullptr, "params is NULL");
    const uint32_t n = p->n, ng = p->n_groups;
    c->gpu_seconds = 0.0;
    if (n == 0 || ng == 0) return;
    require(bodies && cells && groups && out, "array is NULL");
    require(n_cell_list == 0 || cell_list, "cell_list is NULL");
    require(n_body_list == 0 || body_list, "body_list is NULL");
    // metal_backend.mm:218-223: the arrays must agree with the parameters, and
    // every group must fit one wave; every index must be in range.
    for (uint32_t g = 0; g < ng; ++g) {
      const gt_group &G = groups[g];
      require(G.begin <= G.end && G.end <= n, "hip tree: a group's targets are out of range");
      if (G.end - G.begin > static_cast<uint32_t>(kMaxGroup))
        fail(GT_HIP_ERR_ARGUMENT,
             "hip tree: a group has more than " + std::to_string(kMaxGroup) + " targets");
      require(G.cell_begin <= G.cell_end && G.cell_end <= n_cell_list &&
                  G.body_begin <= G.body_end && G.body_end <= n_body_list,
              "hip tree: a group's list range is out of range");
    }
    for (uint64_t k = 0; k < n_cell_list; ++k)
      require(cell_list[k] < n_cells, "hip tree: cell index out of range");
    for (uint64_t k = 0; k < n_body_list; ++k)
      require(body_list[k] < n, "hip tree: body index out of range");
    GTH_CHECK(hipSetDevice(c->device));
    hipStream_t s = c->stream;
    float4 *db = c->raw_bodies.ensure(n), *dout = c->raw_out.ensure(n);
    Cell *dc = c->raw_cells.ensure(std::max<uint64_t>(n_cells, 1));
    uint32_t *dcl = c->raw_cl.ensure(std::max<uint64_t>(n_cell_list, 1));
    uint32_t *dbl = c->raw_bl.ensure(std::max<uint64_t>(n_body_list, 1));
    gt_group *dg = c->raw_groups.ensure(ng);
    upload(db, reinterpret_cast<const float4 *>(bodies), n, s);
    upload(dc, reinterpret_cast<const Cell *>(cells), n_cells, s);
    upload(dcl, cell_list, n_cell_list, s);
    upload(dbl, body_list, n_body_list, s);
    upload(dg, groups, ng, s);
    unsigned long long *rf = c->raw_frame.get();
    launch_fill_u64(rf, frame::kWords, 0ull, s);
    launch_fill_u64(rf + frame::kGroups, 1, ng, s);
    GTH_CHECK(hipEventRecord(c->ev[0], s));
    launch_eval(db, dc, dcl, dbl, dg, rf, ng, p->G, p->eps2, p->quadrupole, exact != 0, 2, dout, s);
    GTH_CHECK(hipEventRecord(c->ev[1], s));
    download(reinterpret_cast<float4 *>(out), dout, n, s);
    GTH_CHECK(hipStreamSynchronize(s));
    c->gpu_seconds = elapsed_ms(c->ev[0], c->ev[1]) * 1e-3;
  });
}

int gt_hip_tree_noalloc(gt_hip_context *c, const gt_float4 *bodies, uint64_t n_cells,
                        const gt_cell *cells, uint64_t n_cell_list, const uint32_t *cell_list,
                        uint64_t n_body_list, const uint32_t *body_list, const gt_group *groups,
                        const gt_params *p, int exact, gt_float4 *out) {
  if (!c) return GT_HIP_ERR_ARGUMENT;
  return guarded(c, [&] {
    require(p != nullptr, "params is NULL");
    const uint32_t n = p->n, ng = p->n_groups;
    c->gpu_seconds = 0.0;
    if (n == 0 || ng == 0) return;
    require(bodies && cells && groups && out, "array is NULL");
    require(n_cell_list == 0 || cell_list, "cell_list is NULL");
    require(n_body_list == 0 || body_list, "body_list is NULL");
    if (n_cells > 0xffffffffull || n_cell_list > 0xffffffffull || n_body_list > 0xffffffffull)
      fail(GT_HIP_ERR_LIMIT, "hip tree: more than 2^32 - 1 cells or list entries");
    GTH_CHECK(hipSetDevice(c->device));
    noalloc_tree(c->noalloc_scratch, reinterpret_cast<const float4 *>(bodies), n,
                 reinterpret_cast<const Cell *>(cells), n_cells, cell_list, n_cell_list,
                 body_list, n_body_list, groups, ng, p->G, p->eps2, p->quadrupole, exact != 0,
                 reinterpret_cast<float4 *>(out), c->stream, c->ev[0], c->ev[1]);
    c->gpu_seconds = elapsed_ms(c->ev[0], c->ev[1]) * 1e-3;
  });
}

int gt_hip_group_compute(gt_hip_context *c, const gt_group_config *cfg, uint64_t n,
                         const double *pos, const double *mass, double *acc, double *pot) {
  if (!c) return GT_HIP_ERR_ARGUMENT;
  return guarded(c, [&] {
    check_config(cfg);
    if (n > 0xffffffffull) fail(GT_HIP_ERR_LIMIT, "too many particles for the 32 bit GPU indices");
    if (n == 0) {
      c->info = gt_lists_info{};
      c->tm = gt_timings{};
      return;
    }
    require(pos && mass && acc && pot, "array is NULL");
    GTH_CHECK(hipSetDevice(c->device));
    hipStream_t s = c->stream;
    double *dp = c->in_pos.ensure(3 * n), *dm = c->in_mass.ensure(n);
    double *da = c->in_acc.ensure(3 * n), *dq = c->in_pot.ensure(n);
    upload(dp, pos, 3 * n, s);
    upload(dm, mass, n, s);
    c->systems = 1;
    c->sys_off = nullptr;
    force_eval(c, cfg, dp, dm, static_cast<uint32_t>(n), da, dq);
    download(acc, da, 3 * n, s);
    download(pot, dq, n, s);
    GTH_CHECK(hipStreamSynchronize(s));
  });
}

namespace {

// The FP64 tree solver of n particles, for the particles themselves (points
// null) or at np points; outputs in the caller's order (n or np rows).
void tree_solve(gt_hip_context *c, const gt_group_config *cfg, uint64_t n, const double *pos,
                const double *mass, uint64_t np, const double *points, double *acc,
                double *pot, uint64_t *stats) {
  check_config(cfg);
  if (n > 0xffffffffull || np > 0xffffffffull)
    fail(GT_HIP_ERR_LIMIT, "too many particles or points for the 32 bit GPU indices");
  c->tm = gt_timings{};
  const uint64_t targets = points ? np : n;
  if (n == 0 || targets == 0) {
    if (stats) stats[0] = stats[1] = 0;
    for (uint64_t k = 0; k < targets; ++k) {  // no particles: nothing attracts
      if (acc) acc[3 * k] = acc[3 * k + 1] = acc[3 * k + 2] = 0.0;
      if (pot) pot[k] = 0.0;
    }
    return;
  }
  require(pos && mass && acc && pot, "array is NULL");
  GTH_CHECK(hipSetDevice(c->device));
  hipStream_t s = c->stream;
  double *dp = c->in_pos.ensure(3 * n), *dm = c->in_mass.ensure(n);
  double *da = c->in_acc.ensure(3 * targets), *dq = c->in_pot.ensure(targets);
  double *dt = points ? c->pt_pos.ensure(3 * np) : nullptr;
  unsigned long long *st_dev = c->misc.ensure(4);
  upload(dp, pos, 3 * n, s);
  upload(dm, mass, n, s);
  if (points) upload(dt, points, 3 * np, s);
  const uint32_t n32 = static_cast<uint32_t>(n);
  unsigned long long *frame = c->frame.get();
  const Plan p = build_single_tree(c, cfg, n32, dp, dm);
  launch_tree_gravity(c->tree.geom(), c->tree.links(), c->tree.skips(), c->tree.depths(),
                      c->tree.tree_pos(), c->tree.tree_mass(), c->tree.order(), n32, frame,
                      p.node_cap, cfg->G, cfg->eps, cfg->theta, cfg->quadrupole != 0, da, dq,
                      st_dev, s, dt, static_cast<uint32_t>(np));
  GTH_CHECK(hipEventRecord(c->ev[2], s));
  download(acc, da, 3 * targets, s);
  download(pot, dq, targets, s);
  unsigned long long counts[2] = {0, 0};
  download(counts, st_dev, 2, s);
  GTH_CHECK(hipStreamSynchronize(s));
  if (stats) {
    stats[0] = counts[0];
    stats[1] = counts[1];
  }
  c->tm.tree = elapsed_ms(c->ev[0], c->ev[1]);
  c->tm.evaluate = elapsed_ms(c->ev[1], c->ev[2]);
  c->tm.total = elapsed_ms(c->ev[0], c->ev[2]);
  c->gpu_seconds = c->tm.total * 1e-3;
}

}  // namespace

int gt_hip_tree_gravity(gt_hip_context *c, const gt_group_config *cfg, uint64_t n,
                        const double *pos, const double *mass, double *acc, double *pot,
                        uint64_t *stats) {
  if (!c) return GT_HIP_ERR_ARGUMENT;
  return guarded(c, [&] { tree_solve(c, cfg, n, pos, mass, 0, nullptr, acc, pot, stats); });
}

int gt_hip_tree_gravity_at(gt_hip_context *c, const gt_group_config *cfg, uint64_t n,
                           const double *pos, const double *mass, uint64_t np,
                           const double *points, double *acc, double *pot, uint64_t *stats) {
  if (!c) return GT_HIP_ERR_ARGUMENT;
  return guarded(c, [&] {
    require(np == 0 || points != nullptr, "points is NULL");
    if (np == 0) {
      if (stats) stats[0] = stats[1] = 0;
      return;
    }
    tree_solve(c, cfg, n, pos, mass, np, points, acc, pot, stats);
  });
}

constexpr uint64_t kNbStageBytes = 4ull << 30;  // slot memory of a one-walk neighbour fill
constexpr uint32_t kNbStageMax = 256;

int gt_hip_neighbours(gt_hip_context *c, int leaf_size, uint64_t n, const double *pos,
                      const double *mass, uint64_t nq, const double *centres,
                      const double *radii, uint64_t n_radii, uint64_t max_out,
                      uint64_t *offsets, uint64_t *index, double *dist2, uint64_t *total) {
  if (!c) return GT_HIP_ERR_ARGUMENT;
  return guarded(c, [&] {
    require(leaf_size >= 1, "leaf_size must be at least 1");
    if (n > 0xffffffffull || nq > 0xffffffffull)
      fail(GT_HIP_ERR_LIMIT, "too many particles or queries for the 32 bit GPU indices");
    require(offsets != nullptr, "offsets is NULL");
    require(n == 0 || pos != nullptr, "pos is NULL");
    require(nq == 0 || (centres != nullptr && radii != nullptr), "centres or radii is NULL");
    require(nq == 0 || n_radii == 1 || n_radii == nq, "n_radii must be 1 or nq");
    c->tm = gt_timings{};
    if (nq == 0 || n == 0) {  // no queries, or an empty tree: nothing found
      std::fill(offsets, offsets + nq + 1, 0ull);
      if (total) *total = 0;
      return;
    }
    GTH_CHECK(hipSetDevice(c->device));
    hipStream_t s = c->stream;
    const uint32_t n32 = static_cast<uint32_t>(n), nq32 = static_cast<uint32_t>(nq);
    double *dp = c->in_pos.ensure(3 * n), *dm = c->in_mass.ensure(n);
    upload(dp, pos, 3 * n, s);
    if (mass) {
      upload(dm, mass, n, s);
    } else {  // the tree's shape does not depend on the masses: unit masses
      launch_fill_u64(reinterpret_cast<unsigned long long *>(dm), n, 0x3ff0000000000000ull, s);
    }
    gt_group_config cfg{};
    cfg.G = 1.0;
    cfg.theta = 1.0;
    cfg.group_size = 32;
    cfg.leaf_size = leaf_size;
    double *dc = c->nb_centres.ensure(3 * nq);
    double *dr = c->nb_radii.ensure(n_radii);
    upload(dc, centres, 3 * nq, s);
    upload(dr, radii, n_radii, s);
    const Plan p = build_single_tree(c, &cfg, n32, dp, dm);  // (times from here on)
    uint32_t *counts = c->nb_counts.ensure(nq), *offs = c->nb_offs.ensure(nq);
    unsigned long long *tmp = c->nb_tmp.ensure(scan_tmp_words(nq));
    unsigned long long *dev = c->misc.ensure(4);  // [0] error bits, [1] total
    unsigned long long *frame = c->frame.get();
    const uint32_t stride = n_radii == 1 ? 0u : 1u;
    // lanes take the queries in Morton order of their centres
    uint32_t *perm = c->nb_perm.ensure(nq);
    uint64_t *keys = c->nb_keys.ensure(nq);
    launch_query_keys(c->tree.geom(), frame, dc, nq32, keys, perm, s);
    radix_sort_u64(keys, perm, c->nb_keys_alt.ensure(nq), c->nb_perm_alt.ensure(nq), nq,
                   kQueryKeyBits, c->nb_sort_tmp.ensure(radix_tmp_words(nq)), s,
                   frame + frame::kSortCtl);
    void *recs = c->nb_nodes.ensure((nb_record_bytes(p.node_cap) + 15) / 16);
    launch_nb_records(c->tree.geom(), c->tree.links(), c->tree.skips(), c->tree.depths(), frame,
                      p.node_cap, recs, s);
    launch_fill_u64(dev, 3, 0ull, s);  // [2]: queries whose lists outgrew their slots
    // A call that wants the lists walks once, writing each query's first
    // nb_stage neighbours into a slot of its own and counting them all; the
    // lists that fit are then moved to their CSR places and only the others
    // are walked again.  Counting calls, and slots past kNbStageBytes, take a
    // counting walk and a filling walk instead.
    const bool want = index || dist2;
    uint32_t stage = want ? c->nb_stage : 0u;
    if (static_cast<uint64_t>(stage) * nq * 16 > kNbStageBytes) stage = 0;
    unsigned long long *si = nullptr;
    double *sd = nullptr;
    if (stage) {
      si = c->nb_stage_index.ensure(static_cast<size_t>(stage) * nq);
      sd = c->nb_stage_dist2.ensure(static_cast<size_t>(stage) * nq);
    }
    launch_within(recs, c->tree.tree_pos(), c->tree.order(), n32, frame, p.node_cap, dc, dr,
                  stride, nq32, nq32, perm, counts, nullptr, si, sd, stage, dev, s);
    exclusive_scan_u32(counts, offs, nq, dev + 1, tmp, s);
    GTH_CHECK(hipEventRecord(c->ev[3], s));
    unsigned long long h[3] = {0, 0, 0};
    download(h, dev, 2, s);
    GTH_CHECK(hipStreamSynchronize(s));
    if (h[0]) fail(GT_HIP_ERR_DEVICE, "neighbour queries: malformed tree");
    if (h[1] > 0xffffffffull)
      fail(GT_HIP_ERR_LIMIT, "more than 2^32 - 1 neighbours in one call: split the queries");
    const bool fill = h[1] && want && h[1] <= max_out;
    GTH_CHECK(hipEventRecord(c->ev[4], s));
    if (fill) {
      unsigned long long *di = c->nb_index.ensure(h[1]);
      double *dd = c->nb_dist2.ensure(h[1]);
      uint32_t *lanes_perm = perm, lanes = nq32;
      if (stage) {
        uint32_t *over = c->nb_over.ensure(nq);
        launch_unstage(counts, offs, nq32, stage, si, sd, di, dd, over, dev + 2, s);
        download(h + 2, dev + 2, 1, s);
        GTH_CHECK(hipStreamSynchronize(s));
        lanes_perm = over;
        lanes = static_cast<uint32_t>(h[2]);
        if (h[2] > nq / 8 && stage < kNbStageMax) c->nb_stage = std::min(2 * stage, kNbStageMax);
      }
      launch_within(recs, c->tree.tree_pos(), c->tree.order(), n32, frame, p.node_cap, dc, dr,
                    stride, nq32, lanes, lanes_perm, counts, offs, di, dd, 0, dev, s);
      GTH_CHECK(hipEventRecord(c->ev[2], s));
      if (index) download(reinterpret_cast<unsigned long long *>(index), di, h[1], s);
      if (dist2) download(dist2, dd, h[1], s);
    } else {
      GTH_CHECK(hipEventRecord(c->ev[2], s));
    }
    std::vector<uint32_t> ho(nq);
    download(ho.data(), offs, nq, s);
    download(h, dev, 1, s);
    GTH_CHECK(hipStreamSynchronize(s));
    if (h[0]) fail(GT_HIP_ERR_DEVICE, "neighbour queries: the count and fill walks differ");
    for (uint64_t q = 0; q < nq; ++q) offsets[q] = ho[q];
    offsets[nq] = h[1];
    if (total) *total = h[1];
    if ((index || dist2) && h[1] > max_out)
      fail(GT_HIP_ERR_LIMIT, "max_out too small: the queries have " + std::to_string(h[1]) +
                                 " neighbours (offsets and total are written)");
    c->tm.tree = elapsed_ms(c->ev[0], c->ev[1]);
    c->tm.lists = elapsed_ms(c->ev[1], c->ev[3]);
    c->tm.evaluate = elapsed_ms(c->ev[4], c->ev[2]);
    c->tm.total = c->tm.tree + c->tm.lists + c->tm.evaluate;
    c->gpu_seconds = c->tm.total * 1e-3;
  });
}

// Checks a batch's offsets (n_systems + 1 of them, from 0 to n, non-decreasing)
// and puts them in device memory.
void upload_systems(gt_hip_context *c, uint64_t n, uint64_t n_systems, const uint64_t *offsets,
                    DBuf<unsigned long long> &dst) {
  require(n_systems >= 1, "n_systems must be at least 1");
  require(offsets != nullptr, "offsets is NULL");
  if (n_systems > (1ull << 24)) fail(GT_HIP_ERR_LIMIT, "at most 2^24 systems in one batch");
  require(offsets[0] == 0 && offsets[n_systems] == n, "offsets must run from 0 to n");
  for (uint64_t k = 0; k < n_systems; ++k)
    require(offsets[k] <= offsets[k + 1], "offsets must not decrease");
  unsigned long long *d = dst.ensure(n_systems + 1);
  upload(d, reinterpret_cast<const unsigned long long *>(offsets), n_systems + 1, c->stream);
}

int gt_hip_group_compute_batch(gt_hip_context *c, const gt_group_config *cfg, uint64_t n_systems,
                               const uint64_t *offsets, const double *pos, const double *mass,
                               double *acc, double *pot) {
  if (!c) return GT_HIP_ERR_ARGUMENT;
  return guarded(c, [&] {
    check_config(cfg);
    require(offsets != nullptr && n_systems >= 1, "offsets is NULL or n_systems is 0");
    const uint64_t n = offsets[n_systems];
    if (n > 0xffffffffull) fail(GT_HIP_ERR_LIMIT, "too many particles for the 32 bit GPU indices");
    GTH_CHECK(hipSetDevice(c->device));
    hipStream_t s = c->stream;
    upload_systems(c, n, n_systems, offsets, c->batch_off);
    if (n == 0) {
      GTH_CHECK(hipStreamSynchronize(s));
      c->info = gt_lists_info{};
      c->tm = gt_timings{};
      return;
    }
    require(pos && mass && acc && pot, "array is NULL");
    double *dp = c->in_pos.ensure(3 * n), *dm = c->in_mass.ensure(n);
    double *da = c->in_acc.ensure(3 * n), *dq = c->in_pot.ensure(n);
    upload(dp, pos, 3 * n, s);
    upload(dm, mass, n, s);
    c->systems = static_cast<uint32_t>(n_systems);
    c->sys_off = n_systems > 1 ? c->batch_off.get() : nullptr;
    if (n_systems == 1) c->systems = 1;
    force_eval(c, cfg, dp, dm, static_cast<uint32_t>(n), da, dq);
    download(acc, da, 3 * n, s);
    download(pot, dq, n, s);
    GTH_CHECK(hipStreamSynchronize(s));
  });
}

namespace {

// The autotuner over one system (offsets NULL) or a batch of independent
// systems (n_systems + 1 offsets): times real evaluations of the caller's
// particles while varying one tuning setting at a time.
void autotune_impl(gt_hip_context *c, const gt_group_config *cfg, uint64_t n_systems,
                   const uint64_t *offsets, uint64_t n, const double *pos, const double *mass,
                   int search_group_size, gt_group_config *best, double *best_ms) {
  check_config(cfg);
  require(best != nullptr, "best is NULL");
  if (n > 0xffffffffull) fail(GT_HIP_ERR_LIMIT, "too many particles for the 32 bit GPU indices");
  *best = *cfg;
  if (best_ms) *best_ms = 0.0;
  GTH_CHECK(hipSetDevice(c->device));
  hipStream_t s = c->stream;
  if (offsets) upload_systems(c, n, n_systems, offsets, c->batch_off);
  if (n == 0) {
    GTH_CHECK(hipStreamSynchronize(s));
    return;
  }
  require(pos && mass, "array is NULL");
  double *dp = c->in_pos.ensure(3 * n), *dm = c->in_mass.ensure(n);
  double *da = c->in_acc.ensure(3 * n), *dq = c->in_pot.ensure(n);
  upload(dp, pos, 3 * n, s);
  upload(dm, mass, n, s);
  c->systems = offsets && n_systems > 1 ? static_cast<uint32_t>(n_systems) : 1;
  c->sys_off = offsets && n_systems > 1 ? c->batch_off.get() : nullptr;
  // One warm-up (bounds settle) then the best of two timed evaluations.
  auto time_of = [&](const gt_group_config &t) {
    force_eval(c, &t, dp, dm, static_cast<uint32_t>(n), da, dq);
    double ms = 1e300;
    for (int r = 0; r < 2; ++r) {
      force_eval(c, &t, dp, dm, static_cast<uint32_t>(n), da, dq);
      ms = std::min(ms, c->tm.total);
    }
    return ms;
  };
  gt_group_config cur = *cfg;
  double cur_ms = time_of(cur);
  auto sweep = [&](int gt_group_config::*field, std::initializer_list<int> values) {
    const int start = cur.*field;
    for (int v : values) {
      if (v == start) continue;
      gt_group_config t = cur;
      t.*field = v;
      const double ms = time_of(t);
      if (ms < 0.99 * cur_ms) {  // a change must beat run-to-run noise
        cur = t;
        cur_ms = ms;
      }
    }
  };
  if (search_group_size) sweep(&gt_group_config::group_size, {32, 64});
  sweep(&gt_group_config::walk_lanes, {2, 4, 16, 32, 64});
  if (!cfg->exact) {
    sweep(&gt_group_config::eval_waves, {2, 3});
    sweep(&gt_group_config::eval_order, {1, 2, 3, 4});
    sweep(&gt_group_config::split_len, {2048, 8192, -1});
    sweep(&gt_group_config::persistent, {-1, 1});
  }
  // The sweep keeps a change on one pair of timings each; noise can make a
  // slower setting look faster.  The winner must beat the starting settings
  // again, timed alternately, best of three each.
  const gt_group_config start = *cfg;
  if (std::memcmp(&cur, &start, sizeof(cur)) != 0) {
    double t_cur = 1e300, t_start = 1e300;
    for (int r = 0; r < 3; ++r) {
      t_start = std::min(t_start, time_of(start));
      t_cur = std::min(t_cur, time_of(cur));
    }
    if (!(t_cur < 0.99 * t_start)) {
      cur = start;
      cur_ms = t_start;
    } else {
      cur_ms = t_cur;
    }
  }
  *best = cur;
  if (best_ms) *best_ms = cur_ms;
}

}  // namespace

int gt_hip_autotune(gt_hip_context *c, const gt_group_config *cfg, uint64_t n, const double *pos,
                    const double *mass, int search_group_size, gt_group_config *best,
                    double *best_ms) {
  if (!c) return GT_HIP_ERR_ARGUMENT;
  return guarded(c, [&] {
    autotune_impl(c, cfg, 1, nullptr, n, pos, mass, search_group_size, best, best_ms);
  });
}

int gt_hip_autotune_batch(gt_hip_context *c, const gt_group_config *cfg, uint64_t n_systems,
                          const uint64_t *offsets, const double *pos, const double *mass,
                          int search_group_size, gt_group_config *best, double *best_ms) {
  if (!c) return GT_HIP_ERR_ARGUMENT;
  return guarded(c, [&] {
    require(offsets != nullptr && n_systems >= 1, "offsets is NULL or n_systems is 0");
    autotune_impl(c, cfg, n_systems, offsets, offsets[n_systems], pos, mass, search_group_size,
                  best, best_ms);
  });
}

int gt_hip_walk_tree(gt_hip_context *c, const gt_group_config *cfg, uint64_t n,
                     uint64_t n_nodes, const double *geom, const int32_t *links,
                     const double *tree_pos, const double *tree_mass, gt_float4 *out) {
  if (!c) return GT_HIP_ERR_ARGUMENT;
  return guarded(c, [&] {
    check_config(cfg);
    if (n > 0xffffffffull || n_nodes > 0x7fffffffull)
      fail(GT_HIP_ERR_LIMIT, "too many particles or nodes for the 32 bit GPU indices");
    require(n == 0 || (n_nodes > 0 && geom && links && tree_pos && tree_mass),
            "tree arrays are NULL");
    if (n_nodes >= kRecIndexMask)
      fail(GT_HIP_ERR_LIMIT, "too many tree nodes for the packed walk records (2^26 - 1)");
    // The walk trusts the links: check them here so a bad tree is an error,
    // not an out-of-bounds access on the GPU.
    for (uint64_t k = 0; k < n_nodes; ++k) {
      const int32_t fc = links[4 * k], nc = links[4 * k + 1];
      const uint32_t b = static_cast<uint32_t>(links[4 * k + 2]),
                     e = static_cast<uint32_t>(links[4 * k + 3]);
      require(b <= e && e <= n, "tree: node range out of bounds");
      require(nc >= 0 && nc <= 8, "tree: bad child count");
      if (nc > 0)
        require(fc > static_cast<int64_t>(k) && static_cast<uint64_t>(fc) + nc <= n_nodes,
                "tree: children out of range");
    }
    require(n_nodes == 0 || (links[2] == 0 && static_cast<uint64_t>(links[3]) == n),
            "tree: the root must hold every body");
    std::vector<uint8_t> seen(n_nodes, 0);
    // depths (children come after their parent, so one pass in node order)
    std::vector<uint8_t> node_depth(n_nodes, 0);
    int tree_depth = 0;
    for (uint64_t k = 0; k < n_nodes; ++k)
      for (int32_t ch = 0; ch < links[4 * k + 1]; ++ch) {
        const uint64_t x = static_cast<uint64_t>(links[4 * k]) + ch;
        require(seen[x] == 0, "tree: a node has two parents");
        seen[x] = 1;
        require(node_depth[k] < 63, "tree: deeper than 63 levels");
        node_depth[x] = static_cast<uint8_t>(node_depth[k] + 1);
        tree_depth = std::max<int>(tree_depth, node_depth[x]);
      }
    GTH_CHECK(hipSetDevice(c->device));
    hipStream_t s = c->stream;
    double *dg = c->up_geom.ensure(static_cast<size_t>(kGeomStride) * std::max<uint64_t>(n_nodes, 1));
    int4 *dl = c->up_lnk.ensure(std::max<uint64_t>(n_nodes, 1));
    double *dp = c->up_tpos.ensure(3 * std::max<uint64_t>(n, 1));
    double *dm = c->up_tmass.ensure(std::max<uint64_t>(n, 1));
    upload(dg, geom, static_cast<size_t>(kGeomStride) * n_nodes, s);
    upload(dl, reinterpret_cast<const int4 *>(links), n_nodes, s);
    upload(dp, tree_pos, 3 * n, s);
    upload(dm, tree_mass, n, s);
    c->info = gt_lists_info{};
    c->tm = gt_timings{};
    if (n == 0) {
      GTH_CHECK(hipStreamSynchronize(s));
      c->lists_n = 0;
      return;
    }
    unsigned long long *frame = c->frame.get();
    c->systems = 1;
    c->sys_off = nullptr;
    Plan p = plan_for(c, static_cast<uint32_t>(n), cfg);
    p.node_cap = static_cast<uint32_t>(n_nodes);
    p.depth_cap = std::max(p.depth_cap, tree_depth);  // the walk's stack has a level each
    for (int attempt = 0;; ++attempt) {
      if (attempt == 4) fail(GT_HIP_ERR_DEVICE, "group walk: bounds did not settle");
      launch_fill_u64(frame + frame::kStatus, 1, 0ull, s);
      clear_counts(frame, s);
      launch_fill_u64(frame + frame::kNodes, 1, n_nodes, s);
      GTH_CHECK(hipEventRecord(c->ev[1], s));
      ListsInput in;
      in.geom = dg;
      in.lnk = dl;
      in.tpos = dp;
      in.tmass = dm;
      enqueue_lists(c, cfg, p, in);
      read_frame(c);
      unsigned long long st = c->fw[frame::kStatus];
      check_walk(c, st, GT_HIP_ERR_ARGUMENT);
      if (st == kAbortLists) {
        learn(c, p, st);
        lists_exact(c, cfg, p, dg, dp);
        read_frame(c);
        st = c->fw[frame::kStatus];
        check_walk(c, st, GT_HIP_ERR_ARGUMENT);
        if (st) fail(GT_HIP_ERR_DEVICE, "group walk: exact lists incomplete");
      }
      if (st == 0) break;
      // more groups than the bound: walk again with room for all of them
      p.group_cap = static_cast<uint32_t>(
          std::min<unsigned long long>(std::max<unsigned long long>(c->fw[frame::kGroups], 1), n));
      p.cap_c = p.cap_b = 0;
      ++c->pstats.reruns;
    }
    finish(c, p, dg);
    GTH_CHECK(hipEventRecord(c->ev[2], s));
    if (out) {
      for (int attempt = 0;; ++attempt) {
        if (attempt == 4) fail(GT_HIP_ERR_DEVICE, "evaluation: bounds did not settle");
        enqueue_items(c, cfg, p);
        enqueue_eval(c, cfg, p);
        enqueue_combine(c, cfg, p);
        read_frame(c);
        if (!(c->fw[frame::kStatus] & kAbortItems)) break;
        // more work items than the bounds: room for what was counted
        learn(c, p, kAbortItems);
        p.split_cap = grown(c->fw[frame::kSplitParts], 64, 1u << 26);
        p.item_cap = static_cast<uint32_t>(
            std::min<uint64_t>(grown(c->fw[frame::kItems], 64, 0xffffffffull), 0xffffffffull));
        launch_fill_u64(frame + frame::kStatus, 1, 0ull, s);
        ++c->pstats.reruns;
      }
      download(reinterpret_cast<float4 *>(out), c->out.get(), n, s);
    }
    // The depth of a caller's tree is not known here: the next GPU build
    // starts from default bounds.
    c->hist.depth = -1;
    GTH_CHECK(hipEventRecord(c->ev[3], s));
    GTH_CHECK(hipStreamSynchronize(s));
    c->tm.lists = elapsed_ms(c->ev[1], c->ev[2]);
    c->tm.evaluate = elapsed_ms(c->ev[2], c->ev[3]);
    c->tm.total = c->tm.lists + c->tm.evaluate;
  });
}

int gt_hip_last_info(const gt_hip_context *c, gt_lists_info *info) {
  if (!c || !info) return GT_HIP_ERR_ARGUMENT;
  *info = c->info;
  return GT_HIP_OK;
}

int gt_hip_last_timings(const gt_hip_context *c, gt_timings *t) {
  if (!c || !t) return GT_HIP_ERR_ARGUMENT;
  *t = c->tm;
  return GT_HIP_OK;
}

int gt_hip_get_tree(gt_hip_context *c, double *geom, int32_t *links, uint64_t *order) {
  if (!c) return GT_HIP_ERR_ARGUMENT;
  return guarded(c, [&] {
    const uint32_t nn = c->tree.n_nodes(), n = c->tree.n();
    hipStream_t s = c->stream;
    if (geom) download(geom, c->tree.geom(), static_cast<size_t>(kGeomStride) * nn, s);
    if (links) download(reinterpret_cast<int4 *>(links), c->tree.links(), nn, s);
    std::vector<uint32_t> o(order ? n : 0), orig;
    if (order) download(o.data(), c->tree.order(), n, s);
    // a tree of the resident state: its positions' particles in the caller's order
    const bool mapped = order && c->tree_on_state && c->permuted && c->n_state == n;
    if (mapped) {
      orig.resize(n);
      download(orig.data(), c->orig.get(), n, s);
    }
    GTH_CHECK(hipStreamSynchronize(s));
    for (uint32_t k = 0; k < o.size(); ++k) order[k] = mapped ? orig[o[k]] : o[k];
  });
}

int gt_hip_get_lists(gt_hip_context *c, gt_group *groups, uint32_t *cell_list,
                     uint32_t *body_list, gt_float4 *bodies, gt_cell *cells, double *origin) {
  if (!c) return GT_HIP_ERR_ARGUMENT;
  return guarded(c, [&] {
    hipStream_t s = c->stream;
    const gt_lists_info &I = c->info;
    if (c->lists_n == 0) return;
    // The device lists may sit in per-group slots with gaps between groups:
    // hand them out packed, as build_group_lists lays them out.
    const uint32_t ng = static_cast<uint32_t>(I.n_groups);
    std::vector<gt_group> g(ng);
    download(g.data(), c->groups.get(), ng, s);
    GTH_CHECK(hipStreamSynchronize(s));
    uint64_t cext = 0, bext = 0;
    for (const auto &G : g) {
      cext = std::max<uint64_t>(cext, G.cell_end);
      bext = std::max<uint64_t>(bext, G.body_end);
    }
    std::vector<uint32_t> cl(cell_list ? cext : 0), bl(body_list ? bext : 0);
    if (cell_list) download(cl.data(), c->cell_list.get(), cext, s);
    if (body_list) download(bl.data(), c->body_list.get(), bext, s);
    if (bodies) download(reinterpret_cast<float4 *>(bodies), c->bodies.get(), c->lists_n, s);
    if (cells) download(reinterpret_cast<Cell *>(cells), c->cells.get(), c->lists_nn, s);
    if (origin) download(origin, c->lists_geom, 3, s);
    GTH_CHECK(hipStreamSynchronize(s));
    uint32_t nc = 0, nb = 0;
    for (uint32_t k = 0; k < ng; ++k) {
      gt_group G = g[k];
      if (cell_list)
        std::copy(cl.begin() + G.cell_begin, cl.begin() + G.cell_end, cell_list + nc);
      if (body_list)
        std::copy(bl.begin() + G.body_begin, bl.begin() + G.body_end, body_list + nb);
      const uint32_t c_len = G.cell_end - G.cell_begin, b_len = G.body_end - G.body_begin;
      G.cell_begin = nc;
      G.cell_end = nc + c_len;
      G.body_begin = nb;
      G.body_end = nb + b_len;
      nc += c_len;
      nb += b_len;
      if (groups) groups[k] = G;
    }
  });
}

int gt_hip_state_set(gt_hip_context *c, uint64_t n, const double *pos, const double *vel,
                     const double *mass) {
  if (!c) return GT_HIP_ERR_ARGUMENT;
  return guarded(c, [&] {
    if (n > 0xffffffffull) fail(GT_HIP_ERR_LIMIT, "too many particles for the 32 bit GPU indices");
    require(n == 0 || (pos && vel && mass), "array is NULL");
    GTH_CHECK(hipSetDevice(c->device));
    hipStream_t s = c->stream;
    const size_t m = std::max<uint64_t>(n, 1);
    c->pos.ensure(3 * m);
    c->vel.ensure(3 * m);
    c->mass.ensure(m);
    c->acc.ensure(3 * m);
    c->pot.ensure(m);
    c->orig.ensure(m);
    upload(c->pos.get(), pos, 3 * n, s);
    upload(c->vel.get(), vel, 3 * n, s);
    upload(c->mass.get(), mass, n, s);
    launch_compose_u32(nullptr, nullptr, n, c->orig.get(), s);  // the identity
    GTH_CHECK(hipStreamSynchronize(s));
    c->permuted = false;
    c->tree_on_state = false;
    c->n_state = n;
    c->state_systems = 1;
    c->state_forces = false;
  });
}

int gt_hip_state_set_systems(gt_hip_context *c, uint64_t n_systems, const uint64_t *offsets) {
  if (!c) return GT_HIP_ERR_ARGUMENT;
  return guarded(c, [&] {
    GTH_CHECK(hipSetDevice(c->device));
    unpermute_state(c);  // the offsets are ranges of the caller's order
    upload_systems(c, c->n_state, n_systems, offsets, c->state_off);
    GTH_CHECK(hipStreamSynchronize(c->stream));
    c->state_systems = static_cast<uint32_t>(n_systems);
    c->state_forces = false;
  });
}

int gt_hip_state_get(gt_hip_context *c, double *pos, double *vel, double *acc, double *pot) {
  if (!c) return GT_HIP_ERR_ARGUMENT;
  return guarded(c, [&] {
    const uint64_t n = c->n_state;
    hipStream_t s = c->stream;
    if (!c->permuted) {
      if (pos) download(pos, c->pos.get(), 3 * n, s);
      if (vel) download(vel, c->vel.get(), 3 * n, s);
      if (acc) download(acc, c->acc.get(), 3 * n, s);
      if (pot) download(pot, c->pot.get(), n, s);
      GTH_CHECK(hipStreamSynchronize(s));
      return;
    }
    // back into the caller's order: row k of the resident state is particle orig[k]
    std::vector<uint32_t> o(n);
    download(o.data(), c->orig.get(), n, s);
    std::vector<double> t;
    auto get = [&](double *out, const double *dev, int w) {
      if (!out) return;
      t.resize(w * n);
      download(t.data(), dev, w * n, s);
      GTH_CHECK(hipStreamSynchronize(s));
      for (uint64_t k = 0; k < n; ++k)
        for (int a = 0; a < w; ++a) out[w * static_cast<uint64_t>(o[k]) + a] = t[w * k + a];
    };
    get(pos, c->pos.get(), 3);
    get(vel, c->vel.get(), 3);
    get(acc, c->acc.get(), 3);
    get(pot, c->pot.get(), 1);
    GTH_CHECK(hipStreamSynchronize(s));
  });
}

int gt_hip_state_diagnostics(gt_hip_context *c, double time, uint64_t max_out,
                             gt_diagnostics *out) {
  if (!c) return GT_HIP_ERR_ARGUMENT;
  return guarded(c, [&] {
    const uint32_t systems = c->state_systems;
    require(out != nullptr, "out is NULL");
    require(max_out >= systems, "out holds fewer entries than the state has systems");
    require(c->n_state == 0 || c->state_forces,
            "potentials are not current: call gt_hip_prime first");
    for (uint32_t k = 0; k < systems; ++k) {
      out[k] = gt_diagnostics{};
      out[k].time = time;
    }
    if (c->n_state == 0) return;
    GTH_CHECK(hipSetDevice(c->device));
    hipStream_t s = c->stream;
    double *partial = c->diag_partial.ensure(diag_partial_words(systems));
    double *total = c->diag_total.ensure(14 * static_cast<size_t>(systems));
    launch_diagnostics(c->pos.get(), c->vel.get(), c->mass.get(), c->pot.get(),
                       systems > 1 ? c->state_off.get() : nullptr, systems, c->n_state, partial,
                       total, s);
    std::vector<double> t(14 * static_cast<size_t>(systems));
    download(t.data(), total, t.size(), s);
    GTH_CHECK(hipStreamSynchronize(s));
    for (uint32_t k = 0; k < systems; ++k) {
      const double *v = t.data() + 14 * static_cast<size_t>(k);
      gt_diagnostics &d = out[k];
      d.kinetic = v[0];
      d.potential = v[1];
      for (int a = 0; a < 3; ++a) {
        d.momentum[a] = v[2 + a];
        d.angular_momentum[a] = v[5 + a];
        d.centre_of_mass[a] = v[11] > 0 ? v[8 + a] / v[11] : 0.0;
      }
    }
  });
}

int gt_hip_prime(gt_hip_context *c, const gt_group_config *cfg) {
  if (!c) return GT_HIP_ERR_ARGUMENT;
  return guarded(c, [&] {
    check_config(cfg);
    GTH_CHECK(hipSetDevice(c->device));
    c->systems = c->state_systems;
    c->sys_off = c->state_systems > 1 ? c->state_off.get() : nullptr;
    if (c->n_state)
      force_eval(c, cfg, c->pos.get(), c->mass.get(), static_cast<uint32_t>(c->n_state),
                 c->acc.get(), c->pot.get(), c->orig.get());
    permute_state(c);
    c->state_forces = true;
  });
}

int gt_hip_steps(gt_hip_context *c, const gt_group_config *cfg, int nsteps, double dt,
                 int use_graph) {
  if (!c) return GT_HIP_ERR_ARGUMENT;
  return guarded(c, [&] {
    check_config(cfg);
    require(nsteps >= 0, "nsteps must be non-negative");
    require(c->state_forces, "forces are not current: call gt_hip_prime first");
    GTH_CHECK(hipSetDevice(c->device));
    const uint64_t n = c->n_state;
    if (n == 0 || nsteps == 0) return;
    // Until the steps have finished the state is between steps: a failure
    // leaves it so, and the forces must be recomputed (gt_hip_prime).
    c->state_forces = false;
    c->systems = c->state_systems;
    c->sys_off = c->state_systems > 1 ? c->state_off.get() : nullptr;
    c->tree_on_state = true;
    if (use_graph) {
      graph_steps(c, cfg, nsteps, dt);
    } else {
      for (int step = 0; step < nsteps; ++step) {
        eager_step(c, cfg, dt);
        if (++c->steps_in_order >= kReorderSteps) permute_state(c);
      }
    }
    GTH_CHECK(hipStreamSynchronize(c->stream));
    c->state_forces = true;
  });
}

int gt_hip_prepare_steps(gt_hip_context *c, const gt_group_config *cfg, int nsteps, double dt) {
  if (!c) return GT_HIP_ERR_ARGUMENT;
  return guarded(c, [&] {
    check_config(cfg);
    require(nsteps >= 0, "nsteps must be non-negative");
    require(c->state_forces, "forces are not current: call gt_hip_prime first");
    GTH_CHECK(hipSetDevice(c->device));
    const uint32_t n = static_cast<uint32_t>(c->n_state);
    if (n == 0 || nsteps == 0) return;
    c->systems = c->state_systems;
    c->sys_off = c->state_systems > 1 ? c->state_off.get() : nullptr;
    // what gt_hip_steps(use_graph = 1) would launch first, made now
    const bool collecting = cfg->predict == 1 && !c->pred.trained;
    if (collecting || !sizes_known(c->hist, n, c->systems)) return;
    prepare_run(c, cfg, plan_for(c, n, cfg), nsteps, dt);
  });
}

int gt_hip_predictor_reset(gt_hip_context *c, int collect) {
  if (!c) return GT_HIP_ERR_ARGUMENT;
  c->pred = Predictor{};
  if (collect > 0) c->pred.collect = static_cast<uint32_t>(collect);
  c->hist.pred_known = false;
  return GT_HIP_OK;
}

int gt_hip_predictor_stats(const gt_hip_context *c, gt_predictor_stats *st) {
  if (!c || !st) return GT_HIP_ERR_ARGUMENT;
  const Predictor &P = c->pred;
  st->samples = P.samples;
  st->trained = P.trained ? 1 : 0;
  st->train_iterations = P.iterations;
  st->train_loss = P.loss;
  st->margin = P.margin;
  st->predicted = P.predicted;
  st->overflows = P.overflows;
  st->mean_abs_log_error = P.err;
  st->slot_entries = P.slot_entries;
  st->list_entries = P.list_entries;
  return GT_HIP_OK;
}

int gt_hip_debug_shrink_slots(gt_hip_context *c, uint32_t shrink) {
  if (!c) return GT_HIP_ERR_ARGUMENT;
  c->debug_shrink = shrink;
  return GT_HIP_OK;
}

int gt_hip_predictor_forward(gt_hip_context *c, uint64_t n, const float *x, const float *params,
                             float *y) {
  if (!c) return GT_HIP_ERR_ARGUMENT;
  return guarded(c, [&] {
    require(n <= 0xffffffffull, "too many rows");
    if (n == 0) return;
    require(x && params && y, "array is NULL");
    GTH_CHECK(hipSetDevice(c->device));
    hipStream_t s = c->stream;
    float *dx = c->tk_in.ensure(kPredFeatures * n), *dp = c->tk_params.ensure(kPredParams);
    float *dy = c->tk_out.ensure(kPredOutputs * n);
    upload(dx, x, kPredFeatures * n, s);
    upload(dp, params, kPredParams, s);
    launch_mlp_forward(dx, static_cast<uint32_t>(n), dp, dy, s);
    download(y, dy, kPredOutputs * n, s);
    GTH_CHECK(hipStreamSynchronize(s));
  });
}

int gt_hip_predictor_train_step(gt_hip_context *c, uint64_t n, const float *samples,
                                const float *stats, float *params, float *m, float *v,
                                uint32_t blocks, uint32_t iter, float lr, float *grad,
                                float *loss) {
  if (!c) return GT_HIP_ERR_ARGUMENT;
  return guarded(c, [&] {
    require(n >= 1 && n <= 0xffffffffull, "n must be 1 .. 2^32 - 1");
    require(blocks >= 1 && blocks <= 4096, "blocks must be 1 .. 4096");
    require(samples && stats && params && m && v && grad && loss, "array is NULL");
    GTH_CHECK(hipSetDevice(c->device));
    hipStream_t s = c->stream;
    const size_t cols = kPredFeatures + kPredOutputs;
    float *ds = c->tk_in.ensure(cols * n), *dst = c->tk_stats.ensure(kPredStats);
    float *dp = c->tk_params.ensure(kPredParams), *dm = c->tk_m.ensure(kPredParams),
          *dv = c->tk_v.ensure(kPredParams);
    float *dg = c->tk_grad.ensure(static_cast<size_t>(blocks) * kPredParams),
          *dl = c->tk_loss.ensure(blocks);
    upload(ds, samples, cols * n, s);
    upload(dst, stats, kPredStats, s);
    upload(dp, params, kPredParams, s);
    upload(dm, m, kPredParams, s);
    upload(dv, v, kPredParams, s);
    launch_train_iteration(ds, static_cast<uint32_t>(n), dst, dp, dm, dv, dg, dl, blocks, iter,
                           lr, s);
    download(params, dp, kPredParams, s);
    download(m, dm, kPredParams, s);
    download(v, dv, kPredParams, s);
    download(grad, dg, static_cast<size_t>(blocks) * kPredParams, s);
    download(loss, dl, blocks, s);
    GTH_CHECK(hipStreamSynchronize(s));
  });
}

int gt_hip_debug_sort_waits(gt_hip_context *c, uint32_t polls) {
  if (!c) return GT_HIP_ERR_ARGUMENT;
  return guarded(c, [&] {
    GTH_CHECK(hipSetDevice(c->device));
    const unsigned long long v = polls;
    upload(c->frame.get() + frame::kSortCtl + 1, &v, 1, c->stream);
    GTH_CHECK(hipStreamSynchronize(c->stream));
  });
}

int gt_hip_debug_valu_rate(gt_hip_context *c, int rsq, double *lane_ops_per_s) {
  if (!c || !lane_ops_per_s) return GT_HIP_ERR_ARGUMENT;
  return guarded(c, [&] {
    GTH_CHECK(hipSetDevice(c->device));
    // eight 256-lane workgroups per CU: eight waves per SIMD
    const uint32_t blocks = static_cast<uint32_t>(c->prop.multiProcessorCount) * 8u;
    const int iters = 20000;
    float *out = c->probe_out.ensure(static_cast<size_t>(blocks) * 256);
    hipStream_t s = c->stream;
    launch_valu_probe(out, blocks, iters / 10, rsq != 0, s);  // warm-up
    double best = 0.0;
    for (int rep = 0; rep < 3; ++rep) {
      GTH_CHECK(hipEventRecord(c->ev[0], s));
      launch_valu_probe(out, blocks, iters, rsq != 0, s);
      GTH_CHECK(hipEventRecord(c->ev[1], s));
      GTH_CHECK(hipEventSynchronize(c->ev[1]));
      const double ms = elapsed_ms(c->ev[0], c->ev[1]);
      if (rep == 0 || ms < best) best = ms;
    }
    *lane_ops_per_s = static_cast<double>(blocks) * 256.0 * 16.0 * iters / (best * 1e-3);
  });
}

int gt_hip_pipeline_stats(const gt_hip_context *c, gt_pipeline_stats *st) {
  if (!c || !st) return GT_HIP_ERR_ARGUMENT;
  *st = c->pstats;
  st->sort_recounts = c->fw[frame::kSortCtl];
  return GT_HIP_OK;
}

int gt_hip_direct_f64_sample(gt_hip_context *c, uint64_t n, const double *pos,
                             const double *mass, double G, double eps, uint64_t ns,
                             const uint64_t *sample, double *acc) {
  if (!c) return GT_HIP_ERR_ARGUMENT;
  return guarded(c, [&] {
    if (ns == 0) return;
    require(pos && mass && sample && acc, "array is NULL");
    for (uint64_t k = 0; k < ns; ++k) require(sample[k] < n, "sample index out of range");
    GTH_CHECK(hipSetDevice(c->device));
    hipStream_t s = c->stream;
    double *dp = c->in_pos.ensure(3 * n), *dm = c->in_mass.ensure(n);
    uint64_t *di = c->sample_idx.ensure(ns);
    double *da = c->sample_acc.ensure(3 * ns);
    upload(dp, pos, 3 * n, s);
    upload(dm, mass, n, s);
    upload(di, sample, ns, s);
    launch_direct_f64_sample(dp, dm, n, G, eps, di, ns, da, s);
    download(acc, da, 3 * ns, s);
    GTH_CHECK(hipStreamSynchronize(s));
  });
}

}  // extern "C"
