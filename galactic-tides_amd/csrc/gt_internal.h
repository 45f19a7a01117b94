// Internal declarations shared by the HIP sources of libgt_hip.so.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <stdexcept>
#include <string>

#include "gt_hip.h"
#include "gt_walk_rec.h"

namespace gth {

// ---- errors -------------------------------------------------------------------

// Carries a gt_hip status code to the C-ABI layer, which turns it into the
// return value and the context's last-error message.
struct Error : std::runtime_error {
  int code;
  Error(int c, const std::string &m) : std::runtime_error(m), code(c) {}
};

[[noreturn]] inline void fail(int code, const std::string &msg) { throw Error(code, msg); }

#define GTH_CHECK(call)                                                                  \
  do {                                                                                   \
    hipError_t gth_e_ = (call);                                                          \
    if (gth_e_ != hipSuccess)                                                            \
      ::gth::fail(gth_e_ == hipErrorOutOfMemory ? GT_HIP_ERR_MEMORY : GT_HIP_ERR_DEVICE, \
                  std::string(#call) + ": " + hipGetErrorString(gth_e_));                \
  } while (0)

#define GTH_LAUNCH_CHECK() GTH_CHECK(hipGetLastError())

// ---- device buffers -----------------------------------------------------------

// Bumped by every device (re)allocation of a DBuf: a captured graph holds raw
// pointers, so it is valid only while this has not changed since its capture.
inline uint64_t alloc_generation = 0;

// Device array that keeps its storage between calls and only grows
// (geometrically), so steady-state steps allocate nothing.
template <class T>
class DBuf {
 public:
  DBuf() = default;
  DBuf(const DBuf &) = delete;
  DBuf &operator=(const DBuf &) = delete;
  ~DBuf() {
    if (p_) (void)hipFree(p_);
  }
  T *ensure(size_t n) {
    if (n > cap_) {
      size_t c = cap_ + cap_ / 2;
      if (c < n) c = n;
      if (c < 64) c = 64;
      if (p_) GTH_CHECK(hipFree(p_));
      p_ = nullptr;
      cap_ = 0;
      GTH_CHECK(hipMalloc(reinterpret_cast<void **>(&p_), c * sizeof(T)));
      cap_ = c;
      ++alloc_generation;
    }
    return p_;
  }
  T *get() const { return p_; }
  size_t capacity() const { return cap_; }

 private:
  T *p_ = nullptr;
  size_t cap_ = 0;
};

// ---- device-side layouts -------------------------------------------------------

// kern::GTCell / GTGroup / GTParams with HIP vector types.
struct Cell {
  float4 cm;  // com - origin, mass
  float4 q0;  // xx yy zz xy
  float4 q1;  // xz yz 0 0
};
static_assert(sizeof(Cell) == 48, "GTCell layout");
static_assert(sizeof(gt_cell) == 48 && sizeof(gt_group) == 32 && sizeof(gt_params) == 32 &&
                  sizeof(gt_float4) == 16,
              "reference struct layouts");

// One node as the group walk reads it (64 bytes, half a cache line): FP64
// centre, centre of mass and opening distance, and `bits` packing the first
// child (bits 0-25), the skip node (26-51, all ones = none), the child count
// (52-55), the depth (56-61; half = root half / 2^depth) and whether the node
// holds at most kBodiesIfAtMost bodies (62).
struct WalkRec {
  double cx, cy, cz;
  double mx, my, mz;
  double crit;
  uint64_t bits;
};
static_assert(sizeof(WalkRec) == 64, "WalkRec layout");
constexpr uint32_t kRecIndexMask = (1u << 26) - 1;  // node indices in WalkRec::bits
constexpr int kWalkExact = 0;     // lists written at scanned offsets
constexpr int kWalkCapacity = 1;  // lists written into fixed per-group slots

constexpr uint32_t kNone = 0xffffffffu;
constexpr int kGeomStride = 16;  // doubles per node in the host-compatible geometry rows
constexpr int kMaxGroup = 64;    // targets per group: one wave64
constexpr uint32_t kBodiesIfAtMost = 2;  // accepted cells this small are listed as bodies

__device__ inline uint64_t pack_walk_bits(uint32_t first_child, uint32_t skip, uint32_t nch,
                                          uint32_t depth, bool small) {
  return static_cast<uint64_t>(first_child & kRecIndexMask) |
         (static_cast<uint64_t>(skip & kRecIndexMask) << 26) |
         (static_cast<uint64_t>(nch & 15u) << 52) | (static_cast<uint64_t>(depth & 63u) << 56) |
         (static_cast<uint64_t>(small ? 1u : 0u) << 62);
}

// The walk's and the evaluation's records of a tree (reference: the FP32
// packing of build_group_lists, src/group_lists.cpp:40-63, and its walk
// nodes, :59-61): per node the FP64 walk record, its FP32 image, its body
// span and its FP32 cell (with the low part of its centre of mass for hi/lo
// positions); per body its FP32 position and mass (and low part).  Positions
// are relative to the system's root centre.  Written by prep_lists, or by the
// tree build itself for a tree it builds (cells null: none).
struct ListRecords {
  double inv_theta = 0.0;
  WalkRec *recs = nullptr;
  WalkRec32 *images = nullptr;
  uint2 *span = nullptr;
  Cell *cells = nullptr;
  float4 *cells_lo = nullptr;    // (null: no hi/lo parts)
  float4 *bodies = nullptr;
  float4 *bodies_lo = nullptr;   // (null: no hi/lo parts)
};

// Group starts flagged on tree positions by the tree build (emit_levels
// clears the n flags, node_links flags each node's starts): group size,
// packing, flags, and (may be null) the node each start is taken from.
struct GroupFlags {
  uint32_t gs = 0;
  int pack = 1;
  uint32_t *flags = nullptr;
  uint32_t *node_of = nullptr;
};

// The group starts node x contributes (reference: visit/chunk,
// src/group_lists.cpp:65-113, as flags on tree positions):
// every node holding more than group_size bodies, and every root, is one the
// reference's depth-first visit reaches; a small one or a leaf is chunked
// (starts every group_size bodies), a larger one flags the starts of its runs
// of small children (packed greedily while they fit) and leaves its large
// children to flag their own.  L: x's links (end known); child(c) gives child
// c's links.  flags must start cleared.
template <typename ChildLinks>
__device__ inline void flag_group_starts(uint32_t x, int4 L, bool root, uint32_t gs, int pack,
                                         uint32_t *__restrict__ flags,
                                         uint32_t *__restrict__ node_of, ChildLinks child) {
  const uint32_t b = static_cast<uint32_t>(L.z), e = static_cast<uint32_t>(L.w);
  // not reached by visit(): a small child of its parent (roots always are)
  if (!root && e - b <= gs) return;
  if (e - b <= gs || L.y == 0) {  // chunk(begin, end)
    for (uint32_t s = b; s < e; s += gs) {
      flags[s] = 1u;
      if (node_of) node_of[s] = x;
    }
    return;
  }
  bool open = false;
  uint32_t run = 0, run_start = 0;
  for (int c = 0; c < L.y; ++c) {
    const int4 C = child(c);
    const uint32_t cc = static_cast<uint32_t>(C.w - C.z);
    if (cc <= gs) {
      if (pack && open && run + cc <= gs) {
        run += cc;
        if (node_of) node_of[run_start] = x;  // siblings packed: their parent holds them
      } else {
        flags[C.z] = 1u;
        if (node_of) node_of[C.z] = static_cast<uint32_t>(L.x + c);
        run = cc;
        run_start = static_cast<uint32_t>(C.z);
        open = true;
      }
    } else {
      open = false;  // visited in its own right: it flags its own starts
    }
  }
}

#pragma clang fp contract(off)

// Node k's records from its geometry row g (centre, half, centre of mass,
// mass, quadrupole, delta: the first 15 doubles of a geometry row), links L,
// skip and depth; o: its system's root centre and half size.  A half size
// that is not the root's halved depth times is reported in walk_error (bit 2).
__device__ inline void write_node_records(const ListRecords &R, uint32_t k, const double *g,
                                          int4 L, uint32_t skip, uint32_t depth, double ox,
                                          double oy, double oz, double root_half,
                                          unsigned long long *walk_error) {
  Cell c;
  const double mx = g[4] - ox, my = g[5] - oy, mz = g[6] - oz;
  c.cm = make_float4(static_cast<float>(mx), static_cast<float>(my), static_cast<float>(mz),
                     static_cast<float>(g[7]));
  if (R.cells_lo)
    R.cells_lo[k] = make_float4(static_cast<float>(mx - c.cm.x), static_cast<float>(my - c.cm.y),
                                static_cast<float>(mz - c.cm.z), 0.0f);
  c.q0 = make_float4(static_cast<float>(g[8]), static_cast<float>(g[9]),
                     static_cast<float>(g[10]), static_cast<float>(g[11]));
  c.q1 = make_float4(static_cast<float>(g[12]), static_cast<float>(g[13]), 0.0f, 0.0f);
  R.cells[k] = c;
  // The packed record keeps half as root_half / 2^depth: check that it is.
  if (depth > 63 || ldexp(root_half, -static_cast<int>(depth)) != g[3])
    atomicOr(walk_error, 2ull);
  WalkRec w;
  w.cx = g[0];
  w.cy = g[1];
  w.cz = g[2];
  w.mx = g[4];
  w.my = g[5];
  w.mz = g[6];
  w.crit = 2.0 * g[3] * R.inv_theta + g[14];
  const bool small = static_cast<uint32_t>(L.w - L.z) <= kBodiesIfAtMost;
  w.bits = pack_walk_bits(static_cast<uint32_t>(L.x), skip, static_cast<uint32_t>(L.y), depth,
                          small);
  R.recs[k] = w;
  R.images[k] = make_rec32(w.cx, w.cy, w.cz, w.mx, w.my, w.mz, w.crit, ox, oy, oz, root_half,
                           image_meta(static_cast<uint32_t>(L.x), static_cast<uint32_t>(L.y)),
                           small);
  R.span[k] = make_uint2(static_cast<uint32_t>(L.z), static_cast<uint32_t>(L.w));
}

// Body k's FP32 position and mass relative to the root centre o (p: its FP64
// position, m its mass).
__device__ inline void write_body_record(const ListRecords &R, uint32_t k, double x, double y,
                                         double z, double m, double ox, double oy, double oz) {
  const double px = x - ox, py = y - oy, pz = z - oz;
  const float hx = static_cast<float>(px), hy = static_cast<float>(py),
              hz = static_cast<float>(pz);
  R.bodies[k] = make_float4(hx, hy, hz, static_cast<float>(m));
  if (R.bodies_lo)
    R.bodies_lo[k] = make_float4(static_cast<float>(px - hx), static_cast<float>(py - hy),
                                 static_cast<float>(pz - hz), 0.0f);
}

#pragma clang fp contract(fast)

// ---- XCD-aware block order ---------------------------------------------------------

// Workgroups of a launch are dealt to the 8 XCDs in turn.  Launching
// xcd_grid(n) blocks and renumbering with xcd_block gives XCD x the logical
// blocks [x * q, (x + 1) * q), q = ceil(n / 8): a contiguous run of the work,
// so neighbouring groups (which read the same tree nodes and cells) share an
// XCD's L2.  Logical blocks >= n must exit.
constexpr uint32_t kXcds = 8;
inline uint32_t xcd_grid(uint32_t n) { return (n + kXcds - 1) / kXcds * kXcds; }
__device__ inline uint32_t xcd_block(uint32_t b, uint32_t grid) {
  return (b % kXcds) * (grid / kXcds) + b / kXcds;
}

// The same when only `used` of the launched blocks have work, a number known on
// the device alone (the grid was sized from a host-side bound): the work is
// still split evenly over the XCDs.  Returns kNone for blocks with no work.
// Needs a grid of at least xcd_grid(used) blocks.
__device__ inline uint32_t xcd_block_used(uint32_t b, uint32_t used) {
  const uint32_t q = (used + kXcds - 1) / kXcds, i = b / kXcds;
  const uint32_t logical = (b % kXcds) * q + i;
  return i < q && logical < used ? logical : 0xffffffffu;
}

// ---- device-side counts of one evaluation -------------------------------------

// The counts that later launches of an evaluation depend on (nodes, groups,
// list lengths) stay in device memory, in a frame of 64-bit words, so an
// evaluation is enqueued with no host round trip and can be captured in a
// graph.  Grids are sized from host-side bounds (Plan); kernels read the
// true count and idle past it.  A bound that proves too small sets a status
// bit; kernels whose inputs would then be incomplete see a count of zero or
// return at once, and the host runs the evaluation again with larger bounds.
namespace frame {
constexpr int kStatus = 0;      // kAbort* bits
constexpr int kNodes = 1;       // nodes the tree needs (scan total)
constexpr int kUnresolved = 2;  // bit 0: a leaf needs more octant-path words
constexpr int kGroups = 3;      // groups (scan total)
constexpr int kDepth = 4;       // level of the deepest node
constexpr int kWalkError = 5;   // bit 0: malformed tree; bit 1: half != root half / 2^depth
constexpr int kStats = 6;       // 8 words, see list_stats
constexpr int kItems = 14;      // evaluation work items (groups, long lists in parts)
constexpr int kSplitParts = 15; // parts of the groups whose lists were split
constexpr int kSlotOverflow = 16;  // a group's lists outgrew their slots (the walk sets it)
constexpr int kPredCells = 17;  // sum of the predicted cell-list slots
constexpr int kPredBodies = 18; // sum of the predicted body-list slots
constexpr int kPredError = 19;  // 1e6 x sum over groups of |log(1 + length) - prediction|
constexpr int kExactTests = 20; // opening tests the FP32 filter left to the FP64 test
constexpr int kCountsEnd = 21;  // [kNodes, kCountsEnd): cleared before each evaluation
constexpr int kSortCtl = 22;    // 2 words: the radix sorts' ctl (recount count, test setting)
constexpr int kStepsDone = 30;  // leapfrog steps completed since the counter was cleared
constexpr int kStepInteractions = 31;  // their cell + body interactions
constexpr int kLevelNodes = 32;  // kLevels words: nodes on each tree level 0 .. kLevels - 1
constexpr int kLevels = 61;
constexpr int kWords = kLevelNodes + kLevels;
}  // namespace frame

constexpr uint32_t kAbortNodes = 1u;   // more nodes than Plan::node_cap
constexpr uint32_t kAbortPaths = 2u;   // more octant-path words needed than Plan::words
constexpr uint32_t kAbortDepth = 4u;   // deeper than Plan::depth_cap
constexpr uint32_t kAbortGroups = 8u;  // more groups than Plan::group_cap
constexpr uint32_t kAbortLists = 16u;  // a group's lists do not fit the slots
constexpr uint32_t kAbortWalk = 32u;   // malformed tree
constexpr uint32_t kAbortItems = 64u;  // more work items or split parts than the plan's bounds
constexpr uint32_t kAbortSlots = 128u; // predicted slots add up to more than the list buffers
constexpr uint32_t kAbortInput = 256u; // caller-supplied device inputs failed their check

// Host-side bounds of one evaluation, fixed before it is enqueued.
struct Plan {
  uint32_t n = 0;          // particles
  uint32_t systems = 1;    // independent systems (a batch), each with its own tree
  int words = 1;           // 21-level octant-path words computed per particle
  // Levels of the path the sort orders by when words == 1 (at most 21): the
  // deepest leaf's depth and some headroom; a deeper leaf is reported like one
  // past the last word.
  int path_levels = 21;
  uint32_t node_cap = 0;   // node arrays
  int depth_cap = 60;      // deepest level whose moments are computed
  uint32_t group_cap = 0;  // groups covered by the walk and evaluation grids
  uint32_t cap_c = 0, cap_b = 0;  // per-group list slots (0: lists always overflow)
  uint32_t split_len = 0;   // list entries per evaluation part (0: no work items)
  uint32_t item_cap = 0;    // evaluation work items (split_len > 0)
  uint32_t split_cap = 0;   // parts of split groups (split_len > 0)
  bool longest_first = false;  // each XCD's items by decreasing cost
  int cost_buckets = 0;        // > 0: by decreasing cost class (1/cost_buckets octave)
  bool hilo = false;           // hi/lo positions relative to group centres
  // Predicted per-group slots (instead of cap_c / cap_b for every group): the
  // list buffers' sizes, which the predicted slots must add up to at most.
  bool predicted = false;
  uint64_t slot_total_c = 0, slot_total_b = 0;
  uint32_t chunks = 1;  // walk ranges whose evaluation runs alongside the walk
  // Workgroups of each level's moments launch, sized from the previous tree's
  // level populations (levels_sized; else kMomentBlocks for every level).
  bool levels_sized = false;
  uint16_t level_blocks[frame::kLevels] = {};
};
// The persistent evaluation's per-XCD item counters sit this many uint32 apart
// (128 bytes: each on a cache line of its own, so the XCDs' atomics do not
// contend for one line).
constexpr uint32_t kQueueStride = 32;
constexpr uint16_t kMomentBlocks = 1024;
// Levels sized at most this many workgroups get no launch of their own: runs
// of them are done by one workgroup, level after level.
constexpr uint16_t kSerialLevelBlocks = 2;
// Two consecutive levels sized at most this many workgroups share a launch.
constexpr uint16_t kPairLevelBlocks = 64;

// A device count that must not exceed cap; past it (an overflow the status
// reports) the count reads as zero, so dependent kernels do nothing.
__device__ inline uint32_t bounded_count(const unsigned long long *c, uint32_t cap) {
  const unsigned long long v = *c;
  return v <= cap ? static_cast<uint32_t>(v) : 0u;
}

// ---- launchers (defined in the .hip sources) -------------------------------------

// gt_scan.hip: exclusive prefix sum of n uint32 values into out (may alias in),
// total written to *total (device).  tmp must hold scan_tmp_words(n) words.
// compact (may be null; in holds 0/1 flags, out must not alias in): also
// compact[out[p]] = p for every flagged p.
size_t scan_tmp_words(size_t n);
void exclusive_scan_u32(const uint32_t *in, uint32_t *out, size_t n, unsigned long long *total,
                        unsigned long long *tmp, hipStream_t s, uint32_t *compact = nullptr);

// gt_sort.hip: stable LSD radix sort of (key, value) pairs on key bits [0, bits).
// The sorted pairs end in (keys, vals); keys_alt/vals_alt are scratch of n
// entries; tmp must hold radix_tmp_words(n) 64-bit words.  ctl (may be null):
// two device words, [0] counts digit totals recounted because a bounded wait
// gave up (the result is right either way), [1] a test setting (0 normally;
// k > 0: every such wait gives up after k polls).
size_t radix_tmp_words(size_t n);
void radix_sort_u64(uint64_t *keys, uint32_t *vals, uint64_t *keys_alt, uint32_t *vals_alt,
                    size_t n, int bits, unsigned long long *tmp, hipStream_t s,
                    unsigned long long *ctl = nullptr);
// The exclusive scan of a row-major table of `rows` (at most 256) rows of
// ntiles counts, the radix sort's digit scan: offsets[r * ntiles + t] = all
// counts of rows before r plus row r's counts of tiles before t.  ticket (one
// word) and status (rows words) must be zero and epoch nonzero; ctl as above.
void launch_row_scan(uint32_t rows, uint32_t ntiles, uint32_t epoch, uint32_t *ticket,
                     unsigned long long *status, unsigned long long *ctl, const uint32_t *counts,
                     uint32_t *offsets, hipStream_t s);
// The same for the first bounded_count(n_dev, cap) pairs of arrays of cap entries.
void radix_sort_u64_dev(uint64_t *keys, uint32_t *vals, uint64_t *keys_alt, uint32_t *vals_alt,
                        const unsigned long long *n_dev, size_t cap, int bits,
                        unsigned long long *tmp, hipStream_t s,
                        unsigned long long *ctl = nullptr);

// gt_eval.hip
// Long lists evaluated in parts (fast evaluation only): per group its part
// count and its parts' first slot among split groups' parts; the work items;
// one kMaxGroup float4 slot of partial sums per part of a split group.
// With longest_first, each XCD evaluates its share of the items in order of
// decreasing cost (a stable sort of item indices into `order`; keys and the
// *_alt arrays hold item_cap entries, sort_tmp radix_tmp_words(item_cap)).
struct EvalSplit {
  uint32_t split_len = 0;  // list entries per part
  uint32_t item_cap = 0, split_cap = 0;
  uint32_t *parts = nullptr, *split_count = nullptr, *item_off = nullptr, *split_off = nullptr;
  uint2 *items = nullptr;
  float4 *partial = nullptr;
  bool longest_first = false;
  int cost_buckets = 0;  // > 0: order by cost classes of 1/cost_buckets octave
  uint64_t *keys = nullptr, *keys_alt = nullptr;
  uint32_t *order = nullptr, *order_alt = nullptr;
  unsigned long long *sort_tmp = nullptr;
};

// Work items for the groups' lists (frame[kItems], frame[kSplitParts]);
// sets kAbortItems past split.item_cap / split.split_cap.
// (Groups g_begin .. g_begin + g_span, g_span 0: all; the per-group arrays of
// split are indexed from g_begin.)
void launch_eval_items(const gt_group *groups, unsigned long long *frame, uint32_t group_cap,
                       const EvalSplit &split, unsigned long long *scan_tmp, hipStream_t s,
                       uint32_t g_begin = 0, uint32_t g_span = 0);
// Evaluates the first frame[kGroups] groups (at most group_cap) unless
// frame[kStatus] is set; with split (fast evaluation), the work items that
// launch_eval_items made (split groups' parts then need launch_combine_parts);
// without, lists longer than inline_split entries (0: none) are evaluated in
// the same parts by one wave each, with the same results.
// Hi/lo precision mode: lo parts of the bodies' and cells' FP32 positions
// (prep_lists), and each group's centre as hi, lo float4 pairs.
struct EvalHiLo {
  const float4 *bodies_lo = nullptr, *cells_lo = nullptr, *group_hilo = nullptr;
};
void launch_group_centres(const gt_group *groups, const unsigned long long *frame,
                          uint32_t group_cap, const double *tpos, const uint32_t *tsid,
                          const double *roots, const double *geom, float4 *group_hilo,
                          hipStream_t s);
// gt_noalloc.hip: the evaluation scratch of gt_hip_tree_noalloc, a fixed
// device array (noalloc_scratch_bytes()); noalloc_scratch(device) is its
// address on that device (looked up once per device; call with it current).
// noalloc_tree checks the caller's device arrays on the device, evaluates them
// into out, waits, and fails with GT_HIP_ERR_ARGUMENT when the check did;
// t0 / t1 are recorded around the evaluation.
void *noalloc_scratch(int device);
size_t noalloc_scratch_bytes();
void noalloc_tree(void *scratch_ptr, const float4 *bodies, uint32_t n, const Cell *cells,
                  uint64_t n_cells, const uint32_t *cell_list, uint64_t n_cell_list,
                  const uint32_t *body_list, uint64_t n_body_list, const gt_group *groups,
                  uint32_t n_groups, float G, float eps2, uint32_t quad, bool exact,
                  float4 *out, hipStream_t s, hipEvent_t t0, hipEvent_t t1);

void launch_eval(const float4 *bodies, const Cell *cells, const uint32_t *cell_list,
                 const uint32_t *body_list, const gt_group *groups,
                 const unsigned long long *frame, uint32_t group_cap, float G, float eps2,
                 uint32_t quad, bool exact, int waves, float4 *out, hipStream_t s,
                 const EvalSplit *split = nullptr, const EvalHiLo *hilo = nullptr,
                 uint32_t g_begin = 0, uint32_t g_span = 0, uint32_t *queue = nullptr,
                 uint32_t persistent_blocks = 0, uint32_t inline_split = 0);
// Adds the parts of split groups into out, in part order.
void launch_combine_parts(const gt_group *groups, const unsigned long long *frame,
                          uint32_t group_cap, const EvalSplit &split, float4 *out, hipStream_t s,
                          uint32_t g_begin = 0, uint32_t g_span = 0);

// gt_direct.hip
// The VALU throughput probe: blocks x 256 lanes, iters x 16 operations each
// (out: blocks x 256 floats).
void launch_valu_probe(float *out, uint32_t blocks, int iters, bool rsq, hipStream_t s);
void launch_direct(const float4 *bodies, uint32_t n, float G, float eps2, float4 *out,
                   hipStream_t s);
void launch_direct_f64_sample(const double *pos, const double *mass, uint64_t n, double G,
                              double eps, const uint64_t *sample, uint64_t ns, double *acc,
                              hipStream_t s);

// gt_walk.hip
void launch_fill_u32(uint32_t *p, size_t n, uint32_t v, hipStream_t s);
void launch_fill_u64(unsigned long long *p, size_t n, unsigned long long v, hipStream_t s);
// Kernels over nodes take the node count as frame + kNodes (bounded by the
// node capacity the grid was sized for), over groups as frame + kGroups.
void launch_link_up(const int4 *lnk, const unsigned long long *frame, uint32_t node_cap,
                    uint32_t *parent, uint32_t *skip, uint32_t *depth, hipStream_t s);
// Batches of independent systems pass tsid (system of each tree position),
// roots (4 doubles per system: root centre, half) and root_node (root node
// of each system); for one system they are null (the root is node 0).
// The records go to rec's arrays (bodies_lo / cells_lo may be null: no lo
// parts of the FP32 positions).
void launch_prep_lists(const double *geom, const int4 *lnk, const uint32_t *skip,
                       const uint32_t *depth, const double *tpos, const double *tmass, uint32_t n,
                       unsigned long long *frame, uint32_t node_cap, const uint32_t *tsid,
                       const double *roots, const ListRecords &rec, hipStream_t s);
// node_of (may be null): at each flagged start, the node the group is taken
// from (the chunked node, a lone packed child, or the parent of packed
// siblings); make_groups copies it per group into group_node (may be null).
void launch_group_starts(const int4 *lnk, const uint32_t *parent, const unsigned long long *frame,
                         uint32_t node_cap, uint32_t n, uint32_t gs, int pack, uint32_t *flags,
                         hipStream_t s, uint32_t *node_of = nullptr);
// starts: the flagged positions in order (the compaction of the flags scan).
void launch_make_groups(uint32_t n, const unsigned long long *frame, uint32_t group_cap,
                        const uint32_t *tsid, const uint32_t *root_node, const uint32_t *starts,
                        gt_group *groups, uint32_t *group_root, hipStream_t s,
                        const uint32_t *node_of = nullptr, uint32_t *group_node = nullptr);
// group_root: the node each group's walk starts from (null: node 0); geom
// gives the root half sizes.  Coherent walks (lanes > 1) need a single tree.
// Capacity mode: per-group slots of cap_c / cap_b entries at g * cap, or (with
// slot lengths given) slot_c[g] / slot_b[g] entries at cell_off[g] /
// body_off[g]; a group whose lists do not fit sets frame[kSlotOverflow] and
// its recorded ranges are cut at its slots.  Any status set before the walk
// (kAbortSlots) stops it.  Groups g_begin .. g_begin + g_span only (g_span 0:
// all); coherent walks (lanes > 2) take all groups.
// lanes: 3 (8 lanes per group) or 4 (4 lanes per group) the walk over child
// sets (images, levels_cap levels of stack per group; 3 is the default), 2 one
// record per step, 1 the same fetching the next records ahead, 16/32/64
// coherent walks.
void launch_walk(int mode, int lanes, const WalkRec *recs, const WalkRec32 *images,
                 uint32_t levels_cap, const uint2 *span, const double *tpos,
                 gt_group *groups, unsigned long long *frame, uint32_t group_cap,
                 uint32_t node_cap, const double *geom, const uint32_t *group_root, uint32_t cap_c,
                 uint32_t cap_b, const uint32_t *cell_off, const uint32_t *body_off,
                 uint32_t *cell_count, uint32_t *body_count, uint32_t *cell_list,
                 uint32_t *body_list, hipStream_t s, const uint32_t *slot_c = nullptr,
                 const uint32_t *slot_b = nullptr, uint32_t g_begin = 0, uint32_t g_span = 0,
                 float4 *group_hilo = nullptr);
// The stats words must start cleared (they are among the counts every
// evaluation clears first).  check (may be null): also set the status as
// launch_check_frame(frame, *check, true) does, in the same launch.
void launch_list_stats(const uint32_t *cell_count, const uint32_t *body_count,
                       const gt_group *groups, unsigned long long *frame, uint32_t group_cap,
                       hipStream_t s, const Plan *check = nullptr);
// Sets frame[kStatus] from the counts against the plan's bounds (list slots
// are checked only when check_slots).
void launch_check_frame(unsigned long long *frame, const Plan &plan, bool check_slots,
                        hipStream_t s);

// gt_treegrav.hip: the per-particle FP64 tree solver over a built tree (one
// system, root node 0; skip and depth from launch_link_up): acc (N x 3) and
// pot in original order, stats[0..1] = cell and particle interactions.
void launch_tree_gravity(const double *geom, const int4 *lnk, const uint32_t *skip,
                         const uint32_t *depth, const double *tpos, const double *tmass,
                         const uint32_t *order, uint32_t n, const unsigned long long *frame,
                         uint32_t node_cap, double G, double eps, double theta, int quadrupole,
                         double *acc, double *pot, unsigned long long *stats, hipStream_t s,
                         const double *points = nullptr, uint32_t n_points = 0);

// gt_neighbours.hip: fixed-radius queries over a built tree (one system, root
// node 0).  launch_nb_records packs each node's cube, children, body range
// and skip link into `records` (nb_record_bytes(node_cap) bytes).  Query q
// has centre centres[3q..3q+2] and radius radii[q * radius_stride]; lane k <
// lanes takes query perm[k].  offs null: counts[q] =
// its neighbours.  Otherwise its neighbours (original index, squared
// distance) go to index / dist2 from offs[q], counts[q] of them.  error: bit 0
// malformed tree, bit 1 the passes disagree or perm is out of range.
size_t nb_record_bytes(uint32_t node_cap);
void launch_nb_records(const double *geom, const int4 *lnk, const uint32_t *skip,
                       const uint32_t *depth, const unsigned long long *frame, uint32_t node_cap,
                       void *records, hipStream_t s);
void launch_within(const void *records, const double *tpos, const uint32_t *order, uint32_t n,
                   const unsigned long long *frame, uint32_t node_cap, const double *centres,
                   const double *radii, uint32_t radius_stride, uint32_t nq, uint32_t lanes,
                   const uint32_t *perm, uint32_t *counts, const uint32_t *offs,
                   unsigned long long *index, double *dist2, uint32_t stage,
                   unsigned long long *error, hipStream_t s);
// stage > 0 (launch_within): the first `stage` neighbours of query q go to
// index / dist2 from q * stage and counts[q] gets all of them; launch_unstage
// then moves every list of at most `stage` entries to its CSR place (offs)
// and appends the other queries to `over` (count in *n_over).
void launch_unstage(const uint32_t *counts, const uint32_t *offs, uint32_t nq, uint32_t stage,
                    const unsigned long long *s_index, const double *s_dist2,
                    unsigned long long *index, double *dist2, uint32_t *over,
                    unsigned long long *n_over, hipStream_t s);

// Sort keys of the queries: 30-bit Morton keys of the centres in the root cube
// of the tree whose node 0 is geom's first row (0 for an empty tree), and
// vals[q] = q.
void launch_query_keys(const double *geom, const unsigned long long *frame,
                       const double *centres, uint32_t nq, uint64_t *keys, uint32_t *vals,
                       hipStream_t s);
constexpr int kQueryKeyBits = 30;

// gt_step.hip: each does nothing once status (a frame's kStatus word) is set.
// acc/pot rows i from the tree-ordered results out[inv[i]] (inv: each
// particle's tree position).
void launch_scatter(const float4 *out, const uint32_t *inv, uint32_t n, double *acc,
                    double *pot, const unsigned long long *status, hipStream_t s);
// launch_scatter then launch_kick(vel, acc, n, h) as one kernel; with
// step_frame (the frame) also launch_count_step's count, in the same launch.
void launch_scatter_kick(const float4 *out, const uint32_t *inv, uint32_t n, double *acc,
                         double *pot, double *vel, double h, const unsigned long long *status,
                         hipStream_t s, unsigned long long *step_frame = nullptr);
void launch_kick_drift(double *pos, double *vel, const double *acc, size_t n, double h,
                       double dt, const unsigned long long *status, hipStream_t s);
// Permutations of the resident state: out[k] = in[order[k]] (rows of w
// doubles); out[idx[k]] = in[k]; out[k] = a[b[k]] (a or b null: the identity).
void launch_gather_rows(const double *in, const uint32_t *order, uint64_t n, int w, double *out,
                        hipStream_t s);
void launch_scatter_rows(const double *in, const uint32_t *idx, uint64_t n, int w, double *out,
                         hipStream_t s);
void launch_compose_u32(const uint32_t *a, const uint32_t *b, uint64_t n, uint32_t *out,
                        hipStream_t s);
void launch_kick(double *vel, const double *acc, size_t n, double h,
                 const unsigned long long *status, hipStream_t s);
// Unless frame[kStatus] is set: frame[kStepsDone] += 1 and the step's
// interactions are added to frame[kStepInteractions].
void launch_count_step(unsigned long long *frame, hipStream_t s);
// Per system (offsets: systems + 1 particle offsets, or null for one system
// of n): 14 sums each into total (K, W, P, L, sum m x, M, 2 unused); partial
// holds diag_partial_words(systems) doubles.
size_t diag_partial_words(uint32_t systems);
void launch_diagnostics(const double *pos, const double *vel, const double *mass,
                        const double *pot, const unsigned long long *offsets, uint32_t systems,
                        uint64_t n, double *partial, double *total, hipStream_t s);

}  // namespace gth
