// Interaction lists on the GPU (reference: build_group_lists,
// src/group_lists.cpp:18-207).
//
//  link_up       parent of every node, then each node's depth and `skip`: the
//                node a LIFO depth-first walk pops after finishing the node's
//                subtree (previous sibling, else the parent's skip).  With it
//                the walk needs no stack: open = go to the last child, done =
//                go to skip.
//  prep_lists    FP32 bodies and cells relative to the root centre (:40-63),
//                and one 64-byte walk record per node: FP64 centre, centre of
//                mass and crit = 2 half / theta + delta, plus one word packing
//                first child, skip, child count, depth and "at most two bodies".
//                A node visit of the walk reads exactly one record (half a
//                cache line); half is root_half / 2^depth, which is exactly the
//                reference's repeated halving.
//  group_starts  group selection (:65-113) as flags on tree positions: every
//                node holding more than group_size bodies is one the
//                reference's depth-first `visit` reaches, and each such node
//                marks where its runs of small children (packed greedily while
//                they fit) and its chunks start.  A scan turns the flags into
//                groups in tree order.
//  walk<MODE>    one thread per group: group bounding box, then the depth-
//                first walk with the FP64 opening test (:147-154) evaluated
//                in the reference's operation order with no FMA contraction,
//                so accept/open decisions, and hence the lists, are identical.
//                MODE kWalkCapacity writes each group's lists into fixed-size
//                slots (sized from the previous evaluation) and counts; groups
//                that do not fit are caught by the counts, and then the lists
//                are rewritten by MODE kWalkExact at scanned offsets.
//  list_stats    interaction counts, longest list and largest group (:187-206).
#include "gt_internal.h"

namespace gth {
namespace {

constexpr int kBlock = 256;
constexpr int kWalkBlock = 64;

inline uint32_t blocks_for(size_t n, int b = kBlock) {
  return static_cast<uint32_t>((n + b - 1) / b);
}

__global__ __launch_bounds__(kBlock) void set_parents(const int4 *__restrict__ lnk,
                                                      const unsigned long long *__restrict__ frame,
                                                      uint32_t node_cap,
                                                      uint32_t *__restrict__ parent) {
  const uint32_t n_nodes = bounded_count(frame + frame::kNodes, node_cap);
  const uint32_t p = blockIdx.x * kBlock + threadIdx.x;
  if (p >= n_nodes) return;
  const int4 L = lnk[p];  // (roots keep the kNone the array was filled with)
  for (int c = 0; c < L.y; ++c) parent[L.x + c] = p;
}

__global__ __launch_bounds__(kBlock) void set_skip_depth(const int4 *__restrict__ lnk,
                                                         const uint32_t *__restrict__ parent,
                                                         const unsigned long long *__restrict__ frame,
                                                         uint32_t node_cap,
                                                         uint32_t *__restrict__ skip,
                                                         uint32_t *__restrict__ depth) {
  const uint32_t n_nodes = bounded_count(frame + frame::kNodes, node_cap);
  const uint32_t x0 = blockIdx.x * kBlock + threadIdx.x;
  if (x0 >= n_nodes) return;
  uint32_t d = 0;
  for (uint32_t x = x0; parent[x] != kNone && d <= 64; x = parent[x]) ++d;
  depth[x0] = d;
  // Climb while x is the first child of its parent: its subtree ends its parent's.
  uint32_t x = x0;
  for (;;) {
    const uint32_t p = parent[x];
    if (p == kNone) {
      skip[x0] = kNone;
      return;
    }
    if (static_cast<uint32_t>(lnk[p].x) != x) {
      skip[x0] = x - 1;
      return;
    }
    x = p;
  }
}

#pragma clang fp contract(off)

__global__ __launch_bounds__(kBlock) void prep_lists(
    const double *__restrict__ geom, const int4 *__restrict__ lnk,
    const uint32_t *__restrict__ skip, const uint32_t *__restrict__ depth,
    const double *__restrict__ tpos, const double *__restrict__ tmass, uint32_t n,
    unsigned long long *__restrict__ frame, uint32_t node_cap,
    const uint32_t *__restrict__ tsid, const double *__restrict__ roots, ListRecords R) {
  const uint32_t n_nodes = bounded_count(frame + frame::kNodes, node_cap);
  const uint32_t k = blockIdx.x * kBlock + threadIdx.x;
  // origin: the root centre of the body's (node's) system (tsid null: one
  // system, whose root is node 0)
  if (k < n) {
    const double *o = tsid ? roots + 4 * static_cast<size_t>(tsid[k]) : geom;
    write_body_record(R, k, tpos[3 * k], tpos[3 * k + 1], tpos[3 * k + 2], tmass[k], o[0], o[1],
                      o[2]);
  }
  if (k < n_nodes) {
    const double *g = geom + static_cast<size_t>(kGeomStride) * k;
    const int4 L = lnk[k];
    const double *o = tsid ? roots + 4 * static_cast<size_t>(tsid[L.z]) : geom;
    double row[15];
    for (int j = 0; j < 15; ++j) row[j] = g[j];
    write_node_records(R, k, row, L, skip[k], depth[k], o[0], o[1], o[2], o[3],
                       frame + frame::kWalkError);
  }
}

template <int MODE, bool PREFETCH>
__global__ __launch_bounds__(kWalkBlock) void walk(
    const WalkRec *__restrict__ recs, const uint2 *__restrict__ span,
    const double *__restrict__ tpos, gt_group *__restrict__ groups,
    unsigned long long *__restrict__ frame, uint32_t group_cap, uint32_t node_cap,
    const double *__restrict__ geom, const uint32_t *__restrict__ group_root, uint32_t cap_c,
    uint32_t cap_b, const uint32_t *__restrict__ cell_off, const uint32_t *__restrict__ body_off,
    uint32_t *__restrict__ cell_count, uint32_t *__restrict__ body_count,
    uint32_t *__restrict__ cell_list, uint32_t *__restrict__ body_list,
    const uint32_t *__restrict__ slot_c, const uint32_t *__restrict__ slot_b, uint32_t g_begin,
    uint32_t g_span) {
  if (frame[frame::kStatus]) return;  // the slots do not fit the buffers
  // groups g_begin .. g_begin + g_span of the first frame[kGroups]
  const uint32_t ng = bounded_count(frame + frame::kGroups, group_cap);
  const uint32_t nn = bounded_count(frame + frame::kNodes, node_cap);
  const uint32_t count = ng > g_begin ? min(ng - g_begin, g_span) : 0u;
  const uint32_t blk = xcd_block_used(blockIdx.x, (count + kWalkBlock - 1) / kWalkBlock);
  if (blk == kNone) return;
  const uint32_t g = g_begin + blk * kWalkBlock + threadIdx.x;
  if (g >= g_begin + count) return;
  const uint32_t gb = groups[g].begin, ge = groups[g].end;
  // the walk starts at the root of the group's system (node 0 for one system)
  const uint32_t start = group_root ? group_root[g] : 0u;
  const double root_half = geom[static_cast<size_t>(kGeomStride) * (start < nn ? start : 0u) + 3];
  double lo[3], hi[3];
  for (int a = 0; a < 3; ++a) lo[a] = hi[a] = tpos[3 * static_cast<size_t>(gb) + a];
  for (uint32_t k = gb; k < ge; ++k)
    for (int a = 0; a < 3; ++a) {
      const double v = tpos[3 * static_cast<size_t>(k) + a];
      lo[a] = (v < lo[a]) ? v : lo[a];  // std::min / std::max semantics
      hi[a] = (hi[a] < v) ? v : hi[a];
    }
  const double gcx = 0.5 * (lo[0] + hi[0]), gcy = 0.5 * (lo[1] + hi[1]),
               gcz = 0.5 * (lo[2] + hi[2]);
  const double ghx = 0.5 * (hi[0] - lo[0]), ghy = 0.5 * (hi[1] - lo[1]),
               ghz = 0.5 * (hi[2] - lo[2]);
  uint32_t cpos, bpos;
  if (MODE == kWalkExact) {
    cpos = cell_off[g];
    bpos = body_off[g];
  } else if (slot_c) {  // predicted slots, packed at scanned offsets
    cpos = cell_off[g];
    bpos = body_off[g];
    cap_c = slot_c[g];
    cap_b = slot_b[g];
  } else {
    cpos = g * cap_c;
    bpos = g * cap_b;
  }
  uint32_t nc = 0, nb = 0;
  uint32_t cbuf[4] = {0, 0, 0, 0};  // pending cell ids, indexed by list position mod 4
  uint32_t cur = start;
  // A depth-first walk visits each node at most once: more visits than nodes,
  // or an index past the end, means a malformed tree.  Stop and report it.
  uint32_t budget = nn;
  if (cur >= nn) {
    atomicOr(frame + frame::kWalkError, 1ull);
    cur = kNone;
  }
  // The node's record and body range are loaded together (the range does not
  // depend on the record).  With PREFETCH, the records of both possible next
  // nodes (last child, skip node) are requested as soon as the current record
  // is in, so they are in flight while the FP64 test runs.
  WalkRec r{};
  uint2 sp{};
  if (cur != kNone) {
    r = recs[cur];
    sp = span[cur];
  }
  while (cur != kNone) {
    if (budget-- == 0) {
      atomicOr(frame + frame::kWalkError, 1ull);
      break;
    }
    const uint32_t fc = static_cast<uint32_t>(r.bits) & kRecIndexMask;
    uint32_t sk = static_cast<uint32_t>(r.bits >> 26) & kRecIndexMask;
    sk = sk == kRecIndexMask ? kNone : sk;
    const uint32_t nch = static_cast<uint32_t>(r.bits >> 52) & 15u;
    const int depth = static_cast<int>((r.bits >> 56) & 63u);
    const bool small = (r.bits >> 62) & 1u;
    const uint32_t down = nch ? fc + nch - 1 : kNone;  // the last child
    WalkRec r_down{}, r_skip{};
    uint2 sp_down{}, sp_skip{};
    if (PREFETCH) {
      if (down < nn) {
        r_down = recs[down];
        sp_down = span[down];
      }
      if (sk < nn) {
        r_skip = recs[sk];
        sp_skip = span[sk];
      }
    }
    const double half = ldexp(root_half, -depth);
    bool overlap = true;
    double d2 = 0.0;
    {
      if (fabs(r.cx - gcx) > half * 1.0001 + ghx) overlap = false;
      const double d = fabs(r.mx - gcx) - ghx;
      if (d > 0) d2 += d * d;
    }
    {
      if (fabs(r.cy - gcy) > half * 1.0001 + ghy) overlap = false;
      const double d = fabs(r.my - gcy) - ghy;
      if (d > 0) d2 += d * d;
    }
    {
      if (fabs(r.cz - gcz) > half * 1.0001 + ghz) overlap = false;
      const double d = fabs(r.mz - gcz) - ghz;
      if (d > 0) d2 += d * d;
    }
    const bool accept = !overlap && d2 > r.crit * r.crit;
    uint32_t next;
    if (accept && !small) {
      // Cell ids go out four at a time (one 16-byte store) once the slot is
      // 16-byte aligned: fewer partial cache-line writes.
      const uint32_t at = cpos + nc;
      if (MODE == kWalkExact || nc < cap_c) {
        cbuf[at & 3u] = cur;
        if ((at & 3u) == 3u) {
          if (at >= 3u && at - 3u >= cpos)
            *reinterpret_cast<uint4 *>(cell_list + (at - 3u)) =
                make_uint4(cbuf[0], cbuf[1], cbuf[2], cbuf[3]);
          else
            for (uint32_t q = cpos; q <= at; ++q) cell_list[q] = cbuf[q & 3u];
        }
      }
      ++nc;
      next = sk;
    } else if (accept || nch == 0) {
      for (uint32_t j = sp.x; j < sp.y; ++j) {
        const uint32_t at = nb + (j - sp.x);
        if (MODE == kWalkExact || at < cap_b) body_list[bpos + at] = j;
      }
      nb += sp.y - sp.x;
      next = sk;
    } else {
      next = down;
    }
    if (next == kNone) break;
    if (next >= nn) {
      atomicOr(frame + frame::kWalkError, 1ull);
      break;
    }
    if (PREFETCH) {
      const bool d = next == down;
      r = d ? r_down : r_skip;
      sp = d ? sp_down : sp_skip;
    } else {
      r = recs[next];
      sp = span[next];
    }
    cur = next;
  }
  // the ids still pending (the last partial group of four)
  {
    const uint32_t stored = MODE == kWalkExact ? nc : min(nc, cap_c);
    const uint32_t end = cpos + stored;
    uint32_t first = end & ~3u;
    if (first < cpos) first = cpos;
    for (uint32_t q = first; q < end; ++q) cell_list[q] = cbuf[q & 3u];
  }
  if (MODE == kWalkCapacity && (nc > cap_c || nb > cap_b))
    atomicOr(frame + frame::kSlotOverflow, 1ull);
  cell_count[g] = nc;
  body_count[g] = nb;
  // The ranges a group's evaluation reads stay inside its slots even when the
  // lists outgrew them (an evaluation may run before the overflow is seen;
  // its results are then discarded).
  gt_group G = groups[g];
  G.cell_begin = cpos;
  G.cell_end = cpos + (MODE == kWalkExact ? nc : min(nc, cap_c));
  G.body_begin = bpos;
  G.body_end = bpos + (MODE == kWalkExact ? nb : min(nb, cap_b));
  groups[g] = G;
}

// Sums over the GL lanes that serve one group (GL = 4 or 8 consecutive lanes,
// inside one row of 16) by DPP moves: the group's total (quad swaps, then the
// other quad of the eight by the half-row mirror), and the sum over this lane
// (k of the group) and the lanes above it (row shifts, masked at the group's
// end).
template <int CTRL>
__device__ inline uint32_t dpp_u32(uint32_t v) {
  return static_cast<uint32_t>(
      __builtin_amdgcn_update_dpp(0, static_cast<int>(v), CTRL, 0xf, 0xf, false));
}
template <int GL>
__device__ inline uint32_t lanes_total(uint32_t v) {
  static_assert(GL == 4 || GL == 8, "4 or 8 lanes per group");
  v += dpp_u32<0xb1>(v);  // quad_perm [1,0,3,2]
  v += dpp_u32<0x4e>(v);  // quad_perm [2,3,0,1]
  if (GL == 8) v += dpp_u32<0x141>(v);  // row_half_mirror
  return v;
}
template <int GL>
__device__ inline uint32_t lanes_suffix(uint32_t v, uint32_t k) {
  static_assert(GL == 4 || GL == 8, "4 or 8 lanes per group");
  const uint32_t u1 = dpp_u32<0x101>(v);  // row_shl:1
  if (k + 1 < GL) v += u1;
  const uint32_t u2 = dpp_u32<0x102>(v);  // row_shl:2
  if (k + 2 < GL) v += u2;
  if (GL == 8) {
    const uint32_t u4 = dpp_u32<0x104>(v);  // row_shl:4
    if (k + 4 < GL) v += u4;
  }
  return v;
}

// Walk over child sets, eight lanes per group.  The nodes of one set (a
// node's children, first to last) are tested together, one per lane of the
// group's octet: their 32-byte images come in one round of independent loads,
// and the FP64 record is read only for a node whose image cannot decide.  The
// set is then taken in the reference's order, last child first: the nodes
// after the last one that opens (in child order) go to the lists at once, each
// lane writing its own entries at positions counted across the octet, and
// the node that opens makes its children the next set; the nodes before it
// wait as one entry (first node, how many, depth) on the octet's stack in
// LDS, one entry per tree level, and are loaded and tested again when the
// walk returns to them.  Decisions and order are those of the reference's
// stack walk, so the lists are identical.  A round of loads serves a whole
// set instead of one node, and eight lanes per group give the GPU eight times
// the waves of a walk with one lane per group to hide memory latency with.
constexpr int kCellCode = 0, kBodiesCode = 1, kOpenCode = 2;
constexpr uint32_t kSetMax = 8;  // children of a node at most

// GL lanes per group (8 or 4): lane k of a group tests children k, k + GL, ...
// of the current set.  With 4, a wave serves 16 groups and each lane tests two
// children; the bookkeeping, which every lane of a group repeats, costs half
// as much per group.
template <int MODE, int GL>
__global__ __launch_bounds__(kWalkBlock) void walk_sets(
    const WalkRec *__restrict__ recs, const WalkRec32 *__restrict__ images, uint32_t levels_cap,
    const uint2 *__restrict__ span, const double *__restrict__ tpos,
    gt_group *__restrict__ groups, unsigned long long *__restrict__ frame, uint32_t group_cap,
    uint32_t node_cap, const double *__restrict__ geom, const uint32_t *__restrict__ group_root,
    uint32_t cap_c, uint32_t cap_b, const uint32_t *__restrict__ cell_off,
    const uint32_t *__restrict__ body_off, uint32_t *__restrict__ cell_count,
    uint32_t *__restrict__ body_count, uint32_t *__restrict__ cell_list,
    uint32_t *__restrict__ body_list, const uint32_t *__restrict__ slot_c,
    const uint32_t *__restrict__ slot_b, uint32_t g_begin, uint32_t g_span,
    float4 *__restrict__ group_hilo) {
  constexpr uint32_t kPer = kSetMax / GL;            // children per lane
  constexpr uint32_t kGroups = kWalkBlock / GL;      // groups per block (one wave)
  constexpr uint32_t kLaneMask = (1u << GL) - 1u;
  extern __shared__ uint2 levels[];  // [kGroups][levels_cap]
  if (frame[frame::kStatus]) return;  // the slots do not fit the buffers
  // groups g_begin .. g_begin + g_span of the first frame[kGroups]
  const uint32_t ng = bounded_count(frame + frame::kGroups, group_cap);
  const uint32_t nn = bounded_count(frame + frame::kNodes, node_cap);
  const uint32_t count = ng > g_begin ? min(ng - g_begin, g_span) : 0u;
  const uint32_t blk = xcd_block_used(blockIdx.x, (count + kGroups - 1) / kGroups);
  if (blk == kNone) return;
  const uint32_t k = threadIdx.x % GL, grp = threadIdx.x / GL;
  const uint32_t g = g_begin + blk * kGroups + grp;
  const bool valid = g < g_begin + count;  // (the same for the group's lanes)
  uint2 *stack = levels + grp * levels_cap;
  uint32_t gb = 0, ge = 0, start = 0;
  if (valid) {
    gb = groups[g].begin;
    ge = groups[g].end;
    // the walk starts at the root of the group's system (node 0 for one system)
    start = group_root ? group_root[g] : 0u;
  }
  const double *origin = geom + static_cast<size_t>(kGeomStride) * (start < nn ? start : 0u);
  const double root_half = origin[3];
  // group box: the lanes take every GL-th body; min and max do not depend on
  // the order they are taken in
  double lo[3] = {INFINITY, INFINITY, INFINITY}, hi[3] = {-INFINITY, -INFINITY, -INFINITY};
  for (uint32_t j = gb + k; j < ge; j += GL)
    for (int a = 0; a < 3; ++a) {
      const double v = tpos[3 * static_cast<size_t>(j) + a];
      lo[a] = (v < lo[a]) ? v : lo[a];  // std::min / std::max semantics
      hi[a] = (hi[a] < v) ? v : hi[a];
    }
  for (int off = GL / 2; off > 0; off >>= 1)
    for (int a = 0; a < 3; ++a) {
      const double l = __shfl_xor(lo[a], off, GL), h = __shfl_xor(hi[a], off, GL);
      lo[a] = (l < lo[a]) ? l : lo[a];
      hi[a] = (hi[a] < h) ? h : hi[a];
    }
  const double gcx = 0.5 * (lo[0] + hi[0]), gcy = 0.5 * (lo[1] + hi[1]),
               gcz = 0.5 * (lo[2] + hi[2]);
  const double ghx = 0.5 * (hi[0] - lo[0]), ghy = 0.5 * (hi[1] - lo[1]),
               ghz = 0.5 * (hi[2] - lo[2]);
  const GroupImage gim =
      group_image(origin[0], origin[1], origin[2], root_half, gcx, gcy, gcz, ghx, ghy, ghz);
  const float root_half32 = static_cast<float>(root_half);
  // Hi/lo precision mode: the centre of the group's box relative to its
  // system's origin (the root node's centre) as FP32 hi + lo, for the
  // evaluation (group_hilo[2g], group_hilo[2g + 1]).
  if (group_hilo && valid && k == 0) {
    // (rounded subtractions, as group_centres forms them)
    const double c[3] = {__dsub_rn(gcx, origin[0]), __dsub_rn(gcy, origin[1]),
                         __dsub_rn(gcz, origin[2])};
    float hv[3], lv[3];
    for (int a = 0; a < 3; ++a) {
      hv[a] = static_cast<float>(c[a]);
      lv[a] = static_cast<float>(c[a] - static_cast<double>(hv[a]));
    }
    group_hilo[2 * static_cast<size_t>(g)] = make_float4(hv[0], hv[1], hv[2], 0.0f);
    group_hilo[2 * static_cast<size_t>(g) + 1] = make_float4(lv[0], lv[1], lv[2], 0.0f);
  }
  uint32_t cpos = 0, bpos = 0;
  if (valid) {
    if (MODE == kWalkExact) {
      cpos = cell_off[g];
      bpos = body_off[g];
    } else if (slot_c) {  // predicted slots, packed at scanned offsets
      cpos = cell_off[g];
      bpos = body_off[g];
      cap_c = slot_c[g];
      cap_b = slot_b[g];
    } else {
      cpos = g * cap_c;
      bpos = g * cap_b;
    }
  }
  // the current set: nodes fc .. fc + n - 1 at depth d (first: the root alone)
  uint32_t fc = start, n = 1, d = 0, top = 0, nc = 0, nb = 0, exact_tests = 0;
  bool done = !valid;
  bool failed = valid && start >= nn;
  done = done || failed;
  // Each round descends one level or returns to a waiting entry: a well-formed
  // tree takes fewer rounds than twice its node count.
  uint32_t budget = 2 * nn + 2;
  const uint32_t shift = threadIdx.x & ~(GL - 1u);
  for (;;) {
    if (!done &&
        (budget-- == 0 || n == 0 || n > kSetMax || fc >= nn || n > nn - fc || d > 63))
      failed = done = true;
    if (!__any(!done)) break;
    if (done) continue;
    WalkRec32 im[kPer];
    uint2 bs[kPer];
    int code[kPer];
    // lanes past the end of the set read its last node again (no branch, no
    // defaults to set; their results are masked below)
#pragma unroll
    for (uint32_t r = 0; r < kPer; ++r) {
      const uint32_t at = fc + min(k + GL * r, n - 1u);
      im[r] = images[at];
      bs[r] = span[at];
    }
    uint32_t open_bits = 0, cell_bits = 0;
#pragma unroll
    for (uint32_t r = 0; r < kPer; ++r) {
      const uint32_t c = k + GL * r;
      const bool in_set = c < n;
      int dec = gim.usable
                    ? accept_filter(im[r], half_image(root_half32, static_cast<int>(d)), gim)
                    : kUndecided;
      if (in_set && dec == kUndecided) {
        const WalkRec x = recs[fc + c];
        dec = accept_exact(x.cx, x.cy, x.cz, x.mx, x.my, x.mz,
                           ldexp(root_half, -static_cast<int>(d)), x.crit, gcx, gcy, gcz, ghx,
                           ghy, ghz);
        ++exact_tests;
      }
      const bool small = signbit(im[r].crit), leaf = (im[r].meta & kMetaLeaf) != 0u;
      code[r] = !in_set                  ? -1
                : (dec == 1 && !small)   ? kCellCode
                : (dec == 1 || leaf)     ? kBodiesCode
                                         : kOpenCode;
      open_bits |= (static_cast<uint32_t>(__ballot(code[r] == kOpenCode) >> shift) & kLaneMask)
                   << (GL * r);
      cell_bits |= (static_cast<uint32_t>(__ballot(code[r] == kCellCode) >> shift) & kLaneMask)
                   << (GL * r);
    }
    // the last node (in child order) that opens; the nodes after it go out now
    const int open = open_bits ? 31 - __clz(static_cast<int>(open_bits)) : -1;
    const uint32_t emit = (0xffu << (open + 1)) & 0xffu;
    // bodies: per slot, this lane's and the higher lanes' counts
    uint32_t cnt[kPer], suffix[kPer], slot_total[kPer];
#pragma unroll
    for (uint32_t r = 0; r < kPer; ++r) {
      const uint32_t c = k + GL * r;
      const bool mine = c < n && ((emit >> c) & 1u);
      cnt[r] = (mine && code[r] == kBodiesCode) ? bs[r].y - bs[r].x : 0u;
      suffix[r] = lanes_suffix<GL>(cnt[r], k);
      slot_total[r] = lanes_total<GL>(cnt[r]);
      if (mine && code[r] == kCellCode) {
        const uint32_t at = nc + __popc(cell_bits & emit & (0xfeu << c));
        if (MODE == kWalkExact || at < cap_c) cell_list[cpos + at] = fc + c;
      }
    }
    // Each lane writes its own node's bodies: up to eight as straight-line
    // predicated stores at immediate offsets (a leaf holds at most leaf_size
    // bodies, eight by default), the rest of a larger leaf in a loop.
    uint32_t higher_slots = 0;  // bodies of the slots above
#pragma unroll
    for (int r = static_cast<int>(kPer) - 1; r >= 0; --r) {
      const uint32_t first = nb + higher_slots + (suffix[r] - cnt[r]);
      uint32_t m = cnt[r];
      if (MODE != kWalkExact) m = min(m, __builtin_elementwise_sub_sat(cap_b, first));  // room left
      if (m) {
        uint32_t *const out = body_list + (bpos + first);
        const uint32_t x = bs[r].x;
#pragma unroll
        for (uint32_t j = 0; j < 8; ++j)
          if (j < m) out[j] = x + j;
        for (uint32_t j = 8; j < m; ++j) out[j] = x + j;
      }
      higher_slots += slot_total[r];
    }
    nc += __popc(cell_bits & emit);
    nb += higher_slots;
    // the next set
    const uint32_t open_slot = open < 0 ? 0u : static_cast<uint32_t>(open) / GL;
    uint32_t meta_sel = im[0].meta;
#pragma unroll
    for (uint32_t r = 1; r < kPer; ++r) meta_sel = open_slot == r ? im[r].meta : meta_sel;
    const uint32_t open_meta = __shfl(meta_sel, open < 0 ? 0 : open % GL, GL);
    if (open >= 0) {
      if (open > 0) {  // the nodes before it wait
        if (top >= levels_cap) {
          failed = done = true;
        } else {
          stack[top] = make_uint2(fc | (static_cast<uint32_t>(open) << 26), d);
          ++top;
        }
      }
      fc = meta_first_child(open_meta);
      n = meta_children(open_meta);
      ++d;
    } else if (top > 0) {
      --top;
      const uint2 e = stack[top];
      fc = e.x & kMetaIndexMask;
      n = e.x >> 26;
      d = e.y;
    } else {
      done = true;
    }
  }
  // (the whole wave is here: one atomic for its count)
  for (int off = 32; off > 0; off >>= 1) exact_tests += __shfl_xor(exact_tests, off, 64);
  if (threadIdx.x == 0 && exact_tests)
    atomicAdd(frame + frame::kExactTests, static_cast<unsigned long long>(exact_tests));
  if (!valid || k != 0) return;
  if (failed) atomicOr(frame + frame::kWalkError, 1ull);
  if (MODE == kWalkCapacity && (nc > cap_c || nb > cap_b))
    atomicOr(frame + frame::kSlotOverflow, 1ull);
  cell_count[g] = nc;
  body_count[g] = nb;
  // The ranges a group's evaluation reads stay inside its slots even when the
  // lists outgrew them (an evaluation may run before the overflow is seen;
  // its results are then discarded).
  gt_group G = groups[g];
  G.cell_begin = cpos;
  G.cell_end = cpos + (MODE == kWalkExact ? nc : min(nc, cap_c));
  G.body_begin = bpos;
  G.body_end = bpos + (MODE == kWalkExact ? nb : min(nb, cap_b));
  groups[g] = G;
}

// Wave-coherent walk: LANES consecutive (spatially adjacent) groups share one
// depth-first traversal of the union of their trees.  The node being visited
// is the same for the whole wave, so its record comes through the scalar cache
// into SGPRs, together with the records of both possible next nodes (skip and
// last child).  Each lane applies its own group's FP64 test; a lane that
// accepts a node (or takes a leaf's bodies) sleeps until the traversal leaves
// that node's subtree (reaches its skip node); the wave descends when any
// awake lane opens the node.  A lane is awake exactly at the nodes its own
// depth-first walk visits, in the same order, so each lane's lists are the
// reference's.
constexpr uint32_t kAwake = 0xfffffffeu;

template <int MODE, int LANES>
__global__ __launch_bounds__(kBlock) void walk_coherent(
    const WalkRec *__restrict__ recs, const uint2 *__restrict__ span,
    const double *__restrict__ tpos, gt_group *__restrict__ groups,
    unsigned long long *__restrict__ frame, uint32_t group_cap, uint32_t node_cap,
    const double *__restrict__ root, const uint32_t *__restrict__ group_root, uint32_t cap_c,
    uint32_t cap_b, const uint32_t *__restrict__ cell_off, const uint32_t *__restrict__ body_off,
    uint32_t *__restrict__ cell_count, uint32_t *__restrict__ body_count,
    uint32_t *__restrict__ cell_list, uint32_t *__restrict__ body_list,
    const uint32_t *__restrict__ slot_c, const uint32_t *__restrict__ slot_b, uint32_t g_begin,
    uint32_t g_span) {
  (void)g_begin;  // all groups at once
  (void)g_span;
  // One tree and uniform slots only: the shared traversal starts at node 0
  // (group_root unused; batches and predicted slots walk one lane per group).
  (void)group_root;
  (void)slot_c;
  (void)slot_b;
  if (frame[frame::kStatus]) return;
  const uint32_t ng = bounded_count(frame + frame::kGroups, group_cap);
  const uint32_t nn = bounded_count(frame + frame::kNodes, node_cap);
  const uint32_t waves = (ng + LANES - 1) / LANES;
  const uint32_t blk = xcd_block_used(blockIdx.x, (waves + kBlock / 64 - 1) / (kBlock / 64));
  if (blk == kNone) return;
  const uint32_t lane = threadIdx.x % 64;
  const uint32_t wave = __builtin_amdgcn_readfirstlane(blk * (kBlock / 64) + threadIdx.x / 64);
  if (wave * LANES >= ng) return;  // the whole wave leaves together
  const uint32_t g = wave * LANES + lane;
  const bool mine = lane < static_cast<uint32_t>(LANES) && g < ng;
  const double root_half = root[3];
  double gcx = 0.0, gcy = 0.0, gcz = 0.0, ghx = 0.0, ghy = 0.0, ghz = 0.0;
  uint32_t cpos = 0, bpos = 0;
  if (mine) {
    const uint32_t gb = groups[g].begin, ge = groups[g].end;
    double lo[3], hi[3];
    for (int a = 0; a < 3; ++a) lo[a] = hi[a] = tpos[3 * static_cast<size_t>(gb) + a];
    for (uint32_t k = gb; k < ge; ++k)
      for (int a = 0; a < 3; ++a) {
        const double v = tpos[3 * static_cast<size_t>(k) + a];
        lo[a] = (v < lo[a]) ? v : lo[a];  // std::min / std::max semantics
        hi[a] = (hi[a] < v) ? v : hi[a];
      }
    gcx = 0.5 * (lo[0] + hi[0]);
    gcy = 0.5 * (lo[1] + hi[1]);
    gcz = 0.5 * (lo[2] + hi[2]);
    ghx = 0.5 * (hi[0] - lo[0]);
    ghy = 0.5 * (hi[1] - lo[1]);
    ghz = 0.5 * (hi[2] - lo[2]);
    if (MODE == kWalkExact) {
      cpos = cell_off[g];
      bpos = body_off[g];
    } else {
      cpos = g * cap_c;
      bpos = g * cap_b;
    }
  }
  uint32_t resume = mine ? kAwake : kNone;
  uint32_t nc = 0, nb = 0;
  uint32_t cur = 0;
  WalkRec r = recs[0];
  uint32_t budget = nn;
  for (;;) {
    if (budget-- == 0) {
      atomicOr(frame + frame::kWalkError, 1ull);
      break;
    }
    const uint32_t fc = static_cast<uint32_t>(r.bits) & kRecIndexMask;
    uint32_t sk = static_cast<uint32_t>(r.bits >> 26) & kRecIndexMask;
    sk = sk == kRecIndexMask ? kNone : sk;
    const uint32_t nch = static_cast<uint32_t>(r.bits >> 52) & 15u;
    const int depth = static_cast<int>((r.bits >> 56) & 63u);
    const bool small = (r.bits >> 62) & 1u;
    const uint32_t last = fc + nch - 1;
    // Both possible next records are requested before this node is tested.
    WalkRec rs = r, rc = r;
    if (sk < nn) rs = recs[sk];
    if (nch != 0 && last < nn) rc = recs[last];
    if (resume == cur) resume = kAwake;
    bool opens = false;
    if (resume == kAwake) {
      const double half = ldexp(root_half, -depth);
      bool overlap = true;
      double d2 = 0.0;
      {
        if (fabs(r.cx - gcx) > half * 1.0001 + ghx) overlap = false;
        const double d = fabs(r.mx - gcx) - ghx;
        if (d > 0) d2 += d * d;
      }
      {
        if (fabs(r.cy - gcy) > half * 1.0001 + ghy) overlap = false;
        const double d = fabs(r.my - gcy) - ghy;
        if (d > 0) d2 += d * d;
      }
      {
        if (fabs(r.cz - gcz) > half * 1.0001 + ghz) overlap = false;
        const double d = fabs(r.mz - gcz) - ghz;
        if (d > 0) d2 += d * d;
      }
      const bool accept = !overlap && d2 > r.crit * r.crit;
      if (accept && !small) {
        if (MODE == kWalkExact || nc < cap_c) cell_list[cpos + nc] = cur;
        ++nc;
        resume = sk;
      } else if (accept || nch == 0) {
        const uint2 sp = span[cur];
        for (uint32_t j = sp.x; j < sp.y; ++j) {
          const uint32_t at = nb + (j - sp.x);
          if (MODE == kWalkExact || at < cap_b) body_list[bpos + at] = j;
        }
        nb += sp.y - sp.x;
        resume = sk;
      } else {
        opens = true;
      }
    }
    const bool descend = __ballot(opens) != 0;
    const uint32_t next = __builtin_amdgcn_readfirstlane(descend ? last : sk);
    if (next == kNone) break;
    if (next >= nn) {
      atomicOr(frame + frame::kWalkError, 1ull);
      break;
    }
    r = descend ? rc : rs;
    cur = next;
  }
  if (mine) {
    if (MODE == kWalkCapacity && (nc > cap_c || nb > cap_b))
      atomicOr(frame + frame::kSlotOverflow, 1ull);
    cell_count[g] = nc;
    body_count[g] = nb;
    gt_group G = groups[g];
    G.cell_begin = cpos;
    G.cell_end = cpos + (MODE == kWalkExact ? nc : min(nc, cap_c));
    G.body_begin = bpos;
    G.body_end = bpos + (MODE == kWalkExact ? nb : min(nb, cap_b));
    groups[g] = G;
  }
}

#pragma clang fp contract(fast)

__global__ __launch_bounds__(kBlock) void group_starts(const int4 *__restrict__ lnk,
                                                       const uint32_t *__restrict__ parent,
                                                       const unsigned long long *__restrict__ frame,
                                                       uint32_t node_cap, uint32_t gs, int pack,
                                                       uint32_t *__restrict__ flags,
                                                       uint32_t *__restrict__ node_of) {
  const uint32_t n_nodes = bounded_count(frame + frame::kNodes, node_cap);
  const uint32_t x = blockIdx.x * kBlock + threadIdx.x;
  if (x >= n_nodes) return;
  const int4 L = lnk[x];
  flag_group_starts(x, L, parent[x] == kNone, gs, pack, flags, node_of,
                    [&](int c) { return lnk[L.x + c]; });
}

__global__ __launch_bounds__(kBlock) void make_groups(const uint32_t *__restrict__ starts,
                                                      const unsigned long long *__restrict__ frame,
                                                      uint32_t group_cap, uint32_t n,
                                                      const uint32_t *__restrict__ tsid,
                                                      const uint32_t *__restrict__ root_node,
                                                      gt_group *__restrict__ groups,
                                                      uint32_t *__restrict__ group_root,
                                                      const uint32_t *__restrict__ node_of,
                                                      uint32_t *__restrict__ group_node) {
  const uint32_t ng = bounded_count(frame + frame::kGroups, group_cap);
  const uint32_t g = blockIdx.x * kBlock + threadIdx.x;
  if (g >= ng) return;
  gt_group G;
  G.begin = starts[g];
  G.end = g + 1 < ng ? starts[g + 1] : n;
  G.cell_begin = G.cell_end = G.body_begin = G.body_end = 0;
  G.pad0 = G.pad1 = 0;
  groups[g] = G;
  if (tsid) group_root[g] = root_node[tsid[G.begin]];
  if (group_node) group_node[g] = node_of[G.begin];
}

// The status of an evaluation from its counts and the plan's bounds (one thread).
__device__ inline void check_counts(unsigned long long *frame, uint32_t node_cap,
                                    uint32_t depth_cap, uint32_t group_cap, int check_slots) {
  unsigned long long st = 0;
  if (frame[frame::kNodes] > node_cap) st |= kAbortNodes;
  if (frame[frame::kUnresolved] & 1ull) st |= kAbortPaths;
  if (frame[frame::kDepth] > depth_cap) st |= kAbortDepth;
  if (frame[frame::kGroups] > group_cap) st |= kAbortGroups;
  if (check_slots && frame[frame::kSlotOverflow]) st |= kAbortLists;
  if (frame[frame::kWalkError] & 1ull) st |= kAbortWalk;
  if (st) atomicOr(frame + frame::kStatus, st);  // (another stream may set bits too)
}

// stats: [0] cell interactions, [1] body interactions, [2] max list, [3] max group,
// [4] most cells of a group, [5] most bodies of a group, [6] all cells, [7] all bodies.
// A few hundred blocks stride over the groups and each adds its totals with one
// set of atomics (one set per wave made the atomics the kernel's cost).
constexpr uint32_t kStatsBlocks = 256;

// With check set, thread 0 of workgroup 0 also sets the status from the
// counts and the plan's bounds (check_frame's test with the slots checked:
// the counts it reads are final once the walk has ended).
__global__ __launch_bounds__(kBlock) void list_stats(const uint32_t *__restrict__ cell_count,
                                                     const uint32_t *__restrict__ body_count,
                                                     const gt_group *__restrict__ groups,
                                                     unsigned long long *__restrict__ frame,
                                                     uint32_t group_cap, int check,
                                                     uint32_t node_cap, uint32_t depth_cap) {
  if (check && blockIdx.x == 0 && threadIdx.x == 0)
    check_counts(frame, node_cap, depth_cap, group_cap, 1);
  __shared__ unsigned long long part[kBlock / 64][8];
  const uint32_t ng = bounded_count(frame + frame::kGroups, group_cap);
  unsigned long long *stats = frame + frame::kStats;
  unsigned long long v[8] = {0, 0, 0, 0, 0, 0, 0, 0};  // ci, bi, ml, mg, mc, mb, sc, sb
  for (uint32_t g = blockIdx.x * kBlock + threadIdx.x; g < ng; g += gridDim.x * kBlock) {
    const unsigned long long t = groups[g].end - groups[g].begin, c = cell_count[g],
                             b = body_count[g];
    v[0] += t * c;
    v[1] += t * b - t;
    v[2] = max(v[2], c + b);
    v[3] = max(v[3], t);
    v[4] = max(v[4], c);
    v[5] = max(v[5], b);
    v[6] += c;
    v[7] += b;
  }
  for (int off = 32; off > 0; off >>= 1)
    for (int i = 0; i < 8; ++i) {
      const unsigned long long o = __shfl_xor(v[i], off, 64);
      v[i] = (i >= 2 && i <= 5) ? max(v[i], o) : v[i] + o;
    }
  if ((threadIdx.x & 63) == 0)
    for (int i = 0; i < 8; ++i) part[threadIdx.x / 64][i] = v[i];
  __syncthreads();
  if (threadIdx.x < 8) {
    const int i = threadIdx.x;
    unsigned long long a = part[0][i];
    for (int w = 1; w < kBlock / 64; ++w)
      a = (i >= 2 && i <= 5) ? max(a, part[w][i]) : a + part[w][i];
    if (a != 0) {
      if (i >= 2 && i <= 5)
        atomicMax(&stats[i], a);
      else
        atomicAdd(&stats[i], a);
    }
  }
}

__global__ void check_frame(unsigned long long *frame, uint32_t node_cap, uint32_t depth_cap,
                            uint32_t group_cap, int check_slots) {
  check_counts(frame, node_cap, depth_cap, group_cap, check_slots);
}

__global__ void fill_u32(uint32_t *p, size_t n, uint32_t v) {
  const size_t i = static_cast<size_t>(blockIdx.x) * kBlock + threadIdx.x;
  if (i < n) p[i] = v;
}

__global__ void fill_u64(unsigned long long *p, size_t n, unsigned long long v) {
  const size_t i = static_cast<size_t>(blockIdx.x) * kBlock + threadIdx.x;
  if (i < n) p[i] = v;
}

}  // namespace

void launch_fill_u32(uint32_t *p, size_t n, uint32_t v, hipStream_t s) {
  if (!n) return;
  hipLaunchKernelGGL(fill_u32, dim3(blocks_for(n)), dim3(kBlock), 0, s, p, n, v);
  GTH_LAUNCH_CHECK();
}

void launch_fill_u64(unsigned long long *p, size_t n, unsigned long long v, hipStream_t s) {
  if (!n) return;
  hipLaunchKernelGGL(fill_u64, dim3(blocks_for(n)), dim3(kBlock), 0, s, p, n, v);
  GTH_LAUNCH_CHECK();
}

void launch_link_up(const int4 *lnk, const unsigned long long *frame, uint32_t node_cap,
                    uint32_t *parent, uint32_t *skip, uint32_t *depth, hipStream_t s) {
  if (!node_cap) return;
  launch_fill_u32(parent, node_cap, kNone, s);
  hipLaunchKernelGGL(set_parents, dim3(blocks_for(node_cap)), dim3(kBlock), 0, s, lnk, frame,
                     node_cap, parent);
  GTH_LAUNCH_CHECK();
  hipLaunchKernelGGL(set_skip_depth, dim3(blocks_for(node_cap)), dim3(kBlock), 0, s, lnk, parent,
                     frame, node_cap, skip, depth);
  GTH_LAUNCH_CHECK();
}

void launch_prep_lists(const double *geom, const int4 *lnk, const uint32_t *skip,
                       const uint32_t *depth, const double *tpos, const double *tmass, uint32_t n,
                       unsigned long long *frame, uint32_t node_cap, const uint32_t *tsid,
                       const double *roots, const ListRecords &rec, hipStream_t s) {
  const uint32_t m = n > node_cap ? n : node_cap;
  if (!m) return;
  hipLaunchKernelGGL(prep_lists, dim3(blocks_for(m)), dim3(kBlock), 0, s, geom, lnk, skip, depth,
                     tpos, tmass, n, frame, node_cap, tsid, roots, rec);
  GTH_LAUNCH_CHECK();
}

void launch_group_starts(const int4 *lnk, const uint32_t *parent, const unsigned long long *frame,
                         uint32_t node_cap, uint32_t n, uint32_t gs, int pack, uint32_t *flags,
                         hipStream_t s, uint32_t *node_of) {
  launch_fill_u32(flags, n, 0u, s);
  if (!node_cap) return;
  hipLaunchKernelGGL(group_starts, dim3(blocks_for(node_cap)), dim3(kBlock), 0, s, lnk, parent,
                     frame, node_cap, gs, pack, flags, node_of);
  GTH_LAUNCH_CHECK();
}

void launch_make_groups(uint32_t n, const unsigned long long *frame, uint32_t group_cap,
                        const uint32_t *tsid, const uint32_t *root_node, const uint32_t *starts,
                        gt_group *groups, uint32_t *group_root, hipStream_t s,
                        const uint32_t *node_of, uint32_t *group_node) {
  if (!n || !group_cap) return;
  hipLaunchKernelGGL(make_groups, dim3(blocks_for(group_cap)), dim3(kBlock), 0, s, starts, frame,
                     group_cap, n, tsid, root_node, groups, group_root, node_of, group_node);
  GTH_LAUNCH_CHECK();
}

void launch_walk(int mode, int lanes, const WalkRec *recs, const WalkRec32 *images,
                 uint32_t levels_cap, const uint2 *span, const double *tpos,
                 gt_group *groups, unsigned long long *frame, uint32_t group_cap,
                 uint32_t node_cap, const double *geom, const uint32_t *group_root, uint32_t cap_c,
                 uint32_t cap_b,
                 const uint32_t *cell_off, const uint32_t *body_off, uint32_t *cell_count,
                 uint32_t *body_count, uint32_t *cell_list, uint32_t *body_list, hipStream_t s,
                 const uint32_t *slot_c, const uint32_t *slot_b, uint32_t g_begin,
                 uint32_t g_span, float4 *group_hilo) {
  if (!group_cap) return;
  if (group_hilo && lanes != 3 && lanes != 4)
    fail(GT_HIP_ERR_ARGUMENT, "walk: only the child-set walks write group centres");
  if (g_span == 0 || g_span > group_cap) g_span = group_cap;
  const bool coherent = lanes >= 16;
  if (coherent && (g_begin != 0 || g_span != group_cap))
    fail(GT_HIP_ERR_ARGUMENT, "coherent walks take all groups at once");
  if ((group_root || slot_c) && coherent)
    fail(GT_HIP_ERR_ARGUMENT, "coherent walks need a single tree and uniform slots");
#define GTH_WALK_ARGS                                                                   \
  recs, span, tpos, groups, frame, group_cap, node_cap, geom, group_root, cap_c, cap_b, \
      cell_off, body_off, cell_count, body_count, cell_list, body_list, slot_c, slot_b, g_begin, \
      g_span
  if (lanes == 3 || lanes == 4) {  // 8 (lanes 3) or 4 (lanes 4) lanes per group, over child sets
    if (levels_cap == 0 || levels_cap > 64) fail(GT_HIP_ERR_ARGUMENT, "walk: levels_cap out of range");
    const uint32_t gl = lanes == 3 ? 8u : 4u, per_block = kWalkBlock / gl;
    const dim3 grid(xcd_grid(blocks_for(g_span, per_block))), block(kWalkBlock);
    const size_t lds = static_cast<size_t>(levels_cap) * per_block * sizeof(uint2);
#define GTH_SETS_ARGS                                                                        \
  recs, images, levels_cap, span, tpos, groups, frame, group_cap, node_cap, geom, group_root, \
      cap_c, cap_b, cell_off, body_off, cell_count, body_count, cell_list, body_list, slot_c,  \
      slot_b, g_begin, g_span, group_hilo
    if (mode == kWalkExact && gl == 8)
      hipLaunchKernelGGL((walk_sets<kWalkExact, 8>), grid, block, lds, s, GTH_SETS_ARGS);
    else if (mode == kWalkExact)
      hipLaunchKernelGGL((walk_sets<kWalkExact, 4>), grid, block, lds, s, GTH_SETS_ARGS);
    else if (gl == 8)
      hipLaunchKernelGGL((walk_sets<kWalkCapacity, 8>), grid, block, lds, s, GTH_SETS_ARGS);
    else
      hipLaunchKernelGGL((walk_sets<kWalkCapacity, 4>), grid, block, lds, s, GTH_SETS_ARGS);
#undef GTH_SETS_ARGS
    GTH_LAUNCH_CHECK();
    return;
  }
  if (lanes == 1 || lanes == 2) {  // one walk per lane: 1 prefetching, 2 plain
    const dim3 grid(xcd_grid(blocks_for(g_span, kWalkBlock))), block(kWalkBlock);
    if (mode == kWalkExact && lanes == 1)
      hipLaunchKernelGGL((walk<kWalkExact, true>), grid, block, 0, s, GTH_WALK_ARGS);
    else if (mode == kWalkExact)
      hipLaunchKernelGGL((walk<kWalkExact, false>), grid, block, 0, s, GTH_WALK_ARGS);
    else if (lanes == 1)
      hipLaunchKernelGGL((walk<kWalkCapacity, true>), grid, block, 0, s, GTH_WALK_ARGS);
    else
      hipLaunchKernelGGL((walk<kWalkCapacity, false>), grid, block, 0, s, GTH_WALK_ARGS);
    GTH_LAUNCH_CHECK();
    return;
  }
  const uint32_t waves = static_cast<uint32_t>((group_cap + lanes - 1) / lanes);
  const dim3 grid(xcd_grid((waves + kBlock / 64 - 1) / (kBlock / 64))), block(kBlock);
  if (mode == kWalkExact) {
    if (lanes == 16)
      hipLaunchKernelGGL((walk_coherent<kWalkExact, 16>), grid, block, 0, s, GTH_WALK_ARGS);
    else if (lanes == 32)
      hipLaunchKernelGGL((walk_coherent<kWalkExact, 32>), grid, block, 0, s, GTH_WALK_ARGS);
    else
      hipLaunchKernelGGL((walk_coherent<kWalkExact, 64>), grid, block, 0, s, GTH_WALK_ARGS);
  } else {
    if (lanes == 16)
      hipLaunchKernelGGL((walk_coherent<kWalkCapacity, 16>), grid, block, 0, s, GTH_WALK_ARGS);
    else if (lanes == 32)
      hipLaunchKernelGGL((walk_coherent<kWalkCapacity, 32>), grid, block, 0, s, GTH_WALK_ARGS);
    else
      hipLaunchKernelGGL((walk_coherent<kWalkCapacity, 64>), grid, block, 0, s, GTH_WALK_ARGS);
  }
#undef GTH_WALK_ARGS
  GTH_LAUNCH_CHECK();
}

void launch_list_stats(const uint32_t *cell_count, const uint32_t *body_count,
                       const gt_group *groups, unsigned long long *frame, uint32_t group_cap,
                       hipStream_t s, const Plan *check) {
  if (!group_cap) {
    if (check) launch_check_frame(frame, *check, true, s);
    return;
  }
  const uint32_t blocks = std::min<uint32_t>(blocks_for(group_cap), kStatsBlocks);
  hipLaunchKernelGGL(list_stats, dim3(blocks), dim3(kBlock), 0, s, cell_count, body_count, groups,
                     frame, group_cap, check ? 1 : 0, check ? check->node_cap : 0u,
                     check ? static_cast<uint32_t>(check->depth_cap) : 0u);
  GTH_LAUNCH_CHECK();
}

void launch_check_frame(unsigned long long *frame, const Plan &plan, bool check_slots,
                        hipStream_t s) {
  hipLaunchKernelGGL(check_frame, dim3(1), dim3(1), 0, s, frame, plan.node_cap,
                     static_cast<uint32_t>(plan.depth_cap), plan.group_cap, check_slots ? 1 : 0);
  GTH_LAUNCH_CHECK();
}

}  // namespace gth
