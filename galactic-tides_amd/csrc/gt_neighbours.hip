// Fixed-radius neighbour queries on the GPU octree (reference:
// Octree::for_each_within, include/galactic_tides/octree.h:103-127).
//
// Each query walks the tree in the reference's LIFO order (children last to
// first) without a stack, through the skip links (`skip` is the node the LIFO
// walk pops after a subtree), and takes a leaf's particles in tree order: the
// order in which the reference's callback receives them.
// The box test and the squared distances are the reference's FP64 expressions
// in its operation order with no contraction, so the neighbour lists and the
// distances equal the reference's bit for bit.  One walk counts every query's
// neighbours and writes the first ones into a slot per query; after an
// exclusive scan of the counts the lists that fit are moved to their CSR
// places and the queries with more are walked again straight into theirs
// (a count pass and a fill pass when the slots would not fit).
//
// Lanes take the queries in the order of a 30-bit Morton key of their centres
// in the root cube (query_keys + a radix sort), so the lanes of a wave walk
// nearby parts of the tree; results stay in the caller's query order.  A node
// visit reads one 64-byte record (centre, half, children, body range, skip
// link) instead of three arrays.
#include <algorithm>

#include "gt_internal.h"

namespace gth {
namespace {

constexpr int kBlock = 256;

// What a query walk reads of a node, in one 64-byte record: the cube (FP64
// centre and half size), the first child and child count (0: a leaf), the
// body range, the skip link and the skip node's depth.
struct NbNode {
  double cx, cy, cz, half;
  uint32_t first, children, begin, end;
  uint32_t skip, skip_depth, pad1, pad2;
};
static_assert(sizeof(NbNode) == 64, "NbNode layout");

__global__ __launch_bounds__(kBlock) void nb_records(const double *__restrict__ geom,
                                                     const int4 *__restrict__ lnk,
                                                     const uint32_t *__restrict__ skip,
                                                     const uint32_t *__restrict__ depth,
                                                     const unsigned long long *__restrict__ frame,
                                                     uint32_t node_cap, NbNode *__restrict__ out) {
  const uint32_t nn = bounded_count(frame + frame::kNodes, node_cap);
  const uint32_t x = blockIdx.x * kBlock + threadIdx.x;
  if (x >= nn) return;
  const double *g = geom + static_cast<size_t>(kGeomStride) * x;
  const int4 L = lnk[x];
  NbNode r;
  r.cx = g[0];
  r.cy = g[1];
  r.cz = g[2];
  r.half = g[3];
  r.first = static_cast<uint32_t>(L.x);
  r.children = static_cast<uint32_t>(L.y);
  r.begin = static_cast<uint32_t>(L.z);
  r.end = static_cast<uint32_t>(L.w);
  r.skip = skip[x];
  r.skip_depth = r.skip < nn ? depth[r.skip] : 0u;
  r.pad1 = r.pad2 = 0;
  out[x] = r;
}

#pragma clang fp contract(off)

// Squared distance from c to the node's cube, as the reference sums it
// (axes in order, only the axes where c is outside the slab).
__device__ inline double box_d2(const NbNode &g, double cx, double cy, double cz) {
  double d2 = 0.0;
  const double dx = fabs(cx - g.cx) - g.half;
  if (dx > 0) d2 += dx * dx;
  const double dy = fabs(cy - g.cy) - g.half;
  if (dy > 0) d2 += dy * dy;
  const double dz = fabs(cz - g.cz) - g.half;
  if (dz > 0) d2 += dz * dz;
  return d2;
}

// kCount: counts[q] = neighbours of query q.  kFill: query q's neighbours
// into index / dist2 from offs[q], at most counts[q] of them.  kStage: the
// first `stride` of them into index / dist2 from q * stride, and counts[q] =
// all of them.  Lane k < lanes takes query perm[k] (< nq).
//
// One wave walks for its 64 queries together (a packet walk): it visits, in
// the depth-first order above, every node that any of its lanes would visit,
// and a lane takes part at a node only while every ancestor's cube was within
// its radius (its active mask for each depth is kept in LDS), so each lane
// sees exactly the nodes of its own walk, in order.  Node records and
// particles are read at wave-uniform addresses.  Against one independent walk
// per lane, a counting and a filling walk for 2M queries of 32 neighbours took
// 9.1 instead of 11.8 ms.
constexpr int kMaxLevels = 64;

constexpr int kCount = 0, kFill = 1, kStage = 2;

template <int kMode>
__global__ __launch_bounds__(kBlock) void within(
    const NbNode *__restrict__ nodes, const double *__restrict__ tpos,
    const uint32_t *__restrict__ order, uint32_t n, const unsigned long long *__restrict__ frame,
    uint32_t node_cap, const double *__restrict__ centres, const double *__restrict__ radii,
    uint32_t radius_stride, uint32_t nq, uint32_t lanes, const uint32_t *__restrict__ perm,
    uint32_t *__restrict__ counts, const uint32_t *__restrict__ offs,
    unsigned long long *__restrict__ index, double *__restrict__ dist2, uint32_t stride,
    unsigned long long *__restrict__ error) {
  __shared__ unsigned long long masks[kBlock / 64][kMaxLevels];
  const uint32_t lane = threadIdx.x % 64, wave = threadIdx.x / 64;
  const uint32_t k = blockIdx.x * kBlock + threadIdx.x;
  const bool valid = k < lanes;
  const uint32_t q = valid ? perm[k] : 0u;
  const bool ok = valid && q < nq;
  if (valid && !ok) atomicOr(error, 2ull);
  const uint32_t nn = bounded_count(frame + frame::kNodes, node_cap);
  double cx = 0, cy = 0, cz = 0, r2 = 0;
  uint32_t found = 0;
  size_t at = 0, end = 0;
  if (ok) {
    cx = centres[3 * static_cast<size_t>(q)];
    cy = centres[3 * static_cast<size_t>(q) + 1];
    cz = centres[3 * static_cast<size_t>(q) + 2];
    const double r = radii[static_cast<size_t>(q) * radius_stride];
    r2 = r * r;
    if (kMode == kFill) {
      at = offs[q];
      end = at + counts[q];
    } else if (kMode == kStage) {
      at = static_cast<size_t>(q) * stride;
      end = at + stride;
    }
  }
  const unsigned long long me = 1ull << lane;
  unsigned long long *stack = masks[wave];
  stack[0] = __ballot(ok);
  uint32_t cur = (nn > 0 && stack[0] != 0) ? 0u : kNone, d = 0, budget = nn;
  while (cur != kNone) {
    if (cur >= nn || budget-- == 0 || d >= kMaxLevels - 1) {
      if (lane == 0) atomicOr(error, 1ull);  // malformed tree (never for a built one)
      break;
    }
    const NbNode g = nodes[__builtin_amdgcn_readfirstlane(cur)];
    const bool act = (stack[d] & me) != 0;
    const bool pass = act && !(box_d2(g, cx, cy, cz) > r2);
    const unsigned long long passed = __ballot(pass);
    if (g.children == 0) {
      if (passed) {
        const uint32_t b = g.begin, e = min(g.end, n);
        for (uint32_t j = b; j < e; ++j) {
          const double dx = tpos[3 * static_cast<size_t>(j)] - cx;
          const double dy = tpos[3 * static_cast<size_t>(j) + 1] - cy;
          const double dz = tpos[3 * static_cast<size_t>(j) + 2] - cz;
          const double dd = dx * dx + dy * dy + dz * dz;
          if (pass && dd <= r2) {
            if (kMode != kCount) {
              if (at < end) {
                index[at] = order[j];
                dist2[at] = dd;
              }
              ++at;
            }
            ++found;
          }
        }
      }
      cur = g.skip;
      d = g.skip_depth;
    } else if (passed) {
      stack[d + 1] = passed;  // lanes active below this node
      cur = g.first + g.children - 1;  // the last child first
      ++d;
    } else {
      cur = g.skip;
      d = g.skip_depth;
    }
  }
  if (!ok) return;
  if (kMode == kFill) {
    if (at != end) atomicOr(error, 2ull);  // the passes disagree (never)
  } else {
    counts[q] = found;
  }
}

// Staged lists to their CSR places: thread (q, i) copies entry i of query q
// when q has at most `stride` entries; queries with more are appended to
// `over` (their lists are written by a kFill walk).
__global__ __launch_bounds__(kBlock) void unstage(const uint32_t *__restrict__ counts,
                                                  const uint32_t *__restrict__ offs, uint32_t nq,
                                                  uint32_t stride,
                                                  const unsigned long long *__restrict__ s_index,
                                                  const double *__restrict__ s_dist2,
                                                  unsigned long long *__restrict__ index,
                                                  double *__restrict__ dist2,
                                                  uint32_t *__restrict__ over,
                                                  unsigned long long *__restrict__ n_over) {
  const size_t total = static_cast<size_t>(nq) * stride;
  for (size_t t = static_cast<size_t>(blockIdx.x) * kBlock + threadIdx.x; t < total;
       t += static_cast<size_t>(gridDim.x) * kBlock) {
    const uint32_t q = static_cast<uint32_t>(t / stride), i = static_cast<uint32_t>(t % stride);
    const uint32_t c = counts[q];
    if (c > stride) {
      if (i == 0) over[atomicAdd(n_over, 1ull)] = q;
    } else if (i < c) {
      index[static_cast<size_t>(offs[q]) + i] = s_index[t];
      dist2[static_cast<size_t>(offs[q]) + i] = s_dist2[t];
    }
  }
}

#pragma clang fp contract(fast)

// Spreads the low 10 bits of v to every third bit.
__device__ inline uint64_t spread3(uint32_t v) {
  uint64_t x = v & 0x3ffu;
  x = (x | (x << 16)) & 0x30000ffull;
  x = (x | (x << 8)) & 0x300f00full;
  x = (x | (x << 4)) & 0x30c30c3ull;
  x = (x | (x << 2)) & 0x9249249ull;
  return x;
}

// 10 bits of a coordinate's place in [lo, lo + 2 half] (clamped; NaN gives 0).
__device__ inline uint32_t cell10(double x, double lo, double inv_side) {
  const double t = (x - lo) * inv_side * 1024.0;
  return t > 0.0 ? static_cast<uint32_t>(fmin(t, 1023.0)) : 0u;
}

__global__ __launch_bounds__(kBlock) void query_keys(const double *__restrict__ geom,
                                                     const unsigned long long *__restrict__ frame,
                                                     const double *__restrict__ centres,
                                                     uint32_t nq, uint64_t *__restrict__ keys,
                                                     uint32_t *__restrict__ vals) {
  const uint32_t q = blockIdx.x * kBlock + threadIdx.x;
  if (q >= nq) return;
  uint64_t key = 0;
  if (frame[frame::kNodes] > 0) {
    const double half = geom[3];
    const double inv = half > 0.0 ? 0.5 / half : 0.0;
    const uint32_t x = cell10(centres[3 * static_cast<size_t>(q)], geom[0] - half, inv);
    const uint32_t y = cell10(centres[3 * static_cast<size_t>(q) + 1], geom[1] - half, inv);
    const uint32_t z = cell10(centres[3 * static_cast<size_t>(q) + 2], geom[2] - half, inv);
    key = spread3(x) | (spread3(y) << 1) | (spread3(z) << 2);
  }
  keys[q] = key;
  vals[q] = q;
}

}  // namespace

void launch_query_keys(const double *geom, const unsigned long long *frame,
                       const double *centres, uint32_t nq, uint64_t *keys, uint32_t *vals,
                       hipStream_t s) {
  if (!nq) return;
  hipLaunchKernelGGL(query_keys, dim3((nq + kBlock - 1) / kBlock), dim3(kBlock), 0, s, geom,
                     frame, centres, nq, keys, vals);
  GTH_LAUNCH_CHECK();
}

size_t nb_record_bytes(uint32_t node_cap) {
  return static_cast<size_t>(node_cap) * sizeof(NbNode);
}

void launch_nb_records(const double *geom, const int4 *lnk, const uint32_t *skip,
                       const uint32_t *depth, const unsigned long long *frame, uint32_t node_cap,
                       void *records, hipStream_t s) {
  if (!node_cap) return;
  hipLaunchKernelGGL(nb_records, dim3((node_cap + kBlock - 1) / kBlock), dim3(kBlock), 0, s, geom,
                     lnk, skip, depth, frame, node_cap, static_cast<NbNode *>(records));
  GTH_LAUNCH_CHECK();
}

void launch_within(const void *records, const double *tpos, const uint32_t *order, uint32_t n,
                   const unsigned long long *frame, uint32_t node_cap, const double *centres,
                   const double *radii, uint32_t radius_stride, uint32_t nq, uint32_t lanes,
                   const uint32_t *perm, uint32_t *counts, const uint32_t *offs,
                   unsigned long long *index, double *dist2, uint32_t stage,
                   unsigned long long *error, hipStream_t s) {
  if (!lanes) return;
  const dim3 grid((lanes + kBlock - 1) / kBlock);
  const NbNode *nodes = static_cast<const NbNode *>(records);
#define GTH_WITHIN(M)                                                                          \
  hipLaunchKernelGGL(within<M>, grid, dim3(kBlock), 0, s, nodes, tpos, order, n, frame, node_cap, \
                     centres, radii, radius_stride, nq, lanes, perm, counts, offs, index, dist2,  \
                     stage,                                                                     \
                     error)
  if (stage)
    GTH_WITHIN(kStage);
  else if (offs)
    GTH_WITHIN(kFill);
  else
    GTH_WITHIN(kCount);
#undef GTH_WITHIN
  GTH_LAUNCH_CHECK();
}

void launch_unstage(const uint32_t *counts, const uint32_t *offs, uint32_t nq, uint32_t stage,
                    const unsigned long long *s_index, const double *s_dist2,
                    unsigned long long *index, double *dist2, uint32_t *over,
                    unsigned long long *n_over, hipStream_t s) {
  if (!nq || !stage) return;
  const size_t total = static_cast<size_t>(nq) * stage;
  const uint32_t grid = static_cast<uint32_t>(std::min<size_t>((total + kBlock - 1) / kBlock, 65536));
  hipLaunchKernelGGL(unstage, dim3(grid), dim3(kBlock), 0, s, counts, offs, nq, stage, s_index,
                     s_dist2, index, dist2, over, n_over);
  GTH_LAUNCH_CHECK();
}

}  // namespace gth
