// Interaction lists on the GPU (reference: build_group_lists,
// src/group_lists.cpp:121-310).
//
//  prep_lists    FP32 bodies and cells relative to the root centre, and the
//                walk nodes with crit = 2 half / theta + delta (:143-166).
//  link_up       parent of every node, then `skip`: the node a LIFO depth-
//                first walk pops after finishing a node's subtree (previous
//                sibling, else the parent's skip).  With it the walk needs
//                no stack: open = go to the last child, done = go to skip.
//  group_starts  group selection (:168-216) as flags on tree positions: every
//                node holding more than group_size bodies is one the
//                reference's depth-first `visit` reaches, and each such node
//                marks where its runs of small children (packed greedily while
//                they fit) and its chunks start.  A scan turns the flags into
//                groups in tree order.
//  walk<FILL>    one thread per group: group bounding box, then the depth-
//                first walk with the FP64 opening test (:227-268) evaluated
//                in the reference's operation order with no FMA contraction,
//                so accept/open decisions, and hence the lists, are identical.
//                A counting pass sizes the lists, a scan places them and a
//                filling pass writes them.
//  list_stats    interaction counts, longest list and largest group (:290-309).
#include "gt_internal.h"

namespace gth {
namespace {

constexpr int kBlock = 256;

inline uint32_t blocks_for(size_t n) { return static_cast<uint32_t>((n + kBlock - 1) / kBlock); }

#pragma clang fp contract(off)

__global__ __launch_bounds__(kBlock) void prep_lists(
    const double *__restrict__ geom, const double *__restrict__ tpos,
    const double *__restrict__ tmass, uint32_t n, uint32_t n_nodes, double inv_theta,
    WalkNode *__restrict__ wn, Cell *__restrict__ cells, float4 *__restrict__ bodies) {
  const uint32_t k = blockIdx.x * kBlock + threadIdx.x;
  const double ox = geom[0], oy = geom[1], oz = geom[2];
  if (k < n) {
    bodies[k] = make_float4(static_cast<float>(tpos[3 * k] - ox),
                            static_cast<float>(tpos[3 * k + 1] - oy),
                            static_cast<float>(tpos[3 * k + 2] - oz),
                            static_cast<float>(tmass[k]));
  }
  if (k < n_nodes) {
    const double *g = geom + static_cast<size_t>(kGeomStride) * k;
    Cell c;
    c.cm = make_float4(static_cast<float>(g[4] - ox), static_cast<float>(g[5] - oy),
                       static_cast<float>(g[6] - oz), static_cast<float>(g[7]));
    c.q0 = make_float4(static_cast<float>(g[8]), static_cast<float>(g[9]),
                       static_cast<float>(g[10]), static_cast<float>(g[11]));
    c.q1 = make_float4(static_cast<float>(g[12]), static_cast<float>(g[13]), 0.0f, 0.0f);
    cells[k] = c;
    WalkNode w;
    w.cx = g[0];
    w.cy = g[1];
    w.cz = g[2];
    w.half = g[3];
    w.mx = g[4];
    w.my = g[5];
    w.mz = g[6];
    w.crit = 2.0 * g[3] * inv_theta + g[14];
    wn[k] = w;
  }
}

__global__ __launch_bounds__(kBlock) void set_parents(const int4 *__restrict__ lnk,
                                                      uint32_t n_nodes,
                                                      uint32_t *__restrict__ parent) {
  const uint32_t p = blockIdx.x * kBlock + threadIdx.x;
  if (p >= n_nodes) return;
  if (p == 0) parent[0] = kNone;
  const int4 L = lnk[p];
  for (int c = 0; c < L.y; ++c) parent[L.x + c] = p;
}

__global__ __launch_bounds__(kBlock) void set_skip(const int4 *__restrict__ lnk,
                                                   const uint32_t *__restrict__ parent,
                                                   uint32_t n_nodes, uint32_t *__restrict__ skip) {
  const uint32_t x0 = blockIdx.x * kBlock + threadIdx.x;
  if (x0 >= n_nodes) return;
  uint32_t x = x0;
  // Climb while x is the first child of its parent: its subtree ends its parent's.
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

__global__ __launch_bounds__(kBlock) void group_starts(const int4 *__restrict__ lnk,
                                                       uint32_t n_nodes, uint32_t gs, int pack,
                                                       uint32_t *__restrict__ flags) {
  const uint32_t x = blockIdx.x * kBlock + threadIdx.x;
  if (x >= n_nodes) return;
  const int4 L = lnk[x];
  const uint32_t b = static_cast<uint32_t>(L.z), e = static_cast<uint32_t>(L.w);
  if (x != 0 && e - b <= gs) return;  // not reached by visit(): a small child of its parent
  if (e - b <= gs || L.y == 0) {      // chunk(begin, end)
    for (uint32_t s = b; s < e; s += gs) flags[s] = 1u;
    return;
  }
  bool open = false;
  uint32_t run = 0;
  for (int c = 0; c < L.y; ++c) {
    const int4 C = lnk[L.x + c];
    const uint32_t cc = static_cast<uint32_t>(C.w - C.z);
    if (cc <= gs) {
      if (pack && open && run + cc <= gs) {
        run += cc;
      } else {
        flags[C.z] = 1u;
        run = cc;
        open = true;
      }
    } else {
      open = false;  // visited in its own right: it flags its own starts
    }
  }
}

__global__ __launch_bounds__(kBlock) void compact_starts(const uint32_t *__restrict__ flags,
                                                         const uint32_t *__restrict__ offs,
                                                         uint32_t n, uint32_t *__restrict__ starts) {
  const uint32_t p = blockIdx.x * kBlock + threadIdx.x;
  if (p < n && flags[p]) starts[offs[p]] = p;
}

__global__ __launch_bounds__(kBlock) void make_groups(const uint32_t *__restrict__ starts,
                                                      uint32_t ng, uint32_t n,
                                                      gt_group *__restrict__ groups) {
  const uint32_t g = blockIdx.x * kBlock + threadIdx.x;
  if (g >= ng) return;
  gt_group G;
  G.begin = starts[g];
  G.end = g + 1 < ng ? starts[g + 1] : n;
  G.cell_begin = G.cell_end = G.body_begin = G.body_end = 0;
  G.pad0 = G.pad1 = 0;
  groups[g] = G;
}

template <bool FILL>
__global__ __launch_bounds__(kBlock) void walk(const WalkNode *__restrict__ wn,
                                               const int4 *__restrict__ lnk,
                                               const uint32_t *__restrict__ skip,
                                               const double *__restrict__ tpos,
                                               gt_group *__restrict__ groups, uint32_t ng,
                                               const uint32_t *__restrict__ cell_off,
                                               const uint32_t *__restrict__ body_off,
                                               uint32_t *__restrict__ cell_count,
                                               uint32_t *__restrict__ body_count,
                                               uint32_t *__restrict__ cell_list,
                                               uint32_t *__restrict__ body_list, uint32_t nn,
                                               uint32_t *__restrict__ error) {
  const uint32_t g = blockIdx.x * kBlock + threadIdx.x;
  if (g >= ng) return;
  const uint32_t gb = groups[g].begin, ge = groups[g].end;
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
  uint32_t nc = 0, nb = 0;
  uint32_t cpos = 0, bpos = 0;
  if (FILL) {
    cpos = cell_off[g];
    bpos = body_off[g];
  }
  uint32_t cur = 0;
  // A depth-first walk visits each node at most once: more visits than nodes,
  // or an index past the end, means a malformed tree.  Stop and report it.
  uint32_t budget = nn;
  while (cur != kNone) {
    if (cur >= nn || budget-- == 0) {
      atomicOr(error, 1u);
      break;
    }
    const WalkNode nd = wn[cur];
    const int4 L = lnk[cur];
    bool overlap = true;
    double d2 = 0.0;
    {
      if (fabs(nd.cx - gcx) > nd.half * 1.0001 + ghx) overlap = false;
      const double d = fabs(nd.mx - gcx) - ghx;
      if (d > 0) d2 += d * d;
    }
    {
      if (fabs(nd.cy - gcy) > nd.half * 1.0001 + ghy) overlap = false;
      const double d = fabs(nd.my - gcy) - ghy;
      if (d > 0) d2 += d * d;
    }
    {
      if (fabs(nd.cz - gcz) > nd.half * 1.0001 + ghz) overlap = false;
      const double d = fabs(nd.mz - gcz) - ghz;
      if (d > 0) d2 += d * d;
    }
    const uint32_t b = static_cast<uint32_t>(L.z), e = static_cast<uint32_t>(L.w);
    const bool accept = !overlap && d2 > nd.crit * nd.crit;
    if (accept && e - b > kBodiesIfAtMost) {
      if (FILL) cell_list[cpos + nc] = cur;
      ++nc;
      cur = skip[cur];
    } else if (accept || L.y == 0) {
      if (FILL)
        for (uint32_t j = b; j < e; ++j) body_list[bpos + nb + (j - b)] = j;
      nb += e - b;
      cur = skip[cur];
    } else {
      cur = static_cast<uint32_t>(L.x + L.y - 1);
    }
  }
  if (!FILL) {
    cell_count[g] = nc;
    body_count[g] = nb;
  } else {
    gt_group G = groups[g];
    G.cell_begin = cpos;
    G.cell_end = cpos + nc;
    G.body_begin = bpos;
    G.body_end = bpos + nb;
    groups[g] = G;
  }
}

#pragma clang fp contract(fast)

// stats: [0] cell interactions, [1] body interactions, [2] max list, [3] max group
__global__ __launch_bounds__(kBlock) void list_stats(const gt_group *__restrict__ groups,
                                                     uint32_t ng,
                                                     unsigned long long *__restrict__ stats) {
  const uint32_t g = blockIdx.x * kBlock + threadIdx.x;
  unsigned long long ci = 0, bi = 0, ml = 0, mg = 0;
  if (g < ng) {
    const gt_group G = groups[g];
    const unsigned long long t = G.end - G.begin, c = G.cell_end - G.cell_begin,
                             b = G.body_end - G.body_begin;
    ci = t * c;
    bi = t * b - t;
    ml = c + b;
    mg = t;
  }
  for (int off = 32; off > 0; off >>= 1) {
    ci += __shfl_xor(ci, off, 64);
    bi += __shfl_xor(bi, off, 64);
    ml = max(ml, __shfl_xor(ml, off, 64));
    mg = max(mg, __shfl_xor(mg, off, 64));
  }
  if ((threadIdx.x & 63) == 0) {
    atomicAdd(&stats[0], ci);
    atomicAdd(&stats[1], bi);
    atomicMax(&stats[2], ml);
    atomicMax(&stats[3], mg);
  }
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

void launch_prep_lists(const double *geom, const double *tpos, const double *tmass, uint32_t n,
                       uint32_t n_nodes, double theta, WalkNode *wn, Cell *cells, float4 *bodies,
                       hipStream_t s) {
  const uint32_t m = n > n_nodes ? n : n_nodes;
  if (!m) return;
  hipLaunchKernelGGL(prep_lists, dim3(blocks_for(m)), dim3(kBlock), 0, s, geom, tpos, tmass, n,
                     n_nodes, 1.0 / theta, wn, cells, bodies);
  GTH_LAUNCH_CHECK();
}

void launch_link_up(const int4 *lnk, uint32_t n_nodes, uint32_t *parent, uint32_t *skip,
                    hipStream_t s) {
  if (!n_nodes) return;
  hipLaunchKernelGGL(set_parents, dim3(blocks_for(n_nodes)), dim3(kBlock), 0, s, lnk, n_nodes,
                     parent);
  GTH_LAUNCH_CHECK();
  hipLaunchKernelGGL(set_skip, dim3(blocks_for(n_nodes)), dim3(kBlock), 0, s, lnk, parent, n_nodes,
                     skip);
  GTH_LAUNCH_CHECK();
}

void launch_group_starts(const int4 *lnk, uint32_t n_nodes, uint32_t n, uint32_t gs, int pack,
                         uint32_t *flags, hipStream_t s) {
  launch_fill_u32(flags, n, 0u, s);
  if (!n_nodes) return;
  hipLaunchKernelGGL(group_starts, dim3(blocks_for(n_nodes)), dim3(kBlock), 0, s, lnk, n_nodes, gs,
                     pack, flags);
  GTH_LAUNCH_CHECK();
}

void launch_make_groups(const uint32_t *flags, const uint32_t *offs, uint32_t n, uint32_t ng,
                        uint32_t *starts, gt_group *groups, hipStream_t s) {
  if (!n) return;
  hipLaunchKernelGGL(compact_starts, dim3(blocks_for(n)), dim3(kBlock), 0, s, flags, offs, n,
                     starts);
  GTH_LAUNCH_CHECK();
  if (!ng) return;
  hipLaunchKernelGGL(make_groups, dim3(blocks_for(ng)), dim3(kBlock), 0, s, starts, ng, n, groups);
  GTH_LAUNCH_CHECK();
}

void launch_walk(bool fill, const WalkNode *wn, const int4 *lnk, const uint32_t *skip,
                 const double *tpos, gt_group *groups, uint32_t ng, const uint32_t *cell_off,
                 const uint32_t *body_off, uint32_t *cell_count, uint32_t *body_count,
                 uint32_t *cell_list, uint32_t *body_list, uint32_t nn, uint32_t *error,
                 hipStream_t s) {
  if (!ng) return;
  if (fill)
    hipLaunchKernelGGL(walk<true>, dim3(blocks_for(ng)), dim3(kBlock), 0, s, wn, lnk, skip, tpos,
                       groups, ng, cell_off, body_off, cell_count, body_count, cell_list,
                       body_list, nn, error);
  else
    hipLaunchKernelGGL(walk<false>, dim3(blocks_for(ng)), dim3(kBlock), 0, s, wn, lnk, skip, tpos,
                       groups, ng, cell_off, body_off, cell_count, body_count, cell_list,
                       body_list, nn, error);
  GTH_LAUNCH_CHECK();
}

void launch_list_stats(const gt_group *groups, uint32_t ng, unsigned long long *stats,
                       hipStream_t s) {
  launch_fill_u64(stats, 4, 0ull, s);
  if (!ng) return;
  hipLaunchKernelGGL(list_stats, dim3(blocks_for(ng)), dim3(kBlock), 0, s, groups, ng, stats);
  GTH_LAUNCH_CHECK();
}

}  // namespace gth
