// Octree construction on the GPU (reference: Octree::build, src/octree.cpp:169-273).
//
// The reference splits each node by a stable counting sort into octants,
// recursively, until a node has at most leaf_size bodies, is 60 levels deep,
// or is too small to split (:40-44).  The same tree falls out of a sort:
//
//  1. root cube from the bounding box (:184-205);
//  2. for every particle, its octant path: 3-bit octants, 21 levels per word, found
//     by comparing against node centres that are updated with exactly the
//     reference's FP64 operations (centre +- half/2, :131-137), so every
//     comparison matches the reference's; also the first depth at which the
//     path's node is degenerate (:43);
//  3. a stable radix sort of the 63-bit paths (equal paths keep particle
//     order, as the counting sorts do); when a cluster of more than leaf_size
//     bodies shares all the levels computed so far, the next 21 levels are
//     computed and the order refined (rare: coincident or near-coincident
//     particles; at most three words cover the 60-level limit);
//  4. lcp[i] = levels shared by sorted neighbours i, i+1.  The depth-d node
//     containing p holds more than leaf_size bodies iff some window of
//     leaf_size + 1 consecutive particles around p shares >= d levels, so p's
//     leaf depth is 1 + max over those windows of their min lcp (capped by
//     the degenerate depth and 60);
//  5. the nodes that begin at sorted position p are depths lcp[p-1]+1 ..
//     leaf_depth(p): a scan numbers them in preorder, a stable sort by depth
//     lays them out level by level (children of a node are contiguous and in
//     octant order, as the reference requires);
//  6. parents by binary search in the level above, node ends, and bodies in
//     each leaf put back into particle order (the reference's order inside a
//     leaf is the input order, as no split has reordered them);
//  7. FP64 moments level by level from the deepest: leaves sum their bodies
//     in tree order (:46-72), inner nodes combine their children in order
//     (:74-102), with no FMA contraction, so masses, centres of mass,
//     quadrupoles and deltas are bit-identical to the reference's.
//
// A batch of independent systems (particles offsets[s] .. offsets[s + 1])
// builds one forest: each system gets its own root cube, the system id is the
// most significant sort key, and sorted neighbours of different systems share
// no level, so every system's tree is exactly the one it would get alone.
// All of it is enqueued with no host round trip (counts stay in the frame).
#include "gt_build.h"

namespace gth {
namespace {

constexpr int kBlock = 256;
constexpr int kLevels = 21;  // octant digits per 64-bit path
constexpr int kDepthLimit = 60;
constexpr uint8_t kNoDegenerate = 0xff;
constexpr uint8_t kNoShare = 0xff;  // lcp of neighbours in different systems (no common root)

// Shared levels as a signed number: -1 for neighbours of different systems.
__device__ inline int shared(uint8_t v) { return v == kNoShare ? -1 : static_cast<int>(v); }

// Depth of the first node that begins at sorted position p.
__device__ inline int first_level(const uint8_t *lcp, uint32_t p) {
  return p == 0 ? 0 : shared(lcp[p - 1]) + 1;
}

inline uint32_t blocks_for(size_t n) { return static_cast<uint32_t>((n + kBlock - 1) / kBlock); }

#pragma clang fp contract(off)

// ---- 1. root cube ---------------------------------------------------------------

__global__ __launch_bounds__(kBlock) void bbox_partial(const double *__restrict__ pos, uint32_t n,
                                                       double *__restrict__ part) {
  __shared__ double s[6][kBlock];
  double lo[3], hi[3];
  const uint32_t first = blockIdx.x * kBlock + threadIdx.x;
  const uint32_t i0 = first < n ? first : 0;
  for (int a = 0; a < 3; ++a) lo[a] = hi[a] = pos[3 * static_cast<size_t>(i0) + a];
  for (uint32_t i = first; i < n; i += gridDim.x * kBlock)
    for (int a = 0; a < 3; ++a) {
      const double v = pos[3 * static_cast<size_t>(i) + a];
      lo[a] = v < lo[a] ? v : lo[a];
      hi[a] = hi[a] < v ? v : hi[a];
    }
  for (int a = 0; a < 3; ++a) {
    s[a][threadIdx.x] = lo[a];
    s[3 + a][threadIdx.x] = hi[a];
  }
  __syncthreads();
  for (int w = kBlock / 2; w > 0; w >>= 1) {
    if (threadIdx.x < static_cast<unsigned>(w))
      for (int a = 0; a < 3; ++a) {
        const double l = s[a][threadIdx.x + w], h = s[3 + a][threadIdx.x + w];
        if (l < s[a][threadIdx.x]) s[a][threadIdx.x] = l;
        if (s[3 + a][threadIdx.x] < h) s[3 + a][threadIdx.x] = h;
      }
    __syncthreads();
  }
  if (threadIdx.x < 6) part[6 * blockIdx.x + threadIdx.x] = s[threadIdx.x][0];
}

// root[0..2] centre, root[3] half (src/octree.cpp:184-205).
__global__ __launch_bounds__(kBlock) void bbox_final(const double *__restrict__ part,
                                                     uint32_t nparts, double *__restrict__ root) {
  __shared__ double r[6][kBlock];
  double lo[3], hi[3];
  for (int a = 0; a < 3; ++a) {
    lo[a] = part[a];
    hi[a] = part[3 + a];
  }
  for (uint32_t b = threadIdx.x; b < nparts; b += kBlock)
    for (int a = 0; a < 3; ++a) {
      const double l = part[6 * b + a], h = part[6 * b + 3 + a];
      lo[a] = l < lo[a] ? l : lo[a];
      hi[a] = hi[a] < h ? h : hi[a];
    }
  for (int a = 0; a < 3; ++a) {
    r[a][threadIdx.x] = lo[a];
    r[3 + a][threadIdx.x] = hi[a];
  }
  __syncthreads();
  for (int w = kBlock / 2; w > 0; w >>= 1) {
    if (threadIdx.x < static_cast<unsigned>(w))
      for (int a = 0; a < 3; ++a) {
        const double l = r[a][threadIdx.x + w], h = r[3 + a][threadIdx.x + w];
        if (l < r[a][threadIdx.x]) r[a][threadIdx.x] = l;
        if (r[3 + a][threadIdx.x] < h) r[3 + a][threadIdx.x] = h;
      }
    __syncthreads();
  }
  if (threadIdx.x != 0) return;
  double half = 0.0;
  for (int a = 0; a < 3; ++a) {
    const double e = 0.5 * (r[3 + a][0] - r[a][0]);
    half = half < e ? e : half;
  }
  half = half * (1.0 + 1e-9) + 1e-300;
  for (int a = 0; a < 3; ++a) root[a] = 0.5 * (r[a][0] + r[3 + a][0]);
  root[3] = half;
}

// Systems of a batch: particles offsets[s] .. offsets[s + 1] form system s.
__global__ __launch_bounds__(kBlock) void system_ids(const unsigned long long *__restrict__ offsets,
                                                     uint32_t n_sys, uint32_t n,
                                                     uint32_t *__restrict__ sid) {
  const uint32_t i = blockIdx.x * kBlock + threadIdx.x;
  if (i >= n) return;
  uint32_t lo = 0, hi = n_sys - 1;  // last system whose first particle is <= i
  while (lo < hi) {
    const uint32_t mid = (lo + hi + 1) / 2;
    if (offsets[mid] <= i)
      lo = mid;
    else
      hi = mid - 1;
  }
  sid[i] = lo;
}

// One workgroup per system: its root cube, as bbox_final computes it for one.
__global__ __launch_bounds__(kBlock) void bbox_systems(const double *__restrict__ pos,
                                                       const unsigned long long *__restrict__ offsets,
                                                       double *__restrict__ roots) {
  __shared__ double r[6][kBlock];
  const uint32_t s = blockIdx.x;
  const uint32_t b = static_cast<uint32_t>(offsets[s]), e = static_cast<uint32_t>(offsets[s + 1]);
  if (b >= e) {  // empty system: no tree
    if (threadIdx.x < 4) roots[4 * s + threadIdx.x] = 0.0;
    return;
  }
  double lo[3], hi[3];
  for (int a = 0; a < 3; ++a) lo[a] = hi[a] = pos[3 * static_cast<size_t>(b) + a];
  for (uint32_t i = b + threadIdx.x; i < e; i += kBlock)
    for (int a = 0; a < 3; ++a) {
      const double v = pos[3 * static_cast<size_t>(i) + a];
      lo[a] = v < lo[a] ? v : lo[a];
      hi[a] = hi[a] < v ? v : hi[a];
    }
  for (int a = 0; a < 3; ++a) {
    r[a][threadIdx.x] = lo[a];
    r[3 + a][threadIdx.x] = hi[a];
  }
  __syncthreads();
  for (int w = kBlock / 2; w > 0; w >>= 1) {
    if (threadIdx.x < static_cast<unsigned>(w))
      for (int a = 0; a < 3; ++a) {
        const double l = r[a][threadIdx.x + w], h = r[3 + a][threadIdx.x + w];
        if (l < r[a][threadIdx.x]) r[a][threadIdx.x] = l;
        if (r[3 + a][threadIdx.x] < h) r[3 + a][threadIdx.x] = h;
      }
    __syncthreads();
  }
  if (threadIdx.x != 0) return;
  double half = 0.0;
  for (int a = 0; a < 3; ++a) {
    const double ext = 0.5 * (r[3 + a][0] - r[a][0]);
    half = half < ext ? ext : half;
  }
  half = half * (1.0 + 1e-9) + 1e-300;
  for (int a = 0; a < 3; ++a) roots[4 * s + a] = 0.5 * (r[a][0] + r[3 + a][0]);
  roots[4 * s + 3] = half;
}

// ---- 2. octant paths ---------------------------------------------------------------

__device__ inline bool degenerate(double cx, double cy, double cz, double half) {
  return half < 1e-12 * (fabs(cx) + fabs(cy) + fabs(cz) + 1e-300);
}

// Digits of levels [21 word, 21 word + 21) of every particle's octant path
// (the descent from its system's root replays all earlier levels), and the
// first depth <= 21 (word + 1) at which the path's node is degenerate.  Word 0
// also seeds the sort: keys[i] = path >> key_shift (its first levels only),
// vals[i] = i; then only those levels are computed (the word's others are 0,
// and a degenerate depth past them is not looked for: the tree is cut there).
// sid may be null (one system).
__global__ __launch_bounds__(kBlock) void octant_paths(const double *__restrict__ pos, uint32_t n,
                                                       const double *__restrict__ roots,
                                                       const uint32_t *__restrict__ sid, int word,
                                                       uint64_t *__restrict__ wkey, int key_shift,
                                                       uint64_t *__restrict__ keys,
                                                       uint32_t *__restrict__ vals,
                                                       uint8_t *__restrict__ dg) {
  const uint32_t i = blockIdx.x * kBlock + threadIdx.x;
  if (i >= n) return;
  const double x = pos[3 * static_cast<size_t>(i)], y = pos[3 * static_cast<size_t>(i) + 1],
               z = pos[3 * static_cast<size_t>(i) + 2];
  const double *root = roots + 4 * static_cast<size_t>(sid ? sid[i] : 0u);
  double cx = root[0], cy = root[1], cz = root[2], h = root[3];
  uint64_t key = 0;
  uint8_t first_dg = kNoDegenerate;
  // (with key_shift > 0 only the levels the sort orders by: the first
  // kLevels - key_shift / 3)
  const int first = kLevels * word, last = first + kLevels - key_shift / 3;
  for (int l = 0; l < last; ++l) {
    if (first_dg == kNoDegenerate && degenerate(cx, cy, cz, h)) first_dg = static_cast<uint8_t>(l);
    const uint32_t o = (x >= cx ? 1u : 0u) | (y >= cy ? 2u : 0u) | (z >= cz ? 4u : 0u);
    if (l >= first) key = (key << 3) | o;
    const double h2 = 0.5 * h;
    cx = cx + ((o & 1u) ? h2 : -h2);
    cy = cy + ((o & 2u) ? h2 : -h2);
    cz = cz + ((o & 4u) ? h2 : -h2);
    h = h2;
  }
  if (first_dg == kNoDegenerate && degenerate(cx, cy, cz, h)) first_dg = static_cast<uint8_t>(last);
  wkey[i] = key << key_shift;
  dg[i] = first_dg;
  if (keys) {
    keys[i] = key;
    vals[i] = i;
  }
}

__global__ __launch_bounds__(kBlock) void gather_keys(const uint64_t *__restrict__ wkey,
                                                      const uint32_t *__restrict__ vals, uint32_t n,
                                                      uint64_t *__restrict__ keys) {
  const uint32_t p = blockIdx.x * kBlock + threadIdx.x;
  if (p < n) keys[p] = wkey[vals[p]];
}

__global__ __launch_bounds__(kBlock) void gather_sid_keys(const uint32_t *__restrict__ sid,
                                                          const uint32_t *__restrict__ vals,
                                                          uint32_t n,
                                                          uint64_t *__restrict__ keys) {
  const uint32_t p = blockIdx.x * kBlock + threadIdx.x;
  if (p < n) keys[p] = sid[vals[p]];
}

// ---- 4. shared prefixes and leaf depths ---------------------------------------------

// Levels shared by the paths of sorted neighbours p, p+1 (at most 21 words),
// at most cap (the levels the sort ordered by).  With one path word and one
// system (sorted_keys set) the sorted keys are those paths' first cap levels,
// in order: read consecutively instead of through vals.
__global__ __launch_bounds__(kBlock) void shared_levels(const uint64_t *__restrict__ w0,
                                                        const uint64_t *__restrict__ w1,
                                                        const uint64_t *__restrict__ w2, int words,
                                                        int cap, const uint32_t *__restrict__ vals,
                                                        const uint32_t *__restrict__ sid,
                                                        const uint64_t *__restrict__ sorted_keys,
                                                        int key_shift, uint32_t n,
                                                        uint8_t *__restrict__ lcp) {
  const uint32_t p = blockIdx.x * kBlock + threadIdx.x;
  if (p + 1 >= n) return;
  if (sorted_keys) {
    const uint64_t d = (sorted_keys[p] ^ sorted_keys[p + 1]) << key_shift;
    const int common = d ? (__clzll(d) - 1) / 3 : kLevels;
    lcp[p] = static_cast<uint8_t>(min(common, cap));
    return;
  }
  const uint32_t a = vals[p], b = vals[p + 1];
  if (sid && sid[a] != sid[b]) {
    lcp[p] = kNoShare;
    return;
  }
  const uint64_t *w[3] = {w0, w1, w2};
  int common = kLevels * words;
  for (int k = 0; k < words; ++k) {
    const uint64_t d = w[k][a] ^ w[k][b];
    if (d) {
      common = kLevels * k + (__clzll(d) - 1) / 3;
      break;
    }
  }
  lcp[p] = static_cast<uint8_t>(min(common, cap));
}

// Leaf depth of sorted position p, and the tile's count of the nodes that
// begin in it per depth (depth-major, for the row scan; workgroup 0 also
// clears the scan's clear_words status words).  best = max over the windows
// j .. j+leaf of leaf+1 particles that contain p and fit in [0, n) of the
// window's min lcp.  Up to kTileLeaf the block puts its lcp values (with leaf
// on either side) in LDS and takes the minima of all windows starting in its
// range once (W), then each position's maximum over its leaf+1 windows; for
// larger leaves, the windows directly.
constexpr uint32_t kTileLeaf = 64;

// Positions per workgroup of leaf_depths and emit_levels (a tile of the
// per-depth node counts): four per thread, thread k's being base + i * kBlock
// + k.
constexpr uint32_t kLevelItems = 4;
constexpr uint32_t kLevelTile = kBlock * kLevelItems;

__global__ __launch_bounds__(kBlock) void leaf_depths(const uint8_t *__restrict__ lcp,
                                                      const uint32_t *__restrict__ vals,
                                                      const uint8_t *__restrict__ dg, uint32_t n,
                                                      uint32_t leaf, int cap,
                                                      uint8_t *__restrict__ ldep, uint32_t rows,
                                                      uint32_t *__restrict__ level_counts,
                                                      unsigned long long *__restrict__ clear,
                                                      uint32_t clear_words,
                                                      uint32_t *__restrict__ unresolved) {
  __shared__ int8_t sl[kLevelTile + 2 * kTileLeaf];  // shared(lcp[k]) for k = base - leaf ..
  __shared__ int8_t win[kLevelTile + kTileLeaf];     // W(j) for j = base - leaf ..
  __shared__ uint32_t hist[kDepthLimit + 2];         // the tile's nodes per depth
  if (blockIdx.x == 0)
    for (uint32_t t = threadIdx.x; t < clear_words; t += kBlock) clear[t] = 0ull;
  for (uint32_t t = threadIdx.x; t < rows; t += kBlock) hist[t] = 0u;
  const uint32_t base = blockIdx.x * kLevelTile;
  int best[kLevelItems];
#pragma unroll
  for (uint32_t i = 0; i < kLevelItems; ++i) best[i] = -1;
  if (leaf <= kTileLeaf) {
    // windows start at j in [base - leaf, base + kLevelTile) and read lcp[j .. j + leaf - 1]
    const int64_t k0 = static_cast<int64_t>(base) - leaf;
    const uint32_t nk = kLevelTile + 2 * leaf;
    for (uint32_t t = threadIdx.x; t < nk; t += kBlock) {
      const int64_t k = k0 + t;
      sl[t] = static_cast<int8_t>(k >= 0 && k + 1 < static_cast<int64_t>(n) ? shared(lcp[k]) : 0);
    }
    __syncthreads();
    const uint32_t nj = kLevelTile + leaf;
    for (uint32_t t = threadIdx.x; t < nj; t += kBlock) {
      int w = cap;
      for (uint32_t i = 0; i < leaf; ++i) w = min(w, static_cast<int>(sl[t + i]));
      win[t] = static_cast<int8_t>(w);
    }
    __syncthreads();
#pragma unroll
    for (uint32_t i = 0; i < kLevelItems; ++i) {
      const uint32_t p = base + i * kBlock + threadIdx.x;
      if (p < n && n > leaf) {
        const uint32_t jlo = p >= leaf ? p - leaf : 0u;
        const uint32_t jhi = p < n - 1 - leaf ? p : n - 1 - leaf;
        for (uint32_t j = jlo; j <= jhi; ++j)
          best[i] = max(best[i], static_cast<int>(win[j - (base - leaf)]));
      }
    }
  } else {
#pragma unroll
    for (uint32_t i = 0; i < kLevelItems; ++i) {
      const uint32_t p = base + i * kBlock + threadIdx.x;
      if (p < n && n > leaf) {
        const uint32_t jlo = p >= leaf ? p - leaf : 0u;
        const uint32_t jhi = p < n - 1 - leaf ? p : n - 1 - leaf;
        for (uint32_t j = jlo; j <= jhi; ++j) {
          int w = cap;
          for (uint32_t k = j; k < j + leaf; ++k) w = min(w, shared(lcp[k]));
          best[i] = max(best[i], w);
        }
      }
    }
  }
  __syncthreads();  // (hist cleared)
#pragma unroll
  for (uint32_t i = 0; i < kLevelItems; ++i) {
    const uint32_t p = base + i * kBlock + threadIdx.x;
    if (p >= n) continue;
    int d = best[i] + 1;
    d = min(d, static_cast<int>(dg[vals[p]]));
    d = min(d, kDepthLimit);
    // Deeper than the path words computed: reported, and the tree is cut at
    // `cap` levels (nodes exist only where their octant digits do), to be
    // built again with another word.
    if (d > cap) atomicOr(unresolved, 1u);
    d = min(d, cap);
    ldep[p] = static_cast<uint8_t>(d);
    // the nodes beginning at p: depths first .. d
    for (int t = first_level(lcp, p); t <= d; ++t) atomicAdd(&hist[t], 1u);
  }
  __syncthreads();
  for (uint32_t t = threadIdx.x; t < rows; t += kBlock)
    level_counts[static_cast<size_t>(t) * gridDim.x + blockIdx.x] = hist[t];
}

// ---- 5-6. nodes in level order, links, geometry ---------------------------------------

// Nodes straight into level order: position p (sorted) begins the nodes of
// depths first_level(p) .. ldep[p] (none when first > ldep), and a level's
// nodes are in order of the position they begin at, which is the reference's
// child order (octant order) under every parent.  level_offsets, the row scan
// of the tiles' node counts per depth (rows = deepest possible level + 2; the
// last row is empty, so its start is the node total), gives the index of the
// tile's first node of each depth; a node's rank among its tile's nodes of
// that depth comes from wave ballots over the depths the wave's nodes span.
// Per node it writes its depth, its begin, its parent (p's node one level up,
// or for the first node at p the last node one level up that begins before
// p: the one containing p), its first child (p's node one level down; a node
// at p's leaf depth is a leaf) and its skip, the node a LIFO depth-first walk
// takes after finishing the node's subtree (its previous sibling, else its
// parent's skip).  Position p also ends its nodes deeper than the levels it
// shares with p + 1 (the nodes of those depths that contain p: the one that
// begins at p, or else the last one before it), and writes their end, p + 1.
// Workgroup 0 writes the level starts
// (lstart[d]; levels past the deepest 0, lstart[deepest + 1] = the total), the
// level sizes, the node total and the depth into the frame.  A tree with more
// nodes than node_cap is not written.
constexpr int kBuildWaves = kBlock / 64;
constexpr int kRankSlots = kLevelItems * kBuildWaves;  // (item, wave) in position order

__device__ inline int wave_min(int v) {
  for (int off = 32; off > 0; off >>= 1) v = min(v, __shfl_xor(v, off, 64));
  return v;
}
__device__ inline int wave_max(int v) {
  for (int off = 32; off > 0; off >>= 1) v = max(v, __shfl_xor(v, off, 64));
  return v;
}

// A node's first child, child count (0 here) and begin, leaving its end.
__device__ inline void put_link(int4 *l, int first_child, int begin) {
  int *w = reinterpret_cast<int *>(l);
  w[0] = first_child;
  w[1] = 0;
  w[2] = begin;
}

__global__ __launch_bounds__(kBlock) void emit_levels(
    const uint8_t *__restrict__ lcp, const uint8_t *__restrict__ ldep, uint32_t n,
    uint32_t rows, const uint32_t *__restrict__ level_offsets,
    unsigned long long *__restrict__ frame, uint32_t node_cap, uint32_t *__restrict__ lstart,
    int4 *__restrict__ lnk, uint32_t *__restrict__ parent, uint32_t *__restrict__ node_begin,
    uint32_t *__restrict__ node_depth, uint32_t *__restrict__ skip,
    uint32_t *__restrict__ group_flags) {
  __shared__ uint32_t rank0[kRankSlots][kDepthLimit + 1];  // counts, then their prefix
  const uint32_t ntiles = gridDim.x, tile = blockIdx.x;
  const int tmax = static_cast<int>(rows) - 2;  // deepest possible level
  const uint32_t total = level_offsets[static_cast<size_t>(rows - 1) * ntiles];
  if (tile == 0) {
    // the deepest level with nodes (every thread finds it; rows <= 62)
    int deepest = 0;
    for (int t = 1; t <= tmax; ++t)
      if (level_offsets[static_cast<size_t>(t + 1) * ntiles] >
          level_offsets[static_cast<size_t>(t) * ntiles])
        deepest = t;
    for (int t = static_cast<int>(threadIdx.x); t < kDepthLimit + 2; t += kBlock)
      lstart[t] = t <= deepest + 1 ? level_offsets[static_cast<size_t>(t) * ntiles] : 0u;
    for (int t = static_cast<int>(threadIdx.x); t < frame::kLevels; t += kBlock)
      frame[frame::kLevelNodes + t] =
          t <= deepest ? level_offsets[static_cast<size_t>(t + 1) * ntiles] -
                             level_offsets[static_cast<size_t>(t) * ntiles]
                       : 0u;
    if (threadIdx.x == 0) {
      frame[frame::kNodes] = total;
      frame[frame::kDepth] = static_cast<unsigned long long>(deepest);
    }
  }
  if (group_flags)  // (cleared for node_links to flag the group starts)
    for (uint32_t i = 0; i < kLevelItems; ++i) {
      const uint32_t p = tile * kLevelTile + i * kBlock + threadIdx.x;
      if (p < n) group_flags[p] = 0u;
    }
  if (total > node_cap) return;  // over capacity: no tree (the whole grid)
  for (uint32_t k = threadIdx.x; k < kRankSlots * (kDepthLimit + 1); k += kBlock)
    (&rank0[0][0])[k] = 0u;
  const uint32_t lane = threadIdx.x % 64, wave = threadIdx.x / 64;
  const uint32_t base = tile * kLevelTile + threadIdx.x;
  int first[kLevelItems], d[kLevelItems], last[kLevelItems];
  int lo = kDepthLimit + 1, hi = -1;  // the depths of the nodes that begin in the wave
  int end_lo = kDepthLimit + 1, end_hi = -1;  // and of those that end in it
#pragma unroll
  for (uint32_t i = 0; i < kLevelItems; ++i) {
    const uint32_t p = base + i * kBlock;
    first[i] = 1;
    d[i] = 0;  // (no nodes past n)
    last[i] = 1;
    if (p < n) {
      first[i] = first_level(lcp, p);
      d[i] = ldep[p];
      // p is the last body of its nodes deeper than the levels it shares
      // with p + 1 (all of them at the end, or at a system's end)
      last[i] = p + 1 < n ? shared(lcp[p]) + 1 : 0;
    }
    if (first[i] <= d[i]) {
      lo = min(lo, first[i]);
      hi = max(hi, d[i]);
    }
    if (last[i] <= d[i]) {
      end_lo = min(end_lo, last[i]);
      end_hi = max(end_hi, d[i]);
    }
  }
  lo = wave_min(lo);
  hi = wave_max(hi);
  end_lo = wave_min(end_lo);
  end_hi = wave_max(end_hi);
  __syncthreads();  // (counts cleared)
#pragma unroll
  for (uint32_t i = 0; i < kLevelItems; ++i)
    for (int t = lo; t <= hi; ++t) {
      const uint64_t m = __ballot(first[i] <= t && t <= d[i]);
      if (lane == 0) rank0[i * kBuildWaves + wave][t] = static_cast<uint32_t>(__popcll(m));
    }
  __syncthreads();
  for (int t = static_cast<int>(threadIdx.x); t <= tmax; t += kBlock) {
    uint32_t run = 0;
    for (int k = 0; k < kRankSlots; ++k) {
      const uint32_t c = rank0[k][t];
      rank0[k][t] = run;
      run += c;
    }
  }
  __syncthreads();
  if (hi < 0 && end_hi < 0) return;  // no nodes begin or end in this wave
  const uint64_t below = (1ull << lane) - 1ull;  // lanes before this one
  // from the level above the first nodes that begin (their parents), or the
  // first level of a node that ends
  const int t0 = min(lo > 0 ? lo - 1 : 0, end_lo), t1 = max(hi, end_hi);
#pragma unroll
  for (uint32_t i = 0; i < kLevelItems; ++i) {
    const uint32_t p = base + i * kBlock;
    uint32_t above = kNone;  // the last node of depth t - 1 that begins at or before p
    uint32_t prev = kNone;   // p's node of depth t - 1
    uint32_t after = kNone;  // what a LIFO walk takes after the subtree of p's nodes
    for (int t = t0; t <= t1; ++t) {
      const bool in = first[i] <= t && t <= d[i];
      const uint64_t m = __ballot(in);
      const uint32_t at = level_offsets[static_cast<size_t>(t) * ntiles + tile] +
                          rank0[i * kBuildWaves + wave][t] +
                          static_cast<uint32_t>(__popcll(m & below));
      if (in) {  // p's node of depth t is node `at`
        // p's deeper nodes are first children (they begin where their parents
        // do), so the walk leaves all of them for the same node: the previous
        // sibling of p's first node (node at - 1), none for a root
        if (t == first[i]) after = t == 0 ? kNone : at - 1u;
        parent[at] = t == 0 ? kNone : above;
        node_begin[at] = p;
        node_depth[at] = static_cast<uint32_t>(t);
        skip[at] = after;
        // (first child, child count, begin: the end is another position's)
        if (t > first[i]) put_link(lnk + prev, static_cast<int>(at), static_cast<int>(p));
        if (t == d[i]) put_link(lnk + at, -1, static_cast<int>(p));
        prev = at;
      }
      // p's node of depth t ends after p: the one that begins at p, or else
      // the last that began before it
      if (last[i] <= t && t <= d[i])
        reinterpret_cast<int *>(lnk + (in ? at : at - 1u))[3] = static_cast<int>(p + 1);
      above = in ? at : at - 1u;  // (at - 1: this tile's or an earlier tile's last before p)
    }
  }
}

// Bodies of a leaf back into particle order (insertion sort; leaves are small).
constexpr uint32_t kLeafRegs = 16;  // leaves up to this size are sorted in registers
// orig (may be null: the identity) maps the positions in pos to the caller's
// particle indices, which decide the order.
__device__ inline void sort_leaf(uint32_t *__restrict__ vals, const uint32_t *__restrict__ orig,
                                 uint32_t b, uint32_t e) {
  if (e - b <= kLeafRegs) {  // the leaf in registers: its loads issue together
    uint32_t v[kLeafRegs], key[kLeafRegs];
#pragma unroll
    for (uint32_t i = 0; i < kLeafRegs; ++i)
      if (b + i < e) v[i] = vals[b + i];
#pragma unroll
    for (uint32_t i = 0; i < kLeafRegs; ++i)
      if (b + i < e) key[i] = orig ? orig[v[i]] : v[i];
    bool moved = false;
#pragma unroll
    for (uint32_t i = 1; i < kLeafRegs; ++i) {  // insertion sort by key, stable
#pragma unroll
      for (uint32_t j = i; j > 0; --j) {
        if (b + i < e && key[j - 1] > key[j]) {
          const uint32_t tk = key[j - 1], tv = v[j - 1];
          key[j - 1] = key[j];
          v[j - 1] = v[j];
          key[j] = tk;
          v[j] = tv;
          moved = true;
        }
      }
    }
    if (moved) {
#pragma unroll
      for (uint32_t i = 0; i < kLeafRegs; ++i)
        if (b + i < e) vals[b + i] = v[i];
    }
    return;
  }
  for (uint32_t k = b + 1; k < e; ++k) {
    const uint32_t v = vals[k], key = orig ? orig[v] : v;
    uint32_t j = k;
    while (j > b && (orig ? orig[vals[j - 1]] : vals[j - 1]) > key) {
      vals[j] = vals[j - 1];
      --j;
    }
    vals[j] = v;
  }
}

// Per node: its centre and half size (the octant path of its first body
// replayed from its system's root with the reference's FP64 centre updates:
// with one path word and one system, sorted_keys set, the path is the sorted
// key at the node's begin shifted back into place, else the path words of
// particle vals[begin]), its child count (its children are consecutive on the
// level below, from its first child on while their parent is the node: at
// most eight, read together as whole aligned 16-byte words), and each
// system's root node; then a leaf's bodies are put back into particle order
// (sort_leaf; the leaf's end is in its link already).  Reading vals[begin]
// for an inner node's path while a leaf that begins there is sorted is
// harmless: every body of the leaf has the node's path.
__global__ __launch_bounds__(kBlock) void node_links(
    const uint32_t *__restrict__ parent, const uint32_t *__restrict__ node_begin,
    const uint32_t *__restrict__ node_depth, const unsigned long long *__restrict__ frame, uint32_t node_cap, uint32_t n,
    const uint64_t *__restrict__ w0, const uint64_t *__restrict__ w1,
    const uint64_t *__restrict__ w2, uint32_t *__restrict__ vals,
    const uint64_t *__restrict__ sorted_keys, int key_shift, const double *__restrict__ roots,
    const uint32_t *__restrict__ sid, double *__restrict__ geom, int4 *__restrict__ lnk,
    uint32_t *__restrict__ root_node, const uint32_t *__restrict__ orig, GroupFlags groups) {
  const uint32_t n_nodes = bounded_count(frame + frame::kNodes, node_cap);
  const uint32_t x = blockIdx.x * kBlock + threadIdx.x;
  if (x >= n_nodes) return;
  const uint32_t d = node_depth[x], b = node_begin[x];
  int4 L = lnk[x];  // (first child, -, begin, end)
  const int first = L.x;
  const bool is_root = parent[x] == kNone;
  // centre and half size
  uint64_t kw[3] = {0ull, 0ull, 0ull};
  uint32_t system = 0;
  if (sorted_keys) {
    kw[0] = sorted_keys[b] << key_shift;
  } else {
    const uint32_t i = vals[b];
    system = sid ? sid[i] : 0u;
    kw[0] = w0[i];
    kw[1] = d > static_cast<uint32_t>(kLevels) ? w1[i] : 0ull;
    kw[2] = d > 2u * kLevels ? w2[i] : 0ull;
  }
  const double *root = roots + 4 * static_cast<size_t>(system);
  double cx = root[0], cy = root[1], cz = root[2], h = root[3];
  for (uint32_t l = 0; l < d; ++l) {
    const uint64_t key = kw[l / kLevels];
    const uint32_t o = static_cast<uint32_t>(key >> (3 * (kLevels - 1 - l % kLevels))) & 7u;
    const double h2 = 0.5 * h;
    cx = cx + ((o & 1u) ? h2 : -h2);
    cy = cy + ((o & 2u) ? h2 : -h2);
    cz = cz + ((o & 4u) ? h2 : -h2);
    h = h2;
  }
  double *g = geom + static_cast<size_t>(kGeomStride) * x;
  g[0] = cx;
  g[1] = cy;
  g[2] = cz;
  g[3] = h;
  g[15] = 0.0;
  if (d == 0 && root_node) root_node[system] = x;
  if (first >= 0) {  // children: parents of nodes c0 + 1 .. c0 + 7, read as three
                     // aligned 16-byte words (parent is padded past node_cap)
    const uint32_t c0 = static_cast<uint32_t>(first), a = (c0 + 1u) & ~3u;
    const uint4 *pw = reinterpret_cast<const uint4 *>(parent + a);
    const uint4 w0 = pw[0], w1 = pw[1], w2 = pw[2];
    const uint32_t up[12] = {w0.x, w0.y, w0.z, w0.w, w1.x, w1.y,
                             w1.z, w1.w, w2.x, w2.y, w2.z, w2.w};
    int children = 1;
#pragma unroll
    for (uint32_t k = 0; k < 7; ++k) {
      const uint32_t c = c0 + 1u + k;
      uint32_t pc = up[0];
#pragma unroll
      for (uint32_t j = 1; j < 12; ++j) pc = c - a == j ? up[j] : pc;
      if (children == static_cast<int>(k) + 1 && c < n_nodes && pc == x) ++children;
    }
    lnk[x].y = children;
    L.y = children;
    // the group starts it flags (its children's begins and ends are theirs
    // from emit_levels already)
    if (groups.flags)
      flag_group_starts(x, L, is_root, groups.gs, groups.pack, groups.flags, groups.node_of,
                        [&](int c) { return lnk[first + c]; });
    return;
  }
  // a leaf (lnk[x].y stays 0): its group starts, then its bodies back into
  // particle order
  if (groups.flags)
    flag_group_starts(x, L, is_root, groups.gs, groups.pack, groups.flags, groups.node_of,
                      [&](int c) { return lnk[first + c]; });
  sort_leaf(vals, orig, b, static_cast<uint32_t>(L.w));
}

// Positions and masses into tree order (tpos/tmass, and the system of each
// tree position for a batch), each particle's tree position (inv; null: not
// needed), and with records the FP32 bodies.  The positions go out through LDS: a block's 256
// rows of three doubles are written as 768 consecutive doubles, one per lane
// per store.
__global__ __launch_bounds__(kBlock) void gather_tree_order(const uint32_t *__restrict__ vals,
                                                            const double *__restrict__ pos,
                                                            const double *__restrict__ mass,
                                                            const uint32_t *__restrict__ sid,
                                                            uint32_t n, double *__restrict__ tpos,
                                                            double *__restrict__ tmass,
                                                            uint32_t *__restrict__ tsid,
                                                            uint32_t *__restrict__ inv,
                                                            const double *__restrict__ roots,
                                                            ListRecords R) {
  __shared__ double rows[3 * kBlock];
  const uint32_t k0 = blockIdx.x * kBlock, k = k0 + threadIdx.x;
  if (k < n) {
    const uint32_t i = vals[k];
    if (inv) inv[i] = k;
    const uint32_t system = sid ? sid[i] : 0u;
    if (sid) tsid[k] = system;
    const double x = pos[3 * static_cast<size_t>(i)], y = pos[3 * static_cast<size_t>(i) + 1],
                 z = pos[3 * static_cast<size_t>(i) + 2], m = mass[i];
    rows[3 * threadIdx.x] = x;
    rows[3 * threadIdx.x + 1] = y;
    rows[3 * threadIdx.x + 2] = z;
    tmass[k] = m;
    if (R.bodies) {
      const double *o = roots + 4 * static_cast<size_t>(system);
      write_body_record(R, k, x, y, z, m, o[0], o[1], o[2]);
    }
  }
  __syncthreads();
  const uint32_t in_block = n - k0 < kBlock ? n - k0 : kBlock;
  for (uint32_t j = threadIdx.x; j < 3 * in_block; j += kBlock)
    tpos[3 * static_cast<size_t>(k0) + j] = rows[j];
}

// ---- 7. moments, one level at a time -------------------------------------------------

// Moments of node x from its bodies (leaf) or its children, which are on the
// level below and already done.  The data of up to eight bodies or children
// are loaded together before they are summed (one round of loads instead of
// one per term: the level launches are latency-bound), and summed one by one
// in order, so every sum is the reference's sequence.
constexpr int kChunk = 8;

// With records (R.cells set: a build for the lists) the node's walk and cell
// records are written too (Ext: what they need besides the moments).
struct RecordExt {
  ListRecords R;
  const uint32_t *skip, *depth, *tsid;
  const double *roots;
  unsigned long long *walk_error;
};

__device__ __attribute__((always_inline)) inline void node_moments(uint32_t x, uint32_t n_nodes, int4 *__restrict__ lnk,
                                    const double *__restrict__ tpos,
                                    const double *__restrict__ tmass,
                                    double *__restrict__ geom, const RecordExt &E) {
  int4 L = lnk[x];
  double *g = geom + static_cast<size_t>(kGeomStride) * x;
  const double ccx = g[0], ccy = g[1], ccz = g[2];
  // the records' other inputs, loaded before this node's stores (which the
  // compiler may not move them past)
  uint32_t skip_x = 0, depth_x = 0;
  double o[4] = {0.0, 0.0, 0.0, 0.0};
  if (E.R.cells) {
    skip_x = E.skip[x];
    depth_x = E.depth[x];
    const double *r = E.roots;  // one system: its root, not waiting for the links
    if (E.tsid) r += 4 * static_cast<size_t>(E.tsid[L.z]);
    for (int j = 0; j < 4; ++j) o[j] = r[j];
  }
  double m = 0.0, cx = 0.0, cy = 0.0, cz = 0.0;
  double q[6] = {0.0, 0.0, 0.0, 0.0, 0.0, 0.0};
  if (L.y == 0) {
    const uint32_t b = static_cast<uint32_t>(L.z), e = static_cast<uint32_t>(L.w);
    for (uint32_t k0 = b; k0 < e; k0 += kChunk) {
      double mi[kChunk], px[kChunk], py[kChunk], pz[kChunk];
#pragma unroll
      for (int i = 0; i < kChunk; ++i) {
        const uint32_t k = k0 + i;
        if (k < e) {
          mi[i] = tmass[k];
          px[i] = tpos[3 * static_cast<size_t>(k)];
          py[i] = tpos[3 * static_cast<size_t>(k) + 1];
          pz[i] = tpos[3 * static_cast<size_t>(k) + 2];
        }
      }
#pragma unroll
      for (int i = 0; i < kChunk; ++i) {
        if (k0 + i < e) {
          m += mi[i];
          cx += mi[i] * px[i];
          cy += mi[i] * py[i];
          cz += mi[i] * pz[i];
        }
      }
    }
    if (m > 0) {
      cx = __ddiv_rn(cx, m);
      cy = __ddiv_rn(cy, m);
      cz = __ddiv_rn(cz, m);
    } else {
      cx = ccx;
      cy = ccy;
      cz = ccz;
    }
    for (uint32_t k0 = b; k0 < e; k0 += kChunk) {
      double mi[kChunk], px[kChunk], py[kChunk], pz[kChunk];
#pragma unroll
      for (int i = 0; i < kChunk; ++i) {
        const uint32_t k = k0 + i;
        if (k < e) {
          mi[i] = tmass[k];
          px[i] = tpos[3 * static_cast<size_t>(k)];
          py[i] = tpos[3 * static_cast<size_t>(k) + 1];
          pz[i] = tpos[3 * static_cast<size_t>(k) + 2];
        }
      }
#pragma unroll
      for (int i = 0; i < kChunk; ++i) {
        if (k0 + i < e) {
          const double dx = px[i] - cx, dy = py[i] - cy, dz = pz[i] - cz;
          const double r2 = dx * dx + dy * dy + dz * dz;
          q[0] += mi[i] * (3 * dx * dx - r2);
          q[1] += mi[i] * (3 * dy * dy - r2);
          q[2] += mi[i] * (3 * dz * dz - r2);
          q[3] += mi[i] * 3 * dx * dy;
          q[4] += mi[i] * 3 * dx * dz;
          q[5] += mi[i] * 3 * dy * dz;
        }
      }
    }
  } else {
    const uint32_t c0 = static_cast<uint32_t>(L.x), nc = static_cast<uint32_t>(L.y);
    if (L.x < 0 || nc > kChunk || c0 + nc > n_nodes) return;  // malformed links: leave the row unset
    double hm[kChunk], hx[kChunk], hy[kChunk], hz[kChunk], hq[kChunk][6];
#pragma unroll
    for (int i = 0; i < kChunk; ++i) {
      if (static_cast<uint32_t>(i) < nc) {
        const double *h = geom + static_cast<size_t>(kGeomStride) * (c0 + i);
        hm[i] = h[7];
        hx[i] = h[4];
        hy[i] = h[5];
        hz[i] = h[6];
#pragma unroll
        for (int j = 0; j < 6; ++j) hq[i][j] = h[8 + j];
      }
    }
#pragma unroll
    for (int i = 0; i < kChunk; ++i) {
      if (static_cast<uint32_t>(i) < nc) {
        m += hm[i];
        cx += hm[i] * hx[i];
        cy += hm[i] * hy[i];
        cz += hm[i] * hz[i];
      }
    }
    if (m > 0) {
      cx = __ddiv_rn(cx, m);
      cy = __ddiv_rn(cy, m);
      cz = __ddiv_rn(cz, m);
    } else {
      cx = ccx;
      cy = ccy;
      cz = ccz;
    }
#pragma unroll
    for (int i = 0; i < kChunk; ++i) {
      if (static_cast<uint32_t>(i) < nc) {
        const double mc = hm[i];
        const double dx = hx[i] - cx, dy = hy[i] - cy, dz = hz[i] - cz;
        const double r2 = dx * dx + dy * dy + dz * dz;
        q[0] += hq[i][0] + mc * (3 * dx * dx - r2);
        q[1] += hq[i][1] + mc * (3 * dy * dy - r2);
        q[2] += hq[i][2] + mc * (3 * dz * dz - r2);
        q[3] += hq[i][3] + mc * 3 * dx * dy;
        q[4] += hq[i][4] + mc * 3 * dx * dz;
        q[5] += hq[i][5] + mc * 3 * dy * dz;
      }
    }
  }
  g[4] = cx;
  g[5] = cy;
  g[6] = cz;
  g[7] = m;
  for (int k = 0; k < 6; ++k) g[8 + k] = q[k];
  const double ex = cx - ccx, ey = cy - ccy, ez = cz - ccz;
  const double delta = __dsqrt_rn(ex * ex + ey * ey + ez * ez);
  g[14] = delta;
  if (E.R.cells) {
    const double row[15] = {ccx, ccy, ccz, g[3], cx,   cy,   cz,   m,
                            q[0], q[1], q[2], q[3], q[4], q[5], delta};
    write_node_records(E.R, x, row, L, skip_x, depth_x, o[0], o[1], o[2], o[3], E.walk_error);
  }
}

// Every node of level d: lstart[d] .. lstart[d + 1] (empty for levels past the
// deepest, whose lstart entries are 0).
__global__ __launch_bounds__(kBlock) void level_moments(int d, const uint32_t *__restrict__ lstart,
                                                        const unsigned long long *__restrict__ frame,
                                                        uint32_t node_cap,
                                                        int4 *__restrict__ lnk,
                                                        const double *__restrict__ tpos,
                                                        const double *__restrict__ tmass,
                                                        double *__restrict__ geom, RecordExt E) {
  const uint32_t n_nodes = bounded_count(frame + frame::kNodes, node_cap);
  const uint32_t first = lstart[d], last = min(lstart[d + 1], n_nodes);
  for (uint32_t x = first + blockIdx.x * kBlock + threadIdx.x; x < last; x += gridDim.x * kBlock)
    node_moments(x, n_nodes, lnk, tpos, tmass, geom, E);
}

// Levels d and d - 1 in one launch: eight lanes per node of level d - 1, lane
// j doing the node's child j (on level d), then, after a barrier, the node's
// first lane the node itself.  For small levels, where a launch costs more
// than the level's work.
__global__ __launch_bounds__(kBlock) void level_pair_moments(
    int d, const uint32_t *__restrict__ lstart, const unsigned long long *__restrict__ frame,
    uint32_t node_cap, int4 *__restrict__ lnk, const double *__restrict__ tpos,
    const double *__restrict__ tmass, double *__restrict__ geom, RecordExt E) {
  const uint32_t n_nodes = bounded_count(frame + frame::kNodes, node_cap);
  const uint32_t first = lstart[d - 1], last = min(lstart[d], n_nodes);
  const uint32_t lanes = last > first ? 8 * (last - first) : 0u;
  // every thread of a workgroup takes the same number of rounds (barriers)
  for (uint32_t base = blockIdx.x * kBlock; base < lanes; base += gridDim.x * kBlock) {
    const uint32_t t = base + threadIdx.x;
    const uint32_t x = first + t / 8, j = t % 8;
    if (t < lanes) {
      const int4 L = lnk[x];
      if (L.x >= 0 && j < static_cast<uint32_t>(L.y) && static_cast<uint32_t>(L.x) + j < n_nodes)
        node_moments(static_cast<uint32_t>(L.x) + j, n_nodes, lnk, tpos, tmass, geom, E);
    }
    __syncthreads();
    if (t < lanes && j == 0) node_moments(x, n_nodes, lnk, tpos, tmass, geom, E);
  }
}

// Levels d_hi down to d_lo by one workgroup, with a barrier between levels:
// for the run of small levels at the bottom of the tree, where a launch per
// level costs more than the level's work.  Any level size is
// handled (the loop strides), only more slowly.
constexpr int kSerialBlock = 512;

__global__ __launch_bounds__(kSerialBlock) void serial_level_moments(
    int d_hi, int d_lo, const uint32_t *__restrict__ lstart,
    const unsigned long long *__restrict__ frame, uint32_t node_cap, int4 *__restrict__ lnk,
    const double *__restrict__ tpos, const double *__restrict__ tmass, double *__restrict__ geom,
    RecordExt E) {
  const uint32_t n_nodes = bounded_count(frame + frame::kNodes, node_cap);
  for (int d = d_hi; d >= d_lo; --d) {
    const uint32_t first = lstart[d], last = min(lstart[d + 1], n_nodes);
    for (uint32_t x = first + threadIdx.x; x < last; x += kSerialBlock)
      node_moments(x, n_nodes, lnk, tpos, tmass, geom, E);
    __syncthreads();
  }
}

#pragma clang fp contract(fast)

}  // namespace

void TreeBuilder::reserve(const Plan &p) {
  const size_t n = std::max<uint32_t>(p.n, 1), nn = std::max<uint32_t>(p.node_cap, 1);
  const uint32_t nparts = std::min<uint32_t>(blocks_for(n), 1024u);
  const size_t ns = std::max<uint32_t>(p.systems, 1);
  part_.ensure(6 * static_cast<size_t>(nparts));
  root_.ensure(4 * ns);
  if (p.systems > 1) {
    sid_.ensure(n);
    tsid_.ensure(n);
    root_node_.ensure(ns);
  }
  keys_.ensure(n);
  vals_.ensure(n);
  inv_.ensure(n);
  keys_alt_.ensure(n);
  vals_alt_.ensure(n);
  dg_.ensure(n);
  word0_.ensure(n);
  if (p.words > 1) word1_.ensure(n);
  if (p.words > 2) word2_.ensure(n);
  sort_tmp_.ensure(radix_tmp_words(n));
  lcp_.ensure(n);
  ldep_.ensure(n);
  const size_t level_cells =
      static_cast<size_t>(kDepthLimit + 2) * ((n + kLevelTile - 1) / kLevelTile);
  level_counts_.ensure(level_cells);
  level_offsets_.ensure(level_cells);
  level_scan_.ensure(kDepthLimit + 3);
  nbeg_.ensure(nn);
  node_depth_.ensure(nn);
  skip_.ensure(nn);
  lstart_.ensure(kDepthLimit + 2);
  geom_.ensure(static_cast<size_t>(kGeomStride) * nn);
  lnk_.ensure(nn);
  parent_.ensure(nn + 16);  // (node_links reads whole aligned 16-byte words)
  tpos_.ensure(3 * n);
  tmass_.ensure(n);
}

void TreeBuilder::enqueue(const double *pos, const double *mass, const Plan &p, int leaf_size,
                          const unsigned long long *sys_off, unsigned long long *frame,
                          hipStream_t s, const uint32_t *orig, const ListRecords *rec,
                          const GroupFlags *groups) {
  wrote_records_ = false;
  flagged_groups_ = false;
  const uint32_t n = p.n;
  n_ = n;
  node_cap_ = p.node_cap;
  systems_ = p.systems;
  if (p.systems > 1 && !sys_off) fail(GT_HIP_ERR_ARGUMENT, "octree build: system offsets missing");
  if (n == 0) return;
  if (p.words < 1 || p.words > 3) fail(GT_HIP_ERR_ARGUMENT, "octree build: 1 to 3 path words");
  if (p.depth_cap < 0 || p.depth_cap > kDepthLimit)
    fail(GT_HIP_ERR_ARGUMENT, "octree build: depth bound out of range");
  reserve(p);
  const uint32_t leaf = static_cast<uint32_t>(leaf_size < 1 ? 1 : leaf_size);
  const uint32_t cap = p.node_cap;

  // 1. root cube (of every system of a batch)
  const bool batch = p.systems > 1;
  uint32_t *sid = batch ? sid_.get() : nullptr;
  uint32_t *root_node = batch ? root_node_.get() : nullptr;
  if (batch) {
    hipLaunchKernelGGL(system_ids, dim3(blocks_for(n)), dim3(kBlock), 0, s, sys_off, p.systems, n,
                       sid);
    GTH_LAUNCH_CHECK();
    hipLaunchKernelGGL(bbox_systems, dim3(p.systems), dim3(kBlock), 0, s, pos, sys_off,
                       root_.get());
    GTH_LAUNCH_CHECK();
  } else {
    const uint32_t nparts = std::min<uint32_t>(blocks_for(n), 1024u);
    double *part = part_.get();
    hipLaunchKernelGGL(bbox_partial, dim3(nparts), dim3(kBlock), 0, s, pos, n, part);
    GTH_LAUNCH_CHECK();
    hipLaunchKernelGGL(bbox_final, dim3(1), dim3(kBlock), 0, s, part, nparts, root_.get());
    GTH_LAUNCH_CHECK();
  }

  // 2-3. octant paths (p.words 21-level words of each) and their stable sort:
  // by the last word first, then by each earlier one.  One word is enough
  // unless a cluster of more than leaf_size bodies shares all 21 levels; the
  // leaf depths report it, and the host builds again with another word.  With
  // one word the sort orders by its first p.path_levels levels only: the tree
  // needs the bodies grouped down to the deepest leaf, and the bodies of a
  // leaf go back into particle order anyway (sort_leaf).  A leaf deeper than
  // that is reported the same way, and the host builds again with all 21.
  const int levels = p.words == 1 ? std::min(std::max(p.path_levels, 1), kLevels) : kLevels * p.words;
  uint64_t *keys = keys_.get();
  uint32_t *vals = vals_.get();
  uint8_t *dg = dg_.get();
  uint64_t *w[3] = {word0_.get(), p.words > 1 ? word1_.get() : nullptr,
                    p.words > 2 ? word2_.get() : nullptr};
  hipLaunchKernelGGL(octant_paths, dim3(blocks_for(n)), dim3(kBlock), 0, s, pos, n, root_.get(),
                     sid, 0, w[0], p.words == 1 ? 3 * (kLevels - levels) : 0, keys, vals, dg);
  GTH_LAUNCH_CHECK();
  for (int k = 1; k < p.words; ++k) {
    hipLaunchKernelGGL(octant_paths, dim3(blocks_for(n)), dim3(kBlock), 0, s, pos, n, root_.get(),
                       sid, k, w[k], 0, static_cast<uint64_t *>(nullptr),
                       static_cast<uint32_t *>(nullptr), dg);
    GTH_LAUNCH_CHECK();
  }
  for (int k = p.words - 1; k >= 0; --k) {
    if (p.words > 1) {
      hipLaunchKernelGGL(gather_keys, dim3(blocks_for(n)), dim3(kBlock), 0, s, w[k], vals, n, keys);
      GTH_LAUNCH_CHECK();
    }
    radix_sort_u64(keys, vals, keys_alt_.get(), vals_alt_.get(), n,
                   p.words == 1 ? 3 * levels : 3 * kLevels, sort_tmp_.get(), s,
                   frame + frame::kSortCtl);
  }
  if (batch) {  // systems apart, most significant: one stable pass more
    int bits = 1;
    while (bits < 32 && (p.systems - 1) >> bits) ++bits;
    hipLaunchKernelGGL(gather_sid_keys, dim3(blocks_for(n)), dim3(kBlock), 0, s, sid, vals, n,
                       keys);
    GTH_LAUNCH_CHECK();
    radix_sort_u64(keys, vals, keys_alt_.get(), vals_alt_.get(), n, bits, sort_tmp_.get(), s,
                   frame + frame::kSortCtl);
  }

  // 4. leaf depths, and each tile's node count per depth
  uint8_t *lcp = lcp_.get();
  uint8_t *ldep = ldep_.get();
  hipLaunchKernelGGL(shared_levels, dim3(blocks_for(n)), dim3(kBlock), 0, s, w[0], w[1], w[2],
                     p.words, levels, vals, sid,
                     p.words == 1 && !batch ? keys : static_cast<const uint64_t *>(nullptr),
                     p.words == 1 ? 3 * (kLevels - levels) : 0, n, lcp);
  GTH_LAUNCH_CHECK();
  const uint32_t ntiles = static_cast<uint32_t>((n + kLevelTile - 1) / kLevelTile);
  const uint32_t rows = static_cast<uint32_t>(std::min(levels, kDepthLimit)) + 2u;
  uint32_t *level_counts = level_counts_.get();
  uint32_t *level_offsets = level_offsets_.get();
  unsigned long long *scan_words = level_scan_.get();  // status words, then the ticket
  hipLaunchKernelGGL(leaf_depths, dim3(ntiles), dim3(kBlock), 0, s, lcp, vals, dg, n, leaf, levels,
                     ldep, rows, level_counts, scan_words, rows + 1u,
                     reinterpret_cast<uint32_t *>(frame + frame::kUnresolved));
  GTH_LAUNCH_CHECK();

  // 5-6. nodes in level order, with parents, first children, begins and
  // ends; then per node its geometry and child count, and a leaf's bodies in
  // particle order
  launch_row_scan(rows, ntiles, 1u, reinterpret_cast<uint32_t *>(scan_words + rows), scan_words,
                  frame + frame::kSortCtl, level_counts, level_offsets, s);
  uint32_t *lstart = lstart_.get();
  double *geom = geom_.get();
  int4 *lnk = lnk_.get();
  uint32_t *parent = parent_.get();
  uint32_t *nbeg = nbeg_.get();
  uint32_t *ndep = node_depth_.get();
  hipLaunchKernelGGL(emit_levels, dim3(ntiles), dim3(kBlock), 0, s, lcp, ldep, n, rows,
                     level_offsets, frame, cap, lstart, lnk, parent, nbeg, ndep, skip_.get(),
                     groups ? groups->flags : static_cast<uint32_t *>(nullptr));
  GTH_LAUNCH_CHECK();
  hipLaunchKernelGGL(node_links, dim3(blocks_for(cap)), dim3(kBlock), 0, s, parent, nbeg, ndep,
                     frame, cap, n, w[0], w[1], w[2], vals,
                     p.words == 1 && !batch ? keys : static_cast<const uint64_t *>(nullptr),
                     p.words == 1 ? 3 * (kLevels - levels) : 0, root_.get(), sid, geom, lnk,
                     root_node, orig, groups ? *groups : GroupFlags{});
  GTH_LAUNCH_CHECK();
  flagged_groups_ = groups && groups->flags;
  double *tpos = tpos_.get();
  double *tmass = tmass_.get();
  // the walk's and evaluation's records with the moments and the gather
  RecordExt E{};
  if (rec && rec->cells) {
    E.R = *rec;
    E.skip = skip_.get();
    E.depth = node_depth_.get();
    E.tsid = batch ? tsid_.get() : nullptr;
    E.roots = root_.get();
    E.walk_error = frame + frame::kWalkError;
    wrote_records_ = true;
  }
  hipLaunchKernelGGL(gather_tree_order, dim3(blocks_for(n)), dim3(kBlock), 0, s, vals, pos, mass,
                     sid, n, tpos, tmass, batch ? tsid_.get() : static_cast<uint32_t *>(nullptr),
                     rec ? inv_.get() : static_cast<uint32_t *>(nullptr),
                     static_cast<const double *>(root_.get()), E.R);
  GTH_LAUNCH_CHECK();

  // 7. moments, deepest level first, down from the deepest level the plan
  // allows (a deeper tree is reported by the depth in the frame).  Each level's
  // grid is sized from the previous tree's level (the launches stride, so any
  // grid is correct): most levels are small, and a full-size grid of idle
  // workgroups per level made the moments launch-bound.  The deep tail of
  // levels that the previous tree had small (with the empty levels the plan
  // allows below the deepest) goes to one workgroup, level after level: 5 us
  // instead of about 19 us for four launches.  (The small top levels are not
  // done that way: their nodes have eight children each, and one workgroup
  // took 61 us where five launches take 37 us.)
  const uint32_t mgrid = std::min<uint32_t>(blocks_for(cap), kMomentBlocks);
  auto small = [&p](int d) {
    return p.levels_sized && d < frame::kLevels && p.level_blocks[d] <= kSerialLevelBlocks;
  };
  int lo = p.depth_cap;  // levels (lo, depth_cap]: the small deep tail
  while (lo >= 0 && small(lo)) --lo;
  if (lo < p.depth_cap) {
    hipLaunchKernelGGL(serial_level_moments, dim3(1), dim3(kSerialBlock), 0, s, p.depth_cap,
                       lo + 1, lstart, frame, cap, lnk, tpos, tmass, geom, E);
    GTH_LAUNCH_CHECK();
  }
  auto paired = [&p](int d) {
    return p.levels_sized && d < frame::kLevels && p.level_blocks[d] <= kPairLevelBlocks;
  };
  for (int d = lo; d >= 0;) {
    if (d >= 1 && paired(d) && paired(d - 1)) {  // levels d and d - 1 together
      const uint32_t blocks = std::min<uint32_t>(8u * p.level_blocks[d - 1], 8u * mgrid);
      hipLaunchKernelGGL(level_pair_moments, dim3(std::max<uint32_t>(blocks, 1u)), dim3(kBlock), 0,
                         s, d, lstart, frame, cap, lnk, tpos, tmass, geom, E);
      GTH_LAUNCH_CHECK();
      d -= 2;
      continue;
    }
    const uint32_t blocks =
        p.levels_sized && d < frame::kLevels ? std::min<uint32_t>(p.level_blocks[d], mgrid) : mgrid;
    hipLaunchKernelGGL(level_moments, dim3(std::max<uint32_t>(blocks, 1u)), dim3(kBlock), 0, s, d,
                       lstart, frame, cap, lnk, tpos, tmass, geom, E);
    GTH_LAUNCH_CHECK();
    --d;
  }
}

}  // namespace gth
