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
// at most cap (the levels the sort ordered by).
__global__ __launch_bounds__(kBlock) void shared_levels(const uint64_t *__restrict__ w0,
                                                        const uint64_t *__restrict__ w1,
                                                        const uint64_t *__restrict__ w2, int words,
                                                        int cap, const uint32_t *__restrict__ vals,
                                                        const uint32_t *__restrict__ sid,
                                                        uint32_t n, uint8_t *__restrict__ lcp) {
  const uint32_t p = blockIdx.x * kBlock + threadIdx.x;
  if (p + 1 >= n) return;
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

// Leaf depth of sorted position p and the number of nodes beginning at p.
// best = max over the windows j .. j+leaf of leaf+1 particles that contain p
// and fit in [0, n) of the window's min lcp.  Up to kTileLeaf the block puts
// its lcp values (with leaf on either side) in LDS and takes the minima of
// all windows starting in its range once (W), then each position's maximum
// over its leaf+1 windows; for larger leaves, the windows directly.
constexpr uint32_t kTileLeaf = 64;

__global__ __launch_bounds__(kBlock) void leaf_depths(const uint8_t *__restrict__ lcp,
                                                      const uint32_t *__restrict__ vals,
                                                      const uint8_t *__restrict__ dg, uint32_t n,
                                                      uint32_t leaf, int cap,
                                                      uint8_t *__restrict__ ldep,
                                                      uint32_t *__restrict__ starts,
                                                      uint32_t *__restrict__ unresolved) {
  __shared__ int8_t sl[kBlock + 2 * kTileLeaf];  // shared(lcp[k]) for k = base - leaf ..
  __shared__ int8_t win[kBlock + kTileLeaf];     // W(j) for j = base - leaf ..
  const uint32_t base = blockIdx.x * kBlock;
  const uint32_t p = base + threadIdx.x;
  int best = -1;
  if (leaf <= kTileLeaf) {
    // windows start at j in [base - leaf, base + kBlock) and read lcp[j .. j + leaf - 1]
    const int64_t k0 = static_cast<int64_t>(base) - leaf;
    const uint32_t nk = kBlock + 2 * leaf;
    for (uint32_t t = threadIdx.x; t < nk; t += kBlock) {
      const int64_t k = k0 + t;
      sl[t] = static_cast<int8_t>(k >= 0 && k + 1 < static_cast<int64_t>(n) ? shared(lcp[k]) : 0);
    }
    __syncthreads();
    const uint32_t nj = kBlock + leaf;
    for (uint32_t t = threadIdx.x; t < nj; t += kBlock) {
      int w = cap;
      for (uint32_t i = 0; i < leaf; ++i) w = min(w, static_cast<int>(sl[t + i]));
      win[t] = static_cast<int8_t>(w);
    }
    __syncthreads();
    if (p < n && n > leaf) {
      const uint32_t jlo = p >= leaf ? p - leaf : 0u;
      const uint32_t jhi = p < n - 1 - leaf ? p : n - 1 - leaf;
      for (uint32_t j = jlo; j <= jhi; ++j)
        best = max(best, static_cast<int>(win[j - (base - leaf)]));
    }
  } else if (p < n && n > leaf) {
    const uint32_t jlo = p >= leaf ? p - leaf : 0u;
    const uint32_t jhi = p < n - 1 - leaf ? p : n - 1 - leaf;
    for (uint32_t j = jlo; j <= jhi; ++j) {
      int w = cap;
      for (uint32_t k = j; k < j + leaf; ++k) w = min(w, shared(lcp[k]));
      best = max(best, w);
    }
  }
  if (p >= n) return;
  int d = best + 1;
  d = min(d, static_cast<int>(dg[vals[p]]));
  d = min(d, kDepthLimit);
  // Deeper than the path words computed: reported, and the tree is cut at
  // `cap` levels (nodes exist only where their octant digits do), to be built
  // again with another word.
  if (d > cap) atomicOr(unresolved, 1u);
  d = min(d, cap);
  ldep[p] = static_cast<uint8_t>(d);
  const int first = first_level(lcp, p);
  starts[p] = d >= first ? static_cast<uint32_t>(d - first + 1) : 0u;
}

// ---- 5. preorder enumeration ---------------------------------------------------------

__global__ __launch_bounds__(kBlock) void emit_nodes(const uint8_t *__restrict__ lcp,
                                                     const uint32_t *__restrict__ starts,
                                                     const uint32_t *__restrict__ offs, uint32_t n,
                                                     const unsigned long long *__restrict__ frame,
                                                     uint32_t node_cap,
                                                     uint64_t *__restrict__ node_depth,
                                                     uint32_t *__restrict__ node_begin) {
  const uint32_t p = blockIdx.x * kBlock + threadIdx.x;
  if (p >= n || frame[frame::kNodes] > node_cap) return;  // over capacity: no tree
  const uint32_t c = starts[p];
  const uint32_t first = static_cast<uint32_t>(first_level(lcp, p));
  for (uint32_t t = 0; t < c; ++t) {
    node_depth[offs[p] + t] = first + t;
    node_begin[offs[p] + t] = p;
  }
}

// lstart[d] = first node of level d (levels past the deepest stay 0), and the
// depth of the tree into the frame.
// frame[kLevelNodes + d] = nodes on level d (the host sizes the next moments
// launches from them).  One workgroup of 64.
__global__ void level_counts(const uint32_t *__restrict__ lstart,
                             unsigned long long *__restrict__ frame) {
  for (int d = threadIdx.x; d < frame::kLevels; d += blockDim.x) {
    const uint32_t a = lstart[d], b = lstart[d + 1];
    frame[frame::kLevelNodes + d] = b > a ? b - a : 0u;
  }
}

__global__ __launch_bounds__(kBlock) void level_starts(const uint64_t *__restrict__ node_depth,
                                                       unsigned long long *__restrict__ frame,
                                                       uint32_t node_cap,
                                                       uint32_t *__restrict__ lstart) {
  const uint32_t n_nodes = bounded_count(frame + frame::kNodes, node_cap);
  const uint32_t x = blockIdx.x * kBlock + threadIdx.x;
  if (x >= n_nodes) return;
  const uint32_t d = static_cast<uint32_t>(node_depth[x]);
  if (x == 0 || static_cast<uint32_t>(node_depth[x - 1]) != d) lstart[d] = x;
  if (x == n_nodes - 1) {
    lstart[d + 1] = n_nodes;
    frame[frame::kDepth] = d;
  }
}

// ---- 6. links, geometry, leaf order ------------------------------------------------

// Largest i in [lo, hi] with begin[i] <= b (begin[lo] <= b, begins ascending).
__device__ inline uint32_t last_begin_at_or_before(const uint32_t *begin, uint32_t lo,
                                                   uint32_t hi, uint32_t b) {
  while (lo < hi) {
    const uint32_t mid = (lo + hi + 1) / 2;
    if (begin[mid] <= b)
      lo = mid;
    else
      hi = mid - 1;
  }
  return lo;
}

// Node x's centre and half size, from the octant path of its first body
// replayed from its system's root, and its parent: the last node of the level
// above beginning at or before x's begin.  With one path word and one system
// (sorted_keys set) the path is the sorted key at the node's begin, shifted
// back into place (it holds the levels the sort ordered by, which the tree
// never exceeds), instead of the first body's path word gathered by particle
// index.  A workgroup whose nodes are all on one level finds the parents of
// its first and last node in global memory and the others in an LDS copy of
// the begins between them.
constexpr uint32_t kParentWindow = 2048;

__global__ __launch_bounds__(kBlock) void link_nodes(
    const uint64_t *__restrict__ node_depth, const uint32_t *__restrict__ node_begin,
    const uint32_t *__restrict__ lstart, const uint64_t *__restrict__ w0,
    const uint64_t *__restrict__ w1, const uint64_t *__restrict__ w2,
    const uint32_t *__restrict__ vals, const uint64_t *__restrict__ sorted_keys, int key_shift,
    const double *__restrict__ roots, const uint32_t *__restrict__ sid,
    const unsigned long long *__restrict__ frame, uint32_t node_cap, double *__restrict__ geom,
    int4 *__restrict__ lnk, uint32_t *__restrict__ parent, uint32_t *__restrict__ root_node) {
  __shared__ uint32_t sbegin[kParentWindow];
  __shared__ uint32_t bracket[2];
  const uint32_t n_nodes = bounded_count(frame + frame::kNodes, node_cap);
  const uint32_t x0 = blockIdx.x * kBlock;
  if (x0 >= n_nodes) return;  // the whole workgroup
  const uint32_t x = x0 + threadIdx.x;
  const uint32_t xl = min(x0 + kBlock, n_nodes) - 1;
  const bool valid = x < n_nodes;
  uint32_t d = 0, b = 0, system = 0;
  if (valid) {
    d = static_cast<uint32_t>(node_depth[x]);
    b = node_begin[x];
    // (the path words this depth needs, loaded once)
    uint64_t kw[3] = {0ull, 0ull, 0ull};
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
    lnk[x] = make_int4(-1, 0, static_cast<int>(b), 0);
    if (d == 0) {
      parent[x] = kNone;
      if (root_node) root_node[system] = x;
    }
  }
  const uint32_t dfirst = static_cast<uint32_t>(node_depth[x0]);
  const bool one_level = dfirst > 0 && dfirst == static_cast<uint32_t>(node_depth[xl]);
  uint32_t up = kNone;
  if (one_level) {
    const uint32_t lo = lstart[dfirst - 1], hi = lstart[dfirst] - 1;
    if (threadIdx.x == 0) bracket[0] = last_begin_at_or_before(node_begin, lo, hi, b);
    if (x == xl) bracket[1] = last_begin_at_or_before(node_begin, lo, hi, b);
    __syncthreads();
    const uint32_t p0 = bracket[0], p1 = bracket[1];
    if (p1 - p0 < kParentWindow) {
      for (uint32_t t = threadIdx.x; t <= p1 - p0; t += kBlock) sbegin[t] = node_begin[p0 + t];
      __syncthreads();
      if (valid) up = p0 + last_begin_at_or_before(sbegin, 0, p1 - p0, b);
    } else if (valid) {
      up = last_begin_at_or_before(node_begin, p0, p1, b);
    }
  } else if (valid && d > 0) {
    up = last_begin_at_or_before(node_begin, lstart[d - 1], lstart[d] - 1, b);
  }
  if (up != kNone) parent[x] = up;
}

// Children of a node are consecutive on the level below: the first child
// (the one beginning where its parent begins) goes to its parent's link, the
// last one (whose successor has another parent) to last_child[parent].
__global__ __launch_bounds__(kBlock) void finish_links(const uint32_t *__restrict__ parent,
                                                       const uint32_t *__restrict__ node_begin,
                                                       const unsigned long long *__restrict__ frame,
                                                       uint32_t node_cap, int4 *__restrict__ lnk,
                                                       uint32_t *__restrict__ last_child) {
  const uint32_t n_nodes = bounded_count(frame + frame::kNodes, node_cap);
  const uint32_t x = blockIdx.x * kBlock + threadIdx.x;
  if (x >= n_nodes) return;
  const uint32_t p = parent[x];
  if (p == kNone) return;
  if (node_begin[p] == node_begin[x]) lnk[p].x = static_cast<int>(x);
  if (x + 1 >= n_nodes || parent[x + 1] != p) last_child[p] = x;
}

__global__ __launch_bounds__(kBlock) void node_ends(const uint32_t *__restrict__ parent,
                                                    const uint32_t *__restrict__ node_begin,
                                                    const uint64_t *__restrict__ node_depth,
                                                    const uint32_t *__restrict__ last_child,
                                                    const unsigned long long *__restrict__ frame,
                                                    uint32_t node_cap, uint32_t n,
                                                    int4 *__restrict__ lnk) {
  const uint32_t n_nodes = bounded_count(frame + frame::kNodes, node_cap);
  const uint32_t x0 = blockIdx.x * kBlock + threadIdx.x;
  if (x0 >= n_nodes) return;
  const int first = lnk[x0].x;  // children: first .. last_child[x0]
  lnk[x0].y = first >= 0 ? static_cast<int>(last_child[x0]) - first + 1 : 0;
  uint32_t x = x0;
  for (;;) {
    const uint32_t p = parent[x];
    if (p == kNone) {  // a root ends where the next system's root begins
      const bool next_root = x + 1 < n_nodes && node_depth[x + 1] == 0;
      lnk[x0].w = static_cast<int>(next_root ? node_begin[x + 1] : n);
      return;
    }
    if (x + 1 < n_nodes && parent[x + 1] == p) {  // not its parent's last child
      lnk[x0].w = static_cast<int>(x + 1 < n_nodes ? node_begin[x + 1] : n);
      return;
    }
    x = p;
  }
}

// Bodies of a leaf back into particle order (insertion sort; leaves are small).
constexpr uint32_t kLeafRegs = 16;  // leaves up to this size are sorted in registers
// orig (may be null: the identity) maps the positions in pos to the caller's
// particle indices, which decide the order.
__global__ __launch_bounds__(kBlock) void leaf_order(const int4 *__restrict__ lnk,
                                                     const unsigned long long *__restrict__ frame,
                                                     uint32_t node_cap,
                                                     const uint32_t *__restrict__ orig,
                                                     uint32_t *__restrict__ vals) {
  const uint32_t n_nodes = bounded_count(frame + frame::kNodes, node_cap);
  const uint32_t x = blockIdx.x * kBlock + threadIdx.x;
  if (x >= n_nodes) return;
  const int4 L = lnk[x];
  if (L.y != 0) return;
  const uint32_t b = static_cast<uint32_t>(L.z), e = static_cast<uint32_t>(L.w);
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

__global__ __launch_bounds__(kBlock) void gather_tree_order(const uint32_t *__restrict__ vals,
                                                            const double *__restrict__ pos,
                                                            const double *__restrict__ mass,
                                                            const uint32_t *__restrict__ sid,
                                                            uint32_t n, double *__restrict__ tpos,
                                                            double *__restrict__ tmass,
                                                            uint32_t *__restrict__ tsid) {
  const uint32_t k = blockIdx.x * kBlock + threadIdx.x;
  if (k >= n) return;
  const uint32_t i = vals[k];
  if (sid) tsid[k] = sid[i];
  for (int a = 0; a < 3; ++a) tpos[3 * static_cast<size_t>(k) + a] = pos[3 * static_cast<size_t>(i) + a];
  tmass[k] = mass[i];
}

// ---- 7. moments, one level at a time -------------------------------------------------

// Moments of node x from its bodies (leaf) or its children, which are on the
// level below and already done.  The data of up to eight bodies or children
// are loaded together before they are summed (one round of loads instead of
// one per term: the level launches are latency-bound), and summed one by one
// in order, so every sum is the reference's sequence.
constexpr int kChunk = 8;

__device__ __attribute__((always_inline)) inline void node_moments(uint32_t x, uint32_t n_nodes, const int4 *__restrict__ lnk,
                                    const double *__restrict__ tpos,
                                    const double *__restrict__ tmass,
                                    double *__restrict__ geom) {
  const int4 L = lnk[x];
  double *g = geom + static_cast<size_t>(kGeomStride) * x;
  const double ccx = g[0], ccy = g[1], ccz = g[2];
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
  g[14] = __dsqrt_rn(ex * ex + ey * ey + ez * ez);
}

// Every node of level d: lstart[d] .. lstart[d + 1] (empty for levels past the
// deepest, whose lstart entries are 0).
__global__ __launch_bounds__(kBlock) void level_moments(int d, const uint32_t *__restrict__ lstart,
                                                        const unsigned long long *__restrict__ frame,
                                                        uint32_t node_cap,
                                                        const int4 *__restrict__ lnk,
                                                        const double *__restrict__ tpos,
                                                        const double *__restrict__ tmass,
                                                        double *__restrict__ geom) {
  const uint32_t n_nodes = bounded_count(frame + frame::kNodes, node_cap);
  const uint32_t first = lstart[d], last = min(lstart[d + 1], n_nodes);
  for (uint32_t x = first + blockIdx.x * kBlock + threadIdx.x; x < last; x += gridDim.x * kBlock)
    node_moments(x, n_nodes, lnk, tpos, tmass, geom);
}

// Levels d and d - 1 in one launch: eight lanes per node of level d - 1, lane
// j doing the node's child j (on level d), then, after a barrier, the node's
// first lane the node itself.  For small levels, where a launch costs more
// than the level's work.
__global__ __launch_bounds__(kBlock) void level_pair_moments(
    int d, const uint32_t *__restrict__ lstart, const unsigned long long *__restrict__ frame,
    uint32_t node_cap, const int4 *__restrict__ lnk, const double *__restrict__ tpos,
    const double *__restrict__ tmass, double *__restrict__ geom) {
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
        node_moments(static_cast<uint32_t>(L.x) + j, n_nodes, lnk, tpos, tmass, geom);
    }
    __syncthreads();
    if (t < lanes && j == 0) node_moments(x, n_nodes, lnk, tpos, tmass, geom);
  }
}

// Levels d_hi down to d_lo by one workgroup, with a barrier between levels:
// for the run of small levels at the bottom of the tree, where a launch per
// level costs more than the level's work.  Any level size is
// handled (the loop strides), only more slowly.
constexpr int kSerialBlock = 512;

__global__ __launch_bounds__(kSerialBlock) void serial_level_moments(
    int d_hi, int d_lo, const uint32_t *__restrict__ lstart,
    const unsigned long long *__restrict__ frame, uint32_t node_cap, const int4 *__restrict__ lnk,
    const double *__restrict__ tpos, const double *__restrict__ tmass, double *__restrict__ geom) {
  const uint32_t n_nodes = bounded_count(frame + frame::kNodes, node_cap);
  for (int d = d_hi; d >= d_lo; --d) {
    const uint32_t first = lstart[d], last = min(lstart[d + 1], n_nodes);
    for (uint32_t x = first + threadIdx.x; x < last; x += kSerialBlock)
      node_moments(x, n_nodes, lnk, tpos, tmass, geom);
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
  keys_alt_.ensure(n);
  vals_alt_.ensure(n);
  dg_.ensure(n);
  word0_.ensure(n);
  if (p.words > 1) word1_.ensure(n);
  if (p.words > 2) word2_.ensure(n);
  sort_tmp_.ensure(radix_tmp_words(std::max(n, nn)));
  lcp_.ensure(n);
  ldep_.ensure(n);
  starts_.ensure(n);
  offs_.ensure(n);
  scan_tmp_.ensure(scan_tmp_words(n));
  ndep_.ensure(nn);
  nbeg_.ensure(nn);
  ndep_alt_.ensure(nn);
  nbeg_alt_.ensure(nn);
  lstart_.ensure(kDepthLimit + 2);
  geom_.ensure(static_cast<size_t>(kGeomStride) * nn);
  lnk_.ensure(nn);
  parent_.ensure(nn);
  last_child_.ensure(nn);
  tpos_.ensure(3 * n);
  tmass_.ensure(n);
}

void TreeBuilder::enqueue(const double *pos, const double *mass, const Plan &p, int leaf_size,
                          const unsigned long long *sys_off, unsigned long long *frame,
                          hipStream_t s, const uint32_t *orig) {
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
  // leaf go back into particle order anyway (leaf_order).  A leaf deeper than
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

  // 4. leaf depths and node counts per sorted position; the node total goes
  // to frame[kNodes]
  uint8_t *lcp = lcp_.get();
  uint8_t *ldep = ldep_.get();
  uint32_t *starts = starts_.get();
  uint32_t *offs = offs_.get();
  hipLaunchKernelGGL(shared_levels, dim3(blocks_for(n)), dim3(kBlock), 0, s, w[0], w[1], w[2],
                     p.words, levels, vals, sid, n, lcp);
  GTH_LAUNCH_CHECK();
  hipLaunchKernelGGL(leaf_depths, dim3(blocks_for(n)), dim3(kBlock), 0, s, lcp, vals, dg, n, leaf,
                     levels, ldep, starts,
                     reinterpret_cast<uint32_t *>(frame + frame::kUnresolved));
  GTH_LAUNCH_CHECK();
  exclusive_scan_u32(starts, offs, n, frame + frame::kNodes, scan_tmp_.get(), s);

  // 5. nodes in preorder, then stably by depth (level order)
  uint64_t *ndep = ndep_.get();
  uint32_t *nbeg = nbeg_.get();
  hipLaunchKernelGGL(emit_nodes, dim3(blocks_for(n)), dim3(kBlock), 0, s, lcp, starts, offs, n,
                     frame, cap, ndep, nbeg);
  GTH_LAUNCH_CHECK();
  radix_sort_u64_dev(ndep, nbeg, ndep_alt_.get(), nbeg_alt_.get(), frame + frame::kNodes, cap, 6,
                     sort_tmp_.get(), s, frame + frame::kSortCtl);
  uint32_t *lstart = lstart_.get();
  launch_fill_u32(lstart, kDepthLimit + 2, 0u, s);
  hipLaunchKernelGGL(level_starts, dim3(blocks_for(cap)), dim3(kBlock), 0, s, ndep, frame, cap,
                     lstart);
  GTH_LAUNCH_CHECK();
  hipLaunchKernelGGL(level_counts, dim3(1), dim3(64), 0, s, lstart, frame);
  GTH_LAUNCH_CHECK();

  // 6. links and geometry
  double *geom = geom_.get();
  int4 *lnk = lnk_.get();
  uint32_t *parent = parent_.get();
  uint32_t *last_child = last_child_.get();
  hipLaunchKernelGGL(link_nodes, dim3(blocks_for(cap)), dim3(kBlock), 0, s, ndep, nbeg, lstart,
                     w[0], w[1], w[2], vals,
                     p.words == 1 && !batch ? keys : static_cast<const uint64_t *>(nullptr),
                     p.words == 1 ? 3 * (kLevels - levels) : 0, root_.get(), sid, frame, cap,
                     geom, lnk, parent, root_node);
  GTH_LAUNCH_CHECK();
  hipLaunchKernelGGL(finish_links, dim3(blocks_for(cap)), dim3(kBlock), 0, s, parent, nbeg, frame,
                     cap, lnk, last_child);
  GTH_LAUNCH_CHECK();
  hipLaunchKernelGGL(node_ends, dim3(blocks_for(cap)), dim3(kBlock), 0, s, parent, nbeg, ndep,
                     last_child, frame, cap, n, lnk);
  GTH_LAUNCH_CHECK();
  hipLaunchKernelGGL(leaf_order, dim3(blocks_for(cap)), dim3(kBlock), 0, s, lnk, frame, cap, orig,
                     vals);
  GTH_LAUNCH_CHECK();
  double *tpos = tpos_.get();
  double *tmass = tmass_.get();
  hipLaunchKernelGGL(gather_tree_order, dim3(blocks_for(n)), dim3(kBlock), 0, s, vals, pos, mass,
                     sid, n, tpos, tmass, batch ? tsid_.get() : static_cast<uint32_t *>(nullptr));
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
                       lo + 1, lstart, frame, cap, lnk, tpos, tmass, geom);
    GTH_LAUNCH_CHECK();
  }
  auto paired = [&p](int d) {
    return p.levels_sized && d < frame::kLevels && p.level_blocks[d] <= kPairLevelBlocks;
  };
  for (int d = lo; d >= 0;) {
    if (d >= 1 && paired(d) && paired(d - 1)) {  // levels d and d - 1 together
      const uint32_t blocks = std::min<uint32_t>(8u * p.level_blocks[d - 1], 8u * mgrid);
      hipLaunchKernelGGL(level_pair_moments, dim3(std::max<uint32_t>(blocks, 1u)), dim3(kBlock), 0,
                         s, d, lstart, frame, cap, lnk, tpos, tmass, geom);
      GTH_LAUNCH_CHECK();
      d -= 2;
      continue;
    }
    const uint32_t blocks =
        p.levels_sized && d < frame::kLevels ? std::min<uint32_t>(p.level_blocks[d], mgrid) : mgrid;
    hipLaunchKernelGGL(level_moments, dim3(std::max<uint32_t>(blocks, 1u)), dim3(kBlock), 0, s, d,
                       lstart, frame, cap, lnk, tpos, tmass, geom);
    GTH_LAUNCH_CHECK();
    --d;
  }
}

}  // namespace gth
