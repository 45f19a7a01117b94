// The per-particle FP64 tree solver on the GPU (reference: Octree::accel_pot,
// src/octree.cpp:275-344, driven by compute_gravity_tree, src/gravity.cpp:7-26;
// the reference runs it on the CPU only).
//
// One lane per particle, particles in tree order (neighbouring lanes visit the
// same nodes); or one lane per caller-given point, nobody left out (the
// reference's accel_pot with self = Octree::kNoSelf, which its ICs generator
// samples the potential with, src/hernquist_ics.cpp:206-330).  Each lane walks the tree depth-first with the reference's LIFO
// order (children last to first) without a stack: `skip` is the node the LIFO
// walk pops after a subtree.  Every opening test, multipole and body
// interaction is the reference's FP64 expression in its operation order, with
// no contraction and correctly rounded sqrt and division, and each lane sums
// its interactions in walk order, so accelerations and potentials equal the
// reference's bit for bit on the same tree.
#include "gt_internal.h"

namespace gth {
namespace {

constexpr int kBlock = 256;

#pragma clang fp contract(off)

// One wave walks for 64 consecutive particles (in tree order, so their walks
// share most nodes): it visits every node any lane would visit, in the LIFO
// order, and a lane takes part at a node only while every ancestor was opened
// for it (active masks per depth in LDS).  Each lane therefore meets exactly
// the nodes of its own walk, in order, and sums the same interactions in the
// same order as a walk of its own.  Node data are read at wave-uniform
// addresses and the wave's control flow is uniform: at 2M particles (theta
// 0.75) 12.6 ms, against 28.2 ms with one independent walk per lane.
constexpr int kMaxLevels = 64;

__global__ __launch_bounds__(kBlock) void tree_gravity(
    const double *__restrict__ geom, const int4 *__restrict__ lnk,
    const uint32_t *__restrict__ skip, const uint32_t *__restrict__ depth,
    const double *__restrict__ tpos, const double *__restrict__ tmass,
    const uint32_t *__restrict__ order, uint32_t n, const unsigned long long *__restrict__ frame,
    uint32_t node_cap, double G, double eps2, double inv_theta, int quadrupole,
    const double *__restrict__ points, uint32_t n_points, double *__restrict__ acc,
    double *__restrict__ pot, unsigned long long *__restrict__ stats) {
  __shared__ unsigned long long masks[kBlock / 64][kMaxLevels];
  const uint32_t nn = bounded_count(frame + frame::kNodes, node_cap);
  const uint32_t lane = threadIdx.x % 64, wave = threadIdx.x / 64;
  const uint32_t k = blockIdx.x * kBlock + threadIdx.x;
  // targets: the particles in tree order (each leaves itself out), or, with
  // points, the points in the given order (nobody left out: self = kNone)
  const bool valid = k < (points ? n_points : n);
  unsigned long long ncell = 0, npart = 0;
  uint32_t self = kNone;
  double x = 0, y = 0, z = 0;
  if (valid) {
    const double *t = points ? points + 3 * static_cast<size_t>(k) : tpos + 3 * static_cast<size_t>(k);
    if (!points) self = order[k];
    x = t[0];
    y = t[1];
    z = t[2];
  }
  double ax = 0, ay = 0, az = 0, phi = 0;
  const unsigned long long me = 1ull << lane;
  unsigned long long *stack = masks[wave];
  stack[0] = __ballot(valid);
  uint32_t cur = (nn > 0 && stack[0] != 0) ? 0u : kNone, d = 0, budget = nn;
  bool broken = false;
  while (cur != kNone) {
    if (cur >= nn || budget-- == 0 || d >= kMaxLevels - 1) {
      broken = true;  // malformed tree (never for a built one)
      break;
    }
    const uint32_t c = __builtin_amdgcn_readfirstlane(cur);
    const double *g = geom + static_cast<size_t>(kGeomStride) * c;
    const int4 L = lnk[c];
    const bool act = (stack[d] & me) != 0;
    const double dx = x - g[4], dy = y - g[5], dz = z - g[6];
    const double r2 = dx * dx + dy * dy + dz * dz;
    const double half = g[3];
    const double crit = 2.0 * half * inv_theta + g[14];
    const double slack = half * 1.0001;
    const bool inside = fabs(x - g[0]) <= slack && fabs(y - g[1]) <= slack &&
                        fabs(z - g[2]) <= slack;
    const bool accept = act && !inside && r2 > crit * crit;
    const bool open = act && !accept;
    if (__ballot(accept)) {
      if (accept) {  // the cell as a multipole
        const double m = g[7];
        const double r2s = r2 + eps2;
        const double inv = __ddiv_rn(1.0, __dsqrt_rn(r2s));
        const double inv3 = inv * inv * inv;
        const double gm3 = G * m * inv3;
        ax -= gm3 * dx;
        ay -= gm3 * dy;
        az -= gm3 * dz;
        phi -= G * m * inv;
        if (quadrupole) {
          const double *q = g + 8;
          const double qx = q[0] * dx + q[3] * dy + q[4] * dz;
          const double qy = q[3] * dx + q[1] * dy + q[5] * dz;
          const double qz = q[4] * dx + q[5] * dy + q[2] * dz;
          const double rqr = dx * qx + dy * qy + dz * qz;
          const double inv5 = inv3 * inv * inv;
          const double inv7 = inv5 * inv * inv;
          ax += G * (qx * inv5 - 2.5 * rqr * dx * inv7);
          ay += G * (qy * inv5 - 2.5 * rqr * dy * inv7);
          az += G * (qz * inv5 - 2.5 * rqr * dz * inv7);
          phi -= 0.5 * G * rqr * inv5;
        }
        ++ncell;
      }
    }
    const unsigned long long opened = __ballot(open);
    if (L.y == 0) {
      if (opened) {
        const uint32_t b = static_cast<uint32_t>(L.z), e = min(static_cast<uint32_t>(L.w), n);
        for (uint32_t j = b; j < e; ++j) {
          const uint32_t oj = order[j];
          const double ex = x - tpos[3 * static_cast<size_t>(j)],
                       ey = y - tpos[3 * static_cast<size_t>(j) + 1],
                       ez = z - tpos[3 * static_cast<size_t>(j) + 2];
          if (open && oj != self) {
            const double r2s = ex * ex + ey * ey + ez * ez + eps2;
            const double inv = __ddiv_rn(1.0, __dsqrt_rn(r2s));
            const double mj = tmass[j];
            const double gm3 = G * mj * inv * inv * inv;
            ax -= gm3 * ex;
            ay -= gm3 * ey;
            az -= gm3 * ez;
            phi -= G * mj * inv;
            ++npart;
          }
        }
      }
      cur = skip[c];
      d = cur < nn ? depth[cur] : 0u;
    } else if (opened) {
      stack[d + 1] = opened;  // lanes active below this node
      cur = static_cast<uint32_t>(L.x + L.y - 1);  // the last child first
      ++d;
    } else {
      cur = skip[c];
      d = cur < nn ? depth[cur] : 0u;
    }
  }
  if (broken) phi = __builtin_nan("");
  if (valid) {
    const size_t o = points ? k : self;  // particles: in the caller's order
    acc[3 * o] = ax;
    acc[3 * o + 1] = ay;
    acc[3 * o + 2] = az;
    pot[o] = phi;
  }
  for (int off = 32; off > 0; off >>= 1) {
    ncell += __shfl_xor(ncell, off, 64);
    npart += __shfl_xor(npart, off, 64);
  }
  if (lane == 0 && (ncell | npart)) {
    atomicAdd(stats, ncell);
    atomicAdd(stats + 1, npart);
  }
}

#pragma clang fp contract(fast)

}  // namespace

void launch_tree_gravity(const double *geom, const int4 *lnk, const uint32_t *skip,
                         const uint32_t *depth, const double *tpos, const double *tmass,
                         const uint32_t *order, uint32_t n, const unsigned long long *frame,
                         uint32_t node_cap, double G, double eps, double theta, int quadrupole,
                         double *acc, double *pot, unsigned long long *stats, hipStream_t s,
                         const double *points, uint32_t n_points) {
  launch_fill_u64(stats, 2, 0ull, s);
  const uint32_t targets = points ? n_points : n;
  if (!n || !targets) return;
  hipLaunchKernelGGL(tree_gravity, dim3((targets + kBlock - 1) / kBlock), dim3(kBlock), 0, s, geom,
                     lnk, skip, depth, tpos, tmass, order, n, frame, node_cap, G, eps * eps,
                     1.0 / theta, quadrupole, points, n_points, acc, pot, stats);
  GTH_LAUNCH_CHECK();
}

}  // namespace gth
