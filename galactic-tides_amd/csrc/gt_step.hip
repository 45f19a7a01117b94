// Per-particle FP64 kernels of a leapfrog step and the scatter from tree
// order back to particle order (reference: LeapfrogKDK::step,
// src/integrator.cpp:5-16, and GroupTreeGravity::compute's scatter,
// src/gravity_backends.cpp:132-136).  Bandwidth-bound: each is one pass over
// the particle arrays.
#include "gt_internal.h"

namespace gth {
namespace {

constexpr int kBlock = 256;

inline uint32_t blocks_for(size_t n) { return static_cast<uint32_t>((n + kBlock - 1) / kBlock); }

#pragma clang fp contract(off)

// acc[i] = out[inv[i]] widened to double (the FP32 results are in tree order;
// inv[i] is particle i's tree position): one pass over the rows in their own
// order, the results gathered.
__global__ __launch_bounds__(kBlock) void scatter_forces(const float4 *__restrict__ out,
                                                         const uint32_t *__restrict__ inv,
                                                         uint32_t n, double *__restrict__ acc,
                                                         double *__restrict__ pot,
                                                         const unsigned long long *status) {
  const uint32_t i = blockIdx.x * kBlock + threadIdx.x;
  if (i >= n || *status) return;
  const float4 a = out[inv[i]];
  acc[3 * static_cast<size_t>(i)] = a.x;
  acc[3 * static_cast<size_t>(i) + 1] = a.y;
  acc[3 * static_cast<size_t>(i) + 2] = a.z;
  pot[i] = a.w;
}

// The scatter and the step's closing kick in one pass: acc and pot as
// scatter_forces writes them, then v += h a for the same row, the same
// operations as kick (so the result is bit for bit the two kernels').
__global__ __launch_bounds__(kBlock) void scatter_kick(const float4 *__restrict__ out,
                                                       const uint32_t *__restrict__ inv,
                                                       uint32_t n, double *__restrict__ acc,
                                                       double *__restrict__ pot,
                                                       double *__restrict__ vel, double h,
                                                       const unsigned long long *status,
                                                       unsigned long long *step_frame) {
  const uint32_t i = blockIdx.x * kBlock + threadIdx.x;
  if (i >= n || *status) return;
  if (step_frame && i == 0) {  // the step is done: count it and its interactions
    step_frame[frame::kStepsDone] += 1;
    step_frame[frame::kStepInteractions] +=
        step_frame[frame::kStats] + step_frame[frame::kStats + 1];
  }
  const float4 a = out[inv[i]];
  const double ax = a.x, ay = a.y, az = a.z;
  const size_t r = 3 * static_cast<size_t>(i);
  acc[r] = ax;
  acc[r + 1] = ay;
  acc[r + 2] = az;
  pot[i] = a.w;
  vel[r] = vel[r] + h * ax;
  vel[r + 1] = vel[r + 1] + h * ay;
  vel[r + 2] = vel[r + 2] + h * az;
}

// v += h a; x += dt v   (first half of the step)
__global__ __launch_bounds__(kBlock) void kick_drift(double *__restrict__ pos,
                                                     double *__restrict__ vel,
                                                     const double *__restrict__ acc, size_t m,
                                                     double h, double dt,
                                                     const unsigned long long *status) {
  const size_t k = static_cast<size_t>(blockIdx.x) * kBlock + threadIdx.x;
  if (k >= m || *status) return;
  const double v = vel[k] + h * acc[k];
  vel[k] = v;
  pos[k] = pos[k] + dt * v;
}

// v += h a   (second half)
__global__ __launch_bounds__(kBlock) void kick(double *__restrict__ vel,
                                               const double *__restrict__ acc, size_t m,
                                               double h, const unsigned long long *status) {
  const size_t k = static_cast<size_t>(blockIdx.x) * kBlock + threadIdx.x;
  if (k >= m || *status) return;
  vel[k] = vel[k] + h * acc[k];
}

#pragma clang fp contract(fast)

// Diagnostics (reference: compute_diagnostics, src/integrator.cpp:18-33): per
// particle 0.5 m |v|^2, 0.5 m phi, m v, m (x cross v), m x and m, summed over
// each system.  Block (j, s) sums the j-th of kDiagChunks equal pieces of
// system s (threads stride the piece, then a fixed-shape tree over the
// block); diag_total adds the pieces in order.  The sums are deterministic,
// though not in the reference's sequential order.
constexpr int kDiagChunks = 256;
constexpr int kDiagTerms = 14;  // K, W, P[3], L[3], M x[3], M, (2 unused)

__global__ __launch_bounds__(kBlock) void diag_partial(
    const double *__restrict__ pos, const double *__restrict__ vel,
    const double *__restrict__ mass, const double *__restrict__ pot,
    const unsigned long long *__restrict__ offsets, uint64_t n, double *__restrict__ partial) {
  __shared__ double red[kBlock];
  const uint32_t s = blockIdx.y, j = blockIdx.x;
  const uint64_t a = offsets ? offsets[s] : 0, b = offsets ? offsets[s + 1] : n;
  const uint64_t len = b > a ? b - a : 0;
  const uint64_t lo = a + len * j / kDiagChunks, hi = a + len * (j + 1) / kDiagChunks;
  double t[kDiagTerms];
  for (int q = 0; q < kDiagTerms; ++q) t[q] = 0.0;
  for (uint64_t i = lo + threadIdx.x; i < hi; i += kBlock) {
    const double m = mass[i];
    const double x = pos[3 * i], y = pos[3 * i + 1], z = pos[3 * i + 2];
    const double vx = vel[3 * i], vy = vel[3 * i + 1], vz = vel[3 * i + 2];
    t[0] += 0.5 * m * (vx * vx + vy * vy + vz * vz);
    t[1] += 0.5 * m * pot[i];
    t[2] += m * vx;
    t[3] += m * vy;
    t[4] += m * vz;
    t[5] += m * (y * vz - z * vy);
    t[6] += m * (z * vx - x * vz);
    t[7] += m * (x * vy - y * vx);
    t[8] += m * x;
    t[9] += m * y;
    t[10] += m * z;
    t[11] += m;
  }
  double *out = partial + (static_cast<size_t>(s) * kDiagChunks + j) * kDiagTerms;
  for (int q = 0; q < kDiagTerms; ++q) {
    red[threadIdx.x] = t[q];
    __syncthreads();
    for (int w = kBlock / 2; w > 0; w >>= 1) {
      if (threadIdx.x < w) red[threadIdx.x] += red[threadIdx.x + w];
      __syncthreads();
    }
    if (threadIdx.x == 0) out[q] = red[0];
    __syncthreads();
  }
}

__global__ void diag_total(const double *__restrict__ partial, uint32_t systems,
                           double *__restrict__ total) {
  const uint32_t s = blockIdx.x * blockDim.x + threadIdx.x;
  if (s >= systems) return;
  for (int q = 0; q < kDiagTerms; ++q) {
    double v = 0.0;
    for (int j = 0; j < kDiagChunks; ++j)
      v += partial[(static_cast<size_t>(s) * kDiagChunks + j) * kDiagTerms + q];
    total[static_cast<size_t>(s) * kDiagTerms + q] = v;
  }
}

__global__ void count_step(unsigned long long *frame) {
  if (frame[frame::kStatus] != 0) return;
  frame[frame::kStepsDone] += 1;
  frame[frame::kStepInteractions] += frame[frame::kStats] + frame[frame::kStats + 1];
}

// out[k] = in[order[k]] for rows of w doubles (the resident state into the
// order of a tree: order gives the position in `in` of tree position k).
__global__ __launch_bounds__(kBlock) void gather_rows(const double *__restrict__ in,
                                                      const uint32_t *__restrict__ order,
                                                      uint64_t n, int w,
                                                      double *__restrict__ out) {
  const uint64_t k = static_cast<uint64_t>(blockIdx.x) * kBlock + threadIdx.x;
  if (k >= n) return;
  const uint64_t i = order[k];
  for (int a = 0; a < w; ++a) out[w * k + a] = in[w * i + a];
}

// out[idx[k]] = in[k] for rows of w doubles (idx a permutation).
__global__ __launch_bounds__(kBlock) void scatter_rows(const double *__restrict__ in,
                                                       const uint32_t *__restrict__ idx,
                                                       uint64_t n, int w,
                                                       double *__restrict__ out) {
  const uint64_t k = static_cast<uint64_t>(blockIdx.x) * kBlock + threadIdx.x;
  if (k >= n) return;
  const uint64_t i = idx[k];
  for (int a = 0; a < w; ++a) out[w * i + a] = in[w * k + a];
}

// out[k] = (a ? a[b[k]] : b[k]); iota when b is null too.
__global__ __launch_bounds__(kBlock) void compose_u32(const uint32_t *__restrict__ a,
                                                      const uint32_t *__restrict__ b, uint64_t n,
                                                      uint32_t *__restrict__ out) {
  const uint64_t k = static_cast<uint64_t>(blockIdx.x) * kBlock + threadIdx.x;
  if (k >= n) return;
  const uint32_t j = b ? b[k] : static_cast<uint32_t>(k);
  out[k] = a ? a[j] : j;
}

}  // namespace

void launch_scatter(const float4 *out, const uint32_t *inv, uint32_t n, double *acc,
                    double *pot, const unsigned long long *status, hipStream_t s) {
  if (!n) return;
  hipLaunchKernelGGL(scatter_forces, dim3(blocks_for(n)), dim3(kBlock), 0, s, out, inv, n, acc,
                     pot, status);
  GTH_LAUNCH_CHECK();
}

void launch_scatter_kick(const float4 *out, const uint32_t *inv, uint32_t n, double *acc,
                         double *pot, double *vel, double h, const unsigned long long *status,
                         hipStream_t s, unsigned long long *step_frame) {
  if (!n) {
    if (step_frame) launch_count_step(step_frame, s);
    return;
  }
  hipLaunchKernelGGL(scatter_kick, dim3(blocks_for(n)), dim3(kBlock), 0, s, out, inv, n, acc,
                     pot, vel, h, status, step_frame);
  GTH_LAUNCH_CHECK();
}

void launch_kick_drift(double *pos, double *vel, const double *acc, size_t n, double h,
                       double dt, const unsigned long long *status, hipStream_t s) {
  if (!n) return;
  hipLaunchKernelGGL(kick_drift, dim3(blocks_for(3 * n)), dim3(kBlock), 0, s, pos, vel, acc, 3 * n,
                     h, dt, status);
  GTH_LAUNCH_CHECK();
}

void launch_kick(double *vel, const double *acc, size_t n, double h,
                 const unsigned long long *status, hipStream_t s) {
  if (!n) return;
  hipLaunchKernelGGL(kick, dim3(blocks_for(3 * n)), dim3(kBlock), 0, s, vel, acc, 3 * n, h,
                     status);
  GTH_LAUNCH_CHECK();
}

void launch_gather_rows(const double *in, const uint32_t *order, uint64_t n, int w, double *out,
                        hipStream_t s) {
  if (!n) return;
  hipLaunchKernelGGL(gather_rows, dim3(blocks_for(n)), dim3(kBlock), 0, s, in, order, n, w, out);
  GTH_LAUNCH_CHECK();
}

void launch_scatter_rows(const double *in, const uint32_t *idx, uint64_t n, int w, double *out,
                         hipStream_t s) {
  if (!n) return;
  hipLaunchKernelGGL(scatter_rows, dim3(blocks_for(n)), dim3(kBlock), 0, s, in, idx, n, w, out);
  GTH_LAUNCH_CHECK();
}

void launch_compose_u32(const uint32_t *a, const uint32_t *b, uint64_t n, uint32_t *out,
                        hipStream_t s) {
  if (!n) return;
  hipLaunchKernelGGL(compose_u32, dim3(blocks_for(n)), dim3(kBlock), 0, s, a, b, n, out);
  GTH_LAUNCH_CHECK();
}

size_t diag_partial_words(uint32_t systems) {
  return static_cast<size_t>(systems) * kDiagChunks * kDiagTerms;
}

void launch_diagnostics(const double *pos, const double *vel, const double *mass,
                        const double *pot, const unsigned long long *offsets, uint32_t systems,
                        uint64_t n, double *partial, double *total, hipStream_t s) {
  if (!systems) return;
  hipLaunchKernelGGL(diag_partial, dim3(kDiagChunks, systems), dim3(kBlock), 0, s, pos, vel, mass,
                     pot, offsets, n, partial);
  GTH_LAUNCH_CHECK();
  hipLaunchKernelGGL(diag_total, dim3((systems + 63) / 64), dim3(64), 0, s, partial, systems,
                     total);
  GTH_LAUNCH_CHECK();
}

void launch_count_step(unsigned long long *frame, hipStream_t s) {
  hipLaunchKernelGGL(count_step, dim3(1), dim3(1), 0, s, frame);
  GTH_LAUNCH_CHECK();
}

}  // namespace gth
