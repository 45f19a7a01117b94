// Per-particle FP64 kernels of a leapfrog step and the scatter from tree
// order back to particle order (reference: LeapfrogKDK::step,
// src/integrator.cpp:5-16, and GroupTreeGravity::compute's scatter,
// src/gravity_backends.cpp:229-233).  Bandwidth-bound: each is one pass over
// the particle arrays.
#include "gt_internal.h"

namespace gth {
namespace {

constexpr int kBlock = 256;

inline uint32_t blocks_for(size_t n) { return static_cast<uint32_t>((n + kBlock - 1) / kBlock); }

#pragma clang fp contract(off)

// acc[order[k]] = out[k] widened to double (the FP32 results are in tree order).
__global__ __launch_bounds__(kBlock) void scatter_forces(const float4 *__restrict__ out,
                                                         const uint32_t *__restrict__ order,
                                                         uint32_t n, double *__restrict__ acc,
                                                         double *__restrict__ pot,
                                                         const unsigned long long *status) {
  const uint32_t k = blockIdx.x * kBlock + threadIdx.x;
  if (k >= n || *status) return;
  const float4 a = out[k];
  const size_t i = order[k];
  acc[3 * i] = a.x;
  acc[3 * i + 1] = a.y;
  acc[3 * i + 2] = a.z;
  pot[i] = a.w;
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

__global__ void count_step(unsigned long long *frame) {
  if (frame[frame::kStatus] != 0) return;
  frame[frame::kStepsDone] += 1;
  frame[frame::kStepInteractions] += frame[frame::kStats] + frame[frame::kStats + 1];
}

}  // namespace

void launch_scatter(const float4 *out, const uint32_t *order, uint32_t n, double *acc,
                    double *pot, const unsigned long long *status, hipStream_t s) {
  if (!n) return;
  hipLaunchKernelGGL(scatter_forces, dim3(blocks_for(n)), dim3(kBlock), 0, s, out, order, n, acc,
                     pot, status);
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

void launch_count_step(unsigned long long *frame, hipStream_t s) {
  hipLaunchKernelGGL(count_step, dim3(1), dim3(1), 0, s, frame);
  GTH_LAUNCH_CHECK();
}

}  // namespace gth
