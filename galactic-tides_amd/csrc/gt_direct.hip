// O(N^2) kernels.
//
//  * direct: FP32 softened body-body sum for every target (reference:
//    gt_direct, src/metal/gravity_kernels.metal:18-43, and gt_direct_accel,
//    gravity_kernels.h:173-182).  Lane = target; the source loop is wave-
//    uniform, so every source body arrives through the scalar cache straight
//    into SGPRs and is a scalar operand of the VALU ops: no LDS staging, no
//    barriers, no broadcast instructions.  v_rsq_f32 on the VALU; pairwise
//    gravity is not a dense contraction, so there is no MFMA here.
//  * direct_f64_sample: FP64 reference forces for a sample of targets
//    (apps/galactic_tides_bench.cpp:80-95), used to measure force errors.
//  * valu_probe: the GPU's sustained FP32 VALU rate (non-packed, as all the
//    kernels here are built), the practical ceiling of the VALU-bound
//    kernels: 16 independent FMA chains per lane, eight waves per SIMD.
#include "gt_internal.h"

namespace gth {
namespace {

constexpr int kBlock = 256;

// Sources j0 .. j1 against this lane's target, into the sums (of m d / r^3
// and m / r with d = source - target, so the signs and G come at the end).
// kSelf: the range may hold the target itself, which is skipped.
template <bool kSelf>
__device__ inline void direct_range(const float4 *__restrict__ bodies, uint32_t j0, uint32_t j1,
                                    uint32_t i, float4 bi, float eps2, float &sx, float &sy,
                                    float &sz, float &sp) {
#pragma unroll 8
  for (uint32_t j = j0; j < j1; ++j) {
    const float4 b = bodies[j];
    const float dx = b.x - bi.x, dy = b.y - bi.y, dz = b.z - bi.z;
    float r2 = fmaf(dx, dx, fmaf(dy, dy, fmaf(dz, dz, eps2)));
    if (kSelf) r2 = (j == i) ? __builtin_inff() : r2;
    const float inv = __builtin_amdgcn_rsqf(r2);
    const float t = b.w * inv;  // m / r
    const float t3 = t * (inv * inv);
    sx = fmaf(t3, dx, sx);
    sy = fmaf(t3, dy, sy);
    sz = fmaf(t3, dz, sz);
    sp += t;
  }
}

// The source loop is wave-uniform; only the 64 sources with the wave's own
// indices can be a lane's target, so only that stretch tests for it.
__global__ __launch_bounds__(kBlock) void direct_fast(const float4 *__restrict__ bodies,
                                                      uint32_t n, float G, float eps2,
                                                      float4 *__restrict__ out) {
  const uint32_t i = blockIdx.x * kBlock + threadIdx.x;
  const float4 bi = bodies[i < n ? i : n - 1];
  // (wave-uniform: the sources then come through the scalar cache)
  const uint32_t w0 =
      __builtin_amdgcn_readfirstlane(min(blockIdx.x * kBlock + (threadIdx.x & ~63u), n));
  const uint32_t w1 = min(w0 + 64u, n);
  float sx = 0.0f, sy = 0.0f, sz = 0.0f, sp = 0.0f;
  direct_range<false>(bodies, 0, w0, i, bi, eps2, sx, sy, sz, sp);
  direct_range<true>(bodies, w0, w1, i, bi, eps2, sx, sy, sz, sp);
  direct_range<false>(bodies, w1, n, i, bi, eps2, sx, sy, sz, sp);
  if (i < n) out[i] = make_float4(G * sx, G * sy, G * sz, -G * sp);
}

#pragma clang fp contract(off)
// One workgroup per sampled target; threads stride over sources and the
// partial sums are combined in a fixed tree order (deterministic).
__global__ __launch_bounds__(kBlock) void direct_f64_sample(const double *__restrict__ pos,
                                                            const double *__restrict__ mass,
                                                            uint64_t n, double G, double eps2,
                                                            const uint64_t *__restrict__ sample,
                                                            double *__restrict__ acc) {
  __shared__ double red[3][kBlock];
  const uint64_t i = sample[blockIdx.x];
  const double px = pos[3 * i], py = pos[3 * i + 1], pz = pos[3 * i + 2];
  double ax = 0.0, ay = 0.0, az = 0.0;
  for (uint64_t j = threadIdx.x; j < n; j += kBlock) {
    if (j == i) continue;
    const double dx = px - pos[3 * j], dy = py - pos[3 * j + 1], dz = pz - pos[3 * j + 2];
    const double inv = 1.0 / sqrt(dx * dx + dy * dy + dz * dz + eps2);
    const double f = G * mass[j] * inv * inv * inv;
    ax -= f * dx;
    ay -= f * dy;
    az -= f * dz;
  }
  red[0][threadIdx.x] = ax;
  red[1][threadIdx.x] = ay;
  red[2][threadIdx.x] = az;
  __syncthreads();
  for (int w = kBlock / 2; w > 0; w >>= 1) {
    if (threadIdx.x < static_cast<unsigned>(w))
      for (int a = 0; a < 3; ++a) red[a][threadIdx.x] += red[a][threadIdx.x + w];
    __syncthreads();
  }
  if (threadIdx.x < 3) acc[3 * blockIdx.x + threadIdx.x] = red[threadIdx.x][0];
}
#pragma clang fp contract(fast)

// iters x 16 operations per lane: FMAs, or with RSQ every fourth one a
// v_rsq_f32 (the evaluation's transcendental).  The sums are written so the
// chains are not dead code.
template <bool RSQ>
__global__ __launch_bounds__(kBlock) void valu_probe(float *__restrict__ out, int iters, float x,
                                                     float y) {
  const uint32_t t = blockIdx.x * kBlock + threadIdx.x;
  float a[16];
#pragma unroll
  for (int i = 0; i < 16; ++i) a[i] = static_cast<float>(t) * 1e-7f + static_cast<float>(i);
  for (int it = 0; it < iters; ++it) {
#pragma unroll
    for (int i = 0; i < 16; ++i)
      a[i] = (RSQ && (i & 3) == 0) ? __builtin_amdgcn_rsqf(a[i] + y) : fmaf(a[i], x, y);
  }
  float sum = 0.0f;
#pragma unroll
  for (int i = 0; i < 16; ++i) sum += a[i];
  out[t] = sum;
}

}  // namespace

void launch_valu_probe(float *out, uint32_t blocks, int iters, bool rsq, hipStream_t s) {
  if (rsq)
    hipLaunchKernelGGL(valu_probe<true>, dim3(blocks), dim3(kBlock), 0, s, out, iters, 0.999f,
                       1e-3f);
  else
    hipLaunchKernelGGL(valu_probe<false>, dim3(blocks), dim3(kBlock), 0, s, out, iters, 0.999f,
                       1e-3f);
  GTH_LAUNCH_CHECK();
}

void launch_direct(const float4 *bodies, uint32_t n, float G, float eps2, float4 *out,
                   hipStream_t s) {
  if (n == 0) return;
  hipLaunchKernelGGL(direct_fast, dim3((n + kBlock - 1) / kBlock), dim3(kBlock), 0, s, bodies, n,
                     G, eps2, out);
  GTH_LAUNCH_CHECK();
}

void launch_direct_f64_sample(const double *pos, const double *mass, uint64_t n, double G,
                              double eps, const uint64_t *sample, uint64_t ns, double *acc,
                              hipStream_t s) {
  if (ns == 0) return;
  hipLaunchKernelGGL(direct_f64_sample, dim3(static_cast<uint32_t>(ns)), dim3(kBlock), 0, s, pos,
                     mass, n, G, eps * eps, sample, acc);
  GTH_LAUNCH_CHECK();
}

}  // namespace gth
