// Grouped interaction-list evaluation on gfx950 (the hot loop of
// GroupTreeGravity::compute; reference: gt_tree, src/metal/gravity_kernels.metal:53-112,
// and evaluate_group_lists_cpu, src/gravity_backends.cpp:167-181).
//
// Two kernels share the list format (GTGroup / GTCell / uint32 lists):
//
//  * eval_exact: one wave64 per group, lane = target, list entries read
//    wave-uniformly (scalar loads), cells then bodies in list order, IEEE
//    1/sqrt and no FMA contraction.  Each target's sum is formed exactly as
//    gt_group_accel (gravity_kernels.h:162-170) forms it, so results are bit-
//    identical to the CPU evaluation.
//
//  * eval_fast: one wave64 per group, lane = list entry, targets wave-uniform.
//    Each lane loads one list entry (coalesced index load plus one 48 B cell
//    or 16 B body gather) and evaluates it against every target of the group,
//    keeping per-target partial sums in registers; a cross-lane reduction
//    ends the group.  Lanes are busy whenever the list is longer than 64
//    entries, whatever the group's fill, and the targets' coordinates are
//    scalar operands.  v_rsq_f32 and FMA contraction; G is folded into the
//    entry's mass and quadrupole once per entry.
#include "gt_internal.h"

namespace gth {
namespace {

constexpr int kWave = 64;
constexpr int kBlock = 256;  // four independent waves per workgroup

__device__ inline uint32_t uniform_wave_id() {
  const uint32_t w = (blockIdx.x * blockDim.x + threadIdx.x) / kWave;
  return __builtin_amdgcn_readfirstlane(w);
}

// ---- exact: the reference's arithmetic, op for op ---------------------------------

struct Acc {
  float ax, ay, az, pot;
};

// 1.0f / sqrtf(v) with both operations correctly rounded, as on the CPU.  Each
// is done in FP64 and rounded to FP32: with 53 >= 2 * 24 + 2 significand bits
// the double rounding cannot change the result of a sqrt or a division.
__device__ inline float exact_rsqrt(float v) {
  const float s = static_cast<float>(__dsqrt_rn(static_cast<double>(v)));
  return static_cast<float>(__ddiv_rn(1.0, static_cast<double>(s)));
}

#pragma clang fp contract(off)
// gravity_kernels.h:71-83
__device__ inline Acc exact_body(Acc a, float px, float py, float pz, float4 b, float G,
                                 float eps2) {
  const float dx = px - b.x;
  const float dy = py - b.y;
  const float dz = pz - b.z;
  const float inv = exact_rsqrt(dx * dx + dy * dy + dz * dz + eps2);
  const float gm = G * b.w * inv;
  const float gm3 = gm * inv * inv;
  a.ax -= gm3 * dx;
  a.ay -= gm3 * dy;
  a.az -= gm3 * dz;
  a.pot -= gm;
  return a;
}

// gravity_kernels.h:109-135
__device__ inline Acc exact_cell(Acc a, float px, float py, float pz, float4 cm, float4 q0,
                                 float4 q1, float G, float eps2, uint32_t quad) {
  const float dx = px - cm.x;
  const float dy = py - cm.y;
  const float dz = pz - cm.z;
  const float inv = exact_rsqrt(dx * dx + dy * dy + dz * dz + eps2);
  const float inv2 = inv * inv;
  const float inv3 = inv * inv2;
  const float gm = G * cm.w;
  a.ax -= gm * inv3 * dx;
  a.ay -= gm * inv3 * dy;
  a.az -= gm * inv3 * dz;
  a.pot -= gm * inv;
  if (quad != 0u) {
    const float qx = q0.x * dx + q0.w * dy + q1.x * dz;
    const float qy = q0.w * dx + q0.y * dy + q1.y * dz;
    const float qz = q1.x * dx + q1.y * dy + q0.z * dz;
    const float rqr = dx * qx + dy * qy + dz * qz;
    const float inv5 = inv3 * inv2;
    const float inv7 = inv5 * inv2;
    a.ax += G * (qx * inv5 - 2.5f * rqr * dx * inv7);
    a.ay += G * (qy * inv5 - 2.5f * rqr * dy * inv7);
    a.az += G * (qz * inv5 - 2.5f * rqr * dz * inv7);
    a.pot -= 0.5f * G * rqr * inv5;
  }
  return a;
}

__global__ __launch_bounds__(kBlock) void eval_exact(
    const float4 *__restrict__ bodies, const Cell *__restrict__ cells,
    const uint32_t *__restrict__ cell_list, const uint32_t *__restrict__ body_list,
    const gt_group *__restrict__ groups, uint32_t n_groups, float G, float eps2, uint32_t quad,
    float4 *__restrict__ out) {
  const uint32_t gi = uniform_wave_id();
  if (gi >= n_groups) return;  // the whole wave leaves together
  const uint32_t lane = threadIdx.x % kWave;
  const gt_group g = groups[gi];
  const uint32_t i = g.begin + lane;
  const bool active = i < g.end;
  float4 bi = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
  if (active) bi = bodies[i];
  Acc a = {0.0f, 0.0f, 0.0f, 0.0f};
  for (uint32_t k = g.cell_begin; k < g.cell_end; ++k) {
    const Cell c = cells[cell_list[k]];
    a = exact_cell(a, bi.x, bi.y, bi.z, c.cm, c.q0, c.q1, G, eps2, quad);
  }
  for (uint32_t k = g.body_begin; k < g.body_end; ++k) {
    const uint32_t j = body_list[k];
    const float4 bj = bodies[j];
    if (j != i) a = exact_body(a, bi.x, bi.y, bi.z, bj, G, eps2);
  }
  if (active) out[i] = make_float4(a.ax, a.ay, a.az, a.pot);
}
#pragma clang fp contract(fast)

// ---- fast: entries across lanes, targets uniform ------------------------------------

__device__ inline float wave_sum(float v) {
#pragma unroll
  for (int off = kWave / 2; off > 0; off >>= 1) v += __shfl_xor(v, off, kWave);
  return v;
}

// Position used for lanes past the end of a list chunk: far enough that the
// zero-mass entry contributes exactly zero without overflowing r^2.
constexpr float kFar = 1.0e18f;

template <int TB>
__global__ __launch_bounds__(kBlock) void eval_fast(
    const float4 *__restrict__ bodies, const Cell *__restrict__ cells,
    const uint32_t *__restrict__ cell_list, const uint32_t *__restrict__ body_list,
    const gt_group *__restrict__ groups, uint32_t n_groups, float G, float eps2, uint32_t quad,
    float4 *__restrict__ out) {
  const uint32_t gi = uniform_wave_id();
  if (gi >= n_groups) return;
  const uint32_t lane = threadIdx.x % kWave;
  const gt_group g = groups[gi];
  const uint32_t ntot = g.end - g.begin;
  const float inf = __builtin_inff();

  for (uint32_t t0 = 0; t0 < ntot; t0 += TB) {
    const uint32_t nt = min(static_cast<uint32_t>(TB), ntot - t0);
    const uint32_t tb = g.begin + t0;
    float ax[TB], ay[TB], az[TB], ap[TB];
#pragma unroll
    for (int t = 0; t < TB; ++t) ax[t] = ay[t] = az[t] = ap[t] = 0.0f;

    for (uint32_t base = g.cell_begin; base < g.cell_end; base += kWave) {
      const uint32_t k = base + lane;
      float cx = kFar, cy = kFar, cz = kFar, gm = 0.0f;
      float q0 = 0.0f, q1 = 0.0f, q2 = 0.0f, q3 = 0.0f, q4 = 0.0f, q5 = 0.0f;
      if (k < g.cell_end) {
        const Cell c = cells[cell_list[k]];
        cx = c.cm.x;
        cy = c.cm.y;
        cz = c.cm.z;
        gm = G * c.cm.w;
        if (quad) {
          q0 = G * c.q0.x;
          q1 = G * c.q0.y;
          q2 = G * c.q0.z;
          q3 = G * c.q0.w;
          q4 = G * c.q1.x;
          q5 = G * c.q1.y;
        }
      }
#pragma unroll
      for (int t = 0; t < TB; ++t) {
        if (t < static_cast<int>(nt)) {
          const float4 p = bodies[tb + t];
          const float dx = p.x - cx, dy = p.y - cy, dz = p.z - cz;
          const float r2 = dx * dx + dy * dy + dz * dz + eps2;
          const float inv = __builtin_amdgcn_rsqf(r2);
          const float inv2 = inv * inv;
          const float inv3 = inv * inv2;
          const float m3 = gm * inv3;
          const float qx = q0 * dx + q3 * dy + q4 * dz;
          const float qy = q3 * dx + q1 * dy + q5 * dz;
          const float qz = q4 * dx + q5 * dy + q2 * dz;
          const float rqr = dx * qx + dy * qy + dz * qz;
          const float inv5 = inv3 * inv2;
          const float h = rqr * inv5;
          const float s = -2.5f * h * inv2 - m3;
          ax[t] += qx * inv5 + s * dx;
          ay[t] += qy * inv5 + s * dy;
          az[t] += qz * inv5 + s * dz;
          ap[t] -= gm * inv + 0.5f * h;
        }
      }
    }

    for (uint32_t base = g.body_begin; base < g.body_end; base += kWave) {
      const uint32_t k = base + lane;
      float bx = kFar, by = kFar, bz = kFar, gm = 0.0f;
      uint32_t j = kNone;
      if (k < g.body_end) {
        j = body_list[k];
        const float4 b = bodies[j];
        bx = b.x;
        by = b.y;
        bz = b.z;
        gm = G * b.w;
      }
#pragma unroll
      for (int t = 0; t < TB; ++t) {
        if (t < static_cast<int>(nt)) {
          const float4 p = bodies[tb + t];
          const float dx = p.x - bx, dy = p.y - by, dz = p.z - bz;
          float r2 = dx * dx + dy * dy + dz * dz + eps2;
          r2 = (j == tb + t) ? inf : r2;  // a body never feels itself
          const float inv = __builtin_amdgcn_rsqf(r2);
          const float gi1 = gm * inv;
          const float g3 = gi1 * inv * inv;
          ax[t] -= g3 * dx;
          ay[t] -= g3 * dy;
          az[t] -= g3 * dz;
          ap[t] -= gi1;
        }
      }
    }

    float rx = 0.0f, ry = 0.0f, rz = 0.0f, rp = 0.0f;
#pragma unroll
    for (int t = 0; t < TB; ++t) {
      if (t < static_cast<int>(nt)) {
        const float sx = wave_sum(ax[t]), sy = wave_sum(ay[t]), sz = wave_sum(az[t]),
                    sp = wave_sum(ap[t]);
        if (lane == static_cast<uint32_t>(t)) {
          rx = sx;
          ry = sy;
          rz = sz;
          rp = sp;
        }
      }
    }
    if (lane < nt) out[tb + lane] = make_float4(rx, ry, rz, rp);
  }
}

}  // namespace

void launch_eval(const float4 *bodies, const Cell *cells, const uint32_t *cell_list,
                 const uint32_t *body_list, const gt_group *groups, uint32_t n_groups, float G,
                 float eps2, uint32_t quad, bool exact, float4 *out, hipStream_t s) {
  if (n_groups == 0) return;
  const uint32_t waves_per_block = kBlock / kWave;
  const uint32_t blocks = (n_groups + waves_per_block - 1) / waves_per_block;
  if (exact)
    hipLaunchKernelGGL(eval_exact, dim3(blocks), dim3(kBlock), 0, s, bodies, cells, cell_list,
                       body_list, groups, n_groups, G, eps2, quad, out);
  else
    hipLaunchKernelGGL(eval_fast<32>, dim3(blocks), dim3(kBlock), 0, s, bodies, cells,
                       cell_list, body_list, groups, n_groups, G, eps2, quad, out);
  GTH_LAUNCH_CHECK();
}

}  // namespace gth
