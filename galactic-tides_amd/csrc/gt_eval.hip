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
//    read from LDS as broadcasts.  Group sizes are bucketed (multiples of 4
//    targets per batch) so each path is branch-free over its targets.
//    v_rsq_f32 and FMA chains; G is folded into the entry's mass and
//    quadrupole once per entry; per-target sums are reduced with DPP.
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
    const gt_group *__restrict__ groups, const unsigned long long *__restrict__ frame,
    uint32_t group_cap, float G, float eps2, uint32_t quad, float4 *__restrict__ out) {
  if (frame[frame::kStatus]) return;  // lists incomplete: evaluate nothing
  const uint32_t n_groups = bounded_count(frame + frame::kGroups, group_cap);
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

// Sum over the 64 lanes with DPP moves (no LDS traffic): pairs, quads, half
// rows and rows by permutes within a row, then rows combined with the row
// broadcasts; the total ends in lane 63.  Fixed order, so deterministic.
template <int CTRL, int ROW_MASK>
__device__ inline float dpp_add(float v) {
  const int moved = __builtin_amdgcn_update_dpp(0, __float_as_int(v), CTRL, ROW_MASK, 0xf, false);
  return v + __int_as_float(moved);
}

__device__ inline float wave_total(float v) {
  v = dpp_add<0xb1, 0xf>(v);   // quad_perm [1,0,3,2]
  v = dpp_add<0x4e, 0xf>(v);   // quad_perm [2,3,0,1]
  v = dpp_add<0x141, 0xf>(v);  // row_half_mirror
  v = dpp_add<0x140, 0xf>(v);  // row_mirror
  v = dpp_add<0x142, 0xa>(v);  // row_bcast:15 into rows 1 and 3
  v = dpp_add<0x143, 0xc>(v);  // row_bcast:31 into rows 2 and 3
  return __int_as_float(__builtin_amdgcn_readlane(__float_as_int(v), 63));
}

// Position used for lanes past the end of a list chunk: far enough that the
// zero-mass entry contributes exactly zero without overflowing r^2.
constexpr float kFar = 1.0e18f;

// A list entry as one lane holds it: position, -G m and (cells) G Q.
struct Entry {
  float x, y, z, ngm;
  float q0, q1, q2, q3, q4, q5;
};

template <bool QUAD>
__device__ inline Entry load_cell(const Cell *__restrict__ cells, uint32_t idx, bool valid, float G) {
  Entry e = {kFar, kFar, kFar, 0.0f, 0.0f, 0.0f, 0.0f, 0.0f, 0.0f, 0.0f};
  if (valid) {
    const float4 cm = cells[idx].cm;
    e.x = cm.x;
    e.y = cm.y;
    e.z = cm.z;
    e.ngm = -G * cm.w;
    if (QUAD) {
      const float4 a = cells[idx].q0;
      const float2 b = *reinterpret_cast<const float2 *>(&cells[idx].q1);
      e.q0 = G * a.x;
      e.q1 = G * a.y;
      e.q2 = G * a.z;
      e.q3 = G * a.w;
      e.q4 = G * b.x;
      e.q5 = G * b.y;
    }
  }
  return e;
}

// Keeps the compiler from hoisting the LDS target reads out of the chunk
// loops (which would hold every target's coordinates in registers): the
// offset is zero, but only known to be so at run time.
__device__ inline uint32_t opaque_zero() {
  uint32_t z = 0;
  asm volatile("" : "+s"(z));
  return z;
}

// One batch of TB targets (the group's targets tb .. tb+nt-1, nt <= TB) against
// the whole list.  Targets sit in LDS and are read as broadcasts; lane k holds
// list entry base + k.  Targets past nt repeat the last one and are not stored.
// The list gathers are software-pipelined: the index of chunk c+2 and the
// entry of chunk c+1 are in flight while chunk c is evaluated.
template <int TB, bool QUAD>
__device__ inline void eval_batch(const float4 *__restrict__ bodies, const Cell *__restrict__ cells,
                                  const uint32_t *__restrict__ cell_list,
                                  const uint32_t *__restrict__ body_list, const gt_group &g,
                                  uint32_t tb, uint32_t nt, float G, float eps2, float4 *tgt,
                                  uint32_t lane, float4 *__restrict__ out) {
  if (lane < TB) tgt[lane] = bodies[tb + min(lane, nt - 1)];
  __builtin_amdgcn_wave_barrier();
  float ax[TB], ay[TB], az[TB], ap[TB];
#pragma unroll
  for (int t = 0; t < TB; ++t) ax[t] = ay[t] = az[t] = ap[t] = 0.0f;

  {
    const uint32_t cb = g.cell_begin, ce = g.cell_end;
    uint32_t k = cb + lane;
    uint32_t idx = k < ce ? cell_list[k] : 0u;
    uint32_t idx_next = k + kWave < ce ? cell_list[k + kWave] : 0u;
    Entry e = load_cell<QUAD>(cells, idx, k < ce, G);
    for (uint32_t base = cb; base < ce; base += kWave) {
      const uint32_t kn = base + kWave + lane;
      const Entry en = load_cell<QUAD>(cells, idx_next, kn < ce, G);
      const uint32_t k2 = kn + kWave;
      idx_next = k2 < ce ? cell_list[k2] : 0u;
      const uint32_t z = opaque_zero();
#pragma unroll
      for (int t = 0; t < TB; ++t) {
        const float4 p = tgt[t + z];
        const float dx = p.x - e.x, dy = p.y - e.y, dz = p.z - e.z;
        const float r2 = fmaf(dx, dx, fmaf(dy, dy, fmaf(dz, dz, eps2)));
        const float inv = __builtin_amdgcn_rsqf(r2);
        const float inv2 = inv * inv;
        const float inv3 = inv * inv2;
        const float m3 = e.ngm * inv3;  // -G m / r^3
        if (QUAD) {
          const float qx = fmaf(e.q0, dx, fmaf(e.q3, dy, e.q4 * dz));
          const float qy = fmaf(e.q3, dx, fmaf(e.q1, dy, e.q5 * dz));
          const float qz = fmaf(e.q4, dx, fmaf(e.q5, dy, e.q2 * dz));
          const float rqr = fmaf(dx, qx, fmaf(dy, qy, dz * qz));
          const float inv5 = inv3 * inv2;
          const float h = rqr * inv5;
          const float s = fmaf(-2.5f * h, inv2, m3);
          ax[t] = fmaf(s, dx, fmaf(qx, inv5, ax[t]));
          ay[t] = fmaf(s, dy, fmaf(qy, inv5, ay[t]));
          az[t] = fmaf(s, dz, fmaf(qz, inv5, az[t]));
          ap[t] = fmaf(-0.5f, h, fmaf(e.ngm, inv, ap[t]));
        } else {
          ax[t] = fmaf(m3, dx, ax[t]);
          ay[t] = fmaf(m3, dy, ay[t]);
          az[t] = fmaf(m3, dz, az[t]);
          ap[t] = fmaf(e.ngm, inv, ap[t]);
        }
      }
      e = en;
    }
  }

  {
    const float inf = __builtin_inff();
    const uint32_t bb = g.body_begin, be = g.body_end;
    uint32_t k = bb + lane;
    uint32_t j = k < be ? body_list[k] : kNone;
    uint32_t j_next = k + kWave < be ? body_list[k + kWave] : kNone;
    float4 b = j != kNone ? bodies[j] : make_float4(kFar, kFar, kFar, 0.0f);
    for (uint32_t base = bb; base < be; base += kWave) {
      const float4 bn = j_next != kNone ? bodies[j_next] : make_float4(kFar, kFar, kFar, 0.0f);
      const uint32_t k2 = base + 2 * kWave + lane;
      const uint32_t j2 = k2 < be ? body_list[k2] : kNone;
      const float ngm = -G * b.w;
      const uint32_t z = opaque_zero();
#pragma unroll
      for (int t = 0; t < TB; ++t) {
        const float4 p = tgt[t + z];
        const float dx = p.x - b.x, dy = p.y - b.y, dz = p.z - b.z;
        float r2 = fmaf(dx, dx, fmaf(dy, dy, fmaf(dz, dz, eps2)));
        r2 = (j == tb + t) ? inf : r2;  // a body never feels itself
        const float inv = __builtin_amdgcn_rsqf(r2);
        const float gi = ngm * inv;
        const float g3 = gi * inv * inv;
        ax[t] = fmaf(g3, dx, ax[t]);
        ay[t] = fmaf(g3, dy, ay[t]);
        az[t] = fmaf(g3, dz, az[t]);
        ap[t] += gi;
      }
      b = bn;
      j = j_next;
      j_next = j2;
    }
  }

  float rx = 0.0f, ry = 0.0f, rz = 0.0f, rp = 0.0f;
#pragma unroll
  for (int t = 0; t < TB; ++t) {
    const float sx = wave_total(ax[t]), sy = wave_total(ay[t]), sz = wave_total(az[t]),
                sp = wave_total(ap[t]);
    if (lane == static_cast<uint32_t>(t)) {
      rx = sx;
      ry = sy;
      rz = sz;
      rp = sp;
    }
  }
  if (lane < nt) out[tb + lane] = make_float4(rx, ry, rz, rp);
  __builtin_amdgcn_wave_barrier();  // targets of the next batch overwrite tgt
}

// WAVES: waves per SIMD the register budget is cut for (2: no spills; 3: a
// few spilled registers for more latency hiding).  A tuning parameter.
template <bool QUAD, int WAVES>
__global__ __launch_bounds__(kBlock, WAVES) void eval_fast(
    const float4 *__restrict__ bodies, const Cell *__restrict__ cells,
    const uint32_t *__restrict__ cell_list, const uint32_t *__restrict__ body_list,
    const gt_group *__restrict__ groups, const unsigned long long *__restrict__ frame,
    uint32_t group_cap, float G, float eps2, float4 *__restrict__ out) {
  __shared__ float4 targets[kBlock / kWave][32];
  if (frame[frame::kStatus]) return;  // lists incomplete: evaluate nothing
  const uint32_t n_groups = bounded_count(frame + frame::kGroups, group_cap);
  // XCD-aware: each XCD evaluates a contiguous run of (spatially adjacent)
  // groups, whose lists gather the same cells, so its L2 holds them.
  const uint32_t waves_per_block = kBlock / kWave;
  const uint32_t blk =
      xcd_block_used(blockIdx.x, (n_groups + waves_per_block - 1) / waves_per_block);
  if (blk == kNone) return;
  const uint32_t gi = __builtin_amdgcn_readfirstlane(blk * waves_per_block + threadIdx.x / kWave);
  if (gi >= n_groups) return;
  const uint32_t lane = threadIdx.x % kWave;
  float4 *tgt = targets[threadIdx.x / kWave];
  const gt_group g = groups[gi];
  const uint32_t ntot = g.end - g.begin;
  for (uint32_t t0 = 0; t0 < ntot; t0 += 32) {
    const uint32_t nt = min(32u, ntot - t0);
    const uint32_t tb = g.begin + t0;
#define GTH_BATCH(TB) \
  eval_batch<TB, QUAD>(bodies, cells, cell_list, body_list, g, tb, nt, G, eps2, tgt, lane, out)
    // Targets per batch rounded up to a multiple of 4: each path is
    // branch-free over its targets, and at most 3 targets are padding.
    switch ((nt + 3) / 4) {
      case 1: GTH_BATCH(4); break;
      case 2: GTH_BATCH(8); break;
      case 3: GTH_BATCH(12); break;
      case 4: GTH_BATCH(16); break;
      case 5: GTH_BATCH(20); break;
      case 6: GTH_BATCH(24); break;
      case 7: GTH_BATCH(28); break;
      default: GTH_BATCH(32); break;
    }
#undef GTH_BATCH
  }
}

}  // namespace

void launch_eval(const float4 *bodies, const Cell *cells, const uint32_t *cell_list,
                 const uint32_t *body_list, const gt_group *groups,
                 const unsigned long long *frame, uint32_t group_cap, float G, float eps2,
                 uint32_t quad, bool exact, int waves, float4 *out, hipStream_t s) {
  if (group_cap == 0) return;
  const uint32_t waves_per_block = kBlock / kWave;
  const uint32_t blocks = (group_cap + waves_per_block - 1) / waves_per_block;
  const dim3 grid(exact ? blocks : xcd_grid(blocks)), block(kBlock);
#define GTH_EVAL_ARGS bodies, cells, cell_list, body_list, groups, frame, group_cap, G, eps2, out
  if (exact)
    hipLaunchKernelGGL(eval_exact, grid, block, 0, s, bodies, cells, cell_list, body_list, groups,
                       frame, group_cap, G, eps2, quad, out);
  else if (quad && waves == 3)
    hipLaunchKernelGGL((eval_fast<true, 3>), grid, block, 0, s, GTH_EVAL_ARGS);
  else if (quad)
    hipLaunchKernelGGL((eval_fast<true, 2>), grid, block, 0, s, GTH_EVAL_ARGS);
  else if (waves == 3)
    hipLaunchKernelGGL((eval_fast<false, 3>), grid, block, 0, s, GTH_EVAL_ARGS);
  else
    hipLaunchKernelGGL((eval_fast<false, 2>), grid, block, 0, s, GTH_EVAL_ARGS);
#undef GTH_EVAL_ARGS
  GTH_LAUNCH_CHECK();
}

}  // namespace gth
