// Grouped interaction-list evaluation on gfx950 (the hot loop of
// GroupTreeGravity::compute; reference: gt_tree, src/metal/gravity_kernels.metal:53-112,
// and evaluate_group_lists_cpu, src/gravity_backends.cpp:70-84).
//
// Two kernels share the list format (GTGroup / GTCell / uint32 lists):
//
//  * eval_exact: one wave64 per group, lane = target, list entries read
//    wave-uniformly (scalar loads), cells then bodies in list order, IEEE
//    1/sqrt and no FMA contraction.  Each target's sum is formed exactly as
//    gt_group_accel (gravity_kernels.h:162-170) forms it, so results are bit-
//    identical to the CPU evaluation.
//
//  * eval_fast: one wave64 per work item (a group, or one part of a group
//    whose lists are longer than the split length: the parts run on
//    different waves and their per-target sums are added in part order by
//    combine_parts), lane = list entry, targets wave-uniform.
//    Each lane loads one list entry (coalesced index load plus one 48 B cell
//    or 16 B body gather) and evaluates it against every target of the group,
//    keeping per-target partial sums in registers; a cross-lane reduction
//    ends the group.  Lanes are busy whenever the list is longer than 64
//    entries, whatever the group's fill, and the targets' coordinates are
//    read from LDS as broadcasts.  Group sizes are bucketed (multiples of 4
//    targets per batch) so each path is branch-free over its targets.
//    v_rsq_f32 and FMA chains; G is folded into the entry's mass and
//    quadrupole once per entry; the per-target sums are reduced all together
//    by a reduce-scatter butterfly (permlane swaps and DPP).
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

// Wave sums of many per-lane values at once, scattered over the lanes (a
// reduce-scatter butterfly, no LDS traffic).  Each step pairs every lane
// with a partner whose lane number differs in bit log2(D) and halves the
// values each lane holds: the lane keeps the lower or the upper half (by that
// bit), adds the partner's copy of it, and hands over the other half.  Lanes
// 32 and 16 apart trade halves with v_permlane32_swap / v_permlane16_swap
// (one instruction per pair of values); within a row the partner's value
// comes by DPP (row_mirror, row_half_mirror, quad permutes).  Once a lane
// holds an odd number of values the remaining steps add the partner's copy
// of every value (an all-reduce), so 2^r lanes end with the same sums.
// Fixed pairing and order, so deterministic.

// The value of partner lane (lane ^ ...) in a step of distance D within a row.
template <int D>
__device__ inline float partner_in_row(float v) {
  // D = 8: row_mirror (i <-> 15 - i), 4: row_half_mirror (i <-> 7 - i),
  // 2: quad_perm [2, 3, 0, 1], 1: quad_perm [1, 0, 3, 2]
  constexpr int ctrl = D == 8 ? 0x140 : D == 4 ? 0x141 : D == 2 ? 0x4e : 0xb1;
  return __int_as_float(__builtin_amdgcn_update_dpp(0, __float_as_int(v), ctrl, 0xf, 0xf, false));
}

// Both values of a pair after a swap of lanes D = 32 or 16 apart: the lane
// holds its own value of the half it keeps and its partner's.
template <int D>
__device__ inline float swapped_sum(float keep_lo, float keep_hi) {
  const auto r =
      D == 32 ? __builtin_amdgcn_permlane32_swap(__float_as_uint(keep_lo), __float_as_uint(keep_hi),
                                                 false, false)
              : __builtin_amdgcn_permlane16_swap(__float_as_uint(keep_lo), __float_as_uint(keep_hi),
                                                 false, false);
  return __uint_as_float(r[0]) + __uint_as_float(r[1]);
}

// Halving steps a butterfly over n values makes (while n is even, at most 6).
constexpr int rs_halvings(int n) {
  int h = 0;
  while (h < 6 && n % 2 == 0) {
    n /= 2;
    ++h;
  }
  return h;
}

// Step S (distance 32 >> S) of the butterfly over the M values acc[0 .. M);
// base: the index of acc[0] among the wave's values so far.
template <int M, int S>
__device__ inline void reduce_scatter_step(float *acc, uint32_t lane, uint32_t &base) {
  if constexpr (S < 6) {
    constexpr int D = 32 >> S;
    if constexpr (M % 2 == 0) {
      constexpr int H = M / 2;
      const bool hi = (lane & D) != 0;
#pragma unroll
      for (int i = 0; i < H; ++i) {
        if constexpr (D >= 16) {
          acc[i] = swapped_sum<D>(acc[i], acc[i + H]);
        } else {
          const float send = hi ? acc[i] : acc[i + H];
          const float keep = hi ? acc[i + H] : acc[i];
          acc[i] = keep + partner_in_row<D>(send);
        }
      }
      base += hi ? H : 0;
      reduce_scatter_step<H, S + 1>(acc, lane, base);
    } else {
#pragma unroll
      for (int j = 0; j < M; ++j) {
        if constexpr (D >= 16)
          acc[j] = swapped_sum<D>(acc[j], acc[j]);
        else
          acc[j] = acc[j] + partner_in_row<D>(acc[j]);
      }
      reduce_scatter_step<M, S + 1>(acc, lane, base);
    }
  }
}

// Position used for lanes past the end of a list chunk: far enough that the
// zero-mass entry contributes exactly zero without overflowing r^2.
constexpr float kFar = 1.0e18f;

// A list entry as one lane holds it: position, -G m and (cells) G Q.
struct Entry {
  float x, y, z, ngm;
  float q0, q1, q2, q3, q4, q5;
};

// Hi/lo positions (precision mode): every position p relative to its system's
// origin is hi + lo (two FP32, hi = fl(p), lo = fl(p - hi)); the wave's group
// centre likewise is (gx, gy, gz) + (lx, ly, lz).  Positions are then taken
// relative to the group centre, (hi - g_hi) + (lo - g_lo), which keeps the
// differences of nearby positions exact to FP32 rounding at their own scale.
struct HiLo {
  const float4 *__restrict__ bodies_lo;
  const float4 *__restrict__ cells_lo;
  float gx, gy, gz, lx, ly, lz;
};

template <bool HILO>
__device__ inline void relative(float &x, float &y, float &z, const float4 *__restrict__ lo_arr,
                                uint32_t idx, const HiLo &h) {
  if (HILO) {
    const float4 lo = lo_arr[idx];
    x = (x - h.gx) + (lo.x - h.lx);
    y = (y - h.gy) + (lo.y - h.ly);
    z = (z - h.gz) + (lo.z - h.lz);
  }
}

// Cell idx of a list.  A lane past the end of the list (valid false) holds
// the list's last cell (idx is that cell) with its mass and quadrupole set
// to zero: an accepted cell is apart from every target of the group, so its
// separation is nonzero and a massless copy adds exactly nothing.  (No
// branch: the entry's registers are the same on every lane.)
template <bool QUAD, bool HILO>
__device__ inline Entry load_cell(const Cell *__restrict__ cells, uint32_t idx, bool valid, float G,
                                  const HiLo &h) {
  Entry e;
  const float4 cm = cells[idx].cm;
  e.x = cm.x;
  e.y = cm.y;
  e.z = cm.z;
  relative<HILO>(e.x, e.y, e.z, h.cells_lo, idx, h);
  const float gv = valid ? G : 0.0f;
  e.ngm = -gv * cm.w;
  e.q0 = e.q1 = e.q2 = e.q3 = e.q4 = e.q5 = 0.0f;
  if (QUAD) {
    const float4 a = cells[idx].q0;
    const float2 b = *reinterpret_cast<const float2 *>(&cells[idx].q1);
    e.q0 = gv * a.x;
    e.q1 = gv * a.y;
    e.q2 = gv * a.z;
    e.q3 = gv * a.w;
    e.q4 = gv * b.x;
    e.q5 = gv * b.y;
  }
  return e;
}

// Body j of a list (kNone: a lane past the end, far away and massless).
template <bool HILO>
__device__ inline float4 load_body(const float4 *__restrict__ bodies, uint32_t j, const HiLo &h) {
  if (j == kNone) return make_float4(kFar, kFar, kFar, 0.0f);
  float4 b = bodies[j];
  relative<HILO>(b.x, b.y, b.z, h.bodies_lo, j, h);
  return b;
}

// The list entries one wave evaluates: cells [cb, ce), bodies [bb, be).
struct Range {
  uint32_t cb, ce, bb, be;
};

// Part `part` of `parts` of a group's lists: the combined sequence (cells,
// then bodies) cut into pieces of whole 64-entry chunks.
__device__ inline Range part_range(const gt_group &g, uint32_t part, uint32_t parts) {
  if (parts <= 1) return Range{g.cell_begin, g.cell_end, g.body_begin, g.body_end};
  const uint32_t nc = g.cell_end - g.cell_begin, len = nc + (g.body_end - g.body_begin);
  const uint32_t per = ((len + parts - 1) / parts + kWave - 1) / kWave * kWave;
  const uint32_t lo = min(part * per, len), hi = min(lo + per, len);
  Range r;
  r.cb = g.cell_begin + min(lo, nc);
  r.ce = g.cell_begin + min(hi, nc);
  r.bb = g.body_begin + (lo > nc ? lo - nc : 0u);
  r.be = g.body_begin + (hi > nc ? hi - nc : 0u);
  return r;
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
// entry of chunk c+1 are in flight while chunk c is evaluated.  accumulate:
// the batch's sums are added to what dst holds (dst + sums, as combine_parts
// adds a part to the parts before it) instead of stored.
template <int TB, bool QUAD, bool HILO>
__device__ inline void eval_batch(const float4 *__restrict__ bodies, const Cell *__restrict__ cells,
                                  const uint32_t *__restrict__ cell_list,
                                  const uint32_t *__restrict__ body_list, const Range &r,
                                  uint32_t tb, uint32_t nt, float G, float eps2, float4 *tgt,
                                  uint32_t lane, float4 *__restrict__ dst, const HiLo &h,
                                  bool accumulate) {
  if (lane < TB) {
    const uint32_t t = tb + min(lane, nt - 1);
    float4 p = bodies[t];
    relative<HILO>(p.x, p.y, p.z, h.bodies_lo, t, h);
    tgt[lane] = p;
  }
  __builtin_amdgcn_wave_barrier();
  // per-target partial sums: acc[4 t + c], c = x, y, z, potential
  float acc[4 * TB];
#pragma unroll
  for (int k = 0; k < 4 * TB; ++k) acc[k] = 0.0f;

  {
    // (zb keeps the compiler from hoisting the cell list's addresses out of
    // the batch loop, where they would be spilled: recomputing them is cheaper
    // and takes ~110 MB of spill writes off a 2M evaluation; the same for the
    // body list made the evaluation slower)
    const uint32_t zb = opaque_zero();
    const uint32_t cb = r.cb + zb, ce = r.ce;
    uint32_t k = cb + lane;
    // lanes past the end take the list's last cell (see load_cell)
    const uint32_t last = ce > cb ? cell_list[ce - 1] : 0u;
    uint32_t idx = k < ce ? cell_list[k] : last;
    uint32_t idx_next = k + kWave < ce ? cell_list[k + kWave] : last;
    Entry e = load_cell<QUAD, HILO>(cells, idx, k < ce, G, h);
    for (uint32_t base = cb; base < ce; base += kWave) {
      const uint32_t kn = base + kWave + lane;
      const Entry en = load_cell<QUAD, HILO>(cells, idx_next, kn < ce, G, h);
      const uint32_t k2 = kn + kWave;
      idx_next = k2 < ce ? cell_list[k2] : last;
      const uint32_t z = opaque_zero();
#pragma unroll
      for (int t = 0; t < TB; ++t) {
        const float4 p = tgt[t + z];
        const float dx = p.x - e.x, dy = p.y - e.y, dz = p.z - e.z;
        const float r2 = fmaf(dx, dx, fmaf(dy, dy, fmaf(dz, dz, eps2)));
        const float inv = __builtin_amdgcn_rsqf(r2);
        const float inv2 = inv * inv;
        if (QUAD) {
          // The reference's monopole and quadrupole terms (gravity_kernels.h:
          // 109-135) with the powers of 1/r shared, 33 operations:
          //   a   = r^-5 (u r^2 d + q),   u = -G m - 2.5 rqr r^-4
          //   pot = r^-1 (-G m - 0.5 rqr r^-4),   q = G Q d, rqr = d . q
          const float qx = fmaf(e.q0, dx, fmaf(e.q3, dy, e.q4 * dz));
          const float qy = fmaf(e.q3, dx, fmaf(e.q1, dy, e.q5 * dz));
          const float qz = fmaf(e.q4, dx, fmaf(e.q5, dy, e.q2 * dz));
          const float rqr = fmaf(dx, qx, fmaf(dy, qy, dz * qz));
          const float inv4 = inv2 * inv2;
          const float inv5 = inv4 * inv;
          const float t4 = rqr * inv4;
          const float w = fmaf(-2.5f, t4, e.ngm) * r2;
          acc[4 * t] = fmaf(fmaf(w, dx, qx), inv5, acc[4 * t]);
          acc[4 * t + 1] = fmaf(fmaf(w, dy, qy), inv5, acc[4 * t + 1]);
          acc[4 * t + 2] = fmaf(fmaf(w, dz, qz), inv5, acc[4 * t + 2]);
          acc[4 * t + 3] = fmaf(fmaf(-0.5f, t4, e.ngm), inv, acc[4 * t + 3]);
        } else {
          const float m3 = e.ngm * (inv * inv2);  // -G m / r^3
          acc[4 * t] = fmaf(m3, dx, acc[4 * t]);
          acc[4 * t + 1] = fmaf(m3, dy, acc[4 * t + 1]);
          acc[4 * t + 2] = fmaf(m3, dz, acc[4 * t + 2]);
          acc[4 * t + 3] = fmaf(e.ngm, inv, acc[4 * t + 3]);
        }
      }
      e = en;
    }
  }

  {
    const float inf = __builtin_inff();
    const uint32_t bb = r.bb, be = r.be;
    uint32_t k = bb + lane;
    uint32_t j = k < be ? body_list[k] : kNone;
    uint32_t j_next = k + kWave < be ? body_list[k + kWave] : kNone;
    float4 b = load_body<HILO>(bodies, j, h);
    for (uint32_t base = bb; base < be; base += kWave) {
      const float4 bn = load_body<HILO>(bodies, j_next, h);
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
        acc[4 * t] = fmaf(g3, dx, acc[4 * t]);
        acc[4 * t + 1] = fmaf(g3, dy, acc[4 * t + 1]);
        acc[4 * t + 2] = fmaf(g3, dz, acc[4 * t + 2]);
        acc[4 * t + 3] += gi;
      }
      b = bn;
      j = j_next;
      j_next = j2;
    }
  }

  // The wave's sums, scattered: after H halvings lane l holds M = 4 TB >> H
  // sums, of values base .. base + M - 1, shared by R = 64 >> H lanes; each
  // of those lanes stores every R-th of them (value k = 4 t + c is target
  // t's component c), targets past nt are not stored.
  constexpr int H = rs_halvings(4 * TB), M = (4 * TB) >> H, R = 64 >> H;
  uint32_t base = 0;
  reduce_scatter_step<4 * TB, 0>(acc, lane, base);
  float *dstf = reinterpret_cast<float *>(dst);
  const uint32_t rho = lane & (R - 1);
#pragma unroll
  for (int j = 0; j < M; ++j) {
    const uint32_t k = base + j;
    if ((static_cast<uint32_t>(j) & (R - 1)) == rho && (k >> 2) < nt) {
      const float v = acc[j];
      dstf[k] = accumulate ? dstf[k] + v : v;
    }
  }
  __builtin_amdgcn_wave_barrier();  // targets of the next batch overwrite tgt
}

// WAVES: waves per SIMD the register budget is cut for (2: no spills; 3: a
// few spilled registers for more latency hiding).  A tuning parameter.
//
// Work items: without `items`, item k is group k; with them, item k is part
// (items[k].y >> 16) of (items[k].y & 0xffff) of group items[k].x, taken in
// the order `order` gives (null: as they are).  A group
// in one part writes its targets' results to out; a part of a split group
// writes them to its slot of `partial` (kMaxGroup float4 each, slot
// split_off[group] + part), which combine_parts adds up in part order.
// Without items a group whose list is longer than split_len entries (0:
// none) is cut into the same parts, which this wave evaluates one after
// another, adding each part's sums to the results in part order: the sums
// combine_parts forms from parts evaluated on different waves, bit for bit.
template <bool QUAD, int WAVES, bool HILO>
__global__ __launch_bounds__(kBlock, WAVES) void eval_fast(
    const float4 *__restrict__ bodies, const Cell *__restrict__ cells,
    const uint32_t *__restrict__ cell_list, const uint32_t *__restrict__ body_list,
    const gt_group *__restrict__ groups, const unsigned long long *__restrict__ frame,
    uint32_t item_cap, const uint2 *__restrict__ items, const uint32_t *__restrict__ order,
    const uint32_t *__restrict__ split_off, float4 *__restrict__ partial, float G, float eps2,
    const float4 *__restrict__ bodies_lo, const float4 *__restrict__ cells_lo,
    const float4 *__restrict__ group_hilo, float4 *__restrict__ out, uint32_t group_cap,
    uint32_t g_begin, uint32_t g_span, uint32_t *__restrict__ queue, uint32_t split_len) {
  __shared__ float4 targets[kBlock / kWave][32];
  if (frame[frame::kStatus]) return;  // lists incomplete: evaluate nothing
  // items: this range's work items; else groups g_begin .. g_begin + g_span
  uint32_t n_items;
  if (items) {
    n_items = bounded_count(frame + frame::kItems, item_cap);
  } else {
    const uint32_t ng = bounded_count(frame + frame::kGroups, group_cap);
    n_items = ng > g_begin ? min(ng - g_begin, g_span) : 0u;
  }
  // XCD-aware: each XCD evaluates a contiguous run of (spatially adjacent)
  // groups, whose lists gather the same cells, so its L2 holds them.
  const uint32_t waves_per_block = kBlock / kWave;
  const uint32_t used = (n_items + waves_per_block - 1) / waves_per_block;
  // queue: a persistent grid; the waves on each XCD take that XCD's share of
  // the items (the same shares as the static mapping) one at a time from the
  // XCD's counter, in order.  A wave whose XCD's share is used up takes the
  // next items of the other XCDs' shares, nearest XCD number first, so no XCD
  // idles while another still has a tail of items.
  const uint32_t q_items =
      queue ? (used + kXcds - 1) / kXcds * waves_per_block : 0u;  // items per share
  const uint32_t home = blockIdx.x % kXcds;
  uint32_t share = 0;  // home + share (mod 8): the share being taken from
  for (uint32_t step = 0;; ++step) {
  uint32_t wi;
  if (queue) {
    wi = n_items;
    while (share < kXcds) {
      const uint32_t x = (home + share) % kXcds;
      uint32_t t = 0;
      if (threadIdx.x % kWave == 0) t = atomicAdd(queue + x * kQueueStride, 1u);
      t = __builtin_amdgcn_readfirstlane(t);
      const uint32_t lo = min(x * q_items, n_items), hi = min(lo + q_items, n_items);
      if (t < hi - lo) {
        wi = lo + t;
        break;
      }
      ++share;  // (this share is used up: its counter only grows from here)
    }
    if (wi >= n_items) break;
  } else {
    if (step) break;
    const uint32_t blk = xcd_block_used(blockIdx.x, used);
    if (blk == kNone) break;
    wi = __builtin_amdgcn_readfirstlane(blk * waves_per_block + threadIdx.x / kWave);
    if (wi >= n_items) break;
  }
  uint32_t gi = g_begin + wi, part = 0, parts = 1;
  if (items) {
    const uint2 it = items[order ? order[wi] : wi];
    gi = __builtin_amdgcn_readfirstlane(it.x);
    part = __builtin_amdgcn_readfirstlane(it.y >> 16);
    parts = __builtin_amdgcn_readfirstlane(it.y & 0xffffu);
  }
  const uint32_t lane = threadIdx.x % kWave;
  float4 *tgt = targets[threadIdx.x / kWave];
  const gt_group g = groups[gi];
  uint32_t inline_parts = 1;
  if (!items && split_len) {
    const uint32_t len = (g.cell_end - g.cell_begin) + (g.body_end - g.body_begin);
    inline_parts = len > split_len ? min((len + split_len - 1) / split_len, 0xffffu) : 1u;
    inline_parts = __builtin_amdgcn_readfirstlane(inline_parts);
  }
  HiLo h = {bodies_lo, cells_lo, 0.0f, 0.0f, 0.0f, 0.0f, 0.0f, 0.0f};
  if (HILO) {
    const float4 hi = group_hilo[2 * static_cast<size_t>(gi)];
    const float4 lo = group_hilo[2 * static_cast<size_t>(gi) + 1];
    h.gx = hi.x;
    h.gy = hi.y;
    h.gz = hi.z;
    h.lx = lo.x;
    h.ly = lo.y;
    h.lz = lo.z;
  }
  float4 *dst0 = parts > 1 ? partial + static_cast<size_t>(split_off[gi - g_begin] + part) *
                                          kMaxGroup
                           : out + g.begin;
  const uint32_t ntot = g.end - g.begin;
  for (uint32_t ip = 0; ip < inline_parts; ++ip) {
  const Range r = inline_parts > 1 ? part_range(g, ip, inline_parts) : part_range(g, part, parts);
  for (uint32_t t0 = 0; t0 < ntot; t0 += 32) {
    const uint32_t nt = min(32u, ntot - t0);
    const uint32_t tb = g.begin + t0;
    float4 *dst = dst0 + t0;
#define GTH_BATCH(TB)                                                                      \
  eval_batch<TB, QUAD, HILO>(bodies, cells, cell_list, body_list, r, tb, nt, G, eps2, tgt, lane, \
                             dst, h, ip > 0)
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
  }
  // The last wave out of the persistent grid clears the counters for the next
  // launch (every wave has taken its last item by then: its taking waited for
  // the counter's value), so no fill launch is needed between evaluations.
  // queue[kXcds * kQueueStride] counts the waves out.
  if (queue && threadIdx.x % kWave == 0) {
    uint32_t *const out_count = queue + kXcds * kQueueStride;
    if (atomicAdd(out_count, 1u) == gridDim.x * waves_per_block - 1u) {
      for (uint32_t x = 0; x < kXcds; ++x)
        __hip_atomic_store(queue + x * kQueueStride, 0u, __ATOMIC_RELAXED,
                           __HIP_MEMORY_SCOPE_AGENT);
      __hip_atomic_store(out_count, 0u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    }
  }
}

// ---- long lists in parts -----------------------------------------------------------

// Parts of every group (lists longer than split_len entries are cut into
// ceil(len / split_len) parts), and the parts of split groups only.  Entries
// past the group count are zero, so the scans over group_cap entries add up.
__global__ __launch_bounds__(kBlock) void count_parts(const gt_group *__restrict__ groups,
                                                      const unsigned long long *__restrict__ frame,
                                                      uint32_t group_cap, uint32_t g_begin,
                                                      uint32_t g_span, uint32_t split_len,
                                                      uint32_t *__restrict__ parts,
                                                      uint32_t *__restrict__ split_parts) {
  const uint32_t ng = bounded_count(frame + frame::kGroups, group_cap);
  const uint32_t count = ng > g_begin ? min(ng - g_begin, g_span) : 0u;
  const uint32_t i = blockIdx.x * kBlock + threadIdx.x;  // index within the range
  if (i >= g_span) return;
  uint32_t p = 0;
  if (i < count) {
    const gt_group G = groups[g_begin + i];
    const uint32_t len = (G.cell_end - G.cell_begin) + (G.body_end - G.body_begin);
    p = len > split_len ? min((len + split_len - 1) / split_len, 0xffffu) : 1u;
  }
  parts[i] = p;
  split_parts[i] = p > 1 ? p : 0u;
}

// The work items, in group order (a split group's parts are adjacent, so they
// run at the same time on one XCD); sets kAbortItems past the plan's bounds.
__global__ __launch_bounds__(kBlock) void make_items(unsigned long long *__restrict__ frame,
                                                     uint32_t group_cap, uint32_t g_begin,
                                                     uint32_t g_span, uint32_t item_cap,
                                                     uint32_t split_cap,
                                                     const uint32_t *__restrict__ parts,
                                                     const uint32_t *__restrict__ item_off,
                                                     uint2 *__restrict__ items) {
  const bool fits = frame[frame::kItems] <= item_cap && frame[frame::kSplitParts] <= split_cap;
  if (blockIdx.x == 0 && threadIdx.x == 0 && !fits) atomicOr(frame + frame::kStatus, kAbortItems);
  const uint32_t ng = bounded_count(frame + frame::kGroups, group_cap);
  const uint32_t count = ng > g_begin ? min(ng - g_begin, g_span) : 0u;
  const uint32_t i = blockIdx.x * kBlock + threadIdx.x;
  if (!fits || i >= count) return;
  const uint32_t p = parts[i], o = item_off[i];
  for (uint32_t j = 0; j < p; ++j) items[o + j] = make_uint2(g_begin + i, (j << 16) | p);
}

// Sort keys for evaluating each XCD's items longest first: the XCD chunk an
// item's position falls in (most significant: items stay on their XCD, so
// its L2 keeps serving the same region of the tree), then decreasing cost
// (targets rounded up to the batch granularity times list entries).
__global__ __launch_bounds__(kBlock) void item_keys(const unsigned long long *__restrict__ frame,
                                                    uint32_t item_cap, int buckets,
                                                    const uint2 *__restrict__ items,
                                                    const gt_group *__restrict__ groups,
                                                    uint64_t *__restrict__ keys,
                                                    uint32_t *__restrict__ order) {
  // items were not written when they did not fit (make_items said so)
  if (frame[frame::kStatus] & kAbortItems) return;
  const uint32_t n_items = bounded_count(frame + frame::kItems, item_cap);
  const uint32_t k = blockIdx.x * kBlock + threadIdx.x;
  if (k >= n_items) return;
  const uint32_t used = (n_items + kBlock / kWave - 1) / (kBlock / kWave);
  const uint32_t q = (used + kXcds - 1) / kXcds;
  const uint32_t xcd = (k / (kBlock / kWave)) / q;
  const uint2 it = items[k];
  const gt_group g = groups[it.x];
  const Range r = part_range(g, it.y >> 16, it.y & 0xffffu);
  const uint64_t cost = static_cast<uint64_t>((g.end - g.begin + 3) & ~3u) *
                        ((r.ce - r.cb) + (r.be - r.bb));
  // buckets > 0: cost classes of 1/buckets octave (items of a class keep
  // their tree order, which keeps neighbouring groups together); else the
  // cost itself.  Classes fit 8 bits (costs stay below 2^38: 64 targets x
  // 2^32 entries), so the keys of classes are 11 bits: two sort passes.
  if (buckets > 0) {
    const uint64_t cq = min(
        static_cast<uint64_t>(buckets * log2f(static_cast<float>(cost) + 1.0f)), 0xffull);
    keys[k] = (static_cast<uint64_t>(xcd) << 8) | (0xffull - cq);
  } else {
    const uint64_t cq = min(cost >> 4, 0xffffffull);
    keys[k] = (static_cast<uint64_t>(xcd) << 24) | (0xffffffull - cq);
  }
  order[k] = k;
}

// Hi/lo group centres: the centre of each group's FP64 bounding box (as the
// walk computes it) relative to its system's origin, split into hi + lo
// FP32 (group_hilo[2g], group_hilo[2g + 1]).
__global__ __launch_bounds__(kBlock) void group_centres(const gt_group *__restrict__ groups,
                                                        const unsigned long long *__restrict__ frame,
                                                        uint32_t group_cap,
                                                        const double *__restrict__ tpos,
                                                        const uint32_t *__restrict__ tsid,
                                                        const double *__restrict__ roots,
                                                        const double *__restrict__ geom,
                                                        float4 *__restrict__ group_hilo) {
  const uint32_t ng = bounded_count(frame + frame::kGroups, group_cap);
  const uint32_t g = blockIdx.x * kBlock + threadIdx.x;
  if (g >= ng) return;
  const gt_group G = groups[g];
  double lo[3], hi[3];
  for (int a = 0; a < 3; ++a) lo[a] = hi[a] = tpos[3 * static_cast<size_t>(G.begin) + a];
  for (uint32_t k = G.begin; k < G.end; ++k)
    for (int a = 0; a < 3; ++a) {
      const double v = tpos[3 * static_cast<size_t>(k) + a];
      lo[a] = v < lo[a] ? v : lo[a];
      hi[a] = hi[a] < v ? v : hi[a];
    }
  const double *o = tsid ? roots + 4 * static_cast<size_t>(tsid[G.begin]) : geom;
  float h[3], l[3];
  for (int a = 0; a < 3; ++a) {
    // each operation rounded (no contraction), as the walk forms the centre
    const double c = __dsub_rn(__dmul_rn(0.5, __dadd_rn(lo[a], hi[a])), o[a]);
    h[a] = static_cast<float>(c);
    l[a] = static_cast<float>(c - static_cast<double>(h[a]));
  }
  group_hilo[2 * static_cast<size_t>(g)] = make_float4(h[0], h[1], h[2], 0.0f);
  group_hilo[2 * static_cast<size_t>(g) + 1] = make_float4(l[0], l[1], l[2], 0.0f);
}

// Split groups: the parts' results added in part order (deterministic).
__global__ __launch_bounds__(kBlock) void combine_parts(const gt_group *__restrict__ groups,
                                                        const unsigned long long *__restrict__ frame,
                                                        uint32_t group_cap, uint32_t g_begin,
                                                        uint32_t g_span,
                                                        const uint32_t *__restrict__ parts,
                                                        const uint32_t *__restrict__ split_off,
                                                        const float4 *__restrict__ partial,
                                                        float4 *__restrict__ out) {
  if (frame[frame::kStatus]) return;
  const uint32_t ng = bounded_count(frame + frame::kGroups, group_cap);
  const uint32_t count = ng > g_begin ? min(ng - g_begin, g_span) : 0u;
  const uint32_t i = blockIdx.x * kBlock + threadIdx.x;
  if (i >= count) return;
  const uint32_t p = parts[i];
  if (p <= 1) return;
  const gt_group G = groups[g_begin + i];
  const float4 *slot = partial + static_cast<size_t>(split_off[i]) * kMaxGroup;
  for (uint32_t t = 0; t < G.end - G.begin; ++t) {
    float4 a = slot[t];
    for (uint32_t j = 1; j < p; ++j) {
      const float4 b = slot[static_cast<size_t>(j) * kMaxGroup + t];
      a.x += b.x;
      a.y += b.y;
      a.z += b.z;
      a.w += b.w;
    }
    out[G.begin + t] = a;
  }
}

}  // namespace

void launch_eval(const float4 *bodies, const Cell *cells, const uint32_t *cell_list,
                 const uint32_t *body_list, const gt_group *groups,
                 const unsigned long long *frame, uint32_t group_cap, float G, float eps2,
                 uint32_t quad, bool exact, int waves, float4 *out, hipStream_t s,
                 const EvalSplit *split, const EvalHiLo *hilo, uint32_t g_begin,
                 uint32_t g_span, uint32_t *queue, uint32_t persistent_blocks,
                 uint32_t inline_split) {
  if (group_cap == 0) return;
  if (g_span == 0 || g_span > group_cap) g_span = group_cap;
  if (exact && (g_begin != 0 || g_span != group_cap))
    fail(GT_HIP_ERR_ARGUMENT, "exact evaluation takes all groups at once");
  const uint32_t waves_per_block = kBlock / kWave;
  const bool parts = split && !exact;
  const uint32_t item_cap = parts ? split->item_cap : g_span;
  const uint32_t blocks = (item_cap + waves_per_block - 1) / waves_per_block;
  if (exact || !persistent_blocks) queue = nullptr;
  // (queue: kXcds + 1 counters kQueueStride apart, zero at the first launch;
  // each launch leaves them zero)
  const dim3 grid(exact ? blocks : queue ? xcd_grid(persistent_blocks) : xcd_grid(blocks)),
      block(kBlock);
  const uint2 *items = parts ? split->items : nullptr;
  const uint32_t *order = parts && split->longest_first ? split->order : nullptr;
  const uint32_t *split_off = parts ? split->split_off : nullptr;
  float4 *partial = parts ? split->partial : nullptr;
  const bool hl = hilo != nullptr && !exact;
  const float4 *blo = hl ? hilo->bodies_lo : nullptr, *clo = hl ? hilo->cells_lo : nullptr,
               *ghl = hl ? hilo->group_hilo : nullptr;
#define GTH_EVAL_ARGS                                                                \
  bodies, cells, cell_list, body_list, groups, frame, item_cap, items, order, split_off, \
      partial, G, eps2, blo, clo, ghl, out, group_cap, g_begin, g_span, queue, \
      parts ? 0u : inline_split
#define GTH_EVAL(Q, W)                                                                    \
  if (hl)                                                                                \
    hipLaunchKernelGGL((eval_fast<Q, W, true>), grid, block, 0, s, GTH_EVAL_ARGS);  \
  else                                                                                   \
    hipLaunchKernelGGL((eval_fast<Q, W, false>), grid, block, 0, s, GTH_EVAL_ARGS)
  if (exact) {
    hipLaunchKernelGGL(eval_exact, grid, block, 0, s, bodies, cells, cell_list, body_list, groups,
                       frame, group_cap, G, eps2, quad, out);
  } else if (quad && waves == 3) {
    GTH_EVAL(true, 3);
  } else if (quad) {
    GTH_EVAL(true, 2);
  } else if (waves == 3) {
    GTH_EVAL(false, 3);
  } else {
    GTH_EVAL(false, 2);
  }
#undef GTH_EVAL
#undef GTH_EVAL_ARGS
  GTH_LAUNCH_CHECK();
}

void launch_combine_parts(const gt_group *groups, const unsigned long long *frame,
                          uint32_t group_cap, const EvalSplit &split, float4 *out, hipStream_t s,
                          uint32_t g_begin, uint32_t g_span) {
  if (group_cap == 0) return;
  if (g_span == 0 || g_span > group_cap) g_span = group_cap;
  hipLaunchKernelGGL(combine_parts, dim3((g_span + kBlock - 1) / kBlock), dim3(kBlock), 0, s,
                     groups, frame, group_cap, g_begin, g_span, split.parts, split.split_off,
                     split.partial, out);
  GTH_LAUNCH_CHECK();
}

void launch_group_centres(const gt_group *groups, const unsigned long long *frame,
                          uint32_t group_cap, const double *tpos, const uint32_t *tsid,
                          const double *roots, const double *geom, float4 *group_hilo,
                          hipStream_t s) {
  if (group_cap == 0) return;
  hipLaunchKernelGGL(group_centres, dim3((group_cap + kBlock - 1) / kBlock), dim3(kBlock), 0, s,
                     groups, frame, group_cap, tpos, tsid, roots, geom, group_hilo);
  GTH_LAUNCH_CHECK();
}

void launch_eval_items(const gt_group *groups, unsigned long long *frame, uint32_t group_cap,
                       const EvalSplit &split, unsigned long long *scan_tmp, hipStream_t s,
                       uint32_t g_begin, uint32_t g_span) {
  if (group_cap == 0) return;
  if (g_span == 0 || g_span > group_cap) g_span = group_cap;
  const dim3 grid((g_span + kBlock - 1) / kBlock), block(kBlock);
  hipLaunchKernelGGL(count_parts, grid, block, 0, s, groups, frame, group_cap, g_begin, g_span,
                     split.split_len, split.parts, split.split_count);
  GTH_LAUNCH_CHECK();
  exclusive_scan_u32(split.parts, split.item_off, g_span, frame + frame::kItems, scan_tmp, s);
  exclusive_scan_u32(split.split_count, split.split_off, g_span, frame + frame::kSplitParts,
                     scan_tmp, s);
  hipLaunchKernelGGL(make_items, grid, block, 0, s, frame, group_cap, g_begin, g_span,
                     split.item_cap, split.split_cap, split.parts, split.item_off, split.items);
  GTH_LAUNCH_CHECK();
  if (split.longest_first && split.item_cap) {
    hipLaunchKernelGGL(item_keys, dim3((split.item_cap + kBlock - 1) / kBlock), block, 0, s, frame,
                       split.item_cap, split.cost_buckets, split.items, groups, split.keys,
                       split.order);
    GTH_LAUNCH_CHECK();
    radix_sort_u64_dev(split.keys, split.order, split.keys_alt, split.order_alt,
                       frame + frame::kItems, split.item_cap, split.cost_buckets > 0 ? 11 : 27,
                       split.sort_tmp, s, frame + frame::kSortCtl);
  }
}

}  // namespace gth
