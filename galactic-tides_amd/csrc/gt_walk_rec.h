// The group walk's opening test, exact and as an FP32 filter (reference:
// build_group_lists' acceptance test, src/group_lists.cpp:147-154).
//
// The exact test runs in FP64 in the reference's operation order with no FMA
// contraction, so its decisions are the reference's bit for bit.  The filter
// runs the same test in FP32 on 32-byte node images (positions relative to
// the system's root centre) and returns a decision only when the FP32 result
// is farther from the threshold than a bound on the error of both
// computations; otherwise it returns kUndecided and the walk runs the exact
// test on the node's FP64 record.  Decisions are therefore always the exact
// test's; the filter only saves loading 64 bytes per node.
//
// Error bounds.  Every quantity is within the root cube of half size H around
// the origin O: |c - O|, |m - O|, |gc - O| <= H, half, gh <= H (checked, with
// 1% to spare, when the images are made: a node outside takes the exact test,
// a group outside takes it for every node).  Each FP32
// image of an FP64 value v is v rounded once (|err| <= u|v|, u = 2^-24), and
// each FP32 operation adds u times its result.  Per axis:
//   a = |c - gc| - (half * 1.0001 + gh): images of c, gc (2uH), the
//       difference (2uH), half and 1.0001f (1.3uH), the product and gh (2uH),
//       the sum (2uH) and the last difference (4uH): under 14 uH; eo = 16 uH.
//   D = |m - gc| - gh: images (2uH), difference (2uH), gh (uH), result (2uH):
//       under 8 uH = eD.
//   s = sum(D>0 ? D^2 : 0) - crit^2: a term whose sign the images may flip
//       has |D| <= eD and adds at most eD^2; otherwise |D'^2 - D^2| <=
//       (2|D'| + eD) eD + u D'^2; the two sums and crit's image and square
//       add 4u (d2 + crit^2).  es = 2 eD sum|D'| + 6 eD^2 + 4u (d2 + crit^2),
//       taken 1.5 times over to cover the rounding of its own evaluation.
// The FP64 test's own rounding (2^-53 relative) is covered by the same
// slack.  The bounds assume FP32 values that are normal and finite: the
// filter is used only for 2^-40 < H < 2^40 (half sizes stay normal to depth
// 63, squares stay finite), and a node whose crit^2 image is below 1e-30
// (where subnormal squares would break the relative bounds) is left to the
// exact test.
#pragma once

#include <stdint.h>

#ifdef __HIPCC__
#define GTH_HD __host__ __device__
#else
#define GTH_HD
#include <cmath>
#endif

namespace gth {

// 32-byte node image: centre and centre of mass relative to the system's
// root centre, crit = 2 half / theta + delta (its sign bit set when the node
// holds at most two bodies), and meta = first child (bits 0-25) | child count
// - 1 (26-28) | leaf (29).
struct WalkRec32 {
  float cx, cy, cz, crit;
  float mx, my, mz;
  uint32_t meta;
};
static_assert(sizeof(WalkRec32) == 32, "WalkRec32 layout");

constexpr uint32_t kMetaIndexMask = (1u << 26) - 1;
constexpr uint32_t kMetaLeaf = 1u << 29;

GTH_HD inline uint32_t image_meta(uint32_t first_child, uint32_t n_children) {
  return n_children == 0 ? kMetaLeaf
                         : (first_child & kMetaIndexMask) | (((n_children - 1) & 7u) << 26);
}
GTH_HD inline uint32_t meta_first_child(uint32_t meta) { return meta & kMetaIndexMask; }
GTH_HD inline uint32_t meta_children(uint32_t meta) {
  return (meta & kMetaLeaf) ? 0u : ((meta >> 26) & 7u) + 1u;
}
constexpr int kUndecided = -1;

// The group's box as the filter needs it: FP32 images relative to the
// origin, the root half size H and the two error bounds.
struct GroupImage {
  float cx, cy, cz, hx, hy, hz;
  float eo, ed;
  float es1, es0;  // the distance test's bound: es1 sum |d| + es0 + ... (accept_filter)
  bool usable;     // H in the filter's range
};

#pragma clang fp contract(off)

// The reference's acceptance test in FP64 (no contraction).
GTH_HD inline bool accept_exact(double cx, double cy, double cz, double mx, double my, double mz,
                                double half, double crit, double gcx, double gcy, double gcz,
                                double ghx, double ghy, double ghz) {
  bool overlap = true;
  double d2 = 0.0;
  {
    if (fabs(cx - gcx) > half * 1.0001 + ghx) overlap = false;
    const double d = fabs(mx - gcx) - ghx;
    if (d > 0) d2 += d * d;
  }
  {
    if (fabs(cy - gcy) > half * 1.0001 + ghy) overlap = false;
    const double d = fabs(my - gcy) - ghy;
    if (d > 0) d2 += d * d;
  }
  {
    if (fabs(cz - gcz) > half * 1.0001 + ghz) overlap = false;
    const double d = fabs(mz - gcz) - ghz;
    if (d > 0) d2 += d * d;
  }
  return !overlap && d2 > crit * crit;
}

GTH_HD inline GroupImage group_image(double ox, double oy, double oz, double H, double gcx,
                                     double gcy, double gcz, double ghx, double ghy, double ghz) {
  GroupImage g;
  g.cx = static_cast<float>(gcx - ox);
  g.cy = static_cast<float>(gcy - oy);
  g.cz = static_cast<float>(gcz - oz);
  g.hx = static_cast<float>(ghx);
  g.hy = static_cast<float>(ghy);
  g.hz = static_cast<float>(ghz);
  const double u = 5.9604644775390625e-08;  // 2^-24
  g.eo = static_cast<float>(16.0 * u * H);
  g.ed = static_cast<float>(8.0 * u * H);
  // 1.5 (2 ed) and 1.5 (6 ed^2), a thousandth larger (the bound is far from
  // tight; the margin covers their own rounding)
  g.es1 = static_cast<float>(1.001 * 3.0 * 8.0 * u * H);
  g.es0 = static_cast<float>(1.001 * 9.0 * (8.0 * u * H) * (8.0 * u * H));
  const double lim = 1.01 * H;
  g.usable = H > 9.094947017729282e-13 && H < 1099511627776.0 &&  // 2^-40, 2^40
             fabs(gcx - ox) <= lim && fabs(gcy - oy) <= lim && fabs(gcz - oz) <= lim &&
             ghx <= lim && ghy <= lim && ghz <= lim;
  return g;
}

// The image of a node (centre c, centre of mass m, opening distance crit,
// meta, "at most two bodies") for origin o and root half H.  A node outside
// the root cube gets a NaN opening distance: the filter leaves it to the
// exact test.
GTH_HD inline WalkRec32 make_rec32(double cx, double cy, double cz, double mx, double my,
                                   double mz, double crit, double ox, double oy, double oz,
                                   double H, uint32_t meta, bool small) {
  WalkRec32 r;
  r.cx = static_cast<float>(cx - ox);
  r.cy = static_cast<float>(cy - oy);
  r.cz = static_cast<float>(cz - oz);
  const double lim = 1.01 * H;
  const bool inside = fabs(cx - ox) <= lim && fabs(cy - oy) <= lim && fabs(cz - oz) <= lim &&
                      fabs(mx - ox) <= lim && fabs(my - oy) <= lim && fabs(mz - oz) <= lim;
  const float c = inside ? static_cast<float>(crit) : __builtin_nanf("");
  r.crit = small ? -c : c;
  r.mx = static_cast<float>(mx - ox);
  r.my = static_cast<float>(my - oy);
  r.mz = static_cast<float>(mz - oz);
  r.meta = meta;
  return r;
}

// The half size image of a node at depth d: the root half's image times 2^-d
// (exact while normal).
GTH_HD inline float half_image(float root_half32, int depth) { return ldexpf(root_half32, -depth); }

// 1 accept, 0 no, kUndecided: the exact test decides.  half32 is the node's
// half size image (half_image).
GTH_HD inline int accept_filter(const WalkRec32 &r, float half32, const GroupImage &g) {
  const float hb = half32 * 1.0001f;
  const float ax = fabsf(r.cx - g.cx) - (hb + g.hx);
  const float ay = fabsf(r.cy - g.cy) - (hb + g.hy);
  const float az = fabsf(r.cz - g.cz) - (hb + g.hz);
  // overlap fails on an axis where a > 0: sure when a > eo on some axis,
  // sure not when a < -eo on every axis (both read off the largest a)
  const float amax = fmaxf(fmaxf(ax, ay), az);
  if (amax < -g.eo) return 0;
  const bool apart = amax > g.eo;
  const float dx = fabsf(r.mx - g.cx) - g.hx;
  const float dy = fabsf(r.my - g.cy) - g.hy;
  const float dz = fabsf(r.mz - g.cz) - g.hz;
  // the squares of the positive parts (FMAs: one rounding each, which keeps
  // within the error bound below)
  const float px = fmaxf(dx, 0.0f), py = fmaxf(dy, 0.0f), pz = fmaxf(dz, 0.0f);
  const float d2 = fmaf(pz, pz, fmaf(py, py, px * px));
  const float crit = fabsf(r.crit);
  const float c2 = crit * crit;
  if (!(c2 > 1e-30f)) return kUndecided;
  const float s = d2 - c2;
  const float sum = fabsf(dx) + fabsf(dy) + fabsf(dz);
  // 1.5 (2 ed sum + 6 ed^2 + 2^-22 (d2 + c2)), with the per-group parts
  // folded into es1 and es0
  const float es = fmaf(g.es1, sum, fmaf(3.5762786865234375e-07f, d2 + c2, g.es0));
  if (s < -es) return 0;
  if (apart && s > es) return 1;
  return kUndecided;
}

#pragma clang fp contract(fast)

}  // namespace gth
