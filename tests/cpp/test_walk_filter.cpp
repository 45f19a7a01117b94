// The walk's FP32 opening-test filter against the exact FP64 test: wherever
// the filter decides, it must decide as the exact test does.  Random nodes and
// groups at depths 0-40 in root cubes from 2^-39 to 2^39, plus cases placed on
// the two thresholds (overlap edge and opening distance) to within a few
// ulps.  Prints the undecided fraction; exits non-zero on any disagreement.
#include <cstdio>
#include <cstdint>
#include <cmath>
#include <random>

#include "gt_walk_rec.h"

using namespace gth;

int main() {
  std::mt19937_64 rng(12345);
  std::uniform_real_distribution<double> U(-1.0, 1.0), P(0.0, 1.0);
  const double Hs[] = {std::ldexp(1.5, -39), 1e-3, 0.7, 13.3, 97.0, 1e3, std::ldexp(1.25, 38)};
  long decided = 0, undecided = 0, wrong = 0, accepts = 0;
  long edge_cases = 0, edge_undecided = 0;
  for (double H : Hs) {
    for (int trial = 0; trial < 400000; ++trial) {
      const double O[3] = {U(rng) * 10 * H, U(rng) * 10 * H, U(rng) * 10 * H};
      const int depth = static_cast<int>(P(rng) * (trial % 8 == 0 ? 41 : 22));
      const double half = std::ldexp(H, -depth);
      const double theta = 0.3 + 0.9 * P(rng);
      double c[3], m[3], gc[3], gh[3];
      for (int a = 0; a < 3; ++a) {
        c[a] = O[a] + U(rng) * (H - half);
        m[a] = c[a] + U(rng) * half;
      }
      const double delta = std::sqrt((m[0] - c[0]) * (m[0] - c[0]) + (m[1] - c[1]) * (m[1] - c[1]) +
                                     (m[2] - c[2]) * (m[2] - c[2]));
      const double crit = 2.0 * half / theta + delta;
      // group: near the node most of the time (where the walk tests it)
      const double reach = (trial % 3 == 0 ? 1.0 : 6.0) * crit;
      const double gscale = half * std::ldexp(1.0, static_cast<int>(P(rng) * 6) - 3);
      for (int a = 0; a < 3; ++a) {
        gh[a] = P(rng) * gscale;
        gc[a] = m[a] + U(rng) * reach;
        // inside the root cube
        const double lim = H - gh[a];
        gc[a] = std::fmin(std::fmax(gc[a], O[a] - lim), O[a] + lim);
      }
      const int mode = trial % 5;
      if (mode == 1) {
        // on the overlap edge of axis 0: |c - gc| = half * 1.0001 + gh, a few ulps off
        const double t = half * 1.0001 + gh[0];
        const double side = U(rng) < 0 ? -1.0 : 1.0;
        gc[0] = c[0] + side * t;
        for (int k = static_cast<int>(U(rng) * 4); k != 0; k += (k < 0 ? 1 : -1))
          gc[0] = std::nextafter(gc[0], k > 0 ? INFINITY : -INFINITY);
        ++edge_cases;
      } else if (mode == 2) {
        // on the opening distance: the box gc +- gh at distance crit from m
        double dir[3], n = 0;
        for (int a = 0; a < 3; ++a) {
          dir[a] = U(rng);
          n += dir[a] * dir[a];
        }
        n = std::sqrt(n);
        for (int a = 0; a < 3; ++a) {
          const double d = crit * std::fabs(dir[a]) / n;
          gc[a] = m[a] + (dir[a] < 0 ? -1 : 1) * (d + gh[a]);
          for (int k = static_cast<int>(U(rng) * 3); k != 0; k += (k < 0 ? 1 : -1))
            gc[a] = std::nextafter(gc[a], k > 0 ? INFINITY : -INFINITY);
        }
        ++edge_cases;
      }
      const bool want = accept_exact(c[0], c[1], c[2], m[0], m[1], m[2], half, crit, gc[0], gc[1],
                                     gc[2], gh[0], gh[1], gh[2]);
      const GroupImage g = group_image(O[0], O[1], O[2], H, gc[0], gc[1], gc[2], gh[0], gh[1],
                                       gh[2]);
      const WalkRec32 r = make_rec32(c[0], c[1], c[2], m[0], m[1], m[2], crit, O[0], O[1], O[2], H,
                                     image_meta(0, 0), false);
      const int got = g.usable ? accept_filter(r, half_image(static_cast<float>(H), depth), g)
                               : kUndecided;
      if (got == kUndecided) {
        ++undecided;
        if (mode == 1 || mode == 2) ++edge_undecided;
        continue;
      }
      ++decided;
      accepts += got;
      if ((got == 1) != want) {
        if (wrong < 10)
          std::printf("MISMATCH H=%g depth=%d mode=%d want=%d got=%d\n", H, depth, mode,
                      static_cast<int>(want), got);
        ++wrong;
      }
    }
  }
  std::printf("decided %ld (accepted %ld), undecided %ld (%.4f%%), threshold cases %ld of which "
              "undecided %ld, mismatches %ld\n",
              decided, accepts, undecided, 100.0 * undecided / (decided + undecided), edge_cases,
              edge_undecided, wrong);
  return wrong == 0 && decided > 0 ? 0 : 1;
}
