// The HIP backend through its C++ face, in the style of the reference's
// tests/test_metal.cpp: brute force and grouped tree on the GPU against double
// precision direct summation, fast against exact evaluation, and refusal of
// groups wider than a wave.  Exits 0 when all checks pass.
#include <algorithm>
#include <cmath>
#include <cstdio>
#include <random>
#include <set>
#include <vector>

#include "hip_backend.h"

namespace {

int failures = 0;
#define CHECK(cond)                                                      \
  do {                                                                   \
    if (!(cond)) {                                                       \
      std::fprintf(stderr, "%s:%d: CHECK failed: %s\n", __FILE__, __LINE__, #cond); \
      ++failures;                                                        \
    }                                                                    \
  } while (0)

// Hernquist sphere with two satellites (the reference tests' system).
void make_system(size_t n, unsigned seed, std::vector<double> &pos, std::vector<double> &mass) {
  std::mt19937_64 rng(seed);
  std::uniform_real_distribution<double> u(0.0, 1.0);
  pos.resize(3 * n);
  mass.assign(n, 1.0 / n);
  for (size_t i = 0; i < n; ++i) {
    double c[3] = {0, 0, 0}, a = 1.0;
    if (i % 10 == 0) {
      if (i % 20 == 0) c[0] = 4; else { c[0] = -1; c[1] = 3; c[2] = 2; }
      a = 0.1;
    }
    const double x = std::sqrt(u(rng) * 0.95), r = a * x / (1 - x);
    const double cz = 2 * u(rng) - 1, ph = 2 * M_PI * u(rng), s = std::sqrt(1 - cz * cz);
    pos[3 * i] = c[0] + r * s * std::cos(ph);
    pos[3 * i + 1] = c[1] + r * s * std::sin(ph);
    pos[3 * i + 2] = c[2] + r * cz;
  }
}

void direct64(const std::vector<double> &pos, const std::vector<double> &mass, double eps,
              std::vector<double> &acc, std::vector<double> &pot) {
  const size_t n = mass.size();
  acc.assign(3 * n, 0.0);
  pot.assign(n, 0.0);
  for (size_t i = 0; i < n; ++i)
    for (size_t j = 0; j < n; ++j) {
      if (i == j) continue;
      const double d[3] = {pos[3 * i] - pos[3 * j], pos[3 * i + 1] - pos[3 * j + 1],
                           pos[3 * i + 2] - pos[3 * j + 2]};
      const double inv = 1.0 / std::sqrt(d[0] * d[0] + d[1] * d[1] + d[2] * d[2] + eps * eps);
      for (int a = 0; a < 3; ++a) acc[3 * i + a] -= mass[j] * inv * inv * inv * d[a];
      pot[i] -= mass[j] * inv;
    }
}

struct Errors {
  double median, max;
};

Errors compare(const std::vector<double> &a, const std::vector<double> &b) {
  std::vector<double> e;
  for (size_t i = 0; i < b.size() / 3; ++i) {
    double dn = 0, bn = 0;
    for (int k = 0; k < 3; ++k) {
      dn += (a[3 * i + k] - b[3 * i + k]) * (a[3 * i + k] - b[3 * i + k]);
      bn += b[3 * i + k] * b[3 * i + k];
    }
    e.push_back(std::sqrt(dn / bn));
  }
  std::sort(e.begin(), e.end());
  return {e[e.size() / 2], e.back()};
}

}  // namespace

int main() {
  gt::hip::Context ctx;
  std::printf("device: %s\n", ctx.device_name().c_str());
  const size_t n = 5003;  // odd, so the last workgroup is partly empty
  const double eps = 0.01;
  std::vector<double> pos, mass, ref_acc, ref_pot;
  make_system(n, 5, pos, mass);
  direct64(pos, mass, eps, ref_acc, ref_pot);

  // brute force (test_metal.cpp:78-92)
  {
    std::vector<gt_float4> bodies(n), out(n);
    for (size_t i = 0; i < n; ++i)
      bodies[i] = gt_float4{float(pos[3 * i]), float(pos[3 * i + 1]), float(pos[3 * i + 2]),
                            float(mass[i])};
    gt_params p{};
    p.n = static_cast<unsigned>(n);
    p.G = 1.0f;
    p.eps2 = float(eps * eps);
    ctx.direct(bodies.data(), p, out.data());
    std::vector<double> acc(3 * n);
    double pot_max = 0;
    for (size_t i = 0; i < n; ++i) {
      acc[3 * i] = out[i].x;
      acc[3 * i + 1] = out[i].y;
      acc[3 * i + 2] = out[i].z;
      pot_max = std::max(pot_max, std::fabs(out[i].w / ref_pot[i] - 1));
    }
    const Errors e = compare(acc, ref_acc);
    std::printf("hip direct vs double: median %g max %g potential max %g (%.3f ms)\n", e.median,
                e.max, pot_max, ctx.last_gpu_seconds() * 1e3);
    CHECK(e.median < 1e-5);
    CHECK(e.max < 1e-3);
    CHECK(pot_max < 1e-5);
  }
  // grouped tree (test_metal.cpp:96-119)
  for (int gs : {8, 32, 64}) {
    for (double theta : {0.5, 0.8}) {
      std::vector<double> fast(3 * n), exact(3 * n), pf(n), pe(n);
      ctx.group_compute(gt::hip::config(1.0, theta, eps, true, gs, 8, true, true), n, pos.data(),
                        mass.data(), exact.data(), pe.data());
      ctx.group_compute(gt::hip::config(1.0, theta, eps, true, gs, 8, true, false), n,
                        pos.data(), mass.data(), fast.data(), pf.data());
      const gt_timings t = ctx.timings();
      const Errors same = compare(fast, exact), acc = compare(fast, ref_acc);
      std::printf("hip tree, groups of %d, theta %g: vs exact median %g max %g | vs direct "
                  "median %g max %g (%.3f ms: tree %.3f lists %.3f evaluate %.3f)\n",
                  gs, theta, same.median, same.max, acc.median, acc.max, t.total, t.tree,
                  t.lists, t.evaluate);
      CHECK(same.median < 1e-5);
      CHECK(same.max < 1e-3);
      CHECK(acc.median < 3e-3 * theta * theta);
    }
  }
  // neighbour search (test_octree.cpp:116-128): the same sets as a scan of
  // every particle, for 50 queries at particles with radii 0.05 .. 0.5
  {
    std::mt19937_64 rng(7);
    std::uniform_int_distribution<size_t> pick(0, n - 1);
    std::vector<double> centres, radii;
    for (int trial = 0; trial < 50; ++trial) {
      const size_t i = pick(rng);
      centres.insert(centres.end(), {pos[3 * i], pos[3 * i + 1], pos[3 * i + 2]});
      radii.push_back(0.05 * (1 + trial % 10));
    }
    std::vector<std::set<size_t>> got(50);
    ctx.for_each_within(8, n, pos.data(), mass.data(), 50, centres.data(), radii.data(), 50,
                        [&](size_t q, uint64_t i, double) { got[q].insert(i); });
    int same = 0;
    for (int q = 0; q < 50; ++q) {
      std::set<size_t> want;
      for (size_t i = 0; i < n; ++i) {
        const double dx = pos[3 * i] - centres[3 * q], dy = pos[3 * i + 1] - centres[3 * q + 1],
                     dz = pos[3 * i + 2] - centres[3 * q + 2];
        if (dx * dx + dy * dy + dz * dz <= radii[q] * radii[q]) want.insert(i);
      }
      same += got[q] == want;
    }
    std::printf("hip neighbour search: %d of 50 queries match the scan\n", same);
    CHECK(same == 50);
  }
  // groups wider than a wave are refused (test_metal.cpp:121-130)
  {
    bool threw = false;
    std::vector<double> a(3 * n), p(n);
    try {
      ctx.group_compute(gt::hip::config(1.0, 0.7, eps, true, 65, 8, true), n, pos.data(),
                        mass.data(), a.data(), p.data());
    } catch (const std::invalid_argument &) {
      threw = true;
    }
    CHECK(threw);
  }
  if (failures) {
    std::fprintf(stderr, "%d failure(s)\n", failures);
    return 1;
  }
  std::printf("all checks passed\n");
  return 0;
}
