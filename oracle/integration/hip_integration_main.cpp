// Driver of the reference built with the INTEGRATION.md edits
// (oracle/_ref/gt_hip_integration): the reference's own make_gravity,
// GroupTreeGravity, DirectGravity and LeapfrogKDK with device = hip, next to
// the same solvers on its cpu device.  Test infrastructure only (built by
// `make -C oracle integration`, run by tests/test_integration_glue.py and
// tests/test_gpu_integration.py).  Prints one JSON object per line.
//
//   gt_hip_integration errors   the argument and device errors (no GPU needed)
//   gt_hip_integration run N    a galaxy of N particles: group hip vs group cpu
//                               (packed, the same lists), direct hip vs FP64
//                               direct, 10 leapfrog steps hip vs cpu
#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <exception>
#include <memory>
#include <stdexcept>
#include <string>
#include <vector>

#include "galactic_tides/gravity.h"
#include "galactic_tides/gravity_backends.h"
#include "galactic_tides/hernquist_ics.h"
#include "galactic_tides/integrator.h"
#include "galactic_tides/params.h"

using namespace galactic_tides;

namespace {

std::string quoted(const std::string &s) {
  std::string o = "\"";
  for (char ch : s) {
    if (ch == '"' || ch == '\\') o += '\\';
    o += ch;
  }
  return o + "\"";
}

// make_gravity with "key=value" overrides; the error or the summary.
void try_make(const char *name, const std::vector<std::string> &kv) {
  Params params;
  params.apply_overrides(kv);
  GravityParams gp;
  std::string summary, kind = "ok", what;
  try {
    auto model = make_gravity(params, gp, &summary);
    what = summary;
  } catch (const std::invalid_argument &e) {
    kind = "invalid_argument";
    what = e.what();
  } catch (const std::runtime_error &e) {
    kind = "runtime_error";
    what = e.what();
  }
  std::printf("{\"case\": %s, \"result\": %s, \"message\": %s}\n", quoted(name).c_str(),
              quoted(kind).c_str(), quoted(what).c_str());
}

// |a - b| / |b| per particle: median and maximum.
void errors(const std::vector<Vec3> &a, const std::vector<Vec3> &b, double &med, double &mx) {
  std::vector<double> e(a.size());
  for (std::size_t i = 0; i < a.size(); ++i) e[i] = norm(a[i] - b[i]) / norm(b[i]);
  std::sort(e.begin(), e.end());
  med = e.empty() ? 0.0 : e[e.size() / 2];
  mx = e.empty() ? 0.0 : e.back();
}

int run(std::size_t n) {
  HernquistModel m;
  m.disk_n = n * 2 / 5;
  m.bulge_n = n / 10;
  m.halo_n = n - m.disk_n - m.bulge_n;
  m.bulge_mass = 0.333;
  m.bulge_rmax = 2.0;
  m.seed = 7;
  m.gravity.eps = 0.025;
  m.gravity.theta = 0.75;
  const Particles base = make_hernquist_galaxy(m);
  GravityParams gp;
  gp.eps = 0.025;
  gp.theta = 0.75;

  // make_gravity's group solver on hip: its summary and forces, against the
  // same solver on the cpu device with the same (packed) lists
  Params params;
  params.apply_overrides({"gravity=group", "device=hip"});
  std::string summary;
  auto hip = make_gravity(params, gp, &summary);
  Particles ph = base, pc = base;
  hip->compute(ph);
  GroupTreeGravity cpu(gp, Device::CPU, 32, 8, 1);
  cpu.compute(pc);
  double med, mx;
  errors(ph.acc, pc.acc, med, mx);
  const WalkStats sh = hip->stats(), sc = cpu.stats();
  const auto &tm = static_cast<GroupTreeGravity &>(*hip).timings();
  std::printf("{\"case\": \"group\", \"n\": %zu, \"summary\": %s, \"median\": %.3e, \"max\": %.3e, "
              "\"cells_hip\": %llu, \"cells_cpu\": %llu, \"bodies_hip\": %llu, \"bodies_cpu\": "
              "%llu, \"ms_total\": %.3f, \"ms_tree\": %.3f, \"ms_lists\": %.3f, "
              "\"ms_evaluate\": %.3f}\n",
              n, quoted(summary).c_str(), med, mx, (unsigned long long)sh.cell_interactions,
              (unsigned long long)sc.cell_interactions,
              (unsigned long long)sh.particle_interactions,
              (unsigned long long)sc.particle_interactions, tm.gpu * 1e3, tm.tree * 1e3,
              tm.lists * 1e3, tm.evaluate * 1e3);

  // DirectGravity on hip (FP32) against the cpu device (FP64)
  Particles dh = base, dc = base;
  DirectGravity dhip(gp, Device::Hip), dcpu(gp, Device::CPU);
  dhip.compute(dh);
  dcpu.compute(dc);
  errors(dh.acc, dc.acc, med, mx);
  std::printf("{\"case\": \"direct\", \"median\": %.3e, \"max\": %.3e}\n", med, mx);

  // LeapfrogKDK with the hip solver against the cpu solver: 10 steps
  Particles lh = base, lc = base;
  GroupTreeGravity gh(gp, Device::Hip, 32, 8, 1), gc(gp, Device::CPU, 32, 8, 1);
  LeapfrogKDK kh(gh), kc(gc);
  kh.prime(lh);
  kc.prime(lc);
  for (int s = 0; s < 10; ++s) {
    kh.step(lh, 0.01);
    kc.step(lc, 0.01);
  }
  double dx = 0.0, scale = 0.0;
  for (std::size_t i = 0; i < lh.size(); ++i) {
    dx = std::max(dx, norm(lh.pos[i] - lc.pos[i]));
    scale = std::max(scale, norm(lc.pos[i]));
  }
  const Diagnostics d = compute_diagnostics(lh, 0.1);
  std::printf("{\"case\": \"leapfrog\", \"steps\": 10, \"max_position_difference\": %.3e, "
              "\"extent\": %.3e, \"energy\": %.10e}\n",
              dx, scale, d.total());
  return 0;
}

}  // namespace

int main(int argc, char **argv) {
  const std::string mode = argc > 1 ? argv[1] : "errors";
  try {
    if (mode == "errors") {
      try_make("bad_device", {"gravity=group", "device=vulkan"});
      try_make("group_size_65", {"gravity=group", "device=hip", "group_size=65"});
      try_make("tree_on_hip", {"gravity=tree", "device=hip"});
      try_make("group_on_hip", {"gravity=group", "device=hip"});
      try_make("direct_on_hip", {"gravity=direct", "device=hip"});
      try_make("group_on_cpu", {"gravity=group", "device=cpu"});
      return 0;
    }
    if (mode == "run") return run(argc > 2 ? std::strtoul(argv[2], nullptr, 10) : 50000);
  } catch (const std::exception &e) {
    std::printf("{\"case\": \"failure\", \"message\": %s}\n", quoted(e.what()).c_str());
    return 1;
  }
  std::fprintf(stderr, "usage: %s errors | run [N]\n", argv[0]);
  return 2;
}
