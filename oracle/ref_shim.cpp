// C interface over the reference library, built from the reference's own
// sources by oracle/Makefile into oracle/_ref/libgt_ref.so.  Test and
// baseline infrastructure only: it lets tests check the CPU oracle
// (oracle/gt_oracle.cpp) against the reference bit for bit, lets
// tests/golden/make_fixtures.py write golden vectors, and lets bench.py time
// the reference CPU solver.  It exports the gto_* interface of
// oracle/gt_oracle.h under the prefix gtref_, plus the ICs generator.

#include <algorithm>
#include <cmath>
#include <chrono>
#include <cstdint>
#include <cstring>
#include <memory>
#include <random>
#include <vector>

#ifdef _OPENMP
#include <omp.h>
#endif

#include "galactic_tides/gravity.h"
#include "galactic_tides/gravity_backends.h"
#include "galactic_tides/group_lists.h"
#include "galactic_tides/hernquist_ics.h"
#include "galactic_tides/integrator.h"
#include "galactic_tides/octree.h"

using namespace galactic_tides;

struct gtref_run {
  Octree tree;
  GroupLists lists;
};

namespace {

void to_vec3(uint64_t n, const double *pos, std::vector<Vec3> &out) {
  out.resize(n);
  for (uint64_t i = 0; i < n; ++i) out[i] = Vec3(pos[3 * i], pos[3 * i + 1], pos[3 * i + 2]);
}

Particles make_particles(uint64_t n, const double *pos, const double *vel, const double *mass) {
  Particles p;
  p.resize(n);
  for (uint64_t i = 0; i < n; ++i) {
    p.pos[i] = Vec3(pos[3 * i], pos[3 * i + 1], pos[3 * i + 2]);
    if (vel) p.vel[i] = Vec3(vel[3 * i], vel[3 * i + 1], vel[3 * i + 2]);
    p.mass[i] = mass[i];
    p.id[i] = i;
    p.type[i] = ParticleType::Halo;
  }
  return p;
}

GravityParams gparams(double G, double theta, double eps, int quad) {
  GravityParams gp;
  gp.G = G;
  gp.theta = theta;
  gp.eps = eps;
  gp.quadrupole = quad != 0;
  return gp;
}

}  // namespace

extern "C" {

gtref_run *gtref_new(void) { return new gtref_run(); }
void gtref_free(gtref_run *r) { delete r; }

int gtref_build(gtref_run *r, uint64_t n, const double *pos, const double *mass, int leaf_size,
                uint64_t parallel_threshold) {
  try {
    std::vector<Vec3> p;
    to_vec3(n, pos, p);
    std::vector<double> m(mass, mass + n);
    r->tree.set_parallel_threshold(parallel_threshold ? parallel_threshold : 20000);
    r->tree.build(p, m, leaf_size);
  } catch (...) {
    return -1;
  }
  return 0;
}

int gtref_lists(gtref_run *r, double theta, int group_size, int pack) {
  try {
    GravityParams gp;
    gp.theta = theta;
    build_group_lists(r->tree, gp, group_size, r->lists, pack != 0);
  } catch (...) {
    return -1;
  }
  return 0;
}

void gtref_info(const gtref_run *r, uint64_t *info) {
  info[0] = r->tree.size();
  info[1] = r->tree.nodes().size();
  info[2] = static_cast<uint64_t>(r->tree.depth());
  info[3] = r->lists.groups.size();
  info[4] = r->lists.cell_list.size();
  info[5] = r->lists.body_list.size();
  info[6] = r->lists.cell_interactions;
  info[7] = r->lists.body_interactions;
  info[8] = r->lists.max_list;
  info[9] = static_cast<uint64_t>(r->lists.max_group);
}

void gtref_nodes(const gtref_run *r, double *geom, int32_t *links) {
  const auto &nodes = r->tree.nodes();
  for (size_t k = 0; k < nodes.size(); ++k) {
    const auto &nd = nodes[k];
    double *g = geom + 16 * k;
    g[0] = nd.center.x;
    g[1] = nd.center.y;
    g[2] = nd.center.z;
    g[3] = nd.half;
    g[4] = nd.com.x;
    g[5] = nd.com.y;
    g[6] = nd.com.z;
    g[7] = nd.mass;
    for (int a = 0; a < 6; ++a) g[8 + a] = nd.quad[a];
    g[14] = nd.delta;
    g[15] = 0.0;
    links[4 * k] = nd.first_child;
    links[4 * k + 1] = nd.n_children;
    links[4 * k + 2] = static_cast<int32_t>(nd.begin);
    links[4 * k + 3] = static_cast<int32_t>(nd.end);
  }
}

void gtref_order(const gtref_run *r, uint64_t *order) {
  const auto &o = r->tree.order();
  for (size_t k = 0; k < o.size(); ++k) order[k] = o[k];
}

void gtref_tree_positions(const gtref_run *r, double *pos, double *mass) {
  const auto &p = r->tree.tree_positions();
  for (size_t k = 0; k < p.size(); ++k) {
    pos[3 * k] = p[k].x;
    pos[3 * k + 1] = p[k].y;
    pos[3 * k + 2] = p[k].z;
    mass[k] = r->tree.tree_masses()[k];
  }
}

void gtref_lists_get(const gtref_run *r, uint32_t *groups, uint32_t *cell_list,
                     uint32_t *body_list, float *bodies, float *cells, double *origin) {
  const GroupLists &L = r->lists;
  if (groups) std::memcpy(groups, L.groups.data(), L.groups.size() * sizeof(kern::GTGroup));
  if (cell_list) std::memcpy(cell_list, L.cell_list.data(), L.cell_list.size() * 4);
  if (body_list) std::memcpy(body_list, L.body_list.data(), L.body_list.size() * 4);
  if (bodies) std::memcpy(bodies, L.bodies.data(), L.bodies.size() * sizeof(kern::float4));
  if (cells) std::memcpy(cells, L.cells.data(), L.cells.size() * sizeof(kern::GTCell));
  if (origin) {
    origin[0] = L.origin.x;
    origin[1] = L.origin.y;
    origin[2] = L.origin.z;
  }
}

void gtref_eval(const gtref_run *r, double G, double eps, int quadrupole, float *out) {
  GravityParams gp = gparams(G, 0.6, eps, quadrupole);
  kern::GTParams kp = make_kernel_params(gp, r->lists.bodies.size());
  kp.n_groups = static_cast<kern::uint>(r->lists.groups.size());
  gpu_vector<kern::float4> o;
  evaluate_group_lists_cpu(r->lists, kp, o);
  std::memcpy(out, o.data(), o.size() * sizeof(kern::float4));
}

void gtref_tree_gravity(gtref_run *r, uint64_t n, const double *pos, const double *mass,
                        double G, double theta, double eps, int quadrupole, int leaf_size,
                        double *acc, double *pot, uint64_t *stats) {
  Particles p = make_particles(n, pos, nullptr, mass);
  const WalkStats s = compute_gravity_tree(p, gparams(G, theta, eps, quadrupole), r->tree,
                                           leaf_size);
  for (uint64_t i = 0; i < n; ++i) {
    acc[3 * i] = p.acc[i].x;
    acc[3 * i + 1] = p.acc[i].y;
    acc[3 * i + 2] = p.acc[i].z;
    pot[i] = p.pot[i];
  }
  if (stats) {
    stats[0] = s.cell_interactions;
    stats[1] = s.particle_interactions;
  }
}

void gtref_accel_at(const gtref_run *r, uint64_t np, const double *points, double G,
                    double theta, double eps, int quadrupole, double *acc, double *pot) {
  const GravityParams gp = gparams(G, theta, eps, quadrupole);
  for (uint64_t k = 0; k < np; ++k) {
    Vec3 a;
    double phi;
    r->tree.accel_pot(Vec3(points[3 * k], points[3 * k + 1], points[3 * k + 2]),
                      Octree::kNoSelf, gp, a, phi);
    acc[3 * k] = a.x;
    acc[3 * k + 1] = a.y;
    acc[3 * k + 2] = a.z;
    pot[k] = phi;
  }
}

void gtref_direct_f64(uint64_t n, const double *pos, const double *mass, double G, double eps,
                      double *acc, double *pot) {
  Particles p = make_particles(n, pos, nullptr, mass);
  compute_gravity_direct(p, gparams(G, 0.6, eps, 1));
  for (uint64_t i = 0; i < n; ++i) {
    acc[3 * i] = p.acc[i].x;
    acc[3 * i + 1] = p.acc[i].y;
    acc[3 * i + 2] = p.acc[i].z;
    pot[i] = p.pot[i];
  }
}

// The bench's FP64 sample reference (apps/galactic_tides_bench.cpp:80-95), which
// lives in the app rather than the library, written with the library's Vec3.
void gtref_direct_f64_sample(uint64_t n, const double *pos, const double *mass, double G,
                             double eps, uint64_t ns, const uint64_t *sample, double *acc) {
  std::vector<Vec3> p;
  to_vec3(n, pos, p);
  const double eps2 = eps * eps;
#pragma omp parallel for schedule(dynamic)
  for (long s = 0; s < static_cast<long>(ns); ++s) {
    const std::size_t i = sample[s];
    Vec3 a;
    for (std::size_t j = 0; j < n; ++j) {
      if (j == i) continue;
      const Vec3 d = p[i] - p[j];
      const double inv = 1.0 / std::sqrt(norm2(d) + eps2);
      a -= (G * mass[j] * inv * inv * inv) * d;
    }
    acc[3 * s] = a.x;
    acc[3 * s + 1] = a.y;
    acc[3 * s + 2] = a.z;
  }
}

void gtref_direct_f32(uint64_t n, const float *bodies, double G, double eps, float *out) {
  GravityParams gp = gparams(G, 0.6, eps, 1);
  const kern::GTParams kp = make_kernel_params(gp, n);
  const kern::float4 *b = reinterpret_cast<const kern::float4 *>(bodies);
  for (uint64_t i = 0; i < n; ++i) {
    const kern::GTAcc a = kern::gt_direct_accel(static_cast<kern::uint>(i), b, kp);
    out[4 * i] = a.ax;
    out[4 * i + 1] = a.ay;
    out[4 * i + 2] = a.az;
    out[4 * i + 3] = a.pot;
  }
}

void gtref_pack_bodies(uint64_t n, const double *pos, const double *mass, const double *origin,
                       float *out) {
  std::vector<Vec3> p;
  to_vec3(n, pos, p);
  std::vector<double> m(mass, mass + n);
  gpu_vector<kern::float4> o;
  pack_bodies(p, m, Vec3(origin[0], origin[1], origin[2]), o);
  std::memcpy(out, o.data(), n * sizeof(kern::float4));
}

int gtref_group_gravity(gtref_run *, uint64_t n, const double *pos, const double *mass, double G,
                        double theta, double eps, int quadrupole, int group_size, int leaf_size,
                        int pack, double *acc, double *pot) {
  try {
    Particles p = make_particles(n, pos, nullptr, mass);
    GroupTreeGravity g(gparams(G, theta, eps, quadrupole), Device::CPU, group_size, leaf_size,
                       pack);
    g.compute(p);
    for (uint64_t i = 0; i < n; ++i) {
      acc[3 * i] = p.acc[i].x;
      acc[3 * i + 1] = p.acc[i].y;
      acc[3 * i + 2] = p.acc[i].z;
      pot[i] = p.pot[i];
    }
  } catch (...) {
    return -1;
  }
  return 0;
}

int gtref_kdk_group(gtref_run *, uint64_t n, double *pos, double *vel, const double *mass,
                    double G, double theta, double eps, int quadrupole, int group_size,
                    int leaf_size, int pack, int nsteps, double dt, double *acc, double *pot) {
  try {
    Particles p = make_particles(n, pos, vel, mass);
    GroupTreeGravity g(gparams(G, theta, eps, quadrupole), Device::CPU, group_size, leaf_size,
                       pack);
    LeapfrogKDK kdk(g);
    kdk.prime(p);
    for (int s = 0; s < nsteps; ++s) kdk.step(p, dt);
    for (uint64_t i = 0; i < n; ++i) {
      for (int a = 0; a < 3; ++a) {
        pos[3 * i + a] = p.pos[i][a];
        vel[3 * i + a] = p.vel[i][a];
        acc[3 * i + a] = p.acc[i][a];
      }
      pot[i] = p.pot[i];
    }
  } catch (...) {
    return -1;
  }
  return 0;
}

int gtref_time_group(gtref_run *, uint64_t n, const double *pos, const double *mass, double G,
                     double theta, double eps, int quadrupole, int group_size, int leaf_size,
                     int pack, int repeat, double *best4) {
  try {
    Particles p = make_particles(n, pos, nullptr, mass);
    GroupTreeGravity g(gparams(G, theta, eps, quadrupole), Device::CPU, group_size, leaf_size,
                       pack);
    g.compute(p);  // warm-up
    best4[0] = 1e300;
    for (int rep = 0; rep < repeat; ++rep) {
      const auto t0 = std::chrono::steady_clock::now();
      g.compute(p);
      const double t =
          std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
      if (t < best4[0]) {
        best4[0] = t;
        best4[1] = g.timings().tree;
        best4[2] = g.timings().lists;
        best4[3] = g.timings().evaluate;
        best4[4] = static_cast<double>(g.stats().cell_interactions);
        best4[5] = static_cast<double>(g.stats().particle_interactions);
      }
    }
  } catch (...) {
    return -1;
  }
  return 0;
}

// Wall time of `repeat` compute_gravity_tree calls after a warm-up (best, seconds).
int gtref_time_tree(uint64_t n, const double *pos, const double *mass, double G, double theta,
                    double eps, int quadrupole, int leaf_size, int repeat, double *best) {
  try {
    Particles p = make_particles(n, pos, nullptr, mass);
    TreeGravity t(gparams(G, theta, eps, quadrupole), leaf_size);
    t.compute(p);
    *best = 1e300;
    for (int rep = 0; rep < repeat; ++rep) {
      const auto t0 = std::chrono::steady_clock::now();
      t.compute(p);
      *best = std::min(*best, std::chrono::duration<double>(std::chrono::steady_clock::now() - t0)
                                  .count());
    }
  } catch (...) {
    return -1;
  }
  return 0;
}

uint64_t gtref_within(const gtref_run *r, uint64_t nq, const double *centres,
                      const double *radii, uint64_t n_radii, uint64_t *offsets, uint64_t cap,
                      uint64_t *index, double *dist2) {
  std::vector<std::vector<std::pair<uint64_t, double>>> found(nq);
  for (uint64_t q = 0; q < nq; ++q) {
    const Vec3 c(centres[3 * q], centres[3 * q + 1], centres[3 * q + 2]);
    r->tree.for_each_within(c, radii[n_radii == 1 ? 0 : q],
                            [&](std::size_t i, const Vec3 &, double dd) {
                              found[q].emplace_back(i, dd);
                            });
  }
  uint64_t total = 0;
  for (uint64_t q = 0; q < nq; ++q) {
    offsets[q] = total;
    total += found[q].size();
  }
  offsets[nq] = total;
  if (index && dist2 && total <= cap) {
    for (uint64_t q = 0; q < nq; ++q)
      for (std::size_t k = 0; k < found[q].size(); ++k) {
        index[offsets[q] + k] = found[q][k].first;
        dist2[offsets[q] + k] = found[q][k].second;
      }
  }
  return total;
}

// The force-error sample of apps/galactic_tides_bench.cpp:74-78 (that file is a
// main(), not part of the library): 0 .. n-1 shuffled by std::shuffle with
// std::mt19937_64(seed), cut to nsample.
uint64_t gtref_bench_sample(uint64_t n, uint64_t seed, uint64_t nsample, uint64_t *out) {
  std::mt19937_64 rng(seed);
  std::vector<std::size_t> sample(n);
  for (std::size_t i = 0; i < sample.size(); ++i) sample[i] = i;
  std::shuffle(sample.begin(), sample.end(), rng);
  sample.resize(std::min<std::size_t>(sample.size(), static_cast<std::size_t>(nsample)));
  for (std::size_t i = 0; i < sample.size(); ++i) out[i] = sample[i];
  return sample.size();
}

int gtref_max_threads(void) {
#ifdef _OPENMP
  return omp_get_max_threads();
#else
  return 1;
#endif
}

// make_hernquist_galaxy (src/hernquist_ics.cpp:142-419) with the model's
// numeric fields in `model` (order: G, disk_n, disk_mass, disk_h, disk_z0,
// disk_rmax, disk_Q, disk_Q_radius, bulge_n, bulge_mass, bulge_a, bulge_rmax,
// halo_n, halo_mass, halo_core, halo_cutoff, halo_rmax, vesc_fraction, seed,
// eps, theta), then orient(incl, omega).  Outputs are n x 3 / n.
uint64_t gtref_galaxy(const double *model, double incl, double omega, double *pos, double *vel,
                      double *mass, uint8_t *type) {
  HernquistModel m;
  m.G = model[0];
  m.disk_n = static_cast<std::size_t>(model[1]);
  m.disk_mass = model[2];
  m.disk_h = model[3];
  m.disk_z0 = model[4];
  m.disk_rmax = model[5];
  m.disk_Q = model[6];
  m.disk_Q_radius = model[7];
  m.bulge_n = static_cast<std::size_t>(model[8]);
  m.bulge_mass = model[9];
  m.bulge_a = model[10];
  m.bulge_rmax = model[11];
  m.halo_n = static_cast<std::size_t>(model[12]);
  m.halo_mass = model[13];
  m.halo_core = model[14];
  m.halo_cutoff = model[15];
  m.halo_rmax = model[16];
  m.vesc_fraction = model[17];
  m.seed = static_cast<std::uint64_t>(model[18]);
  m.gravity.G = m.G;
  m.gravity.eps = model[19];
  m.gravity.theta = model[20];
  Particles p = make_hernquist_galaxy(m);
  orient(p, incl, omega);
  if (pos) {
    for (size_t i = 0; i < p.size(); ++i) {
      for (int a = 0; a < 3; ++a) {
        pos[3 * i + a] = p.pos[i][a];
        vel[3 * i + a] = p.vel[i][a];
      }
      mass[i] = p.mass[i];
      type[i] = static_cast<uint8_t>(p.type[i]);
    }
  }
  return p.size();
}

// apps/galactic_tides_ics.cpp:78-97 encounter mode: galaxy 2 uses seed + 1.
uint64_t gtref_encounter(const double *model, double g1_incl, double g1_omega, double g2_incl,
                         double g2_omega, double rp, double r0, double *pos, double *vel,
                         double *mass, uint8_t *type) {
  std::vector<double> m2(model, model + 21);
  m2[18] = model[18] + 1;
  auto build = [](const double *mm) {
    HernquistModel m;
    m.G = mm[0];
    m.disk_n = static_cast<std::size_t>(mm[1]);
    m.disk_mass = mm[2];
    m.disk_h = mm[3];
    m.disk_z0 = mm[4];
    m.disk_rmax = mm[5];
    m.disk_Q = mm[6];
    m.disk_Q_radius = mm[7];
    m.bulge_n = static_cast<std::size_t>(mm[8]);
    m.bulge_mass = mm[9];
    m.bulge_a = mm[10];
    m.bulge_rmax = mm[11];
    m.halo_n = static_cast<std::size_t>(mm[12]);
    m.halo_mass = mm[13];
    m.halo_core = mm[14];
    m.halo_cutoff = mm[15];
    m.halo_rmax = mm[16];
    m.vesc_fraction = mm[17];
    m.seed = static_cast<std::uint64_t>(mm[18]);
    m.gravity.G = m.G;
    m.gravity.eps = mm[19];
    m.gravity.theta = mm[20];
    return m;
  };
  const HernquistModel a = build(model), b = build(m2.data());
  Particles g1 = make_hernquist_galaxy(a);
  Particles g2 = make_hernquist_galaxy(b, nullptr, g1.size());
  orient(g1, g1_incl, g1_omega);
  orient(g2, g2_incl, g2_omega);
  const double M1 = a.total_mass(), M2 = b.total_mass();
  Vec3 dx, dv;
  parabolic_orbit(a.G, M1, M2, rp, r0, dx, dv);
  shift(g1, -(M2 / (M1 + M2)) * dx, -(M2 / (M1 + M2)) * dv);
  shift(g2, (M1 / (M1 + M2)) * dx, (M1 / (M1 + M2)) * dv);
  Particles all;
  append(all, g1);
  append(all, g2);
  if (pos) {
    for (size_t i = 0; i < all.size(); ++i) {
      for (int k = 0; k < 3; ++k) {
        pos[3 * i + k] = all.pos[i][k];
        vel[3 * i + k] = all.vel[i][k];
      }
      mass[i] = all.mass[i];
      type[i] = static_cast<uint8_t>(all.type[i]);
    }
  }
  return all.size();
}

}  // extern "C"
