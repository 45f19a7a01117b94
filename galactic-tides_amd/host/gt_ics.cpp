// Initial conditions for the benchmark configurations: Hernquist (1993)
// disk + bulge + halo galaxies and two-galaxy parabolic encounters.
//
// Positions follow the reference's sampler draw for draw (same std::mt19937_64
// stream, same tables and interpolation, same operation order:
// src/hernquist_ics.cpp:16-64,142-204,380-388,421-460 and
// apps/galactic_tides_ics.cpp:78-97), so a galaxy built here has exactly the
// reference's particle positions and the octree, groups and interaction lists
// built from it can be compared with the reference's entry for entry.
//
// Velocities use a simpler, analytic recipe (no tree): isotropic Gaussians
// from the spherical Jeans equation in G M(<r) / r^2 for the bulge and halo,
// and circular rotation plus a small dispersion for the disk.  They are not
// the reference's tree-measured moments; only the dynamics of a run depend on
// them, not a force evaluation.
#include "gt_ics.h"

#include <algorithm>
#include <cmath>
#include <random>
#include <vector>

namespace {

struct V3 {
  double x = 0, y = 0, z = 0;
};

double table_interp(const std::vector<double> &x, const std::vector<double> &y, double xv) {
  if (xv <= x.front()) return y.front();
  if (xv >= x.back()) return y.back();
  const std::size_t j = static_cast<std::size_t>(std::upper_bound(x.begin(), x.end(), xv) - x.begin());
  const double t = (xv - x[j - 1]) / (x[j] - x[j - 1]);
  return y[j - 1] + t * (y[j] - y[j - 1]);
}

std::vector<double> log_spaced(double lo, double hi, std::size_t n) {
  std::vector<double> g(n);
  const double a = std::log(lo), b = std::log(hi);
  for (std::size_t i = 0; i < n; ++i) g[i] = std::exp(a + (b - a) * double(i) / double(n - 1));
  return g;
}

// Truncated spherical profile with a trapezoid-rule cumulative mass table.
template <class Shape>
struct Sphere {
  Shape shape;
  double rmax, mass, norm = 1.0;
  std::vector<double> r, m;
  Sphere(Shape s, double rmax_, double mass_, double rmin) : shape(s), rmax(rmax_), mass(mass_) {
    r.push_back(0.0);
    for (double v : log_spaced(rmin, rmax, 4000)) r.push_back(v);
    m.assign(r.size(), 0.0);
    auto f = [&](double x) { return x > 0 ? 4 * M_PI * x * x * shape(x) : 0.0; };
    for (std::size_t i = 1; i < r.size(); ++i)
      m[i] = m[i - 1] + 0.5 * (f(r[i - 1]) + f(r[i])) * (r[i] - r[i - 1]);
    norm = mass / m.back();
    for (auto &v : m) v *= norm;
  }
  double rho(double x) const { return (x > rmax || x <= 0) ? 0.0 : norm * shape(x); }
  double within(double x) const { return x >= rmax ? mass : table_interp(r, m, x); }
  double radius_at(double u) const { return table_interp(m, r, u * mass); }
};

double exp_disk_fraction(double x) { return 1.0 - (1.0 + x) * std::exp(-x); }

struct Particles {
  std::vector<V3> pos, vel;
  std::vector<double> mass;
  std::vector<uint8_t> type;
  void add(const V3 &p, double m, uint8_t t) {
    pos.push_back(p);
    vel.push_back(V3{});
    mass.push_back(m);
    type.push_back(t);
  }
};

void make_galaxy(const gt_ics_model &M, Particles &P) {
  std::mt19937_64 rng(M.seed);
  std::uniform_real_distribution<double> uni(0.0, 1.0);
  std::normal_distribution<double> gauss(0.0, 1.0);
  auto u_open = [&]() {
    double u;
    do u = uni(rng);
    while (u <= 0.0 || u >= 1.0);
    return u;
  };
  auto direction = [&]() {
    const double cz = 2 * uni(rng) - 1, ph = 2 * M_PI * uni(rng), s = std::sqrt(1 - cz * cz);
    return V3{s * std::cos(ph), s * std::sin(ph), cz};
  };
  const bool disk = M.disk_n > 0 && M.disk_mass > 0;
  const bool bulge = M.bulge_n > 0 && M.bulge_mass > 0;
  const bool halo = M.halo_n > 0 && M.halo_mass > 0;
  const double h = M.disk_h, gam = M.halo_core, rc = M.halo_cutoff, a = M.bulge_a;
  const bool hern_halo = M.halo_profile == 1;
  const double ha = M.halo_a;
  auto halo_shape = [=](double r) {
    return hern_halo ? ha / (r * std::pow(r + ha, 3))
                     : std::exp(-r * r / (rc * rc)) / (r * r + gam * gam);
  };
  const double halo_rmin =
      hern_halo ? 1e-4 * ha : 1e-4 * (gam > 0 ? std::min(gam, rc) : rc);
  Sphere<decltype(halo_shape)> H(halo_shape, M.halo_rmax, halo ? M.halo_mass : 0.0, halo_rmin);
  auto bulge_shape = [=](double r) { return a / (r * std::pow(r + a, 3)); };
  Sphere<decltype(bulge_shape)> B(bulge_shape, M.bulge_rmax, bulge ? M.bulge_mass : 0.0, 1e-4 * a);
  const double xmax = M.disk_rmax / h, dnorm = exp_disk_fraction(xmax);

  const std::size_t first = P.pos.size();
  // Radius * direction: the reference's expression evaluates the direction
  // (the right operand of operator*) before the radius when built with g++.
  auto spheroid_point = [&](auto &comp) {
    const V3 d = direction();
    const double r = comp.radius_at(u_open());
    return V3{d.x * r, d.y * r, d.z * r};
  };
  if (halo)
    for (uint64_t i = 0; i < M.halo_n; ++i) P.add(spheroid_point(H), M.halo_mass / M.halo_n, 1);
  if (disk)
    for (uint64_t i = 0; i < M.disk_n; ++i) {
      const double target = u_open() * dnorm;
      double lo = 0.0, hi = xmax;
      for (int it = 0; it < 60; ++it) {
        const double mid = 0.5 * (lo + hi);
        (exp_disk_fraction(mid) < target ? lo : hi) = mid;
      }
      const double R = 0.5 * (lo + hi) * h;
      const double phi = 2 * M_PI * uni(rng);
      const double z = M.disk_z0 * std::atanh(2 * u_open() - 1);
      P.add(V3{R * std::cos(phi), R * std::sin(phi), z}, M.disk_mass / M.disk_n, 2);
    }
  if (bulge)
    for (uint64_t i = 0; i < M.bulge_n; ++i) P.add(spheroid_point(B), M.bulge_mass / M.bulge_n, 3);

  // Velocities (analytic recipe, see the file comment).
  auto disk_within = [&](double r) {
    return disk ? M.disk_mass * exp_disk_fraction(std::min(r, M.disk_rmax) / h) / dnorm : 0.0;
  };
  auto menc = [&](double r) { return H.within(r) + B.within(r) + disk_within(r); };
  const double rlo = 1e-4 * std::min({gam > 0 ? gam : rc, a, h});
  const double rhi = std::max({M.halo_rmax, M.bulge_rmax, M.disk_rmax});
  const std::vector<double> rg = log_spaced(rlo, rhi, 2000);
  auto jeans = [&](auto &comp) {
    std::vector<double> s2(rg.size(), 0.0);
    double integral = 0.0;
    auto f = [&](std::size_t i) { return comp.rho(rg[i]) * M.G * menc(rg[i]) / (rg[i] * rg[i]); };
    for (std::size_t i = rg.size() - 1; i-- > 0;) {
      integral += 0.5 * (f(i) + f(i + 1)) * (rg[i + 1] - rg[i]);
      const double rho = comp.rho(rg[i]);
      s2[i] = rho > 0 ? integral / rho : 0.0;
    }
    return s2;
  };
  const std::vector<double> s2h = jeans(H), s2b = jeans(B);
  for (std::size_t i = first; i < P.pos.size(); ++i) {
    const V3 &p = P.pos[i];
    if (P.type[i] == 2) {
      const double R = std::hypot(p.x, p.y);
      const double vc = std::sqrt(M.G * menc(std::max(R, rlo)) / std::max(R, rlo));
      const double sig = 0.1 * vc;
      const double vR = sig * gauss(rng), vphi = vc + sig * gauss(rng), vz = 0.5 * sig * gauss(rng);
      const double c = R > 0 ? p.x / R : 1.0, s = R > 0 ? p.y / R : 0.0;
      P.vel[i] = V3{vR * c - vphi * s, vR * s + vphi * c, vz};
    } else {
      const double r = std::sqrt(p.x * p.x + p.y * p.y + p.z * p.z);
      const auto &s2 = P.type[i] == 1 ? s2h : s2b;
      const double sig = std::sqrt(std::max(0.0, table_interp(rg, s2, std::max(r, rlo))));
      const double gx = gauss(rng), gy = gauss(rng), gz = gauss(rng);
      P.vel[i] = V3{sig * gx, sig * gy, sig * gz};
    }
  }

  // Centre of mass frame (src/hernquist_ics.cpp:381-388).
  double mt = 0.0;
  V3 cm, cv;
  for (std::size_t i = first; i < P.pos.size(); ++i) {
    const double m = P.mass[i];
    mt += m;
    cm.x += P.pos[i].x * m;
    cm.y += P.pos[i].y * m;
    cm.z += P.pos[i].z * m;
    cv.x += P.vel[i].x * m;
    cv.y += P.vel[i].y * m;
    cv.z += P.vel[i].z * m;
  }
  const V3 dx{-cm.x / mt, -cm.y / mt, -cm.z / mt}, dv{-cv.x / mt, -cv.y / mt, -cv.z / mt};
  for (std::size_t i = first; i < P.pos.size(); ++i) {
    P.pos[i].x += dx.x;
    P.pos[i].y += dx.y;
    P.pos[i].z += dx.z;
    P.vel[i].x += dv.x;
    P.vel[i].y += dv.y;
    P.vel[i].z += dv.z;
  }
}

// orient (src/hernquist_ics.cpp:421-432): inclination about x, then argument about z.
void rotate(Particles &P, std::size_t first, std::size_t last, double incl_deg, double omega_deg) {
  const double i = incl_deg * M_PI / 180.0, w = omega_deg * M_PI / 180.0;
  const double ci = std::cos(i), si = std::sin(i), cw = std::cos(w), sw = std::sin(w);
  auto rot = [&](const V3 &v) {
    const V3 a{v.x, ci * v.y - si * v.z, si * v.y + ci * v.z};
    return V3{cw * a.x - sw * a.y, sw * a.x + cw * a.y, a.z};
  };
  for (std::size_t k = first; k < last; ++k) {
    P.pos[k] = rot(P.pos[k]);
    P.vel[k] = rot(P.vel[k]);
  }
}

void shift(Particles &P, std::size_t first, std::size_t last, const V3 &dx, const V3 &dv) {
  for (std::size_t k = first; k < last; ++k) {
    P.pos[k].x += dx.x;
    P.pos[k].y += dx.y;
    P.pos[k].z += dx.z;
    P.vel[k].x += dv.x;
    P.vel[k].y += dv.y;
    P.vel[k].z += dv.z;
  }
}

uint64_t emit(const Particles &P, double *pos, double *vel, double *mass, uint8_t *type) {
  if (pos)
    for (std::size_t i = 0; i < P.pos.size(); ++i) {
      pos[3 * i] = P.pos[i].x;
      pos[3 * i + 1] = P.pos[i].y;
      pos[3 * i + 2] = P.pos[i].z;
      if (vel) {
        vel[3 * i] = P.vel[i].x;
        vel[3 * i + 1] = P.vel[i].y;
        vel[3 * i + 2] = P.vel[i].z;
      }
      if (mass) mass[i] = P.mass[i];
      if (type) type[i] = P.type[i];
    }
  return P.pos.size();
}

}  // namespace

extern "C" {

void gt_ics_default_model(gt_ics_model *m) {
  *m = gt_ics_model{};
  m->G = 1.0;
  m->disk_n = 20000;
  m->disk_mass = 1.0;
  m->disk_h = 1.0;
  m->disk_z0 = 0.2;
  m->disk_rmax = 10.0;
  m->disk_Q = 1.5;
  m->disk_Q_radius = 2.43;
  m->bulge_n = 5000;
  m->bulge_mass = 1.0 / 3.0;
  m->bulge_a = 0.2;
  m->bulge_rmax = 4.0;
  m->halo_n = 60000;
  m->halo_mass = 5.8;
  m->halo_core = 1.0;
  m->halo_cutoff = 10.0;
  m->halo_rmax = 30.0;
  m->halo_profile = 0;
  m->halo_a = 4.0;
  m->seed = 12345;
}

uint64_t gt_ics_galaxy(const gt_ics_model *m, double incl, double omega, double *pos,
                       double *vel, double *mass, uint8_t *type) {
  Particles P;
  make_galaxy(*m, P);
  rotate(P, 0, P.pos.size(), incl, omega);
  return emit(P, pos, vel, mass, type);
}

uint64_t gt_ics_encounter(const gt_ics_model *m, double g1_incl, double g1_omega,
                          double g2_incl, double g2_omega, double rp, double r0, double *pos,
                          double *vel, double *mass, uint8_t *type) {
  if (r0 < rp) return 0;
  Particles P;
  make_galaxy(*m, P);
  const std::size_t n1 = P.pos.size();
  gt_ics_model m2 = *m;
  m2.seed = m->seed + 1;
  make_galaxy(m2, P);
  rotate(P, 0, n1, g1_incl, g1_omega);
  rotate(P, n1, P.pos.size(), g2_incl, g2_omega);
  const double M1 = m->disk_mass + m->bulge_mass + m->halo_mass, M2 = M1;
  // parabolic_orbit (src/hernquist_ics.cpp:450-460)
  const double Mt = M1 + M2;
  const double f = -std::acos(std::clamp(2 * rp / r0 - 1, -1.0, 1.0));
  const double hmom = std::sqrt(2 * m->G * Mt * rp);
  const double vr = m->G * Mt / hmom * std::sin(f);
  const double vt = hmom / r0;
  const V3 rhat{std::cos(f), std::sin(f), 0}, fhat{-std::sin(f), std::cos(f), 0};
  const V3 dx{r0 * rhat.x, r0 * rhat.y, r0 * rhat.z};
  const V3 dv{vr * rhat.x + vt * fhat.x, vr * rhat.y + vt * fhat.y, vr * rhat.z + vt * fhat.z};
  const double w1 = M2 / (M1 + M2), w2 = M1 / (M1 + M2);
  const V3 d1{-w1 * dx.x, -w1 * dx.y, -w1 * dx.z}, v1{-w1 * dv.x, -w1 * dv.y, -w1 * dv.z};
  const V3 d2{w2 * dx.x, w2 * dx.y, w2 * dx.z}, v2{w2 * dv.x, w2 * dv.y, w2 * dv.z};
  shift(P, 0, n1, d1, v1);
  shift(P, n1, P.pos.size(), d2, v2);
  return emit(P, pos, vel, mass, type);
}

uint64_t gt_ics_bench_sample(uint64_t n, uint64_t seed, uint64_t nsample, uint64_t *out) {
  // apps/galactic_tides_bench.cpp:74-78: the indices 0 .. n-1 shuffled by
  // std::shuffle with a std::mt19937_64 seeded with `seed`, cut to nsample.
  std::mt19937_64 rng(seed);
  std::vector<uint64_t> idx(n);
  for (uint64_t i = 0; i < n; ++i) idx[i] = i;
  std::shuffle(idx.begin(), idx.end(), rng);
  const uint64_t m = std::min(n, nsample);
  if (out) std::copy(idx.begin(), idx.begin() + static_cast<std::ptrdiff_t>(m), out);
  return m;
}

}  // extern "C"
