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
// Velocities are either the reference's (measured_velocities below,
// src/hernquist_ics.cpp:206-388: moments from the realised potential, which
// a tree solver supplied by the caller measures; every expression and random
// draw in the reference's order, so with its tree they are its velocities bit
// for bit), or a simpler analytic recipe that needs no tree: isotropic
// Gaussians from the spherical Jeans equation in G M(<r) / r^2 for the bulge
// and halo, and circular rotation plus a small dispersion for the disk.
#include "gt_ics.h"

#include <algorithm>
#include <cmath>
#include <functional>
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

// <v^2> / sigma^2 of an isotropic Gaussian truncated at speed x sigma
// (src/hernquist_ics.cpp:67-73).
double truncated_maxwell_moment(double x) {
  const double e = std::exp(-0.5 * x * x);
  const double n = std::sqrt(M_PI / 2) * std::erf(x / std::sqrt(2.0)) - x * e;
  if (n <= 0) return 0.6 * x * x;
  return 3.0 - x * x * x * e / n;
}

// The width whose Gaussian truncated at vmax keeps <v^2> = 3 sigma^2
// (src/hernquist_ics.cpp:78-96): doubling then bisection.
double inflated_sigma(double sigma_jeans, double vmax) {
  if (sigma_jeans <= 0 || vmax <= 0) return sigma_jeans;
  const double target = 3 * sigma_jeans * sigma_jeans;
  auto moment = [&](double s) { return s * s * truncated_maxwell_moment(vmax / s); };
  double hi = sigma_jeans;
  for (int k = 0; k < 60 && moment(hi) < target; ++k) {
    hi *= 2;
    if (hi > 100 * vmax) return hi;
  }
  double lo = sigma_jeans;
  for (int it = 0; it < 60; ++it) {
    const double mid = 0.5 * (lo + hi);
    (moment(mid) < target ? lo : hi) = mid;
  }
  return 0.5 * (lo + hi);
}

// What the velocity stage needs of the galaxy's mass model.
struct MassModel {
  std::function<double(double)> halo_rho, halo_within, bulge_rho, bulge_within, disk_within,
      surface_density;
  bool halo, bulge, disk;
  double h, gam, rc, a;
};

// The reference's velocity stage (src/hernquist_ics.cpp:206-378) for the
// particles first .. P.pos.size() - 1 of one galaxy (before its centre-of-
// mass shift): the realised potential from the caller's tree solver, then the
// spheroid dispersions (Jeans) and the disk's rotation curve, epicyclic
// frequency and dispersions, and the draws in particle order from rng/gauss.
template <class Rng, class Gauss>
bool measured_velocities(const gt_ics_model &M, const MassModel &mm, Particles &P,
                         std::size_t first, Rng &rng, Gauss &gauss, gt_ics_tree_fn tree,
                         void *user) {
  const std::size_t n = P.pos.size() - first;
  std::vector<double> xyz(3 * n), mass(n), pacc(3 * n), ppot(n);
  for (std::size_t k = 0; k < n; ++k) {
    xyz[3 * k] = P.pos[first + k].x;
    xyz[3 * k + 1] = P.pos[first + k].y;
    xyz[3 * k + 2] = P.pos[first + k].z;
    mass[k] = P.mass[first + k];
  }
  // the potential at every particle (compute_gravity_tree)
  if (tree(user, &M, n, xyz.data(), mass.data(), 0, nullptr, pacc.data(), ppot.data()) != 0)
    return false;

  // the points the realised potential is sampled at: spheroid radii over a
  // Fibonacci sphere of directions, the disk midplane, and (Jeans sigma_z)
  // heights above it
  const double rlo = 1e-4 * std::min({mm.gam > 0 ? mm.gam : mm.rc, mm.a, mm.h});
  const double rhi = std::max({M.halo_rmax, M.bulge_rmax, M.disk_rmax});
  const std::vector<double> rg = log_spaced(rlo, rhi, 2000);
  const std::size_t ng = 120, ndir = 64;
  const std::vector<double> rt = log_spaced(10 * rlo, rhi, ng);
  std::vector<V3> dirs(ndir);
  const double golden = M_PI * (3.0 - std::sqrt(5.0));
  for (std::size_t k = 0; k < ndir; ++k) {
    const double z = 1.0 - (2.0 * double(k) + 1.0) / double(ndir);
    const double sz = std::sqrt(1.0 - z * z);
    dirs[k] = V3{sz * std::cos(golden * double(k)), sz * std::sin(golden * double(k)), z};
  }
  const std::size_t nR = 200, nphi = 64, nz = 32, nphi_z = 16;
  const double dR = M.disk_rmax / double(nR);
  std::vector<double> Rg(nR);
  for (std::size_t i = 0; i < nR; ++i) Rg[i] = (double(i) + 1.0) * dR;
  const double zmax = 6.0 * M.disk_z0, dz = zmax / double(nz);
  std::vector<double> pts;
  auto add = [&](double x, double y, double z) {
    pts.push_back(x);
    pts.push_back(y);
    pts.push_back(z);
  };
  for (std::size_t i = 0; i < ng; ++i)
    for (const V3 &d : dirs) add(d.x * rt[i], d.y * rt[i], d.z * rt[i]);
  const std::size_t disk_at = pts.size() / 3;
  if (mm.disk) {
    for (std::size_t i = 0; i < nR; ++i)
      for (std::size_t k = 0; k < nphi; ++k) {
        const double phi = 2 * M_PI * (double(k) + 0.5 * static_cast<double>(i % 2)) / double(nphi);
        add(Rg[i] * std::cos(phi), Rg[i] * std::sin(phi), 0.0);
      }
    if (M.disk_sigma_z_jeans)
      for (std::size_t i = 0; i < nR; ++i)
        for (std::size_t j = 0; j < nz; ++j) {
          const double z = (double(j) + 0.5) * dz;
          for (std::size_t k = 0; k < nphi_z; ++k) {
            const double phi = 2 * M_PI * (double(k) + 0.5) / double(nphi_z);
            add(Rg[i] * std::cos(phi), Rg[i] * std::sin(phi), z);
          }
        }
  }
  const std::size_t npts = pts.size() / 3;
  std::vector<double> acc(3 * npts), pot(npts);
  if (tree(user, &M, n, xyz.data(), mass.data(), npts, pts.data(), acc.data(), pot.data()) != 0)
    return false;

  // ---- spheroids: spherical isotropic Jeans equation
  std::vector<double> menc(rg.size()), gr(rg.size()), sig2_halo(rg.size()), sig2_bulge(rg.size());
  for (std::size_t i = 0; i < rg.size(); ++i)
    menc[i] = mm.halo_within(rg[i]) + mm.bulge_within(rg[i]) + mm.disk_within(rg[i]);
  {
    std::vector<double> gt(ng);
    for (std::size_t i = 0; i < ng; ++i) {
      double sum = 0.0;
      for (std::size_t k = 0; k < ndir; ++k) {
        const double *a = &acc[3 * (i * ndir + k)];
        const V3 &d = dirs[k];
        sum += -(a[0] * d.x + a[1] * d.y + a[2] * d.z);
      }
      gt[i] = std::max(0.0, sum / double(ndir));
    }
    for (std::size_t i = 0; i < rg.size(); ++i) {
      if (rg[i] < rt.front()) gr[i] = gt.front() * rg[i] / rt.front();  // harmonic core
      else if (rg[i] > rt.back()) gr[i] = M.G * menc[i] / (rg[i] * rg[i]);
      else gr[i] = table_interp(rt, gt, rg[i]);
    }
  }
  auto jeans = [&](const std::function<double(double)> &rho, std::vector<double> &sig2) {
    double integral = 0.0;
    auto f = [&](std::size_t i) { return rho(rg[i]) * gr[i]; };
    sig2.back() = 0.0;
    for (std::size_t i = rg.size() - 1; i-- > 0;) {
      integral += 0.5 * (f(i) + f(i + 1)) * (rg[i + 1] - rg[i]);
      const double r = rho(rg[i]);
      sig2[i] = r > 0 ? integral / r : 0.0;
    }
  };
  if (mm.halo) jeans(mm.halo_rho, sig2_halo);
  if (mm.bulge) jeans(mm.bulge_rho, sig2_bulge);
  auto spheroid_velocity = [&](std::size_t i, const std::vector<double> &sig2) {
    const V3 &p = P.pos[i];
    const double r = std::sqrt(p.x * p.x + p.y * p.y + p.z * p.z);
    double sigma = std::sqrt(std::max(0.0, table_interp(rg, sig2, std::max(r, rlo))));
    const double vmax = M.vesc_fraction * std::sqrt(std::max(0.0, -2.0 * ppot[i - first]));
    if (M.preserve_dispersion) sigma = inflated_sigma(sigma, vmax);
    V3 v;
    for (int tries = 0;; ++tries) {
      // the reference's Vec3(gauss, gauss, gauss): g++ draws the arguments
      // last to first
      const double gz = gauss(rng), gy = gauss(rng), gx = gauss(rng);
      v = V3{gx * sigma, gy * sigma, gz * sigma};
      const double sp = std::sqrt(v.x * v.x + v.y * v.y + v.z * v.z);
      if (sp <= vmax) break;
      if (tries > 1000) {  // sigma far above escape: keep the direction, cap the speed
        const double s = vmax / sp;
        v = V3{v.x * s, v.y * s, v.z * s};
        break;
      }
    }
    return v;
  };

  // ---- disk: rotation curve and epicyclic frequency from the tree
  std::vector<double> F(nR), omega2(nR), kappa2(nR), sigz2(nR, 0.0);
  if (mm.disk) {
    for (std::size_t i = 0; i < nR; ++i) {
      double sum = 0.0;
      for (std::size_t k = 0; k < nphi; ++k) {
        const double phi = 2 * M_PI * (double(k) + 0.5 * static_cast<double>(i % 2)) / double(nphi);
        const double *a = &acc[3 * (disk_at + i * nphi + k)];
        sum += -(a[0] * std::cos(phi) + a[1] * std::sin(phi));
      }
      F[i] = sum / double(nphi);
    }
    if (M.disk_sigma_z_jeans) {
      const std::size_t base = disk_at + nR * nphi;
      for (std::size_t i = 0; i < nR; ++i) {
        double integral = 0.0;
        for (std::size_t j = 0; j < nz; ++j) {
          const double z = (double(j) + 0.5) * dz;
          double az = 0.0;
          for (std::size_t k = 0; k < nphi_z; ++k) az += -acc[3 * (base + (i * nz + j) * nphi_z + k) + 2];
          const double sech = 1.0 / std::cosh(z / M.disk_z0);
          integral += sech * sech * (az / double(nphi_z)) * dz;
        }
        sigz2[i] = std::max(0.0, integral);
      }
    }
    for (std::size_t i = 0; i < nR; ++i) {
      double dF;
      if (i == 0) dF = (F[1] - F[0]) / dR;
      else if (i == nR - 1) dF = (F[i] - F[i - 1]) / dR;
      else dF = (F[i + 1] - F[i - 1]) / (2 * dR);
      omega2[i] = std::max(F[i] / Rg[i], 1e-30);
      kappa2[i] = std::clamp(dF + 3 * F[i] / Rg[i], omega2[i], 4 * omega2[i]);
    }
  }
  auto disk_moments = [&](double R, double &vc2, double &kap2, double &sR2, double &sP2,
                          double &sZ2, double &vbar2) {
    const double o2 = table_interp(Rg, omega2, R);
    kap2 = table_interp(Rg, kappa2, R);
    vc2 = o2 * R * R;
    const double sig = mm.surface_density(R);
    sZ2 = M.disk_sigma_z_jeans ? (R > M.disk_rmax ? 0.0 : table_interp(Rg, sigz2, R))
                               : M_PI * M.G * sig * M.disk_z0;
    const double kQ = std::sqrt(table_interp(Rg, kappa2, M.disk_Q_radius));
    const double sRQ = M.disk_Q * 3.36 * M.G * mm.surface_density(M.disk_Q_radius) / kQ;
    sR2 = sRQ * sRQ * std::exp(-(R - M.disk_Q_radius) / mm.h);
    const double ratio = kap2 / (4 * o2);
    sP2 = sR2 * ratio;
    vbar2 = std::max(0.0, vc2 + sR2 * (1 - ratio - 2 * R / mm.h));
  };

  // ---- velocities, in particle order
  for (std::size_t i = first; i < P.pos.size(); ++i) {
    if (P.type[i] == 1) {
      P.vel[i] = spheroid_velocity(i, sig2_halo);
    } else if (P.type[i] == 3) {
      P.vel[i] = spheroid_velocity(i, sig2_bulge);
    } else {
      const V3 &p = P.pos[i];
      const double R = std::hypot(p.x, p.y);
      double vc2, kap2, sR2, sP2, sZ2, vbar2;
      disk_moments(R, vc2, kap2, sR2, sP2, sZ2, vbar2);
      const double vR = std::sqrt(sR2) * gauss(rng);
      const double vphi = std::sqrt(vbar2) + std::sqrt(sP2) * gauss(rng);
      const double vz = std::sqrt(sZ2) * gauss(rng);
      const double c = R > 0 ? p.x / R : 1.0, s = R > 0 ? p.y / R : 0.0;
      P.vel[i] = V3{vR * c - vphi * s, vR * s + vphi * c, vz};
    }
  }
  return true;
}

bool make_galaxy(const gt_ics_model &M, Particles &P, gt_ics_tree_fn tree = nullptr,
                 void *user = nullptr) {
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

  auto disk_within = [&](double r) {
    return disk ? M.disk_mass * exp_disk_fraction(std::min(r, M.disk_rmax) / h) / dnorm : 0.0;
  };
  if (tree) {
    // the reference's velocities, measured in the realised potential
    const double sigma0 = disk ? M.disk_mass / (2 * M_PI * h * h * dnorm) : 0.0;
    MassModel mm;
    mm.halo_rho = [&](double r) { return H.rho(r); };
    mm.halo_within = [&](double r) { return H.within(r); };
    mm.bulge_rho = [&](double r) { return B.rho(r); };
    mm.bulge_within = [&](double r) { return B.within(r); };
    mm.disk_within = disk_within;
    mm.surface_density = [=](double R) {
      return R > M.disk_rmax ? 0.0 : sigma0 * std::exp(-R / h);
    };
    mm.halo = halo;
    mm.bulge = bulge;
    mm.disk = disk;
    mm.h = h;
    mm.gam = gam;
    mm.rc = rc;
    mm.a = a;
    if (!measured_velocities(M, mm, P, first, rng, gauss, tree, user)) return false;
  } else {
  // Velocities (analytic recipe, see the file comment).
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
  return true;
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
  // HernquistModel (include/galactic_tides/hernquist_ics.h:74-78) and its
  // GravityParams (include/galactic_tides/octree.h:13-18)
  m->vesc_fraction = 0.95;
  m->preserve_dispersion = 0;
  m->disk_sigma_z_jeans = 0;
  m->eps = 0.05;
  m->theta = 0.6;
  m->quadrupole = 1;
}

uint64_t gt_ics_galaxy(const gt_ics_model *m, double incl, double omega, double *pos,
                       double *vel, double *mass, uint8_t *type) {
  Particles P;
  make_galaxy(*m, P);
  rotate(P, 0, P.pos.size(), incl, omega);
  return emit(P, pos, vel, mass, type);
}

uint64_t gt_ics_galaxy_measured(const gt_ics_model *m, double incl, double omega,
                                gt_ics_tree_fn tree, void *user, double *pos, double *vel,
                                double *mass, uint8_t *type) {
  if (!tree) return 0;
  Particles P;
  if (!make_galaxy(*m, P, tree, user)) return 0;
  rotate(P, 0, P.pos.size(), incl, omega);
  return emit(P, pos, vel, mass, type);
}

namespace {

// Encounter mode (apps/galactic_tides_ics.cpp:78-97) with either velocity stage.
uint64_t encounter(const gt_ics_model *m, double g1_incl, double g1_omega, double g2_incl,
                   double g2_omega, double rp, double r0, gt_ics_tree_fn tree, void *user,
                   double *pos, double *vel, double *mass, uint8_t *type) {
  if (r0 < rp) return 0;
  Particles P;
  if (!make_galaxy(*m, P, tree, user)) return 0;
  const std::size_t n1 = P.pos.size();
  gt_ics_model m2 = *m;
  m2.seed = m->seed + 1;
  if (!make_galaxy(m2, P, tree, user)) return 0;
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

}  // namespace

uint64_t gt_ics_encounter(const gt_ics_model *m, double g1_incl, double g1_omega,
                          double g2_incl, double g2_omega, double rp, double r0, double *pos,
                          double *vel, double *mass, uint8_t *type) {
  return encounter(m, g1_incl, g1_omega, g2_incl, g2_omega, rp, r0, nullptr, nullptr, pos, vel,
                   mass, type);
}

uint64_t gt_ics_encounter_measured(const gt_ics_model *m, double g1_incl, double g1_omega,
                                   double g2_incl, double g2_omega, double rp, double r0,
                                   gt_ics_tree_fn tree, void *user, double *pos, double *vel,
                                   double *mass, uint8_t *type) {
  if (!tree) return 0;
  return encounter(m, g1_incl, g1_omega, g2_incl, g2_omega, rp, r0, tree, user, pos, vel, mass,
                   type);
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
