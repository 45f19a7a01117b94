// C++ face of libgt_hip.so for the reference's host code: the role
// metal::Context plays (include/galactic_tides/metal_backend.h:25-54), over the
// C ABI in gt_hip.h.  Failures become the reference's exception types:
// std::invalid_argument for bad arguments, std::runtime_error otherwise.
#pragma once

#include <cstddef>
#include <cstdint>
#include <stdexcept>
#include <string>
#include <vector>

#include "gt_hip.h"

namespace gt {
namespace hip {

[[noreturn]] inline void raise(int rc, const char *msg) {
  const std::string m = (msg && *msg) ? msg : "hip: error " + std::to_string(rc);
  if (rc == GT_HIP_ERR_ARGUMENT) throw std::invalid_argument(m);
  throw std::runtime_error(m);
}

// GravityParams + GroupTreeGravity's arguments as the ABI's config.
inline gt_group_config config(double G, double theta, double eps, bool quadrupole,
                              int group_size, int leaf_size, bool pack, bool exact = false) {
  gt_group_config c{};
  c.G = G;
  c.theta = theta;
  c.eps = eps;
  c.quadrupole = quadrupole ? 1 : 0;
  c.group_size = group_size;
  c.leaf_size = leaf_size;
  c.pack = pack ? 1 : 0;
  c.exact = exact ? 1 : 0;
  return c;
}

// The same from a GravityParams (include/galactic_tides/octree.h:13-18: G,
// theta, eps, quadrupole) and GroupTreeGravity's own arguments
// (gravity_backends.h:62-63), as GroupTreeGravity::compute has them.
template <class GravityParams>
inline gt_group_config config(const GravityParams &gp, int group_size, int leaf_size, bool pack,
                              bool exact = false) {
  return config(gp.G, gp.theta, gp.eps, gp.quadrupole, group_size, leaf_size, pack, exact);
}

class Context {
 public:
  explicit Context(int device = 0) {
    const int rc = gt_hip_create(&c_, device);
    if (rc != GT_HIP_OK) raise(rc, gt_hip_last_error(nullptr));
  }
  ~Context() { gt_hip_destroy(c_); }
  Context(const Context &) = delete;
  Context &operator=(const Context &) = delete;

  std::string device_name() const {
    char buf[256];
    check(gt_hip_device_name(c_, buf, sizeof(buf)));
    return buf;
  }
  double last_gpu_seconds() const { return gt_hip_last_gpu_seconds(c_); }

  // metal::Context::direct (metal_backend.h:35-36)
  void direct(const gt_float4 *bodies, const gt_params &p, gt_float4 *out) {
    check(gt_hip_direct(c_, bodies, &p, out));
  }
  // metal::Context::tree (metal_backend.h:41-44)
  void tree(const gt_float4 *bodies, size_t n_cells, const gt_cell *cells, size_t n_cell_list,
            const unsigned *cell_list, size_t n_body_list, const unsigned *body_list,
            const gt_group *groups, const gt_params &p, gt_float4 *out, bool exact = false) {
    check(gt_hip_tree(c_, bodies, n_cells, cells, n_cell_list, cell_list, n_body_list, body_list,
                      groups, &p, exact ? 1 : 0, out));
  }
  // GroupTreeGravity::compute, all on the GPU
  void group_compute(const gt_group_config &cfg, size_t n, const double *pos, const double *mass,
                     double *acc, double *pot) {
    check(gt_hip_group_compute(c_, &cfg, n, pos, mass, acc, pot));
  }
  gt_lists_info info() const {
    gt_lists_info i{};
    check(gt_hip_last_info(c_, &i));
    return i;
  }
  gt_timings timings() const {
    gt_timings t{};
    check(gt_hip_last_timings(c_, &t));
    return t;
  }
  // Octree::build + Octree::for_each_within (include/galactic_tides/octree.h:73-75)
  // for a batch of queries: CSR lists in the reference's callback order.
  struct Neighbours {
    std::vector<uint64_t> offsets, index;
    std::vector<double> dist2;
  };
  Neighbours neighbours(int leaf_size, size_t n, const double *pos, const double *mass,
                        size_t nq, const double *centres, const double *radii,
                        size_t n_radii) {
    Neighbours r;
    r.offsets.resize(nq + 1);
    uint64_t total = 0;
    check(gt_hip_neighbours(c_, leaf_size, n, pos, mass, nq, centres, radii, n_radii, 0,
                            r.offsets.data(), nullptr, nullptr, &total));
    r.index.resize(total);
    r.dist2.resize(total);
    check(gt_hip_neighbours(c_, leaf_size, n, pos, mass, nq, centres, radii, n_radii, total,
                            r.offsets.data(), r.index.data(), r.dist2.data(), &total));
    return r;
  }
  // f(query, index, distance_squared) for every neighbour, queries in order and
  // each query's neighbours in the order Octree::for_each_within calls f.
  template <class F>
  void for_each_within(int leaf_size, size_t n, const double *pos, const double *mass,
                       size_t nq, const double *centres, const double *radii, size_t n_radii,
                       F &&f) {
    const Neighbours r = neighbours(leaf_size, n, pos, mass, nq, centres, radii, n_radii);
    for (size_t q = 0; q < nq; ++q)
      for (uint64_t k = r.offsets[q]; k < r.offsets[q + 1]; ++k) f(q, r.index[k], r.dist2[k]);
  }
  // resident leapfrog
  void state_set(size_t n, const double *pos, const double *vel, const double *mass) {
    check(gt_hip_state_set(c_, n, pos, vel, mass));
  }
  void state_get(double *pos, double *vel, double *acc, double *pot) {
    check(gt_hip_state_get(c_, pos, vel, acc, pot));
  }
  void prime(const gt_group_config &cfg) { check(gt_hip_prime(c_, &cfg)); }
  void steps(const gt_group_config &cfg, int n, double dt) {
    check(gt_hip_steps(c_, &cfg, n, dt, 0));
  }

 private:
  void check(int rc) const {
    if (rc != GT_HIP_OK) raise(rc, gt_hip_last_error(c_));
  }
  gt_hip_context *c_ = nullptr;
};

}  // namespace hip
}  // namespace gt
