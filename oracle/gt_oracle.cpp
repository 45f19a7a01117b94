// CPU oracle: a plain restatement of the reference's grouped Barnes-Hut path
// (tree build, group selection and walk, FP32 list evaluation, FP64 direct
// summation, leapfrog KDK).  TEST INFRASTRUCTURE ONLY: it is the checker that
// tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg compare the
// HIP library against.  The product library never links or calls it.
//
// Every function follows the reference's operation order so that results are
// bit-identical to it (checked against the reference's own sources built by
// oracle/Makefile, see tests/test_oracle_vs_reference.py).  Build with
// -ffp-contract=off: the reference's x86-64 build has no FMA contraction.

#include "gt_oracle.h"

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstring>
#include <functional>
#include <random>
#include <vector>

#ifdef _OPENMP
#include <omp.h>
#endif

namespace {

// ---- tree (src/octree.cpp) ----------------------------------------------------

struct Node {
  double c[3] = {0, 0, 0};
  double half = 0;
  double com[3] = {0, 0, 0};
  double mass = 0;
  double q[6] = {0, 0, 0, 0, 0, 0};
  double delta = 0;
  int32_t first_child = -1;
  int32_t n_children = 0;
  uint32_t begin = 0, end = 0;
};

constexpr int kDepthLimit = 60;  // src/octree.cpp:18

// src/octree.cpp:40-44: small enough, deep enough, or a cube too small to split.
bool leaf_rule(const Node &nd, uint32_t leaf, int depth) {
  const double scale = std::fabs(nd.c[0]) + std::fabs(nd.c[1]) + std::fabs(nd.c[2]) + 1e-300;
  return nd.end - nd.begin <= leaf || depth >= kDepthLimit || nd.half < 1e-12 * scale;
}

struct Tree {
  std::vector<Node> nodes;
  std::vector<uint64_t> idx;  // tree order -> particle
  std::vector<double> tpos;   // tree order, x y z
  std::vector<double> tmass;
  int depth = 0;
};

struct Builder {
  const double *pos;
  const double *mass;
  uint32_t leaf;
  std::vector<uint64_t> &idx;
  std::vector<uint64_t> tmp;
  std::vector<Node> &nodes;

  int octant_of(uint64_t i, const double *c) const {
    const double *p = pos + 3 * i;
    return (p[0] >= c[0] ? 1 : 0) | (p[1] >= c[1] ? 2 : 0) | (p[2] >= c[2] ? 4 : 0);
  }

  // src/octree.cpp:46-72, bodies summed in tree order.
  void leaf_moments(Node &nd) const {
    double m = 0, cx = 0, cy = 0, cz = 0;
    for (uint32_t k = nd.begin; k < nd.end; ++k) {
      const uint64_t i = idx[k];
      const double mi = mass[i];
      m += mi;
      cx += mi * pos[3 * i];
      cy += mi * pos[3 * i + 1];
      cz += mi * pos[3 * i + 2];
    }
    if (m > 0) {
      cx /= m;
      cy /= m;
      cz /= m;
    } else {
      cx = nd.c[0];
      cy = nd.c[1];
      cz = nd.c[2];
    }
    double q[6] = {0, 0, 0, 0, 0, 0};
    for (uint32_t k = nd.begin; k < nd.end; ++k) {
      const uint64_t i = idx[k];
      const double mi = mass[i];
      const double dx = pos[3 * i] - cx, dy = pos[3 * i + 1] - cy, dz = pos[3 * i + 2] - cz;
      const double r2 = dx * dx + dy * dy + dz * dz;
      q[0] += mi * (3 * dx * dx - r2);
      q[1] += mi * (3 * dy * dy - r2);
      q[2] += mi * (3 * dz * dz - r2);
      q[3] += mi * 3 * dx * dy;
      q[4] += mi * 3 * dx * dz;
      q[5] += mi * 3 * dy * dz;
    }
    nd.mass = m;
    nd.com[0] = cx;
    nd.com[1] = cy;
    nd.com[2] = cz;
    for (int k = 0; k < 6; ++k) nd.q[k] = q[k];
    const double ex = cx - nd.c[0], ey = cy - nd.c[1], ez = cz - nd.c[2];
    nd.delta = std::sqrt(ex * ex + ey * ey + ez * ez);
  }

  // src/octree.cpp:74-102, children in stored order (parallel axis theorem).
  static void combine(std::vector<Node> &nodes, size_t ni) {
    Node &nd = nodes[ni];
    const size_t a = static_cast<size_t>(nd.first_child), b = a + nd.n_children;
    double m = 0, cx = 0, cy = 0, cz = 0;
    for (size_t k = a; k < b; ++k) {
      const Node &ch = nodes[k];
      m += ch.mass;
      cx += ch.mass * ch.com[0];
      cy += ch.mass * ch.com[1];
      cz += ch.mass * ch.com[2];
    }
    if (m > 0) {
      cx /= m;
      cy /= m;
      cz /= m;
    } else {
      cx = nd.c[0];
      cy = nd.c[1];
      cz = nd.c[2];
    }
    double q[6] = {0, 0, 0, 0, 0, 0};
    for (size_t k = a; k < b; ++k) {
      const Node &ch = nodes[k];
      const double dx = ch.com[0] - cx, dy = ch.com[1] - cy, dz = ch.com[2] - cz;
      const double r2 = dx * dx + dy * dy + dz * dz;
      q[0] += ch.q[0] + ch.mass * (3 * dx * dx - r2);
      q[1] += ch.q[1] + ch.mass * (3 * dy * dy - r2);
      q[2] += ch.q[2] + ch.mass * (3 * dz * dz - r2);
      q[3] += ch.q[3] + ch.mass * 3 * dx * dy;
      q[4] += ch.q[4] + ch.mass * 3 * dx * dz;
      q[5] += ch.q[5] + ch.mass * 3 * dy * dz;
    }
    nd.mass = m;
    nd.com[0] = cx;
    nd.com[1] = cy;
    nd.com[2] = cz;
    for (int k = 0; k < 6; ++k) nd.q[k] = q[k];
    const double ex = cx - nd.c[0], ey = cy - nd.c[1], ez = cz - nd.c[2];
    nd.delta = std::sqrt(ex * ex + ey * ey + ez * ez);
  }

  // src/octree.cpp:104-140: stable counting sort into octants; the non-empty
  // octants become children, appended to the node array in octant order.
  void split(size_t ni) {
    const uint32_t b = nodes[ni].begin, e = nodes[ni].end;
    double c[3] = {nodes[ni].c[0], nodes[ni].c[1], nodes[ni].c[2]};
    const double h = nodes[ni].half;
    uint32_t cnt[8] = {0, 0, 0, 0, 0, 0, 0, 0};
    for (uint32_t k = b; k < e; ++k) ++cnt[octant_of(idx[k], c)];
    uint32_t start[9];
    start[0] = b;
    for (int o = 0; o < 8; ++o) start[o + 1] = start[o] + cnt[o];
    uint32_t put[8];
    for (int o = 0; o < 8; ++o) put[o] = start[o];
    for (uint32_t k = b; k < e; ++k) tmp[put[octant_of(idx[k], c)]++] = idx[k];
    for (uint32_t k = b; k < e; ++k) idx[k] = tmp[k];
    int kids = 0;
    for (int o = 0; o < 8; ++o) kids += cnt[o] != 0;
    const size_t first = nodes.size();
    nodes.resize(first + kids);
    nodes[ni].first_child = static_cast<int32_t>(first);
    nodes[ni].n_children = kids;
    const double h2 = 0.5 * h;
    size_t at = first;
    for (int o = 0; o < 8; ++o) {
      if (!cnt[o]) continue;
      Node &ch = nodes[at++];
      ch.half = h2;
      ch.c[0] = c[0] + ((o & 1) ? h2 : -h2);
      ch.c[1] = c[1] + ((o & 2) ? h2 : -h2);
      ch.c[2] = c[2] + ((o & 4) ? h2 : -h2);
      ch.begin = start[o];
      ch.end = start[o + 1];
    }
  }

  int build(size_t ni, int depth) {
    if (leaf_rule(nodes[ni], leaf, depth)) {
      leaf_moments(nodes[ni]);
      return depth;
    }
    split(ni);
    const size_t first = static_cast<size_t>(nodes[ni].first_child);
    const int kids = nodes[ni].n_children;
    int deepest = depth;
    for (int k = 0; k < kids; ++k) deepest = std::max(deepest, build(first + k, depth + 1));
    combine(nodes, ni);
    return deepest;
  }
};

// src/octree.cpp:169-208 root cube, then the serial recursive build.
void build_tree(Tree &t, uint64_t n, const double *pos, const double *mass, int leaf_size) {
  t.nodes.clear();
  t.idx.resize(n);
  t.tpos.resize(3 * n);
  t.tmass.resize(n);
  t.depth = 0;
  if (n == 0) return;
  double lo[3] = {pos[0], pos[1], pos[2]}, hi[3] = {pos[0], pos[1], pos[2]};
  for (uint64_t i = 0; i < n; ++i) {
    t.idx[i] = i;
    for (int a = 0; a < 3; ++a) {
      lo[a] = std::min(lo[a], pos[3 * i + a]);
      hi[a] = std::max(hi[a], pos[3 * i + a]);
    }
  }
  double half = 0;
  for (int a = 0; a < 3; ++a) half = std::max(half, 0.5 * (hi[a] - lo[a]));
  half = half * (1.0 + 1e-9) + 1e-300;
  t.nodes.reserve(2 * n / static_cast<size_t>(std::max(1, leaf_size)) + 16);
  t.nodes.emplace_back();
  for (int a = 0; a < 3; ++a) t.nodes[0].c[a] = 0.5 * (lo[a] + hi[a]);
  t.nodes[0].half = half;
  t.nodes[0].begin = 0;
  t.nodes[0].end = static_cast<uint32_t>(n);
  Builder b{pos, mass, static_cast<uint32_t>(std::max(1, leaf_size)), t.idx, {}, t.nodes};
  b.tmp.resize(n);
  t.depth = b.build(0, 0);
  for (uint64_t k = 0; k < n; ++k) {
    const uint64_t i = t.idx[k];
    for (int a = 0; a < 3; ++a) t.tpos[3 * k + a] = pos[3 * i + a];
    t.tmass[k] = mass[i];
  }
}

// ---- FP32 kernel arithmetic (include/galactic_tides/kernels/gravity_kernels.h) --

struct F4 {
  float x, y, z, w;
};
struct Acc {
  float ax, ay, az, pot;
};
struct Cell {
  F4 cm, q0, q1;
};
struct Group {
  uint32_t begin, end, cell_begin, cell_end, body_begin, body_end, pad0, pad1;
};

inline float inv_sqrt(float v) { return 1.0f / std::sqrt(v); }

// gravity_kernels.h:71-83
inline void body_term(Acc &a, float px, float py, float pz, const F4 &b, float G, float eps2) {
  const float dx = px - b.x, dy = py - b.y, dz = pz - b.z;
  const float inv = inv_sqrt(dx * dx + dy * dy + dz * dz + eps2);
  const float gm = G * b.w * inv;
  const float gm3 = gm * inv * inv;
  a.ax -= gm3 * dx;
  a.ay -= gm3 * dy;
  a.az -= gm3 * dz;
  a.pot -= gm;
}

// gravity_kernels.h:109-135
inline void cell_term(Acc &a, float px, float py, float pz, const Cell &c, float G, float eps2,
                      bool quad) {
  const float dx = px - c.cm.x, dy = py - c.cm.y, dz = pz - c.cm.z;
  const float inv = inv_sqrt(dx * dx + dy * dy + dz * dz + eps2);
  const float inv2 = inv * inv;
  const float inv3 = inv * inv2;
  const float gm = G * c.cm.w;
  a.ax -= gm * inv3 * dx;
  a.ay -= gm * inv3 * dy;
  a.az -= gm * inv3 * dz;
  a.pot -= gm * inv;
  if (quad) {
    const float qx = c.q0.x * dx + c.q0.w * dy + c.q1.x * dz;
    const float qy = c.q0.w * dx + c.q0.y * dy + c.q1.y * dz;
    const float qz = c.q1.x * dx + c.q1.y * dy + c.q0.z * dz;
    const float rqr = dx * qx + dy * qy + dz * qz;
    const float inv5 = inv3 * inv2;
    const float inv7 = inv5 * inv2;
    a.ax += G * (qx * inv5 - 2.5f * rqr * dx * inv7);
    a.ay += G * (qy * inv5 - 2.5f * rqr * dy * inv7);
    a.az += G * (qz * inv5 - 2.5f * rqr * dz * inv7);
    a.pot -= 0.5f * G * rqr * inv5;
  }
}

// ---- group lists (src/group_lists.cpp) ---------------------------------------

struct Lists {
  double origin[3] = {0, 0, 0};
  std::vector<F4> bodies;
  std::vector<Cell> cells;
  std::vector<uint32_t> cell_list, body_list;
  std::vector<Group> groups;
  uint64_t cell_int = 0, body_int = 0, max_list = 0;
  int max_group = 0;
};

struct WalkNode {
  double c[3], half, com[3], crit;
  int32_t first_child, n_children;
  uint32_t begin, end;
};

constexpr uint32_t kBodiesIfAtMost = 2;  // group_lists.h:58

void make_lists(const Tree &t, double theta, int group_size, bool pack, Lists &L) {
  const auto &nodes = t.nodes;
  const uint64_t n = t.tmass.size();
  L.cell_int = L.body_int = L.max_list = 0;
  L.max_group = 0;
  L.bodies.resize(n);
  L.cells.resize(nodes.size());
  L.groups.clear();
  L.cell_list.clear();
  L.body_list.clear();
  if (n == 0) return;
  for (int a = 0; a < 3; ++a) L.origin[a] = nodes[0].c[a];
  const double inv_theta = 1.0 / theta;
  // group_lists.cpp:40-63: FP32 bodies and cells relative to the root centre.
  std::vector<WalkNode> wn(nodes.size());
  for (uint64_t k = 0; k < n; ++k) {
    L.bodies[k] = F4{float(t.tpos[3 * k] - L.origin[0]), float(t.tpos[3 * k + 1] - L.origin[1]),
                     float(t.tpos[3 * k + 2] - L.origin[2]), float(t.tmass[k])};
  }
  for (size_t k = 0; k < nodes.size(); ++k) {
    const Node &nd = nodes[k];
    Cell &c = L.cells[k];
    c.cm = F4{float(nd.com[0] - L.origin[0]), float(nd.com[1] - L.origin[1]),
              float(nd.com[2] - L.origin[2]), float(nd.mass)};
    c.q0 = F4{float(nd.q[0]), float(nd.q[1]), float(nd.q[2]), float(nd.q[3])};
    c.q1 = F4{float(nd.q[4]), float(nd.q[5]), 0.0f, 0.0f};
    WalkNode &w = wn[k];
    for (int a = 0; a < 3; ++a) {
      w.c[a] = nd.c[a];
      w.com[a] = nd.com[a];
    }
    w.half = nd.half;
    w.crit = 2.0 * nd.half * inv_theta + nd.delta;
    w.first_child = nd.first_child;
    w.n_children = nd.n_children;
    w.begin = nd.begin;
    w.end = nd.end;
  }

  // group_lists.cpp:65-113: groups in tree order.
  const uint32_t gs = static_cast<uint32_t>(group_size);
  std::vector<std::pair<uint32_t, uint32_t>> specs;
  auto cut = [&](uint32_t b, uint32_t e) {
    for (uint32_t s = b; s < e; s += gs) specs.push_back({s, std::min(e, s + gs)});
  };
  std::function<void(size_t)> visit = [&](size_t ni) {
    const Node &nd = nodes[ni];
    if (nd.end - nd.begin <= gs || nd.n_children == 0) {
      cut(nd.begin, nd.end);
      return;
    }
    bool open = false;
    std::pair<uint32_t, uint32_t> run{0, 0};
    for (int k = 0; k < nd.n_children; ++k) {
      const size_t ci = static_cast<size_t>(nd.first_child + k);
      const Node &ch = nodes[ci];
      const uint32_t cc = ch.end - ch.begin;
      if (cc <= gs) {
        if (pack && open && (run.second - run.first) + cc <= gs) {
          run.second = ch.end;
        } else {
          if (open) specs.push_back(run);
          run = {ch.begin, ch.end};
          open = true;
        }
      } else {
        if (open) specs.push_back(run);
        open = false;
        visit(ci);
      }
    }
    if (open) specs.push_back(run);
  };
  visit(0);

  // group_lists.cpp:115-166: one FP64 walk per group.
  const long ng = static_cast<long>(specs.size());
  std::vector<std::vector<uint32_t>> gc_list(ng), gb_list(ng);
#pragma omp parallel for schedule(dynamic, 16)
  for (long g = 0; g < ng; ++g) {
    const uint32_t gb = specs[g].first, ge = specs[g].second;
    double lo[3], hi[3];
    for (int a = 0; a < 3; ++a) lo[a] = hi[a] = t.tpos[3 * gb + a];
    for (uint32_t k = gb; k < ge; ++k)
      for (int a = 0; a < 3; ++a) {
        lo[a] = std::min(lo[a], t.tpos[3 * k + a]);
        hi[a] = std::max(hi[a], t.tpos[3 * k + a]);
      }
    double gc[3], gh[3];
    for (int a = 0; a < 3; ++a) {
      gc[a] = 0.5 * (lo[a] + hi[a]);
      gh[a] = 0.5 * (hi[a] - lo[a]);
    }
    auto &cl = gc_list[g];
    auto &bl = gb_list[g];
    cl.clear();
    bl.clear();
    auto bodies_of = [&](uint32_t b, uint32_t e) {
      for (uint32_t j = b; j < e; ++j) bl.push_back(j);
    };
    uint32_t stack[1024];
    int sp = 0;
    stack[sp++] = 0;
    while (sp > 0) {
      const uint32_t ni = stack[--sp];
      const WalkNode &nd = wn[ni];
      bool overlap = true;
      double d2 = 0.0;
      for (int a = 0; a < 3; ++a) {
        if (std::fabs(nd.c[a] - gc[a]) > nd.half * 1.0001 + gh[a]) overlap = false;
        const double d = std::fabs(nd.com[a] - gc[a]) - gh[a];
        if (d > 0) d2 += d * d;
      }
      if (!overlap && d2 > nd.crit * nd.crit) {
        if (nd.end - nd.begin <= kBodiesIfAtMost)
          bodies_of(nd.begin, nd.end);
        else
          cl.push_back(ni);
      } else if (nd.n_children == 0) {
        bodies_of(nd.begin, nd.end);
      } else {
        for (int k = 0; k < nd.n_children; ++k)
          stack[sp++] = static_cast<uint32_t>(nd.first_child + k);
      }
    }
  }

  // group_lists.cpp:168-206: concatenation and statistics.
  uint64_t nc = 0, nb = 0;
  L.groups.resize(ng);
  for (long g = 0; g < ng; ++g) {
    Group &G = L.groups[g];
    G = Group{specs[g].first, specs[g].second, 0, 0, 0, 0, 0, 0};
    G.cell_begin = static_cast<uint32_t>(nc);
    nc += gc_list[g].size();
    G.cell_end = static_cast<uint32_t>(nc);
    G.body_begin = static_cast<uint32_t>(nb);
    nb += gb_list[g].size();
    G.body_end = static_cast<uint32_t>(nb);
  }
  L.cell_list.resize(nc);
  L.body_list.resize(nb);
  for (long g = 0; g < ng; ++g) {
    const Group &G = L.groups[g];
    std::copy(gc_list[g].begin(), gc_list[g].end(), L.cell_list.begin() + G.cell_begin);
    std::copy(gb_list[g].begin(), gb_list[g].end(), L.body_list.begin() + G.body_begin);
    const uint64_t tg = G.end - G.begin;
    L.cell_int += tg * gc_list[g].size();
    L.body_int += tg * gb_list[g].size() - tg;
    L.max_list = std::max<uint64_t>(L.max_list, gc_list[g].size() + gb_list[g].size());
    L.max_group = std::max(L.max_group, static_cast<int>(tg));
  }
}

// gravity_kernels.h:138-170 + gravity_backends.cpp:70-84
void eval_lists(const Lists &L, float G, float eps2, bool quad, float *out) {
  const long ng = static_cast<long>(L.groups.size());
#pragma omp parallel for schedule(dynamic, 8)
  for (long g = 0; g < ng; ++g) {
    const Group grp = L.groups[g];
    for (uint32_t i = grp.begin; i < grp.end; ++i) {
      const F4 bi = L.bodies[i];
      Acc a{0.0f, 0.0f, 0.0f, 0.0f};
      for (uint32_t k = grp.cell_begin; k < grp.cell_end; ++k)
        cell_term(a, bi.x, bi.y, bi.z, L.cells[L.cell_list[k]], G, eps2, quad);
      for (uint32_t k = grp.body_begin; k < grp.body_end; ++k) {
        const uint32_t j = L.body_list[k];
        if (j == i) continue;
        body_term(a, bi.x, bi.y, bi.z, L.bodies[j], G, eps2);
      }
      out[4 * i] = a.ax;
      out[4 * i + 1] = a.ay;
      out[4 * i + 2] = a.az;
      out[4 * i + 3] = a.pot;
    }
  }
}

// src/octree.cpp:275-344 (Barnes 1994 offset criterion, FP64).
void tree_accel(const Tree &t, const double *x, uint64_t self, double G, double theta,
                double eps, bool quad, double *acc, double &pot, uint64_t &ncell,
                uint64_t &npart) {
  acc[0] = acc[1] = acc[2] = 0;
  pot = 0;
  if (t.nodes.empty()) return;
  const double eps2 = eps * eps, inv_theta = 1.0 / theta;
  size_t stack[1024];
  int sp = 0;
  stack[sp++] = 0;
  double ax = 0, ay = 0, az = 0, phi = 0;
  while (sp > 0) {
    const Node &n = t.nodes[stack[--sp]];
    const double dx = x[0] - n.com[0], dy = x[1] - n.com[1], dz = x[2] - n.com[2];
    const double r2 = dx * dx + dy * dy + dz * dz;
    const double crit = 2.0 * n.half * inv_theta + n.delta;
    const double slack = n.half * 1.0001;
    const bool inside = std::fabs(x[0] - n.c[0]) <= slack && std::fabs(x[1] - n.c[1]) <= slack &&
                        std::fabs(x[2] - n.c[2]) <= slack;
    if (!inside && r2 > crit * crit) {
      const double inv = 1.0 / std::sqrt(r2 + eps2);
      const double inv3 = inv * inv * inv;
      const double gm3 = G * n.mass * inv3;
      ax -= gm3 * dx;
      ay -= gm3 * dy;
      az -= gm3 * dz;
      phi -= G * n.mass * inv;
      if (quad) {
        const double *q = n.q;
        const double qx = q[0] * dx + q[3] * dy + q[4] * dz;
        const double qy = q[3] * dx + q[1] * dy + q[5] * dz;
        const double qz = q[4] * dx + q[5] * dy + q[2] * dz;
        const double rqr = dx * qx + dy * qy + dz * qz;
        const double inv5 = inv3 * inv * inv;
        const double inv7 = inv5 * inv * inv;
        ax += G * (qx * inv5 - 2.5 * rqr * dx * inv7);
        ay += G * (qy * inv5 - 2.5 * rqr * dy * inv7);
        az += G * (qz * inv5 - 2.5 * rqr * dz * inv7);
        phi -= 0.5 * G * rqr * inv5;
      }
      ++ncell;
    } else if (n.n_children == 0) {
      for (uint32_t k = n.begin; k < n.end; ++k) {
        if (t.idx[k] == self) continue;
        const double ex = x[0] - t.tpos[3 * k], ey = x[1] - t.tpos[3 * k + 1],
                     ez = x[2] - t.tpos[3 * k + 2];
        const double inv = 1.0 / std::sqrt(ex * ex + ey * ey + ez * ez + eps2);
        const double gm3 = G * t.tmass[k] * inv * inv * inv;
        ax -= gm3 * ex;
        ay -= gm3 * ey;
        az -= gm3 * ez;
        phi -= G * t.tmass[k] * inv;
        ++npart;
      }
    } else {
      for (int k = 0; k < n.n_children; ++k) stack[sp++] = n.first_child + k;
    }
  }
  acc[0] = ax;
  acc[1] = ay;
  acc[2] = az;
  pot = phi;
}

}  // namespace

struct gto_run {
  Tree tree;
  Lists lists;
};

extern "C" {

gto_run *gto_new(void) { return new gto_run(); }
void gto_free(gto_run *r) { delete r; }

int gto_build(gto_run *r, uint64_t n, const double *pos, const double *mass, int leaf_size,
              uint64_t) {
  if (n > 0xffffffffull) return -1;
  build_tree(r->tree, n, pos, mass, leaf_size);
  return 0;
}

int gto_lists(gto_run *r, double theta, int group_size, int pack) {
  if (group_size < 1) return -1;
  make_lists(r->tree, theta, group_size, pack != 0, r->lists);
  return 0;
}

void gto_info(const gto_run *r, uint64_t *info) {
  info[0] = r->tree.tmass.size();
  info[1] = r->tree.nodes.size();
  info[2] = static_cast<uint64_t>(r->tree.depth);
  info[3] = r->lists.groups.size();
  info[4] = r->lists.cell_list.size();
  info[5] = r->lists.body_list.size();
  info[6] = r->lists.cell_int;
  info[7] = r->lists.body_int;
  info[8] = r->lists.max_list;
  info[9] = static_cast<uint64_t>(r->lists.max_group);
}

void gto_nodes(const gto_run *r, double *geom, int32_t *links) {
  const auto &nodes = r->tree.nodes;
  for (size_t k = 0; k < nodes.size(); ++k) {
    const Node &nd = nodes[k];
    double *g = geom + 16 * k;
    for (int a = 0; a < 3; ++a) {
      g[a] = nd.c[a];
      g[4 + a] = nd.com[a];
    }
    g[3] = nd.half;
    g[7] = nd.mass;
    for (int a = 0; a < 6; ++a) g[8 + a] = nd.q[a];
    g[14] = nd.delta;
    g[15] = 0.0;
    links[4 * k] = nd.first_child;
    links[4 * k + 1] = nd.n_children;
    links[4 * k + 2] = static_cast<int32_t>(nd.begin);
    links[4 * k + 3] = static_cast<int32_t>(nd.end);
  }
}

void gto_order(const gto_run *r, uint64_t *order) {
  std::copy(r->tree.idx.begin(), r->tree.idx.end(), order);
}

void gto_tree_positions(const gto_run *r, double *pos, double *mass) {
  std::copy(r->tree.tpos.begin(), r->tree.tpos.end(), pos);
  std::copy(r->tree.tmass.begin(), r->tree.tmass.end(), mass);
}

void gto_lists_get(const gto_run *r, uint32_t *groups, uint32_t *cell_list, uint32_t *body_list,
                   float *bodies, float *cells, double *origin) {
  const Lists &L = r->lists;
  if (groups) std::memcpy(groups, L.groups.data(), L.groups.size() * sizeof(Group));
  if (cell_list) std::memcpy(cell_list, L.cell_list.data(), L.cell_list.size() * 4);
  if (body_list) std::memcpy(body_list, L.body_list.data(), L.body_list.size() * 4);
  if (bodies) std::memcpy(bodies, L.bodies.data(), L.bodies.size() * sizeof(F4));
  if (cells) std::memcpy(cells, L.cells.data(), L.cells.size() * sizeof(Cell));
  if (origin)
    for (int a = 0; a < 3; ++a) origin[a] = L.origin[a];
}

void gto_eval(const gto_run *r, double G, double eps, int quadrupole, float *out) {
  eval_lists(r->lists, static_cast<float>(G), static_cast<float>(eps * eps), quadrupole != 0, out);
}

void gto_tree_gravity(gto_run *r, uint64_t n, const double *pos, const double *mass, double G,
                      double theta, double eps, int quadrupole, int leaf_size, double *acc,
                      double *pot, uint64_t *stats) {
  build_tree(r->tree, n, pos, mass, leaf_size);
  uint64_t nc = 0, np = 0;
  const long nl = static_cast<long>(n);
  const Tree &t = r->tree;
#pragma omp parallel for schedule(dynamic, 256) reduction(+ : nc, np)
  for (long k = 0; k < nl; ++k) {
    const uint64_t i = t.idx[k];
    uint64_t c = 0, p = 0;
    tree_accel(t, pos + 3 * i, i, G, theta, eps, quadrupole != 0, acc + 3 * i, pot[i], c, p);
    nc += c;
    np += p;
  }
  if (stats) {
    stats[0] = nc;
    stats[1] = np;
  }
}

// Octree::accel_pot (src/octree.cpp:275-344) at np points with nobody left out
// (self = Octree::kNoSelf) on the last built tree.
void gto_accel_at(const gto_run *r, uint64_t np, const double *points, double G, double theta,
                  double eps, int quadrupole, double *acc, double *pot) {
  const long nl = static_cast<long>(np);
#pragma omp parallel for schedule(dynamic, 64)
  for (long k = 0; k < nl; ++k) {
    uint64_t c = 0, p = 0;
    tree_accel(r->tree, points + 3 * k, UINT64_MAX, G, theta, eps, quadrupole != 0, acc + 3 * k,
               pot[k], c, p);
  }
}

void gto_direct_f64(uint64_t n, const double *pos, const double *mass, double G, double eps,
                    double *acc, double *pot) {
  const long nl = static_cast<long>(n);
  const double eps2 = eps * eps;
#pragma omp parallel for schedule(static)
  for (long i = 0; i < nl; ++i) {
    double ax = 0, ay = 0, az = 0, phi = 0;
    for (long j = 0; j < nl; ++j) {
      if (j == i) continue;
      const double dx = pos[3 * i] - pos[3 * j], dy = pos[3 * i + 1] - pos[3 * j + 1],
                   dz = pos[3 * i + 2] - pos[3 * j + 2];
      const double inv = 1.0 / std::sqrt(dx * dx + dy * dy + dz * dz + eps2);
      const double s = G * mass[j] * inv * inv * inv;
      ax -= s * dx;
      ay -= s * dy;
      az -= s * dz;
      phi -= G * mass[j] * inv;
    }
    acc[3 * i] = ax;
    acc[3 * i + 1] = ay;
    acc[3 * i + 2] = az;
    pot[i] = phi;
  }
}

void gto_direct_f64_sample(uint64_t n, const double *pos, const double *mass, double G,
                           double eps, uint64_t ns, const uint64_t *sample, double *acc) {
  const double eps2 = eps * eps;
  const long nsl = static_cast<long>(ns);
#pragma omp parallel for schedule(dynamic)
  for (long s = 0; s < nsl; ++s) {
    const uint64_t i = sample[s];
    double ax = 0, ay = 0, az = 0;
    for (uint64_t j = 0; j < n; ++j) {
      if (j == i) continue;
      const double dx = pos[3 * i] - pos[3 * j], dy = pos[3 * i + 1] - pos[3 * j + 1],
                   dz = pos[3 * i + 2] - pos[3 * j + 2];
      const double inv = 1.0 / std::sqrt(dx * dx + dy * dy + dz * dz + eps2);
      const double f = G * mass[j] * inv * inv * inv;
      ax -= f * dx;
      ay -= f * dy;
      az -= f * dz;
    }
    acc[3 * s] = ax;
    acc[3 * s + 1] = ay;
    acc[3 * s + 2] = az;
  }
}

void gto_direct_f32(uint64_t n, const float *bodies, double G, double eps, float *out) {
  const float g = static_cast<float>(G), e2 = static_cast<float>(eps * eps);
  const F4 *b = reinterpret_cast<const F4 *>(bodies);
  const long nl = static_cast<long>(n);
#pragma omp parallel for schedule(static)
  for (long i = 0; i < nl; ++i) {
    Acc a{0.0f, 0.0f, 0.0f, 0.0f};
    const F4 bi = b[i];
    for (long j = 0; j < nl; ++j) {
      if (j == i) continue;
      body_term(a, bi.x, bi.y, bi.z, b[j], g, e2);
    }
    out[4 * i] = a.ax;
    out[4 * i + 1] = a.ay;
    out[4 * i + 2] = a.az;
    out[4 * i + 3] = a.pot;
  }
}

void gto_pack_bodies(uint64_t n, const double *pos, const double *mass, const double *origin,
                     float *out) {
  for (uint64_t i = 0; i < n; ++i) {
    out[4 * i] = static_cast<float>(pos[3 * i] - origin[0]);
    out[4 * i + 1] = static_cast<float>(pos[3 * i + 1] - origin[1]);
    out[4 * i + 2] = static_cast<float>(pos[3 * i + 2] - origin[2]);
    out[4 * i + 3] = static_cast<float>(mass[i]);
  }
}

int gto_group_gravity(gto_run *r, uint64_t n, const double *pos, const double *mass, double G,
                      double theta, double eps, int quadrupole, int group_size, int leaf_size,
                      int pack, double *acc, double *pot) {
  if (group_size < 1 || n > 0xffffffffull) return -1;
  build_tree(r->tree, n, pos, mass, leaf_size);
  make_lists(r->tree, theta, group_size, pack != 0, r->lists);
  std::vector<float> out(4 * n);
  eval_lists(r->lists, static_cast<float>(G), static_cast<float>(eps * eps), quadrupole != 0,
             out.data());
  for (uint64_t k = 0; k < n; ++k) {
    const uint64_t i = r->tree.idx[k];
    acc[3 * i] = out[4 * k];
    acc[3 * i + 1] = out[4 * k + 1];
    acc[3 * i + 2] = out[4 * k + 2];
    pot[i] = out[4 * k + 3];
  }
  return 0;
}

int gto_kdk_group(gto_run *r, uint64_t n, double *pos, double *vel, const double *mass,
                  double G, double theta, double eps, int quadrupole, int group_size,
                  int leaf_size, int pack, int nsteps, double dt, double *acc, double *pot) {
  int rc = gto_group_gravity(r, n, pos, mass, G, theta, eps, quadrupole, group_size, leaf_size,
                             pack, acc, pot);
  if (rc) return rc;
  const double h = 0.5 * dt;
  const long nl = static_cast<long>(n);
  for (int s = 0; s < nsteps; ++s) {
#pragma omp parallel for schedule(static)
    for (long i = 0; i < nl; ++i)
      for (int a = 0; a < 3; ++a) {
        vel[3 * i + a] += h * acc[3 * i + a];
        pos[3 * i + a] += dt * vel[3 * i + a];
      }
    rc = gto_group_gravity(r, n, pos, mass, G, theta, eps, quadrupole, group_size, leaf_size,
                           pack, acc, pot);
    if (rc) return rc;
#pragma omp parallel for schedule(static)
    for (long i = 0; i < nl; ++i)
      for (int a = 0; a < 3; ++a) vel[3 * i + a] += h * acc[3 * i + a];
  }
  return 0;
}

int gto_time_group(gto_run *r, uint64_t n, const double *pos, const double *mass, double G,
                   double theta, double eps, int quadrupole, int group_size, int leaf_size,
                   int pack, int repeat, double *best4) {
  using clk = std::chrono::steady_clock;
  auto sec = [](clk::time_point a, clk::time_point b) {
    return std::chrono::duration<double>(b - a).count();
  };
  std::vector<double> acc(3 * n), pot(n);
  std::vector<float> out(4 * n);
  best4[0] = 1e300;
  for (int rep = 0; rep <= repeat; ++rep) {
    const auto t0 = clk::now();
    build_tree(r->tree, n, pos, mass, leaf_size);
    const auto t1 = clk::now();
    make_lists(r->tree, theta, group_size, pack != 0, r->lists);
    const auto t2 = clk::now();
    eval_lists(r->lists, static_cast<float>(G), static_cast<float>(eps * eps), quadrupole != 0,
               out.data());
    const auto t3 = clk::now();
    for (uint64_t k = 0; k < n; ++k) {
      const uint64_t i = r->tree.idx[k];
      acc[3 * i] = out[4 * k];
      acc[3 * i + 1] = out[4 * k + 1];
      acc[3 * i + 2] = out[4 * k + 2];
      pot[i] = out[4 * k + 3];
    }
    const auto t4 = clk::now();
    if (rep > 0 && sec(t0, t4) < best4[0]) {
      best4[0] = sec(t0, t4);
      best4[1] = sec(t0, t1);
      best4[2] = sec(t1, t2);
      best4[3] = sec(t2, t3);
      best4[4] = static_cast<double>(r->lists.cell_int);
      best4[5] = static_cast<double>(r->lists.body_int);
    }
  }
  return 0;
}

void gto_test_system(int kind, uint64_t n, uint64_t seed, double *pos, double *mass) {
  std::mt19937_64 rng(seed);
  std::uniform_real_distribution<double> u(0.0, 1.0);
  for (uint64_t i = 0; i < n; ++i) {
    double cx = 0, cy = 0, cz = 0, a = 1.0;
    if (kind == 0 && i % 10 == 0) {
      if (i % 20 == 0) {
        cx = 4;
      } else {
        cx = -1;
        cy = 3;
        cz = 2;
      }
      a = 0.1;
    }
    if (kind == 1 && i % 7 == 0) {
      cx = 3;
      cy = 1;
    }
    const double x = std::sqrt(u(rng) * 0.95);
    const double r = a * x / (1 - x);
    const double mu = 2 * u(rng) - 1, ph = 2 * M_PI * u(rng), s = std::sqrt(1 - mu * mu);
    const double ux = s * std::cos(ph), uy = s * std::sin(ph), uz = mu;
    pos[3 * i] = cx + r * ux;
    pos[3 * i + 1] = cy + r * uy;
    pos[3 * i + 2] = cz + r * uz;
    mass[i] = 1.0 / static_cast<double>(n);
  }
}

uint64_t gto_within(const gto_run *r, uint64_t nq, const double *centres, const double *radii,
                    uint64_t n_radii, uint64_t *offsets, uint64_t cap, uint64_t *index,
                    double *dist2) {
  const Tree &t = r->tree;
  const long nl = static_cast<long>(nq);
  std::vector<uint64_t> count(nq, 0);
  // The two passes run the same walk: counting, then writing from the offsets.
  auto walk = [&](uint64_t q, uint64_t *out_i, double *out_d) {
    uint64_t found = 0;
    if (t.nodes.empty()) return found;
    const double *c = centres + 3 * q;
    const double rad = radii[n_radii == 1 ? 0 : q];
    const double r2 = rad * rad;
    std::vector<uint32_t> stack{0};  // LIFO: the last child pushed is taken first
    while (!stack.empty()) {
      const Node &nd = t.nodes[stack.back()];
      stack.pop_back();
      double d2 = 0.0;
      for (int k = 0; k < 3; ++k) {
        const double d = std::fabs(c[k] - nd.c[k]) - nd.half;
        if (d > 0) d2 += d * d;
      }
      if (d2 > r2) continue;
      if (nd.n_children == 0) {
        for (uint32_t j = nd.begin; j < nd.end; ++j) {
          const double *p = &t.tpos[3 * static_cast<size_t>(j)];
          const double dx = p[0] - c[0], dy = p[1] - c[1], dz = p[2] - c[2];
          const double dd = dx * dx + dy * dy + dz * dz;
          if (dd <= r2) {
            if (out_i) {
              out_i[found] = t.idx[j];
              out_d[found] = dd;
            }
            ++found;
          }
        }
      } else {
        for (int k = 0; k < nd.n_children; ++k)
          stack.push_back(static_cast<uint32_t>(nd.first_child + k));
      }
    }
    return found;
  };
#pragma omp parallel for schedule(dynamic, 64)
  for (long q = 0; q < nl; ++q) count[q] = walk(static_cast<uint64_t>(q), nullptr, nullptr);
  uint64_t total = 0;
  for (uint64_t q = 0; q < nq; ++q) {
    offsets[q] = total;
    total += count[q];
  }
  offsets[nq] = total;
  if (index && dist2 && total <= cap) {
#pragma omp parallel for schedule(dynamic, 64)
    for (long q = 0; q < nl; ++q)
      walk(static_cast<uint64_t>(q), index + offsets[q], dist2 + offsets[q]);
  }
  return total;
}

int gto_max_threads(void) {
#ifdef _OPENMP
  return omp_get_max_threads();
#else
  return 1;
#endif
}

}  // extern "C"
