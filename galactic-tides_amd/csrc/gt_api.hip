// C-ABI entry points of libgt_hip.so (declared in include/gt_hip.h): the
// context, the reference's two GPU calls (direct, tree), the whole force
// evaluation on the GPU, and the device-resident leapfrog.
#include <algorithm>
#include <cstring>
#include <new>
#include <string>
#include <vector>

#include "gt_build.h"
#include "gt_internal.h"

using namespace gth;

struct gt_hip_context {
  int device = 0;
  hipStream_t stream = nullptr;
  hipEvent_t ev[6] = {};
  std::string err;
  double gpu_seconds = 0.0;
  gt_lists_info info{};
  gt_timings tm{};
  hipDeviceProp_t prop{};

  TreeBuilder tree;
  // caller-supplied tree (gt_hip_walk_tree)
  DBuf<double> up_geom, up_tpos, up_tmass;
  DBuf<int4> up_lnk;
  // lists of the last walk
  DBuf<WalkRec> recs;
  DBuf<uint2> span;
  DBuf<Cell> cells;
  DBuf<float4> bodies, out;
  DBuf<uint32_t> parent, skip, depth, flags, offs, starts, ccount, bcount, coff, boff, cell_list,
      body_list;
  // per-group list slots of the one-pass walk, from the previous evaluation's
  // longest lists (0 before the first)
  uint32_t cap_cells = 0, cap_bodies = 0;
  uint64_t walk_passes = 0, walk_fallbacks = 0;
  DBuf<gt_group> groups;
  DBuf<unsigned long long> scan_tmp, misc;
  uint32_t lists_n = 0, lists_nn = 0;
  const double *lists_geom = nullptr;  // geometry rows the last lists refer to
  // particle state (resident leapfrog) and inputs of single evaluations
  DBuf<double> pos, vel, mass, acc, pot;
  DBuf<double> in_pos, in_mass, in_acc, in_pot;
  uint64_t n_state = 0;
  bool state_forces = false;
  // raw-array entry points and sampling
  DBuf<float4> raw_bodies, raw_out;
  DBuf<Cell> raw_cells;
  DBuf<uint32_t> raw_cl, raw_bl;
  DBuf<gt_group> raw_groups;
  DBuf<uint64_t> sample_idx;
  DBuf<double> sample_acc;
};

namespace {

thread_local std::string g_create_error;

template <class F>
int guarded(gt_hip_context *ctx, F &&f) {
  try {
    f();
    if (ctx) ctx->err.clear();
    return GT_HIP_OK;
  } catch (const Error &e) {
    if (ctx) ctx->err = e.what();
    return e.code;
  } catch (const std::bad_alloc &) {
    if (ctx) ctx->err = "host allocation failed";
    return GT_HIP_ERR_MEMORY;
  } catch (const std::exception &e) {
    if (ctx) ctx->err = e.what();
    return GT_HIP_ERR_DEVICE;
  } catch (...) {
    if (ctx) ctx->err = "unknown error";
    return GT_HIP_ERR_DEVICE;
  }
}

void require(bool ok, const char *msg) {
  if (!ok) fail(GT_HIP_ERR_ARGUMENT, msg);
}

void check_config(const gt_group_config *cfg) {
  require(cfg != nullptr, "config is NULL");
  require(cfg->group_size >= 1, "group_size must be at least 1");
  if (cfg->group_size > kMaxGroup)
    fail(GT_HIP_ERR_ARGUMENT,
         "group_size can be at most " + std::to_string(kMaxGroup) + " on hip (one wave64)");
  require(cfg->leaf_size >= 1, "leaf_size must be at least 1");
  require(cfg->theta > 0.0, "theta must be positive");
}

template <class T>
void upload(T *dst, const T *src, size_t n, hipStream_t s) {
  if (n) GTH_CHECK(hipMemcpyAsync(dst, src, n * sizeof(T), hipMemcpyHostToDevice, s));
}

template <class T>
void download(T *dst, const T *src, size_t n, hipStream_t s) {
  if (n) GTH_CHECK(hipMemcpyAsync(dst, src, n * sizeof(T), hipMemcpyDeviceToHost, s));
}

float elapsed_ms(hipEvent_t a, hipEvent_t b) {
  float ms = 0.0f;
  GTH_CHECK(hipEventElapsedTime(&ms, a, b));
  return ms;
}

// Group selection, walk and list statistics for a tree held in device memory.
// Leaves groups, lists, bodies and cells in the context.  The walk writes each
// group's lists into slots sized from the previous evaluation (one pass); if a
// group's lists do not fit, its counts say so and the lists are rewritten at
// exact scanned offsets (the same walk again, so the contents are the same).
void run_lists(gt_hip_context *c, const gt_group_config *cfg, const double *geom,
               const int4 *lnk, const double *tpos, const double *tmass, uint32_t n,
               uint32_t nn) {
  hipStream_t s = c->stream;
  c->lists_n = n;
  c->lists_nn = nn;
  c->lists_geom = geom;
  c->info = gt_lists_info{};
  c->info.n = n;
  c->info.n_nodes = nn;
  if (n == 0) return;
  if (nn >= kRecIndexMask)
    fail(GT_HIP_ERR_LIMIT, "too many tree nodes for the packed walk records (2^26 - 1)");
  unsigned long long *misc = c->misc.ensure(16);  // [0..2] totals, [3] error, [4..11] stats
  launch_fill_u64(misc, 16, 0ull, s);
  uint32_t *werr = reinterpret_cast<uint32_t *>(misc + 3);
  uint32_t *parent = c->parent.ensure(nn), *skip = c->skip.ensure(nn), *depth = c->depth.ensure(nn);
  launch_link_up(lnk, nn, parent, skip, depth, s);
  WalkRec *recs = c->recs.ensure(nn);
  uint2 *span = c->span.ensure(nn);
  Cell *cells = c->cells.ensure(nn);
  float4 *bodies = c->bodies.ensure(n);
  launch_prep_lists(geom, lnk, skip, depth, tpos, tmass, n, nn, cfg->theta, recs, span, cells,
                    bodies, werr, s);

  const int pack = cfg->pack < 0 ? 1 : (cfg->pack != 0);
  uint32_t *flags = c->flags.ensure(n), *offs = c->offs.ensure(n);
  unsigned long long *stmp = c->scan_tmp.ensure(scan_tmp_words(n) + 8);
  launch_group_starts(lnk, nn, n, static_cast<uint32_t>(cfg->group_size), pack, flags, s);
  exclusive_scan_u32(flags, offs, n, misc, stmp, s);
  unsigned long long hm[12];
  download(hm, misc, 4, s);
  GTH_CHECK(hipStreamSynchronize(s));
  if (hm[3] & 2u) fail(GT_HIP_ERR_ARGUMENT, "tree: a node's half size is not root half / 2^depth");
  const uint32_t ng = static_cast<uint32_t>(hm[0]);
  gt_group *groups = c->groups.ensure(ng);
  uint32_t *starts = c->starts.ensure(ng);
  launch_make_groups(flags, offs, n, ng, starts, groups, s);

  uint32_t *cc = c->ccount.ensure(ng), *bc = c->bcount.ensure(ng);
  uint32_t cap_c = c->cap_cells, cap_b = c->cap_bodies;
  if (static_cast<uint64_t>(ng) * cap_c > 0xffffffffull ||
      static_cast<uint64_t>(ng) * cap_b > 0xffffffffull)
    cap_c = cap_b = 0;
  uint32_t *cl = c->cell_list.ensure(std::max<uint64_t>(static_cast<uint64_t>(ng) * cap_c, 1));
  uint32_t *bl = c->body_list.ensure(std::max<uint64_t>(static_cast<uint64_t>(ng) * cap_b, 1));
  const int lanes = cfg->walk_lanes == 16 || cfg->walk_lanes == 32 || cfg->walk_lanes == 64
                        ? cfg->walk_lanes
                        : 1;
  launch_walk(kWalkCapacity, lanes, recs, span, tpos, groups, ng, nn, geom, cap_c, cap_b, nullptr,
              nullptr, cc, bc, cl, bl, werr, s);
  launch_list_stats(cc, bc, groups, ng, misc + 4, s);
  download(hm, misc, 12, s);
  GTH_CHECK(hipStreamSynchronize(s));
  ++c->walk_passes;
  if (hm[3] & 1u) fail(GT_HIP_ERR_DEVICE, "group walk: malformed tree (walk did not terminate)");
  const unsigned long long most_c = hm[8], most_b = hm[9], all_c = hm[10], all_b = hm[11];
  if (all_c > 0xffffffffull || all_b > 0xffffffffull)
    fail(GT_HIP_ERR_LIMIT, "interaction lists too long for 32 bit offsets");
  if (most_c > cap_c || most_b > cap_b) {
    // Some group did not fit its slots: place every group at scanned offsets.
    ++c->walk_fallbacks;
    uint32_t *co = c->coff.ensure(ng), *bo = c->boff.ensure(ng);
    stmp = c->scan_tmp.ensure(scan_tmp_words(ng) + 8);
    exclusive_scan_u32(cc, co, ng, misc + 1, stmp, s);
    exclusive_scan_u32(bc, bo, ng, misc + 2, stmp, s);
    cl = c->cell_list.ensure(std::max<unsigned long long>(all_c, 1));
    bl = c->body_list.ensure(std::max<unsigned long long>(all_b, 1));
    launch_walk(kWalkExact, lanes, recs, span, tpos, groups, ng, nn, geom, 0, 0, co, bo, cc, bc, cl,
                bl, werr, s);
  }
  // Slots for the next evaluation: a quarter more than this one's longest lists.
  // (multiples of 4, so that slots start 16-byte aligned for the walk's vector stores)
  c->cap_cells = static_cast<uint32_t>(
      std::min<unsigned long long>((most_c + most_c / 4 + 32 + 3) & ~3ull, 1u << 30));
  c->cap_bodies = static_cast<uint32_t>(
      std::min<unsigned long long>((most_b + most_b / 4 + 32 + 3) & ~3ull, 1u << 30));
  c->info.n_groups = ng;
  c->info.n_cell_list = all_c;
  c->info.n_body_list = all_b;
  c->info.cell_interactions = hm[4];
  c->info.body_interactions = hm[5];
  c->info.max_list = hm[6];
  c->info.max_group = hm[7];
}

void run_eval(gt_hip_context *c, const gt_group_config *cfg) {
  float4 *out = c->out.ensure(std::max<uint32_t>(c->lists_n, 1));
  const float G = static_cast<float>(cfg->G);
  const float eps2 = static_cast<float>(cfg->eps * cfg->eps);
  launch_eval(c->bodies.get(), c->cells.get(), c->cell_list.get(), c->body_list.get(),
              c->groups.get(), static_cast<uint32_t>(c->info.n_groups), G, eps2,
              cfg->quadrupole ? 1u : 0u, cfg->exact != 0, cfg->eval_waves == 2 ? 2 : 3, out,
              c->stream);
}

// One force evaluation of particles already in device memory: build, lists,
// evaluate, scatter.  Phase times go to c->tm.
void force_eval(gt_hip_context *c, const gt_group_config *cfg, const double *pos,
                const double *mass, uint32_t n, double *acc, double *pot) {
  hipStream_t s = c->stream;
  GTH_CHECK(hipEventRecord(c->ev[0], s));
  c->tree.build(pos, mass, n, cfg->leaf_size, s);
  GTH_CHECK(hipEventRecord(c->ev[1], s));
  run_lists(c, cfg, c->tree.geom(), c->tree.links(), c->tree.tree_pos(), c->tree.tree_mass(), n,
            c->tree.n_nodes());
  c->info.depth = static_cast<uint64_t>(c->tree.depth());
  GTH_CHECK(hipEventRecord(c->ev[2], s));
  run_eval(c, cfg);
  GTH_CHECK(hipEventRecord(c->ev[3], s));
  launch_scatter(c->out.get(), c->tree.order(), n, acc, pot, s);
  GTH_CHECK(hipEventRecord(c->ev[4], s));
  GTH_CHECK(hipStreamSynchronize(s));
  c->tm.tree = elapsed_ms(c->ev[0], c->ev[1]);
  c->tm.lists = elapsed_ms(c->ev[1], c->ev[2]);
  c->tm.evaluate = elapsed_ms(c->ev[2], c->ev[3]);
  c->tm.scatter = elapsed_ms(c->ev[3], c->ev[4]);
  c->tm.total = elapsed_ms(c->ev[0], c->ev[4]);
  c->gpu_seconds = c->tm.total * 1e-3;
}

}  // namespace

extern "C" {

int gt_hip_compiled_in(void) { return 1; }

int gt_hip_create(gt_hip_context **out, int device) {
  if (!out) return GT_HIP_ERR_ARGUMENT;
  *out = nullptr;
  gt_hip_context *c = nullptr;
  int rc = GT_HIP_OK;
  try {
    int count = 0;
    if (hipGetDeviceCount(&count) != hipSuccess || count == 0)
      fail(GT_HIP_ERR_DEVICE, "hip: no GPU device available");
    if (device < 0 || device >= count)
      fail(GT_HIP_ERR_ARGUMENT, "hip: device " + std::to_string(device) + " does not exist");
    GTH_CHECK(hipSetDevice(device));
    c = new gt_hip_context();
    c->device = device;
    GTH_CHECK(hipGetDeviceProperties(&c->prop, device));
    GTH_CHECK(hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking));
    for (auto &e : c->ev) GTH_CHECK(hipEventCreate(&e));
    g_create_error.clear();
  } catch (const Error &e) {
    g_create_error = e.what();
    rc = e.code;
  } catch (const std::exception &e) {
    g_create_error = e.what();
    rc = GT_HIP_ERR_DEVICE;
  }
  if (rc != GT_HIP_OK) {
    delete c;
    return rc;
  }
  *out = c;
  return GT_HIP_OK;
}

void gt_hip_destroy(gt_hip_context *c) {
  if (!c) return;
  (void)hipSetDevice(c->device);
  if (c->stream) (void)hipStreamSynchronize(c->stream);
  for (auto &e : c->ev)
    if (e) (void)hipEventDestroy(e);
  if (c->stream) (void)hipStreamDestroy(c->stream);
  delete c;
}

const char *gt_hip_last_error(const gt_hip_context *c) {
  return c ? c->err.c_str() : g_create_error.c_str();
}

int gt_hip_device_name(const gt_hip_context *c, char *buf, size_t len) {
  if (!c || !buf || !len) return GT_HIP_ERR_ARGUMENT;
  std::snprintf(buf, len, "%s (%s, %d CUs)", c->prop.name, c->prop.gcnArchName,
                c->prop.multiProcessorCount);
  return GT_HIP_OK;
}

double gt_hip_last_gpu_seconds(const gt_hip_context *c) { return c ? c->gpu_seconds : 0.0; }

int gt_hip_direct(gt_hip_context *c, const gt_float4 *bodies, const gt_params *p,
                  gt_float4 *out) {
  if (!c) return GT_HIP_ERR_ARGUMENT;
  return guarded(c, [&] {
    require(p != nullptr, "params is NULL");
    const uint32_t n = p->n;
    c->gpu_seconds = 0.0;
    if (n == 0) return;
    require(bodies && out, "bodies/out is NULL");
    GTH_CHECK(hipSetDevice(c->device));
    float4 *db = c->raw_bodies.ensure(n), *dout = c->raw_out.ensure(n);
    upload(db, reinterpret_cast<const float4 *>(bodies), n, c->stream);
    GTH_CHECK(hipEventRecord(c->ev[0], c->stream));
    launch_direct(db, n, p->G, p->eps2, dout, c->stream);
    GTH_CHECK(hipEventRecord(c->ev[1], c->stream));
    download(reinterpret_cast<float4 *>(out), dout, n, c->stream);
    GTH_CHECK(hipStreamSynchronize(c->stream));
    c->gpu_seconds = elapsed_ms(c->ev[0], c->ev[1]) * 1e-3;
  });
}

int gt_hip_tree(gt_hip_context *c, const gt_float4 *bodies, uint64_t n_cells,
                const gt_cell *cells, uint64_t n_cell_list, const uint32_t *cell_list,
                uint64_t n_body_list, const uint32_t *body_list, const gt_group *groups,
                const gt_params *p, int exact, gt_float4 *out) {
  if (!c) return GT_HIP_ERR_ARGUMENT;
  return guarded(c, [&] {
    require(p != nullptr, "params is NULL");
    const uint32_t n = p->n, ng = p->n_groups;
    c->gpu_seconds = 0.0;
    if (n == 0 || ng == 0) return;
    require(bodies && cells && groups && out, "array is NULL");
    require(n_cell_list == 0 || cell_list, "cell_list is NULL");
    require(n_body_list == 0 || body_list, "body_list is NULL");
    // metal_backend.mm:218-223: the arrays must agree with the parameters, and
    // every group must fit one wave; every index must be in range.
    for (uint32_t g = 0; g < ng; ++g) {
      const gt_group &G = groups[g];
      require(G.begin <= G.end && G.end <= n, "hip tree: a group's targets are out of range");
      if (G.end - G.begin > static_cast<uint32_t>(kMaxGroup))
        fail(GT_HIP_ERR_ARGUMENT,
             "hip tree: a group has more than " + std::to_string(kMaxGroup) + " targets");
      require(G.cell_begin <= G.cell_end && G.cell_end <= n_cell_list &&
                  G.body_begin <= G.body_end && G.body_end <= n_body_list,
              "hip tree: a group's list range is out of range");
    }
    for (uint64_t k = 0; k < n_cell_list; ++k)
      require(cell_list[k] < n_cells, "hip tree: cell index out of range");
    for (uint64_t k = 0; k < n_body_list; ++k)
      require(body_list[k] < n, "hip tree: body index out of range");
    GTH_CHECK(hipSetDevice(c->device));
    hipStream_t s = c->stream;
    float4 *db = c->raw_bodies.ensure(n), *dout = c->raw_out.ensure(n);
    Cell *dc = c->raw_cells.ensure(std::max<uint64_t>(n_cells, 1));
    uint32_t *dcl = c->raw_cl.ensure(std::max<uint64_t>(n_cell_list, 1));
    uint32_t *dbl = c->raw_bl.ensure(std::max<uint64_t>(n_body_list, 1));
    gt_group *dg = c->raw_groups.ensure(ng);
    upload(db, reinterpret_cast<const float4 *>(bodies), n, s);
    upload(dc, reinterpret_cast<const Cell *>(cells), n_cells, s);
    upload(dcl, cell_list, n_cell_list, s);
    upload(dbl, body_list, n_body_list, s);
    upload(dg, groups, ng, s);
    GTH_CHECK(hipEventRecord(c->ev[0], s));
    launch_eval(db, dc, dcl, dbl, dg, ng, p->G, p->eps2, p->quadrupole, exact != 0, 2, dout, s);
    GTH_CHECK(hipEventRecord(c->ev[1], s));
    download(reinterpret_cast<float4 *>(out), dout, n, s);
    GTH_CHECK(hipStreamSynchronize(s));
    c->gpu_seconds = elapsed_ms(c->ev[0], c->ev[1]) * 1e-3;
  });
}

int gt_hip_group_compute(gt_hip_context *c, const gt_group_config *cfg, uint64_t n,
                         const double *pos, const double *mass, double *acc, double *pot) {
  if (!c) return GT_HIP_ERR_ARGUMENT;
  return guarded(c, [&] {
    check_config(cfg);
    if (n > 0xffffffffull) fail(GT_HIP_ERR_LIMIT, "too many particles for the 32 bit GPU indices");
    if (n == 0) {
      c->info = gt_lists_info{};
      c->tm = gt_timings{};
      return;
    }
    require(pos && mass && acc && pot, "array is NULL");
    GTH_CHECK(hipSetDevice(c->device));
    hipStream_t s = c->stream;
    double *dp = c->in_pos.ensure(3 * n), *dm = c->in_mass.ensure(n);
    double *da = c->in_acc.ensure(3 * n), *dq = c->in_pot.ensure(n);
    upload(dp, pos, 3 * n, s);
    upload(dm, mass, n, s);
    force_eval(c, cfg, dp, dm, static_cast<uint32_t>(n), da, dq);
    download(acc, da, 3 * n, s);
    download(pot, dq, n, s);
    GTH_CHECK(hipStreamSynchronize(s));
  });
}

int gt_hip_walk_tree(gt_hip_context *c, const gt_group_config *cfg, uint64_t n,
                     uint64_t n_nodes, const double *geom, const int32_t *links,
                     const double *tree_pos, const double *tree_mass, gt_float4 *out) {
  if (!c) return GT_HIP_ERR_ARGUMENT;
  return guarded(c, [&] {
    check_config(cfg);
    if (n > 0xffffffffull || n_nodes > 0x7fffffffull)
      fail(GT_HIP_ERR_LIMIT, "too many particles or nodes for the 32 bit GPU indices");
    require(n == 0 || (n_nodes > 0 && geom && links && tree_pos && tree_mass),
            "tree arrays are NULL");
    // The walk trusts the links: check them here so a bad tree is an error,
    // not an out-of-bounds access on the GPU.
    for (uint64_t k = 0; k < n_nodes; ++k) {
      const int32_t fc = links[4 * k], nc = links[4 * k + 1];
      const uint32_t b = static_cast<uint32_t>(links[4 * k + 2]),
                     e = static_cast<uint32_t>(links[4 * k + 3]);
      require(b <= e && e <= n, "tree: node range out of bounds");
      require(nc >= 0 && nc <= 8, "tree: bad child count");
      if (nc > 0)
        require(fc > static_cast<int64_t>(k) && static_cast<uint64_t>(fc) + nc <= n_nodes,
                "tree: children out of range");
    }
    require(n_nodes == 0 || (links[2] == 0 && static_cast<uint64_t>(links[3]) == n),
            "tree: the root must hold every body");
    std::vector<uint8_t> seen(n_nodes, 0);
    for (uint64_t k = 0; k < n_nodes; ++k)
      for (int32_t ch = 0; ch < links[4 * k + 1]; ++ch) {
        const uint64_t x = static_cast<uint64_t>(links[4 * k]) + ch;
        require(seen[x] == 0, "tree: a node has two parents");
        seen[x] = 1;
      }
    GTH_CHECK(hipSetDevice(c->device));
    hipStream_t s = c->stream;
    double *dg = c->up_geom.ensure(static_cast<size_t>(kGeomStride) * std::max<uint64_t>(n_nodes, 1));
    int4 *dl = c->up_lnk.ensure(std::max<uint64_t>(n_nodes, 1));
    double *dp = c->up_tpos.ensure(3 * std::max<uint64_t>(n, 1));
    double *dm = c->up_tmass.ensure(std::max<uint64_t>(n, 1));
    upload(dg, geom, static_cast<size_t>(kGeomStride) * n_nodes, s);
    upload(dl, reinterpret_cast<const int4 *>(links), n_nodes, s);
    upload(dp, tree_pos, 3 * n, s);
    upload(dm, tree_mass, n, s);
    GTH_CHECK(hipEventRecord(c->ev[1], s));
    run_lists(c, cfg, dg, dl, dp, dm, static_cast<uint32_t>(n), static_cast<uint32_t>(n_nodes));
    GTH_CHECK(hipEventRecord(c->ev[2], s));
    if (out && n) {
      run_eval(c, cfg);
      download(reinterpret_cast<float4 *>(out), c->out.get(), n, s);
    }
    GTH_CHECK(hipEventRecord(c->ev[3], s));
    GTH_CHECK(hipStreamSynchronize(s));
    c->tm = gt_timings{};
    c->tm.lists = elapsed_ms(c->ev[1], c->ev[2]);
    c->tm.evaluate = elapsed_ms(c->ev[2], c->ev[3]);
    c->tm.total = c->tm.lists + c->tm.evaluate;
  });
}

int gt_hip_last_info(const gt_hip_context *c, gt_lists_info *info) {
  if (!c || !info) return GT_HIP_ERR_ARGUMENT;
  *info = c->info;
  return GT_HIP_OK;
}

int gt_hip_last_timings(const gt_hip_context *c, gt_timings *t) {
  if (!c || !t) return GT_HIP_ERR_ARGUMENT;
  *t = c->tm;
  return GT_HIP_OK;
}

int gt_hip_get_tree(gt_hip_context *c, double *geom, int32_t *links, uint64_t *order) {
  if (!c) return GT_HIP_ERR_ARGUMENT;
  return guarded(c, [&] {
    const uint32_t nn = c->tree.n_nodes(), n = c->tree.n();
    hipStream_t s = c->stream;
    if (geom) download(geom, c->tree.geom(), static_cast<size_t>(kGeomStride) * nn, s);
    if (links) download(reinterpret_cast<int4 *>(links), c->tree.links(), nn, s);
    std::vector<uint32_t> o(order ? n : 0);
    if (order) download(o.data(), c->tree.order(), n, s);
    GTH_CHECK(hipStreamSynchronize(s));
    for (uint32_t k = 0; k < o.size(); ++k) order[k] = o[k];
  });
}

int gt_hip_get_lists(gt_hip_context *c, gt_group *groups, uint32_t *cell_list,
                     uint32_t *body_list, gt_float4 *bodies, gt_cell *cells, double *origin) {
  if (!c) return GT_HIP_ERR_ARGUMENT;
  return guarded(c, [&] {
    hipStream_t s = c->stream;
    const gt_lists_info &I = c->info;
    if (c->lists_n == 0) return;
    // The device lists may sit in per-group slots with gaps between groups:
    // hand them out packed, as build_group_lists lays them out.
    const uint32_t ng = static_cast<uint32_t>(I.n_groups);
    std::vector<gt_group> g(ng);
    download(g.data(), c->groups.get(), ng, s);
    GTH_CHECK(hipStreamSynchronize(s));
    uint64_t cext = 0, bext = 0;
    for (const auto &G : g) {
      cext = std::max<uint64_t>(cext, G.cell_end);
      bext = std::max<uint64_t>(bext, G.body_end);
    }
    std::vector<uint32_t> cl(cell_list ? cext : 0), bl(body_list ? bext : 0);
    if (cell_list) download(cl.data(), c->cell_list.get(), cext, s);
    if (body_list) download(bl.data(), c->body_list.get(), bext, s);
    if (bodies) download(reinterpret_cast<float4 *>(bodies), c->bodies.get(), c->lists_n, s);
    if (cells) download(reinterpret_cast<Cell *>(cells), c->cells.get(), c->lists_nn, s);
    if (origin) download(origin, c->lists_geom, 3, s);
    GTH_CHECK(hipStreamSynchronize(s));
    uint32_t nc = 0, nb = 0;
    for (uint32_t k = 0; k < ng; ++k) {
      gt_group G = g[k];
      if (cell_list)
        std::copy(cl.begin() + G.cell_begin, cl.begin() + G.cell_end, cell_list + nc);
      if (body_list)
        std::copy(bl.begin() + G.body_begin, bl.begin() + G.body_end, body_list + nb);
      const uint32_t c_len = G.cell_end - G.cell_begin, b_len = G.body_end - G.body_begin;
      G.cell_begin = nc;
      G.cell_end = nc + c_len;
      G.body_begin = nb;
      G.body_end = nb + b_len;
      nc += c_len;
      nb += b_len;
      if (groups) groups[k] = G;
    }
  });
}

int gt_hip_state_set(gt_hip_context *c, uint64_t n, const double *pos, const double *vel,
                     const double *mass) {
  if (!c) return GT_HIP_ERR_ARGUMENT;
  return guarded(c, [&] {
    if (n > 0xffffffffull) fail(GT_HIP_ERR_LIMIT, "too many particles for the 32 bit GPU indices");
    require(n == 0 || (pos && vel && mass), "array is NULL");
    GTH_CHECK(hipSetDevice(c->device));
    hipStream_t s = c->stream;
    const size_t m = std::max<uint64_t>(n, 1);
    c->pos.ensure(3 * m);
    c->vel.ensure(3 * m);
    c->mass.ensure(m);
    c->acc.ensure(3 * m);
    c->pot.ensure(m);
    upload(c->pos.get(), pos, 3 * n, s);
    upload(c->vel.get(), vel, 3 * n, s);
    upload(c->mass.get(), mass, n, s);
    GTH_CHECK(hipStreamSynchronize(s));
    c->n_state = n;
    c->state_forces = false;
  });
}

int gt_hip_state_get(gt_hip_context *c, double *pos, double *vel, double *acc, double *pot) {
  if (!c) return GT_HIP_ERR_ARGUMENT;
  return guarded(c, [&] {
    const uint64_t n = c->n_state;
    hipStream_t s = c->stream;
    if (pos) download(pos, c->pos.get(), 3 * n, s);
    if (vel) download(vel, c->vel.get(), 3 * n, s);
    if (acc) download(acc, c->acc.get(), 3 * n, s);
    if (pot) download(pot, c->pot.get(), n, s);
    GTH_CHECK(hipStreamSynchronize(s));
  });
}

int gt_hip_prime(gt_hip_context *c, const gt_group_config *cfg) {
  if (!c) return GT_HIP_ERR_ARGUMENT;
  return guarded(c, [&] {
    check_config(cfg);
    GTH_CHECK(hipSetDevice(c->device));
    if (c->n_state)
      force_eval(c, cfg, c->pos.get(), c->mass.get(), static_cast<uint32_t>(c->n_state),
                 c->acc.get(), c->pot.get());
    c->state_forces = true;
  });
}

int gt_hip_steps(gt_hip_context *c, const gt_group_config *cfg, int nsteps, double dt,
                 int use_graph) {
  if (!c) return GT_HIP_ERR_ARGUMENT;
  return guarded(c, [&] {
    check_config(cfg);
    require(nsteps >= 0, "nsteps must be non-negative");
    require(c->state_forces, "forces are not current: call gt_hip_prime first");
    (void)use_graph;
    GTH_CHECK(hipSetDevice(c->device));
    const uint64_t n = c->n_state;
    const double h = 0.5 * dt;
    for (int step = 0; step < nsteps && n; ++step) {
      launch_kick_drift(c->pos.get(), c->vel.get(), c->acc.get(), n, h, dt, c->stream);
      force_eval(c, cfg, c->pos.get(), c->mass.get(), static_cast<uint32_t>(n), c->acc.get(),
                 c->pot.get());
      launch_kick(c->vel.get(), c->acc.get(), n, h, c->stream);
    }
    GTH_CHECK(hipStreamSynchronize(c->stream));
  });
}

int gt_hip_direct_f64_sample(gt_hip_context *c, uint64_t n, const double *pos,
                             const double *mass, double G, double eps, uint64_t ns,
                             const uint64_t *sample, double *acc) {
  if (!c) return GT_HIP_ERR_ARGUMENT;
  return guarded(c, [&] {
    if (ns == 0) return;
    require(pos && mass && sample && acc, "array is NULL");
    for (uint64_t k = 0; k < ns; ++k) require(sample[k] < n, "sample index out of range");
    GTH_CHECK(hipSetDevice(c->device));
    hipStream_t s = c->stream;
    double *dp = c->in_pos.ensure(3 * n), *dm = c->in_mass.ensure(n);
    uint64_t *di = c->sample_idx.ensure(ns);
    double *da = c->sample_acc.ensure(3 * ns);
    upload(dp, pos, 3 * n, s);
    upload(dm, mass, n, s);
    upload(di, sample, ns, s);
    launch_direct_f64_sample(dp, dm, n, G, eps, di, ns, da, s);
    download(acc, da, 3 * ns, s);
    GTH_CHECK(hipStreamSynchronize(s));
  });
}

}  // extern "C"
