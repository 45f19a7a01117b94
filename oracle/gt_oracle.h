/* CPU oracle for the grouped Barnes-Hut gravity path: a plain restatement of
 * the reference algorithms, used ONLY as a checker by tests/, by
 * __graft_entry__.smoke() and by bench.py's cpu_baseline leg.  Nothing in the
 * product library links or calls it.
 *
 * The same C interface is exported by two libraries:
 *   oracle/libgt_oracle.so   prefix gto_   (this restatement, oracle/gt_oracle.cpp)
 *   oracle/_ref/libgt_ref.so prefix gtref_ (the reference's own sources, built by
 *                                           oracle/Makefile, plus oracle/ref_shim.cpp)
 * so a test can run both on the same input and compare results bit for bit.
 *
 * Layouts (all caller-allocated):
 *   pos      n x 3 doubles (x, y, z), particle order
 *   geom     n_nodes x 16 doubles: center xyz, half, com xyz, mass,
 *            quad xx yy zz xy xz yz, delta, 0
 *   links    n_nodes x 4 int32: first_child, n_children, begin, end
 *   groups   n_groups x 8 uint32 (GTGroup: begin end cell_begin cell_end
 *            body_begin body_end 0 0)
 *   info     10 uint64: n, n_nodes, depth, n_groups, n_cell_list, n_body_list,
 *            cell_interactions, body_interactions, max_list, max_group
 */
#ifndef GT_ORACLE_H
#define GT_ORACLE_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct gto_run gto_run;

gto_run *gto_new(void);
void gto_free(gto_run *r);
/* Octree::build (src/octree.cpp:169-273) with serial node numbering. */
int gto_build(gto_run *r, uint64_t n, const double *pos, const double *mass, int leaf_size,
              uint64_t parallel_threshold);
/* build_group_lists (src/group_lists.cpp:18-207) on the tree held by r. */
int gto_lists(gto_run *r, double theta, int group_size, int pack);
void gto_info(const gto_run *r, uint64_t *info);
void gto_nodes(const gto_run *r, double *geom, int32_t *links);
void gto_order(const gto_run *r, uint64_t *order);
void gto_tree_positions(const gto_run *r, double *pos, double *mass);
void gto_lists_get(const gto_run *r, uint32_t *groups, uint32_t *cell_list, uint32_t *body_list,
                   float *bodies, float *cells, double *origin);
/* evaluate_group_lists_cpu (src/gravity_backends.cpp:70-84): out n x 4 floats, tree order. */
void gto_eval(const gto_run *r, double G, double eps, int quadrupole, float *out);
/* Octree::accel_pot per particle = compute_gravity_tree (src/gravity.cpp:7-26). */
void gto_tree_gravity(gto_run *r, uint64_t n, const double *pos, const double *mass, double G,
                      double theta, double eps, int quadrupole, int leaf_size, double *acc,
                      double *pot, uint64_t *stats);
/* Octree::accel_pot at np points, nobody left out, on the last built tree. */
void gto_accel_at(const gto_run *r, uint64_t np, const double *points, double G, double theta,
                  double eps, int quadrupole, double *acc, double *pot);

/* compute_gravity_direct (src/gravity.cpp:28-45), FP64. */
void gto_direct_f64(uint64_t n, const double *pos, const double *mass, double G, double eps,
                    double *acc, double *pot);
/* FP64 direct on a sample of targets (apps/galactic_tides_bench.cpp:80-95): acc only. */
void gto_direct_f64_sample(uint64_t n, const double *pos, const double *mass, double G,
                           double eps, uint64_t ns, const uint64_t *sample, double *acc);
/* gt_direct_accel for every target (include/.../gravity_kernels.h:173-182). */
void gto_direct_f32(uint64_t n, const float *bodies, double G, double eps, float *out);
/* pack_bodies (src/gravity_backends.cpp:15-23). */
void gto_pack_bodies(uint64_t n, const double *pos, const double *mass, const double *origin,
                     float *out);
/* GroupTreeGravity::compute on the cpu device (src/gravity_backends.cpp:109-137). */
int gto_group_gravity(gto_run *r, uint64_t n, const double *pos, const double *mass, double G,
                      double theta, double eps, int quadrupole, int group_size, int leaf_size,
                      int pack, double *acc, double *pot);
/* LeapfrogKDK::prime + nsteps x LeapfrogKDK::step (src/integrator.cpp:5-16) with
 * GroupTreeGravity on the cpu device.  pos/vel are updated in place. */
int gto_kdk_group(gto_run *r, uint64_t n, double *pos, double *vel, const double *mass,
                  double G, double theta, double eps, int quadrupole, int group_size,
                  int leaf_size, int pack, int nsteps, double dt, double *acc, double *pot);
/* Wall time of `repeat` GroupTreeGravity::compute calls after one warm-up: best4[0]
 * the best, [1..3] its tree / lists / evaluate split (seconds), [4..5] its cell and
 * body interactions (best4 holds 6 doubles). */
int gto_time_group(gto_run *r, uint64_t n, const double *pos, const double *mass, double G,
                   double theta, double eps, int quadrupole, int group_size, int leaf_size,
                   int pack, int repeat, double *best4);
/* Octree::for_each_within (include/galactic_tides/octree.h:103-127) on the tree
 * of the last gto_build, for nq queries: query q has centre centres[3q..3q+2] and
 * radius radii[q] (n_radii = nq) or radii[0] (n_radii = 1).  offsets (nq + 1)
 * gets the CSR offsets; the neighbours (original index, squared distance) in
 * the order the reference's callback receives them go to index / dist2 when
 * both are non-NULL and the total fits cap.  Returns the total. */
uint64_t gto_within(const gto_run *r, uint64_t nq, const double *centres, const double *radii,
                    uint64_t n_radii, uint64_t *offsets, uint64_t cap, uint64_t *index,
                    double *dist2);
int gto_max_threads(void);
/* Seeded test systems of the reference's tests (std::mt19937_64 +
 * uniform_real_distribution<double>):
 *   kind 0: Hernquist sphere with two satellites (tests/test_gpu_kernels.cpp:29-47,
 *           tests/test_octree.cpp:19-41, tests/test_metal.cpp:22-40), mass 1/n;
 *   kind 1: Hernquist sphere with every 7th particle offset by (3,1,0)
 *           (tests/test_metal_mock.cpp:19-31), mass 1/n. */
void gto_test_system(int kind, uint64_t n, uint64_t seed, double *pos, double *mass);

#ifdef __cplusplus
}
#endif
#endif
