/* Initial conditions for the benchmark configurations (host C, no GPU):
 * Hernquist (1993) galaxies and parabolic two-galaxy encounters.  Positions
 * are those of the reference's generator (src/hernquist_ics.cpp).  Velocities
 * are either measured as the reference measures them, with a tree solver the
 * caller supplies (gt_ics_*_measured; the library itself needs no GPU), or
 * follow a simpler analytic recipe (gt_ics_galaxy / gt_ics_encounter). */
#ifndef GT_ICS_H
#define GT_ICS_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* HernquistModel (include/galactic_tides/hernquist_ics.h:52-82), numeric fields. */
typedef struct {
  double G;
  uint64_t disk_n;
  double disk_mass, disk_h, disk_z0, disk_rmax, disk_Q, disk_Q_radius;
  uint64_t bulge_n;
  double bulge_mass, bulge_a, bulge_rmax;
  uint64_t halo_n;
  double halo_mass, halo_core, halo_cutoff, halo_rmax;
  int halo_profile; /* 0: the reference's cored, cut-off halo; 1: Hernquist halo of scale halo_a */
  double halo_a;
  uint64_t seed;
  /* The velocity stage (measured velocities only): */
  double vesc_fraction;    /* spheroid speeds above this fraction of escape are redrawn */
  int preserve_dispersion; /* widen truncated spheroid Gaussians to keep <v^2> */
  int disk_sigma_z_jeans;  /* disk sigma_z from the vertical Jeans equation (else the sheet) */
  double eps, theta;       /* softening and opening angle of the tree that measures the */
  int quadrupole;          /* realised potential (HernquistModel::gravity; G is G above) */
} gt_ics_model;

/* The FP64 tree solver the velocity stage measures the realised potential
 * with (compute_gravity_tree and Octree::accel_pot, src/gravity.cpp:7-26,
 * src/octree.cpp:275-344), supplied by the caller: on the tree of the n
 * particles (leaf size 8; G, eps, theta and quadrupole of the model), either
 * (np = 0) the potential at every particle, the particle itself left out, into
 * pot (n) and acc (n x 3), or (np > 0) at the np points with nobody left out,
 * into acc (np x 3) and pot (np).  Returns 0 on success. */
typedef int (*gt_ics_tree_fn)(void *user, const gt_ics_model *m, uint64_t n, const double *pos,
                              const double *mass, uint64_t np, const double *points,
                              double *acc, double *pot);

void gt_ics_default_model(gt_ics_model *m);
/* make_hernquist_galaxy + orient (apps/galactic_tides_ics.cpp:73-77).  Writes
 * disk_n + bulge_n + halo_n particles (halo, then disk, then bulge) when pos is
 * not NULL; returns the count.  type: 1 halo, 2 disk, 3 bulge (GADGET). */
uint64_t gt_ics_galaxy(const gt_ics_model *m, double incl, double omega, double *pos,
                       double *vel, double *mass, uint8_t *type);
/* Encounter mode (apps/galactic_tides_ics.cpp:78-97): galaxy 2 uses seed + 1. */
uint64_t gt_ics_encounter(const gt_ics_model *m, double g1_incl, double g1_omega,
                          double g2_incl, double g2_omega, double rp, double r0, double *pos,
                          double *vel, double *mass, uint8_t *type);
/* The same with the reference's velocities (src/hernquist_ics.cpp:206-388):
 * spheroids from the spherical Jeans equation in the realised potential
 * (tree accelerations averaged over a Fibonacci sphere of directions), with
 * escape-speed rejection against each particle's tree potential, and the disk
 * from the tree's rotation curve and epicyclic frequency (Toomre Q at
 * disk_Q_radius), drawn from the same random stream as the positions: with
 * the reference's tree solver bit for bit, the particles equal
 * make_hernquist_galaxy's (positions, velocities, masses).  Returns 0 when the
 * tree callback fails. */
uint64_t gt_ics_galaxy_measured(const gt_ics_model *m, double incl, double omega,
                                gt_ics_tree_fn tree, void *user, double *pos, double *vel,
                                double *mass, uint8_t *type);
uint64_t gt_ics_encounter_measured(const gt_ics_model *m, double g1_incl, double g1_omega,
                                   double g2_incl, double g2_omega, double rp, double r0,
                                   gt_ics_tree_fn tree, void *user, double *pos, double *vel,
                                   double *mass, uint8_t *type);
/* The force-error sample of the reference's benchmark
 * (apps/galactic_tides_bench.cpp:74-78): the indices 0 .. n-1 put through
 * std::shuffle with std::mt19937_64(seed), the first min(n, nsample) of them
 * written to out (when not NULL); returns that count.  The benchmark uses seed 1. */
uint64_t gt_ics_bench_sample(uint64_t n, uint64_t seed, uint64_t nsample, uint64_t *out);

#ifdef __cplusplus
}
#endif
#endif
