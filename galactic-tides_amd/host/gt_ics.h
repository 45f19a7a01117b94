/* Initial conditions for the benchmark configurations (host C, no GPU):
 * Hernquist (1993) galaxies and parabolic two-galaxy encounters.  Positions
 * are those of the reference's generator (src/hernquist_ics.cpp); velocities
 * follow a simpler analytic recipe (see gt_ics.cpp). */
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
} gt_ics_model;

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
/* The force-error sample of the reference's benchmark
 * (apps/galactic_tides_bench.cpp:74-78): the indices 0 .. n-1 put through
 * std::shuffle with std::mt19937_64(seed), the first min(n, nsample) of them
 * written to out (when not NULL); returns that count.  The benchmark uses seed 1. */
uint64_t gt_ics_bench_sample(uint64_t n, uint64_t seed, uint64_t nsample, uint64_t *out);

#ifdef __cplusplus
}
#endif
#endif
