/*
 * driver_loop.c -- the oracle's restatement of flow's time loop (main.c:64-115)
 * without its I/O. TEST INFRASTRUCTURE ONLY (see oracle.h).
 */
#include <omp.h>
#include <stdio.h>

#include "oracle.h"

/* dt_seq receives 2*(niters+1) values: (dt, dt_h) after the pre-loop
 * set_timestep (main.c:64-66), then (dt, dt_h) after each step. mass_seq and
 * energy_seq receive the per-step totals print_conservation would print
 * (main.c:90-91). The wall time covers solve + conservation of every step,
 * as main.c:80-93 times it. */
int orc_run(OrcMesh* m, double** a, int niters, double* dt_seq,
            double* mass_seq, double* energy_seq, double* wall_seconds,
            int* undefined_seq) {
  double *P = a[0], *rho = a[1], *rho_old = a[2], *e = a[3], *u = a[4],
         *v = a[5], *mom_x = a[6], *mom_y = a[7], *Qxx = a[8], *Qyy = a[9],
         *F_x = a[10], *F_y = a[11], *uF_x = a[12], *uF_y = a[13],
         *vF_x = a[14], *vF_y = a[15];
  orc_set_timestep(m->nx, m->ny, Qxx, Qyy, rho, e, m, 0);
  if (dt_seq) {
    dt_seq[0] = m->dt;
    dt_seq[1] = m->dt_h;
  }
  if (undefined_seq) undefined_seq[0] = orc_last_undefined();
  double wall = 0.0, elapsed = 0.0;
  int tt;
  for (tt = 0; tt < niters; ++tt) {
    if (m->verbose) printf("\nIteration %d\n", tt + 1);
    const double w0 = omp_get_wtime();
    orc_solve_hydro_2d(m, tt, P, rho, rho_old, e, u, v, mom_x, mom_y, Qxx,
                       Qyy, F_x, F_y, uF_x, uF_y, vF_x, vF_y);
    double mass, energy;
    orc_conservation(m->nx, m->ny, m->pad, rho, e, &mass, &energy);
    if (m->verbose) {
      printf("Total mass:      %.12e\n", mass);
      printf("Total energy:    %.12e\n", energy);
    }
    wall += omp_get_wtime() - w0;
    if (dt_seq) {
      dt_seq[2 * (tt + 1)] = m->dt;
      dt_seq[2 * (tt + 1) + 1] = m->dt_h;
    }
    if (mass_seq) mass_seq[tt] = mass;
    if (undefined_seq) undefined_seq[tt + 1] = orc_last_undefined();
    if (energy_seq) energy_seq[tt] = energy;
    elapsed += m->dt;
    if (elapsed >= m->sim_end) {
      ++tt;
      break;
    }
  }
  if (wall_seconds) *wall_seconds = wall;
  return tt;
}
