/*
 * hydro_omp3.c -- the oracle's restatement of flow's omp3 hydro timestep
 * (omp3/flow.c). TEST INFRASTRUCTURE ONLY (see oracle.h).
 *
 * Every expression keeps the reference's association order, every division
 * stays a division, and the file is compiled with -ffp-contract=off, so each
 * value rounds exactly as the reference's omp3 build with contraction off.
 * Loop domains and branch conditions follow the cited lines one for one.
 */
#include <math.h>
#include <stdio.h>

#include "oracle.h"

#define GAM 1.4           /* flow_interface.h:7 */
#define C_Q 3.0           /* flow_interface.h:8 */
#define C_T 0.3           /* flow_data.h:6 */
#define C_M (1.5 / C_T)   /* flow_interface.h:9 */

#define MINMOD(a, b) orc_minmod((a), (b))
#define MIN(a, b) orc_min((a), (b))

/* omp3/flow.c:55-70 -- ideal-gas pressure on every cell, ghosts included. */
void orc_equation_of_state(int nx, int ny, double* P, const double* rho,
                           const double* e) {
#pragma omp parallel for
  for (int i = 0; i < ny; ++i)
    for (int j = 0; j < nx; ++j) {
      const int c = i * nx + j;
      P[c] = (GAM - 1.0) * rho[c] * e[c];
    }
}

/* Edge-centred densities, shared by PA, AV and the momentum updates
 * (omp3/flow.c:134-141, 222-229, 585-588, 638-641). */
static inline double rho_edge_x(const double* rho, const OrcMesh* m, int nx,
                                int i, int j) {
  const int c = i * nx + j;
  return (rho[c] * m->celldx[j] * m->celldy[i] +
          rho[c - 1] * m->celldx[j - 1] * m->celldy[i]) /
         (2.0 * m->edgedx[j] * m->celldy[i]);
}
static inline double rho_edge_y(const double* rho, const OrcMesh* m, int nx,
                                int i, int j) {
  const int c = i * nx + j;
  return (rho[c] * m->celldx[j] * m->celldy[i] +
          rho[c - nx] * m->celldx[j] * m->celldy[i - 1]) /
         (2.0 * m->celldx[j] * m->edgedy[i]);
}

/* omp3/flow.c:110-158 */
void orc_pressure_acceleration(int nx, int ny, const OrcMesh* m, double dt,
                               double* mom_x, double* mom_y, double* u,
                               double* v, const double* P, const double* rho) {
  const int pad = m->pad, sx = nx + 1;
#pragma omp parallel for
  for (int i = pad; i < (ny + 1) - pad; ++i)
    for (int j = pad; j < (nx + 1) - pad; ++j) {
      const int c = i * nx + j, f = i * sx + j;
      mom_x[f] -= dt * (P[c] - P[c - 1]) / m->edgedx[j];
      mom_y[c] -= dt * (P[c] - P[c - nx]) / m->edgedy[i];
      const double rex = rho_edge_x(rho, m, nx, i, j);
      const double rey = rho_edge_y(rho, m, nx, i, j);
      u[f] = (rex == 0.0) ? 0.0 : mom_x[f] / rex;
      v[c] = (rey == 0.0) ? 0.0 : mom_y[c] / rey;
    }
  orc_handle_boundary_2d(nx + 1, ny, pad, u, ORC_INVERT_X);
  orc_handle_boundary_2d(nx, ny + 1, pad, v, ORC_INVERT_Y);
}

/* omp3/flow.c:160-245 */
void orc_artificial_viscosity(int nx, int ny, const OrcMesh* m, double dt,
                              double* Qxx, double* Qyy, double* u, double* v,
                              double* mom_x, double* mom_y, const double* rho) {
  const int pad = m->pad, sx = nx + 1;
#pragma omp parallel for
  for (int i = pad; i < ny - pad; ++i)
    for (int j = pad; j < nx - pad; ++j) {
      const int c = i * nx + j, f = i * sx + j;
      /* x: faces j-1 .. j+2 of row i */
      const double dum = MIN(0.0, u[f] - u[f - 1]);
      const double du0 = MIN(0.0, u[f + 1] - u[f]);
      const double dup = MIN(0.0, u[f + 2] - u[f + 1]);
      const double u_i = du0;
      const double u_ii = 0.5 * (fabs(dup - du0) + fabs(du0 - dum));
      /* y: faces i-1 .. i+2 of column j */
      const double dvm = MIN(0.0, v[c] - v[c - nx]);
      const double dv0 = MIN(0.0, v[c + nx] - v[c]);
      const double dvp = MIN(0.0, v[c + 2 * nx] - v[c + nx]);
      const double v_i = dv0;
      const double v_ii = 0.5 * (fabs(dvp - dv0) + fabs(dv0 - dvm));
      Qxx[c] = -C_Q * rho[c] * u_i * u_ii;
      Qyy[c] = -C_Q * rho[c] * v_i * v_ii;
    }
  orc_handle_boundary_2d(nx, ny, pad, Qxx, ORC_NO_INVERT);
  orc_handle_boundary_2d(nx, ny, pad, Qyy, ORC_NO_INVERT);

#pragma omp parallel for
  for (int i = pad; i < (ny + 1) - pad; ++i)
    for (int j = pad; j < (nx + 1) - pad; ++j) {
      const int c = i * nx + j, f = i * sx + j;
      mom_x[f] -= dt * (Qxx[c] - Qxx[c - 1]) / m->celldx[j];
      mom_y[c] -= dt * (Qyy[c] - Qyy[c - nx]) / m->celldy[i];
      const double rex = rho_edge_x(rho, m, nx, i, j);
      const double rey = rho_edge_y(rho, m, nx, i, j);
      u[f] = (rex == 0.0) ? 0.0 : mom_x[f] / rex;
      v[c] = (rey == 0.0) ? 0.0 : mom_y[c] / rey;
    }
  orc_handle_boundary_2d(nx + 1, ny, pad, u, ORC_INVERT_X);
  orc_handle_boundary_2d(nx, ny + 1, pad, v, ORC_INVERT_Y);
}

/* omp3/flow.c:248-290 */
void orc_shock_heating_and_work(int nx, int ny, const OrcMesh* m, double dt_h,
                                double* e, const double* P, const double* u,
                                const double* v, const double* rho,
                                const double* Qxx, const double* Qyy) {
  const int pad = m->pad, sx = nx + 1;
#pragma omp parallel for
  for (int i = pad; i < ny - pad; ++i)
    for (int j = pad; j < nx - pad; ++j) {
      const int c = i * nx + j, f = i * sx + j;
      const double dvx = (u[f + 1] - u[f]) / m->celldx[j];
      const double dvy = (v[c + nx] - v[c]) / m->celldy[i];
      const double div_dt = (dvx + dvy) * dt_h;
      const double e_q = e[c] - dt_h * (Qxx[c] * dvx + Qyy[c] * dvy) / rho[c];
      const double rho_c = rho[c] / (1.0 + div_dt);
      const double e_c = e_q - (P[c] * div_dt) / rho[c];
      const double work =
          0.5 * div_dt * (P[c] + (GAM - 1.0) * e_c * rho_c) / rho[c];
      e[c] = (rho[c] == 0.0) ? 0.0 : e_q - work;
    }
  orc_handle_boundary_2d(nx, ny, pad, e, ORC_NO_INVERT);
}

/* The CFL candidate of one cell, omp3/flow.c:87-94. A cell with negative
 * energy (or zero density) gives NaN: sqrt of a negative, or 0/0. */
static inline double cfl_candidate(const double* Qxx, const double* Qyy,
                                   const double* rho, const double* e,
                                   const OrcMesh* m, int nx, int i, int j) {
  const int c = i * nx + j;
  const double c_s = sqrt(GAM * (GAM - 1.0) * e[c]);
  const double dtx = m->celldx[j] / sqrt(c_s * c_s + 2.0 * Qxx[c] / rho[c]);
  const double dty = m->celldy[i] / sqrt(c_s * c_s + 2.0 * Qyy[c] / rho[c]);
  return MIN(dtx, dty);
}

static int cfl_rule = ORC_CFL_DEFINED;
static int last_undefined = 0;
void orc_set_cfl_rule(int rule) { cfl_rule = rule; }
int orc_get_cfl_rule(void) { return cfl_rule; }
int orc_last_undefined(void) { return last_undefined; }

/* omp3/flow.c:100-105: the clamp and the (dt, dt_h) recurrence. */
static void timestep_update(OrcMesh* m, double global_min, int first_step) {
  const double final_min = MIN(global_min, C_M * m->dt_h);
  m->dt = 0.5 * (C_T * final_min + m->dt_h);
  m->dt_h = first_step ? m->dt : C_T * final_min;
}

/* omp3/flow.c:73-106. One rank: reduce_all_min is the identity. Returns the
 * minimum before the clamp.
 *
 * The two rules differ only where a candidate is NaN. omp3's loop folds the
 * candidates with min(a, b) = a < b ? a : b. A NaN candidate makes the
 * running value NaN (a < NaN is false), and the next candidate replaces that
 * NaN (NaN < b is false), so every candidate before the last NaN is
 * forgotten, and a NaN in the last cell leaves NaN. Under
 * reduction(min : local_min_dt) each OpenMP thread folds its own rows from
 * +inf, and the threads' values are merged into local_min_dt = max_dt by
 * OpenMP's min combiner, out = in < out ? in : out, which passes over a
 * thread's NaN (gcc's merge does so when built for AVX, as the reference's
 * -march=native and this file's -march=x86-64-v3 build it; built for
 * baseline SSE2 it propagates the NaN instead). So with NaN candidates the
 * result depends on the thread count (orc_set_timestep_literal).
 *
 *   ORC_CFL_OMP3_SERIAL: the fold as one thread computes it: every interior
 *     candidate in row-major order from +inf, then merged with max_dt by the
 *     combiner (a fold that ends on a NaN leaves max_dt).
 *   ORC_CFL_DEFINED: the minimum over the candidates that are not NaN, and
 *     max_dt; independent of order and thread count, and equal to the serial
 *     fold whenever no candidate is NaN. */
double orc_set_timestep(int nx, int ny, const double* Qxx, const double* Qyy,
                        const double* rho, const double* e, OrcMesh* m,
                        int first_step) {
  const int pad = m->pad;
  double local_min;
  int undefined = 0;
  if (cfl_rule == ORC_CFL_OMP3_SERIAL) {
    double fold = INFINITY;
    for (int i = pad; i < ny - pad; ++i)
      for (int j = pad; j < nx - pad; ++j) {
        const double cell_min = cfl_candidate(Qxx, Qyy, rho, e, m, nx, i, j);
        undefined += (cell_min != cell_min);
        fold = MIN(fold, cell_min);
      }
    local_min = MIN(fold, m->max_dt);
  } else {
    local_min = m->max_dt;
#pragma omp parallel for reduction(min : local_min) reduction(+ : undefined)
    for (int i = pad; i < ny - pad; ++i)
      for (int j = pad; j < nx - pad; ++j) {
        const double cell_min = cfl_candidate(Qxx, Qyy, rho, e, m, nx, i, j);
        if (cell_min == cell_min)
          local_min = MIN(local_min, cell_min);
        else
          ++undefined;
      }
  }
  last_undefined = undefined;
  timestep_update(m, local_min, first_step);
  return local_min;
}

/* TEST ONLY: omp3/flow.c:77-105 with its loop as the reference writes it, a
 * min reduction over nthreads OpenMP threads with the ternary min in the
 * body. With one thread GCC's OpenMP gives ORC_CFL_OMP3_SERIAL; with more,
 * and NaN candidates, the result depends on the thread count. */
double orc_set_timestep_literal(int nx, int ny, const double* Qxx,
                                const double* Qyy, const double* rho,
                                const double* e, OrcMesh* m, int first_step,
                                int nthreads) {
  const int pad = m->pad;
  double local_min = m->max_dt;
#pragma omp parallel for num_threads(nthreads) reduction(min : local_min)
  for (int i = pad; i < ny - pad; ++i)
    for (int j = pad; j < nx - pad; ++j)
      local_min = MIN(local_min, cfl_candidate(Qxx, Qyy, rho, e, m, nx, i, j));
  timestep_update(m, local_min, first_step);
  return local_min;
}

/* omp3/flow.c:326-439 */
void orc_x_mass_and_energy_flux(int nx, int ny, int first, const OrcMesh* m,
                                double dt, double dt_h, double* rho,
                                double* rho_old, double* e, const double* u,
                                double* F_x, double* eF_x) {
  const int pad = m->pad, sx = nx + 1;
  const double *edx = m->edgedx, *cdx = m->celldx;
  const double *edy = m->edgedy, *cdy = m->celldy;
#pragma omp parallel for
  for (int i = pad; i < ny - pad; ++i)
    for (int j = pad; j < (nx + 1) - pad; ++j) {
      const int c = i * nx + j, f = i * sx + j;
      /* second-order-in-time face velocity (:344-353) */
      const double invdx = 1.0 / edx[j];
      const double s_c = 0.5 * invdx * (u[f + 1] - u[f - 1]);
      const double s_r = 2.0 * invdx * (u[f] - u[f - 1]);
      const double s_l = 2.0 * invdx * (u[f + 1] - u[f]);
      const double u_tc = u[f] - 0.5 * u[f] * dt * MINMOD(s_c, MINMOD(s_r, s_l));
      /* Van Leer limited mass flux (:356-375) */
      double limiter = 0.0;
      const double drho = (rho[c] - rho[c - 1]);
      if (drho) {
        const double r = (u_tc >= 0.0) ? (rho[c - 1] - rho[c - 2]) / drho
                                       : (rho[c + 1] - rho[c]) / drho;
        limiter = (r + fabs(r)) / (1.0 + fabs(r));
      }
      const double rho_up = (u_tc >= 0.0) ? rho[c - 1] : rho[c];
      F_x[f] = (u_tc * rho_up +
                0.5 * fabs(u_tc) * (1.0 - fabs((u_tc * dt_h) / cdx[j])) *
                    limiter * drho);
      /* MC-limited energy at the face (:378-399) */
      const double a0 = 0.5 * invdx * (e[c] - e[c - 2]);
      const double b0 = 2.0 * invdx * (e[c - 1] - e[c - 2]);
      const double c0 = 2.0 * invdx * (e[c] - e[c - 1]);
      const double a1 = 0.5 * invdx * (e[c + 1] - e[c - 1]);
      const double b1 = 2.0 * invdx * (e[c] - e[c - 1]);
      const double c1 = 2.0 * invdx * (e[c + 1] - e[c]);
      const double e_face =
          (u_tc > 0.0)
              ? e[c - 1] + 0.5 * MINMOD(MINMOD(a0, b0), c0) *
                               (cdx[j - 1] - u_tc * dt_h)
              : e[c] - 0.5 * MINMOD(MINMOD(a1, b1), c1) *
                           (cdx[j] + u_tc * dt_h);
      eF_x[f] = edy[i] * e_face * F_x[f];
    }
  orc_handle_boundary_2d(nx + 1, ny, pad, F_x, ORC_INVERT_X);

  /* conservative update (:412-434) */
#pragma omp parallel for
  for (int i = pad; i < ny - pad; ++i)
    for (int j = pad; j < nx - pad; ++j) {
      const int c = i * nx + j, f = i * sx + j;
      rho[c] -= dt_h * (edy[i + 1] * F_x[f + 1] - edy[i] * F_x[f]) /
                (cdx[j] * cdy[i]);
      const double rho_e = (rho_old[c] * e[c] -
                            (dt_h * (eF_x[f + 1] - eF_x[f])) / (cdx[j] * cdy[i]));
      e[c] = first ? ((rho_old[c] == 0.0) ? 0.0 : rho_e / rho_old[c])
                   : ((rho[c] == 0.0) ? 0.0 : rho_e / rho[c]);
    }
  orc_handle_boundary_2d(nx, ny, pad, rho, ORC_NO_INVERT);
  orc_handle_boundary_2d(nx, ny, pad, e, ORC_NO_INVERT);
}

/* omp3/flow.c:442-554. Note the Van Leer branch tests the face velocity v,
 * not v_tc (:477), and the update uses edgedx[j+1] (:533). */
void orc_y_mass_and_energy_flux(int nx, int ny, int first, const OrcMesh* m,
                                double dt, double dt_h, double* rho,
                                double* rho_old, double* e, const double* v,
                                double* F_y, double* eF_y) {
  const int pad = m->pad;
  const double *edx = m->edgedx, *cdx = m->celldx;
  const double *edy = m->edgedy, *cdy = m->celldy;
#pragma omp parallel for
  for (int i = pad; i < (ny + 1) - pad; ++i)
    for (int j = pad; j < nx - pad; ++j) {
      const int c = i * nx + j;
      const double invdy = 1.0 / edy[i];
      const double s_c = 0.5 * invdy * (v[c + nx] - v[c - nx]);
      const double s_r = 2.0 * invdy * (v[c] - v[c - nx]);
      const double s_l = 2.0 * invdy * (v[c + nx] - v[c]);
      const double v_tc = v[c] - 0.5 * v[c] * dt * MINMOD(s_c, MINMOD(s_r, s_l));
      const double drho = (rho[c] - rho[c - nx]);
      double limiter = 0.0;
      if (drho) {
        const double r = (v[c] >= 0.0) ? (rho[c - nx] - rho[c - 2 * nx]) / drho
                                       : (rho[c + nx] - rho[c]) / drho;
        limiter = (r + fabs(r)) / (1.0 + fabs(r));
      }
      const double rho_up = (v_tc >= 0.0) ? rho[c - nx] : rho[c];
      F_y[c] = (v_tc * rho_up +
                0.5 * fabs(v_tc) * (1.0 - fabs((v_tc * dt_h) / cdy[i])) *
                    limiter * drho);
      const double a0 = 0.5 * invdy * (e[c] - e[c - 2 * nx]);
      const double b0 = 2.0 * invdy * (e[c - nx] - e[c - 2 * nx]);
      const double c0 = 2.0 * invdy * (e[c] - e[c - nx]);
      const double a1 = 0.5 * invdy * (e[c + nx] - e[c - nx]);
      const double b1 = 2.0 * invdy * (e[c] - e[c - nx]);
      const double c1 = 2.0 * invdy * (e[c + nx] - e[c]);
      const double e_face =
          (v_tc > 0.0)
              ? e[c - nx] + 0.5 * MINMOD(MINMOD(a0, b0), c0) *
                                (cdy[i - 1] - v_tc * dt_h)
              : e[c] - 0.5 * MINMOD(MINMOD(a1, b1), c1) *
                           (cdy[i] + v_tc * dt_h);
      eF_y[c] = edx[j] * e_face * F_y[c];
    }
  orc_handle_boundary_2d(nx, ny + 1, pad, F_y, ORC_INVERT_Y);

#pragma omp parallel for
  for (int i = pad; i < ny - pad; ++i)
    for (int j = pad; j < nx - pad; ++j) {
      const int c = i * nx + j;
      rho[c] -= dt_h * (edx[j + 1] * F_y[c + nx] - edx[j] * F_y[c]) /
                (cdx[j] * cdy[i]);
      const double rho_e = (rho_old[c] * e[c] -
                            (dt_h * (eF_y[c + nx] - eF_y[c])) / (cdx[j] * cdy[i]));
      e[c] = first ? ((rho_old[c] == 0.0) ? 0.0 : rho_e / rho_old[c])
                   : ((rho[c] == 0.0) ? 0.0 : rho_e / rho[c]);
    }
  orc_handle_boundary_2d(nx, ny, pad, rho, ORC_NO_INVERT);
  orc_handle_boundary_2d(nx, ny, pad, e, ORC_NO_INVERT);
}

/* omp3/flow.c:293-323: copy rho over every cell, then split sweeps whose
 * order alternates with the parity of tt; the first sweep has first=1. */
void orc_advect_mass_and_energy(int nx, int ny, const OrcMesh* m, int tt,
                                double dt, double dt_h, double* rho, double* e,
                                double* rho_old, double* F_x, double* F_y,
                                double* eF_x, double* eF_y, const double* u,
                                const double* v) {
#pragma omp parallel for
  for (int k = 0; k < nx * ny; ++k) rho_old[k] = rho[k];
  if (tt % 2 == 0) {
    orc_x_mass_and_energy_flux(nx, ny, 1, m, dt, dt_h, rho, rho_old, e, u, F_x, eF_x);
    orc_y_mass_and_energy_flux(nx, ny, 0, m, dt, dt_h, rho, rho_old, e, v, F_y, eF_y);
  } else {
    orc_y_mass_and_energy_flux(nx, ny, 1, m, dt, dt_h, rho, rho_old, e, v, F_y, eF_y);
    orc_x_mass_and_energy_flux(nx, ny, 0, m, dt, dt_h, rho, rho_old, e, u, F_x, eF_x);
  }
}

/* omp3/flow.c:778-823 -- cell-centred x-momentum flux in x (layout C). */
void orc_momentum_x_flux_in_x(int nx, int ny, const OrcMesh* m, double dt_h,
                              const double* u, double* uF_x, const double* F_x) {
  const int pad = m->pad, sx = nx + 1;
#pragma omp parallel for
  for (int i = pad; i < ny - pad; ++i)
    for (int j = pad; j < nx - pad; ++j) {
      const int c = i * nx + j, f = i * sx + j;
      const double invdx = 1.0 / m->edgedx[j];
      const double a0 = 0.5 * invdx * (u[f + 1] - u[f - 1]);
      const double b0 = 2.0 * invdx * (u[f] - u[f - 1]);
      const double c0 = 2.0 * invdx * (u[f + 1] - u[f]);
      const double a1 = 0.5 * invdx * (u[f + 2] - u[f]);
      const double b1 = 2.0 * invdx * (u[f + 1] - u[f]);
      const double c1 = 2.0 * invdx * (u[f + 2] - u[f + 1]);
      const double u_cell = 0.5 * (u[f] + u[f + 1]);
      const double Fc = m->edgedy[i] * 0.5 * (F_x[f] + F_x[f + 1]);
      const double u_int =
          (u_cell > 0.0)
              ? u[f] + 0.5 * MINMOD(MINMOD(a0, b0), c0) *
                           (m->celldx[j - 1] - u_cell * dt_h)
              : u[f + 1] - 0.5 * MINMOD(MINMOD(a1, b1), c1) *
                               (m->celldx[j] + u_cell * dt_h);
      uF_x[c] = Fc * u_int;
    }
}

/* omp3/flow.c:826-874 -- corner x-momentum flux in y (layout K). */
void orc_momentum_x_flux_in_y(int nx, int ny, const OrcMesh* m, double dt_h,
                              const double* u, const double* v, double* uF_y,
                              const double* F_y) {
  const int pad = m->pad, sx = nx + 1;
#pragma omp parallel for
  for (int i = pad; i < (ny + 1) - pad; ++i)
    for (int j = pad; j < (nx + 1) - pad; ++j) {
      const int c = i * nx + j, f = i * sx + j;
      const double invdy = 1.0 / m->edgedy[i];
      const double a0 = 0.5 * invdy * (u[f] - u[f - 2 * sx]);
      const double b0 = 2.0 * invdy * (u[f - sx] - u[f - 2 * sx]);
      const double c0 = 2.0 * invdy * (u[f] - u[f - sx]);
      const double a1 = 0.5 * invdy * (u[f + sx] - u[f - sx]);
      const double b1 = 2.0 * invdy * (u[f] - u[f - sx]);
      const double c1 = 2.0 * invdy * (u[f + sx] - u[f]);
      const double v_cell = 0.5 * (v[c - 1] + v[c]);
      const double Fc = m->edgedx[j] * 0.5 * (F_y[c] + F_y[c - 1]);
      const double u_corner =
          (v_cell > 0.0)
              ? u[f - sx] + 0.5 * MINMOD(MINMOD(a0, b0), c0) *
                                (m->celldy[i - 1] - v_cell * dt_h)
              : u[f] - 0.5 * MINMOD(MINMOD(a1, b1), c1) *
                           (m->celldy[i] + v_cell * dt_h);
      uF_y[f] = Fc * u_corner;
    }
}

/* omp3/flow.c:877-925 -- corner y-momentum flux in x (layout K). */
void orc_momentum_y_flux_in_x(int nx, int ny, const OrcMesh* m, double dt_h,
                              const double* u, const double* v, double* vF_x,
                              const double* F_x) {
  const int pad = m->pad, sx = nx + 1;
#pragma omp parallel for
  for (int i = pad; i < (ny + 1) - pad; ++i)
    for (int j = pad; j < (nx + 1) - pad; ++j) {
      const int c = i * nx + j, f = i * sx + j;
      const double invdx = 1.0 / m->edgedx[j];
      const double a0 = 0.5 * invdx * (v[c] - v[c - 2]);
      const double b0 = 2.0 * invdx * (v[c - 1] - v[c - 2]);
      const double c0 = 2.0 * invdx * (v[c] - v[c - 1]);
      const double a1 = 0.5 * invdx * (v[c + 1] - v[c - 1]);
      const double b1 = 2.0 * invdx * (v[c] - v[c - 1]);
      const double c1 = 2.0 * invdx * (v[c + 1] - v[c]);
      const double Fc = m->celldy[i] * 0.5 * (F_x[f] + F_x[f - sx]);
      const double u_cell = 0.5 * (u[f] + u[f - sx]);
      const double v_int =
          (u_cell > 0.0)
              ? v[c - 1] + 0.5 * MINMOD(MINMOD(a0, b0), c0) *
                               (m->celldx[j - 1] - u_cell * dt_h)
              : v[c] - 0.5 * MINMOD(MINMOD(a1, b1), c1) *
                           (m->celldx[j] + u_cell * dt_h);
      vF_x[f] = Fc * v_int;
    }
}

/* omp3/flow.c:928-970 -- cell-centred y-momentum flux in y (layout C). */
void orc_momentum_y_flux_in_y(int nx, int ny, const OrcMesh* m, double dt_h,
                              const double* v, double* vF_y, const double* F_y) {
  const int pad = m->pad;
#pragma omp parallel for
  for (int i = pad; i < ny - pad; ++i)
    for (int j = pad; j < nx - pad; ++j) {
      const int c = i * nx + j;
      const double invdy = 1.0 / m->edgedy[i];
      const double a0 = 0.5 * invdy * (v[c + nx] - v[c - nx]);
      const double b0 = 2.0 * invdy * (v[c] - v[c - nx]);
      const double c0 = 2.0 * invdy * (v[c + nx] - v[c]);
      const double a1 = 0.5 * invdy * (v[c + 2 * nx] - v[c]);
      const double b1 = 2.0 * invdy * (v[c + nx] - v[c]);
      const double c1 = 2.0 * invdy * (v[c + 2 * nx] - v[c + nx]);
      const double Fc = m->celldx[j] * 0.5 * (F_y[c] + F_y[c + nx]);
      const double v_cell = 0.5 * (v[c] + v[c + nx]);
      const double v_int =
          (v_cell > 0.0)
              ? v[c] + 0.5 * MINMOD(MINMOD(a0, b0), c0) *
                           (m->celldy[i - 1] - v_cell * dt_h)
              : v[c + nx] - 0.5 * MINMOD(MINMOD(a1, b1), c1) *
                                (m->celldy[i] + v_cell * dt_h);
      vF_y[c] = Fc * v_int;
    }
}

/* The four momentum sub-updates of omp3/flow.c:557-775, named by what they
 * apply. Each is a loop body shared by both parities. */
static void mom_x_from_uF_x(int nx, int ny, const OrcMesh* m, double dt_h,
                            double* mom_x, double* u, const double* uF_x,
                            const double* rho, int recompute_u) {
  const int pad = m->pad, sx = nx + 1;
#pragma omp parallel for
  for (int i = pad; i < ny - pad; ++i)
    for (int j = pad; j < (nx + 1) - pad; ++j) {
      const int c = i * nx + j, f = i * sx + j;
      mom_x[f] -= dt_h * (uF_x[c] - uF_x[c - 1]) / (m->edgedx[j] * m->celldy[i]);
      if (recompute_u) {
        const double rex = rho_edge_x(rho, m, nx, i, j);
        u[f] = (rex == 0.0) ? 0.0 : mom_x[f] / rex;
      }
    }
}
static void mom_x_from_uF_y(int nx, int ny, const OrcMesh* m, double dt_h,
                            double* mom_x, double* u, const double* uF_y,
                            const double* rho, int recompute_u) {
  const int pad = m->pad, sx = nx + 1;
#pragma omp parallel for
  for (int i = pad; i < ny - pad; ++i)
    for (int j = pad; j < (nx + 1) - pad; ++j) {
      const int f = i * sx + j;
      mom_x[f] -= dt_h * (uF_y[f + sx] - uF_y[f]) / (m->celldx[j] * m->edgedy[i]);
      if (recompute_u) {
        const double rex = rho_edge_x(rho, m, nx, i, j);
        u[f] = (rex == 0.0) ? 0.0 : mom_x[f] / rex;
      }
    }
}
static void mom_y_from_vF_x(int nx, int ny, const OrcMesh* m, double dt_h,
                            double* mom_y, double* v, const double* vF_x,
                            const double* rho, int recompute_v) {
  const int pad = m->pad, sx = nx + 1;
#pragma omp parallel for
  for (int i = pad; i < (ny + 1) - pad; ++i)
    for (int j = pad; j < nx - pad; ++j) {
      const int c = i * nx + j, f = i * sx + j;
      mom_y[c] -= dt_h * (vF_x[f + 1] - vF_x[f]) / (m->edgedx[j] * m->celldy[i]);
      if (recompute_v) {
        const double rey = rho_edge_y(rho, m, nx, i, j);
        v[c] = (rey == 0.0) ? 0.0 : mom_y[c] / rey;
      }
    }
}
static void mom_y_from_vF_y(int nx, int ny, const OrcMesh* m, double dt_h,
                            double* mom_y, double* v, const double* vF_y,
                            const double* rho, int recompute_v) {
  const int pad = m->pad;
#pragma omp parallel for
  for (int i = pad; i < (ny + 1) - pad; ++i)
    for (int j = pad; j < nx - pad; ++j) {
      const int c = i * nx + j;
      mom_y[c] -= dt_h * (vF_y[c] - vF_y[c - nx]) / (m->celldx[j] * m->edgedy[i]);
      if (recompute_v) {
        const double rey = rho_edge_y(rho, m, nx, i, j);
        v[c] = (rey == 0.0) ? 0.0 : mom_y[c] / rey;
      }
    }
}

/* omp3/flow.c:557-775. Odd tt: x-flux-in-x first; even tt: x-flux-in-y
 * first. Velocities are recomputed only after the first sub-update of each
 * momentum component, and every flux and recomputed velocity is reflected. */
void orc_advect_momentum(int nx, int ny, int tt, const OrcMesh* m, double dt_h,
                         double dt, double* u, double* v, double* uF_x,
                         double* uF_y, double* vF_x, double* vF_y,
                         double* mom_x, double* mom_y, const double* rho,
                         const double* F_x, const double* F_y) {
  (void)dt;
  const int pad = m->pad;
  if (tt % 2) {
    orc_momentum_x_flux_in_x(nx, ny, m, dt_h, u, uF_x, F_x);
    orc_handle_boundary_2d(nx, ny, pad, uF_x, ORC_NO_INVERT);
    mom_x_from_uF_x(nx, ny, m, dt_h, mom_x, u, uF_x, rho, 1);
    orc_handle_boundary_2d(nx + 1, ny, pad, u, ORC_INVERT_X);
    orc_momentum_x_flux_in_y(nx, ny, m, dt_h, u, v, uF_y, F_y);
    orc_handle_boundary_2d(nx + 1, ny + 1, pad, uF_y, ORC_NO_INVERT);
    mom_x_from_uF_y(nx, ny, m, dt_h, mom_x, u, uF_y, rho, 0);
    orc_momentum_y_flux_in_x(nx, ny, m, dt_h, u, v, vF_x, F_x);
    orc_handle_boundary_2d(nx + 1, ny + 1, pad, vF_x, ORC_NO_INVERT);
    mom_y_from_vF_x(nx, ny, m, dt_h, mom_y, v, vF_x, rho, 1);
    orc_handle_boundary_2d(nx, ny + 1, pad, v, ORC_INVERT_Y);
    orc_momentum_y_flux_in_y(nx, ny, m, dt_h, v, vF_y, F_y);
    orc_handle_boundary_2d(nx, ny, pad, vF_y, ORC_NO_INVERT);
    mom_y_from_vF_y(nx, ny, m, dt_h, mom_y, v, vF_y, rho, 0);
  } else {
    orc_momentum_x_flux_in_y(nx, ny, m, dt_h, u, v, uF_y, F_y);
    orc_handle_boundary_2d(nx + 1, ny + 1, pad, uF_y, ORC_NO_INVERT);
    mom_x_from_uF_y(nx, ny, m, dt_h, mom_x, u, uF_y, rho, 1);
    orc_handle_boundary_2d(nx + 1, ny, pad, u, ORC_INVERT_X);
    orc_momentum_x_flux_in_x(nx, ny, m, dt_h, u, uF_x, F_x);
    orc_handle_boundary_2d(nx, ny, pad, uF_x, ORC_NO_INVERT);
    mom_x_from_uF_x(nx, ny, m, dt_h, mom_x, u, uF_x, rho, 0);
    orc_momentum_y_flux_in_y(nx, ny, m, dt_h, v, vF_y, F_y);
    orc_handle_boundary_2d(nx, ny, pad, vF_y, ORC_NO_INVERT);
    mom_y_from_vF_y(nx, ny, m, dt_h, mom_y, v, vF_y, rho, 1);
    orc_handle_boundary_2d(nx, ny + 1, pad, v, ORC_INVERT_Y);
    orc_momentum_y_flux_in_x(nx, ny, m, dt_h, u, v, vF_x, F_x);
    orc_handle_boundary_2d(nx + 1, ny + 1, pad, vF_x, ORC_NO_INVERT);
    mom_y_from_vF_x(nx, ny, m, dt_h, mom_y, v, vF_x, rho, 0);
  }
}

/* omp3/flow.c:973-995 -- interior mass and rho*e totals. */
void orc_conservation(int nx, int ny, int pad, const double* rho,
                      const double* e, double* mass, double* energy) {
  double mt = 0.0, et = 0.0;
#pragma omp parallel for reduction(+ : mt, et)
  for (int i = pad; i < ny - pad; ++i)
    for (int j = pad; j < nx - pad; ++j) {
      mt += rho[i * nx + j];
      et += rho[i * nx + j] * e[i * nx + j];
    }
  *mass = mt;
  *energy = et;
}

/* omp3/flow.c:10-52. eF_x/eF_y borrow the uF_x/uF_y buffers (:43). */
void orc_solve_hydro_2d(OrcMesh* m, int tt, double* P, double* rho,
                        double* rho_old, double* e, double* u, double* v,
                        double* mom_x, double* mom_y, double* Qxx, double* Qyy,
                        double* F_x, double* F_y, double* uF_x, double* uF_y,
                        double* vF_x, double* vF_y) {
  const int nx = m->nx, ny = m->ny;
  if (m->verbose) printf("Timestep:        %.12e\n", m->dt);
  orc_equation_of_state(nx, ny, P, rho, e);
  orc_pressure_acceleration(nx, ny, m, m->dt, mom_x, mom_y, u, v, P, rho);
  orc_artificial_viscosity(nx, ny, m, m->dt, Qxx, Qyy, u, v, mom_x, mom_y, rho);
  orc_shock_heating_and_work(nx, ny, m, m->dt_h, e, P, u, v, rho, Qxx, Qyy);
  orc_set_timestep(nx, ny, Qxx, Qyy, rho, e, m, tt == 0);
  orc_advect_mass_and_energy(nx, ny, m, tt, m->dt, m->dt_h, rho, e, rho_old,
                             F_x, F_y, uF_x, uF_y, u, v);
  orc_advect_momentum(nx, ny, tt, m, m->dt_h, m->dt, u, v, uF_x, uF_y, vF_x,
                      vF_y, mom_x, mom_y, rho, F_x, F_y);
}
