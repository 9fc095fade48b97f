/*
 * fused.c -- the fused schedule: one step of omp3/flow.c:10-52 in four
 * kernels. Arrays omp3 uses only inside a step serve as the ping-pong
 * buffers that keep each kernel from writing what it reads:
 *   P        energy after shock heating
 *   rho_old  density before advection (omp3/flow.c:302-305 copies it too)
 *   Qxx, Qyy momenta after the viscous step
 *   uF_x, uF_y velocities after the viscous step
 * After the step rho, e, u, v, mom_x, mom_y, F_x, F_y and rho_old hold what
 * omp3 leaves in them; P, Qxx, Qyy and the four momentum-flux arrays hold
 * the values above instead of omp3's intermediates.
 */
#include "schedules.h"

void flowhip_step_fused(FlowHipCtx* h, const FlowHipStep* st) {
  const FkMesh m = st->m;
  hipStream_t s = h->stream;
  double* e1 = st->P;
  double* rho0 = st->rho_old;
  double* mx1 = st->Qxx;
  double* my1 = st->Qyy;
  double* u0 = st->uF_x;
  double* v0 = st->uF_y;
  /* under the serial CFL rule the Lagrangian kernel stores its candidates
   * and fk_timestep_serial folds them in omp3's order */
  const FkReduce lr = flowhip_lagrangian_red(h, m.nx, m.ny);
  if (h->schedule == FLOWHIP_MARCH)
    FK_LAUNCH(h, "lagrangian",
              fk_march_lagrangian(m, st->rho, st->e, st->mom_x, st->mom_y, mx1,
                                  my1, u0, v0, e1, rho0, lr, st->max_dt,
                                  st->tt == 0, flowhip_rows(h, FLOWHIP_K_LAGRANGIAN, 0),
                                  s));
  else
    FK_LAUNCH(h, "lagrangian",
              fk_lagrangian(m, st->rho, st->e, st->mom_x, st->mom_y, mx1, my1,
                            u0, v0, e1, rho0, lr, st->max_dt, st->tt == 0, s));
  if (lr.cfl)
    FK_LAUNCH(h, "timestep_serial",
              fk_timestep_serial(m, lr.cfl, NULL, NULL, NULL, NULL, lr, st->max_dt,
                                 st->tt == 0, s));
  if (h->schedule == FLOWHIP_MARCH)
    FK_LAUNCH(h, "advect_mass_energy",
              fk_march_mass_energy(m, st->tt, u0, v0, rho0, e1, st->rho, st->e,
                                   st->F_x, st->F_y, h->red,
                                   flowhip_rows(h, FLOWHIP_K_MASS, 0), s));
  else
    FK_LAUNCH(h, "advect_mass_energy",
              fk_advect_mass_energy(m, st->tt, u0, v0, rho0, e1, st->rho,
                                    st->e, st->F_x, st->F_y, h->red, s));
  if (h->schedule == FLOWHIP_MARCH) {
    FK_LAUNCH(h, "advect_momentum",
              fk_march_momentum(m, st->tt, u0, v0, st->F_x, st->F_y, st->rho, mx1,
                                my1, st->mom_x, st->mom_y, st->u, st->v,
                                h->red, flowhip_rows(h, FLOWHIP_K_MOMENTUM, 0), s));
  } else {
    FK_LAUNCH(h, "advect_momentum_x",
              fk_advect_momentum_x(m, st->tt, u0, v0, st->F_x, st->F_y, st->rho,
                                   mx1, st->mom_x, st->u, h->red.dtp, s));
    FK_LAUNCH(h, "advect_momentum_y",
              fk_advect_momentum_y(m, st->tt, st->u, v0, st->F_x, st->F_y,
                                   st->rho, my1, st->mom_y, st->v, h->red.dtp, s));
  }
  /* the advection kernel left the totals of the new rho, e in h->red.sums */
  h->sums_valid = 1;
  h->sums_rho = st->rho;
  h->sums_e = st->e;
}
