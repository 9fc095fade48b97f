/*
 * stagewise.c -- the stagewise schedule: the kernel sequence of one step in
 * omp3's loop order (omp3/flow.c:10-52), with an on-device boundary kernel
 * at each of omp3's 19 handle_boundary_2d sites.
 */
#include "schedules.h"

#define BND(h, nxp, nyp, arr, inv)                                        \
  FK_LAUNCH(h, "boundary",                                                \
            fk_boundary((nxp), (nyp), st->m.pad, (arr), (inv), h->stream))

/* omp3/flow.c:326-439 (x) or 442-554 (y) */
static void mass_energy_sweep(FlowHipCtx* h, const FlowHipStep* st, int ysweep,
                              int first) {
  const FkMesh m = st->m;
  const int nx = m.nx, ny = m.ny;
  const double* dtp = h->red.dtp;
  hipStream_t s = h->stream;
  if (!ysweep) {
    FK_LAUNCH(h, "x_flux",
              fk_x_flux(m, dtp, st->rho, st->e, st->u, st->F_x, st->uF_x, s));
    BND(h, nx + 1, ny, st->F_x, INVERT_X);
    FK_LAUNCH(h, "x_update", fk_x_update(m, dtp, first, st->rho, st->rho_old,
                                         st->e, st->F_x, st->uF_x, s));
  } else {
    FK_LAUNCH(h, "y_flux",
              fk_y_flux(m, dtp, st->rho, st->e, st->v, st->F_y, st->uF_y, s));
    BND(h, nx, ny + 1, st->F_y, INVERT_Y);
    FK_LAUNCH(h, "y_update", fk_y_update(m, dtp, first, st->rho, st->rho_old,
                                         st->e, st->F_y, st->uF_y, s));
  }
  BND(h, nx, ny, st->rho, NO_INVERT);
  BND(h, nx, ny, st->e, NO_INVERT);
}

void flowhip_step_stagewise(FlowHipCtx* h, const FlowHipStep* st) {
  const FkMesh m = st->m;
  const int nx = m.nx, ny = m.ny;
  const double* dtp = h->red.dtp;
  hipStream_t s = h->stream;

  /* Lagrangian phase with the previous (dt, dt_h) */
  FK_LAUNCH(h, "eos", fk_eos(m, st->P, st->rho, st->e, s));
  FK_LAUNCH(h, "pressure_accel",
            fk_pressure_accel(m, dtp, st->mom_x, st->mom_y, st->u, st->v,
                              st->P, st->rho, s));
  BND(h, nx + 1, ny, st->u, INVERT_X);
  BND(h, nx, ny + 1, st->v, INVERT_Y);
  FK_LAUNCH(h, "av_stress",
            fk_av_stress(m, st->Qxx, st->Qyy, st->u, st->v, st->rho, s));
  BND(h, nx, ny, st->Qxx, NO_INVERT);
  BND(h, nx, ny, st->Qyy, NO_INVERT);
  FK_LAUNCH(h, "av_accel",
            fk_av_accel(m, dtp, st->mom_x, st->mom_y, st->u, st->v, st->Qxx,
                        st->Qyy, st->rho, s));
  BND(h, nx + 1, ny, st->u, INVERT_X);
  BND(h, nx, ny + 1, st->v, INVERT_Y);
  FK_LAUNCH(h, "shock_heating",
            fk_shock_heating(m, dtp, st->e, st->P, st->u, st->v, st->rho,
                             st->Qxx, st->Qyy, s));
  BND(h, nx, ny, st->e, NO_INVERT);

  /* the only global dependency: the new (dt, dt_h), left on the device */
  FK_LAUNCH(h, "timestep",
            fk_timestep(m, st->Qxx, st->Qyy, st->rho, st->e, h->red,
                        st->max_dt, st->tt == 0, s));

  /* advection of mass and energy (omp3/flow.c:293-323) */
  FK_LAUNCH(h, "copy_rho_old",
            fk_copy((long)nx * ny, st->rho_old, st->rho, s));
  if (st->tt % 2 == 0) {
    mass_energy_sweep(h, st, 0, 1);
    mass_energy_sweep(h, st, 1, 0);
  } else {
    mass_energy_sweep(h, st, 1, 1);
    mass_energy_sweep(h, st, 0, 0);
  }

  /* advection of momentum (omp3/flow.c:557-775) */
  if (st->tt % 2) {
    FK_LAUNCH(h, "mom_flux_xx",
              fk_mom_flux_xx(m, dtp, st->u, st->uF_x, st->F_x, s));
    BND(h, nx, ny, st->uF_x, NO_INVERT);
    FK_LAUNCH(h, "mom_x_update_u", fk_mom_x_update(m, dtp, 0, 1, st->mom_x,
                                                 st->u, st->uF_x, st->rho, s));
    BND(h, nx + 1, ny, st->u, INVERT_X);
    FK_LAUNCH(h, "mom_flux_xy",
              fk_mom_flux_xy(m, dtp, st->u, st->v, st->uF_y, st->F_y, s));
    BND(h, nx + 1, ny + 1, st->uF_y, NO_INVERT);
    FK_LAUNCH(h, "mom_x_update", fk_mom_x_update(m, dtp, 1, 0, st->mom_x,
                                                 st->u, st->uF_y, st->rho, s));
    FK_LAUNCH(h, "mom_flux_yx",
              fk_mom_flux_yx(m, dtp, st->u, st->v, st->vF_x, st->F_x, s));
    BND(h, nx + 1, ny + 1, st->vF_x, NO_INVERT);
    FK_LAUNCH(h, "mom_y_update_v", fk_mom_y_update(m, dtp, 0, 1, st->mom_y,
                                                 st->v, st->vF_x, st->rho, s));
    BND(h, nx, ny + 1, st->v, INVERT_Y);
    FK_LAUNCH(h, "mom_flux_yy",
              fk_mom_flux_yy(m, dtp, st->v, st->vF_y, st->F_y, s));
    BND(h, nx, ny, st->vF_y, NO_INVERT);
    FK_LAUNCH(h, "mom_y_update", fk_mom_y_update(m, dtp, 1, 0, st->mom_y,
                                                 st->v, st->vF_y, st->rho, s));
  } else {
    FK_LAUNCH(h, "mom_flux_xy",
              fk_mom_flux_xy(m, dtp, st->u, st->v, st->uF_y, st->F_y, s));
    BND(h, nx + 1, ny + 1, st->uF_y, NO_INVERT);
    FK_LAUNCH(h, "mom_x_update_u", fk_mom_x_update(m, dtp, 1, 1, st->mom_x,
                                                 st->u, st->uF_y, st->rho, s));
    BND(h, nx + 1, ny, st->u, INVERT_X);
    FK_LAUNCH(h, "mom_flux_xx",
              fk_mom_flux_xx(m, dtp, st->u, st->uF_x, st->F_x, s));
    BND(h, nx, ny, st->uF_x, NO_INVERT);
    FK_LAUNCH(h, "mom_x_update", fk_mom_x_update(m, dtp, 0, 0, st->mom_x,
                                                 st->u, st->uF_x, st->rho, s));
    FK_LAUNCH(h, "mom_flux_yy",
              fk_mom_flux_yy(m, dtp, st->v, st->vF_y, st->F_y, s));
    BND(h, nx, ny, st->vF_y, NO_INVERT);
    FK_LAUNCH(h, "mom_y_update_v", fk_mom_y_update(m, dtp, 1, 1, st->mom_y,
                                                 st->v, st->vF_y, st->rho, s));
    BND(h, nx, ny + 1, st->v, INVERT_Y);
    FK_LAUNCH(h, "mom_flux_yx",
              fk_mom_flux_yx(m, dtp, st->u, st->v, st->vF_x, st->F_x, s));
    BND(h, nx + 1, ny + 1, st->vF_x, NO_INVERT);
    FK_LAUNCH(h, "mom_y_update", fk_mom_y_update(m, dtp, 0, 0, st->mom_y,
                                                 st->v, st->vF_x, st->rho, s));
  }
}

void flowhip_enqueue_step(FlowHipCtx* h, const FlowHipStep* st) {
  if (h->schedule == FLOWHIP_FUSED || h->schedule == FLOWHIP_MARCH)
    flowhip_step_fused(h, st);
  else
    flowhip_step_stagewise(h, st);
}
