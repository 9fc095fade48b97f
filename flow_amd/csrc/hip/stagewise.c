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
  if (h->cfl_rule == FLOWHIP_CFL_OMP3_SERIAL)
    FK_LAUNCH(h, "timestep_serial",
              fk_timestep_serial(m, NULL, st->Qxx, st->Qyy, st->rho, st->e, h->red,
                                 st->max_dt, st->tt == 0, s));
  else
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

static void enqueue_direct(FlowHipCtx* h, const FlowHipStep* st) {
  if (h->schedule == FLOWHIP_FUSED || h->schedule == FLOWHIP_MARCH)
    flowhip_step_fused(h, st);
  else
    flowhip_step_stagewise(h, st);
}

static int same_step(const FlowHipStep* a, const FlowHipStep* b) {
  const double* pa[16] = {a->P,   a->rho,  a->rho_old, a->e,    a->u,    a->v,
                          a->mom_x, a->mom_y, a->Qxx, a->Qyy,  a->F_x,  a->F_y,
                          a->uF_x, a->uF_y, a->vF_x, a->vF_y};
  const double* pb[16] = {b->P,   b->rho,  b->rho_old, b->e,    b->u,    b->v,
                          b->mom_x, b->mom_y, b->Qxx, b->Qyy,  b->F_x,  b->F_y,
                          b->uF_x, b->uF_y, b->vF_x, b->vF_y};
  for (int k = 0; k < 16; ++k)
    if (pa[k] != pb[k]) return 0;
  return a->tt == b->tt && a->max_dt == b->max_dt && a->m.nx == b->m.nx &&
         a->m.ny == b->m.ny && a->m.pad == b->m.pad &&
         a->m.edgedx == b->m.edgedx && a->m.edgedy == b->m.edgedy &&
         a->m.celldx == b->m.celldx && a->m.celldy == b->m.celldy &&
         a->m.inv_edgedx == b->m.inv_edgedx && a->m.inv_edgedy == b->m.inv_edgedy &&
         a->m.pow2 == b->m.pow2 && a->m.uniform == b->m.uniform;
}

void flowhip_enqueue_step(FlowHipCtx* h, const FlowHipStep* st) {
  /* timing records events around each launch: always direct */
  if (!h->use_graphs || h->timing) {
    enqueue_direct(h, st);
    return;
  }
  /* the step's kind: sweep parity, and tt == 0's first_step dt update */
  const int slot = st->tt == 0 ? 2 : (st->tt & 1);
  FlowHipStep k = *st;
  k.tt = slot == 2 ? 0 : (slot == 1 ? 1 : 2);
  if (!(h->graph_ok[slot] && same_step(&h->graph_step[slot], &k) &&
        h->graph_sched[slot] == h->schedule &&
        h->graph_rows[slot] == h->march_rows &&
        h->graph_rule[slot] == h->cfl_rule)) {
    if (!h->graph_warm) {
      /* the first step runs directly: launch shapes are sized then */
      enqueue_direct(h, st);
      h->graph_warm = 1;
      return;
    }
    if (h->graph_ok[slot]) HIPCHECK(hipGraphExecDestroy(h->graph[slot]));
    h->graph_ok[slot] = 0;
    /* no allocation may happen inside the capture */
    (void)flowhip_lagrangian_red(h, st->m.nx, st->m.ny);
    hipGraph_t g;
    HIPCHECK(hipStreamBeginCapture(h->stream, hipStreamCaptureModeThreadLocal));
    enqueue_direct(h, &k);
    HIPCHECK(hipStreamEndCapture(h->stream, &g));
    HIPCHECK(hipGraphInstantiate(&h->graph[slot], g, NULL, NULL, 0));
    HIPCHECK(hipGraphDestroy(g));
    h->graph_ok[slot] = 1;
    h->graph_step[slot] = k;
    h->graph_sched[slot] = h->schedule;
    h->graph_rows[slot] = h->march_rows;
    h->graph_rule[slot] = h->cfl_rule;
  }
  HIPCHECK(hipGraphLaunch(h->graph[slot], h->stream));
  /* host-side state of a direct launch: the graph's advection kernel left
   * the totals of the new rho and e on the device for print_conservation */
  if (h->schedule == FLOWHIP_FUSED || h->schedule == FLOWHIP_MARCH) {
    h->sums_valid = 1;
    h->sums_rho = st->rho;
    h->sums_e = st->e;
  }
}
