/*
 * flow.c -- the hip kernel set's implementation of flow_interface.h.
 *
 * C host code that drives hand-written HIP kernels through the thin C ABI of
 * flow_hip_kernels.h. Every array stays in device memory; the timestep pair
 * lives on the device between kernels, and each call ends with one
 * device-to-host transfer of the scalars the caller reads (mesh->dt and
 * mesh->dt_h, or the conservation totals), as main.c:95 requires.
 */
#include <math.h>
#include <stdio.h>

#include "../../../include/flow_interface.h"
#include "flowhip_ctx.h"
#include "schedules.h"

/* Upload the host timestep pair (the caller may have set it). */
static void push_dt(FlowHipCtx* h, const Mesh* mesh) {
  h->h_stage[0] = mesh->dt;
  h->h_stage[1] = mesh->dt_h;
  HIPCHECK(hipMemcpyAsync(h->red.dtp, h->h_stage, 2 * sizeof(double),
                          hipMemcpyHostToDevice, h->stream));
}

/* Fetch the device timestep pair and the status flags into the mesh. */
static void pull_dt(FlowHipCtx* h, Mesh* mesh) {
  HIPCHECK(hipMemcpyAsync(h->h_stage, h->red.dtp, 2 * sizeof(double),
                          hipMemcpyDeviceToHost, h->stream));
  HIPCHECK(hipMemcpyAsync(h->h_status, h->red.status, 2 * sizeof(int),
                          hipMemcpyDeviceToHost, h->stream));
  HIPCHECK(hipStreamSynchronize(h->stream));
  flowhip_timer_collect(h);
  mesh->dt = h->h_stage[0];
  mesh->dt_h = h->h_stage[1];
  if (!isfinite(mesh->dt) || !(mesh->dt > 0.0) || !isfinite(mesh->dt_h))
    TERMINATE("hip kernel set: non-finite or non-positive timestep "
              "(dt=%.12e dt_h=%.12e, %d undefined CFL candidates): the "
              "state has diverged",
              mesh->dt, mesh->dt_h, h->h_status[1]);
}

/* omp3/flow.c:10-52 */
void solve_hydro_2d(Mesh* mesh, int tt, double* P, double* rho,
                    double* rho_old, double* e, double* u, double* v,
                    double* rho_u, double* rho_v, double* Qxx, double* Qyy,
                    double* F_x, double* F_y, double* uF_x, double* uF_y,
                    double* vF_x, double* vF_y, double* reduce_array) {
  (void)reduce_array;
  FlowHipCtx* h = flowhip_ctx(mesh);
  /* a pad-2 reflection needs two interior cells to mirror along each axis */
  if (mesh->local_nx - 2 * mesh->pad < 2 || mesh->local_ny - 2 * mesh->pad < 2)
    TERMINATE("solve_hydro_2d: the mesh needs at least 2x2 interior cells "
              "(got %dx%d)", mesh->local_nx - 2 * mesh->pad,
              mesh->local_ny - 2 * mesh->pad);
  if (mesh->rank == MASTER && mesh->verbose)
    printf("Timestep:        %.12e\n", mesh->dt);

  FlowHipStep st;
  st.m = flowhip_fkmesh(mesh);
  st.tt = tt;
  st.max_dt = mesh->max_dt;
  st.P = P;
  st.rho = rho;
  st.rho_old = rho_old;
  st.e = e;
  st.u = u;
  st.v = v;
  st.mom_x = rho_u;
  st.mom_y = rho_v;
  st.Qxx = Qxx;
  st.Qyy = Qyy;
  st.F_x = F_x;
  st.F_y = F_y;
  st.uF_x = uF_x;
  st.uF_y = uF_y;
  st.vF_x = vF_x;
  st.vF_y = vF_y;

  h->sums_valid = 0;
  push_dt(h, mesh);
  flowhip_enqueue_step(h, &st);
  pull_dt(h, mesh);
}

/* omp3/flow.c:73-106 (reduce_all_min is the identity on one GPU) */
void set_timestep(const int nx, const int ny, double* Qxx, double* Qyy,
                  const double* rho, const double* e, Mesh* mesh,
                  double* reduce_array, const int first_step,
                  const double* celldx, const double* celldy) {
  (void)reduce_array;
  FlowHipCtx* h = flowhip_ctx(mesh);
  FkMesh m = flowhip_fkmesh(mesh);
  m.nx = nx;
  m.ny = ny;
  m.celldx = celldx;
  m.celldy = celldy;
  push_dt(h, mesh);
  if (h->cfl_rule == FLOWHIP_CFL_OMP3_SERIAL)
    FK_LAUNCH(h, "timestep_serial",
              fk_timestep_serial(m, NULL, Qxx, Qyy, rho, e, h->red, mesh->max_dt,
                                 first_step, h->stream));
  else
    FK_LAUNCH(h, "timestep",
              fk_timestep(m, Qxx, Qyy, rho, e, h->red, mesh->max_dt, first_step,
                          h->stream));
  pull_dt(h, mesh);
}

/* omp3/flow.c:973-995 (reduce_to_master is the identity on one GPU) */
void print_conservation(const int nx, const int ny, double* rho, double* e,
                        double* reduce_array, Mesh* mesh) {
  (void)reduce_array;
  FlowHipCtx* h = flowhip_ctx(mesh);
  if (!(h->sums_valid && h->sums_rho == rho && h->sums_e == e &&
        nx == mesh->local_nx && ny == mesh->local_ny)) {
    FkMesh m = flowhip_fkmesh(mesh);
    m.nx = nx;
    m.ny = ny;
    FK_LAUNCH(h, "conservation",
              fk_conservation(m, rho, e, h->red, h->stream));
  }
  HIPCHECK(hipMemcpyAsync(h->h_stage + 2, h->red.sums, 2 * sizeof(double),
                          hipMemcpyDeviceToHost, h->stream));
  HIPCHECK(hipStreamSynchronize(h->stream));
  flowhip_timer_collect(h);
  h->sums_valid = 0; /* fused totals are used at most once */
  if (mesh->rank == MASTER && mesh->verbose) {
    printf("Total mass:      %.12e\n", h->h_stage[2]);
    printf("Total energy:    %.12e\n", h->h_stage[3]);
  }
}

/* Totals from the last print_conservation, for callers of flowhip_run. */
void flowhip_last_sums(Mesh* mesh, double* mass, double* energy) {
  FlowHipCtx* h = flowhip_ctx(mesh);
  *mass = h->h_stage[2];
  *energy = h->h_stage[3];
}
