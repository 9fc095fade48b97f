/*
 * f32.c -- the FP32-state precision form of the timestep.
 *
 * The 16 arrays of solve_hydro_2d are held as FP32 device arrays and the
 * march kernels run on them in FP32. Three things stay FP64: the CFL
 * candidates and the dt min-reduction (omp3/flow.c:77-105; the state is
 * widened before the sound speed is formed), the mass and energy totals
 * (omp3/flow.c:973-995; compensated FP64 sums of the widened state), and
 * the EOS pressure of near-vacuum cells (rho < 1e-6). The dt pair stays on
 * the device in FP64 and is rounded to FP32 inside each kernel.
 */
#include <math.h>
#include <stdio.h>
#include <time.h>

#include "schedules.h"

int flowhip_f32_alloc(float** buf, size_t len) {
  if (!buf) return -1;
  HIPCHECK(hipMalloc((void**)buf, sizeof(float) * (len ? len : 1)));
  HIPCHECK(hipMemset(*buf, 0, sizeof(float) * (len ? len : 1)));
  return 0;
}

void flowhip_f32_free(float* buf) {
  if (buf) HIPCHECK(hipFree(buf));
}

int flowhip_f32_zero(float* buf, size_t len) {
  HIPCHECK(hipMemset(buf, 0, sizeof(float) * len));
  return 0;
}

int flowhip_f32_copy(float* dev, float* host, size_t len, int dir) {
  HIPCHECK(hipDeviceSynchronize());
  if (dir == SEND)
    HIPCHECK(hipMemcpy(dev, host, sizeof(float) * len, hipMemcpyHostToDevice));
  else
    HIPCHECK(hipMemcpy(host, dev, sizeof(float) * len, hipMemcpyDeviceToHost));
  return 0;
}

int flowhip_f32_narrow(float* dst, const double* src, size_t len) {
  HIPCHECK(fk_narrow((long)len, dst, src, 0));
  HIPCHECK(hipDeviceSynchronize());
  return 0;
}

int flowhip_f32_widen(double* dst, const float* src, size_t len) {
  HIPCHECK(fk_widen((long)len, dst, src, 0));
  HIPCHECK(hipDeviceSynchronize());
  return 0;
}

int flowhip_f32_boundary(Mesh* mesh, float* arr, int nxp, int nyp, int invert) {
  FlowHipCtx* h = flowhip_ctx(mesh);
  HIPCHECK(fk_boundary_f32(nxp, nyp, mesh->pad, arr, invert, h->stream));
  HIPCHECK(hipStreamSynchronize(h->stream));
  return 0;
}

static void push_dt32(FlowHipCtx* h, const Mesh* mesh) {
  h->h_stage[0] = mesh->dt;
  h->h_stage[1] = mesh->dt_h;
  HIPCHECK(hipMemcpyAsync(h->red.dtp, h->h_stage, 2 * sizeof(double),
                          hipMemcpyHostToDevice, h->stream));
}

/* dt pair, and the totals when with_sums, in one transfer + sync */
static void pull32(FlowHipCtx* h, Mesh* mesh, int with_sums) {
  HIPCHECK(hipMemcpyAsync(h->h_stage, h->red.dtp, 2 * sizeof(double),
                          hipMemcpyDeviceToHost, h->stream));
  if (with_sums)
    HIPCHECK(hipMemcpyAsync(h->h_stage + 2, h->red.sums, 2 * sizeof(double),
                            hipMemcpyDeviceToHost, h->stream));
  HIPCHECK(hipMemcpyAsync(h->h_status, h->red.status, 2 * sizeof(int),
                          hipMemcpyDeviceToHost, h->stream));
  HIPCHECK(hipStreamSynchronize(h->stream));
  flowhip_timer_collect(h);
  mesh->dt = h->h_stage[0];
  mesh->dt_h = h->h_stage[1];
  if (!isfinite(mesh->dt) || !(mesh->dt > 0.0) || !isfinite(mesh->dt_h))
    TERMINATE("hip kernel set (FP32 state): non-finite or non-positive "
              "timestep (dt=%.12e dt_h=%.12e, %d undefined CFL candidates): "
              "the state has diverged",
              mesh->dt, mesh->dt_h, h->h_status[1]);
}

/* One step of omp3/flow.c:10-52 on FP32 state, in the march schedule's four
 * kernels (the same ping-pong use of P, Qxx, Qyy, uF_x, uF_y as the FP64
 * fused schedule). */
static void step_f32(FlowHipCtx* h, Mesh* mesh, FkMesh32 m, int tt, float** a) {
  float *P = a[0], *rho = a[1], *rho_old = a[2], *e = a[3], *u = a[4],
        *v = a[5], *mom_x = a[6], *mom_y = a[7], *Qxx = a[8], *Qyy = a[9],
        *F_x = a[10], *F_y = a[11], *uF_x = a[12], *uF_y = a[13];
  hipStream_t s = h->stream;
  FK_LAUNCH(h, "lagrangian_f32",
            fk_march_lagrangian_f32(m, rho, e, mom_x, mom_y, Qxx, Qyy, uF_x,
                                    uF_y, P, rho_old, h->red, mesh->max_dt,
                                    tt == 0, flowhip_rows(h, FLOWHIP_K_LAGRANGIAN, 1), s));
  FK_LAUNCH(h, "advect_mass_energy_f32",
            fk_march_mass_energy_f32(m, tt, uF_x, uF_y, rho_old, P, rho, e,
                                     F_x, F_y, h->red, flowhip_rows(h, FLOWHIP_K_MASS, 1),
                                     s));
  FK_LAUNCH(h, "advect_momentum_f32",
            fk_march_momentum_f32(m, tt, uF_x, uF_y, F_x, F_y, rho, Qxx, Qyy,
                                  mom_x, mom_y, u, v, h->red,
                                  flowhip_rows(h, FLOWHIP_K_MOMENTUM, 1), s));
}

static double now_seconds(void) {
  struct timespec ts;
  clock_gettime(CLOCK_MONOTONIC, &ts);
  return (double)ts.tv_sec + 1e-9 * (double)ts.tv_nsec;
}

/* main.c:64-115 on FP32 state; the arguments mean what flowhip_run's do. */
int flowhip_run_f32(Mesh* mesh, float** a, int niters, int tt0,
                    int pre_timestep, double* dt_pairs, double* mass,
                    double* energy, double* wall_seconds, int* undefined) {
  if (niters < 0 || !a) return -1;
  if (mesh->local_nx - 2 * mesh->pad < 2 || mesh->local_ny - 2 * mesh->pad < 2)
    TERMINATE("flowhip_run_f32: the mesh needs at least 2x2 interior cells");
  FlowHipCtx* h = flowhip_ctx(mesh);
  const FkMesh32 m = flowhip_fkmesh32(mesh);
  h->sums_valid = 0;
  if (pre_timestep) {
    push_dt32(h, mesh);
    FK_LAUNCH(h, "timestep_f32",
              fk_timestep_f32(m, a[8], a[9], a[1], a[3], h->red, mesh->max_dt,
                              0, h->stream));
    pull32(h, mesh, 0);
  }
  if (dt_pairs) {
    dt_pairs[0] = mesh->dt;
    dt_pairs[1] = mesh->dt_h;
  }
  if (undefined) undefined[0] = pre_timestep ? h->h_status[1] : 0;
  double wall = 0.0, elapsed = 0.0;
  int k;
  for (k = 0; k < niters; ++k) {
    const int tt = tt0 + k;
    const double w0 = now_seconds();
    push_dt32(h, mesh);
    step_f32(h, mesh, m, tt, a);
    pull32(h, mesh, 1);
    wall += now_seconds() - w0;
    if (mesh->rank == MASTER && mesh->verbose) {
      printf("\nIteration %d\n", tt + 1);
      printf("Total mass:      %.12e\n", h->h_stage[2]);
      printf("Total energy:    %.12e\n", h->h_stage[3]);
    }
    if (dt_pairs) {
      dt_pairs[2 * (k + 1)] = mesh->dt;
      dt_pairs[2 * (k + 1) + 1] = mesh->dt_h;
    }
    if (mass) mass[k] = h->h_stage[2];
    if (energy) energy[k] = h->h_stage[3];
    if (undefined) undefined[k + 1] = h->h_status[1];
    elapsed += mesh->dt;
    if (elapsed >= mesh->sim_end) {
      ++k;
      break;
    }
  }
  if (wall_seconds) *wall_seconds = wall;
  return k;
}
