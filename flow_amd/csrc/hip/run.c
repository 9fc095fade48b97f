/*
 * run.c -- flow_hip.h: schedule and graph controls, the restated driver
 * loop, and small device helpers for callers that hold their own buffers.
 */
#include <stdio.h>
#include <string.h>
#include <time.h>

#include "../../../include/flow_interface.h"
#include "flowhip_ctx.h"
#include "schedules.h"

int flowhip_set_schedule(Mesh* mesh, int schedule) {
  if (schedule != FLOWHIP_STAGEWISE && schedule != FLOWHIP_FUSED &&
      schedule != FLOWHIP_MARCH)
    return -1;
  flowhip_ctx(mesh)->schedule = schedule;
  return 0;
}

int flowhip_get_schedule(Mesh* mesh) { return flowhip_ctx(mesh)->schedule; }

int flowhip_set_march_rows(Mesh* mesh, int rows) {
  if (rows != 0 && (rows < 4 || rows > 120)) return -1;
  flowhip_ctx(mesh)->march_rows = rows;
  return 0;
}

int flowhip_set_graphs(Mesh* mesh, int on) {
  flowhip_ctx(mesh)->use_graphs = on ? 1 : 0;
  return 0;
}

static double now_seconds(void) {
  struct timespec ts;
  clock_gettime(CLOCK_MONOTONIC, &ts);
  return (double)ts.tv_sec + 1e-9 * (double)ts.tv_nsec;
}

/* main.c:64-115 */
int flowhip_run(Mesh* mesh, double** a, int niters, int tt0, int pre_timestep,
                double* dt_pairs, double* mass, double* energy,
                double* wall_seconds, int* undefined) {
  if (niters < 0 || !a) return -1;
  const int nx = mesh->local_nx, ny = mesh->local_ny;
  double *P = a[0], *rho = a[1], *rho_old = a[2], *e = a[3], *u = a[4],
         *v = a[5], *mom_x = a[6], *mom_y = a[7], *Qxx = a[8], *Qyy = a[9],
         *F_x = a[10], *F_y = a[11], *uF_x = a[12], *uF_y = a[13],
         *vF_x = a[14], *vF_y = a[15];
  if (pre_timestep)
    set_timestep(nx, ny, Qxx, Qyy, rho, e, mesh, NULL, 0, mesh->celldx,
                 mesh->celldy);
  if (dt_pairs) {
    dt_pairs[0] = mesh->dt;
    dt_pairs[1] = mesh->dt_h;
  }
  if (undefined) undefined[0] = pre_timestep ? flowhip_undefined_candidates(mesh) : 0;
  double wall = 0.0, elapsed = 0.0;
  int k;
  for (k = 0; k < niters; ++k) {
    const int tt = tt0 + k;
    if (mesh->rank == MASTER && mesh->verbose)
      printf("\nIteration %d\n", tt + 1);
    const double w0 = now_seconds();
    solve_hydro_2d(mesh, tt, P, rho, rho_old, e, u, v, mom_x, mom_y, Qxx, Qyy,
                   F_x, F_y, uF_x, uF_y, vF_x, vF_y, NULL);
    print_conservation(nx, ny, rho, e, NULL, mesh);
    wall += now_seconds() - w0;
    if (dt_pairs) {
      dt_pairs[2 * (k + 1)] = mesh->dt;
      dt_pairs[2 * (k + 1) + 1] = mesh->dt_h;
    }
    if (undefined) undefined[k + 1] = flowhip_undefined_candidates(mesh);
    double ms, es;
    flowhip_last_sums(mesh, &ms, &es);
    if (mass) mass[k] = ms;
    if (energy) energy[k] = es;
    elapsed += mesh->dt;
    if (elapsed >= mesh->sim_end) {
      ++k;
      break;
    }
  }
  if (wall_seconds) *wall_seconds = wall;
  return k;
}

int flowhip_device_sync(void) {
  HIPCHECK(hipDeviceSynchronize());
  return 0;
}

int flowhip_memset_zero(double* dev, size_t len) {
  HIPCHECK(hipMemset(dev, 0, sizeof(double) * len));
  return 0;
}

int flowhip_copy(double* dev, double* host, size_t len, int dir) {
  HIPCHECK(hipDeviceSynchronize());
  if (dir == SEND)
    HIPCHECK(hipMemcpy(dev, host, sizeof(double) * len, hipMemcpyHostToDevice));
  else
    HIPCHECK(hipMemcpy(host, dev, sizeof(double) * len, hipMemcpyDeviceToHost));
  return 0;
}

int flowhip_device_name(char* name, int name_len) {
  int dev = 0;
  hipDeviceProp_t prop;
  HIPCHECK(hipGetDevice(&dev));
  HIPCHECK(hipGetDeviceProperties(&prop, dev));
  snprintf(name, (size_t)name_len, "%s (%s, %d CUs)", prop.name,
           prop.gcnArchName, prop.multiProcessorCount);
  return 0;
}

int flowhip_mesh_size_ok(int global_nx, int global_ny) {
  return global_nx >= 2 && global_ny >= 2 && global_nx <= FLOWHIP_MAX_CELLS_SIDE &&
         global_ny <= FLOWHIP_MAX_CELLS_SIDE;
}

Mesh* flowhip_mesh_new(int global_nx, int global_ny, double width,
                       double height, double max_dt, double sim_end,
                       double dt_param) {
  if (!flowhip_mesh_size_ok(global_nx, global_ny)) return NULL;
  Mesh* m = (Mesh*)calloc(1, sizeof(Mesh));
  if (!m) return NULL;
  m->global_nx = global_nx;
  m->global_ny = global_ny;
  m->pad = 2;
  m->local_nx = global_nx + 2 * m->pad;
  m->local_ny = global_ny + 2 * m->pad;
  m->width = width;
  m->height = height;
  m->max_dt = max_dt;
  m->sim_end = sim_end;
  m->dt = C_T * dt_param;
  m->dt_h = m->dt;
  initialise_mpi(0, NULL, &m->rank, &m->nranks);
  initialise_comms(m);
  initialise_mesh_2d(m);
  return m;
}

void flowhip_mesh_delete(Mesh* mesh) {
  if (!mesh) return;
  finalise_mesh(mesh);
  free(mesh);
}
