/*
 * state.c -- the bit-exact state file of the hip kernel set.
 *
 * A state file holds what one timestep carries to the next: the cell fields
 * rho and e, the face velocities u and v, the momenta mom_x and mom_y (each
 * at its layout's exact extent in rows x columns, C: Ly x Lx, X: Ly x
 * (Lx+1), Y: (Ly+1) x Lx with Lx = nx + 2*pad, Ly = ny + 2*pad), the
 * timestep pair (dt, dt_h) the next solve_hydro_2d call
 * starts from, the number of steps completed and the simulated time. It is
 * the on-disk form of a workload's initial state (step 0) and a restart
 * point (step > 0); flow_amd/statefile.py reads and writes the same format
 * for the oracle and the tests.
 *
 * Layout, little-endian, no padding between fields:
 *   char   magic[8]    "FLOWSTAT"
 *   u32    version     1
 *   u32    narrays     6
 *   i32    global_nx, global_ny, pad, step
 *   f64    width, height, dt, dt_h, sim_time
 *   then per array, in the order rho, e, u, v, mom_x, mom_y:
 *   char   name[16]    NUL-padded
 *   char   layout      'C', 'X' or 'Y'
 *   char   reserved[7]
 *   u64    count       rows * cols of the layout
 *   f64    data[count] row-major
 */
#include <stdint.h>
#include <stdio.h>
#include <string.h>
#include <hip/hip_runtime.h>
#include "../../../include/flow_hip.h"

#define STATE_MAGIC "FLOWSTAT"
#define STATE_VERSION 1u
#define STATE_NARRAYS 6

static const char* const kNames[STATE_NARRAYS] = {"rho", "e", "u", "v", "mom_x", "mom_y"};
static const char kLayouts[STATE_NARRAYS] = {'C', 'C', 'X', 'Y', 'X', 'Y'};

/* Elements of one array at its layout's extent; rows are contiguous at the
 * layout's stride, so this is also the prefix of the device array. */
static size_t layout_count(char layout, int Lx, int Ly) {
  switch (layout) {
    case 'C': return (size_t)Ly * Lx;
    case 'X': return (size_t)Ly * (Lx + 1);
    default: return (size_t)(Ly + 1) * Lx;
  }
}

typedef struct {
  char magic[8];
  uint32_t version, narrays;
  int32_t global_nx, global_ny, pad, step;
  double width, height, dt, dt_h, sim_time;
} StateHeader;

typedef struct {
  char name[16];
  char layout;
  char reserved[7];
  uint64_t count;
} ArrayHeader;

static int fail(FILE* f, const char* path, const char* what) {
  fprintf(stderr, "state file %s: %s\n", path, what);
  if (f) fclose(f);
  return -1;
}

static int read_header(FILE* f, const char* path, StateHeader* h) {
  if (fread(h, sizeof(*h), 1, f) != 1) return fail(f, path, "truncated header");
  if (memcmp(h->magic, STATE_MAGIC, 8) != 0) return fail(f, path, "not a state file");
  if (h->version != STATE_VERSION) return fail(f, path, "unsupported version");
  if (h->narrays != STATE_NARRAYS) return fail(f, path, "unexpected array count");
  if (!flowhip_mesh_size_ok(h->global_nx, h->global_ny) || h->pad != 2)
    return fail(f, path, "unsupported mesh (2 to FLOWHIP_MAX_CELLS_SIDE interior "
                         "cells a side, pad 2)");
  return 0;
}

int flowhip_state_info(const char* path, int* global_nx, int* global_ny, int* step,
                       double* sim_time) {
  FILE* f = fopen(path, "rb");
  if (!f) return fail(NULL, path, "cannot open");
  StateHeader h;
  if (read_header(f, path, &h)) return -1;
  fclose(f);
  if (global_nx) *global_nx = h.global_nx;
  if (global_ny) *global_ny = h.global_ny;
  if (step) *step = h.step;
  if (sim_time) *sim_time = h.sim_time;
  return 0;
}

int flowhip_state_save(const char* path, const Mesh* mesh, int step, double sim_time,
                       const double* rho, const double* e, const double* u,
                       const double* v, const double* mom_x, const double* mom_y) {
  const double* dev[STATE_NARRAYS] = {rho, e, u, v, mom_x, mom_y};
  const int Lx = mesh->local_nx, Ly = mesh->local_ny;
  FILE* f = fopen(path, "wb");
  if (!f) return fail(NULL, path, "cannot create");
  StateHeader h;
  memset(&h, 0, sizeof(h));
  memcpy(h.magic, STATE_MAGIC, 8);
  h.version = STATE_VERSION;
  h.narrays = STATE_NARRAYS;
  h.global_nx = mesh->global_nx;
  h.global_ny = mesh->global_ny;
  h.pad = mesh->pad;
  h.step = step;
  h.width = mesh->width;
  h.height = mesh->height;
  h.dt = mesh->dt;
  h.dt_h = mesh->dt_h;
  h.sim_time = sim_time;
  if (fwrite(&h, sizeof(h), 1, f) != 1) return fail(f, path, "write failed");
  double* host = NULL;
  size_t cap = 0;
  for (int k = 0; k < STATE_NARRAYS; ++k) {
    ArrayHeader a;
    memset(&a, 0, sizeof(a));
    strncpy(a.name, kNames[k], sizeof(a.name) - 1);
    a.layout = kLayouts[k];
    a.count = layout_count(a.layout, Lx, Ly);
    if (a.count > cap) {
      free(host);
      cap = a.count;
      host = (double*)malloc(cap * sizeof(double));
      if (!host) return fail(f, path, "out of host memory");
    }
    if (hipMemcpy(host, dev[k], a.count * sizeof(double), hipMemcpyDeviceToHost) !=
        hipSuccess) {
      free(host);
      return fail(f, path, "device-to-host copy failed");
    }
    if (fwrite(&a, sizeof(a), 1, f) != 1 ||
        fwrite(host, sizeof(double), a.count, f) != a.count) {
      free(host);
      return fail(f, path, "write failed");
    }
  }
  free(host);
  if (fclose(f) != 0) return fail(NULL, path, "close failed");
  return 0;
}

int flowhip_state_load(const char* path, Mesh* mesh, int* step, double* sim_time,
                       double* rho, double* e, double* u, double* v, double* mom_x,
                       double* mom_y) {
  double* dev[STATE_NARRAYS] = {rho, e, u, v, mom_x, mom_y};
  FILE* f = fopen(path, "rb");
  if (!f) return fail(NULL, path, "cannot open");
  StateHeader h;
  if (read_header(f, path, &h)) return -1;
  if (h.global_nx != mesh->global_nx || h.global_ny != mesh->global_ny ||
      h.pad != mesh->pad)
    return fail(f, path, "mesh size differs from the run's");
  if (h.width != mesh->width || h.height != mesh->height)
    return fail(f, path, "domain size differs from the run's");
  const int Lx = mesh->local_nx, Ly = mesh->local_ny;
  double* host = NULL;
  size_t cap = 0;
  for (int k = 0; k < STATE_NARRAYS; ++k) {
    ArrayHeader a;
    if (fread(&a, sizeof(a), 1, f) != 1) {
      free(host);
      return fail(f, path, "truncated array header");
    }
    if (strncmp(a.name, kNames[k], sizeof(a.name)) != 0 || a.layout != kLayouts[k] ||
        a.count != layout_count(kLayouts[k], Lx, Ly)) {
      free(host);
      return fail(f, path, "unexpected array record");
    }
    if (a.count > cap) {
      free(host);
      cap = a.count;
      host = (double*)malloc(cap * sizeof(double));
      if (!host) return fail(f, path, "out of host memory");
    }
    if (fread(host, sizeof(double), a.count, f) != a.count) {
      free(host);
      return fail(f, path, "truncated array data");
    }
    if (hipMemcpy(dev[k], host, a.count * sizeof(double), hipMemcpyHostToDevice) !=
        hipSuccess) {
      free(host);
      return fail(f, path, "host-to-device copy failed");
    }
  }
  free(host);
  fclose(f);
  mesh->dt = h.dt;
  mesh->dt_h = h.dt_h;
  if (step) *step = h.step;
  if (sim_time) *sim_time = h.sim_time;
  return 0;
}
