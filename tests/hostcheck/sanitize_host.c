/*
 * sanitize_host.c -- host-only harness that runs the CPU-side C of the
 * build under AddressSanitizer and UndefinedBehaviorSanitizer (gcc): the
 * oracle over a few steps of two problems under both CFL rules, the
 * parameter-file readers (arch/params.c) on a given file, and the state-file
 * reader and writer (hip/state.c) on given files. The state files' device
 * copies go through the stub hipMemcpy below (host buffers stand in for
 * device memory), so no GPU runtime is linked.
 *
 *   sanitize_host oracle
 *   sanitize_host params <file> <int key> <double key> <specifier>
 *   sanitize_host state-write <file> <nx> <ny>
 *   sanitize_host state-read <file> <nx> <ny>     (prints the load status)
 */
#include <hip/hip_runtime_api.h>
#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include "../../include/flow_hip.h"
#include "../../oracle/oracle.h"

/* ---- stand-ins for the HIP runtime calls state.c makes ---- */
hipError_t hipMemcpy(void* dst, const void* src, size_t bytes, hipMemcpyKind kind) {
  (void)kind;
  memmove(dst, src, bytes);
  return hipSuccess;
}
int flowhip_mesh_size_ok(int global_nx, int global_ny) {
  return global_nx >= 2 && global_ny >= 2 && global_nx <= FLOWHIP_MAX_CELLS_SIDE &&
         global_ny <= FLOWHIP_MAX_CELLS_SIDE;
}

static int run_oracle(void) {
  const int sizes[2][2] = {{24, 20}, {17, 31}};
  for (int rule = 0; rule < 2; ++rule) {
    orc_set_cfl_rule(rule);
    for (int p = 0; p < 2; ++p) {
      OrcMesh m;
      orc_mesh_init(&m, sizes[p][0], sizes[p][1], 1.0, 1.0, 1.0e30, 1.0e30, 1.0e-4);
      const size_t n = (size_t)(m.nx + 1) * (m.ny + 1);
      double* a[16];
      for (int k = 0; k < 16; ++k) a[k] = (double*)calloc(n, sizeof(double));
      orc_rasterise_box(&m, a[1], a[3], 0.125, 2.0, 0.0, 0.0, 1.0, 1.0);
      orc_rasterise_box(&m, a[1], a[3], 1.0, 250.0, 0.3, 0.4, 0.3, 0.2);
      orc_handle_boundary_2d(m.nx, m.ny, m.pad, a[1], ORC_NO_INVERT);
      orc_handle_boundary_2d(m.nx, m.ny, m.pad, a[3], ORC_NO_INVERT);
      double dts[2 * 9], mass[8], energy[8], wall = 0.0;
      int und[9];
      const int steps = orc_run(&m, a, 8, dts, mass, energy, &wall, und);
      printf("oracle rule=%d mesh=%dx%d steps=%d dt=%.6e mass=%.12e undefined=%d\n",
             rule, sizes[p][0], sizes[p][1], steps, dts[2 * steps], mass[steps - 1],
             und[steps]);
      for (int k = 0; k < 16; ++k) free(a[k]);
      orc_mesh_free(&m);
    }
  }
  return 0;
}

static int run_params(const char* file, const char* ikey, const char* dkey,
                      const char* spec) {
  const int iv = get_int_parameter(ikey, file);
  const double dv = get_double_parameter(dkey, file);
  char* keys = (char*)malloc((size_t)MAX_KEYS * MAX_STR_LEN);
  double* values = (double*)malloc(sizeof(double) * MAX_KEYS);
  int nkeys = 0;
  const int found = get_key_value_parameter(spec, file, keys, values, &nkeys);
  printf("params %s=%d %s=%.17g %s found=%d nkeys=%d", ikey, iv, dkey, dv, spec,
         found, nkeys);
  for (int k = 0; k < nkeys; ++k) printf(" %s=%.17g", keys + (size_t)k * MAX_STR_LEN,
                                         values[k]);
  printf("\n");
  free(keys);
  free(values);
  return 0;
}

static void mesh_of(Mesh* m, int nx, int ny) {
  memset(m, 0, sizeof(*m));
  m->global_nx = nx;
  m->global_ny = ny;
  m->pad = 2;
  m->local_nx = nx + 4;
  m->local_ny = ny + 4;
  m->width = m->height = 1.0;
  m->dt = 1.5e-4;
  m->dt_h = 1.25e-4;
}

static int run_state(int write, const char* file, int nx, int ny) {
  Mesh m;
  mesh_of(&m, nx, ny);
  const size_t n = (size_t)(m.local_nx + 1) * (m.local_ny + 1);
  double* a[6];
  for (int k = 0; k < 6; ++k) {
    a[k] = (double*)malloc(sizeof(double) * n);
    for (size_t i = 0; i < n; ++i) a[k][i] = write ? sin(0.1 * (double)i + k) : -1.0;
  }
  int rc, step = 0;
  double t = 0.0;
  if (write) {
    rc = flowhip_state_save(file, &m, 7, 0.25, a[0], a[1], a[2], a[3], a[4], a[5]);
  } else {
    rc = flowhip_state_load(file, &m, &step, &t, a[0], a[1], a[2], a[3], a[4], a[5]);
    int gx = 0, gy = 0;
    const int irc = flowhip_state_info(file, &gx, &gy, NULL, NULL);
    printf("info rc=%d nx=%d ny=%d\n", irc, gx, gy);
  }
  /* over each array's layout extent (rho, e: cells; u, mom_x: x faces;
   * v, mom_y: y faces), what the file holds */
  const size_t Lx = (size_t)m.local_nx, Ly = (size_t)m.local_ny;
  const size_t extent[6] = {Lx * Ly, Lx * Ly, Ly * (Lx + 1), (Ly + 1) * Lx,
                            Ly * (Lx + 1), (Ly + 1) * Lx};
  double sum = 0.0;
  for (int k = 0; k < 6; ++k)
    for (size_t i = 0; i < extent[k]; ++i) sum += a[k][i];
  printf("state %s rc=%d step=%d t=%g dt=%.6e checksum=%.12e\n",
         write ? "write" : "read", rc, step, t, m.dt, sum);
  for (int k = 0; k < 6; ++k) free(a[k]);
  return 0;
}

int main(int argc, char** argv) {
  if (argc >= 2 && !strcmp(argv[1], "oracle")) return run_oracle();
  if (argc == 6 && !strcmp(argv[1], "params"))
    return run_params(argv[2], argv[3], argv[4], argv[5]);
  if (argc == 5 && (!strcmp(argv[1], "state-write") || !strcmp(argv[1], "state-read")))
    return run_state(!strcmp(argv[1], "state-write"), argv[2], atoi(argv[3]),
                     atoi(argv[4]));
  fprintf(stderr, "usage: %s oracle | params f ikey dkey spec | "
                  "state-write|state-read f nx ny\n", argv[0]);
  return 2;
}
