/*
 * arch_hip.c -- the arch-compatible layer on one GPU: device memory,
 * transfers, the on-device reflective boundary, single-rank comms, uniform
 * mesh geometry and problem set-up (the behaviours main.c:20-62 and
 * flow_data.c:8-26 rely on; arch itself is not part of the reference).
 */
#include <hip/hip_runtime_api.h>
#include <stdio.h>
#include <string.h>

#include "../../../include/flow_arch.h"
#include "../../../include/flow_hip.h"
#include "../hip/flowhip_ctx.h"

struct Profile compute_profile;
struct Profile comms_profile;

/* ---- memory ---- */

size_t allocate_data(double** buf, size_t len) {
  const size_t bytes = sizeof(double) * (len ? len : 1);
  HIPCHECK(hipMalloc((void**)buf, bytes));
  HIPCHECK(hipMemset(*buf, 0, bytes));
  return bytes;
}

void deallocate_data(double* buf) {
  if (buf) HIPCHECK(hipFree(buf));
}

size_t allocate_host_data(double** buf, size_t len) {
  *buf = (double*)calloc(len ? len : 1, sizeof(double));
  if (!*buf) TERMINATE("allocate_host_data: out of memory (%zu doubles)", len);
  return sizeof(double) * len;
}

void deallocate_host_data(double* buf) { free(buf); }

void copy_buffer(const size_t len, double** src, double** dst, int send) {
  HIPCHECK(hipDeviceSynchronize());
  HIPCHECK(hipMemcpy(*dst, *src, sizeof(double) * len,
                     send == SEND ? hipMemcpyHostToDevice
                                  : hipMemcpyDeviceToHost));
}

/* ---- comms on one GPU ---- */

void initialise_mpi(int argc, char** argv, int* rank, int* nranks) {
  (void)argc;
  (void)argv;
  *rank = MASTER;
  *nranks = 1;
}

void initialise_comms(Mesh* mesh) {
  for (int k = 0; k < NNEIGHBOURS; ++k) mesh->neighbours[k] = EDGE;
  mesh->x_off = 0;
  mesh->y_off = 0;
}

void initialise_devices(int rank) {
  int count = 0;
  HIPCHECK(hipGetDeviceCount(&count));
  if (count < 1) TERMINATE("initialise_devices: no GPU visible");
  HIPCHECK(hipSetDevice(rank % count));
}

void finalise_comms(void) {}

void barrier(void) { HIPCHECK(hipDeviceSynchronize()); }

double reduce_all_min(double local_val) { return local_val; }
double reduce_all_sum(double local_val) { return local_val; }
double reduce_to_master(double local_val) { return local_val; }

void handle_boundary_2d(const int nx, const int ny, Mesh* mesh, double* arr,
                        const int invert, const int pack) {
  (void)pack;
  FlowHipCtx* h = flowhip_ctx(mesh);
  HIPCHECK(fk_boundary(nx, ny, mesh->pad, arr, invert, h->stream));
  HIPCHECK(hipStreamSynchronize(h->stream));
}

/* ---- mesh ---- */

/* Uniform rectilinear geometry: every spacing is width/global_nx and
 * edgex[k] the left edge of local cell k (one rank, no offset). The edge
 * positions are only used for problem set-up and stay on the host; the
 * spacings are device arrays read by the kernels. */
void initialise_mesh_2d(Mesh* mesh) {
  if (!flowhip_mesh_size_ok(mesh->global_nx, mesh->global_ny))
    TERMINATE("initialise_mesh_2d: a %dx%d mesh is outside what the hip kernel set "
              "indexes (2..%d cells a side: every field's (n+5)^2 elements must "
              "have 32-bit offsets)",
              mesh->global_nx, mesh->global_ny, FLOWHIP_MAX_CELLS_SIDE);
  const int lx = mesh->local_nx + 1, ly = mesh->local_ny + 1;
  double* hx = (double*)calloc((size_t)lx, sizeof(double));
  double* hy = (double*)calloc((size_t)ly, sizeof(double));
  mesh->edgex = (double*)calloc((size_t)lx, sizeof(double));
  mesh->edgey = (double*)calloc((size_t)ly, sizeof(double));
  if (!hx || !hy || !mesh->edgex || !mesh->edgey)
    TERMINATE("initialise_mesh_2d: out of memory");
  for (int k = 0; k < lx; ++k) {
    hx[k] = mesh->width / mesh->global_nx;
    mesh->edgex[k] = hx[k] * (double)(mesh->x_off + k - mesh->pad);
  }
  for (int k = 0; k < ly; ++k) {
    hy[k] = mesh->height / mesh->global_ny;
    mesh->edgey[k] = hy[k] * (double)(mesh->y_off + k - mesh->pad);
  }
  allocate_data(&mesh->edgedx, (size_t)lx);
  allocate_data(&mesh->celldx, (size_t)lx);
  allocate_data(&mesh->edgedy, (size_t)ly);
  allocate_data(&mesh->celldy, (size_t)ly);
  HIPCHECK(hipMemcpy(mesh->edgedx, hx, sizeof(double) * lx,
                     hipMemcpyHostToDevice));
  HIPCHECK(hipMemcpy(mesh->celldx, hx, sizeof(double) * lx,
                     hipMemcpyHostToDevice));
  HIPCHECK(hipMemcpy(mesh->edgedy, hy, sizeof(double) * ly,
                     hipMemcpyHostToDevice));
  HIPCHECK(hipMemcpy(mesh->celldy, hy, sizeof(double) * ly,
                     hipMemcpyHostToDevice));
  free(hx);
  free(hy);
  mesh->verbose = 1;
  mesh->hip = NULL;
  flowhip_ctx(mesh);
}

void finalise_mesh(Mesh* mesh) {
  flowhip_ctx_destroy(mesh);
  deallocate_data(mesh->edgedx);
  deallocate_data(mesh->edgedy);
  deallocate_data(mesh->celldx);
  deallocate_data(mesh->celldy);
  free(mesh->edgex);
  free(mesh->edgey);
  mesh->edgedx = mesh->edgedy = mesh->celldx = mesh->celldy = NULL;
  mesh->edgex = mesh->edgey = NULL;
}

/* ---- problem set-up ---- */

/* Apply every "problem_N" line of the file in order: bounds are fractions
 * of the domain, and a local cell takes the listed values when its
 * left/bottom edge lies in [pos, pos+extent) (flow.tests:1 requires this
 * rule: 819x819 cells for flow.params' problem_1). */
static void set_problem_2d(int local_nx, int local_ny, double mesh_width,
                           double mesh_height, const double* edgex,
                           const double* edgey, const char* filename,
                           double* density, double* energy) {
  char* keys = (char*)malloc((size_t)MAX_KEYS * MAX_STR_LEN);
  double* values = (double*)malloc(sizeof(double) * MAX_KEYS);
  if (!keys || !values) TERMINATE("set_problem_2d: out of memory");
  for (int n = 0;; ++n) {
    char spec[64];
    snprintf(spec, sizeof(spec), "problem_%d", n);
    int nkeys = 0;
    if (!get_key_value_parameter(spec, filename, keys, values, &nkeys)) break;
    if (nkeys < 4)
      TERMINATE("set_problem_2d: %s in %s needs xpos ypos width height",
                spec, filename);
    const double xpos = values[nkeys - 4] * mesh_width;
    const double ypos = values[nkeys - 3] * mesh_height;
    const double w = values[nkeys - 2] * mesh_width;
    const double hgt = values[nkeys - 1] * mesh_height;
    for (int i = 0; i < local_ny; ++i) {
      for (int j = 0; j < local_nx; ++j) {
        const double gx = edgex[j], gy = edgey[i];
        if (!(gx >= xpos && gy >= ypos && gx < xpos + w && gy < ypos + hgt))
          continue;
        for (int k = 0; k < nkeys - 4; ++k) {
          const char* key = keys + (size_t)k * MAX_STR_LEN;
          if (strmatch(key, "density"))
            density[i * local_nx + j] = values[k];
          else if (strmatch(key, "energy"))
            energy[i * local_nx + j] = values[k];
          else
            TERMINATE("set_problem_2d: unrecognised key %s in %s", key,
                      filename);
        }
      }
    }
  }
  free(keys);
  free(values);
}

void initialise_shared_data_2d(const int local_nx, const int local_ny,
                               const int pad, const double mesh_width,
                               const double mesh_height,
                               const char* problem_def_filename,
                               const double* edgex, const double* edgey,
                               SharedData* sd) {
  (void)pad;
  const size_t n = (size_t)(local_nx + 1) * (local_ny + 1);
  allocate_data(&sd->density, n);
  allocate_data(&sd->energy, n);
  allocate_data(&sd->density_old, n);
  allocate_data(&sd->pressure, n);
  allocate_data(&sd->u, n);
  allocate_data(&sd->v, n);
  allocate_data(&sd->Qxx, n);
  allocate_data(&sd->Qyy, n);
  allocate_data(&sd->reduce_array0, n);
  allocate_data(&sd->reduce_array1, n);
  double *hd, *he;
  allocate_host_data(&hd, n);
  allocate_host_data(&he, n);
  set_problem_2d(local_nx, local_ny, mesh_width, mesh_height, edgex, edgey,
                 problem_def_filename, hd, he);
  copy_buffer(n, &hd, &sd->density, SEND);
  copy_buffer(n, &he, &sd->energy, SEND);
  deallocate_host_data(hd);
  deallocate_host_data(he);
}

void finalise_shared_data(SharedData* sd) {
  deallocate_data(sd->density);
  deallocate_data(sd->energy);
  deallocate_data(sd->density_old);
  deallocate_data(sd->pressure);
  deallocate_data(sd->u);
  deallocate_data(sd->v);
  deallocate_data(sd->Qxx);
  deallocate_data(sd->Qyy);
  deallocate_data(sd->reduce_array0);
  deallocate_data(sd->reduce_array1);
  memset(sd, 0, sizeof(*sd));
}

/* ---- output ---- */

/* A BOV header plus the interior of the array as raw little-endian
 * doubles, named <name><tt>.bov / .dat. */
void write_all_ranks_to_visit(const int global_nx, const int global_ny,
                              const int local_nx, const int local_ny,
                              const int pad, const int x_off, const int y_off,
                              const int rank, const int nranks,
                              int* neighbours, double* local_arr,
                              const char* name, const int tt,
                              const double elapsed_sim_time) {
  (void)global_nx;
  (void)global_ny;
  (void)x_off;
  (void)y_off;
  (void)nranks;
  (void)neighbours;
  if (rank != MASTER) return;
  const size_t n = (size_t)local_nx * local_ny;
  double* h;
  allocate_host_data(&h, n);
  copy_buffer(n, &local_arr, &h, RECV);
  char datname[512], bovname[512];
  snprintf(datname, sizeof(datname), "%s%d.dat", name, tt);
  snprintf(bovname, sizeof(bovname), "%s%d.bov", name, tt);
  FILE* f = fopen(datname, "wb");
  if (!f) TERMINATE("write_all_ranks_to_visit: cannot open %s", datname);
  for (int i = pad; i < local_ny - pad; ++i)
    fwrite(h + (size_t)i * local_nx + pad, sizeof(double),
           (size_t)(local_nx - 2 * pad), f);
  fclose(f);
  FILE* b = fopen(bovname, "w");
  if (!b) TERMINATE("write_all_ranks_to_visit: cannot open %s", bovname);
  fprintf(b, "TIME: %.12e\nDATA_FILE: %s\nDATA_SIZE: %d %d 1\n",
          elapsed_sim_time, datname, local_nx - 2 * pad, local_ny - 2 * pad);
  fprintf(b, "DATA_FORMAT: DOUBLE\nVARIABLE: %s\nDATA_ENDIAN: LITTLE\n", name);
  fprintf(b, "CENTERING: zonal\nBRICK_ORIGIN: 0. 0. 0.\n");
  fprintf(b, "BRICK_SIZE: %d %d 1\n", local_nx - 2 * pad, local_ny - 2 * pad);
  fclose(b);
  deallocate_host_data(h);
}
