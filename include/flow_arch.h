/*
 * flow_arch.h -- the arch-compatible layer of the hip kernel set.
 *
 * UoB-HPC/flow is a sub-project of UoB-HPC/arch and compiles against arch's
 * comms.h, mesh.h, params.h, shared.h and shared_data.h (main.c:1-4,
 * flow_interface.h:2-4, omp3/flow.c:2-3). Those files are not part of the
 * reference, so this header supplies the subset that main.c, flow_data.c and
 * the hot path use, with the same names, argument meanings and error
 * behaviour (fatal conditions print and exit). Data arrays live in device
 * memory (HBM) for the whole run, as in the reference's GPU kernel sets.
 * One process drives one GPU; the MPI halo exchange of handle_boundary_2d is
 * replaced by an on-device reflective boundary kernel.
 */
#ifndef FLOW_ARCH_H
#define FLOW_ARCH_H

#include <math.h>
#include <stddef.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---- constants and helper macros (arch shared.h) ---- */
#define MASTER 0
#define EDGE -1
#define NORTH 0
#define EAST 1
#define SOUTH 2
#define WEST 3
#define NNEIGHBOURS 4
#define NO_INVERT 0
#define INVERT_X 1
#define INVERT_Y 2
#define NO_PACK 0
#define PACK 1
#define RECV 0
#define SEND 1
#define MAX_STR_LEN 1024
#define MAX_KEYS 256

#define TERMINATE(...)                  \
  do {                                  \
    fprintf(stderr, __VA_ARGS__);       \
    fprintf(stderr, "\n");              \
    exit(EXIT_FAILURE);                 \
  } while (0)

#ifndef min
#define min(a, b) (((a) < (b)) ? (a) : (b))
#endif
#ifndef max
#define max(a, b) (((a) > (b)) ? (a) : (b))
#endif
#define samesign(a, b) (((a) * (b)) > 0.0)
#define absmin(a, b) ((fabs(a) < fabs(b)) ? (a) : (b))
#define minmod(a, b) (samesign((a), (b)) ? (absmin((a), (b))) : (0.0))
#define strmatch(a, b) (strcmp((a), (b)) == 0)
/* Relative tolerance test used by validate() (main.c:197-200). */
#define within_tolerance(a, b, eps)                                     \
  (!isnan(a) && !isnan(b) &&                                            \
   fabs((a) - (b)) <= ((eps) * (min(fabs(a), fabs(b)))))

/* ---- profiling (arch profiler.h). START/STOP mark omp3's CPU loops and
 * compile to nothing here: the kernel set times its own launches with
 * hipEvents on every mesh created while FLOWHIP_PROFILE=1 is set in the
 * environment (flowhip_set_timing). Built with -DENABLE_PROFILING, as
 * arch's profiler is switched on, the driver's PRINT_PROFILING_RESULTS
 * prints those times per kernel for compute_profile, and for comms_profile
 * that one GPU exchanges nothing. */
struct Profile {
  int unused;
};
extern struct Profile compute_profile;
extern struct Profile comms_profile;
void flowhip_print_profile(const struct Profile* p);
#define START_PROFILING(p)
#define STOP_PROFILING(p, name)
#ifdef ENABLE_PROFILING
#define PRINT_PROFILING_RESULTS(p) flowhip_print_profile(p)
#else
#define PRINT_PROFILING_RESULTS(p)
#endif

/* ---- mesh (arch mesh.h). Same field names as the reference uses
 * (main.c:18-40, omp3/flow.c). */
typedef struct Mesh {
  int local_nx, local_ny;   /* interior + 2*pad */
  int global_nx, global_ny; /* interior cells */
  int niters;
  int pad;
  int rank, nranks;
  int x_off, y_off;
  int neighbours[NNEIGHBOURS];
  double width, height;
  double max_dt, sim_end;
  double dt, dt_h;
  double *edgex, *edgey;   /* left/bottom edge of each local cell, L+1;
                              host memory (used only by problem set-up) */
  double *edgedx, *edgedy; /* face spacing, L+1, device memory */
  double *celldx, *celldy; /* cell spacing, L+1, device memory */
  /* Fields below have no reference counterpart. initialise_mesh_2d sets
   * them. `verbose` gates the per-step prints that rank MASTER makes;
   * `hip` is the kernel set's private per-mesh state (stream, device dt,
   * reduction scratch, captured graphs). */
  int verbose;
  void* hip;
} Mesh;

/* ---- shared data (arch shared_data.h): the arrays flow shares with other
 * arch applications. All device arrays of (local_nx+1)*(local_ny+1). */
typedef struct SharedData {
  double* density;
  double* energy;
  double* density_old;
  double* pressure;
  double* u;
  double* v;
  double* Qxx;
  double* Qyy;
  double* reduce_array0;
  double* reduce_array1;
} SharedData;

/* ---- memory (arch shared.h). Device allocations are zero-initialised. */
size_t allocate_data(double** buf, size_t len);
void deallocate_data(double* buf);
size_t allocate_host_data(double** buf, size_t len);
void deallocate_host_data(double* buf);
/* send != 0 copies host *src to device *dst, RECV copies device to host. */
void copy_buffer(const size_t len, double** src, double** dst, int send);

/* ---- comms (arch comms.h), one process per GPU ---- */
void initialise_mpi(int argc, char** argv, int* rank, int* nranks);
void initialise_comms(Mesh* mesh);
void initialise_devices(int rank);
void finalise_comms(void);
void barrier(void);
double reduce_all_min(double local_val);
double reduce_all_sum(double local_val);
double reduce_to_master(double local_val);
/* Reflective boundary of pad-deep ghost strips on an nx x ny array of device
 * memory (omp3 calls it with staggered extents, e.g. omp3/flow.c:156-157). */
void handle_boundary_2d(const int nx, const int ny, Mesh* mesh, double* arr,
                        const int invert, const int pack);

/* ---- mesh and problem set-up (arch mesh.c, shared_data.c) ---- */
void initialise_mesh_2d(Mesh* mesh);
void finalise_mesh(Mesh* mesh);
void initialise_shared_data_2d(const int local_nx, const int local_ny,
                               const int pad, const double mesh_width,
                               const double mesh_height,
                               const char* problem_def_filename,
                               const double* edgex, const double* edgey,
                               SharedData* shared_data);
void finalise_shared_data(SharedData* shared_data);

/* ---- parameters (arch params.h) ---- */
int get_int_parameter(const char* param_name, const char* filename);
double get_double_parameter(const char* param_name, const char* filename);
/* Finds the line whose first token is `specifier` and returns its
 * key=value pairs: key k at keys + k*MAX_STR_LEN. Returns 0 if absent. */
int get_key_value_parameter(const char* specifier, const char* filename,
                            char* keys, double* values, int* nkeys);

/* ---- output (arch shared.c): writes a BOV header + raw density dump ---- */
void write_all_ranks_to_visit(const int global_nx, const int global_ny,
                              const int local_nx, const int local_ny,
                              const int pad, const int x_off, const int y_off,
                              const int rank, const int nranks,
                              int* neighbours, double* local_arr,
                              const char* name, const int tt,
                              const double elapsed_sim_time);

#ifdef __cplusplus
}
#endif
#endif
