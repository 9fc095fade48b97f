/*
 * flow_hip.h -- controls of the hip kernel set that have no counterpart in
 * flow_interface.h: schedule choice, graph replay, per-kernel timing, and a
 * C restatement of the driver's time loop (main.c:64-115) for callers that
 * drive many steps without per-step host work of their own.
 *
 * Everything here operates on a Mesh set up by initialise_mesh_2d and on
 * device arrays from allocate_data. Functions return 0 on success and a
 * negative value on bad arguments; HIP failures are fatal, as in the
 * reference's GPU kernel sets (cuda/flow.cu:24).
 */
#ifndef FLOW_HIP_H
#define FLOW_HIP_H

#include "flow_arch.h"

#ifdef __cplusplus
extern "C" {
#endif

/* Schedules of solve_hydro_2d. */
#define FLOWHIP_STAGEWISE 0 /* one kernel per omp3 loop; every intermediate
                               array (P, Q, rho_old, fluxes) is written */
#define FLOWHIP_FUSED 1     /* fused Lagrangian, advection and momentum
                               kernels; only the carried state and the
                               fluxes the momentum step reads are written */
#define FLOWHIP_MARCH 2     /* the fused schedule with row-marching kernels
                               (default) */
/* The schedule of a new Mesh may also be set from the environment, for
 * callers that only see flow_interface.h: FLOWHIP_SCHEDULE=stagewise|fused|
 * march and FLOWHIP_MARCH_ROWS=<rows>. */

int flowhip_set_schedule(Mesh* mesh, int schedule);
int flowhip_get_schedule(Mesh* mesh);
/* Rows per chunk of the row-marching kernels: 4..120, default 32. */
int flowhip_set_march_rows(Mesh* mesh, int rows);
/* Replay each step from a captured hipGraph (one per sweep-order parity and
 * for the first step). */
int flowhip_set_graphs(Mesh* mesh, int on);

/* Per-kernel timing with hipEvents on the kernel set's stream. */
int flowhip_set_timing(Mesh* mesh, int on);
int flowhip_timing_count(Mesh* mesh);
/* Entry k: kernel name, summed milliseconds and number of launches. */
int flowhip_timing_entry(Mesh* mesh, int k, char* name, int name_len,
                         double* total_ms, long* launches);
int flowhip_timing_reset(Mesh* mesh);

/* main.c:64-115 without its prints: pre-loop set_timestep(first_step=0),
 * then up to niters x (solve_hydro_2d + print_conservation), stopping at
 * mesh->sim_end. arrays: P rho rho_old e u v mom_x mom_y Qxx Qyy F_x F_y
 * uF_x uF_y vF_x vF_y (device). dt_pairs receives 2*(niters+1) values
 * ((dt, dt_h) after the pre-loop call, then after each step); mass/energy
 * receive the per-step totals; wall_seconds the time spent in the steps.
 * Any output pointer may be NULL. Returns the number of steps taken.
 * pre_timestep=0 skips the pre-loop call (to continue a run). */
int flowhip_run(Mesh* mesh, double** arrays, int niters, int tt0,
                int pre_timestep, double* dt_pairs, double* mass,
                double* energy, double* wall_seconds);

/* ---- FP32-state precision form ----
 * The 16 arrays of solve_hydro_2d (same roles, layouts and (L+1)^2 sizes)
 * held as FP32 device arrays; the march kernels run on them in FP32. The
 * CFL candidates and the dt reduction, the mass/energy totals (compensated
 * sums) and the EOS at near-vacuum cells (rho < 1e-6) are FP64; the dt pair
 * is FP64 on the device. Geometry is the mesh's, rounded to FP32. */
int flowhip_f32_alloc(float** buf, size_t len);  /* zeroed device floats */
void flowhip_f32_free(float* buf);
int flowhip_f32_zero(float* buf, size_t len);
/* Host<->device copies of len floats (dir: SEND = host to device). */
int flowhip_f32_copy(float* dev, float* host, size_t len, int dir);
/* Device-to-device conversions (round to nearest when narrowing). */
int flowhip_f32_narrow(float* dst, const double* src, size_t len);
int flowhip_f32_widen(double* dst, const float* src, size_t len);
/* handle_boundary_2d on an FP32 device array. */
int flowhip_f32_boundary(Mesh* mesh, float* arr, int nxp, int nyp, int invert);
/* flowhip_run on FP32 state: arrays in solve_hydro_2d order; dt pairs and
 * totals are FP64. */
int flowhip_run_f32(Mesh* mesh, float** arrays, int niters, int tt0,
                    int pre_timestep, double* dt_pairs, double* mass,
                    double* energy, double* wall_seconds);

/* Device helpers for callers that manage their own buffers. */
int flowhip_device_sync(void);
int flowhip_memset_zero(double* dev, size_t len);
/* Raw host<->device copies of len doubles (dir: SEND = host to device). */
int flowhip_copy(double* dev, double* host, size_t len, int dir);
/* Name of the GPU in use. */
int flowhip_device_name(char* name, int name_len);

/* Mesh objects for callers that cannot place a Mesh on their stack: reads
 * no files; width/height/max_dt/sim_end are the arch.params values and
 * dt_param the .params dt (main.c:20-31). */
Mesh* flowhip_mesh_new(int global_nx, int global_ny, double width,
                       double height, double max_dt, double sim_end,
                       double dt_param);
void flowhip_mesh_delete(Mesh* mesh);

#ifdef __cplusplus
}
#endif
#endif
