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
/* Rows per chunk of the row-marching kernels: 4..120 for all of them, or 0
 * (the default) for each kernel's own height, measured fastest on MI355X at
 * 8192^2: 24 / 20 / 16 rows for the Lagrangian / mass-energy / momentum
 * marches in FP64, 24 / 16 / 16 in the FP32-state form; when the
 * Lagrangian and mass/energy marches hold two columns per lane (uniform
 * power-of-two meshes), 12 / 12 rows in FP64 and 10 / 10 in FP32. Fields
 * and dt are bitwise the same for every height. */
int flowhip_set_march_rows(Mesh* mesh, int rows);
/* Replay each step from a captured hipGraph: one per sweep-order parity
 * and one for tt == 0, captured on first use and recaptured when the
 * arrays, geometry, schedule or chunk height change. Per-kernel timing
 * (flowhip_set_timing) launches directly. */
int flowhip_set_graphs(Mesh* mesh, int on);

/* Rules for set_timestep's undefined CFL candidates (cells with negative
 * energy or zero density, whose candidate omp3/flow.c:87-94 makes NaN).
 * omp3 folds the candidates with min(a, b) = a < b ? a : b under an OpenMP
 * min reduction: a NaN makes the running value NaN and the next candidate
 * replaces it, so each thread forgets its candidates before its last NaN,
 * and the result depends on OMP_NUM_THREADS.
 *   FLOWHIP_CFL_DEFINED (default): the minimum over the defined candidates
 *     and max_dt; order- and thread-independent, and equal to omp3's
 *     result whenever no candidate is undefined.
 *   FLOWHIP_CFL_OMP3_SERIAL: omp3 on one thread, bit for bit: the fold in
 *     row-major order from +inf, merged into max_dt by OpenMP's combiner
 *     out = in < out ? in : out. Costs a candidate store and one more
 *     launch per step (FP64 single-mesh forms only).
 * The environment may set it for callers that only see flow_interface.h:
 * FLOWHIP_CFL_RULE=defined|omp3-serial. */
#define FLOWHIP_CFL_DEFINED 0
#define FLOWHIP_CFL_OMP3_SERIAL 1
int flowhip_set_cfl_rule(Mesh* mesh, int rule);
int flowhip_get_cfl_rule(Mesh* mesh);
/* Undefined CFL candidates met by the last set_timestep or solve_hydro_2d
 * call (each is left out of, or restarts, the minimum as the rule says). */
int flowhip_undefined_candidates(Mesh* mesh);

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
 * undefined receives niters+1 counts of undefined CFL candidates (the
 * pre-loop call's, then each step's; 0 when pre_timestep is 0). Any output
 * pointer may be NULL. Returns the number of steps taken. pre_timestep=0
 * skips the pre-loop call (to continue a run). */
int flowhip_run(Mesh* mesh, double** arrays, int niters, int tt0,
                int pre_timestep, double* dt_pairs, double* mass,
                double* energy, double* wall_seconds, int* undefined);

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
                    double* energy, double* wall_seconds, int* undefined);

/* ---- ensemble (batching) form ----
 * ninst independent meshes of uneven sizes (global_nx[i] x global_ny[i]
 * interior cells, arch.params width/height/max_dt and .params dt per
 * instance), FP64 or FP32 state (fp32 = 1: the precision form's rules).
 * Every stage is one launch for the whole ensemble: persistent kernels
 * walk one work list of row-march items over all instances; each instance
 * keeps its own dt pair and totals and computes bit for bit what it
 * computes alone. Steps can be replayed from captured hipGraphs. */
typedef struct FlowHipEnsemble FlowHipEnsemble;
FlowHipEnsemble* flowhip_ensemble_new(int ninst, const int* global_nx,
                                      const int* global_ny, const double* width,
                                      const double* height, const double* max_dt,
                                      const double* dt_param, int fp32);
void flowhip_ensemble_delete(FlowHipEnsemble* ens);
int flowhip_ensemble_size(FlowHipEnsemble* ens);
/* Local sizes (pad included) of instance inst. */
int flowhip_ensemble_local_size(FlowHipEnsemble* ens, int inst, int* nx, int* ny);
/* Device pointer of array `which` (solve_hydro_2d order, 0..15) of an
 * instance: (nx+1)*(ny+1) doubles, or floats in an FP32 ensemble. */
void* flowhip_ensemble_array(FlowHipEnsemble* ens, int inst, int which);
/* Initial density and energy (host FP64, (nx+1)*(ny+1) each) of one
 * instance, reflected as main.c:56-59 does. */
int flowhip_ensemble_set_state(FlowHipEnsemble* ens, int inst, const double* rho,
                               const double* e);
/* One array of one instance, widened to FP64, into host memory. */
int flowhip_ensemble_get_array(FlowHipEnsemble* ens, int inst, int which,
                               double* host);
int flowhip_ensemble_set_dt(FlowHipEnsemble* ens, int inst, double dt, double dt_h);
int flowhip_ensemble_set_graphs(FlowHipEnsemble* ens, int on);
/* Chunk height (4..120), work order (0: by instance, 1: largest instance
 * first, 2: one chunk row of every instance in turn) and momentum schedule
 * (1: both directions in one march, 0: x then y). */
int flowhip_ensemble_set_schedule(FlowHipEnsemble* ens, int chunk_rows, int order,
                                  int fused_momentum);
int flowhip_ensemble_get_schedule(FlowHipEnsemble* ens, int* chunk_rows, int* order,
                                  int* fused_momentum);
/* Times trial_steps steps of each chunk height, work order and momentum
 * schedule from the current state, keeps the fastest, restores the state. */
int flowhip_ensemble_autotune(FlowHipEnsemble* ens, int trial_steps, int tt0,
                              double* best_ms_per_step);
/* niters steps of every instance; dt_pairs [inst][niters+1][2], mass and
 * energy [inst][niters], undefined CFL candidates per step [inst][niters]
 * (any may be NULL). Returns niters. */
int flowhip_ensemble_run(FlowHipEnsemble* ens, int niters, int tt0,
                         int pre_timestep, double* dt_pairs, double* mass,
                         double* energy, double* wall_seconds, int* undefined);

/* ---- one step without allocation ----
 * No counterpart in the reference. For callers that cannot allocate device
 * memory (tight serving loops, harnesses that forbid hipMalloc): one
 * solve_hydro_2d step (omp3/flow.c:10-52) in the march schedule, bit for
 * bit the step solve_hydro_2d takes, plus that step's totals (what
 * print_conservation prints) in *mass and *energy (either may be NULL).
 * It calls no hipMalloc, hipFree or other allocating routine and creates no
 * per-Mesh state: its scratch (reduction partials, counters, dt pair,
 * totals, reciprocal spacings) is one static device array of 1008896 bytes
 * per GPU, held until the process ends. Meshes up to FLOWHIP_MAX_CELLS_SIDE
 * interior cells a side; the CFL rule is FLOWHIP_CFL_DEFINED. Reads
 * mesh->dt and dt_h, writes the next pair into them (as solve_hydro_2d
 * does), and works on `stream`, a hipStream_t (NULL: the default stream),
 * synchronising it before it returns. Every call on one GPU shares that
 * GPU's scratch: calls that overlap on one GPU (other threads, other
 * streams) race.
 * Returns 0; -1 for bad arguments, -2 for a mesh beyond the scratch, -3 if
 * the scratch cannot be found, -4 for a non-finite or non-positive new dt
 * (each with a message on stderr). */
int flowhip_solve_hydro_2d_static(Mesh* mesh, int tt, double* P, double* rho,
                                  double* rho_old, double* e, double* u, double* v,
                                  double* mom_x, double* mom_y, double* Qxx,
                                  double* Qyy, double* F_x, double* F_y,
                                  double* uF_x, double* uF_y, double* vF_x,
                                  double* vF_y, double* mass, double* energy,
                                  void* stream);

/* Device helpers for callers that manage their own buffers. */
int flowhip_device_sync(void);
int flowhip_memset_zero(double* dev, size_t len);
/* Raw host<->device copies of len doubles (dir: SEND = host to device). */
int flowhip_copy(double* dev, double* host, size_t len, int dir);
/* Name of the GPU in use. */
int flowhip_device_name(char* name, int name_len);

/* Bit-exact state files (format in flow_amd/csrc/hip/state.c): the carried
 * fields rho, e, u, v, mom_x and mom_y (device arrays), the timestep pair,
 * the steps completed and the simulated time. The reference has no
 * counterpart; its driver keeps no state between runs. A file with step 0
 * is an initial state; a later step is a restart point that continues the
 * run bit for bit (mesh->dt and dt_h are restored on load). Both return 0,
 * or -1 after printing what failed (an unreadable file, another mesh size
 * or domain). */
int flowhip_state_save(const char* path, const Mesh* mesh, int step, double sim_time,
                       const double* rho, const double* e, const double* u,
                       const double* v, const double* mom_x, const double* mom_y);
int flowhip_state_load(const char* path, Mesh* mesh, int* step, double* sim_time,
                       double* rho, double* e, double* u, double* v, double* mom_x,
                       double* mom_y);
/* Header fields of a state file (any output may be NULL). */
int flowhip_state_info(const char* path, int* global_nx, int* global_ny, int* step,
                       double* sim_time);

/* Mesh sizes the kernel set can index: 2 to FLOWHIP_MAX_CELLS_SIDE interior
 * cells a side. Kernels address a field's (n+5)^2 elements with 32-bit
 * offsets; at n = 46335 one FP64 field is 17.2 GB and the 16 arrays of a
 * step 275 GB, within the 288 GB of one MI355X. Larger meshes are refused
 * (initialise_mesh_2d exits with a message, flowhip_mesh_new returns NULL)
 * instead of overflowing an index. */
#define FLOWHIP_MAX_CELLS_SIDE 46335
int flowhip_mesh_size_ok(int global_nx, int global_ny);

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
