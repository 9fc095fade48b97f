/*
 * flow_hip_kernels.h -- thin C ABI between the hip kernel set's C host code
 * (flow_amd/csrc/hip/flow.c) and its HIP kernels (flow_amd/csrc/kernels/).
 *
 * Plain device pointers, sizes and a hipStream_t; every launcher returns 0
 * or a hipError_t. Each launcher names the omp3 loop it replaces. Arrays use
 * the reference layouts (cell i*nx+j; x-face i*(nx+1)+j; y-face i*nx+j with
 * ny+1 rows; corner i*(nx+1)+j with ny+1 rows), nx/ny local sizes with pad.
 *
 * Timestep scalars stay on the device: `dtp` points at {dt, dt_h}. Kernels
 * launched before fk_timestep read the previous pair, kernels after it the
 * new one, by stream order.
 */
#ifndef FLOW_HIP_KERNELS_H
#define FLOW_HIP_KERNELS_H

#include <hip/hip_runtime_api.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct FkMesh {
  int nx, ny, pad;
  const double *edgedx, *edgedy, *celldx, *celldy; /* device, L+1 each */
  /* 1.0 / edgedx and 1.0 / edgedy (the reference's invdx / invdy,
   * omp3/flow.c:344, 460), computed once per mesh by fk_reciprocal; read
   * by the fused kernels only. */
  const double *inv_edgedx, *inv_edgedy;
  /* 1 when every spacing above is a power of two (the march kernels then
   * divide by spacings as products with exact reciprocals), else 0 */
  int pow2;
  /* 1 when, besides, each of the four spacing arrays holds one value
   * throughout (the mass/energy and momentum marches then read the row
   * geometry once per work item instead of once per row), else 0 */
  int uniform;
  /* Whether the Lagrangian and mass/energy marches hold two columns per
   * lane on uniform power-of-two meshes (each can when nx is even and the
   * cell and y-face arrays it reads and writes are 16-byte aligned):
   * 0 never, 1 on meshes of at least FK_PAIR_MIN_CELLS interior cells,
   * 2 on every mesh */
  int pairs;
} FkMesh;

/* Interior cells from which the two-column marches are the faster form
 * (measured on MI355X: slower at 1024^2, where the mesh does not fill the
 * GPU and a two-column item is the longer critical path; faster from
 * 2048^2) */
#define FK_PAIR_MIN_CELLS (2048L * 2048L)

/* Geometry of the FP32-state precision form: the same arrays rounded to
 * FP32 (the reciprocals are rounded from the FP64 reciprocals). */
typedef struct FkMesh32 {
  int nx, ny, pad;
  const float *edgedx, *edgedy, *celldx, *celldy;
  const float *inv_edgedx, *inv_edgedy;
  int pow2;
  int uniform; /* as in FkMesh */
  int pairs;   /* as in FkMesh (8-byte alignment) */
} FkMesh32;

/* Device scratch for the single-pass grid reductions. */
typedef struct FkReduce {
  double* partials;      /* fk_reduce_blocks() doubles */
  unsigned int* ticket;  /* one counter, zero between launches */
  double* dtp;           /* {dt, dt_h} */
  double* sums;          /* {mass, energy, unclamped min dt, spare} */
  int* status;           /* [0] NaN CFL candidates being counted, [1] the
                            count of the last fk_timestep launch */
  int max_blocks;
  /* FK_MARCH_QUEUE_WORDS words holding the counters from which the waves
   * of a march kernel claim work items; zero between launches (the last
   * wave resets them). */
  unsigned int* queue;
  /* FK_TOTALS_WORDS exact fixed-point accumulators of the mass/energy
   * march's totals; zero between launches (the last block clears them). */
  unsigned long long* totals;
  /* NULL, or a cell-layout array (nx * ny doubles) that the FP64
   * Lagrangian kernels fill with every interior CFL candidate. When set,
   * those kernels leave the (dt, dt_h) update and the status count to a
   * following fk_timestep_serial launch that reads the candidates. */
  double* cfl;
} FkReduce;

#define FK_MARCH_QUEUE_WORDS 512
#define FK_TOTALS_WORDS 256

/* Number of doubles FkReduce.partials must hold. */
int fk_reduce_blocks(void);

/* arch handle_boundary_2d, on-device (omp3/flow.c:156 etc.). */
hipError_t fk_boundary(int nxp, int nyp, int pad, double* arr, int invert,
                       hipStream_t s);
hipError_t fk_copy(long n, double* dst, const double* src, hipStream_t s);
/* dst[k] = 1.0 / src[k] for k < n (correctly rounded division). */
hipError_t fk_reciprocal(int n, double* dst, const double* src,
                         hipStream_t s);

/* ---- stagewise schedule: one kernel per omp3 loop ---- */
/* omp3/flow.c:59-67 */
hipError_t fk_eos(FkMesh m, double* P, const double* rho, const double* e,
                  hipStream_t s);
/* omp3/flow.c:121-152 (dt = dtp[0]) */
hipError_t fk_pressure_accel(FkMesh m, const double* dtp, double* mom_x,
                             double* mom_y, double* u, double* v,
                             const double* P, const double* rho, hipStream_t s);
/* omp3/flow.c:173-202 */
hipError_t fk_av_stress(FkMesh m, double* Qxx, double* Qyy, const double* u,
                        const double* v, const double* rho, hipStream_t s);
/* omp3/flow.c:212-240 (dt = dtp[0]) */
hipError_t fk_av_accel(FkMesh m, const double* dtp, double* mom_x,
                       double* mom_y, double* u, double* v, const double* Qxx,
                       const double* Qyy, const double* rho, hipStream_t s);
/* omp3/flow.c:258-285 (dt_h = dtp[1]) */
hipError_t fk_shock_heating(FkMesh m, const double* dtp, double* e,
                            const double* P, const double* u, const double* v,
                            const double* rho, const double* Qxx,
                            const double* Qyy, hipStream_t s);
/* omp3/flow.c:77-105: grid minimum of the CFL candidates, seeded with
 * max_dt, then the (dt, dt_h) update written to r->dtp. One launch. */
hipError_t fk_timestep(FkMesh m, const double* Qxx, const double* Qyy,
                       const double* rho, const double* e, FkReduce r,
                       double max_dt, int first_step, hipStream_t s);
/* omp3/flow.c:77-105 with the candidates folded as one OpenMP thread folds
 * them: every interior candidate in row-major order with min(a, b) = a < b ?
 * a : b from +inf (a NaN candidate makes the running value NaN and the next
 * candidate replaces it), then merged into max_dt by OpenMP's min combiner
 * out = in < out ? in : out; then the (dt, dt_h) update. Candidates come from
 * cfl (cell layout, as left by a Lagrangian kernel with FkReduce.cfl set),
 * or, with cfl NULL, are computed from e, Qxx, Qyy and rho; undefined
 * candidates are counted into r.status in the second case only. */
hipError_t fk_timestep_serial(FkMesh m, const double* cfl, const double* Qxx,
                              const double* Qyy, const double* rho,
                              const double* e, FkReduce r, double max_dt,
                              int first_step, hipStream_t s);
/* omp3/flow.c:338-404 / 454-521 (dt = dtp[0], dt_h = dtp[1]) */
hipError_t fk_x_flux(FkMesh m, const double* dtp, const double* rho,
                     const double* e, const double* u, double* F_x,
                     double* eF_x, hipStream_t s);
hipError_t fk_y_flux(FkMesh m, const double* dtp, const double* rho,
                     const double* e, const double* v, double* F_y,
                     double* eF_y, hipStream_t s);
/* omp3/flow.c:412-434 / 528-549 */
hipError_t fk_x_update(FkMesh m, const double* dtp, int first, double* rho,
                       const double* rho_old, double* e, const double* F_x,
                       const double* eF_x, hipStream_t s);
hipError_t fk_y_update(FkMesh m, const double* dtp, int first, double* rho,
                       const double* rho_old, double* e, const double* F_y,
                       const double* eF_y, hipStream_t s);
/* omp3/flow.c:778-823, 826-874, 877-925, 928-970 (dt_h = dtp[1]) */
hipError_t fk_mom_flux_xx(FkMesh m, const double* dtp, const double* u,
                          double* uF_x, const double* F_x, hipStream_t s);
hipError_t fk_mom_flux_xy(FkMesh m, const double* dtp, const double* u,
                          const double* v, double* uF_y, const double* F_y,
                          hipStream_t s);
hipError_t fk_mom_flux_yx(FkMesh m, const double* dtp, const double* u,
                          const double* v, double* vF_x, const double* F_x,
                          hipStream_t s);
hipError_t fk_mom_flux_yy(FkMesh m, const double* dtp, const double* v,
                          double* vF_y, const double* F_y, hipStream_t s);
/* Momentum sub-updates of omp3/flow.c:577-594 (flux_dir 0: uF_x),
 * :610-618 and :683-700 (flux_dir 1: uF_y); recompute selects the velocity
 * recompute that follows the first sub-update of each parity. */
hipError_t fk_mom_x_update(FkMesh m, const double* dtp, int flux_dir,
                           int recompute, double* mom_x, double* u,
                           const double* flux, const double* rho,
                           hipStream_t s);
/* omp3/flow.c:630-647, 660-669, 731-748, 763-771 (flux_dir 0: vF_x,
 * flux_dir 1: vF_y) */
hipError_t fk_mom_y_update(FkMesh m, const double* dtp, int flux_dir,
                           int recompute, double* mom_y, double* v,
                           const double* flux, const double* rho,
                           hipStream_t s);
/* omp3/flow.c:979-986: compensated FP64 interior sums of rho and rho*e
 * into r->sums[0..1]. One launch. */
hipError_t fk_conservation(FkMesh m, const double* rho, const double* e,
                           FkReduce r, hipStream_t s);

/* ---- fused schedule: four kernels per step (fused.hip) ----
 * Each kernel reads only arrays it does not write; the host schedule routes
 * the intermediate fields through scratch arrays. */
/* omp3/flow.c:22-38 (EOS, pressure acceleration, artificial viscosity,
 * shock heating, set_timestep): reads rho, e, mom_x, mom_y and the old
 * (dt, dt_h) in r.dtp; writes the momenta after the viscous step (mx1, my1,
 * interior faces), the velocities after it (u0, v0, with ghosts), the
 * energy after shock heating (e1, with ghosts), rho_old = rho, and the new
 * (dt, dt_h) into r.dtp. One launch. */
hipError_t fk_lagrangian(FkMesh m, const double* rho, const double* e,
                         const double* mom_x, const double* mom_y, double* mx1,
                         double* my1, double* u0, double* v0, double* e1,
                         double* rho_old, FkReduce r, double max_dt,
                         int first_step, hipStream_t s);
/* omp3/flow.c:293-554: both directional sweeps (x first for even tt),
 * rho0 = density before advection, e1 = energy after shock heating; writes
 * rho, e, F_x, F_y (with ghosts) and the conservation totals (omp3/flow.c:
 * 979-986) into r.sums. */
hipError_t fk_advect_mass_energy(FkMesh m, int tt, const double* u0,
                                 const double* v0, const double* rho0,
                                 const double* e1, double* rho, double* e,
                                 double* F_x, double* F_y, FkReduce r,
                                 hipStream_t s);
/* omp3/flow.c:570-619 (odd tt) / 673-722 (even tt): x momentum mx1 ->
 * mom_x, and the recomputed x velocity u (with ghosts). */
hipError_t fk_advect_momentum_x(FkMesh m, int tt, const double* u0,
                                const double* v0, const double* F_x,
                                const double* F_y, const double* rho,
                                const double* mx1, double* mom_x, double* u,
                                const double* dtp, hipStream_t s);
/* omp3/flow.c:621-669 (odd tt) / 724-774 (even tt): y momentum my1 ->
 * mom_y, and the recomputed y velocity v (with ghosts); u1 is the x
 * velocity fk_advect_momentum_x wrote. */
hipError_t fk_advect_momentum_y(FkMesh m, int tt, const double* u1,
                                const double* v0, const double* F_x,
                                const double* F_y, const double* rho,
                                const double* my1, double* mom_y, double* v,
                                const double* dtp, hipStream_t s);

/* ---- row-marching forms (march.hip) ----
 * Same contracts as the fused kernels above; one wave marches a strip of
 * columns down chunks of chunk_rows rows. */
hipError_t fk_march_mass_energy(FkMesh m, int tt, const double* u0,
                                const double* v0, const double* rho0,
                                const double* e1, double* rho, double* e,
                                double* F_x, double* F_y, FkReduce r,
                                int chunk_rows, hipStream_t s);
hipError_t fk_march_lagrangian(FkMesh m, const double* rho, const double* e,
                               const double* mom_x, const double* mom_y, double* mx1,
                               double* my1, double* u0, double* v0, double* e1,
                               double* rho_old, FkReduce r, double max_dt,
                               int first_step, int chunk_rows, hipStream_t s);
/* omp3/flow.c:557-775: both momentum directions in one row march (the y
 * direction reads the u the x direction has just computed in the same
 * wave); the result equals fk_march_momentum_x then _y. */
hipError_t fk_march_momentum(FkMesh m, int tt, const double* u0, const double* v0,
                             const double* F_x, const double* F_y, const double* rho,
                             const double* mx1, const double* my1, double* mom_x,
                             double* mom_y, double* u, double* v, FkReduce r,
                             int chunk_rows, hipStream_t s);
hipError_t fk_march_momentum_x(FkMesh m, int tt, const double* u0,
                               const double* v0, const double* F_x,
                               const double* F_y, const double* rho,
                               const double* mx1, double* mom_x, double* u,
                               const double* dtp, int chunk_rows, hipStream_t s);
hipError_t fk_march_momentum_y(FkMesh m, int tt, const double* u1,
                               const double* v0, const double* F_x,
                               const double* F_y, const double* rho,
                               const double* my1, double* mom_y, double* v,
                               const double* dtp, int chunk_rows, hipStream_t s);

/* ---- FP32-state precision form (march kernels over float arrays) ----
 * Same contracts as the FP64 launchers. The CFL candidates, the dt
 * reduction and the conservation totals are computed in FP64 from the FP32
 * state, the EOS in FP64 at near-vacuum cells (rho < 1e-6). */
hipError_t fk_timestep_f32(FkMesh32 m, const float* Qxx, const float* Qyy,
                           const float* rho, const float* e, FkReduce r,
                           double max_dt, int first_step, hipStream_t s);
hipError_t fk_conservation_f32(FkMesh32 m, const float* rho, const float* e,
                               FkReduce r, hipStream_t s);
hipError_t fk_march_lagrangian_f32(FkMesh32 m, const float* rho, const float* e,
                                   const float* mom_x, const float* mom_y,
                                   float* mx1, float* my1, float* u0, float* v0,
                                   float* e1, float* rho_old, FkReduce r,
                                   double max_dt, int first_step, int chunk_rows,
                                   hipStream_t s);
hipError_t fk_march_mass_energy_f32(FkMesh32 m, int tt, const float* u0,
                                    const float* v0, const float* rho0,
                                    const float* e1, float* rho, float* e,
                                    float* F_x, float* F_y, FkReduce r,
                                    int chunk_rows, hipStream_t s);
hipError_t fk_march_momentum_f32(FkMesh32 m, int tt, const float* u0, const float* v0,
                                 const float* F_x, const float* F_y, const float* rho,
                                 const float* mx1, const float* my1, float* mom_x,
                                 float* mom_y, float* u, float* v, FkReduce r,
                                 int chunk_rows, hipStream_t s);
hipError_t fk_march_momentum_x_f32(FkMesh32 m, int tt, const float* u0,
                                   const float* v0, const float* F_x,
                                   const float* F_y, const float* rho,
                                   const float* mx1, float* mom_x, float* u,
                                   const double* dtp, int chunk_rows,
                                   hipStream_t s);
hipError_t fk_march_momentum_y_f32(FkMesh32 m, int tt, const float* u1,
                                   const float* v0, const float* F_x,
                                   const float* F_y, const float* rho,
                                   const float* my1, float* mom_y, float* v,
                                   const double* dtp, int chunk_rows,
                                   hipStream_t s);
/* ---- ensemble (batching) form ----
 * One descriptor per instance; arrays and geometry are FP64 or FP32 (one
 * type for the whole ensemble). Work-list entries are (instance, item)
 * pairs, items numbered strip-fastest within their instance. */
typedef struct FkEnsInst {
  int nx, ny;              /* local sizes, pad included */
  const void* geo[6];      /* edgedx, edgedy, celldx, celldy, 1/edgedx, 1/edgedy */
  void* a[16];             /* solve_hydro_2d's arrays, in its order */
  double* dtp;             /* {dt, dt_h} of this instance (device) */
  double max_dt;
  int nstrips, nchunks, nitems;
  int item0;               /* first partial slot of this instance */
} FkEnsInst;

typedef struct FkEnsPartials {
  double* min;   /* per item: CFL minimum */
  int* nan;      /* per item: undefined CFL candidates */
  double* sums;  /* per item: compensated mass and energy (4 doubles) */
} FkEnsPartials;

typedef struct FkEnsHistory {
  double* dt;    /* [step][instance][2]: dt pair after the step */
  double* sums;  /* [step][instance][2]: mass and energy totals */
  int* nan;      /* [step][instance]: undefined CFL candidates */
  int* counter;  /* device step counter (the history row) */
} FkEnsHistory;

typedef struct FkEnsLaunch {
  const FkEnsInst* insts;  /* device, ninst descriptors */
  const int* work;         /* device, nwork (instance, item) pairs */
  int ninst, nwork, chunk_rows;
  int pow2;                /* every instance's spacings are powers of two */
  int uniform;             /* read each instance's spacings once per item
                              (they hold one value throughout) */
  int fused_momentum;      /* both momentum directions in one march */
  unsigned int* queue;     /* FK_MARCH_QUEUE_WORDS work-queue words, zero
                              between launches (each kernel resets them) */
  FkEnsPartials part;
  FkEnsHistory hist;
} FkEnsLaunch;

/* Strip width of the ensemble work items. */
int fk_ens_strip_width(void);
/* One step (omp3/flow.c:10-52) of every instance: Lagrangian stage, dt,
 * mass/energy advection, totals, x and y momentum, counter advance. */
hipError_t fk_ens_step(FkEnsLaunch L, int tt, int fp32, hipStream_t s);
/* The pre-loop set_timestep (first_step = 0) of every instance. */
hipError_t fk_ens_timestep(FkEnsLaunch L, int fp32, hipStream_t s);
hipError_t fk_ens_set_counter(int* counter, int value, hipStream_t s);

/* ---- the no-allocation step (noalloc.hip) ----
 * fk_static_scratch fills *r and the two reciprocal-spacing buffers (L+1
 * doubles each) from the current device's static scratch array, whose
 * device address it looks up once per device; max_side is the largest
 * interior side those buffers serve, bytes the array's size. Every caller
 * on one device gets the same memory. fk_static_begin zeroes r's counter
 * words (ticket, status, queue, totals) and sets r.dtp = {dt, dt_h}. */
hipError_t fk_static_scratch(FkReduce* r, double** inv_edgedx, double** inv_edgedy,
                             int* max_side, size_t* bytes);
hipError_t fk_static_begin(FkReduce r, double dt, double dt_h, hipStream_t s);

/* dst[k] = (float)src[k] / (double)src[k], k < n (round to nearest). */
hipError_t fk_narrow(long n, float* dst, const double* src, hipStream_t s);
hipError_t fk_widen(long n, double* dst, const float* src, hipStream_t s);
/* arch handle_boundary_2d on an FP32 array. */
hipError_t fk_boundary_f32(int nxp, int nyp, int pad, float* arr, int invert,
                           hipStream_t s);

#ifdef __cplusplus
}
#endif
#endif
