/*
 * flowhip_ctx.h -- private per-Mesh state of the hip kernel set: its stream,
 * the device-resident timestep pair, reduction scratch, pinned staging for
 * the one device-to-host transfer per step, captured step graphs and the
 * per-kernel event timers.
 */
#ifndef FLOWHIP_CTX_H
#define FLOWHIP_CTX_H

#include <hip/hip_runtime_api.h>

#include "../../../include/flow_arch.h"
#include "../../../include/flow_hip.h"
#include "../../../include/flow_hip_kernels.h"

#define FLOWHIP_MAX_TIMERS 48
#define FLOWHIP_MAX_PENDING 256

#define HIPCHECK(call)                                                     \
  do {                                                                     \
    hipError_t err_ = (call);                                              \
    if (err_ != hipSuccess)                                                \
      TERMINATE("hip kernel set: %s failed at %s:%d: %s", #call, __FILE__, \
                __LINE__, hipGetErrorString(err_));                        \
  } while (0)

typedef struct FlowHipTimer {
  const char* name;
  double total_ms;
  long launches;
} FlowHipTimer;

/* One step's launch arguments (omp3/flow.c:10-52's arguments on the
 * device side). */
typedef struct FlowHipStep {
  FkMesh m;
  int tt;
  double max_dt;
  double *P, *rho, *rho_old, *e, *u, *v, *mom_x, *mom_y, *Qxx, *Qyy;
  double *F_x, *F_y, *uF_x, *uF_y, *vF_x, *vF_y;
} FlowHipStep;

typedef struct FlowHipCtx {
  int device;
  hipStream_t stream;
  FkReduce red;         /* device scratch */
  double *inv_edgedx, *inv_edgedy; /* 1/edgedx, 1/edgedy of the mesh */
  const double *inv_src_x, *inv_src_y, *inv_src_cx, *inv_src_cy;
  int inv_nx, inv_ny;
  int pow2, pow2_f32; /* every spacing a power of two (FP64 / FP32 range) */
  int uniform;        /* each spacing array holds one value throughout */
  int no_pow2;        /* keep the quotients anyway */
  int no_uniform;     /* keep the per-row geometry loads anyway */
  int pairs;          /* FkMesh.pairs: 0 never, 1 large meshes, 2 always */
  /* FP32 copies of the geometry for the precision form: edgedx, edgedy,
   * celldx, celldy, 1/edgedx, 1/edgedy, each (L+1) floats */
  float* geo32;
  const double* geo32_src[4];
  int geo32_nx, geo32_ny;
  double* h_stage;      /* pinned: [0..1] dt pair, [2..5] sums */
  int* h_status;        /* pinned */
  int schedule;
  int march_rows;       /* chunk height of every march kernel; 0: each its own */
  /* rule for undefined CFL candidates (FLOWHIP_CFL_*); under the serial
   * rule the fused Lagrangian kernels store every candidate in cfl_buf
   * (cfl_len doubles, cell layout) for fk_timestep_serial */
  int cfl_rule;
  double* cfl_buf;
  size_t cfl_len;
  int use_graphs;
  /* captured steps: [0] even tt, [1] odd tt, [2] tt == 0; each valid for
   * the launch arguments, schedule and chunk height it was captured with */
  hipGraphExec_t graph[3];
  int graph_ok[3];
  FlowHipStep graph_step[3];
  int graph_sched[3], graph_rows[3], graph_rule[3];
  int graph_warm;
  /* conservation sums fused into the last advection kernel, valid while
   * the arrays they were taken from are unchanged */
  int sums_valid;
  const double* sums_rho;
  const double* sums_e;
  /* timing */
  int timing;
  int ntimers;
  FlowHipTimer timers[FLOWHIP_MAX_TIMERS];
  int npending;
  int pending_timer[FLOWHIP_MAX_PENDING];
  hipEvent_t ev_start[FLOWHIP_MAX_PENDING], ev_stop[FLOWHIP_MAX_PENDING];
  int events_created;
} FlowHipCtx;

FlowHipCtx* flowhip_ctx(Mesh* mesh);
void flowhip_ctx_destroy(Mesh* mesh);
FkMesh flowhip_fkmesh(const Mesh* mesh);
FkMesh32 flowhip_fkmesh32(Mesh* mesh);

/* Timing brackets around one launch; no-ops unless timing is on. */
int flowhip_timer_begin(FlowHipCtx* h, const char* name);
void flowhip_timer_end(FlowHipCtx* h, int slot);
/* Folds completed event pairs into the timers; call after a sync. */
void flowhip_timer_collect(FlowHipCtx* h);

/* Launch with the kernel set's error check and optional timing. */
#define FK_LAUNCH(h, name, call)                  \
  do {                                            \
    const int slot_ = flowhip_timer_begin(h, name); \
    HIPCHECK(call);                               \
    flowhip_timer_end(h, slot_);                  \
  } while (0)

/* Invalidate cached conservation sums (any write to device state). */
void flowhip_invalidate_sums(Mesh* mesh);

/* The reduction scratch for an FP64 Lagrangian launch on an nx x ny
 * (local) mesh: h->red, with .cfl pointing at the candidate buffer (grown
 * to nx * ny doubles) under the serial CFL rule. */
FkReduce flowhip_lagrangian_red(FlowHipCtx* h, int nx, int ny);

/* Chunk height of one march kernel. An explicit height (flowhip_set_march_rows,
 * FLOWHIP_MARCH_ROWS) gives uniform chunks of that height to all of them;
 * 0, the default, lets each kernel use the chunking measured fastest for it
 * on MI355X (march.hip). Every chunking gives bitwise the same result. */
enum { FLOWHIP_K_LAGRANGIAN = 0, FLOWHIP_K_MASS = 1, FLOWHIP_K_MOMENTUM = 2 };
static inline int flowhip_rows(const FlowHipCtx* h, int kernel, int fp32) {
  (void)kernel;
  (void)fp32;
  return h->march_rows;
}

#endif
