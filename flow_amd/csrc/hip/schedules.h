/*
 * schedules.h -- the kernel sequences that make up one solve_hydro_2d step.
 */
#ifndef FLOWHIP_SCHEDULES_H
#define FLOWHIP_SCHEDULES_H

#include "flowhip_ctx.h"

typedef struct FlowHipStep {
  FkMesh m;
  int tt;
  double max_dt;
  double *P, *rho, *rho_old, *e, *u, *v, *mom_x, *mom_y, *Qxx, *Qyy;
  double *F_x, *F_y, *uF_x, *uF_y, *vF_x, *vF_y;
} FlowHipStep;

/* Enqueue one step on h->stream (device dt pair already current). */
void flowhip_enqueue_step(FlowHipCtx* h, const FlowHipStep* st);

/* The two schedules. */
void flowhip_step_stagewise(FlowHipCtx* h, const FlowHipStep* st);
void flowhip_step_fused(FlowHipCtx* h, const FlowHipStep* st);

void flowhip_last_sums(Mesh* mesh, double* mass, double* energy);

#endif
