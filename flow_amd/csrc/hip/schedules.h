/*
 * schedules.h -- the kernel sequences that make up one solve_hydro_2d step.
 */
#ifndef FLOWHIP_SCHEDULES_H
#define FLOWHIP_SCHEDULES_H

#include "flowhip_ctx.h"

/* Enqueue one step on h->stream (device dt pair already current); with
 * graphs on, replayed from the step graph of its kind. */
void flowhip_enqueue_step(FlowHipCtx* h, const FlowHipStep* st);

/* The two schedules. */
void flowhip_step_stagewise(FlowHipCtx* h, const FlowHipStep* st);
void flowhip_step_fused(FlowHipCtx* h, const FlowHipStep* st);

void flowhip_last_sums(Mesh* mesh, double* mass, double* energy);

#endif
