/* Placeholder until the fused kernels are built: selecting the fused
 * schedule is a fatal error, never a silent fallback. */
#include "schedules.h"

void flowhip_step_fused(FlowHipCtx* h, const FlowHipStep* st) {
  (void)h;
  (void)st;
  TERMINATE("hip kernel set: fused schedule not available in this build");
}
