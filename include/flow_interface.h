/*
 * flow_interface.h -- the drop-in boundary of the hip kernel set.
 *
 * Same three C-linkage entry points, argument meanings and constants as the
 * reference's flow_interface.h:7-30. The hip kernel set (libflowhip.so)
 * implements them on device arrays; the driver (main.c:64-91) calls them
 * unchanged. nx/ny are local sizes including 2*pad.
 */
#ifndef FLOW_INTERFACE_H
#define FLOW_INTERFACE_H

#include "flow_arch.h"

#define GAM 1.4          /* flow_interface.h:7 */
#define C_Q 3.0          /* flow_interface.h:8 */
#ifndef C_T
#define C_T 0.3          /* flow_data.h:6 */
#endif
#define C_M (1.5 / C_T)  /* flow_interface.h:9 */

#ifdef __cplusplus
extern "C" {
#endif

/* Replaces flow_interface.h:16-20 (omp3/flow.c:10-52). Advances every
 * array one timestep in place and leaves the new (dt, dt_h) in mesh on the
 * host. Parity of tt selects the sweep order. reduce_array is unused
 * (the kernel set keeps its own reduction scratch). */
void solve_hydro_2d(Mesh* mesh, int tt, double* P, double* rho,
                    double* rho_old, double* e, double* u, double* v,
                    double* rho_u, double* rho_v, double* Qxx, double* Qyy,
                    double* F_x, double* F_y, double* uF_x, double* uF_y,
                    double* vF_x, double* vF_y, double* reduce_array);

/* Replaces flow_interface.h:23-26 (omp3/flow.c:73-106). Grid-wide CFL
 * minimum on the device, then the (dt, dt_h) update on the host copy in
 * mesh. celldx/celldy are device pointers. */
void set_timestep(const int nx, const int ny, double* Qxx, double* Qyy,
                  const double* rho, const double* e, Mesh* mesh,
                  double* reduce_array, const int first_step,
                  const double* celldx, const double* celldy);

/* Replaces flow_interface.h:29-30 (omp3/flow.c:973-995). Interior sums of
 * rho and rho*e (FP64, compensated), printed by rank MASTER. */
void print_conservation(const int nx, const int ny, double* rho, double* e,
                        double* reduce_array, Mesh* mesh);

#ifdef __cplusplus
}
#endif
#endif
