/*
 * noalloc.c -- flowhip_solve_hydro_2d_static: one solve_hydro_2d step
 * (omp3/flow.c:10-52) in the march schedule that allocates nothing. Its
 * scratch is the file-scope device array of kernels/noalloc.hip, shared by
 * every call on one GPU; the kernels and their order are those of
 * solve_hydro_2d's march schedule, so the results are the same bits.
 */
#include <math.h>
#include <stdio.h>
#include <string.h>

#include "schedules.h"

int flowhip_solve_hydro_2d_static(Mesh* mesh, int tt, double* P, double* rho,
                                  double* rho_old, double* e, double* u, double* v,
                                  double* mom_x, double* mom_y, double* Qxx,
                                  double* Qyy, double* F_x, double* F_y,
                                  double* uF_x, double* uF_y, double* vF_x,
                                  double* vF_y, double* mass, double* energy,
                                  void* stream_arg) {
  hipStream_t stream = (hipStream_t)stream_arg;
  const double* arrays[16] = {P, rho, rho_old, e, u, v, mom_x, mom_y,
                              Qxx, Qyy, F_x, F_y, uF_x, uF_y, vF_x, vF_y};
  if (!mesh) {
    fprintf(stderr, "flowhip_solve_hydro_2d_static: no mesh\n");
    return -1;
  }
  for (int k = 0; k < 16; ++k)
    if (!arrays[k]) {
      fprintf(stderr, "flowhip_solve_hydro_2d_static: array %d is NULL\n", k);
      return -1;
    }
  const int nx = mesh->local_nx, ny = mesh->local_ny, pad = mesh->pad;
  if (pad != 2 || nx - 2 * pad < 2 || ny - 2 * pad < 2) {
    fprintf(stderr, "flowhip_solve_hydro_2d_static: the mesh needs pad 2 and at "
                    "least 2x2 interior cells (got %dx%d, pad %d)\n",
            nx - 2 * pad, ny - 2 * pad, pad);
    return -1;
  }

  FlowHipCtx h;
  memset(&h, 0, sizeof(h));
  double *inv_x = NULL, *inv_y = NULL;
  int max_side = 0;
  hipError_t err = fk_static_scratch(&h.red, &inv_x, &inv_y, &max_side, NULL);
  if (err != hipSuccess) {
    fprintf(stderr, "flowhip_solve_hydro_2d_static: no device scratch: %s\n",
            hipGetErrorString(err));
    return -3;
  }
  if (nx - 2 * pad > max_side || ny - 2 * pad > max_side) {
    fprintf(stderr, "flowhip_solve_hydro_2d_static: %dx%d interior cells exceed "
                    "the static scratch's %d a side\n",
            nx - 2 * pad, ny - 2 * pad, max_side);
    return -2;
  }
  if (mesh->rank == MASTER && mesh->verbose) printf("Timestep:        %.12e\n", mesh->dt);

  /* the march schedule with its default chunk heights, the defined CFL
   * rule and IEEE quotients (no power-of-two probe: bit for bit the same) */
  h.stream = stream;
  h.schedule = FLOWHIP_MARCH;
  h.cfl_rule = FLOWHIP_CFL_DEFINED;
  FlowHipStep st;
  st.m.nx = nx;
  st.m.ny = ny;
  st.m.pad = pad;
  st.m.edgedx = mesh->edgedx;
  st.m.edgedy = mesh->edgedy;
  st.m.celldx = mesh->celldx;
  st.m.celldy = mesh->celldy;
  st.m.inv_edgedx = inv_x;
  st.m.inv_edgedy = inv_y;
  st.m.pow2 = 0;
  st.m.uniform = 0;
  st.m.pairs = 0;
  st.tt = tt;
  st.max_dt = mesh->max_dt;
  st.P = P;
  st.rho = rho;
  st.rho_old = rho_old;
  st.e = e;
  st.u = u;
  st.v = v;
  st.mom_x = mom_x;
  st.mom_y = mom_y;
  st.Qxx = Qxx;
  st.Qyy = Qyy;
  st.F_x = F_x;
  st.F_y = F_y;
  st.uF_x = uF_x;
  st.uF_y = uF_y;
  st.vF_x = vF_x;
  st.vF_y = vF_y;

  HIPCHECK(fk_static_begin(h.red, mesh->dt, mesh->dt_h, stream));
  HIPCHECK(fk_reciprocal(nx + 1, inv_x, mesh->edgedx, stream));
  HIPCHECK(fk_reciprocal(ny + 1, inv_y, mesh->edgedy, stream));
  flowhip_step_fused(&h, &st);

  /* dt pair and totals are adjacent in the scratch: one copy */
  double out[4];
  HIPCHECK(hipMemcpyAsync(out, h.red.dtp, sizeof(out), hipMemcpyDeviceToHost, stream));
  HIPCHECK(hipStreamSynchronize(stream));
  mesh->dt = out[0];
  mesh->dt_h = out[1];
  if (mass) *mass = out[2];
  if (energy) *energy = out[3];
  if (!isfinite(mesh->dt) || !(mesh->dt > 0.0) || !isfinite(mesh->dt_h)) {
    fprintf(stderr, "flowhip_solve_hydro_2d_static: non-finite or non-positive "
                    "timestep (dt=%.12e dt_h=%.12e): the state has diverged\n",
            mesh->dt, mesh->dt_h);
    return -4;
  }
  return 0;
}
