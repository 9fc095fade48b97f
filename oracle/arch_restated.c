/*
 * arch_restated.c -- the UoB-HPC/arch behaviours that flow's omp3 path calls,
 * restated for the oracle. TEST INFRASTRUCTURE ONLY (see oracle.h).
 *
 * UoB-HPC/arch is an un-vendored sibling project (flow README.md:10-17 clones
 * its default branch; no version is pinned anywhere). Only its call sites are
 * in the reference: minmod/min (omp3/flow.c:94-95,177-197,353,395,...),
 * handle_boundary_2d (19 sites per step, e.g. omp3/flow.c:156-157), mesh
 * geometry (main.c:40), problem rasterisation (main.c:52-54). The behaviours
 * below are arch's published algorithm as used by those call sites; the
 * rasterisation rule is pinned by flow.tests:1 (the interior mass of
 * flow.params requires 819x819 cells for problem_1).
 */
#include <math.h>
#include <stdlib.h>

#include "oracle.h"

/* arch shared.h: minmod(a,b) = samesign(a,b) ? absmin(a,b) : 0.0, with
 * samesign(a,b) = (a*b) > 0.0 and absmin(a,b) = |a| < |b| ? a : b. */
double orc_minmod(double a, double b) {
  if (!(a * b > 0.0)) return 0.0;
  return (fabs(a) < fabs(b)) ? a : b;
}

/* arch shared.h: min(a,b) = a < b ? a : b (NaN only propagates from b). */
double orc_min(double a, double b) { return (a < b) ? a : b; }

/* arch mesh.c, uniform rectilinear geometry (main.c:40): every spacing is
 * width/global_nx; edgex[k] is the left edge of local cell k, offset by pad.
 * One rank, so the x/y offsets are zero. */
void orc_mesh_init(OrcMesh* m, int global_nx, int global_ny, double width,
                   double height, double max_dt, double sim_end,
                   double dt_param) {
  m->pad = 2;
  m->global_nx = global_nx;
  m->global_ny = global_ny;
  m->nx = global_nx + 2 * m->pad;
  m->ny = global_ny + 2 * m->pad;
  m->width = width;
  m->height = height;
  m->max_dt = max_dt;
  m->sim_end = sim_end;
  m->dt = 0.3 * dt_param; /* main.c:30: C_T * dt */
  m->dt_h = m->dt;        /* main.c:31 */
  m->verbose = 0;
  const int lx = m->nx + 1, ly = m->ny + 1;
  m->edgex = (double*)calloc(lx, sizeof(double));
  m->edgedx = (double*)calloc(lx, sizeof(double));
  m->celldx = (double*)calloc(lx, sizeof(double));
  m->edgey = (double*)calloc(ly, sizeof(double));
  m->edgedy = (double*)calloc(ly, sizeof(double));
  m->celldy = (double*)calloc(ly, sizeof(double));
  for (int k = 0; k < lx; ++k) {
    m->edgedx[k] = width / global_nx;
    m->celldx[k] = width / global_nx;
    m->edgex[k] = m->edgedx[k] * (double)(k - m->pad);
  }
  for (int k = 0; k < ly; ++k) {
    m->edgedy[k] = height / global_ny;
    m->celldy[k] = height / global_ny;
    m->edgey[k] = m->edgedy[k] * (double)(k - m->pad);
  }
}

void orc_mesh_free(OrcMesh* m) {
  free(m->edgex);
  free(m->edgey);
  free(m->edgedx);
  free(m->edgedy);
  free(m->celldx);
  free(m->celldy);
  m->edgex = m->edgey = m->edgedx = m->edgedy = m->celldx = m->celldy = 0;
}

/* arch handle_boundary_2d on one rank (all four neighbours are the domain
 * edge): pad-deep mirror strips. North/south strips span columns
 * [pad, nx-pad), east/west strips span rows [pad, ny-pad); corners are not
 * touched. INVERT_X negates the east/west copies, INVERT_Y the north/south
 * copies. nx, ny are the staggered extents of the array being reflected. */
void orc_handle_boundary_2d(int nx, int ny, int pad, double* arr, int invert) {
  const double sx = (invert == ORC_INVERT_X) ? -1.0 : 1.0;
  const double sy = (invert == ORC_INVERT_Y) ? -1.0 : 1.0;
  for (int d = 0; d < pad; ++d) {
    for (int j = pad; j < nx - pad; ++j) {
      arr[(ny - pad + d) * nx + j] = sy * arr[(ny - 1 - pad - d) * nx + j];
      arr[(pad - 1 - d) * nx + j] = sy * arr[(pad + d) * nx + j];
    }
  }
  for (int i = pad; i < ny - pad; ++i) {
    for (int d = 0; d < pad; ++d) {
      arr[i * nx + (nx - pad + d)] = sx * arr[i * nx + (nx - 1 - pad - d)];
      arr[i * nx + (pad - 1 - d)] = sx * arr[i * nx + (pad + d)];
    }
  }
}

/* arch set_problem_2d for one "problem_N" entry (main.c:52-54): bounds are
 * fractions of the domain; a local cell (ghosts included) takes the values
 * when its left/bottom edge lies in [pos, pos+extent). */
void orc_rasterise_box(const OrcMesh* m, double* density, double* energy,
                       double rho, double e, double xpos, double ypos,
                       double w, double h) {
  const double x0 = xpos * m->width, y0 = ypos * m->height;
  const double wx = w * m->width, hy = h * m->height;
  for (int i = 0; i < m->ny; ++i) {
    for (int j = 0; j < m->nx; ++j) {
      const double gx = m->edgex[j], gy = m->edgey[i];
      if (gx >= x0 && gy >= y0 && gx < x0 + wx && gy < y0 + hy) {
        density[i * m->nx + j] = rho;
        energy[i * m->nx + j] = e;
      }
    }
  }
}
