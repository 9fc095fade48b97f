/*
 * stagewise.hip -- the hip kernel set's stagewise schedule: one kernel per
 * omp3 loop, one thread per point, plus the on-device reflective boundary.
 *
 * This schedule keeps every intermediate array of omp3 (P, Q, rho_old,
 * fluxes) in device memory. It is the reference-shaped path that the fused
 * schedule (fused_*.hip) is checked against, and it serves any caller that
 * needs those intermediates.
 */
#include "common.h"
#include "hydro_math.h"

namespace flowhip {
namespace {

#define SW_INDEX(d)                                         \
  const int j = (d).c0 + blockIdx.x * kSwX + threadIdx.x;   \
  const int i = (d).r0 + blockIdx.y * kSwY + threadIdx.y;   \
  if (i >= (d).r1 || j >= (d).c1) return;

/* arch handle_boundary_2d on one rank: pad-deep mirror strips; north/south
 * over columns [pad, nxp-pad), east/west over rows [pad, nyp-pad); corners
 * untouched. One thread per ghost element, all four strips in one grid. */
template <typename T>
__global__ void k_boundary(T* __restrict__ a, int nxp, int nyp, int pad, T sx,
                           T sy) {
  const int wx = nxp - 2 * pad, wy = nyp - 2 * pad;
  const int n_ns = pad * wx, n_ew = pad * wy;
  int t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t < 2 * n_ns) {
    const bool north = t < n_ns;
    if (!north) t -= n_ns;
    const int d = t / wx, j = pad + t % wx;
    if (north)
      a[(nyp - pad + d) * nxp + j] = sy * a[(nyp - 1 - pad - d) * nxp + j];
    else
      a[(pad - 1 - d) * nxp + j] = sy * a[(pad + d) * nxp + j];
    return;
  }
  t -= 2 * n_ns;
  if (t < 2 * n_ew) {
    const bool east = t < n_ew;
    if (!east) t -= n_ew;
    const int i = pad + t / pad, d = t % pad;
    if (east)
      a[i * nxp + (nxp - pad + d)] = sx * a[i * nxp + (nxp - 1 - pad - d)];
    else
      a[i * nxp + (pad - 1 - d)] = sx * a[i * nxp + (pad + d)];
  }
}

__global__ void k_reciprocal(int n, double* __restrict__ dst,
                             const double* __restrict__ src) {
  const int k = blockIdx.x * blockDim.x + threadIdx.x;
  if (k < n) dst[k] = 1.0 / src[k];
}

/* dst = src converted to the destination type (round to nearest when
 * narrowing); a plain copy when the types agree. */
template <typename D, typename S>
__global__ void k_copy(long n, D* __restrict__ dst, const S* __restrict__ src) {
  const long k = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (k < n) dst[k] = D(src[k]);
}

__global__ void k_eos(Domain d, int nx, double* __restrict__ P,
                      const double* __restrict__ rho,
                      const double* __restrict__ e) {
  SW_INDEX(d);
  const int c = i * nx + j;
  P[c] = eos_pressure(rho[c], e[c]);
}

/* Face velocities from momenta on the FXY domain (omp3/flow.c:133-150). */
__device__ __forceinline__ void face_velocities(FkMesh m, int i, int j,
                                                const double* rho,
                                                const double* mom_x,
                                                const double* mom_y, double* u,
                                                double* v) {
  const int nx = m.nx, c = i * nx + j, f = i * (nx + 1) + j;
  const double rex = rho_edge_x(rho[c], rho[c - 1], m.celldx[j],
                                m.celldx[j - 1], m.celldy[i], m.edgedx[j]);
  const double rey = rho_edge_y(rho[c], rho[c - nx], m.celldx[j],
                                m.celldy[i], m.celldy[i - 1], m.edgedy[i]);
  u[f] = velocity(mom_x[f], rex);
  v[c] = velocity(mom_y[c], rey);
}

__global__ void k_pressure_accel(FkMesh m, Domain d, const double* dtp,
                                 double* mom_x, double* mom_y, double* u,
                                 double* v, const double* __restrict__ P,
                                 const double* __restrict__ rho) {
  SW_INDEX(d);
  const double dt = dtp[0];
  const int nx = m.nx, c = i * nx + j, f = i * (nx + 1) + j;
  mom_x[f] = accelerate(mom_x[f], dt, P[c], P[c - 1], m.edgedx[j]);
  mom_y[c] = accelerate(mom_y[c], dt, P[c], P[c - nx], m.edgedy[i]);
  face_velocities(m, i, j, rho, mom_x, mom_y, u, v);
}

__global__ void k_av_stress(FkMesh m, Domain d, double* __restrict__ Qxx,
                            double* __restrict__ Qyy,
                            const double* __restrict__ u,
                            const double* __restrict__ v,
                            const double* __restrict__ rho) {
  SW_INDEX(d);
  const int nx = m.nx, c = i * nx + j, f = i * (nx + 1) + j;
  Qxx[c] = viscous_stress(rho[c], u[f - 1], u[f], u[f + 1], u[f + 2]);
  Qyy[c] = viscous_stress(rho[c], v[c - nx], v[c], v[c + nx], v[c + 2 * nx]);
}

__global__ void k_av_accel(FkMesh m, Domain d, const double* dtp,
                           double* mom_x, double* mom_y, double* u, double* v,
                           const double* __restrict__ Qxx,
                           const double* __restrict__ Qyy,
                           const double* __restrict__ rho) {
  SW_INDEX(d);
  const double dt = dtp[0];
  const int nx = m.nx, c = i * nx + j, f = i * (nx + 1) + j;
  mom_x[f] = accelerate(mom_x[f], dt, Qxx[c], Qxx[c - 1], m.celldx[j]);
  mom_y[c] = accelerate(mom_y[c], dt, Qyy[c], Qyy[c - nx], m.celldy[i]);
  face_velocities(m, i, j, rho, mom_x, mom_y, u, v);
}

__global__ void k_shock_heating(FkMesh m, Domain d, const double* dtp,
                                double* __restrict__ e,
                                const double* __restrict__ P,
                                const double* __restrict__ u,
                                const double* __restrict__ v,
                                const double* __restrict__ rho,
                                const double* __restrict__ Qxx,
                                const double* __restrict__ Qyy) {
  SW_INDEX(d);
  const double dt_h = dtp[1];
  const int nx = m.nx, c = i * nx + j, f = i * (nx + 1) + j;
  e[c] = shock_heating(e[c], P[c], rho[c], Qxx[c], Qyy[c], u[f], u[f + 1],
                       v[c], v[c + nx], m.celldx[j], m.celldy[i], dt_h);
}

__global__ void k_x_flux(FkMesh m, Domain d, const double* dtp,
                         const double* __restrict__ rho,
                         const double* __restrict__ e,
                         const double* __restrict__ u,
                         double* __restrict__ F_x, double* __restrict__ eF_x) {
  SW_INDEX(d);
  const double dt = dtp[0], dt_h = dtp[1];
  const int nx = m.nx, c = i * nx + j, f = i * (nx + 1) + j;
  const FaceFlux<double> fl = mass_energy_flux<double, false>(
      u[f - 1], u[f], u[f + 1], rho[c - 2], rho[c - 1], rho[c], rho[c + 1],
      e[c - 2], e[c - 1], e[c], e[c + 1], 1.0 / m.edgedx[j], m.celldx[j - 1],
      m.celldx[j], m.edgedy[i], dt, dt_h);
  F_x[f] = fl.mass;
  eF_x[f] = fl.energy;
}

__global__ void k_y_flux(FkMesh m, Domain d, const double* dtp,
                         const double* __restrict__ rho,
                         const double* __restrict__ e,
                         const double* __restrict__ v,
                         double* __restrict__ F_y, double* __restrict__ eF_y) {
  SW_INDEX(d);
  const double dt = dtp[0], dt_h = dtp[1];
  const int nx = m.nx, c = i * nx + j;
  const FaceFlux<double> fl = mass_energy_flux<double, true>(
      v[c - nx], v[c], v[c + nx], rho[c - 2 * nx], rho[c - nx], rho[c],
      rho[c + nx], e[c - 2 * nx], e[c - nx], e[c], e[c + nx], 1.0 / m.edgedy[i],
      m.celldy[i - 1], m.celldy[i], m.edgedx[j], dt, dt_h);
  F_y[c] = fl.mass;
  eF_y[c] = fl.energy;
}

__global__ void k_x_update(FkMesh m, Domain d, const double* dtp, int first,
                           double* __restrict__ rho,
                           const double* __restrict__ rho_old,
                           double* __restrict__ e,
                           const double* __restrict__ F_x,
                           const double* __restrict__ eF_x) {
  SW_INDEX(d);
  const double dt_h = dtp[1];
  const int nx = m.nx, c = i * nx + j, f = i * (nx + 1) + j;
  const CellState<double> s = sweep_update(
      rho[c], rho_old[c], e[c], F_x[f], F_x[f + 1], eF_x[f], eF_x[f + 1],
      m.edgedy[i], m.edgedy[i + 1], m.celldx[j], m.celldy[i], dt_h, first);
  rho[c] = s.rho;
  e[c] = s.e;
}

__global__ void k_y_update(FkMesh m, Domain d, const double* dtp, int first,
                           double* __restrict__ rho,
                           const double* __restrict__ rho_old,
                           double* __restrict__ e,
                           const double* __restrict__ F_y,
                           const double* __restrict__ eF_y) {
  SW_INDEX(d);
  const double dt_h = dtp[1];
  const int nx = m.nx, c = i * nx + j;
  const CellState<double> s = sweep_update(
      rho[c], rho_old[c], e[c], F_y[c], F_y[c + nx], eF_y[c], eF_y[c + nx],
      m.edgedx[j], m.edgedx[j + 1], m.celldx[j], m.celldy[i], dt_h, first);
  rho[c] = s.rho;
  e[c] = s.e;
}

/* omp3/flow.c:790-820: cell (i, j), faces j-1..j+2 of row i. */
__global__ void k_mom_flux_xx(FkMesh m, Domain d, const double* dtp,
                              const double* __restrict__ u,
                              double* __restrict__ uF_x,
                              const double* __restrict__ F_x) {
  SW_INDEX(d);
  const double dt_h = dtp[1];
  const int nx = m.nx, c = i * nx + j, f = i * (nx + 1) + j;
  const double u_cell = 0.5 * (u[f] + u[f + 1]);
  uF_x[c] = momentum_flux(m.edgedy[i], F_x[f], F_x[f + 1], u[f - 1], u[f],
                          u[f + 1], u[f + 2], 1.0 / m.edgedx[j], u_cell,
                          m.celldx[j - 1], m.celldx[j], dt_h);
}

/* omp3/flow.c:837-871: corner (i, j), rows i-2..i+1 of face column j. */
__global__ void k_mom_flux_xy(FkMesh m, Domain d, const double* dtp,
                              const double* __restrict__ u,
                              const double* __restrict__ v,
                              double* __restrict__ uF_y,
                              const double* __restrict__ F_y) {
  SW_INDEX(d);
  const double dt_h = dtp[1];
  const int nx = m.nx, sx = nx + 1, c = i * nx + j, f = i * sx + j;
  const double v_cell = 0.5 * (v[c - 1] + v[c]);
  uF_y[f] = momentum_flux(m.edgedx[j], F_y[c], F_y[c - 1], u[f - 2 * sx],
                          u[f - sx], u[f], u[f + sx], 1.0 / m.edgedy[i], v_cell,
                          m.celldy[i - 1], m.celldy[i], dt_h);
}

/* omp3/flow.c:890-922: corner (i, j), cells j-2..j+1 of row i. */
__global__ void k_mom_flux_yx(FkMesh m, Domain d, const double* dtp,
                              const double* __restrict__ u,
                              const double* __restrict__ v,
                              double* __restrict__ vF_x,
                              const double* __restrict__ F_x) {
  SW_INDEX(d);
  const double dt_h = dtp[1];
  const int nx = m.nx, sx = nx + 1, c = i * nx + j, f = i * sx + j;
  const double u_cell = 0.5 * (u[f] + u[f - sx]);
  vF_x[f] = momentum_flux(m.celldy[i], F_x[f], F_x[f - sx], v[c - 2],
                          v[c - 1], v[c], v[c + 1], 1.0 / m.edgedx[j], u_cell,
                          m.celldx[j - 1], m.celldx[j], dt_h);
}

/* omp3/flow.c:939-967: cell (i, j), y faces i-1..i+2 of column j. */
__global__ void k_mom_flux_yy(FkMesh m, Domain d, const double* dtp,
                              const double* __restrict__ v,
                              double* __restrict__ vF_y,
                              const double* __restrict__ F_y) {
  SW_INDEX(d);
  const double dt_h = dtp[1];
  const int nx = m.nx, c = i * nx + j;
  const double v_cell = 0.5 * (v[c] + v[c + nx]);
  vF_y[c] = momentum_flux(m.celldx[j], F_y[c], F_y[c + nx], v[c - nx], v[c],
                          v[c + nx], v[c + 2 * nx], 1.0 / m.edgedy[i], v_cell,
                          m.celldy[i - 1], m.celldy[i], dt_h);
}

__global__ void k_mom_x_update(FkMesh m, Domain d, const double* dtp,
                               int flux_dir, int recompute,
                               double* __restrict__ mom_x,
                               double* __restrict__ u,
                               const double* __restrict__ flux,
                               const double* __restrict__ rho) {
  SW_INDEX(d);
  const double dt_h = dtp[1];
  const int nx = m.nx, sx = nx + 1, c = i * nx + j, f = i * sx + j;
  double mx;
  if (flux_dir == 0)
    mx = advect_update(mom_x[f], dt_h, flux[c], flux[c - 1], m.edgedx[j],
                       m.celldy[i]);
  else
    mx = advect_update(mom_x[f], dt_h, flux[f + sx], flux[f], m.celldx[j],
                       m.edgedy[i]);
  mom_x[f] = mx;
  if (recompute) {
    const double rex = rho_edge_x(rho[c], rho[c - 1], m.celldx[j],
                                  m.celldx[j - 1], m.celldy[i], m.edgedx[j]);
    u[f] = velocity(mx, rex);
  }
}

__global__ void k_mom_y_update(FkMesh m, Domain d, const double* dtp,
                               int flux_dir, int recompute,
                               double* __restrict__ mom_y,
                               double* __restrict__ v,
                               const double* __restrict__ flux,
                               const double* __restrict__ rho) {
  SW_INDEX(d);
  const double dt_h = dtp[1];
  const int nx = m.nx, sx = nx + 1, c = i * nx + j, f = i * sx + j;
  double my;
  if (flux_dir == 0)
    my = advect_update(mom_y[c], dt_h, flux[f + 1], flux[f], m.edgedx[j],
                       m.celldy[i]);
  else
    my = advect_update(mom_y[c], dt_h, flux[c], flux[c - nx], m.celldx[j],
                       m.edgedy[i]);
  mom_y[c] = my;
  if (recompute) {
    const double rey = rho_edge_y(rho[c], rho[c - nx], m.celldx[j],
                                  m.celldy[i], m.celldy[i - 1], m.edgedy[i]);
    v[c] = velocity(my, rey);
  }
}

/* Loop domains of omp3: interior cells I = [pad, L-pad)^2, x faces FX
 * (columns up to L+1-pad), y faces FY (rows up to L+1-pad), and
 * FXY = [pad, L+1-pad)^2. */
inline Domain dom_all(FkMesh m) { return Domain{0, m.ny, 0, m.nx}; }
inline Domain dom_I(FkMesh m) {
  return Domain{m.pad, m.ny - m.pad, m.pad, m.nx - m.pad};
}
inline Domain dom_FX(FkMesh m) {
  return Domain{m.pad, m.ny - m.pad, m.pad, m.nx + 1 - m.pad};
}
inline Domain dom_FY(FkMesh m) {
  return Domain{m.pad, m.ny + 1 - m.pad, m.pad, m.nx - m.pad};
}
inline Domain dom_FXY(FkMesh m) {
  return Domain{m.pad, m.ny + 1 - m.pad, m.pad, m.nx + 1 - m.pad};
}

const dim3 kSwBlock(kSwX, kSwY, 1);

}  // namespace
}  // namespace flowhip

using namespace flowhip;

extern "C" {

hipError_t fk_boundary(int nxp, int nyp, int pad, double* arr, int invert,
                       hipStream_t s) {
  const int n = 2 * pad * (nxp - 2 * pad) + 2 * pad * (nyp - 2 * pad);
  if (n <= 0) return hipSuccess;
  const double sx = (invert == 1) ? -1.0 : 1.0;
  const double sy = (invert == 2) ? -1.0 : 1.0;
  hipLaunchKernelGGL(k_boundary<double>, dim3((n + 255) / 256), dim3(256), 0, s,
                     arr, nxp, nyp, pad, sx, sy);
  return hipGetLastError();
}

hipError_t fk_boundary_f32(int nxp, int nyp, int pad, float* arr, int invert,
                           hipStream_t s) {
  const int n = 2 * pad * (nxp - 2 * pad) + 2 * pad * (nyp - 2 * pad);
  if (n <= 0) return hipSuccess;
  const float sx = (invert == 1) ? -1.0f : 1.0f;
  const float sy = (invert == 2) ? -1.0f : 1.0f;
  hipLaunchKernelGGL(k_boundary<float>, dim3((n + 255) / 256), dim3(256), 0, s,
                     arr, nxp, nyp, pad, sx, sy);
  return hipGetLastError();
}

hipError_t fk_narrow(long n, float* dst, const double* src, hipStream_t s) {
  if (n <= 0) return hipSuccess;
  hipLaunchKernelGGL((k_copy<float, double>), dim3((unsigned)((n + 255) / 256)),
                     dim3(256), 0, s, n, dst, src);
  return hipGetLastError();
}

hipError_t fk_widen(long n, double* dst, const float* src, hipStream_t s) {
  if (n <= 0) return hipSuccess;
  hipLaunchKernelGGL((k_copy<double, float>), dim3((unsigned)((n + 255) / 256)),
                     dim3(256), 0, s, n, dst, src);
  return hipGetLastError();
}

hipError_t fk_reciprocal(int n, double* dst, const double* src,
                         hipStream_t s) {
  if (n <= 0) return hipSuccess;
  hipLaunchKernelGGL(k_reciprocal, dim3((n + 255) / 256), dim3(256), 0, s, n,
                     dst, src);
  return hipGetLastError();
}

hipError_t fk_copy(long n, double* dst, const double* src, hipStream_t s) {
  if (n <= 0) return hipSuccess;
  hipLaunchKernelGGL((k_copy<double, double>), dim3((unsigned)((n + 255) / 256)),
                     dim3(256), 0, s, n, dst, src);
  return hipGetLastError();
}

hipError_t fk_eos(FkMesh m, double* P, const double* rho, const double* e,
                  hipStream_t s) {
  const Domain d = dom_all(m);
  hipLaunchKernelGGL(k_eos, sw_grid(d), kSwBlock, 0, s, d, m.nx, P, rho, e);
  return hipGetLastError();
}

hipError_t fk_pressure_accel(FkMesh m, const double* dtp, double* mom_x,
                             double* mom_y, double* u, double* v,
                             const double* P, const double* rho,
                             hipStream_t s) {
  const Domain d = dom_FXY(m);
  hipLaunchKernelGGL(k_pressure_accel, sw_grid(d), kSwBlock, 0, s, m, d, dtp,
                     mom_x, mom_y, u, v, P, rho);
  return hipGetLastError();
}

hipError_t fk_av_stress(FkMesh m, double* Qxx, double* Qyy, const double* u,
                        const double* v, const double* rho, hipStream_t s) {
  const Domain d = dom_I(m);
  hipLaunchKernelGGL(k_av_stress, sw_grid(d), kSwBlock, 0, s, m, d, Qxx, Qyy,
                     u, v, rho);
  return hipGetLastError();
}

hipError_t fk_av_accel(FkMesh m, const double* dtp, double* mom_x,
                       double* mom_y, double* u, double* v, const double* Qxx,
                       const double* Qyy, const double* rho, hipStream_t s) {
  const Domain d = dom_FXY(m);
  hipLaunchKernelGGL(k_av_accel, sw_grid(d), kSwBlock, 0, s, m, d, dtp, mom_x,
                     mom_y, u, v, Qxx, Qyy, rho);
  return hipGetLastError();
}

hipError_t fk_shock_heating(FkMesh m, const double* dtp, double* e,
                            const double* P, const double* u, const double* v,
                            const double* rho, const double* Qxx,
                            const double* Qyy, hipStream_t s) {
  const Domain d = dom_I(m);
  hipLaunchKernelGGL(k_shock_heating, sw_grid(d), kSwBlock, 0, s, m, d, dtp, e,
                     P, u, v, rho, Qxx, Qyy);
  return hipGetLastError();
}

hipError_t fk_x_flux(FkMesh m, const double* dtp, const double* rho,
                     const double* e, const double* u, double* F_x,
                     double* eF_x, hipStream_t s) {
  const Domain d = dom_FX(m);
  hipLaunchKernelGGL(k_x_flux, sw_grid(d), kSwBlock, 0, s, m, d, dtp, rho, e,
                     u, F_x, eF_x);
  return hipGetLastError();
}

hipError_t fk_y_flux(FkMesh m, const double* dtp, const double* rho,
                     const double* e, const double* v, double* F_y,
                     double* eF_y, hipStream_t s) {
  const Domain d = dom_FY(m);
  hipLaunchKernelGGL(k_y_flux, sw_grid(d), kSwBlock, 0, s, m, d, dtp, rho, e,
                     v, F_y, eF_y);
  return hipGetLastError();
}

hipError_t fk_x_update(FkMesh m, const double* dtp, int first, double* rho,
                       const double* rho_old, double* e, const double* F_x,
                       const double* eF_x, hipStream_t s) {
  const Domain d = dom_I(m);
  hipLaunchKernelGGL(k_x_update, sw_grid(d), kSwBlock, 0, s, m, d, dtp, first,
                     rho, rho_old, e, F_x, eF_x);
  return hipGetLastError();
}

hipError_t fk_y_update(FkMesh m, const double* dtp, int first, double* rho,
                       const double* rho_old, double* e, const double* F_y,
                       const double* eF_y, hipStream_t s) {
  const Domain d = dom_I(m);
  hipLaunchKernelGGL(k_y_update, sw_grid(d), kSwBlock, 0, s, m, d, dtp, first,
                     rho, rho_old, e, F_y, eF_y);
  return hipGetLastError();
}

hipError_t fk_mom_flux_xx(FkMesh m, const double* dtp, const double* u,
                          double* uF_x, const double* F_x, hipStream_t s) {
  const Domain d = dom_I(m);
  hipLaunchKernelGGL(k_mom_flux_xx, sw_grid(d), kSwBlock, 0, s, m, d, dtp, u,
                     uF_x, F_x);
  return hipGetLastError();
}

hipError_t fk_mom_flux_xy(FkMesh m, const double* dtp, const double* u,
                          const double* v, double* uF_y, const double* F_y,
                          hipStream_t s) {
  const Domain d = dom_FXY(m);
  hipLaunchKernelGGL(k_mom_flux_xy, sw_grid(d), kSwBlock, 0, s, m, d, dtp, u,
                     v, uF_y, F_y);
  return hipGetLastError();
}

hipError_t fk_mom_flux_yx(FkMesh m, const double* dtp, const double* u,
                          const double* v, double* vF_x, const double* F_x,
                          hipStream_t s) {
  const Domain d = dom_FXY(m);
  hipLaunchKernelGGL(k_mom_flux_yx, sw_grid(d), kSwBlock, 0, s, m, d, dtp, u,
                     v, vF_x, F_x);
  return hipGetLastError();
}

hipError_t fk_mom_flux_yy(FkMesh m, const double* dtp, const double* v,
                          double* vF_y, const double* F_y, hipStream_t s) {
  const Domain d = dom_I(m);
  hipLaunchKernelGGL(k_mom_flux_yy, sw_grid(d), kSwBlock, 0, s, m, d, dtp, v,
                     vF_y, F_y);
  return hipGetLastError();
}

hipError_t fk_mom_x_update(FkMesh m, const double* dtp, int flux_dir,
                           int recompute, double* mom_x, double* u,
                           const double* flux, const double* rho,
                           hipStream_t s) {
  const Domain d = dom_FX(m);
  hipLaunchKernelGGL(k_mom_x_update, sw_grid(d), kSwBlock, 0, s, m, d, dtp,
                     flux_dir, recompute, mom_x, u, flux, rho);
  return hipGetLastError();
}

hipError_t fk_mom_y_update(FkMesh m, const double* dtp, int flux_dir,
                           int recompute, double* mom_y, double* v,
                           const double* flux, const double* rho,
                           hipStream_t s) {
  const Domain d = dom_FY(m);
  hipLaunchKernelGGL(k_mom_y_update, sw_grid(d), kSwBlock, 0, s, m, d, dtp,
                     flux_dir, recompute, mom_y, v, flux, rho);
  return hipGetLastError();
}

}  // extern "C"
