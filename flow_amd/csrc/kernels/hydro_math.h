/*
 * hydro_math.h -- the pointwise arithmetic of flow's hydro timestep, shared by
 * every kernel of the hip set (stagewise and fused, FP64 and FP32 state).
 *
 * Each function evaluates one reference expression with its association
 * order and its divisions kept as written (omp3/flow.c lines cited). The
 * kernels are compiled with -ffp-contract=off and without fast-math, so in
 * FP64 every result rounds exactly as the omp3 loop it follows.
 */
#pragma once

#include <hip/hip_runtime.h>

namespace flowhip {

constexpr double kGam = 1.4;          /* flow_interface.h:7 */
constexpr double kCQ = 3.0;           /* flow_interface.h:8 */
constexpr double kCT = 0.3;           /* flow_data.h:6 */
constexpr double kCM = 1.5 / kCT;     /* flow_interface.h:9 */

__device__ __forceinline__ double vabs(double x) { return __builtin_fabs(x); }
__device__ __forceinline__ float vabs(float x) { return __builtin_fabsf(x); }

/* a / b. FP64: the IEEE quotient, as omp3 computes it. FP32 state: a times
 * the hardware reciprocal of b (within 1 ulp of the reciprocal); the
 * precision form trades the last bits of each quotient for the ten-odd
 * instructions of a correctly rounded FP32 division. */
template <typename T>
__device__ __forceinline__ T quot(T a, T b) {
  return a / b;
}
template <>
__device__ __forceinline__ float quot<float>(float a, float b) {
  return a * __builtin_amdgcn_rcpf(b);
}

/* x / d for a divisor d made of mesh spacings. When every spacing of the
 * mesh is a power of two (P2), d is one too and rd = 1/d is exact, so x * rd
 * is the IEEE quotient bit for bit (division by a power of two only
 * rescales) at the cost of one multiply; otherwise the quotient itself. */
template <bool P2, typename T>
__device__ __forceinline__ T gquot(T x, T d, T rd) {
  if constexpr (P2)
    return x * rd;
  else
    return quot(x, d);
}

/* arch min(a, b) = a < b ? a : b */
template <typename T>
__host__ __device__ __forceinline__ T lesser(T a, T b) {
  return (a < b) ? a : b;
}

/* arch minmod(a, b) = (a*b > 0) ? (|a| < |b| ? a : b) : 0 */
template <typename T>
__device__ __forceinline__ T minmod(T a, T b) {
  return (a * b > T(0.0)) ? ((vabs(a) < vabs(b)) ? a : b) : T(0.0);
}

/* omp3/flow.c:64-65 */
template <typename T>
__device__ __forceinline__ T eos_pressure(T rho, T e) {
  return T(kGam - 1.0) * rho * e;
}

/* Density below which the FP32-state form evaluates the EOS in FP64: the
 * product of a near-vacuum density and its energy loses its low bits (or
 * underflows) in FP32 before it is rounded back. */
constexpr float kNearVacuumRho = 1.0e-6f;

template <>
__device__ __forceinline__ float eos_pressure<float>(float rho, float e) {
  if (rho < kNearVacuumRho)
    return float((kGam - 1.0) * double(rho) * double(e));
  return float(kGam - 1.0) * rho * e;
}

/* Edge-centred density on an x face between cells j-1 (w) and j (c):
 * omp3/flow.c:134-137. */
template <typename T, bool P2 = false>
__device__ __forceinline__ T rho_edge_x(T rho_c, T rho_w, T cdx_c, T cdx_w,
                                        T cdy, T edx, T r_den = T(0)) {
  return gquot<P2>(rho_c * cdx_c * cdy + rho_w * cdx_w * cdy, T(2.0) * edx * cdy,
                   r_den);
}

/* Edge-centred density on a y face between cells i-1 (s) and i (c):
 * omp3/flow.c:138-141. */
template <typename T, bool P2 = false>
__device__ __forceinline__ T rho_edge_y(T rho_c, T rho_s, T cdx, T cdy_c,
                                        T cdy_s, T edy, T r_den = T(0)) {
  return gquot<P2>(rho_c * cdx * cdy_c + rho_s * cdx * cdy_s, T(2.0) * cdx * edy,
                   r_den);
}

/* Velocity from momentum with the zero-density guard, omp3/flow.c:144-150. */
template <typename T>
__device__ __forceinline__ T velocity(T mom, T rho_edge) {
  return (rho_edge == T(0.0)) ? T(0.0) : quot(mom, rho_edge);
}

/* mom -= dt * (a - b) / d: the pressure (omp3/flow.c:126-131) and viscous
 * (:216-219) accelerations. */
template <typename T, bool P2 = false>
__device__ __forceinline__ T accelerate(T mom, T dt, T a, T b, T d, T rd = T(0)) {
  return mom - gquot<P2>(dt * (a - b), d, rd);
}

/* mom -= dt_h * (a - b) / (d1 * d2): the momentum advection updates
 * (omp3/flow.c:581-584, 614-617, 634-637, 664-667). */
template <typename T, bool P2 = false>
__device__ __forceinline__ T advect_update(T mom, T dt_h, T a, T b, T d1,
                                           T d2, T rd12 = T(0)) {
  return mom - gquot<P2>(dt_h * (a - b), d1 * d2, rd12);
}

/* Artificial viscous stress from the four face velocities w0..w3 around a
 * cell (faces c-1, c, c+1, c+2 along the stress direction):
 * omp3/flow.c:177-200. */
template <typename T>
__device__ __forceinline__ T viscous_stress(T rho, T w0, T w1, T w2, T w3) {
  const T d_m = lesser(T(0.0), w1 - w0);
  const T d_0 = lesser(T(0.0), w2 - w1);
  const T d_p = lesser(T(0.0), w3 - w2);
  const T q_ii = T(0.5) * (vabs(d_p - d_0) + vabs(d_0 - d_m));
  return T(-kCQ) * rho * d_0 * q_ii;
}

/* Shock heating and work on one cell, omp3/flow.c:262-283. u0/u1 are the
 * west/east face velocities, v0/v1 the south/north ones. */
template <typename T, bool P2 = false>
__device__ __forceinline__ T shock_heating(T e, T P, T rho, T qxx, T qyy,
                                           T u0, T u1, T v0, T v1, T cdx,
                                           T cdy, T dt_h, T rcdx = T(0),
                                           T rcdy = T(0)) {
  const T dvx = gquot<P2>(u1 - u0, cdx, rcdx);
  const T dvy = gquot<P2>(v1 - v0, cdy, rcdy);
  const T div_dt = (dvx + dvy) * dt_h;
  const T e_q = e - quot(dt_h * (qxx * dvx + qyy * dvy), rho);
  const T rho_c = quot(rho, T(1.0) + div_dt);
  const T e_c = e_q - quot(P * div_dt, rho);
  const T work = quot(T(0.5) * div_dt * (P + T(kGam - 1.0) * e_c * rho_c), rho);
  return (rho == T(0.0)) ? T(0.0) : e_q - work;
}

/* The FP32-state form's CFL candidate, still in FP64: the square of
 * omp3's min(dtx, dty), cdx^2 rho / (c^2 rho + 2 Qxx) with c^2 =
 * gamma (gamma-1) e, so that no square root is taken per cell (the
 * reduction takes one of the minimum). Cells with e < 0, where omp3's
 * sound speed is undefined, give NaN as there. */
__host__ __device__ __forceinline__ double cell_timestep_sq(double e, double qxx,
                                                            double qyy, double rho,
                                                            double cdx, double cdy) {
  const double c2 = kGam * (kGam - 1.0) * e;
  if (!(c2 >= 0.0)) return __builtin_nan("");
  const double dx2 = (cdx * cdx * rho) / (c2 * rho + 2.0 * qxx);
  const double dy2 = (cdy * cdy * rho) / (c2 * rho + 2.0 * qyy);
  return lesser(dx2, dy2);
}

/* Whether a cell's CFL candidate is certain to be defined (not NaN) and
 * not below a running minimum whose square is lim2 (pass the minimum
 * itself in the FP32-state form, whose candidates are squares already),
 * so that leaving it out changes neither the minimum nor the count of
 * undefined candidates. With c^2 = gamma (gamma-1) e >= 0, Q >= 0 and
 * rho, cdx > 0, dt_x >= L iff cdx^2 rho >= L^2 (c^2 rho + 2 Qxx), which needs no
 * quotient or root. The 1e-12 margin covers this test's rounding (a few
 * ulps) and that of the candidate itself (at most 6 ulps in cell_timestep,
 * one quotient in cell_timestep_sq), so a cell that passes has a candidate
 * above the minimum. Negative or non-finite inputs and products outside
 * [1e-280, 1e280] fail the test and are evaluated. */
__host__ __device__ __forceinline__ bool cfl_above(double e, double qxx, double qyy,
                                                   double rho, double cdx, double cdy,
                                                   double lim2) {
  const double c2 = kGam * (kGam - 1.0) * e;
  const double m = lim2 * (1.0 + 1.0e-12);
  const double rx = m * (c2 * rho + 2.0 * qxx), ry = m * (c2 * rho + 2.0 * qyy);
  const double lx = cdx * cdx * rho, ly = cdy * cdy * rho;
  return c2 >= 0.0 && qxx >= 0.0 && qyy >= 0.0 && cdx > 0.0 && cdy > 0.0 &&
         rho >= 1.0e-250 && rho <= 1.0e250 && rx >= 1.0e-280 && ry >= 1.0e-280 &&
         lx <= 1.0e280 && ly <= 1.0e280 && lx >= rx && ly >= ry;
}

/* CFL candidate of one cell, omp3/flow.c:87-94. Always FP64. */
__host__ __device__ __forceinline__ double cell_timestep(double e, double qxx,
                                                         double qyy, double rho,
                                                         double cdx, double cdy) {
  const double c_s = __builtin_sqrt(kGam * (kGam - 1.0) * e);
  const double dtx = cdx / __builtin_sqrt(c_s * c_s + 2.0 * qxx / rho);
  const double dty = cdy / __builtin_sqrt(c_s * c_s + 2.0 * qyy / rho);
  return lesser(dtx, dty);
}

/* Monotonised-central upwind interpolation from four consecutive samples
 * w0..w3, the point lying between w1 and w2. This one pattern is the energy
 * face value of the mass/energy fluxes (omp3/flow.c:378-399, 494-515) and
 * the four momentum fluxes (:793-819, :840-870, :894-921, :942-966). */
template <typename T>
__device__ __forceinline__ T mc_upwind(T w0, T w1, T w2, T w3, T inv, T vel,
                                       bool upwind_low, T cd_lo, T cd_hi,
                                       T dt_h) {
  /* The reference's branch computes one of two mirror-image candidates,
   *   low  = w1 + 0.5 m(w0, w1, w2) (cd_lo - vel dt_h)
   *   high = w2 - 0.5 m(w1, w2, w3) (cd_hi + vel dt_h),
   *   m(p0, p1, p2) = minmod(minmod(0.5 inv (p2 - p0), 2 inv (p1 - p0)),
   *                          2 inv (p2 - p1)).
   * The three samples are selected first and one candidate is evaluated, so
   * the wave stays convergent at half the arithmetic. high is formed as
   * w2 + (0.5 m) * -(cd_hi + vel dt_h): x - y is x + (-y) in IEEE
   * arithmetic and a rounded product only changes sign with an operand, so
   * both candidates round exactly as the reference's expressions. */
  const T p0 = upwind_low ? w0 : w1;
  const T p1 = upwind_low ? w1 : w2;
  const T p2 = upwind_low ? w2 : w3;
  const T a = T(0.5) * inv * (p2 - p0);
  const T b = T(2.0) * inv * (p1 - p0);
  const T c = T(2.0) * inv * (p2 - p1);
  const T span = upwind_low ? (cd_lo - vel * dt_h) : -(cd_hi + vel * dt_h);
  return p1 + T(0.5) * minmod(minmod(a, b), c) * span;
}

template <typename T>
struct FaceFlux {
  T mass;
  T energy;
};

/* Mass and energy flux through one face of a directional sweep,
 * omp3/flow.c:344-403 (x) and :460-519 (y).
 *   vm, v0, vp   face velocities at f-1, f, f+1
 *   rmm..rpp     densities of cells f-2, f-1, f, f+1
 *   emm..epp     specific energies of the same cells
 *   inv          1.0 / face spacing at f (the reference's invdx, :344);
 *                cd_lo, cd_hi cell spacings of f-1, f
 *   et           transverse edge length (edgedy[i] in x, edgedx[j] in y)
 * The y sweep selects the Van Leer smoothness by the face velocity v0
 * instead of the time-centred one (omp3/flow.c:477 vs :361). */
template <typename T, bool kYSweep, bool P2 = false>
__device__ __forceinline__ FaceFlux<T> mass_energy_flux(
    T vm, T v0, T vp, T rmm, T rm, T rp, T rpp, T emm, T em, T ep, T epp,
    T inv, T cd_lo, T cd_hi, T et, T dt, T dt_h, T r_cd_hi = T(0)) {
  const T s_c = T(0.5) * inv * (vp - vm);
  const T s_r = T(2.0) * inv * (v0 - vm);
  const T s_l = T(2.0) * inv * (vp - v0);
  const T v_tc = v0 - T(0.5) * v0 * dt * minmod(s_c, minmod(s_r, s_l));

  /* Van Leer limiter. The smoothness numerator is selected before the one
   * division; where drho == 0 the quotient is discarded and the limiter is
   * 0, as in the reference's if (density_diff). */
  const T drho = (rp - rm);
  const T sel = kYSweep ? v0 : v_tc;
  const T r = quot((sel >= T(0.0)) ? (rm - rmm) : (rpp - rp), drho);
  const T limiter =
      (drho != T(0.0)) ? quot(r + vabs(r), T(1.0) + vabs(r)) : T(0.0);
  const T rho_up = (v_tc >= T(0.0)) ? rm : rp;
  FaceFlux<T> out;
  out.mass = (v_tc * rho_up + T(0.5) * vabs(v_tc) *
                                  (T(1.0) - vabs(gquot<P2>(v_tc * dt_h, cd_hi, r_cd_hi))) *
                                  limiter * drho);
  const T e_face = mc_upwind(emm, em, ep, epp, inv, v_tc, v_tc > T(0.0),
                             cd_lo, cd_hi, dt_h);
  out.energy = et * e_face * out.mass;
  return out;
}

/* Conservative update of one cell by a directional sweep,
 * omp3/flow.c:416-432 (x) and :532-547 (y). f_lo/f_hi are the fluxes
 * through the low/high faces and et_lo/et_hi their edge lengths. */
template <typename T>
struct CellState {
  T rho;
  T e;
};
template <typename T, bool P2 = false>
__device__ __forceinline__ CellState<T> sweep_update(
    T rho, T rho_old, T e, T f_lo, T f_hi, T ef_lo, T ef_hi, T et_lo, T et_hi,
    T cdx, T cdy, T dt_h, bool first, T r_cxy = T(0)) {
  CellState<T> out;
  out.rho = rho - gquot<P2>(dt_h * (et_hi * f_hi - et_lo * f_lo), cdx * cdy, r_cxy);
  const T rho_e = (rho_old * e - gquot<P2>(dt_h * (ef_hi - ef_lo), cdx * cdy, r_cxy));
  out.e = first ? ((rho_old == T(0.0)) ? T(0.0) : quot(rho_e, rho_old))
                : ((out.rho == T(0.0)) ? T(0.0) : quot(rho_e, out.rho));
  return out;
}

/* Momentum flux = advected mass flux x interpolated velocity:
 * F = len * 0.5 * (Fa + Fb); flux = F * w_interp (omp3/flow.c:809-820 etc.).
 * inv is 1.0 / the spacing along the flux direction (invdx, :792). */
template <typename T>
__device__ __forceinline__ T momentum_flux(T len, T fa, T fb, T w0, T w1,
                                           T w2, T w3, T inv, T vel, T cd_lo,
                                           T cd_hi, T dt_h) {
  const T Fc = len * T(0.5) * (fa + fb);
  const T w = mc_upwind(w0, w1, w2, w3, inv, vel, vel > T(0.0), cd_lo, cd_hi,
                        dt_h);
  return Fc * w;
}

}  // namespace flowhip
