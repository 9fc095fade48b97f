/*
 * fused.hip -- the fused schedule of one hydro timestep in four kernels:
 *
 *   lagrangian        EOS + pressure acceleration + artificial viscosity +
 *                     shock heating + the CFL minimum and (dt, dt_h) update
 *                     (omp3/flow.c:22-38)
 *   advect_mass_energy  both directional sweeps of mass and energy, in the
 *                     order the step parity selects, + conservation totals
 *                     (omp3/flow.c:293-554, 973-995)
 *   advect_momentum_x the two x-momentum sub-updates (omp3/flow.c:570-619 /
 *                     673-722)
 *   advect_momentum_y the two y-momentum sub-updates (omp3/flow.c:621-669 /
 *                     724-774)
 *
 * No kernel writes an array it reads, so a tile can read its neighbours'
 * cells in its halo while they are being advanced; intermediate fields go
 * through the interface's scratch arrays (the host schedule picks them).
 * Every value is computed by the same pointwise functions (hydro_math.h) and
 * in the same order as the stagewise kernels, so in FP64 the results are
 * bitwise those of omp3.
 *
 * Index conventions: cell (i, j) at i*nx + j; x face (i, f) at i*(nx+1) + f;
 * y face (g, j) at g*nx + j; corner (g, f) at g*(nx+1) + f. Interior cells
 * [pad, n-pad), interior faces [pad, n-pad].
 */
#include <type_traits>

#include "reduce_common.h"
#include "tile.h"

namespace flowhip {

template <typename T>
struct Geo {
  int nx, ny;
  const T* __restrict__ edx;
  const T* __restrict__ edy;
  const T* __restrict__ cdx;
  const T* __restrict__ cdy;
  const T* __restrict__ iedx; /* 1.0 / edx, computed once per mesh */
  const T* __restrict__ iedy; /* 1.0 / edy */
};

constexpr int kTX = 32;
constexpr int kTY = 16;
constexpr int kNT = 256;
/* Minimum resident waves per SIMD the register allocator must allow
 * (build-time tuning knob; 1 leaves the allocator unconstrained). */
#ifndef FLOWHIP_MIN_WAVES
#define FLOWHIP_MIN_WAVES 1
#endif

/* ---- writers of pad-deep ghost strips, as handle_boundary_2d fills them ---- */

/* x-face field (extent nx+1 along x) at interior face f of interior row i:
 * west/east ghost faces get sx * value, south/north ghost rows get value. */
template <bool EDGE, typename T>
__device__ __forceinline__ void put_xface(T* a, int nx, int ny, int i, int f,
                                          T val, T sx) {
  const int s = nx + 1;
  a[i * s + f] = val;
  if (!EDGE) return;
  if (f == kPad) a[i * s + kPad - 1] = sx * val;
  if (f == kPad + 1) a[i * s + kPad - 2] = sx * val;
  if (f == nx - kPad) a[i * s + nx - kPad + 1] = sx * val;
  if (f == nx - kPad - 1) a[i * s + nx - kPad + 2] = sx * val;
  if (i == kPad) a[(kPad - 1) * s + f] = val;
  if (i == kPad + 1) a[(kPad - 2) * s + f] = val;
  if (i == ny - kPad - 1) a[(ny - kPad) * s + f] = val;
  if (i == ny - kPad - 2) a[(ny - kPad + 1) * s + f] = val;
}

/* y-face field (extent ny+1 along y) at interior face g of interior column
 * j: south/north ghost faces get sy * value, west/east ghost columns get
 * value. */
template <bool EDGE, typename T>
__device__ __forceinline__ void put_yface(T* a, int nx, int ny, int g, int j,
                                          T val, T sy) {
  a[g * nx + j] = val;
  if (!EDGE) return;
  if (g == kPad) a[(kPad - 1) * nx + j] = sy * val;
  if (g == kPad + 1) a[(kPad - 2) * nx + j] = sy * val;
  if (g == ny - kPad) a[(ny - kPad + 1) * nx + j] = sy * val;
  if (g == ny - kPad - 1) a[(ny - kPad + 2) * nx + j] = sy * val;
  if (j == kPad) a[g * nx + kPad - 1] = val;
  if (j == kPad + 1) a[g * nx + kPad - 2] = val;
  if (j == nx - kPad - 1) a[g * nx + nx - kPad] = val;
  if (j == nx - kPad - 2) a[g * nx + nx - kPad + 1] = val;
}

/* cell field at interior cell (i, j), no inversion. */
template <bool EDGE, typename T>
__device__ __forceinline__ void put_cell(T* a, int nx, int ny, int i, int j,
                                         T val) {
  a[i * nx + j] = val;
  if (!EDGE) return;
  if (j == kPad) a[i * nx + kPad - 1] = val;
  if (j == kPad + 1) a[i * nx + kPad - 2] = val;
  if (j == nx - kPad - 1) a[i * nx + nx - kPad] = val;
  if (j == nx - kPad - 2) a[i * nx + nx - kPad + 1] = val;
  if (i == kPad) a[(kPad - 1) * nx + j] = val;
  if (i == kPad + 1) a[(kPad - 2) * nx + j] = val;
  if (i == ny - kPad - 1) a[(ny - kPad) * nx + j] = val;
  if (i == ny - kPad - 2) a[(ny - kPad + 1) * nx + j] = val;
}

/* ======================================================================= */
/* lagrangian                                                              */
/* ======================================================================= */

template <typename T>
struct LagIO {
  const T* __restrict__ rho;
  const T* __restrict__ e;
  const T* __restrict__ mom_x;
  const T* __restrict__ mom_y;
  T* __restrict__ mx1;     /* x momentum after the viscous step (interior) */
  T* __restrict__ my1;     /* y momentum after the viscous step (interior) */
  T* __restrict__ u0;      /* x velocity after the viscous step, + ghosts */
  T* __restrict__ v0;      /* y velocity after the viscous step, + ghosts */
  T* __restrict__ e1;      /* energy after shock heating, + ghosts */
  T* __restrict__ rho_old; /* copy of rho over every cell (omp3/flow.c:302-305) */
  double* __restrict__ cfl; /* NULL, or every interior CFL candidate */
};

template <typename T>
__global__ __launch_bounds__(kNT, FLOWHIP_MIN_WAVES) void k_lagrangian(Geo<T> g, LagIO<T> io,
                                                    const double* dtp,
                                                    FkReduce red, double max_dt,
                                                    int first_step,
                                                    TileList tl) {
  constexpr int TX = kTX, TY = kTY, NT = kNT;
  const int nx = g.nx, ny = g.ny, sx = nx + 1;
  const T dt = T(dtp[0]), dt_h = T(dtp[1]);
  __shared__ T s_u[TY][TX + 5];  /* u_PA on faces [x0-2, x1+2]; then u_AV on [x0, x1] */
  __shared__ T s_mx[TY][TX + 1]; /* pressure-accelerated x momentum, faces [x0, x1] */
  __shared__ T s_rx[TY][TX + 1]; /* x edge density, faces [x0, x1] */
  __shared__ T s_v[TY + 5][TX];  /* v_PA on faces [y0-2, y1+2]; then v_AV on [y0, y1] */
  __shared__ T s_my[TY + 1][TX];
  __shared__ T s_ry[TY + 1][TX];
  __shared__ T s_qx[TY][TX + 2]; /* Qxx on cells [x0-1, x1] */
  __shared__ T s_qy[TY + 2][TX]; /* Qyy on cells [y0-1, y1] */

  auto pres = [&](int k) { return eos_pressure(io.rho[k], io.e[k]); };
  double local_min = max_dt;
  int nan_count = 0;
  TileWalk walk(tl);
  int tile;
  while (walk.next(tile)) {
    const int x0 = kPad + (tile % tl.ntx) * TX, y0 = kPad + (tile / tl.ntx) * TY;
    const int x1 = min(x0 + TX, nx - kPad), y1 = min(y0 + TY, ny - kPad);
    auto body = [&](auto edge_tag) {
      constexpr bool EDGE = decltype(edge_tag)::value;

      /* pressure acceleration, omp3/flow.c:121-152 (ghosts: INVERT_X / _Y) */
      region<NT, TY, TX + 5>([&](int rr, int cc) {
        const int i = y0 + rr, f = x0 - 2 + cc;
        if (i >= y1 || f > x1 + 2) return;
        const bool gh = ghost<EDGE>(f, sx);
        const int fm = gh ? mir<EDGE>(f, sx) : f;
        const int c = i * nx + fm;
        const T mom = accelerate(io.mom_x[i * sx + fm], dt, pres(c), pres(c - 1), g.edx[fm]);
        const T rex = rho_edge_x(io.rho[c], io.rho[c - 1], g.cdx[fm], g.cdx[fm - 1],
                                 g.cdy[i], g.edx[fm]);
        const T u = velocity(mom, rex);
        s_u[rr][cc] = gh ? -u : u;
        if (!gh && f >= x0 && f <= x1) {
          s_mx[rr][f - x0] = mom;
          s_rx[rr][f - x0] = rex;
        }
      });
      region<NT, TY + 5, TX>([&](int rr, int cc) {
        const int gf = y0 - 2 + rr, j = x0 + cc;
        if (j >= x1 || gf > y1 + 2) return;
        const bool gh = ghost<EDGE>(gf, ny + 1);
        const int gm = gh ? mir<EDGE>(gf, ny + 1) : gf;
        const int c = gm * nx + j;
        const T mom = accelerate(io.mom_y[c], dt, pres(c), pres(c - nx), g.edy[gm]);
        const T rey = rho_edge_y(io.rho[c], io.rho[c - nx], g.cdx[j], g.cdy[gm],
                                 g.cdy[gm - 1], g.edy[gm]);
        const T v = velocity(mom, rey);
        s_v[rr][cc] = gh ? -v : v;
        if (!gh && gf >= y0 && gf <= y1) {
          s_my[gf - y0][cc] = mom;
          s_ry[gf - y0][cc] = rey;
        }
      });
      __syncthreads();

      /* viscous stresses, omp3/flow.c:173-202 (ghosts: NO_INVERT) */
      region<NT, TY, TX + 2>([&](int rr, int cc) {
        const int i = y0 + rr, c = x0 - 1 + cc;
        if (i >= y1 || c > x1) return;
        const int cm = mir<EDGE>(c, nx);
        const int o = cm - (x0 - 2);
        s_qx[rr][cc] = viscous_stress(io.rho[i * nx + cm], s_u[rr][o - 1], s_u[rr][o],
                                      s_u[rr][o + 1], s_u[rr][o + 2]);
      });
      region<NT, TY + 2, TX>([&](int rr, int cc) {
        const int r = y0 - 1 + rr, j = x0 + cc;
        if (j >= x1 || r > y1) return;
        const int rm = mir<EDGE>(r, ny);
        const int o = rm - (y0 - 2);
        s_qy[rr][cc] = viscous_stress(io.rho[rm * nx + j], s_v[o - 1][cc], s_v[o][cc],
                                      s_v[o + 1][cc], s_v[o + 2][cc]);
      });
      __syncthreads();

      /* viscous acceleration, omp3/flow.c:212-240 */
      region<NT, TY, TX + 1>([&](int rr, int cc) {
        const int i = y0 + rr, f = x0 + cc;
        if (i >= y1 || f > x1) return;
        const T mom = accelerate(s_mx[rr][cc], dt, s_qx[rr][cc + 1], s_qx[rr][cc], g.cdx[f]);
        const T u = velocity(mom, s_rx[rr][cc]);
        s_u[rr][cc] = u;
        if (f < x1 || x1 == nx - kPad) {
          io.mx1[i * sx + f] = mom;
          put_xface<EDGE>(io.u0, nx, ny, i, f, u, T(-1.0));
        }
      });
      region<NT, TY + 1, TX>([&](int rr, int cc) {
        const int gf = y0 + rr, j = x0 + cc;
        if (j >= x1 || gf > y1) return;
        const T mom = accelerate(s_my[rr][cc], dt, s_qy[rr + 1][cc], s_qy[rr][cc], g.cdy[gf]);
        const T v = velocity(mom, s_ry[rr][cc]);
        s_v[rr][cc] = v;
        if (gf < y1 || y1 == ny - kPad) {
          io.my1[gf * nx + j] = mom;
          put_yface<EDGE>(io.v0, nx, ny, gf, j, v, T(-1.0));
        }
      });
      __syncthreads();

      /* shock heating (omp3/flow.c:258-285), CFL candidate (:83-96), and the
       * copy of rho (:302-305) */
      region<NT, TY, TX>([&](int rr, int cc) {
        const int i = y0 + rr, j = x0 + cc;
        if (i >= y1 || j >= x1) return;
        const int c = i * nx + j;
        const T rho = io.rho[c], e = io.e[c];
        const T qxx = s_qx[rr][cc + 1], qyy = s_qy[rr + 1][cc];
        const T es = shock_heating(e, eos_pressure(rho, e), rho, qxx, qyy, s_u[rr][cc],
                                   s_u[rr][cc + 1], s_v[rr][cc], s_v[rr + 1][cc],
                                   g.cdx[j], g.cdy[i], dt_h);
        put_cell<EDGE>(io.e1, nx, ny, i, j, es);
        io.rho_old[c] = rho;
        /* ghost and corner cells of the copy */
        const bool w = (j == kPad), ea = (j == nx - kPad - 1);
        const bool s = (i == kPad), no = (i == ny - kPad - 1);
        for (int d = 1; d <= kPad; ++d) {
          if (w) io.rho_old[c - d] = io.rho[c - d];
          if (ea) io.rho_old[c + d] = io.rho[c + d];
          if (s) io.rho_old[c - d * nx] = io.rho[c - d * nx];
          if (no) io.rho_old[c + d * nx] = io.rho[c + d * nx];
          for (int d2 = 1; d2 <= kPad; ++d2) {
            if (s && w) io.rho_old[c - d * nx - d2] = io.rho[c - d * nx - d2];
            if (s && ea) io.rho_old[c - d * nx + d2] = io.rho[c - d * nx + d2];
            if (no && w) io.rho_old[c + d * nx - d2] = io.rho[c + d * nx - d2];
            if (no && ea) io.rho_old[c + d * nx + d2] = io.rho[c + d * nx + d2];
          }
        }
        const double cand = cell_timestep(double(es), double(qxx), double(qyy),
                                          double(rho), double(g.cdx[j]), double(g.cdy[i]));
        if (cand == cand)
          local_min = lesser(local_min, cand);
        else
          ++nan_count;
        if (io.cfl) io.cfl[c] = cand;
      });
      __syncthreads();
    };
    if (x0 - 3 < kPad || x1 + 3 > nx - kPad || y0 - 3 < kPad || y1 + 3 > ny - kPad)
      body(std::true_type{});
    else
      body(std::false_type{});
  }
  finish_timestep<NT>(local_min, nan_count, red, max_dt, first_step);
}

/* ======================================================================= */
/* advect_mass_energy                                                      */
/* ======================================================================= */

template <typename T>
struct MassIO {
  const T* __restrict__ u0;   /* face velocities (after the viscous step) */
  const T* __restrict__ v0;
  const T* __restrict__ rho0; /* density before advection */
  const T* __restrict__ e1;   /* energy after shock heating */
  T* __restrict__ rho;        /* advected density, + ghosts */
  T* __restrict__ e;          /* advected energy, + ghosts */
  T* __restrict__ F_x;        /* mass fluxes, + ghosts */
  T* __restrict__ F_y;
};

/* x-sweep first (even tt, omp3/flow.c:308-314). */
template <typename T>
__global__ __launch_bounds__(kNT, FLOWHIP_MIN_WAVES) void k_mass_energy_xy(Geo<T> g, MassIO<T> io,
                                                        const double* dtp,
                                                        FkReduce red,
                                                        TileList tl) {
  constexpr int TX = kTX, TY = kTY, NT = kNT;
  const int nx = g.nx, ny = g.ny, sx = nx + 1;
  const T dt = T(dtp[0]), dt_h = T(dtp[1]);
  __shared__ T s_fx[TY + 4][TX + 1], s_efx[TY + 4][TX + 1]; /* rows [y0-2, y1+1] */
  __shared__ T s_r1[TY + 4][TX], s_e1[TY + 4][TX];
  __shared__ T s_fy[TY + 1][TX], s_efy[TY + 1][TX];          /* faces [y0, y1] */
  CSum ms{0.0, 0.0}, es{0.0, 0.0};
  TileWalk walk(tl);
  int tile;
  while (walk.next(tile)) {
    const int x0 = kPad + (tile % tl.ntx) * TX, y0 = kPad + (tile / tl.ntx) * TY;
    const int x1 = min(x0 + TX, nx - kPad), y1 = min(y0 + TY, ny - kPad);
    auto body = [&](auto edge_tag) {
      constexpr bool EDGE = decltype(edge_tag)::value;

      /* x fluxes (first sweep) on rows [y0-2, y1+1]; a ghost row is its
       * mirror row (omp3/flow.c:338-404, boundary :408) */
      region<NT, TY + 4, TX + 1>([&](int rr, int cc) {
        const int r = y0 - 2 + rr, f = x0 + cc;
        if (r > y1 + 1 || f > x1) return;
        const int i = mir<EDGE>(r, ny);
        const int c = i * nx + f, fi = i * sx + f;
        const FaceFlux<T> fl = mass_energy_flux<T, false>(
            io.u0[fi - 1], io.u0[fi], io.u0[fi + 1], io.rho0[c - 2], io.rho0[c - 1],
            io.rho0[c], io.rho0[c + 1], io.e1[c - 2], io.e1[c - 1], io.e1[c],
            io.e1[c + 1], g.iedx[f], g.cdx[f - 1], g.cdx[f], g.edy[i], dt, dt_h);
        s_fx[rr][cc] = fl.mass;
        s_efx[rr][cc] = fl.energy;
        if (r == i && r >= y0 && r < y1 && (f < x1 || x1 == nx - kPad))
          put_xface<EDGE>(io.F_x, nx, ny, i, f, fl.mass, T(-1.0));
      });
      __syncthreads();
      /* x update, first = 1 (omp3/flow.c:412-434) */
      region<NT, TY + 4, TX>([&](int rr, int cc) {
        const int r = y0 - 2 + rr, j = x0 + cc;
        if (r > y1 + 1 || j >= x1) return;
        const int i = mir<EDGE>(r, ny);
        const int c = i * nx + j;
        const CellState<T> st = sweep_update(
            io.rho0[c], io.rho0[c], io.e1[c], s_fx[rr][cc], s_fx[rr][cc + 1],
            s_efx[rr][cc], s_efx[rr][cc + 1], g.edy[i], g.edy[i + 1], g.cdx[j],
            g.cdy[i], dt_h, true);
        s_r1[rr][cc] = st.rho;
        s_e1[rr][cc] = st.e;
      });
      __syncthreads();
      /* y fluxes (second sweep) on faces [y0, y1] (omp3/flow.c:454-521) */
      region<NT, TY + 1, TX>([&](int rr, int cc) {
        const int gf = y0 + rr, j = x0 + cc;
        if (gf > y1 || j >= x1) return;
        const int c = gf * nx + j;
        const FaceFlux<T> fl = mass_energy_flux<T, true>(
            io.v0[c - nx], io.v0[c], io.v0[c + nx], s_r1[rr][cc], s_r1[rr + 1][cc],
            s_r1[rr + 2][cc], s_r1[rr + 3][cc], s_e1[rr][cc], s_e1[rr + 1][cc],
            s_e1[rr + 2][cc], s_e1[rr + 3][cc], g.iedy[gf], g.cdy[gf - 1], g.cdy[gf],
            g.edx[j], dt, dt_h);
        s_fy[rr][cc] = fl.mass;
        s_efy[rr][cc] = fl.energy;
        if (gf < y1 || y1 == ny - kPad) put_yface<EDGE>(io.F_y, nx, ny, gf, j, fl.mass, T(-1.0));
      });
      __syncthreads();
      /* y update, first = 0 (omp3/flow.c:528-549), and the totals */
      region<NT, TY, TX>([&](int rr, int cc) {
        const int i = y0 + rr, j = x0 + cc;
        if (i >= y1 || j >= x1) return;
        const int c = i * nx + j;
        const CellState<T> st = sweep_update(
            s_r1[rr + 2][cc], io.rho0[c], s_e1[rr + 2][cc], s_fy[rr][cc],
            s_fy[rr + 1][cc], s_efy[rr][cc], s_efy[rr + 1][cc], g.edx[j],
            g.edx[j + 1], g.cdx[j], g.cdy[i], dt_h, false);
        put_cell<EDGE>(io.rho, nx, ny, i, j, st.rho);
        put_cell<EDGE>(io.e, nx, ny, i, j, st.e);
        ms.add(double(st.rho));
        es.add(double(st.rho) * double(st.e));
      });
      __syncthreads();
    };
    if (x0 - 3 < kPad || x1 + 3 > nx - kPad || y0 - 3 < kPad || y1 + 3 > ny - kPad)
      body(std::true_type{});
    else
      body(std::false_type{});
  }
  finish_sums<NT>(ms, es, red);
}

/* y-sweep first (odd tt, omp3/flow.c:315-322). */
template <typename T>
__global__ __launch_bounds__(kNT, FLOWHIP_MIN_WAVES) void k_mass_energy_yx(Geo<T> g, MassIO<T> io,
                                                        const double* dtp,
                                                        FkReduce red,
                                                        TileList tl) {
  constexpr int TX = kTX, TY = kTY, NT = kNT;
  const int nx = g.nx, ny = g.ny, sx = nx + 1;
  const T dt = T(dtp[0]), dt_h = T(dtp[1]);
  __shared__ T s_fy[TY + 1][TX + 4], s_efy[TY + 1][TX + 4]; /* cols [x0-2, x1+1] */
  __shared__ T s_r1[TY][TX + 4], s_e1[TY][TX + 4];
  __shared__ T s_fx[TY][TX + 1], s_efx[TY][TX + 1];
  CSum ms{0.0, 0.0}, es{0.0, 0.0};
  TileWalk walk(tl);
  int tile;
  while (walk.next(tile)) {
    const int x0 = kPad + (tile % tl.ntx) * TX, y0 = kPad + (tile / tl.ntx) * TY;
    const int x1 = min(x0 + TX, nx - kPad), y1 = min(y0 + TY, ny - kPad);
    auto body = [&](auto edge_tag) {
      constexpr bool EDGE = decltype(edge_tag)::value;

      /* y fluxes (first sweep) on faces [y0, y1] of columns [x0-2, x1+1]; a
       * ghost column is its mirror column */
      region<NT, TY + 1, TX + 4>([&](int rr, int cc) {
        const int gf = y0 + rr, q = x0 - 2 + cc;
        if (gf > y1 || q > x1 + 1) return;
        const int j = mir<EDGE>(q, nx);
        const int c = gf * nx + j;
        const FaceFlux<T> fl = mass_energy_flux<T, true>(
            io.v0[c - nx], io.v0[c], io.v0[c + nx], io.rho0[c - 2 * nx],
            io.rho0[c - nx], io.rho0[c], io.rho0[c + nx], io.e1[c - 2 * nx],
            io.e1[c - nx], io.e1[c], io.e1[c + nx], g.iedy[gf], g.cdy[gf - 1],
            g.cdy[gf], g.edx[j], dt, dt_h);
        s_fy[rr][cc] = fl.mass;
        s_efy[rr][cc] = fl.energy;
        if (q == j && j >= x0 && j < x1 && (gf < y1 || y1 == ny - kPad))
          put_yface<EDGE>(io.F_y, nx, ny, gf, j, fl.mass, T(-1.0));
      });
      __syncthreads();
      /* y update, first = 1 */
      region<NT, TY, TX + 4>([&](int rr, int cc) {
        const int i = y0 + rr, q = x0 - 2 + cc;
        if (i >= y1 || q > x1 + 1) return;
        const int j = mir<EDGE>(q, nx);
        const int c = i * nx + j;
        const CellState<T> st = sweep_update(
            io.rho0[c], io.rho0[c], io.e1[c], s_fy[rr][cc], s_fy[rr + 1][cc],
            s_efy[rr][cc], s_efy[rr + 1][cc], g.edx[j], g.edx[j + 1], g.cdx[j],
            g.cdy[i], dt_h, true);
        s_r1[rr][cc] = st.rho;
        s_e1[rr][cc] = st.e;
      });
      __syncthreads();
      /* x fluxes (second sweep) on faces [x0, x1] */
      region<NT, TY, TX + 1>([&](int rr, int cc) {
        const int i = y0 + rr, f = x0 + cc;
        if (i >= y1 || f > x1) return;
        const int fi = i * sx + f;
        const FaceFlux<T> fl = mass_energy_flux<T, false>(
            io.u0[fi - 1], io.u0[fi], io.u0[fi + 1], s_r1[rr][cc], s_r1[rr][cc + 1],
            s_r1[rr][cc + 2], s_r1[rr][cc + 3], s_e1[rr][cc], s_e1[rr][cc + 1],
            s_e1[rr][cc + 2], s_e1[rr][cc + 3], g.iedx[f], g.cdx[f - 1], g.cdx[f],
            g.edy[i], dt, dt_h);
        s_fx[rr][cc] = fl.mass;
        s_efx[rr][cc] = fl.energy;
        if (f < x1 || x1 == nx - kPad) put_xface<EDGE>(io.F_x, nx, ny, i, f, fl.mass, T(-1.0));
      });
      __syncthreads();
      /* x update, first = 0, and the totals */
      region<NT, TY, TX>([&](int rr, int cc) {
        const int i = y0 + rr, j = x0 + cc;
        if (i >= y1 || j >= x1) return;
        const int c = i * nx + j;
        const CellState<T> st = sweep_update(
            s_r1[rr][cc + 2], io.rho0[c], s_e1[rr][cc + 2], s_fx[rr][cc],
            s_fx[rr][cc + 1], s_efx[rr][cc], s_efx[rr][cc + 1], g.edy[i],
            g.edy[i + 1], g.cdx[j], g.cdy[i], dt_h, false);
        put_cell<EDGE>(io.rho, nx, ny, i, j, st.rho);
        put_cell<EDGE>(io.e, nx, ny, i, j, st.e);
        ms.add(double(st.rho));
        es.add(double(st.rho) * double(st.e));
      });
      __syncthreads();
    };
    if (x0 - 3 < kPad || x1 + 3 > nx - kPad || y0 - 3 < kPad || y1 + 3 > ny - kPad)
      body(std::true_type{});
    else
      body(std::false_type{});
  }
  finish_sums<NT>(ms, es, red);
}

/* ======================================================================= */
/* advect_momentum_x                                                       */
/* ======================================================================= */

template <typename T>
struct MomIO {
  const T* __restrict__ u;    /* x velocity read (u0 for x, u1 for y) */
  const T* __restrict__ v0;   /* y velocity after the viscous step */
  const T* __restrict__ F_x;
  const T* __restrict__ F_y;
  const T* __restrict__ rho;  /* advected density */
  const T* __restrict__ mom1; /* momentum after the viscous step */
  T* __restrict__ mom;        /* advected momentum */
  T* __restrict__ vel;        /* recomputed velocity, + ghosts */
};

/* odd tt: x flux in x, recompute u, then x flux in y
 * (omp3/flow.c:570-619). */
template <typename T>
__global__ __launch_bounds__(kNT, FLOWHIP_MIN_WAVES) void k_momentum_x_xy(Geo<T> g, MomIO<T> io,
                                                       const double* dtp,
                                                       TileList tl) {
  constexpr int TX = kTX, TY = kTY, NT = kNT;
  const int nx = g.nx, ny = g.ny, sx = nx + 1;
  const T dt_h = T(dtp[1]);
  __shared__ T s_ufx[TY + 4][TX + 2]; /* rows [y0-2, y1+1], cells [x0-1, x1] */
  __shared__ T s_u1[TY + 4][TX + 1];  /* rows [y0-2, y1+1], faces [x0, x1] */
  __shared__ T s_m1[TY][TX + 1];
  __shared__ T s_ufy[TY + 1][TX + 1]; /* corner rows [y0, y1], faces [x0, x1] */
  TileWalk walk(tl);
  int tile;
  while (walk.next(tile)) {
    const int x0 = kPad + (tile % tl.ntx) * TX, y0 = kPad + (tile / tl.ntx) * TY;
    const int x1 = min(x0 + TX, nx - kPad), y1 = min(y0 + TY, ny - kPad);
    auto body = [&](auto edge_tag) {
      constexpr bool EDGE = decltype(edge_tag)::value;

      /* momentum_x_flux_in_x, omp3/flow.c:790-820 (ghosts NO_INVERT) */
      region<NT, TY + 4, TX + 2>([&](int rr, int cc) {
        const int r = y0 - 2 + rr, q = x0 - 1 + cc;
        if (r > y1 + 1 || q > x1) return;
        const int i = mir<EDGE>(r, ny), c = mir<EDGE>(q, nx);
        const int fi = i * sx + c;
        const T u_cell = T(0.5) * (io.u[fi] + io.u[fi + 1]);
        s_ufx[rr][cc] = momentum_flux(g.edy[i], io.F_x[fi], io.F_x[fi + 1], io.u[fi - 1],
                                      io.u[fi], io.u[fi + 1], io.u[fi + 2], g.iedx[c],
                                      u_cell, g.cdx[c - 1], g.cdx[c], dt_h);
      });
      __syncthreads();
      /* first sub-update and velocity recompute, omp3/flow.c:578-594 */
      region<NT, TY + 4, TX + 1>([&](int rr, int cc) {
        const int r = y0 - 2 + rr, f = x0 + cc;
        if (r > y1 + 1 || f > x1) return;
        const int i = mir<EDGE>(r, ny);
        const T m1 = advect_update(io.mom1[i * sx + f], dt_h, s_ufx[rr][cc + 1],
                                   s_ufx[rr][cc], g.edx[f], g.cdy[i]);
        const int c = i * nx + f;
        const T rex = rho_edge_x(io.rho[c], io.rho[c - 1], g.cdx[f], g.cdx[f - 1],
                                 g.cdy[i], g.edx[f]);
        s_u1[rr][cc] = velocity(m1, rex);
        if (r >= y0 && r < y1) s_m1[r - y0][cc] = m1;
      });
      __syncthreads();
      /* momentum_x_flux_in_y on corners, omp3/flow.c:837-871 */
      region<NT, TY + 1, TX + 1>([&](int rr, int cc) {
        const int gf = y0 + rr, f = x0 + cc;
        if (gf > y1 || f > x1) return;
        const int c = gf * nx + f;
        const T v_cell = T(0.5) * (io.v0[c - 1] + io.v0[c]);
        s_ufy[rr][cc] = momentum_flux(g.edx[f], io.F_y[c], io.F_y[c - 1], s_u1[rr][cc],
                                      s_u1[rr + 1][cc], s_u1[rr + 2][cc], s_u1[rr + 3][cc],
                                      g.iedy[gf], v_cell, g.cdy[gf - 1], g.cdy[gf], dt_h);
      });
      __syncthreads();
      /* second sub-update, omp3/flow.c:611-618 */
      region<NT, TY, TX + 1>([&](int rr, int cc) {
        const int i = y0 + rr, f = x0 + cc;
        if (i >= y1 || f > x1 || !(f < x1 || x1 == nx - kPad)) return;
        io.mom[i * sx + f] = advect_update(s_m1[rr][cc], dt_h, s_ufy[rr + 1][cc],
                                           s_ufy[rr][cc], g.cdx[f], g.edy[i]);
        put_xface<EDGE>(io.vel, nx, ny, i, f, s_u1[rr + 2][cc], T(-1.0));
      });
      __syncthreads();
    };
    if (x0 - 3 < kPad || x1 + 3 > nx - kPad || y0 - 3 < kPad || y1 + 3 > ny - kPad)
      body(std::true_type{});
    else
      body(std::false_type{});
  }
}

/* even tt: x flux in y, recompute u, then x flux in x
 * (omp3/flow.c:673-722). */
template <typename T>
__global__ __launch_bounds__(kNT, FLOWHIP_MIN_WAVES) void k_momentum_x_yx(Geo<T> g, MomIO<T> io,
                                                       const double* dtp,
                                                       TileList tl) {
  constexpr int TX = kTX, TY = kTY, NT = kNT;
  const int nx = g.nx, ny = g.ny, sx = nx + 1;
  const T dt_h = T(dtp[1]);
  __shared__ T s_ufy[TY + 1][TX + 5]; /* corner rows [y0, y1], faces [x0-2, x1+2] */
  __shared__ T s_u1[TY][TX + 5];      /* rows [y0, y1), faces [x0-2, x1+2] */
  __shared__ T s_m1[TY][TX + 1];
  __shared__ T s_ufx[TY][TX + 2];     /* cells [x0-1, x1] */
  TileWalk walk(tl);
  int tile;
  while (walk.next(tile)) {
    const int x0 = kPad + (tile % tl.ntx) * TX, y0 = kPad + (tile / tl.ntx) * TY;
    const int x1 = min(x0 + TX, nx - kPad), y1 = min(y0 + TY, ny - kPad);
    auto body = [&](auto edge_tag) {
      constexpr bool EDGE = decltype(edge_tag)::value;

      /* momentum_x_flux_in_y on corners of interior faces */
      region<NT, TY + 1, TX + 5>([&](int rr, int cc) {
        const int gf = y0 + rr, f = x0 - 2 + cc;
        if (gf > y1 || f > x1 + 2 || ghost<EDGE>(f, sx)) return;
        const int c = gf * nx + f, fi = gf * sx + f;
        const T v_cell = T(0.5) * (io.v0[c - 1] + io.v0[c]);
        s_ufy[rr][cc] = momentum_flux(g.edx[f], io.F_y[c], io.F_y[c - 1], io.u[fi - 2 * sx],
                                      io.u[fi - sx], io.u[fi], io.u[fi + sx], g.iedy[gf],
                                      v_cell, g.cdy[gf - 1], g.cdy[gf], dt_h);
      });
      __syncthreads();
      /* first sub-update and velocity recompute, omp3/flow.c:684-700; ghost
       * faces are the negated mirror (INVERT_X) */
      region<NT, TY, TX + 5>([&](int rr, int cc) {
        const int i = y0 + rr, f = x0 - 2 + cc;
        if (i >= y1 || f > x1 + 2) return;
        const bool gh = ghost<EDGE>(f, sx);
        const int fm = gh ? mir<EDGE>(f, sx) : f;
        const int o = fm - (x0 - 2);
        const T m1 = advect_update(io.mom1[i * sx + fm], dt_h, s_ufy[rr + 1][o],
                                   s_ufy[rr][o], g.cdx[fm], g.edy[i]);
        const int c = i * nx + fm;
        const T rex = rho_edge_x(io.rho[c], io.rho[c - 1], g.cdx[fm], g.cdx[fm - 1],
                                 g.cdy[i], g.edx[fm]);
        const T u1 = velocity(m1, rex);
        s_u1[rr][cc] = gh ? -u1 : u1;
        if (!gh && f >= x0 && f <= x1) s_m1[rr][f - x0] = m1;
      });
      __syncthreads();
      /* momentum_x_flux_in_x on cells [x0-1, x1] (ghosts NO_INVERT) */
      region<NT, TY, TX + 2>([&](int rr, int cc) {
        const int i = y0 + rr, q = x0 - 1 + cc;
        if (i >= y1 || q > x1) return;
        const int c = mir<EDGE>(q, nx);
        const int o = c - (x0 - 2);
        const int fi = i * sx + c;
        const T u_cell = T(0.5) * (s_u1[rr][o] + s_u1[rr][o + 1]);
        s_ufx[rr][cc] = momentum_flux(g.edy[i], io.F_x[fi], io.F_x[fi + 1], s_u1[rr][o - 1],
                                      s_u1[rr][o], s_u1[rr][o + 1], s_u1[rr][o + 2],
                                      g.iedx[c], u_cell, g.cdx[c - 1], g.cdx[c], dt_h);
      });
      __syncthreads();
      /* second sub-update, omp3/flow.c:714-721 */
      region<NT, TY, TX + 1>([&](int rr, int cc) {
        const int i = y0 + rr, f = x0 + cc;
        if (i >= y1 || f > x1 || !(f < x1 || x1 == nx - kPad)) return;
        io.mom[i * sx + f] = advect_update(s_m1[rr][cc], dt_h, s_ufx[rr][cc + 1],
                                           s_ufx[rr][cc], g.edx[f], g.cdy[i]);
        put_xface<EDGE>(io.vel, nx, ny, i, f, s_u1[rr][cc + 2], T(-1.0));
      });
      __syncthreads();
    };
    if (x0 - 3 < kPad || x1 + 3 > nx - kPad || y0 - 3 < kPad || y1 + 3 > ny - kPad)
      body(std::true_type{});
    else
      body(std::false_type{});
  }
}

/* ======================================================================= */
/* advect_momentum_y                                                       */
/* ======================================================================= */

/* odd tt: y flux in x, recompute v, then y flux in y
 * (omp3/flow.c:621-669). io.u holds u after the x sub-updates. */
template <typename T>
__global__ __launch_bounds__(kNT, FLOWHIP_MIN_WAVES) void k_momentum_y_xy(Geo<T> g, MomIO<T> io,
                                                       const double* dtp,
                                                       TileList tl) {
  constexpr int TX = kTX, TY = kTY, NT = kNT;
  const int nx = g.nx, ny = g.ny, sx = nx + 1;
  const T dt_h = T(dtp[1]);
  __shared__ T s_vfx[TY + 5][TX + 1]; /* corner rows [y0-2, y1+2], faces [x0, x1] */
  __shared__ T s_v1[TY + 5][TX];      /* y faces [y0-2, y1+2] */
  __shared__ T s_m1[TY + 1][TX];      /* y faces [y0, y1] */
  __shared__ T s_vfy[TY + 2][TX];     /* cell rows [y0-1, y1] */
  TileWalk walk(tl);
  int tile;
  while (walk.next(tile)) {
    const int x0 = kPad + (tile % tl.ntx) * TX, y0 = kPad + (tile / tl.ntx) * TY;
    const int x1 = min(x0 + TX, nx - kPad), y1 = min(y0 + TY, ny - kPad);
    auto body = [&](auto edge_tag) {
      constexpr bool EDGE = decltype(edge_tag)::value;

      /* momentum_y_flux_in_x on corners of interior y faces, omp3/flow.c:890-922 */
      region<NT, TY + 5, TX + 1>([&](int rr, int cc) {
        const int gf = y0 - 2 + rr, f = x0 + cc;
        if (gf > y1 + 2 || f > x1 || ghost<EDGE>(gf, ny + 1)) return;
        const int c = gf * nx + f, fi = gf * sx + f;
        const T u_cell = T(0.5) * (io.u[fi] + io.u[fi - sx]);
        s_vfx[rr][cc] = momentum_flux(g.cdy[gf], io.F_x[fi], io.F_x[fi - sx], io.v0[c - 2],
                                      io.v0[c - 1], io.v0[c], io.v0[c + 1], g.iedx[f],
                                      u_cell, g.cdx[f - 1], g.cdx[f], dt_h);
      });
      __syncthreads();
      /* first sub-update and velocity recompute, omp3/flow.c:631-647; ghost y
       * faces are the negated mirror (INVERT_Y) */
      region<NT, TY + 5, TX>([&](int rr, int cc) {
        const int gf = y0 - 2 + rr, j = x0 + cc;
        if (gf > y1 + 2 || j >= x1) return;
        const bool gh = ghost<EDGE>(gf, ny + 1);
        const int gm = gh ? mir<EDGE>(gf, ny + 1) : gf;
        const int o = gm - (y0 - 2);
        const T m1 = advect_update(io.mom1[gm * nx + j], dt_h, s_vfx[o][cc + 1],
                                   s_vfx[o][cc], g.edx[j], g.cdy[gm]);
        const int c = gm * nx + j;
        const T rey = rho_edge_y(io.rho[c], io.rho[c - nx], g.cdx[j], g.cdy[gm],
                                 g.cdy[gm - 1], g.edy[gm]);
        const T v1 = velocity(m1, rey);
        s_v1[rr][cc] = gh ? -v1 : v1;
        if (!gh && gf >= y0 && gf <= y1) s_m1[gf - y0][cc] = m1;
      });
      __syncthreads();
      /* momentum_y_flux_in_y on cell rows [y0-1, y1] (ghosts NO_INVERT),
       * omp3/flow.c:939-967 */
      region<NT, TY + 2, TX>([&](int rr, int cc) {
        const int r = y0 - 1 + rr, j = x0 + cc;
        if (r > y1 || j >= x1) return;
        const int i = mir<EDGE>(r, ny);
        const int o = i - (y0 - 2);
        const int c = i * nx + j;
        const T v_cell = T(0.5) * (s_v1[o][cc] + s_v1[o + 1][cc]);
        s_vfy[rr][cc] = momentum_flux(g.cdx[j], io.F_y[c], io.F_y[c + nx], s_v1[o - 1][cc],
                                      s_v1[o][cc], s_v1[o + 1][cc], s_v1[o + 2][cc],
                                      g.iedy[i], v_cell, g.cdy[i - 1], g.cdy[i], dt_h);
      });
      __syncthreads();
      /* second sub-update, omp3/flow.c:660-669 */
      region<NT, TY + 1, TX>([&](int rr, int cc) {
        const int gf = y0 + rr, j = x0 + cc;
        if (gf > y1 || j >= x1 || !(gf < y1 || y1 == ny - kPad)) return;
        io.mom[gf * nx + j] = advect_update(s_m1[rr][cc], dt_h, s_vfy[rr + 1][cc],
                                            s_vfy[rr][cc], g.cdx[j], g.edy[gf]);
        put_yface<EDGE>(io.vel, nx, ny, gf, j, s_v1[rr + 2][cc], T(-1.0));
      });
      __syncthreads();
    };
    if (x0 - 3 < kPad || x1 + 3 > nx - kPad || y0 - 3 < kPad || y1 + 3 > ny - kPad)
      body(std::true_type{});
    else
      body(std::false_type{});
  }
}

/* even tt: y flux in y, recompute v, then y flux in x
 * (omp3/flow.c:724-774). io.u holds u after the x sub-updates. */
template <typename T>
__global__ __launch_bounds__(kNT, FLOWHIP_MIN_WAVES) void k_momentum_y_yx(Geo<T> g, MomIO<T> io,
                                                       const double* dtp,
                                                       TileList tl) {
  constexpr int TX = kTX, TY = kTY, NT = kNT;
  const int nx = g.nx, ny = g.ny, sx = nx + 1;
  const T dt_h = T(dtp[1]);
  __shared__ T s_vfy[TY + 2][TX + 4]; /* cell rows [y0-1, y1], cols [x0-2, x1+1] */
  __shared__ T s_v1[TY + 1][TX + 4];  /* y faces [y0, y1], cols [x0-2, x1+1] */
  __shared__ T s_m1[TY + 1][TX];
  __shared__ T s_vfx[TY + 1][TX + 1]; /* corners: y faces [y0, y1], faces [x0, x1] */
  TileWalk walk(tl);
  int tile;
  while (walk.next(tile)) {
    const int x0 = kPad + (tile % tl.ntx) * TX, y0 = kPad + (tile / tl.ntx) * TY;
    const int x1 = min(x0 + TX, nx - kPad), y1 = min(y0 + TY, ny - kPad);
    auto body = [&](auto edge_tag) {
      constexpr bool EDGE = decltype(edge_tag)::value;

      /* momentum_y_flux_in_y on cell rows [y0-1, y1] of interior columns */
      region<NT, TY + 2, TX + 4>([&](int rr, int cc) {
        const int r = y0 - 1 + rr, q = x0 - 2 + cc;
        if (r > y1 || q > x1 + 1 || ghost<EDGE>(q, nx)) return;
        const int i = mir<EDGE>(r, ny);
        const int c = i * nx + q;
        const T v_cell = T(0.5) * (io.v0[c] + io.v0[c + nx]);
        s_vfy[rr][cc] = momentum_flux(g.cdx[q], io.F_y[c], io.F_y[c + nx], io.v0[c - nx],
                                      io.v0[c], io.v0[c + nx], io.v0[c + 2 * nx], g.iedy[i],
                                      v_cell, g.cdy[i - 1], g.cdy[i], dt_h);
      });
      __syncthreads();
      /* first sub-update and velocity recompute, omp3/flow.c:732-748; a ghost
       * column is its mirror column (INVERT_Y leaves columns unsigned) */
      region<NT, TY + 1, TX + 4>([&](int rr, int cc) {
        const int gf = y0 + rr, q = x0 - 2 + cc;
        if (gf > y1 || q > x1 + 1) return;
        const int j = mir<EDGE>(q, nx);
        const int o = j - (x0 - 2);
        const T m1 = advect_update(io.mom1[gf * nx + j], dt_h, s_vfy[rr + 1][o],
                                   s_vfy[rr][o], g.cdx[j], g.edy[gf]);
        const int c = gf * nx + j;
        const T rey = rho_edge_y(io.rho[c], io.rho[c - nx], g.cdx[j], g.cdy[gf],
                                 g.cdy[gf - 1], g.edy[gf]);
        s_v1[rr][cc] = velocity(m1, rey);
        if (q == j && j >= x0 && j < x1) s_m1[rr][j - x0] = m1;
      });
      __syncthreads();
      /* momentum_y_flux_in_x on corners, omp3/flow.c:890-922 */
      region<NT, TY + 1, TX + 1>([&](int rr, int cc) {
        const int gf = y0 + rr, f = x0 + cc;
        if (gf > y1 || f > x1) return;
        const int fi = gf * sx + f;
        const T u_cell = T(0.5) * (io.u[fi] + io.u[fi - sx]);
        s_vfx[rr][cc] = momentum_flux(g.cdy[gf], io.F_x[fi], io.F_x[fi - sx], s_v1[rr][cc],
                                      s_v1[rr][cc + 1], s_v1[rr][cc + 2], s_v1[rr][cc + 3],
                                      g.iedx[f], u_cell, g.cdx[f - 1], g.cdx[f], dt_h);
      });
      __syncthreads();
      /* second sub-update, omp3/flow.c:764-771 */
      region<NT, TY + 1, TX>([&](int rr, int cc) {
        const int gf = y0 + rr, j = x0 + cc;
        if (gf > y1 || j >= x1 || !(gf < y1 || y1 == ny - kPad)) return;
        io.mom[gf * nx + j] = advect_update(s_m1[rr][cc], dt_h, s_vfx[rr][cc + 1],
                                            s_vfx[rr][cc], g.edx[j], g.cdy[gf]);
        put_yface<EDGE>(io.vel, nx, ny, gf, j, s_v1[rr][cc + 2], T(-1.0));
      });
      __syncthreads();
    };
    if (x0 - 3 < kPad || x1 + 3 > nx - kPad || y0 - 3 < kPad || y1 + 3 > ny - kPad)
      body(std::true_type{});
    else
      body(std::false_type{});
  }
}

}  // namespace flowhip

/* ======================================================================= */
/* launchers                                                               */
/* ======================================================================= */

namespace flowhip {
namespace {

template <typename T>
Geo<T> geo_of(int nx, int ny, const T* edx, const T* edy, const T* cdx,
              const T* cdy, const T* iedx, const T* iedy) {
  Geo<T> g;
  g.nx = nx;
  g.ny = ny;
  g.edx = edx;
  g.edy = edy;
  g.cdx = cdx;
  g.cdy = cdy;
  g.iedx = iedx;
  g.iedy = iedy;
  return g;
}

/* Persistent grid: as many blocks as the CUs hold at once for this kernel
 * (queried once per kernel), capped by the tile count and the reduction
 * scratch. */
template <typename K>
int grid_for(K kernel, int& per_cu, int ntiles) {
  static int cus = 0;
  if (cus == 0) {
    int dev = 0;
    hipDeviceProp_t p;
    if (hipGetDevice(&dev) != hipSuccess ||
        hipGetDeviceProperties(&p, dev) != hipSuccess)
      cus = 256;
    else
      cus = p.multiProcessorCount;
  }
  if (per_cu == 0 &&
      (hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, kernel, kNT, 0) !=
           hipSuccess ||
       per_cu < 1))
    per_cu = 1;
  long g = (long)cus * per_cu;
  if (g > ntiles) g = ntiles;
  if (g > kMaxReduceBlocks) g = kMaxReduceBlocks;
  return g < 1 ? 1 : (int)g;
}

}  // namespace

template <typename T>
hipError_t launch_lagrangian(Geo<T> g, LagIO<T> io, FkReduce r, double max_dt,
                             int first_step, hipStream_t s) {
  const TileList tl = make_tiles(g.nx, g.ny, kTX, kTY);
  if (tl.ntiles < 1) return hipErrorInvalidValue;
  static int per_cu = 0;
  const int grid = grid_for(k_lagrangian<T>, per_cu, tl.ntiles);
  hipLaunchKernelGGL(k_lagrangian<T>, dim3(grid), dim3(kNT), 0, s, g, io,
                     (const double*)r.dtp, r, max_dt, first_step, tl);
  return hipGetLastError();
}

template <typename T>
hipError_t launch_mass_energy(Geo<T> g, MassIO<T> io, int tt, FkReduce r,
                              hipStream_t s) {
  const TileList tl = make_tiles(g.nx, g.ny, kTX, kTY);
  if (tl.ntiles < 1) return hipErrorInvalidValue;
  if (tt % 2 == 0) {
    static int per_cu = 0;
    const int grid = grid_for(k_mass_energy_xy<T>, per_cu, tl.ntiles);
    hipLaunchKernelGGL(k_mass_energy_xy<T>, dim3(grid), dim3(kNT), 0, s, g, io,
                       (const double*)r.dtp, r, tl);
  } else {
    static int per_cu = 0;
    const int grid = grid_for(k_mass_energy_yx<T>, per_cu, tl.ntiles);
    hipLaunchKernelGGL(k_mass_energy_yx<T>, dim3(grid), dim3(kNT), 0, s, g, io,
                       (const double*)r.dtp, r, tl);
  }
  return hipGetLastError();
}

template <typename T>
hipError_t launch_momentum(Geo<T> g, MomIO<T> io, int tt, int ydir,
                           const double* dtp, hipStream_t s) {
  const TileList tl = make_tiles(g.nx, g.ny, kTX, kTY);
  if (tl.ntiles < 1) return hipErrorInvalidValue;
  const bool odd = (tt % 2) != 0;
  if (!ydir && odd) {
    static int per_cu = 0;
    const int grid = grid_for(k_momentum_x_xy<T>, per_cu, tl.ntiles);
    hipLaunchKernelGGL(k_momentum_x_xy<T>, dim3(grid), dim3(kNT), 0, s, g, io, dtp, tl);
  } else if (!ydir) {
    static int per_cu = 0;
    const int grid = grid_for(k_momentum_x_yx<T>, per_cu, tl.ntiles);
    hipLaunchKernelGGL(k_momentum_x_yx<T>, dim3(grid), dim3(kNT), 0, s, g, io, dtp, tl);
  } else if (odd) {
    static int per_cu = 0;
    const int grid = grid_for(k_momentum_y_xy<T>, per_cu, tl.ntiles);
    hipLaunchKernelGGL(k_momentum_y_xy<T>, dim3(grid), dim3(kNT), 0, s, g, io, dtp, tl);
  } else {
    static int per_cu = 0;
    const int grid = grid_for(k_momentum_y_yx<T>, per_cu, tl.ntiles);
    hipLaunchKernelGGL(k_momentum_y_yx<T>, dim3(grid), dim3(kNT), 0, s, g, io, dtp, tl);
  }
  return hipGetLastError();
}

}  // namespace flowhip

using namespace flowhip;

static Geo<double> geo64(FkMesh m) {
  return geo_of<double>(m.nx, m.ny, m.edgedx, m.edgedy, m.celldx, m.celldy,
                        m.inv_edgedx, m.inv_edgedy);
}

extern "C" {

hipError_t fk_lagrangian(FkMesh m, const double* rho, const double* e,
                         const double* mom_x, const double* mom_y, double* mx1,
                         double* my1, double* u0, double* v0, double* e1,
                         double* rho_old, FkReduce r, double max_dt,
                         int first_step, hipStream_t s) {
  LagIO<double> io{rho, e, mom_x, mom_y, mx1, my1, u0, v0, e1, rho_old, r.cfl};
  return launch_lagrangian<double>(geo64(m), io, r, max_dt, first_step, s);
}

hipError_t fk_advect_mass_energy(FkMesh m, int tt, const double* u0,
                                 const double* v0, const double* rho0,
                                 const double* e1, double* rho, double* e,
                                 double* F_x, double* F_y, FkReduce r,
                                 hipStream_t s) {
  MassIO<double> io{u0, v0, rho0, e1, rho, e, F_x, F_y};
  return launch_mass_energy<double>(geo64(m), io, tt, r, s);
}

hipError_t fk_advect_momentum_x(FkMesh m, int tt, const double* u0,
                                const double* v0, const double* F_x,
                                const double* F_y, const double* rho,
                                const double* mx1, double* mom_x, double* u,
                                const double* dtp, hipStream_t s) {
  MomIO<double> io{u0, v0, F_x, F_y, rho, mx1, mom_x, u};
  return launch_momentum<double>(geo64(m), io, tt, 0, dtp, s);
}

hipError_t fk_advect_momentum_y(FkMesh m, int tt, const double* u1,
                                const double* v0, const double* F_x,
                                const double* F_y, const double* rho,
                                const double* my1, double* mom_y, double* v,
                                const double* dtp, hipStream_t s) {
  MomIO<double> io{u1, v0, F_x, F_y, rho, my1, mom_y, v};
  return launch_momentum<double>(geo64(m), io, tt, 1, dtp, s);
}

}  // extern "C"
