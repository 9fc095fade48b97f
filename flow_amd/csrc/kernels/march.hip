/*
 * march.hip -- row-marching forms of the fused kernels.
 *
 * One wave owns a strip of W columns and marches down a chunk of H rows.
 * Along x a wave covers the strip plus its stencil halo in its 64 lanes, and
 * x neighbours come from cross-lane shuffles; along y every lane keeps the
 * rows its column still needs in a register window, so nothing along y is
 * recomputed and no block barrier is taken. Input rows are loaded one step
 * ahead of use. Waves of a block are independent.
 *
 * The arithmetic is the same pointwise code (hydro_math.h) in the same
 * order, so results are bitwise those of the tile kernels and of omp3;
 * ghost rows and columns are the mirrored values handle_boundary_2d would
 * copy there.
 */
#include <type_traits>

#include "reduce_common.h"
#include "tile.h"

namespace flowhip {

template <typename T>
struct MGeo {
  int nx, ny;
  const T* __restrict__ edx;
  const T* __restrict__ edy;
  const T* __restrict__ cdx;
  const T* __restrict__ cdy;
  const T* __restrict__ iedx;
  const T* __restrict__ iedy;
};

template <typename T>
struct MMassIO {
  const T* __restrict__ u0;
  const T* __restrict__ v0;
  const T* __restrict__ rho0;
  const T* __restrict__ e1;
  T* __restrict__ rho;
  T* __restrict__ e;
  T* __restrict__ F_x;
  T* __restrict__ F_y;
};

constexpr int kMNT = 256;            /* threads per block: 4 independent waves */
constexpr int kMWaves = kMNT / kWave;

struct MarchList {
  int W, H;           /* strip width (output columns), chunk height (rows) */
  int nstrips, nchunks, nitems;
};

inline MarchList make_march(int nx, int ny, int W, int H) {
  MarchList m;
  m.W = W;
  m.H = H;
  m.nstrips = (nx - 2 * kPad + W - 1) / W;
  m.nchunks = (ny - 2 * kPad + H - 1) / H;
  m.nitems = m.nstrips * m.nchunks;
  return m;
}

template <typename T>
__device__ __forceinline__ T up(T v, int d) {
  return __shfl_up(v, d, kWave);
}
template <typename T>
__device__ __forceinline__ T dn(T v, int d) {
  return __shfl_down(v, d, kWave);
}


/* ---- per-row geometry of a chunk held across the lanes ----
 * Lane k holds rows base+k and base+64+k; row(k) broadcasts one row's value
 * with v_readlane (k is uniform), so the row loop issues no memory load for
 * geometry. Chunks are at most kMaxChunkRows rows. */
constexpr int kMaxChunkRows = 120;

__device__ __forceinline__ double lane_value(double v, int k) {
  const long long b = __double_as_longlong(v);
  const int lo = __builtin_amdgcn_readlane((int)b, k);
  const int hi = __builtin_amdgcn_readlane((int)(b >> 32), k);
  return __longlong_as_double(((long long)hi << 32) | (unsigned int)lo);
}
__device__ __forceinline__ float lane_value(float v, int k) {
  return __int_as_float(__builtin_amdgcn_readlane(__float_as_int(v), k));
}

template <typename T>
struct RowGeo {
  T a, b;
  int base;
  __device__ RowGeo(const T* arr, int base_row, int last_row, int lane) : base(base_row) {
    a = arr[min(base_row + lane, last_row)];
    b = arr[min(base_row + kWave + lane, last_row)];
  }
  /* value at row r (uniform, base <= r < base + 128) */
  __device__ __forceinline__ T operator()(int r) const {
    const int k = r - base;
    return k < kWave ? lane_value(a, k) : lane_value(b, k - kWave);
  }
};

/* ---- boundary-strip writers (ghost writes only in edge items) ---- */
template <typename T>
__device__ __forceinline__ void mput_cell(bool edge, T* a, int nx, int ny, int i,
                                          int j, T v) {
  a[i * nx + j] = v;
  if (!edge) return;
  if (j == kPad) a[i * nx + kPad - 1] = v;
  if (j == kPad + 1) a[i * nx + kPad - 2] = v;
  if (j == nx - kPad - 1) a[i * nx + nx - kPad] = v;
  if (j == nx - kPad - 2) a[i * nx + nx - kPad + 1] = v;
  if (i == kPad) a[(kPad - 1) * nx + j] = v;
  if (i == kPad + 1) a[(kPad - 2) * nx + j] = v;
  if (i == ny - kPad - 1) a[(ny - kPad) * nx + j] = v;
  if (i == ny - kPad - 2) a[(ny - kPad + 1) * nx + j] = v;
}
template <typename T>
__device__ __forceinline__ void mput_xface(bool edge, T* a, int nx, int ny, int i,
                                           int f, T v, T sx) {
  const int s = nx + 1;
  a[i * s + f] = v;
  if (!edge) return;
  if (f == kPad) a[i * s + kPad - 1] = sx * v;
  if (f == kPad + 1) a[i * s + kPad - 2] = sx * v;
  if (f == nx - kPad) a[i * s + nx - kPad + 1] = sx * v;
  if (f == nx - kPad - 1) a[i * s + nx - kPad + 2] = sx * v;
  if (i == kPad) a[(kPad - 1) * s + f] = v;
  if (i == kPad + 1) a[(kPad - 2) * s + f] = v;
  if (i == ny - kPad - 1) a[(ny - kPad) * s + f] = v;
  if (i == ny - kPad - 2) a[(ny - kPad + 1) * s + f] = v;
}
template <typename T>
__device__ __forceinline__ void mput_yface(bool edge, T* a, int nx, int ny, int g,
                                           int j, T v, T sy) {
  a[g * nx + j] = v;
  if (!edge) return;
  if (g == kPad) a[(kPad - 1) * nx + j] = sy * v;
  if (g == kPad + 1) a[(kPad - 2) * nx + j] = sy * v;
  if (g == ny - kPad) a[(ny - kPad + 1) * nx + j] = sy * v;
  if (g == ny - kPad - 1) a[(ny - kPad + 2) * nx + j] = sy * v;
  if (j == kPad) a[g * nx + kPad - 1] = v;
  if (j == kPad + 1) a[g * nx + kPad - 2] = v;
  if (j == nx - kPad - 1) a[g * nx + nx - kPad] = v;
  if (j == nx - kPad - 2) a[g * nx + nx - kPad + 1] = v;
}

/* ======================================================================= */
/* mass and energy, x sweep first (even tt)                                 */
/* ======================================================================= */

/* Lane l holds column q = x0 - 2 + l. Per row r: the x fluxes at faces
 * [x0, x1] and the first (x) update of cells [x0, x1); per lane a window of
 * the x-updated rows r-3..r feeds the y flux at face r-1 and the second (y)
 * update of row r-2. */
template <typename T>
__global__ __launch_bounds__(kMNT) void k_march_mass_xy(MGeo<T> g, MMassIO<T> io,
                                                        const double* dtp,
                                                        FkReduce red,
                                                        MarchList ml) {
  const int nx = g.nx, ny = g.ny, sx = nx + 1;
  const T dt = T(dtp[0]), dt_h = T(dtp[1]);
  const int lane = threadIdx.x % kWave;
  CSum ms{0.0, 0.0}, es{0.0, 0.0};
  for (int item = blockIdx.x * kMWaves + threadIdx.x / kWave; item < ml.nitems;
       item += gridDim.x * kMWaves) {
    const int strip = item % ml.nstrips, chunk = item / ml.nstrips;
    const int x0 = kPad + strip * ml.W, x1 = min(x0 + ml.W, nx - kPad);
    const int y0 = kPad + chunk * ml.H, y1 = min(y0 + ml.H, ny - kPad);
    const bool edge = (x0 == kPad) || (x1 + 3 > nx - kPad) || (y0 == kPad) ||
                      (y1 + 3 > ny - kPad);
    const int q = x0 - 2 + lane;
    const bool inq = q < nx;                     /* column exists */
    const bool face = q >= x0 && q <= x1;        /* lane computes face q */
    const bool own = q >= x0 && q < x1;          /* lane outputs cell q */
    const bool own_face = face && (q < x1 || x1 == nx - kPad);
    const int qc = inq ? q : nx - 1;             /* a valid column for loads */
    const int qf = q <= nx ? q : nx;             /* a valid x face for loads */
    /* per-lane geometry, loaded once per item: no load inside the row loop
     * other than the next row's prefetch */
    const T cdx_q = g.cdx[qc], edx_q = g.edx[qc], edx_q1 = g.edx[qc + 1];
    const int qg = face ? q : x0;
    const T cdx_qm1 = g.cdx[qg - 1], iedx_q = g.iedx[qg];
    const RowGeo<T> edy(g.edy, y0 - 4, ny, lane), cdy(g.cdy, y0 - 4, ny, lane);
    const RowGeo<T> iedy(g.iedy, y0 - 4, ny, lane);

    /* rows of the x-updated fields for this lane's column: w[0] oldest */
    T wr[4], we[4];
    T vy[3];   /* v0 at faces r-2 .. r */
    T r0[3];   /* rho0 at rows r-2 .. r */
    T fy_prev = T(0), efy_prev = T(0);
    const bool south = (y0 == kPad);
    const int rb = south ? kPad : y0 - 2;
#pragma unroll
    for (int k = 0; k < 4; ++k) wr[k] = we[k] = T(0);
    vy[0] = io.v0[(rb - 2) * nx + qc];
    vy[1] = io.v0[(rb - 1) * nx + qc];
    r0[0] = io.rho0[(rb - 2) * nx + qc];
    r0[1] = io.rho0[(rb - 1) * nx + qc];
    /* row rb's inputs; later rows are loaded one step ahead */
    T a_r = io.rho0[rb * nx + qc], a_e = io.e1[rb * nx + qc];
    T a_u = io.u0[rb * sx + qf], a_v = io.v0[rb * nx + qc];
    for (int r = rb; r <= y1 + 1; ++r) {
      const bool ghost_row = r >= ny - kPad;
      /* prefetch the next row */
      const int rn = min(r + 1, ny - kPad - 1);
      const T n_r = io.rho0[rn * nx + qc], n_e = io.e1[rn * nx + qc];
      const T n_u = io.u0[rn * sx + qf];
      const T n_v = io.v0[min(r + 1, ny) * nx + qc];
      vy[2] = a_v;
      r0[2] = a_r;

      T xr, xe;
      if (!ghost_row) {
        /* x fluxes at face q from the row in the lanes around it */
        const T rm2 = up(a_r, 2), rm1 = up(a_r, 1), rp1 = dn(a_r, 1);
        const T em2 = up(a_e, 2), em1 = up(a_e, 1), ep1 = dn(a_e, 1);
        const T um1 = up(a_u, 1), up1 = dn(a_u, 1);
        FaceFlux<T> fl{T(0), T(0)};
        if (face)
          fl = mass_energy_flux<T, false>(um1, a_u, up1, rm2, rm1, a_r, rp1, em2, em1,
                                          a_e, ep1, iedx_q, cdx_qm1, cdx_q,
                                          edy(r), dt, dt_h);
        if (own_face && r >= y0 && r < y1)
          mput_xface(edge, io.F_x, nx, ny, r, q, fl.mass, T(-1.0));
        const T f_hi = dn(fl.mass, 1), ef_hi = dn(fl.energy, 1);
        const CellState<T> st =
            sweep_update(a_r, a_r, a_e, fl.mass, f_hi, fl.energy, ef_hi, edy(r),
                         edy(r + 1), cdx_q, cdy(r), dt_h, true);
        xr = st.rho;
        xe = st.e;
      } else {
        /* ghost row past the north wall: the mirror row's x update */
        xr = (r == ny - kPad) ? wr[3] : wr[1];
        xe = (r == ny - kPad) ? we[3] : we[1];
      }
#pragma unroll
      for (int k = 0; k < 3; ++k) {
        wr[k] = wr[k + 1];
        we[k] = we[k + 1];
      }
      wr[3] = xr;
      we[3] = xe;
      if (south && r == kPad + 1) {
        /* ghost rows below the south wall mirror rows pad, pad+1 */
        wr[0] = wr[3];
        we[0] = we[3];
        wr[1] = wr[2];
        we[1] = we[2];
      }
      /* y flux at face r-1 from rows r-3 .. r */
      const int gf = r - 1;
      if (gf >= y0 && gf <= y1) {
        const FaceFlux<T> fl = mass_energy_flux<T, true>(
            vy[0], vy[1], vy[2], wr[0], wr[1], wr[2], wr[3], we[0], we[1], we[2],
            we[3], iedy(gf), cdy(gf - 1), cdy(gf), edx_q, dt, dt_h);
        if (own && (gf < y1 || y1 == ny - kPad))
          mput_yface(edge, io.F_y, nx, ny, gf, q, fl.mass, T(-1.0));
        /* y update of row r-2 */
        const int i = r - 2;
        if (i >= y0 && own) {
          const CellState<T> st =
              sweep_update(wr[1], r0[0], we[1], fy_prev, fl.mass, efy_prev,
                           fl.energy, edx_q, edx_q1, cdx_q, cdy(i), dt_h, false);
          mput_cell(edge, io.rho, nx, ny, i, q, st.rho);
          mput_cell(edge, io.e, nx, ny, i, q, st.e);
          ms.add(double(st.rho));
          es.add(double(st.rho) * double(st.e));
        }
        fy_prev = fl.mass;
        efy_prev = fl.energy;
      }
      vy[0] = vy[1];
      vy[1] = vy[2];
      r0[0] = r0[1];
      r0[1] = r0[2];
      a_r = n_r;
      a_e = n_e;
      a_u = n_u;
      a_v = n_v;
    }
  }
  finish_sums<kMNT>(ms, es, red);
}

/* ======================================================================= */
/* mass and energy, y sweep first (odd tt)                                  */
/* ======================================================================= */

/* Lane l holds column q = x0 - 2 + l and runs the y sweep at its mirror
 * column (the same column inside the domain). Per row r: the y flux at face
 * r+1 and the first (y) update of row r; then across the lanes the x fluxes
 * at faces [x0, x1] and the second (x) update of cells [x0, x1). */
template <typename T>
__global__ __launch_bounds__(kMNT) void k_march_mass_yx(MGeo<T> g, MMassIO<T> io,
                                                        const double* dtp,
                                                        FkReduce red,
                                                        MarchList ml) {
  const int nx = g.nx, ny = g.ny, sx = nx + 1;
  const T dt = T(dtp[0]), dt_h = T(dtp[1]);
  const int lane = threadIdx.x % kWave;
  CSum ms{0.0, 0.0}, es{0.0, 0.0};
  for (int item = blockIdx.x * kMWaves + threadIdx.x / kWave; item < ml.nitems;
       item += gridDim.x * kMWaves) {
    const int strip = item % ml.nstrips, chunk = item / ml.nstrips;
    const int x0 = kPad + strip * ml.W, x1 = min(x0 + ml.W, nx - kPad);
    const int y0 = kPad + chunk * ml.H, y1 = min(y0 + ml.H, ny - kPad);
    const bool edge = (x0 == kPad) || (x1 + 3 > nx - kPad) || (y0 == kPad) ||
                      (y1 + 3 > ny - kPad);
    const int q = x0 - 2 + lane;
    const bool inq = q < nx;
    const bool face = q >= x0 && q <= x1;
    const bool own = q >= x0 && q < x1;
    const bool own_face = face && (q < x1 || x1 == nx - kPad);
    /* the y sweep of a ghost column is that of its mirror column */
    const int qm = inq ? mirror(q, nx) : nx - kPad - 1;
    const int qc = inq ? q : nx - 1;
    const int qf = q <= nx ? q : nx;
    const T cdx_m = g.cdx[qm], edx_m = g.edx[qm], edx_m1 = g.edx[qm + 1];
    const T cdx_q = g.cdx[qc];
    const int qg = face ? q : x0;
    const T cdx_qm1 = g.cdx[qg - 1], iedx_q = g.iedx[qg];
    const RowGeo<T> edy(g.edy, y0 - 4, ny, lane), cdy(g.cdy, y0 - 4, ny, lane);
    const RowGeo<T> iedy(g.iedy, y0 - 4, ny, lane);

    /* rho0, e1 at rows r-1 .. r+2 and v0 at faces r .. r+2 of column qm */
    T wr[4], we[4], wv[3];
#pragma unroll
    for (int k = 0; k < 4; ++k) {
      wr[k] = io.rho0[(y0 - 2 + k) * nx + qm];
      we[k] = io.e1[(y0 - 2 + k) * nx + qm];
    }
#pragma unroll
    for (int k = 0; k < 3; ++k) wv[k] = io.v0[(y0 - 1 + k) * nx + qm];
    /* y flux at face y0 (rows y0-2 .. y0+1) */
    FaceFlux<T> fprev = mass_energy_flux<T, true>(
        wv[0], wv[1], wv[2], wr[0], wr[1], wr[2], wr[3], we[0], we[1], we[2], we[3],
        iedy(y0), cdy(y0 - 1), cdy(y0), edx_m, dt, dt_h);
    if (own && q == qm) mput_yface(edge, io.F_y, nx, ny, y0, q, fprev.mass, T(-1.0));
    /* shift: windows now hold rows r-2 .. r+1 for r = y0; load row r+2 */
#pragma unroll
    for (int k = 0; k < 3; ++k) {
      wr[k] = wr[k + 1];
      we[k] = we[k + 1];
    }
    wv[0] = wv[1];
    wv[1] = wv[2];
    T n_r = io.rho0[(y0 + 2) * nx + qm], n_e = io.e1[(y0 + 2) * nx + qm];
    T n_v = io.v0[(y0 + 2) * nx + qm];
    T a_u = io.u0[y0 * sx + qf];
    T a_r0 = io.rho0[y0 * nx + qc];
    for (int r = y0; r < y1; ++r) {
      wr[3] = n_r;
      we[3] = n_e;
      wv[2] = n_v;
      const T c_u = a_u, c_r0 = a_r0;
      /* prefetch the next row's inputs */
      {
        const int rn = min(r + 3, ny);
        n_r = io.rho0[rn * nx + qm];
        n_e = io.e1[rn * nx + qm];
        n_v = io.v0[rn * nx + qm];
        const int rr = min(r + 1, ny - 1);
        a_u = io.u0[rr * sx + qf];
        a_r0 = io.rho0[rr * nx + qc];
      }
      /* y flux at face r+1 (rows r-1 .. r+2) and the y update of row r */
      const int gf = r + 1;
      const FaceFlux<T> fl = mass_energy_flux<T, true>(
          wv[0], wv[1], wv[2], wr[0], wr[1], wr[2], wr[3], we[0], we[1], we[2],
          we[3], iedy(gf), cdy(gf - 1), cdy(gf), edx_m, dt, dt_h);
      if (own && q == qm && (gf < y1 || y1 == ny - kPad))
        mput_yface(edge, io.F_y, nx, ny, gf, q, fl.mass, T(-1.0));
      const CellState<T> ys =
          sweep_update(wr[1], wr[1], we[1], fprev.mass, fl.mass, fprev.energy,
                       fl.energy, edx_m, edx_m1, cdx_m, cdy(r), dt_h, true);
      fprev = fl;
      /* x fluxes at face q of the y-updated row, then the x update */
      const T rm2 = up(ys.rho, 2), rm1 = up(ys.rho, 1), rp1 = dn(ys.rho, 1);
      const T em2 = up(ys.e, 2), em1 = up(ys.e, 1), ep1 = dn(ys.e, 1);
      const T um1 = up(c_u, 1), up1 = dn(c_u, 1);
      FaceFlux<T> fx{T(0), T(0)};
      if (face)
        fx = mass_energy_flux<T, false>(um1, c_u, up1, rm2, rm1, ys.rho, rp1, em2, em1,
                                        ys.e, ep1, iedx_q, cdx_qm1, cdx_q,
                                        edy(r), dt, dt_h);
      if (own_face) mput_xface(edge, io.F_x, nx, ny, r, q, fx.mass, T(-1.0));
      const T f_hi = dn(fx.mass, 1), ef_hi = dn(fx.energy, 1);
      if (own) {
        const CellState<T> st =
            sweep_update(ys.rho, c_r0, ys.e, fx.mass, f_hi, fx.energy, ef_hi,
                         edy(r), edy(r + 1), cdx_q, cdy(r), dt_h, false);
        mput_cell(edge, io.rho, nx, ny, r, q, st.rho);
        mput_cell(edge, io.e, nx, ny, r, q, st.e);
        ms.add(double(st.rho));
        es.add(double(st.rho) * double(st.e));
      }
#pragma unroll
      for (int k = 0; k < 3; ++k) {
        wr[k] = wr[k + 1];
        we[k] = we[k + 1];
      }
      wv[0] = wv[1];
      wv[1] = wv[2];
    }
  }
  finish_sums<kMNT>(ms, es, red);
}


template <typename K>
int march_grid(K kernel, int& per_cu, int nitems) {
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
      (hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, kernel, kMNT, 0) !=
           hipSuccess ||
       per_cu < 1))
    per_cu = 1;
  long gsz = (long)cus * per_cu;
  const long need = (nitems + kMWaves - 1) / kMWaves;
  if (gsz > need) gsz = need;
  if (gsz > kMaxReduceBlocks) gsz = kMaxReduceBlocks;
  return gsz < 1 ? 1 : (int)gsz;
}

/* Strip width: the 64 lanes hold the strip and its x halo of 2 + 2 columns. */
constexpr int kMassW = kWave - 4;

template <typename T>
hipError_t launch_march_mass(MGeo<T> g, MMassIO<T> io, int tt, FkReduce r,
                             int chunk_rows, hipStream_t s) {
  if (chunk_rows > kMaxChunkRows) chunk_rows = kMaxChunkRows;
  const MarchList ml = make_march(g.nx, g.ny, kMassW, chunk_rows);
  if (ml.nitems < 1) return hipErrorInvalidValue;
  if (tt % 2 == 0) {
    static int per_cu = 0;
    const int grid = march_grid(k_march_mass_xy<T>, per_cu, ml.nitems);
    hipLaunchKernelGGL(k_march_mass_xy<T>, dim3(grid), dim3(kMNT), 0, s, g, io,
                       (const double*)r.dtp, r, ml);
  } else {
    static int per_cu = 0;
    const int grid = march_grid(k_march_mass_yx<T>, per_cu, ml.nitems);
    hipLaunchKernelGGL(k_march_mass_yx<T>, dim3(grid), dim3(kMNT), 0, s, g, io,
                       (const double*)r.dtp, r, ml);
  }
  return hipGetLastError();
}

}  // namespace flowhip

using namespace flowhip;

extern "C" {

hipError_t fk_march_mass_energy(FkMesh m, int tt, const double* u0,
                                const double* v0, const double* rho0,
                                const double* e1, double* rho, double* e,
                                double* F_x, double* F_y, FkReduce r,
                                int chunk_rows, hipStream_t s) {
  MGeo<double> g{m.nx, m.ny, m.edgedx, m.edgedy, m.celldx, m.celldy,
                 m.inv_edgedx, m.inv_edgedy};
  MMassIO<double> io{u0, v0, rho0, e1, rho, e, F_x, F_y};
  return launch_march_mass<double>(g, io, tt, r, chunk_rows, s);
}

}  // extern "C"
