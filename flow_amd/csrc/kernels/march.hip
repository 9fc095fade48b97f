/*
 * march.hip -- launchers of the row-marching kernels on one mesh (FP64 and
 * the FP32-state form). The kernels are in march.h.
 */
#include "march.h"

namespace flowhip {

/* Chunking of a march kernel when no chunk height is given (chunk_rows 0):
 * units of H2 rows, chunks of k units, and nsmall one-unit chunks at the end
 * of each of the 8 XCD row bands (MarchList). Measured on MI355X at 8192^2,
 * uniform chunks (k = 1, nsmall = 0) of 24 / 20 / 16 rows (Lagrangian /
 * mass-energy / momentum; 24 / 16 / 16 in FP32; with two columns per lane
 * 12 / 12 for the FP64 Lagrangian and mass/energy marches, 10 / 10 in
 * FP32) are fastest (FP32 momentum: 16 rows 2.8% faster than 20 once the
 * last row of a chunk loads nothing, 12 and 14 the same as 16): tapering from
 * tall chunks (48-80 rows) down to 12-24-row ones cut the re-read prologue
 * rows but made the Lagrangian and mass/energy marches 5-10% slower. An
 * explicit chunk_rows > 0 gives uniform chunks of that height. Every
 * chunking gives bitwise the same result. */
struct Taper {
  int H2, k, nsmall;
};
#ifndef FLOWHIP_TAPER_LAG
#define FLOWHIP_TAPER_LAG 24, 1, 0
#endif
#ifndef FLOWHIP_TAPER_MASS
#define FLOWHIP_TAPER_MASS 20, 1, 0
#endif
#ifndef FLOWHIP_TAPER_MOM
#define FLOWHIP_TAPER_MOM 16, 1, 0
#endif
#ifndef FLOWHIP_TAPER_LAG32
#define FLOWHIP_TAPER_LAG32 24, 1, 0
#endif
#ifndef FLOWHIP_TAPER_MASS32
#define FLOWHIP_TAPER_MASS32 16, 1, 0
#endif
#ifndef FLOWHIP_TAPER_MOM32
#define FLOWHIP_TAPER_MOM32 16, 1, 0
#endif
/* the FP32 mass/energy march with two columns per lane (three waves per
 * SIMD): 10 rows (measured 8 to 32; 14 rows +0.3%, 20 +3%) */
#ifndef FLOWHIP_TAPER_MASS32P
#define FLOWHIP_TAPER_MASS32P 10, 1, 0
#endif
/* the FP32 Lagrangian march with two columns per lane (three waves per
 * SIMD): 10 rows (measured 4 to 24; 12 rows +0.5%, 16 +4%, 20 +5%) */
#ifndef FLOWHIP_TAPER_LAG32P
#define FLOWHIP_TAPER_LAG32P 10, 1, 0
#endif
/* the FP64 Lagrangian and mass/energy marches with two columns per lane
 * (two waves per SIMD): 12 rows each (measured 8 to 24; 8 rows +2-5%,
 * 24 +3%, 10 to 20 within 1%) */
#ifndef FLOWHIP_TAPER_LAG64P
#define FLOWHIP_TAPER_LAG64P 12, 1, 0
#endif
#ifndef FLOWHIP_TAPER_MASS64P
#define FLOWHIP_TAPER_MASS64P 12, 1, 0
#endif
enum { kKLag = 0, kKMass = 1, kKMom = 2 };
template <typename T>
constexpr Taper taper_of(int kernel) {
  return sizeof(T) == 8 ? (kernel == kKLag    ? Taper{FLOWHIP_TAPER_LAG}
                           : kernel == kKMass ? Taper{FLOWHIP_TAPER_MASS}
                                              : Taper{FLOWHIP_TAPER_MOM})
                        : (kernel == kKLag    ? Taper{FLOWHIP_TAPER_LAG32}
                           : kernel == kKMass ? Taper{FLOWHIP_TAPER_MASS32}
                                              : Taper{FLOWHIP_TAPER_MOM32});
}

inline MarchList march_items(int nx, int ny, int W, int chunk_rows, Taper t) {
  if (chunk_rows > 0)
    return make_march(nx, ny, W, chunk_rows < kMaxChunkRows ? chunk_rows : kMaxChunkRows);
  const int k = t.k * t.H2 <= kMaxChunkRows ? t.k : kMaxChunkRows / t.H2;
  return make_march(nx, ny, W, t.H2, k, t.nsmall, kMarchBands);
}

template <typename T, bool P2, bool UNI, bool PAIR = false>
hipError_t launch_march_mass_p(MGeo<T> g, MMassIO<T> io, int tt, FkReduce r,
                               int chunk_rows, hipStream_t s) {
  const MarchList ml = march_items(
      g.nx, g.ny, PAIR ? kMassW2 : kMassW, chunk_rows,
      PAIR ? (sizeof(T) == 8 ? Taper{FLOWHIP_TAPER_MASS64P} : Taper{FLOWHIP_TAPER_MASS32P})
           : taper_of<T>(kKMass));
  if (ml.nitems < 1 || !r.queue || !r.totals) return hipErrorInvalidValue;
  if (tt % 2 == 0) {
    static int per_cu = 0;
    const int grid = march_grid(k_march_mass_xy<T, P2, UNI, PAIR>, per_cu, ml.nitems,
                                PAIR ? kMassPairWaves<T> : kMarchMaxWaves<T>);
    hipLaunchKernelGGL((k_march_mass_xy<T, P2, UNI, PAIR>), dim3(grid), dim3(kMNT), 0, s,
                       g, io, (const double*)r.dtp, r, ml);
  } else {
    static int per_cu = 0;
    const int grid = march_grid(k_march_mass_yx<T, P2, UNI, PAIR>, per_cu, ml.nitems,
                                PAIR ? kMassPairWaves<T> : kMarchMaxWaves<T>);
    hipLaunchKernelGGL((k_march_mass_yx<T, P2, UNI, PAIR>), dim3(grid), dim3(kMNT), 0, s,
                       g, io, (const double*)r.dtp, r, ml);
  }
  return hipGetLastError();
}

/* Whether an FP32 mass/energy march can hold two columns per lane: rows of
 * the cell and y-face arrays it reads or writes as column pairs start on
 * 8-byte boundaries (nx even, arrays 8-byte aligned; strips start at even
 * columns). */
/* Whether the two-column form runs: FkMesh.pairs 2 always, 1 on meshes of
 * at least FK_PAIR_MIN_CELLS interior cells. */
inline bool pairs_wanted(int pairs, int nx, int ny) {
  return pairs == 2 ||
         (pairs == 1 && (long)(nx - 2 * kPad) * (ny - 2 * kPad) >= FK_PAIR_MIN_CELLS);
}

template <typename T>
inline bool pairs_fit(const MGeo<T>& g, const MMassIO<T>& io) {
  const unsigned long long bits =
      (unsigned long long)io.v0 | (unsigned long long)io.rho0 | (unsigned long long)io.e1 |
      (unsigned long long)io.rho | (unsigned long long)io.e | (unsigned long long)io.F_y;
  return g.nx % 2 == 0 && (bits & (2 * sizeof(T) - 1)) == 0;
}

template <typename T, bool P2, bool UNI, bool PAIR = false>
hipError_t launch_march_lagrangian_p(MGeo<T> g, MLagIO<T> io, FkReduce r, double max_dt,
                                     int first_step, int chunk_rows, hipStream_t s) {
  const MarchList ml = march_items(
      g.nx, g.ny, PAIR ? kLagW2 : kLagW, chunk_rows,
      PAIR ? (sizeof(T) == 8 ? Taper{FLOWHIP_TAPER_LAG64P} : Taper{FLOWHIP_TAPER_LAG32P})
           : taper_of<T>(kKLag));
  if (ml.nitems < 1 || !r.queue) return hipErrorInvalidValue;
  static int per_cu = 0;
  const int grid = march_grid(k_march_lagrangian<T, P2, UNI, PAIR>, per_cu, ml.nitems,
                              PAIR ? kLagPairWaves<T> : kLagMaxWaves<T>);
  hipLaunchKernelGGL((k_march_lagrangian<T, P2, UNI, PAIR>), dim3(grid), dim3(kMNT), 0, s,
                     g, io, (const double*)r.dtp, r, max_dt, first_step, ml);
  return hipGetLastError();
}

/* Whether a Lagrangian march can hold two columns per lane (as pairs_fit,
 * for the arrays it reads or writes in column pairs: aligned to two
 * values). */
template <typename T>
inline bool lag_pairs_fit(const MGeo<T>& g, const MLagIO<T>& io) {
  const unsigned long long bits =
      (unsigned long long)io.rho | (unsigned long long)io.e | (unsigned long long)io.mom_y |
      (unsigned long long)io.my1 | (unsigned long long)io.v0 | (unsigned long long)io.e1 |
      (unsigned long long)io.rho_old;
  return g.nx % 2 == 0 && (bits & (2 * sizeof(T) - 1)) == 0;
}

/* Even-tt x momentum needs faces [x0-2, x1+1] and one lane below them. */
constexpr int kMomXW = kWave - 5;
constexpr int kMomW = kWave - 4;

template <typename T, bool P2>
hipError_t launch_march_momentum_p(MGeo<T> g, MMomIO<T> io, int tt, int ydir,
                                   const double* dtp, int chunk_rows, hipStream_t s) {
  const bool odd = (tt % 2) != 0;
  const MarchList ml =
      march_items(g.nx, g.ny, (!ydir && !odd) ? kMomXW : kMomW, chunk_rows,
                  taper_of<T>(kKMom));
  if (ml.nitems < 1) return hipErrorInvalidValue;
  if (!ydir && odd) {
    static int per_cu = 0;
    const int grid = march_grid(k_march_momx_xy<T, P2>, per_cu, ml.nitems);
    hipLaunchKernelGGL((k_march_momx_xy<T, P2>), dim3(grid), dim3(kMNT), 0, s, g, io, dtp, ml);
  } else if (!ydir) {
    static int per_cu = 0;
    const int grid = march_grid(k_march_momx_yx<T, P2>, per_cu, ml.nitems);
    hipLaunchKernelGGL((k_march_momx_yx<T, P2>), dim3(grid), dim3(kMNT), 0, s, g, io, dtp, ml);
  } else if (odd) {
    static int per_cu = 0;
    const int grid = march_grid(k_march_momy_xy<T, P2>, per_cu, ml.nitems);
    hipLaunchKernelGGL((k_march_momy_xy<T, P2>), dim3(grid), dim3(kMNT), 0, s, g, io, dtp, ml);
  } else {
    static int per_cu = 0;
    const int grid = march_grid(k_march_momy_yx<T, P2>, per_cu, ml.nitems);
    hipLaunchKernelGGL((k_march_momy_yx<T, P2>), dim3(grid), dim3(kMNT), 0, s, g, io, dtp, ml);
  }
  return hipGetLastError();
}


/* both momentum directions in one march: strips of 59 columns */
constexpr int kMomXYW = kWave - 5;

template <typename T, bool P2, bool UNI>
hipError_t launch_march_momxy_p(MGeo<T> g, MMom2IO<T> io, int tt, const double* dtp,
                                unsigned int* queue, int chunk_rows, hipStream_t s) {
  const MarchList ml = march_items(g.nx, g.ny, kMomXYW, chunk_rows, taper_of<T>(kKMom));
  if (ml.nitems < 1 || !queue) return hipErrorInvalidValue;
  if (tt % 2) {
    static int per_cu = 0;
    const int grid =
        march_grid(k_march_momxy_xy<T, P2, UNI>, per_cu, ml.nitems, kMarchMaxWaves<T>);
    hipLaunchKernelGGL((k_march_momxy_xy<T, P2, UNI>), dim3(grid), dim3(kMNT), 0, s, g,
                       io, dtp, queue, ml);
  } else {
    static int per_cu = 0;
    const int grid =
        march_grid(k_march_momxy_yx<T, P2, UNI>, per_cu, ml.nitems, kMarchMaxWaves<T>);
    hipLaunchKernelGGL((k_march_momxy_yx<T, P2, UNI>), dim3(grid), dim3(kMNT), 0, s, g,
                       io, dtp, queue, ml);
  }
  return hipGetLastError();
}

/* Dispatch on whether every mesh spacing is a power of two (quotients by
 * spacings become products with exact reciprocals) and, for the mass/energy
 * and fused momentum marches, whether such a mesh is also uniform (row
 * geometry read once per item). */
template <typename T>
hipError_t launch_march_mass(MGeo<T> g, MMassIO<T> io, int tt, FkReduce r,
                             int chunk_rows, int pow2, int uniform, int pairs, hipStream_t s) {
  if (pairs_wanted(pairs, g.nx, g.ny) && pow2 && uniform && pairs_fit(g, io))
    return launch_march_mass_p<T, true, true, true>(g, io, tt, r, chunk_rows, s);
  if (pow2 && uniform) return launch_march_mass_p<T, true, true>(g, io, tt, r, chunk_rows, s);
  return pow2 ? launch_march_mass_p<T, true, false>(g, io, tt, r, chunk_rows, s)
              : launch_march_mass_p<T, false, false>(g, io, tt, r, chunk_rows, s);
}

/* The Lagrangian march takes the uniform-mesh form in FP32 only (in FP64 it
 * measured 1% slower than the per-row loads). */
template <typename T>
hipError_t launch_march_lagrangian(MGeo<T> g, MLagIO<T> io, FkReduce r, double max_dt,
                                   int first_step, int chunk_rows, int pow2, int uniform,
                                   int pairs, hipStream_t s) {
  if (pairs_wanted(pairs, g.nx, g.ny) && pow2 && uniform && lag_pairs_fit(g, io))
    return launch_march_lagrangian_p<T, true, true, true>(g, io, r, max_dt, first_step,
                                                          chunk_rows, s);
  if constexpr (sizeof(T) == 4) {
    if (pow2 && uniform)
      return launch_march_lagrangian_p<T, true, true>(g, io, r, max_dt, first_step,
                                                      chunk_rows, s);
  }
  return pow2 ? launch_march_lagrangian_p<T, true, false>(g, io, r, max_dt, first_step,
                                                          chunk_rows, s)
              : launch_march_lagrangian_p<T, false, false>(g, io, r, max_dt, first_step,
                                                           chunk_rows, s);
}

template <typename T>
hipError_t launch_march_momentum(MGeo<T> g, MMomIO<T> io, int tt, int ydir,
                                 const double* dtp, int chunk_rows, int pow2,
                                 hipStream_t s) {
  return pow2 ? launch_march_momentum_p<T, true>(g, io, tt, ydir, dtp, chunk_rows, s)
              : launch_march_momentum_p<T, false>(g, io, tt, ydir, dtp, chunk_rows, s);
}

template <typename T>
hipError_t launch_march_momxy(MGeo<T> g, MMom2IO<T> io, int tt, const double* dtp,
                              unsigned int* queue, int chunk_rows, int pow2, int uniform,
                              hipStream_t s) {
  if (pow2 && uniform)
    return launch_march_momxy_p<T, true, true>(g, io, tt, dtp, queue, chunk_rows, s);
  return pow2 ? launch_march_momxy_p<T, true, false>(g, io, tt, dtp, queue, chunk_rows, s)
              : launch_march_momxy_p<T, false, false>(g, io, tt, dtp, queue, chunk_rows, s);
}

}  // namespace flowhip

using namespace flowhip;

static MGeo<double> mgeo64(FkMesh m) {
  MGeo<double> g{m.nx, m.ny, m.edgedx, m.edgedy, m.celldx, m.celldy,
                 m.inv_edgedx, m.inv_edgedy};
  return g;
}

extern "C" {

hipError_t fk_march_lagrangian(FkMesh m, const double* rho, const double* e,
                               const double* mom_x, const double* mom_y, double* mx1,
                               double* my1, double* u0, double* v0, double* e1,
                               double* rho_old, FkReduce r, double max_dt,
                               int first_step, int chunk_rows, hipStream_t s) {
  MLagIO<double> io{rho, e, mom_x, mom_y, mx1, my1, u0, v0, e1, rho_old, r.cfl};
  return launch_march_lagrangian<double>(mgeo64(m), io, r, max_dt, first_step,
                                         chunk_rows, m.pow2, m.uniform, m.pairs, s);
}

hipError_t fk_march_momentum(FkMesh m, int tt, const double* u0, const double* v0,
                             const double* F_x, const double* F_y, const double* rho,
                             const double* mx1, const double* my1, double* mom_x,
                             double* mom_y, double* u, double* v, FkReduce r,
                             int chunk_rows, hipStream_t s) {
  MMom2IO<double> io{u0, v0, F_x, F_y, rho, mx1, my1, mom_x, mom_y, u, v};
  return launch_march_momxy<double>(mgeo64(m), io, tt, r.dtp, r.queue, chunk_rows,
                                    m.pow2, m.uniform, s);
}

hipError_t fk_march_momentum_x(FkMesh m, int tt, const double* u0,
                               const double* v0, const double* F_x,
                               const double* F_y, const double* rho,
                               const double* mx1, double* mom_x, double* u,
                               const double* dtp, int chunk_rows, hipStream_t s) {
  MMomIO<double> io{u0, v0, F_x, F_y, rho, mx1, mom_x, u};
  return launch_march_momentum<double>(mgeo64(m), io, tt, 0, dtp, chunk_rows, m.pow2, s);
}

hipError_t fk_march_momentum_y(FkMesh m, int tt, const double* u1,
                               const double* v0, const double* F_x,
                               const double* F_y, const double* rho,
                               const double* my1, double* mom_y, double* v,
                               const double* dtp, int chunk_rows, hipStream_t s) {
  MMomIO<double> io{u1, v0, F_x, F_y, rho, my1, mom_y, v};
  return launch_march_momentum<double>(mgeo64(m), io, tt, 1, dtp, chunk_rows, m.pow2, s);
}

hipError_t fk_march_mass_energy(FkMesh m, int tt, const double* u0,
                                const double* v0, const double* rho0,
                                const double* e1, double* rho, double* e,
                                double* F_x, double* F_y, FkReduce r,
                                int chunk_rows, hipStream_t s) {
  MMassIO<double> io{u0, v0, rho0, e1, rho, e, F_x, F_y};
  return launch_march_mass<double>(mgeo64(m), io, tt, r, chunk_rows, m.pow2, m.uniform,
                                   m.pairs, s);
}

static MGeo<float> mgeo32(FkMesh32 m) {
  MGeo<float> g{m.nx, m.ny, m.edgedx, m.edgedy, m.celldx, m.celldy,
                m.inv_edgedx, m.inv_edgedy};
  return g;
}

hipError_t fk_march_lagrangian_f32(FkMesh32 m, const float* rho, const float* e,
                                   const float* mom_x, const float* mom_y,
                                   float* mx1, float* my1, float* u0, float* v0,
                                   float* e1, float* rho_old, FkReduce r,
                                   double max_dt, int first_step, int chunk_rows,
                                   hipStream_t s) {
  MLagIO<float> io{rho, e, mom_x, mom_y, mx1, my1, u0, v0, e1, rho_old};
  return launch_march_lagrangian<float>(mgeo32(m), io, r, max_dt, first_step,
                                        chunk_rows, m.pow2, m.uniform, m.pairs, s);
}

hipError_t fk_march_mass_energy_f32(FkMesh32 m, int tt, const float* u0,
                                    const float* v0, const float* rho0,
                                    const float* e1, float* rho, float* e,
                                    float* F_x, float* F_y, FkReduce r,
                                    int chunk_rows, hipStream_t s) {
  MMassIO<float> io{u0, v0, rho0, e1, rho, e, F_x, F_y};
  return launch_march_mass<float>(mgeo32(m), io, tt, r, chunk_rows, m.pow2, m.uniform,
                                  m.pairs, s);
}

hipError_t fk_march_momentum_f32(FkMesh32 m, int tt, const float* u0, const float* v0,
                                 const float* F_x, const float* F_y, const float* rho,
                                 const float* mx1, const float* my1, float* mom_x,
                                 float* mom_y, float* u, float* v, FkReduce r,
                                 int chunk_rows, hipStream_t s) {
  MMom2IO<float> io{u0, v0, F_x, F_y, rho, mx1, my1, mom_x, mom_y, u, v};
  return launch_march_momxy<float>(mgeo32(m), io, tt, r.dtp, r.queue, chunk_rows,
                                   m.pow2, m.uniform, s);
}

hipError_t fk_march_momentum_x_f32(FkMesh32 m, int tt, const float* u0,
                                   const float* v0, const float* F_x,
                                   const float* F_y, const float* rho,
                                   const float* mx1, float* mom_x, float* u,
                                   const double* dtp, int chunk_rows,
                                   hipStream_t s) {
  MMomIO<float> io{u0, v0, F_x, F_y, rho, mx1, mom_x, u};
  return launch_march_momentum<float>(mgeo32(m), io, tt, 0, dtp, chunk_rows, m.pow2, s);
}

hipError_t fk_march_momentum_y_f32(FkMesh32 m, int tt, const float* u1,
                                   const float* v0, const float* F_x,
                                   const float* F_y, const float* rho,
                                   const float* my1, float* mom_y, float* v,
                                   const double* dtp, int chunk_rows,
                                   hipStream_t s) {
  MMomIO<float> io{u1, v0, F_x, F_y, rho, my1, mom_y, v};
  return launch_march_momentum<float>(mgeo32(m), io, tt, 1, dtp, chunk_rows, m.pow2, s);
}

}  // extern "C"
