/*
 * reduce.hip -- single-pass grid reductions of the stagewise schedule: the
 * CFL timestep minimum (omp3/flow.c:77-105) and the conservation totals
 * (omp3/flow.c:973-995).
 *
 * Each block reduces its share through wave64 shuffles and LDS, writes one
 * partial, and takes a ticket; the last block to arrive combines the
 * partials, writes the result and resets the ticket to zero, so no memset is
 * needed between launches (or inside a captured graph).
 */
#include "reduce_common.h"

namespace flowhip {
namespace {

constexpr int kRedThreads = 256;
constexpr int kRedBlocks = 1024;

/* CFL minimum over the interior, then the dt pair update in the last block.
 * min is exact, so the visiting order does not change the result. NaN
 * candidates (cells with negative energy: sqrt of a negative) are left out,
 * as the reference's min(a, b) = a < b ? a : b leaves them out in effect,
 * and counted. */
template <typename M, typename T>
__global__ __launch_bounds__(kRedThreads) void k_timestep(
    M m, const T* __restrict__ Qxx, const T* __restrict__ Qyy,
    const T* __restrict__ rho, const T* __restrict__ e, FkReduce r,
    double max_dt, int first_step) {
  const int pad = m.pad, wx = m.nx - 2 * pad, wy = m.ny - 2 * pad;
  const long n = (long)wx * wy;
  double local = max_dt;
  int bad = 0;
  for (long k = (long)blockIdx.x * blockDim.x + threadIdx.x; k < n;
       k += (long)gridDim.x * blockDim.x) {
    const int i = pad + (int)(k / wx), j = pad + (int)(k % wx);
    const int c = i * m.nx + j;
    const double cand = cell_timestep(double(e[c]), double(Qxx[c]), double(Qyy[c]),
                                      double(rho[c]), double(m.celldx[j]),
                                      double(m.celldy[i]));
    if (cand == cand)
      local = lesser(local, cand);
    else
      ++bad;
  }
  finish_timestep<kRedThreads>(local, bad, r, max_dt, first_step);
}

/* State of omp3's serial CFL fold over a run of consecutive candidates
 * (row-major order): p is the fold from +inf with min(a, b) = a < b ? a : b,
 * so NaN when the run ends on a NaN candidate and otherwise the minimum
 * over the candidates after the run's last NaN (all of them if none); nan
 * records that the run holds a NaN, any that it holds a candidate. Joining
 * two adjacent runs gives the state of their concatenation, so the grid
 * can fold contiguous ranges independently and join them in order. */
struct SerialFold {
  double p;
  int nan, any;
};

__device__ __forceinline__ SerialFold fold_join(SerialFold a, SerialFold b) {
  if (!b.any) return a;
  if (!a.any) return b;
  /* b's last NaN makes every earlier candidate irrelevant; otherwise b's
   * first candidate replaces a NaN in a's fold and the rest are minima */
  SerialFold o{b.p, a.nan | b.nan, 1};
  if (!b.nan && a.p == a.p) o.p = lesser(a.p, b.p);
  return o;
}

/* The fold of 64 consecutive candidates, one per lane (valid lanes only),
 * the same in every lane. */
__device__ __forceinline__ SerialFold wave_fold(double cand, bool valid) {
  const unsigned long long vmask = __ballot(valid);
  const unsigned long long nmask = __ballot(valid && cand != cand);
  const int lane = threadIdx.x % kWave;
  SerialFold f{__builtin_inf(), nmask != 0ull, vmask != 0ull};
  unsigned long long tail = vmask;
  if (nmask) {
    const int last = 63 - __builtin_clzll(nmask);
    tail = vmask & ~((2ull << last) - 1ull); /* lanes after the last NaN */
  }
  const double x = ((tail >> lane) & 1ull) ? cand : __builtin_inf();
  f.p = wave_min(x);
  if (nmask && !tail) f.p = __builtin_nan("");
  return f;
}

/* omp3/flow.c:77-105 read as one OpenMP thread executes it (see
 * fk_timestep_serial): block b folds a contiguous range of the interior in
 * row-major order, its waves contiguous sub-ranges, 64 cells at a time; the
 * last block joins the blocks' folds in order. */
template <bool kFromArrays>
__global__ __launch_bounds__(kRedThreads) void k_timestep_serial(
    FkMesh m, const double* __restrict__ cfl, const double* __restrict__ Qxx,
    const double* __restrict__ Qyy, const double* __restrict__ rho,
    const double* __restrict__ e, FkReduce r, double max_dt, int first_step) {
  constexpr int kWaves = kRedThreads / kWave;
  __shared__ SerialFold s_fold[kWaves];
  const int pad = m.pad, wx = m.nx - 2 * pad, wy = m.ny - 2 * pad;
  const long n = (long)wx * wy;
  const int lane = threadIdx.x % kWave, w = threadIdx.x / kWave;
  const long b0 = n * blockIdx.x / gridDim.x, b1 = n * (blockIdx.x + 1) / gridDim.x;
  const long k0 = b0 + (b1 - b0) * w / kWaves, k1 = b0 + (b1 - b0) * (w + 1) / kWaves;
  SerialFold acc{__builtin_inf(), 0, 0};
  int bad = 0;
  for (long base = k0; base < k1; base += kWave) {
    const long k = base + lane;
    const bool valid = k < k1;
    double cand = 0.0;
    if (valid) {
      const int i = pad + (int)(k / wx), j = pad + (int)(k % wx);
      const int c = i * m.nx + j;
      cand = kFromArrays ? cell_timestep(e[c], Qxx[c], Qyy[c], rho[c], m.celldx[j],
                                         m.celldy[i])
                         : cfl[c];
    }
    const SerialFold f = wave_fold(cand, valid);
    if (kFromArrays && f.nan) bad += __popcll(__ballot(valid && cand != cand));
    acc = fold_join(acc, f);
  }
  if (kFromArrays && lane == 0 && bad) atomicAdd(r.status, bad);
  if (lane == 0) s_fold[w] = acc;
  __syncthreads();
  if (threadIdx.x == 0) {
    SerialFold blk = s_fold[0];
    for (int q = 1; q < kWaves; ++q) blk = fold_join(blk, s_fold[q]);
    r.partials[2 * blockIdx.x] = blk.p;
    r.partials[2 * blockIdx.x + 1] = (double)(blk.nan + 2 * blk.any);
  }
  if (!last_block_arrives(r.ticket)) return;
  if (threadIdx.x == 0) {
    SerialFold all{__builtin_inf(), 0, 0};
    for (int b = 0; b < (int)gridDim.x; ++b) {
      const int flags = (int)r.partials[2 * b + 1];
      all = fold_join(all, SerialFold{r.partials[2 * b], flags & 1, flags >> 1});
    }
    /* OpenMP's min combiner merges the thread's fold into max_dt: a fold
     * that ends on a NaN leaves max_dt */
    const double g = lesser(all.p, max_dt);
    const double dt_h = r.dtp[1];
    const double final_min = lesser(g, kCM * dt_h);
    const double dt = 0.5 * (kCT * final_min + dt_h);
    r.dtp[0] = dt;
    r.dtp[1] = first_step ? dt : kCT * final_min;
    r.sums[2] = g;
    r.status[1] = atomicExch(r.status, 0);
    *r.ticket = 0u;
  }
}

template <typename M, typename T>
__global__ __launch_bounds__(kRedThreads) void k_conservation(
    M m, const T* __restrict__ rho, const T* __restrict__ e, FkReduce r) {
  const int pad = m.pad, wx = m.nx - 2 * pad, wy = m.ny - 2 * pad;
  const long n = (long)wx * wy;
  CSum ms{0.0, 0.0}, es{0.0, 0.0};
  for (long k = (long)blockIdx.x * blockDim.x + threadIdx.x; k < n;
       k += (long)gridDim.x * blockDim.x) {
    const int i = pad + (int)(k / wx), j = pad + (int)(k % wx);
    const int c = i * m.nx + j;
    ms.add(double(rho[c]));
    es.add(double(rho[c]) * double(e[c]));
  }
  finish_sums<kRedThreads>(ms, es, r);
}

int red_blocks(long n) {
  long b = (n + kRedThreads - 1) / kRedThreads;
  if (b > kRedBlocks) b = kRedBlocks;
  return b < 1 ? 1 : (int)b;
}

}  // namespace
}  // namespace flowhip

using namespace flowhip;

extern "C" {

int fk_reduce_blocks(void) { return 4 * kMaxReduceBlocks; }

hipError_t fk_timestep(FkMesh m, const double* Qxx, const double* Qyy,
                       const double* rho, const double* e, FkReduce r,
                       double max_dt, int first_step, hipStream_t s) {
  const long n = (long)(m.nx - 2 * m.pad) * (m.ny - 2 * m.pad);
  hipLaunchKernelGGL((k_timestep<FkMesh, double>), dim3(red_blocks(n)),
                     dim3(kRedThreads), 0, s, m, Qxx, Qyy, rho, e, r, max_dt,
                     first_step);
  return hipGetLastError();
}

hipError_t fk_timestep_serial(FkMesh m, const double* cfl, const double* Qxx,
                              const double* Qyy, const double* rho,
                              const double* e, FkReduce r, double max_dt,
                              int first_step, hipStream_t s) {
  const long n = (long)(m.nx - 2 * m.pad) * (m.ny - 2 * m.pad);
  /* two partials per block, within fk_reduce_blocks() */
  const int grid = red_blocks(n);
  if (cfl)
    hipLaunchKernelGGL(k_timestep_serial<false>, dim3(grid), dim3(kRedThreads), 0, s, m,
                       cfl, Qxx, Qyy, rho, e, r, max_dt, first_step);
  else
    hipLaunchKernelGGL(k_timestep_serial<true>, dim3(grid), dim3(kRedThreads), 0, s, m,
                       cfl, Qxx, Qyy, rho, e, r, max_dt, first_step);
  return hipGetLastError();
}

hipError_t fk_conservation(FkMesh m, const double* rho, const double* e,
                           FkReduce r, hipStream_t s) {
  const long n = (long)(m.nx - 2 * m.pad) * (m.ny - 2 * m.pad);
  hipLaunchKernelGGL((k_conservation<FkMesh, double>), dim3(red_blocks(n)),
                     dim3(kRedThreads), 0, s, m, rho, e, r);
  return hipGetLastError();
}

hipError_t fk_timestep_f32(FkMesh32 m, const float* Qxx, const float* Qyy,
                           const float* rho, const float* e, FkReduce r,
                           double max_dt, int first_step, hipStream_t s) {
  const long n = (long)(m.nx - 2 * m.pad) * (m.ny - 2 * m.pad);
  hipLaunchKernelGGL((k_timestep<FkMesh32, float>), dim3(red_blocks(n)),
                     dim3(kRedThreads), 0, s, m, Qxx, Qyy, rho, e, r, max_dt,
                     first_step);
  return hipGetLastError();
}

hipError_t fk_conservation_f32(FkMesh32 m, const float* rho, const float* e,
                               FkReduce r, hipStream_t s) {
  const long n = (long)(m.nx - 2 * m.pad) * (m.ny - 2 * m.pad);
  hipLaunchKernelGGL((k_conservation<FkMesh32, float>), dim3(red_blocks(n)),
                     dim3(kRedThreads), 0, s, m, rho, e, r);
  return hipGetLastError();
}

}  // extern "C"
