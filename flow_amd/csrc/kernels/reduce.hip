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
