/*
 * reduce.hip -- single-pass grid reductions of the hip kernel set: the CFL
 * timestep minimum (omp3/flow.c:77-105) and the conservation totals
 * (omp3/flow.c:973-995).
 *
 * Each block reduces its share through wave64 shuffles and LDS, writes one
 * partial, and takes a ticket; the last block to arrive combines the
 * partials, writes the result and resets the ticket to zero, so no memset is
 * needed between launches (or inside a captured graph).
 */
#include "common.h"
#include "hydro_math.h"

namespace flowhip {
namespace {

constexpr int kRedThreads = 256;
constexpr int kRedBlocks = 1024;

/* Returns true in the one block that arrives last. */
__device__ bool last_block_arrives(unsigned int* ticket) {
  __shared__ bool last;
  __threadfence();
  __syncthreads();
  if (threadIdx.x == 0) {
    const unsigned int prev = atomicAdd(ticket, 1u);
    last = (prev == gridDim.x - 1);
  }
  __syncthreads();
  /* acquire: the other blocks' partials become visible to this block */
  if (last) __threadfence();
  return last;
}

__device__ double block_min(double x) {
  __shared__ double red[kRedThreads / kWave];
  x = wave_min(x);
  const int lane = threadIdx.x % kWave, w = threadIdx.x / kWave;
  if (lane == 0) red[w] = x;
  __syncthreads();
  if (w == 0) {
    x = (lane < kRedThreads / kWave) ? red[lane] : red[0];
    x = wave_min(x);
  }
  return x; /* valid in wave 0 */
}

__device__ CSum block_csum(CSum v) {
  __shared__ double rs[kRedThreads / kWave], rc[kRedThreads / kWave];
  v = wave_csum(v);
  const int lane = threadIdx.x % kWave, w = threadIdx.x / kWave;
  if (lane == 0) {
    rs[w] = v.s;
    rc[w] = v.c;
  }
  __syncthreads();
  if (w == 0) {
    CSum t{0.0, 0.0};
    if (lane < kRedThreads / kWave) t = CSum{rs[lane], rc[lane]};
    v = wave_csum(t);
  }
  return v; /* valid in wave 0 */
}

/* CFL minimum over the interior, then the dt pair update of
 * omp3/flow.c:103-105 in the last block. Cells are visited in a grid-stride
 * loop; min is exact, so the visiting order does not change the result. */
__global__ __launch_bounds__(kRedThreads) void k_timestep(
    FkMesh m, const double* __restrict__ Qxx, const double* __restrict__ Qyy,
    const double* __restrict__ rho, const double* __restrict__ e, FkReduce r,
    double max_dt, int first_step) {
  const int pad = m.pad, wx = m.nx - 2 * pad, wy = m.ny - 2 * pad;
  const long n = (long)wx * wy;
  double local = max_dt;
  int bad = 0;
  for (long k = (long)blockIdx.x * blockDim.x + threadIdx.x; k < n;
       k += (long)gridDim.x * blockDim.x) {
    const int i = pad + (int)(k / wx), j = pad + (int)(k % wx);
    const int c = i * m.nx + j;
    const double cand =
        cell_timestep(e[c], Qxx[c], Qyy[c], rho[c], m.celldx[j], m.celldy[i]);
    /* NaN candidates (cells with negative energy) are left out, as the
     * reference's min(a, b) leaves them out in effect; they are counted. */
    if (cand == cand)
      local = lesser(local, cand);
    else
      ++bad;
  }
  if (bad) atomicAdd(r.status, bad);
  const double bmin = block_min(local);
  if (threadIdx.x == 0) r.partials[blockIdx.x] = bmin;
  if (!last_block_arrives(r.ticket)) return;
  double g = max_dt;
  for (int b = threadIdx.x; b < (int)gridDim.x; b += blockDim.x)
    g = lesser(g, r.partials[b]);
  g = block_min(g);
  if (threadIdx.x == 0) {
    const double dt_h = r.dtp[1];
    const double final_min = lesser(g, kCM * dt_h);
    const double dt = 0.5 * (kCT * final_min + dt_h);
    r.dtp[0] = dt;
    r.dtp[1] = first_step ? dt : kCT * final_min;
    r.sums[2] = g; /* unclamped grid minimum, kept for tests */
    /* NaN-candidate count of this launch moves to status[1] */
    r.status[1] = atomicExch(r.status, 0);
    *r.ticket = 0u;
  }
}

__global__ __launch_bounds__(kRedThreads) void k_conservation(
    FkMesh m, const double* __restrict__ rho, const double* __restrict__ e,
    FkReduce r) {
  const int pad = m.pad, wx = m.nx - 2 * pad, wy = m.ny - 2 * pad;
  const long n = (long)wx * wy;
  CSum ms{0.0, 0.0}, es{0.0, 0.0};
  for (long k = (long)blockIdx.x * blockDim.x + threadIdx.x; k < n;
       k += (long)gridDim.x * blockDim.x) {
    const int i = pad + (int)(k / wx), j = pad + (int)(k % wx);
    const int c = i * m.nx + j;
    ms.add(rho[c]);
    es.add(rho[c] * e[c]);
  }
  ms = block_csum(ms);
  __syncthreads();
  es = block_csum(es);
  if (threadIdx.x == 0) {
    double* p = r.partials + 4 * blockIdx.x;
    p[0] = ms.s;
    p[1] = ms.c;
    p[2] = es.s;
    p[3] = es.c;
  }
  if (!last_block_arrives(r.ticket)) return;
  CSum gm{0.0, 0.0}, ge{0.0, 0.0};
  for (int b = threadIdx.x; b < (int)gridDim.x; b += blockDim.x) {
    const double* p = r.partials + 4 * b;
    gm.merge(CSum{p[0], p[1]});
    ge.merge(CSum{p[2], p[3]});
  }
  gm = block_csum(gm);
  __syncthreads();
  ge = block_csum(ge);
  if (threadIdx.x == 0) {
    r.sums[0] = gm.s + gm.c;
    r.sums[1] = ge.s + ge.c;
    *r.ticket = 0u;
  }
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

int fk_reduce_blocks(void) { return 4 * kRedBlocks; }

hipError_t fk_timestep(FkMesh m, const double* Qxx, const double* Qyy,
                       const double* rho, const double* e, FkReduce r,
                       double max_dt, int first_step, hipStream_t s) {
  const long n = (long)(m.nx - 2 * m.pad) * (m.ny - 2 * m.pad);
  hipLaunchKernelGGL(k_timestep, dim3(red_blocks(n)), dim3(kRedThreads), 0, s,
                     m, Qxx, Qyy, rho, e, r, max_dt, first_step);
  return hipGetLastError();
}

hipError_t fk_conservation(FkMesh m, const double* rho, const double* e,
                           FkReduce r, hipStream_t s) {
  const long n = (long)(m.nx - 2 * m.pad) * (m.ny - 2 * m.pad);
  hipLaunchKernelGGL(k_conservation, dim3(red_blocks(n)), dim3(kRedThreads), 0,
                     s, m, rho, e, r);
  return hipGetLastError();
}

}  // extern "C"
