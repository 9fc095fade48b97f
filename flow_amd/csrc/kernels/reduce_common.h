/*
 * reduce_common.h -- block-level reductions and the single-pass epilogues
 * (last block to arrive combines the per-block partials) shared by the
 * reduction kernels and the fused kernels.
 */
#pragma once

#include "common.h"
#include "hydro_math.h"

namespace flowhip {

/* Upper bound on blocks of any launch that writes reduction partials; the
 * partials buffer holds 4 doubles per block. */
constexpr int kMaxReduceBlocks = 8192;

/* Returns true in the one block that arrives last (all others have
 * finished their work and published their partials). */
__device__ __forceinline__ bool last_block_arrives(unsigned int* ticket) {
  __shared__ bool last;
  __threadfence(); /* release this block's partial */
  __syncthreads();
  if (threadIdx.x == 0) {
    const unsigned int prev = atomicAdd(ticket, 1u);
    last = (prev == gridDim.x - 1);
  }
  __syncthreads();
  if (last) __threadfence(); /* acquire the other blocks' partials */
  return last;
}

template <int NT>
__device__ __forceinline__ double block_min(double x) {
  __shared__ double red[NT / kWave];
  x = wave_min(x);
  const int lane = threadIdx.x % kWave, w = threadIdx.x / kWave;
  __syncthreads();
  if (lane == 0) red[w] = x;
  __syncthreads();
  if (w == 0) {
    x = (lane < NT / kWave) ? red[lane] : red[0];
    x = wave_min(x);
  }
  return x; /* valid in wave 0 */
}

template <int NT>
__device__ __forceinline__ CSum block_csum(CSum v) {
  __shared__ double rs[NT / kWave], rc[NT / kWave];
  v = wave_csum(v);
  const int lane = threadIdx.x % kWave, w = threadIdx.x / kWave;
  __syncthreads();
  if (lane == 0) {
    rs[w] = v.s;
    rc[w] = v.c;
  }
  __syncthreads();
  if (w == 0) {
    CSum t{0.0, 0.0};
    if (lane < NT / kWave) t = CSum{rs[lane], rc[lane]};
    v = wave_csum(t);
  }
  return v; /* valid in wave 0 */
}

/* Per-block CFL minimum -> last block -> (dt, dt_h) update of
 * omp3/flow.c:103-105 and the undefined-candidate count. */
/* kSquared: the candidates are squared time steps (the FP32-state form);
 * the minimum's square root is taken once, here. */
template <int NT, bool kSquared = false>
__device__ void finish_timestep(double local_min, int nan_count, FkReduce r,
                                double max_dt, int first_step) {
  if (nan_count) atomicAdd(r.status, nan_count);
  const double bmin = block_min<NT>(local_min);
  if (threadIdx.x == 0) r.partials[blockIdx.x] = bmin;
  if (!last_block_arrives(r.ticket)) return;
  double g = kSquared ? max_dt * max_dt : max_dt;
  for (int b = threadIdx.x; b < (int)gridDim.x; b += NT)
    g = lesser(g, r.partials[b]);
  g = block_min<NT>(g);
  if (kSquared) g = __builtin_sqrt(g);
  if (threadIdx.x == 0) {
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

/* Per-block compensated sums -> last block -> r.sums[0..1]. */
template <int NT>
__device__ void finish_sums(CSum ms, CSum es, FkReduce r) {
  ms = block_csum<NT>(ms);
  es = block_csum<NT>(es);
  if (threadIdx.x == 0) {
    double* p = r.partials + 4 * blockIdx.x;
    p[0] = ms.s;
    p[1] = ms.c;
    p[2] = es.s;
    p[3] = es.c;
  }
  if (!last_block_arrives(r.ticket)) return;
  CSum gm{0.0, 0.0}, ge{0.0, 0.0};
  for (int b = threadIdx.x; b < (int)gridDim.x; b += NT) {
    const double* p = r.partials + 4 * b;
    gm.merge(CSum{p[0], p[1]});
    ge.merge(CSum{p[2], p[3]});
  }
  gm = block_csum<NT>(gm);
  ge = block_csum<NT>(ge);
  if (threadIdx.x == 0) {
    r.sums[0] = gm.s + gm.c;
    r.sums[1] = ge.s + ge.c;
    *r.ticket = 0u;
  }
}

}  // namespace flowhip
