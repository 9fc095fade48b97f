/*
 * reduce_common.h -- block-level reductions and the single-pass epilogues
 * (the last block to arrive combines what the others left) shared by the
 * reduction kernels, the fused kernels and the march kernels: the CFL
 * minimum, compensated per-block totals, and exact fixed-point totals.
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
 * omp3/flow.c:103-105 and the undefined-candidate count. With r.cfl set the
 * candidates were also stored for k_timestep_serial, which then does the
 * update and publishes the count; this only leaves the minimum over the
 * defined candidates in r.sums[2]. */
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
    if (!r.cfl) {
      const double dt_h = r.dtp[1];
      const double final_min = lesser(g, kCM * dt_h);
      const double dt = 0.5 * (kCT * final_min + dt_h);
      r.dtp[0] = dt;
      r.dtp[1] = first_step ? dt : kCT * final_min;
      r.status[1] = atomicExch(r.status, 0);
    }
    r.sums[2] = g;
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

/* Exact accumulator for totals: a sum of doubles, each first truncated
 * toward zero to a multiple of 2^-64, held as a 192-bit two's complement
 * integer in units of 2^-64 (w0 the low word). Integer addition is
 * associative, so the sum is the same bits whatever the order or grouping
 * of the additions. Values of magnitude below 2^127 are taken exactly down
 * to 2^-64; a larger or non-finite value marks the sum bad. */
struct FixSum {
  unsigned long long w0, w1, w2;
  unsigned int bad;

  __host__ __device__ void add(double v) {
    const unsigned long long u = __builtin_bit_cast(unsigned long long, v);
    const unsigned int E = (unsigned int)(u >> 52) & 0x7ffu;
    if (E >= 1023u + 127u) { /* |v| >= 2^127, inf or NaN */
      bad = 1u;
      return;
    }
    const unsigned long long mant = u & 0xfffffffffffffull;
    const unsigned long long M = E ? (mant | (1ull << 52)) : mant;
    /* v = M 2^(s - 64) */
    const int s = (E ? (int)E : 1) - 1075 + 64;
    unsigned long long a0 = 0, a1 = 0, a2 = 0;
    if (s < 0) {
      a0 = s > -64 ? (M >> -s) : 0ull;
    } else if (s < 64) {
      a0 = M << s;
      a1 = s ? (M >> (64 - s)) : 0ull;
    } else if (s < 128) {
      const int t = s - 64;
      a1 = M << t;
      a2 = t ? (M >> (64 - t)) : 0ull;
    } else {
      a2 = M << (s - 128); /* s <= 138: below 2^63 */
    }
    if (u >> 63) {
      const unsigned long long b0 = w0 < a0;
      w0 -= a0;
      const unsigned long long t1 = w1 - a1, b1 = (w1 < a1) + (t1 < b0);
      w1 = t1 - b0;
      w2 -= a2 + b1;
    } else {
      w0 += a0;
      const unsigned long long c0 = w0 < a0;
      const unsigned long long t1 = w1 + a1, c1a = t1 < a1;
      w1 = t1 + c0;
      w2 += a2 + c1a + (w1 < c0);
    }
  }
};

__device__ __forceinline__ unsigned long long wave_sum_u64(unsigned long long x) {
#pragma unroll
  for (int off = kWave / 2; off > 0; off >>= 1) x += __shfl_xor(x, off, kWave);
  return x;
}

/* Accumulator word k of total q lives at r.totals[(8 q + k) * kTotalsStride],
 * one per 128-byte line, so the blocks' atomics on different words do not
 * queue on one line. */
constexpr int kTotalsStride = 16;
static_assert(16 * kTotalsStride <= FK_TOTALS_WORDS, "totals words");

/* The double nearest (to within an ulp; the same bits every time) to the
 * fixed-point total held in 6 limb counters t[0..5] (t[k] counts units of
 * 2^(32k - 64), any size below 2^63) and its bad count t[6]. */
__host__ __device__ inline double fixed_total(const unsigned long long* t) {
  if (t[6]) return __builtin_nan("");
  unsigned int d[6];
  unsigned long long c = 0;
  for (int k = 0; k < 6; ++k) {
    const unsigned long long x = t[k] + c;
    d[k] = (unsigned int)x;
    c = x >> 32;
  }
  const bool neg = (d[5] >> 31) != 0;
  if (neg) { /* two's complement negation of the 192-bit value */
    unsigned long long carry = 1;
    for (int k = 0; k < 6; ++k) {
      const unsigned long long x = (unsigned long long)(~d[k]) + carry;
      d[k] = (unsigned int)x;
      carry = x >> 32;
    }
  }
  /* exact parts, largest first, into a compensated sum */
  CSum acc{0.0, 0.0};
  for (int k = 5; k >= 0; --k) acc.add(__builtin_ldexp((double)d[k], 32 * k - 64));
  const double r = acc.s + acc.c;
  return neg ? -r : r;
}

/* Per-lane exact totals (rho in fm, rho * e in fe) -> r.totals -> last
 * block -> r.sums[0..1]. The block sums its lanes' 32-bit limbs (integer
 * sums, in any order) and its first 14 threads add them to the counters
 * with 64-bit atomics; the last block reads and clears the counters. The
 * totals are the same bits whichever waves and blocks added which values.
 * (One atomic per wave and limb, on counters sharing a line, cost the
 * mass/energy march 2-3%.) */
template <int NT>
__device__ void finish_fixed_sums(const FixSum& fm, const FixSum& fe, FkReduce r) {
  __shared__ unsigned long long part[NT / kWave][14];
  const int lane = threadIdx.x % kWave, w = threadIdx.x / kWave;
  const FixSum* f[2] = {&fm, &fe};
#pragma unroll
  for (int q = 0; q < 2; ++q) {
    const unsigned long long limb[7] = {
        f[q]->w0 & 0xffffffffull, f[q]->w0 >> 32, f[q]->w1 & 0xffffffffull,
        f[q]->w1 >> 32,           f[q]->w2 & 0xffffffffull, f[q]->w2 >> 32,
        (unsigned long long)f[q]->bad};
#pragma unroll
    for (int k = 0; k < 7; ++k) {
      const unsigned long long v = wave_sum_u64(limb[k]);
      if (lane == 0) part[w][7 * q + k] = v;
    }
  }
  __syncthreads();
  if (threadIdx.x < 14) {
    unsigned long long v = 0;
    for (int b = 0; b < NT / kWave; ++b) v += part[b][threadIdx.x];
    const int q = threadIdx.x / 7, k = threadIdx.x % 7;
    if (v) atomicAdd(&r.totals[(8 * q + k) * kTotalsStride], v);
  }
  if (!last_block_arrives(r.ticket)) return;
  if (threadIdx.x == 0) {
    for (int q = 0; q < 2; ++q) {
      unsigned long long t[7];
      for (int k = 0; k < 7; ++k)
        t[k] = atomicExch(&r.totals[(8 * q + k) * kTotalsStride], 0ull);
      r.sums[q] = fixed_total(t);
    }
    *r.ticket = 0u;
  }
}

}  // namespace flowhip
