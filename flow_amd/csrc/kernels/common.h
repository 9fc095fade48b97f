/*
 * common.h -- launch geometry, index helpers and wave64 reductions shared by
 * the hip kernel set's translation units.
 */
#pragma once

#include <hip/hip_runtime.h>

#include "../../../include/flow_hip_kernels.h"

namespace flowhip {

constexpr int kWave = 64;

/* A rectangular loop domain [r0, r1) x [c0, c1) of a 2-D array. */
struct Domain {
  int r0, r1, c0, c1;
  __host__ __device__ int rows() const { return r1 - r0; }
  __host__ __device__ int cols() const { return c1 - c0; }
};

/* Stagewise kernels: 64 x 4 threads, one point per thread. */
constexpr int kSwX = 64;
constexpr int kSwY = 4;

inline dim3 sw_grid(const Domain& d) {
  const int gx = (d.cols() + kSwX - 1) / kSwX;
  const int gy = (d.rows() + kSwY - 1) / kSwY;
  return dim3(gx > 0 ? gx : 1, gy > 0 ? gy : 1, 1);
}

/* Wave64 reductions through cross-lane shuffles. */
__device__ __forceinline__ double wave_min(double x) {
#pragma unroll
  for (int off = kWave / 2; off > 0; off >>= 1) {
    const double y = __shfl_xor(x, off, kWave);
    x = (y < x) ? y : x;
  }
  return x;
}

/* 1/x for a positive power of two x with a normal reciprocal, exactly: 2^-k
 * has biased exponent 2 * bias - (bias + k), so the reciprocal's bits are a
 * constant minus x's bits. */
__host__ __device__ __forceinline__ double recip_pow2(double x) {
  return __builtin_bit_cast(double, 0x7fe0000000000000ull -
                                        __builtin_bit_cast(unsigned long long, x));
}
__host__ __device__ __forceinline__ float recip_pow2(float x) {
  return __builtin_bit_cast(float, 0x7f000000u - __builtin_bit_cast(unsigned int, x));
}

/* Error-free transformation a + b = s + err (Knuth TwoSum). */
__host__ __device__ __forceinline__ void two_sum(double a, double b, double& s,
                                                 double& err) {
  s = a + b;
  const double bp = s - a;
  const double ap = s - bp;
  err = (a - ap) + (b - bp);
}

/* Compensated accumulator (sum, compensation), merged with TwoSum. */
struct CSum {
  double s, c;
  __host__ __device__ void add(double x) {
    double t, e;
    two_sum(s, x, t, e);
    s = t;
    c += e;
  }
  __host__ __device__ void merge(const CSum& o) {
    double t, e;
    two_sum(s, o.s, t, e);
    s = t;
    c += e + o.c;
  }
};

__device__ __forceinline__ CSum wave_csum(CSum v) {
#pragma unroll
  for (int off = kWave / 2; off > 0; off >>= 1) {
    CSum o;
    o.s = __shfl_xor(v.s, off, kWave);
    o.c = __shfl_xor(v.c, off, kWave);
    v.merge(o);
  }
  return v;
}

}  // namespace flowhip
