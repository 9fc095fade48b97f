/*
 * noalloc.hip -- the device scratch of flowhip_solve_hydro_2d_static: one
 * file-scope device array holding every buffer the march step's kernels use
 * beyond the caller's arrays (the reduction partials and ticket, the dt
 * pair and totals, the status counts, the work-queue counters, the
 * fixed-point accumulators and the reciprocal edge spacings), so that the
 * entry point allocates nothing. Each GPU holds its own copy, at its own
 * address, until the process ends.
 */
#include "reduce_common.h"

namespace flowhip {
namespace {

/* Byte layout of the scratch, every part 256-byte aligned. The counter
 * words (ticket, status, queue, totals) are contiguous so one launch clears
 * them; the dt pair and the totals are contiguous so one copy reads both. */
constexpr size_t align256(size_t b) { return (b + 255) / 256 * 256; }
constexpr size_t kOffPartials = 0;
constexpr size_t kPartialsBytes = align256(sizeof(double) * 4 * kMaxReduceBlocks);
constexpr size_t kOffTicket = kOffPartials + kPartialsBytes;
constexpr size_t kOffStatus = kOffTicket + 256;
constexpr size_t kOffQueue = kOffStatus + 256;
constexpr size_t kOffTotals = kOffQueue + align256(sizeof(unsigned int) * FK_MARCH_QUEUE_WORDS);
constexpr size_t kOffPair = kOffTotals + align256(sizeof(unsigned long long) * FK_TOTALS_WORDS);
constexpr size_t kCounterWords = (kOffPair - kOffTicket) / sizeof(unsigned int);
/* the largest mesh the library indexes: FLOWHIP_MAX_CELLS_SIDE interior
 * cells a side, L + 1 = side + 5 edge spacings */
constexpr int kMaxSide = 46335;
constexpr size_t kInvBytes = align256(sizeof(double) * (kMaxSide + 5));
constexpr size_t kOffInvX = kOffPair + 256;
constexpr size_t kOffInvY = kOffInvX + kInvBytes;
constexpr size_t kScratchBytes = kOffInvY + kInvBytes;  /* 1008896 B, < 1 MiB */
static_assert(kScratchBytes < (1u << 20), "scratch stays below 1 MiB");

constexpr int kMaxDevices = 64;

__global__ void k_static_begin(unsigned int* __restrict__ words, int nwords,
                               double* __restrict__ pair, double dt, double dt_h) {
  for (int k = blockIdx.x * blockDim.x + threadIdx.x; k < nwords;
       k += gridDim.x * blockDim.x)
    words[k] = 0u;
  if (blockIdx.x == 0 && threadIdx.x == 0) {
    pair[0] = dt;
    pair[1] = dt_h;
  }
}

}  // namespace
}  // namespace flowhip

using namespace flowhip;

/* In host code this name stands for the array; the kernels get the
 * device address looked up below, never the name. */
static __device__ __align__(256) unsigned char scratch[kScratchBytes];

/* The device address of `scratch`, looked up once per device (with that
 * device current) and kept. */
static void* g_scratch[kMaxDevices];

extern "C" {

hipError_t fk_static_scratch(FkReduce* r, double** inv_edgedx, double** inv_edgedy,
                             int* max_side, size_t* bytes) {
  int dev = 0;
  hipError_t e = hipGetDevice(&dev);
  if (e != hipSuccess) return e;
  if (dev < 0 || dev >= kMaxDevices) return hipErrorInvalidDevice;
  if (!g_scratch[dev]) {
    void* ptr = nullptr;
    e = hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch));
    if (e != hipSuccess) return e;
    if (!ptr) return hipErrorInvalidSymbol;
    g_scratch[dev] = ptr;
  }
  unsigned char* base = static_cast<unsigned char*>(g_scratch[dev]);
  FkReduce q;
  q.partials = reinterpret_cast<double*>(base + kOffPartials);
  q.ticket = reinterpret_cast<unsigned int*>(base + kOffTicket);
  q.status = reinterpret_cast<int*>(base + kOffStatus);
  q.queue = reinterpret_cast<unsigned int*>(base + kOffQueue);
  q.totals = reinterpret_cast<unsigned long long*>(base + kOffTotals);
  q.dtp = reinterpret_cast<double*>(base + kOffPair);
  q.sums = q.dtp + 2; /* right after the dt pair: one copy reads both */
  q.max_blocks = kMaxReduceBlocks;
  q.cfl = nullptr;
  *r = q;
  *inv_edgedx = reinterpret_cast<double*>(base + kOffInvX);
  *inv_edgedy = reinterpret_cast<double*>(base + kOffInvY);
  *max_side = kMaxSide;
  if (bytes) *bytes = kScratchBytes;
  return hipSuccess;
}

/* Zero the counter words of r (ticket, status, queue, totals: the kernels
 * count on them starting at zero) and set the dt pair, in one launch. */
hipError_t fk_static_begin(FkReduce r, double dt, double dt_h, hipStream_t s) {
  hipLaunchKernelGGL(k_static_begin, dim3(4), dim3(256), 0, s, r.ticket,
                     (int)kCounterWords, r.dtp, dt, dt_h);
  return hipGetLastError();
}

}  // extern "C"
