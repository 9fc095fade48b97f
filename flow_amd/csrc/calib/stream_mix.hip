/*
 * stream_mix.hip -- the streaming floor of each march kernel's array mix.
 *
 * A march kernel's compulsory traffic is NR whole arrays read and NW whole
 * arrays written. This program times kernels that do only that: each
 * element i of the NW outputs is a function of element i of the NR inputs,
 * 16-byte accesses, a grid-stride loop, no stencil and no halo. The best
 * time over the launch shapes and store forms tried is the floor the HBM
 * gives that read/write mix on this GPU, which the whole-array plain copy
 * (1 read, 1 write) overstates for write-heavy mixes.
 *
 *   flowhip_stream_mix [n]      arrays of n^2 elements (default 8192)
 *
 * Prints, per mix and element size, the best ms per launch, the bytes moved
 * and the rate, and the launch shape that gave it. Needs (NRmax + NWmax) n^2
 * x 8 bytes of device memory (11 arrays, 5.9 GB at n = 8192).
 */
#include <hip/hip_runtime.h>

#include <cstdio>
#include <cstdlib>
#include <vector>

#define CK(x)                                                                     \
  do {                                                                            \
    hipError_t e_ = (x);                                                          \
    if (e_ != hipSuccess) {                                                       \
      std::fprintf(stderr, "HIP error %s at line %d\n", hipGetErrorString(e_),    \
                   __LINE__);                                                     \
      std::exit(1);                                                               \
    }                                                                             \
  } while (0)

typedef double V2 __attribute__((ext_vector_type(2)));

/* out[k][i] = sum of in[.][i] + k for every 16-byte element i < n16 */
template <int NR, int NW, bool NT>
__global__ __launch_bounds__(256) void k_mix(const V2* const* __restrict__ in,
                                             V2* const* __restrict__ out, long n16) {
  const long stride = (long)gridDim.x * blockDim.x;
  for (long i = (long)blockIdx.x * blockDim.x + threadIdx.x; i < n16; i += stride) {
    V2 s = {0.0, 0.0};
#pragma unroll
    for (int k = 0; k < NR; ++k) s += in[k][i];
#pragma unroll
    for (int k = 0; k < NW; ++k) {
      const V2 v = s + (double)k;
      if (NT)
        __builtin_nontemporal_store(v, &out[k][i]);
      else
        out[k][i] = v;
    }
  }
}

struct Best {
  float ms;
  int grid;
  bool nt;
};

template <int NR, int NW, bool NT>
static void time_one(const V2** din, V2** dout, long n16, Best& best) {
  hipEvent_t a, b;
  CK(hipEventCreate(&a));
  CK(hipEventCreate(&b));
  for (int grid : {512, 1024, 2048, 4096, 8192}) {
    for (int w = 0; w < 3; ++w)
      hipLaunchKernelGGL((k_mix<NR, NW, NT>), dim3(grid), dim3(256), 0, 0, din, dout, n16);
    CK(hipGetLastError());
    CK(hipDeviceSynchronize());
    const int reps = 20;
    CK(hipEventRecord(a));
    for (int r = 0; r < reps; ++r)
      hipLaunchKernelGGL((k_mix<NR, NW, NT>), dim3(grid), dim3(256), 0, 0, din, dout, n16);
    CK(hipEventRecord(b));
    CK(hipEventSynchronize(b));
    float ms = 0.f;
    CK(hipEventElapsedTime(&ms, a, b));
    ms /= reps;
    if (ms < best.ms) best = Best{ms, grid, NT};
  }
  CK(hipEventDestroy(a));
  CK(hipEventDestroy(b));
}

template <int NR, int NW>
static void run(const char* name, long array_bytes, const std::vector<void*>& bufs) {
  const long n16 = array_bytes / 16;
  const V2** din;
  V2** dout;
  CK(hipMalloc(&din, NR * sizeof(void*)));
  CK(hipMalloc(&dout, NW * sizeof(void*)));
  std::vector<const V2*> hi(NR);
  std::vector<V2*> ho(NW);
  for (int k = 0; k < NR; ++k) hi[k] = (const V2*)bufs[k];
  for (int k = 0; k < NW; ++k) ho[k] = (V2*)bufs[NR + k];
  CK(hipMemcpy(din, hi.data(), NR * sizeof(void*), hipMemcpyHostToDevice));
  CK(hipMemcpy(dout, ho.data(), NW * sizeof(void*), hipMemcpyHostToDevice));
  Best best{1e30f, 0, false};
  time_one<NR, NW, false>(din, dout, n16, best);
  time_one<NR, NW, true>(din, dout, n16, best);
  const double gb = (double)(NR + NW) * array_bytes / 1e9;
  std::printf("%-22s %2dr %2dw  %8.4f ms  %6.2f GB  %6.3f TB/s  (grid %d x 256, %s stores)\n",
              name, NR, NW, best.ms, gb, gb / best.ms, best.grid,
              best.nt ? "non-temporal" : "plain");
  std::fflush(stdout);
  CK(hipFree(din));
  CK(hipFree(dout));
}

int main(int argc, char** argv) {
  const long n = argc > 1 ? std::atol(argv[1]) : 8192;
  if (n < 64 || n > 16384 || n % 64) {
    std::fprintf(stderr, "n must be a multiple of 64 in [64, 16384]\n");
    return 2;
  }
  const long cells = n * n;
  std::vector<void*> bufs(11);
  for (auto& p : bufs) {
    CK(hipMalloc(&p, cells * 8));
    CK(hipMemset(p, 0, cells * 8));
  }
  for (int es : {8, 4}) {
    std::printf("arrays of %ld^2 x %d B\n", n, es);
    const long ab = cells * es;
    run<1, 1>("copy", ab, bufs);
    run<4, 6>("lagrangian", ab, bufs);
    run<4, 4>("mass/energy", ab, bufs);
    run<7, 4>("momentum", ab, bufs);
  }
  for (auto p : bufs) CK(hipFree(p));
  return 0;
}
