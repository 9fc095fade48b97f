/*
 * ensemble.hip -- the batching form: many independent meshes of uneven
 * sizes advanced together, one launch per stage for the whole ensemble.
 *
 * Every instance's interior is cut into the same row-march items (strips
 * of kEnsW columns, chunks of H rows), and one work list holds the items
 * of all instances, in the order the host (or its autotuner) chose. The
 * persistent kernels below walk that list: each wave takes an item, reads
 * its instance's descriptor (geometry, arrays, dt pair) with scalar loads,
 * and runs the same per-item body as the single-mesh march kernels, so a
 * batched instance computes bit for bit what it computes alone.
 *
 * Per-instance reductions go through per-item partials, indexed by the
 * item's position inside its instance, and small per-instance kernels that
 * combine them in item order: the CFL minimum and dt update of
 * omp3/flow.c:77-105 (one dt pair per instance) and the compensated mass
 * and energy totals of omp3/flow.c:973-995. Per-step dt pairs and totals
 * are appended to device history buffers at a device-side step counter,
 * so a step can be captured once in a hipGraph and replayed.
 */
#include "march.h"

namespace flowhip {

/* Strip width of every ensemble kernel: the narrowest any march kernel
 * allows (the Lagrangian kernel's), so one work list serves all stages.
 * Momentum runs as one fused march or as the x then y marches, a schedule
 * choice of the host (and its autotuner). */
constexpr int kEnsW = kWave - 6;

template <typename T>
__device__ __forceinline__ MGeo<T> ens_geo(const FkEnsInst& in) {
  MGeo<T> g;
  g.nx = in.nx;
  g.ny = in.ny;
  g.edx = (const T*)in.geo[0];
  g.edy = (const T*)in.geo[1];
  g.cdx = (const T*)in.geo[2];
  g.cdy = (const T*)in.geo[3];
  g.iedx = (const T*)in.geo[4];
  g.iedy = (const T*)in.geo[5];
  return g;
}

template <typename T>
__device__ __forceinline__ T* arr(const FkEnsInst& in, int k) {
  return (T*)in.a[k];
}

/* solve_hydro_2d's array order */
enum { kP, kRho, kRhoOld, kE, kU, kV, kMomX, kMomY, kQxx, kQyy, kFx, kFy,
       kUFx, kUFy, kVFx, kVFy };

__device__ __forceinline__ MarchList ens_list(const FkEnsInst& in, int H) {
  MarchList ml;
  ml.W = kEnsW;
  ml.H = H;
  ml.nstrips = in.nstrips;
  ml.nchunks = in.nchunks;
  ml.nitems = in.nitems;
  /* uniform chunks of H rows */
  ml.G = 1;
  ml.cpb = in.nchunks;
  ml.H2 = H;
  ml.k = 1;
  ml.nsmall = 0;
  ml.R = in.ny - 2 * kPad;
  return ml;
}

__device__ __forceinline__ int wave_sum_int(int x) {
#pragma unroll
  for (int off = kWave / 2; off > 0; off >>= 1) x += __shfl_xor(x, off, kWave);
  return x;
}

/* ---- the four stages ---- */

template <typename T, bool P2, bool UNI>
__global__ __launch_bounds__(kMNT, FLOWHIP_MARCH_WAVES) void k_ens_lagrangian(
    const FkEnsInst* __restrict__ insts, const int2* __restrict__ work, int nwork,
    unsigned int* __restrict__ queue,
    int H, FkEnsPartials part) {
  const int lane = threadIdx.x % kWave;
  MarchQueue<true> mq(queue, nwork);
  for (int w = mq.next(); w < mq.end; w = mq.next()) {
    const int2 wi = work[w];
    const FkEnsInst& in = insts[wi.x];
    const MLagIO<T> io{arr<T>(in, kRho),   arr<T>(in, kE),   arr<T>(in, kMomX),
                       arr<T>(in, kMomY),  arr<T>(in, kQxx), arr<T>(in, kQyy),
                       arr<T>(in, kUFx),   arr<T>(in, kUFy), arr<T>(in, kP),
                       arr<T>(in, kRhoOld)};
    const T dt = T(in.dtp[0]), dt_h = T(in.dtp[1]);
    /* FP32 state: squared candidates (cell_timestep_sq) */
    double local_min = std::is_same<T, float>::value ? in.max_dt * in.max_dt : in.max_dt;
    int nan_count = 0;
    const MGeo<T> g = ens_geo<T>(in);
    const MarchList ml = ens_list(in, H);
    /* interior items take the straight-line form, as on a single mesh; the
     * uniform-geometry form in FP64 only (0.3% faster there, 1% slower in
     * FP32, where the per-row reciprocals are cheap) */
    constexpr bool kUni = UNI && sizeof(T) == 8;
    if (!item_edge(ml, g.nx, g.ny, wi.y))
      lagrangian_item<T, P2, 0, false, kUni>(g, io, dt, dt_h, ml, wi.y, lane, local_min,
                                             nan_count);
    else
      lagrangian_item<T, P2, -1, true, kUni>(g, io, dt, dt_h, ml, wi.y, lane, local_min,
                                             nan_count);
    local_min = wave_min(local_min);
    nan_count = wave_sum_int(nan_count);
    if (lane == 0) {
      part.min[in.item0 + wi.y] = local_min;
      part.nan[in.item0 + wi.y] = nan_count;
    }
  }
  mq.finish();
}

template <typename T, bool P2, bool UNI>
__global__ __launch_bounds__(kMNT, FLOWHIP_MARCH_WAVES) void k_ens_mass(
    const FkEnsInst* __restrict__ insts, const int2* __restrict__ work, int nwork,
    unsigned int* __restrict__ queue,
    int H, int tt, FkEnsPartials part) {
  const int lane = threadIdx.x % kWave;
  MarchQueue<true> mq(queue, nwork);
  for (int w = mq.next(); w < mq.end; w = mq.next()) {
    const int2 wi = work[w];
    const FkEnsInst& in = insts[wi.x];
    const MMassIO<T> io{arr<T>(in, kUFx), arr<T>(in, kUFy), arr<T>(in, kRhoOld),
                        arr<T>(in, kP),   arr<T>(in, kRho), arr<T>(in, kE),
                        arr<T>(in, kFx),  arr<T>(in, kFy)};
    const T dt = T(in.dtp[0]), dt_h = T(in.dtp[1]);
    CSum ms{0.0, 0.0}, es{0.0, 0.0};
    const MGeo<T> g = ens_geo<T>(in);
    const MarchList ml = ens_list(in, H);
    const bool edge = item_edge(ml, g.nx, g.ny, wi.y);
    if (tt % 2 == 0) {
      if (edge)
        mass_xy_item<T, P2, true, UNI>(g, io, dt, dt_h, ml, wi.y, lane, ms, es);
      else
        mass_xy_item<T, P2, false, UNI>(g, io, dt, dt_h, ml, wi.y, lane, ms, es);
    } else {
      if (edge)
        mass_yx_item<T, P2, true, UNI>(g, io, dt, dt_h, ml, wi.y, lane, ms, es);
      else
        mass_yx_item<T, P2, false, UNI>(g, io, dt, dt_h, ml, wi.y, lane, ms, es);
    }
    ms = wave_csum(ms);
    es = wave_csum(es);
    if (lane == 0) {
      double* p = part.sums + 4 * (in.item0 + wi.y);
      p[0] = ms.s;
      p[1] = ms.c;
      p[2] = es.s;
      p[3] = es.c;
    }
  }
  mq.finish();
}

/* ydir 0: x momentum (u0 = uF_x), 1: y momentum (io.u = the u just written) */
template <typename T, bool P2>
__global__ __launch_bounds__(kMNT, FLOWHIP_MARCH_WAVES) void k_ens_momentum(
    const FkEnsInst* __restrict__ insts, const int2* __restrict__ work, int nwork,
    unsigned int* __restrict__ queue,
    int H, int tt, int ydir) {
  const int lane = threadIdx.x % kWave;
  MarchQueue<true> mq(queue, nwork);
  for (int w = mq.next(); w < mq.end; w = mq.next()) {
    const int2 wi = work[w];
    const FkEnsInst& in = insts[wi.x];
    const T dt = T(in.dtp[0]), dt_h = T(in.dtp[1]);
    const MGeo<T> g = ens_geo<T>(in);
    const MarchList ml = ens_list(in, H);
    if (!ydir) {
      const MMomIO<T> io{arr<T>(in, kUFx), arr<T>(in, kUFy), arr<T>(in, kFx),
                         arr<T>(in, kFy),  arr<T>(in, kRho), arr<T>(in, kQxx),
                         arr<T>(in, kMomX), arr<T>(in, kU)};
      if (tt % 2)
        momx_xy_item<T, P2>(g, io, dt, dt_h, ml, wi.y, lane);
      else
        momx_yx_item<T, P2>(g, io, dt, dt_h, ml, wi.y, lane);
    } else {
      const MMomIO<T> io{arr<T>(in, kU),   arr<T>(in, kUFy), arr<T>(in, kFx),
                         arr<T>(in, kFy),  arr<T>(in, kRho), arr<T>(in, kQyy),
                         arr<T>(in, kMomY), arr<T>(in, kV)};
      if (tt % 2)
        momy_xy_item<T, P2>(g, io, dt, dt_h, ml, wi.y, lane);
      else
        momy_yx_item<T, P2>(g, io, dt, dt_h, ml, wi.y, lane);
    }
  }
  mq.finish();
}

/* Both momentum directions in one march (momxy_*_item). */
template <typename T, bool P2, bool UNI>
__global__ __launch_bounds__(kMNT, FLOWHIP_MARCH_WAVES) void k_ens_momentum_xy(
    const FkEnsInst* __restrict__ insts, const int2* __restrict__ work, int nwork,
    unsigned int* __restrict__ queue,
    int H, int tt) {
  const int lane = threadIdx.x % kWave;
  /* claimed per block, as the single-mesh fused momentum march */
  const int wave = threadIdx.x / kWave;
  MarchBlockQueue mq(queue, nwork);
  for (int base = mq.next(); base < mq.end; base = mq.next()) {
    const int w = __builtin_amdgcn_readfirstlane(base + wave);
    if (w >= mq.end) continue;
    const int2 wi = work[w];
    const FkEnsInst& in = insts[wi.x];
    const T dt = T(in.dtp[0]), dt_h = T(in.dtp[1]);
    const MMom2IO<T> io{arr<T>(in, kUFx), arr<T>(in, kUFy), arr<T>(in, kFx),
                        arr<T>(in, kFy),  arr<T>(in, kRho), arr<T>(in, kQxx),
                        arr<T>(in, kQyy), arr<T>(in, kMomX), arr<T>(in, kMomY),
                        arr<T>(in, kU),   arr<T>(in, kV)};
    const MGeo<T> g = ens_geo<T>(in);
    const MarchList ml = ens_list(in, H);
    const bool edge = item_edge(ml, g.nx, g.ny, wi.y);
    if (tt % 2) {
      if (edge)
        momxy_xy_item<T, P2, -1, true, UNI>(g, io, dt, dt_h, ml, wi.y, lane);
      else
        momxy_xy_item<T, P2, 0, false, UNI>(g, io, dt, dt_h, ml, wi.y, lane);
    } else {
      if (edge)
        momxy_yx_item<T, P2, -1, true, UNI>(g, io, dt, dt_h, ml, wi.y, lane);
      else
        momxy_yx_item<T, P2, 0, false, UNI>(g, io, dt, dt_h, ml, wi.y, lane);
    }
  }
  mq.finish();
}

/* CFL candidates of one item's cells (the pre-loop set_timestep). */
template <typename T>
__global__ __launch_bounds__(kMNT) void k_ens_timestep(
    const FkEnsInst* __restrict__ insts, const int2* __restrict__ work, int nwork,
    unsigned int* __restrict__ queue,
    int H, FkEnsPartials part) {
  const int lane = threadIdx.x % kWave;
  MarchQueue<true> mq(queue, nwork);
  for (int w = mq.next(); w < mq.end; w = mq.next()) {
    const int2 wi = work[w];
    const FkEnsInst& in = insts[wi.x];
    const MGeo<T> g = ens_geo<T>(in);
    const int nx = in.nx, ny = in.ny;
    const int strip = wi.y % in.nstrips, chunk = wi.y / in.nstrips;
    const int x0 = kPad + strip * kEnsW, x1 = min(x0 + kEnsW, nx - kPad);
    const int y0 = kPad + chunk * H, y1 = min(y0 + H, ny - kPad);
    const int j = x0 + lane;
    const T *e = arr<T>(in, kE), *qxx = arr<T>(in, kQxx), *qyy = arr<T>(in, kQyy),
            *rho = arr<T>(in, kRho);
    double local_min = in.max_dt;
    int nan_count = 0;
    if (j < x1) {
      for (int i = y0; i < y1; ++i) {
        const int c = i * nx + j;
        const double cand =
            cell_timestep(double(e[c]), double(qxx[c]), double(qyy[c]),
                          double(rho[c]), double(g.cdx[j]), double(g.cdy[i]));
        if (cand == cand)
          local_min = lesser(local_min, cand);
        else
          ++nan_count;
      }
    }
    local_min = wave_min(local_min);
    nan_count = wave_sum_int(nan_count);
    if (lane == 0) {
      part.min[in.item0 + wi.y] = local_min;
      part.nan[in.item0 + wi.y] = nan_count;
    }
  }
  mq.finish();
}

/* ---- per-instance epilogues: one wave per instance ---- */

/* dt pair of omp3/flow.c:103-105 from the instance's item minima; history
 * row *counter when hist != null. */
__global__ __launch_bounds__(kWave) void k_ens_dt(const FkEnsInst* __restrict__ insts,
                                                  int ninst, FkEnsPartials part,
                                                  int first_step, int squared,
                                                  FkEnsHistory hist) {
  const int inst = blockIdx.x;
  if (inst >= ninst) return;
  const FkEnsInst& in = insts[inst];
  const int lane = threadIdx.x;
  double g = squared ? in.max_dt * in.max_dt : in.max_dt;
  int bad = 0;
  for (int k = lane; k < in.nitems; k += kWave) {
    g = lesser(g, part.min[in.item0 + k]);
    bad += part.nan[in.item0 + k];
  }
  g = wave_min(g);
  if (squared) g = __builtin_sqrt(g);
  bad = wave_sum_int(bad);
  if (lane == 0) {
    const double dt_h = in.dtp[1];
    const double final_min = lesser(g, kCM * dt_h);
    const double dt = 0.5 * (kCT * final_min + dt_h);
    in.dtp[0] = dt;
    in.dtp[1] = first_step ? dt : kCT * final_min;
    if (hist.dt) {
      const int row = *hist.counter;
      hist.dt[2 * ((long)row * ninst + inst)] = in.dtp[0];
      hist.dt[2 * ((long)row * ninst + inst) + 1] = in.dtp[1];
      hist.nan[(long)row * ninst + inst] = bad;
    }
  }
}

/* Mass and energy totals from the instance's item partials, merged in item
 * order; history row *counter. */
__global__ __launch_bounds__(kWave) void k_ens_sums(const FkEnsInst* __restrict__ insts,
                                                    int ninst, FkEnsPartials part,
                                                    FkEnsHistory hist) {
  const int inst = blockIdx.x;
  if (inst >= ninst) return;
  const FkEnsInst& in = insts[inst];
  const int lane = threadIdx.x;
  CSum gm{0.0, 0.0}, ge{0.0, 0.0};
  for (int k = lane; k < in.nitems; k += kWave) {
    const double* p = part.sums + 4 * (in.item0 + k);
    gm.merge(CSum{p[0], p[1]});
    ge.merge(CSum{p[2], p[3]});
  }
  gm = wave_csum(gm);
  ge = wave_csum(ge);
  if (lane == 0) {
    const int row = *hist.counter;
    hist.sums[2 * ((long)row * ninst + inst)] = gm.s + gm.c;
    hist.sums[2 * ((long)row * ninst + inst) + 1] = ge.s + ge.c;
  }
}

__global__ void k_ens_advance(int* counter) {
  if (threadIdx.x == 0 && blockIdx.x == 0) *counter += 1;
}

__global__ void k_ens_set_counter(int* counter, int v) {
  if (threadIdx.x == 0 && blockIdx.x == 0) *counter = v;
}

/* ---- launch helpers ---- */

template <typename T, bool P2, bool UNI>
hipError_t ens_step_p(const FkEnsLaunch& L, int tt, hipStream_t s) {
  static int pc_lag = 0, pc_mass = 0, pc_mom = 0;
  const int nw = L.nwork;
  if (nw < 1 || L.ninst < 1 || !L.queue) return hipErrorInvalidValue;
  const int2* work = (const int2*)L.work;
  hipLaunchKernelGGL((k_ens_lagrangian<T, P2, UNI>),
                     dim3(march_grid(k_ens_lagrangian<T, P2, UNI>, pc_lag, nw, kMarchMaxWaves<T>)),
                     dim3(kMNT), 0, s, L.insts, work, nw, L.queue, L.chunk_rows, L.part);
  hipLaunchKernelGGL(k_ens_dt, dim3(L.ninst), dim3(kWave), 0, s, L.insts, L.ninst,
                     L.part, tt == 0 ? 1 : 0, std::is_same<T, float>::value ? 1 : 0,
                     L.hist);
  hipLaunchKernelGGL((k_ens_mass<T, P2, UNI>),
                     dim3(march_grid(k_ens_mass<T, P2, UNI>, pc_mass, nw, kMarchMaxWaves<T>)),
                     dim3(kMNT), 0, s, L.insts, work, nw, L.queue, L.chunk_rows, tt, L.part);
  hipLaunchKernelGGL(k_ens_sums, dim3(L.ninst), dim3(kWave), 0, s, L.insts, L.ninst,
                     L.part, L.hist);
  if (L.fused_momentum) {
    static int pc_momxy = 0;
    hipLaunchKernelGGL((k_ens_momentum_xy<T, P2, UNI>),
                       dim3(march_grid(k_ens_momentum_xy<T, P2, UNI>, pc_momxy, nw,
                                       kMarchMaxWaves<T>)),
                       dim3(kMNT), 0, s, L.insts, work, nw, L.queue, L.chunk_rows, tt);
  } else {
    const int gm = march_grid(k_ens_momentum<T, P2>, pc_mom, nw);
    hipLaunchKernelGGL((k_ens_momentum<T, P2>), dim3(gm), dim3(kMNT), 0, s, L.insts,
                       work, nw, L.queue, L.chunk_rows, tt, 0);
    hipLaunchKernelGGL((k_ens_momentum<T, P2>), dim3(gm), dim3(kMNT), 0, s, L.insts,
                       work, nw, L.queue, L.chunk_rows, tt, 1);
  }
  hipLaunchKernelGGL(k_ens_advance, dim3(1), dim3(1), 0, s, L.hist.counter);
  return hipGetLastError();
}

template <typename T>
hipError_t ens_step(const FkEnsLaunch& L, int tt, hipStream_t s) {
  /* every instance's spacings are uniform (width / n): the mass/energy and
   * momentum marches read their row geometry once per item unless
   * FLOWHIP_UNIFORM=0 (the same bits either way) */
  if (L.uniform)
    return L.pow2 ? ens_step_p<T, true, true>(L, tt, s) : ens_step_p<T, false, true>(L, tt, s);
  return L.pow2 ? ens_step_p<T, true, false>(L, tt, s) : ens_step_p<T, false, false>(L, tt, s);
}

template <typename T>
hipError_t ens_timestep(const FkEnsLaunch& L, hipStream_t s) {
  static int pc = 0;
  const int nw = L.nwork;
  if (nw < 1 || L.ninst < 1 || !L.queue) return hipErrorInvalidValue;
  FkEnsHistory none = L.hist;
  none.dt = nullptr;
  hipLaunchKernelGGL(k_ens_timestep<T>, dim3(march_grid(k_ens_timestep<T>, pc, nw)),
                     dim3(kMNT), 0, s, L.insts, (const int2*)L.work, nw, L.queue,
                     L.chunk_rows, L.part);
  hipLaunchKernelGGL(k_ens_dt, dim3(L.ninst), dim3(kWave), 0, s, L.insts, L.ninst,
                     L.part, 0, 0, none);
  return hipGetLastError();
}

}  // namespace flowhip

using namespace flowhip;

extern "C" {

int fk_ens_strip_width(void) { return kEnsW; }

hipError_t fk_ens_step(FkEnsLaunch L, int tt, int fp32, hipStream_t s) {
  if (L.chunk_rows < 4 || L.chunk_rows > kMaxChunkRows) return hipErrorInvalidValue;
  return fp32 ? ens_step<float>(L, tt, s) : ens_step<double>(L, tt, s);
}

hipError_t fk_ens_timestep(FkEnsLaunch L, int fp32, hipStream_t s) {
  if (L.chunk_rows < 4 || L.chunk_rows > kMaxChunkRows) return hipErrorInvalidValue;
  return fp32 ? ens_timestep<float>(L, s) : ens_timestep<double>(L, s);
}

hipError_t fk_ens_set_counter(int* counter, int value, hipStream_t s) {
  hipLaunchKernelGGL(k_ens_set_counter, dim3(1), dim3(1), 0, s, counter, value);
  return hipGetLastError();
}

}  // extern "C"
