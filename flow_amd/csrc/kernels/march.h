/*
 * march.h -- row-marching forms of the fused kernels: the per-item bodies
 * shared by the single-mesh kernels (march.hip) and the ensemble kernels
 * (ensemble.hip), and the single-mesh kernels themselves.
 *
 * One wave owns a strip of W columns and marches down a chunk of H rows.
 * Along x a wave covers the strip plus its stencil halo in its 64 lanes, and
 * x neighbours come from cross-lane shuffles; along y every lane keeps the
 * rows its column still needs in a register window, so nothing along y is
 * recomputed and no block barrier is taken. Input rows are loaded one step
 * ahead of use. Waves of a block are independent.
 *
 * The arithmetic is the same pointwise code (hydro_math.h) in the same
 * order, so results are bitwise those of the tile kernels and of omp3;
 * ghost rows and columns are the mirrored values handle_boundary_2d would
 * copy there.
 */
#pragma once

#include <type_traits>

#include "reduce_common.h"
#include "tile.h"

namespace flowhip {

template <typename T>
struct MGeo {
  int nx, ny;
  const T* __restrict__ edx;
  const T* __restrict__ edy;
  const T* __restrict__ cdx;
  const T* __restrict__ cdy;
  const T* __restrict__ iedx;
  const T* __restrict__ iedy;
};

template <typename T>
struct MMassIO {
  const T* __restrict__ u0;
  const T* __restrict__ v0;
  const T* __restrict__ rho0;
  const T* __restrict__ e1;
  T* __restrict__ rho;
  T* __restrict__ e;
  T* __restrict__ F_x;
  T* __restrict__ F_y;
};

constexpr int kMNT = 256;            /* threads per block: 4 independent waves */
/* Minimum resident waves per SIMD the register allocation must allow
 * (the second __launch_bounds__ argument); 1 leaves the compiler free. */
#ifndef FLOWHIP_MARCH_WAVES
#define FLOWHIP_MARCH_WAVES 1
#endif
constexpr int kMWaves = kMNT / kWave;
/* Resident blocks per CU (= waves per SIMD) of the fused march kernels:
 * three in FP64, five in FP32. Their registers would let some of them fit
 * more (FP64 momentum 111-119 VGPRs, Lagrangian 121, mass/energy 126-129:
 * four; FP32 momentum 66-70: seven), and each extra wave measured slower on
 * MI355X at 8192^2 (Lagrangian +2.5%, momentum +3% FP64 and +4% FP32, with
 * six FP32 waves +3%); with four FP32 waves the momentum march was 2.5%
 * slower. */
template <typename T>
constexpr int kMarchMaxWaves = sizeof(T) == 8 ? 3 : 5;
/* The FP32 Lagrangian march: four. In its uniform-mesh form (95 VGPRs) five
 * would fit, and made it 8% slower. */
template <typename T>
constexpr int kLagMaxWaves = sizeof(T) == 8 ? 3 : 4;
/* The marches with two columns per lane: in FP32 three (the Lagrangian's
 * 124 VGPRs and the mass/energy march's 107-109 allow four, 3-7% slower at
 * every chunk height measured; two waves for the Lagrangian: 3% slower);
 * in FP64 two, all that their 192 and 173-177 VGPRs allow. */
#ifndef FLOWHIP_LAG_PAIR_WAVES
#define FLOWHIP_LAG_PAIR_WAVES 3
#endif
#ifndef FLOWHIP_MASS_PAIR_WAVES
#define FLOWHIP_MASS_PAIR_WAVES 3
#endif
template <typename T>
constexpr int kLagPairWaves = sizeof(T) == 8 ? 2 : FLOWHIP_LAG_PAIR_WAVES;
template <typename T>
constexpr int kMassPairWaves = sizeof(T) == 8 ? 2 : FLOWHIP_MASS_PAIR_WAVES;

/* The work items of a march kernel: strips of W columns by chunks of rows.
 * The interior rows are cut into G row bands (one per XCD band of the
 * item list, MarchQueue), each into units of H2 rows. A band's chunks are
 * k units tall (H = k * H2 rows) except its last `nsmall`, one unit each:
 * the waves claim tall chunks first, which re-read few prologue rows per
 * row marched, and finish on short ones, which keeps the band's tail short.
 * Every band has cpb chunk slots (a shorter band leaves its last slot
 * empty); items run strip-fastest within a chunk, band by band. Every chunk
 * but the mesh's first starts at least H2 rows above the south wall. With
 * G = 1, k = 1 and nsmall = 0 the chunks are uniform, H2 rows each. */
struct MarchList {
  int W, H;           /* strip width (output columns), tall chunk height (rows) */
  int nstrips, nchunks, nitems;
  int G, cpb;         /* row bands, chunk slots per band */
  int H2, k, nsmall;  /* unit rows, units per tall chunk, short chunks per band */
  int R;              /* interior rows */
};

/* Columns and rows of one item; y0 == y1 for an empty slot. */
struct ItemBox {
  int x0, x1, y0, y1;
};

__device__ __forceinline__ ItemBox item_box(const MarchList& ml, int nx, int item) {
  const int strip = item % ml.nstrips, slot = item / ml.nstrips;
  const int g = slot / ml.cpb, c = slot - g * ml.cpb;
  const int r0 = kPad + g * ml.R / ml.G, r1 = kPad + (g + 1) * ml.R / ml.G;
  const int units = (r1 - r0 + ml.H2 - 1) / ml.H2;
  const int ub = max(units - ml.nsmall, 0);           /* units in tall chunks */
  const int nbig = (ub + ml.k - 1) / ml.k;
  ItemBox b;
  b.x0 = kPad + strip * ml.W;
  b.x1 = min(b.x0 + ml.W, nx - kPad);
  if (c < nbig) {
    b.y0 = r0 + c * ml.k * ml.H2;
    b.y1 = min(r0 + min((c + 1) * ml.k, ub) * ml.H2, r1);
  } else {
    const int u = ub + (c - nbig);
    b.y0 = min(r0 + u * ml.H2, r1);
    b.y1 = u < units ? min(b.y0 + ml.H2, r1) : b.y0;
  }
  return b;
}

/* A wave's work items under the static banded walk: first, first + stride,
 * ... below end. The item list is cut into B = min(kMarchBands, gridDim)
 * contiguous bands; band b = blockIdx % B is walked by the blocks with that
 * residue, kMWaves items per block at a time. Blocks reach the 8 XCDs
 * round-robin, so with 8 bands a band is one XCD's share: neighbouring
 * strips and chunks, which read the same halo cache lines and rows, meet in
 * the same L2. Every item is visited exactly once for any grid size. The
 * split momentum marches walk this way; the other marches claim the items
 * of the same bands dynamically (MarchQueue below).
 *
 * UNIFORM makes the item wave-uniform: the item, its bounds and the row
 * counter then live in scalar registers, row tests become scalar branches
 * and per-row geometry is read with a scalar lane index. That pays in the
 * Lagrangian and mass/energy marches; in the momentum marches the extra
 * scalar registers it costs made them slower (measured on MI355X), so those
 * keep the per-lane index. */
constexpr int kMarchBands = kXcds;

struct MarchWalk {
  int first, end, stride;
};

template <bool UNIFORM>
__device__ __forceinline__ MarchWalk march_walk(int nitems) {
  const int nb = gridDim.x, b = blockIdx.x;
  const int bands = nb < kMarchBands ? nb : kMarchBands;
  const int band = b % bands, local = b / bands;
  const int nb_band = (nb - band + bands - 1) / bands;
  MarchWalk w;
  const int first = (int)((long)band * nitems / bands) + local * kMWaves + threadIdx.x / kWave;
  w.first = UNIFORM ? __builtin_amdgcn_readfirstlane(first) : first;
  w.end = (int)((long)(band + 1) * nitems / bands);
  w.stride = nb_band * kMWaves;
  return w;
}

/* Dynamic form of the banded walk: the waves of band b claim its items one
 * at a time from band b's counter (atomicAdd by lane 0) as they finish the
 * previous one, so a wave that draws short or fast items takes more of them
 * and the band's waves finish together. Claiming the next item when
 * starting the current one, to hide the claim's round trip, was slower on
 * MI355X: a wave then holds an item it cannot start, and the band's tail
 * grows by up to an item. The last wave of the grid to run out (counter
 * kMarchBands) resets every counter, so they are zero between launches
 * without a memset (the launches stay capturable in a graph). Every wave
 * reaches finish(); no wave waits on another. UNIFORM as for march_walk.
 *
 * Counter k lives at queue[k * kQueueStride], one per 128-byte line: with
 * all of them in one line the claims of the eight bands serialise on it,
 * and the kernels ran 2-3% slower on MI355X. */
constexpr int kQueueStride = 32;
static_assert((kMarchBands + 1) * kQueueStride <= FK_MARCH_QUEUE_WORDS, "queue counters");

template <bool UNIFORM, int BANDS = kMarchBands>
struct MarchQueue {
  static_assert(BANDS >= 1 && BANDS <= kMarchBands, "bands");
  unsigned int* ctr;
  int lo, end, band;
  __device__ MarchQueue(unsigned int* queue, int nitems) : ctr(queue) {
    const int nb = gridDim.x;
    const int bands = nb < BANDS ? nb : BANDS;
    band = blockIdx.x % bands;
    lo = (int)((long)band * nitems / bands);
    end = (int)((long)(band + 1) * nitems / bands);
  }
  /* the next item of the band; end or beyond once the band is drained */
  __device__ int next() {
    unsigned int v = 0;
    if (threadIdx.x % kWave == 0) v = atomicAdd(&ctr[band * kQueueStride], 1u);
    if (UNIFORM) return lo + (int)__builtin_amdgcn_readfirstlane(v);
    return lo + (int)__shfl(v, 0, kWave);
  }
  __device__ void finish() {
    if (threadIdx.x % kWave == 0) {
      const unsigned int waves = gridDim.x * kMWaves;
      if (atomicAdd(&ctr[kMarchBands * kQueueStride], 1u) == waves - 1)
        for (int k = 0; k <= kMarchBands; ++k) atomicExch(&ctr[k * kQueueStride], 0u);
    }
  }
};

/* Block form of the queue, for the Lagrangian and fused momentum marches:
 * thread 0 claims kMWaves consecutive items of the band for the whole block
 * and wave w marches base + w, so a block's waves march adjacent strips of
 * one chunk row side by side, share the halo columns between them in cache
 * and read and write each row's neighbouring segments together (measured on
 * MI355X at 8192^2: the Lagrangian march 2.5-3% faster than with per-wave
 * claims; the mass/energy marches 1% slower, waiting at the barriers, so
 * they keep the per-wave queue). Every wave of the block takes part in
 * every claim, and the loop ends for the whole block at once:
 *
 *   MarchBlockQueue mq(queue, nitems);
 *   for (int base = mq.next(); base < mq.end; base = mq.next()) {
 *     const int item = base + wave;
 *     if (item < mq.end) ... march item ...
 *   }
 *   mq.finish();
 */
struct MarchBlockQueue : MarchQueue<true> {
  __device__ MarchBlockQueue(unsigned int* queue, int nitems)
      : MarchQueue<true>(queue, nitems) {}
  /* the block's next base item; the same in every wave */
  __device__ int next() {
    __shared__ int s_base;
    __syncthreads(); /* every wave has read the previous base */
    if (threadIdx.x == 0)
      s_base = lo + (int)atomicAdd(&ctr[band * kQueueStride], (unsigned int)kMWaves);
    __syncthreads();
    return __builtin_amdgcn_readfirstlane(s_base);
  }
};

/* Items of a mesh: H2-row units in tall chunks of k units, the last nsmall
 * chunks of each of G row bands one unit each (G is lowered until every
 * band holds at least 4 units; k = 1, nsmall = 0 gives uniform chunks). */
inline MarchList make_march(int nx, int ny, int W, int H2, int k = 1, int nsmall = 0,
                            int G = 1) {
  MarchList m;
  m.W = W;
  m.H2 = H2;
  m.k = k;
  m.nsmall = nsmall;
  m.H = k * H2;
  m.R = ny - 2 * kPad;
  m.G = G;
  while (m.G > 1 && m.R / m.G < 4 * H2) --m.G;
  const int rmax = (m.R + m.G - 1) / m.G;
  const int units = (rmax + H2 - 1) / H2;
  const int ub = units > nsmall ? units - nsmall : 0;
  m.cpb = (ub + k - 1) / k + (units < nsmall ? units : nsmall);
  m.nstrips = (nx - 2 * kPad + W - 1) / W;
  m.nchunks = m.G * m.cpb;
  m.nitems = m.nstrips * m.nchunks;
  return m;
}

/* Whole-wave shifts by one lane as DPP moves (wave_shr:1 / wave_shl:1):
 * VALU operations on registers, with none of the LDS round trip of a
 * ds_bpermute. Lane 0 of a shift up and lane 63 of a shift down read zero
 * (bound_ctrl), so a move needs no copy of its destination's old value; no
 * kernel reads those lanes' results. */
constexpr int kDppWaveShl1 = 0x130; /* lane i <- lane i+1 */
constexpr int kDppWaveShr1 = 0x138; /* lane i <- lane i-1 */

template <int CTRL>
__device__ __forceinline__ int dpp_move(int v) {
  return __builtin_amdgcn_mov_dpp(v, CTRL, 0xf, 0xf, true);
}
template <int CTRL>
__device__ __forceinline__ double dpp_move(double v) {
  const long long b = __double_as_longlong(v);
  const int lo = dpp_move<CTRL>((int)b);
  const int hi = dpp_move<CTRL>((int)(b >> 32));
  return __longlong_as_double(((long long)hi << 32) | (unsigned int)lo);
}
template <int CTRL>
__device__ __forceinline__ float dpp_move(float v) {
  return __int_as_float(dpp_move<CTRL>(__float_as_int(v)));
}

/* value of the lane d below (up) or above (dn), d = 1 or 2 */
template <typename T>
__device__ __forceinline__ T up(T v, int d) {
  const T v1 = dpp_move<kDppWaveShr1>(v);
  return d == 1 ? v1 : dpp_move<kDppWaveShr1>(v1);
}
template <typename T>
__device__ __forceinline__ T dn(T v, int d) {
  const T v1 = dpp_move<kDppWaveShl1>(v);
  return d == 1 ? v1 : dpp_move<kDppWaveShl1>(v1);
}


/* Whether a work item's strip lies within three columns of a wall, where
 * some of its lanes are ghost columns. */
__device__ __forceinline__ bool item_xedge(const MarchList& ml, int nx, int item) {
  const int x0 = kPad + (item % ml.nstrips) * ml.W, x1 = min(x0 + ml.W, nx - kPad);
  return (x0 - 3 < kPad) || (x1 + 3 > nx - kPad);
}

/* Whether an item's stencils reach a ghost row or column anywhere (the
 * x test above, or a chunk within three rows of a north/south wall); the
 * other items, all but the mesh's outer ring of items, run the
 * straight-line interior form of their kernel. */
__device__ __forceinline__ bool item_edge(const MarchList& ml, int nx, int ny, int item) {
  const ItemBox b = item_box(ml, nx, item);
  return item_xedge(ml, nx, item) || (b.y0 == kPad) || (b.y1 + 3 > ny - kPad);
}

/* ---- per-row geometry of a chunk held across the lanes ----
 * Lane k holds rows base+k and base+64+k; row(k) broadcasts one row's value
 * with v_readlane (k is uniform), so the row loop issues no memory load for
 * geometry. Chunks are at most kMaxChunkRows rows. */
constexpr int kMaxChunkRows = 120;

__device__ __forceinline__ double lane_value(double v, int k) {
  const long long b = __double_as_longlong(v);
  const int lo = __builtin_amdgcn_readlane((int)b, k);
  const int hi = __builtin_amdgcn_readlane((int)(b >> 32), k);
  return __longlong_as_double(((long long)hi << 32) | (unsigned int)lo);
}
__device__ __forceinline__ float lane_value(float v, int k) {
  return __int_as_float(__builtin_amdgcn_readlane(__float_as_int(v), k));
}

template <typename T>
struct RowGeo {
  T a, b;
  int base;
  __device__ RowGeo(const T* arr, int base_row, int last_row, int lane,
                    bool recip = false)
      : base(base_row) {
    /* rows outside [0, last_row] are never asked for; clamp the loads */
    a = arr[max(0, min(base_row + lane, last_row))];
    b = arr[max(0, min(base_row + kWave + lane, last_row))];
    if (recip) {
      a = T(1.0) / a;
      b = T(1.0) / b;
    }
  }
  /* value at row r (uniform, base <= r < base + 128) */
  __device__ __forceinline__ T operator()(int r) const {
    const int k = r - base;
    return k < kWave ? lane_value(a, k) : lane_value(b, k - kWave);
  }
};

/* Per-row geometry read straight from memory at a wave-uniform row: the
 * geometry arrays are not written while a march kernel runs, so they are
 * read through the constant address space and the compiler issues scalar
 * loads (no vector issue slots, no lane reads). Same values as RowGeo for
 * the same rows, clamped the same way. Used where the row is wave-uniform
 * (every fused march; the items are claimed per wave); measured on MI355X
 * at 8192^2, 1-2.5% off the Lagrangian and mass/energy marches. */
template <typename T>
struct RowScalar {
  const __attribute__((address_space(4))) T* p;
  int last;
  __device__ RowScalar(const T* arr, int, int last_row, int)
      : p((const __attribute__((address_space(4))) T*)(unsigned long long)arr),
        last(last_row) {}
  /* value at row r (wave-uniform) */
  __device__ __forceinline__ T operator()(int r) const { return p[max(0, min(r, last))]; }
};

/* Reciprocal cell heights at a wave-uniform row. On power-of-two meshes
 * (P2: every spacing a power of two, exponents well inside the normal range)
 * from the scalar-loaded height by exponent arithmetic, bit for bit the IEEE
 * quotient 1/cdy; otherwise the quotient (only P2 code reads these). */
template <typename T, bool P2>
struct RowRecip {
  RowScalar<T> d;
  /* (the flag matches RowGeo's constructor; these are always reciprocals) */
  __device__ RowRecip(const T* arr, int base_row, int last_row, int lane, bool = true)
      : d(arr, base_row, last_row, lane) {}
  __device__ __forceinline__ T operator()(int r) const {
    if constexpr (P2)
      return recip_pow2(d(r));
    else
      return T(1.0) / d(r);
  }
};

/* Row geometry of a uniform mesh (each spacing array holds one value
 * throughout): the value is read once per work item and every row reads it
 * from a scalar register. Same values as RowScalar / RowRecip at every row.
 * Measured on MI355X at 8192^2: the fused momentum march 5% faster, the
 * mass/energy march 0.5-1%; the Lagrangian march was not (it keeps the
 * per-row loads). */
template <typename T>
struct RowConst {
  T v;
  __device__ RowConst(const T* arr, int, int, int, bool = false)
      : v(((const __attribute__((address_space(4))) T*)(unsigned long long)arr)[0]) {}
  __device__ __forceinline__ T operator()(int) const { return v; }
};
template <typename T, bool P2>
struct RowConstRecip {
  T v;
  __device__ RowConstRecip(const T* arr, int, int, int, bool = true) {
    const T d = ((const __attribute__((address_space(4))) T*)(unsigned long long)arr)[0];
    if constexpr (P2)
      v = recip_pow2(d);
    else
      v = T(1.0) / d;
  }
  __device__ __forceinline__ T operator()(int) const { return v; }
};
/* UNI selects the uniform-mesh forms */
template <typename T, bool UNI>
using RowS = typename std::conditional<UNI, RowConst<T>, RowScalar<T>>::type;
template <typename T, bool P2, bool UNI>
using RowR = typename std::conditional<UNI, RowConstRecip<T, P2>, RowRecip<T, P2>>::type;

/* ---- boundary-strip writers (ghost writes only in edge items) ---- */
template <typename T>
__device__ __forceinline__ void mput_cell(bool edge, T* a, int nx, int ny, int i,
                                          int j, T v) {
  a[i * nx + j] = v;
  if (!edge) return;
  if (j == kPad) a[i * nx + kPad - 1] = v;
  if (j == kPad + 1) a[i * nx + kPad - 2] = v;
  if (j == nx - kPad - 1) a[i * nx + nx - kPad] = v;
  if (j == nx - kPad - 2) a[i * nx + nx - kPad + 1] = v;
  if (i == kPad) a[(kPad - 1) * nx + j] = v;
  if (i == kPad + 1) a[(kPad - 2) * nx + j] = v;
  if (i == ny - kPad - 1) a[(ny - kPad) * nx + j] = v;
  if (i == ny - kPad - 2) a[(ny - kPad + 1) * nx + j] = v;
}
template <typename T>
__device__ __forceinline__ void mput_xface(bool edge, T* a, int nx, int ny, int i,
                                           int f, T v, T sx) {
  const int s = nx + 1;
  a[i * s + f] = v;
  if (!edge) return;
  if (f == kPad) a[i * s + kPad - 1] = sx * v;
  if (f == kPad + 1) a[i * s + kPad - 2] = sx * v;
  if (f == nx - kPad) a[i * s + nx - kPad + 1] = sx * v;
  if (f == nx - kPad - 1) a[i * s + nx - kPad + 2] = sx * v;
  if (i == kPad) a[(kPad - 1) * s + f] = v;
  if (i == kPad + 1) a[(kPad - 2) * s + f] = v;
  if (i == ny - kPad - 1) a[(ny - kPad) * s + f] = v;
  if (i == ny - kPad - 2) a[(ny - kPad + 1) * s + f] = v;
}
template <typename T>
__device__ __forceinline__ void mput_yface(bool edge, T* a, int nx, int ny, int g,
                                           int j, T v, T sy) {
  a[g * nx + j] = v;
  if (!edge) return;
  if (g == kPad) a[(kPad - 1) * nx + j] = sy * v;
  if (g == kPad + 1) a[(kPad - 2) * nx + j] = sy * v;
  if (g == ny - kPad) a[(ny - kPad + 1) * nx + j] = sy * v;
  if (g == ny - kPad - 1) a[(ny - kPad + 2) * nx + j] = sy * v;
  if (j == kPad) a[g * nx + kPad - 1] = v;
  if (j == kPad + 1) a[g * nx + kPad - 2] = v;
  if (j == nx - kPad - 1) a[g * nx + nx - kPad] = v;
  if (j == nx - kPad - 2) a[g * nx + nx - kPad + 1] = v;
}

/* ---- two columns per lane: the FP32 mass/energy and Lagrangian marches'
 * interior items ----
 * Lane l holds columns qb + c, c = 0, 1, with qb = x0 - 2 + 2l (mass/energy:
 * 124 owned columns, two halo columns on each side) or x0 - 4 + 2l
 * (Lagrangian: 120 owned, four halo columns each side). A row of any array
 * is one 8-byte access per lane (512 bytes per wave, as a row of the FP64
 * march), and the row's addresses, lane shifts, row tests and loop control
 * serve two columns. Each column's arithmetic is the one-column march's, in
 * the same order, so the fields are bitwise the same. */
constexpr int kMassW = kWave - 4;       /* one column per lane */
constexpr int kMassW2 = 2 * kWave - 4;  /* two */
/* Two consecutive values of a row: V aligned to its size, for cell and
 * y-face arrays; Vx aligned to one value, for the x-face arrays, whose rows
 * (odd stride nx+1) start on odd columns every other row
 * (global_load_dwordx2/x4 and global_store_dwordx2/x4 need only dword
 * alignment). */
template <typename T>
struct PairVec;
template <>
struct PairVec<float> {
  typedef float V __attribute__((ext_vector_type(2)));
  typedef V Vx __attribute__((aligned(4)));
};
template <>
struct PairVec<double> {
  typedef double V __attribute__((ext_vector_type(2)));
  typedef V Vx __attribute__((aligned(8)));
};

/* the value at column qb + c + d of a pair array, d in -2..2.
 * Take every shift in wave-uniform control flow, before any branch on the
 * lane's columns: a lane shift reads zero from a lane the branch has
 * switched off. */
template <typename T>
__device__ __forceinline__ T pshift(const T (&x)[2], int c, int d) {
  const int k = c + d;
  if (k < 0) return dpp_move<kDppWaveShr1>(x[k + 2]);
  if (k > 1) return dpp_move<kDppWaveShl1>(x[k - 2]);
  return x[k];
}
/* a row's two columns of a cell or y-face array (idx even) */
template <typename T>
__device__ __forceinline__ void ld_pair(const T* a, long idx, T (&x)[2]) {
  const typename PairVec<T>::V v = *reinterpret_cast<const typename PairVec<T>::V*>(a + idx);
  x[0] = v.x;
  x[1] = v.y;
}
template <typename T>
__device__ __forceinline__ void st_pair(T* a, long idx, const T (&x)[2]) {
  typename PairVec<T>::V v;
  v.x = x[0];
  v.y = x[1];
  *reinterpret_cast<typename PairVec<T>::V*>(a + idx) = v;
}
/* the same for an x-face array (any idx) */
template <typename T>
__device__ __forceinline__ void ld_xpair(const T* a, long idx, T (&x)[2]) {
  const typename PairVec<T>::V v = *reinterpret_cast<const typename PairVec<T>::Vx*>(a + idx);
  x[0] = v.x;
  x[1] = v.y;
}
template <typename T>
__device__ __forceinline__ void st_xpair(T* a, long idx, const T (&x)[2]) {
  typename PairVec<T>::V v;
  v.x = x[0];
  v.y = x[1];
  *reinterpret_cast<typename PairVec<T>::Vx*>(a + idx) = v;
}
template <typename T>
__device__ __forceinline__ void copy_pair(T (&d)[2], const T (&s)[2]) {
  d[0] = s[0];
  d[1] = s[1];
}

/* A lane's sums of an item's cells for the mass and energy totals, rounded
 * once into the exact fixed-point accumulators. FP64 state: compensated
 * (CSum). FP32 state: rho and rho*e formed in FP64 are exact (24-bit
 * mantissas, a 48-bit product), and a lane adds at most 2 x 120 of them
 * per item, so a plain FP64 sum is within 240 x 2^-53 of the exact one
 * (relative), far inside the FP32 state's own rounding, at 4 FP64
 * operations a cell where the compensated sums take 17. */
struct PlainSum {
  double s, c; /* c stays 0: the same interface as CSum */
  __device__ void add(double x) { s += x; }
};
template <typename T>
using ItemSum = typename std::conditional<sizeof(T) == 4, PlainSum, CSum>::type;

/* ======================================================================= */
/* mass and energy, x sweep first (even tt)                                 */
/* ======================================================================= */

/* Lane l holds column q = x0 - 2 + l. Per row r: the x fluxes at faces
 * [x0, x1] and the first (x) update of cells [x0, x1); per lane a window of
 * the x-updated rows r-3..r feeds the y flux at face r-1 and the second (y)
 * update of row r-2. */
template <typename T, bool P2, bool EDGE = true, bool UNI = false, typename Acc = CSum>
__device__ __forceinline__ void mass_xy_box(const MGeo<T>& g, const MMassIO<T>& io, T dt,
                                            T dt_h, const ItemBox box, int lane, Acc& ms,
                                            Acc& es) {
  const int nx = g.nx, ny = g.ny, sx = nx + 1;
  (void)dt;
  (void)sx;
  const int x0 = box.x0, x1 = box.x1, y0 = box.y0, y1 = box.y1;
  if (y0 >= y1) return;
  const bool edge = EDGE && ((x0 == kPad) || (x1 + 3 > nx - kPad) || (y0 == kPad) ||
                             (y1 + 3 > ny - kPad));
  const int q = x0 - 2 + lane;
  const bool inq = q < nx;                     /* column exists */
  const bool face = q >= x0 && q <= x1;        /* lane computes face q */
  const bool own = q >= x0 && q < x1;          /* lane outputs cell q */
  const bool own_face = face && (q < x1 || x1 == nx - kPad);
  const int qc = inq ? q : nx - 1;             /* a valid column for loads */
  const int qf = q <= nx ? q : nx;             /* a valid x face for loads */
  /* per-lane geometry, loaded once per item: no load inside the row loop
   * other than the next row's prefetch */
  const T cdx_q = g.cdx[qc], edx_q = g.edx[qc], edx_q1 = g.edx[qc + 1];
  const int qg = face ? q : x0;
  const T cdx_qm1 = g.cdx[qg - 1], iedx_q = g.iedx[qg];
  const RowS<T, UNI> edy(g.edy, y0 - 4, ny, lane), cdy(g.cdy, y0 - 4, ny, lane);
  const RowS<T, UNI> iedy(g.iedy, y0 - 4, ny, lane);
  /* reciprocal spacings, used in place of quotients on power-of-two meshes */
  const T rcdx_q = T(1.0) / cdx_q;
  const RowR<T, P2, UNI> rcdy(g.cdy, y0 - 4, ny, lane);

  /* rows of the x-updated fields for this lane's column: w[0] oldest */
  T wr[4], we[4];
  T vy[3];   /* v0 at faces r-2 .. r */
  T r0[3];   /* rho0 at rows r-2 .. r */
  T fy_prev = T(0), efy_prev = T(0);
  const bool south = EDGE && (y0 == kPad);
  const int rb = south ? kPad : y0 - 2;
#pragma unroll
  for (int k = 0; k < 4; ++k) wr[k] = we[k] = T(0);
  /* rows rb-2 and rb-1 feed only fluxes and updates above y0, which are not
   * computed, except v0 of row y0-1 at the south wall (rb = y0): the flux
   * at face y0 takes it */
  vy[0] = r0[0] = r0[1] = T(0);
  vy[1] = south ? io.v0[(rb - 1) * nx + qc] : T(0);
  /* row rb's inputs; later rows are loaded one step ahead */
  T a_r = io.rho0[rb * nx + qc], a_e = io.e1[rb * nx + qc];
  T a_u = io.u0[rb * sx + qf], a_v = io.v0[rb * nx + qc];
  for (int r = rb; r <= y1 + 1; ++r) {
    const bool ghost_row = EDGE && r >= ny - kPad;
    /* prefetch the next row (the last row, y1+1, has none) */
    T n_r = T(0), n_e = T(0), n_u = T(0), n_v = T(0);
    if (r <= y1) {
      const int rn = min(r + 1, ny - kPad - 1);
      n_r = io.rho0[rn * nx + qc];
      n_e = io.e1[rn * nx + qc];
      n_u = io.u0[rn * sx + qf];
      n_v = io.v0[min(r + 1, ny) * nx + qc];
    }
    vy[2] = a_v;
    r0[2] = a_r;

    T xr, xe;
    if (!ghost_row) {
      /* x fluxes at face q from the row in the lanes around it */
      const T rm2 = up(a_r, 2), rm1 = up(a_r, 1), rp1 = dn(a_r, 1);
      const T em2 = up(a_e, 2), em1 = up(a_e, 1), ep1 = dn(a_e, 1);
      const T um1 = up(a_u, 1), up1 = dn(a_u, 1);
      FaceFlux<T> fl{T(0), T(0)};
      if (face)
        fl = mass_energy_flux<T, false, P2>(um1, a_u, up1, rm2, rm1, a_r, rp1, em2,
                                            em1, a_e, ep1, iedx_q, cdx_qm1, cdx_q,
                                            edy(r), dt, dt_h, rcdx_q);
      if (own_face && r >= y0 && r < y1)
        mput_xface(edge, io.F_x, nx, ny, r, q, fl.mass, T(-1.0));
      const T f_hi = dn(fl.mass, 1), ef_hi = dn(fl.energy, 1);
      const CellState<T> st =
          sweep_update<T, P2>(a_r, a_r, a_e, fl.mass, f_hi, fl.energy, ef_hi, edy(r),
                              edy(r + 1), cdx_q, cdy(r), dt_h, true,
                              rcdx_q * rcdy(r));
      xr = st.rho;
      xe = st.e;
    } else {
      /* ghost row past the north wall: the mirror row's x update */
      xr = (r == ny - kPad) ? wr[3] : wr[1];
      xe = (r == ny - kPad) ? we[3] : we[1];
    }
#pragma unroll
    for (int k = 0; k < 3; ++k) {
      wr[k] = wr[k + 1];
      we[k] = we[k + 1];
    }
    wr[3] = xr;
    we[3] = xe;
    if (south && r == kPad + 1) {
      /* ghost rows below the south wall mirror rows pad, pad+1 */
      wr[0] = wr[3];
      we[0] = we[3];
      wr[1] = wr[2];
      we[1] = we[2];
    }
    /* y flux at face r-1 from rows r-3 .. r */
    const int gf = r - 1;
    if (gf >= y0 && gf <= y1) {
      const FaceFlux<T> fl = mass_energy_flux<T, true, P2>(
          vy[0], vy[1], vy[2], wr[0], wr[1], wr[2], wr[3], we[0], we[1], we[2],
          we[3], iedy(gf), cdy(gf - 1), cdy(gf), edx_q, dt, dt_h, rcdy(gf));
      if (own && (gf < y1 || y1 == ny - kPad))
        mput_yface(edge, io.F_y, nx, ny, gf, q, fl.mass, T(-1.0));
      /* y update of row r-2 */
      const int i = r - 2;
      if (i >= y0 && own) {
        const CellState<T> st =
            sweep_update<T, P2>(wr[1], r0[0], we[1], fy_prev, fl.mass, efy_prev,
                                fl.energy, edx_q, edx_q1, cdx_q, cdy(i), dt_h, false,
                                rcdx_q * rcdy(i));
        mput_cell(edge, io.rho, nx, ny, i, q, st.rho);
        mput_cell(edge, io.e, nx, ny, i, q, st.e);
        ms.add(double(st.rho));
        es.add(double(st.rho) * double(st.e));
      }
      fy_prev = fl.mass;
      efy_prev = fl.energy;
    }
    vy[0] = vy[1];
    vy[1] = vy[2];
    r0[0] = r0[1];
    r0[1] = r0[2];
    a_r = n_r;
    a_e = n_e;
    a_u = n_u;
    a_v = n_v;
  }
}

template <typename T, bool P2, bool EDGE = true, bool UNI = false, typename Acc = CSum>
__device__ __forceinline__ void mass_xy_item(const MGeo<T>& g, const MMassIO<T>& io, T dt,
                                             T dt_h, const MarchList& ml, int item, int lane,
                                             Acc& ms, Acc& es) {
  mass_xy_box<T, P2, EDGE, UNI, Acc>(g, io, dt, dt_h, item_box(ml, g.nx, item), lane, ms, es);
}

/* The x-first mass/energy march of an interior item, two columns per lane
 * (mass_xy_box's interior form, column by column). */
template <typename T, bool P2, bool UNI, typename Acc>
__device__ __forceinline__ void mass_xy_pair(const MGeo<T>& g, const MMassIO<T>& io, T dt,
                                             T dt_h, const ItemBox box, int lane, Acc& ms,
                                             Acc& es) {
  const int nx = g.nx, ny = g.ny, sx = nx + 1;
  const int x0 = box.x0, x1 = box.x1, y0 = box.y0, y1 = box.y1;
  if (y0 >= y1) return;
  const int qb = x0 - 2 + 2 * lane; /* the lane's first column */
  /* x0 and the strip width are even, so a lane owns both columns or neither */
  const bool own = qb >= x0 && qb < x1;
  bool face[2];
  T cdx_q[2], edx_q[2], edx_q1[2], cdx_qm1[2], iedx_q[2], rcdx_q[2];
#pragma unroll
  for (int c = 0; c < 2; ++c) {
    const int q = qb + c;
    face[c] = q >= x0 && q <= x1;
    cdx_q[c] = g.cdx[q];
    edx_q[c] = g.edx[q];
    edx_q1[c] = g.edx[q + 1];
    const int qg = face[c] ? q : x0;
    cdx_qm1[c] = g.cdx[qg - 1];
    iedx_q[c] = g.iedx[qg];
    rcdx_q[c] = T(1.0) / cdx_q[c];
  }
  const RowS<T, UNI> edy(g.edy, y0 - 4, ny, lane), cdy(g.cdy, y0 - 4, ny, lane);
  const RowS<T, UNI> iedy(g.iedy, y0 - 4, ny, lane);
  const RowR<T, P2, UNI> rcdy(g.cdy, y0 - 4, ny, lane);
  auto ldc = [&](const T* a, int row, T(&x)[2]) { ld_pair(a, (long)row * nx + qb, x); };
  auto ldx = [&](const T* a, int row, T(&x)[2]) { ld_xpair(a, (long)row * sx + qb, x); };

  /* Row rings of four slots, row k in slot (k - rb) & 3: rho0 and v0 rows
   * r-2 .. r+1, the x-updated rho and e rows r-3 .. r; e1 and u0 rows r,
   * r+1 in two slots, the y fluxes of faces r-2, r-1 likewise. The row loop
   * is unrolled by four with the slot of each row fixed at compile time,
   * so no window is copied from row to row. */
  T r0[4][2], vy[4][2], wr[4][2], we[4][2], eb[2][2], ub[2][2];
  T fyb[2][2], efyb[2][2];
  const int rb = y0 - 2;
#pragma unroll
  for (int k = 0; k < 4; ++k)
    wr[k][0] = wr[k][1] = we[k][0] = we[k][1] = T(0);
#pragma unroll
  for (int k = 0; k < 2; ++k)
    fyb[k][0] = fyb[k][1] = efyb[k][0] = efyb[k][1] = T(0);
  /* rows rb-2 and rb-1 of rho0 and v0 would only feed the y fluxes and
   * updates of faces and rows above y0, which are not computed */
#pragma unroll
  for (int k = 2; k < 4; ++k) r0[k][0] = r0[k][1] = vy[k][0] = vy[k][1] = T(0);
  ldc(io.rho0, rb, r0[0]);
  ldc(io.e1, rb, eb[0]);
  ldx(io.u0, rb, ub[0]);
  ldc(io.v0, rb, vy[0]);
  auto row = [&](auto phase, int r) {
    constexpr int P = decltype(phase)::value;
    constexpr int S0 = P, S1 = (P + 1) & 3, S2 = (P + 2) & 3, S3 = (P + 3) & 3;
    constexpr int A = P & 1, B = (P + 1) & 1;
    /* the next row (an interior item's rows all lie inside the mesh); the
     * last row, y1+1, has none to load. FP32 loads it anyway: the branch
     * measured 1% slower there, and 0.5% faster in FP64. */
    if (sizeof(T) == 4 || r <= y1) {
      ldc(io.rho0, r + 1, r0[S1]);
      ldc(io.e1, r + 1, eb[B]);
      ldx(io.u0, r + 1, ub[B]);
      ldc(io.v0, r + 1, vy[S1]);
    }
    const T(&a_r)[2] = r0[S0];
    const T(&a_e)[2] = eb[A];
    const T(&a_u)[2] = ub[A];

    /* x fluxes and the x update of row r */
    T fm[2], fe[2];
#pragma unroll
    for (int c = 0; c < 2; ++c) {
      const T um1 = pshift(a_u, c, -1), up1 = pshift(a_u, c, 1);
      const T rm2 = pshift(a_r, c, -2), rm1 = pshift(a_r, c, -1), rp1 = pshift(a_r, c, 1);
      const T em2 = pshift(a_e, c, -2), em1 = pshift(a_e, c, -1), ep1 = pshift(a_e, c, 1);
      FaceFlux<T> fl{T(0), T(0)};
      if (face[c])
        fl = mass_energy_flux<T, false, P2>(um1, a_u[c], up1, rm2, rm1, a_r[c], rp1, em2,
                                            em1, a_e[c], ep1, iedx_q[c], cdx_qm1[c],
                                            cdx_q[c], edy(r), dt, dt_h, rcdx_q[c]);
      fm[c] = fl.mass;
      fe[c] = fl.energy;
    }
    if (own && r >= y0 && r < y1) st_xpair(io.F_x, (long)r * sx + qb, fm);
#pragma unroll
    for (int c = 0; c < 2; ++c) {
      const CellState<T> st =
          sweep_update<T, P2>(a_r[c], a_r[c], a_e[c], fm[c], pshift(fm, c, 1), fe[c],
                              pshift(fe, c, 1), edy(r), edy(r + 1), cdx_q[c], cdy(r), dt_h,
                              true, rcdx_q[c] * rcdy(r));
      wr[S0][c] = st.rho;
      we[S0][c] = st.e;
    }
    /* y flux at face r-1 from x-updated rows r-3 .. r, and the y update of
     * row r-2 */
    const int gf = r - 1;
    if (gf >= y0 && gf <= y1) {
#pragma unroll
      for (int c = 0; c < 2; ++c) {
        const FaceFlux<T> fl = mass_energy_flux<T, true, P2>(
            vy[S2][c], vy[S3][c], vy[S0][c], wr[S1][c], wr[S2][c], wr[S3][c], wr[S0][c],
            we[S1][c], we[S2][c], we[S3][c], we[S0][c], iedy(gf), cdy(gf - 1), cdy(gf),
            edx_q[c], dt, dt_h, rcdy(gf));
        fyb[A][c] = fl.mass;
        efyb[A][c] = fl.energy;
      }
      if (own && gf < y1) st_pair(io.F_y, (long)gf * nx + qb, fyb[A]);
      const int i = r - 2;
      if (i >= y0 && own) {
        T nr[2], ne[2];
#pragma unroll
        for (int c = 0; c < 2; ++c) {
          const CellState<T> st = sweep_update<T, P2>(
              wr[S2][c], r0[S2][c], we[S2][c], fyb[B][c], fyb[A][c], efyb[B][c], efyb[A][c],
              edx_q[c], edx_q1[c], cdx_q[c], cdy(i), dt_h, false, rcdx_q[c] * rcdy(i));
          nr[c] = st.rho;
          ne[c] = st.e;
        }
        st_pair(io.rho, (long)i * nx + qb, nr);
        st_pair(io.e, (long)i * nx + qb, ne);
#pragma unroll
        for (int c = 0; c < 2; ++c) {
          ms.add(double(nr[c]));
          es.add(double(nr[c]) * double(ne[c]));
        }
      }
    }
  };
  using P0 = std::integral_constant<int, 0>;
  using P1 = std::integral_constant<int, 1>;
  using P2c = std::integral_constant<int, 2>;
  using P3 = std::integral_constant<int, 3>;
  int r = rb;
  for (; r + 3 <= y1 + 1; r += 4) {
    row(P0{}, r);
    row(P1{}, r + 1);
    row(P2c{}, r + 2);
    row(P3{}, r + 3);
  }
  if (r <= y1 + 1) row(P0{}, r);
  if (r + 1 <= y1 + 1) row(P1{}, r + 1);
  if (r + 2 <= y1 + 1) row(P2c{}, r + 2);
}

/* PAIR: interior items two columns per lane (FP32, strips of kMassW2
 * columns; edge items go through the one-column form in sub-strips of
 * kMassW). */
template <typename T, bool P2, bool UNI = false, bool PAIR = false>
__global__ __launch_bounds__(kMNT, FLOWHIP_MARCH_WAVES) void k_march_mass_xy(
    MGeo<T> g, MMassIO<T> io, const double* dtp, FkReduce red, MarchList ml) {
  const T dt = T(dtp[0]), dt_h = T(dtp[1]);
  const int lane = threadIdx.x % kWave;
  /* each item's per-lane compensated sums, rounded once, go into exact
   * accumulators: the totals do not depend on which wave takes which item */
  FixSum fm{0, 0, 0, 0}, fe{0, 0, 0, 0};
  MarchQueue<true> mq(red.queue, ml.nitems);
  for (int item = mq.next(); item < mq.end; item = mq.next()) {
    ItemSum<T> ms{0.0, 0.0}, es{0.0, 0.0};
    if constexpr (PAIR) {
      const ItemBox b = item_box(ml, g.nx, item);
      if (item_edge(ml, g.nx, g.ny, item)) {
        for (int sx0 = b.x0; sx0 < b.x1; sx0 += kMassW) {
          const ItemBox sb{sx0, min(sx0 + kMassW, b.x1), b.y0, b.y1};
          mass_xy_box<T, P2, true, UNI, ItemSum<T>>(g, io, dt, dt_h, sb, lane, ms, es);
        }
      } else {
        mass_xy_pair<T, P2, UNI>(g, io, dt, dt_h, b, lane, ms, es);
      }
    } else if (item_edge(ml, g.nx, g.ny, item)) {
      mass_xy_item<T, P2, true, UNI>(g, io, dt, dt_h, ml, item, lane, ms, es);
    } else {
      mass_xy_item<T, P2, false, UNI>(g, io, dt, dt_h, ml, item, lane, ms, es);
    }
    fm.add(ms.s + ms.c);
    fe.add(es.s + es.c);
  }
  mq.finish();
  finish_fixed_sums<kMNT>(fm, fe, red);
}

/* ======================================================================= */
/* mass and energy, y sweep first (odd tt)                                  */
/* ======================================================================= */

/* Lane l holds column q = x0 - 2 + l and runs the y sweep at its mirror
 * column (the same column inside the domain). Per row r: the y flux at face
 * r+1 and the first (y) update of row r; then across the lanes the x fluxes
 * at faces [x0, x1] and the second (x) update of cells [x0, x1). */
template <typename T, bool P2, bool EDGE = true, bool UNI = false, typename Acc = CSum>
__device__ __forceinline__ void mass_yx_box(const MGeo<T>& g, const MMassIO<T>& io, T dt,
                                            T dt_h, const ItemBox box, int lane, Acc& ms,
                                            Acc& es) {
  const int nx = g.nx, ny = g.ny, sx = nx + 1;
  (void)dt;
  (void)sx;
  const int x0 = box.x0, x1 = box.x1, y0 = box.y0, y1 = box.y1;
  if (y0 >= y1) return;
  const bool edge = EDGE && ((x0 == kPad) || (x1 + 3 > nx - kPad) || (y0 == kPad) ||
                             (y1 + 3 > ny - kPad));
  const int q = x0 - 2 + lane;
  const bool inq = q < nx;
  const bool face = q >= x0 && q <= x1;
  const bool own = q >= x0 && q < x1;
  const bool own_face = face && (q < x1 || x1 == nx - kPad);
  /* the y sweep of a ghost column is that of its mirror column */
  const int qm = !EDGE ? q : (inq ? mirror(q, nx) : nx - kPad - 1);
  const int qc = inq ? q : nx - 1;
  const int qf = q <= nx ? q : nx;
  const T cdx_m = g.cdx[qm], edx_m = g.edx[qm], edx_m1 = g.edx[qm + 1];
  const T cdx_q = g.cdx[qc];
  const int qg = face ? q : x0;
  const T cdx_qm1 = g.cdx[qg - 1], iedx_q = g.iedx[qg];
  const RowS<T, UNI> edy(g.edy, y0 - 4, ny, lane), cdy(g.cdy, y0 - 4, ny, lane);
  const RowS<T, UNI> iedy(g.iedy, y0 - 4, ny, lane);
  /* reciprocal spacings, used in place of quotients on power-of-two meshes */
  const T rcdx_q = T(1.0) / cdx_q, rcdx_m = T(1.0) / cdx_m;
  const RowR<T, P2, UNI> rcdy(g.cdy, y0 - 4, ny, lane);

  /* rho0, e1 at rows r-1 .. r+2 and v0 at faces r .. r+2 of column qm */
  T wr[4], we[4], wv[3];
#pragma unroll
  for (int k = 0; k < 4; ++k) {
    wr[k] = io.rho0[(y0 - 2 + k) * nx + qm];
    we[k] = io.e1[(y0 - 2 + k) * nx + qm];
  }
#pragma unroll
  for (int k = 0; k < 3; ++k) wv[k] = io.v0[(y0 - 1 + k) * nx + qm];
  /* y flux at face y0 (rows y0-2 .. y0+1) */
  FaceFlux<T> fprev = mass_energy_flux<T, true, P2>(
      wv[0], wv[1], wv[2], wr[0], wr[1], wr[2], wr[3], we[0], we[1], we[2], we[3],
      iedy(y0), cdy(y0 - 1), cdy(y0), edx_m, dt, dt_h, rcdy(y0));
  if (own && q == qm) mput_yface(edge, io.F_y, nx, ny, y0, q, fprev.mass, T(-1.0));
  /* shift: windows now hold rows r-2 .. r+1 for r = y0; load row r+2 */
#pragma unroll
  for (int k = 0; k < 3; ++k) {
    wr[k] = wr[k + 1];
    we[k] = we[k + 1];
  }
  wv[0] = wv[1];
  wv[1] = wv[2];
  T n_r = io.rho0[(y0 + 2) * nx + qm], n_e = io.e1[(y0 + 2) * nx + qm];
  T n_v = io.v0[(y0 + 2) * nx + qm];
  T a_u = io.u0[y0 * sx + qf];
  T a_r0 = io.rho0[y0 * nx + qc];
  for (int r = y0; r < y1; ++r) {
    wr[3] = n_r;
    we[3] = n_e;
    wv[2] = n_v;
    const T c_u = a_u, c_r0 = a_r0;
    /* prefetch the next row's inputs (the last row has none) */
    if (r + 1 < y1) {
      const int rn = min(r + 3, ny);
      n_r = io.rho0[rn * nx + qm];
      n_e = io.e1[rn * nx + qm];
      n_v = io.v0[rn * nx + qm];
      const int rr = min(r + 1, ny - 1);
      a_u = io.u0[rr * sx + qf];
      a_r0 = io.rho0[rr * nx + qc];
    }
    /* y flux at face r+1 (rows r-1 .. r+2) and the y update of row r */
    const int gf = r + 1;
    const FaceFlux<T> fl = mass_energy_flux<T, true, P2>(
        wv[0], wv[1], wv[2], wr[0], wr[1], wr[2], wr[3], we[0], we[1], we[2],
        we[3], iedy(gf), cdy(gf - 1), cdy(gf), edx_m, dt, dt_h, rcdy(gf));
    if (own && q == qm && (gf < y1 || y1 == ny - kPad))
      mput_yface(edge, io.F_y, nx, ny, gf, q, fl.mass, T(-1.0));
    const CellState<T> ys =
        sweep_update<T, P2>(wr[1], wr[1], we[1], fprev.mass, fl.mass, fprev.energy,
                            fl.energy, edx_m, edx_m1, cdx_m, cdy(r), dt_h, true,
                            rcdx_m * rcdy(r));
    fprev = fl;
    /* x fluxes at face q of the y-updated row, then the x update */
    const T rm2 = up(ys.rho, 2), rm1 = up(ys.rho, 1), rp1 = dn(ys.rho, 1);
    const T em2 = up(ys.e, 2), em1 = up(ys.e, 1), ep1 = dn(ys.e, 1);
    const T um1 = up(c_u, 1), up1 = dn(c_u, 1);
    FaceFlux<T> fx{T(0), T(0)};
    if (face)
      fx = mass_energy_flux<T, false, P2>(um1, c_u, up1, rm2, rm1, ys.rho, rp1, em2,
                                          em1, ys.e, ep1, iedx_q, cdx_qm1, cdx_q,
                                          edy(r), dt, dt_h, rcdx_q);
    if (own_face) mput_xface(edge, io.F_x, nx, ny, r, q, fx.mass, T(-1.0));
    const T f_hi = dn(fx.mass, 1), ef_hi = dn(fx.energy, 1);
    if (own) {
      const CellState<T> st =
          sweep_update<T, P2>(ys.rho, c_r0, ys.e, fx.mass, f_hi, fx.energy, ef_hi,
                              edy(r), edy(r + 1), cdx_q, cdy(r), dt_h, false,
                              rcdx_q * rcdy(r));
      mput_cell(edge, io.rho, nx, ny, r, q, st.rho);
      mput_cell(edge, io.e, nx, ny, r, q, st.e);
      ms.add(double(st.rho));
      es.add(double(st.rho) * double(st.e));
    }
#pragma unroll
    for (int k = 0; k < 3; ++k) {
      wr[k] = wr[k + 1];
      we[k] = we[k + 1];
    }
    wv[0] = wv[1];
    wv[1] = wv[2];
  }
}

template <typename T, bool P2, bool EDGE = true, bool UNI = false, typename Acc = CSum>
__device__ __forceinline__ void mass_yx_item(const MGeo<T>& g, const MMassIO<T>& io, T dt,
                                             T dt_h, const MarchList& ml, int item, int lane,
                                             Acc& ms, Acc& es) {
  mass_yx_box<T, P2, EDGE, UNI, Acc>(g, io, dt, dt_h, item_box(ml, g.nx, item), lane, ms, es);
}

/* The y-first mass/energy march of an interior item, two columns per lane
 * (mass_yx_box's interior form, column by column). */
template <typename T, bool P2, bool UNI, typename Acc>
__device__ __forceinline__ void mass_yx_pair(const MGeo<T>& g, const MMassIO<T>& io, T dt,
                                             T dt_h, const ItemBox box, int lane, Acc& ms,
                                             Acc& es) {
  const int nx = g.nx, ny = g.ny, sx = nx + 1;
  const int x0 = box.x0, x1 = box.x1, y0 = box.y0, y1 = box.y1;
  if (y0 >= y1) return;
  const int qb = x0 - 2 + 2 * lane;
  const bool own = qb >= x0 && qb < x1;
  bool face[2];
  T cdx_q[2], edx_q[2], edx_q1[2], cdx_qm1[2], iedx_q[2], rcdx_q[2];
#pragma unroll
  for (int c = 0; c < 2; ++c) {
    const int q = qb + c;
    face[c] = q >= x0 && q <= x1;
    cdx_q[c] = g.cdx[q];
    edx_q[c] = g.edx[q];
    edx_q1[c] = g.edx[q + 1];
    const int qg = face[c] ? q : x0;
    cdx_qm1[c] = g.cdx[qg - 1];
    iedx_q[c] = g.iedx[qg];
    rcdx_q[c] = T(1.0) / cdx_q[c];
  }
  const RowS<T, UNI> edy(g.edy, y0 - 4, ny, lane), cdy(g.cdy, y0 - 4, ny, lane);
  const RowS<T, UNI> iedy(g.iedy, y0 - 4, ny, lane);
  const RowR<T, P2, UNI> rcdy(g.cdy, y0 - 4, ny, lane);
  auto ldc = [&](const T* a, int row, T(&x)[2]) { ld_pair(a, (long)row * nx + qb, x); };
  auto ldx = [&](const T* a, int row, T(&x)[2]) { ld_xpair(a, (long)row * sx + qb, x); };

  /* Rings of four slots: rho0 and e1 row k in slot (k - y0 + 2) & 3 (rows
   * r-1 .. r+2 in use at row r), v0 face k in slot (k - y0) & 3 (faces
   * r .. r+3); u0 rows and the y fluxes at faces r, r+1 in two slots. The
   * row loop is unrolled by four with every slot fixed at compile time, so
   * no window is copied from row to row. */
  T wr[4][2], we[4][2], wv[4][2], ub[2][2], fyb[2][2], efyb[2][2];
#pragma unroll
  for (int k = 0; k < 4; ++k) {
    ldc(io.rho0, y0 - 2 + k, wr[k]);
    ldc(io.e1, y0 - 2 + k, we[k]);
  }
  ldc(io.v0, y0 - 1, wv[3]);
  ldc(io.v0, y0, wv[0]);
  ldc(io.v0, y0 + 1, wv[1]);
  /* y flux at face y0 (rows y0-2 .. y0+1) */
#pragma unroll
  for (int c = 0; c < 2; ++c) {
    const FaceFlux<T> fl = mass_energy_flux<T, true, P2>(
        wv[3][c], wv[0][c], wv[1][c], wr[0][c], wr[1][c], wr[2][c], wr[3][c], we[0][c],
        we[1][c], we[2][c], we[3][c], iedy(y0), cdy(y0 - 1), cdy(y0), edx_q[c], dt, dt_h,
        rcdy(y0));
    fyb[0][c] = fl.mass;
    efyb[0][c] = fl.energy;
  }
  if (own) st_pair(io.F_y, (long)y0 * nx + qb, fyb[0]);
  ldc(io.rho0, y0 + 2, wr[0]);
  ldc(io.e1, y0 + 2, we[0]);
  ldc(io.v0, y0 + 2, wv[2]);
  ldx(io.u0, y0, ub[0]);
  auto row = [&](auto phase, int r) {
    constexpr int P = decltype(phase)::value;
    /* rho0, e1 rows r-1, r, r+1, r+2; v0 faces r .. r+3 */
    constexpr int R1 = (P + 1) & 3, R2 = (P + 2) & 3, R3 = (P + 3) & 3, R0 = P;
    constexpr int V0 = P, V1 = (P + 1) & 3, V2 = (P + 2) & 3, V3 = (P + 3) & 3;
    constexpr int A = P & 1, B = (P + 1) & 1;
    /* the next row's inputs (an interior item's rows all lie inside); the
     * last row, y1-1, has none to load (loaded anyway in FP32, as in the
     * x-first march) */
    const bool more = sizeof(T) == 4 || r + 1 < y1;
    if (more) {
      ldc(io.v0, r + 3, wv[V3]);
      ldx(io.u0, r + 1, ub[B]);
    }
    const T(&c_u)[2] = ub[A];
    /* y flux at face r+1 (rows r-1 .. r+2) and the y update of row r */
    const int gf = r + 1;
    T yr[2], ye[2];
#pragma unroll
    for (int c = 0; c < 2; ++c) {
      const FaceFlux<T> fl = mass_energy_flux<T, true, P2>(
          wv[V0][c], wv[V1][c], wv[V2][c], wr[R1][c], wr[R2][c], wr[R3][c], wr[R0][c],
          we[R1][c], we[R2][c], we[R3][c], we[R0][c], iedy(gf), cdy(gf - 1), cdy(gf),
          edx_q[c], dt, dt_h, rcdy(gf));
      fyb[B][c] = fl.mass;
      efyb[B][c] = fl.energy;
      const CellState<T> ys =
          sweep_update<T, P2>(wr[R2][c], wr[R2][c], we[R2][c], fyb[A][c], fl.mass,
                              efyb[A][c], fl.energy, edx_q[c], edx_q1[c], cdx_q[c], cdy(r),
                              dt_h, true, rcdx_q[c] * rcdy(r));
      yr[c] = ys.rho;
      ye[c] = ys.e;
    }
    /* row r+3 into the slot of row r-1 */
    if (more) {
      ldc(io.rho0, r + 3, wr[R1]);
      ldc(io.e1, r + 3, we[R1]);
    }
    if (own && gf < y1) st_pair(io.F_y, (long)gf * nx + qb, fyb[B]);
    /* x fluxes of the y-updated row, then the x update */
    T xm[2], xe[2];
#pragma unroll
    for (int c = 0; c < 2; ++c) {
      const T um1 = pshift(c_u, c, -1), up1 = pshift(c_u, c, 1);
      const T rm2 = pshift(yr, c, -2), rm1 = pshift(yr, c, -1), rp1 = pshift(yr, c, 1);
      const T em2 = pshift(ye, c, -2), em1 = pshift(ye, c, -1), ep1 = pshift(ye, c, 1);
      FaceFlux<T> fx{T(0), T(0)};
      if (face[c])
        fx = mass_energy_flux<T, false, P2>(um1, c_u[c], up1, rm2, rm1, yr[c], rp1, em2, em1,
                                            ye[c], ep1, iedx_q[c], cdx_qm1[c], cdx_q[c],
                                            edy(r), dt, dt_h, rcdx_q[c]);
      xm[c] = fx.mass;
      xe[c] = fx.energy;
    }
    if (own) st_xpair(io.F_x, (long)r * sx + qb, xm);
    const T xm_hi[2] = {pshift(xm, 0, 1), pshift(xm, 1, 1)};
    const T xe_hi[2] = {pshift(xe, 0, 1), pshift(xe, 1, 1)};
    if (own) {
      T nr[2], ne[2];
#pragma unroll
      for (int c = 0; c < 2; ++c) {
        const CellState<T> st = sweep_update<T, P2>(
            yr[c], wr[R2][c], ye[c], xm[c], xm_hi[c], xe[c], xe_hi[c], edy(r), edy(r + 1),
            cdx_q[c], cdy(r), dt_h, false, rcdx_q[c] * rcdy(r));
        nr[c] = st.rho;
        ne[c] = st.e;
      }
      st_pair(io.rho, (long)r * nx + qb, nr);
      st_pair(io.e, (long)r * nx + qb, ne);
#pragma unroll
      for (int c = 0; c < 2; ++c) {
        ms.add(double(nr[c]));
        es.add(double(nr[c]) * double(ne[c]));
      }
    }
  };
  using P0 = std::integral_constant<int, 0>;
  using P1 = std::integral_constant<int, 1>;
  using P2c = std::integral_constant<int, 2>;
  using P3 = std::integral_constant<int, 3>;
  int r = y0;
  for (; r + 3 < y1; r += 4) {
    row(P0{}, r);
    row(P1{}, r + 1);
    row(P2c{}, r + 2);
    row(P3{}, r + 3);
  }
  if (r < y1) row(P0{}, r);
  if (r + 1 < y1) row(P1{}, r + 1);
  if (r + 2 < y1) row(P2c{}, r + 2);
}

template <typename T, bool P2, bool UNI = false, bool PAIR = false>
__global__ __launch_bounds__(kMNT, FLOWHIP_MARCH_WAVES) void k_march_mass_yx(
    MGeo<T> g, MMassIO<T> io, const double* dtp, FkReduce red, MarchList ml) {
  const T dt = T(dtp[0]), dt_h = T(dtp[1]);
  const int lane = threadIdx.x % kWave;
  /* each item's per-lane compensated sums, rounded once, go into exact
   * accumulators: the totals do not depend on which wave takes which item */
  FixSum fm{0, 0, 0, 0}, fe{0, 0, 0, 0};
  MarchQueue<true> mq(red.queue, ml.nitems);
  for (int item = mq.next(); item < mq.end; item = mq.next()) {
    ItemSum<T> ms{0.0, 0.0}, es{0.0, 0.0};
    if constexpr (PAIR) {
      const ItemBox b = item_box(ml, g.nx, item);
      if (item_edge(ml, g.nx, g.ny, item)) {
        for (int sx0 = b.x0; sx0 < b.x1; sx0 += kMassW) {
          const ItemBox sb{sx0, min(sx0 + kMassW, b.x1), b.y0, b.y1};
          mass_yx_box<T, P2, true, UNI, ItemSum<T>>(g, io, dt, dt_h, sb, lane, ms, es);
        }
      } else {
        mass_yx_pair<T, P2, UNI>(g, io, dt, dt_h, b, lane, ms, es);
      }
    } else if (item_edge(ml, g.nx, g.ny, item)) {
      mass_yx_item<T, P2, true, UNI>(g, io, dt, dt_h, ml, item, lane, ms, es);
    } else {
      mass_yx_item<T, P2, false, UNI>(g, io, dt, dt_h, ml, item, lane, ms, es);
    }
    fm.add(ms.s + ms.c);
    fe.add(es.s + es.c);
  }
  mq.finish();
  finish_fixed_sums<kMNT>(fm, fe, red);
}


/* ======================================================================= */
/* momentum                                                                */
/* ======================================================================= */

template <typename T>
struct MMomIO {
  const T* __restrict__ u;    /* x velocity read (u0 for x, u1 for y) */
  const T* __restrict__ v0;   /* y velocity after the viscous step */
  const T* __restrict__ F_x;
  const T* __restrict__ F_y;
  const T* __restrict__ rho;  /* advected density */
  const T* __restrict__ mom1; /* momentum after the viscous step */
  T* __restrict__ mom;        /* advected momentum */
  T* __restrict__ vel;        /* recomputed velocity, + ghosts */
};

/* value of v at lane src of the wave */
template <typename T>
__device__ __forceinline__ T from_lane(T v, int src) {
  return __shfl(v, src, kWave);
}

/* A ghost lane's value: sign times its mirror lane's. Only strips within
 * three columns of a wall hold ghost lanes (xedge, uniform over the wave),
 * so only they pay for the cross-lane permute. */
template <typename T>
__device__ __forceinline__ T mirror_fix(bool xedge, T v, int src, bool ghost, T sign) {
  if (!xedge) return v;
  const T m = from_lane(v, src);
  return ghost ? sign * m : v;
}

/* lane holding the mirror of index q (lane holds q), kept inside the wave */
__device__ __forceinline__ int mirror_lane(int lane, int q, int n) {
  return min(max(lane + (mirror(q, n) - q), 0), kWave - 1);
}

/* x momentum, odd tt (omp3/flow.c:570-619): x flux in x, u recompute, then
 * x flux in y. Lane l holds x face / cell column q = x0 - 2 + l. Per row k
 * the cell fluxes, the first sub-update and u1 are computed two rows ahead
 * of the output row r, which takes the corner fluxes at rows r and r+1 from
 * a window of u1 rows r-1 .. r+2. Ghost rows are the mirror row's values
 * (u is copied unsigned across the south/north walls). */
template <typename T, bool P2>
__device__ __forceinline__ void momx_xy_item(const MGeo<T>& g, const MMomIO<T>& io, T dt,
                                             T dt_h, const MarchList& ml, int item, int lane) {
  const int nx = g.nx, ny = g.ny, sx = nx + 1;
  (void)dt;
  (void)sx;
  const ItemBox box = item_box(ml, nx, item);
  const int x0 = box.x0, x1 = box.x1, y0 = box.y0, y1 = box.y1;
  if (y0 >= y1) return;
  const bool edge = (x0 == kPad) || (x1 + 3 > nx - kPad) || (y0 == kPad) ||
                    (y1 + 3 > ny - kPad);
  const bool xedge = (x0 - 3 < kPad) || (x1 + 3 > nx - kPad);
  const int q = x0 - 2 + lane;
  const int qf = min(q, nx), qc = min(q, nx - 1), qg = min(max(q, 1), nx);
  const bool cell_ghost = is_ghost(q, nx);
  const int mc_lane = mirror_lane(lane, q, nx);
  const bool own = q >= x0 && (q < x1 || (q == x1 && x1 == nx - kPad));
  const T edx_q = g.edx[qg], cdx_q = g.cdx[qg], cdx_qm = g.cdx[qg - 1];
  const T iedx_q = g.iedx[qg];
  const RowGeo<T> edy(g.edy, y0 - 4, ny, lane), cdy(g.cdy, y0 - 4, ny, lane);
  const RowGeo<T> iedy(g.iedy, y0 - 4, ny, lane);
  /* reciprocal spacings, used in place of quotients on power-of-two meshes */
  const T rcdx_q = T(1.0) / cdx_q;
  const RowGeo<T> rcdy(g.cdy, y0 - 4, ny, lane, true);

  /* u1 and m1 of row i from that row's u0, F_x, mom1 and rho */
  auto xrow = [&](int i, T a_u, T a_fx, T a_m1, T a_r, T& m1, T& u1) {
    const T up1 = dn(a_u, 1);
    T ufx = momentum_flux(edy(i), a_fx, dn(a_fx, 1), up(a_u, 1), a_u, up1,
                          dn(a_u, 2), iedx_q, T(0.5) * (a_u + up1), cdx_qm,
                          cdx_q, dt_h);
    ufx = mirror_fix(xedge, ufx, mc_lane, cell_ghost, T(1.0));
    m1 = advect_update<T, P2>(a_m1, dt_h, ufx, up(ufx, 1), edx_q, cdy(i),
                              iedx_q * rcdy(i));
    const T rex = rho_edge_x<T, P2>(a_r, up(a_r, 1), cdx_q, cdx_qm, cdy(i), edx_q,
                                    T(0.5) * iedx_q * rcdy(i));
    u1 = velocity(m1, rex);
  };
  /* x flux in y at corner row gf from u1 rows gf-2 .. gf+1 */
  auto corner = [&](int gf, T w0, T w1, T w2, T w3, T a_v, T a_fy) {
    const T v_cell = T(0.5) * (up(a_v, 1) + a_v);
    return momentum_flux(edx_q, a_fy, up(a_fy, 1), w0, w1, w2, w3, iedy(gf),
                         v_cell, cdy(gf - 1), cdy(gf), dt_h);
  };

  T wu[4], wm[3];
#pragma unroll
  for (int k = 0; k < 4; ++k) {
    const int i = mirror(y0 - 2 + k, ny);
    T m1, u1;
    xrow(i, io.u[i * sx + qf], io.F_x[i * sx + qf], io.mom1[i * sx + qf],
         io.rho[i * nx + qc], m1, u1);
    wu[k] = u1;
    if (k >= 2) wm[k - 2] = m1;
  }
  T uf_prev = corner(y0, wu[0], wu[1], wu[2], wu[3], io.v0[y0 * nx + qc],
                     io.F_y[y0 * nx + qc]);
#pragma unroll
  for (int k = 0; k < 3; ++k) wu[k] = wu[k + 1];
  int ia = mirror(y0 + 2, ny);
  T a_u = io.u[ia * sx + qf], a_fx = io.F_x[ia * sx + qf];
  T a_m1 = io.mom1[ia * sx + qf], a_r = io.rho[ia * nx + qc];
  T a_v = io.v0[(y0 + 1) * nx + qc], a_fy = io.F_y[(y0 + 1) * nx + qc];
  for (int r = y0; r < y1; ++r) {
    const T c_u = a_u, c_fx = a_fx, c_m1 = a_m1, c_r = a_r, c_v = a_v, c_fy = a_fy;
    {
      const int in = mirror(r + 3, ny), rv = min(r + 2, ny);
      a_u = io.u[in * sx + qf];
      a_fx = io.F_x[in * sx + qf];
      a_m1 = io.mom1[in * sx + qf];
      a_r = io.rho[in * nx + qc];
      a_v = io.v0[rv * nx + qc];
      a_fy = io.F_y[rv * nx + qc];
    }
    T m1, u1;
    xrow(mirror(r + 2, ny), c_u, c_fx, c_m1, c_r, m1, u1);
    wu[3] = u1;
    wm[2] = m1;
    const T uf = corner(r + 1, wu[0], wu[1], wu[2], wu[3], c_v, c_fy);
    if (own) {
      io.mom[r * sx + q] = advect_update<T, P2>(wm[0], dt_h, uf, uf_prev, cdx_q, edy(r),
                                                rcdx_q * iedy(r));
      mput_xface(edge, io.vel, nx, ny, r, q, wu[1], T(-1.0));
    }
    uf_prev = uf;
#pragma unroll
    for (int k = 0; k < 3; ++k) wu[k] = wu[k + 1];
    wm[0] = wm[1];
    wm[1] = wm[2];
  }
}

template <typename T, bool P2>
__global__ __launch_bounds__(kMNT, FLOWHIP_MARCH_WAVES) void k_march_momx_xy(
    MGeo<T> g, MMomIO<T> io, const double* dtp, MarchList ml) {
  const T dt = T(dtp[0]), dt_h = T(dtp[1]);
  const int lane = threadIdx.x % kWave;
  const MarchWalk mw = march_walk<false>(ml.nitems);
  for (int item = mw.first; item < mw.end; item += mw.stride) {
    momx_xy_item<T, P2>(g, io, dt, dt_h, ml, item, lane);
  }
}

/* x momentum, even tt (omp3/flow.c:673-722): x flux in y, u recompute, then
 * x flux in x. Lane l holds x face / cell column q = x0 - 3 + l (strips of
 * 59 columns, so faces [x0-2, x1+1] and the cells between them fit the
 * wave). Output row r takes the corner fluxes at rows r and r+1, the second
 * from a window of u0 rows r-1 .. r+2; ghost faces hold the negated mirror
 * face's u1 (INVERT_X) and ghost cells the mirror cell's flux. */
template <typename T, bool P2>
__device__ __forceinline__ void momx_yx_item(const MGeo<T>& g, const MMomIO<T>& io, T dt,
                                             T dt_h, const MarchList& ml, int item, int lane) {
  const int nx = g.nx, ny = g.ny, sx = nx + 1;
  (void)dt;
  (void)sx;
  const ItemBox box = item_box(ml, nx, item);
  const int x0 = box.x0, x1 = box.x1, y0 = box.y0, y1 = box.y1;
  if (y0 >= y1) return;
  const bool edge = (x0 == kPad) || (x1 + 3 > nx - kPad) || (y0 == kPad) ||
                    (y1 + 3 > ny - kPad);
  const bool xedge = (x0 - 3 < kPad) || (x1 + 3 > nx - kPad);
  const int q = x0 - 3 + lane;
  const int qf = min(max(q, 0), nx), qc = min(max(q, 0), nx - 1);
  const int qg = min(max(q, 1), nx);
  const bool face_ghost = is_ghost(q, sx), cell_ghost = is_ghost(q, nx);
  const int mf_lane = mirror_lane(lane, q, sx), mc_lane = mirror_lane(lane, q, nx);
  const bool own = q >= x0 && (q < x1 || (q == x1 && x1 == nx - kPad));
  const T edx_q = g.edx[qg], cdx_q = g.cdx[qg], cdx_qm = g.cdx[qg - 1];
  const T iedx_q = g.iedx[qg];
  const RowGeo<T> edy(g.edy, y0 - 4, ny, lane), cdy(g.cdy, y0 - 4, ny, lane);
  const RowGeo<T> iedy(g.iedy, y0 - 4, ny, lane);
  /* reciprocal spacings, used in place of quotients on power-of-two meshes */
  const T rcdx_q = T(1.0) / cdx_q;
  const RowGeo<T> rcdy(g.cdy, y0 - 4, ny, lane, true);

  auto corner = [&](int gf, T w0, T w1, T w2, T w3, T a_v, T a_fy) {
    const T v_cell = T(0.5) * (up(a_v, 1) + a_v);
    return momentum_flux(edx_q, a_fy, up(a_fy, 1), w0, w1, w2, w3, iedy(gf),
                         v_cell, cdy(gf - 1), cdy(gf), dt_h);
  };

  /* u0 rows r-1 .. r+2 of face q (ghost rows are in memory) */
  T wu[4];
#pragma unroll
  for (int k = 0; k < 4; ++k) wu[k] = io.u[(y0 - 2 + k) * sx + qf];
  T uf_prev = corner(y0, wu[0], wu[1], wu[2], wu[3], io.v0[y0 * nx + qc],
                     io.F_y[y0 * nx + qc]);
#pragma unroll
  for (int k = 0; k < 3; ++k) wu[k] = wu[k + 1];
  T a_u = io.u[(y0 + 2) * sx + qf];
  T a_v = io.v0[(y0 + 1) * nx + qc], a_fy = io.F_y[(y0 + 1) * nx + qc];
  T a_m1 = io.mom1[y0 * sx + qf], a_r = io.rho[y0 * nx + qc];
  T a_fx = io.F_x[y0 * sx + qf];
  for (int r = y0; r < y1; ++r) {
    wu[3] = a_u;
    const T c_v = a_v, c_fy = a_fy, c_m1 = a_m1, c_r = a_r, c_fx = a_fx;
    {
      const int ru = min(r + 3, ny - 1), rv = min(r + 2, ny), rc = min(r + 1, ny - 1);
      a_u = io.u[ru * sx + qf];
      a_v = io.v0[rv * nx + qc];
      a_fy = io.F_y[rv * nx + qc];
      a_m1 = io.mom1[rc * sx + qf];
      a_r = io.rho[rc * nx + qc];
      a_fx = io.F_x[rc * sx + qf];
    }
    const T uf = corner(r + 1, wu[0], wu[1], wu[2], wu[3], c_v, c_fy);
    const T m1 = advect_update<T, P2>(c_m1, dt_h, uf, uf_prev, cdx_q, edy(r),
                                      rcdx_q * iedy(r));
    const T rex = rho_edge_x<T, P2>(c_r, up(c_r, 1), cdx_q, cdx_qm, cdy(r), edx_q,
                                    T(0.5) * iedx_q * rcdy(r));
    T u1 = velocity(m1, rex);
    u1 = mirror_fix(xedge, u1, mf_lane, face_ghost, T(-1.0));
    const T up1 = dn(u1, 1);
    T ufx = momentum_flux(edy(r), c_fx, dn(c_fx, 1), up(u1, 1), u1, up1, dn(u1, 2),
                          iedx_q, T(0.5) * (u1 + up1), cdx_qm, cdx_q, dt_h);
    ufx = mirror_fix(xedge, ufx, mc_lane, cell_ghost, T(1.0));
    const T ufx_lo = up(ufx, 1);
    if (own) {
      io.mom[r * sx + q] = advect_update<T, P2>(m1, dt_h, ufx, ufx_lo, edx_q, cdy(r),
                                                iedx_q * rcdy(r));
      mput_xface(edge, io.vel, nx, ny, r, q, u1, T(-1.0));
    }
    uf_prev = uf;
#pragma unroll
    for (int k = 0; k < 3; ++k) wu[k] = wu[k + 1];
  }
}

template <typename T, bool P2>
__global__ __launch_bounds__(kMNT, FLOWHIP_MARCH_WAVES) void k_march_momx_yx(
    MGeo<T> g, MMomIO<T> io, const double* dtp, MarchList ml) {
  const T dt = T(dtp[0]), dt_h = T(dtp[1]);
  const int lane = threadIdx.x % kWave;
  const MarchWalk mw = march_walk<false>(ml.nitems);
  for (int item = mw.first; item < mw.end; item += mw.stride) {
    momx_yx_item<T, P2>(g, io, dt, dt_h, ml, item, lane);
  }
}

/* y momentum, odd tt (omp3/flow.c:621-669): y flux in x, v recompute, then
 * y flux in y. Lane l holds column q = x0 - 2 + l. Per y-face row k (two
 * rows ahead of the output row g) the corner fluxes along the row, the first
 * sub-update and v1; ghost y faces past the north wall are the negated
 * mirror (INVERT_Y). The cell fluxes at rows g-1 and g come from a window
 * of v1 rows g-1 .. g+2; a ghost cell row's flux is its mirror row's. io.u
 * is u after the x sub-updates. */
template <typename T, bool P2>
__device__ __forceinline__ void momy_xy_item(const MGeo<T>& g, const MMomIO<T>& io, T dt,
                                             T dt_h, const MarchList& ml, int item, int lane) {
  const int nx = g.nx, ny = g.ny, sx = nx + 1;
  (void)dt;
  (void)sx;
  const ItemBox box = item_box(ml, nx, item);
  const int x0 = box.x0, x1 = box.x1, y0 = box.y0, y1 = box.y1;
  if (y0 >= y1) return;
  const bool edge = (x0 == kPad) || (x1 + 3 > nx - kPad) || (y0 == kPad) ||
                    (y1 + 3 > ny - kPad);
  const int gend = (y1 == ny - kPad) ? y1 + 1 : y1;
  const int q = x0 - 2 + lane;
  const int qf = min(q, nx), qc = min(q, nx - 1), qg = min(max(q, 1), nx);
  const bool own = q >= x0 && q < x1;
  const T edx_q = g.edx[qg], cdx_q = g.cdx[qg], cdx_qm = g.cdx[qg - 1];
  const T iedx_q = g.iedx[qg];
  const RowGeo<T> edy(g.edy, y0 - 4, ny, lane), cdy(g.cdy, y0 - 4, ny, lane);
  const RowGeo<T> iedy(g.iedy, y0 - 4, ny, lane);
  /* reciprocal spacings, used in place of quotients on power-of-two meshes */
  const T rcdx_q = T(1.0) / cdx_q;
  const RowGeo<T> rcdy(g.cdy, y0 - 4, ny, lane, true);

  /* v1 and m1 of interior y-face row k from row k's v0, mom1 and rows
   * k, k-1 of F_x, u and rho */
  auto yrow = [&](int k, T a_v, T a_m1, T a_fx, T p_fx, T a_u, T p_u, T a_r,
                  T p_r, T& m1, T& v1) {
    const T vm1 = up(a_v, 1);
    const T vfx = momentum_flux(cdy(k), a_fx, p_fx, up(a_v, 2), vm1, a_v,
                                dn(a_v, 1), iedx_q, T(0.5) * (a_u + p_u),
                                cdx_qm, cdx_q, dt_h);
    m1 = advect_update<T, P2>(a_m1, dt_h, dn(vfx, 1), vfx, edx_q, cdy(k),
                              iedx_q * rcdy(k));
    const T rey = rho_edge_y<T, P2>(a_r, p_r, cdx_q, cdy(k), cdy(k - 1), edy(k),
                                    T(0.5) * rcdx_q * iedy(k));
    v1 = velocity(m1, rey);
  };
  /* y flux in y at cell row i from v1 rows i-1 .. i+2 */
  auto cellflux = [&](int i, T w0, T w1, T w2, T w3, T fy_lo, T fy_hi) {
    return momentum_flux(cdx_q, fy_lo, fy_hi, w0, w1, w2, w3, iedy(i),
                         T(0.5) * (w1 + w2), cdy(i - 1), cdy(i), dt_h);
  };

  T wv[4], wm[3];
  wm[0] = wm[1] = wm[2] = T(0);
  /* rows y0-2 .. y0+1; below the south wall, the negated mirror rows */
  T p_fx = T(0), p_u = T(0), p_r = T(0);
  {
    const int k0 = max(y0 - 2, kPad);
    p_fx = io.F_x[(k0 - 1) * sx + qf];
    p_u = io.u[(k0 - 1) * sx + qf];
    p_r = io.rho[(k0 - 1) * nx + qc];
#pragma unroll
    for (int s = 0; s < 4; ++s) {
      const int k = y0 - 2 + s;
      if (k < kPad) continue;
      const T c_fx = io.F_x[k * sx + qf], c_u = io.u[k * sx + qf];
      const T c_r = io.rho[k * nx + qc];
      T m1, v1;
      yrow(k, io.v0[k * nx + qc], io.mom1[k * nx + qc], c_fx, p_fx, c_u, p_u,
           c_r, p_r, m1, v1);
      wv[s] = v1;
      if (s >= 2) wm[s - 2] = m1;
      p_fx = c_fx;
      p_u = c_u;
      p_r = c_r;
    }
  }
  const bool south = (y0 == kPad);
  if (south) {
    wv[0] = -wv[3];
    wv[1] = -wv[2];
  }
  T p_fy = io.F_y[y0 * nx + qc];
  T vfy_prev = south ? T(0)
                     : cellflux(y0 - 1, wv[0], wv[1], wv[2], wv[3],
                                io.F_y[(y0 - 1) * nx + qc], p_fy);
#pragma unroll
  for (int s = 0; s < 3; ++s) wv[s] = wv[s + 1];
  /* lookahead row y0+2 and the F_y row g+1 */
  int kn = min(y0 + 2, ny - 1);
  T a_v = io.v0[(y0 + 2) * nx + qc], a_m1 = io.mom1[(y0 + 2) * nx + qc];
  T a_fx = io.F_x[kn * sx + qf], a_u = io.u[kn * sx + qf], a_r = io.rho[kn * nx + qc];
  T a_fy = io.F_y[(y0 + 1) * nx + qc];
  for (int gr = y0; gr < gend; ++gr) {
    const int k = gr + 2;
    const T c_v = a_v, c_m1 = a_m1, c_fx = a_fx, c_u = a_u, c_r = a_r, c_fy = a_fy;
    {
      const int ky = min(k + 1, ny), kx = min(k + 1, ny - 1);
      a_v = io.v0[ky * nx + qc];
      a_m1 = io.mom1[ky * nx + qc];
      a_fx = io.F_x[kx * sx + qf];
      a_u = io.u[kx * sx + qf];
      a_r = io.rho[kx * nx + qc];
      a_fy = io.F_y[min(gr + 2, ny) * nx + qc];
    }
    T m1 = T(0), v1;
    if (k < ny - 1) {
      yrow(k, c_v, c_m1, c_fx, p_fx, c_u, p_u, c_r, p_r, m1, v1);
    } else {
      /* ghost y face: mirror of face ny-1 is ny-2 (k-1), of ny is ny-3 (k-3) */
      v1 = -((k == ny - 1) ? wv[2] : wv[0]);
    }
    wv[3] = v1;
    wm[2] = m1;
    T vfy = (gr >= ny - kPad)
                ? vfy_prev
                : cellflux(gr, wv[0], wv[1], wv[2], wv[3], p_fy, c_fy);
    if (south && gr == y0) vfy_prev = vfy;
    if (own) {
      io.mom[gr * nx + q] = advect_update<T, P2>(wm[0], dt_h, vfy, vfy_prev, cdx_q,
                                                 edy(gr), rcdx_q * iedy(gr));
      mput_yface(edge, io.vel, nx, ny, gr, q, wv[1], T(-1.0));
    }
    vfy_prev = vfy;
    p_fx = c_fx;
    p_u = c_u;
    p_r = c_r;
    p_fy = c_fy;
#pragma unroll
    for (int s = 0; s < 3; ++s) wv[s] = wv[s + 1];
    wm[0] = wm[1];
    wm[1] = wm[2];
  }
}

template <typename T, bool P2>
__global__ __launch_bounds__(kMNT, FLOWHIP_MARCH_WAVES) void k_march_momy_xy(
    MGeo<T> g, MMomIO<T> io, const double* dtp, MarchList ml) {
  const T dt = T(dtp[0]), dt_h = T(dtp[1]);
  const int lane = threadIdx.x % kWave;
  const MarchWalk mw = march_walk<false>(ml.nitems);
  for (int item = mw.first; item < mw.end; item += mw.stride) {
    momy_xy_item<T, P2>(g, io, dt, dt_h, ml, item, lane);
  }
}

/* y momentum, even tt (omp3/flow.c:724-774): y flux in y, v recompute, then
 * y flux in x. Lane l holds column q = x0 - 2 + l. Output y-face row g takes
 * the cell fluxes at rows g-1 and g (the second from a window of v0 rows
 * g-1 .. g+2; a ghost cell row's flux is its mirror row's), then v1 along the
 * row (ghost columns: the mirror column, unsigned) and the corner fluxes
 * across the lanes. io.u is u after the x sub-updates. */
template <typename T, bool P2>
__device__ __forceinline__ void momy_yx_item(const MGeo<T>& g, const MMomIO<T>& io, T dt,
                                             T dt_h, const MarchList& ml, int item, int lane) {
  const int nx = g.nx, ny = g.ny, sx = nx + 1;
  (void)dt;
  (void)sx;
  const ItemBox box = item_box(ml, nx, item);
  const int x0 = box.x0, x1 = box.x1, y0 = box.y0, y1 = box.y1;
  if (y0 >= y1) return;
  const bool edge = (x0 == kPad) || (x1 + 3 > nx - kPad) || (y0 == kPad) ||
                    (y1 + 3 > ny - kPad);
  const bool xedge = (x0 - 3 < kPad) || (x1 + 3 > nx - kPad);
  const int gend = (y1 == ny - kPad) ? y1 + 1 : y1;
  const int q = x0 - 2 + lane;
  const int qf = min(q, nx), qc = min(q, nx - 1), qg = min(max(q, 1), nx);
  const bool col_ghost = is_ghost(q, nx);
  const int mc_lane = mirror_lane(lane, q, nx);
  const bool own = q >= x0 && q < x1;
  const T edx_q = g.edx[qg], cdx_q = g.cdx[qg], cdx_qm = g.cdx[qg - 1];
  const T iedx_q = g.iedx[qg];
  const RowGeo<T> edy(g.edy, y0 - 4, ny, lane), cdy(g.cdy, y0 - 4, ny, lane);
  const RowGeo<T> iedy(g.iedy, y0 - 4, ny, lane);
  /* reciprocal spacings, used in place of quotients on power-of-two meshes */
  const T rcdx_q = T(1.0) / cdx_q;
  const RowGeo<T> rcdy(g.cdy, y0 - 4, ny, lane, true);

  auto cellflux = [&](int i, T w0, T w1, T w2, T w3, T fy_lo, T fy_hi) {
    return momentum_flux(cdx_q, fy_lo, fy_hi, w0, w1, w2, w3, iedy(i),
                         T(0.5) * (w1 + w2), cdy(i - 1), cdy(i), dt_h);
  };

  /* v0 rows g-1 .. g+2 (ghost rows are in memory) */
  T wv[4];
#pragma unroll
  for (int s = 0; s < 4; ++s) wv[s] = io.v0[(y0 - 2 + s) * nx + qc];
  T p_fy = io.F_y[y0 * nx + qc];
  const bool south = (y0 == kPad);
  T vfy_prev = south ? T(0)
                     : cellflux(y0 - 1, wv[0], wv[1], wv[2], wv[3],
                                io.F_y[(y0 - 1) * nx + qc], p_fy);
#pragma unroll
  for (int s = 0; s < 3; ++s) wv[s] = wv[s + 1];
  T p_fx = io.F_x[(y0 - 1) * sx + qf], p_u = io.u[(y0 - 1) * sx + qf];
  T p_r = io.rho[(y0 - 1) * nx + qc];
  T a_v = io.v0[(y0 + 2) * nx + qc], a_fy = io.F_y[(y0 + 1) * nx + qc];
  T a_fx = io.F_x[y0 * sx + qf], a_u = io.u[y0 * sx + qf];
  T a_r = io.rho[y0 * nx + qc], a_m1 = io.mom1[y0 * nx + qc];
  for (int gr = y0; gr < gend; ++gr) {
    wv[3] = a_v;
    const T c_fy = a_fy, c_fx = a_fx, c_u = a_u, c_r = a_r, c_m1 = a_m1;
    {
      const int rx = min(gr + 1, ny - 1);
      a_v = io.v0[min(gr + 3, ny) * nx + qc];
      a_fy = io.F_y[min(gr + 2, ny) * nx + qc];
      a_fx = io.F_x[rx * sx + qf];
      a_u = io.u[rx * sx + qf];
      a_r = io.rho[rx * nx + qc];
      a_m1 = io.mom1[min(gr + 1, ny) * nx + qc];
    }
    T vfy = (gr >= ny - kPad)
                ? vfy_prev
                : cellflux(gr, wv[0], wv[1], wv[2], wv[3], p_fy, c_fy);
    if (south && gr == y0) vfy_prev = vfy;
    const T m1 = advect_update<T, P2>(c_m1, dt_h, vfy, vfy_prev, cdx_q, edy(gr),
                                      rcdx_q * iedy(gr));
    const T rey = rho_edge_y<T, P2>(c_r, p_r, cdx_q, cdy(gr), cdy(gr - 1), edy(gr),
                                    T(0.5) * rcdx_q * iedy(gr));
    T v1 = velocity(m1, rey);
    v1 = mirror_fix(xedge, v1, mc_lane, col_ghost, T(1.0));
    const T vm1 = up(v1, 1);
    const T vfx = momentum_flux(cdy(gr), c_fx, p_fx, up(v1, 2), vm1, v1, dn(v1, 1),
                                iedx_q, T(0.5) * (c_u + p_u), cdx_qm, cdx_q, dt_h);
    const T vfx_hi = dn(vfx, 1);
    if (own) {
      io.mom[gr * nx + q] = advect_update<T, P2>(m1, dt_h, vfx_hi, vfx, edx_q, cdy(gr),
                                                 iedx_q * rcdy(gr));
      mput_yface(edge, io.vel, nx, ny, gr, q, v1, T(-1.0));
    }
    vfy_prev = vfy;
    p_fy = c_fy;
    p_fx = c_fx;
    p_u = c_u;
    p_r = c_r;
#pragma unroll
    for (int s = 0; s < 3; ++s) wv[s] = wv[s + 1];
  }
}

template <typename T, bool P2>
__global__ __launch_bounds__(kMNT, FLOWHIP_MARCH_WAVES) void k_march_momy_yx(
    MGeo<T> g, MMomIO<T> io, const double* dtp, MarchList ml) {
  const T dt = T(dtp[0]), dt_h = T(dtp[1]);
  const int lane = threadIdx.x % kWave;
  const MarchWalk mw = march_walk<false>(ml.nitems);
  for (int item = mw.first; item < mw.end; item += mw.stride) {
    momy_yx_item<T, P2>(g, io, dt, dt_h, ml, item, lane);
  }
}

/* ======================================================================= */
/* Lagrangian step                                                          */
/* ======================================================================= */

template <typename T>
struct MLagIO {
  const T* __restrict__ rho;
  const T* __restrict__ e;
  const T* __restrict__ mom_x;
  const T* __restrict__ mom_y;
  T* __restrict__ mx1;     /* x momentum after the viscous step */
  T* __restrict__ my1;     /* y momentum after the viscous step */
  T* __restrict__ u0;      /* x velocity after the viscous step, + ghosts */
  T* __restrict__ v0;      /* y velocity after the viscous step, + ghosts */
  T* __restrict__ e1;      /* energy after shock heating, + ghosts */
  T* __restrict__ rho_old; /* copy of rho over every cell */
  double* __restrict__ cfl; /* NULL, or every interior CFL candidate (FP64) */
};

/* EOS, pressure acceleration, artificial viscosity, shock heating, the CFL
 * candidates and the copy of rho (omp3/flow.c:55-70, 110-290, 73-96,
 * 302-305). Lane l holds x face / cell column q = x0 - 3 + l (strips of 58
 * columns: shock heating on cells [x0, x1) needs u after the viscous step on
 * faces [x0, x1], hence Qxx on cells [x0-1, x1], u after pressure
 * acceleration on faces [x0-2, x1+2] and the pressure one cell further
 * west). Output row i
 * takes, along x, u after pressure acceleration on the row's faces, Qxx on
 * its cells and u after the viscous acceleration; along y, the lane's
 * column carries v after pressure acceleration on y faces i .. i+3 (face
 * i+3 computed this step), Qyy on cell rows i and i+1 and v after the
 * viscous acceleration on faces i and i+1. Ghost faces hold the negated
 * mirror (INVERT_X / INVERT_Y), ghost Q cells the mirror cell.
 * EDGE = false is the interior form, for items that item_edge clears: no
 * ghost store, mirror or wall test is compiled, so each row is one
 * straight-line block. */
template <typename T, bool P2, int XE = -1, bool EDGE = true, bool UNI = false>
__device__ __forceinline__ void lagrangian_box(const MGeo<T>& g, const MLagIO<T>& io, T dt,
                                               T dt_h, const ItemBox box, int lane,
                                               double& local_min, int& nan_count) {
  const int nx = g.nx, ny = g.ny, sx = nx + 1;
  (void)dt;
  (void)sx;
  const int x0 = box.x0, x1 = box.x1, y0 = box.y0, y1 = box.y1;
  if (y0 >= y1) return;
  const bool edge = EDGE && ((x0 == kPad) || (x1 + 3 > nx - kPad) || (y0 == kPad) ||
                             (y1 + 3 > ny - kPad));
  /* strips within three columns of a wall hold ghost lanes; XE fixes that at
   * compile time (0: none, 1: some), -1 tests it here */
  const bool xedge =
      EDGE && (XE >= 0 ? XE != 0 : ((x0 - 3 < kPad) || (x1 + 3 > nx - kPad)));
  const bool last_chunk = EDGE && (y1 == ny - kPad);
  const int q = x0 - 3 + lane;
  const int qf = min(max(q, 0), nx), qc = min(max(q, 0), nx - 1);
  const int qg = min(max(q, 1), nx);
  const bool face_ghost = is_ghost(q, sx), cell_ghost = is_ghost(q, nx);
  const int mf_lane = mirror_lane(lane, q, sx), mc_lane = mirror_lane(lane, q, nx);
  const bool own_face = q >= x0 && (q < x1 || (q == x1 && x1 == nx - kPad));
  const bool own = q >= x0 && q < x1;
  const T edx_q = g.edx[qg], cdx_q = g.cdx[qg], cdx_qm = g.cdx[qg - 1];
  const RowS<T, UNI> edy(g.edy, y0 - 4, ny, lane), cdy(g.cdy, y0 - 4, ny, lane);
  /* reciprocal spacings, used in place of quotients on power-of-two meshes */
  const T iedx_q = g.iedx[qg], rcdx_q = T(1.0) / cdx_q;
  const RowS<T, UNI> iedy(g.iedy, y0 - 4, ny, lane);
  /* lane-held reciprocals in FP32: 2% faster there than RowRecip */
  using Recip = typename std::conditional<
      UNI, RowConstRecip<T, P2>,
      typename std::conditional<std::is_same<T, float>::value, RowGeo<T>,
                                RowRecip<T, P2>>::type>::type;
  const Recip rcdy(g.cdy, y0 - 4, ny, lane, true);

  /* y face k of the lane's column after pressure acceleration, from cell
   * rows k (rk, ek) and k-1 (rs, es) and mom_y at face k */
  auto ypa = [&](int k, T rk, T ek, T rs, T es, T my, T& mom, T& rey) {
    mom = accelerate<T, P2>(my, dt, eos_pressure(rk, ek), eos_pressure(rs, es), edy(k),
                            iedy(k));
    rey = rho_edge_y<T, P2>(rk, rs, cdx_q, cdy(k), cdy(k - 1), edy(k),
                            T(0.5) * rcdx_q * iedy(k));
    return velocity(mom, rey);
  };

  /* ---- prologue: y faces y0-2 .. y0+2 and v after the viscous step at y0 */
  T wr[4], we[4];          /* rho, e at cell rows i .. i+3 */
  T vp[5], pm[5], pr[5];   /* y faces y0-2 .. y0+2: v, mom, rho edge */
  T qy_prev, qy_m1;        /* Qyy at cell rows i and i-1 */
  {
    T rr[6], ee[6];        /* cell rows y0-3 .. y0+2 */
#pragma unroll
    for (int s = 0; s < 6; ++s) {
      const int row = max(y0 - 3 + s, 0);
      rr[s] = io.rho[row * nx + qc];
      ee[s] = io.e[row * nx + qc];
    }
#pragma unroll
    for (int s = 0; s < 5; ++s) {
      const int f = y0 - 2 + s;
      pm[s] = pr[s] = T(0);
      if (!EDGE || (f >= kPad && f < ny - 1))
        vp[s] = ypa(f, rr[s + 1], ee[s + 1], rr[s], ee[s], io.mom_y[f * nx + qc],
                    pm[s], pr[s]);
      else if (f == ny - 1)
        vp[s] = -vp[s - 1 > 0 ? s - 1 : 0];
      else if (f == ny)
        vp[s] = -vp[s - 3 > 0 ? s - 3 : 0];
      else
        vp[s] = T(0);
    }
    if (EDGE && y0 == kPad) {
      vp[0] = -vp[3];
      vp[1] = -vp[2];
    }
#pragma unroll
    for (int s = 0; s < 3; ++s) {
      wr[s] = rr[s + 3];
      we[s] = ee[s + 3];
    }
    /* Qyy at cell rows y0 and y0-1 (below the south wall: row y0's) */
    qy_prev = viscous_stress(rr[3], vp[1], vp[2], vp[3], vp[4]);
    qy_m1 = (EDGE && y0 == kPad) ? qy_prev
                                 : viscous_stress(rr[2], vp[0], vp[1], vp[2], vp[3]);
  }
  T vav_prev;
  {
    const T mom = accelerate<T, P2>(pm[2], dt, qy_prev, qy_m1, cdy(y0), rcdy(y0));
    vav_prev = velocity(mom, pr[2]);
    if (own) {
      io.my1[y0 * nx + q] = mom;
      mput_yface(edge, io.v0, nx, ny, y0, q, vav_prev, T(-1.0));
    }
  }
  /* windows entering row y0: v at faces y0 .. y0+2; mom, rho edge at faces
   * y0+1, y0+2 */
  T wv[4] = {vp[2], vp[3], vp[4], T(0)};
  T wm[3] = {pm[3], pm[4], T(0)}, wx[3] = {pr[3], pr[4], T(0)};
  /* row y0+3 and the row-y0 x inputs */
  {
    const int r3 = min(y0 + 3, ny - 1);
    wr[3] = io.rho[r3 * nx + qc];
    we[3] = io.e[r3 * nx + qc];
  }
  T a_my = io.mom_y[min(y0 + 3, ny) * nx + qc];
  T a_mx = io.mom_x[y0 * sx + qf];
  T a_r, a_e;
  {
    const int r4 = min(y0 + 4, ny - 1);
    a_r = io.rho[r4 * nx + qc];
    a_e = io.e[r4 * nx + qc];
  }

  for (int i = y0; i < y1; ++i) {
    const T c_my = a_my, c_mx = a_mx;
    const T n_r = a_r, n_e = a_e;
    /* the next row's momenta while there is a next row; rho and e of row
     * i+5, which row i+2 takes for y face i+5, while there is a row i+2 */
    if (i + 1 < y1) {
      a_my = io.mom_y[min(i + 4, ny) * nx + qc];
      a_mx = io.mom_x[min(i + 1, ny - 1) * sx + qf];
    }
    if (i + 2 < y1) {
      const int r5 = min(i + 5, ny - 1);
      a_r = io.rho[r5 * nx + qc];
      a_e = io.e[r5 * nx + qc];
    }
    /* y face i+3 after pressure acceleration (an interior item's faces all
     * lie below the north ghost faces) */
    const int k = i + 3;
    T vk;
    if (!EDGE || k < ny - 1) {
      vk = ypa(k, wr[3], we[3], wr[2], we[2], c_my, wm[2], wx[2]);
    } else {
      vk = -((k == ny - 1) ? wv[2] : wv[0]);
      wm[2] = wx[2] = T(0);
    }
    wv[3] = vk;
    /* Qyy at cell row i+1 (a ghost row is its mirror, row i) */
    const T qy = (EDGE && i + 1 >= ny - kPad)
                     ? qy_prev
                     : viscous_stress(wr[1], wv[0], wv[1], wv[2], wv[3]);
    /* y face i+1 after the viscous acceleration */
    const T my_av = accelerate<T, P2>(wm[0], dt, qy, qy_prev, cdy(i + 1), rcdy(i + 1));
    const T vav = velocity(my_av, wx[0]);
    if (own && (i + 1 < y1 || last_chunk)) {
      io.my1[(i + 1) * nx + q] = my_av;
      mput_yface(edge, io.v0, nx, ny, i + 1, q, vav, T(-1.0));
    }

    /* along x on row i */
    const T ri = wr[0], ei = we[0];
    const T P = eos_pressure(ri, ei);
    const T mx_pa = accelerate<T, P2>(c_mx, dt, P, up(P, 1), edx_q, iedx_q);
    const T rex = rho_edge_x<T, P2>(ri, up(ri, 1), cdx_q, cdx_qm, cdy(i), edx_q,
                                    T(0.5) * iedx_q * rcdy(i));
    T u_pa = velocity(mx_pa, rex);
    u_pa = mirror_fix(xedge, u_pa, mf_lane, face_ghost, T(-1.0));
    T qx = viscous_stress(ri, up(u_pa, 1), u_pa, dn(u_pa, 1), dn(u_pa, 2));
    qx = mirror_fix(xedge, qx, mc_lane, cell_ghost, T(1.0));
    const T mx_av = accelerate<T, P2>(mx_pa, dt, qx, up(qx, 1), cdx_q, rcdx_q);
    const T uav = velocity(mx_av, rex);
    if (own_face) {
      io.mx1[i * sx + q] = mx_av;
      mput_xface(edge, io.u0, nx, ny, i, q, uav, T(-1.0));
    }

    /* shock heating, CFL candidate and the copy of rho on row i */
    const T uav_e = dn(uav, 1);
    if (own) {
      const int c = i * nx + q;
      const T es = shock_heating<T, P2>(ei, P, ri, qx, qy_prev, uav, uav_e, vav_prev,
                                        vav, cdx_q, cdy(i), dt_h, rcdx_q, rcdy(i));
      mput_cell(edge, io.e1, nx, ny, i, q, es);
      io.rho_old[c] = ri;
      if (edge) {
        const bool w = (q == kPad), ea = (q == nx - kPad - 1);
        const bool s = (i == kPad), no = (i == ny - kPad - 1);
        for (int d = 1; d <= kPad; ++d) {
          if (w) io.rho_old[c - d] = io.rho[c - d];
          if (ea) io.rho_old[c + d] = io.rho[c + d];
          if (s) io.rho_old[c - d * nx] = io.rho[c - d * nx];
          if (no) io.rho_old[c + d * nx] = io.rho[c + d * nx];
          for (int d2 = 1; d2 <= kPad; ++d2) {
            if (s && w) io.rho_old[c - d * nx - d2] = io.rho[c - d * nx - d2];
            if (s && ea) io.rho_old[c - d * nx + d2] = io.rho[c - d * nx + d2];
            if (no && w) io.rho_old[c + d * nx - d2] = io.rho[c + d * nx - d2];
            if (no && ea) io.rho_old[c + d * nx + d2] = io.rho[c + d * nx + d2];
          }
        }
      }
      /* FP64 CFL candidate (squared in the FP32-state form), unless it
       * certainly cannot lower the lane's minimum (and none is dumped) */
      constexpr bool kSq = std::is_same<T, float>::value;
      if (io.cfl || !cfl_above(double(es), double(qx), double(qy_prev), double(ri),
                               double(cdx_q), double(cdy(i)),
                               kSq ? local_min : local_min * local_min)) {
        const double cand = kSq ? cell_timestep_sq(double(es), double(qx), double(qy_prev),
                                                   double(ri), double(cdx_q), double(cdy(i)))
                                : cell_timestep(double(es), double(qx), double(qy_prev),
                                                double(ri), double(cdx_q), double(cdy(i)));
        if (cand == cand)
          local_min = lesser(local_min, cand);
        else
          ++nan_count;
        if (io.cfl) io.cfl[c] = cand;
      }
    }

    /* advance the windows */
    qy_prev = qy;
    vav_prev = vav;
#pragma unroll
    for (int s = 0; s < 3; ++s) {
      wv[s] = wv[s + 1];
      wr[s] = wr[s + 1];
      we[s] = we[s + 1];
    }
    wr[3] = n_r;
    we[3] = n_e;
    wm[0] = wm[1];
    wm[1] = wm[2];
    wx[0] = wx[1];
    wx[1] = wx[2];
  }
}

template <typename T, bool P2, int XE = -1, bool EDGE = true, bool UNI = false>
__device__ __forceinline__ void lagrangian_item(const MGeo<T>& g, const MLagIO<T>& io, T dt,
                                                T dt_h, const MarchList& ml, int item,
                                                int lane, double& local_min, int& nan_count) {
  lagrangian_box<T, P2, XE, EDGE, UNI>(g, io, dt, dt_h, item_box(ml, g.nx, item), lane,
                                       local_min, nan_count);
}

/* ---- the Lagrangian march, two columns per lane (FP32 interior items) ----
 * Lane l holds columns qb + c, c = 0, 1, with qb = x0 - 4 + 2l: strips of
 * 120 columns with four halo columns on each side (the one-column form
 * needs three). Every row access is one 8-byte access per lane. Each
 * column's arithmetic is lagrangian_box's interior form in the same order,
 * so the fields are bitwise the same; every lane shift is taken in
 * wave-uniform control flow. */
constexpr int kLagW = kWave - 6;       /* one column per lane */
constexpr int kLagW2 = 2 * kWave - 8;  /* two */

template <typename T, bool P2, bool UNI>
__device__ __forceinline__ void lagrangian_pair(const MGeo<T>& g, const MLagIO<T>& io, T dt,
                                                T dt_h, const ItemBox box, int lane,
                                                double& local_min, int& nan_count) {
  using P = T[2];
  const int nx = g.nx, ny = g.ny, sx = nx + 1;
  const int x0 = box.x0, x1 = box.x1, y0 = box.y0, y1 = box.y1;
  if (y0 >= y1) return;
  const int qb = x0 - 4 + 2 * lane;
  const bool own = qb >= x0 && qb < x1; /* both columns: x0 and the width even */
  T edx_q[2], cdx_q[2], cdx_qm[2], iedx_q[2], rcdx_q[2];
#pragma unroll
  for (int c = 0; c < 2; ++c) {
    const int q = qb + c;
    edx_q[c] = g.edx[q];
    cdx_q[c] = g.cdx[q];
    cdx_qm[c] = g.cdx[q - 1];
    iedx_q[c] = g.iedx[q];
    rcdx_q[c] = T(1.0) / cdx_q[c];
  }
  const RowS<T, UNI> edy(g.edy, y0 - 4, ny, lane), cdy(g.cdy, y0 - 4, ny, lane);
  const RowS<T, UNI> iedy(g.iedy, y0 - 4, ny, lane);
  using Recip = typename std::conditional<
      UNI, RowConstRecip<T, P2>,
      typename std::conditional<std::is_same<T, float>::value, RowGeo<T>,
                                RowRecip<T, P2>>::type>::type;
  const Recip rcdy(g.cdy, y0 - 4, ny, lane, true);
  auto ldc = [&](const T* a, int row, P& x) { ld_pair(a, (long)row * nx + qb, x); };
  auto ldx = [&](const T* a, int row, P& x) { ld_xpair(a, (long)row * sx + qb, x); };

  /* y face k after pressure acceleration, from cell rows k (rk, ek) and
   * k-1 (rs, es) and mom_y at face k */
  auto ypa = [&](int k, const P& rk, const P& ek, const P& rs, const P& es, const P& my,
                 P& mom, P& rey, P& v) {
#pragma unroll
    for (int c = 0; c < 2; ++c) {
      mom[c] = accelerate<T, P2>(my[c], dt, eos_pressure(rk[c], ek[c]),
                                 eos_pressure(rs[c], es[c]), edy(k), iedy(k));
      rey[c] = rho_edge_y<T, P2>(rk[c], rs[c], cdx_q[c], cdy(k), cdy(k - 1), edy(k),
                                 T(0.5) * rcdx_q[c] * iedy(k));
      v[c] = velocity(mom[c], rey[c]);
    }
  };

  /* ---- prologue: y faces y0-2 .. y0+2 and v after the viscous step at y0 */
  T wr[4][2], we[4][2];
  T vp[5][2], pm[5][2], pr[5][2];
  T qy_prev[2], vav_prev[2];
  {
    T rr[6][2], ee[6][2];
#pragma unroll
    for (int s = 0; s < 6; ++s) {
      ldc(io.rho, y0 - 3 + s, rr[s]);
      ldc(io.e, y0 - 3 + s, ee[s]);
    }
#pragma unroll
    for (int s = 0; s < 5; ++s) {
      T my[2];
      ldc(io.mom_y, y0 - 2 + s, my);
      ypa(y0 - 2 + s, rr[s + 1], ee[s + 1], rr[s], ee[s], my, pm[s], pr[s], vp[s]);
    }
#pragma unroll
    for (int s = 0; s < 3; ++s) {
      copy_pair(wr[s], rr[s + 3]);
      copy_pair(we[s], ee[s + 3]);
    }
    /* Qyy at cell rows y0 and y0-1, v after the viscous step at face y0 */
    T mom0[2];
#pragma unroll
    for (int c = 0; c < 2; ++c) {
      qy_prev[c] = viscous_stress(rr[3][c], vp[1][c], vp[2][c], vp[3][c], vp[4][c]);
      const T qy_m1 = viscous_stress(rr[2][c], vp[0][c], vp[1][c], vp[2][c], vp[3][c]);
      mom0[c] = accelerate<T, P2>(pm[2][c], dt, qy_prev[c], qy_m1, cdy(y0), rcdy(y0));
      vav_prev[c] = velocity(mom0[c], pr[2][c]);
    }
    if (own) {
      st_pair(io.my1, (long)y0 * nx + qb, mom0);
      st_pair(io.v0, (long)y0 * nx + qb, vav_prev);
    }
  }
  /* Rings of four slots, row or face k in slot (k - y0) & 3: rho and e rows
   * i .. i+3, v at y faces i .. i+3, mom and rho edge at faces i+1 .. i+3.
   * Two slots (k - y0) & 1: Qyy of rows i, i+1, v after the viscous step at
   * faces i, i+1, mom_y at faces i+3, i+4, mom_x at rows i, i+1, and the
   * prefetched rho and e rows i+4, i+5 (copied into the row ring when row i
   * leaves it). The row loop is unrolled by four with every slot fixed at
   * compile time. */
  T wv[4][2], wm[4][2], wx[4][2], qyb[2][2], vavb[2][2], myb[2][2], mxb[2][2];
  T pfr[2][2], pfe[2][2];
  copy_pair(wv[0], vp[2]);
  copy_pair(wv[1], vp[3]);
  copy_pair(wv[2], vp[4]);
  copy_pair(wm[1], pm[3]);
  copy_pair(wm[2], pm[4]);
  copy_pair(wx[1], pr[3]);
  copy_pair(wx[2], pr[4]);
  copy_pair(qyb[0], qy_prev);
  copy_pair(vavb[0], vav_prev);
  ldc(io.rho, y0 + 3, wr[3]);
  ldc(io.e, y0 + 3, we[3]);
  ldc(io.mom_y, y0 + 3, myb[1]);
  ldx(io.mom_x, y0, mxb[0]);
  ldc(io.rho, y0 + 4, pfr[0]);
  ldc(io.e, y0 + 4, pfe[0]);

  auto row = [&](auto phase, int i) {
    constexpr int P = decltype(phase)::value;
    constexpr int S0 = P, S1 = (P + 1) & 3, S2 = (P + 2) & 3, S3 = (P + 3) & 3;
    constexpr int A = P & 1, B = (P + 1) & 1;
    /* the next rows (an interior item's rows all lie inside the mesh): the
     * next row's momenta while there is a next row; rho and e of row i+5,
     * which row i+2 takes for y face i+5, while there is a row i+2 (the
     * last row, y1-1, reaches y face y1+2) */
    if (i + 1 < y1) {
      ldc(io.mom_y, i + 4, myb[A]);
      ldx(io.mom_x, i + 1, mxb[B]);
    }
    if (i + 2 < y1) {
      ldc(io.rho, i + 5, pfr[B]);
      ldc(io.e, i + 5, pfe[B]);
    }
    /* y face i+3 after pressure acceleration */
    ypa(i + 3, wr[S3], we[S3], wr[S2], we[S2], myb[B], wm[S3], wx[S3], wv[S3]);
    /* Qyy at cell row i+1, then y face i+1 after the viscous acceleration */
    T my_av[2];
#pragma unroll
    for (int c = 0; c < 2; ++c) {
      qyb[B][c] = viscous_stress(wr[S1][c], wv[S0][c], wv[S1][c], wv[S2][c], wv[S3][c]);
      my_av[c] = accelerate<T, P2>(wm[S1][c], dt, qyb[B][c], qyb[A][c], cdy(i + 1),
                                   rcdy(i + 1));
      vavb[B][c] = velocity(my_av[c], wx[S1][c]);
    }
    if (own && i + 1 < y1) {
      st_pair(io.my1, (long)(i + 1) * nx + qb, my_av);
      st_pair(io.v0, (long)(i + 1) * nx + qb, vavb[B]);
    }

    /* along x on row i */
    const T(&ri)[2] = wr[S0];
    const T(&ei)[2] = we[S0];
    T Pr[2], mx_pa[2], rex[2], u_pa[2];
#pragma unroll
    for (int c = 0; c < 2; ++c) Pr[c] = eos_pressure(ri[c], ei[c]);
#pragma unroll
    for (int c = 0; c < 2; ++c) {
      mx_pa[c] = accelerate<T, P2>(mxb[A][c], dt, Pr[c], pshift(Pr, c, -1), edx_q[c],
                                   iedx_q[c]);
      rex[c] = rho_edge_x<T, P2>(ri[c], pshift(ri, c, -1), cdx_q[c], cdx_qm[c], cdy(i),
                                 edx_q[c], T(0.5) * iedx_q[c] * rcdy(i));
      u_pa[c] = velocity(mx_pa[c], rex[c]);
    }
    T qx[2];
#pragma unroll
    for (int c = 0; c < 2; ++c)
      qx[c] = viscous_stress(ri[c], pshift(u_pa, c, -1), u_pa[c], pshift(u_pa, c, 1),
                             pshift(u_pa, c, 2));
    T mx_av[2], uav[2];
#pragma unroll
    for (int c = 0; c < 2; ++c) {
      mx_av[c] = accelerate<T, P2>(mx_pa[c], dt, qx[c], pshift(qx, c, -1), cdx_q[c],
                                   rcdx_q[c]);
      uav[c] = velocity(mx_av[c], rex[c]);
    }
    const T uav_e[2] = {pshift(uav, 0, 1), pshift(uav, 1, 1)};
    if (own) {
      st_xpair(io.mx1, (long)i * sx + qb, mx_av);
      st_xpair(io.u0, (long)i * sx + qb, uav);
      /* shock heating, CFL candidate and the copy of rho on row i */
      T es[2];
#pragma unroll
      for (int c = 0; c < 2; ++c)
        es[c] = shock_heating<T, P2>(ei[c], Pr[c], ri[c], qx[c], qyb[A][c], uav[c],
                                     uav_e[c], vavb[A][c], vavb[B][c], cdx_q[c], cdy(i),
                                     dt_h, rcdx_q[c], rcdy(i));
      st_pair(io.e1, (long)i * nx + qb, es);
      st_pair(io.rho_old, (long)i * nx + qb, ri);
#pragma unroll
      for (int c = 0; c < 2; ++c) {
        /* FP64 CFL candidate (squared in the FP32-state form), unless it
         * certainly cannot lower the lane's minimum (and none is dumped) */
        constexpr bool kSq = std::is_same<T, float>::value;
        if (io.cfl || !cfl_above(double(es[c]), double(qx[c]), double(qyb[A][c]),
                                 double(ri[c]), double(cdx_q[c]), double(cdy(i)),
                                 kSq ? local_min : local_min * local_min)) {
          const double cand =
              kSq ? cell_timestep_sq(double(es[c]), double(qx[c]), double(qyb[A][c]),
                                     double(ri[c]), double(cdx_q[c]), double(cdy(i)))
                  : cell_timestep(double(es[c]), double(qx[c]), double(qyb[A][c]),
                                  double(ri[c]), double(cdx_q[c]), double(cdy(i)));
          if (cand == cand)
            local_min = lesser(local_min, cand);
          else
            ++nan_count;
          if (io.cfl) io.cfl[(long)i * nx + qb + c] = cand;
        }
      }
    }
    /* row i+4 takes row i's slot */
    copy_pair(wr[S0], pfr[A]);
    copy_pair(we[S0], pfe[A]);
  };
  using P0 = std::integral_constant<int, 0>;
  using P1 = std::integral_constant<int, 1>;
  using P2c = std::integral_constant<int, 2>;
  using P3 = std::integral_constant<int, 3>;
  int i = y0;
  for (; i + 3 < y1; i += 4) {
    row(P0{}, i);
    row(P1{}, i + 1);
    row(P2c{}, i + 2);
    row(P3{}, i + 3);
  }
  if (i < y1) row(P0{}, i);
  if (i + 1 < y1) row(P1{}, i + 1);
  if (i + 2 < y1) row(P2c{}, i + 2);
}

/* Waves per SIMD the register allocation of a Lagrangian march must allow:
 * the FP32 two-column form runs three (its loop, unrolled, would
 * otherwise take 170 VGPRs and leave room for two). */
template <typename T, bool PAIR>
constexpr int kLagMinWaves = (PAIR && sizeof(T) == 4) ? kLagPairWaves<T> : 1;

template <typename T, bool P2, bool UNI = false, bool PAIR = false>
__global__ __launch_bounds__(kMNT) __attribute__((amdgpu_waves_per_eu(kLagMinWaves<T, PAIR>)))
void k_march_lagrangian(MGeo<T> g, MLagIO<T> io, const double* dtp, FkReduce red,
                        double max_dt, int first_step, MarchList ml) {
  const T dt = T(dtp[0]), dt_h = T(dtp[1]);
  const int lane = threadIdx.x % kWave;
  constexpr bool kSq = std::is_same<T, float>::value;
  double local_min = kSq ? max_dt * max_dt : max_dt;
  int nan_count = 0;
  MarchBlockQueue mq(red.queue, ml.nitems);
  const int wave = threadIdx.x / kWave;
  for (int base = mq.next(); base < mq.end; base = mq.next()) {
    const int item = __builtin_amdgcn_readfirstlane(base + wave);
    if (item >= mq.end) continue;
    if constexpr (PAIR) {
      /* interior items two columns per lane; edge items in one-column
       * sub-strips */
      const ItemBox b = item_box(ml, g.nx, item);
      if (!item_edge(ml, g.nx, g.ny, item)) {
        lagrangian_pair<T, P2, UNI>(g, io, dt, dt_h, b, lane, local_min, nan_count);
      } else {
        for (int sx0 = b.x0; sx0 < b.x1; sx0 += kLagW) {
          const ItemBox sb{sx0, min(sx0 + kLagW, b.x1), b.y0, b.y1};
          lagrangian_box<T, P2, -1, true, UNI>(g, io, dt, dt_h, sb, lane, local_min,
                                               nan_count);
        }
      }
    } else if (!item_edge(ml, g.nx, g.ny, item)) {
      lagrangian_item<T, P2, 0, false, UNI>(g, io, dt, dt_h, ml, item, lane, local_min,
                                            nan_count);
    } else if (item_xedge(ml, g.nx, item)) {
      lagrangian_item<T, P2, 1, true, UNI>(g, io, dt, dt_h, ml, item, lane, local_min,
                                           nan_count);
    } else {
      lagrangian_item<T, P2, 0, true, UNI>(g, io, dt, dt_h, ml, item, lane, local_min,
                                           nan_count);
    }
  }
  mq.finish();
  finish_timestep<kMNT, kSq>(local_min, nan_count, red, max_dt, first_step);
}

/* ======================================================================= */
/* momentum, both directions in one pass                                    */
/* ======================================================================= */

template <typename T>
struct MMom2IO {
  const T* __restrict__ u0;   /* x velocity after the viscous step */
  const T* __restrict__ v0;   /* y velocity after the viscous step */
  const T* __restrict__ F_x;
  const T* __restrict__ F_y;
  const T* __restrict__ rho;  /* advected density */
  const T* __restrict__ mx1;  /* momenta after the viscous step */
  const T* __restrict__ my1;
  T* __restrict__ mom_x;      /* advected momenta */
  T* __restrict__ mom_y;
  T* __restrict__ u;          /* recomputed velocities, + ghosts */
  T* __restrict__ v;
};

/* Odd tt (omp3/flow.c:570-669): x momentum (x flux in x, u recompute, x flux
 * in y) and y momentum (y flux in x, v recompute, y flux in y) of the same
 * strip in one march. The y stage reads u after the x stage, which this
 * wave has just computed for its lanes: row k's u, two rows ahead of the
 * output row r, feeds both the x stage's corner window and the y stage's
 * corner fluxes. Lane l holds x face / column q = x0 - 2 + l (strips of 59
 * columns: the y stage's corner flux at face x1 needs u at x1, hence the x
 * stage's cell flux at x1 and u0 up to x1+2). The per-row arithmetic is the
 * two separate kernels', so the results are theirs bit for bit. */
template <typename T, bool P2, int XE = -1, bool EDGE = true, bool UNI = false>
__device__ __forceinline__ void momxy_xy_item(const MGeo<T>& g, const MMom2IO<T>& io,
                                              T dt, T dt_h, const MarchList& ml,
                                              int item, int lane) {
  const int nx = g.nx, ny = g.ny, sx = nx + 1;
  (void)dt;
  const ItemBox box = item_box(ml, nx, item);
  const int x0 = box.x0, x1 = box.x1, y0 = box.y0, y1 = box.y1;
  if (y0 >= y1) return;
  const bool edge = EDGE && ((x0 == kPad) || (x1 + 3 > nx - kPad) || (y0 == kPad) ||
                             (y1 + 3 > ny - kPad));
  /* strips within three columns of a wall hold ghost lanes; XE fixes that at
   * compile time (0: none, 1: some), -1 tests it here */
  const bool xedge =
      EDGE && (XE >= 0 ? XE != 0 : ((x0 - 3 < kPad) || (x1 + 3 > nx - kPad)));
  const bool last_chunk = EDGE && (y1 == ny - kPad);
  /* the row a ghost row mirrors (an interior item has none) */
  auto ymir = [&](int r) { return EDGE ? mirror(r, ny) : r; };
  const int gend = last_chunk ? y1 + 1 : y1;
  const int q = x0 - 2 + lane;
  const int qf = min(q, nx), qc = min(q, nx - 1), qg = min(max(q, 1), nx);
  const bool cell_ghost = is_ghost(q, nx);
  const int mc_lane = mirror_lane(lane, q, nx);
  const bool own_x = q >= x0 && (q < x1 || (q == x1 && x1 == nx - kPad));
  const bool own_y = q >= x0 && q < x1;
  const T edx_q = g.edx[qg], cdx_q = g.cdx[qg], cdx_qm = g.cdx[qg - 1];
  const T iedx_q = g.iedx[qg];
  const RowS<T, UNI> edy(g.edy, y0 - 4, ny, lane), cdy(g.cdy, y0 - 4, ny, lane);
  const RowS<T, UNI> iedy(g.iedy, y0 - 4, ny, lane);
  /* reciprocal spacings, used in place of quotients on power-of-two meshes */
  const T rcdx_q = T(1.0) / cdx_q;
  const RowR<T, P2, UNI> rcdy(g.cdy, y0 - 4, ny, lane);

  /* x stage: u1 and m1 of cell row i from that row's u0, F_x, mx1, rho */
  auto xrow = [&](int i, T a_u, T a_fx, T a_m1, T a_r, T& m1, T& u1) {
    const T up1 = dn(a_u, 1);
    T ufx = momentum_flux(edy(i), a_fx, dn(a_fx, 1), up(a_u, 1), a_u, up1,
                          dn(a_u, 2), iedx_q, T(0.5) * (a_u + up1), cdx_qm,
                          cdx_q, dt_h);
    ufx = mirror_fix(xedge, ufx, mc_lane, cell_ghost, T(1.0));
    m1 = advect_update<T, P2>(a_m1, dt_h, ufx, up(ufx, 1), edx_q, cdy(i),
                              iedx_q * rcdy(i));
    const T rex = rho_edge_x<T, P2>(a_r, up(a_r, 1), cdx_q, cdx_qm, cdy(i), edx_q,
                                    T(0.5) * iedx_q * rcdy(i));
    u1 = velocity(m1, rex);
  };
  /* x stage: x flux in y at corner row gf from u1 rows gf-2 .. gf+1 */
  auto corner = [&](int gf, T w0, T w1, T w2, T w3, T a_v, T a_fy) {
    const T v_cell = T(0.5) * (up(a_v, 1) + a_v);
    return momentum_flux(edx_q, a_fy, up(a_fy, 1), w0, w1, w2, w3, iedy(gf),
                         v_cell, cdy(gf - 1), cdy(gf), dt_h);
  };
  /* y stage: v1 and m1 of y-face row k from v0, my1 at k and F_x, u, rho
   * at rows k, k-1 */
  auto yrow = [&](int k, T a_v, T a_m1, T a_fx, T p_fx, T a_u, T p_u, T a_r,
                  T p_r, T& m1, T& v1) {
    const T vm1 = up(a_v, 1);
    const T vfx = momentum_flux(cdy(k), a_fx, p_fx, up(a_v, 2), vm1, a_v,
                                dn(a_v, 1), iedx_q, T(0.5) * (a_u + p_u),
                                cdx_qm, cdx_q, dt_h);
    m1 = advect_update<T, P2>(a_m1, dt_h, dn(vfx, 1), vfx, edx_q, cdy(k),
                              iedx_q * rcdy(k));
    const T rey = rho_edge_y<T, P2>(a_r, p_r, cdx_q, cdy(k), cdy(k - 1), edy(k),
                                    T(0.5) * rcdx_q * iedy(k));
    v1 = velocity(m1, rey);
  };
  /* y stage: y flux in y at cell row i from v1 rows i-1 .. i+2 */
  auto cellflux = [&](int i, T w0, T w1, T w2, T w3, T fy_lo, T fy_hi) {
    return momentum_flux(cdx_q, fy_lo, fy_hi, w0, w1, w2, w3, iedy(i),
                         T(0.5) * (w1 + w2), cdy(i - 1), cdy(i), dt_h);
  };

  /* prologue. x stage on cell rows y0-3 .. y0+1 (a ghost row is its mirror
   * row, as the boundary copies it); F_x and rho of those rows as the y
   * stage reads them (their ghost rows are mirror copies too) */
  T uu[5], ff[5], rr[5], wm[3];
#pragma unroll
  for (int s = 0; s < 5; ++s) {
    const int i = ymir(y0 - 3 + s);
    const T a_fx = io.F_x[i * sx + qf], a_r = io.rho[i * nx + qc];
    T m1, u1;
    xrow(i, io.u0[i * sx + qf], a_fx, io.mx1[i * sx + qf], a_r, m1, u1);
    uu[s] = u1;
    ff[s] = a_fx;
    rr[s] = a_r;
    if (s >= 3) wm[s - 3] = m1;
  }
  T wu[4] = {uu[1], uu[2], uu[3], uu[4]};
  T uf_prev = corner(y0, wu[0], wu[1], wu[2], wu[3], io.v0[y0 * nx + qc],
                     io.F_y[y0 * nx + qc]);
#pragma unroll
  for (int k = 0; k < 3; ++k) wu[k] = wu[k + 1];
  /* y stage on y faces y0-2 .. y0+1; below the south wall the negated
   * mirror faces */
  T wv[4], wn[3];
  wn[0] = wn[1] = wn[2] = T(0);
#pragma unroll
  for (int s = 0; s < 4; ++s) {
    const int k = y0 - 2 + s;
    wv[s] = T(0);
    if (EDGE && k < kPad) continue;
    T m1, v1;
    yrow(k, io.v0[k * nx + qc], io.my1[k * nx + qc], ff[s + 1], ff[s], uu[s + 1],
         uu[s], rr[s + 1], rr[s], m1, v1);
    wv[s] = v1;
    if (s >= 2) wn[s - 2] = m1;
  }
  const bool south = EDGE && (y0 == kPad);
  if (south) {
    wv[0] = -wv[3];
    wv[1] = -wv[2];
  }
  T p_fy = io.F_y[y0 * nx + qc];
  T vfy_prev = south ? T(0)
                     : cellflux(y0 - 1, wv[0], wv[1], wv[2], wv[3],
                                io.F_y[(y0 - 1) * nx + qc], p_fy);
#pragma unroll
  for (int s = 0; s < 3; ++s) wv[s] = wv[s + 1];
  T p_fx = ff[4], p_u = uu[4], p_r = rr[4];
  T p_v = io.v0[(y0 + 1) * nx + qc];
  /* inputs of the lookahead row y0+2 */
  int ia = ymir(y0 + 2);
  T a_u = io.u0[ia * sx + qf], a_fx = io.F_x[ia * sx + qf];
  T a_mx = io.mx1[ia * sx + qf], a_r = io.rho[ia * nx + qc];
  T a_v = io.v0[min(y0 + 2, ny) * nx + qc], a_my = io.my1[min(y0 + 2, ny) * nx + qc];
  T a_fy = io.F_y[(y0 + 1) * nx + qc];
  for (int r = y0; r < gend; ++r) {
    const T c_u = a_u, c_fx = a_fx, c_mx = a_mx, c_r = a_r, c_v = a_v, c_my = a_my,
            c_fy = a_fy;
    /* the next row's inputs; the last row has none to load */
    if (r + 1 < gend) {
      const int in = ymir(r + 3), ry = min(r + 3, ny);
      a_u = io.u0[in * sx + qf];
      a_fx = io.F_x[in * sx + qf];
      a_mx = io.mx1[in * sx + qf];
      a_r = io.rho[in * nx + qc];
      a_v = io.v0[ry * nx + qc];
      a_my = io.my1[ry * nx + qc];
      a_fy = io.F_y[min(r + 2, ny) * nx + qc];
    }
    const int k = r + 2;
    /* x stage on lookahead row k: u there feeds both stages */
    T m1x, u1;
    xrow(ymir(k), c_u, c_fx, c_mx, c_r, m1x, u1);
    wu[3] = u1;
    wm[2] = m1x;
    /* y stage on lookahead face k */
    T m1y = T(0), v1;
    if (!EDGE || k < ny - 1)
      yrow(k, c_v, c_my, c_fx, p_fx, u1, p_u, c_r, p_r, m1y, v1);
    else
      /* ghost y face: mirror of face ny-1 is ny-2 (k-1), of ny is ny-3 (k-3) */
      v1 = -((k == ny - 1) ? wv[2] : wv[0]);
    wv[3] = v1;
    wn[2] = m1y;
    /* x stage output: cell row r */
    if (r < y1) {
      const T uf = corner(r + 1, wu[0], wu[1], wu[2], wu[3], p_v, c_fy);
      if (own_x) {
        io.mom_x[r * sx + q] = advect_update<T, P2>(wm[0], dt_h, uf, uf_prev, cdx_q,
                                                    edy(r), rcdx_q * iedy(r));
        mput_xface(edge, io.u, nx, ny, r, q, wu[1], T(-1.0));
      }
      uf_prev = uf;
    }
    /* y stage output: y-face row r */
    const T vfy = (EDGE && r >= ny - kPad)
                      ? vfy_prev
                      : cellflux(r, wv[0], wv[1], wv[2], wv[3], p_fy, c_fy);
    if (south && r == y0) vfy_prev = vfy;
    if (own_y) {
      io.mom_y[r * nx + q] = advect_update<T, P2>(wn[0], dt_h, vfy, vfy_prev, cdx_q,
                                                  edy(r), rcdx_q * iedy(r));
      mput_yface(edge, io.v, nx, ny, r, q, wv[1], T(-1.0));
    }
    vfy_prev = vfy;
    p_fx = c_fx;
    p_u = u1;
    p_r = c_r;
    p_fy = c_fy;
    p_v = c_v;
#pragma unroll
    for (int s = 0; s < 3; ++s) {
      wu[s] = wu[s + 1];
      wv[s] = wv[s + 1];
    }
    wm[0] = wm[1];
    wm[1] = wm[2];
    wn[0] = wn[1];
    wn[1] = wn[2];
  }
}

/* Even tt (omp3/flow.c:673-774): x momentum (x flux in y, u recompute, x
 * flux in x) and y momentum (y flux in y, v recompute, y flux in x) in one
 * march; the y stage at row r reads u of rows r and r-1 from the x stage.
 * Lane l holds x face / column q = x0 - 3 + l (strips of 59 columns). */
template <typename T, bool P2, int XE = -1, bool EDGE = true, bool UNI = false>
__device__ __forceinline__ void momxy_yx_item(const MGeo<T>& g, const MMom2IO<T>& io,
                                              T dt, T dt_h, const MarchList& ml,
                                              int item, int lane) {
  const int nx = g.nx, ny = g.ny, sx = nx + 1;
  (void)dt;
  const ItemBox box = item_box(ml, nx, item);
  const int x0 = box.x0, x1 = box.x1, y0 = box.y0, y1 = box.y1;
  if (y0 >= y1) return;
  const bool edge = EDGE && ((x0 == kPad) || (x1 + 3 > nx - kPad) || (y0 == kPad) ||
                             (y1 + 3 > ny - kPad));
  /* strips within three columns of a wall hold ghost lanes; XE fixes that at
   * compile time (0: none, 1: some), -1 tests it here */
  const bool xedge =
      EDGE && (XE >= 0 ? XE != 0 : ((x0 - 3 < kPad) || (x1 + 3 > nx - kPad)));
  const bool last_chunk = EDGE && (y1 == ny - kPad);
  const int gend = last_chunk ? y1 + 1 : y1;
  const int q = x0 - 3 + lane;
  const int qf = min(max(q, 0), nx), qc = min(max(q, 0), nx - 1);
  const int qg = min(max(q, 1), nx);
  const bool face_ghost = is_ghost(q, sx), cell_ghost = is_ghost(q, nx);
  const int mf_lane = mirror_lane(lane, q, sx), mc_lane = mirror_lane(lane, q, nx);
  const bool own_x = q >= x0 && (q < x1 || (q == x1 && x1 == nx - kPad));
  const bool own_y = q >= x0 && q < x1;
  const T edx_q = g.edx[qg], cdx_q = g.cdx[qg], cdx_qm = g.cdx[qg - 1];
  const T iedx_q = g.iedx[qg];
  const RowS<T, UNI> edy(g.edy, y0 - 4, ny, lane), cdy(g.cdy, y0 - 4, ny, lane);
  const RowS<T, UNI> iedy(g.iedy, y0 - 4, ny, lane);
  /* reciprocal spacings, used in place of quotients on power-of-two meshes */
  const T rcdx_q = T(1.0) / cdx_q;
  const RowR<T, P2, UNI> rcdy(g.cdy, y0 - 4, ny, lane);

  /* x stage: x flux in y at corner row gf from u0 rows gf-2 .. gf+1 */
  auto corner = [&](int gf, T w0, T w1, T w2, T w3, T a_v, T a_fy) {
    const T v_cell = T(0.5) * (up(a_v, 1) + a_v);
    return momentum_flux(edx_q, a_fy, up(a_fy, 1), w0, w1, w2, w3, iedy(gf),
                         v_cell, cdy(gf - 1), cdy(gf), dt_h);
  };
  /* x stage: u1 of cell row i (ghost faces: the negated mirror) */
  auto urow = [&](int i, T a_m1, T a_r, T uf_hi, T uf_lo, T& m1) {
    m1 = advect_update<T, P2>(a_m1, dt_h, uf_hi, uf_lo, cdx_q, edy(i),
                              rcdx_q * iedy(i));
    const T rex = rho_edge_x<T, P2>(a_r, up(a_r, 1), cdx_q, cdx_qm, cdy(i), edx_q,
                                    T(0.5) * iedx_q * rcdy(i));
    T u1 = velocity(m1, rex);
    return mirror_fix(xedge, u1, mf_lane, face_ghost, T(-1.0));
  };
  /* y stage: y flux in y at cell row i from v0 rows i-1 .. i+2 */
  auto cellflux = [&](int i, T w0, T w1, T w2, T w3, T fy_lo, T fy_hi) {
    return momentum_flux(cdx_q, fy_lo, fy_hi, w0, w1, w2, w3, iedy(i),
                         T(0.5) * (w1 + w2), cdy(i - 1), cdy(i), dt_h);
  };

  /* prologue: u0 and v0 rows y0-2 .. y0+1 (ghost rows are in memory) */
  T wu[4], wv[4];
#pragma unroll
  for (int s = 0; s < 4; ++s) {
    wu[s] = io.u0[(y0 - 2 + s) * sx + qf];
    wv[s] = io.v0[(y0 - 2 + s) * nx + qc];
  }
  const T fy0 = io.F_y[y0 * nx + qc];
  T uf_prev = corner(y0, wu[0], wu[1], wu[2], wu[3], wv[2], fy0);
  const bool south = EDGE && (y0 == kPad);
  T vfy_prev = south ? T(0)
                     : cellflux(y0 - 1, wv[0], wv[1], wv[2], wv[3],
                                io.F_y[(y0 - 1) * nx + qc], fy0);
  T p_fx = io.F_x[(y0 - 1) * sx + qf], p_r = io.rho[(y0 - 1) * nx + qc];
  /* u of cell row y0-1 for the y stage's first row: the x stage there
   * (below the south wall it is row y0's, taken in the first step) */
  T p_u = T(0);
  if (!south) {
    const T um3 = io.u0[(y0 - 3) * sx + qf];
    const T uf_m = corner(y0 - 1, um3, wu[0], wu[1], wu[2], wv[1],
                          io.F_y[(y0 - 1) * nx + qc]);
    T m1;
    p_u = urow(y0 - 1, io.mx1[(y0 - 1) * sx + qf], p_r, uf_prev, uf_m, m1);
  }
#pragma unroll
  for (int s = 0; s < 3; ++s) {
    wu[s] = wu[s + 1];
    wv[s] = wv[s + 1];
  }
  T p_fy = fy0;
  T a_u = io.u0[min(y0 + 2, ny - 1) * sx + qf], a_v = io.v0[(y0 + 2) * nx + qc];
  T a_fy = io.F_y[(y0 + 1) * nx + qc];
  T a_mx = io.mx1[y0 * sx + qf], a_my = io.my1[y0 * nx + qc];
  T a_r = io.rho[y0 * nx + qc], a_fx = io.F_x[y0 * sx + qf];
  for (int r = y0; r < gend; ++r) {
    wu[3] = a_u;
    wv[3] = a_v;
    const T c_fy = a_fy, c_mx = a_mx, c_my = a_my, c_r = a_r, c_fx = a_fx;
    /* the next row's inputs; the last row has none to load */
    if (r + 1 < gend) {
      const int rc = min(r + 1, ny - 1);
      a_u = io.u0[min(r + 3, ny - 1) * sx + qf];
      a_v = io.v0[min(r + 3, ny) * nx + qc];
      a_fy = io.F_y[min(r + 2, ny) * nx + qc];
      a_mx = io.mx1[rc * sx + qf];
      a_my = io.my1[min(r + 1, ny) * nx + qc];
      a_r = io.rho[rc * nx + qc];
      a_fx = io.F_x[rc * sx + qf];
    }
    /* x stage: cell row r (past the north wall, row y1 is a ghost row:
     * its u is row y1-1's) */
    T u1 = p_u;
    if (r < y1) {
      const T uf = corner(r + 1, wu[0], wu[1], wu[2], wu[3], wv[2], c_fy);
      T m1;
      u1 = urow(r, c_mx, c_r, uf, uf_prev, m1);
      const T up1 = dn(u1, 1);
      T ufx = momentum_flux(edy(r), c_fx, dn(c_fx, 1), up(u1, 1), u1, up1,
                            dn(u1, 2), iedx_q, T(0.5) * (u1 + up1), cdx_qm, cdx_q,
                            dt_h);
      ufx = mirror_fix(xedge, ufx, mc_lane, cell_ghost, T(1.0));
      const T ufx_lo = up(ufx, 1);
      if (own_x) {
        io.mom_x[r * sx + q] = advect_update<T, P2>(m1, dt_h, ufx, ufx_lo, edx_q,
                                                    cdy(r), iedx_q * rcdy(r));
        mput_xface(edge, io.u, nx, ny, r, q, u1, T(-1.0));
      }
      uf_prev = uf;
    }
    if (south && r == y0) p_u = u1;
    /* y stage: y-face row r */
    const T vfy = (EDGE && r >= ny - kPad)
                      ? vfy_prev
                      : cellflux(r, wv[0], wv[1], wv[2], wv[3], p_fy, c_fy);
    if (south && r == y0) vfy_prev = vfy;
    const T m1y = advect_update<T, P2>(c_my, dt_h, vfy, vfy_prev, cdx_q, edy(r),
                                       rcdx_q * iedy(r));
    const T rey = rho_edge_y<T, P2>(c_r, p_r, cdx_q, cdy(r), cdy(r - 1), edy(r),
                                    T(0.5) * rcdx_q * iedy(r));
    T v1 = velocity(m1y, rey);
    v1 = mirror_fix(xedge, v1, mc_lane, cell_ghost, T(1.0));
    const T vm1 = up(v1, 1);
    const T vfx = momentum_flux(cdy(r), c_fx, p_fx, up(v1, 2), vm1, v1, dn(v1, 1),
                                iedx_q, T(0.5) * (u1 + p_u), cdx_qm, cdx_q, dt_h);
    const T vfx_hi = dn(vfx, 1);
    if (own_y) {
      io.mom_y[r * nx + q] = advect_update<T, P2>(m1y, dt_h, vfx_hi, vfx, edx_q,
                                                  cdy(r), iedx_q * rcdy(r));
      mput_yface(edge, io.v, nx, ny, r, q, v1, T(-1.0));
    }
    vfy_prev = vfy;
    p_fy = c_fy;
    p_fx = c_fx;
    p_u = u1;
    p_r = c_r;
#pragma unroll
    for (int s = 0; s < 3; ++s) {
      wu[s] = wu[s + 1];
      wv[s] = wv[s + 1];
    }
  }
}

template <typename T, bool P2, bool UNI = false>
__global__ __launch_bounds__(kMNT, FLOWHIP_MARCH_WAVES) void k_march_momxy_xy(
    MGeo<T> g, MMom2IO<T> io, const double* dtp, unsigned int* queue, MarchList ml) {
  const T dt = T(dtp[0]), dt_h = T(dtp[1]);
  const int lane = threadIdx.x % kWave;
  const int wave = threadIdx.x / kWave;
  MarchBlockQueue mq(queue, ml.nitems);
  for (int base = mq.next(); base < mq.end; base = mq.next()) {
    const int item = __builtin_amdgcn_readfirstlane(base + wave);
    if (item >= mq.end) continue;
    if (!item_edge(ml, g.nx, g.ny, item))
      momxy_xy_item<T, P2, 0, false, UNI>(g, io, dt, dt_h, ml, item, lane);
    else if (item_xedge(ml, g.nx, item))
      momxy_xy_item<T, P2, 1, true, UNI>(g, io, dt, dt_h, ml, item, lane);
    else
      momxy_xy_item<T, P2, 0, true, UNI>(g, io, dt, dt_h, ml, item, lane);
  }
  mq.finish();
}

template <typename T, bool P2, bool UNI = false>
__global__ __launch_bounds__(kMNT, FLOWHIP_MARCH_WAVES) void k_march_momxy_yx(
    MGeo<T> g, MMom2IO<T> io, const double* dtp, unsigned int* queue, MarchList ml) {
  const T dt = T(dtp[0]), dt_h = T(dtp[1]);
  const int lane = threadIdx.x % kWave;
  const int wave = threadIdx.x / kWave;
  MarchBlockQueue mq(queue, ml.nitems);
  for (int base = mq.next(); base < mq.end; base = mq.next()) {
    const int item = __builtin_amdgcn_readfirstlane(base + wave);
    if (item >= mq.end) continue;
    if (!item_edge(ml, g.nx, g.ny, item))
      momxy_yx_item<T, P2, 0, false, UNI>(g, io, dt, dt_h, ml, item, lane);
    else if (item_xedge(ml, g.nx, item))
      momxy_yx_item<T, P2, 1, true, UNI>(g, io, dt, dt_h, ml, item, lane);
    else
      momxy_yx_item<T, P2, 0, true, UNI>(g, io, dt, dt_h, ml, item, lane);
  }
  mq.finish();
}

/* Persistent grid of a march kernel: as many blocks as stay resident on the
 * device (per_cu per CU, from the occupancy query on first use, at most
 * max_per_cu when that is positive), and no more than the items need. */
template <typename K>
int march_grid(K kernel, int& per_cu, int nitems, int max_per_cu = 0) {
  static int cus = 0;
  if (cus == 0) {
    int dev = 0;
    hipDeviceProp_t p;
    if (hipGetDevice(&dev) != hipSuccess ||
        hipGetDeviceProperties(&p, dev) != hipSuccess)
      cus = 256;
    else
      cus = p.multiProcessorCount;
  }
  if (per_cu == 0 &&
      (hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, kernel, kMNT, 0) !=
           hipSuccess ||
       per_cu < 1))
    per_cu = 1;
  if (max_per_cu > 0 && per_cu > max_per_cu) per_cu = max_per_cu;
  long gsz = (long)cus * per_cu;
  const long need = (nitems + kMWaves - 1) / kMWaves;
  if (gsz > need) gsz = need;
  if (gsz > kMaxReduceBlocks) gsz = kMaxReduceBlocks;
  return gsz < 1 ? 1 : (int)gsz;
}

}  // namespace flowhip
