/*
 * tile.h -- the tiling framework of the fused kernels.
 *
 * A fused kernel walks a list of TY x TX cell tiles covering the interior.
 * Launches are persistent: a bounded grid of blocks loops over the tile
 * list, so the per-block reduction partials stay few. Blocks are dispatched
 * round-robin over the 8 XCDs of the chip; each XCD's blocks walk one
 * contiguous band of tile rows, so the halo rows one tile re-reads are the
 * rows its neighbours just read into the same XCD's L2.
 *
 * Inside a tile every intermediate stage lives in LDS over a rectangle that
 * includes the halo the next stage reads. A value at a ghost index of a
 * stage is that stage evaluated at the mirrored interior index, times the
 * boundary's sign: exactly what handle_boundary_2d would have copied there
 * (pad-deep mirror strips, corners untouched; handle_boundary_2d's call
 * sites in omp3/flow.c fix the sign of each field).
 */
#pragma once

#include "common.h"

namespace flowhip {

constexpr int kXcds = 8;
constexpr int kPad = 2;

/* Mirror of index k along an axis of extent n (cells: n = local size;
 * faces: n = local size + 1): handle_boundary_2d copies index
 * 2*pad-1-k into ghost k < pad and 2*(n-pad)-1-k into ghost k >= n-pad. */
__device__ __forceinline__ int mirror(int k, int n) {
  return (k < kPad) ? (2 * kPad - 1 - k) : ((k >= n - kPad) ? (2 * (n - kPad) - 1 - k) : k);
}
__device__ __forceinline__ bool is_ghost(int k, int n) {
  return k < kPad || k >= n - kPad;
}

/* Mirror / ghost tests specialised on whether the tile touches the domain
 * edge: an interior tile's halo is all interior. */
template <bool EDGE>
__device__ __forceinline__ int mir(int k, int n) {
  return EDGE ? mirror(k, n) : k;
}
template <bool EDGE>
__device__ __forceinline__ bool ghost(int k, int n) {
  return EDGE ? is_ghost(k, n) : false;
}

struct TileList {
  int ntx, nty;  /* tiles along x and y */
  int ntiles;
};

__host__ __device__ inline TileList make_tiles(int nx, int ny, int tx, int ty) {
  TileList t;
  t.ntx = (nx - 2 * kPad + tx - 1) / tx;
  t.nty = (ny - 2 * kPad + ty - 1) / ty;
  t.ntiles = t.ntx * t.nty;
  return t;
}

/* Iterates the tiles of this block. The tile list is cut into
 * B = min(8, gridDim) contiguous bands; band b = blockIdx % B is dealt to
 * the blocks with that residue in turn. With 8 or more blocks a band is
 * one XCD's share (blocks reach the XCDs round-robin), and every tile is
 * visited exactly once for any grid size. */
struct TileWalk {
  int t, hi, stride;
  __device__ TileWalk(const TileList& tl) {
    const int nb = gridDim.x, b = blockIdx.x;
    const int bands = nb < kXcds ? nb : kXcds;
    const int band = b % bands, local = b / bands;
    const int nb_band = (nb - band + bands - 1) / bands;
    const int lo = (int)((long)band * tl.ntiles / bands);
    hi = (int)((long)(band + 1) * tl.ntiles / bands);
    t = lo + local;
    stride = nb_band;
  }
  __device__ bool next(int& tile) {
    if (t >= hi) return false;
    tile = t;
    t += stride;
    return true;
  }
};

/* Calls f(r, c) for every element of a ROWS x COLS region, spread over the
 * block's NT threads. */
template <int NT, int ROWS, int COLS, class F>
__device__ __forceinline__ void region(F&& f) {
  constexpr int kIters = (ROWS * COLS + NT - 1) / NT;
#pragma unroll
  for (int it = 0; it < kIters; ++it) {
    const int k = threadIdx.x + it * NT;
    if (k < ROWS * COLS) f(k / COLS, k % COLS);
  }
}

inline int persistent_grid(int ntiles, int blocks_per_cu) {
  int g = 256 * blocks_per_cu;
  if (g > ntiles) g = ntiles;
  if (g > 8192) g = 8192;
  return g < 1 ? 1 : g;
}

}  // namespace flowhip
