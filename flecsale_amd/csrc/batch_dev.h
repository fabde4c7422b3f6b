// Device helpers for the persistent, tile-driven kernels of a ragged batch.
#pragma once

#include "internal.h"
#include "kernels_common.h"

namespace fsale {

// Persistent loop over a device tile list.  The hardware deals blocks
// round-robin to the 8 XCDs; XCD x = blockIdx % 8 takes tiles
// [x*chunk, (x+1)*chunk) so each XCD (and its own L2) works on one contiguous
// range of the batch and neighbour gathers stay in that L2.  `body(tile)` is
// called uniformly by the whole block (it may use __syncthreads).
template <class F>
__device__ inline void for_each_tile(const TileList tl, F &&body) {
  const int T = *tl.n;
  const int chunk = (T + 7) >> 3;
  const int x = blockIdx.x & 7, b = blockIdx.x >> 3, G8 = gridDim.x >> 3;
  for (int i = b; i < chunk; i += G8) {
    const int t = x * chunk + i;
    if (t >= T) break;
    body(tl.tiles[t]);
  }
}

// A problem's step flag, read once per block (another block may clear it
// while this one runs, and the block reductions need every lane).
__device__ inline bool block_active(const DevScalars *s, int prob) {
  __shared__ int sh;
  __syncthreads();
  if (threadIdx.x == 0) sh = s[prob].active;
  __syncthreads();
  return sh != 0;
}

// block max of a non-negative value, result valid in thread 0
__device__ inline double block_max(double v) {
  __shared__ double red[kBlock / 64];
  v = wave_max(v);
  __syncthreads();
  if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = v;
  __syncthreads();
  double b = red[0];
  for (int w = 1; w < kBlock / 64; ++w) b = smax(red[w], b);
  return b;
}

__device__ inline void flag_negative(DevScalars *s, int prob, int local_cell) {
  atomicCAS(&s[prob].err_code, 0, (int)FSALE_ERR_NEG_STATE);
  atomicMin(reinterpret_cast<unsigned long long *>(&s[prob].err_cell), (unsigned long long)local_cell);
  s[prob].active = 0;
}

}  // namespace fsale
