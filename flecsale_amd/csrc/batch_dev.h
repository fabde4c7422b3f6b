// Device helpers for the tile-driven kernels of a ragged batch.
#pragma once

#include "internal.h"
#include "kernels_common.h"

namespace fsale {

// Iterates a block over its tiles.  A direct list (one block per tile)
// holds the tile of every block of the grid in launch order, arranged on the
// host so that XCD x = blockIdx % 8 (the hardware deals blocks round-robin
// to the 8 XCDs) gets one contiguous range of the batch and neighbour
// gathers stay in that XCD's L2; entries past the live tiles have count 0.
// Otherwise the block loops: XCD x takes tiles [x*chunk, (x+1)*chunk) in
// steps of gridDim/8.  `body(tile)` is called uniformly by the whole block.
template <class F>
__device__ inline void for_each_tile(const TileList tl, F &&body) {
  if (tl.direct) {
    if ((int)blockIdx.x < tl.ndirect) {
      const Tile t = tl.direct[blockIdx.x];
      if (t.count > 0) body(t);
    }
    return;
  }
  const int T = *tl.n;
  const int chunk = (T + 7) >> 3;
  const int x = blockIdx.x & 7, b = blockIdx.x >> 3, G8 = gridDim.x >> 3;
  for (int i = b; i < chunk; i += G8) {
    const int t = x * chunk + i;
    if (t >= T) break;
    body(tl.tiles[t]);
  }
}

// Whether a wave works on problem `prob` in a launch identified by `token`:
// the step is live and no earlier launch of it raised an error (launches
// that can raise the negative-state error pass their own non-zero token, so
// all of their waves finish even after one of them failed, and the reported
// cell is the lowest failing one of that launch, as in the reference's
// serial loop).  Each wave decides for itself: the kernels on the step path
// take no block-wide barrier, so waves of one block may disagree.
__device__ inline bool wave_live(const DevScalars *s, int prob, int token) {
  const DevScalars &S = s[prob];
  const int et = S.err_token;
  return S.active && (et == 0 || et == token);
}

// A problem's step flag, agreed on by the whole block (for the one-block-
// per-problem finalize kernels, which use block reductions).
__device__ inline bool block_active(const DevScalars *s, int prob) {
  __shared__ int sh;
  __syncthreads();
  if (threadIdx.x == 0) sh = s[prob].active;
  __syncthreads();
  return sh != 0;
}

// Max over the time-step partials of one problem, computed by one block of
// kFinalBlock threads: n entries of S interleaved non-negative values; four
// independent loads in flight per thread.  Max is exact, so the order does
// not matter.  Result (per component) valid in every thread.
constexpr int kFinalBlock = 1024;

template <int S, int B = kFinalBlock>
__device__ inline void final_max(const double *a, int n, double (&r)[S]) {
  // U independent loads in flight per thread: a smaller block keeps more
  constexpr int U = B >= kFinalBlock ? 4 : 16;
  __shared__ double red[S][B / 64];
  double acc[U][S];
  for (int k = 0; k < U; ++k)
    for (int j = 0; j < S; ++j) acc[k][j] = 0;
  int i = threadIdx.x;
  for (; i + (U - 1) * B < n; i += U * B)
#pragma unroll
    for (int k = 0; k < U; ++k)
      for (int j = 0; j < S; ++j) acc[k][j] = smax(a[(size_t)(i + k * B) * S + j], acc[k][j]);
  for (; i < n; i += B)
    for (int j = 0; j < S; ++j) acc[0][j] = smax(a[(size_t)i * S + j], acc[0][j]);
  __syncthreads();
  for (int j = 0; j < S; ++j) {
    double v = acc[0][j];
    for (int k = 1; k < U; ++k) v = smax(v, acc[k][j]);
    v = wave_max(v);
    if ((threadIdx.x & 63) == 0) red[j][threadIdx.x >> 6] = v;
  }
  __syncthreads();
  for (int j = 0; j < S; ++j) {
    double b = red[j][0];
    for (int w = 1; w < B / 64; ++w) b = smax(red[j][w], b);
    r[j] = b;
  }
}

// negative density or internal energy in cell `local_cell` of `prob`,
// raised by launch `token`
__device__ inline void flag_negative(DevScalars *s, int prob, int local_cell, int token) {
  DevScalars &S = s[prob];
  const int old = atomicCAS(&S.err_token, 0, token);
  if (old == 0 || old == token) {
    atomicCAS(&S.err_code, 0, (int)FSALE_ERR_NEG_STATE);
    atomicMin(reinterpret_cast<unsigned long long *>(&S.err_cell), (unsigned long long)local_cell);
  }
}

}  // namespace fsale
