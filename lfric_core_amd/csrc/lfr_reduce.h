// Deterministic FP64 block reductions (wavefront shuffles + fixed-order LDS
// combine).  A reduction over n items always uses the same grid for the same
// n, and partials are combined in a fixed order, so results are bitwise
// reproducible run to run.
#pragma once

#include "lfr_internal.h"

namespace lfr {

constexpr int kRedThreads = 256;
constexpr int kRedMaxBlocks = 1024;

// Grid size of the first reduction stage for n items (function of n only).
static inline int red_blocks(int64_t n) {
  int64_t b = (n + (int64_t)kRedThreads * 16 - 1) / ((int64_t)kRedThreads * 16);
  if (b < 1) b = 1;
  if (b > kRedMaxBlocks) b = kRedMaxBlocks;
  return (int)b;
}

__device__ __forceinline__ double wave_sum(double v) {
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) v += __shfl_down(v, off, 64);
  return v;
}

// Sum of one value per thread over a block of BS threads; the result is
// valid in thread 0.  `lds` must hold BS/64 doubles.
template <int BS = kRedThreads>
__device__ __forceinline__ double block_sum(double v, double* lds) {
  v = wave_sum(v);
  const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
  if (lane == 0) lds[wid] = v;
  __syncthreads();
  double s = 0.0;
  if (threadIdx.x == 0) {
#pragma unroll
    for (int w = 0; w < BS / 64; ++w) s += lds[w];
  }
  __syncthreads();
  return s;
}

// sum of p[threadIdx.x + k*BS] over k in increasing order (the thread's
// share of a block-wide sum of n values over BS threads); eight loads are
// issued before they are added, the addition order is unchanged.  The last
// (at most seven) values are loaded together too, at clamped addresses, and
// a slot past the end adds +0: s starts at +0 and, rounding to nearest,
// never becomes -0, so adding +0 leaves it bitwise unchanged (added one a
// trip, each value waited for its own load: a round trip per value).
template <int BS = kRedThreads>
__device__ __forceinline__ double strided_sum(const double* __restrict__ p, int n) {
  double s = 0.0;
  int b = threadIdx.x;
  for (; b + 7 * BS < n; b += 8 * BS) {
    double v[8];
#pragma unroll
    for (int q = 0; q < 8; ++q) v[q] = p[b + q * BS];
#pragma unroll
    for (int q = 0; q < 8; ++q) s += v[q];
  }
  if (n > 0) {
    double v[7];
#pragma unroll
    for (int q = 0; q < 7; ++q) {
      const int i = b + q * BS;
      const double w = p[i < n ? i : n - 1];
      v[q] = i < n ? w : 0.0;
    }
#pragma unroll
    for (int q = 0; q < 7; ++q) s += v[q];
  }
  return s;
}

// Final stage: out[q] = sum over b of part[q*stride + b], b < nb, for
// q < nq (one block per q).
static __global__ void __launch_bounds__(kRedThreads)
k_reduce_final(const double* __restrict__ part, int nb, int stride, double* __restrict__ out) {
  __shared__ double lds[kRedThreads / 64];
  const double* p = part + (int64_t)blockIdx.x * stride;
  double s = strided_sum(p, nb);
  s = block_sum(s, lds);
  if (threadIdx.x == 0) out[blockIdx.x] = s;
}

}  // namespace lfr
