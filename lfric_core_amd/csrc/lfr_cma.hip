// Column-wise (CMA) operators: the tridiagonal column solve
// columnwise_op_appinv and the banded apply columnwise_op_app.
//
// appinv (components/science/source/kernel/columnwise_operators/
// sci_columnwise_op_appinv_kernel_mod.F90:77-145): one Thomas solve per
// column, overwriting lhs(map(1) + indirection_dofmap(i) - 1).  Bands (1,2,3)
// = (sub, diag, super); one reciprocal of the pivot per row, then the products
// inv*c and inv*(x - a*d') exactly as written there.
//
// Mapping: one lane per column, 64 columns per workgroup.  The band matrix
// of the workgroup's 64 columns (contiguous in columnwise_matrix(bw, nrow,
// ncell)) and their right-hand sides are staged through LDS in chunks of
// CH rows with coalesced loads; each lane runs its recurrence out of LDS and
// keeps c' and d' in registers (NROW is a compile-time constant and every
// loop over rows is unrolled, so the arrays never spill to scratch).  The
// solution goes back through LDS for a coalesced store.
#include "lfr_internal.h"

namespace {

constexpr int kCols = 64;   // columns (lanes) per workgroup
constexpr int kCh = 10;     // rows per LDS stage

template <typename T, typename C, int NROW>
__global__ void __launch_bounds__(kCols)
k_appinv(T* __restrict__ lhs, const T* __restrict__ x, const T* __restrict__ cm,
         const int32_t* __restrict__ map0, int ndf, const int32_t* __restrict__ ind,
         int ncell) {
  constexpr int NST = (NROW + kCh - 1) / kCh;
  constexpr int PITCH_B = kCh * 3 + 1;     // odd pitch: conflict-free lane-strided reads
  constexpr int PITCH_X = kCh + 1;
  __shared__ T sb[kCols * PITCH_B];
  __shared__ T sx[kCols * PITCH_X];
  __shared__ int32_t sbase[kCols];
  const int lane = threadIdx.x;
  const int c0 = blockIdx.x * kCols;
  const int ncols = min(kCols, ncell - c0);
  if (lane < ncols) sbase[lane] = map0[(int64_t)(c0 + lane) * ndf];
  C cp[NROW], dp[NROW];
#pragma unroll
  for (int s = 0; s < NST; ++s) {
    const int r0 = s * kCh;
    const int nr = (NROW - r0) < kCh ? (NROW - r0) : kCh;
    __syncthreads();
    // coalesced staging: element q of column cc's chunk
    const int nb = ncols * nr * 3;
    for (int e = lane; e < nb; e += kCols) {
      const int cc = e / (nr * 3);
      const int q = e - cc * (nr * 3);
      sb[cc * PITCH_B + q] = cm[(int64_t)(c0 + cc) * 3 * NROW + (int64_t)r0 * 3 + q];
    }
    const int nx = ncols * nr;
    for (int e = lane; e < nx; e += kCols) {
      const int cc = e / nr;
      const int q = e - cc * nr;
      sx[cc * PITCH_X + q] = x[sbase[cc] + ind[r0 + q] - 1];
    }
    __syncthreads();
    if (lane < ncols) {
#pragma unroll
      for (int q = 0; q < kCh; ++q) {
        const int i = r0 + q;
        if (i < NROW) {
          const C a = (C)sb[lane * PITCH_B + 3 * q + 0];
          const C d = (C)sb[lane * PITCH_B + 3 * q + 1];
          const C u = (C)sb[lane * PITCH_B + 3 * q + 2];
          const C xv = (C)sx[lane * PITCH_X + q];
          if (i == 0) {
            const C inv = C(1) / d;
            cp[i] = inv * u;
            dp[i] = inv * xv;
          } else {
            const C ac = a * cp[i - 1];
            const C inv = C(1) / (d - ac);
            if (i < NROW - 1) cp[i] = inv * u;
            const C ad = a * dp[i - 1];
            dp[i] = inv * (xv - ad);
          }
        }
      }
    }
  }
  // backward substitution in registers
#pragma unroll
  for (int i = NROW - 2; i >= 0; --i) {
    const C cd = cp[i] * dp[i + 1];
    dp[i] = dp[i] - cd;
  }
  // stage the solution back through LDS for the store
#pragma unroll
  for (int s = 0; s < NST; ++s) {
    const int r0 = s * kCh;
    const int nr = (NROW - r0) < kCh ? (NROW - r0) : kCh;
    __syncthreads();
    if (lane < ncols) {
#pragma unroll
      for (int q = 0; q < kCh; ++q)
        if (r0 + q < NROW) sx[lane * PITCH_X + q] = (T)dp[r0 + q];
    }
    __syncthreads();
    const int nx = ncols * nr;
    for (int e = lane; e < nx; e += kCols) {
      const int cc = e / nr;
      const int q = e - cc * nr;
      lhs[sbase[cc] + ind[r0 + q] - 1] = sx[cc * PITCH_X + q];
    }
  }
}

// Generic row count: same recurrence with c' and d' in a global scratch
// array of the compute kind (2*nrow per column).
template <typename T, typename C>
__global__ void __launch_bounds__(kCols)
k_appinv_generic(T* __restrict__ lhs, const T* __restrict__ x, const T* __restrict__ cm,
                 const int32_t* __restrict__ map0, int ndf, const int32_t* __restrict__ ind,
                 int ncell, int nrow, C* __restrict__ scratch) {
  const int c = blockIdx.x * kCols + threadIdx.x;
  if (c >= ncell) return;
  const T* m = cm + (int64_t)c * 3 * nrow;
  C* cpc = scratch + (int64_t)c * 2 * nrow;
  C* dpc = cpc + nrow;
  const int base = map0[(int64_t)c * ndf];
  for (int i = 0; i < nrow; ++i) {
    const C a = (C)m[3 * i], d = (C)m[3 * i + 1], u = (C)m[3 * i + 2];
    const C xv = (C)x[base + ind[i] - 1];
    if (i == 0) {
      const C inv = C(1) / d;
      cpc[i] = inv * u;
      dpc[i] = inv * xv;
    } else {
      const C ac = a * cpc[i - 1];
      const C inv = C(1) / (d - ac);
      if (i < nrow - 1) cpc[i] = inv * u;
      const C ad = a * dpc[i - 1];
      dpc[i] = inv * (xv - ad);
    }
  }
  for (int i = nrow - 1; i >= 0; --i) {
    if (i < nrow - 1) {
      const C cd = cpc[i] * dpc[i + 1];
      dpc[i] = dpc[i] - cd;
    }
    lhs[base + ind[i] - 1] = (T)dpc[i];
  }
}

template <typename T, typename C>
int launch_appinv(lfr_ctx* ctx, T* lhs, const T* x, const T* cm, const lfr_space* sp,
                  const int32_t* ind, int nrow, int ncell) {
  const unsigned nb = (unsigned)((ncell + kCols - 1) / kCols);
  switch (nrow) {
#define CASE(NR)                                                                              \
  case NR:                                                                                    \
    hipLaunchKernelGGL((k_appinv<T, C, NR>), dim3(nb), dim3(kCols), 0, ctx->stream, lhs, x, cm, \
                       sp->map0_d, sp->ndf, ind, ncell);                                      \
    break;
    CASE(3) CASE(4) CASE(10) CASE(11) CASE(20) CASE(21) CASE(30) CASE(31) CASE(50) CASE(51)
    CASE(70) CASE(71)
#undef CASE
    default: {
      // c', d' scratch for row counts without a compiled kernel
      int rc = lfr_work_reserve(ctx, (size_t)ncell * 2 * nrow * sizeof(C) + 4096);
      if (rc) return rc;
      hipLaunchKernelGGL((k_appinv_generic<T, C>), dim3(nb), dim3(kCols), 0, ctx->stream, lhs, x,
                         cm, sp->map0_d, sp->ndf, ind, ncell, nrow,
                         (C*)((char*)ctx->work + 4096));
    }
  }
  LFR_CHECK_LAUNCH(ctx);
  return LFR_OK;
}

}  // namespace

extern "C" int lfr_columnwise_op_appinv(lfr_ctx* ctx, int prec, void* lhs, const void* x,
                                        const void* cm, int bandwidth, int nrow, int ncell_2d,
                                        const lfr_space* space, const int32_t* indirection_d,
                                        int fp64_pivots) {
  if (!ctx) return LFR_ERR_ARG;
  LFR_REQUIRE(ctx, prec == 4 || prec == 8, "lfr_columnwise_op_appinv: prec %d", prec);
  LFR_REQUIRE(ctx, lhs && x && cm && space && indirection_d, "lfr_columnwise_op_appinv: null argument");
  LFR_REQUIRE(ctx, bandwidth == 3, "lfr_columnwise_op_appinv: tridiagonal only (bandwidth %d)", bandwidth);
  LFR_REQUIRE(ctx, ncell_2d == space->ncell && nrow > 0,
              "lfr_columnwise_op_appinv: ncell_2d %d vs space %d, nrow %d", ncell_2d, space->ncell, nrow);
  if (prec == 4) {
    if (fp64_pivots)
      return launch_appinv<float, double>(ctx, (float*)lhs, (const float*)x, (const float*)cm,
                                          space, indirection_d, nrow, ncell_2d);
    return launch_appinv<float, float>(ctx, (float*)lhs, (const float*)x, (const float*)cm, space,
                                       indirection_d, nrow, ncell_2d);
  }
  return launch_appinv<double, double>(ctx, (double*)lhs, (const double*)x, (const double*)cm,
                                       space, indirection_d, nrow, ncell_2d);
}
