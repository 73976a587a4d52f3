// Column-wise (CMA) operators: the tridiagonal column solve
// columnwise_op_appinv, the banded apply columnwise_op_app and the assembly
// columnwise_op_asm (components/science/source/kernel/columnwise_operators/
// sci_columnwise_op_{app,asm}_kernel_mod.F90), one thread per (column, row)
// for the latter two, each row's sums formed in the reference order.
//
// appinv (components/science/source/kernel/columnwise_operators/
// sci_columnwise_op_appinv_kernel_mod.F90:77-145): one Thomas solve per
// column, overwriting lhs(map(1) + indirection_dofmap(i) - 1).  Bands (1,2,3)
// = (sub, diag, super); one reciprocal of the pivot per row, then the products
// inv*c and inv*(x - a*d') exactly as written there.
//
// Mapping: one lane per column, 64 columns per workgroup.  A lane reads its
// column's bands (contiguous in columnwise_matrix(bw, nrow, ncell)) and
// right-hand side straight from global memory, every load a constant offset
// from one per-lane base, issued kPf rows ahead of the recurrence (rows are
// unrolled, NROW a compile-time constant, and scheduling barriers keep each
// group of loads one group ahead); c' goes to LDS, d' stays in registers,
// and the solution goes back through LDS for a coalesced store.
#include <algorithm>
#include <cmath>
#include <cstdint>
#include <vector>

#include "lfr_internal.h"

namespace {

constexpr int kCols = 64;   // columns (lanes) per workgroup
constexpr int kCh = 14;     // rows per LDS stage of the solution store
constexpr int kPf = 10;     // rows of loads ahead of a column recurrence (<= 63 loads in flight)

// indirection_dofmap(i) - 1, clamped into the column's rows 0..nrow-1 so that
// no access leaves the column (lfr_cma_create validates its maps; the raw
// pointer given to lfr_columnwise_op_appinv is not read on the host)
__device__ __forceinline__ int ind0(const int32_t* __restrict__ ind, int i, int nrow) {
  const int v = ind[i] - 1;
  return v < 0 ? 0 : (v >= nrow ? nrow - 1 : v);
}

template <typename T, typename C, int NROW>
__global__ void __launch_bounds__(kCols)
k_appinv(T* __restrict__ lhs, const T* __restrict__ x, const T* __restrict__ cm,
         const int32_t* __restrict__ map0, int ndf, const int32_t* __restrict__ ind,
         int ncell) {
  constexpr int NST = (NROW + kCh - 1) / kCh;
  constexpr int PITCH_X = kCh + 1;         // odd pitch: conflict-free lane-strided access
  __shared__ T sx[kCols * PITCH_X];
  __shared__ int32_t sbase[kCols];
  __shared__ int32_t sind[NROW];   // indirection_dofmap - 1, read once
  constexpr int PC = NROW | 1;     // odd pitch of the c' rows
  __shared__ C scp[kCols * PC];    // c' of every lane's column (registers hold d')
  const int lane = threadIdx.x;
  const int c0 = blockIdx.x * kCols;
  const int ncols = min(kCols, ncell - c0);
  if (lane < ncols) sbase[lane] = map0[(int64_t)(c0 + lane) * ndf];
  for (int i = lane; i < NROW; i += kCols) sind[i] = ind0(ind, i, NROW);
  C* cp = scp + lane * PC;
  C dp[NROW];
  // forward sweep straight from global memory: lane = column, its bands
  // are contiguous (3 NROW values) and so are its W3 dofs, so each load is a
  // constant offset from one per-lane base.  Loads run kPf rows ahead of the
  // recurrence: scheduling region r holds the loads of rows [r kPf, (r+1) kPf)
  // and the recurrence of the kPf rows before them, and no instruction
  // crosses a region boundary, so about 4 kPf loads are in flight while the
  // recurrence runs (a partial last workgroup clamps idle lanes to a real
  // column).
  {
    const int cl = min(lane, ncols - 1);
    const T* mb = cm + (int64_t)(c0 + cl) * 3 * NROW;
    const T* xb = x + map0[(int64_t)(c0 + cl) * ndf];
    T la[NROW], ld[NROW], lu[NROW], lx[NROW];
#pragma unroll
    for (int t = 0; t < NROW + kPf; ++t) {
      if (t < NROW) {
        la[t] = mb[3 * t + 0];
        ld[t] = mb[3 * t + 1];
        lu[t] = mb[3 * t + 2];
        lx[t] = xb[ind0(ind, t, NROW)];
      }
      if (t >= kPf) {
        const int i = t - kPf;
        const C a = (C)la[i], d = (C)ld[i], u = (C)lu[i], xv = (C)lx[i];
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
      if ((t + 1) % kPf == 0) __builtin_amdgcn_sched_barrier(0);
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
      lhs[sbase[cc] + sind[r0 + q]] = sx[cc * PITCH_X + q];
    }
  }
}

// columnwise_op_appinv with LDS-staged loads (FP32 throughout): the bands
// of a workgroup's 64 columns are one contiguous block (3 NROW values per
// column), so each kCh-row stage is read by instructions covering a few
// columns' contiguous runs (the per-lane loads of k_appinv touch 64 lines
// per instruction), all of a stage's loads issued before its LDS writes; the
// recurrence is k_appinv's term for term with c' and d' in registers.
template <typename T, int NROW>
__global__ void __launch_bounds__(kCols)
k_appinv_staged(T* __restrict__ lhs, const T* __restrict__ x, const T* __restrict__ cm,
                const int32_t* __restrict__ map0, int ndf, const int32_t* __restrict__ ind, int ncell) {
  constexpr int KCH = kCh;
  constexpr int NST = (NROW + KCH - 1) / KCH;
  constexpr int NB = 3 * KCH;              // band values of one column per stage
  constexpr int PITCH_B = NB + 1;          // odd
  constexpr int PITCH_X = KCH + 1;         // odd
  __shared__ T sb[kCols * PITCH_B];
  __shared__ T sx[kCols * PITCH_X];
  __shared__ int32_t sbase[kCols];
  __shared__ int32_t sind[NROW];   // indirection_dofmap - 1, read once
  const int lane = threadIdx.x;
  const int c0 = blockIdx.x * kCols;
  const int ncols = min(kCols, ncell - c0);
  if (lane < ncols) sbase[lane] = map0[(int64_t)(c0 + lane) * ndf];
  for (int i = lane; i < NROW; i += kCols) sind[i] = ind0(ind, i, NROW);
  T cp[NROW], dp[NROW];
#pragma unroll
  for (int s = 0; s < NST; ++s) {
    const int r0 = s * KCH;
    const int nr = (NROW - r0) < KCH ? (NROW - r0) : KCH;
    __syncthreads();
    // band slot it of this lane: value qs of column cs's stage (e = lane +
    // it kCols = cs NB + qs); x slot: row qs of column cs (e = cs KCH + qs);
    // both clamped into the workgroup's columns and the stage's rows
    T rb[NB], rx[KCH];
#pragma unroll
    for (int it = 0; it < NB; ++it) {   // kCols NB slots: NB per lane
      const int e = lane + it * kCols;
      const int cs = e / NB;
      const int qs = e - cs * NB;
      rb[it] = cm[(int64_t)(c0 + min(cs, ncols - 1)) * 3 * NROW + 3 * r0 + min(qs, 3 * nr - 1)];
    }
#pragma unroll
    for (int it = 0; it < KCH; ++it) {
      const int e = lane + it * kCols;
      const int cs = e / KCH;
      const int qs = e - cs * KCH;
      rx[it] = x[sbase[min(cs, ncols - 1)] + sind[r0 + min(qs, nr - 1)]];
    }
#pragma unroll
    for (int it = 0; it < NB; ++it) {
      const int e = lane + it * kCols;
      const int cs = e / NB;
      sb[cs * PITCH_B + (e - cs * NB)] = rb[it];
    }
#pragma unroll
    for (int it = 0; it < KCH; ++it) {
      const int e = lane + it * kCols;
      const int cs = e / KCH;
      sx[cs * PITCH_X + (e - cs * KCH)] = rx[it];
    }
    __syncthreads();
    if (lane < ncols) {
#pragma unroll
      for (int q = 0; q < KCH; ++q) {
        const int i = r0 + q;
        if (i < NROW) {
          const T* v = sb + lane * PITCH_B + 3 * q;
          const T a = v[0], d = v[1], u = v[2], xv = sx[lane * PITCH_X + q];
          if (i == 0) {
            const T inv = T(1) / d;
            cp[i] = inv * u;
            dp[i] = inv * xv;
          } else {
            const T ac = a * cp[i - 1];
            const T inv = T(1) / (d - ac);
            if (i < NROW - 1) cp[i] = inv * u;
            const T ad = a * dp[i - 1];
            dp[i] = inv * (xv - ad);
          }
        }
      }
    }
  }
#pragma unroll
  for (int i = NROW - 2; i >= 0; --i) {
    const T cd = cp[i] * dp[i + 1];
    dp[i] = dp[i] - cd;
  }
#pragma unroll
  for (int s = 0; s < NST; ++s) {
    const int r0 = s * KCH;
    const int nr = (NROW - r0) < KCH ? (NROW - r0) : KCH;
    __syncthreads();
    if (lane < ncols) {
#pragma unroll
      for (int q = 0; q < KCH; ++q)
        if (r0 + q < NROW) sx[lane * PITCH_X + q] = dp[r0 + q];
    }
    __syncthreads();
    const int nx = ncols * nr;
    for (int e = lane; e < nx; e += kCols) {
      const int cc = e / nr;
      const int q = e - cc * nr;
      lhs[sbase[cc] + sind[r0 + q]] = sx[cc * PITCH_X + q];
    }
  }
}

// columnwise_op_appinv, FP32, every load issued at entry.  A C96 W3 solve
// has about one wave per SIMD, so a wave that loads a stage, waits, solves
// it and loads the next spends its life on memory latency (measured: each
// stage's loads cost ~2 us, the whole kernel ran as long as one wave's chain).
// Here a wave owns COLS columns (COLS of its 64 lanes run the recurrence) and
// issues all its loads up front: the bands of every KCH-row stage into
// registers as 8-byte pairs (NST * SLOTS per lane; a column's bands are
// 3 NROW contiguous values, so stage s of column cs is the run
// [3 KCH s, 3 KCH (s+1)) of it) and the right-hand side of all NROW rows of
// its columns in flat order (64 consecutive dofs per instruction when the
// columns are contiguous).  The right-hand side goes to LDS, each band stage
// goes to LDS before its rows are solved (odd pitches: conflict-free per-lane
// reads), the solution goes back through LDS and out in the same flat order
// with streaming stores.  The recurrence is k_appinv's term for term, so the
// bits are the same.
// DPL: d' lives in the column's right-hand-side slots in LDS (each row reads
// its right-hand side once, then holds d'), so a lane keeps only c' in
// registers (about 70 fewer VGPRs, more waves per SIMD) and the backward
// sweep leaves the solution in place for the store.
template <int NROW, int KCH, int COLS, bool DPL = false>
__global__ void __launch_bounds__(kCols)
k_appinv_deep(float* __restrict__ lhs, const float* __restrict__ x, const float* __restrict__ cm,
              const int32_t* __restrict__ map0, int ndf, const int32_t* __restrict__ ind, int ncell) {
  static_assert(NROW % KCH == 0 && (3 * KCH) % 2 == 0 && COLS <= kCols, "stage shape");
  constexpr int NST = NROW / KCH;
  constexpr int NB = 3 * KCH;                        // band values of one column per stage
  constexpr int NBV = NB / 2;                        // 8-byte pairs of them
  constexpr int PB = NB + 1;                         // odd pitches
  constexpr int PX = NROW | 1;
  constexpr int SLOTS = (COLS * NBV + kCols - 1) / kCols;
  constexpr int NXF = (COLS * NROW + kCols - 1) / kCols;
  __shared__ float sb[COLS * PB];
  __shared__ float sxa[COLS * PX];
  __shared__ int32_t sbase[COLS];
  __shared__ int32_t sind[NROW];
  const int lane = threadIdx.x;
  const int c0 = blockIdx.x * COLS;
  const int ncols = min(COLS, ncell - c0);
  // the column bases and the row indirection first, unconditionally at
  // clamped (in-bounds) indices: the right-hand side's addresses depend on
  // them, and loaded after the bands (or behind a branch) their wait would
  // also wait for every band load
  static_assert(NROW <= 2 * kCols, "two indirection entries a lane");
  const int32_t mbv = map0[(int64_t)(c0 + min(lane, ncols - 1)) * ndf];
  const int32_t ir0 = ind[min(lane, NROW - 1)];
  const int32_t ir1 = ind[min(lane + kCols, NROW - 1)];
  // bands: pair slot it of stage s is pair qv of column cs (lane + it kCols =
  // cs NBV + qv), clamped into the workgroup's columns
  const float* cmw = cm + (int64_t)c0 * 3 * NROW;
  float2 rb[NST][SLOTS];
#pragma unroll
  for (int s = 0; s < NST; ++s)
#pragma unroll
    for (int it = 0; it < SLOTS; ++it) {
      const int e = min(lane + it * kCols, COLS * NBV - 1);
      const int cs = e / NBV, qv = e - cs * NBV;
      rb[s][it] = *reinterpret_cast<const float2*>(cmw + min(cs, ncols - 1) * 3 * NROW + s * NB + 2 * qv);
    }
  // an unconditional use of the three values after the band loads, so that
  // the compiler does not move their loads down into the branches below
  // (behind the band loads); the statement emits no instruction
  asm volatile("" ::"v"(mbv), "v"(ir0), "v"(ir1));
  if (lane < ncols) sbase[lane] = mbv;
  if (lane < NROW) sind[lane] = min(max(ir0 - 1, 0), NROW - 1);   // as ind0
  if (lane + kCols < NROW) sind[lane + kCols] = min(max(ir1 - 1, 0), NROW - 1);
  __syncthreads();
  // right-hand side, flat slot e = column c row r (e = c NROW + r)
  const int nflat = ncols * NROW;
  {
    float rf[NXF];
#pragma unroll
    for (int it = 0; it < NXF; ++it) {
      const int e = min(lane + it * kCols, nflat - 1);
      const int c = e / NROW, r = e - c * NROW;
      rf[it] = x[sbase[c] + sind[r]];
    }
#pragma unroll
    for (int it = 0; it < NXF; ++it) {
      const int e = lane + it * kCols;
      if (e < nflat) {
        const int c = e / NROW;
        sxa[c * PX + (e - c * NROW)] = rf[it];
      }
    }
  }
  float cp[NROW], dp[DPL ? 1 : NROW];
  float dprev = 0.f;   // DPL: d' of the previous row
#pragma unroll
  for (int s = 0; s < NST; ++s) {
    __syncthreads();   // the previous stage's reads are done (and, first time, the rhs is in)
#pragma unroll
    for (int it = 0; it < SLOTS; ++it) {
      const int e = lane + it * kCols;
      if (e < COLS * NBV) {
        const int cs = e / NBV, qv = e - cs * NBV;
        sb[cs * PB + 2 * qv] = rb[s][it].x;
        sb[cs * PB + 2 * qv + 1] = rb[s][it].y;
      }
    }
    __syncthreads();
    if (lane < COLS) {
#pragma unroll
      for (int q = 0; q < KCH; ++q) {
        const int i = s * KCH + q;
        const float* v = sb + lane * PB + 3 * q;
        const float a = v[0], d = v[1], u = v[2], xv = sxa[lane * PX + i];
        float di;
        if (i == 0) {
          const float inv = 1.0f / d;
          cp[i] = inv * u;
          di = inv * xv;
        } else {
          const float ac = a * cp[i - 1];
          const float inv = 1.0f / (d - ac);
          if (i < NROW - 1) cp[i] = inv * u;
          float dprv;
          if constexpr (DPL) dprv = dprev;
          else dprv = dp[i - 1];
          const float ad = a * dprv;
          di = inv * (xv - ad);
        }
        if constexpr (DPL) {
          sxa[lane * PX + i] = di;
          dprev = di;
        } else {
          dp[i] = di;
        }
      }
    }
  }
  if constexpr (DPL) {
    if (lane < COLS) {
      float dnext = dprev;   // d'(NROW-1), already the solution there
#pragma unroll
      for (int i = NROW - 2; i >= 0; --i) {
        const float cd = cp[i] * dnext;
        const float di = sxa[lane * PX + i] - cd;
        sxa[lane * PX + i] = di;
        dnext = di;
      }
    }
    __syncthreads();
  } else {
#pragma unroll
    for (int i = NROW - 2; i >= 0; --i) {
      const float cd = cp[i] * dp[i + 1];
      dp[i] = dp[i] - cd;
    }
    __syncthreads();
    if (lane < COLS) {
#pragma unroll
      for (int i = 0; i < NROW; ++i) sxa[lane * PX + i] = dp[i];
    }
    __syncthreads();
  }
#pragma unroll
  for (int it = 0; it < NXF; ++it) {
    const int e = lane + it * kCols;
    if (e < nflat) {
      const int c = e / NROW, r = e - c * NROW;
      __builtin_nontemporal_store(sxa[c * PX + r], lhs + sbase[c] + sind[r]);
    }
  }
}

// FP32 appinv kernel choice: 1 the unpipelined staged kernel, 0 the per-lane
// k_appinv (LFR_APPINV_STAGED=0/1); otherwise the up-front-load kernel.
int appinv_variant() {
  static const int v = [] {
    const char* e = getenv("LFR_APPINV_STAGED");
    if (e && e[0] == '0' && !e[1]) return 0;
    if (e && e[0] == '1' && !e[1]) return 1;
    return 2;
  }();
  return v;
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
    const C xv = (C)x[base + ind0(ind, i, nrow)];
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
    lhs[base + ind0(ind, i, nrow)] = (T)dpc[i];
  }
}

template <typename T, typename C>
int launch_appinv(lfr_ctx* ctx, T* lhs, const T* x, const T* cm, const lfr_space* sp,
                  const int32_t* ind, int nrow, int ncell) {
  const unsigned nb = (unsigned)((ncell + kCols - 1) / kCols);
  if constexpr (sizeof(T) == 4 && sizeof(C) == 4) {
    const int v = nrow == 70 ? appinv_variant() : 0;
    if (v == 1) {
      hipLaunchKernelGGL((k_appinv_staged<T, 70>), dim3(nb), dim3(kCols), 0, ctx->stream, lhs, x, cm, sp->map0_d,
                         sp->ndf, ind, ncell);
      LFR_CHECK_LAUNCH(ctx);
      return LFR_OK;
    }
    // 8-byte band pairs need an 8-byte aligned matrix (columns are 840 bytes apart)
    if (v == 2 && ((uintptr_t)cm & 7) == 0) {
      // result-identical variants: columns a wave x rows a stage, and the
      // staged kernel.  Default: 32 columns a wave while the grid fits two
      // waves per SIMD (all resident at once), 16 beyond.  The autotuner
      // times them on the first call of a shape when lhs does not alias x
      // (every variant then rewrites the same lhs from the same inputs).
      const auto launch = [&](int c) -> int {
        switch (c) {
          case 0:
            hipLaunchKernelGGL((k_appinv_deep<70, 14, 32>), dim3((ncell + 31) / 32), dim3(kCols), 0, ctx->stream,
                               lhs, x, cm, sp->map0_d, sp->ndf, ind, ncell);
            break;
          case 1:
            hipLaunchKernelGGL((k_appinv_deep<70, 10, 16>), dim3((ncell + 15) / 16), dim3(kCols), 0, ctx->stream,
                               lhs, x, cm, sp->map0_d, sp->ndf, ind, ncell);
            break;
          case 2:
            hipLaunchKernelGGL((k_appinv_deep<70, 14, 16>), dim3((ncell + 15) / 16), dim3(kCols), 0, ctx->stream,
                               lhs, x, cm, sp->map0_d, sp->ndf, ind, ncell);
            break;
          case 3:
            hipLaunchKernelGGL((k_appinv_deep<70, 10, 32>), dim3((ncell + 31) / 32), dim3(kCols), 0, ctx->stream,
                               lhs, x, cm, sp->map0_d, sp->ndf, ind, ncell);
            break;
          case 5:   // d' in LDS
            hipLaunchKernelGGL((k_appinv_deep<70, 10, 16, true>), dim3((ncell + 15) / 16), dim3(kCols), 0,
                               ctx->stream, lhs, x, cm, sp->map0_d, sp->ndf, ind, ncell);
            break;
          case 6:
            hipLaunchKernelGGL((k_appinv_deep<70, 14, 32, true>), dim3((ncell + 31) / 32), dim3(kCols), 0,
                               ctx->stream, lhs, x, cm, sp->map0_d, sp->ndf, ind, ncell);
            break;
          case 7:
            hipLaunchKernelGGL((k_appinv_deep<70, 14, 16, true>), dim3((ncell + 15) / 16), dim3(kCols), 0,
                               ctx->stream, lhs, x, cm, sp->map0_d, sp->ndf, ind, ncell);
            break;
          default:
            hipLaunchKernelGGL((k_appinv_staged<T, 70>), dim3(nb), dim3(kCols), 0, ctx->stream, lhs, x, cm,
                               sp->map0_d, sp->ndf, ind, ncell);
        }
        return hipGetLastError() == hipSuccess ? LFR_OK : LFR_ERR_HIP;
      };
      const int dflt = (ncell + 31) / 32 <= 2048 ? 0 : 7;
      int c = dflt;
      if ((const void*)lhs != (const void*)x) {
        char key[96];
        snprintf(key, sizeof key, "appinv f32 nrow=70 ncell=%d", ncell);
        c = lfr_tune_pick(ctx, key, 8, dflt, launch);
      }
      if (launch(c) != LFR_OK) return lfr_set_error(ctx, LFR_ERR_HIP, "appinv: launch failed");
      return LFR_OK;
    }
  }
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
  // every column's rows map(1) .. map(1)+nrow-1 (1-based) must be dofs of the space
  LFR_REQUIRE(ctx, (int64_t)space->max_first + nrow - 1 <= space->undf,
              "lfr_columnwise_op_appinv: nrow %d runs past undf %d from a column starting at dof %d", nrow,
              space->undf, space->max_first);
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

// ---------------------------------------------------------------------------
// Banded (CMA) geometry handle, assembly and apply
// ---------------------------------------------------------------------------
struct lfr_cma {
  lfr_ctx* ctx = nullptr;
  int nrow = 0, ncol = 0, bw = 0, alpha = 0, beta = 0, gamma_m = 0, gamma_p = 0;
  int ndf_to = 0, ndf_from = 0, nlayers = 0;
  int32_t* ind_to_d = nullptr;      // indirection_dofmap_to(nrow), 1-based
  int32_t* ind_from_d = nullptr;    // indirection_dofmap_from(ncol), 1-based
  int32_t* banded_from_d = nullptr; // column_banded_dofmap_from(ndf_from, nlayers)
  int32_t* row_ptr_d = nullptr;     // per row i: its (k, df1) entries, in (k, df1) order
  int32_t* row_ent_d = nullptr;     // k * ndf_to + df1
  int32_t* asm_tab_d = nullptr;     // per row: count, then (k, df1 + ndf_to df2, band slot) per contribution
  int32_t* app_tab_d = nullptr;     // per row: output offset, then bw input offsets (-1 outside [lo, hi])
};

namespace {

// band limit of row i (1-based), formed in the operator's real kind as the
// reference does: ceiling((alpha*i - gamma_p) / (1.0_r_solver*beta))
template <typename T>
__device__ __forceinline__ int band_lo(int i, int alpha, int beta, int gamma_p) {
  return (int)ceil((T)(alpha * i - gamma_p) / ((T)1 * beta));
}
template <typename T>
__device__ __forceinline__ int band_hi(int i, int alpha, int beta, int gamma_m) {
  return (int)floor((T)(alpha * i + gamma_m) / ((T)1 * beta));
}

// columnwise_op_asm_kernel_code, thread = (column, row i): cm(:, i, cell)
// accumulates its (k, df1, df2) contributions in the reference loop order.
template <typename T, int BW>
__global__ void __launch_bounds__(256)
k_cma_asm(T* __restrict__ cm, const T* __restrict__ ls, const int32_t* __restrict__ row_ptr,
          const int32_t* __restrict__ row_ent, const int32_t* __restrict__ banded_from, int ncell, int nl,
          int nrow, int ndf_to, int ndf_from, int alpha, int beta, int gamma_p) {
  const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= (int64_t)ncell * nrow) return;
  const int cell = (int)(t / nrow), i0 = (int)(t - (int64_t)cell * nrow);
  const int64_t n3 = (int64_t)ncell * nl;
  const int jm = band_lo<T>(i0 + 1, alpha, beta, gamma_p);
  T acc[BW];
#pragma unroll
  for (int b = 0; b < BW; ++b) acc[b] = T(0);
  for (int e = row_ptr[i0]; e < row_ptr[i0 + 1]; ++e) {
    const int k = row_ent[e] / ndf_to, df1 = row_ent[e] - k * ndf_to;
    const int64_t ik = (int64_t)cell * nl + k;
    for (int df2 = 0; df2 < ndf_from; ++df2) {
      const int b = banded_from[df2 + ndf_from * k] - jm;
      const T v = ls[ik + n3 * (df1 + (int64_t)ndf_to * df2)];
#pragma unroll
      for (int q = 0; q < BW; ++q)
        if (q == b) acc[q] = acc[q] + v;
    }
  }
  T* out = cm + ((int64_t)cell * nrow + i0) * BW;
#pragma unroll
  for (int b = 0; b < BW; ++b) out[b] = acc[b];
}

// (cell, row) of item b + threadIdx.x, items numbered row-fastest, with b
// uniform across the workgroup: one division per workgroup (scalar) and,
// for the lane's offset, a multiply by minv = ceil(2^20 / nrow) and a shift,
// exact for offsets below 256 + nrow when (256 + nrow) nrow < 2^20 (checked
// for every nrow <= kSplitMaxRow).  A division per lane (64-bit for the
// assembly) cost more time than the kernel's memory traffic.
constexpr int kSplitMaxRow = 768;
__device__ __forceinline__ void split_item(unsigned b, unsigned nrow, unsigned minv, int& cell, int& row) {
  const unsigned c0 = b / nrow;
  const unsigned local = threadIdx.x + (b - c0 * nrow);
  const unsigned dq = (local * minv) >> 20;
  cell = (int)(c0 + dq);
  row = (int)(local - dq * nrow);
}
unsigned split_minv(int nrow) { return (unsigned)(((1u << 20) + (unsigned)nrow - 1) / (unsigned)nrow); }

// The same assembly from a per-row table (built at creation) of at most
// kAsmMc contributions in the reference loop order (entries (k, df1), then
// df2): one table load, then every stencil load at once (unused slots read
// layer 0, plane 0 and are not added), then the sums in table order.  The loop form above waits for an index load before each stencil
// load.
constexpr int kAsmMc = 4;
constexpr int kAsmStride = 16;   // 1 + 3 kAsmMc, padded to whole int4
template <typename T, int BW, int IT>
__global__ void __launch_bounds__(256)
k_cma_asm_tab(T* __restrict__ cm, const T* __restrict__ ls, const int32_t* __restrict__ tab, int ncell, int nl,
              int nrow, unsigned minv) {
  // IT items a grid stride apart per thread: both table records, then all
  // their stencil loads, are in flight together (each item's stencil loads
  // wait for its record)
  const unsigned n = (unsigned)ncell * (unsigned)nrow;
  const unsigned stride = gridDim.x * 256u;
  const unsigned n3u = (unsigned)ncell * (unsigned)nl;
  int cell[IT], i0[IT];
  bool ok[IT];
  int r[IT][kAsmStride];
#pragma unroll
  for (int u = 0; u < IT; ++u) {
    const unsigned b = blockIdx.x * 256u + u * stride;
    ok[u] = b + threadIdx.x < n;
    split_item(b, (unsigned)nrow, minv, cell[u], i0[u]);
    if (!ok[u]) cell[u] = i0[u] = 0;
    const int4* r4 = reinterpret_cast<const int4*>(tab + (int64_t)i0[u] * kAsmStride);
#pragma unroll
    for (int q = 0; q < kAsmStride / 4; ++q) {
      const int4 v = r4[q];
      r[u][4 * q] = v.x;
      r[u][4 * q + 1] = v.y;
      r[u][4 * q + 2] = v.z;
      r[u][4 * q + 3] = v.w;
    }
  }
  // 32-bit offsets: the launch checks that the stencil and band arrays have
  // fewer than 2^31 entries; unused slots hold layer 0, plane 0 (in bounds;
  // their values are not added)
  T v[IT][kAsmMc];
#pragma unroll
  for (int u = 0; u < IT; ++u) {
    const unsigned base = (unsigned)cell[u] * (unsigned)nl;
#pragma unroll
    for (int m = 0; m < kAsmMc; ++m)
      v[u][m] = ls[base + (unsigned)r[u][1 + 3 * m] + n3u * (unsigned)r[u][2 + 3 * m]];
  }
  // branch-free sums: a slot that the reference loop does not reach adds
  // +0, which leaves the sum's bits unchanged (it starts at +0 and, rounding
  // to nearest, never becomes -0)
#pragma unroll
  for (int u = 0; u < IT; ++u) {
    const int cnt = r[u][0];
    T acc[BW];
#pragma unroll
    for (int b = 0; b < BW; ++b) acc[b] = T(0);
#pragma unroll
    for (int m = 0; m < kAsmMc; ++m) {
      const int b = m < cnt ? r[u][3 + 3 * m] : -1;
#pragma unroll
      for (int q = 0; q < BW; ++q) acc[q] = acc[q] + (q == b ? v[u][m] : T(0));
    }
    if (ok[u]) {
      T* out = cm + ((unsigned)cell[u] * (unsigned)nrow + (unsigned)i0[u]) * (unsigned)BW;
#pragma unroll
      for (int b = 0; b < BW; ++b) out[b] = acc[b];
    }
  }
}

// columnwise_op_app_kernel_code with per-row tables: row i0 of every column
// adds its band entries b = j - jm for the columns j in [max(1, jm),
// min(ncol, jp)] in ascending order, as the reference loop does; the band
// limits and the indirection offsets depend only on the row, so the host
// forms them once (tab[i0 * (BW + 1)] = ind_to(i) - 1, then for each b the
// input offset ind_from(j) - 1 or -1 outside the loop range).  A thread
// takes IT (column, row) items a grid stride apart and issues all of their
// loads before any arithmetic; each item's sum is the reference's.
template <typename T, int BW, int IT>
__global__ void __launch_bounds__(256)
k_cma_app_tab(T* __restrict__ lhs, const T* __restrict__ x, const T* __restrict__ cm,
              const int32_t* __restrict__ map_to, const int32_t* __restrict__ map_from,
              const int32_t* __restrict__ tab, int ncell, int nrow, int ndf_to, int ndf_from, unsigned minv) {
  const int n = ncell * nrow;
  const int stride = gridDim.x * 256;
  for (int base = blockIdx.x * 256; base < n; base += stride * IT) {
    T acc[IT], cv[IT][BW], xv[IT][BW];
    int64_t mu[IT];
    int xo[IT][BW];
    bool ok[IT];
#pragma unroll
    for (int u = 0; u < IT; ++u) {
      const int idx = base + u * stride + (int)threadIdx.x;
      ok[u] = idx < n;
      int cell, i0;
      split_item((unsigned)(base + u * stride), (unsigned)nrow, minv, cell, i0);
      if (!ok[u]) cell = i0 = 0;
      const int32_t* r = tab + i0 * (BW + 1);
      mu[u] = (int64_t)map_to[(int64_t)cell * ndf_to] + r[0];
      const int64_t bf = map_from[(int64_t)cell * ndf_from];
      const T* c = cm + ((int64_t)cell * nrow + i0) * BW;
      // every load is unconditional at an in-bounds address (item 0 past
      // the end, input offset 0 outside the loop range); the values that
      // the reference loop does not reach are never added.  Conditional
      // loads would each wait for all loads before them.
      acc[u] = lhs[mu[u]];
#pragma unroll
      for (int b = 0; b < BW; ++b) {
        xo[u][b] = r[1 + b];
        cv[u][b] = c[b];
        xv[u][b] = x[bf + max(xo[u][b], 0)];
      }
    }
#pragma unroll
    for (int u = 0; u < IT; ++u) {
      if (!ok[u]) continue;
      T a = acc[u];
#pragma unroll
      for (int b = 0; b < BW; ++b)
        if (xo[u][b] >= 0) {
          const T prod = cv[u][b] * xv[u][b];
          a = a + prod;
        }
      lhs[mu[u]] = a;
    }
  }
}

// columnwise_op_app_kernel_code, thread = (column, row i): lhs(mu_i) += sum_j
template <typename T>
__global__ void __launch_bounds__(256)
k_cma_app(T* __restrict__ lhs, const T* __restrict__ x, const T* __restrict__ cm,
          const int32_t* __restrict__ map_to, const int32_t* __restrict__ map_from,
          const int32_t* __restrict__ ind_to, const int32_t* __restrict__ ind_from, int ncell, int nrow,
          int ncol, int bw, int ndf_to, int ndf_from, int alpha, int beta, int gamma_m, int gamma_p) {
  const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= (int64_t)ncell * nrow) return;
  const int cell = (int)(t / nrow), i0 = (int)(t - (int64_t)cell * nrow);
  const int i = i0 + 1;
  const int64_t mu_i = (int64_t)map_to[(int64_t)cell * ndf_to] + ind_to[i0] - 1;
  const int64_t base_from = map_from[(int64_t)cell * ndf_from];
  const int jm = band_lo<T>(i, alpha, beta, gamma_p);
  const int jp = band_hi<T>(i, alpha, beta, gamma_m);
  const int lo = jm > 1 ? jm : 1, hi = jp < ncol ? jp : ncol;
  const T* c = cm + ((int64_t)cell * nrow + i0) * bw;
  T acc = lhs[mu_i];
  for (int j = lo; j <= hi; ++j) {
    const T prod = c[j - jm] * x[base_from + ind_from[j - 1] - 1];
    acc = acc + prod;
  }
  lhs[mu_i] = acc;
}

void cma_free(lfr_cma* h) {
  if (!h) return;
  for (void* p : {(void*)h->ind_to_d, (void*)h->ind_from_d, (void*)h->banded_from_d, (void*)h->row_ptr_d,
                  (void*)h->row_ent_d, (void*)h->app_tab_d, (void*)h->asm_tab_d})
    if (p) (void)hipFree(p);
  delete h;
}

}  // namespace

extern "C" {

int lfr_cma_create(lfr_ctx* ctx, int nlayers, int nrow, int ncol, int bandwidth, int alpha, int beta,
                   int gamma_m, int gamma_p, int ndf_to, const int32_t* banded_to_h, const int32_t* ind_to_h,
                   int ndf_from, const int32_t* banded_from_h, const int32_t* ind_from_h, lfr_cma** out) {
  if (!ctx || !out) return LFR_ERR_ARG;
  *out = nullptr;
  LFR_REQUIRE(ctx, nlayers > 0 && nrow > 0 && ncol > 0 && bandwidth >= 1 && bandwidth <= 8 && beta > 0 &&
                       ndf_to > 0 && ndf_from > 0 && banded_to_h && ind_to_h && banded_from_h && ind_from_h,
              "lfr_cma_create: bad arguments");
  // validate the maps against nrow/ncol and the band
  for (int q = 0; q < nrow; ++q)
    LFR_REQUIRE(ctx, ind_to_h[q] >= 1 && ind_to_h[q] <= nrow, "lfr_cma_create: indirection_dofmap_to out of range");
  for (int q = 0; q < ncol; ++q)
    LFR_REQUIRE(ctx, ind_from_h[q] >= 1 && ind_from_h[q] <= ncol,
                "lfr_cma_create: indirection_dofmap_from out of range");
  std::vector<std::vector<int32_t>> rows(nrow);
  for (int k = 0; k < nlayers; ++k)
    for (int df1 = 0; df1 < ndf_to; ++df1) {
      const int i = banded_to_h[df1 + ndf_to * k];
      LFR_REQUIRE(ctx, i >= 1 && i <= nrow, "lfr_cma_create: column_banded_dofmap_to out of range");
      const int jm = (int)std::ceil((double)(alpha * i - gamma_p) / (double)beta);
      for (int df2 = 0; df2 < ndf_from; ++df2) {
        const int j = banded_from_h[df2 + ndf_from * k];
        LFR_REQUIRE(ctx, j >= 1 && j <= ncol && j - jm >= 0 && j - jm < bandwidth,
                    "lfr_cma_create: entry (row %d, column %d) outside the band", i, j);
      }
      rows[i - 1].push_back(k * ndf_to + df1);   // appended in (k, df1) order
    }
  std::vector<int32_t> ptr(nrow + 1, 0), ent;
  for (int i = 0; i < nrow; ++i) {
    ptr[i + 1] = ptr[i] + (int)rows[i].size();
    ent.insert(ent.end(), rows[i].begin(), rows[i].end());
  }
  lfr_cma* h = new lfr_cma();
  h->ctx = ctx;
  h->nrow = nrow;
  h->ncol = ncol;
  h->bw = bandwidth;
  h->alpha = alpha;
  h->beta = beta;
  h->gamma_m = gamma_m;
  h->gamma_p = gamma_p;
  h->ndf_to = ndf_to;
  h->ndf_from = ndf_from;
  h->nlayers = nlayers;
  const size_t nbf = (size_t)ndf_from * nlayers;
  const size_t nent = ent.empty() ? 1 : ent.size();
  if (hipMalloc(&h->ind_to_d, nrow * 4) != hipSuccess || hipMalloc(&h->ind_from_d, ncol * 4) != hipSuccess ||
      hipMalloc(&h->banded_from_d, nbf * 4) != hipSuccess || hipMalloc(&h->row_ptr_d, (nrow + 1) * 4) != hipSuccess ||
      hipMalloc(&h->row_ent_d, nent * 4) != hipSuccess) {
    cma_free(h);
    return lfr_set_error(ctx, LFR_ERR_NOMEM, "lfr_cma_create: allocation failed");
  }
  hipError_t e = hipMemcpy(h->ind_to_d, ind_to_h, nrow * 4, hipMemcpyHostToDevice);
  if (e == hipSuccess) e = hipMemcpy(h->ind_from_d, ind_from_h, ncol * 4, hipMemcpyHostToDevice);
  if (e == hipSuccess) e = hipMemcpy(h->banded_from_d, banded_from_h, nbf * 4, hipMemcpyHostToDevice);
  if (e == hipSuccess) e = hipMemcpy(h->row_ptr_d, ptr.data(), (nrow + 1) * 4, hipMemcpyHostToDevice);
  if (e == hipSuccess && !ent.empty()) e = hipMemcpy(h->row_ent_d, ent.data(), ent.size() * 4, hipMemcpyHostToDevice);
  if (e != hipSuccess) {
    cma_free(h);
    return lfr_set_error(ctx, LFR_ERR_HIP, "lfr_cma_create: %s", hipGetErrorString(e));
  }
  // per-row table of the assembly (k_cma_asm_tab) when no row has more than
  // kAsmMc contributions; unused slots name layer 0, plane 0 (in bounds)
  {
    bool fits = true;
    for (int i0 = 0; i0 < nrow && fits; ++i0) fits = (int)rows[i0].size() * ndf_from <= kAsmMc;
    if (fits) {
      std::vector<int32_t> at((size_t)nrow * kAsmStride, 0);
      for (int i0 = 0; i0 < nrow; ++i0) {
        int32_t* r = &at[(size_t)i0 * kAsmStride];
        const int jm = (int)std::ceil((double)(alpha * (i0 + 1) - gamma_p) / (double)beta);
        int m = 0;
        for (int ent_v : rows[i0]) {
          const int k = ent_v / ndf_to, df1 = ent_v - k * ndf_to;
          for (int df2 = 0; df2 < ndf_from; ++df2, ++m) {
            r[1 + 3 * m] = k;
            r[2 + 3 * m] = df1 + ndf_to * df2;
            r[3 + 3 * m] = banded_from_h[df2 + ndf_from * k] - jm;
          }
        }
        r[0] = m;
      }
      if (hipMalloc(&h->asm_tab_d, at.size() * 4) != hipSuccess ||
          hipMemcpy(h->asm_tab_d, at.data(), at.size() * 4, hipMemcpyHostToDevice) != hipSuccess) {
        cma_free(h);
        return lfr_set_error(ctx, LFR_ERR_NOMEM, "lfr_cma_create: assembly table allocation failed");
      }
    }
  }
  // per-row table of the apply: band limits formed as the reference does
  // (exact for these small integers), usable when every column of a row's
  // loop range falls inside the band
  {
    const int S = bandwidth + 1;
    std::vector<int32_t> tab((size_t)nrow * S, -1);
    bool ok = true;
    for (int i0 = 0; i0 < nrow && ok; ++i0) {
      const int i = i0 + 1;
      const int jm = (int)std::ceil((double)(alpha * i - gamma_p) / (double)beta);
      const int jp = (int)std::floor((double)(alpha * i + gamma_m) / (double)beta);
      const int lo = std::max(jm, 1), hi = std::min(jp, ncol);
      tab[(size_t)i0 * S] = ind_to_h[i0] - 1;
      for (int j = lo; j <= hi; ++j) {
        if (j - jm < 0 || j - jm >= bandwidth) {
          ok = false;
          break;
        }
        tab[(size_t)i0 * S + 1 + (j - jm)] = ind_from_h[j - 1] - 1;
      }
    }
    if (ok && (bandwidth <= 5 || bandwidth == 8)) {
      if (hipMalloc(&h->app_tab_d, tab.size() * 4) != hipSuccess ||
          hipMemcpy(h->app_tab_d, tab.data(), tab.size() * 4, hipMemcpyHostToDevice) != hipSuccess) {
        cma_free(h);
        return lfr_set_error(ctx, LFR_ERR_NOMEM, "lfr_cma_create: apply table allocation failed");
      }
    }
  }
  *out = h;
  return LFR_OK;
}

int lfr_cma_destroy(lfr_cma* h) {
  cma_free(h);
  return LFR_OK;
}

int lfr_columnwise_op_asm(lfr_ctx* ctx, int prec, const lfr_cma* h, const void* local_stencil_d, void* cm_d,
                          int ncell_2d) {
  if (!ctx || !h) return LFR_ERR_ARG;
  LFR_REQUIRE(ctx, (prec == 4 || prec == 8) && local_stencil_d && cm_d && ncell_2d > 0,
              "lfr_columnwise_op_asm: bad arguments");
  const int64_t n = (int64_t)ncell_2d * h->nrow;
  const unsigned g = (unsigned)((n + 255) / 256);
  constexpr int kAsmIt = 2;   // 4 measured the same (32.9 against 32.8 µs)
  const unsigned gi = (unsigned)std::max<int64_t>(1, (n + 256 * kAsmIt - 1) / (256 * kAsmIt));
  hipStream_t s = ctx->stream;
#define LFR_ASM(T, BWV)                                                                                   \
  if (h->asm_tab_d && h->nrow <= kSplitMaxRow && n * 8 < ((int64_t)1 << 31) &&                             \
      (int64_t)ncell_2d * h->nlayers * h->ndf_to * h->ndf_from < ((int64_t)1 << 31))                          \
    hipLaunchKernelGGL((k_cma_asm_tab<T, BWV, kAsmIt>), dim3(gi), dim3(256), 0, s, (T*)cm_d,                \
                       (const T*)local_stencil_d, h->asm_tab_d, ncell_2d, h->nlayers, h->nrow,              \
                       split_minv(h->nrow));                                                                \
  else                                                                                                      \
    hipLaunchKernelGGL((k_cma_asm<T, BWV>), dim3(g), dim3(256), 0, s, (T*)cm_d, (const T*)local_stencil_d,   \
                       h->row_ptr_d, h->row_ent_d, h->banded_from_d, ncell_2d, h->nlayers, h->nrow, h->ndf_to, \
                       h->ndf_from, h->alpha, h->beta, h->gamma_p)
#define LFR_ASM_BW(T)                      \
  switch (h->bw) {                         \
    case 1: LFR_ASM(T, 1); break;          \
    case 2: LFR_ASM(T, 2); break;          \
    case 3: LFR_ASM(T, 3); break;          \
    case 4: LFR_ASM(T, 4); break;          \
    case 5: LFR_ASM(T, 5); break;          \
    default: LFR_ASM(T, 8); break;         \
  }
  if (h->bw > 5 && h->bw != 8) return lfr_set_error(ctx, LFR_ERR_ARG, "lfr_columnwise_op_asm: bandwidth %d", h->bw);
  if (prec == 4) {
    LFR_ASM_BW(float)
  } else {
    LFR_ASM_BW(double)
  }
#undef LFR_ASM_BW
#undef LFR_ASM
  LFR_CHECK_LAUNCH(ctx);
  return LFR_OK;
}

int lfr_columnwise_op_app(lfr_ctx* ctx, int prec, const lfr_cma* h, void* lhs_d, const void* x_d,
                          const void* cm_d, const lfr_space* to_space, const lfr_space* from_space) {
  if (!ctx || !h) return LFR_ERR_ARG;
  LFR_REQUIRE(ctx, (prec == 4 || prec == 8) && lhs_d && x_d && cm_d && to_space && from_space,
              "lfr_columnwise_op_app: bad arguments");
  LFR_REQUIRE(ctx, to_space->ndf == h->ndf_to && from_space->ndf == h->ndf_from &&
                       to_space->ncell == from_space->ncell && to_space->nlayers == h->nlayers,
              "lfr_columnwise_op_app: spaces do not match the operator geometry");
  const int ncell = to_space->ncell;
  const int64_t n = (int64_t)ncell * h->nrow;
  if (h->app_tab_d && h->nrow <= kSplitMaxRow && n < ((int64_t)1 << 31)) {
    // two items per thread (C96L70 Wtheta FP32: 37.1 µs at one, 32.1 at
    // two, 33.2 at four; 38.3 with the per-thread band arithmetic)
    constexpr int IT = 2;
    const unsigned gt = (unsigned)std::max<int64_t>(1, (n + 256 * IT - 1) / (256 * IT));
#define LFR_APP(T, BWV)                                                                                             \
  hipLaunchKernelGGL((k_cma_app_tab<T, BWV, IT>), dim3(gt), dim3(256), 0, ctx->stream, (T*)lhs_d, (const T*)x_d, \
                     (const T*)cm_d, to_space->map0_d, from_space->map0_d, h->app_tab_d, ncell, h->nrow, h->ndf_to, \
                     h->ndf_from, split_minv(h->nrow))
#define LFR_APP_BW(T)                 \
  switch (h->bw) {                    \
    case 1: LFR_APP(T, 1); break;     \
    case 2: LFR_APP(T, 2); break;     \
    case 3: LFR_APP(T, 3); break;     \
    case 4: LFR_APP(T, 4); break;     \
    case 5: LFR_APP(T, 5); break;     \
    default: LFR_APP(T, 8); break;    \
  }
    if (prec == 4) {
      LFR_APP_BW(float)
    } else {
      LFR_APP_BW(double)
    }
#undef LFR_APP_BW
#undef LFR_APP
    LFR_CHECK_LAUNCH(ctx);
    return LFR_OK;
  }
  const unsigned g = (unsigned)((n + 255) / 256);
  if (prec == 4)
    hipLaunchKernelGGL(k_cma_app<float>, dim3(g), dim3(256), 0, ctx->stream, (float*)lhs_d, (const float*)x_d,
                       (const float*)cm_d, to_space->map0_d, from_space->map0_d, h->ind_to_d, h->ind_from_d, ncell,
                       h->nrow, h->ncol, h->bw, h->ndf_to, h->ndf_from, h->alpha, h->beta, h->gamma_m, h->gamma_p);
  else
    hipLaunchKernelGGL(k_cma_app<double>, dim3(g), dim3(256), 0, ctx->stream, (double*)lhs_d, (const double*)x_d,
                       (const double*)cm_d, to_space->map0_d, from_space->map0_d, h->ind_to_d, h->ind_from_d, ncell,
                       h->nrow, h->ncol, h->bw, h->ndf_to, h->ndf_from, h->alpha, h->beta, h->gamma_m, h->gamma_p);
  LFR_CHECK_LAUNCH(ctx);
  return LFR_OK;
}

}  // extern "C"

// ---------------------------------------------------------------------------
// Vertical tridiagonal solves on fields:
//   tri_solve_code (components/science/source/kernel/solver/sci_tri_solve_kernel_mod.f90:95-118)
//   operator_tri_solve_code (.../solver/sci_operator_tri_solve_kernel_mod.f90:65-120)
// One lane per column; the forward sweep keeps the eliminated right-hand side
// in y and the eliminated super-diagonal in a scratch field, the backward
// sweep finishes y in place.  Operation order as the reference: one
// reciprocal of the pivot per row, multiplied in.
// ---------------------------------------------------------------------------
namespace {

template <typename T>
__global__ void __launch_bounds__(256)
k_tri_solve(T* __restrict__ y, const T* __restrict__ x, const T* __restrict__ t0, const T* __restrict__ tp,
            const T* __restrict__ tm, T* __restrict__ tpn, const int32_t* __restrict__ map0, int ncell, int ndf,
            int nl) {
  const int c = blockIdx.x * 256 + threadIdx.x;
  if (c >= ncell) return;
  const int64_t ij = map0[(int64_t)c * ndf];
  T denom = T(1) / t0[ij];
  T tpk = tp[ij] * denom;
  T xk = x[ij] * denom;
  tpn[ij] = tpk;
  y[ij] = xk;
  for (int k = 1; k < nl; ++k) {
    const T pm = tm[ij + k] * tpk;
    denom = T(1) / (t0[ij + k] - pm);
    const T xm = tm[ij + k] * xk;
    tpk = tp[ij + k] * denom;
    xk = (x[ij + k] - xm) * denom;
    tpn[ij + k] = tpk;
    y[ij + k] = xk;
  }
  T yk = xk;   // y(nl-1)
  for (int k = nl - 2; k >= 0; --k) {
    const T py = tpn[ij + k] * yk;
    yk = y[ij + k] - py;
    y[ij + k] = yk;
  }
}

// The column's tridiagonal from the 2x2 local matrices m(ik, i, j) at
// ik + n3 * (i + 2 j) (0-based): diag(k) = m(k,0,0) + m(k-1,1,1),
// sub(k) = m(k-1,1,0), super(k) = m(k,0,1), for the nl+1 dofs of the column.
template <typename T>
__global__ void __launch_bounds__(256)
k_operator_tri_solve(T* __restrict__ y, const T* __restrict__ x, const T* __restrict__ m, T* __restrict__ tpn,
                     const int32_t* __restrict__ map0, int ncell, int ndf, int nl) {
  const int c = blockIdx.x * 256 + threadIdx.x;
  if (c >= ncell) return;
  const int64_t n3 = (int64_t)ncell * nl;
  const int64_t ij = map0[(int64_t)c * ndf];
  const int64_t ik0 = (int64_t)c * nl;
  // k = 0
  T d = m[ik0] + T(0);
  T denom = T(1) / d;
  T tpk = m[ik0 + 2 * n3] * denom;
  T xk = x[ij] * denom;
  tpn[ij] = tpk;
  y[ij] = xk;
  for (int k = 1; k <= nl; ++k) {
    const int64_t ikm = ik0 + k - 1;
    const T sub = m[ikm + n3];                          // (k-1; 2,1)
    const T t0 = k < nl ? m[ikm + 1] + m[ikm + 3 * n3]  // (k; 1,1) + (k-1; 2,2)
                        : m[ikm + 3 * n3];
    const T sup = k < nl ? m[ikm + 1 + 2 * n3] : T(0);  // (k; 1,2)
    const T pm = sub * tpk;
    denom = T(1) / (t0 - pm);
    const T xm = sub * xk;
    tpk = sup * denom;
    xk = (x[ij + k] - xm) * denom;
    tpn[ij + k] = tpk;
    y[ij + k] = xk;
  }
  T yk = xk;   // y(nl)
  for (int k = nl - 1; k >= 0; --k) {
    const T py = tpn[ij + k] * yk;
    yk = y[ij + k] - py;
    y[ij + k] = yk;
  }
}

// tri_solve in the form of k_appinv: one lane per column, 64 columns per
// workgroup, the column's bands and right-hand side loaded straight from
// global memory (each a constant offset from one per-lane base: the W3 dofs
// of a column are contiguous) kPf rows ahead of the recurrence, c' in LDS
// and d' in registers (NROW a compile-time constant), the solution stored
// through LDS.  The arithmetic is k_tri_solve's term for term (products
// commute exactly), so the bits are the same.
template <typename T, int NROW>
__global__ void __launch_bounds__(kCols)
k_thomas(T* __restrict__ y, const T* __restrict__ x, const T* __restrict__ t0, const T* __restrict__ tp,
         const T* __restrict__ tm, const int32_t* __restrict__ map0, int ndf, int ncell) {
  constexpr int NST = (NROW + kCh - 1) / kCh;
  constexpr int PITCH_X = kCh + 1;                 // odd pitch: conflict-free lane-strided access
  __shared__ T sx[kCols * PITCH_X];
  __shared__ int32_t sbase[kCols];
  constexpr int PC = NROW | 1;     // odd pitch of the c' rows
  __shared__ T scp[kCols * PC];    // c' of every lane's column (registers hold d')
  const int lane = threadIdx.x;
  const int c0 = blockIdx.x * kCols;
  const int ncols = min(kCols, ncell - c0);
  if (lane < ncols) sbase[lane] = map0[(int64_t)(c0 + lane) * ndf];
  T* cp = scp + lane * PC;
  T dp[NROW];
  {
    const int cl = min(lane, ncols - 1);   // idle lanes of a partial workgroup repeat a real column
    const int64_t base = map0[(int64_t)(c0 + cl) * ndf];
    T la[NROW], ld[NROW], lu[NROW], lx[NROW];
#pragma unroll
    for (int t = 0; t < NROW + kPf; ++t) {
      if (t < NROW) {
        la[t] = tm[base + t];
        ld[t] = t0[base + t];
        lu[t] = tp[base + t];
        lx[t] = x[base + t];
      }
      if (t >= kPf) {
        const int i = t - kPf;
        const T a = la[i], dg = ld[i], u = lu[i], xv = lx[i];
        if (i == 0) {
          const T inv = T(1) / dg;
          cp[i] = u * inv;
          dp[i] = xv * inv;
        } else {
          const T ac = a * cp[i - 1];
          const T inv = T(1) / (dg - ac);
          const T ad = a * dp[i - 1];
          if (i < NROW - 1) cp[i] = u * inv;
          dp[i] = (xv - ad) * inv;
        }
      }
      // loads of rows [r kPf, (r+1) kPf) share a scheduling region with the
      // recurrence of the kPf rows before them
      if ((t + 1) % kPf == 0) __builtin_amdgcn_sched_barrier(0);
    }
  }
#pragma unroll
  for (int i = NROW - 2; i >= 0; --i) {
    const T cd = cp[i] * dp[i + 1];
    dp[i] = dp[i] - cd;
  }
#pragma unroll
  for (int s = 0; s < NST; ++s) {
    const int r0 = s * kCh;
    const int nr = (NROW - r0) < kCh ? (NROW - r0) : kCh;
    __syncthreads();
    if (lane < ncols) {
#pragma unroll
      for (int q = 0; q < kCh; ++q)
        if (r0 + q < NROW) sx[lane * PITCH_X + q] = dp[r0 + q];
    }
    __syncthreads();
    const int nx = ncols * nr;
    for (int e = lane; e < nx; e += kCols) {
      const int cc = e / nr;
      const int q = e - cc * nr;
      y[(int64_t)sbase[cc] + r0 + q] = sx[cc * PITCH_X + q];
    }
  }
}

// tri_solve, FP32, every load issued at entry (the scheme of
// k_appinv_deep): a wave owns COLS columns, loads the KCH-row stages of
// tri_minus, tri_0, tri_plus and x of all its columns into registers up
// front (slot it of stage s, array v: row s KCH + qs of column cs with
// lane + it kCols = cs KCH + qs; a column's W3 dofs are contiguous), moves
// one stage at a time to LDS for the recurrence, and writes the solution
// back through LDS in flat order (64 consecutive dofs per instruction when
// the columns are contiguous) with streaming stores.  Same arithmetic as
// k_tri_solve.
template <int NROW, int KCH, int COLS, bool DPL = false>
__global__ void __launch_bounds__(kCols)
k_thomas_deep(float* __restrict__ y, const float* __restrict__ x, const float* __restrict__ t0,
              const float* __restrict__ tp, const float* __restrict__ tm, const int32_t* __restrict__ map0,
              int ndf, int ncell, int cstride) {
  static_assert(NROW % KCH == 0 && COLS <= kCols, "stage shape");
  constexpr int NST = NROW / KCH;
  constexpr int SLOTS = (COLS * KCH + kCols - 1) / kCols;
  constexpr int PB = 4 * KCH + 1;                    // odd pitches
  constexpr int PX = NROW | 1;
  __shared__ float sb[COLS * PB];
  __shared__ float sxa[COLS * PX];
  __shared__ int32_t sbase[COLS];
  const int lane = threadIdx.x;
  const int c0 = blockIdx.x * COLS;
  const int ncols = min(COLS, ncell - c0);
  // every load addresses through the column base: formed from the column
  // stride when the map has one (no round trip before the loads), read
  // from the map otherwise
  const int cl = c0 + min(lane, ncols - 1);
  const int32_t mybase = cstride >= 0 ? cl * cstride : map0[(int64_t)cl * ndf];
  if (lane < ncols) sbase[lane] = mybase;
  __syncthreads();
  const float* src[4] = {tm, t0, tp, x};
  float rv[NST][4][SLOTS];
#pragma unroll
  for (int s = 0; s < NST; ++s)
#pragma unroll
    for (int it = 0; it < SLOTS; ++it) {
      const int e = min(lane + it * kCols, COLS * KCH - 1);
      const int cs = e / KCH, qs = e - cs * KCH;
      const int64_t a = (int64_t)sbase[min(cs, ncols - 1)] + s * KCH + qs;
#pragma unroll
      for (int v = 0; v < 4; ++v) rv[s][v][it] = src[v][a];
    }
  float cp[NROW], dp[DPL ? 1 : NROW];
  float dprev = 0.f;   // DPL: d' of the previous row
#pragma unroll
  for (int s = 0; s < NST; ++s) {
    __syncthreads();
#pragma unroll
    for (int it = 0; it < SLOTS; ++it) {
      const int e = lane + it * kCols;
      if (e < COLS * KCH) {
        const int cs = e / KCH, qs = e - cs * KCH;
#pragma unroll
        for (int v = 0; v < 4; ++v) sb[cs * PB + 4 * qs + v] = rv[s][v][it];
      }
    }
    __syncthreads();
    if (lane < COLS) {
#pragma unroll
      for (int q = 0; q < KCH; ++q) {
        const int i = s * KCH + q;
        const float* w = sb + lane * PB + 4 * q;
        const float a = w[0], dg = w[1], u = w[2], xv = w[3];
        float di;
        if (i == 0) {
          const float inv = 1.0f / dg;
          cp[i] = u * inv;
          di = xv * inv;
        } else {
          const float ac = a * cp[i - 1];
          const float inv = 1.0f / (dg - ac);
          float dprv;
          if constexpr (DPL) dprv = dprev;
          else dprv = dp[i - 1];
          const float ad = a * dprv;
          if (i < NROW - 1) cp[i] = u * inv;
          di = (xv - ad) * inv;
        }
        if constexpr (DPL) {
          sxa[lane * PX + i] = di;
          dprev = di;
        } else {
          dp[i] = di;
        }
      }
    }
  }
  if constexpr (DPL) {
    if (lane < COLS) {
      float dnext = dprev;   // d'(NROW-1), already the solution there
#pragma unroll
      for (int i = NROW - 2; i >= 0; --i) {
        const float cd = cp[i] * dnext;
        const float di = sxa[lane * PX + i] - cd;
        sxa[lane * PX + i] = di;
        dnext = di;
      }
    }
    __syncthreads();
  } else {
#pragma unroll
    for (int i = NROW - 2; i >= 0; --i) {
      const float cd = cp[i] * dp[i + 1];
      dp[i] = dp[i] - cd;
    }
    __syncthreads();
    if (lane < COLS) {
#pragma unroll
      for (int i = 0; i < NROW; ++i) sxa[lane * PX + i] = dp[i];
    }
    __syncthreads();
  }
  const int nflat = ncols * NROW;
  constexpr int NXF = (COLS * NROW + kCols - 1) / kCols;
#pragma unroll
  for (int it = 0; it < NXF; ++it) {
    const int e = lane + it * kCols;
    if (e < nflat) {
      const int c = e / NROW, r = e - c * NROW;
      __builtin_nontemporal_store(sxa[c * PX + r], y + (int64_t)sbase[c] + r);
    }
  }
}

// operator_tri_solve with LDS staging: the 2x2 local matrices of the 64
// columns of a workgroup come in KCH-row stages with coalesced, branch-free
// loads (clamped addresses; slots past the last column or row hold
// duplicates and are never read), all of a stage's loads issued before any
// of its LDS writes (52 µs at C96L70 FP32 against 59 µs with each load
// followed by its write), the recurrence runs from LDS with c' and d' in
// registers.  For these four strided planes this beats the per-lane
// direct loads of k_thomas.  Same arithmetic as k_operator_tri_solve.
template <typename T, int NROW, int KCH>
__global__ void __launch_bounds__(kCols)
k_optri_staged(T* __restrict__ y, const T* __restrict__ x, const T* __restrict__ m,
               const int32_t* __restrict__ map0, int ndf, int ncell, int nl) {
  constexpr int NST = (NROW + KCH - 1) / KCH;
  constexpr int NV = 4;
  constexpr int PITCH_B = KCH * NV + 1;             // odd
  constexpr int PITCH_X = KCH + 1 + (KCH + 1) % 2;  // odd pitch: conflict-free lane-strided access
  __shared__ T sb[kCols * PITCH_B];
  __shared__ T sx[kCols * PITCH_X];
  __shared__ int32_t sbase[kCols];
  const int lane = threadIdx.x;
  const int c0 = blockIdx.x * kCols;
  const int ncols = min(kCols, ncell - c0);
  const int64_t n3 = (int64_t)ncell * nl;
  if (lane < ncols) sbase[lane] = map0[(int64_t)(c0 + lane) * ndf];
  T cp[NROW], dp[NROW];
  // one stage's loads: slot it of this lane is row r0 + qs of column cs
  // (e = lane + it kCols = cs KCH + qs), its address clamped into the
  // workgroup's columns and the stage's rows
  T ra[KCH], rb[KCH], rs[KCH], ru[KCH], rx[KCH];
  auto gather = [&](int s) {
    const int r0 = s * KCH;
    const int nr = (NROW - r0) < KCH ? (NROW - r0) : KCH;
#pragma unroll
    for (int it = 0; it < KCH; ++it) {
      const int e = lane + it * kCols;
      const int cs = e / KCH;
      const int qs = e - cs * KCH;
      const int cc = min(cs, ncols - 1);
      const int r = r0 + min(qs, nr - 1);
      // m(ik, i, j) at ik + n3 (i + 2 j); rows 0 and nl read a clamped
      // neighbour and select zero
      const int64_t ik = (int64_t)(c0 + cc) * nl + r;
      const int64_t ia = r < nl ? ik : ik - 1;
      const int64_t ib = r >= 1 ? ik - 1 : ik;
      ra[it] = m[ia];
      rb[it] = m[ib + 3 * n3];
      rs[it] = m[ib + n3];
      ru[it] = m[ia + 2 * n3];
      rx[it] = x[(int64_t)sbase[cc] + r];
    }
  };
  auto put = [&](int s) {
    const int r0 = s * KCH;
    const int nr = (NROW - r0) < KCH ? (NROW - r0) : KCH;
#pragma unroll
    for (int it = 0; it < KCH; ++it) {
      const int e = lane + it * kCols;
      const int cs = e / KCH;
      const int qs = e - cs * KCH;
      const int r = r0 + min(qs, nr - 1);
      T* o = sb + cs * PITCH_B + NV * qs;
      o[0] = r < nl ? ra[it] : T(0);    // (k; 1,1)
      o[1] = r >= 1 ? rb[it] : T(0);    // (k-1; 2,2)
      o[2] = r >= 1 ? rs[it] : T(0);    // (k-1; 2,1)
      o[3] = r < nl ? ru[it] : T(0);    // (k; 1,2)
      sx[cs * PITCH_X + qs] = rx[it];
    }
  };
#pragma unroll
  for (int s = 0; s < NST; ++s) {
    const int r0 = s * KCH;
    __syncthreads();
    gather(s);
    put(s);
    __syncthreads();
    if (lane < ncols) {
#pragma unroll
      for (int q = 0; q < KCH; ++q) {
        const int i = r0 + q;
        if (i < NROW) {
          const T* v = sb + lane * PITCH_B + NV * q;
          const T a = v[2];
          const T dg = i == 0 ? v[0] + T(0) : (i < nl ? v[0] + v[1] : v[1]);
          const T u = i < nl ? v[3] : T(0);
          const T xv = sx[lane * PITCH_X + q];
          if (i == 0) {
            const T inv = T(1) / dg;
            cp[i] = u * inv;
            dp[i] = xv * inv;
          } else {
            const T ac = a * cp[i - 1];
            const T inv = T(1) / (dg - ac);
            const T ad = a * dp[i - 1];
            if (i < NROW - 1) cp[i] = u * inv;
            dp[i] = (xv - ad) * inv;
          }
        }
      }
    }
  }
#pragma unroll
  for (int i = NROW - 2; i >= 0; --i) {
    const T cd = cp[i] * dp[i + 1];
    dp[i] = dp[i] - cd;
  }
#pragma unroll
  for (int s = 0; s < NST; ++s) {
    const int r0 = s * KCH;
    const int nr = (NROW - r0) < KCH ? (NROW - r0) : KCH;
    __syncthreads();
    if (lane < ncols) {
#pragma unroll
      for (int q = 0; q < KCH; ++q)
        if (r0 + q < NROW) sx[lane * PITCH_X + q] = dp[r0 + q];
    }
    __syncthreads();
    const int nx = ncols * nr;
    for (int e = lane; e < nx; e += kCols) {
      const int cc = e / nr;
      const int q = e - cc * nr;
      y[(int64_t)sbase[cc] + r0 + q] = sx[cc * PITCH_X + q];
    }
  }
}

// operator_tri_solve, FP32, 70 layers (71 dofs a column), every load issued
// at entry (the scheme of k_thomas_deep): a wave owns COLS columns and
// loads, for every KCH-row stage, the four planes of the 2x2 local matrices
// m(ik, i, j) at ik + n3 (i + 2 j) and x for the stage's rows into
// registers up front (slot it of stage s: row s KCH + qs of column cs with
// lane + it kCols = cs KCH + qs, rows past the column clamped; layer r of
// planes (0,0), (0,1) and layer r - 1 of planes (1,0), (1,1), clamped into
// the column, masked to zero where the reference has no term), moves one
// stage at a time to LDS for the recurrence and writes the solution back in
// flat order with streaming stores.  Same arithmetic as k_operator_tri_solve.
template <int NROW, int KCH, int COLS>
__global__ void __launch_bounds__(kCols)
k_optri_deep(float* __restrict__ y, const float* __restrict__ x, const float* __restrict__ m,
             const int32_t* __restrict__ map0, int ndf, int ncell, int cstride) {
  static_assert(COLS <= kCols, "stage shape");
  constexpr int NL = NROW - 1;                      // layers
  constexpr int NST = (NROW + KCH - 1) / KCH;
  constexpr int SLOTS = (COLS * KCH + kCols - 1) / kCols;
  constexpr int PB = 4 * KCH + 1;                   // odd pitches
  constexpr int PX = NROW | 1;
  __shared__ float sb[COLS * PB];
  __shared__ float sxa[COLS * PX];
  __shared__ int32_t sbase[COLS];
  const int lane = threadIdx.x;
  const int c0 = blockIdx.x * COLS;
  const int ncols = min(COLS, ncell - c0);
  const int64_t n3 = (int64_t)ncell * NL;
  // the column bases first (formed from the column stride when the map has
  // one, else read unconditionally at a clamped index), then the matrix
  // planes, which are addressed by cell and do not wait for them; x, whose
  // addresses do, after (a base loaded behind a branch and waited for before
  // the matrix loads cost a whole round trip)
  const int cl = c0 + min(lane, ncols - 1);
  const int32_t mbv = cstride >= 0 ? cl * cstride : map0[(int64_t)cl * ndf];
  // per slot: a = m(r-1;1,0), d0 = m(r;0,0), d1 = m(r-1;1,1), u = m(r;0,1), x(r)
  float rv[NST][5][SLOTS];
#pragma unroll
  for (int s = 0; s < NST; ++s)
#pragma unroll
    for (int it = 0; it < SLOTS; ++it) {
      const int e = min(lane + it * kCols, COLS * KCH - 1);
      const int cs = min(e / KCH, ncols - 1);
      const int r = min(s * KCH + (e % KCH), NROW - 1);
      const int64_t ik = (int64_t)(c0 + cs) * NL + r;
      const int64_t ia = r < NL ? ik : ik - 1;          // layer r (clamped at the top)
      const int64_t ib = r >= 1 ? ik - 1 : ik;          // layer r - 1 (clamped at the bottom)
      rv[s][0][it] = m[ib + n3];
      rv[s][1][it] = m[ia];
      rv[s][2][it] = m[ib + 3 * n3];
      rv[s][3][it] = m[ia + 2 * n3];
    }
  // an unconditional use after the matrix loads, so that the base's load is
  // not moved down into the branch below; the statement emits no instruction
  asm volatile("" ::"v"(mbv));
  if (lane < ncols) sbase[lane] = mbv;
  __syncthreads();
#pragma unroll
  for (int s = 0; s < NST; ++s)
#pragma unroll
    for (int it = 0; it < SLOTS; ++it) {
      const int e = min(lane + it * kCols, COLS * KCH - 1);
      const int cs = min(e / KCH, ncols - 1);
      const int r = min(s * KCH + (e % KCH), NROW - 1);
      rv[s][4][it] = x[(int64_t)sbase[cs] + r];
    }
  float cp[NROW], dp[NROW];
#pragma unroll
  for (int s = 0; s < NST; ++s) {
    __syncthreads();
#pragma unroll
    for (int it = 0; it < SLOTS; ++it) {
      const int e = lane + it * kCols;
      if (e < COLS * KCH) {
        const int cs = e / KCH, qs = e - cs * KCH;
        const int r = min(s * KCH + qs, NROW - 1);
        float* o = sb + cs * PB + 4 * qs;
        o[0] = r >= 1 ? rv[s][0][it] : 0.0f;                                   // sub
        o[1] = r == 0 ? rv[s][1][it] + 0.0f
                      : (r < NL ? rv[s][1][it] + rv[s][2][it] : rv[s][2][it]);  // diag
        o[2] = r < NL ? rv[s][3][it] : 0.0f;                                   // super
        o[3] = rv[s][4][it];                                                   // x
      }
    }
    __syncthreads();
    if (lane < COLS) {
#pragma unroll
      for (int q = 0; q < KCH; ++q) {
        const int i = s * KCH + q;
        if (i < NROW) {
          const float* w = sb + lane * PB + 4 * q;
          const float a = w[0], dg = w[1], u = w[2], xv = w[3];
          if (i == 0) {
            const float inv = 1.0f / dg;
            cp[i] = u * inv;
            dp[i] = xv * inv;
          } else {
            const float ac = a * cp[i - 1];
            const float inv = 1.0f / (dg - ac);
            const float ad = a * dp[i - 1];
            if (i < NROW - 1) cp[i] = u * inv;
            dp[i] = (xv - ad) * inv;
          }
        }
      }
    }
  }
#pragma unroll
  for (int i = NROW - 2; i >= 0; --i) {
    const float cd = cp[i] * dp[i + 1];
    dp[i] = dp[i] - cd;
  }
  __syncthreads();
  if (lane < COLS) {
#pragma unroll
    for (int i = 0; i < NROW; ++i) sxa[lane * PX + i] = dp[i];
  }
  __syncthreads();
  const int nflat = ncols * NROW;
  constexpr int NXF = (COLS * NROW + kCols - 1) / kCols;
#pragma unroll
  for (int it = 0; it < NXF; ++it) {
    const int e = lane + it * kCols;
    if (e < nflat) {
      const int c = e / NROW, r = e - c * NROW;
      __builtin_nontemporal_store(sxa[c * PX + r], y + (int64_t)sbase[c] + r);
    }
  }
}

template <typename T, int MODE>
bool launch_thomas(lfr_ctx* ctx, hipStream_t s, T* y, const T* x, const T* b0, const T* b1, const T* b2,
                   const lfr_space* sp) {
  const unsigned nb = (unsigned)((sp->ncell + kCols - 1) / kCols);
  if constexpr (MODE == 0) {   // tri_solve: b0 = tri_0, b1 = tri_plus, b2 = tri_minus
    if constexpr (sizeof(T) == 4) {
      // every load at entry (LFR_TRI_DEEP=0 keeps k_thomas); result-identical
      // variants (columns a wave x rows a stage, k_thomas) chosen by the
      // autotuner on the first call of a shape (y never aliases an input)
      static const bool deep = [] {
        const char* e = getenv("LFR_TRI_DEEP");
        return !(e && e[0] == '0');
      }();
      if (deep && sp->nlayers == 70) {
        const int ncell = sp->ncell;
        const auto launch = [&](int c) -> int {
          switch (c) {
            case 0:
              hipLaunchKernelGGL((k_thomas_deep<70, 14, 32>), dim3((ncell + 31) / 32), dim3(kCols), 0, s, y, x, b0,
                                 b1, b2, sp->map0_d, sp->ndf, ncell, sp->col_stride);
              break;
            case 1:
              hipLaunchKernelGGL((k_thomas_deep<70, 10, 16>), dim3((ncell + 15) / 16), dim3(kCols), 0, s, y, x, b0,
                                 b1, b2, sp->map0_d, sp->ndf, ncell, sp->col_stride);
              break;
            case 2:
              hipLaunchKernelGGL((k_thomas_deep<70, 14, 16>), dim3((ncell + 15) / 16), dim3(kCols), 0, s, y, x, b0,
                                 b1, b2, sp->map0_d, sp->ndf, ncell, sp->col_stride);
              break;
            case 4:   // d' in LDS
              hipLaunchKernelGGL((k_thomas_deep<70, 10, 16, true>), dim3((ncell + 15) / 16), dim3(kCols), 0, s, y,
                                 x, b0, b1, b2, sp->map0_d, sp->ndf, ncell, sp->col_stride);
              break;
            case 5:
              hipLaunchKernelGGL((k_thomas_deep<70, 14, 16, true>), dim3((ncell + 15) / 16), dim3(kCols), 0, s, y,
                                 x, b0, b1, b2, sp->map0_d, sp->ndf, ncell, sp->col_stride);
              break;
            default:
              hipLaunchKernelGGL((k_thomas<T, 70>), dim3(nb), dim3(kCols), 0, s, y, x, b0, b1, b2, sp->map0_d,
                                 sp->ndf, ncell);
          }
          return hipGetLastError() == hipSuccess ? LFR_OK : LFR_ERR_HIP;
        };
        const int dflt = (ncell + 31) / 32 <= 2048 ? 0 : 5;
        char key[96];
        snprintf(key, sizeof key, "tri_solve f32 nl=70 ncell=%d", ncell);
        const int c = ctx ? lfr_tune_pick(ctx, key, 6, dflt, launch) : dflt;
        (void)launch(c);
        return true;
      }
    }
    switch (sp->nlayers) {
#define CASE(NR)                                                                                          \
  case NR:                                                                                                \
    hipLaunchKernelGGL((k_thomas<T, NR>), dim3(nb), dim3(kCols), 0, s, y, x, b0, b1, b2, sp->map0_d, sp->ndf, \
                       sp->ncell);                                                                        \
    return true;
      CASE(10) CASE(70)
#undef CASE
      default: return false;
    }
  } else {                     // operator_tri_solve: b0 = the 2x2 local matrices
    if (sp->nlayers != 70) return false;
    if constexpr (sizeof(T) == 4) {
      // result-identical variants (columns a wave x rows a stage, the staged
      // kernel) chosen by the autotuner on the first call of a shape
      const int ncell = sp->ncell;
      const auto launch = [&](int c) -> int {
        switch (c) {
          case 0:
            hipLaunchKernelGGL((k_optri_deep<71, 12, 16>), dim3((ncell + 15) / 16), dim3(kCols), 0, s, y, x, b0,
                               sp->map0_d, sp->ndf, ncell, sp->col_stride);
            break;
          case 1:
            hipLaunchKernelGGL((k_optri_deep<71, 8, 32>), dim3((ncell + 31) / 32), dim3(kCols), 0, s, y, x, b0,
                               sp->map0_d, sp->ndf, ncell, sp->col_stride);
            break;
          default:
            hipLaunchKernelGGL((k_optri_staged<T, 71, kCh>), dim3(nb), dim3(kCols), 0, s, y, x, b0, sp->map0_d,
                               sp->ndf, ncell, sp->nlayers);
        }
        return hipGetLastError() == hipSuccess ? LFR_OK : LFR_ERR_HIP;
      };
      char key[96];
      snprintf(key, sizeof key, "operator_tri_solve f32 nl=70 ncell=%d", ncell);
      // untuned default: 32 columns a wave while that still gives few
      // enough waves for one pass over the chip, else 16
      const int dflt = (ncell + 31) / 32 <= 2048 ? 1 : 0;
      const int c = ctx ? lfr_tune_pick(ctx, key, 3, dflt, launch) : dflt;
      (void)launch(c);
      return true;
    }
    hipLaunchKernelGGL((k_optri_staged<T, 71, kCh>), dim3(nb), dim3(kCols), 0, s, y, x, b0, sp->map0_d,
                       sp->ndf, sp->ncell, sp->nlayers);
    return true;
  }
}

}  // namespace

extern "C" {

int lfr_tri_solve(lfr_ctx* ctx, int prec, void* y, const void* x, const void* tri_0, const void* tri_plus,
                  const void* tri_minus, const lfr_space* space) {
  if (!ctx) return LFR_ERR_ARG;
  LFR_REQUIRE(ctx, prec == 4 || prec == 8, "lfr_tri_solve: prec must be 4 or 8, got %d", prec);
  LFR_REQUIRE(ctx, y && x && tri_0 && tri_plus && tri_minus && space, "lfr_tri_solve: null argument");
  LFR_REQUIRE(ctx, space->ndf == 1, "lfr_tri_solve: lowest-order W3 only (ndf %d, expected 1)", space->ndf);
  LFR_REQUIRE(ctx, y != x && y != tri_0 && y != tri_plus && y != tri_minus,
              "lfr_tri_solve: y must not alias an input");
  const bool staged =
      prec == 4 ? launch_thomas<float, 0>(ctx, ctx->stream, (float*)y, (const float*)x, (const float*)tri_0,
                                          (const float*)tri_plus, (const float*)tri_minus, space)
                : launch_thomas<double, 0>(ctx, ctx->stream, (double*)y, (const double*)x, (const double*)tri_0,
                                           (const double*)tri_plus, (const double*)tri_minus, space);
  if (staged) {
    LFR_CHECK_LAUNCH(ctx);
    return LFR_OK;
  }
  const size_t sb = (size_t)space->undf * prec;
  int rc = lfr_work_reserve(ctx, sb);
  if (rc) return rc;
  const unsigned nb = (unsigned)((space->ncell + 255) / 256);
  if (prec == 4)
    hipLaunchKernelGGL(k_tri_solve<float>, dim3(nb), dim3(256), 0, ctx->stream, (float*)y, (const float*)x,
                       (const float*)tri_0, (const float*)tri_plus, (const float*)tri_minus, (float*)ctx->work,
                       space->map0_d, space->ncell, space->ndf, space->nlayers);
  else
    hipLaunchKernelGGL(k_tri_solve<double>, dim3(nb), dim3(256), 0, ctx->stream, (double*)y, (const double*)x,
                       (const double*)tri_0, (const double*)tri_plus, (const double*)tri_minus,
                       (double*)ctx->work, space->map0_d, space->ncell, space->ndf, space->nlayers);
  LFR_CHECK_LAUNCH(ctx);
  return LFR_OK;
}

int lfr_operator_tri_solve(lfr_ctx* ctx, int prec, void* y, const void* x, const void* matrix,
                           const lfr_space* space) {
  if (!ctx) return LFR_ERR_ARG;
  LFR_REQUIRE(ctx, prec == 4 || prec == 8, "lfr_operator_tri_solve: prec must be 4 or 8, got %d", prec);
  LFR_REQUIRE(ctx, y && x && matrix && space, "lfr_operator_tri_solve: null argument");
  LFR_REQUIRE(ctx, y != x, "lfr_operator_tri_solve: y must not alias x");
  LFR_REQUIRE(ctx, space->ndf == 2,
              "lfr_operator_tri_solve: lowest-order Wtheta/W2v only (ndf %d, expected 2)", space->ndf);
  if (space->vstack < 0) {   // checked once per space
    int ok = 1;
    for (int c = 0; c < space->ncell && ok; ++c)
      if (space->map_h[(size_t)c * 2 + 1] != space->map_h[(size_t)c * 2] + 1) ok = 0;
    space->vstack = ok;
  }
  LFR_REQUIRE(ctx, space->vstack == 1,
              "lfr_operator_tri_solve: the top dof of a cell does not follow its bottom dof");
  const bool staged =
      prec == 4 ? launch_thomas<float, 1>(ctx, ctx->stream, (float*)y, (const float*)x, (const float*)matrix, nullptr,
                                          nullptr, space)
                : launch_thomas<double, 1>(ctx, ctx->stream, (double*)y, (const double*)x, (const double*)matrix,
                                           nullptr, nullptr, space);
  if (staged) {
    LFR_CHECK_LAUNCH(ctx);
    return LFR_OK;
  }
  const size_t sb = (size_t)space->undf * prec;
  int rc = lfr_work_reserve(ctx, sb);
  if (rc) return rc;
  const unsigned nb = (unsigned)((space->ncell + 255) / 256);
  if (prec == 4)
    hipLaunchKernelGGL(k_operator_tri_solve<float>, dim3(nb), dim3(256), 0, ctx->stream, (float*)y,
                       (const float*)x, (const float*)matrix, (float*)ctx->work, space->map0_d, space->ncell,
                       space->ndf, space->nlayers);
  else
    hipLaunchKernelGGL(k_operator_tri_solve<double>, dim3(nb), dim3(256), 0, ctx->stream, (double*)y,
                       (const double*)x, (const double*)matrix, (double*)ctx->work, space->map0_d, space->ncell,
                       space->ndf, space->nlayers);
  LFR_CHECK_LAUNCH(ctx);
  return LFR_OK;
}

}  // extern "C"
