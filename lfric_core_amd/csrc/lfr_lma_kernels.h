// LMA apply kernels shared by the apply entry points and the fused solvers.
#pragma once

#include "lfr_internal.h"
#include "lfr_reduce.h"

namespace lfr {

constexpr int kThreads = 256;
static_assert(kThreads == kRedThreads, "block reductions assume 256 threads");

// ---------------------------------------------------------------------------
// Owner-computes gather for a continuous output space.
//
// Thread <-> (column, level L).  The column's contributions come in ngroup
// cell groups (colour order); each group holds GS local dofs (i, shift s) of
// that cell; entry (i, s) adds the cell's layer k = L - s when 0 <= k < nl.
//  KIND 0 (matrix_vector): per group, for j (df2) outer, for entries (df)
//         inner: acc = acc + M(ik,i,j)*x(map_in(j)+k)  (matrix_vector_kernel_mod.F90:130-138)
//  KIND 1 (transpose):     per group, for entries in layer order: the full
//         dot over j is formed first, then added
//         (transpose_matrix_vector_kernel_mod.F90:92-101).
// ---------------------------------------------------------------------------
template <typename T, int NDF_IN, int NG, int GS, int KIND, bool DOT = false>
__global__ void __launch_bounds__(kThreads)
k_gather(const T* __restrict__ M, const T* __restrict__ x, T* __restrict__ y,
         const int32_t* __restrict__ map_in, const int32_t* __restrict__ col_start,
         const int32_t* __restrict__ ent, int ncol, int H, int ngroup_rt, int nl, int64_t n3,
         int ndf_out, int ndf_in_rt, int overwrite, unsigned nb_pad,
         const T* __restrict__ dotv = nullptr, double* __restrict__ part = nullptr,
         const double* __restrict__ stop = nullptr) {
  // NG > 0: compile-time group count (all loads issued before the ordered
  // accumulation); NG == 0: runtime group count.
  if (DOT && stop && stop[0] != 0.0) return;   // solver already converged (uniform)
  const unsigned lb = lfr_xcd_block(blockIdx.x, nb_pad);
  const int64_t t = (int64_t)lb * kThreads + threadIdx.x;
  const bool valid = t < (int64_t)ncol * H;
  const int ngroup = NG > 0 ? NG : ngroup_rt;
  const int ndf_in = NDF_IN > 0 ? NDF_IN : ndf_in_rt;
  // Invalid (padding) lanes compute column 0 level 0 and discard the result,
  // so every address below stays inside the operands.
  const int64_t tt = valid ? t : 0;
  const int col = (int)(tt / H);
  const int L = (int)(tt - (int64_t)col * H);
  const int32_t g0 = col_start[col];
  T acc = (overwrite || !valid) ? T(0) : y[g0 + L];
  const int32_t* e = ent + (int64_t)col * ngroup * GS * 2;
  constexpr int NGU = NG > 0 ? NG : 1;          // unrolled groups per pass
  for (int gb = 0; gb < ngroup; gb += NGU) {
    int cc[NGU], ii[NGU][GS], kk[NGU][GS];
    bool ok[NGU][GS];
#pragma unroll
    for (int g = 0; g < NGU; ++g) {
      const int32_t* eg = e + (gb + g) * GS * 2;
      cc[g] = eg[0];
#pragma unroll
      for (int q = 0; q < GS; ++q) {
        const int code = eg[2 * q + 1];
        const int k = L - (code & 1);
        ok[g][q] = code >= 0 && k >= 0 && k < nl;
        ii[g][q] = code >= 0 ? (code >> 1) : 0;
        kk[g][q] = k < 0 ? 0 : (k >= nl ? nl - 1 : k);   // clamped: always in bounds
      }
    }
    if constexpr (NDF_IN > 0) {
      T mv[NGU][GS][NDF_IN], xv[NGU][GS][NDF_IN];
#pragma unroll
      for (int g = 0; g < NGU; ++g) {
        const int32_t* mi = map_in + (int64_t)cc[g] * NDF_IN;
        int xb[NDF_IN];
#pragma unroll
        for (int j = 0; j < NDF_IN; ++j) xb[j] = mi[j];
#pragma unroll
        for (int q = 0; q < GS; ++q) {
          const int64_t ik = (int64_t)cc[g] * nl + kk[g][q];
#pragma unroll
          for (int j = 0; j < NDF_IN; ++j) {
            const int64_t mo = KIND == 0 ? (int64_t)ii[g][q] + (int64_t)ndf_out * j
                                         : (int64_t)j + (int64_t)NDF_IN * ii[g][q];
            mv[g][q][j] = M[ik + n3 * mo];
            xv[g][q][j] = x[xb[j] + kk[g][q]];
          }
        }
      }
#pragma unroll
      for (int g = 0; g < NGU; ++g) {
        if (KIND == 0) {
          // matrix_vector order inside a cell: df2 outer, df inner
#pragma unroll
          for (int j = 0; j < NDF_IN; ++j)
#pragma unroll
            for (int q = 0; q < GS; ++q) {
              const T p = mv[g][q][j] * xv[g][q][j];
              const T s = acc + p;
              acc = ok[g][q] ? s : acc;
            }
        } else {
          // transpose order: layer outer (entries stored lower layer first),
          // the dot over df2 formed before it is added
#pragma unroll
          for (int q = 0; q < GS; ++q) {
            T dot = T(0);
#pragma unroll
            for (int j = 0; j < NDF_IN; ++j) {
              const T p = mv[g][q][j] * xv[g][q][j];
              dot = dot + p;
            }
            const T s = acc + dot;
            acc = ok[g][q] ? s : acc;
          }
        }
      }
    } else {
      // runtime ndf_in: same order, loads interleaved with the accumulation
#pragma unroll
      for (int g = 0; g < NGU; ++g) {
        const int32_t* mi = map_in + (int64_t)cc[g] * ndf_in;
        if (KIND == 0) {
          for (int j = 0; j < ndf_in; ++j) {
            const int xbj = mi[j];
#pragma unroll
            for (int q = 0; q < GS; ++q) {
              const int64_t ik = (int64_t)cc[g] * nl + kk[g][q];
              const T p = M[ik + n3 * ((int64_t)ii[g][q] + (int64_t)ndf_out * j)] * x[xbj + kk[g][q]];
              const T s = acc + p;
              acc = ok[g][q] ? s : acc;
            }
          }
        } else {
#pragma unroll
          for (int q = 0; q < GS; ++q) {
            const int64_t ik = (int64_t)cc[g] * nl + kk[g][q];
            T dot = T(0);
            for (int j = 0; j < ndf_in; ++j) {
              const T p = M[ik + n3 * ((int64_t)j + (int64_t)ndf_in * ii[g][q])] * x[mi[j] + kk[g][q]];
              dot = dot + p;
            }
            const T s = acc + dot;
            acc = ok[g][q] ? s : acc;
          }
        }
      }
    }
  }
  if (valid) y[g0 + L] = acc;
  if (DOT) {
    __shared__ double lds[kThreads / 64];
    double v = valid ? (double)acc * (double)dotv[g0 + L] : 0.0;
    v = block_sum(v, lds);
    if (threadIdx.x == 0) part[lb] = v;
  }
}

// ---------------------------------------------------------------------------
// Gather with compiled column records.  For one (output space, input space)
// pair the plan is flattened into one record per column:
//   [ g0 | per group: c*nl, code[GS], map_in(c, 0..NDF_IN-1) ] padded to x4
// so a thread fetches its whole column description with STRIDE/4 aligned
// 16-byte loads and issues every matrix and operand load one memory latency
// later (no dependent dofmap lookup).  Summation order is identical to
// k_gather.  The matrix is streamed exactly once, so it is read with
// non-temporal loads that do not displace the gathered operand from cache.
// ---------------------------------------------------------------------------
template <int NDF_IN, int NG, int GS>
struct RecShape {
  static constexpr int kGroup = 1 + GS + NDF_IN;
  static constexpr int kInts = 1 + NG * kGroup;
  static constexpr int kStride = (kInts + 3) & ~3;
};

template <typename T>
__device__ __forceinline__ T ldnt(const T* p) { return __builtin_nontemporal_load(p); }

// One thread's work of the record gather for logical block lb; returns
// acc * dotv (FP64) for the fused inner product (0 when no dotv or invalid).
template <typename T, int NDF_IN, int NG, int GS, int KIND, bool NT>
__device__ __forceinline__ double gather_rec_body(const T* __restrict__ M, const T* __restrict__ x,
                                                  T* __restrict__ y, const int32_t* __restrict__ rec, int ncol,
                                                  int H, int nl, int64_t n3, int ndf_out, int overwrite,
                                                  int64_t pcs, unsigned lb, const T* __restrict__ dotv) {
  using S = RecShape<NDF_IN, NG, GS>;
  const int64_t t = (int64_t)lb * kThreads + threadIdx.x;
  const bool valid = t < (int64_t)ncol * H;
  const int64_t tt = valid ? t : 0;
  const int col = (int)(tt / H);
  const int L = (int)(tt - (int64_t)col * H);
  int r[S::kStride];
  const int4* r4 = reinterpret_cast<const int4*>(rec + (int64_t)col * S::kStride);
#pragma unroll
  for (int q = 0; q < S::kStride / 4; ++q) {
    const int4 v = r4[q];
    r[4 * q + 0] = v.x;
    r[4 * q + 1] = v.y;
    r[4 * q + 2] = v.z;
    r[4 * q + 3] = v.w;
  }
  const int32_t g0 = r[0];
  T acc = (overwrite || !valid) ? T(0) : y[g0 + L];
  T mv[NG][GS][NDF_IN], xv[NG][GS][NDF_IN];
  bool ok[NG][GS];
  int iis[NG][GS];
#pragma unroll
  for (int g = 0; g < NG; ++g) {
    const int* rg = r + 1 + g * S::kGroup;
    const int64_t mbase = (int64_t)rg[0];
#pragma unroll
    for (int q = 0; q < GS; ++q) {
      const int code = rg[1 + q];
      const int k = L - (code & 1);
      ok[g][q] = code >= 0 && k >= 0 && k < nl;
      const int ii = code >= 0 ? (code >> 1) : 0;
      iis[g][q] = ii;
      const int kk = k < 0 ? 0 : (k >= nl ? nl - 1 : k);
      // reference layout: M(cell*nl + k, row, column) planes n3 apart;
      // packed (pcs > 0): row (entry, j) of this column at level L
      const int64_t eb = pcs > 0 ? (int64_t)col * pcs + (int64_t)(g * GS + q) * NDF_IN * H + L
                                 : mbase + kk + n3 * (KIND == 0 ? (int64_t)ii : (int64_t)NDF_IN * ii);
      const int64_t js = pcs > 0 ? (int64_t)H : (KIND == 0 ? n3 * ndf_out : n3);
#pragma unroll
      for (int j = 0; j < NDF_IN; ++j) {
        mv[g][q][j] = NT ? ldnt(M + eb + js * j) : M[eb + js * j];
        xv[g][q][j] = x[rg[1 + GS + j] + kk];
      }
    }
  }
#pragma unroll
  for (int g = 0; g < NG; ++g) {
    if (KIND == 0) {
#pragma unroll
      for (int j = 0; j < NDF_IN; ++j)
#pragma unroll
        for (int q = 0; q < GS; ++q) {
          const T p = mv[g][q][j] * xv[g][q][j];
          const T s = acc + p;
          acc = ok[g][q] ? s : acc;
        }
    } else {
#pragma unroll
      for (int q = 0; q < GS; ++q) {
        T dot = T(0);
#pragma unroll
        for (int j = 0; j < NDF_IN; ++j) {
          const T p = mv[g][q][j] * xv[g][q][j];
          dot = dot + p;
        }
        const T s = acc + dot;
        acc = ok[g][q] ? s : acc;
      }
    }
  }
  if (valid) y[g0 + L] = acc;
  if (!dotv || !valid) return 0.0;
  // Fused <y, x> of the solvers (dotv == x, square operator on one space):
  // x at the output dof is the operand of any valid entry's own local dof
  // (input dof ii of the cell at layer k is output dof g0 + L), so it is
  // taken from the registers instead of being loaded again.
  T self = T(0);
  bool found = false;
#pragma unroll
  for (int g = 0; g < NG; ++g)
#pragma unroll
    for (int q = 0; q < GS; ++q) {
      T v = xv[g][q][0];
#pragma unroll
      for (int j = 1; j < NDF_IN; ++j) v = iis[g][q] == j ? xv[g][q][j] : v;
      self = (ok[g][q] && !found) ? v : self;
      found = found || ok[g][q];
    }
  return (double)acc * (double)self;
}

template <typename T, int NDF_IN, int NG, int GS, int KIND, bool DOT = false, bool NT = true>
__global__ void __launch_bounds__(kThreads)
k_gather_rec(const T* __restrict__ M, const T* __restrict__ x, T* __restrict__ y,
             const int32_t* __restrict__ rec, int ncol, int H, int nl, int64_t n3, int ndf_out,
             int overwrite, int64_t pcs, unsigned nb_pad, const T* __restrict__ dotv = nullptr,
             double* __restrict__ part = nullptr, const double* __restrict__ stop = nullptr) {
  if (DOT && stop && stop[0] != 0.0) return;
  const unsigned lb = lfr_xcd_block(blockIdx.x, nb_pad);
  double v = gather_rec_body<T, NDF_IN, NG, GS, KIND, NT>(M, x, y, rec, ncol, H, nl, n3, ndf_out, overwrite, pcs,
                                                          lb, DOT ? dotv : nullptr);
  if (DOT) {
    __shared__ double lds[kThreads / 64];
    v = block_sum(v, lds);
    if (threadIdx.x == 0) part[lb] = v;
  }
}

// ---------------------------------------------------------------------------
// Record gather, LPT consecutive levels per thread, for classes whose columns
// have exactly nl levels (no bottom/top shift: every entry reads layer k = L).
// Each matrix plane segment, operand column and output column is then read
// or written LPT values per instruction (multi-dword loads; the hardware
// needs only 4-byte alignment), which divides the vector-memory instruction
// count by LPT.  Threads per column = ceil(nl / LPT); the last one starts at
// nl - LPT so that its window stays inside the column, recomputes the levels
// it shares with its neighbour (identical operations, identical bits) and
// writes only the levels that neighbour does not own.  The arithmetic per
// output is exactly that of k_gather_rec.
// ---------------------------------------------------------------------------
template <typename T, int N>
struct __attribute__((aligned(4))) LVec {
  T v[N];
};

template <typename T, int N, bool NT>
__device__ __forceinline__ LVec<T, N> ldvec(const T* p) {
  if constexpr (NT) {
    // element-wise non-temporal loads; the compiler merges them into one
    // multi-dword load with the streaming cache policy
    LVec<T, N> r;
#pragma unroll
    for (int i = 0; i < N; ++i) r.v[i] = __builtin_nontemporal_load(p + i);
    return r;
  } else {
    return *reinterpret_cast<const LVec<T, N>*>(p);
  }
}

template <typename T, int NDF_IN, int NG, int GS, int KIND, int LPT, bool NT = false, int BS = kThreads>
__device__ __forceinline__ double gather_vec_body(const T* __restrict__ M, const T* __restrict__ x,
                                                  T* __restrict__ y, const int32_t* __restrict__ rec, int ncol,
                                                  int nl, int64_t n3, int ndf_out, int overwrite, int64_t pcs,
                                                  unsigned lb, const T* __restrict__ dotv) {
  using S = RecShape<NDF_IN, NG, GS>;
  const int tpc = (nl + LPT - 1) / LPT;
  const int64_t t = (int64_t)lb * BS + threadIdx.x;
  const bool valid = t < (int64_t)ncol * tpc;
  const int64_t tt = valid ? t : 0;
  const int col = (int)(tt / tpc);
  const int p = (int)(tt - (int64_t)col * tpc);
  const int L0 = min(p * LPT, nl - LPT);      // window [L0, L0 + LPT) inside the column
  const int own = p * LPT - L0;               // levels below L0 + own belong to thread p - 1
  int r[S::kStride];
  const int4* r4 = reinterpret_cast<const int4*>(rec + (int64_t)col * S::kStride);
#pragma unroll
  for (int q = 0; q < S::kStride / 4; ++q) {
    const int4 v = r4[q];
    r[4 * q + 0] = v.x;
    r[4 * q + 1] = v.y;
    r[4 * q + 2] = v.z;
    r[4 * q + 3] = v.w;
  }
  const int32_t g0 = r[0];
  T acc[LPT];
  if (overwrite || !valid) {
#pragma unroll
    for (int u = 0; u < LPT; ++u) acc[u] = T(0);
  } else {
    const LVec<T, LPT> yv = ldvec<T, LPT, false>(y + g0 + L0);
#pragma unroll
    for (int u = 0; u < LPT; ++u) acc[u] = yv.v[u];
  }
  LVec<T, LPT> mv[NG][GS][NDF_IN], xv[NG][GS][NDF_IN];
  bool ok[NG][GS];
  int iis[NG][GS];
#pragma unroll
  for (int g = 0; g < NG; ++g) {
    const int* rg = r + 1 + g * S::kGroup;
    const int64_t mbase = (int64_t)rg[0];
#pragma unroll
    for (int q = 0; q < GS; ++q) {
      const int code = rg[1 + q];
      ok[g][q] = code >= 0;
      const int ii = code >= 0 ? (code >> 1) : 0;
      iis[g][q] = ii;
      const int64_t eb = pcs > 0 ? (int64_t)col * pcs + (int64_t)(g * GS + q) * NDF_IN * nl + L0
                                 : mbase + L0 + n3 * (KIND == 0 ? (int64_t)ii : (int64_t)NDF_IN * ii);
      const int64_t js = pcs > 0 ? (int64_t)nl : (KIND == 0 ? n3 * ndf_out : n3);
#pragma unroll
      for (int j = 0; j < NDF_IN; ++j) {
        mv[g][q][j] = ldvec<T, LPT, NT>(M + eb + js * j);
        xv[g][q][j] = ldvec<T, LPT, false>(x + rg[1 + GS + j] + L0);
      }
    }
  }
#pragma unroll
  for (int g = 0; g < NG; ++g) {
    if (KIND == 0) {
#pragma unroll
      for (int j = 0; j < NDF_IN; ++j)
#pragma unroll
        for (int q = 0; q < GS; ++q)
#pragma unroll
          for (int u = 0; u < LPT; ++u) {
            const T pr = mv[g][q][j].v[u] * xv[g][q][j].v[u];
            const T s = acc[u] + pr;
            acc[u] = ok[g][q] ? s : acc[u];
          }
    } else {
#pragma unroll
      for (int q = 0; q < GS; ++q)
#pragma unroll
        for (int u = 0; u < LPT; ++u) {
          T dot = T(0);
#pragma unroll
          for (int j = 0; j < NDF_IN; ++j) {
            const T pr = mv[g][q][j].v[u] * xv[g][q][j].v[u];
            dot = dot + pr;
          }
          const T s = acc[u] + dot;
          acc[u] = ok[g][q] ? s : acc[u];
        }
    }
  }
  if (valid) {
    if (own == 0) {
      LVec<T, LPT> o;
#pragma unroll
      for (int u = 0; u < LPT; ++u) o.v[u] = acc[u];
      *reinterpret_cast<LVec<T, LPT>*>(y + g0 + L0) = o;
    } else {
#pragma unroll
      for (int u = 0; u < LPT; ++u)
        if (u >= own) y[g0 + L0 + u] = acc[u];
    }
  }
  double v = 0.0;
  if (dotv && valid) {
    // fused <y, x> with dotv == x: x at the output dofs is the own-dof
    // operand of the first valid entry (see gather_rec_body)
    T self[LPT];
#pragma unroll
    for (int u = 0; u < LPT; ++u) self[u] = T(0);
    bool found = false;
#pragma unroll
    for (int g = 0; g < NG; ++g)
#pragma unroll
      for (int q = 0; q < GS; ++q) {
#pragma unroll
        for (int u = 0; u < LPT; ++u) {
          T w = xv[g][q][0].v[u];
#pragma unroll
          for (int j = 1; j < NDF_IN; ++j) w = iis[g][q] == j ? xv[g][q][j].v[u] : w;
          self[u] = (ok[g][q] && !found) ? w : self[u];
        }
        found = found || ok[g][q];
      }
#pragma unroll
    for (int u = 0; u < LPT; ++u)
      if (u >= own) v += (double)acc[u] * (double)self[u];
  }
  return v;
}

template <typename T, int NDF_IN, int NG, int GS, int KIND, bool DOT, int LPT, bool NT = false, int BS = kThreads>
__global__ void __launch_bounds__(BS)
k_gather_vec(const T* __restrict__ M, const T* __restrict__ x, T* __restrict__ y,
             const int32_t* __restrict__ rec, int ncol, int nl, int64_t n3, int ndf_out,
             int overwrite, int64_t pcs, unsigned nb_pad, const T* __restrict__ dotv = nullptr,
             double* __restrict__ part = nullptr, const double* __restrict__ stop = nullptr) {
  if (DOT && stop && stop[0] != 0.0) return;
  const unsigned lb = lfr_xcd_block(blockIdx.x, nb_pad);
  double v = gather_vec_body<T, NDF_IN, NG, GS, KIND, LPT, NT, BS>(M, x, y, rec, ncol, nl, n3, ndf_out, overwrite,
                                                                   pcs, lb, DOT ? dotv : nullptr);
  if (DOT) {
    __shared__ double lds[BS / 64];
    v = block_sum<BS>(v, lds);
    if (threadIdx.x == 0) part[lb] = v;
  }
}

// ---------------------------------------------------------------------------
// k_gather for nf fields at once (a multidata field, "ndata first": element
// (dof d, field r) at d*nf + r).  Thread <-> (column, level, FPT fields); the
// threads of one output dof sit side by side in a wavefront, so a matrix
// element is one load shared by nf/FPT lanes and each operand dof is one
// nf-wide contiguous read.  The terms of every field are added in exactly the
// order of k_gather's runtime-ndf path (bit-identical to NF single-field
// applies).
// ---------------------------------------------------------------------------
template <typename T, int GS, int KIND, int FPT>
__global__ void __launch_bounds__(kThreads)
k_gather_nf(const T* __restrict__ M, const T* __restrict__ x, T* __restrict__ y,
            const int32_t* __restrict__ map_in, const int32_t* __restrict__ col_start,
            const int32_t* __restrict__ ent, int ncol, int H, int ngroup, int nl, int64_t n3, int ndf_out,
            int ndf_in, int nf, int overwrite, unsigned nb_pad) {
  // FPT fields per thread (nf % FPT == 0): thread <-> (column, level, field group)
  const unsigned lb = lfr_xcd_block(blockIdx.x, nb_pad);
  const int ng = nf / FPT;
  const int64_t t = (int64_t)lb * kThreads + threadIdx.x;
  const bool valid = t < (int64_t)ncol * H * ng;
  const int64_t tt = valid ? t : 0;
  const int64_t slot = tt / ng;
  const int r0 = (int)(tt - slot * ng) * FPT;
  const int col = (int)(slot / H);
  const int L = (int)(slot - (int64_t)col * H);
  const int32_t g0 = col_start[col];
  T acc[FPT];
  if (overwrite || !valid) {
#pragma unroll
    for (int f = 0; f < FPT; ++f) acc[f] = T(0);
  } else {
    const LVec<T, FPT> yv = ldvec<T, FPT, false>(y + (int64_t)(g0 + L) * nf + r0);
#pragma unroll
    for (int f = 0; f < FPT; ++f) acc[f] = yv.v[f];
  }
  const int32_t* e = ent + (int64_t)col * ngroup * GS * 2;
  for (int g = 0; g < ngroup; ++g) {
    const int32_t* eg = e + g * GS * 2;
    const int cc = eg[0];
    int ii[GS], kk[GS];
    bool ok[GS];
#pragma unroll
    for (int q = 0; q < GS; ++q) {
      const int code = eg[2 * q + 1];
      const int k = L - (code & 1);
      ok[q] = code >= 0 && k >= 0 && k < nl;
      ii[q] = code >= 0 ? (code >> 1) : 0;
      kk[q] = k < 0 ? 0 : (k >= nl ? nl - 1 : k);
    }
    const int32_t* mi = map_in + (int64_t)cc * ndf_in;
    if (KIND == 0) {
      for (int j = 0; j < ndf_in; ++j) {
        const int xbj = mi[j];
#pragma unroll
        for (int q = 0; q < GS; ++q) {
          const int64_t ik = (int64_t)cc * nl + kk[q];
          const T m = M[ik + n3 * ((int64_t)ii[q] + (int64_t)ndf_out * j)];
          const LVec<T, FPT> xv = ldvec<T, FPT, false>(x + (int64_t)(xbj + kk[q]) * nf + r0);
#pragma unroll
          for (int f = 0; f < FPT; ++f) {
            const T p = m * xv.v[f];
            const T s = acc[f] + p;
            acc[f] = ok[q] ? s : acc[f];
          }
        }
      }
    } else {
#pragma unroll
      for (int q = 0; q < GS; ++q) {
        const int64_t ik = (int64_t)cc * nl + kk[q];
        T dot[FPT];
#pragma unroll
        for (int f = 0; f < FPT; ++f) dot[f] = T(0);
        for (int j = 0; j < ndf_in; ++j) {
          const T m = M[ik + n3 * ((int64_t)j + (int64_t)ndf_in * ii[q])];
          const LVec<T, FPT> xv = ldvec<T, FPT, false>(x + (int64_t)(mi[j] + kk[q]) * nf + r0);
#pragma unroll
          for (int f = 0; f < FPT; ++f) {
            const T p = m * xv.v[f];
            dot[f] = dot[f] + p;
          }
        }
#pragma unroll
        for (int f = 0; f < FPT; ++f) {
          const T s = acc[f] + dot[f];
          acc[f] = ok[q] ? s : acc[f];
        }
      }
    }
  }
  if (valid) {
    LVec<T, FPT> o;
#pragma unroll
    for (int f = 0; f < FPT; ++f) o.v[f] = acc[f];
    *reinterpret_cast<LVec<T, FPT>*>(y + (int64_t)(g0 + L) * nf + r0) = o;
  }
}

// ---------------------------------------------------------------------------
// Unified gather over every output column of a plan in memory order, for
// plans whose columns all have two entries and nl or nl+1 levels (W2 lowest
// order: side-face columns shared by two cells, and bottom/top face stacks
// touched by one cell through two local dofs whose layers are one apart).
// One launch writes y front to back, so the lines a column shares with its
// neighbours are completed while they are still in L2, and the operand is
// fetched once.  Record per column (ints):
//   [ g0 | H | jouter | per entry e = 0, 1: cell*nl, code, map_in(cell, 0..NDF_IN-1) ]
// padded to x4.  A thread takes LPT consecutive output levels [L0, L0+LPT) of
// one column (ceil((nl+1)/LPT) threads per column; windows clamped inside the
// column, overlapping levels written by their owner only).  Entry e reads
// cell layers k = L - shift: its loads start at kst = L0 - shift clamped to
// [0, nl-LPT], and lane u takes element u + (kst - start) of what was loaded
// (a rotation by -1, 0 or +1); levels whose k falls outside [0, nl) are
// masked.  Summation order per output: kind 0 adds the twelve products entry
// by entry (two cells, `jouter` = 0) or input dof by input dof (one cell with
// two local dofs, `jouter` = 1), as k_gather_rec does for each class; kind 1
// adds one dot product per entry.
// ---------------------------------------------------------------------------
template <int NDF_IN>
struct UniRec {
  static constexpr int kEntry = 2 + NDF_IN;
  static constexpr int kInts = 3 + 2 * kEntry;
  static constexpr int kStride = (kInts + 3) & ~3;
};

template <typename T, int LPT>
__device__ __forceinline__ T rot_pick(const LVec<T, LPT>& v, int u, int d) {
  const T a = v.v[u];
  const T b = v.v[u + 1 < LPT ? u + 1 : LPT - 1];
  const T c = v.v[u > 0 ? u - 1 : 0];
  return d == 0 ? a : (d > 0 ? b : c);
}

template <typename T, int NDF_IN, int KIND, bool DOT, int LPT, bool NT, bool STAGE = false>
__global__ void __launch_bounds__(kThreads)
k_gather_uni(const T* __restrict__ M, const T* __restrict__ x, T* __restrict__ y,
             const int32_t* __restrict__ urec, int ncol, int nl, int64_t n3, int ndf_out, int overwrite,
             int64_t pcs, unsigned nb_pad, const T* __restrict__ dotv = nullptr,
             double* __restrict__ part = nullptr, const double* __restrict__ stop = nullptr) {
  if (DOT && stop && stop[0] != 0.0) return;
  using S = UniRec<NDF_IN>;
  const unsigned lb = lfr_xcd_block(blockIdx.x, nb_pad);
  const int tpc = (nl + LPT) / LPT;  // ceil((nl + 1) / LPT)
  const int64_t t = (int64_t)lb * kThreads + threadIdx.x;
  const bool valid = t < (int64_t)ncol * tpc;
  const int64_t tt = valid ? t : 0;
  const int col = (int)(tt / tpc);
  const int p = (int)(tt - (int64_t)col * tpc);
  int r[S::kStride];
  const int4* r4 = reinterpret_cast<const int4*>(urec + (int64_t)col * S::kStride);
#pragma unroll
  for (int q = 0; q < S::kStride / 4; ++q) {
    const int4 v = r4[q];
    r[4 * q + 0] = v.x;
    r[4 * q + 1] = v.y;
    r[4 * q + 2] = v.z;
    r[4 * q + 3] = v.w;
  }
  const int32_t g0 = r[0];
  const int H = r[1];
  const bool jouter = r[2] != 0;
  const int L0 = min(p * LPT, H - LPT);
  const int own = p * LPT - L0;  // levels below L0 + own belong to the previous thread
  T acc[LPT];
  if (overwrite || !valid) {
#pragma unroll
    for (int u = 0; u < LPT; ++u) acc[u] = T(0);
  } else {
    const LVec<T, LPT> yv = ldvec<T, LPT, false>(y + g0 + L0);
#pragma unroll
    for (int u = 0; u < LPT; ++u) acc[u] = yv.v[u];
  }
  T mv[2][NDF_IN][LPT], xv[2][NDF_IN][LPT];
  bool ok[2][LPT];
  int iis[2];
#pragma unroll
  for (int e = 0; e < 2; ++e) {
    const int* re = r + 3 + e * S::kEntry;
    const int64_t cb = (int64_t)re[0];
    const int code = re[1];
    const int ii = code >= 0 ? (code >> 1) : 0;
    iis[e] = ii;
    const int kst = L0 - (code & 1);
    const int xs = min(max(kst, 0), nl - LPT);
    const int d = kst - xs;
#pragma unroll
    for (int u = 0; u < LPT; ++u) ok[e][u] = code >= 0 && kst + u >= 0 && kst + u < nl;
    const int64_t js = pcs > 0 ? (int64_t)(nl + 1) : (KIND == 0 ? n3 * ndf_out : n3);
    const int64_t eb = pcs > 0 ? (int64_t)col * pcs + (int64_t)e * NDF_IN * (nl + 1) + L0
                               : cb + xs + n3 * (KIND == 0 ? (int64_t)ii : (int64_t)NDF_IN * ii);
    LVec<T, LPT> lm[NDF_IN], lx[NDF_IN];
#pragma unroll
    for (int j = 0; j < NDF_IN; ++j) {
      lm[j] = ldvec<T, LPT, NT>(M + eb + js * j);
      lx[j] = ldvec<T, LPT, false>(x + re[2 + j] + xs);
    }
#pragma unroll
    for (int j = 0; j < NDF_IN; ++j)
#pragma unroll
      for (int u = 0; u < LPT; ++u) {
        // the packed copy is stored by output level: no rotation
        mv[e][j][u] = pcs > 0 ? lm[j].v[u] : rot_pick<T, LPT>(lm[j], u, d);
        xv[e][j][u] = rot_pick<T, LPT>(lx[j], u, d);
      }
  }
#pragma unroll
  for (int u = 0; u < LPT; ++u) {
    if (KIND == 0) {
      T a = acc[u];
#pragma unroll
      for (int e = 0; e < 2; ++e)
#pragma unroll
        for (int j = 0; j < NDF_IN; ++j) {
          const T pr = mv[e][j][u] * xv[e][j][u];
          const T s = a + pr;
          a = ok[e][u] ? s : a;
        }
      T b = acc[u];
#pragma unroll
      for (int j = 0; j < NDF_IN; ++j)
#pragma unroll
        for (int e = 0; e < 2; ++e) {
          const T pr = mv[e][j][u] * xv[e][j][u];
          const T s = b + pr;
          b = ok[e][u] ? s : b;
        }
      acc[u] = jouter ? b : a;
    } else {
#pragma unroll
      for (int e = 0; e < 2; ++e) {
        T dot = T(0);
#pragma unroll
        for (int j = 0; j < NDF_IN; ++j) {
          const T pr = mv[e][j][u] * xv[e][j][u];
          dot = dot + pr;
        }
        const T s = acc[u] + dot;
        acc[u] = ok[e][u] ? s : acc[u];
      }
    }
  }
  if (STAGE) {
    // The workgroup's threads own one contiguous dof range [wlo, whi)
    // (columns in memory order, levels in thread order): stage it in LDS
    // and write it with aligned 16-byte stores.
    __shared__ T buf[kThreads * LPT];
    __shared__ int64_t range[2];
    const int64_t nthr = (int64_t)ncol * tpc;
    const int64_t tlast = min((int64_t)lb * kThreads + kThreads, nthr) - 1;
    if (threadIdx.x == 0) range[0] = (int64_t)g0 + min(p * LPT, H);
    if (t == tlast) range[1] = (int64_t)g0 + min((p + 1) * LPT, H);
    __syncthreads();
    const int64_t wlo = range[0], whi = range[1];
    if (valid)
#pragma unroll
      for (int u = 0; u < LPT; ++u)
        if (u >= own) buf[(int64_t)g0 + L0 + u - wlo] = acc[u];
    __syncthreads();
    if ((int64_t)lb * kThreads < nthr) {
      constexpr int V = 16 / (int)sizeof(T);
      const int64_t a = (wlo + V - 1) & ~(int64_t)(V - 1);
      const int64_t bnd = whi & ~(int64_t)(V - 1);
      if (a >= bnd) {
        for (int64_t i = wlo + threadIdx.x; i < whi; i += kThreads) y[i] = buf[i - wlo];
      } else {
        if (wlo + threadIdx.x < a) y[wlo + threadIdx.x] = buf[threadIdx.x];
        if (bnd + threadIdx.x < whi) y[bnd + threadIdx.x] = buf[bnd + threadIdx.x - wlo];
        for (int64_t q = a + (int64_t)threadIdx.x * V; q < bnd; q += (int64_t)kThreads * V) {
          LVec<T, V> o;
#pragma unroll
          for (int u = 0; u < V; ++u) o.v[u] = buf[q - wlo + u];
          // streaming (non-temporal) stores: y is written once per apply
#pragma unroll
          for (int u = 0; u < V; ++u) __builtin_nontemporal_store(o.v[u], y + q + u);
        }
      }
    }
  } else if (valid) {
    if (own == 0) {
      LVec<T, LPT> o;
#pragma unroll
      for (int u = 0; u < LPT; ++u) o.v[u] = acc[u];
      *reinterpret_cast<LVec<T, LPT>*>(y + g0 + L0) = o;
    } else {
#pragma unroll
      for (int u = 0; u < LPT; ++u)
        if (u >= own) y[g0 + L0 + u] = acc[u];
    }
  }
  if (DOT) {
    // fused <y, x> with dotv == x: x at the output dof is the operand of the
    // own local dof of the first entry valid at that level
    double v = 0.0;
    if (valid) {
#pragma unroll
      for (int u = 0; u < LPT; ++u) {
        T self = T(0);
        bool found = false;
#pragma unroll
        for (int e = 0; e < 2; ++e) {
          T w = xv[e][0][u];
#pragma unroll
          for (int j = 1; j < NDF_IN; ++j) w = iis[e] == j ? xv[e][j][u] : w;
          self = (ok[e][u] && !found) ? w : self;
          found = found || ok[e][u];
        }
        if (u >= own) v += (double)acc[u] * (double)self;
      }
    }
    __shared__ double lds[kThreads / 64];
    v = block_sum(v, lds);
    if (threadIdx.x == 0) part[lb] = v;
  }
}

// ---------------------------------------------------------------------------
// Cell-centric apply for a discontinuous output space: thread <-> (cell, k).
// Each output dof has exactly one (cell, df) contributor, so per output the
// sum over df2 in ascending order is the reference order
// (dg_matrix_vector_kernel_mod.F90:94-107, dg_inc_matrix_vector_kernel_mod.F90:134-141).
// ---------------------------------------------------------------------------
template <typename T, int NDF_OUT, int NDF_IN, int LPT = 1>
__global__ void __launch_bounds__(kThreads)
k_cells(const T* __restrict__ M, const T* __restrict__ x, T* __restrict__ y,
        const int32_t* __restrict__ map_out, const int32_t* __restrict__ map_in, int ncell,
        int nl, int64_t n3, int ndf_out_rt, int ndf_in_rt, int inc, unsigned nb_pad) {
  const unsigned lb = lfr_xcd_block(blockIdx.x, nb_pad);
  const int64_t t = (int64_t)lb * kThreads + threadIdx.x;
  const int ndf_in = NDF_IN > 0 ? NDF_IN : ndf_in_rt;
  const int ndf_out = NDF_OUT > 0 ? NDF_OUT : ndf_out_rt;
  if constexpr (NDF_OUT > 0 && NDF_IN > 0) {
    // thread <-> (cell, j): levels j, j + S, ... (S = ceil(nl / LPT)), so one
    // load instruction still reads contiguous runs of a column across the
    // wavefront and the cell's map entries are read once for LPT levels
    const int S = (nl + LPT - 1) / LPT;
    if (t >= (int64_t)ncell * S) return;
    const int c = (int)(t / S);
    const int j = (int)(t - (int64_t)c * S);
    const int32_t* mo = map_out + (int64_t)c * ndf_out;
    const int32_t* mi = map_in + (int64_t)c * ndf_in;
    const T* Mc = M + (int64_t)c * nl;
    // levels past the column are clamped to its top for the loads (always in
    // bounds, so they issue without branches) and dropped at the store
    int kk[LPT];
    bool ok[LPT];
#pragma unroll
    for (int q = 0; q < LPT; ++q) {
      const int k = j + q * S;
      ok[q] = k < nl;
      kk[q] = ok[q] ? k : nl - 1;
    }
    T acc[LPT][NDF_OUT];
#pragma unroll
    for (int q = 0; q < LPT; ++q)
#pragma unroll
      for (int i = 0; i < NDF_OUT; ++i) acc[q][i] = inc ? y[mo[i] + kk[q]] : T(0);
#pragma unroll
    for (int jj = 0; jj < NDF_IN; ++jj) {
#pragma unroll
      for (int q = 0; q < LPT; ++q) {
        const T xv = x[mi[jj] + kk[q]];
#pragma unroll
        for (int i = 0; i < NDF_OUT; ++i) {
          const T p = Mc[kk[q] + n3 * ((int64_t)i + (int64_t)NDF_OUT * jj)] * xv;
          acc[q][i] = acc[q][i] + p;
        }
      }
    }
#pragma unroll
    for (int q = 0; q < LPT; ++q)
      if (ok[q])
#pragma unroll
        for (int i = 0; i < NDF_OUT; ++i) y[mo[i] + kk[q]] = acc[q][i];
  } else {
    if (t >= (int64_t)ncell * nl) return;
    const int c = (int)(t / nl);
    const int k = (int)(t - (int64_t)c * nl);
    const int64_t ik = t;
    const int32_t* mo = map_out + (int64_t)c * ndf_out;
    const int32_t* mi = map_in + (int64_t)c * ndf_in;
    for (int i = 0; i < ndf_out; ++i) {
      T acc = inc ? y[mo[i] + k] : T(0);
      for (int j = 0; j < ndf_in; ++j) {
        const T p = M[ik + n3 * ((int64_t)i + (int64_t)ndf_out * j)] * x[mi[j] + k];
        acc = acc + p;
      }
      y[mo[i] + k] = acc;
    }
  }
}

}  // namespace lfr
