// Intergrid transfers of lowest-order scalar fields between two meshes of a
// multigrid hierarchy (W3: ndf 1, levels 0..nl-1; Wtheta: ndf 2, levels
// 0..nl, a cell's two dofs one apart in its column):
//   restrict_scalar_unweighted_code
//   (components/science/source/kernel/inter_mesh/sci_restrict_scalar_unweighted_kernel_mod.F90:71-125):
//     coarse(k) = sum over the children, y outer and x inner, of
//     fine(child, k) * denom, denom = 1 / (nfx nfy), from zero;
//   prolong_scalar_unweighted_code
//   (components/science/source/kernel/inter_mesh/sci_prolong_scalar_unweighted_kernel_mod.F90):
//     fine(child, k) += coarse(k).
// One thread per coarse cell and LPT of its levels, strided by S =
// ceil(levels / LPT) (thread j of a column takes levels j, j + S, ...), so
// the loads of one instruction are contiguous runs of a column across a
// wavefront, a thread's map entries are read once for LPT levels and LPT
// loads per child are in flight; each output is one thread's sum in the
// reference order.  The children's first dofs are resolved once, when the
// intergrid handle is made, so a thread's loads do not wait on the map.
#include <cstdlib>
#include <type_traits>
#include <vector>

#include "lfr_internal.h"

namespace {

// fbase(child, coarse cell): first (0-based) dof of the child's column, in
// the map's order (x fastest, then y).  Levels past the column are clamped
// to its top for the loads (always in bounds, so every load issues without a
// branch) and masked out of the sums and stores.  NCH > 0: that many
// children, unrolled, every child's loads issued before the first sum.
template <typename T, int LPT>
__device__ __forceinline__ void level_slots(int j, int S, int nlev, int (&kk)[LPT], bool (&ok)[LPT]) {
#pragma unroll
  for (int q = 0; q < LPT; ++q) {
    const int k = j + q * S;
    ok[q] = k < nlev;
    kk[q] = ok[q] ? k : nlev - 1;
  }
}

template <typename T, int LPT, int NCH>
__global__ void __launch_bounds__(256)
k_restrict_scalar(T* __restrict__ coarse, const T* __restrict__ fine, const int32_t* __restrict__ map_c,
                  const int32_t* __restrict__ fbase, int ncc, int ndf, int nlev, int S, int nch_rt, T denom) {
  const int64_t t = (int64_t)blockIdx.x * 256 + threadIdx.x;
  if (t >= (int64_t)ncc * S) return;
  const int c = (int)(t / S);
  const int j = (int)(t - (int64_t)c * S);
  const int nch = NCH > 0 ? NCH : nch_rt;
  const int32_t* fb = fbase + (int64_t)c * nch;
  int kk[LPT];
  bool ok[LPT];
  level_slots<T, LPT>(j, S, nlev, kk, ok);
  T v[LPT];
#pragma unroll
  for (int q = 0; q < LPT; ++q) v[q] = T(0);
  if constexpr (NCH > 0) {
    int fbv[NCH];
#pragma unroll
    for (int i = 0; i < NCH; ++i) fbv[i] = fb[i];
    T val[NCH][LPT];
#pragma unroll
    for (int i = 0; i < NCH; ++i)
#pragma unroll
      for (int q = 0; q < LPT; ++q) val[i][q] = fine[fbv[i] + kk[q]];
#pragma unroll
    for (int i = 0; i < NCH; ++i)   // y outer, x inner
#pragma unroll
      for (int q = 0; q < LPT; ++q) {
        const T p = val[i][q] * denom;
        if (ok[q]) v[q] = v[q] + p;
      }
  } else {
    for (int i = 0; i < nch; ++i) {   // y outer, x inner
      const T* f = fine + fb[i];
      T val[LPT];
#pragma unroll
      for (int q = 0; q < LPT; ++q) val[q] = f[kk[q]];
#pragma unroll
      for (int q = 0; q < LPT; ++q) {
        const T p = val[q] * denom;
        if (ok[q]) v[q] = v[q] + p;
      }
    }
  }
  T* o = coarse + map_c[(int64_t)c * ndf];
#pragma unroll
  for (int q = 0; q < LPT; ++q)
    if (ok[q]) o[kk[q]] = v[q];
}

template <typename T, int LPT, int NCH>
__global__ void __launch_bounds__(256)
k_prolong_scalar(T* __restrict__ fine, const T* __restrict__ coarse, const int32_t* __restrict__ map_c,
                 const int32_t* __restrict__ fbase, int ncc, int ndf, int nlev, int S, int nch_rt) {
  const int64_t t = (int64_t)blockIdx.x * 256 + threadIdx.x;
  if (t >= (int64_t)ncc * S) return;
  const int c = (int)(t / S);
  const int j = (int)(t - (int64_t)c * S);
  const int nch = NCH > 0 ? NCH : nch_rt;
  const int32_t* fb = fbase + (int64_t)c * nch;
  const T* cc = coarse + map_c[(int64_t)c * ndf];
  int kk[LPT];
  bool ok[LPT];
  level_slots<T, LPT>(j, S, nlev, kk, ok);
  T cv[LPT];
#pragma unroll
  for (int q = 0; q < LPT; ++q) cv[q] = cc[kk[q]];
  if constexpr (NCH > 0) {
    int fbv[NCH];
#pragma unroll
    for (int i = 0; i < NCH; ++i) fbv[i] = fb[i];
    T val[NCH][LPT];
#pragma unroll
    for (int i = 0; i < NCH; ++i)
#pragma unroll
      for (int q = 0; q < LPT; ++q) val[i][q] = fine[fbv[i] + kk[q]];
#pragma unroll
    for (int i = 0; i < NCH; ++i)
#pragma unroll
      for (int q = 0; q < LPT; ++q)
        if (ok[q]) fine[fbv[i] + kk[q]] = val[i][q] + cv[q];
  } else {
    for (int i = 0; i < nch; ++i) {
      T* f = fine + fb[i];
      T val[LPT];
#pragma unroll
      for (int q = 0; q < LPT; ++q) val[q] = f[kk[q]];
#pragma unroll
      for (int q = 0; q < LPT; ++q)
        if (ok[q]) f[kk[q]] = val[q] + cv[q];
    }
  }
}

// restrict_w2_code
// (components/science/source/kernel/inter_mesh/sci_restrict_w2_kernel_mod.F90):
// a coarse side face (W, S, E, N) is the sum of the fine faces on it (nl
// levels), the B face column the sum over every child (nl + 1 levels), each
// summed y outer, x inner.  One thread per coarse cell and LPT of its nl + 1
// levels (strided by S); wmask bit f says this cell writes face f (the last
// writer in cell order of the reference's loop, whose faces come from the
// face selectors), and fface(child, face, coarse cell) is the first dof
// (0-based) of the fine face column.
template <typename T, int LPT, int N>
__global__ void __launch_bounds__(256)
k_restrict_w2(T* __restrict__ coarse, const T* __restrict__ fine, const int32_t* __restrict__ map_c,
              const int32_t* __restrict__ fface, const uint8_t* __restrict__ wmask, int ncc, int nl, int S,
              int nfx, int nfy) {
  const int64_t t = (int64_t)blockIdx.x * 256 + threadIdx.x;
  if (t >= (int64_t)ncc * S) return;
  const int c = (int)(t / S);
  const int j = (int)(t - (int64_t)c * S);
  const int nch = nfx * nfy;
  const int32_t* ff = fface + (int64_t)c * nch * 5;
  const int32_t* mc = map_c + (int64_t)c * 6;
  const unsigned m = wmask[c];
  // side faces have nl levels, the B face column nl + 1 (loads clamped into
  // the column, masked out of sums and stores)
  int ks[LPT], kb[LPT];
  bool oks[LPT], okb[LPT];
  level_slots<T, LPT>(j, S, nl, ks, oks);
  level_slots<T, LPT>(j, S, nl + 1, kb, okb);
  T v[LPT];
#pragma unroll
  for (int df = 0; df < 4; ++df) {   // W, S, E, N
    if (!((m >> df) & 1u)) continue;
    int x0 = 0, x1 = nfx - 1, y0 = 0, y1 = nfy - 1;
    if (df == 3) y1 = 0;              // N: first row
    else if (df == 1) y0 = nfy - 1;   // S: last row
    else if (df == 0) x1 = 0;         // W: first column
    else x0 = nfx - 1;                // E: last column
#pragma unroll
    for (int q = 0; q < LPT; ++q) v[q] = T(0);
    if constexpr (N > 0) {
      // the N fine faces on this side, y outer and x inner
      const T* fs[N];
#pragma unroll
      for (int i = 0; i < N; ++i) {
        const int x = x0 == x1 ? x0 : x0 + i, y = y0 == y1 ? y0 : y0 + i;
        fs[i] = fine + ff[(x + N * y) * 5 + df];
      }
      T val[N][LPT];
#pragma unroll
      for (int i = 0; i < N; ++i)
#pragma unroll
        for (int q = 0; q < LPT; ++q) val[i][q] = fs[i][ks[q]];
#pragma unroll
      for (int i = 0; i < N; ++i)
#pragma unroll
        for (int q = 0; q < LPT; ++q)
          if (oks[q]) v[q] = v[q] + val[i][q];
    } else {
      for (int y = y0; y <= y1; ++y)
        for (int x = x0; x <= x1; ++x) {
          const T* f = fine + ff[(x + nfx * y) * 5 + df];
          T val[LPT];
#pragma unroll
          for (int q = 0; q < LPT; ++q) val[q] = f[ks[q]];
#pragma unroll
          for (int q = 0; q < LPT; ++q)
            if (oks[q]) v[q] = v[q] + val[q];
        }
    }
    T* o = coarse + mc[df];
#pragma unroll
    for (int q = 0; q < LPT; ++q)
      if (oks[q]) o[ks[q]] = v[q];
  }
#pragma unroll
  for (int q = 0; q < LPT; ++q) v[q] = T(0);
  if constexpr (N > 0) {
    const T* fs[N * N];
#pragma unroll
    for (int i = 0; i < N * N; ++i) fs[i] = fine + ff[i * 5 + 4];
    T val[N * N][LPT];
#pragma unroll
    for (int i = 0; i < N * N; ++i)
#pragma unroll
      for (int q = 0; q < LPT; ++q) val[i][q] = fs[i][kb[q]];
#pragma unroll
    for (int i = 0; i < N * N; ++i)
#pragma unroll
      for (int q = 0; q < LPT; ++q)
        if (okb[q]) v[q] = v[q] + val[i][q];
  } else {
    for (int i = 0; i < nch; ++i) {
      const T* f = fine + ff[i * 5 + 4];
      T val[LPT];
#pragma unroll
      for (int q = 0; q < LPT; ++q) val[q] = f[kb[q]];
#pragma unroll
      for (int q = 0; q < LPT; ++q)
        if (okb[q]) v[q] = v[q] + val[q];
    }
  }
  T* o = coarse + mc[4];
#pragma unroll
  for (int q = 0; q < LPT; ++q)
    if (okb[q]) o[kb[q]] = v[q];
}

// prolong_w2_kernel_code
// (components/science/source/kernel/inter_mesh/sci_prolong_w2_kernel_mod.F90)
// with the weights of weights_prolong_w2_kernel_code
// (.../sci_weights_prolong_w2_kernel_mod.F90) formed in place: a fine face on
// a coarse side face takes 1/n of the coarse flux (n fine faces per coarse
// face), an interior fine W (N) face at column x (row y) the linear blend
// (1/n) (w a + (1 - w) b) of the coarse W and E (N and S) fluxes with
// w = (n + 1 - x) / n, a fine B/T face 1/n^2 of the coarse one.  Square
// children only (nfx = nfy = n): a fine face shared by two coarse cells then
// gets the same weight, and the same value, from both, so only the coarse
// cell that owns the shared coarse side (wmask, as for the restriction)
// writes it: 12 fine face columns per coarse cell instead of 16.  One thread
// per coarse cell and LPT of its nl + 1 levels (strided by S).
template <typename T, int LPT>
__global__ void __launch_bounds__(256)
k_prolong_w2(T* __restrict__ fine, const T* __restrict__ coarse, const int32_t* __restrict__ map_c,
             const int32_t* __restrict__ fface, const uint8_t* __restrict__ wmask, int ncc, int nl, int S, int n) {
  const int64_t t = (int64_t)blockIdx.x * 256 + threadIdx.x;
  if (t >= (int64_t)ncc * S) return;
  const int c = (int)(t / S);
  const int j = (int)(t - (int64_t)c * S);
  const int32_t* ff = fface + (int64_t)c * n * n * 5;
  const int32_t* mc = map_c + (int64_t)c * 6;
  const unsigned m = wmask[c];   // bits W, S, E, N
  // fine face column of child (x, y) (0-based), face df
  auto fcol = [&](int x, int y, int df) { return fine + ff[(x + n * y) * 5 + df]; };
  auto side = [&](T* f, const T* val) {
#pragma unroll
    for (int q = 0; q < LPT; ++q)
      if (j + q * S < nl) f[j + q * S] = val[q];
  };
  const T w1 = T(1) / T(n);
  // loads clamped into the columns (side faces nl levels, B/T nl + 1), so
  // they issue without branches; the stores are masked
  int ks[LPT], kb[LPT];
  bool oks[LPT], okb[LPT];
  level_slots<T, LPT>(j, S, nl, ks, oks);
  level_slots<T, LPT>(j, S, nl + 1, kb, okb);
  T cw[LPT], cs[LPT], ce[LPT], cn[LPT], cb[LPT], tmp[LPT];
#pragma unroll
  for (int q = 0; q < LPT; ++q) {
    cw[q] = coarse[mc[0] + ks[q]];
    cs[q] = coarse[mc[1] + ks[q]];
    ce[q] = coarse[mc[2] + ks[q]];
    cn[q] = coarse[mc[3] + ks[q]];
    cb[q] = coarse[mc[4] + kb[q]];
  }
  if (m & 8u) {
#pragma unroll
    for (int q = 0; q < LPT; ++q) tmp[q] = w1 * cn[q];
    for (int x = 0; x < n; ++x) side(fcol(x, 0, 3), tmp);       // N faces of the first row
  }
  if (m & 2u) {
#pragma unroll
    for (int q = 0; q < LPT; ++q) tmp[q] = w1 * cs[q];
    for (int x = 0; x < n; ++x) side(fcol(x, n - 1, 1), tmp);   // S faces of the last row
  }
  if (m & 1u) {
#pragma unroll
    for (int q = 0; q < LPT; ++q) tmp[q] = w1 * cw[q];
    for (int y = 0; y < n; ++y) side(fcol(0, y, 0), tmp);       // W faces of the first column
  }
  if (m & 4u) {
#pragma unroll
    for (int q = 0; q < LPT; ++q) tmp[q] = w1 * ce[q];
    for (int y = 0; y < n; ++y) side(fcol(n - 1, y, 2), tmp);   // E faces of the last column
  }
  for (int x = 1; x < n; ++x) {                                  // interior W faces
    const T iw = T(n - x) / T(n);                                // (n + 1 - x_1based) / n
#pragma unroll
    for (int q = 0; q < LPT; ++q) {
      const T a = iw * cw[q];
      const T b = (T(1) - iw) * ce[q];
      tmp[q] = w1 * (a + b);
    }
    for (int y = 0; y < n; ++y) side(fcol(x, y, 0), tmp);
  }
  for (int y = 1; y < n; ++y) {                                  // interior N faces
    const T iw = T(n - y) / T(n);
#pragma unroll
    for (int q = 0; q < LPT; ++q) {
      const T a = iw * cn[q];
      const T b = (T(1) - iw) * cs[q];
      tmp[q] = w1 * (a + b);
    }
    for (int x = 0; x < n; ++x) side(fcol(x, y, 3), tmp);
  }
  const T wb = T(1) / T(n * n);
#pragma unroll
  for (int q = 0; q < LPT; ++q) tmp[q] = wb * cb[q];
  for (int i = 0; i < n * n; ++i) {
    T* f = fine + ff[i * 5 + 4];
#pragma unroll
    for (int q = 0; q < LPT; ++q)
      if (j + q * S <= nl) f[j + q * S] = tmp[q];
  }
}

// every cell's dofs 1..ndf are one column (dof j = dof 1 + j - 1)
bool column_dofs(const lfr_space* sp) {
  if (sp->vstack >= 0) return sp->vstack == 1;
  int ok = 1;
  for (int c = 0; c < sp->ncell && ok; ++c)
    for (int j = 1; j < sp->ndf; ++j)
      if (sp->map_h[(size_t)c * sp->ndf + j] != sp->map_h[(size_t)c * sp->ndf] + j) ok = 0;
  sp->vstack = ok;
  return ok;
}

}  // namespace

// The intergrid map of two meshes, checked and uploaded once (the
// mesh-to-mesh connectivity the reference holds on the mesh object).
struct lfr_intergrid {
  lfr_ctx* ctx = nullptr;
  const lfr_space* coarse = nullptr;
  const lfr_space* fine = nullptr;
  int nfx = 0, nfy = 0;
  int32_t* fbase_d = nullptr;   // scalar: (nfx nfy, ncell_coarse) first dof (0-based) of each child's column;
                                // W2: (5, nfx nfy, ncell_coarse) first dof of each child's W, S, E, N, B face column
  uint8_t* wmask_d = nullptr;   // W2: per coarse cell, the faces (bits W, S, E, N) it writes
};

namespace {

// Levels per thread of the transfer kernels (C192 <-> C96 L70 FP32, loads
// issued without branches): the W3 restriction 4 (19.8 µs; 20.9 at 2, 22.0
// at 1), the W2 restriction 2 (48.7 µs; 51.6 at 4), the prolongations 2
// (W3 32.3 µs, 34.8 at 4; W2 59.0 µs, 78.3 at 4: their strided stores touch
// more lines per instruction as LPT grows); LFR_TRANSFER_LPT=1, 2 or 4
// forces one value for all.
int transfer_lpt(bool restrict_, bool w2 = false) {
  static const int forced = [] {
    const char* e = getenv("LFR_TRANSFER_LPT");
    if (e && (e[0] == '1' || e[0] == '2' || e[0] == '4')) return e[0] - '0';
    return 0;
  }();
  return forced ? forced : (restrict_ && !w2 ? 4 : 2);
}

// f(std::integral_constant<int, LPT>) for lpt 1, 2 or 4
template <typename F>
void with_lpt(int lpt, F&& f) {
  if (lpt == 1) f(std::integral_constant<int, 1>{});
  else if (lpt == 2) f(std::integral_constant<int, 2>{});
  else f(std::integral_constant<int, 4>{});
}

int transfer(lfr_ctx* ctx, bool restrict_, int prec, void* out, const void* in, const lfr_intergrid* ig,
             const char* who) {
  if (!ctx) return LFR_ERR_ARG;
  LFR_REQUIRE(ctx, prec == 4 || prec == 8, "%s: prec must be 4 or 8, got %d", who, prec);
  LFR_REQUIRE(ctx, out && in && ig, "%s: null argument", who);
  LFR_REQUIRE(ctx, out != in, "%s: the output must not alias the input", who);
  const lfr_space* cs = ig->coarse;
  const int nfx = ig->nfx, nfy = ig->nfy;
  LFR_REQUIRE(ctx, cs->ndf != 6, "%s: a W2 intergrid map takes lfr_restrict_w2", who);
  const int nlev = cs->nlayers - 1 + cs->ndf;   // top_df + 1
  const int lpt = transfer_lpt(restrict_);
  const int S = (nlev + lpt - 1) / lpt;
  const unsigned nb = (unsigned)(((int64_t)cs->ncell * S + 255) / 256);
  const int32_t* fb = ig->fbase_d;
  const int nch = nfx * nfy;
  with_lpt(lpt, [&](auto L) {
    constexpr int LPT = decltype(L)::value;
    if (restrict_) {
      if (prec == 4) {
        if (nch == 4)
          hipLaunchKernelGGL((k_restrict_scalar<float, LPT, 4>), dim3(nb), dim3(256), 0, ctx->stream, (float*)out,
                             (const float*)in, cs->map0_d, fb, cs->ncell, cs->ndf, nlev, S, nch, 1.0f / (float)nch);
        else
          hipLaunchKernelGGL((k_restrict_scalar<float, LPT, 0>), dim3(nb), dim3(256), 0, ctx->stream, (float*)out,
                             (const float*)in, cs->map0_d, fb, cs->ncell, cs->ndf, nlev, S, nch, 1.0f / (float)nch);
      } else {
        if (nch == 4)
          hipLaunchKernelGGL((k_restrict_scalar<double, LPT, 4>), dim3(nb), dim3(256), 0, ctx->stream,
                             (double*)out, (const double*)in, cs->map0_d, fb, cs->ncell, cs->ndf, nlev, S, nch,
                             1.0 / (double)nch);
        else
          hipLaunchKernelGGL((k_restrict_scalar<double, LPT, 0>), dim3(nb), dim3(256), 0, ctx->stream,
                             (double*)out, (const double*)in, cs->map0_d, fb, cs->ncell, cs->ndf, nlev, S, nch,
                             1.0 / (double)nch);
      }
    } else {
      if (prec == 4) {
        if (nch == 4)
          hipLaunchKernelGGL((k_prolong_scalar<float, LPT, 4>), dim3(nb), dim3(256), 0, ctx->stream, (float*)out,
                             (const float*)in, cs->map0_d, fb, cs->ncell, cs->ndf, nlev, S, nch);
        else
          hipLaunchKernelGGL((k_prolong_scalar<float, LPT, 0>), dim3(nb), dim3(256), 0, ctx->stream, (float*)out,
                             (const float*)in, cs->map0_d, fb, cs->ncell, cs->ndf, nlev, S, nch);
      } else {
        if (nch == 4)
          hipLaunchKernelGGL((k_prolong_scalar<double, LPT, 4>), dim3(nb), dim3(256), 0, ctx->stream,
                             (double*)out, (const double*)in, cs->map0_d, fb, cs->ncell, cs->ndf, nlev, S, nch);
        else
          hipLaunchKernelGGL((k_prolong_scalar<double, LPT, 0>), dim3(nb), dim3(256), 0, ctx->stream,
                             (double*)out, (const double*)in, cs->map0_d, fb, cs->ncell, cs->ndf, nlev, S, nch);
      }
    }
  });
  LFR_CHECK_LAUNCH(ctx);
  return LFR_OK;
}

}  // namespace

extern "C" int lfr_intergrid_create(lfr_ctx* ctx, const lfr_space* coarse_space, const lfr_space* fine_space,
                                    const int32_t* cell_map_h, int nfx, int nfy, lfr_intergrid** out) {
  const char* who = "lfr_intergrid_create";
  if (!ctx || !out) return LFR_ERR_ARG;
  *out = nullptr;
  const lfr_space* cs = coarse_space;
  const lfr_space* fs = fine_space;
  LFR_REQUIRE(ctx, cs && fs && cell_map_h, "%s: null argument", who);
  LFR_REQUIRE(ctx, cs->ndf == fs->ndf && (cs->ndf == 1 || cs->ndf == 2 || cs->ndf == 6),
              "%s: lowest-order W3 (ndf 1), Wtheta (ndf 2) or W2 (ndf 6) spaces expected (ndf %d and %d)", who,
              cs->ndf, fs->ndf);
  LFR_REQUIRE(ctx, cs->nlayers == fs->nlayers, "%s: the meshes have different layer counts", who);
  LFR_REQUIRE(ctx, nfx >= 1 && nfy >= 1 && (int64_t)cs->ncell * nfx * nfy == fs->ncell,
              "%s: %d coarse cells x %d x %d children do not cover %d fine cells", who, cs->ncell, nfx, nfy,
              fs->ncell);
  const bool w2 = cs->ndf == 6;
  if (w2) {
    // lowest-order W2: local dofs W, S, E, N, B, T with T = B + 1 (the B
    // column holds every horizontal face of the cell's column)
    for (const lfr_space* sp : {cs, fs})
      for (int c = 0; c < sp->ncell; ++c)
        LFR_REQUIRE(ctx, sp->map_h[(size_t)c * 6 + 5] == sp->map_h[(size_t)c * 6 + 4] + 1,
                    "%s: cell %d is not a lowest-order W2 dofmap", who, c + 1);
  } else {
    LFR_REQUIRE(ctx, column_dofs(cs) && column_dofs(fs), "%s: a cell's dofs are not one column", who);
  }
  const int64_t nmap = (int64_t)cs->ncell * nfx * nfy;
  for (int64_t i = 0; i < nmap; ++i)
    LFR_REQUIRE(ctx, cell_map_h[i] >= 1 && cell_map_h[i] <= fs->ncell, "%s: cell_map entry %lld is %d (1..%d)",
                who, (long long)i, cell_map_h[i], fs->ncell);
  auto* ig = new lfr_intergrid();
  ig->ctx = ctx;
  ig->coarse = cs;
  ig->fine = fs;
  ig->nfx = nfx;
  ig->nfy = nfy;
  std::vector<int32_t> fb;
  std::vector<uint8_t> wm;
  if (!w2) {
    fb.resize((size_t)nmap);
    for (int64_t i = 0; i < nmap; ++i) fb[(size_t)i] = fs->map_h[(size_t)(cell_map_h[i] - 1) * fs->ndf] - 1;
  } else {
    fb.resize((size_t)nmap * 5);
    for (int64_t i = 0; i < nmap; ++i)
      for (int df = 0; df < 5; ++df)
        fb[(size_t)i * 5 + df] = fs->map_h[(size_t)(cell_map_h[i] - 1) * 6 + df] - 1;
    // face selectors of the coarse mesh, as face_selector_code
    // (components/science/source/kernel/algebra/sci_face_selector_kernel_mod.F90)
    // does over the cells in order: a cell takes its E (N) face unless an
    // earlier cell took that face as its E or N face
    const int ncc = cs->ncell;
    std::vector<char> taken((size_t)cs->undf + 1, 0);
    std::vector<int> ew(ncc, 1), ns(ncc, 1);
    for (int c = 0; c < ncc; ++c) {
      const int32_t* m = &cs->map_h[(size_t)c * 6];
      if (!taken[m[2]]) { ew[c] = 2; taken[m[2]] = 1; }
      if (!taken[m[3]]) { ns[c] = 2; taken[m[3]] = 1; }
    }
    // faces each cell writes (W, S, then E and/or N), and the last writer
    // of every face in cell order (the value the reference's loop leaves)
    std::vector<int> last((size_t)cs->undf + 1, -1);
    std::vector<uint8_t> writes(ncc, 0);
    for (int c = 0; c < ncc; ++c) {
      uint8_t w = 0x3;                          // W, S
      const int nface = ew[c] + ns[c];
      if (nface >= 3) w |= (ew[c] == 2) ? 0x4 : 0x8;   // face 3: E, or N when only N is selected
      if (nface == 4) w |= 0x8;                        // face 4: N
      writes[c] = w;
      for (int df = 0; df < 4; ++df)
        if ((w >> df) & 1) last[cs->map_h[(size_t)c * 6 + df]] = c;
    }
    wm.resize(ncc);
    for (int c = 0; c < ncc; ++c) {
      uint8_t w = 0;
      for (int df = 0; df < 4; ++df)
        if (((writes[c] >> df) & 1) && last[cs->map_h[(size_t)c * 6 + df]] == c) w |= (uint8_t)(1u << df);
      wm[c] = w;
    }
  }
  if (hipMalloc(&ig->fbase_d, fb.size() * sizeof(int32_t)) != hipSuccess ||
      (w2 && hipMalloc(&ig->wmask_d, wm.size()) != hipSuccess)) {
    (void)hipGetLastError();
    if (ig->fbase_d) (void)hipFree(ig->fbase_d);
    delete ig;
    return lfr_set_error(ctx, LFR_ERR_NOMEM, "%s: cannot allocate the map", who);
  }
  hipError_t e = hipMemcpy(ig->fbase_d, fb.data(), fb.size() * sizeof(int32_t), hipMemcpyHostToDevice);
  if (e == hipSuccess && w2) e = hipMemcpy(ig->wmask_d, wm.data(), wm.size(), hipMemcpyHostToDevice);
  if (e != hipSuccess) {
    (void)hipFree(ig->fbase_d);
    if (ig->wmask_d) (void)hipFree(ig->wmask_d);
    delete ig;
    return lfr_set_error(ctx, LFR_ERR_HIP, "%s: %s", who, hipGetErrorString(e));
  }
  *out = ig;
  return LFR_OK;
}

extern "C" int lfr_intergrid_destroy(lfr_intergrid* ig) {
  if (!ig) return LFR_OK;
  if (ig->fbase_d) (void)hipFree(ig->fbase_d);
  if (ig->wmask_d) (void)hipFree(ig->wmask_d);
  delete ig;
  return LFR_OK;
}

extern "C" int lfr_restrict_scalar(lfr_ctx* ctx, int prec, void* coarse_d, const void* fine_d,
                                   const lfr_intergrid* ig) {
  return transfer(ctx, true, prec, coarse_d, fine_d, ig, "lfr_restrict_scalar");
}

extern "C" int lfr_prolong_scalar(lfr_ctx* ctx, int prec, void* fine_d, const void* coarse_d,
                                  const lfr_intergrid* ig) {
  return transfer(ctx, false, prec, fine_d, coarse_d, ig, "lfr_prolong_scalar");
}

extern "C" int lfr_restrict_w2(lfr_ctx* ctx, int prec, void* coarse_d, const void* fine_d, const lfr_intergrid* ig) {
  const char* who = "lfr_restrict_w2";
  if (!ctx) return LFR_ERR_ARG;
  LFR_REQUIRE(ctx, prec == 4 || prec == 8, "%s: prec must be 4 or 8, got %d", who, prec);
  LFR_REQUIRE(ctx, coarse_d && fine_d && ig, "%s: null argument", who);
  LFR_REQUIRE(ctx, coarse_d != fine_d, "%s: the output must not alias the input", who);
  LFR_REQUIRE(ctx, ig->coarse->ndf == 6, "%s: the intergrid map is not between W2 spaces", who);
  const lfr_space* cs = ig->coarse;
  const int nl = cs->nlayers;
  const int lpt = transfer_lpt(true, true);
  const int S = (nl + 1 + lpt - 1) / lpt;
  const unsigned nb = (unsigned)(((int64_t)cs->ncell * S + 255) / 256);
  with_lpt(lpt, [&](auto L) {
    constexpr int LPT = decltype(L)::value;
    const bool two = ig->nfx == 2 && ig->nfy == 2;
    if (prec == 4) {
      if (two)
        hipLaunchKernelGGL((k_restrict_w2<float, LPT, 2>), dim3(nb), dim3(256), 0, ctx->stream, (float*)coarse_d,
                           (const float*)fine_d, cs->map0_d, ig->fbase_d, ig->wmask_d, cs->ncell, nl, S, ig->nfx,
                           ig->nfy);
      else
        hipLaunchKernelGGL((k_restrict_w2<float, LPT, 0>), dim3(nb), dim3(256), 0, ctx->stream, (float*)coarse_d,
                           (const float*)fine_d, cs->map0_d, ig->fbase_d, ig->wmask_d, cs->ncell, nl, S, ig->nfx,
                           ig->nfy);
    } else {
      if (two)
        hipLaunchKernelGGL((k_restrict_w2<double, LPT, 2>), dim3(nb), dim3(256), 0, ctx->stream,
                           (double*)coarse_d, (const double*)fine_d, cs->map0_d, ig->fbase_d, ig->wmask_d, cs->ncell,
                           nl, S, ig->nfx, ig->nfy);
      else
        hipLaunchKernelGGL((k_restrict_w2<double, LPT, 0>), dim3(nb), dim3(256), 0, ctx->stream,
                           (double*)coarse_d, (const double*)fine_d, cs->map0_d, ig->fbase_d, ig->wmask_d, cs->ncell,
                           nl, S, ig->nfx, ig->nfy);
    }
  });
  LFR_CHECK_LAUNCH(ctx);
  return LFR_OK;
}

extern "C" int lfr_prolong_w2(lfr_ctx* ctx, int prec, void* fine_d, const void* coarse_d, const lfr_intergrid* ig) {
  const char* who = "lfr_prolong_w2";
  if (!ctx) return LFR_ERR_ARG;
  LFR_REQUIRE(ctx, prec == 4 || prec == 8, "%s: prec must be 4 or 8, got %d", who, prec);
  LFR_REQUIRE(ctx, coarse_d && fine_d && ig, "%s: null argument", who);
  LFR_REQUIRE(ctx, coarse_d != fine_d, "%s: the output must not alias the input", who);
  LFR_REQUIRE(ctx, ig->coarse->ndf == 6, "%s: the intergrid map is not between W2 spaces", who);
  LFR_REQUIRE(ctx, ig->nfx == ig->nfy, "%s: square children only (%d x %d)", who, ig->nfx, ig->nfy);
  const lfr_space* cs = ig->coarse;
  const int nl = cs->nlayers;
  const int lpt = transfer_lpt(false);
  const int S = (nl + 1 + lpt - 1) / lpt;
  const unsigned nb = (unsigned)(((int64_t)cs->ncell * S + 255) / 256);
  with_lpt(lpt, [&](auto L) {
    constexpr int LPT = decltype(L)::value;
    if (prec == 4)
      hipLaunchKernelGGL((k_prolong_w2<float, LPT>), dim3(nb), dim3(256), 0, ctx->stream, (float*)fine_d,
                         (const float*)coarse_d, cs->map0_d, ig->fbase_d, ig->wmask_d, cs->ncell, nl, S, ig->nfx);
    else
      hipLaunchKernelGGL((k_prolong_w2<double, LPT>), dim3(nb), dim3(256), 0, ctx->stream, (double*)fine_d,
                         (const double*)coarse_d, cs->map0_d, ig->fbase_d, ig->wmask_d, cs->ncell, nl, S, ig->nfx);
  });
  LFR_CHECK_LAUNCH(ctx);
  return LFR_OK;
}
