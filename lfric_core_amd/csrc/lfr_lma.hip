// LMA operator applies: colour-ordered gather plans and the HIP kernels for
// matrix_vector, dg_matrix_vector, dg_inc_matrix_vector and
// transpose_matrix_vector at invoke granularity.
//
// Data layout (unchanged from the reference, so a PSy caller passes its
// arrays straight through):
//   local_stencil(ncell_3d, ndf_to, ndf_from): entry (ik, i, j) at
//     ik + ncell_3d*(i + ndf_to*j) with ik = cell*nl + k (0-based here).  For
//     a fixed (i, j) the whole operator is contiguous over (cell, k), k
//     fastest, so lanes that walk consecutive layers read consecutive
//     addresses: every matrix load of a wavefront is one coalesced 256-byte
//     (fp32) / 512-byte (fp64) request.
//   fields: layer k of the column that starts at map(df) is map(df) + k.
//
// Summation order matches the reference statement order exactly and every
// term is formed as a separate multiply and add (this file is compiled with
// -ffp-contract=off), so results are bit-identical to the reference loops.
#include <algorithm>
#include <cstdlib>
#include <map>
#include <numeric>
#include <tuple>
#include <type_traits>

#include "lfr_internal.h"
#include "lfr_lma_kernels.h"

namespace {

using lfr::kThreads;
using lfr::k_gather;
using lfr::k_cells;

unsigned pad8(unsigned nb) { return (nb + 7u) & ~7u; }

}  // namespace

// Non-temporal (streaming) loads of the matrix in the one-level-per-thread
// record kernels (measured faster in both precisions on the W2 C96L70 face
// stacks); LFR_GATHER_NT=0 turns them off.
bool lfr_gather_nt(int prec) {
  (void)prec;
  static const bool on = [] {
    const char* e = getenv("LFR_GATHER_NT");
    return !(e && e[0] == '0');
  }();
  return on;
}

// Levels per thread of the vectorised record gather: 4 (FP32) or 2 (FP64) by
// default; LFR_GATHER_LPT=1, 2 or 4 forces one value.
int lfr_gather_lpt(int prec) {
  static const int forced = [] {
    const char* e = getenv("LFR_GATHER_LPT");
    if (e && (e[0] == '1' || e[0] == '2' || e[0] == '4')) return e[0] - '0';
    return 0;
  }();
  return forced ? forced : (prec == 4 ? 4 : 2);
}

// Non-temporal matrix loads in the vectorised and unified gathers: off by
// default (measured slower, packed or not); LFR_GATHER_VEC_NT=1 forces.
bool lfr_gather_vec_nt(bool packed) {
  static const int forced = [] {
    const char* e = getenv("LFR_GATHER_VEC_NT");
    return (e && (e[0] == '0' || e[0] == '1')) ? e[0] - '0' : -1;
  }();
  (void)packed;
  return forced == 1;
}

int lfr_class_lpt(const lfr_plan_class& c, int nl, int prec, int ndf_in, bool packed) {
  if (!(packed || lfr_records_enabled()) || !lfr_record_shape(ndf_in, c.ngroup, c.gsize) || c.height != nl)
    return 1;
  const int lpt = lfr_gather_lpt(prec);
  return nl >= lpt ? lpt : 1;
}

unsigned lfr_class_blocks(const lfr_plan_class& c, int lpt) {
  const int64_t threads = (int64_t)c.ncol * ((c.height + lpt - 1) / lpt);
  return lfr_blocks(threads, lfr::kThreads);
}

// LFR_NO_RECORDS=1 switches the record kernels off (plan-entry kernels only).
bool lfr_records_enabled() {
  static const bool on = [] {
    const char* e = getenv("LFR_NO_RECORDS");
    return !(e && e[0] == '1');
  }();
  return on;
}

namespace {

template <typename T, int NDF_IN, int NG, int GS>
void launch_gather_t(hipStream_t s, int kind, unsigned nb, const T* M, const T* x, T* y,
                     const int32_t* mi, const lfr_plan_class& c, int nl, int64_t n3, int ndf_out,
                     int ndf_in, int overwrite) {
  const unsigned nbp = pad8(nb);
  if (kind == 0)
    hipLaunchKernelGGL((k_gather<T, NDF_IN, NG, GS, 0>), dim3(nbp), dim3(kThreads), 0, s, M, x, y,
                       mi, c.col_start_d, c.ent_d, c.ncol, c.height, c.ngroup, nl, n3, ndf_out,
                       ndf_in, overwrite, nbp, (const T*)nullptr, (double*)nullptr,
                       (const double*)nullptr);
  else
    hipLaunchKernelGGL((k_gather<T, NDF_IN, NG, GS, 1>), dim3(nbp), dim3(kThreads), 0, s, M, x, y,
                       mi, c.col_start_d, c.ent_d, c.ncol, c.height, c.ngroup, nl, n3, ndf_out,
                       ndf_in, overwrite, nbp, (const T*)nullptr, (double*)nullptr,
                       (const double*)nullptr);
}

// Compile-time shapes for the common column structures (groups x entries):
// (2,1) shared side faces, (1,2) bottom/top stacks, (1,1) single-cell
// columns, (4,2)/(3,2) vertex columns (3 at cube corners), (2,2)
// horizontal edges and (4,1)/(3,1) vertical edges at element order 1.
template <typename T, int NDF_IN>
void launch_gather_shape(hipStream_t s, int kind, unsigned nb, const T* M, const T* x, T* y,
                         const int32_t* mi, const lfr_plan_class& c, int nl, int64_t n3,
                         int ndf_out, int ndf_in, int overwrite) {
#define SHAPE(NGV, GSV) \
  if (c.ngroup == NGV && c.gsize == GSV) \
    return launch_gather_t<T, NDF_IN, NGV, GSV>(s, kind, nb, M, x, y, mi, c, nl, n3, ndf_out, ndf_in, overwrite);
  SHAPE(2, 1) SHAPE(1, 2) SHAPE(1, 1) SHAPE(4, 2) SHAPE(3, 2) SHAPE(2, 2) SHAPE(4, 1) SHAPE(3, 1)
#undef SHAPE
  if (c.gsize == 1)
    launch_gather_t<T, NDF_IN, 0, 1>(s, kind, nb, M, x, y, mi, c, nl, n3, ndf_out, ndf_in, overwrite);
  else
    launch_gather_t<T, NDF_IN, 0, 2>(s, kind, nb, M, x, y, mi, c, nl, n3, ndf_out, ndf_in, overwrite);
}

template <typename T>
void launch_gather_any(hipStream_t s, int kind, unsigned nb, const T* M, const T* x, T* y,
                       const int32_t* mi, const lfr_plan_class& c, int nl, int64_t n3, int ndf_out,
                       int ndf_in, int overwrite) {
  switch (ndf_in) {
    case 1: launch_gather_shape<T, 1>(s, kind, nb, M, x, y, mi, c, nl, n3, ndf_out, ndf_in, overwrite); break;
    case 6: launch_gather_shape<T, 6>(s, kind, nb, M, x, y, mi, c, nl, n3, ndf_out, ndf_in, overwrite); break;
    case 8: launch_gather_shape<T, 8>(s, kind, nb, M, x, y, mi, c, nl, n3, ndf_out, ndf_in, overwrite); break;
    default:
      if (c.gsize == 1)
        launch_gather_t<T, 0, 0, 1>(s, kind, nb, M, x, y, mi, c, nl, n3, ndf_out, ndf_in, overwrite);
      else
        launch_gather_t<T, 0, 0, 2>(s, kind, nb, M, x, y, mi, c, nl, n3, ndf_out, ndf_in, overwrite);
  }
}

template <typename T, int NDF_IN, int NG, int GS>
void launch_rec_t(hipStream_t s, int kind, unsigned nb, const T* M, const T* x, T* y,
                  const int32_t* rec, const lfr_plan_class& c, int nl, int64_t n3, int ndf_out,
                  int overwrite, int64_t pcs) {
  const unsigned nbp = pad8(nb);
  const bool nt = lfr_gather_nt((int)sizeof(T));
#define LREC(KD, NTV)                                                                             \
  hipLaunchKernelGGL((lfr::k_gather_rec<T, NDF_IN, NG, GS, KD, false, NTV>), dim3(nbp), dim3(kThreads), \
                     0, s, M, x, y, rec, c.ncol, c.height, nl, n3, ndf_out, overwrite, pcs, nbp,  \
                     (const T*)nullptr, (double*)nullptr, (const double*)nullptr)
  if (kind == 0) {
    if (nt) LREC(0, true); else LREC(0, false);
  } else {
    if (nt) LREC(1, true); else LREC(1, false);
  }
#undef LREC
}

template <typename T, int NDF_IN, int NG, int GS, int LPT>
void launch_vec_t(hipStream_t s, int kind, unsigned nb, const T* M, const T* x, T* y, const int32_t* rec,
                  const lfr_plan_class& c, int nl, int64_t n3, int ndf_out, int overwrite, int64_t pcs) {
  const unsigned nbp = pad8(nb);
  const bool nt = lfr_gather_vec_nt(pcs > 0);
#define LVEC(KD, NTV)                                                                                   \
  hipLaunchKernelGGL((lfr::k_gather_vec<T, NDF_IN, NG, GS, KD, false, LPT, NTV>), dim3(nbp), dim3(kThreads), 0, \
                     s, M, x, y, rec, c.ncol, nl, n3, ndf_out, overwrite, pcs, nbp, (const T*)nullptr,  \
                     (double*)nullptr, (const double*)nullptr)
  if (kind == 0) {
    if (nt) LVEC(0, true); else LVEC(0, false);
  } else {
    if (nt) LVEC(1, true); else LVEC(1, false);
  }
#undef LVEC
}

template <typename T, int NDF_IN, int LPT>
bool launch_vec_shape(hipStream_t s, int kind, unsigned nb, const T* M, const T* x, T* y, const int32_t* rec,
                      const lfr_plan_class& c, int nl, int64_t n3, int ndf_out, int overwrite, int64_t pcs) {
#define SHAPE(NGV, GSV)                                                                               \
  if (c.ngroup == NGV && c.gsize == GSV) {                                                            \
    launch_vec_t<T, NDF_IN, NGV, GSV, LPT>(s, kind, nb, M, x, y, rec, c, nl, n3, ndf_out, overwrite, pcs); \
    return true;                                                                                      \
  }
  SHAPE(2, 1) SHAPE(1, 1) SHAPE(4, 1) SHAPE(3, 1) SHAPE(1, 2) SHAPE(2, 2) SHAPE(3, 2) SHAPE(4, 2)
#undef SHAPE
  return false;
}

template <typename T, int NDF_IN>
bool launch_vec_any(hipStream_t s, int lpt, int kind, unsigned nb, const T* M, const T* x, T* y,
                    const int32_t* rec, const lfr_plan_class& c, int nl, int64_t n3, int ndf_out, int overwrite,
                    int64_t pcs) {
  if (lpt == 2) return launch_vec_shape<T, NDF_IN, 2>(s, kind, nb, M, x, y, rec, c, nl, n3, ndf_out, overwrite, pcs);
  if (lpt == 4) return launch_vec_shape<T, NDF_IN, 4>(s, kind, nb, M, x, y, rec, c, nl, n3, ndf_out, overwrite, pcs);
  return false;
}

template <typename T, int NDF_IN>
void launch_rec_shape(hipStream_t s, int kind, unsigned nb, const T* M, const T* x, T* y,
                      const int32_t* rec, const lfr_plan_class& c, int nl, int64_t n3, int ndf_out,
                      int overwrite, int64_t pcs) {
#define SHAPE(NGV, GSV) \
  if (c.ngroup == NGV && c.gsize == GSV) \
    return launch_rec_t<T, NDF_IN, NGV, GSV>(s, kind, nb, M, x, y, rec, c, nl, n3, ndf_out, overwrite, pcs);
  SHAPE(2, 1) SHAPE(1, 2) SHAPE(1, 1) SHAPE(4, 2) SHAPE(3, 2) SHAPE(2, 2) SHAPE(4, 1) SHAPE(3, 1)
#undef SHAPE
}

template <typename T, int NDF_IN, int KIND, int LPT>
void launch_uni_t(hipStream_t s, bool nt, bool st, unsigned nbp, const T* M, const T* x, T* y, const int32_t* ur,
                  int ncol, int nl, int64_t n3, int ndf_out, int overwrite, int64_t pcs) {
#define LUNI(NTV, STV)                                                                                        \
  hipLaunchKernelGGL((lfr::k_gather_uni<T, NDF_IN, KIND, false, LPT, NTV, STV>), dim3(nbp), dim3(kThreads), 0, s, \
                     M, x, y, ur, ncol, nl, n3, ndf_out, overwrite, pcs, nbp, (const T*)nullptr, (double*)nullptr, \
                     (const double*)nullptr)
  if (nt) {
    if (st) LUNI(true, true); else LUNI(true, false);
  } else {
    if (st) LUNI(false, true); else LUNI(false, false);
  }
#undef LUNI
}

template <typename T, int NDF_IN>
void launch_uni_k(hipStream_t s, int kind, int lpt, bool nt, bool st, unsigned nbp, const T* M, const T* x, T* y,
                  const int32_t* ur, int ncol, int nl, int64_t n3, int ndf_out, int overwrite, int64_t pcs) {
  if (kind == 0) {
    if (lpt == 4) launch_uni_t<T, NDF_IN, 0, 4>(s, nt, st, nbp, M, x, y, ur, ncol, nl, n3, ndf_out, overwrite, pcs);
    else launch_uni_t<T, NDF_IN, 0, 2>(s, nt, st, nbp, M, x, y, ur, ncol, nl, n3, ndf_out, overwrite, pcs);
  } else {
    if (lpt == 4) launch_uni_t<T, NDF_IN, 1, 4>(s, nt, st, nbp, M, x, y, ur, ncol, nl, n3, ndf_out, overwrite, pcs);
    else launch_uni_t<T, NDF_IN, 1, 2>(s, nt, st, nbp, M, x, y, ur, ncol, nl, n3, ndf_out, overwrite, pcs);
  }
}

template <typename T>
int run_gather(lfr_ctx* ctx, const lfr_plan* plan, int kind, const T* M, const T* x, T* y,
               const lfr_space* in, int ndf_out, int64_t n3, int overwrite,
               const lfr_lma_packed* pk = nullptr) {
  const int nl = plan->space->nlayers;
  const int lptu = lfr_unified_lpt((int)sizeof(T));
  if (pk ? pk->unified : (lfr_records_enabled() && lfr_plan_unified(plan, in->ndf, lptu))) {
    const int32_t* ur = lfr_plan_urecords(ctx, plan, in);
    if (!ur) return LFR_ERR_NOMEM;
    const unsigned nbp = pad8(lfr_unified_blocks(plan, lptu));
    const T* Mu = pk ? (const T*)pk->data_d : M;
    const int64_t pcs = pk ? pk->pcs[0] : 0;
    const bool nt = pk ? lfr_pk_nt(pk) : lfr_gather_vec_nt(false);
    const bool st = pk ? lfr_pk_stage(pk) : lfr_gather_stage();
    const int ncol = lfr_plan_ncolumns_total(plan);
    if (in->ndf == 6)
      launch_uni_k<T, 6>(ctx->stream, kind, lptu, nt, st, nbp, Mu, x, y, ur, ncol, nl, n3, ndf_out, overwrite, pcs);
    else
      launch_uni_k<T, 1>(ctx->stream, kind, lptu, nt, st, nbp, Mu, x, y, ur, ncol, nl, n3, ndf_out, overwrite, pcs);
    LFR_CHECK_LAUNCH(ctx);
    return LFR_OK;
  }
  for (size_t ci = 0; ci < plan->classes.size(); ++ci) {
    const auto& c = plan->classes[ci];
    if (c.ncol == 0) continue;
    const unsigned nb = lfr_blocks((int64_t)c.ncol * c.height, kThreads);
    int stride = 0;
    const int32_t* rec = (pk || lfr_records_enabled()) ? lfr_plan_records(ctx, plan, ci, in, &stride) : nullptr;
    // packed operator: this class's block, columns pcs elements apart
    const T* Mc = pk ? (const T*)pk->data_d + pk->off[ci] : M;
    const int64_t pcs = pk ? pk->pcs[ci] : 0;
    if (pk && !rec) return lfr_set_error(ctx, LFR_ERR_NOMEM, "packed apply: column records unavailable");
    const int lpt = rec ? lfr_class_lpt(c, nl, (int)sizeof(T), in->ndf, pk != nullptr) : 1;
    if (rec && lpt > 1) {
      const unsigned nbv = lfr_class_blocks(c, lpt);
      bool ok = false;
      switch (in->ndf) {
        case 1: ok = launch_vec_any<T, 1>(ctx->stream, lpt, kind, nbv, Mc, x, y, rec, c, nl, n3, ndf_out, overwrite, pcs); break;
        case 6: ok = launch_vec_any<T, 6>(ctx->stream, lpt, kind, nbv, Mc, x, y, rec, c, nl, n3, ndf_out, overwrite, pcs); break;
        case 8: ok = launch_vec_any<T, 8>(ctx->stream, lpt, kind, nbv, Mc, x, y, rec, c, nl, n3, ndf_out, overwrite, pcs); break;
      }
      if (!ok) return lfr_set_error(ctx, LFR_ERR_ARG, "no vectorised gather for this class shape");
      LFR_CHECK_LAUNCH(ctx);
      continue;
    }
    if (rec) {
      switch (in->ndf) {
        case 1: launch_rec_shape<T, 1>(ctx->stream, kind, nb, Mc, x, y, rec, c, nl, n3, ndf_out, overwrite, pcs); break;
        case 6: launch_rec_shape<T, 6>(ctx->stream, kind, nb, Mc, x, y, rec, c, nl, n3, ndf_out, overwrite, pcs); break;
        case 8: launch_rec_shape<T, 8>(ctx->stream, kind, nb, Mc, x, y, rec, c, nl, n3, ndf_out, overwrite, pcs); break;
      }
      LFR_CHECK_LAUNCH(ctx);
      continue;
    }
    if (c.gsize != 1 && c.gsize != 2)
      return lfr_set_error(ctx, LFR_ERR_ARG, "plan class with %d entries per cell unsupported",
                           c.gsize);
    launch_gather_any<T>(ctx->stream, kind, nb, M, x, y, in->map0_d, c, nl, n3, ndf_out, in->ndf,
                         overwrite);
    LFR_CHECK_LAUNCH(ctx);
  }
  return LFR_OK;
}

// Levels per thread of the cell-centric apply: 1 (C96L70 div W3 <- W2, FP32:
// 31.7 µs at 1, 56.9 µs at 2, 60.9 µs at 4, the strided levels splitting
// each of the six matrix-plane loads into short runs); LFR_CELLS_LPT=2 or 4
// forces another value.
int cells_lpt(int) {
  static const int forced = [] {
    const char* e = getenv("LFR_CELLS_LPT");
    if (e && (e[0] == '1' || e[0] == '2' || e[0] == '4')) return e[0] - '0';
    return 0;
  }();
  return forced ? forced : 1;
}

template <typename T, int NO, int NI>
void launch_cells_t(hipStream_t s, const T* M, const T* x, T* y, const lfr_space* out,
                    const lfr_space* in, int64_t n3, int inc) {
  const int nl = out->nlayers;
  // the runtime-ndf kernel (NO = NI = 0) runs one level per thread
  const int lpt = (NO > 0 && NI > 0) ? cells_lpt((int)sizeof(T)) : 1;
  auto go = [&](auto L) {
    constexpr int LPT = decltype(L)::value;
    const unsigned nb = pad8(lfr_blocks((int64_t)out->ncell * ((nl + LPT - 1) / LPT), kThreads));
    hipLaunchKernelGGL((k_cells<T, NO, NI, LPT>), dim3(nb), dim3(kThreads), 0, s, M, x, y, out->map0_d,
                       in->map0_d, out->ncell, nl, n3, out->ndf, in->ndf, inc, nb);
  };
  if (lpt == 4) go(std::integral_constant<int, 4>{});
  else if (lpt == 2) go(std::integral_constant<int, 2>{});
  else go(std::integral_constant<int, 1>{});
}

template <typename T>
int run_cells(lfr_ctx* ctx, const T* M, const T* x, T* y, const lfr_space* out,
              const lfr_space* in, int64_t n3, int inc) {
  const int no = out->ndf, ni = in->ndf;
  hipStream_t s = ctx->stream;
  if (no == 1 && ni == 1) launch_cells_t<T, 1, 1>(s, M, x, y, out, in, n3, inc);
  else if (no == 1 && ni == 6) launch_cells_t<T, 1, 6>(s, M, x, y, out, in, n3, inc);
  else if (no == 1 && ni == 8) launch_cells_t<T, 1, 8>(s, M, x, y, out, in, n3, inc);
  else if (no == 8 && ni == 8) launch_cells_t<T, 8, 8>(s, M, x, y, out, in, n3, inc);
  else if (no == 8 && ni == 36) launch_cells_t<T, 8, 36>(s, M, x, y, out, in, n3, inc);
  else if (no == 6 && ni == 6) launch_cells_t<T, 6, 6>(s, M, x, y, out, in, n3, inc);
  else if (no == 6 && ni == 1) launch_cells_t<T, 6, 1>(s, M, x, y, out, in, n3, inc);
  else launch_cells_t<T, 0, 0>(s, M, x, y, out, in, n3, inc);
  LFR_CHECK_LAUNCH(ctx);
  return LFR_OK;
}

int check_pair(lfr_ctx* ctx, int prec, const void* lhs, const void* x, const void* M,
               const lfr_space* a, const lfr_space* b, const char* who) {
  LFR_REQUIRE(ctx, prec == 4 || prec == 8, "%s: prec must be 4 or 8, got %d", who, prec);
  LFR_REQUIRE(ctx, lhs && x && M && a && b, "%s: null argument", who);
  LFR_REQUIRE(ctx, a->ncell == b->ncell && a->nlayers == b->nlayers,
              "%s: spaces on different meshes (ncell %d/%d, nlayers %d/%d)", who, a->ncell,
              b->ncell, a->nlayers, b->nlayers);
  return LFR_OK;
}

}  // namespace

// ---------------------------------------------------------------------------
// Plan construction (host)
// ---------------------------------------------------------------------------
extern "C" int lfr_plan_create(lfr_ctx* ctx, const lfr_space* sp, const int32_t* cic,
                               const int32_t* npc, int ncolour, int maxc, int kind,
                               lfr_plan** out) {
  if (!ctx || !out) return LFR_ERR_ARG;
  *out = nullptr;
  LFR_REQUIRE(ctx, sp && cic && npc && ncolour > 0 && maxc > 0 && (kind == 0 || kind == 1),
              "lfr_plan_create: bad arguments");
  const int ncell = sp->ncell, ndf = sp->ndf, nl = sp->nlayers;
  // colour position of every cell, in the coloured PSy loop order
  std::vector<int> pos(ncell, -1);
  int p = 0;
  for (int col = 0; col < ncolour; ++col) {
    LFR_REQUIRE(ctx, npc[col] >= 0 && npc[col] <= maxc, "lfr_plan_create: ncells_per_colour(%d)=%d",
                col + 1, npc[col]);
    for (int idx = 0; idx < npc[col]; ++idx) {
      const int cell = cic[col + (int64_t)ncolour * idx];
      LFR_REQUIRE(ctx, cell >= 1 && cell <= ncell, "lfr_plan_create: cell %d out of range", cell);
      LFR_REQUIRE(ctx, pos[cell - 1] < 0, "lfr_plan_create: cell %d coloured twice", cell);
      pos[cell - 1] = p++;
    }
  }
  LFR_REQUIRE(ctx, p == ncell, "lfr_plan_create: colouring covers %d of %d cells", p, ncell);

  // Column ranges [start, start+nl) touched by every (cell, df); overlapping
  // ranges (bottom/top copies of one entity) merge into one column.
  struct R { int32_t start; int32_t cell; int32_t df; };
  std::vector<R> rs((size_t)ncell * ndf);
  for (int c = 0; c < ncell; ++c)
    for (int i = 0; i < ndf; ++i) rs[(size_t)c * ndf + i] = {sp->map_h[(size_t)c * ndf + i] - 1, c, i};
  std::sort(rs.begin(), rs.end(), [](const R& a, const R& b) {
    return std::tie(a.start, a.cell, a.df) < std::tie(b.start, b.cell, b.df);
  });

  struct Col { int32_t start, height; size_t first, count; };
  std::vector<Col> cols;
  int32_t expect = 0;
  for (size_t q = 0; q < rs.size();) {
    const int32_t cs = rs[q].start;
    LFR_REQUIRE(ctx, cs == expect,
                "lfr_plan_create: dofs %d..%d are not touched by any cell (dofmap not column-contiguous)",
                expect + 1, cs);
    int32_t ce = cs + nl;
    size_t r = q + 1;
    while (r < rs.size() && rs[r].start < ce) {
      ce = std::max(ce, rs[r].start + nl);
      ++r;
    }
    cols.push_back({cs, ce - cs, q, r - q});
    expect = ce;
    q = r;
  }
  LFR_REQUIRE(ctx, expect == sp->undf, "lfr_plan_create: columns cover %d dofs, undf=%d", expect,
              sp->undf);

  // Order entries inside each column and group them by cell.
  struct E { int pos, cell, df, shift; };
  using Sig = std::tuple<int, int, int>;  // height, ngroup, gsize
  std::map<Sig, std::vector<int>> by_sig;  // signature -> column indices
  std::vector<std::vector<E>> col_ent(cols.size());
  for (size_t ci = 0; ci < cols.size(); ++ci) {
    auto& v = col_ent[ci];
    for (size_t q = cols[ci].first; q < cols[ci].first + cols[ci].count; ++q) {
      const R& r = rs[q];
      const int s = r.start - cols[ci].start;
      LFR_REQUIRE(ctx, s == 0 || s == 1,
                  "lfr_plan_create: local dof %d of cell %d starts %d dofs into its column",
                  r.df + 1, r.cell + 1, s);
      v.push_back({pos[r.cell], r.cell, r.df, s});
    }
    // kind 0: df order inside a cell (the kernel's inner df loop);
    // kind 1: layer order inside a cell (k outer loop): larger shift first.
    std::sort(v.begin(), v.end(), [kind](const E& a, const E& b) {
      if (kind == 0) return std::tie(a.pos, a.df) < std::tie(b.pos, b.df);
      return std::make_tuple(a.pos, -a.shift, a.df) < std::make_tuple(b.pos, -b.shift, b.df);
    });
    int ngroup = 0, gmax = 0;
    for (size_t q = 0; q < v.size();) {
      size_t r = q;
      while (r < v.size() && v[r].cell == v[q].cell) ++r;
      // two entries of one cell in one column must differ in shift
      for (size_t a = q; a < r; ++a)
        for (size_t b = a + 1; b < r; ++b)
          LFR_REQUIRE(ctx, v[a].shift != v[b].shift,
                      "lfr_plan_create: cell %d touches a dof twice through dofs %d and %d",
                      v[a].cell + 1, v[a].df + 1, v[b].df + 1);
      ++ngroup;
      gmax = std::max<int>(gmax, (int)(r - q));
      q = r;
    }
    LFR_REQUIRE(ctx, gmax <= 2, "lfr_plan_create: %d entries of one cell in a column", gmax);
    by_sig[Sig(cols[ci].height, ngroup, gmax)].push_back((int)ci);
  }

  lfr_plan* plan = new lfr_plan();
  plan->uid = lfr_next_uid();
  plan->ctx = ctx;
  plan->space = sp;
  plan->kind = kind;
  plan->ncolumns = (int)cols.size();
  for (auto& kv : by_sig) {
    lfr_plan_class pc;
    std::tie(pc.height, pc.ngroup, pc.gsize) = kv.first;
    pc.ncol = (int)kv.second.size();
    const int per = pc.ngroup * pc.gsize * 2;
    std::vector<int32_t> starts(pc.ncol), ent((size_t)pc.ncol * per, -1);
    for (int q = 0; q < pc.ncol; ++q) {
      const int ci = kv.second[q];
      starts[q] = cols[ci].start;
      const auto& v = col_ent[ci];
      int g = -1, slot = 0, last_cell = -1;
      for (const E& e : v) {
        if (e.cell != last_cell) {
          ++g;
          slot = 0;
          last_cell = e.cell;
        }
        int32_t* d = &ent[(size_t)q * per + (size_t)(g * pc.gsize + slot) * 2];
        d[0] = e.cell;
        d[1] = (e.df << 1) | e.shift;
        ++slot;
      }
      // padded slots keep the group's cell id and a negative code
      for (int gg = 0; gg < pc.ngroup; ++gg)
        for (int s = 0; s < pc.gsize; ++s) {
          int32_t* d = &ent[(size_t)q * per + (size_t)(gg * pc.gsize + s) * 2];
          if (d[1] < 0) d[0] = ent[(size_t)q * per + (size_t)(gg * pc.gsize) * 2];
        }
    }
    pc.starts_h = starts;
    pc.ent_h = ent;
    const size_t b1 = starts.size() * sizeof(int32_t), b2 = ent.size() * sizeof(int32_t);
    if (hipMalloc(&pc.col_start_d, b1) != hipSuccess || hipMalloc(&pc.ent_d, b2) != hipSuccess) {
      lfr_plan_destroy(plan);
      return lfr_set_error(ctx, LFR_ERR_NOMEM, "plan allocation failed");
    }
    if (hipMemcpy(pc.col_start_d, starts.data(), b1, hipMemcpyHostToDevice) != hipSuccess ||
        hipMemcpy(pc.ent_d, ent.data(), b2, hipMemcpyHostToDevice) != hipSuccess) {
      plan->classes.push_back(pc);
      lfr_plan_destroy(plan);
      return lfr_set_error(ctx, LFR_ERR_HIP, "plan upload failed");
    }
    plan->nbytes += (int64_t)(b1 + b2);
    plan->classes.push_back(pc);
  }
  *out = plan;
  return LFR_OK;
}

bool lfr_record_shape(int ndf_in, int ng, int gs) {
  if (ndf_in != 1 && ndf_in != 6 && ndf_in != 8) return false;
  return (ng == 2 && gs == 1) || (ng == 1 && gs == 2) || (ng == 1 && gs == 1) ||
         (ng == 4 && gs == 2) || (ng == 3 && gs == 2) || (ng == 2 && gs == 2) ||
         (ng == 4 && gs == 1) || (ng == 3 && gs == 1);
}

const int32_t* lfr_plan_records(lfr_ctx* ctx, const lfr_plan* plan, size_t ci, const lfr_space* in,
                                int* stride) {
  const lfr_plan_class& c = plan->classes[ci];
  const int ndf_in = in->ndf;
  if (!lfr_record_shape(ndf_in, c.ngroup, c.gsize)) return nullptr;
  const int group = 1 + c.gsize + ndf_in;
  const int st = (1 + c.ngroup * group + 3) & ~3;
  *stride = st;
  for (auto& kv : plan->records)
    if (kv.first == in->uid) return kv.second[ci];
  // build every class's records for this input space
  std::vector<int32_t*> recs(plan->classes.size(), nullptr);
  const int nl = plan->space->nlayers;
  for (size_t k = 0; k < plan->classes.size(); ++k) {
    const lfr_plan_class& pc = plan->classes[k];
    if (!lfr_record_shape(ndf_in, pc.ngroup, pc.gsize)) continue;
    const int grp = 1 + pc.gsize + ndf_in;
    const int str = (1 + pc.ngroup * grp + 3) & ~3;
    const int per = pc.ngroup * pc.gsize * 2;
    std::vector<int32_t> rec((size_t)pc.ncol * str, 0);
    for (int q = 0; q < pc.ncol; ++q) {
      int32_t* r = &rec[(size_t)q * str];
      r[0] = pc.starts_h[q];
      for (int g = 0; g < pc.ngroup; ++g) {
        const int32_t* e = &pc.ent_h[(size_t)q * per + (size_t)g * pc.gsize * 2];
        const int cell = e[0];
        int32_t* rg = r + 1 + g * grp;
        rg[0] = cell * nl;
        for (int s = 0; s < pc.gsize; ++s) rg[1 + s] = e[2 * s + 1];
        for (int j = 0; j < ndf_in; ++j) rg[1 + pc.gsize + j] = in->map_h[(size_t)cell * ndf_in + j] - 1;
      }
    }
    if (hipMalloc(&recs[k], rec.size() * sizeof(int32_t)) != hipSuccess ||
        hipMemcpy(recs[k], rec.data(), rec.size() * sizeof(int32_t), hipMemcpyHostToDevice) != hipSuccess) {
      for (int32_t* p : recs)
        if (p) (void)hipFree(p);
      lfr_set_error(ctx, LFR_ERR_NOMEM, "column record allocation failed");
      return nullptr;
    }
  }
  plan->records.emplace_back(in->uid, recs);
  return recs[ci];
}

// ---------------------------------------------------------------------------
// Unified memory-order gather (plans whose columns all have two entries)
// ---------------------------------------------------------------------------
// LFR_GATHER_UNIFIED=0 switches it off (per-class launches instead).
bool lfr_plan_unified(const lfr_plan* plan, int ndf_in, int lpt) {
  static const bool on = [] {
    const char* e = getenv("LFR_GATHER_UNIFIED");
    return !(e && e[0] == '0');
  }();
  if (!on || (ndf_in != 1 && ndf_in != 6)) return false;
  const int nl = plan->space->nlayers;
  if (nl < lpt) return false;
  bool any = false;
  for (const auto& c : plan->classes) {
    if (c.ncol == 0) continue;
    if (!((c.ngroup == 2 && c.gsize == 1) || (c.ngroup == 1 && c.gsize == 2))) return false;
    if (c.height != nl && c.height != nl + 1) return false;
    any = true;
  }
  return any;
}

// Levels per thread of the unified gather (2 or 4; 1 is not offered).
int lfr_unified_lpt(int prec) {
  const int l = lfr_gather_lpt(prec);
  return l >= 2 ? l : 2;
}

// The unified gather writes its output through LDS with aligned stores
// (measured faster); LFR_GATHER_STAGE=0 stores from registers instead.
bool lfr_gather_stage() {
  static const bool on = [] {
    const char* e = getenv("LFR_GATHER_STAGE");
    return !(e && e[0] == '0');
  }();
  return on;
}

bool lfr_pk_stage(const lfr_lma_packed* pk) { return pk && pk->stage >= 0 ? pk->stage != 0 : lfr_gather_stage(); }
bool lfr_pk_nt(const lfr_lma_packed* pk) { return pk && pk->nt >= 0 ? pk->nt != 0 : lfr_gather_vec_nt(true); }

int lfr_plan_ncolumns_total(const lfr_plan* plan) {
  int n = 0;
  for (const auto& c : plan->classes) n += c.ncol;
  return n;
}

unsigned lfr_unified_blocks(const lfr_plan* plan, int lpt) {
  const int tpc = (plan->space->nlayers + lpt) / lpt;
  return lfr_blocks((int64_t)lfr_plan_ncolumns_total(plan) * tpc, lfr::kThreads);
}

namespace {
// (start, class, column-in-class) of every column, in memory order
std::vector<std::tuple<int32_t, int, int>> memory_order(const lfr_plan* plan) {
  std::vector<std::tuple<int32_t, int, int>> v;
  for (size_t k = 0; k < plan->classes.size(); ++k)
    for (int q = 0; q < plan->classes[k].ncol; ++q) v.emplace_back(plan->classes[k].starts_h[q], (int)k, q);
  std::sort(v.begin(), v.end());
  return v;
}
}  // namespace

void lfr_plan_uentries(const lfr_plan* plan, std::vector<int32_t>& ent) {
  const auto order = memory_order(plan);
  ent.assign(order.size() * 4, 0);
  for (size_t i = 0; i < order.size(); ++i) {
    const auto& pc = plan->classes[std::get<1>(order[i])];
    const int32_t* e = &pc.ent_h[(size_t)std::get<2>(order[i]) * 4];  // two (cell, code) pairs
    for (int q = 0; q < 4; ++q) ent[i * 4 + q] = e[q];
  }
}

const int32_t* lfr_plan_urecords(lfr_ctx* ctx, const lfr_plan* plan, const lfr_space* in) {
  for (auto& kv : plan->urecords)
    if (kv.first == in->uid) return kv.second;
  const int ndf_in = in->ndf;
  const int kEntry = 2 + ndf_in;
  const int stride = (3 + 2 * kEntry + 3) & ~3;
  const int nl = plan->space->nlayers;
  const auto order = memory_order(plan);
  std::vector<int32_t> rec(order.size() * stride, 0);
  for (size_t i = 0; i < order.size(); ++i) {
    const auto& pc = plan->classes[std::get<1>(order[i])];
    const int q = std::get<2>(order[i]);
    int32_t* r = &rec[i * stride];
    r[0] = pc.starts_h[q];
    r[1] = pc.height;
    r[2] = pc.gsize == 2 ? 1 : 0;  // one cell, two local dofs: input-dof-outer order (kind 0)
    const int32_t* e = &pc.ent_h[(size_t)q * 4];
    for (int k = 0; k < 2; ++k) {
      const int cell = e[2 * k];  // padded slots carry a valid cell id
      int32_t* re = r + 3 + k * kEntry;
      re[0] = cell * nl;
      re[1] = e[2 * k + 1];
      for (int j = 0; j < ndf_in; ++j) re[2 + j] = in->map_h[(size_t)cell * ndf_in + j] - 1;
    }
  }
  int32_t* d = nullptr;
  if (hipMalloc(&d, std::max<size_t>(rec.size(), 4) * sizeof(int32_t)) != hipSuccess ||
      hipMemcpy(d, rec.data(), rec.size() * sizeof(int32_t), hipMemcpyHostToDevice) != hipSuccess) {
    if (d) (void)hipFree(d);
    (void)hipGetLastError();
    lfr_set_error(ctx, LFR_ERR_NOMEM, "unified column record allocation failed");
    return nullptr;
  }
  plan->urecords.emplace_back(in->uid, d);
  return d;
}

extern "C" int lfr_plan_destroy(lfr_plan* plan) {
  if (!plan) return LFR_OK;
  lfr_graphs_evict(plan->ctx, plan->uid);
  for (auto& kv : plan->records)
    for (int32_t* p : kv.second)
      if (p) (void)hipFree(p);
  for (auto& kv : plan->urecords)
    if (kv.second) (void)hipFree(kv.second);
  for (auto& c : plan->classes) {
    if (c.col_start_d) (void)hipFree(c.col_start_d);
    if (c.ent_d) (void)hipFree(c.ent_d);
  }
  delete plan;
  return LFR_OK;
}

extern "C" int lfr_plan_info(const lfr_plan* plan, int* ncolumns, int* nclasses, int64_t* nbytes) {
  if (!plan) return LFR_ERR_ARG;
  if (ncolumns) *ncolumns = plan->ncolumns;
  if (nclasses) *nclasses = (int)plan->classes.size();
  if (nbytes) *nbytes = plan->nbytes;
  return LFR_OK;
}

// ---------------------------------------------------------------------------
// Entry points
// ---------------------------------------------------------------------------
extern "C" int lfr_matrix_vector(lfr_ctx* ctx, int prec, void* lhs, const void* x, const void* M,
                                 const lfr_space* to, const lfr_space* from, const lfr_plan* plan,
                                 int overwrite) {
  if (!ctx) return LFR_ERR_ARG;
  int rc = check_pair(ctx, prec, lhs, x, M, to, from, "lfr_matrix_vector");
  if (rc) return rc;
  LFR_REQUIRE(ctx, lhs != x, "lfr_matrix_vector: lhs must not alias x");
  LFR_REQUIRE(ctx, plan && plan->space == to && plan->kind == 0,
              "lfr_matrix_vector: plan must be a matrix_vector (kind 0) plan of the output space");
  const int64_t n3 = (int64_t)to->ncell * to->nlayers;
  if (prec == 4)
    return run_gather<float>(ctx, plan, 0, (const float*)M, (const float*)x, (float*)lhs, from,
                             to->ndf, n3, overwrite);
  return run_gather<double>(ctx, plan, 0, (const double*)M, (const double*)x, (double*)lhs, from,
                            to->ndf, n3, overwrite);
}

extern "C" int lfr_transpose_matrix_vector(lfr_ctx* ctx, int prec, void* lhs, const void* x,
                                           const void* M, const lfr_space* lsp,
                                           const lfr_space* xsp, const lfr_plan* plan,
                                           int overwrite) {
  if (!ctx) return LFR_ERR_ARG;
  int rc = check_pair(ctx, prec, lhs, x, M, lsp, xsp, "lfr_transpose_matrix_vector");
  if (rc) return rc;
  LFR_REQUIRE(ctx, lhs != x, "lfr_transpose_matrix_vector: lhs must not alias x");
  LFR_REQUIRE(ctx, plan && plan->space == lsp && plan->kind == 1,
              "lfr_transpose_matrix_vector: plan must be a transpose (kind 1) plan of lhs' space");
  const int64_t n3 = (int64_t)lsp->ncell * lsp->nlayers;
  if (prec == 4)
    return run_gather<float>(ctx, plan, 1, (const float*)M, (const float*)x, (float*)lhs, xsp,
                             lsp->ndf, n3, overwrite);
  return run_gather<double>(ctx, plan, 1, (const double*)M, (const double*)x, (double*)lhs, xsp,
                            lsp->ndf, n3, overwrite);
}

// ---------------------------------------------------------------------------
// Packed operator
// ---------------------------------------------------------------------------
namespace {

// P[col][e][j][L] = the matrix value entry e of column col multiplies with
// input dof j at output level L (the same index arithmetic as the gather
// kernels on the reference layout), 0 where the entry has no layer there.
template <typename T>
__global__ void __launch_bounds__(kThreads)
k_lma_pack(const T* __restrict__ M, T* __restrict__ P, const int32_t* __restrict__ ent, int ne, int ndf_in,
           int H, int nl, int64_t n3, int ndf_out, int kind, int64_t total) {
  for (int64_t idx = (int64_t)blockIdx.x * kThreads + threadIdx.x; idx < total;
       idx += (int64_t)gridDim.x * kThreads) {
    const int L = (int)(idx % H);
    int64_t r = idx / H;
    const int j = (int)(r % ndf_in);
    r /= ndf_in;  // col * ne + e
    const int cell = ent[2 * r];
    const int code = ent[2 * r + 1];
    const int k = L - (code & 1);
    const bool ok = code >= 0 && k >= 0 && k < nl;
    const int ii = code >= 0 ? (code >> 1) : 0;
    const int64_t mo = kind == 0 ? (int64_t)ii + (int64_t)ndf_out * j : (int64_t)j + (int64_t)ndf_in * ii;
    P[idx] = ok ? M[(int64_t)cell * nl + k + n3 * mo] : T(0);
  }
}

}  // namespace

extern "C" int lfr_lma_pack(lfr_ctx* ctx, int prec, const void* M, const lfr_space* in, const lfr_plan* plan,
                            lfr_lma_packed** out) {
  if (!ctx || !out) return LFR_ERR_ARG;
  *out = nullptr;
  LFR_REQUIRE(ctx, prec == 4 || prec == 8, "lfr_lma_pack: prec must be 4 or 8, got %d", prec);
  LFR_REQUIRE(ctx, M && in && plan, "lfr_lma_pack: null argument");
  const lfr_space* sp = plan->space;
  LFR_REQUIRE(ctx, sp->ncell == in->ncell && sp->nlayers == in->nlayers,
              "lfr_lma_pack: spaces on different meshes");
  LFR_REQUIRE(ctx, in->ndf == 1 || in->ndf == 6 || in->ndf == 8,
              "lfr_lma_pack: no packed kernel for %d input dofs per cell", in->ndf);
  for (const auto& c : plan->classes)
    LFR_REQUIRE(ctx, c.ncol == 0 || lfr_record_shape(in->ndf, c.ngroup, c.gsize),
                "lfr_lma_pack: plan class (%d cells x %d dofs) has no packed kernel", c.ngroup, c.gsize);
  auto* pk = new lfr_lma_packed();
  pk->uid = lfr_next_uid();
  pk->ctx = ctx;
  pk->prec = prec;
  pk->kind = plan->kind;
  pk->plan = plan;
  pk->in = in;
  pk->ndf_out = sp->ndf;
  pk->unified = lfr_plan_unified(plan, in->ndf, lfr_unified_lpt(prec));
  int64_t tot = 0;
  if (pk->unified) {
    pk->off.push_back(0);
    pk->pcs.push_back((int64_t)2 * in->ndf * (sp->nlayers + 1));
    tot = ((int64_t)lfr_plan_ncolumns_total(plan) * pk->pcs[0] + 63) & ~(int64_t)63;
  } else {
    for (const auto& c : plan->classes) {
      const int64_t pcs = (int64_t)c.ngroup * c.gsize * in->ndf * c.height;
      pk->off.push_back(tot);
      pk->pcs.push_back(pcs);
      tot += ((int64_t)c.ncol * pcs + 63) & ~(int64_t)63;  // 256-byte aligned blocks
    }
  }
  pk->nbytes = tot * prec;
  if (hipMalloc(&pk->data_d, (size_t)std::max<int64_t>(pk->nbytes, 256)) != hipSuccess) {
    (void)hipGetLastError();
    delete pk;
    return lfr_set_error(ctx, LFR_ERR_NOMEM, "lfr_lma_pack: cannot allocate %lld bytes", (long long)tot * prec);
  }
  const int nl = sp->nlayers;
  const int64_t n3 = (int64_t)sp->ncell * nl;
  if (pk->unified) {
    // entries of every column in memory order, rows of nl+1 levels
    std::vector<int32_t> ent;
    lfr_plan_uentries(plan, ent);
    int32_t* ent_d = nullptr;
    const int64_t total = (int64_t)lfr_plan_ncolumns_total(plan) * pk->pcs[0];
    hipError_t e = hipMalloc(&ent_d, ent.size() * sizeof(int32_t));
    if (e == hipSuccess) e = hipMemcpy(ent_d, ent.data(), ent.size() * sizeof(int32_t), hipMemcpyHostToDevice);
    if (e == hipSuccess) {
      const unsigned nb = lfr_blocks(total, kThreads, 65536);
      if (prec == 4)
        hipLaunchKernelGGL(k_lma_pack<float>, dim3(nb), dim3(kThreads), 0, ctx->stream, (const float*)M,
                           (float*)pk->data_d, ent_d, 2, in->ndf, nl + 1, nl, n3, sp->ndf, plan->kind, total);
      else
        hipLaunchKernelGGL(k_lma_pack<double>, dim3(nb), dim3(kThreads), 0, ctx->stream, (const double*)M,
                           (double*)pk->data_d, ent_d, 2, in->ndf, nl + 1, nl, n3, sp->ndf, plan->kind, total);
      e = hipGetLastError();
      if (e == hipSuccess) e = hipStreamSynchronize(ctx->stream);
    }
    if (ent_d) (void)hipFree(ent_d);
    if (e != hipSuccess) {
      (void)hipFree(pk->data_d);
      delete pk;
      return lfr_set_error(ctx, LFR_ERR_HIP, "lfr_lma_pack: %s", hipGetErrorString(e));
    }
  }
  for (size_t ci = 0; !pk->unified && ci < plan->classes.size(); ++ci) {
    const auto& c = plan->classes[ci];
    const int64_t total = (int64_t)c.ncol * pk->pcs[ci];
    if (total == 0) continue;
    const unsigned nb = lfr_blocks(total, kThreads, 65536);
    if (prec == 4)
      hipLaunchKernelGGL(k_lma_pack<float>, dim3(nb), dim3(kThreads), 0, ctx->stream, (const float*)M,
                         (float*)pk->data_d + pk->off[ci], c.ent_d, c.ngroup * c.gsize, in->ndf, c.height, nl,
                         n3, sp->ndf, plan->kind, total);
    else
      hipLaunchKernelGGL(k_lma_pack<double>, dim3(nb), dim3(kThreads), 0, ctx->stream, (const double*)M,
                         (double*)pk->data_d + pk->off[ci], c.ent_d, c.ngroup * c.gsize, in->ndf, c.height, nl,
                         n3, sp->ndf, plan->kind, total);
    const hipError_t e = hipGetLastError();
    if (e != hipSuccess) {
      (void)hipFree(pk->data_d);
      delete pk;
      return lfr_set_error(ctx, LFR_ERR_HIP, "lfr_lma_pack: launch failed: %s", hipGetErrorString(e));
    }
  }
  // build the column records now so the first apply does no host work
  if (pk->unified && !lfr_plan_urecords(ctx, plan, in)) {
    (void)hipFree(pk->data_d);
    delete pk;
    return LFR_ERR_NOMEM;
  }
  for (size_t ci = 0; !pk->unified && ci < plan->classes.size(); ++ci) {
    int stride = 0;
    if (plan->classes[ci].ncol && !lfr_plan_records(ctx, plan, ci, in, &stride)) {
      (void)hipFree(pk->data_d);
      delete pk;
      return LFR_ERR_NOMEM;  // message set by lfr_plan_records
    }
  }
  // autotuner: time the four unified-gather variants (output through LDS or
  // not, non-temporal matrix loads or not; identical results) on scratch
  // vectors and keep the fastest on the packed operator, unless the
  // variant is forced through LFR_GATHER_STAGE / LFR_GATHER_VEC_NT
  if (pk->unified && lfr_tune_enabled() && !getenv("LFR_GATHER_STAGE") && !getenv("LFR_GATHER_VEC_NT")) {
    void* xs = nullptr;
    void* ys = nullptr;
    if (hipMalloc(&xs, (size_t)in->undf * prec) == hipSuccess && hipMalloc(&ys, (size_t)sp->undf * prec) == hipSuccess &&
        hipMemsetAsync(xs, 0, (size_t)in->undf * prec, ctx->stream) == hipSuccess) {
      char key[128];
      snprintf(key, sizeof key, "packed gather prec=%d ndf=%d kind=%d ncol=%d nl=%d", prec, in->ndf, pk->kind,
               lfr_plan_ncolumns_total(plan), nl);
      const int dflt = (lfr_gather_stage() ? 1 : 0) | (lfr_gather_vec_nt(true) ? 2 : 0);
      const auto run = [&](int c) -> int {
        pk->stage = c & 1;
        pk->nt = (c >> 1) & 1;
        return prec == 4 ? run_gather<float>(ctx, plan, pk->kind, nullptr, (const float*)xs, (float*)ys, in, sp->ndf,
                                              n3, 1, pk)
                         : run_gather<double>(ctx, plan, pk->kind, nullptr, (const double*)xs, (double*)ys, in,
                                              sp->ndf, n3, 1, pk);
      };
      const int c = lfr_tune_pick(ctx, key, 4, dflt, run);
      pk->stage = c & 1;
      pk->nt = (c >> 1) & 1;
      (void)hipStreamSynchronize(ctx->stream);
    }
    (void)hipGetLastError();
    if (xs) (void)hipFree(xs);
    if (ys) (void)hipFree(ys);
  }
  // matrix_d may be released as soon as this returns
  const hipError_t e = hipStreamSynchronize(ctx->stream);
  if (e != hipSuccess) {
    (void)hipFree(pk->data_d);
    delete pk;
    return lfr_set_error(ctx, LFR_ERR_HIP, "lfr_lma_pack: %s", hipGetErrorString(e));
  }
  *out = pk;
  return LFR_OK;
}

extern "C" int lfr_lma_packed_destroy(lfr_lma_packed* pk) {
  if (!pk) return LFR_OK;
  lfr_graphs_evict(pk->ctx, pk->uid);
  if (pk->data_d) (void)hipFree(pk->data_d);
  delete pk;
  return LFR_OK;
}

extern "C" int lfr_lma_packed_info(const lfr_lma_packed* pk, long long* nbytes) {
  if (!pk || !nbytes) return LFR_ERR_ARG;
  *nbytes = (long long)pk->nbytes;
  return LFR_OK;
}

extern "C" int lfr_matrix_vector_packed(lfr_ctx* ctx, const lfr_lma_packed* pk, void* lhs, const void* x,
                                        int overwrite) {
  if (!ctx) return LFR_ERR_ARG;
  LFR_REQUIRE(ctx, pk && lhs && x, "lfr_matrix_vector_packed: null argument");
  LFR_REQUIRE(ctx, pk->ctx == ctx, "lfr_matrix_vector_packed: operator packed on another context");
  LFR_REQUIRE(ctx, lhs != x, "lfr_matrix_vector_packed: lhs must not alias x");
  const lfr_space* sp = pk->plan->space;
  const int64_t n3 = (int64_t)sp->ncell * sp->nlayers;
  if (pk->prec == 4)
    return run_gather<float>(ctx, pk->plan, pk->kind, nullptr, (const float*)x, (float*)lhs, pk->in, sp->ndf, n3,
                             overwrite, pk);
  return run_gather<double>(ctx, pk->plan, pk->kind, nullptr, (const double*)x, (double*)lhs, pk->in, sp->ndf, n3,
                            overwrite, pk);
}

// ---------------------------------------------------------------------------
// Multidata (ndata fields, "ndata first") apply
// ---------------------------------------------------------------------------
namespace {
// fields per thread of the multidata gather: the widest of 16/8/4/1 that
// divides ndata, capped at 4 (measured fastest on the 16-field FP64 W2
// order-1 apply: 13.8 ms against 16.5 with 8 and 23.2 with 16);
// LFR_NDATA_FPT=1/4/8/16 sets another cap
int ndata_fpt(int nd) {
  static const int cap = [] {
    const char* e = getenv("LFR_NDATA_FPT");
    const int v = e ? atoi(e) : 4;
    return (v == 1 || v == 4 || v == 8 || v == 16) ? v : 4;
  }();
  for (int f : {16, 8, 4})
    if (f <= cap && nd % f == 0) return f;
  return 1;
}

template <typename T>
int run_gather_ndata(lfr_ctx* ctx, const lfr_plan* plan, const T* M, const T* x, T* y, const lfr_space* in,
                     int ndf_out, int64_t n3, int nd, int overwrite) {
  const int nl = plan->space->nlayers;
  for (const auto& c : plan->classes) {
    if (c.ncol == 0) continue;
    LFR_REQUIRE(ctx, c.gsize == 1 || c.gsize == 2, "plan class with %d entries per cell unsupported", c.gsize);
    const int fpt = ndata_fpt(nd);
    const unsigned nbp = pad8(lfr_blocks((int64_t)c.ncol * c.height * (nd / fpt), kThreads));
#define LND(GSV, KD, F)                                                                                       \
  hipLaunchKernelGGL((lfr::k_gather_nf<T, GSV, KD, F>), dim3(nbp), dim3(kThreads), 0, ctx->stream, M, x, y,  \
                     in->map0_d, c.col_start_d, c.ent_d, c.ncol, c.height, c.ngroup, nl, n3, ndf_out, in->ndf, nd, \
                     overwrite, nbp)
    if (fpt == 16) {
      if (plan->kind == 0) {
        if (c.gsize == 1) LND(1, 0, 16); else LND(2, 0, 16);
      } else {
        if (c.gsize == 1) LND(1, 1, 16); else LND(2, 1, 16);
      }
    } else if (fpt == 8) {
      if (plan->kind == 0) {
        if (c.gsize == 1) LND(1, 0, 8); else LND(2, 0, 8);
      } else {
        if (c.gsize == 1) LND(1, 1, 8); else LND(2, 1, 8);
      }
    } else if (fpt == 4) {
      if (plan->kind == 0) {
        if (c.gsize == 1) LND(1, 0, 4); else LND(2, 0, 4);
      } else {
        if (c.gsize == 1) LND(1, 1, 4); else LND(2, 1, 4);
      }
    } else {
      if (plan->kind == 0) {
        if (c.gsize == 1) LND(1, 0, 1); else LND(2, 0, 1);
      } else {
        if (c.gsize == 1) LND(1, 1, 1); else LND(2, 1, 1);
      }
    }
#undef LND
    LFR_CHECK_LAUNCH(ctx);
  }
  return LFR_OK;
}
}  // namespace

extern "C" int lfr_matrix_vector_ndata(lfr_ctx* ctx, int prec, void* lhs, const void* x, const void* M,
                                       const lfr_space* lhs_space, const lfr_space* x_space, const lfr_plan* plan,
                                       int ndata, int overwrite) {
  if (!ctx) return LFR_ERR_ARG;
  int rc = check_pair(ctx, prec, lhs, x, M, lhs_space, x_space, "lfr_matrix_vector_ndata");
  if (rc) return rc;
  LFR_REQUIRE(ctx, lhs != x, "lfr_matrix_vector_ndata: lhs must not alias x");
  LFR_REQUIRE(ctx, ndata >= 1 && ndata <= 64, "lfr_matrix_vector_ndata: ndata must be 1..64, got %d", ndata);
  LFR_REQUIRE(ctx, plan && plan->space == lhs_space,
              "lfr_matrix_vector_ndata: plan must be a plan of the output space");
  const int64_t n3 = (int64_t)lhs_space->ncell * lhs_space->nlayers;
  if (prec == 4)
    return run_gather_ndata<float>(ctx, plan, (const float*)M, (const float*)x, (float*)lhs, x_space,
                                   lhs_space->ndf, n3, ndata, overwrite);
  return run_gather_ndata<double>(ctx, plan, (const double*)M, (const double*)x, (double*)lhs, x_space,
                                  lhs_space->ndf, n3, ndata, overwrite);
}

static int dg_common(lfr_ctx* ctx, int prec, void* lhs, const void* x, const void* M,
                     const lfr_space* to, const lfr_space* from, int inc, const char* who) {
  if (!ctx) return LFR_ERR_ARG;
  int rc = check_pair(ctx, prec, lhs, x, M, to, from, who);
  if (rc) return rc;
  const int64_t n3 = (int64_t)to->ncell * to->nlayers;
  if (prec == 4)
    return run_cells<float>(ctx, (const float*)M, (const float*)x, (float*)lhs, to, from, n3, inc);
  return run_cells<double>(ctx, (const double*)M, (const double*)x, (double*)lhs, to, from, n3, inc);
}

extern "C" int lfr_dg_matrix_vector(lfr_ctx* ctx, int prec, void* lhs, const void* x,
                                    const void* M, const lfr_space* to, const lfr_space* from) {
  return dg_common(ctx, prec, lhs, x, M, to, from, 0, "lfr_dg_matrix_vector");
}

extern "C" int lfr_dg_inc_matrix_vector(lfr_ctx* ctx, int prec, void* lhs, const void* x,
                                        const void* M, const lfr_space* to,
                                        const lfr_space* from) {
  return dg_common(ctx, prec, lhs, x, M, to, from, 1, "lfr_dg_inc_matrix_vector");
}

// ---------------------------------------------------------------------------
// Gradient of an LMA apply with respect to its local matrices:
//   y(map_to(i)+k) += M(ik,i,j) x(map_from(j)+k)   =>
//   dL/dM(ik,i,j) = ybar(map_to(i)+k) * x(map_from(j)+k)
// Mbar = alpha * ybar (x) x  (+ Mbar when accumulate).  One thread per
// (cell, layer); every matrix plane is written coalesced.  No counterpart in
// the reference (it has no adjoint with respect to operator coefficients).
// ---------------------------------------------------------------------------
namespace {
template <typename T>
__global__ void __launch_bounds__(kThreads)
k_lma_outer(T* __restrict__ Mbar, const T* __restrict__ ybar, const T* __restrict__ x,
            const int32_t* __restrict__ map_to, const int32_t* __restrict__ map_from, int ncell,
            int nl, int ndf_to, int ndf_from, T alpha, int accumulate) {
  const int64_t t = (int64_t)blockIdx.x * kThreads + threadIdx.x;
  const int64_t n3 = (int64_t)ncell * nl;
  if (t >= n3) return;
  const int c = (int)(t / nl);
  const int k = (int)(t - (int64_t)c * nl);
  const int32_t* mt = map_to + (int64_t)c * ndf_to;
  const int32_t* mf = map_from + (int64_t)c * ndf_from;
  for (int j = 0; j < ndf_from; ++j) {
    const T xv = x[mf[j] + k];
    for (int i = 0; i < ndf_to; ++i) {
      const T g = alpha * (ybar[mt[i] + k] * xv);
      T* dst = Mbar + t + n3 * ((int64_t)i + (int64_t)ndf_to * j);
      *dst = accumulate ? *dst + g : g;
    }
  }
}
}  // namespace

extern "C" int lfr_lma_outer(lfr_ctx* ctx, int prec, void* Mbar, const void* ybar, const void* x,
                             const lfr_space* to, const lfr_space* from, double alpha,
                             int accumulate) {
  if (!ctx) return LFR_ERR_ARG;
  int rc = check_pair(ctx, prec, Mbar, ybar, x, to, from, "lfr_lma_outer");
  if (rc) return rc;
  const int64_t n3 = (int64_t)to->ncell * to->nlayers;
  const unsigned nb = lfr_blocks(n3, kThreads);
  if (prec == 4)
    hipLaunchKernelGGL(k_lma_outer<float>, dim3(nb), dim3(kThreads), 0, ctx->stream, (float*)Mbar,
                       (const float*)ybar, (const float*)x, to->map0_d, from->map0_d, to->ncell,
                       to->nlayers, to->ndf, from->ndf, (float)alpha, accumulate);
  else
    hipLaunchKernelGGL(k_lma_outer<double>, dim3(nb), dim3(kThreads), 0, ctx->stream, (double*)Mbar,
                       (const double*)ybar, (const double*)x, to->map0_d, from->map0_d, to->ncell,
                       to->nlayers, to->ndf, from->ndf, alpha, accumulate);
  LFR_CHECK_LAUNCH(ctx);
  return LFR_OK;
}
