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
// One thread per (coarse cell, level), level fastest, so the reads of a
// child's column and the writes of the coarse column are contiguous across
// a wavefront; each output is one thread's sum in the reference order.  The
// children's first dofs are resolved once, when the intergrid handle is
// made, so a thread's loads do not wait on the map.
#include <vector>

#include "lfr_internal.h"

namespace {

// fbase(child, coarse cell): first (0-based) dof of the child's column, in
// the map's order (x fastest, then y)
template <typename T>
__global__ void __launch_bounds__(256)
k_restrict_scalar(T* __restrict__ coarse, const T* __restrict__ fine, const int32_t* __restrict__ map_c,
                  const int32_t* __restrict__ fbase, int ncc, int ndf, int nlev, int nch, T denom) {
  const int64_t t = (int64_t)blockIdx.x * 256 + threadIdx.x;
  if (t >= (int64_t)ncc * nlev) return;
  const int c = (int)(t / nlev);
  const int k = (int)(t - (int64_t)c * nlev);
  const int32_t* fb = fbase + (int64_t)c * nch;
  T v = T(0);
  for (int i = 0; i < nch; ++i) {   // y outer, x inner
    const T p = fine[(int64_t)fb[i] + k] * denom;
    v = v + p;
  }
  coarse[(int64_t)map_c[(int64_t)c * ndf] + k] = v;
}

template <typename T>
__global__ void __launch_bounds__(256)
k_prolong_scalar(T* __restrict__ fine, const T* __restrict__ coarse, const int32_t* __restrict__ map_c,
                 const int32_t* __restrict__ fbase, int ncc, int ndf, int nlev, int nch) {
  const int64_t t = (int64_t)blockIdx.x * 256 + threadIdx.x;
  if (t >= (int64_t)ncc * nlev) return;
  const int c = (int)(t / nlev);
  const int k = (int)(t - (int64_t)c * nlev);
  const int32_t* fb = fbase + (int64_t)c * nch;
  const T cv = coarse[(int64_t)map_c[(int64_t)c * ndf] + k];
  for (int i = 0; i < nch; ++i) {
    const int64_t d = (int64_t)fb[i] + k;
    fine[d] = fine[d] + cv;
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
  int32_t* fbase_d = nullptr;   // (nfx nfy, ncell_coarse): first dof (0-based) of each child's column
};

namespace {

int transfer(lfr_ctx* ctx, bool restrict_, int prec, void* out, const void* in, const lfr_intergrid* ig,
             const char* who) {
  if (!ctx) return LFR_ERR_ARG;
  LFR_REQUIRE(ctx, prec == 4 || prec == 8, "%s: prec must be 4 or 8, got %d", who, prec);
  LFR_REQUIRE(ctx, out && in && ig, "%s: null argument", who);
  LFR_REQUIRE(ctx, out != in, "%s: the output must not alias the input", who);
  const lfr_space* cs = ig->coarse;
  const lfr_space* fs = ig->fine;
  const int nfx = ig->nfx, nfy = ig->nfy;
  const int nlev = cs->nlayers - 1 + cs->ndf;   // top_df + 1
  const int64_t n = (int64_t)cs->ncell * nlev;
  const unsigned nb = (unsigned)((n + 255) / 256);
  const int32_t* fb = ig->fbase_d;
  const int nch = nfx * nfy;
  if (restrict_) {
    if (prec == 4)
      hipLaunchKernelGGL(k_restrict_scalar<float>, dim3(nb), dim3(256), 0, ctx->stream, (float*)out,
                         (const float*)in, cs->map0_d, fb, cs->ncell, cs->ndf, nlev, nch, 1.0f / (float)nch);
    else
      hipLaunchKernelGGL(k_restrict_scalar<double>, dim3(nb), dim3(256), 0, ctx->stream, (double*)out,
                         (const double*)in, cs->map0_d, fb, cs->ncell, cs->ndf, nlev, nch, 1.0 / (double)nch);
  } else {
    if (prec == 4)
      hipLaunchKernelGGL(k_prolong_scalar<float>, dim3(nb), dim3(256), 0, ctx->stream, (float*)out,
                         (const float*)in, cs->map0_d, fb, cs->ncell, cs->ndf, nlev, nch);
    else
      hipLaunchKernelGGL(k_prolong_scalar<double>, dim3(nb), dim3(256), 0, ctx->stream, (double*)out,
                         (const double*)in, cs->map0_d, fb, cs->ncell, cs->ndf, nlev, nch);
  }
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
  LFR_REQUIRE(ctx, cs->ndf == fs->ndf && (cs->ndf == 1 || cs->ndf == 2),
              "%s: lowest-order W3 (ndf 1) or Wtheta (ndf 2) spaces expected (ndf %d and %d)", who, cs->ndf,
              fs->ndf);
  LFR_REQUIRE(ctx, cs->nlayers == fs->nlayers, "%s: the meshes have different layer counts", who);
  LFR_REQUIRE(ctx, nfx >= 1 && nfy >= 1 && (int64_t)cs->ncell * nfx * nfy == fs->ncell,
              "%s: %d coarse cells x %d x %d children do not cover %d fine cells", who, cs->ncell, nfx, nfy,
              fs->ncell);
  LFR_REQUIRE(ctx, column_dofs(cs) && column_dofs(fs), "%s: a cell's dofs are not one column", who);
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
  std::vector<int32_t> fb((size_t)nmap);
  for (int64_t i = 0; i < nmap; ++i) fb[(size_t)i] = fs->map_h[(size_t)(cell_map_h[i] - 1) * fs->ndf] - 1;
  if (hipMalloc(&ig->fbase_d, (size_t)nmap * sizeof(int32_t)) != hipSuccess) {
    (void)hipGetLastError();
    delete ig;
    return lfr_set_error(ctx, LFR_ERR_NOMEM, "%s: cannot allocate the map", who);
  }
  const hipError_t e = hipMemcpy(ig->fbase_d, fb.data(), (size_t)nmap * sizeof(int32_t), hipMemcpyHostToDevice);
  if (e != hipSuccess) {
    (void)hipFree(ig->fbase_d);
    delete ig;
    return lfr_set_error(ctx, LFR_ERR_HIP, "%s: %s", who, hipGetErrorString(e));
  }
  *out = ig;
  return LFR_OK;
}

extern "C" int lfr_intergrid_destroy(lfr_intergrid* ig) {
  if (!ig) return LFR_OK;
  if (ig->fbase_d) (void)hipFree(ig->fbase_d);
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
