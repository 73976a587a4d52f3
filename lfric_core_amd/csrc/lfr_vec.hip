// Vector algebra of the solver path: the PSyclone built-ins behind
// field_vector_type / r_solver_field_vector_type
// (components/science/source/algorithm/solver/sci_field_vector_mod.x90:132-283,
//  sci_r_solver_field_vector_mod.x90:119-296) and the mixed-precision inner
// products of components/science/source/psy/sci_psykal_light_mod.f90:72-214.
//
// Scalars arrive as double (r_def) and are cast to the field's kind before
// the update, as r_solver_field_vector_type does (:189-195).  Inner products
// of r_solver data accumulate real(x,r_double)*real(y,r_double) in FP64.
#include <mutex>

#include "lfr_internal.h"
#include "lfr_reduce.h"

// Scratch of the allocation-free inner product lfr_vec_dot_dev_noalloc: the
// first-stage partial sums, one slot per reduction block (the grid never has
// more than lfr::kRedMaxBlocks blocks, so 8 KiB serve every length).  Each GPU
// holds its own copy for the life of the process.  Host code never passes
// this name to a kernel: it uses the device address that lfr_noalloc_setup
// resolves once per device.
constexpr int kScratchSlots = lfr::kRedMaxBlocks;
static __device__ __align__(256) double scratch[kScratchSlots];

namespace {

constexpr int kMaxDevices = 64;
std::mutex g_scratch_mu;
double* g_scratch[kMaxDevices] = {};

}  // namespace

// Resolve this device's copy of `scratch` (called by lfr_init with the
// context's device current).
int lfr_noalloc_setup(lfr_ctx* ctx) {
  LFR_REQUIRE(ctx, ctx->device >= 0 && ctx->device < kMaxDevices, "device %d beyond the %d the library tracks",
              ctx->device, kMaxDevices);
  std::lock_guard<std::mutex> lk(g_scratch_mu);
  if (g_scratch[ctx->device]) return LFR_OK;
  double* p = nullptr;
  const hipError_t e = hipGetSymbolAddress((void**)&p, HIP_SYMBOL(scratch));
  if (e != hipSuccess || !p)
    return lfr_set_error(ctx, LFR_ERR_HIP, "hipGetSymbolAddress(scratch) on device %d: %s", ctx->device,
                         hipGetErrorString(e));
  g_scratch[ctx->device] = p;
  return LFR_OK;
}

namespace {

constexpr int kT = 256;

template <typename T>
__global__ void k_setval(T* __restrict__ x, int64_t n, T c) {
  for (int64_t i = (int64_t)blockIdx.x * kT + threadIdx.x; i < n; i += (int64_t)gridDim.x * kT) x[i] = c;
}
template <typename T>
__global__ void k_copy(T* __restrict__ y, const T* __restrict__ x, int64_t n) {
  for (int64_t i = (int64_t)blockIdx.x * kT + threadIdx.x; i < n; i += (int64_t)gridDim.x * kT) y[i] = x[i];
}
// y = y + a*x   (inc_X_plus_bY)
template <typename T>
__global__ void k_axpy(T* __restrict__ y, const T* __restrict__ x, int64_t n, T a) {
  for (int64_t i = (int64_t)blockIdx.x * kT + threadIdx.x; i < n; i += (int64_t)gridDim.x * kT) {
    const T p = a * x[i];
    y[i] = y[i] + p;
  }
}
// y = a*y + x   (inc_aX_plus_Y)
template <typename T>
__global__ void k_aypx(T* __restrict__ y, const T* __restrict__ x, int64_t n, T a) {
  for (int64_t i = (int64_t)blockIdx.x * kT + threadIdx.x; i < n; i += (int64_t)gridDim.x * kT) {
    const T p = a * y[i];
    y[i] = p + x[i];
  }
}
// z = a*x + b*y (aX_plus_bY)
template <typename T>
__global__ void k_axpby(T* __restrict__ z, const T* __restrict__ x, const T* __restrict__ y,
                        int64_t n, T a, T b) {
  for (int64_t i = (int64_t)blockIdx.x * kT + threadIdx.x; i < n; i += (int64_t)gridDim.x * kT) {
    const T p = a * x[i];
    const T q = b * y[i];
    z[i] = p + q;
  }
}
template <typename T>
__global__ void k_scale(T* __restrict__ x, int64_t n, T a) {
  for (int64_t i = (int64_t)blockIdx.x * kT + threadIdx.x; i < n; i += (int64_t)gridDim.x * kT) x[i] = a * x[i];
}
// z = x / y (X_divideby_Y), z = x * y (X_times_Y)
template <typename T>
__global__ void k_divide(T* __restrict__ z, const T* __restrict__ x, const T* __restrict__ y, int64_t n) {
  for (int64_t i = (int64_t)blockIdx.x * kT + threadIdx.x; i < n; i += (int64_t)gridDim.x * kT) z[i] = x[i] / y[i];
}
template <typename T>
__global__ void k_multiply(T* __restrict__ z, const T* __restrict__ x, const T* __restrict__ y, int64_t n) {
  for (int64_t i = (int64_t)blockIdx.x * kT + threadIdx.x; i < n; i += (int64_t)gridDim.x * kT) z[i] = x[i] * y[i];
}
template <typename TO, typename TI>
__global__ void k_convert(TO* __restrict__ y, const TI* __restrict__ x, int64_t n) {
  for (int64_t i = (int64_t)blockIdx.x * kT + threadIdx.x; i < n; i += (int64_t)gridDim.x * kT) y[i] = (TO)x[i];
}

// First stage of <x, y> in FP64: block b handles a fixed interleaved subset.
template <typename T>
__global__ void __launch_bounds__(lfr::kRedThreads)
k_dot_part(const T* __restrict__ x, const T* __restrict__ y, int64_t n, double* __restrict__ part) {
  __shared__ double lds[lfr::kRedThreads / 64];
  double s = 0.0;
  const int64_t st = (int64_t)gridDim.x * lfr::kRedThreads;
  int64_t i = (int64_t)blockIdx.x * lfr::kRedThreads + threadIdx.x;
  // eight elements' loads issued before their products are added, in the
  // same order as one at a time
  for (; i + 7 * st < n; i += 8 * st) {
    T a[8], b[8];
#pragma unroll
    for (int q = 0; q < 8; ++q) {
      a[q] = x[i + q * st];
      b[q] = y[i + q * st];
    }
#pragma unroll
    for (int q = 0; q < 8; ++q) {
      const double p = (double)a[q] * (double)b[q];
      s = s + p;
    }
  }
  // the last (at most seven) elements loaded together at clamped indices; a
  // slot past the end adds +0 (selected, not 0 * x, which may be NaN), which
  // leaves the sum bitwise unchanged: it starts at +0 and, rounding to
  // nearest, never becomes -0
  if (n > 0) {
    T a[7], b[7];
#pragma unroll
    for (int q = 0; q < 7; ++q) {
      const int64_t k = i + q * st < n ? i + q * st : n - 1;
      a[q] = x[k];
      b[q] = y[k];
    }
#pragma unroll
    for (int q = 0; q < 7; ++q) {
      const double p = (double)a[q] * (double)b[q];
      s = s + (i + q * st < n ? p : 0.0);
    }
  }
  s = lfr::block_sum(s, lds);
  if (threadIdx.x == 0) part[blockIdx.x] = s;
}

unsigned grid_for(int64_t n) {
  // enough waves to saturate HBM: up to 8 blocks of 256 per CU
  return lfr_blocks(n, kT, 2048);
}

bool prec_ok(int p) { return p == 4 || p == 8; }

}  // namespace

// Queue <x, y> (FP64 accumulate) into out_d[0] using `part` scratch of
// red_blocks(n) doubles.
int lfr_dot_queue(lfr_ctx* ctx, int prec, int64_t n, const void* x, const void* y, double* part,
                  double* out_d) {
  const int nb = lfr::red_blocks(n);
  if (prec == 4)
    hipLaunchKernelGGL(k_dot_part<float>, dim3(nb), dim3(lfr::kRedThreads), 0, ctx->stream,
                       (const float*)x, (const float*)y, n, part);
  else
    hipLaunchKernelGGL(k_dot_part<double>, dim3(nb), dim3(lfr::kRedThreads), 0, ctx->stream,
                       (const double*)x, (const double*)y, n, part);
  LFR_CHECK_LAUNCH(ctx);
  hipLaunchKernelGGL(lfr::k_reduce_final, dim3(1), dim3(lfr::kRedThreads), 0, ctx->stream, part, nb,
                     0, out_d);
  LFR_CHECK_LAUNCH(ctx);
  return LFR_OK;
}

#define VEC_PRELUDE(name)                                                        \
  if (!ctx) return LFR_ERR_ARG;                                                 \
  LFR_REQUIRE(ctx, prec_ok(prec), name ": prec must be 4 or 8, got %d", prec);  \
  LFR_REQUIRE(ctx, n >= 0, name ": negative length");                          \
  if (n == 0) return LFR_OK;                                                    \
  const unsigned g = grid_for(n);

extern "C" {

int lfr_vec_setval(lfr_ctx* ctx, int prec, int64_t n, void* x, double c) {
  VEC_PRELUDE("lfr_vec_setval");
  LFR_REQUIRE(ctx, x, "lfr_vec_setval: null");
  if (prec == 4) hipLaunchKernelGGL(k_setval<float>, dim3(g), dim3(kT), 0, ctx->stream, (float*)x, n, (float)c);
  else hipLaunchKernelGGL(k_setval<double>, dim3(g), dim3(kT), 0, ctx->stream, (double*)x, n, c);
  LFR_CHECK_LAUNCH(ctx);
  return LFR_OK;
}

int lfr_vec_copy(lfr_ctx* ctx, int prec, int64_t n, void* y, const void* x) {
  VEC_PRELUDE("lfr_vec_copy");
  LFR_REQUIRE(ctx, x && y, "lfr_vec_copy: null");
  if (prec == 4) hipLaunchKernelGGL(k_copy<float>, dim3(g), dim3(kT), 0, ctx->stream, (float*)y, (const float*)x, n);
  else hipLaunchKernelGGL(k_copy<double>, dim3(g), dim3(kT), 0, ctx->stream, (double*)y, (const double*)x, n);
  LFR_CHECK_LAUNCH(ctx);
  return LFR_OK;
}

int lfr_vec_axpy(lfr_ctx* ctx, int prec, int64_t n, double a, const void* x, void* y) {
  VEC_PRELUDE("lfr_vec_axpy");
  LFR_REQUIRE(ctx, x && y, "lfr_vec_axpy: null");
  if (prec == 4) hipLaunchKernelGGL(k_axpy<float>, dim3(g), dim3(kT), 0, ctx->stream, (float*)y, (const float*)x, n, (float)a);
  else hipLaunchKernelGGL(k_axpy<double>, dim3(g), dim3(kT), 0, ctx->stream, (double*)y, (const double*)x, n, a);
  LFR_CHECK_LAUNCH(ctx);
  return LFR_OK;
}

int lfr_vec_aypx(lfr_ctx* ctx, int prec, int64_t n, double a, const void* x, void* y) {
  VEC_PRELUDE("lfr_vec_aypx");
  LFR_REQUIRE(ctx, x && y, "lfr_vec_aypx: null");
  if (prec == 4) hipLaunchKernelGGL(k_aypx<float>, dim3(g), dim3(kT), 0, ctx->stream, (float*)y, (const float*)x, n, (float)a);
  else hipLaunchKernelGGL(k_aypx<double>, dim3(g), dim3(kT), 0, ctx->stream, (double*)y, (const double*)x, n, a);
  LFR_CHECK_LAUNCH(ctx);
  return LFR_OK;
}

int lfr_vec_axpby(lfr_ctx* ctx, int prec, int64_t n, double a, const void* x, double b,
                  const void* y, void* z) {
  VEC_PRELUDE("lfr_vec_axpby");
  LFR_REQUIRE(ctx, x && y && z, "lfr_vec_axpby: null");
  if (prec == 4)
    hipLaunchKernelGGL(k_axpby<float>, dim3(g), dim3(kT), 0, ctx->stream, (float*)z, (const float*)x,
                       (const float*)y, n, (float)a, (float)b);
  else
    hipLaunchKernelGGL(k_axpby<double>, dim3(g), dim3(kT), 0, ctx->stream, (double*)z, (const double*)x,
                       (const double*)y, n, a, b);
  LFR_CHECK_LAUNCH(ctx);
  return LFR_OK;
}

int lfr_vec_scale(lfr_ctx* ctx, int prec, int64_t n, double a, void* x) {
  VEC_PRELUDE("lfr_vec_scale");
  LFR_REQUIRE(ctx, x, "lfr_vec_scale: null");
  if (prec == 4) hipLaunchKernelGGL(k_scale<float>, dim3(g), dim3(kT), 0, ctx->stream, (float*)x, n, (float)a);
  else hipLaunchKernelGGL(k_scale<double>, dim3(g), dim3(kT), 0, ctx->stream, (double*)x, n, a);
  LFR_CHECK_LAUNCH(ctx);
  return LFR_OK;
}

int lfr_vec_divide(lfr_ctx* ctx, int prec, int64_t n, void* z, const void* x, const void* y) {
  VEC_PRELUDE("lfr_vec_divide");
  LFR_REQUIRE(ctx, x && y && z, "lfr_vec_divide: null");
  if (prec == 4) hipLaunchKernelGGL(k_divide<float>, dim3(g), dim3(kT), 0, ctx->stream, (float*)z, (const float*)x, (const float*)y, n);
  else hipLaunchKernelGGL(k_divide<double>, dim3(g), dim3(kT), 0, ctx->stream, (double*)z, (const double*)x, (const double*)y, n);
  LFR_CHECK_LAUNCH(ctx);
  return LFR_OK;
}

int lfr_vec_multiply(lfr_ctx* ctx, int prec, int64_t n, void* z, const void* x, const void* y) {
  VEC_PRELUDE("lfr_vec_multiply");
  LFR_REQUIRE(ctx, x && y && z, "lfr_vec_multiply: null");
  if (prec == 4) hipLaunchKernelGGL(k_multiply<float>, dim3(g), dim3(kT), 0, ctx->stream, (float*)z, (const float*)x, (const float*)y, n);
  else hipLaunchKernelGGL(k_multiply<double>, dim3(g), dim3(kT), 0, ctx->stream, (double*)z, (const double*)x, (const double*)y, n);
  LFR_CHECK_LAUNCH(ctx);
  return LFR_OK;
}

int lfr_vec_convert(lfr_ctx* ctx, int prec_out, void* y, int prec_in, const void* x, int64_t n) {
  if (!ctx) return LFR_ERR_ARG;
  LFR_REQUIRE(ctx, prec_ok(prec_out) && prec_ok(prec_in), "lfr_vec_convert: bad precision");
  LFR_REQUIRE(ctx, x && y && n >= 0, "lfr_vec_convert: bad arguments");
  if (n == 0) return LFR_OK;
  const unsigned g = grid_for(n);
  if (prec_out == 4 && prec_in == 8)
    hipLaunchKernelGGL((k_convert<float, double>), dim3(g), dim3(kT), 0, ctx->stream, (float*)y, (const double*)x, n);
  else if (prec_out == 8 && prec_in == 4)
    hipLaunchKernelGGL((k_convert<double, float>), dim3(g), dim3(kT), 0, ctx->stream, (double*)y, (const float*)x, n);
  else if (prec_out == 4)
    hipLaunchKernelGGL(k_copy<float>, dim3(g), dim3(kT), 0, ctx->stream, (float*)y, (const float*)x, n);
  else
    hipLaunchKernelGGL(k_copy<double>, dim3(g), dim3(kT), 0, ctx->stream, (double*)y, (const double*)x, n);
  LFR_CHECK_LAUNCH(ctx);
  return LFR_OK;
}

int lfr_vec_dot(lfr_ctx* ctx, int prec, int64_t n, const void* x, const void* y, double* out) {
  if (!ctx) return LFR_ERR_ARG;
  LFR_REQUIRE(ctx, prec_ok(prec) && x && y && out && n >= 0, "lfr_vec_dot: bad arguments");
  if (n == 0) {
    *out = 0.0;
    return LFR_OK;
  }
  int rc = lfr_work_reserve(ctx, (lfr::kRedMaxBlocks + 8) * sizeof(double));
  if (rc) return rc;
  double* part = (double*)ctx->work;
  double* res = part + lfr::kRedMaxBlocks;
  rc = lfr_dot_queue(ctx, prec, n, x, y, part, res);
  if (rc) return rc;
  LFR_HIP(ctx, hipMemcpyAsync(ctx->host_scalars, res, sizeof(double), hipMemcpyDeviceToHost, ctx->stream));
  LFR_HIP(ctx, hipStreamSynchronize(ctx->stream));
  *out = ctx->host_scalars[0];
  return LFR_OK;
}

int lfr_vec_dot_dev_noalloc(lfr_ctx* ctx, int prec, int64_t n, const void* x, const void* y, double* out_d) {
  if (!ctx) return LFR_ERR_ARG;
  LFR_REQUIRE(ctx, prec_ok(prec) && x && y && out_d && n > 0, "lfr_vec_dot_dev_noalloc: bad arguments");
  const int nb = lfr::red_blocks(n);
  LFR_REQUIRE(ctx, nb <= kScratchSlots, "lfr_vec_dot_dev_noalloc: %d partial sums exceed the %d-slot scratch", nb,
              kScratchSlots);
  double* part = nullptr;
  {
    std::lock_guard<std::mutex> lk(g_scratch_mu);
    part = g_scratch[ctx->device];
  }
  LFR_REQUIRE(ctx, part, "lfr_vec_dot_dev_noalloc: scratch of device %d not set up", ctx->device);
  // every slot the final stage sums (0 .. nb-1) is written by the first
  // stage of this call, so nothing left by an earlier call is read
  return lfr_dot_queue(ctx, prec, n, x, y, part, out_d);
}

int lfr_vec_dot_dev(lfr_ctx* ctx, int prec, int64_t n, const void* x, const void* y,
                    double* out_d) {
  if (!ctx) return LFR_ERR_ARG;
  LFR_REQUIRE(ctx, prec_ok(prec) && x && y && out_d && n > 0, "lfr_vec_dot_dev: bad arguments");
  int rc = lfr_work_reserve(ctx, (lfr::kRedMaxBlocks + 8) * sizeof(double));
  if (rc) return rc;
  return lfr_dot_queue(ctx, prec, n, x, y, (double*)ctx->work, out_d);
}

}  // extern "C"

// ---------------------------------------------------------------------------
// enforce_bc (components/science/source/kernel/algebra/sci_enforce_bc_kernel_mod.F90:63-91)
// ---------------------------------------------------------------------------
namespace {

constexpr int kBcMaxNdf = 64;
struct BcValues {
  int v[2][kBcMaxNdf];
};

// thread = (cell, df, end): field(map(df) + (end ? nl-1 : 0)) *= value.  Cells
// sharing a dof carry the same flag for it, so concurrent updates agree.
template <typename T>
__global__ void k_enforce_bc(T* __restrict__ field, const int32_t* __restrict__ map0, int ncell, int ndf, int nl,
                             BcValues bv) {
  const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= (int64_t)ncell * ndf * 2) return;
  const int e = (int)(t & 1);
  const int64_t cd = t >> 1;
  const int df = (int)(cd % ndf);
  const int v = bv.v[e][df];
  if (v == 1) return;
  const int64_t d = (int64_t)map0[cd] + (e ? nl - 1 : 0);
  field[d] = field[d] * (T)v;
}

}  // namespace

int lfr_enforce_bc_queue(lfr_ctx* ctx, int prec, void* field, const lfr_space* sp) {
  LFR_REQUIRE(ctx, sp && field && (prec == 4 || prec == 8), "enforce_bc: bad arguments");
  LFR_REQUIRE(ctx, (int)sp->boundary_h.size() == 2 * sp->ndf, "enforce_bc: the space has no boundary values");
  LFR_REQUIRE(ctx, sp->ndf <= kBcMaxNdf, "enforce_bc: ndf %d > %d", sp->ndf, kBcMaxNdf);
  BcValues bv;
  for (int e = 0; e < 2; ++e)
    for (int df = 0; df < sp->ndf; ++df) bv.v[e][df] = sp->boundary_h[df + sp->ndf * e];
  const int64_t n = (int64_t)sp->ncell * sp->ndf * 2;
  const unsigned g = (unsigned)((n + 255) / 256);
  if (prec == 4)
    hipLaunchKernelGGL(k_enforce_bc<float>, dim3(g), dim3(256), 0, ctx->stream, (float*)field, sp->map0_d, sp->ncell,
                       sp->ndf, sp->nlayers, bv);
  else
    hipLaunchKernelGGL(k_enforce_bc<double>, dim3(g), dim3(256), 0, ctx->stream, (double*)field, sp->map0_d,
                       sp->ncell, sp->ndf, sp->nlayers, bv);
  LFR_CHECK_LAUNCH(ctx);
  return LFR_OK;
}

extern "C" int lfr_space_set_boundary(lfr_space* sp, const int32_t* bv) {
  if (!sp || !bv) return LFR_ERR_ARG;
  sp->boundary_h.assign(bv, bv + 2 * sp->ndf);
  return LFR_OK;
}

extern "C" int lfr_enforce_bc(lfr_ctx* ctx, int prec, void* field_d, const lfr_space* space) {
  if (!ctx) return LFR_ERR_ARG;
  return lfr_enforce_bc_queue(ctx, prec, field_d, space);
}
