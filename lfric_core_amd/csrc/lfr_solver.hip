// Krylov solvers over LMA operators, device resident.
//
// conjugate_gradient_type%cg_solve
// (components/science/source/solver/sci_iterative_solver_mod.f90:495-592) with
// the same statement sequence, FP64 (r_def) scalars and the same convergence
// test, restated as a fixed chain of kernels per iteration:
//
//   apply    z = A p            gather pass, fused partial sums of <p, z>
//   alpha    alpha = rz / <p,z> one block, fixed-order combine of partials
//   update   x += alpha p, r -= alpha z (, z = r / diag), partials of <r,r>, <r,z>
//   beta     ||r||, convergence test, beta = rz_new / rz
//   direction p = z + beta p
//
// The scalars never leave the device.  The convergence predicate is
// evaluated on the device; once it holds every later kernel returns at
// entry, so the host launches iterations in captured chunks (a hipGraph of
// `chunk` iterations) and reads the state back once per chunk.  The iterate
// sequence is exactly the reference's: the predicate only stops work that
// the reference would not have done.
//
// With the null preconditioner z = P^{-1} r is a copy of r
// (sci_null_preconditioner_alg_mod.x90:84); r is used in its place.
#include <algorithm>
#include <cmath>
#include <cstdlib>
#include <cstring>
#include <vector>

#include "lfr_internal.h"
#include "lfr_lma_kernels.h"
#include "lfr_reduce.h"

namespace {

using lfr::kThreads;

enum Slot { S_STOP = 0, S_RZ, S_ALPHA, S_BETA, S_RNRM0, S_RNRM, S_ITER, S_NAN, S_RR, S_NSLOT = 16 };
// S_STOP: 0 running, 1 converged, 2 max_iter reached, 3 NaN

constexpr double kEps = 2.220446049250313e-16;  // epsilon(1.0_r_def), constants_mod.F90:167

// ---------------------------------------------------------------------------
// kernels
// ---------------------------------------------------------------------------
// alpha = rz / <p, A p> from the apply's per-workgroup partials (one per
// gather workgroup, ~12,000 at C96L70): 1024 threads share the sum.
constexpr int kAlphaThreads = 1024;

// rz = <r, z> (z = r without a preconditioner) is summed here from the
// partials of the pass that formed r (residual or update), stored in S_RZ
// for this iteration's beta, and used for alpha.
__global__ void __launch_bounds__(kAlphaThreads)
k_cg_alpha(const double* __restrict__ part, int np, const double* __restrict__ part_rz, int nrz,
           double* __restrict__ S, int max_iter) {
  if (S[S_STOP] != 0.0) return;
  if (S[S_ITER] >= (double)max_iter) {
    if (threadIdx.x == 0) S[S_STOP] = 2.0;
    return;
  }
  __shared__ double lds[kAlphaThreads / 64];
  // both threads' shares first, so that their loads are in flight together
  // (each sum's order is unchanged)
  const double sp = lfr::strided_sum<kAlphaThreads>(part, np);
  // rz in the order k_cg_beta_direction sums the same partials (256-thread
  // stride; the extra waves add zeros), so both see the same bits
  const double sz = threadIdx.x < kThreads ? lfr::strided_sum<kThreads>(part_rz, nrz) : 0.0;
  const double s = lfr::block_sum<kAlphaThreads>(sp, lds);
  const double rz = lfr::block_sum<kAlphaThreads>(sz, lds);
  if (threadIdx.x == 0) {
    S[S_RZ] = rz;
    S[S_ALPHA] = rz / s;
    S[S_ITER] = S[S_ITER] + 1.0;
  }
}

template <typename T, bool DIAG>
__global__ void __launch_bounds__(kThreads)
k_cg_update(T* __restrict__ x, T* __restrict__ r, const T* __restrict__ p, T* __restrict__ z,
            const T* __restrict__ diag, int64_t n, const double* __restrict__ S,
            double* __restrict__ part_rr, double* __restrict__ part_rz) {
  if (S[S_STOP] != 0.0) return;
  __shared__ double lds[kThreads / 64];
  const double alpha = S[S_ALPHA];
  const T a = (T)alpha, ma = (T)(-alpha);
  double rr = 0.0, rz = 0.0;
  for (int64_t i = (int64_t)blockIdx.x * kThreads + threadIdx.x; i < n;
       i += (int64_t)gridDim.x * kThreads) {
    const T xp = a * p[i];
    x[i] = x[i] + xp;
    const T rq = ma * z[i];
    const T ri = r[i] + rq;
    r[i] = ri;
    const double rd = (double)ri;
    rr = rr + rd * rd;
    if (DIAG) {
      const T zi = ri / diag[i];
      z[i] = zi;
      rz = rz + rd * (double)zi;
    }
  }
  rr = lfr::block_sum(rr, lds);
  if (DIAG) rz = lfr::block_sum(rz, lds);
  if (threadIdx.x == 0) {
    part_rr[blockIdx.x] = rr;
    if (DIAG) part_rz[blockIdx.x] = rz;
  }
}

// beta and the convergence test, then p = beta * p + z.  Every workgroup
// sums the update's partials in the same order (identical bits); workgroup 0
// records the results; S_RZ (read here) is only written by k_cg_alpha.
template <typename T>
__global__ void __launch_bounds__(kThreads)
k_cg_beta_direction(T* __restrict__ p, const T* __restrict__ z, int64_t n, const double* __restrict__ part_rr,
                    const double* __restrict__ part_rz, int np, double* __restrict__ S, int monitor, double r_tol,
                    double a_tol, int diag) {
  if (S[S_STOP] != 0.0) return;
  __shared__ double lds[kThreads / 64];
  __shared__ double bc[2];
  const double srr = lfr::strided_sum(part_rr, np);
  const double srz = diag ? lfr::strided_sum(part_rz, np) : 0.0;
  const double rr = lfr::block_sum(srr, lds);
  const double rzs = diag ? lfr::block_sum(srz, lds) : 0.0;
  if (threadIdx.x == 0) {
    const double rz = diag ? rzs : rr;
    const double rn = sqrt(rr);
    int stop = 0;
    if (rn != rn) stop = 3;
    else if (monitor && ((rn / S[S_RNRM0] <= r_tol) || (rn <= a_tol))) stop = 1;
    const double beta = rz / S[S_RZ];
    if (blockIdx.x == 0) {
      S[S_RNRM] = rn;
      S[S_RR] = rr;
      if (stop == 3) S[S_NAN] = 1.0;
      if (stop) S[S_STOP] = (double)stop;
      else S[S_BETA] = beta;
    }
    bc[0] = (double)stop;
    bc[1] = beta;
  }
  __syncthreads();
  if (bc[0] != 0.0) return;
  const T b = (T)bc[1];
  for (int64_t i = (int64_t)blockIdx.x * kThreads + threadIdx.x; i < n;
       i += (int64_t)gridDim.x * kThreads) {
    const T bp = b * p[i];
    p[i] = bp + z[i];
  }
}

// r = b - r  (r holds A x): the reference's r%scale(-1); r%axpy(1, b)
// (sci_iterative_solver_mod.f90:518-520), with FP64 partials of <r, r>.
template <typename T>
__global__ void __launch_bounds__(kThreads)
k_residual(T* __restrict__ r, const T* __restrict__ b, int64_t n, double* __restrict__ part) {
  __shared__ double lds[kThreads / 64];
  double rr = 0.0;
  for (int64_t i = (int64_t)blockIdx.x * kThreads + threadIdx.x; i < n;
       i += (int64_t)gridDim.x * kThreads) {
    const T m = (T)(-1) * r[i];
    const T v = m + b[i];
    r[i] = v;
    rr = rr + (double)v * (double)v;
  }
  rr = lfr::block_sum(rr, lds);
  if (threadIdx.x == 0) part[blockIdx.x] = rr;
}

// The start of a solve from x = 0 in one pass: x = 0, r = b - A 0 = b
// (the reference's r%scale(-1); r%axpy(1, b) on A x = 0 gives b exactly),
// p = r when p is given (no preconditioner), and the partials of <r, r> in
// k_residual's grid and order.
template <typename T>
__global__ void __launch_bounds__(kThreads)
k_cg_start_zero(T* __restrict__ x, T* __restrict__ r, T* __restrict__ p, const T* __restrict__ b, int64_t n,
                double* __restrict__ part) {
  __shared__ double lds[kThreads / 64];
  double rr = 0.0;
  for (int64_t i = (int64_t)blockIdx.x * kThreads + threadIdx.x; i < n;
       i += (int64_t)gridDim.x * kThreads) {
    const T v = b[i];
    x[i] = T(0);
    r[i] = v;
    if (p) p[i] = v;
    rr = rr + (double)v * (double)v;
  }
  rr = lfr::block_sum(rr, lds);
  if (threadIdx.x == 0) part[blockIdx.x] = rr;
}

// Refinement residual and the start of its inner FP32 solve in one pass:
// r = b - A x in FP64 from A x (in ax; r itself is not stored), FP64
// partials of <r, r> in part64, and the inner solve's start from zero,
// d = 0 and r32 = p32 = (float) r, with FP64 partials of <r32, r32> in
// part32.  The sums are k_residual's and k_cg_start_zero's (on the FP32
// right-hand side (float) r), in their grid and order.
__global__ void __launch_bounds__(kThreads)
k_residual_start32(const double* __restrict__ ax, const double* __restrict__ b, int64_t n,
                   double* __restrict__ part64, float* __restrict__ d, float* __restrict__ r32,
                   float* __restrict__ p32, double* __restrict__ part32) {
  __shared__ double lds[kThreads / 64];
  double rr = 0.0, ff = 0.0;
  for (int64_t i = (int64_t)blockIdx.x * kThreads + threadIdx.x; i < n;
       i += (int64_t)gridDim.x * kThreads) {
    const double m = -1.0 * ax[i];
    const double v = m + b[i];
    const float f = (float)v;
    d[i] = 0.0f;
    r32[i] = f;
    p32[i] = f;
    rr = rr + v * v;
    ff = ff + (double)f * (double)f;
  }
  rr = lfr::block_sum(rr, lds);
  ff = lfr::block_sum(ff, lds);
  if (threadIdx.x == 0) {
    part64[blockIdx.x] = rr;
    part32[blockIdx.x] = ff;
  }
}

// ||r0|| < epsilon: cg converged in 0 iterations
// (sci_iterative_solver_mod.f90:523-529); every later kernel returns at entry.
__global__ void k_cg_start(double* __restrict__ S, double eps) {
  if (threadIdx.x != 0) return;
  S[S_RNRM] = S[S_RNRM0];
  if (S[S_RNRM0] < eps) S[S_STOP] = 1.0;
}

// S[slot] = (sqrt?) sum(part)
__global__ void __launch_bounds__(kThreads)
k_store_sum(const double* __restrict__ part, int np, double* __restrict__ S, int slot, int take_sqrt) {
  __shared__ double lds[kThreads / 64];
  double s = 0.0;
  s = lfr::strided_sum(part, np);
  s = lfr::block_sum(s, lds);
  if (threadIdx.x == 0) S[slot] = take_sqrt ? sqrt(s) : s;
}

template <typename T>
__global__ void k_fill(T* __restrict__ x, int64_t n, T v) {
  for (int64_t i = (int64_t)blockIdx.x * kThreads + threadIdx.x; i < n; i += (int64_t)gridDim.x * kThreads)
    x[i] = v;
}

template <typename T>
__global__ void k_copy_v(T* __restrict__ y, const T* __restrict__ x, int64_t n) {
  for (int64_t i = (int64_t)blockIdx.x * kThreads + threadIdx.x; i < n; i += (int64_t)gridDim.x * kThreads)
    y[i] = x[i];
}

// z = r / diag with partials of <r, z>
template <typename T>
__global__ void __launch_bounds__(kThreads)
k_prec_diag(T* __restrict__ z, const T* __restrict__ r, const T* __restrict__ diag, int64_t n,
            double* __restrict__ part) {
  __shared__ double lds[kThreads / 64];
  double rz = 0.0;
  for (int64_t i = (int64_t)blockIdx.x * kThreads + threadIdx.x; i < n;
       i += (int64_t)gridDim.x * kThreads) {
    const T zi = r[i] / diag[i];
    z[i] = zi;
    rz = rz + (double)r[i] * (double)zi;
  }
  rz = lfr::block_sum(rz, lds);
  if (threadIdx.x == 0) part[blockIdx.x] = rz;
}

// x64 += (double) d32   (refinement correction)
__global__ void k_add_correction(double* __restrict__ x, const float* __restrict__ d, int64_t n) {
  for (int64_t i = (int64_t)blockIdx.x * kThreads + threadIdx.x; i < n; i += (int64_t)gridDim.x * kThreads)
    x[i] = x[i] + (double)d[i];
}

// ---------------------------------------------------------------------------
// host helpers
// ---------------------------------------------------------------------------
struct Ws {
  void* r = nullptr;
  void* p = nullptr;
  void* z = nullptr;
  void* aux = nullptr;       // refinement: fp64 A x, fp32 correction
  void* aux2 = nullptr;
  double* part_apply = nullptr;
  double* part_rr = nullptr;
  double* part_rz = nullptr;
  double* S = nullptr;
  int napply = 0;
};

size_t align_up(size_t b) { return (b + 255) & ~(size_t)255; }

// Partial sums written by one fused apply: padded block count of every class
// under the kernel apply_op launches for it (depends on the precision, which
// sets the levels per thread, and on whether the operator is packed).
bool apply_unified(const lfr_plan* plan, int prec, const lfr_lma_packed* pk) {
  if (pk) return pk->unified;
  return lfr_records_enabled() && lfr_plan_unified(plan, plan->space->ndf, lfr_unified_lpt(prec));
}

int apply_partials(const lfr_plan* plan, int prec, bool packed) {
  int np = 0;
  const lfr_space* sp = plan->space;
  if (packed ? lfr_plan_unified(plan, sp->ndf, lfr_unified_lpt(prec)) : apply_unified(plan, prec, nullptr))
    return (int)((lfr_unified_blocks(plan, lfr_unified_lpt(prec)) + 7u) & ~7u);
  for (const auto& c : plan->classes) {
    const unsigned nb = lfr_class_blocks(c, lfr_class_lpt(c, sp->nlayers, prec, sp->ndf, packed));
    np += (int)((nb + 7u) & ~7u);
  }
  return np;
}

int apply_partials_max(const lfr_plan* plan) {
  int np = 0;
  for (int prec : {4, 8})
    for (bool packed : {false, true}) np = std::max(np, apply_partials(plan, prec, packed));
  return np;
}

// Carve the solver workspace out of one grow-only device buffer.
int ws_get(lfr_ctx* ctx, int64_t n, int prec, const lfr_plan* plan, int naux, Ws& w) {
  const size_t vb = align_up((size_t)n * prec);
  // room for the partials of an apply in either precision (the refined
  // solver runs FP32 applies in an FP64-sized workspace)
  const int napply = apply_partials_max(plan);
  const size_t need = 3 * vb + (size_t)naux * vb + align_up(napply * sizeof(double)) +
                      2 * align_up(lfr::kRedMaxBlocks * sizeof(double)) +
                      align_up(S_NSLOT * sizeof(double)) + align_up(8 * sizeof(double));
  // the first 8 doubles stay free for lfr_vec_dot's result slot users
  int rc = lfr_work_reserve(ctx, need + align_up((lfr::kRedMaxBlocks + 8) * sizeof(double)));
  if (rc) return rc;
  char* b = (char*)ctx->work + align_up((lfr::kRedMaxBlocks + 8) * sizeof(double));
  w.r = b; b += vb;
  w.p = b; b += vb;
  w.z = b; b += vb;
  if (naux > 0) { w.aux = b; b += vb; }
  if (naux > 1) { w.aux2 = b; b += vb; }
  w.part_apply = (double*)b; b += align_up(napply * sizeof(double));
  w.part_rr = (double*)b; b += align_up(lfr::kRedMaxBlocks * sizeof(double));
  w.part_rz = (double*)b; b += align_up(lfr::kRedMaxBlocks * sizeof(double));
  w.S = (double*)b;
  w.napply = napply;
  return LFR_OK;
}

template <typename T, int NDF_IN, int NG, int GS>
void launch_apply_t(hipStream_t s, const lfr_plan_class& c, const T* M, const T* x, T* y,
                    const int32_t* mi, int nl, int64_t n3, int ndf_out, int ndf_in,
                    const T* dotv, double* part, const double* stop) {
  const unsigned nb = lfr_blocks((int64_t)c.ncol * c.height, kThreads);
  const unsigned nbp = (nb + 7u) & ~7u;
  if (dotv)
    hipLaunchKernelGGL((lfr::k_gather<T, NDF_IN, NG, GS, 0, true>), dim3(nbp), dim3(kThreads), 0, s,
                       M, x, y, mi, c.col_start_d, c.ent_d, c.ncol, c.height, c.ngroup, nl, n3,
                       ndf_out, ndf_in, 1, nbp, dotv, part, stop);
  else
    hipLaunchKernelGGL((lfr::k_gather<T, NDF_IN, NG, GS, 0, false>), dim3(nbp), dim3(kThreads), 0,
                       s, M, x, y, mi, c.col_start_d, c.ent_d, c.ncol, c.height, c.ngroup, nl, n3,
                       ndf_out, ndf_in, 1, nbp, (const T*)nullptr, (double*)nullptr,
                       (const double*)nullptr);
}

template <typename T, int NDF_IN, int NG, int GS>
void launch_apply_rec_t(hipStream_t s, const lfr_plan_class& c, const int32_t* rec, const T* M,
                        const T* x, T* y, int nl, int64_t n3, int ndf_out, int64_t pcs, const T* dotv,
                        double* part, const double* stop) {
  const unsigned nb = lfr_blocks((int64_t)c.ncol * c.height, kThreads);
  const unsigned nbp = (nb + 7u) & ~7u;
  const bool nt = lfr_gather_nt((int)sizeof(T));
  if (dotv) {
    if (nt)
      hipLaunchKernelGGL((lfr::k_gather_rec<T, NDF_IN, NG, GS, 0, true, true>), dim3(nbp), dim3(kThreads),
                         0, s, M, x, y, rec, c.ncol, c.height, nl, n3, ndf_out, 1, pcs, nbp, dotv, part, stop);
    else
      hipLaunchKernelGGL((lfr::k_gather_rec<T, NDF_IN, NG, GS, 0, true, false>), dim3(nbp), dim3(kThreads),
                         0, s, M, x, y, rec, c.ncol, c.height, nl, n3, ndf_out, 1, pcs, nbp, dotv, part, stop);
  } else {
    if (nt)
      hipLaunchKernelGGL((lfr::k_gather_rec<T, NDF_IN, NG, GS, 0, false, true>), dim3(nbp), dim3(kThreads),
                         0, s, M, x, y, rec, c.ncol, c.height, nl, n3, ndf_out, 1, pcs, nbp,
                         (const T*)nullptr, (double*)nullptr, (const double*)nullptr);
    else
      hipLaunchKernelGGL((lfr::k_gather_rec<T, NDF_IN, NG, GS, 0, false, false>), dim3(nbp), dim3(kThreads),
                         0, s, M, x, y, rec, c.ncol, c.height, nl, n3, ndf_out, 1, pcs, nbp,
                         (const T*)nullptr, (double*)nullptr, (const double*)nullptr);
  }
}

template <typename T, int NDF_IN, int NG, int GS, int LPT, bool NT>
void launch_apply_vec_nt(hipStream_t s, const lfr_plan_class& c, const int32_t* rec, const T* M, const T* x, T* y,
                         int nl, int64_t n3, int ndf_out, int64_t pcs, const T* dotv, double* part,
                         const double* stop) {
  const unsigned nbp = (lfr_class_blocks(c, LPT) + 7u) & ~7u;
  if (dotv)
    hipLaunchKernelGGL((lfr::k_gather_vec<T, NDF_IN, NG, GS, 0, true, LPT, NT>), dim3(nbp), dim3(kThreads), 0, s, M,
                       x, y, rec, c.ncol, nl, n3, ndf_out, 1, pcs, nbp, dotv, part, stop);
  else
    hipLaunchKernelGGL((lfr::k_gather_vec<T, NDF_IN, NG, GS, 0, false, LPT, NT>), dim3(nbp), dim3(kThreads), 0, s, M,
                       x, y, rec, c.ncol, nl, n3, ndf_out, 1, pcs, nbp, (const T*)nullptr, (double*)nullptr,
                       (const double*)nullptr);
}

template <typename T, int NDF_IN, int NG, int GS, int LPT>
void launch_apply_vec_t(hipStream_t s, const lfr_plan_class& c, const int32_t* rec, const T* M, const T* x, T* y,
                        int nl, int64_t n3, int ndf_out, int64_t pcs, const T* dotv, double* part,
                        const double* stop) {
  if (lfr_gather_vec_nt(pcs > 0))
    launch_apply_vec_nt<T, NDF_IN, NG, GS, LPT, true>(s, c, rec, M, x, y, nl, n3, ndf_out, pcs, dotv, part, stop);
  else
    launch_apply_vec_nt<T, NDF_IN, NG, GS, LPT, false>(s, c, rec, M, x, y, nl, n3, ndf_out, pcs, dotv, part, stop);
}

template <typename T, int NDF_IN, int LPT>
bool launch_apply_vec(hipStream_t s, const lfr_plan_class& c, const int32_t* rec, const T* M, const T* x, T* y,
                      int nl, int64_t n3, int ndf, int64_t pcs, const T* dotv, double* part, const double* stop) {
#define SHAPE(NGV, GSV)                                                                                   \
  if (c.ngroup == NGV && c.gsize == GSV) {                                                                \
    launch_apply_vec_t<T, NDF_IN, NGV, GSV, LPT>(s, c, rec, M, x, y, nl, n3, ndf, pcs, dotv, part, stop); \
    return true;                                                                                          \
  }
  SHAPE(2, 1) SHAPE(1, 1) SHAPE(4, 1) SHAPE(3, 1) SHAPE(1, 2) SHAPE(2, 2) SHAPE(3, 2) SHAPE(4, 2)
#undef SHAPE
  return false;
}

template <typename T, int NDF_IN>
bool launch_apply_rec(hipStream_t s, const lfr_plan_class& c, const int32_t* rec, const T* M,
                      const T* x, T* y, int nl, int64_t n3, int ndf, int64_t pcs, const T* dotv, double* part,
                      const double* stop) {
#define SHAPE(NGV, GSV) \
  if (c.ngroup == NGV && c.gsize == GSV) { \
    launch_apply_rec_t<T, NDF_IN, NGV, GSV>(s, c, rec, M, x, y, nl, n3, ndf, pcs, dotv, part, stop); \
    return true; }
  SHAPE(2, 1) SHAPE(1, 2) SHAPE(1, 1) SHAPE(4, 2) SHAPE(3, 2) SHAPE(2, 2) SHAPE(4, 1) SHAPE(3, 1)
#undef SHAPE
  return false;
}

template <typename T, int NDF_IN>
void launch_apply_shape(hipStream_t s, const lfr_plan_class& c, const T* M, const T* x, T* y,
                        const int32_t* mi, int nl, int64_t n3, int ndf, const T* dotv,
                        double* part, const double* stop) {
#define SHAPE(NGV, GSV) \
  if (c.ngroup == NGV && c.gsize == GSV) \
    return launch_apply_t<T, NDF_IN, NGV, GSV>(s, c, M, x, y, mi, nl, n3, ndf, ndf, dotv, part, stop);
  SHAPE(2, 1) SHAPE(1, 2) SHAPE(1, 1) SHAPE(4, 2) SHAPE(3, 2) SHAPE(2, 2) SHAPE(4, 1) SHAPE(3, 1)
#undef SHAPE
  if (c.gsize == 1)
    launch_apply_t<T, NDF_IN, 0, 1>(s, c, M, x, y, mi, nl, n3, ndf, ndf, dotv, part, stop);
  else
    launch_apply_t<T, NDF_IN, 0, 2>(s, c, M, x, y, mi, nl, n3, ndf, ndf, dotv, part, stop);
}

template <typename T, int NDF_IN, int LPT, bool NT, bool ST>
void launch_apply_uni_st(hipStream_t s, unsigned nbp, const T* M, const T* x, T* y, const int32_t* ur, int ncol,
                         int nl, int64_t n3, int64_t pcs, const T* dotv, double* part, const double* stop) {
  if (dotv)
    hipLaunchKernelGGL((lfr::k_gather_uni<T, NDF_IN, 0, true, LPT, NT, ST>), dim3(nbp), dim3(kThreads), 0, s, M, x,
                       y, ur, ncol, nl, n3, NDF_IN, 1, pcs, nbp, dotv, part, stop);
  else
    hipLaunchKernelGGL((lfr::k_gather_uni<T, NDF_IN, 0, false, LPT, NT, ST>), dim3(nbp), dim3(kThreads), 0, s, M, x,
                       y, ur, ncol, nl, n3, NDF_IN, 1, pcs, nbp, (const T*)nullptr, (double*)nullptr,
                       (const double*)nullptr);
}

template <typename T, int NDF_IN, int LPT, bool NT>
void launch_apply_uni_nt(hipStream_t s, bool st, unsigned nbp, const T* M, const T* x, T* y, const int32_t* ur,
                         int ncol, int nl, int64_t n3, int64_t pcs, const T* dotv, double* part, const double* stop) {
  if (st)
    launch_apply_uni_st<T, NDF_IN, LPT, NT, true>(s, nbp, M, x, y, ur, ncol, nl, n3, pcs, dotv, part, stop);
  else
    launch_apply_uni_st<T, NDF_IN, LPT, NT, false>(s, nbp, M, x, y, ur, ncol, nl, n3, pcs, dotv, part, stop);
}

template <typename T, int NDF_IN, int LPT>
void launch_apply_uni(hipStream_t s, bool nt, bool st, unsigned nbp, const T* M, const T* x, T* y, const int32_t* ur,
                      int ncol, int nl, int64_t n3, int64_t pcs, const T* dotv, double* part, const double* stop) {
  if (nt) launch_apply_uni_nt<T, NDF_IN, LPT, true>(s, st, nbp, M, x, y, ur, ncol, nl, n3, pcs, dotv, part, stop);
  else launch_apply_uni_nt<T, NDF_IN, LPT, false>(s, st, nbp, M, x, y, ur, ncol, nl, n3, pcs, dotv, part, stop);
}

// y = A x (overwrite) by the gather kernels; with dotv: partials of
// <y, dotv> into part[0..napply)
template <typename T>
int apply_gather(lfr_ctx* ctx, const lfr_lma_op* A, const T* x, T* y, const T* dotv, double* part,
                 const double* stop) {
  // the record kernels form the fused inner product from the operand
  // registers, which holds for <A x, x> only
  LFR_REQUIRE(ctx, !dotv || dotv == x, "apply: fused inner product must be <A x, x>");
  const lfr_space* sp = A->space;
  const int nl = sp->nlayers;
  const int64_t n3 = (int64_t)sp->ncell * nl;
  if (apply_unified(A->plan, (int)sizeof(T), A->packed)) {
    const int32_t* ur = lfr_plan_urecords(ctx, A->plan, sp);
    if (!ur) return LFR_ERR_NOMEM;
    const int lpt = lfr_unified_lpt((int)sizeof(T));
    const unsigned nbp = (lfr_unified_blocks(A->plan, lpt) + 7u) & ~7u;
    const T* M = A->packed ? (const T*)A->packed->data_d : (const T*)A->matrix_d;
    const int64_t pcs = A->packed ? A->packed->pcs[0] : 0;
    const int ncol = lfr_plan_ncolumns_total(A->plan);
    const bool nt = A->packed ? lfr_pk_nt(A->packed) : lfr_gather_vec_nt(false);
    const bool st = A->packed ? lfr_pk_stage(A->packed) : lfr_gather_stage();
    if (sp->ndf == 6) {
      if (lpt == 4) launch_apply_uni<T, 6, 4>(ctx->stream, nt, st, nbp, M, x, y, ur, ncol, nl, n3, pcs, dotv, part, stop);
      else launch_apply_uni<T, 6, 2>(ctx->stream, nt, st, nbp, M, x, y, ur, ncol, nl, n3, pcs, dotv, part, stop);
    } else {
      if (lpt == 4) launch_apply_uni<T, 1, 4>(ctx->stream, nt, st, nbp, M, x, y, ur, ncol, nl, n3, pcs, dotv, part, stop);
      else launch_apply_uni<T, 1, 2>(ctx->stream, nt, st, nbp, M, x, y, ur, ncol, nl, n3, pcs, dotv, part, stop);
    }
    LFR_CHECK_LAUNCH(ctx);
    return LFR_OK;
  }
  int off = 0;
  for (size_t ci = 0; ci < A->plan->classes.size(); ++ci) {
    const auto& c = A->plan->classes[ci];
    const lfr_lma_packed* pk = A->packed;
    const int lpt = lfr_class_lpt(c, nl, (int)sizeof(T), sp->ndf, pk != nullptr);
    const unsigned nb = lfr_class_blocks(c, lpt);
    const unsigned nbp = (nb + 7u) & ~7u;
    double* pp = part ? part + off : nullptr;
    // packed operator: this class's block, columns pcs elements apart
    const T* M = pk ? (const T*)pk->data_d + pk->off[ci] : (const T*)A->matrix_d;
    const int64_t pcs = pk ? pk->pcs[ci] : 0;
    int stride = 0;
    const int32_t* rec = (pk || lfr_records_enabled()) ? lfr_plan_records(ctx, A->plan, ci, sp, &stride) : nullptr;
    if (pk && !rec) return lfr_set_error(ctx, LFR_ERR_NOMEM, "apply: column records unavailable");
    bool done = false;
    if (lpt > 1) {
      if (!rec) return lfr_set_error(ctx, LFR_ERR_NOMEM, "apply: column records unavailable");
      if (sp->ndf == 6) {
        done = lpt == 2 ? launch_apply_vec<T, 6, 2>(ctx->stream, c, rec, M, x, y, nl, n3, 6, pcs, dotv, pp, stop)
                        : launch_apply_vec<T, 6, 4>(ctx->stream, c, rec, M, x, y, nl, n3, 6, pcs, dotv, pp, stop);
      } else if (sp->ndf == 8) {
        done = lpt == 2 ? launch_apply_vec<T, 8, 2>(ctx->stream, c, rec, M, x, y, nl, n3, 8, pcs, dotv, pp, stop)
                        : launch_apply_vec<T, 8, 4>(ctx->stream, c, rec, M, x, y, nl, n3, 8, pcs, dotv, pp, stop);
      } else if (sp->ndf == 1) {
        done = lpt == 2 ? launch_apply_vec<T, 1, 2>(ctx->stream, c, rec, M, x, y, nl, n3, 1, pcs, dotv, pp, stop)
                        : launch_apply_vec<T, 1, 4>(ctx->stream, c, rec, M, x, y, nl, n3, 1, pcs, dotv, pp, stop);
      }
      if (!done) return lfr_set_error(ctx, LFR_ERR_ARG, "apply: no vectorised gather for this class shape");
    } else if (rec) {
      if (sp->ndf == 6) done = launch_apply_rec<T, 6>(ctx->stream, c, rec, M, x, y, nl, n3, 6, pcs, dotv, pp, stop);
      else if (sp->ndf == 8) done = launch_apply_rec<T, 8>(ctx->stream, c, rec, M, x, y, nl, n3, 8, pcs, dotv, pp, stop);
      else if (sp->ndf == 1) done = launch_apply_rec<T, 1>(ctx->stream, c, rec, M, x, y, nl, n3, 1, pcs, dotv, pp, stop);
    }
    if (pk && !done) return lfr_set_error(ctx, LFR_ERR_ARG, "apply: no packed kernel for this class shape");
    if (done) {
      LFR_CHECK_LAUNCH(ctx);
      off += (int)nbp;
      continue;
    }
switch (sp->ndf) {
      case 6: launch_apply_shape<T, 6>(ctx->stream, c, M, x, y, sp->map0_d, nl, n3, sp->ndf, dotv, pp, stop); break;
      case 8: launch_apply_shape<T, 8>(ctx->stream, c, M, x, y, sp->map0_d, nl, n3, sp->ndf, dotv, pp, stop); break;
      default:
        if (c.gsize == 1)
          launch_apply_t<T, 0, 0, 1>(ctx->stream, c, M, x, y, sp->map0_d, nl, n3, sp->ndf, sp->ndf, dotv, pp, stop);
        else
          launch_apply_t<T, 0, 0, 2>(ctx->stream, c, M, x, y, sp->map0_d, nl, n3, sp->ndf, sp->ndf, dotv, pp, stop);
    }
    LFR_CHECK_LAUNCH(ctx);
    off += (int)nbp;
  }
  return LFR_OK;
}

// Partial sums of <a, b> over a grid of exactly nblk blocks (part[0..nblk)).
template <typename T>
__global__ void __launch_bounds__(kThreads)
k_dot_partials(const T* __restrict__ a, const T* __restrict__ b, int64_t n, double* __restrict__ part,
               const double* __restrict__ stop) {
  if (stop && stop[0] != 0.0) return;
  __shared__ double lds[kThreads / 64];
  double v = 0.0;
  for (int64_t i = (int64_t)blockIdx.x * kThreads + threadIdx.x; i < n; i += (int64_t)gridDim.x * kThreads)
    v += (double)a[i] * (double)b[i];
  v = lfr::block_sum(v, lds);
  if (threadIdx.x == 0) part[blockIdx.x] = v;
}

// y = A x as the operator object applies it: the gather apply, then
// enforce_bc when the operator asks for it (the fused dot product is then
// formed after the mask, over the same napply partial slots).
template <typename T>
int apply_op(lfr_ctx* ctx, const lfr_lma_op* A, const T* x, T* y, const T* dotv, double* part,
             const double* stop) {
  if (!A->apply_bc) return apply_gather<T>(ctx, A, x, y, dotv, part, stop);
  int rc = apply_gather<T>(ctx, A, x, y, (const T*)nullptr, nullptr, stop);
  if (rc) return rc;
  rc = lfr_enforce_bc_queue(ctx, (int)sizeof(T), y, A->space);
  if (rc) return rc;
  if (dotv) {
    const int np = apply_partials(A->plan, (int)sizeof(T), A->packed != nullptr);
    hipLaunchKernelGGL(k_dot_partials<T>, dim3(np), dim3(kThreads), 0, ctx->stream, (const T*)y, dotv,
                       (int64_t)A->space->undf, part, stop);
    LFR_CHECK_LAUNCH(ctx);
  }
  return LFR_OK;
}

}  // namespace

int lfr_lma_apply_dot_queue(lfr_ctx* ctx, const lfr_lma_op* A, const void* x, void* y, double* part,
                            int* npart) {
  *npart = apply_partials(A->plan, A->prec, A->packed != nullptr);
  if (A->prec == 4) return apply_op<float>(ctx, A, (const float*)x, (float*)y, (const float*)x, part, nullptr);
  return apply_op<double>(ctx, A, (const double*)x, (double*)y, (const double*)x, part, nullptr);
}

int lfr_lma_apply_partials_max(const lfr_plan* plan) { return apply_partials_max(plan); }

int lfr_lma_apply_queue(lfr_ctx* ctx, const lfr_lma_op* A, const void* x, void* y) {
  if (A->prec == 4) return apply_op<float>(ctx, A, (const float*)x, (float*)y, nullptr, nullptr, nullptr);
  return apply_op<double>(ctx, A, (const double*)x, (double*)y, nullptr, nullptr, nullptr);
}

int lfr_check_op(lfr_ctx* ctx, const lfr_lma_op* A, const char* who) {
  LFR_REQUIRE(ctx, A && A->matrix_d && A->space && A->plan, "%s: incomplete operator", who);
  LFR_REQUIRE(ctx, A->prec == 4 || A->prec == 8, "%s: prec %d", who, A->prec);
  LFR_REQUIRE(ctx, A->plan->space == A->space && A->plan->kind == 0,
              "%s: operator plan must be a matrix_vector plan of its space", who);
  LFR_REQUIRE(ctx, !A->apply_bc || (int)A->space->boundary_h.size() == 2 * A->space->ndf,
              "%s: apply_bc set but the space has no boundary values", who);
  if (A->packed) {
    const lfr_lma_packed* pk = A->packed;
    LFR_REQUIRE(ctx, pk->ctx == ctx && pk->plan == A->plan && pk->in == A->space && pk->prec == A->prec &&
                         pk->kind == 0,
                "%s: packed operator was not packed from this operator's plan, space and precision", who);
  }
  // build the column records now: never inside a captured graph
  if (apply_unified(A->plan, A->prec, A->packed)) {
    ctx->err.clear();
    if (!lfr_plan_urecords(ctx, A->plan, A->space)) return LFR_ERR_NOMEM;
  } else if (A->packed || lfr_records_enabled()) {
    int stride = 0;
    ctx->err.clear();
    for (size_t ci = 0; ci < A->plan->classes.size(); ++ci)
      (void)lfr_plan_records(ctx, A->plan, ci, A->space, &stride);
    if (!ctx->err.empty()) return LFR_ERR_NOMEM;
  }
  return LFR_OK;
}

namespace {

bool use_graphs() {
  const char* e = getenv("LFR_NO_GRAPH");
  return !(e && e[0] == '1');
}

template <typename T>
int cg_iterations(lfr_ctx* ctx, const lfr_lma_op* A, const lfr_solver_params* P, T* x, Ws& w,
                  int64_t n, int chunk, bool sync_end) {
  const bool diag = P->preconditioner == 1;
  T* r = (T*)w.r;
  T* p = (T*)w.p;
  T* z = (T*)w.z;
  const T* dg = (const T*)P->diag_d;
  const int G = lfr::red_blocks(n);
  const int np_apply = apply_partials(A->plan, (int)sizeof(T), A->packed != nullptr);
  auto one_iteration = [&]() -> int {
    int rc = apply_op<T>(ctx, A, p, z, p, w.part_apply, w.S);
    if (rc) return rc;
    hipLaunchKernelGGL(k_cg_alpha, dim3(1), dim3(kAlphaThreads), 0, ctx->stream, w.part_apply,
                       np_apply, diag ? w.part_rz : w.part_rr, G, w.S, P->max_iter);
    if (diag)
      hipLaunchKernelGGL((k_cg_update<T, true>), dim3(G), dim3(kThreads), 0, ctx->stream, x, r, p,
                         z, dg, n, w.S, w.part_rr, w.part_rz);
    else
      hipLaunchKernelGGL((k_cg_update<T, false>), dim3(G), dim3(kThreads), 0, ctx->stream, x, r,
                         p, z, dg, n, w.S, w.part_rr, w.part_rz);
    hipLaunchKernelGGL(k_cg_beta_direction<T>, dim3(G), dim3(kThreads), 0, ctx->stream, p,
                       diag ? (const T*)z : (const T*)r, n, w.part_rr, w.part_rz, G, w.S,
                       P->monitor_convergence, P->r_tol, P->a_tol, (int)diag);
    LFR_CHECK_LAUNCH(ctx);
    return LFR_OK;
  };

  // The captured chunk depends only on the pointers and scalars below, so
  // its instantiation is kept on the context and reused by later solves
  // with the same operator, vectors and settings.
  hipGraphExec_t exec = nullptr;
  if (use_graphs()) {
    uint64_t rt, at;
    memcpy(&rt, &P->r_tol, 8);
    memcpy(&at, &P->a_tol, 8);
    std::vector<uint64_t> key = {
        0xC6ull, sizeof(T), (uint64_t)(uintptr_t)A->matrix_d, A->packed ? A->packed->uid : 0,
        A->plan->uid, A->space->uid, (uint64_t)A->apply_bc,
        (uint64_t)(uintptr_t)x, (uint64_t)(uintptr_t)r, (uint64_t)(uintptr_t)p, (uint64_t)(uintptr_t)z,
        (uint64_t)(uintptr_t)w.part_apply, (uint64_t)(uintptr_t)w.part_rr, (uint64_t)(uintptr_t)w.part_rz,
        (uint64_t)(uintptr_t)w.S, (uint64_t)(uintptr_t)dg, (uint64_t)P->max_iter,
        (uint64_t)P->monitor_convergence, rt, at, (uint64_t)diag, (uint64_t)chunk, (uint64_t)n};
    // enforce_bc takes the boundary values by value: they are part of the graph
    if (A->apply_bc)
      for (int32_t v : A->space->boundary_h) key.push_back((uint64_t)(uint32_t)v);
    for (size_t i = 0; i < ctx->graphs.size(); ++i)
      if (ctx->graphs[i].key == key) {
        lfr_graph_entry e = ctx->graphs[i];
        ctx->graphs.erase(ctx->graphs.begin() + i);
        ctx->graphs.push_back(e);
        exec = e.exec;
        break;
      }
    if (!exec) {
      hipGraph_t graph = nullptr;
      if (hipStreamBeginCapture(ctx->stream, hipStreamCaptureModeRelaxed) == hipSuccess) {
        int rc = LFR_OK;
        for (int i = 0; i < chunk && rc == LFR_OK; ++i) rc = one_iteration();
        hipError_t e = hipStreamEndCapture(ctx->stream, &graph);
        if (rc != LFR_OK) {
          if (graph) (void)hipGraphDestroy(graph);
          return rc;
        }
        if (e == hipSuccess && graph) {
          if (hipGraphInstantiate(&exec, graph, nullptr, nullptr, 0) != hipSuccess) exec = nullptr;
          (void)hipGraphDestroy(graph);
        }
      }
      (void)hipGetLastError();
      if (exec) {
        constexpr size_t kKeep = 8;
        if (ctx->graphs.size() >= kKeep) {
          (void)hipGraphExecDestroy(ctx->graphs.front().exec);
          ctx->graphs.erase(ctx->graphs.begin());
        }
        lfr_graph_entry e;
        e.key = key;
        e.deps = {A->plan->uid, A->space->uid, A->packed ? A->packed->uid : 0};
        e.exec = exec;
        ctx->graphs.push_back(e);
      }
    }
  }
  auto run_chunk = [&]() -> int {
    if (exec) {
      LFR_HIP(ctx, hipGraphLaunch(exec, ctx->stream));
    } else {
      for (int i = 0; i < chunk; ++i) {
        int rc = one_iteration();
        if (rc) return rc;
      }
    }
    return LFR_OK;
  };
  double* Sh = ctx->host_scalars;
  if (!P->monitor_convergence) {
    // a fixed number of iterations: all chunks, then one read-back
    // (iterations past max_iter return at entry)
    for (int done = 0; done < P->max_iter; done += chunk) {
      int rc = run_chunk();
      if (rc) return rc;
    }
  } else {
    // While the host checks chunk k, chunk k+1 is normally queued behind it
    // so that the device never waits for the host.  A chunk that starts
    // after convergence returns at entry, but each of its launches still
    // costs a few microseconds, so chunk k+1 is queued ahead only while the
    // residual history (the geometric rate over the last checked chunk)
    // says chunk k will not converge; a wrong guess costs one host round
    // trip instead of a chunk of empty launches.  The decision changes only
    // when work is queued, never which iterations run.
    auto queue = [&](int slot) -> int {
      int rc = run_chunk();
      if (rc) return rc;
      LFR_HIP(ctx, hipMemcpyAsync(Sh + 16 + 16 * slot, w.S, S_NSLOT * sizeof(double), hipMemcpyDeviceToHost,
                                  ctx->stream));
      LFR_HIP(ctx, hipEventRecord(ctx->ev[slot], ctx->stream));
      return LFR_OK;
    };
    int rc = queue(0);
    if (rc) return rc;
    double it_a = 0.0, rn_a = -1.0, it_b = 0.0, rn_b = -1.0, target = 0.0;
    for (int k = 0;; ++k) {
      const int cur = k & 1, nxt = (k + 1) & 1;
      bool ahead = true;
      if (rn_a > 0.0 && rn_b > 0.0 && it_b > it_a) {
        const double rho = pow(rn_b / rn_a, 1.0 / (it_b - it_a));
        if (rho < 1.0 && rn_b * pow(rho, (double)chunk) <= target) ahead = false;
      }
      if (ahead && (rc = queue(nxt))) return rc;
      LFR_HIP(ctx, hipEventSynchronize(ctx->ev[cur]));
      const double* s = Sh + 16 + 16 * cur;
      if (s[S_STOP] != 0.0) {
        // iterations after the stop return at entry and leave S as it is,
        // so this copy is the final state
        memcpy(Sh, s, S_NSLOT * sizeof(double));
        break;
      }
      if (!ahead && (rc = queue(nxt))) return rc;
      if (rn_b < 0.0) {
        rn_a = s[S_RNRM0];
        target = std::max(P->r_tol * rn_a, P->a_tol);
      } else {
        it_a = it_b;
        rn_a = rn_b;
      }
      it_b = s[S_ITER];
      rn_b = s[S_RNRM];
    }
    if (sync_end) LFR_HIP(ctx, hipStreamSynchronize(ctx->stream));
    return LFR_OK;
  }
  LFR_HIP(ctx, hipMemcpyAsync(Sh, w.S, S_NSLOT * sizeof(double), hipMemcpyDeviceToHost, ctx->stream));
  LFR_HIP(ctx, hipStreamSynchronize(ctx->stream));
  return LFR_OK;
}

template <typename T>
int cg_solve(lfr_ctx* ctx, const lfr_lma_op* A, const lfr_solver_params* P, T* x, const T* b,
             lfr_solver_stats* st, Ws& w, bool sync_end = true, bool started = false) {
  const int64_t n = A->space->undf;
  const int G = lfr::red_blocks(n);
  T* r = (T*)w.r;
  T* p = (T*)w.p;
  T* z = (T*)w.z;
  const bool diag = P->preconditioner == 1;
  // S = 0
  hipLaunchKernelGGL(k_fill<double>, dim3(1), dim3(kThreads), 0, ctx->stream, w.S, (int64_t)S_NSLOT, 0.0);
  // r = b - A x
  if (P->x_is_zero) {
    // x = 0, so A x = 0 exactly: no apply; r = b, and p = r right away
    // without a preconditioner ('started': the caller's pass formed x, r, p
    // and the partials already)
    if (!started)
      hipLaunchKernelGGL(k_cg_start_zero<T>, dim3(G), dim3(kThreads), 0, ctx->stream, x, r, diag ? (T*)nullptr : p,
                       b, n, w.part_rr);
  } else {
    int rc = apply_op<T>(ctx, A, x, r, (const T*)nullptr, nullptr, nullptr);
    if (rc) return rc;
    hipLaunchKernelGGL(k_residual<T>, dim3(G), dim3(kThreads), 0, ctx->stream, r, b, n,
                       w.part_rr);
  }
  hipLaunchKernelGGL(k_store_sum, dim3(1), dim3(kThreads), 0, ctx->stream, w.part_rr, G, w.S,
                     (int)S_RNRM0, 1);
  hipLaunchKernelGGL(k_cg_start, dim3(1), dim3(64), 0, ctx->stream, w.S, kEps);
  LFR_CHECK_LAUNCH(ctx);
  double* Sh = ctx->host_scalars;
  // z = P^{-1} r ; rz = <r, z> ; p = z
  if (diag) {
    // (rz itself is summed by the first k_cg_alpha from these partials)
    hipLaunchKernelGGL(k_prec_diag<T>, dim3(G), dim3(kThreads), 0, ctx->stream, z, r,
                       (const T*)P->diag_d, n, w.part_rz);
    hipLaunchKernelGGL(k_copy_v<T>, dim3(G), dim3(kThreads), 0, ctx->stream, p, (const T*)z, n);
  } else if (!P->x_is_zero) {
    hipLaunchKernelGGL(k_copy_v<T>, dim3(G), dim3(kThreads), 0, ctx->stream, p, (const T*)r, n);
  }
  LFR_CHECK_LAUNCH(ctx);
  if (P->max_iter > 0) {
    // iterations per captured chunk: a monitored solve keeps one chunk queued
    // while the host checks the previous one, so after convergence up to
    // ~1.5 chunks of iterations return at entry; 4 keeps that to a few
    // launches while a chunk still outlasts the host's check
    int cmax = P->monitor_convergence ? 4 : 10;
    if (const char* e = getenv("LFR_CG_CHUNK")) {   // experiments: iterations per captured chunk
      const int v = atoi(e);
      if (v >= 1 && v <= 64) cmax = v;
    }
    const int chunk = P->max_iter < cmax ? P->max_iter : cmax;
    int rc = cg_iterations<T>(ctx, A, P, x, w, n, chunk, sync_end);  // ends with S in Sh
    if (rc) return rc;
  } else {
    LFR_HIP(ctx, hipMemcpyAsync(Sh, w.S, S_NSLOT * sizeof(double), hipMemcpyDeviceToHost, ctx->stream));
    LFR_HIP(ctx, hipStreamSynchronize(ctx->stream));
  }
  st->r_nrm0 = Sh[S_RNRM0];
  st->iterations = (int)Sh[S_ITER];
  st->r_nrm = Sh[S_RNRM];
  st->converged = Sh[S_STOP] == 1.0 ? 1 : 0;
  if (Sh[S_STOP] == 3.0)
    return lfr_set_error(ctx, LFR_ERR_NAN, "cg: NaN residual after %d iterations", st->iterations);
  if (P->monitor_convergence && !st->converged && P->fail_on_non_converged)
    return lfr_set_error(ctx, LFR_ERR_NOT_CONVERGED,
                         "cg failed to converge after %d iterations (||r||/||r0|| = %g)",
                         st->iterations, st->r_nrm / st->r_nrm0);
  return LFR_OK;
}

}  // namespace

extern "C" int lfr_cg(lfr_ctx* ctx, const lfr_lma_op* A, const lfr_solver_params* P, void* x,
                      const void* b, lfr_solver_stats* st) {
  if (!ctx) return LFR_ERR_ARG;
  int rc = lfr_check_op(ctx, A, "lfr_cg");
  if (rc) return rc;
  LFR_REQUIRE(ctx, P && x && b && st, "lfr_cg: null argument");
  LFR_REQUIRE(ctx, P->preconditioner == 0 || (P->preconditioner == 1 && P->diag_d),
              "lfr_cg: preconditioner must be 0 (null) or 1 (diagonal, with diag_d)");
  Ws w;
  rc = ws_get(ctx, A->space->undf, A->prec, A->plan, 0, w);
  if (rc) return rc;
  if (A->prec == 4) return cg_solve<float>(ctx, A, P, (float*)x, (const float*)b, st, w);
  return cg_solve<double>(ctx, A, P, (double*)x, (const double*)b, st, w);
}

// Mixed-precision iterative refinement: FP64 residual and solution, FP32
// inner CG (the r_solver split of the reference: r_solver operators and
// vectors inside an r_def outer loop).
extern "C" int lfr_refined_cg(lfr_ctx* ctx, const lfr_lma_op* A32, const lfr_lma_op* A64,
                              double r_tol, int max_outer, const lfr_solver_params* inner,
                              void* x64_v, const void* b64_v, lfr_refine_stats* st) {
  if (!ctx) return LFR_ERR_ARG;
  int rc = lfr_check_op(ctx, A32, "lfr_refined_cg(A32)");
  if (rc) return rc;
  rc = lfr_check_op(ctx, A64, "lfr_refined_cg(A64)");
  if (rc) return rc;
  LFR_REQUIRE(ctx, A32->prec == 4 && A64->prec == 8, "lfr_refined_cg: need fp32 and fp64 operators");
  LFR_REQUIRE(ctx, A32->space->undf == A64->space->undf, "lfr_refined_cg: operator sizes differ");
  LFR_REQUIRE(ctx, inner && x64_v && b64_v && st && max_outer > 0, "lfr_refined_cg: bad arguments");
  LFR_REQUIRE(ctx, inner->preconditioner == 0, "lfr_refined_cg: inner preconditioner must be null");
  const int64_t n = A64->space->undf;
  double* x = (double*)x64_v;
  const double* b = (const double*)b64_v;
  // workspace: fp64-sized slots; the inner fp32 CG uses r, p, z as fp32 views
  Ws w;
  rc = ws_get(ctx, n, 8, A32->plan, 2, w);
  if (rc) return rc;
  double* r64 = (double*)w.aux;    // fp64 A x (the residual pass does not store b - A x)
  float* d32 = (float*)w.aux2;     // fp32 correction
  const int G = lfr::red_blocks(n);
  double* Sh = ctx->host_scalars;
  st->outer = 0;
  st->inner_total = 0;
  st->converged = 0;
  // ||b||
  hipLaunchKernelGGL(k_fill<double>, dim3(1), dim3(kThreads), 0, ctx->stream, w.S, (int64_t)S_NSLOT, 0.0);
  {
    double bn = 0.0;
    rc = lfr_vec_dot(ctx, 8, n, b, b, &bn);
    if (rc) return rc;
    st->b_nrm = sqrt(bn);
  }
  for (int it = 0; it <= max_outer; ++it) {
    // r = b - A64 x in FP64 and ||r||, with the inner solve's start from
    // zero on (float) r in the same pass (r64 holds A64 x)
    rc = apply_op<double>(ctx, A64, x, r64, (const double*)nullptr, nullptr, nullptr);
    if (rc) return rc;
    hipLaunchKernelGGL(k_residual_start32, dim3(G), dim3(kThreads), 0, ctx->stream, (const double*)r64, b, n,
                       w.part_rz, d32, (float*)w.r, (float*)w.p, w.part_rr);
    hipLaunchKernelGGL(k_store_sum, dim3(1), dim3(kThreads), 0, ctx->stream, w.part_rz, G, w.S,
                       (int)S_RNRM, 1);
    LFR_CHECK_LAUNCH(ctx);
    LFR_HIP(ctx, hipMemcpyAsync(Sh, w.S, S_NSLOT * sizeof(double), hipMemcpyDeviceToHost, ctx->stream));
    LFR_HIP(ctx, hipStreamSynchronize(ctx->stream));
    st->r_nrm = Sh[S_RNRM];
    if (st->r_nrm != st->r_nrm)
      return lfr_set_error(ctx, LFR_ERR_NAN, "refined cg: NaN residual at outer iteration %d", it);
    if (st->r_nrm <= r_tol * st->b_nrm) {
      st->converged = 1;
      break;
    }
    if (it == max_outer) break;
    // inner FP32 CG from zero
    lfr_solver_stats ist;
    lfr_solver_params ip = *inner;
    ip.x_is_zero = 1;
    ip.fail_on_non_converged = 0;
    // no wait at its end: the correction and the next residual queue behind
    // it, and the residual's read-back waits for all of them
    rc = cg_solve<float>(ctx, A32, &ip, d32, (const float*)nullptr, &ist, w, false, true);
    if (rc) return rc;
    st->inner_total += ist.iterations;
    st->outer = it + 1;
    hipLaunchKernelGGL(k_add_correction, dim3(G), dim3(kThreads), 0, ctx->stream, x, d32, n);
    LFR_CHECK_LAUNCH(ctx);
  }
  if (!st->converged)
    return lfr_set_error(ctx, LFR_ERR_NOT_CONVERGED,
                         "refined cg: ||r||/||b|| = %g after %d outer iterations",
                         st->r_nrm / st->b_nrm, st->outer);
  return LFR_OK;
}

// ---------------------------------------------------------------------------
// GMRES(gcrk): gmres_type%gmres_solve
// (components/science/source/solver/sci_iterative_solver_mod.f90:830-1003).
// Left preconditioned, modified Gram-Schmidt Arnoldi.  Inner products are
// written straight into a device copy of the Hessenberg matrix and the
// orthogonalisation axpys read their coefficient from there, so an Arnoldi
// cycle never waits on the host; the (gcrk+1) x gcrk Hessenberg comes back
// once per restart for the Givens rotations and back substitution, done on
// the host in FP64 exactly as the reference writes them.
// ---------------------------------------------------------------------------
int lfr_dot_queue(lfr_ctx* ctx, int prec, int64_t n, const void* x, const void* y, double* part,
                  double* out_d);

namespace {

// y += (T)(sign * a[0]) * x
template <typename T>
__global__ void k_axpy_dev(T* __restrict__ y, const T* __restrict__ x, int64_t n,
                           const double* __restrict__ a, double sign) {
  const T c = (T)(sign * a[0]);
  for (int64_t i = (int64_t)blockIdx.x * kThreads + threadIdx.x; i < n; i += (int64_t)gridDim.x * kThreads) {
    const T p = c * x[i];
    y[i] = y[i] + p;
  }
}

// v = (T)(1 / a[0]) * s   (copy then scale(1/beta))
template <typename T>
__global__ void k_scale_inv_dev(T* __restrict__ v, const T* __restrict__ s, int64_t n,
                                const double* __restrict__ a) {
  const T c = (T)(1.0 / a[0]);
  for (int64_t i = (int64_t)blockIdx.x * kThreads + threadIdx.x; i < n; i += (int64_t)gridDim.x * kThreads)
    v[i] = c * s[i];
}

// Modified Gram-Schmidt step fused with the next inner product: w += (T)(-a[0])
// * vj (k_axpy_dev's arithmetic), then the first stage of <vn, w> (vn == w:
// <w, w>) over the updated w, in the grid and order of the dot kernel
// (red_blocks(n) blocks, interleaved, FP64 block sums), so both results
// have the bits of the separate kernels while w is read once.
template <typename T, bool SELF>
__global__ void __launch_bounds__(lfr::kRedThreads)
k_mgs_axpy_dot(T* __restrict__ w, const T* __restrict__ vj, const T* __restrict__ vn, int64_t n,
               const double* __restrict__ a, double* __restrict__ part) {
  __shared__ double lds[lfr::kRedThreads / 64];
  const T c = (T)(-1.0 * a[0]);
  double s = 0.0;
  for (int64_t i = (int64_t)blockIdx.x * lfr::kRedThreads + threadIdx.x; i < n;
       i += (int64_t)gridDim.x * lfr::kRedThreads) {
    const T p = c * vj[i];
    const T wn = w[i] + p;
    w[i] = wn;
    const double q = SELF ? (double)wn * (double)wn : (double)vn[i] * (double)wn;
    s = s + q;
  }
  s = lfr::block_sum(s, lds);
  if (threadIdx.x == 0) part[blockIdx.x] = s;
}

__global__ void k_sqrt_slot(double* __restrict__ v) {
  if (threadIdx.x == 0 && blockIdx.x == 0) v[0] = sqrt(v[0]);
}

template <typename T>
__global__ void k_divide_v(T* __restrict__ z, const T* __restrict__ x, const T* __restrict__ d, int64_t n) {
  for (int64_t i = (int64_t)blockIdx.x * kThreads + threadIdx.x; i < n; i += (int64_t)gridDim.x * kThreads)
    z[i] = x[i] / d[i];
}

template <typename T>
__global__ void k_axpy_val(T* __restrict__ y, const T* __restrict__ x, int64_t n, T a) {
  for (int64_t i = (int64_t)blockIdx.x * kThreads + threadIdx.x; i < n; i += (int64_t)gridDim.x * kThreads) {
    const T p = a * x[i];
    y[i] = y[i] + p;
  }
}

template <typename T>
int gmres_solve(lfr_ctx* ctx, const lfr_lma_op* A, const lfr_solver_params* P, T* x, const T* b,
                lfr_solver_stats* st) {
  const int64_t n = A->space->undf;
  const int m = P->gcrk;
  const int G = lfr::red_blocks(n);
  const size_t vb = align_up((size_t)n * sizeof(T));
  const size_t hb = align_up((size_t)((m + 1) * m + 4) * sizeof(double));
  const size_t head = align_up((lfr::kRedMaxBlocks + 8) * sizeof(double));
  const size_t part_b = align_up(lfr::kRedMaxBlocks * sizeof(double));
  int rc = lfr_work_reserve(ctx, head + (size_t)(m + 3) * vb + hb + part_b);
  if (rc) return rc;
  char* base = (char*)ctx->work + head;
  T* V = (T*)base;                                   // m vectors
  T* w = (T*)(base + (size_t)m * vb);
  T* s = (T*)(base + (size_t)(m + 1) * vb);
  T* res = (T*)(base + (size_t)(m + 2) * vb);
  double* H = (double*)(base + (size_t)(m + 3) * vb);   // h(i,j) at H[i + (m+1)*j]; H[(m+1)*m] = beta
  double* part = (double*)((char*)H + hb);
  double* beta_d = H + (m + 1) * m;
  const bool diag = P->preconditioner == 1;
  const T* dg = (const T*)P->diag_d;
  auto Vv = [&](int i) { return V + (size_t)i * (vb / sizeof(T)); };
  auto prec = [&](T* out, const T* in) {
    if (diag)
      hipLaunchKernelGGL(k_divide_v<T>, dim3(G), dim3(kThreads), 0, ctx->stream, out, in, dg, n);
    else
      hipLaunchKernelGGL(k_copy_v<T>, dim3(G), dim3(kThreads), 0, ctx->stream, out, in, n);
  };
  auto norm_host = [&](const T* v, double* out) -> int {
    int r = lfr_dot_queue(ctx, sizeof(T), n, v, v, part, beta_d + 1);
    if (r) return r;
    LFR_HIP(ctx, hipMemcpyAsync(ctx->host_scalars, beta_d + 1, sizeof(double), hipMemcpyDeviceToHost, ctx->stream));
    LFR_HIP(ctx, hipStreamSynchronize(ctx->stream));
    *out = sqrt(ctx->host_scalars[0]);
    return LFR_OK;
  };
  // res = b - A x  (res%copy(b); res%axpy(-1, Ax))
  rc = apply_op<T>(ctx, A, x, res, (const T*)nullptr, nullptr, nullptr);
  if (rc) return rc;
  hipLaunchKernelGGL(k_residual<T>, dim3(G), dim3(kThreads), 0, ctx->stream, res, b, n, part);
  double sc_err = 0;
  rc = norm_host(res, &sc_err);
  if (rc) return rc;
  st->r_nrm0 = sc_err;
  st->r_nrm = sc_err;
  st->iterations = 0;
  st->converged = 0;
  if (sc_err < kEps) {
    st->converged = 1;
    return LFR_OK;
  }
  if (P->monitor_convergence) sc_err = sc_err > P->a_tol ? sc_err : P->a_tol;
  std::vector<double> h((size_t)(m + 1) * m), g(m + 1), u(m);
  double err = 0;
  int it;
  for (it = 1; it <= P->max_iter; ++it) {
    prec(s, res);
    rc = lfr_dot_queue(ctx, sizeof(T), n, s, s, part, beta_d);
    if (rc) return rc;
    hipLaunchKernelGGL(k_sqrt_slot, dim3(1), dim3(1), 0, ctx->stream, beta_d);
    hipLaunchKernelGGL(k_scale_inv_dev<T>, dim3(G), dim3(kThreads), 0, ctx->stream, Vv(0), s, n, beta_d);
    for (int iv = 0; iv < m; ++iv) {
      // w = prec(A v_iv): with the null preconditioner the apply writes w
      // directly (the reference's copy gives the same values)
      if (diag) {
        rc = apply_op<T>(ctx, A, Vv(iv), s, (const T*)nullptr, nullptr, nullptr);
        if (rc) return rc;
        prec(w, s);
      } else {
        rc = apply_op<T>(ctx, A, Vv(iv), w, (const T*)nullptr, nullptr, nullptr);
        if (rc) return rc;
      }
      // h(0, iv) = <v_0, w>; then for each j: w -= h(j, iv) v_j fused with
      // h(j+1, iv) = <v_{j+1}, w> (j < iv) or <w, w> (j == iv)
      double* hn = H + (iv + 1) + (size_t)(m + 1) * iv;
      rc = lfr_dot_queue(ctx, sizeof(T), n, Vv(0), w, part, H + (size_t)(m + 1) * iv);
      if (rc) return rc;
      for (int j = 0; j <= iv; ++j) {
        const double* hij = H + j + (size_t)(m + 1) * iv;
        double* out = j < iv ? H + (j + 1) + (size_t)(m + 1) * iv : hn;
        if (j < iv)
          hipLaunchKernelGGL((k_mgs_axpy_dot<T, false>), dim3(G), dim3(lfr::kRedThreads), 0, ctx->stream, w,
                             Vv(j), Vv(j + 1), n, hij, part);
        else
          hipLaunchKernelGGL((k_mgs_axpy_dot<T, true>), dim3(G), dim3(lfr::kRedThreads), 0, ctx->stream, w,
                             Vv(j), (const T*)nullptr, n, hij, part);
        hipLaunchKernelGGL(lfr::k_reduce_final, dim3(1), dim3(lfr::kRedThreads), 0, ctx->stream, part, G, 0, out);
      }
      hipLaunchKernelGGL(k_sqrt_slot, dim3(1), dim3(1), 0, ctx->stream, hn);
      if (iv < m - 1)
        hipLaunchKernelGGL(k_scale_inv_dev<T>, dim3(G), dim3(kThreads), 0, ctx->stream, Vv(iv + 1), w, n, hn);
      LFR_CHECK_LAUNCH(ctx);
    }
    // Hessenberg and beta to the host
    std::vector<double> hb_((size_t)(m + 1) * m + 1);
    LFR_HIP(ctx, hipMemcpyAsync(hb_.data(), H, hb_.size() * sizeof(double), hipMemcpyDeviceToHost, ctx->stream));
    LFR_HIP(ctx, hipStreamSynchronize(ctx->stream));
    auto hh = [&](int i, int j) -> double& { return hb_[(size_t)i + (size_t)(m + 1) * j]; };
    std::fill(g.begin(), g.end(), 0.0);
    g[0] = hb_[(size_t)(m + 1) * m];
    // Givens rotations (:930-944)
    for (int iv = 0; iv < m; ++iv) {
      const double nrm = sqrt(hh(iv, iv) * hh(iv, iv) + hh(iv + 1, iv) * hh(iv + 1, iv));
      const double si = hh(iv + 1, iv) / nrm, ci = hh(iv, iv) / nrm;
      const double p_ = ci * g[iv] + si * g[iv + 1];
      const double q_ = -si * g[iv] + ci * g[iv + 1];
      g[iv] = p_;
      g[iv + 1] = q_;
      for (int j = iv; j < m; ++j) {
        const double h1 = ci * hh(iv, j) + si * hh(iv + 1, j);
        const double h2 = -si * hh(iv, j) + ci * hh(iv + 1, j);
        hh(iv, j) = h1;
        hh(iv + 1, j) = h2;
      }
    }
    // back substitution (:946-953)
    u[m - 1] = g[m - 1] / hh(m - 1, m - 1);
    for (int iv = m - 2; iv >= 0; --iv) {
      u[iv] = g[iv];
      for (int j = iv + 1; j < m; ++j) u[iv] = u[iv] - hh(iv, j) * u[j];
      u[iv] = u[iv] / hh(iv, iv);
    }
    for (int iv = 0; iv < m; ++iv)
      hipLaunchKernelGGL(k_axpy_val<T>, dim3(G), dim3(kThreads), 0, ctx->stream, x, Vv(iv), n, (T)u[iv]);
    // res = -1 * (A x) + b   (res%copy(Ax); res%aypx(-1, b))
    rc = apply_op<T>(ctx, A, x, res, (const T*)nullptr, nullptr, nullptr);
    if (rc) return rc;
    hipLaunchKernelGGL(k_residual<T>, dim3(G), dim3(kThreads), 0, ctx->stream, res, b, n, part);
    LFR_CHECK_LAUNCH(ctx);
    st->iterations = it;
    if (P->monitor_convergence) {
      double rn = 0;
      rc = norm_host(res, &rn);
      if (rc) return rc;
      st->r_nrm = rn;
      err = rn / sc_err;
      if (err != err) return lfr_set_error(ctx, LFR_ERR_NAN, "gmres: NaN residual at restart %d", it);
      if (err < P->r_tol) {
        st->converged = 1;
        return LFR_OK;
      }
    }
  }
  if (P->monitor_convergence && P->fail_on_non_converged)
    return lfr_set_error(ctx, LFR_ERR_NOT_CONVERGED, "GMRES solver_algorithm: NOT converged in %d iters, Res=%g",
                         P->max_iter, err);
  return LFR_OK;
}

}  // namespace

extern "C" int lfr_gmres(lfr_ctx* ctx, const lfr_lma_op* A, const lfr_solver_params* P, void* x,
                         const void* b, lfr_solver_stats* st) {
  if (!ctx) return LFR_ERR_ARG;
  int rc = lfr_check_op(ctx, A, "lfr_gmres");
  if (rc) return rc;
  LFR_REQUIRE(ctx, P && x && b && st && P->gcrk > 0 && P->gcrk <= 512, "lfr_gmres: bad arguments");
  LFR_REQUIRE(ctx, P->preconditioner == 0 || (P->preconditioner == 1 && P->diag_d),
              "lfr_gmres: preconditioner must be 0 (null) or 1 (diagonal, with diag_d)");
  if (A->prec == 4) return gmres_solve<float>(ctx, A, P, (float*)x, (const float*)b, st);
  return gmres_solve<double>(ctx, A, P, (double*)x, (const double*)b, st);
}
