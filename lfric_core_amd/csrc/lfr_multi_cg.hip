// Sixteen simultaneous refined CG solves A x_r = b_r, r = 0..15.
//
// Outer loop in FP64 on the device: residual r_r = b_r - A64 x_r by one
// multidata FP64 gather apply of all fields (lfr_matrix_vector_ndata: each
// field as a single-field apply computes it), convergence when
// ||r_r|| <= r_tol ||b_r|| for every field.  Inner solves: unpreconditioned
// CG on the FP32 multidata field (16 values per dof contiguous) with the
// MFMA apply (lfr_matrix_vector_multi), per-field FP64 scalars
// (conjugate_gradient_type%cg_solve,
// components/science/source/solver/sci_iterative_solver_mod.f90:495-592:
// alpha = <r,z>/<p,Ap>, updates with the scalars cast to r_solver, inner
// products accumulated in FP64), each field frozen (alpha = beta = 0) once
// ||r_r|| <= inner_r_tol ||r0_r||.  Every 16-field inner product uses a
// fixed grid and a fixed combination order, so results are reproducible.
#include <cmath>
#include <vector>

#include "lfr_internal.h"
#include "lfr_reduce.h"

namespace {

constexpr int kR = 16;

// Partial sums per field of a[i]*b[i] over interleaved fields (i % 16 = field).
// Stride of every thread is a multiple of 16, so a thread always sees one field.
template <typename TA, typename TB>
__global__ void __launch_bounds__(256)
k_mdot(const TA* __restrict__ a, const TB* __restrict__ b, int64_t n16, double* __restrict__ part) {
  __shared__ double lds[4][kR];
  double s = 0.0;
  for (int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x; i < n16; i += (int64_t)gridDim.x * 256)
    s += (double)a[i] * (double)b[i];
  s += __shfl_xor(s, 16, 64);
  s += __shfl_xor(s, 32, 64);
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  if (lane < kR) lds[w][lane] = s;
  __syncthreads();
  if (threadIdx.x < kR) {
    const double t = ((lds[0][threadIdx.x] + lds[1][threadIdx.x]) + lds[2][threadIdx.x]) + lds[3][threadIdx.x];
    part[(int64_t)blockIdx.x * kR + threadIdx.x] = t;
  }
}

// out[r] = sum over blocks (fixed order) of part[blk*16 + r]
__global__ void __launch_bounds__(256) k_mdot_final(const double* __restrict__ part, int nb, double* __restrict__ out) {
  __shared__ double lds[16][kR];
  const int r = threadIdx.x & 15, j = threadIdx.x >> 4;
  double s = 0.0;
  for (int bb = j; bb < nb; bb += 16) s += part[(int64_t)bb * kR + r];
  lds[j][r] = s;
  __syncthreads();
  if (threadIdx.x < kR) {
    double t = 0.0;
    for (int q = 0; q < 16; ++q) t += lds[q][threadIdx.x];
    out[threadIdx.x] = t;
  }
}

// S layout (doubles): [0,16) rz; [16,32) pz; [32,48) alpha; [48,64) beta;
// [64,80) r0 norm; [80,96) active flag; [96] active count
constexpr int kRZ = 0, kPZ = 16, kAL = 32, kBE = 48, kR0 = 64, kACT = 80, kNACT = 96;

__global__ void k_mcg_init(double* S) {
  const int r = threadIdx.x;
  if (r < kR) {
    S[kR0 + r] = sqrt(S[kRZ + r]);
    S[kACT + r] = S[kR0 + r] > 0.0 ? 1.0 : 0.0;
  }
}

__global__ void k_mcg_alpha(double* S) {
  const int r = threadIdx.x;
  if (r < kR) S[kAL + r] = S[kACT + r] != 0.0 ? S[kRZ + r] / S[kPZ + r] : 0.0;
}

// x += alpha p ; r -= alpha z   (scalars cast to FP32), fused partials of <r,r>
__global__ void __launch_bounds__(256)
k_mcg_update(float* __restrict__ x, float* __restrict__ res, const float* __restrict__ p,
             const float* __restrict__ z, int64_t n16, const double* __restrict__ S, double* __restrict__ part) {
  __shared__ double lds[4][kR];
  const int r = threadIdx.x & 15;
  const float a = (float)S[kAL + r];
  double s = 0.0;
  for (int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x; i < n16; i += (int64_t)gridDim.x * 256) {
    x[i] = x[i] + a * p[i];
    const float rv = res[i] - a * z[i];
    res[i] = rv;
    s += (double)rv * (double)rv;
  }
  s += __shfl_xor(s, 16, 64);
  s += __shfl_xor(s, 32, 64);
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  if (lane < kR) lds[w][lane] = s;
  __syncthreads();
  if (threadIdx.x < kR) {
    const double t = ((lds[0][threadIdx.x] + lds[1][threadIdx.x]) + lds[2][threadIdx.x]) + lds[3][threadIdx.x];
    part[(int64_t)blockIdx.x * kR + threadIdx.x] = t;
  }
}

// beta = <r,r>_new / <r,r>; freeze converged fields; count active ones
__global__ void k_mcg_beta(double* S, const double* __restrict__ rr_new, double tol) {
  const int r = threadIdx.x;
  if (r < kR) {
    const bool act = S[kACT + r] != 0.0;
    double beta = 0.0;
    if (act) {
      beta = rr_new[r] / S[kRZ + r];
      S[kRZ + r] = rr_new[r];
      if (sqrt(rr_new[r]) <= tol * S[kR0 + r]) {
        S[kACT + r] = 0.0;
        beta = 0.0;
      }
    }
    S[kBE + r] = beta;
  }
  __syncthreads();
  if (r == 0) {
    double cnt = 0.0;
    for (int q = 0; q < kR; ++q) cnt += S[kACT + q];
    S[kNACT] = cnt;
  }
}

// p = r + beta p
__global__ void __launch_bounds__(256)
k_mcg_dir(float* __restrict__ p, const float* __restrict__ res, int64_t n16, const double* __restrict__ S) {
  const int r = threadIdx.x & 15;
  const float bt = (float)S[kBE + r];
  for (int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x; i < n16; i += (int64_t)gridDim.x * 256)
    p[i] = res[i] + bt * p[i];
}

// res64 = b - Y (multidata fields); res32 = (float)res64
__global__ void __launch_bounds__(256) k_residual16(double* __restrict__ res64, float* __restrict__ res32,
                                                    const double* __restrict__ b, const double* __restrict__ Y,
                                                    int64_t n16) {
  for (int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x; i < n16; i += (int64_t)gridDim.x * 256) {
    const double v = b[i] - Y[i];
    res64[i] = v;
    res32[i] = (float)v;
  }
}

// x64 += (double) d32
__global__ void k_add_correction(double* __restrict__ x, const float* __restrict__ d, int64_t n16) {
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n16; i += (int64_t)gridDim.x * blockDim.x)
    x[i] = x[i] + (double)d[i];
}

__global__ void k_fill0(float* __restrict__ v, int64_t n) {
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x)
    v[i] = 0.f;
}

__global__ void k_copyf(float* __restrict__ dst, const float* __restrict__ src, int64_t n) {
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x)
    dst[i] = src[i];
}

size_t up(size_t b) { return (b + 255) & ~(size_t)255; }

}  // namespace

extern "C" int lfr_refined_cg_multi(lfr_ctx* ctx, lfr_multi* A32, const lfr_lma_op* A64, double r_tol,
                                    int max_outer, double inner_r_tol, int inner_max_iter, void* x64_d,
                                    const void* b64_d, int nrhs, lfr_refine_stats* stats,
                                    double* rel_res_h) {
  if (!ctx || !A32 || !A64) return LFR_ERR_ARG;
  LFR_REQUIRE(ctx, nrhs == kR, "lfr_refined_cg_multi: nrhs must be %d", kR);
  LFR_REQUIRE(ctx, A64->prec == 8, "lfr_refined_cg_multi: A64 must be an FP64 operator");
  int rc0 = lfr_check_op(ctx, A64, "lfr_refined_cg_multi(A64)");
  if (rc0) return rc0;
  // the FP64 residual is a multidata apply of A64's matrix in the reference
  // layout; enforce_bc has no multidata form here
  LFR_REQUIRE(ctx, !A64->apply_bc, "lfr_refined_cg_multi: apply_bc is not supported for the 16-field residual");
  int ut = 0, uf = 0;
  lfr_multi_space_undf(A32, &ut, &uf);
  LFR_REQUIRE(ctx, ut == uf && ut == A64->space->undf,
              "lfr_refined_cg_multi: operators on different spaces (%d, %d, %d)", ut, uf, A64->space->undf);
  LFR_REQUIRE(ctx, x64_d && b64_d && max_outer >= 1 && inner_max_iter >= 1, "lfr_refined_cg_multi: bad arguments");
  const int64_t n = ut, n16 = n * kR;
  const int nb = lfr::red_blocks(n16);
  // workspace: 4 FP32 multidata fields (d, res, p, z), FP64 residual field,
  // A x of the 16 fields (FP64, multidata), partials, scalars
  const size_t f32 = up(n16 * 4), f64 = up(n16 * 8), s64 = up(n16 * 8);
  const size_t need = 4 * f32 + f64 + s64 + up((size_t)nb * kR * 8) + up(128 * 8) + up(2 * kR * 8);
  int rc = lfr_work_reserve(ctx, need);
  if (rc) return rc;
  char* w = (char*)ctx->work;
  float* d = (float*)w;
  float* res = (float*)(w + f32);
  float* p = (float*)(w + 2 * f32);
  float* z = (float*)(w + 3 * f32);
  double* r64 = (double*)(w + 4 * f32);
  double* Y16 = (double*)(w + 4 * f32 + f64);
  double* part = (double*)(w + 4 * f32 + f64 + s64);
  double* S = (double*)(w + 4 * f32 + f64 + s64 + up((size_t)nb * kR * 8));
  double* tmp = S + 128;
  hipStream_t s = ctx->stream;
  const unsigned g = lfr_blocks(n16, 256, 8192);
  double* x64 = (double*)x64_d;
  const double* b64 = (const double*)b64_d;
  double* Sh = ctx->host_scalars;   // 64 doubles

  // ||b_r||
  hipLaunchKernelGGL((k_mdot<double, double>), dim3(nb), dim3(256), 0, s, b64, b64, n16, part);
  hipLaunchKernelGGL(k_mdot_final, dim3(1), dim3(256), 0, s, part, nb, tmp);
  LFR_HIP(ctx, hipMemcpyAsync(Sh, tmp, kR * 8, hipMemcpyDeviceToHost, s));
  LFR_HIP(ctx, hipStreamSynchronize(s));
  double bn[kR];
  for (int r = 0; r < kR; ++r) bn[r] = sqrt(Sh[r]);

  int outer = 0, inner_total = 0, conv = 0;
  double rel[kR];
  for (;;) {
    // FP64 residual: one multidata FP64 apply of all 16 fields (the matrix
    // is read once; each field as a single-field apply would compute it)
    rc = lfr_matrix_vector_ndata(ctx, 8, Y16, x64, A64->matrix_d, A64->space, A64->space, A64->plan, kR, 1);
    if (rc) return rc;
    hipLaunchKernelGGL(k_residual16, dim3(g), dim3(256), 0, s, r64, res, b64, (const double*)Y16, n16);
    hipLaunchKernelGGL((k_mdot<double, double>), dim3(nb), dim3(256), 0, s, (const double*)r64,
                       (const double*)r64, n16, part);
    hipLaunchKernelGGL(k_mdot_final, dim3(1), dim3(256), 0, s, part, nb, tmp);
    LFR_CHECK_LAUNCH(ctx);
    LFR_HIP(ctx, hipMemcpyAsync(Sh, tmp, kR * 8, hipMemcpyDeviceToHost, s));
    LFR_HIP(ctx, hipStreamSynchronize(s));
    bool all = true;
    for (int r = 0; r < kR; ++r) {
      rel[r] = bn[r] > 0 ? sqrt(Sh[r]) / bn[r] : sqrt(Sh[r]);
      if (std::isnan(rel[r])) return lfr_set_error(ctx, LFR_ERR_NAN, "lfr_refined_cg_multi: NaN residual");
      if (rel[r] > r_tol) all = false;
    }
    if (all) {
      conv = 1;
      break;
    }
    if (outer >= max_outer) break;
    // inner CG from d = 0 on res (FP32): r = res, p = res, rz = <r,r>
    hipLaunchKernelGGL(k_fill0, dim3(g), dim3(256), 0, s, d, n16);
    hipLaunchKernelGGL(k_copyf, dim3(g), dim3(256), 0, s, p, (const float*)res, n16);
    hipLaunchKernelGGL((k_mdot<float, float>), dim3(nb), dim3(256), 0, s, (const float*)res, (const float*)res,
                       n16, part);
    hipLaunchKernelGGL(k_mdot_final, dim3(1), dim3(256), 0, s, part, nb, S + kRZ);
    hipLaunchKernelGGL(k_mcg_init, dim3(1), dim3(64), 0, s, S);
    LFR_CHECK_LAUNCH(ctx);
    int it = 0;
    for (; it < inner_max_iter;) {
      rc = lfr_matrix_vector_multi(ctx, A32, z, p, kR, 1);
      if (rc) return rc;
      hipLaunchKernelGGL((k_mdot<float, float>), dim3(nb), dim3(256), 0, s, (const float*)p, (const float*)z, n16,
                         part);
      hipLaunchKernelGGL(k_mdot_final, dim3(1), dim3(256), 0, s, part, nb, S + kPZ);
      hipLaunchKernelGGL(k_mcg_alpha, dim3(1), dim3(64), 0, s, S);
      hipLaunchKernelGGL(k_mcg_update, dim3(nb), dim3(256), 0, s, d, res, (const float*)p, (const float*)z, n16,
                         (const double*)S, part);
      hipLaunchKernelGGL(k_mdot_final, dim3(1), dim3(256), 0, s, part, nb, tmp);
      hipLaunchKernelGGL(k_mcg_beta, dim3(1), dim3(64), 0, s, S, (const double*)tmp, inner_r_tol);
      hipLaunchKernelGGL(k_mcg_dir, dim3(g), dim3(256), 0, s, p, (const float*)res, n16, (const double*)S);
      LFR_CHECK_LAUNCH(ctx);
      ++it;
      if (it % 5 == 0 || it == inner_max_iter) {
        LFR_HIP(ctx, hipMemcpyAsync(Sh, S + kNACT, 8, hipMemcpyDeviceToHost, s));
        LFR_HIP(ctx, hipStreamSynchronize(s));
        if (Sh[0] == 0.0) break;
      }
    }
    inner_total += it;
    hipLaunchKernelGGL(k_add_correction, dim3(g), dim3(256), 0, s, x64, (const float*)d, n16);
    LFR_CHECK_LAUNCH(ctx);
    ++outer;
  }
  double worst = 0, bmax = 0;
  for (int r = 0; r < kR; ++r) {
    worst = rel[r] > worst ? rel[r] : worst;
    bmax = bn[r] > bmax ? bn[r] : bmax;
    if (rel_res_h) rel_res_h[r] = rel[r];
  }
  if (stats) {
    stats->outer = outer;
    stats->inner_total = inner_total;
    stats->converged = conv;
    stats->b_nrm = bmax;
    stats->r_nrm = worst;
  }
  if (!conv)
    return lfr_set_error(ctx, LFR_ERR_NOT_CONVERGED,
                         "lfr_refined_cg_multi: not converged after %d corrections (worst ||r||/||b|| = %g)", outer,
                         worst);
  return LFR_OK;
}
