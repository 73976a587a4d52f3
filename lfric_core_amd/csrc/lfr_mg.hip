// Geometric multigrid V-cycle for a W3 Helmholtz-type operator with
// vertical line relaxation (columnwise_op_appinv as smoother).
//
// Operator on level l (lowest-order W3, one dof per cell and layer):
//   (A x)(c,k) = a(c,k) x(c,k-1) + d(c,k) x(c,k) + u(c,k) x(c,k+1)
//              + sigma_l * (4 x(c,k) - sum_{n in W,S,E,N neighbours} x(n,k))
// i.e. a column tridiagonal T (the CMA operator solved by
// sci_columnwise_op_appinv_kernel_mod.F90) plus a horizontal graph Laplacian.
// Smoother: damped vertical line Jacobi, x <- x + omega D^{-1} (b - A x) with
// D = T + 4 sigma I, one Thomas solve per column (the appinv recurrence).
// Transfer: cell-average restriction (sci_restrict_scalar_unweighted_kernel_mod.F90:100-116,
// children visited y outer, x inner, each scaled by 1/4 before it is added)
// and injection prolongation (sci_prolong_scalar_unweighted_kernel_mod.F90:108-116,
// fine += coarse).
//
// Layout: level fields are stored layer-major, f[(rhs*nl + k)*ncell + c], so
// that one thread per (column, right-hand side) walking its column reads
// consecutive cells across a wavefront for every layer (coalesced), and the
// four neighbour reads of a layer hit one contiguous plane.  Band arrays are
// [k][ncell] per band.  The fine-level right-hand side and solution are
// converted from/to LFRic's k-fastest W3 layout once per solve.
//
// One V-cycle over all right-hand sides is captured into a hipGraph on first
// use and replayed; each level's kernels use the workgroup size chosen by a
// short timing sweep at set-up (the autotuner).
#include <algorithm>
#include <cstring>
#include <cmath>
#include <vector>

#include "lfr_internal.h"
#include "lfr_reduce.h"

struct lfr_mg_level {
  int ncell = 0;
  int32_t* cell_next_d = nullptr;   // (ncell, 4) 0-based neighbours
  int32_t* children_d = nullptr;    // (ncell, 4) fine children in y-outer, x-inner order (to next finer level)
  int32_t* parent_d = nullptr;      // (ncell) coarse parent on the next coarser level
  void* a_d = nullptr;              // bands [k][ncell]
  void* d_d = nullptr;
  void* u_d = nullptr;
  void* x_d = nullptr;              // [rhs][k][ncell]
  void* b_d = nullptr;
  void* t_d = nullptr;              // smoother ping-pong buffer
  double sigma = 0;
  int block_smooth = 256;           // autotuned workgroup sizes
  int block_other = 256;
};

struct lfr_mg {
  lfr_ctx* ctx = nullptr;
  int prec = 4, nl = 0, nrhs = 1, nlev = 0;
  int nu1 = 2, nu2 = 2, ncoarse = 8;
  double omega = 0.8;
  std::vector<lfr_mg_level> lev;
  double* part_d = nullptr;         // residual-norm partials (nrhs * red blocks)
  hipGraphExec_t graph = nullptr;
  const void* graph_x = nullptr;
  const void* graph_b = nullptr;
};

namespace {

constexpr int kMaxNl = 128;

template <typename T>
__device__ __forceinline__ T horiz(const T* __restrict__ xk, const int32_t* __restrict__ nb, int c,
                                   T xc, T sigma) {
  const T s4 = xk[nb[4 * c + 0]] + xk[nb[4 * c + 1]];
  const T s5 = s4 + xk[nb[4 * c + 2]];
  const T sn = s5 + xk[nb[4 * c + 3]];
  const T four = (T)4 * xc;
  return sigma * (four - sn);
}

// One damped line-Jacobi sweep: xo = xi + omega * D^{-1} (b - A xi)
template <typename T, int NL>
__global__ void k_mg_smooth(T* __restrict__ xo, const T* __restrict__ xi, const T* __restrict__ b,
                            const T* __restrict__ ba, const T* __restrict__ bd,
                            const T* __restrict__ bu, const int32_t* __restrict__ nb, int ncell,
                            int nl_rt, T sigma, T omega) {
  const int nl = NL > 0 ? NL : nl_rt;
  const int c = blockIdx.x * blockDim.x + threadIdx.x;
  const int rhs = blockIdx.y;
  if (c >= ncell) return;
  const int64_t off = (int64_t)rhs * nl * ncell;
  const T* x = xi + off;
  const T* bb = b + off;
  T cp[NL > 0 ? NL : kMaxNl], dp[NL > 0 ? NL : kMaxNl];
  T xm = 0, x0 = x[c];
#pragma unroll
  for (int k = 0; k < (NL > 0 ? NL : kMaxNl); ++k) {
    if (NL == 0 && k >= nl) break;
    const int64_t ik = (int64_t)k * ncell + c;
    const T xp = (k + 1 < nl) ? x[ik + ncell] : T(0);
    const T ak = ba[ik], dk = bd[ik], uk = bu[ik];
    // residual r = b - A x at (c, k)
    T ax = ak * xm;
    ax = ax + dk * x0;
    ax = ax + uk * xp;
    ax = ax + horiz(x + (int64_t)k * ncell, nb, c, x0, sigma);
    const T r = bb[ik] - ax;
    // Thomas forward sweep on D = tridiag(a, d + 4 sigma, u)
    const T dd = dk + (T)4 * sigma;
    if (k == 0) {
      const T inv = T(1) / dd;
      cp[k] = inv * uk;
      dp[k] = inv * r;
    } else {
      const T inv = T(1) / (dd - ak * cp[k - 1]);
      cp[k] = inv * uk;
      dp[k] = inv * (r - ak * dp[k - 1]);
    }
    xm = x0;
    x0 = xp;
  }
  T* xout = xo + off;
#pragma unroll
  for (int k = (NL > 0 ? NL : kMaxNl) - 1; k >= 0; --k) {
    if (NL == 0 && k >= nl) continue;
    if (k < nl - 1) dp[k] = dp[k] - cp[k] * dp[k + 1];
    const int64_t ik = (int64_t)k * ncell + c;
    xout[ik] = x[ik] + omega * dp[k];
  }
}

// bc(C,k) = sum over the 4 children f (y outer, x inner) of r(f,k) * 1/4,
// r = b - A x on the fine level.  Also zeroes the coarse solution.
template <typename T>
__global__ void k_mg_restrict(T* __restrict__ bc, T* __restrict__ xc, const T* __restrict__ xf,
                              const T* __restrict__ bf, const T* __restrict__ ba,
                              const T* __restrict__ bd, const T* __restrict__ bu,
                              const int32_t* __restrict__ nb, const int32_t* __restrict__ children,
                              int ncell_c, int ncell_f, int nl, T sigma) {
  const int C = blockIdx.x * blockDim.x + threadIdx.x;
  const int rhs = blockIdx.y;
  if (C >= ncell_c) return;
  const int64_t offf = (int64_t)rhs * nl * ncell_f, offc = (int64_t)rhs * nl * ncell_c;
  const T* x = xf + offf;
  const T* b = bf + offf;
  const T denom = T(1) / T(4);
  int ch[4];
#pragma unroll
  for (int q = 0; q < 4; ++q) ch[q] = children[4 * C + q];
  for (int k = 0; k < nl; ++k) {
    T acc = 0;
#pragma unroll
    for (int q = 0; q < 4; ++q) {
      const int f = ch[q];
      const int64_t ik = (int64_t)k * ncell_f + f;
      const T x0 = x[ik];
      const T xm = k > 0 ? x[ik - ncell_f] : T(0);
      const T xp = k + 1 < nl ? x[ik + ncell_f] : T(0);
      T ax = ba[ik] * xm;
      ax = ax + bd[ik] * x0;
      ax = ax + bu[ik] * xp;
      ax = ax + horiz(x + (int64_t)k * ncell_f, nb, f, x0, sigma);
      const T r = b[ik] - ax;
      acc = acc + r * denom;
    }
    bc[offc + (int64_t)k * ncell_c + C] = acc;
    xc[offc + (int64_t)k * ncell_c + C] = T(0);
  }
}

// xf(f,k) += xc(parent(f), k)
template <typename T>
__global__ void k_mg_prolong(T* __restrict__ xf, const T* __restrict__ xc,
                             const int32_t* __restrict__ parent, int ncell_f, int ncell_c, int nl) {
  const int f = blockIdx.x * blockDim.x + threadIdx.x;
  const int rhs = blockIdx.y;
  if (f >= ncell_f) return;
  const int P = parent[f];
  T* x = xf + (int64_t)rhs * nl * ncell_f;
  const T* y = xc + (int64_t)rhs * nl * ncell_c;
  for (int k = 0; k < nl; ++k) x[(int64_t)k * ncell_f + f] += y[(int64_t)k * ncell_c + P];
}

// per-rhs partial sums of r^2, r = b - A x
template <typename T>
__global__ void __launch_bounds__(lfr::kRedThreads)
k_mg_resnorm(const T* __restrict__ xf, const T* __restrict__ bf, const T* __restrict__ ba,
             const T* __restrict__ bd, const T* __restrict__ bu, const int32_t* __restrict__ nb,
             int ncell, int nl, T sigma, double* __restrict__ part, int nblk) {
  __shared__ double lds[lfr::kRedThreads / 64];
  const int rhs = blockIdx.y;
  const int64_t off = (int64_t)rhs * nl * ncell;
  const T* x = xf + off;
  const T* b = bf + off;
  const int64_t n = (int64_t)ncell * nl;
  double s = 0.0;
  for (int64_t i = (int64_t)blockIdx.x * lfr::kRedThreads + threadIdx.x; i < n;
       i += (int64_t)gridDim.x * lfr::kRedThreads) {
    const int k = (int)(i / ncell);
    const int c = (int)(i - (int64_t)k * ncell);
    const T x0 = x[i];
    const T xm = k > 0 ? x[i - ncell] : T(0);
    const T xp = k + 1 < nl ? x[i + ncell] : T(0);
    T ax = ba[i] * xm;
    ax = ax + bd[i] * x0;
    ax = ax + bu[i] * xp;
    ax = ax + horiz(x + (int64_t)k * ncell, nb, c, x0, sigma);
    const double r = (double)(b[i] - ax);
    s += r * r;
  }
  s = lfr::block_sum(s, lds);
  if (threadIdx.x == 0) part[rhs * nblk + blockIdx.x] = s;
}

// LFRic W3 layout (k fastest: dof = c*nl + k) <-> layer-major
template <typename T>
__global__ void k_to_layer_major(T* __restrict__ out, const T* __restrict__ in, int ncell, int nl, int nrhs) {
  const int64_t n = (int64_t)ncell * nl * nrhs;
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x) {
    const int64_t per = (int64_t)ncell * nl;
    const int64_t rhs = i / per, rem = i - rhs * per;
    const int k = (int)(rem / ncell), c = (int)(rem - (int64_t)k * ncell);
    out[i] = in[rhs * per + (int64_t)c * nl + k];
  }
}
template <typename T>
__global__ void k_from_layer_major(T* __restrict__ out, const T* __restrict__ in, int ncell, int nl, int nrhs) {
  const int64_t n = (int64_t)ncell * nl * nrhs;
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x) {
    const int64_t per = (int64_t)ncell * nl;
    const int64_t rhs = i / per, rem = i - rhs * per;
    const int c = (int)(rem / nl), k = (int)(rem - (int64_t)c * nl);
    out[i] = in[rhs * per + (int64_t)k * ncell + c];
  }
}

template <typename T>
void launch_smooth(hipStream_t s, lfr_mg* mg, lfr_mg_level& L, T* xo, const T* xi) {
  const int bs = L.block_smooth;
  dim3 grid((L.ncell + bs - 1) / bs, mg->nrhs);
  const T sig = (T)L.sigma, om = (T)mg->omega;
  if (mg->nl == 70)
    hipLaunchKernelGGL((k_mg_smooth<T, 70>), grid, dim3(bs), 0, s, xo, xi, (const T*)L.b_d,
                       (const T*)L.a_d, (const T*)L.d_d, (const T*)L.u_d, L.cell_next_d, L.ncell, 70, sig, om);
  else if (mg->nl == 10)
    hipLaunchKernelGGL((k_mg_smooth<T, 10>), grid, dim3(bs), 0, s, xo, xi, (const T*)L.b_d,
                       (const T*)L.a_d, (const T*)L.d_d, (const T*)L.u_d, L.cell_next_d, L.ncell, 10, sig, om);
  else
    hipLaunchKernelGGL((k_mg_smooth<T, 0>), grid, dim3(bs), 0, s, xo, xi, (const T*)L.b_d,
                       (const T*)L.a_d, (const T*)L.d_d, (const T*)L.u_d, L.cell_next_d, L.ncell, mg->nl, sig, om);
}

// two sweeps: x -> t -> x
template <typename T>
void smooth2(hipStream_t s, lfr_mg* mg, lfr_mg_level& L) {
  launch_smooth<T>(s, mg, L, (T*)L.t_d, (const T*)L.x_d);
  launch_smooth<T>(s, mg, L, (T*)L.x_d, (const T*)L.t_d);
}

template <typename T>
void vcycle(hipStream_t s, lfr_mg* mg, int l) {
  lfr_mg_level& L = mg->lev[l];
  if (l == mg->nlev - 1) {
    for (int i = 0; i < mg->ncoarse / 2; ++i) smooth2<T>(s, mg, L);
    return;
  }
  for (int i = 0; i < mg->nu1 / 2; ++i) smooth2<T>(s, mg, L);
  lfr_mg_level& C = mg->lev[l + 1];
  {
    const int bs = L.block_other;
    dim3 grid((C.ncell + bs - 1) / bs, mg->nrhs);
    hipLaunchKernelGGL(k_mg_restrict<T>, grid, dim3(bs), 0, s, (T*)C.b_d, (T*)C.x_d, (const T*)L.x_d,
                       (const T*)L.b_d, (const T*)L.a_d, (const T*)L.d_d, (const T*)L.u_d,
                       L.cell_next_d, C.children_d, C.ncell, L.ncell, mg->nl, (T)L.sigma);
  }
  vcycle<T>(s, mg, l + 1);
  {
    const int bs = L.block_other;
    dim3 grid((L.ncell + bs - 1) / bs, mg->nrhs);
    hipLaunchKernelGGL(k_mg_prolong<T>, grid, dim3(bs), 0, s, (T*)L.x_d, (const T*)C.x_d, L.parent_d,
                       L.ncell, C.ncell, mg->nl);
  }
  for (int i = 0; i < mg->nu2 / 2; ++i) smooth2<T>(s, mg, L);
}

size_t esz(int prec) { return prec == 4 ? 4 : 8; }

void mg_free(lfr_mg* mg) {
  if (!mg) return;
  if (mg->graph) (void)hipGraphExecDestroy(mg->graph);
  for (auto& L : mg->lev)
    for (void* p : {(void*)L.cell_next_d, (void*)L.children_d, (void*)L.parent_d, L.a_d, L.d_d, L.u_d,
                    L.x_d, L.b_d, L.t_d})
      if (p) (void)hipFree(p);
  if (mg->part_d) (void)hipFree(mg->part_d);
  delete mg;
}

// Pick the workgroup size of the smoother (the dominant kernel) for each
// level by timing one sweep with each candidate.
template <typename T>
int autotune(lfr_mg* mg) {
  lfr_ctx* ctx = mg->ctx;
  hipEvent_t e0, e1;
  LFR_HIP(ctx, hipEventCreate(&e0));
  LFR_HIP(ctx, hipEventCreate(&e1));
  const int cand[3] = {64, 128, 256};
  for (auto& L : mg->lev) {
    float best = 1e30f;
    int bestb = 256;
    for (int b : cand) {
      L.block_smooth = b;
      launch_smooth<T>(ctx->stream, mg, L, (T*)L.t_d, (const T*)L.x_d);  // warm
      LFR_HIP(ctx, hipEventRecord(e0, ctx->stream));
      for (int r = 0; r < 3; ++r) launch_smooth<T>(ctx->stream, mg, L, (T*)L.t_d, (const T*)L.x_d);
      LFR_HIP(ctx, hipEventRecord(e1, ctx->stream));
      LFR_HIP(ctx, hipEventSynchronize(e1));
      float ms = 0;
      LFR_HIP(ctx, hipEventElapsedTime(&ms, e0, e1));
      if (ms < best) {
        best = ms;
        bestb = b;
      }
    }
    L.block_smooth = bestb;
    L.block_other = 256;
  }
  (void)hipEventDestroy(e0);
  (void)hipEventDestroy(e1);
  return LFR_OK;
}

}  // namespace

extern "C" {

int lfr_mg_create(lfr_ctx* ctx, int prec, int nlev, int nlayers, int nrhs, const int* ncell,
                  const int32_t* const* cell_next_h, const int32_t* const* children_h,
                  const double* const* bands_h, const double* sigma, double omega, int nu1, int nu2,
                  int ncoarse, lfr_mg** out) {
  if (!ctx || !out) return LFR_ERR_ARG;
  *out = nullptr;
  LFR_REQUIRE(ctx, (prec == 4 || prec == 8) && nlev >= 1 && nlayers >= 2 && nlayers <= kMaxNl &&
                       nrhs >= 1 && nrhs <= 64,
              "lfr_mg_create: bad sizes (prec %d nlev %d nlayers %d nrhs %d)", prec, nlev, nlayers, nrhs);
  LFR_REQUIRE(ctx, nu1 % 2 == 0 && nu2 % 2 == 0 && ncoarse % 2 == 0 && ncoarse > 0,
              "lfr_mg_create: smoothing counts must be even (ping-pong sweeps)");
  lfr_mg* mg = new lfr_mg();
  mg->ctx = ctx;
  mg->prec = prec;
  mg->nl = nlayers;
  mg->nrhs = nrhs;
  mg->nlev = nlev;
  mg->nu1 = nu1;
  mg->nu2 = nu2;
  mg->ncoarse = ncoarse;
  mg->omega = omega;
  mg->lev.resize(nlev);
  const size_t s = esz(prec);
  for (int l = 0; l < nlev; ++l) {
    lfr_mg_level& L = mg->lev[l];
    L.ncell = ncell[l];
    L.sigma = sigma[l];
    const int nc = L.ncell;
    // validate connectivity on the host before anything reaches a kernel
    for (int64_t q = 0; q < (int64_t)nc * 4; ++q)
      if (cell_next_h[l][q] < 0 || cell_next_h[l][q] >= nc) {
        mg_free(mg);
        return lfr_set_error(ctx, LFR_ERR_ARG, "lfr_mg_create: level %d neighbour out of range", l);
      }
    const size_t nb = (size_t)nc * nlayers;
    bool ok = hipMalloc(&L.cell_next_d, (size_t)nc * 4 * 4) == hipSuccess &&
              hipMalloc(&L.a_d, nb * s) == hipSuccess && hipMalloc(&L.d_d, nb * s) == hipSuccess &&
              hipMalloc(&L.u_d, nb * s) == hipSuccess &&
              hipMalloc(&L.x_d, nb * nrhs * s) == hipSuccess &&
              hipMalloc(&L.b_d, nb * nrhs * s) == hipSuccess &&
              hipMalloc(&L.t_d, nb * nrhs * s) == hipSuccess;
    if (!ok) {
      mg_free(mg);
      return lfr_set_error(ctx, LFR_ERR_NOMEM, "lfr_mg_create: allocation failed");
    }
    (void)hipMemcpy(L.cell_next_d, cell_next_h[l], (size_t)nc * 4 * 4, hipMemcpyHostToDevice);
    // bands: bands_h[l] is [3][k][ncell] doubles (sub, diag, super)
    std::vector<char> buf(nb * s);
    for (int band = 0; band < 3; ++band) {
      const double* src = bands_h[l] + (size_t)band * nb;
      if (prec == 4)
        for (size_t q = 0; q < nb; ++q) ((float*)buf.data())[q] = (float)src[q];
      else
        memcpy(buf.data(), src, nb * s);
      void* dst = band == 0 ? L.a_d : (band == 1 ? L.d_d : L.u_d);
      (void)hipMemcpy(dst, buf.data(), nb * s, hipMemcpyHostToDevice);
    }
    (void)hipMemset(L.x_d, 0, nb * nrhs * s);
    (void)hipMemset(L.b_d, 0, nb * nrhs * s);
    (void)hipMemset(L.t_d, 0, nb * nrhs * s);
  }
  for (int l = 1; l < nlev; ++l) {
    lfr_mg_level& C = mg->lev[l];
    lfr_mg_level& F = mg->lev[l - 1];
    std::vector<int32_t> parent(F.ncell, -1);
    for (int64_t q = 0; q < (int64_t)C.ncell * 4; ++q) {
      const int f = children_h[l][q];
      if (f < 0 || f >= F.ncell || parent[f] >= 0) {
        mg_free(mg);
        return lfr_set_error(ctx, LFR_ERR_ARG, "lfr_mg_create: level %d children are not a partition", l);
      }
      parent[f] = (int32_t)(q / 4);
    }
    // every fine cell must have a parent (a 2x refinement: 4 * ncoarse == nfine)
    if ((int64_t)C.ncell * 4 != (int64_t)F.ncell) {
      mg_free(mg);
      return lfr_set_error(ctx, LFR_ERR_ARG, "lfr_mg_create: level %d has %d cells, level %d has %d (need 4x)",
                           l - 1, F.ncell, l, C.ncell);
    }
    if (hipMalloc(&C.children_d, (size_t)C.ncell * 16) != hipSuccess ||
        hipMalloc(&F.parent_d, (size_t)F.ncell * 4) != hipSuccess) {
      mg_free(mg);
      return lfr_set_error(ctx, LFR_ERR_NOMEM, "lfr_mg_create: allocation failed");
    }
    (void)hipMemcpy(C.children_d, children_h[l], (size_t)C.ncell * 16, hipMemcpyHostToDevice);
    (void)hipMemcpy(F.parent_d, parent.data(), (size_t)F.ncell * 4, hipMemcpyHostToDevice);
  }
  if (hipMalloc(&mg->part_d, (size_t)nrhs * lfr::kRedMaxBlocks * sizeof(double) + 64 * sizeof(double)) !=
      hipSuccess) {
    mg_free(mg);
    return lfr_set_error(ctx, LFR_ERR_NOMEM, "lfr_mg_create: allocation failed");
  }
  int rc = prec == 4 ? autotune<float>(mg) : autotune<double>(mg);
  if (rc) {
    mg_free(mg);
    return rc;
  }
  *out = mg;
  return LFR_OK;
}

int lfr_mg_destroy(lfr_mg* mg) {
  mg_free(mg);
  return LFR_OK;
}

int lfr_mg_block_sizes(const lfr_mg* mg, int* smooth_blocks) {
  if (!mg || !smooth_blocks) return LFR_ERR_ARG;
  for (int l = 0; l < mg->nlev; ++l) smooth_blocks[l] = mg->lev[l].block_smooth;
  return LFR_OK;
}

int lfr_mg_solve(lfr_mg* mg, void* x_d, const void* b_d, int ncycles, double* res_h) {
  if (!mg) return LFR_ERR_ARG;
  lfr_ctx* ctx = mg->ctx;
  LFR_REQUIRE(ctx, x_d && b_d && ncycles >= 0, "lfr_mg_solve: bad arguments");
  lfr_mg_level& F = mg->lev[0];
  const int64_t n = (int64_t)F.ncell * mg->nl * mg->nrhs;
  const unsigned g = lfr_blocks(n, 256, 4096);
  hipStream_t s = ctx->stream;
  const bool f32 = mg->prec == 4;
  if (f32) {
    hipLaunchKernelGGL(k_to_layer_major<float>, dim3(g), dim3(256), 0, s, (float*)F.b_d, (const float*)b_d, F.ncell, mg->nl, mg->nrhs);
    hipLaunchKernelGGL(k_to_layer_major<float>, dim3(g), dim3(256), 0, s, (float*)F.x_d, (const float*)x_d, F.ncell, mg->nl, mg->nrhs);
  } else {
    hipLaunchKernelGGL(k_to_layer_major<double>, dim3(g), dim3(256), 0, s, (double*)F.b_d, (const double*)b_d, F.ncell, mg->nl, mg->nrhs);
    hipLaunchKernelGGL(k_to_layer_major<double>, dim3(g), dim3(256), 0, s, (double*)F.x_d, (const double*)x_d, F.ncell, mg->nl, mg->nrhs);
  }
  LFR_CHECK_LAUNCH(ctx);
  if (!mg->graph) {
    hipGraph_t graph = nullptr;
    bool captured = false;
    if (hipStreamBeginCapture(s, hipStreamCaptureModeRelaxed) == hipSuccess) {
      if (f32) vcycle<float>(s, mg, 0);
      else vcycle<double>(s, mg, 0);
      captured = hipStreamEndCapture(s, &graph) == hipSuccess && graph;
    }
    if (captured) {
      if (hipGraphInstantiate(&mg->graph, graph, nullptr, nullptr, 0) != hipSuccess) mg->graph = nullptr;
      (void)hipGraphDestroy(graph);
    }
    (void)hipGetLastError();
  }
  const int nblk = lfr::red_blocks((int64_t)F.ncell * mg->nl);
  auto resnorm = [&](double* dst) -> int {
    dim3 grid(nblk, mg->nrhs);
    if (f32)
      hipLaunchKernelGGL(k_mg_resnorm<float>, grid, dim3(lfr::kRedThreads), 0, s, (const float*)F.x_d,
                         (const float*)F.b_d, (const float*)F.a_d, (const float*)F.d_d, (const float*)F.u_d,
                         F.cell_next_d, F.ncell, mg->nl, (float)F.sigma, mg->part_d, nblk);
    else
      hipLaunchKernelGGL(k_mg_resnorm<double>, grid, dim3(lfr::kRedThreads), 0, s, (const double*)F.x_d,
                         (const double*)F.b_d, (const double*)F.a_d, (const double*)F.d_d, (const double*)F.u_d,
                         F.cell_next_d, F.ncell, mg->nl, F.sigma, mg->part_d, nblk);
    double* sums = mg->part_d + (size_t)mg->nrhs * lfr::kRedMaxBlocks;
    hipLaunchKernelGGL(lfr::k_reduce_final, dim3(mg->nrhs), dim3(lfr::kRedThreads), 0, s, mg->part_d,
                       nblk, nblk, sums);
    LFR_CHECK_LAUNCH(ctx);
    LFR_HIP(ctx, hipMemcpyAsync(ctx->host_scalars, sums, mg->nrhs * sizeof(double), hipMemcpyDeviceToHost, s));
    LFR_HIP(ctx, hipStreamSynchronize(s));
    for (int r = 0; r < mg->nrhs; ++r) dst[r] = sqrt(ctx->host_scalars[r]);
    return LFR_OK;
  };
  if (res_h) {
    std::vector<double> r0(mg->nrhs);
    int rc = resnorm(r0.data());
    if (rc) return rc;
    for (int r = 0; r < mg->nrhs; ++r) res_h[r] = r0[r];
  }
  for (int cyc = 0; cyc < ncycles; ++cyc) {
    if (mg->graph) {
      LFR_HIP(ctx, hipGraphLaunch(mg->graph, s));
    } else {
      if (f32) vcycle<float>(s, mg, 0);
      else vcycle<double>(s, mg, 0);
      LFR_CHECK_LAUNCH(ctx);
    }
    if (res_h) {
      int rc = resnorm(res_h + (size_t)(cyc + 1) * mg->nrhs);
      if (rc) return rc;
    }
  }
  if (f32)
    hipLaunchKernelGGL(k_from_layer_major<float>, dim3(g), dim3(256), 0, s, (float*)x_d, (const float*)F.x_d, F.ncell, mg->nl, mg->nrhs);
  else
    hipLaunchKernelGGL(k_from_layer_major<double>, dim3(g), dim3(256), 0, s, (double*)x_d, (const double*)F.x_d, F.ncell, mg->nl, mg->nrhs);
  LFR_CHECK_LAUNCH(ctx);
  return LFR_OK;
}

}  // extern "C"
