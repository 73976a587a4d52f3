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
// Layout: level fields are stored f[(k*ncell + c)*R + r] (layer, cell, then
// the R right-hand sides fastest), so that the lanes of a wavefront cover R
// right-hand sides of consecutive cells: every load of x, b and of a
// neighbour value is a contiguous run, and the band values of a column are
// shared by its R lanes.  Band arrays are [k][ncell].  The fine-level
// right-hand side and solution are converted from/to LFRic's W3 layout
// (field r at r*undf, dof = c*nl + k) once per solve with LDS-tiled
// transposes.
//
// The line smoother never divides: the Thomas factors of D = T + 4 sigma I
// (cp_k and 1/(d_k + 4 sigma - a_k cp_{k-1})) depend only on the operator and
// are computed once per level at set-up with the same operations the
// recurrence would use, so results are bitwise those of the textbook sweep.
// The forward sweep stores d'_k in the output field and the backward sweep
// finishes it in place, so a thread holds no per-layer arrays.
//
// One V-cycle over all right-hand sides is captured into a hipGraph on first
// use and replayed; each level's kernels use the workgroup size chosen by a
// short timing sweep at set-up (the autotuner).
#include <algorithm>
#include <cstdio>
#include <cstdlib>
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
  void* cp_d = nullptr;             // Thomas factors of D, [k][ncell]
  void* iv_d = nullptr;
  void* adui_d = nullptr;           // (a, d, u, iv) interleaved, [k][ncell] records of 4 values
  void* x_d = nullptr;              // [k][ncell][rhs]
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
constexpr int kMaxRhs = 32;

template <typename T>
__device__ __forceinline__ T horiz(const T* __restrict__ xk, const int32_t* __restrict__ nb, int c, int R, int r,
                                   T xc, T sigma) {
  const T s4 = xk[(int64_t)nb[4 * c + 0] * R + r] + xk[(int64_t)nb[4 * c + 1] * R + r];
  const T s5 = s4 + xk[(int64_t)nb[4 * c + 2] * R + r];
  const T sn = s5 + xk[(int64_t)nb[4 * c + 3] * R + r];
  const T four = (T)4 * xc;
  return sigma * (four - sn);
}

// The forward sweep's four per-cell coefficients of one layer in one record,
// so that a wave reads them with one vector load instead of four scalar ones
template <typename T>
struct alignas(4 * sizeof(T)) Band4 {
  T a, d, u, iv;
};

template <typename T>
__global__ void k_mg_interleave(Band4<T>* __restrict__ out, const T* __restrict__ ba, const T* __restrict__ bd,
                                const T* __restrict__ bu, const T* __restrict__ iv, int64_t n) {
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x)
    out[i] = Band4<T>{ba[i], bd[i], bu[i], iv[i]};
}

// Thomas factors of D = tridiag(a, d + 4 sigma, u), per column
template <typename T>
__global__ void k_mg_factor(T* __restrict__ cp, T* __restrict__ iv, const T* __restrict__ ba,
                            const T* __restrict__ bd, const T* __restrict__ bu, int ncell, int nl, T sigma) {
  const int c = blockIdx.x * blockDim.x + threadIdx.x;
  if (c >= ncell) return;
  T cprev = 0;
  for (int k = 0; k < nl; ++k) {
    const int64_t ik = (int64_t)k * ncell + c;
    const T dd = bd[ik] + (T)4 * sigma;
    const T inv = k == 0 ? T(1) / dd : T(1) / (dd - ba[ik] * cprev);
    iv[ik] = inv;
    cprev = inv * bu[ik];
    cp[ik] = cprev;
  }
}

// One damped line-Jacobi sweep: xo = xi + omega * D^{-1} (b - A xi); thread = (cell, rhs)
template <typename T>
__global__ void __launch_bounds__(256) k_mg_smooth(T* __restrict__ xo, const T* __restrict__ xi, const T* __restrict__ b,
                                                   const T* __restrict__ ba, const T* __restrict__ bd,
                                                   const T* __restrict__ bu, const T* __restrict__ cp,
                                                   const T* __restrict__ iv, const int32_t* __restrict__ nb,
                                                   int ncell, int nl, int lgR, T sigma, T omega) {
  const int R = 1 << lgR;
  const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= (int64_t)ncell * R) return;
  const int c = (int)(t >> lgR), r = (int)(t & (R - 1));
  const int64_t plane = (int64_t)ncell * R;
  T xm = 0, x0 = xi[t], dprev = 0;
  for (int k = 0; k < nl; ++k) {
    const int64_t i = (int64_t)k * plane + t;
    const int64_t ik = (int64_t)k * ncell + c;
    const T xp = (k + 1 < nl) ? xi[i + plane] : T(0);
    const T ak = ba[ik], dk = bd[ik], uk = bu[ik];
    T ax = ak * xm;
    ax = ax + dk * x0;
    ax = ax + uk * xp;
    ax = ax + horiz(xi + (int64_t)k * plane, nb, c, R, r, x0, sigma);
    const T res = b[i] - ax;
    const T dp = k == 0 ? iv[ik] * res : iv[ik] * (res - ak * dprev);
    xo[i] = dp;
    dprev = dp;
    xm = x0;
    x0 = xp;
  }
  T dnext = 0;
  for (int k = nl - 1; k >= 0; --k) {
    const int64_t i = (int64_t)k * plane + t;
    T dp = xo[i];
    if (k < nl - 1) dp = dp - cp[(int64_t)k * ncell + c] * dnext;
    xo[i] = xi[i] + omega * dp;
    dnext = dp;
  }
}

// bc(C,k) = sum over the 4 children f (y outer, x inner) of r(f,k) * 1/4,
// r = b - A x on the fine level.  Also zeroes the coarse solution.
template <typename T>
__global__ void __launch_bounds__(256) k_mg_restrict(T* __restrict__ bc, T* __restrict__ xc, const T* __restrict__ xf,
                                                     const T* __restrict__ bf, const Band4<T>* __restrict__ adui,
                                                     const int32_t* __restrict__ nb, const int32_t* __restrict__ children,
                                                     int ncell_c, int ncell_f, int nl, int lgR, T sigma,
                                                     unsigned nb_pad) {
  const int R = 1 << lgR;
  // XCD-contiguous blocks, as the smoother: the fine neighbours across the
  // children's N/S faces are then read on the same XCD
  const int64_t t = (int64_t)lfr_xcd_block(blockIdx.x, nb_pad) * blockDim.x + threadIdx.x;
  if (t >= (int64_t)ncell_c * R) return;
  const int C = (int)(t >> lgR), r = (int)(t & (R - 1));
  const int64_t pf = (int64_t)ncell_f * R, pc = (int64_t)ncell_c * R;
  const T denom = T(1) / T(4);
  // in-plane offsets of each child and of its four neighbours, loaded once
  // (ncell * R < 2^31 is checked at set-up)
  int ch[4], of[4], on[4][4];
#pragma unroll
  for (int q = 0; q < 4; ++q) {
    ch[q] = children[4 * C + q];
    of[q] = ch[q] * R + r;
#pragma unroll
    for (int j = 0; j < 4; ++j) on[q][j] = nb[4 * ch[q] + j] * R + r;
  }
  // one level's operands for the four children; two sets in turn, so that
  // the loads of level k + 1 are in flight while level k is summed.  The
  // load of x above the top layer is clamped to the top layer (in bounds)
  // and its value replaced by 0.
  struct Lv {
    T xp[4], bf[4], h[4][4];
    Band4<T> bv[4];
  };
  auto load = [&](int k, Lv& v) {
    const int kp = min(k + 1, nl - 1);
    const T* xk = xf + (int64_t)k * pf;
#pragma unroll
    for (int q = 0; q < 4; ++q) {
      v.xp[q] = xf[(int64_t)kp * pf + of[q]];
      v.bv[q] = adui[(int64_t)k * ncell_f + ch[q]];
#pragma unroll
      for (int j = 0; j < 4; ++j) v.h[q][j] = xk[on[q][j]];
      v.bf[q] = bf[(int64_t)k * pf + of[q]];
    }
  };
  // each child's x at levels k - 1 and k, carried from level to level
  T xmv[4], x0v[4];
#pragma unroll
  for (int q = 0; q < 4; ++q) {
    xmv[q] = T(0);
    x0v[q] = xf[of[q]];
  }
  auto level = [&](int k, const Lv& v) {
    T acc = 0;
#pragma unroll
    for (int q = 0; q < 4; ++q) {
      const T x0 = x0v[q];
      const T xm = xmv[q];
      const T xp = k + 1 < nl ? v.xp[q] : T(0);
      const Band4<T> bv = v.bv[q];
      T ax = bv.a * xm;
      ax = ax + bv.d * x0;
      ax = ax + bv.u * xp;
      const T s4 = v.h[q][0] + v.h[q][1];
      const T s5 = s4 + v.h[q][2];
      const T sn = s5 + v.h[q][3];
      const T four = (T)4 * x0;
      ax = ax + sigma * (four - sn);
      const T res = v.bf[q] - ax;
      acc = acc + res * denom;
      xmv[q] = x0;
      x0v[q] = xp;
    }
    bc[(int64_t)k * pc + t] = acc;
    xc[(int64_t)k * pc + t] = T(0);
  };
  Lv A, B;
  load(0, A);
  for (int k = 0; k < nl; k += 2) {
    load(min(k + 1, nl - 1), B);
    level(k, A);
    load(min(k + 2, nl - 1), A);
    if (k + 1 < nl) level(k + 1, B);
  }
}

// The same restriction with one lane per fine child and right-hand side
// (4R consecutive lanes per coarse cell, R <= 16): each lane forms its
// child's scaled residual, the child-0 lane gathers the other three by lane
// shuffles and adds them in child order, acc = (((0 + v0) + v1) + v2) + v3,
// the statement order of the cell-at-a-time form, so the bits are the same.
// Four times the lanes of that form, each with a quarter of its loads; for
// a fixed layer count, unrolled with the loads PF levels ahead.
template <typename T, int NL, int PF>
__global__ void __launch_bounds__(256) k_mg_restrict_q(T* __restrict__ bc, T* __restrict__ xc, const T* __restrict__ xf,
                                                       const T* __restrict__ bf, const Band4<T>* __restrict__ adui,
                                                       const int32_t* __restrict__ nb,
                                                       const int32_t* __restrict__ children, int ncell_c, int ncell_f,
                                                       int lgR, T sigma, unsigned nb_pad) {
  const int R = 1 << lgR;
  const int64_t g = (int64_t)lfr_xcd_block(blockIdx.x, nb_pad) * blockDim.x + threadIdx.x;
  // every lane stays for the shuffles; lanes past the last coarse cell work
  // on cell 0 and do not store
  const bool valid = g < (int64_t)ncell_c * 4 * R;
  const int64_t gg = valid ? g : 0;
  const int C = (int)(gg >> (lgR + 2));
  const int q = (int)(gg >> lgR) & 3;
  const int r = (int)gg & (R - 1);
  const int64_t pf = (int64_t)ncell_f * R, pc = (int64_t)ncell_c * R;
  const T denom = T(1) / T(4);
  const int f = children[4 * C + q];
  const int of = f * R + r;   // ncell * R < 2^31 is checked at set-up
  const int o0 = nb[4 * f + 0] * R + r, o1 = nb[4 * f + 1] * R + r;
  const int o2 = nb[4 * f + 2] * R + r, o3 = nb[4 * f + 3] * R + r;
  T xp[PF + 1], bfv[PF + 1], h0[PF + 1], h1[PF + 1], h2[PF + 1], h3[PF + 1];
  Band4<T> bv[PF + 1];
  auto load = [&](int k, int s) {
    const T* xk = xf + (int64_t)k * pf;
    if (k + 1 < NL) xp[s] = xf[(int64_t)(k + 1) * pf + of];
    bv[s] = adui[(int64_t)k * ncell_f + f];
    h0[s] = xk[o0];
    h1[s] = xk[o1];
    h2[s] = xk[o2];
    h3[s] = xk[o3];
    bfv[s] = bf[(int64_t)k * pf + of];
  };
  T xm = T(0), x0 = xf[of];
#pragma unroll
  for (int k = 0; k < PF && k < NL; ++k) load(k, k);
#pragma clang loop unroll(full)
  for (int k = 0; k < NL; ++k) {
    if (k + PF < NL) load(k + PF, (k + PF) % (PF + 1));
    const int s = k % (PF + 1);
    const T xpv = k + 1 < NL ? xp[s] : T(0);
    T ax = bv[s].a * xm;
    ax = ax + bv[s].d * x0;
    ax = ax + bv[s].u * xpv;
    const T s4 = h0[s] + h1[s];
    const T s5 = s4 + h2[s];
    const T sn = s5 + h3[s];
    const T four = (T)4 * x0;
    ax = ax + sigma * (four - sn);
    const T res = bfv[s] - ax;
    const T v = res * denom;
    const T v1 = __shfl_down(v, R);
    const T v2 = __shfl_down(v, 2 * R);
    const T v3 = __shfl_down(v, 3 * R);
    if (valid && q == 0) {
      T acc = 0;
      acc = acc + v;
      acc = acc + v1;
      acc = acc + v2;
      acc = acc + v3;
      const int64_t o = (int64_t)k * pc + (int64_t)C * R + r;
      bc[o] = acc;
      xc[o] = T(0);
    }
    xm = x0;
    x0 = xpv;
    __builtin_amdgcn_sched_barrier(0);
  }
}

// xf(f,k) += xc(parent(f), k)
template <typename T>
__global__ void __launch_bounds__(256) k_mg_prolong(T* __restrict__ xf, const T* __restrict__ xc,
                                                    const int32_t* __restrict__ parent, int ncell_f, int ncell_c,
                                                    int nl, int lgR) {
  const int R = 1 << lgR;
  const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= (int64_t)ncell_f * R) return;
  const int f = (int)(t >> lgR), r = (int)(t & (R - 1));
  const int64_t src = (int64_t)parent[f] * R + r;
  const int64_t pf = (int64_t)ncell_f * R, pc = (int64_t)ncell_c * R;
  // levels in groups of 8: the group's loads are issued before its stores
  int k = 0;
  for (; k + 8 <= nl; k += 8) {
    T a[8], b[8];
#pragma unroll
    for (int u = 0; u < 8; ++u) {
      a[u] = xf[(int64_t)(k + u) * pf + t];
      b[u] = xc[(int64_t)(k + u) * pc + src];
    }
#pragma unroll
    for (int u = 0; u < 8; ++u) xf[(int64_t)(k + u) * pf + t] = a[u] + b[u];
  }
  for (; k < nl; ++k) xf[(int64_t)k * pf + t] += xc[(int64_t)k * pc + src];
}

// per-rhs partial sums of r^2, r = b - A x; part[blk*R + r]
template <typename T>
__global__ void __launch_bounds__(lfr::kRedThreads)
k_mg_resnorm(const T* __restrict__ xf, const T* __restrict__ bf, const T* __restrict__ ba,
             const T* __restrict__ bd, const T* __restrict__ bu, const int32_t* __restrict__ nb,
             int ncell, int nl, int lgR, T sigma, double* __restrict__ part) {
  __shared__ double lds[lfr::kRedThreads / 64][kMaxRhs];
  const int R = 1 << lgR;
  const int64_t plane = (int64_t)ncell * R;
  const int64_t n = plane * nl;
  double s = 0.0;
  for (int64_t i = (int64_t)blockIdx.x * lfr::kRedThreads + threadIdx.x; i < n;
       i += (int64_t)gridDim.x * lfr::kRedThreads) {
    const int k = (int)(i / plane);
    const int64_t t = i - (int64_t)k * plane;
    const int c = (int)(t >> lgR), r = (int)(t & (R - 1));
    const int64_t ik = (int64_t)k * ncell + c;
    const T x0 = xf[i];
    const T xm = k > 0 ? xf[i - plane] : T(0);
    const T xp = k + 1 < nl ? xf[i + plane] : T(0);
    T ax = ba[ik] * xm;
    ax = ax + bd[ik] * x0;
    ax = ax + bu[ik] * xp;
    ax = ax + horiz(xf + (int64_t)k * plane, nb, c, R, r, x0, sigma);
    const double res = (double)(bf[i] - ax);
    s += res * res;
  }
  // lanes l and l + R, l + 2R, ... hold the same right-hand side
  for (int off = 32; off >= R; off >>= 1) s += __shfl_xor(s, off, 64);
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  if (lane < R) lds[w][lane] = s;
  __syncthreads();
  if (threadIdx.x < R) {
    double v = 0.0;
    for (int q = 0; q < lfr::kRedThreads / 64; ++q) v += lds[q][threadIdx.x];
    part[(int64_t)blockIdx.x * R + threadIdx.x] = v;
  }
}

// out[r] = sum over blocks (fixed order) of part[blk*R + r]; one block of 256
__global__ void __launch_bounds__(256) k_mg_resnorm_final(const double* __restrict__ part, int nb, int R,
                                                          double* __restrict__ out) {
  __shared__ double lds[256];
  const int r = threadIdx.x % R, j = threadIdx.x / R, nj = 256 / R;
  double s = 0.0;
  for (int bb = j; bb < nb; bb += nj) s += part[(int64_t)bb * R + r];
  lds[threadIdx.x] = s;
  __syncthreads();
  if (threadIdx.x < R) {
    double t = 0.0;
    for (int q = 0; q < nj; ++q) t += lds[q * R + threadIdx.x];
    out[threadIdx.x] = t;
  }
}

// LFRic W3 fields (field r at r*undf, dof = c*nl + k) -> [k][c][r], a tile of
// TC cells per block staged through LDS (contiguous reads and writes).
template <typename T>
__global__ void __launch_bounds__(256) k_to_mg(T* __restrict__ out, const T* __restrict__ in, int ncell, int nl,
                                               int R, int TC) {
  extern __shared__ char smem_raw[];
  T* tile = (T*)smem_raw;          // [R][TC*nl]
  const int c0 = blockIdx.x * TC;
  const int tc = min(TC, ncell - c0);
  const int64_t undf = (int64_t)ncell * nl;
  const int span = tc * nl;
  for (int idx = threadIdx.x; idx < R * span; idx += blockDim.x) {
    const int r = idx / span, j = idx - r * span;
    tile[r * TC * nl + j] = in[(int64_t)r * undf + (int64_t)c0 * nl + j];
  }
  __syncthreads();
  for (int idx = threadIdx.x; idx < R * span; idx += blockDim.x) {
    const int r = idx % R, rest = idx / R;
    const int cc = rest % tc, k = rest / tc;
    out[((int64_t)k * ncell + c0 + cc) * R + r] = tile[r * TC * nl + cc * nl + k];
  }
}

template <typename T>
__global__ void __launch_bounds__(256) k_from_mg(T* __restrict__ out, const T* __restrict__ in, int ncell, int nl,
                                                 int R, int TC) {
  extern __shared__ char smem_raw[];
  T* tile = (T*)smem_raw;
  const int c0 = blockIdx.x * TC;
  const int tc = min(TC, ncell - c0);
  const int64_t undf = (int64_t)ncell * nl;
  const int span = tc * nl;
  for (int idx = threadIdx.x; idx < R * span; idx += blockDim.x) {
    const int r = idx % R, rest = idx / R;
    const int cc = rest % tc, k = rest / tc;
    tile[r * TC * nl + cc * nl + k] = in[((int64_t)k * ncell + c0 + cc) * R + r];
  }
  __syncthreads();
  for (int idx = threadIdx.x; idx < R * span; idx += blockDim.x) {
    const int r = idx / span, j = idx - r * span;
    out[(int64_t)r * undf + (int64_t)c0 * nl + j] = tile[r * TC * nl + j];
  }
}

// The same sweep for a compile-time layer count: the layers are unrolled,
// every load (x of the layer, its four neighbours, b, the bands and the
// Thomas factor) is issued PF layers ahead of the recurrence (scheduling
// region g holds the loads of layers [g PF, (g+1) PF) and the sweep of the PF
// layers before them), and d' stays in registers, so xo is written once
// instead of written, re-read and rewritten.  Same operations in the same
// order as k_mg_smooth.
template <typename T, int NL, int PF = 4, bool KX = false>
__global__ void __launch_bounds__(256) k_mg_smooth_nl(T* __restrict__ xo, const T* __restrict__ xi,
                                                      const T* __restrict__ b, const Band4<T>* __restrict__ adui,
                                                      const T* __restrict__ cp, const int32_t* __restrict__ nb,
                                                      int ncell, int lgR, T sigma, T omega, unsigned nb_pad) {
  const int R = 1 << lgR;
  // XCD-contiguous blocks: the N and S neighbours of a cell (a panel row
  // away) are then read on the same XCD as the cell, mostly from its L2
  const int64_t t = (int64_t)lfr_xcd_block(blockIdx.x, nb_pad) * blockDim.x + threadIdx.x;
  if (t >= (int64_t)ncell * R) return;
  const int c = (int)(t >> lgR), r = (int)(t & (R - 1));
  const int64_t plane = (int64_t)ncell * R;
  const int64_t o0 = (int64_t)nb[4 * c + 0] * R + r, o1 = (int64_t)nb[4 * c + 1] * R + r;
  const int64_t o2 = (int64_t)nb[4 * c + 2] * R + r, o3 = (int64_t)nb[4 * c + 3] * R + r;
  T lx[NL], lb[NL], la[NL], ld[NL], lu[NL], li[NL], h0[NL], h1[NL], h2[NL], h3[NL];
  T dp[NL];
  T dprev = 0;
#pragma unroll
  for (int q = 0; q < NL + PF; ++q) {
    if (q < NL) {
      const int k = q;
      const T* xk = xi + (int64_t)k * plane;
      const int64_t ik = (int64_t)k * ncell + c;
      lx[k] = xk[t];
      lb[k] = b[(int64_t)k * plane + t];
      const Band4<T> bv = adui[ik];
      la[k] = bv.a;
      ld[k] = bv.d;
      lu[k] = bv.u;
      li[k] = bv.iv;
      h0[k] = xk[o0];
      h1[k] = xk[o1];
      h2[k] = xk[o2];
      h3[k] = xk[o3];
    }
    if (q >= PF) {
      const int k = q - PF;
      const T x0 = lx[k];
      const T xm = k > 0 ? lx[k > 0 ? k - 1 : 0] : T(0);
      const T xp = k + 1 < NL ? lx[k + 1 < NL ? k + 1 : k] : T(0);
      const T ak = la[k];
      T ax = ak * xm;
      ax = ax + ld[k] * x0;
      ax = ax + lu[k] * xp;
      const T s4 = h0[k] + h1[k];
      const T s5 = s4 + h2[k];
      const T sn = s5 + h3[k];
      const T four = (T)4 * x0;
      ax = ax + sigma * (four - sn);
      const T res = lb[k] - ax;
      const T d = k == 0 ? li[k] * res : li[k] * (res - ak * dprev);
      dp[k] = d;
      dprev = d;
    }
    if ((q + 1) % PF == 0) __builtin_amdgcn_sched_barrier(0);
  }
  // Backward sweep.  cp (and x when it is not kept) of a chunk of CH layers
  // is loaded one chunk ahead of its use: loaded just in time, each layer's
  // load would wait for the store issued before it (the vector memory
  // counter covers both), one memory round trip per layer.
  constexpr int CH = sizeof(T) == 4 ? 8 : 4;
  T cpv[NL], xv[NL];
  auto load_chunk = [&](int hi, int lo) {   // layers hi down to lo (inclusive, clipped to >= 0)
#pragma unroll
    for (int kk = hi; kk >= lo && kk >= 0; --kk) {
      if (kk < NL - 1) cpv[kk] = cp[(int64_t)kk * ncell + c];
      if (!KX) xv[kk] = xi[(int64_t)kk * plane + t];
    }
  };
  load_chunk(NL - 1, NL - CH);
  T dnext = 0;
#pragma unroll
  for (int k = NL - 1; k >= 0; --k) {
    if ((NL - 1 - k) % CH == 0) {
      // this chunk is layers k .. k - CH + 1; fetch the next one
      load_chunk(k - CH, k - 2 * CH + 1);
      __builtin_amdgcn_sched_barrier(0);
    }
    const int64_t i = (int64_t)k * plane + t;
    T d = dp[k];
    if (k < NL - 1) d = d - cpv[k] * dnext;
    xo[i] = (KX ? lx[k] : xv[k]) + omega * d;
    dnext = d;
  }
}

int lg2(int v) {
  int l = 0;
  while ((1 << l) < v) ++l;
  return l;
}

template <typename T>
void launch_smooth(hipStream_t s, lfr_mg* mg, lfr_mg_level& L, T* xo, const T* xi) {
  const int bs = L.block_smooth;
  const int64_t n = (int64_t)L.ncell * mg->nrhs;
  if (mg->nl == 70) {
    // FP64: two layers ahead (the registers of four would halve occupancy)
    // and x read again in the backward sweep
    const unsigned nbp = ((unsigned)((n + bs - 1) / bs) + 7u) & ~7u;
    // FP32 keeps every layer's x in registers for the backward sweep
    // instead of reading it again (228 VGPRs; measured 4.19 -> 4.09 ms per
    // C192 V-cycle); LFR_MG_KEEPX=0 reads it again
    static const bool kx = [] {
      const char* e = getenv("LFR_MG_KEEPX");
      return !(e && e[0] == '0');
    }();
    if constexpr (sizeof(T) == 4) {
      if (kx) {
        hipLaunchKernelGGL((k_mg_smooth_nl<T, 70, 4, true>), dim3(nbp), dim3(bs), 0, s, xo, xi, (const T*)L.b_d,
                           (const Band4<T>*)L.adui_d, (const T*)L.cp_d, L.cell_next_d, L.ncell, lg2(mg->nrhs),
                           (T)L.sigma, (T)mg->omega, nbp);
        return;
      }
    }
    hipLaunchKernelGGL((k_mg_smooth_nl<T, 70, sizeof(T) == 4 ? 4 : 2>), dim3(nbp), dim3(bs), 0, s, xo, xi,
                         (const T*)L.b_d, (const Band4<T>*)L.adui_d, (const T*)L.cp_d, L.cell_next_d, L.ncell,
                         lg2(mg->nrhs), (T)L.sigma, (T)mg->omega, nbp);
    return;
  }
  hipLaunchKernelGGL(k_mg_smooth<T>, dim3((unsigned)((n + bs - 1) / bs)), dim3(bs), 0, s, xo, xi, (const T*)L.b_d,
                     (const T*)L.a_d, (const T*)L.d_d, (const T*)L.u_d, (const T*)L.cp_d, (const T*)L.iv_d,
                     L.cell_next_d, L.ncell, mg->nl, lg2(mg->nrhs), (T)L.sigma, (T)mg->omega);
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
  const int lgR = lg2(mg->nrhs);
  {
    const int bs = L.block_other;
    const int64_t n = (int64_t)C.ncell * mg->nrhs;
    const unsigned nbp = ((unsigned)((n + bs - 1) / bs) + 7u) & ~7u;
    // 70 layers and at most 16 right-hand sides: one lane per child (C192
    // FP32 V-cycle 3.03 against 3.05 ms, FP64 7.36 against 7.43 ms for the
    // cell-at-a-time form unrolled the same way); LFR_MG_RESTRICT=0 selects
    // the cell-at-a-time loop
    static const int rq = [] {
      const char* e = getenv("LFR_MG_RESTRICT");
      return e ? atoi(e) : 1;
    }();
    if (mg->nl == 70 && rq == 1 && mg->nrhs <= 16) {
      const unsigned nbq = ((unsigned)((4 * n + bs - 1) / bs) + 7u) & ~7u;
      hipLaunchKernelGGL((k_mg_restrict_q<T, 70, 4>), dim3(nbq), dim3(bs), 0, s, (T*)C.b_d, (T*)C.x_d,
                         (const T*)L.x_d, (const T*)L.b_d, (const Band4<T>*)L.adui_d, L.cell_next_d, C.children_d,
                         C.ncell, L.ncell, lgR, (T)L.sigma, nbq);
    } else
      hipLaunchKernelGGL(k_mg_restrict<T>, dim3(nbp), dim3(bs), 0, s, (T*)C.b_d, (T*)C.x_d, (const T*)L.x_d,
                         (const T*)L.b_d, (const Band4<T>*)L.adui_d, L.cell_next_d, C.children_d, C.ncell, L.ncell,
                         mg->nl, lgR, (T)L.sigma, nbp);
  }
  vcycle<T>(s, mg, l + 1);
  {
    const int bs = L.block_other;
    const int64_t n = (int64_t)L.ncell * mg->nrhs;
    hipLaunchKernelGGL(k_mg_prolong<T>, dim3((unsigned)((n + bs - 1) / bs)), dim3(bs), 0, s, (T*)L.x_d,
                       (const T*)C.x_d, L.parent_d, L.ncell, C.ncell, mg->nl, lgR);
  }
  for (int i = 0; i < mg->nu2 / 2; ++i) smooth2<T>(s, mg, L);
}

size_t esz(int prec) { return prec == 4 ? 4 : 8; }

void mg_free(lfr_mg* mg) {
  if (!mg) return;
  if (mg->graph) (void)hipGraphExecDestroy(mg->graph);
  for (auto& L : mg->lev)
    for (void* p : {(void*)L.cell_next_d, (void*)L.children_d, (void*)L.parent_d, L.a_d, L.d_d, L.u_d,
                    L.cp_d, L.iv_d, L.adui_d, L.x_d, L.b_d, L.t_d})
      if (p) (void)hipFree(p);
  if (mg->part_d) (void)hipFree(mg->part_d);
  delete mg;
}

// Pick the workgroup size of the smoother (the dominant kernel) for each
// level through the autotuner (one sweep per candidate size, identical
// results), keyed by the level's shape so that levels of the same size and
// later hierarchies reuse the choice.
template <typename T>
int autotune(lfr_mg* mg) {
  lfr_ctx* ctx = mg->ctx;
  const int cand[3] = {64, 128, 256};   // launch bound of the smoother is 256
  for (auto& L : mg->lev) {
    char key[128];
    snprintf(key, sizeof key, "mg smooth prec=%d nl=%d ncell=%d rhs=%d", (int)sizeof(T), mg->nl, L.ncell,
             mg->nrhs);
    const auto run = [&](int c) -> int {
      L.block_smooth = cand[c];
      launch_smooth<T>(ctx->stream, mg, L, (T*)L.t_d, (const T*)L.x_d);
      return hipGetLastError() == hipSuccess ? LFR_OK : LFR_ERR_HIP;
    };
    L.block_smooth = cand[lfr_tune_pick(ctx, key, 3, 2, run)];
    L.block_other = 256;
  }
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
                       nrhs >= 1 && nrhs <= kMaxRhs && (nrhs & (nrhs - 1)) == 0,
              "lfr_mg_create: bad sizes (prec %d nlev %d nlayers %d nrhs %d; nrhs a power of 2 <= %d)", prec,
              nlev, nlayers, nrhs, kMaxRhs);
  for (int l = 0; l < nlev; ++l)
    LFR_REQUIRE(ctx, ncell[l] > 0 && (int64_t)ncell[l] * nrhs < ((int64_t)1 << 31),
                "lfr_mg_create: level %d has %d cells (ncell * nrhs must be below 2^31)", l, ncell[l]);
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
              hipMalloc(&L.u_d, nb * s) == hipSuccess && hipMalloc(&L.cp_d, nb * s) == hipSuccess &&
              hipMalloc(&L.iv_d, nb * s) == hipSuccess && hipMalloc(&L.adui_d, nb * 4 * s) == hipSuccess &&
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
    {
      const unsigned g = (unsigned)((nc + 255) / 256);
      if (prec == 4)
        hipLaunchKernelGGL(k_mg_factor<float>, dim3(g), dim3(256), 0, ctx->stream, (float*)L.cp_d, (float*)L.iv_d,
                           (const float*)L.a_d, (const float*)L.d_d, (const float*)L.u_d, nc, nlayers,
                           (float)L.sigma);
      else
        hipLaunchKernelGGL(k_mg_factor<double>, dim3(g), dim3(256), 0, ctx->stream, (double*)L.cp_d,
                           (double*)L.iv_d, (const double*)L.a_d, (const double*)L.d_d, (const double*)L.u_d, nc,
                           nlayers, L.sigma);
      const unsigned gi = (unsigned)std::min<size_t>((nb + 255) / 256, 8192);
      if (prec == 4)
        hipLaunchKernelGGL(k_mg_interleave<float>, dim3(gi), dim3(256), 0, ctx->stream, (Band4<float>*)L.adui_d,
                           (const float*)L.a_d, (const float*)L.d_d, (const float*)L.u_d, (const float*)L.iv_d,
                           (int64_t)nb);
      else
        hipLaunchKernelGGL(k_mg_interleave<double>, dim3(gi), dim3(256), 0, ctx->stream, (Band4<double>*)L.adui_d,
                           (const double*)L.a_d, (const double*)L.d_d, (const double*)L.u_d, (const double*)L.iv_d,
                           (int64_t)nb);
      if (hipGetLastError() != hipSuccess || hipStreamSynchronize(ctx->stream) != hipSuccess) {
        mg_free(mg);
        return lfr_set_error(ctx, LFR_ERR_HIP, "lfr_mg_create: factor kernel failed");
      }
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
  hipStream_t s = ctx->stream;
  const bool f32 = mg->prec == 4;
  const int R = mg->nrhs, lgR = lg2(R);
  const size_t es = f32 ? 4 : 8;
  int TC = (int)(32768 / ((size_t)mg->nl * R * es));
  if (TC > 16) TC = 16;
  if (TC < 1) TC = 1;
  const size_t smem = (size_t)TC * mg->nl * R * es;
  const unsigned gt = (unsigned)((F.ncell + TC - 1) / TC);
  if (f32) {
    hipLaunchKernelGGL(k_to_mg<float>, dim3(gt), dim3(256), smem, s, (float*)F.b_d, (const float*)b_d, F.ncell, mg->nl, R, TC);
    hipLaunchKernelGGL(k_to_mg<float>, dim3(gt), dim3(256), smem, s, (float*)F.x_d, (const float*)x_d, F.ncell, mg->nl, R, TC);
  } else {
    hipLaunchKernelGGL(k_to_mg<double>, dim3(gt), dim3(256), smem, s, (double*)F.b_d, (const double*)b_d, F.ncell, mg->nl, R, TC);
    hipLaunchKernelGGL(k_to_mg<double>, dim3(gt), dim3(256), smem, s, (double*)F.x_d, (const double*)x_d, F.ncell, mg->nl, R, TC);
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
  const int nblk = lfr::red_blocks((int64_t)F.ncell * mg->nl * R);
  auto resnorm = [&](double* dst) -> int {
    if (f32)
      hipLaunchKernelGGL(k_mg_resnorm<float>, dim3(nblk), dim3(lfr::kRedThreads), 0, s, (const float*)F.x_d,
                         (const float*)F.b_d, (const float*)F.a_d, (const float*)F.d_d, (const float*)F.u_d,
                         F.cell_next_d, F.ncell, mg->nl, lgR, (float)F.sigma, mg->part_d);
    else
      hipLaunchKernelGGL(k_mg_resnorm<double>, dim3(nblk), dim3(lfr::kRedThreads), 0, s, (const double*)F.x_d,
                         (const double*)F.b_d, (const double*)F.a_d, (const double*)F.d_d, (const double*)F.u_d,
                         F.cell_next_d, F.ncell, mg->nl, lgR, F.sigma, mg->part_d);
    double* sums = mg->part_d + (size_t)R * lfr::kRedMaxBlocks;
    hipLaunchKernelGGL(k_mg_resnorm_final, dim3(1), dim3(256), 0, s, mg->part_d, nblk, R, sums);
    LFR_CHECK_LAUNCH(ctx);
    LFR_HIP(ctx, hipMemcpyAsync(ctx->host_scalars, sums, R * sizeof(double), hipMemcpyDeviceToHost, s));
    LFR_HIP(ctx, hipStreamSynchronize(s));
    for (int r = 0; r < R; ++r) dst[r] = sqrt(ctx->host_scalars[r]);
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
    hipLaunchKernelGGL(k_from_mg<float>, dim3(gt), dim3(256), smem, s, (float*)x_d, (const float*)F.x_d, F.ncell, mg->nl, R, TC);
  else
    hipLaunchKernelGGL(k_from_mg<double>, dim3(gt), dim3(256), smem, s, (double*)x_d, (const double*)F.x_d, F.ncell, mg->nl, R, TC);
  LFR_CHECK_LAUNCH(ctx);
  return LFR_OK;
}

}  // extern "C"
