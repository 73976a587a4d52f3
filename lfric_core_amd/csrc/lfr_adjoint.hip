// Differentiable K-iteration CG: the map b -> x_K of K unmonitored CG
// iterations from x = 0 with the null preconditioner
// (conjugate_gradient_type%cg_solve,
// components/science/source/solver/sci_iterative_solver_mod.f90:495-592) and
// its exact reverse mode, on the device.
//
// Forward: the reference's statement sequence with FP64 scalars.  The tape
// keeps every scalar of the iteration (alpha_i, beta_i, <r_i,r_i>, <p_i,A p_i>),
// the operator products q_i = A p_i, and the iterates r, p only at the start
// of each segment of `segment` iterations (and throughout the last segment).
//
// Reverse: segment by segment from the last; the iterates r_i, p_i of a
// segment are recomputed from its checkpoint with the stored q_i and scalars
// (r_{i+1} = r_i - alpha_i q_i, p_{i+1} = beta_i p_i + r_{i+1}: the forward's
// arithmetic, so the same bits, and no matrix read), then each step is walked
// backwards:
//
//   rbar += pbar;  betabar = <pbar, p_i>;  pbar = beta_i pbar
//   rzbar_{i+1} = rzbar + betabar / rz_i;  rzbar_i = -betabar rz_{i+1} / rz_i^2
//   rbar += 2 rzbar_{i+1} r_{i+1}
//   alphabar = -<rbar, q_i> + <xbar, p_i>
//   qbar = -alpha_i rbar;  pbar += alpha_i xbar
//   rzbar_i += alphabar / pq_i;  pqbar = -alphabar rz_i / pq_i^2
//   pbar += pqbar q_i;  qbar += pqbar p_i
//   pbar += A^T qbar                    (transpose_matrix_vector)
//   Mbar += qbar (x) p_i per cell-layer  (once per segment, for all its steps)
//
// and finally bbar = rbar + pbar + 2 rzbar_0 b.  Each line keeps the
// reference-order roundings of separate vector operations; the kernels fuse
// neighbouring lines so each vector is read once per step.
#include <cmath>
#include <vector>

#include "lfr_internal.h"
#include "lfr_reduce.h"

namespace {

using lfr::kRedMaxBlocks;
using lfr::kRedThreads;
constexpr int kT = 256;
constexpr int kAlphaThreads = 1024;

// reverse-sweep scalar slots
enum RSlot { R_RZBAR = 0, R_RZBAR_I, R_C, R_XP, R_PQBAR, R_NSLOT = 8 };

size_t align_up(size_t b) { return (b + 255) & ~(size_t)255; }

// r = b, p = b, partials of <b, b>
template <typename T>
__global__ void __launch_bounds__(kRedThreads)
k_tape_init(T* __restrict__ r, T* __restrict__ p, const T* __restrict__ b, int64_t n, double* __restrict__ part) {
  __shared__ double lds[kRedThreads / 64];
  double s = 0.0;
  for (int64_t i = (int64_t)blockIdx.x * kRedThreads + threadIdx.x; i < n; i += (int64_t)gridDim.x * kRedThreads) {
    const T v = b[i];
    r[i] = v;
    p[i] = v;
    const double d = (double)v * (double)v;
    s = s + d;
  }
  s = lfr::block_sum(s, lds);
  if (threadIdx.x == 0) part[blockIdx.x] = s;
}

// out = sum of part[0..np) (one block, fixed order)
template <int BS>
__global__ void __launch_bounds__(BS) k_tape_sum(const double* __restrict__ part, int np, double* __restrict__ out) {
  __shared__ double lds[BS / 64];
  const double s = lfr::block_sum<BS>(lfr::strided_sum<BS>(part, np), lds);
  if (threadIdx.x == 0) out[0] = s;
}

// pq_i = <p_i, q_i> from the apply's partials; alpha_i = rz_i / pq_i
__global__ void __launch_bounds__(kAlphaThreads)
k_tape_alpha(const double* __restrict__ part, int np, double* __restrict__ pq, double* __restrict__ alpha,
             const double* __restrict__ rz, int i) {
  __shared__ double lds[kAlphaThreads / 64];
  const double s = lfr::block_sum<kAlphaThreads>(lfr::strided_sum<kAlphaThreads>(part, np), lds);
  if (threadIdx.x == 0) {
    pq[i] = s;
    alpha[i] = rz[i] / s;
  }
}

// x += alpha p (x may be null); r_out = r_in - alpha q; partials of <r_out, r_out>
template <typename T>
__global__ void __launch_bounds__(kRedThreads)
k_tape_update(T* __restrict__ x, const T* __restrict__ r_in, T* __restrict__ r_out, const T* __restrict__ p,
              const T* __restrict__ q, int64_t n, const double* __restrict__ alpha, int i, double* __restrict__ part) {
  __shared__ double lds[kRedThreads / 64];
  const T a = (T)alpha[i], ma = (T)(-alpha[i]);
  double s = 0.0;
  for (int64_t e = (int64_t)blockIdx.x * kRedThreads + threadIdx.x; e < n; e += (int64_t)gridDim.x * kRedThreads) {
    if (x) {
      const T xp = a * p[e];
      x[e] = x[e] + xp;
    }
    const T rq = ma * q[e];
    const T rv = r_in[e] + rq;
    r_out[e] = rv;
    const double d = (double)rv * (double)rv;
    s = s + d;
  }
  s = lfr::block_sum(s, lds);
  if (threadIdx.x == 0) part[blockIdx.x] = s;
}

// rz_{i+1} from the update's partials (every block, one fixed order), beta_i =
// rz_{i+1} / rz_i (block 0 records both); p_out = beta_i p_in + r
template <typename T>
__global__ void __launch_bounds__(kRedThreads)
k_tape_direction(T* __restrict__ p_out, const T* __restrict__ p_in, const T* __restrict__ r, int64_t n,
                 const double* __restrict__ part, int np, double* __restrict__ rz, double* __restrict__ beta, int i) {
  __shared__ double lds[kRedThreads / 64];
  __shared__ double bsh;
  const double rr = lfr::block_sum(lfr::strided_sum(part, np), lds);
  if (threadIdx.x == 0) {
    const double bt = rr / rz[i];
    bsh = bt;
    if (blockIdx.x == 0) {
      rz[i + 1] = rr;
      beta[i] = bt;
    }
  }
  __syncthreads();
  const T bv = (T)bsh;
  for (int64_t e = (int64_t)blockIdx.x * kRedThreads + threadIdx.x; e < n; e += (int64_t)gridDim.x * kRedThreads) {
    const T pb = bv * p_in[e];
    p_out[e] = pb + r[e];
  }
}

// recompute one step from stored q_i and scalars: r_out = r_in - alpha_i q_i;
// p_out = beta_i p_in + r_out (the forward's expressions)
template <typename T>
__global__ void __launch_bounds__(kT)
k_tape_recompute(const T* __restrict__ r_in, T* __restrict__ r_out, const T* __restrict__ p_in, T* __restrict__ p_out,
                 const T* __restrict__ q, int64_t n, const double* __restrict__ alpha, const double* __restrict__ beta,
                 int i) {
  const T ma = (T)(-alpha[i]);
  const T bv = (T)beta[i];
  for (int64_t e = (int64_t)blockIdx.x * kT + threadIdx.x; e < n; e += (int64_t)gridDim.x * kT) {
    const T rq = ma * q[e];
    const T rv = r_in[e] + rq;
    r_out[e] = rv;
    const T pb = bv * p_in[e];
    p_out[e] = pb + rv;
  }
}

// B1: pbar += t (the previous step's A^T qbar; skipped on the first step);
// partials of <pbar, p_i> and <xbar, p_i>
template <typename T>
__global__ void __launch_bounds__(kRedThreads)
k_rev_dots(T* __restrict__ pbar, const T* __restrict__ t, const T* __restrict__ xbar, const T* __restrict__ p,
           int64_t n, int add_t, double* __restrict__ partA, double* __restrict__ partB) {
  __shared__ double lds[kRedThreads / 64];
  double sa = 0.0, sb = 0.0;
  for (int64_t e = (int64_t)blockIdx.x * kRedThreads + threadIdx.x; e < n; e += (int64_t)gridDim.x * kRedThreads) {
    T pb = pbar[e];
    if (add_t) {
      pb = pb + t[e];
      pbar[e] = pb;
    }
    const double pv = (double)p[e];
    const double da = (double)pb * pv;
    sa = sa + da;
    const double db = (double)xbar[e] * pv;
    sb = sb + db;
  }
  sa = lfr::block_sum(sa, lds);
  sb = lfr::block_sum(sb, lds);
  if (threadIdx.x == 0) {
    partA[blockIdx.x] = sa;
    partB[blockIdx.x] = sb;
  }
}

// B2: betabar, <xbar, p_i>, the rzbar updates and c = 2 rzbar_{i+1}
__global__ void __launch_bounds__(kRedThreads)
k_rev_s1(const double* __restrict__ partA, const double* __restrict__ partB, int np, double* __restrict__ R,
         const double* __restrict__ rz, int i) {
  __shared__ double lds[kRedThreads / 64];
  const double betabar = lfr::block_sum(lfr::strided_sum(partA, np), lds);
  const double xp = lfr::block_sum(lfr::strided_sum(partB, np), lds);
  if (threadIdx.x == 0) {
    const double rzbar_next = R[R_RZBAR] + betabar / rz[i];
    R[R_RZBAR_I] = -betabar * rz[i + 1] / (rz[i] * rz[i]);
    R[R_C] = 2.0 * rzbar_next;
    R[R_XP] = xp;
  }
}

// B3: rbar = (rbar + pbar) + c r_{i+1}; partials of <rbar, q_i>
template <typename T>
__global__ void __launch_bounds__(kRedThreads)
k_rev_rbar(T* __restrict__ rbar, const T* __restrict__ pbar, const T* __restrict__ rn, const T* __restrict__ q,
           int64_t n, const double* __restrict__ R, double* __restrict__ part) {
  __shared__ double lds[kRedThreads / 64];
  const T c = (T)R[R_C];
  double s = 0.0;
  for (int64_t e = (int64_t)blockIdx.x * kRedThreads + threadIdx.x; e < n; e += (int64_t)gridDim.x * kRedThreads) {
    const T a = rbar[e] + pbar[e];
    const T cr = c * rn[e];
    const T rb = a + cr;
    rbar[e] = rb;
    const double d = (double)rb * (double)q[e];
    s = s + d;
  }
  s = lfr::block_sum(s, lds);
  if (threadIdx.x == 0) part[blockIdx.x] = s;
}

// B4: alphabar = -<rbar, q_i> + <xbar, p_i>; rzbar_i += alphabar / pq_i;
// pqbar = -alphabar rz_i / pq_i^2; rzbar = rzbar_i
__global__ void __launch_bounds__(kRedThreads)
k_rev_s2(const double* __restrict__ part, int np, double* __restrict__ R, const double* __restrict__ rz,
         const double* __restrict__ pq, int i) {
  __shared__ double lds[kRedThreads / 64];
  const double rq = lfr::block_sum(lfr::strided_sum(part, np), lds);
  if (threadIdx.x == 0) {
    const double alphabar = -rq + R[R_XP];
    const double rzbar_i = R[R_RZBAR_I] + alphabar / pq[i];
    R[R_PQBAR] = -alphabar * rz[i] / (pq[i] * pq[i]);
    R[R_RZBAR] = rzbar_i;
  }
}

// B5: pbar = ((beta_i pbar) + alpha_i xbar) + pqbar q_i;
//     qbar = (-alpha_i rbar) + pqbar p_i
template <typename T>
__global__ void __launch_bounds__(kT)
k_rev_pq(T* __restrict__ pbar, T* __restrict__ qbar, const T* __restrict__ rbar, const T* __restrict__ xbar,
         const T* __restrict__ q, const T* __restrict__ p, int64_t n, const double* __restrict__ R,
         const double* __restrict__ alpha, const double* __restrict__ beta, int i) {
  const T bv = (T)beta[i], a = (T)alpha[i], ma = (T)(-alpha[i]), pqb = (T)R[R_PQBAR];
  for (int64_t e = (int64_t)blockIdx.x * kT + threadIdx.x; e < n; e += (int64_t)gridDim.x * kT) {
    T pb = bv * pbar[e];
    const T ax = a * xbar[e];
    pb = pb + ax;
    const T pqq = pqb * q[e];
    pbar[e] = pb + pqq;
    const T qb = ma * rbar[e];
    const T pqp = pqb * p[e];
    qbar[e] = qb + pqp;
  }
}

// bbar = (rbar + (pbar + t)) + 2 rzbar_0 b
template <typename T>
__global__ void __launch_bounds__(kT)
k_rev_final(T* __restrict__ bbar, const T* __restrict__ rbar, const T* __restrict__ pbar, const T* __restrict__ t,
            const T* __restrict__ b, int64_t n, const double* __restrict__ R) {
  const T c = (T)(2.0 * R[R_RZBAR]);
  for (int64_t e = (int64_t)blockIdx.x * kT + threadIdx.x; e < n; e += (int64_t)gridDim.x * kT) {
    const T pb = pbar[e] + t[e];
    const T rb = rbar[e] + pb;
    const T cb = c * b[e];
    bbar[e] = rb + cb;
  }
}

constexpr int kMaxSeg = 64;   // steps per segment (the outer product takes their vectors by value)
struct SegPtrs {
  const void* q[kMaxSeg];
  const void* p[kMaxSeg];
};

template <typename T>
__global__ void k_fill0(T* __restrict__ v, int64_t n) {
  for (int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; e < n; e += (int64_t)gridDim.x * blockDim.x)
    v[e] = T(0);
}

template <typename T>
void fill0(lfr_ctx* ctx, T* v, int64_t n) {
  hipLaunchKernelGGL(k_fill0<T>, dim3(lfr_blocks(n, kT, 2048)), dim3(kT), 0, ctx->stream, v, n);
}

// Mbar(ik, a, c) += qbar_t(map(a)+k) * p_t(map(c)+k) for t = nt-1 .. 0 (the
// reverse sweep's order), thread = cell-layer.  NDF > 0: the ndf^2 sums are
// held in registers over all t (Mbar read and written once per segment, each
// vector value loaded once); NDF == 0: entry by entry.
template <typename T, int NDF>
__global__ void __launch_bounds__(kT)
k_tape_outer(T* __restrict__ Mbar, SegPtrs sp, int nt, const int32_t* __restrict__ map0, int ncell, int nl,
             int ndf_rt) {
  const int64_t n3 = (int64_t)ncell * nl;
  const int64_t ik = (int64_t)blockIdx.x * kT + threadIdx.x;
  if (ik >= n3) return;
  const int c = (int)(ik / nl);
  const int k = (int)(ik - (int64_t)c * nl);
  if constexpr (NDF > 0) {
    const int32_t* m = map0 + (int64_t)c * NDF;
    int d[NDF];
#pragma unroll
    for (int a = 0; a < NDF; ++a) d[a] = m[a] + k;
    T acc[NDF][NDF];
#pragma unroll
    for (int col = 0; col < NDF; ++col)
#pragma unroll
      for (int row = 0; row < NDF; ++row) acc[row][col] = Mbar[ik + n3 * (row + NDF * col)];
    for (int tt = nt - 1; tt >= 0; --tt) {
      const T* q = (const T*)sp.q[tt];
      const T* p = (const T*)sp.p[tt];
      T qv[NDF], pv[NDF];
#pragma unroll
      for (int a = 0; a < NDF; ++a) {
        qv[a] = q[d[a]];
        pv[a] = p[d[a]];
      }
#pragma unroll
      for (int col = 0; col < NDF; ++col)
#pragma unroll
        for (int row = 0; row < NDF; ++row) {
          const T g = qv[row] * pv[col];
          acc[row][col] = acc[row][col] + g;
        }
    }
#pragma unroll
    for (int col = 0; col < NDF; ++col)
#pragma unroll
      for (int row = 0; row < NDF; ++row) Mbar[ik + n3 * (row + NDF * col)] = acc[row][col];
  } else {
    const int ndf = ndf_rt;
    const int32_t* m = map0 + (int64_t)c * ndf;
    for (int col = 0; col < ndf; ++col)
      for (int row = 0; row < ndf; ++row) {
        T* dst = Mbar + ik + n3 * ((int64_t)row + (int64_t)ndf * col);
        T acc = *dst;
        for (int tt = nt - 1; tt >= 0; --tt) {
          const T g = ((const T*)sp.q[tt])[m[row] + k] * ((const T*)sp.p[tt])[m[col] + k];
          acc = acc + g;
        }
        *dst = acc;
      }
  }
}

}  // namespace

struct lfr_cg_tape {
  lfr_ctx* ctx = nullptr;
  lfr_lma_op A{};
  const lfr_plan* plan_t = nullptr;
  const lfr_lma_packed* packed_t = nullptr;
  int prec = 8, K = 0, seg = 1, nseg = 1, record = 0;
  int64_t n = 0;
  void* mem = nullptr;
  size_t bytes = 0;
  size_t vstride = 0;
  char* vecs = nullptr;       // vector slots
  double* sc = nullptr;       // alpha[K], beta[K], rz[K+1], pq[K], R[R_NSLOT]
  double* part = nullptr;     // 3 x max(kRedMaxBlocks, apply partials)
  int part_stride = 0;
  // vector slot indices
  int slot_q0 = 0;            // Q[i], K slots (record) or 1
  int slot_cr0 = 0, slot_cp0 = 0;   // checkpoints per segment
  int slot_hr0 = 0, slot_hp0 = 0;   // segment history, seg+1 slots each
  int slot_tr = 0, slot_tp = 0;     // ping-pong r, p (2 each)
  int slot_rbar = 0, slot_pbar = 0, slot_t = 0, slot_qb0 = 0;  // reverse state, qbar per step (seg)
  int nvec = 0;

  char* v(int s) const { return vecs + (size_t)s * vstride; }
  double* alpha() const { return sc; }
  double* beta() const { return sc + K; }
  double* rz() const { return sc + 2 * K; }
  double* pq() const { return sc + 3 * K + 1; }
  double* R() const { return sc + 4 * K + 1; }
  int seg_of(int i) const { return i / seg < nseg ? i / seg : nseg - 1; }
  // where the forward keeps r_i / p_i (i = 0..K)
  // (record: checkpoints at segment starts and the whole last segment in
  // the history slots, ping-pong between; otherwise ping-pong throughout)
  char* rloc(int i) const {
    if (!record) return v(slot_tr + (i & 1));
    const int s = seg_of(i), j = i - s * seg;
    if (s == nseg - 1) return v(slot_hr0 + j);
    if (j == 0) return v(slot_cr0 + s);
    return v(slot_tr + (i & 1));
  }
  char* ploc(int i) const {
    if (!record) return v(slot_tp + (i & 1));
    const int s = seg_of(i), j = i - s * seg;
    if (s == nseg - 1) return v(slot_hp0 + j);
    if (j == 0) return v(slot_cp0 + s);
    return v(slot_tp + (i & 1));
  }
  char* qloc(int i) const { return record ? v(slot_q0 + i) : v(slot_q0); }
};

namespace {

template <typename T>
int tape_forward(lfr_cg_tape* tp, const T* b, T* x) {
  lfr_ctx* ctx = tp->ctx;
  hipStream_t s = ctx->stream;
  const int64_t n = tp->n;
  const int G = lfr::red_blocks(n);
  const unsigned gv = lfr_blocks(n, kT, 2048);
  double* pa = tp->part;
  fill0(ctx, x, n);
  hipLaunchKernelGGL(k_tape_init<T>, dim3(G), dim3(kRedThreads), 0, s, (T*)tp->rloc(0), (T*)tp->ploc(0), b, n, pa);
  hipLaunchKernelGGL(k_tape_sum<kRedThreads>, dim3(1), dim3(kRedThreads), 0, s, (const double*)pa, G, tp->rz());
  LFR_CHECK_LAUNCH(ctx);
  (void)gv;
  for (int i = 0; i < tp->K; ++i) {
    T* p = (T*)tp->ploc(i);
    T* q = (T*)tp->qloc(i);
    int np = 0;
    int rc = lfr_lma_apply_dot_queue(ctx, &tp->A, p, q, pa, &np);
    if (rc) return rc;
    hipLaunchKernelGGL(k_tape_alpha, dim3(1), dim3(kAlphaThreads), 0, s, (const double*)pa, np, tp->pq(),
                       tp->alpha(), (const double*)tp->rz(), i);
    hipLaunchKernelGGL(k_tape_update<T>, dim3(G), dim3(kRedThreads), 0, s, x, (const T*)tp->rloc(i),
                       (T*)tp->rloc(i + 1), (const T*)p, (const T*)q, n, (const double*)tp->alpha(), i, pa);
    if (i + 1 < tp->K || tp->record)
      hipLaunchKernelGGL(k_tape_direction<T>, dim3(G), dim3(kRedThreads), 0, s, (T*)tp->ploc(i + 1), (const T*)p,
                         (const T*)tp->rloc(i + 1), n, (const double*)pa, G, tp->rz(), tp->beta(), i);
    LFR_CHECK_LAUNCH(ctx);
  }
  return LFR_OK;
}

template <typename T, int NDF>
void launch_outer(lfr_cg_tape* tp, T* Mbar, const SegPtrs& ptrs, int nt) {
  const lfr_space* sp = tp->A.space;
  const int64_t n3 = (int64_t)sp->ncell * sp->nlayers;
  hipLaunchKernelGGL((k_tape_outer<T, NDF>), dim3((unsigned)((n3 + kT - 1) / kT)), dim3(kT), 0, tp->ctx->stream,
                     Mbar, ptrs, nt, sp->map0_d, sp->ncell, sp->nlayers, sp->ndf);
}


template <typename T>
int tape_reverse(lfr_cg_tape* tp, const T* b, const T* xbar, T* bbar, T* Mbar, int accumulate) {
  lfr_ctx* ctx = tp->ctx;
  hipStream_t s = ctx->stream;
  const int64_t n = tp->n;
  const int G = lfr::red_blocks(n);
  const unsigned gv = lfr_blocks(n, kT, 2048);
  double* pA = tp->part;
  double* pB = tp->part + tp->part_stride;
  double* pC = tp->part + 2 * tp->part_stride;
  T* rbar = (T*)tp->v(tp->slot_rbar);
  T* pbar = (T*)tp->v(tp->slot_pbar);
  T* t = (T*)tp->v(tp->slot_t);
  fill0(ctx, rbar, n);
  fill0(ctx, pbar, n);
  fill0(ctx, t, n);
  fill0(ctx, tp->R(), (int64_t)R_NSLOT);
  if (Mbar && !accumulate) {
    const lfr_space* sp = tp->A.space;
    fill0(ctx, Mbar, (int64_t)sp->ncell * sp->nlayers * sp->ndf * sp->ndf);
  }
  bool first = true;
  for (int sg = tp->nseg - 1; sg >= 0; --sg) {
    const int i0 = sg * tp->seg;
    const int i1 = sg == tp->nseg - 1 ? tp->K : i0 + tp->seg;
    const int len = i1 - i0;
    // segment history: r_{i0+j} in hr(j), p_{i0+j} in hp(j), j = 0..len
    auto hr = [&](int j) -> T* {
      if (sg == tp->nseg - 1) return (T*)tp->v(tp->slot_hr0 + j);
      return j == 0 ? (T*)tp->v(tp->slot_cr0 + sg) : (T*)tp->v(tp->slot_hr0 + j);
    };
    auto hp = [&](int j) -> T* {
      if (sg == tp->nseg - 1) return (T*)tp->v(tp->slot_hp0 + j);
      return j == 0 ? (T*)tp->v(tp->slot_cp0 + sg) : (T*)tp->v(tp->slot_hp0 + j);
    };
    if (sg != tp->nseg - 1) {
      for (int j = 0; j < len; ++j)
        hipLaunchKernelGGL(k_tape_recompute<T>, dim3(gv), dim3(kT), 0, s, (const T*)hr(j), hr(j + 1),
                           (const T*)hp(j), hp(j + 1), (const T*)tp->qloc(i0 + j), n, (const double*)tp->alpha(),
                           (const double*)tp->beta(), i0 + j);
      LFR_CHECK_LAUNCH(ctx);
    }
    for (int j = len - 1; j >= 0; --j) {
      const int i = i0 + j;
      const T* p_i = hp(j);
      const T* q_i = (const T*)tp->qloc(i);
      const T* r_next = hr(j + 1);
      T* qbar = (T*)tp->v(tp->slot_qb0 + j);
      hipLaunchKernelGGL(k_rev_dots<T>, dim3(G), dim3(kRedThreads), 0, s, pbar, (const T*)t, xbar, p_i, n,
                         first ? 0 : 1, pA, pB);
      hipLaunchKernelGGL(k_rev_s1, dim3(1), dim3(kRedThreads), 0, s, (const double*)pA, (const double*)pB, G, tp->R(),
                         (const double*)tp->rz(), i);
      hipLaunchKernelGGL(k_rev_rbar<T>, dim3(G), dim3(kRedThreads), 0, s, rbar, (const T*)pbar, r_next, q_i, n,
                         (const double*)tp->R(), pC);
      hipLaunchKernelGGL(k_rev_s2, dim3(1), dim3(kRedThreads), 0, s, (const double*)pC, G, tp->R(),
                         (const double*)tp->rz(), (const double*)tp->pq(), i);
      hipLaunchKernelGGL(k_rev_pq<T>, dim3(gv), dim3(kT), 0, s, pbar, qbar, (const T*)rbar, xbar, q_i, p_i, n,
                         (const double*)tp->R(), (const double*)tp->alpha(), (const double*)tp->beta(), i);
      LFR_CHECK_LAUNCH(ctx);
      // t = A^T qbar
      int rc;
      if (tp->packed_t)
        rc = lfr_matrix_vector_packed(ctx, tp->packed_t, t, qbar, 1);
      else
        rc = lfr_transpose_matrix_vector(ctx, tp->prec, t, qbar, tp->A.matrix_d, tp->A.space, tp->A.space,
                                         tp->plan_t, 1);
      if (rc) return rc;
      first = false;
    }
    if (Mbar) {
      SegPtrs ptrs{};
      for (int j = 0; j < len; ++j) {
        ptrs.q[j] = tp->v(tp->slot_qb0 + j);
        ptrs.p[j] = hp(j);
      }
      switch (tp->A.space->ndf) {
        case 6: launch_outer<T, 6>(tp, Mbar, ptrs, len); break;
        case 8: launch_outer<T, 8>(tp, Mbar, ptrs, len); break;
        default: launch_outer<T, 0>(tp, Mbar, ptrs, len);
      }
      LFR_CHECK_LAUNCH(ctx);
    }
  }
  hipLaunchKernelGGL(k_rev_final<T>, dim3(gv), dim3(kT), 0, s, bbar, (const T*)rbar, (const T*)pbar, (const T*)t, b,
                     n, (const double*)tp->R());
  LFR_CHECK_LAUNCH(ctx);
  return LFR_OK;
}

}  // namespace

extern "C" {

int lfr_cg_tape_create(lfr_ctx* ctx, const lfr_lma_op* A, const lfr_plan* plan_t, const lfr_lma_packed* packed_t,
                       int iters, int segment, int record, lfr_cg_tape** out) {
  if (!ctx || !out) return LFR_ERR_ARG;
  *out = nullptr;
  int rc = lfr_check_op(ctx, A, "lfr_cg_tape_create");
  if (rc) return rc;
  LFR_REQUIRE(ctx, !A->apply_bc, "lfr_cg_tape_create: apply_bc is not supported");
  LFR_REQUIRE(ctx, iters >= 1 && iters <= 100000, "lfr_cg_tape_create: iters %d", iters);
  if (record) {
    LFR_REQUIRE(ctx, (plan_t && plan_t->kind == 1 && plan_t->space == A->space) ||
                         (packed_t && packed_t->kind == 1 && packed_t->plan->space == A->space),
                "lfr_cg_tape_create: the reverse sweep needs a kind-1 plan (or kind-1 pack) of the operator's space");
    LFR_REQUIRE(ctx, !packed_t || (packed_t->ctx == ctx && packed_t->prec == A->prec),
                "lfr_cg_tape_create: packed transpose of another context or precision");
  }
  auto* tp = new lfr_cg_tape();
  tp->ctx = ctx;
  tp->A = *A;
  tp->plan_t = plan_t;
  tp->packed_t = packed_t;
  tp->prec = A->prec;
  tp->K = iters;
  tp->n = A->space->undf;
  tp->record = record ? 1 : 0;
  int sgl = segment > 0 ? segment : (int)std::ceil(std::sqrt((double)iters));
  if (sgl > iters) sgl = iters;
  if (sgl > kMaxSeg) {
    delete tp;
    return lfr_set_error(ctx, LFR_ERR_ARG, "lfr_cg_tape_create: segment %d above %d", sgl, kMaxSeg);
  }
  tp->seg = sgl;
  tp->nseg = (iters + sgl - 1) / sgl;
  // vector slots
  int s = 0;
  tp->slot_tr = s;
  s += 2;
  tp->slot_tp = s;
  s += 2;
  tp->slot_q0 = s;
  s += record ? iters : 1;
  if (record) {
    tp->slot_cr0 = s;
    s += tp->nseg;
    tp->slot_cp0 = s;
    s += tp->nseg;
    // the last segment may hold up to seg + (iters - nseg*seg) steps: size for the longest
    const int last = iters - (tp->nseg - 1) * sgl;
    const int hl = (last > sgl ? last : sgl) + 1;
    tp->slot_hr0 = s;
    s += hl;
    tp->slot_hp0 = s;
    s += hl;
    tp->slot_rbar = s++;
    tp->slot_pbar = s++;
    tp->slot_t = s++;
    tp->slot_qb0 = s;
    s += hl;
  }
  tp->nvec = s;
  tp->vstride = align_up((size_t)tp->n * tp->prec);
  const int napply = lfr_lma_apply_partials_max(A->plan);
  tp->part_stride = napply > kRedMaxBlocks ? napply : kRedMaxBlocks;
  const size_t sb = align_up((4 * (size_t)iters + 1 + R_NSLOT) * sizeof(double));
  const size_t pb = align_up((size_t)3 * tp->part_stride * sizeof(double));
  tp->bytes = sb + pb + (size_t)tp->nvec * tp->vstride;
  if (hipMalloc(&tp->mem, tp->bytes) != hipSuccess) {
    (void)hipGetLastError();
    const size_t want = tp->bytes;
    delete tp;
    return lfr_set_error(ctx, LFR_ERR_NOMEM, "lfr_cg_tape_create: cannot allocate %zu bytes", want);
  }
  tp->sc = (double*)tp->mem;
  tp->part = (double*)((char*)tp->mem + sb);
  tp->vecs = (char*)tp->mem + sb + pb;
  *out = tp;
  return LFR_OK;
}

int lfr_cg_tape_destroy(lfr_cg_tape* tp) {
  if (!tp) return LFR_OK;
  if (tp->mem) {
    (void)hipStreamSynchronize(tp->ctx->stream);
    (void)hipFree(tp->mem);
  }
  delete tp;
  return LFR_OK;
}

int lfr_cg_tape_info(const lfr_cg_tape* tp, int64_t* bytes, int* segment) {
  if (!tp) return LFR_ERR_ARG;
  if (bytes) *bytes = (int64_t)tp->bytes;
  if (segment) *segment = tp->seg;
  return LFR_OK;
}

int lfr_cg_tape_forward(lfr_ctx* ctx, lfr_cg_tape* tp, const void* b, void* x) {
  if (!ctx) return LFR_ERR_ARG;
  LFR_REQUIRE(ctx, tp && tp->ctx == ctx && b && x && b != x, "lfr_cg_tape_forward: bad arguments");
  if (tp->prec == 4) return tape_forward<float>(tp, (const float*)b, (float*)x);
  return tape_forward<double>(tp, (const double*)b, (double*)x);
}

int lfr_cg_tape_reverse(lfr_ctx* ctx, lfr_cg_tape* tp, const void* b, const void* xbar, void* bbar, void* Mbar,
                        int accumulate) {
  if (!ctx) return LFR_ERR_ARG;
  LFR_REQUIRE(ctx, tp && tp->ctx == ctx && b && xbar && bbar, "lfr_cg_tape_reverse: bad arguments");
  LFR_REQUIRE(ctx, tp->record, "lfr_cg_tape_reverse: the tape was created without recording");
  if (tp->prec == 4)
    return tape_reverse<float>(tp, (const float*)b, (const float*)xbar, (float*)bbar, (float*)Mbar, accumulate);
  return tape_reverse<double>(tp, (const double*)b, (const double*)xbar, (double*)bbar, (double*)Mbar, accumulate);
}

}  // extern "C"
