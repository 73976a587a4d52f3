// Iterative solvers over abstract operators, preconditioners and n-field
// vectors: the solver types of
// components/science/source/solver/sci_iterative_solver_mod.f90 (CG :495-592,
// BiCGstab :645-779, GMRES :830-1003, FGMRES :1055-1237, GCR :1288-1408,
// precondition_only :1664-1694, Jacobi :1745-1815, Chebyshev :1868-1940) over
// abstract_linear_operator_type / abstract_preconditioner_type objects
// (sci_linear_operator_mod.f90:18-35, sci_preconditioner_mod.f90:19-37) and
// field vectors of several fields (infrastructure/source/field/vector_mod.F90).
//
// Each method keeps the reference's statement sequence.  Vectors live in one
// device buffer per vector (fields end to end); every scalar the reference
// holds in an r_def variable lives in a device slot, is formed there by a
// one-thread kernel in the reference's expression order, and is read by the
// vector kernels, which cast it to the vector kind before the update (the
// r_solver_field_vector_type convention).  Inner products are formed per field
// (FP64 accumulation, a fixed reduction grid per field length) and added in
// field order, as field_vector_type%dot does.  The host waits only where the
// reference takes a decision: the convergence checks when monitoring, the
// r_nrm0 < EPS early exit, and the Givens rotations of the (F)GMRES
// Hessenberg once per restart.
#include <cmath>
#include <cstring>
#include <vector>

#include "lfr_internal.h"
#include "lfr_reduce.h"

struct lfr_linop {
  lfr_ctx* ctx = nullptr;
  int kind = 0;   // 0 LMA operator, 1 block operator, 2 caller callback
  lfr_vector_layout layout{};
  lfr_lma_op lma{};
  std::vector<lfr_block_entry> blocks;
  std::vector<int> single;   // per output field: index of its only entry, or -1
  lfr_apply_fn fn = nullptr;
  void* user = nullptr;
};

struct lfr_precon {
  lfr_ctx* ctx = nullptr;
  int kind = 0;   // 0 null, 1 diagonal, 2 column solve, 3 caller callback
  lfr_vector_layout layout{};
  const void* diag = nullptr;
  const void* cm = nullptr;
  int nrow = 0;
  const lfr_space* space = nullptr;
  const int32_t* ind = nullptr;
  int fp64 = 0;
  lfr_apply_fn fn = nullptr;
  void* user = nullptr;
};

namespace {

using lfr::kRedMaxBlocks;
using lfr::kRedThreads;
constexpr int kT = 256;
constexpr double kEps = 2.220446049250313e-16;  // epsilon(1.0_r_def), constants_mod.F90:167

int64_t layout_len(const lfr_vector_layout& L) {
  int64_t n = 0;
  for (int f = 0; f < L.nfield; ++f) n += L.undf[f];
  return n;
}
int64_t layout_off(const lfr_vector_layout& L, int f) {
  int64_t o = 0;
  for (int q = 0; q < f; ++q) o += L.undf[q];
  return o;
}
bool layout_ok(const lfr_vector_layout* L) {
  if (!L || (L->prec != 4 && L->prec != 8) || L->nfield < 1 || L->nfield > LFR_MAX_FIELDS) return false;
  for (int f = 0; f < L->nfield; ++f)
    if (L->undf[f] <= 0) return false;
  return true;
}
bool layout_eq(const lfr_vector_layout& a, const lfr_vector_layout& b) {
  if (a.prec != b.prec || a.nfield != b.nfield) return false;
  for (int f = 0; f < a.nfield; ++f)
    if (a.undf[f] != b.undf[f]) return false;
  return true;
}
size_t align_up(size_t b) { return (b + 255) & ~(size_t)255; }
unsigned vgrid(int64_t n) { return lfr_blocks(n, kT, 2048); }

// A coefficient read by a vector kernel: mult (slot < 0), mult * S[slot], or
// mult / S[slot] (inv), formed in FP64 and then cast to the vector kind.
struct Coef {
  int slot;
  int inv;
  double mult;
};
Coef cst(double c) { return Coef{-1, 0, c}; }
Coef sl(int slot, double mult = 1.0) { return Coef{slot, 0, mult}; }
Coef inv(int slot) { return Coef{slot, 1, 1.0}; }

__device__ __forceinline__ double coef(const double* __restrict__ S, Coef c) {
  if (c.slot < 0) return c.mult;
  return c.inv ? c.mult / S[c.slot] : c.mult * S[c.slot];
}

// Elementwise vector operations (each the arithmetic of one PSyclone
// built-in, or of a copy followed by one, with the same roundings).
enum VOp {
  V_COPY = 0,    // z = x                                   (setval_X)
  V_SET,         // z = a                                   (setval_c)
  V_AXPY,        // z = z + a x                             (inc_X_plus_bY)
  V_AYPX,        // z = a z + x                             (inc_aX_plus_Y)
  V_SCALE,       // z = a z                                 (inc_a_times_X)
  V_COPY_SCALE,  // z = a x                                 (copy, then scale)
  V_COPY_AXPY,   // z = x + a y                             (copy, then inc_X_plus_bY)
  V_AXPBY,       // z = a x + b y                           (aX_plus_bY)
  V_DIVIDE,      // z = x / y                               (X_divideby_Y)
  V_AXPY2        // z = (z + a x) + b y                     (two inc_X_plus_bY)
};

template <typename T, int OP>
__global__ void __launch_bounds__(kT)
k_vop(T* __restrict__ z, const T* __restrict__ x, const T* __restrict__ y, int64_t n,
      const double* __restrict__ S, Coef ca, Coef cb) {
  const T a = (T)coef(S, ca);
  const T b = (T)coef(S, cb);
  for (int64_t i = (int64_t)blockIdx.x * kT + threadIdx.x; i < n; i += (int64_t)gridDim.x * kT) {
    if (OP == V_COPY) {
      z[i] = x[i];
    } else if (OP == V_SET) {
      z[i] = a;
    } else if (OP == V_AXPY) {
      const T p = a * x[i];
      z[i] = z[i] + p;
    } else if (OP == V_AYPX) {
      const T p = a * z[i];
      z[i] = p + x[i];
    } else if (OP == V_SCALE) {
      z[i] = a * z[i];
    } else if (OP == V_COPY_SCALE) {
      z[i] = a * x[i];
    } else if (OP == V_COPY_AXPY) {
      const T p = a * y[i];
      z[i] = x[i] + p;
    } else if (OP == V_AXPBY) {
      const T p = a * x[i];
      const T q = b * y[i];
      z[i] = p + q;
    } else if (OP == V_DIVIDE) {
      z[i] = x[i] / y[i];
    } else if (OP == V_AXPY2) {
      const T p = a * x[i];
      const T s = z[i] + p;
      const T q = b * y[i];
      z[i] = s + q;
    }
  }
}

// First stage of one field's inner product, on the grid and in the order of
// the single-field dot (red_blocks(n) blocks, interleaved): MODE 0 <x, y>;
// MODE 1: z = z + a w first (inc_X_plus_bY), then <y, z> (y == nullptr:
// <z, z>) over the updated z, so z is read once.
template <typename T, int MODE>
__global__ void __launch_bounds__(kRedThreads)
k_fdot(T* __restrict__ z, const T* __restrict__ x, const T* __restrict__ y, int64_t n,
       const double* __restrict__ S, Coef ca, double* __restrict__ part) {
  __shared__ double lds[kRedThreads / 64];
  const T a = MODE == 1 ? (T)coef(S, ca) : T(0);
  double s = 0.0;
  for (int64_t i = (int64_t)blockIdx.x * kRedThreads + threadIdx.x; i < n; i += (int64_t)gridDim.x * kRedThreads) {
    if (MODE == 0) {
      const double p = (double)x[i] * (double)y[i];
      s = s + p;
    } else {
      const T p = a * x[i];
      const T zn = z[i] + p;
      z[i] = zn;
      const double q = y ? (double)y[i] * (double)zn : (double)zn * (double)zn;
      s = s + q;
    }
  }
  s = lfr::block_sum(s, lds);
  if (threadIdx.x == 0) part[blockIdx.x] = s;
}

struct FieldNb {
  int nb[LFR_MAX_FIELDS];
  int nfield;
};

// S[slot] = sum over fields f, in order, of that field's partials summed as
// the single-field dot's final stage does; optionally its square root.
__global__ void __launch_bounds__(kRedThreads)
k_fields_final(const double* __restrict__ part, FieldNb fn, double* __restrict__ S, int slot, int take_sqrt) {
  __shared__ double lds[kRedThreads / 64];
  double tot = 0.0;
  for (int f = 0; f < fn.nfield; ++f) {
    double s = lfr::strided_sum(part + (int64_t)f * kRedMaxBlocks, fn.nb[f]);
    s = lfr::block_sum(s, lds);
    tot = tot + s;   // meaningful in thread 0
  }
  if (threadIdx.x == 0) S[slot] = take_sqrt ? sqrt(tot) : tot;
}

// One-thread scalar programs: the reference's r_def expressions, in order.
enum SOpKind { SO_DIV = 0, SO_MUL, SO_NEG, SO_MOV, SO_CONST, SO_SQRT, SO_RECIP, SO_MAXC, SO_SUB, SO_ADD };
struct SOp {
  int op, a, b, o;
  double c;
};
struct SProg {
  SOp ops[8];
  int n;
};

__global__ void k_sprog(double* __restrict__ S, SProg p) {
  if (threadIdx.x != 0 || blockIdx.x != 0) return;
  for (int q = 0; q < p.n; ++q) {
    const SOp& o = p.ops[q];
    double v = 0.0;
    switch (o.op) {
      case SO_DIV: v = S[o.a] / S[o.b]; break;
      case SO_MUL: v = S[o.a] * S[o.b]; break;
      case SO_NEG: v = -S[o.a]; break;
      case SO_MOV: v = S[o.a]; break;
      case SO_CONST: v = o.c; break;
      case SO_SQRT: v = sqrt(S[o.a]); break;
      case SO_RECIP: v = 1.0 / S[o.a]; break;
      case SO_MAXC: v = S[o.a] > o.c ? S[o.a] : o.c; break;
      case SO_SUB: v = S[o.a] - S[o.b]; break;
      case SO_ADD: v = S[o.a] + S[o.b]; break;
    }
    S[o.o] = v;
  }
}

// mm_diagonal on a continuous space, thread = (output column, level): the
// entries of each cell in colour order, within a cell by layer then local
// dof, as the reference kernel's k-outer / df-inner loop meets this dof.
template <typename T>
__global__ void __launch_bounds__(kT)
k_mm_diag_plan(T* __restrict__ diag, const T* __restrict__ M, const int32_t* __restrict__ col_start,
               const int32_t* __restrict__ ent, int ncol, int H, int ngroup, int gs, int nl, int64_t n3, int ndf) {
  const int64_t t = (int64_t)blockIdx.x * kT + threadIdx.x;
  if (t >= (int64_t)ncol * H) return;
  const int col = (int)(t / H);
  const int L = (int)(t - (int64_t)col * H);
  const int64_t d = col_start[col] + L;
  T acc = diag[d];
  const int32_t* e = ent + (int64_t)col * ngroup * gs * 2;
  for (int g = 0; g < ngroup; ++g) {
    const int32_t* eg = e + (int64_t)g * gs * 2;
    const int cell = eg[0];
    int code[2] = {eg[1], gs > 1 ? eg[3] : -1};
    // order the (at most two) entries by layer k = L - shift, then local dof
    if (gs > 1 && code[0] >= 0 && code[1] >= 0) {
      const int s0 = code[0] & 1, s1 = code[1] & 1, i0 = code[0] >> 1, i1 = code[1] >> 1;
      if (s0 < s1 || (s0 == s1 && i0 > i1)) {
        const int tmp = code[0];
        code[0] = code[1];
        code[1] = tmp;
      }
    }
    for (int q = 0; q < gs && q < 2; ++q) {
      if (code[q] < 0) continue;
      const int k = L - (code[q] & 1);
      if (k < 0 || k >= nl) continue;
      const int i = code[q] >> 1;
      const T v = M[(int64_t)cell * nl + k + n3 * ((int64_t)i + (int64_t)ndf * i)];
      acc = acc + v;
    }
  }
  diag[d] = acc;
}

// mm_diagonal on a discontinuous space: every dof is touched once.
template <typename T>
__global__ void __launch_bounds__(kT)
k_mm_diag_dg(T* __restrict__ diag, const T* __restrict__ M, const int32_t* __restrict__ map0, int ncell, int nl,
             int ndf, int64_t n3) {
  const int64_t t = (int64_t)blockIdx.x * kT + threadIdx.x;
  if (t >= n3 * ndf) return;
  const int df = (int)(t / n3);
  const int64_t ik = t - (int64_t)df * n3;
  const int cell = (int)(ik / nl);
  const int k = (int)(ik - (int64_t)cell * nl);
  const int64_t d = (int64_t)map0[(int64_t)cell * ndf + df] + k;
  const T v = M[ik + n3 * ((int64_t)df + (int64_t)ndf * df)];
  diag[d] = diag[d] + v;
}

// ---------------------------------------------------------------------------
// The engine: vectors, slots and the operations the solvers are written in
// ---------------------------------------------------------------------------
enum Slots {
  K_R0 = 0, K_RN, K_RZ, K_RZN, K_PZ, K_ALPHA, K_BETA, K_OMEGA, K_RHO, K_RHO_OLD, K_TT, K_TS, K_R0V,
  K_SC, K_T1, K_T2, K_C1, K_C2, K_HN, K_GAM, K_NSLOT = 32
};

// Algorithmic bytes of one operator apply: per LMA matrix its entries and
// the input field read, the output written (and read when incremented);
// -1 for a caller-written operator.
double lma_pass_bytes(const lfr_space* to, const lfr_space* from, int s, bool inc) {
  const double m = (double)to->ncell * to->nlayers * to->ndf * from->ndf * s;
  return m + (double)from->undf * s + (double)to->undf * s * (inc ? 2 : 1);
}
double linop_bytes(const lfr_linop* A) {
  const int s = A->layout.prec;
  switch (A->kind) {
    case 0:
      return lma_pass_bytes(A->lma.space, A->lma.space, s, false);
    case 1: {
      double b = 0.0;
      const lfr_vector_layout& L = A->layout;
      for (int f = 0; f < L.nfield; ++f)
        if (A->single[f] < 0) b += (double)L.undf[f] * s;   // zeroed first
      for (size_t q = 0; q < A->blocks.size(); ++q) {
        const lfr_block_entry& e = A->blocks[q];
        b += lma_pass_bytes(e.to_space, e.from_space, s, A->single[e.row] != (int)q);
      }
      return b;
    }
    default:
      return -1.0;
  }
}
// ... and of one preconditioner apply: copy, divide, the column solve (three
// bands, right-hand side and solution); -1 for a caller-written one
double precon_bytes(const lfr_precon* P) {
  const int s = P->layout.prec;
  const double n = (double)layout_len(P->layout);
  switch (P->kind) {
    case 0: return 2 * n * s;
    case 1: return 3 * n * s;
    case 2: return (3.0 * P->nrow * P->space->ncell + 2 * n) * s;
    default: return -1.0;
  }
}

template <typename T>
struct Eng {
  lfr_ctx* ctx = nullptr;
  const lfr_linop* A = nullptr;
  const lfr_precon* P = nullptr;
  lfr_vector_layout L{};
  int64_t n = 0;
  double* S = nullptr;      // K_NSLOT scalar slots, then the Hessenberg
  double* H = nullptr;      // (m+1) x m, h(i,j) at H[i + (m+1) j]
  double* part = nullptr;   // kRedMaxBlocks per field
  char* vbase = nullptr;
  size_t vstride = 0;
  FieldNb nb{};
  int64_t off[LFR_MAX_FIELDS] = {};
  // algorithmic HBM bytes of the solve: every vector pass (each array read
  // or written once), operator and preconditioner applies (their matrix and
  // vector bytes); unknown when a caller-written operator or preconditioner
  // takes part
  double bytes = 0.0, apply_b = 0.0, precon_b = 0.0;
  bool bytes_known = true;

  int init(lfr_ctx* c, const lfr_linop* a, const lfr_precon* p, int nvec, int m) {
    ctx = c;
    A = a;
    P = p;
    L = a->layout;
    n = layout_len(L);
    apply_b = linop_bytes(a);
    precon_b = precon_bytes(p);
    bytes_known = apply_b >= 0.0 && precon_b >= 0.0;
    nb.nfield = L.nfield;
    for (int f = 0; f < L.nfield; ++f) {
      off[f] = layout_off(L, f);
      nb.nb[f] = lfr::red_blocks(L.undf[f]);
    }
    const size_t sb = align_up((K_NSLOT + (size_t)(m + 1) * (m > 0 ? m : 1) + 8) * sizeof(double));
    const size_t pb = align_up((size_t)L.nfield * kRedMaxBlocks * sizeof(double));
    vstride = align_up((size_t)n * sizeof(T));
    const size_t need = sb + pb + (size_t)nvec * vstride;
    if (need > ctx->kwork_bytes) {
      if (ctx->kwork) {
        LFR_HIP(ctx, hipStreamSynchronize(ctx->stream));
        LFR_HIP(ctx, hipFree(ctx->kwork));
        ctx->kwork = nullptr;
        ctx->kwork_bytes = 0;
      }
      if (hipMalloc(&ctx->kwork, need) != hipSuccess) {
        (void)hipGetLastError();
        return lfr_set_error(ctx, LFR_ERR_NOMEM, "iterative solver: cannot allocate %zu bytes of workspace", need);
      }
      ctx->kwork_bytes = need;
    }
    S = (double*)ctx->kwork;
    H = S + K_NSLOT;
    part = (double*)((char*)ctx->kwork + sb);
    vbase = (char*)ctx->kwork + sb + pb;
    return LFR_OK;
  }
  T* vec(int i) const { return (T*)(vbase + (size_t)i * vstride); }
  hipStream_t st() const { return ctx->stream; }

  template <int OP>
  int vop(T* z, const T* x, const T* y, Coef a, Coef b = cst(0.0)) {
    constexpr int reads = OP == V_SET ? 0
                          : (OP == V_COPY || OP == V_SCALE || OP == V_COPY_SCALE) ? 1
                          : OP == V_AXPY2 ? 3
                                          : 2;
    bytes += (double)n * sizeof(T) * (reads + 1);
    hipLaunchKernelGGL((k_vop<T, OP>), dim3(vgrid(n)), dim3(kT), 0, st(), z, x, y, n, (const double*)S, a, b);
    LFR_CHECK_LAUNCH(ctx);
    return LFR_OK;
  }
  // S[slot] = <x, y> (or its square root)
  int dot(const T* x, const T* y, int slot, bool take_sqrt = false) {
    bytes += (double)n * sizeof(T) * (x == y ? 1 : 2);
    for (int f = 0; f < L.nfield; ++f)
      hipLaunchKernelGGL((k_fdot<T, 0>), dim3(nb.nb[f]), dim3(kRedThreads), 0, st(), (T*)nullptr, x + off[f],
                         y + off[f], (int64_t)L.undf[f], (const double*)S, cst(0.0), part + (int64_t)f * kRedMaxBlocks);
    hipLaunchKernelGGL(k_fields_final, dim3(1), dim3(kRedThreads), 0, st(), (const double*)part, nb, S, slot,
                       (int)take_sqrt);
    LFR_CHECK_LAUNCH(ctx);
    return LFR_OK;
  }
  int norm(const T* x, int slot) { return dot(x, x, slot, true); }
  // z = z + a w, then S[slot] = <y, z> (y == nullptr: <z, z>, sqrt if asked)
  int axpy_dot(T* z, const T* w, Coef a, const T* y, int slot, bool take_sqrt = false) {
    bytes += (double)n * sizeof(T) * (3 + (y && y != z && y != w ? 1 : 0));
    for (int f = 0; f < L.nfield; ++f)
      hipLaunchKernelGGL((k_fdot<T, 1>), dim3(nb.nb[f]), dim3(kRedThreads), 0, st(), z + off[f], w + off[f],
                         y ? y + off[f] : (const T*)nullptr, (int64_t)L.undf[f], (const double*)S, a,
                         part + (int64_t)f * kRedMaxBlocks);
    hipLaunchKernelGGL(k_fields_final, dim3(1), dim3(kRedThreads), 0, st(), (const double*)part, nb, S, slot,
                       (int)take_sqrt);
    LFR_CHECK_LAUNCH(ctx);
    return LFR_OK;
  }
  int sprog(std::initializer_list<SOp> ops) {
    SProg p{};
    p.n = 0;
    for (const SOp& o : ops) p.ops[p.n++] = o;
    hipLaunchKernelGGL(k_sprog, dim3(1), dim3(64), 0, st(), S, p);
    LFR_CHECK_LAUNCH(ctx);
    return LFR_OK;
  }
  int read(int slot, double* out) {
    LFR_HIP(ctx, hipMemcpyAsync(ctx->host_scalars, S + slot, sizeof(double), hipMemcpyDeviceToHost, st()));
    LFR_HIP(ctx, hipStreamSynchronize(st()));
    *out = ctx->host_scalars[0];
    return LFR_OK;
  }
  int apply(const T* x, T* y);
  int precon(const T* x, T* y);
};

#define KTRY(expr)              \
  do {                          \
    int _rc = (expr);           \
    if (_rc != LFR_OK) return _rc; \
  } while (0)

int linop_apply(lfr_ctx* ctx, const lfr_linop* A, const void* x, void* y) {
  const int pr = A->layout.prec;
  switch (A->kind) {
    case 0:
      return lfr_lma_apply_queue(ctx, &A->lma, x, y);
    case 1: {
      const lfr_vector_layout& L = A->layout;
      // fields written by several entries (or none) start from zero
      for (int f = 0; f < L.nfield; ++f)
        if (A->single[f] < 0)
          KTRY(lfr_vec_setval(ctx, pr, L.undf[f], (char*)y + layout_off(L, f) * pr, 0.0));
      for (size_t q = 0; q < A->blocks.size(); ++q) {
        const lfr_block_entry& b = A->blocks[q];
        const void* xf = (const char*)x + layout_off(L, b.col) * pr;
        void* yf = (char*)y + layout_off(L, b.row) * pr;
        const int ow = A->single[b.row] == (int)q ? 1 : 0;
        int rc;
        if (b.transpose)
          rc = lfr_transpose_matrix_vector(ctx, pr, yf, xf, b.matrix_d, b.to_space, b.from_space, b.plan, ow);
        else if (b.plan)
          rc = lfr_matrix_vector(ctx, pr, yf, xf, b.matrix_d, b.to_space, b.from_space, b.plan, ow);
        else if (ow)
          rc = lfr_dg_matrix_vector(ctx, pr, yf, xf, b.matrix_d, b.to_space, b.from_space);
        else
          rc = lfr_dg_inc_matrix_vector(ctx, pr, yf, xf, b.matrix_d, b.to_space, b.from_space);
        if (rc) return rc;
      }
      return LFR_OK;
    }
    default: {
      const int rc = A->fn(A->user, x, y, (void*)ctx->stream);
      if (rc) return lfr_set_error(ctx, LFR_ERR_ARG, "operator callback returned %d", rc);
      return LFR_OK;
    }
  }
}

int precon_apply(lfr_ctx* ctx, const lfr_precon* P, const void* x, void* y) {
  const lfr_vector_layout& L = P->layout;
  const int64_t n = layout_len(L);
  switch (P->kind) {
    case 0:
      return lfr_vec_copy(ctx, L.prec, n, y, x);
    case 1:
      return lfr_vec_divide(ctx, L.prec, n, y, x, P->diag);
    case 2:
      return lfr_columnwise_op_appinv(ctx, L.prec, y, x, P->cm, 3, P->nrow, P->space->ncell, P->space, P->ind,
                                      P->fp64);
    default: {
      const int rc = P->fn(P->user, x, y, (void*)ctx->stream);
      if (rc) return lfr_set_error(ctx, LFR_ERR_ARG, "preconditioner callback returned %d", rc);
      return LFR_OK;
    }
  }
}

template <typename T>
int Eng<T>::apply(const T* x, T* y) {
  bytes += apply_b;
  return linop_apply(ctx, A, x, y);
}
template <typename T>
int Eng<T>::precon(const T* x, T* y) {
  bytes += precon_b;
  return precon_apply(ctx, P, x, y);
}

SOp op(int kind, int o, int a = 0, int b = 0, double c = 0.0) { return SOp{kind, a, b, o, c}; }

int not_converged(lfr_ctx* ctx, const lfr_iterative_solver* s, const char* who, int iters, double err) {
  if (err != err) return lfr_set_error(ctx, LFR_ERR_NAN, "%s: NaN residual after %d iterations", who, iters);
  if (s->monitor_convergence && s->fail_on_non_converged)
    return lfr_set_error(ctx, LFR_ERR_NOT_CONVERGED, "%s: NOT converged in %d iters, Res=%g", who, iters, err);
  return LFR_OK;
}

// ---- conjugate_gradient_type%cg_solve (:495-592) --------------------------
template <typename T>
int cg(Eng<T>& e, const lfr_iterative_solver* s, T* x, const T* b, lfr_solver_stats* st) {
  T* r = e.vec(0);
  T* p = e.vec(1);
  T* z = e.vec(2);
  KTRY(e.apply(x, r));                                        // r = A x
  KTRY(e.template vop<V_AYPX>(r, b, nullptr, cst(-1.0)));     // r = -r ; r = r + b
  KTRY(e.norm(r, K_R0));
  double r0 = 0;
  KTRY(e.read(K_R0, &r0));
  st->r_nrm0 = st->r_nrm = r0;
  if (r0 < kEps) {
    st->converged = 1;
    return LFR_OK;
  }
  KTRY(e.precon(r, z));                                       // z = P^{-1} r
  KTRY(e.dot(r, z, K_RZ));
  KTRY(e.template vop<V_COPY>(p, z, nullptr, cst(0.0)));
  bool conv = false;
  double rn = r0;
  int it;
  for (it = 1; it <= s->max_iter; ++it) {
    KTRY(e.apply(p, z));                                      // z = A p
    KTRY(e.dot(p, z, K_PZ));
    KTRY(e.sprog({op(SO_DIV, K_ALPHA, K_RZ, K_PZ)}));         // alpha = rz / <p, z>
    KTRY(e.template vop<V_AXPY>(x, p, nullptr, sl(K_ALPHA)));
    KTRY(e.template vop<V_AXPY>(r, z, nullptr, sl(K_ALPHA, -1.0)));
    if (s->monitor_convergence) {
      KTRY(e.norm(r, K_RN));
      KTRY(e.read(K_RN, &rn));
      st->r_nrm = rn;
      st->iterations = it;
      if (rn != rn) return not_converged(e.ctx, s, "cg", it, rn);
      if (rn / r0 <= s->r_tol || rn <= s->a_tol) {
        conv = true;
        break;
      }
    }
    KTRY(e.precon(r, z));
    KTRY(e.dot(r, z, K_RZN));
    KTRY(e.sprog({op(SO_DIV, K_BETA, K_RZN, K_RZ), op(SO_MOV, K_RZ, K_RZN)}));
    KTRY(e.template vop<V_AYPX>(p, z, nullptr, sl(K_BETA)));  // p = beta p + z
  }
  st->iterations = conv ? it : s->max_iter;
  st->converged = conv;
  if (s->monitor_convergence && !conv) return not_converged(e.ctx, s, "cg", s->max_iter, rn / r0);
  return LFR_OK;
}

// ---- BiCGstab_type%bicgstab_solve (:645-779) ------------------------------
template <typename T>
int bicgstab(Eng<T>& e, const lfr_iterative_solver* s, T* x, const T* b, lfr_solver_stats* st) {
  T *r = e.vec(0), *r0 = e.vec(1), *p = e.vec(2), *v = e.vec(3), *t = e.vec(4), *sv = e.vec(5), *y = e.vec(6),
    *z = e.vec(7);
  KTRY(e.apply(x, v));                                        // v = A x
  KTRY(e.template vop<V_COPY_AXPY>(r, b, v, cst(-1.0)));      // r = b ; r = r - v
  KTRY(e.template vop<V_COPY>(r0, r, nullptr, cst(0.0)));
  KTRY(e.norm(r, K_SC));
  double sc = 0;
  KTRY(e.read(K_SC, &sc));
  st->r_nrm0 = st->r_nrm = sc;
  if (sc < kEps) {
    st->converged = 1;
    return LFR_OK;
  }
  sc = sc > s->a_tol ? sc : s->a_tol;
  KTRY(e.sprog({op(SO_CONST, K_ALPHA, 0, 0, 1.0), op(SO_CONST, K_OMEGA, 0, 0, 1.0),
                op(SO_CONST, K_RHO_OLD, 0, 0, 1.0)}));
  KTRY(e.template vop<V_SET>(v, nullptr, nullptr, cst(0.0)));
  KTRY(e.template vop<V_SET>(p, nullptr, nullptr, cst(0.0)));
  bool conv = false;
  double err = 0.0;
  int it;
  for (it = 1; it <= s->max_iter; ++it) {
    KTRY(e.dot(r, r0, K_RHO));                                // rho = r.r0
    // beta = (rho/rho_old) * (alpha/omega); c1 = -beta*omega
    KTRY(e.sprog({op(SO_DIV, K_T1, K_RHO, K_RHO_OLD), op(SO_DIV, K_T2, K_ALPHA, K_OMEGA),
                  op(SO_MUL, K_BETA, K_T1, K_T2), op(SO_MUL, K_C1, K_BETA, K_OMEGA), op(SO_NEG, K_C1, K_C1)}));
    KTRY(e.template vop<V_COPY_AXPY>(t, r, v, sl(K_C1)));    // t = r - beta omega v
    KTRY(e.precon(t, y));                                     // y = P^{-1} t
    KTRY(e.template vop<V_AYPX>(p, y, nullptr, sl(K_BETA)));  // p = beta p + y
    KTRY(e.apply(p, v));                                      // v = A p
    KTRY(e.dot(r0, v, K_R0V));
    KTRY(e.sprog({op(SO_DIV, K_ALPHA, K_RHO, K_R0V)}));       // alpha = rho / r0.v
    KTRY(e.template vop<V_COPY_AXPY>(sv, r, v, sl(K_ALPHA, -1.0)));   // s = r - alpha v
    KTRY(e.precon(sv, z));                                    // z = P^{-1} s
    KTRY(e.apply(z, t));                                      // t = A z
    KTRY(e.dot(t, t, K_TT));
    KTRY(e.dot(t, sv, K_TS));
    KTRY(e.sprog({op(SO_DIV, K_OMEGA, K_TS, K_TT)}));         // omega = ts / tt
    KTRY(e.template vop<V_AXPY2>(x, z, p, sl(K_OMEGA), sl(K_ALPHA)));  // x += omega z ; x += alpha p
    KTRY(e.template vop<V_COPY_AXPY>(r, sv, t, sl(K_OMEGA, -1.0)));    // r = s - omega t
    KTRY(e.sprog({op(SO_MOV, K_RHO_OLD, K_RHO)}));
    if (s->monitor_convergence) {
      double rn = 0;
      KTRY(e.norm(r, K_RN));
      KTRY(e.read(K_RN, &rn));
      err = rn / sc;
      st->r_nrm = rn;
      st->iterations = it;
      if (err != err) return not_converged(e.ctx, s, "bicgstab", it, err);
      if (err < s->r_tol) {
        conv = true;
        break;
      }
    }
  }
  st->iterations = conv ? it : s->max_iter;
  st->converged = conv;
  if (s->monitor_convergence && !conv) return not_converged(e.ctx, s, "bicgstab", s->max_iter, err);
  return LFR_OK;
}

// Givens rotations of the (m+1) x m Hessenberg h and right-hand side g, then
// back substitution into u (:930-953 of gmres_solve; the same steps in
// fgmres_solve :1134-1160), in FP64 on the host.
void givens_solve(std::vector<double>& h, int m, std::vector<double>& g, std::vector<double>& u) {
  auto hh = [&](int i, int j) -> double& { return h[(size_t)i + (size_t)(m + 1) * j]; };
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
  u[m - 1] = g[m - 1] / hh(m - 1, m - 1);
  for (int iv = m - 2; iv >= 0; --iv) {
    u[iv] = g[iv];
    for (int j = iv + 1; j < m; ++j) u[iv] = u[iv] - hh(iv, j) * u[j];
    u[iv] = u[iv] / hh(iv, iv);
  }
}

// Arnoldi with modified Gram-Schmidt on w for column iv: h(j, iv) = <v_j, w>,
// w -= h(j, iv) v_j, j = 0..iv, then h(iv+1, iv) = ||w||; each axpy is fused
// with the next inner product (the separate kernels' bits).
template <typename T>
int mgs(Eng<T>& e, int m, int iv, T* w, T* const* V) {
  double* H = e.H;
  auto hslot = [&](int i, int j) { return (int)(K_NSLOT + i + (m + 1) * j); };
  KTRY(e.dot(V[0], w, hslot(0, iv)));
  for (int j = 0; j <= iv; ++j) {
    if (j < iv)
      KTRY(e.axpy_dot(w, V[j], sl(hslot(j, iv), -1.0), V[j + 1], hslot(j + 1, iv)));
    else
      KTRY(e.axpy_dot(w, V[j], sl(hslot(j, iv), -1.0), nullptr, hslot(iv + 1, iv), true));
  }
  (void)H;
  return LFR_OK;
}

// ---- gmres_type%gmres_solve (:830-1003) -----------------------------------
template <typename T>
int gmres(Eng<T>& e, const lfr_iterative_solver* s, T* x, const T* b, lfr_solver_stats* st) {
  const int m = s->gcrk;
  T *sv = e.vec(0), *w = e.vec(1), *res = e.vec(2);
  std::vector<T*> V(m);
  for (int i = 0; i < m; ++i) V[i] = e.vec(3 + i);
  auto hslot = [&](int i, int j) { return (int)(K_NSLOT + i + (m + 1) * j); };
  KTRY(e.apply(x, w));                                        // Ax = A x (w holds it)
  KTRY(e.template vop<V_COPY_AXPY>(res, b, w, cst(-1.0)));    // res = b - Ax
  KTRY(e.norm(res, K_SC));
  double sc = 0;
  KTRY(e.read(K_SC, &sc));
  st->r_nrm0 = st->r_nrm = sc;
  if (sc < kEps) {
    st->converged = 1;
    return LFR_OK;
  }
  if (s->monitor_convergence) sc = sc > s->a_tol ? sc : s->a_tol;
  std::vector<double> h((size_t)(m + 1) * m), g(m + 1), u(m);
  double err = 0.0;
  bool conv = false;
  int it;
  for (it = 1; it <= s->max_iter; ++it) {
    KTRY(e.precon(res, sv));                                  // s = P^{-1} res
    KTRY(e.norm(sv, K_GAM));                                  // beta = ||s||
    KTRY(e.template vop<V_COPY_SCALE>(V[0], sv, nullptr, inv(K_GAM)));
    for (int iv = 0; iv < m; ++iv) {
      // w = v_iv ; s = A w ; w = P^{-1} s
      KTRY(e.apply(V[iv], sv));
      KTRY(e.precon(sv, w));
      KTRY(mgs(e, m, iv, w, V.data()));
      if (iv < m - 1) KTRY(e.template vop<V_COPY_SCALE>(V[iv + 1], w, nullptr, inv(hslot(iv + 1, iv))));
    }
    LFR_HIP(e.ctx, hipMemcpyAsync(h.data(), e.H, h.size() * sizeof(double), hipMemcpyDeviceToHost, e.st()));
    LFR_HIP(e.ctx, hipMemcpyAsync(e.ctx->host_scalars, e.S + K_GAM, sizeof(double), hipMemcpyDeviceToHost, e.st()));
    LFR_HIP(e.ctx, hipStreamSynchronize(e.st()));
    std::fill(g.begin(), g.end(), 0.0);
    g[0] = e.ctx->host_scalars[0];
    givens_solve(h, m, g, u);
    for (int iv = 0; iv < m; ++iv)                            // s = v_iv ; x += u(iv) s
      KTRY(e.template vop<V_AXPY>(x, V[iv], nullptr, cst(u[iv])));
    KTRY(e.apply(x, w));                                      // Ax = A x
    KTRY(e.template vop<V_COPY_AXPY>(res, b, w, cst(-1.0)));  // res = -Ax + b
    st->iterations = it;
    if (s->monitor_convergence) {
      double rn = 0;
      KTRY(e.norm(res, K_RN));
      KTRY(e.read(K_RN, &rn));
      st->r_nrm = rn;
      err = rn / sc;
      if (err != err) return not_converged(e.ctx, s, "gmres", it, err);
      if (err < s->r_tol) {
        conv = true;
        break;
      }
    }
  }
  st->iterations = conv ? it : s->max_iter;
  st->converged = conv;
  if (s->monitor_convergence && !conv) return not_converged(e.ctx, s, "gmres", s->max_iter, err);
  return LFR_OK;
}

// ---- fgmres_type%fgmres_solve (:1055-1237) --------------------------------
// The initial residual is b (the reference assumes a zero initial guess for
// the correction dx); x += dx at the end.
template <typename T>
int fgmres(Eng<T>& e, const lfr_iterative_solver* s, T* x, const T* b, lfr_solver_stats* st) {
  const int m = s->gcrk;
  T *dx = e.vec(0), *ax = e.vec(1), *res = e.vec(2), *sv = e.vec(3), *w = e.vec(4);
  std::vector<T*> V(m), PV(m);
  for (int i = 0; i < m; ++i) {
    V[i] = e.vec(5 + i);
    PV[i] = e.vec(5 + m + i);
  }
  auto hslot = [&](int i, int j) { return (int)(K_NSLOT + i + (m + 1) * j); };
  KTRY(e.template vop<V_SET>(dx, nullptr, nullptr, cst(0.0)));
  KTRY(e.template vop<V_COPY>(res, b, nullptr, cst(0.0)));
  KTRY(e.norm(res, K_SC));
  double sc = 0;
  KTRY(e.read(K_SC, &sc));
  st->r_nrm0 = st->r_nrm = sc;
  if (sc < kEps) {
    st->converged = 1;
    return LFR_OK;
  }
  if (s->monitor_convergence) sc = sc > s->a_tol ? sc : s->a_tol;
  // s = res ; beta = ||s|| ; v1 = s / beta
  KTRY(e.norm(res, K_GAM));
  KTRY(e.template vop<V_COPY_SCALE>(V[0], res, nullptr, inv(K_GAM)));
  std::vector<double> h((size_t)(m + 1) * m), g(m + 1), u(m);
  double err = 0.0;
  bool conv = false;
  int it;
  for (it = 1; it <= s->max_iter; ++it) {
    for (int iv = 0; iv < m; ++iv) {
      KTRY(e.precon(V[iv], PV[iv]));                          // Pv_iv = P^{-1} v_iv
      KTRY(e.apply(PV[iv], w));                               // s = A Pv_iv ; w = s
      KTRY(mgs(e, m, iv, w, V.data()));
      if (iv < m - 1) KTRY(e.template vop<V_COPY_SCALE>(V[iv + 1], w, nullptr, inv(hslot(iv + 1, iv))));
    }
    LFR_HIP(e.ctx, hipMemcpyAsync(h.data(), e.H, h.size() * sizeof(double), hipMemcpyDeviceToHost, e.st()));
    LFR_HIP(e.ctx, hipMemcpyAsync(e.ctx->host_scalars, e.S + K_GAM, sizeof(double), hipMemcpyDeviceToHost, e.st()));
    LFR_HIP(e.ctx, hipStreamSynchronize(e.st()));
    std::fill(g.begin(), g.end(), 0.0);
    g[0] = e.ctx->host_scalars[0];
    givens_solve(h, m, g, u);
    for (int iv = 0; iv < m; ++iv) KTRY(e.template vop<V_AXPY>(dx, PV[iv], nullptr, cst(u[iv])));
    KTRY(e.apply(dx, ax));                                    // Ax = A dx
    KTRY(e.template vop<V_COPY_AXPY>(res, b, ax, cst(-1.0))); // res = -Ax + b
    KTRY(e.norm(res, K_GAM));                                 // beta = ||res||
    st->iterations = it;
    if (s->monitor_convergence) {
      double rn = 0;
      KTRY(e.read(K_GAM, &rn));
      st->r_nrm = rn;
      err = rn / sc;
      if (err != err) break;
      if (err < s->r_tol) {
        conv = true;
        break;
      }
    }
    // v1 = 0 ; v1 += (1/beta) s  with s = res
    KTRY(e.template vop<V_COPY_SCALE>(V[0], res, nullptr, inv(K_GAM)));
  }
  KTRY(e.template vop<V_AXPY>(x, dx, nullptr, cst(1.0)));    // x = x + dx
  st->iterations = conv ? it : (it > s->max_iter ? s->max_iter : it);
  st->converged = conv;
  if (err != err || (s->monitor_convergence && !conv))
    return not_converged(e.ctx, s, "fgmres", st->iterations, err);
  return LFR_OK;
}

// ---- gcr_type%gcr_solve (:1288-1408) --------------------------------------
template <typename T>
int gcr(Eng<T>& e, const lfr_iterative_solver* s, T* x, const T* b, lfr_solver_stats* st) {
  const int m = s->gcrk;
  T *dx = e.vec(0), *res = e.vec(1), *ax = e.vec(2);
  std::vector<T*> V(m), PV(m);
  for (int i = 0; i < m; ++i) {
    V[i] = e.vec(3 + i);
    PV[i] = e.vec(3 + m + i);
  }
  KTRY(e.precon(b, dx));                                      // dx = P^{-1} b
  KTRY(e.apply(dx, ax));                                      // Ax = A dx
  KTRY(e.template vop<V_COPY_AXPY>(res, b, ax, cst(-1.0)));   // res = b - Ax
  KTRY(e.norm(b, K_SC));
  double sc = 0;
  KTRY(e.read(K_SC, &sc));
  st->r_nrm0 = st->r_nrm = sc;
  if (sc < kEps) {
    st->converged = 1;
    return LFR_OK;
  }
  if (s->monitor_convergence) sc = sc > s->a_tol ? sc : s->a_tol;
  double err = 0.0;
  bool conv = false;
  int steps = 0, it;
  for (it = 1; it <= s->max_iter && !conv; ++it) {
    for (int iv = 0; iv < m; ++iv) {
      KTRY(e.precon(res, PV[iv]));                            // Pv_iv = P^{-1} res
      KTRY(e.apply(PV[iv], V[iv]));                           // v_iv = A Pv_iv
      for (int j = 0; j < iv; ++j) {
        KTRY(e.dot(V[iv], V[j], K_ALPHA));                    // alpha = v_iv . v_j
        KTRY(e.template vop<V_AXPY>(V[iv], V[j], nullptr, sl(K_ALPHA, -1.0)));
        KTRY(e.template vop<V_AXPY>(PV[iv], PV[j], nullptr, sl(K_ALPHA, -1.0)));
      }
      KTRY(e.norm(V[iv], K_ALPHA));                           // alpha = ||v_iv||
      KTRY(e.sprog({op(SO_RECIP, K_BETA, K_ALPHA)}));          // beta = 1 / alpha
      KTRY(e.template vop<V_SCALE>(V[iv], nullptr, nullptr, sl(K_BETA)));
      KTRY(e.template vop<V_SCALE>(PV[iv], nullptr, nullptr, sl(K_BETA)));
      KTRY(e.dot(res, V[iv], K_ALPHA));                       // alpha = res . v_iv
      KTRY(e.template vop<V_AXPY>(dx, PV[iv], nullptr, sl(K_ALPHA)));
      KTRY(e.template vop<V_AXPY>(res, V[iv], nullptr, sl(K_ALPHA, -1.0)));
      ++steps;
      if (s->monitor_convergence) {
        double rn = 0;
        KTRY(e.norm(res, K_RN));
        KTRY(e.read(K_RN, &rn));
        st->r_nrm = rn;
        err = rn / sc;
        if (err != err) break;
        if (err < s->r_tol) {
          conv = true;
          break;
        }
      }
    }
    if (err != err) break;
  }
  KTRY(e.template vop<V_AXPY>(x, dx, nullptr, cst(1.0)));    // x = x + dx
  st->iterations = steps;
  st->converged = conv;
  if (err != err || (s->monitor_convergence && !conv)) return not_converged(e.ctx, s, "gcr", steps, err);
  return LFR_OK;
}

// ---- precondition_only_type (:1664-1694) ----------------------------------
template <typename T>
int precondition_only(Eng<T>& e, const lfr_iterative_solver* s, T* x, const T* b, lfr_solver_stats* st) {
  KTRY(e.precon(b, x));                                       // x = P^{-1} b
  st->iterations = 1;
  st->converged = 1;
  if (s->monitor_convergence) {
    T *res = e.vec(0), *ax = e.vec(1);
    KTRY(e.norm(b, K_R0));
    KTRY(e.apply(x, ax));
    KTRY(e.template vop<V_COPY_AXPY>(res, b, ax, cst(-1.0)));
    KTRY(e.norm(res, K_RN));
    KTRY(e.read(K_R0, &st->r_nrm0));
    KTRY(e.read(K_RN, &st->r_nrm));
  }
  return LFR_OK;
}

// ---- jacobi_type%jacobi_solve (:1745-1815) --------------------------------
template <typename T>
int jacobi(Eng<T>& e, const lfr_iterative_solver* s, T* x, const T* b, lfr_solver_stats* st) {
  T *r = e.vec(0), *z = e.vec(1);
  const double cst_relax = -s->rho_relax;
  bool conv = false;
  double r0 = 0, rn = 0;
  int it;
  for (it = 1; it <= s->max_iter; ++it) {
    KTRY(e.apply(x, r));                                      // r = A x
    KTRY(e.template vop<V_AXPY>(r, b, nullptr, cst(-1.0)));   // r = r - b
    if (s->monitor_convergence) {
      KTRY(e.norm(r, K_RN));
      KTRY(e.read(K_RN, &rn));
      if (it == 1) r0 = rn;
      st->r_nrm0 = r0;
      st->r_nrm = rn;
      st->iterations = it;
      if (rn != rn) return not_converged(e.ctx, s, "jacobi", it, rn);
      if (rn / r0 <= s->r_tol || rn <= s->a_tol) {
        conv = true;
        break;
      }
    }
    KTRY(e.precon(r, z));                                     // z = P^{-1} r
    KTRY(e.template vop<V_AXPY>(x, z, nullptr, cst(cst_relax)));
  }
  st->iterations = conv ? it : s->max_iter;
  st->converged = conv;
  if (s->monitor_convergence && !conv) return not_converged(e.ctx, s, "jacobi", s->max_iter, rn / r0);
  return LFR_OK;
}

// ---- chebyshev_type%chebyshev_solve (:1868-1940) --------------------------
template <typename T>
int chebyshev(Eng<T>& e, const lfr_iterative_solver* s, T* x, const T* b, lfr_solver_stats* st) {
  T *z = e.vec(0), *r = e.vec(1), *xp = e.vec(2), *xo = e.vec(3);
  KTRY(e.template vop<V_SET>(x, nullptr, nullptr, cst(0.0)));
  KTRY(e.template vop<V_SET>(xp, nullptr, nullptr, cst(0.0)));
  KTRY(e.template vop<V_SET>(xo, nullptr, nullptr, cst(0.0)));
  double init_norm = 0;
  if (s->monitor_convergence) {
    KTRY(e.norm(b, K_R0));
    KTRY(e.read(K_R0, &init_norm));
    init_norm = init_norm > 1.0 ? init_norm : 1.0;
  }
  const double a1 = 2.0 / (s->lmax - s->lmin);
  const double a2 = (s->lmax + s->lmin) / (s->lmax - s->lmin);
  const double a_over_b = a1 / a2;
  double w = 1.0;
  for (int it = 1; it <= s->max_iter; ++it) {
    KTRY(e.apply(xo, z));                                     // z = A xo
    KTRY(e.template vop<V_AXPBY>(r, b, z, cst(1.0), cst(-1.0)));   // r = b - z
    KTRY(e.precon(r, z));                                     // z = P^{-1} r
    w = 1.0 / (1.0 - w / (4.0 * a2 * a2));
    const double wab = w * a_over_b;
    KTRY(e.template vop<V_AXPBY>(x, z, xo, cst(wab), cst(w)));
    KTRY(e.template vop<V_AXPY>(x, xp, nullptr, cst(1.0 - w)));
    if (it < s->max_iter) {
      KTRY(e.template vop<V_COPY>(xp, xo, nullptr, cst(0.0)));
      KTRY(e.template vop<V_COPY>(xo, x, nullptr, cst(0.0)));
    }
  }
  st->iterations = s->max_iter;
  st->converged = 1;
  if (s->monitor_convergence) {
    KTRY(e.apply(x, z));
    KTRY(e.template vop<V_AXPY>(z, b, nullptr, cst(-1.0)));
    KTRY(e.norm(z, K_RN));
    double fn = 0;
    KTRY(e.read(K_RN, &fn));
    st->r_nrm0 = init_norm;
    st->r_nrm = fn;
    if (fn != fn) return not_converged(e.ctx, s, "chebyshev", s->max_iter, fn);
  }
  return LFR_OK;
}

int nvec_for(const lfr_iterative_solver* s) {
  switch (s->method) {
    case LFR_SOLVER_CG: return 3;
    case LFR_SOLVER_BICGSTAB: return 8;
    case LFR_SOLVER_GMRES: return 3 + s->gcrk;
    case LFR_SOLVER_FGMRES: return 5 + 2 * s->gcrk;
    case LFR_SOLVER_GCR: return 3 + 2 * s->gcrk;
    case LFR_SOLVER_PRECONDITION_ONLY: return 2;
    case LFR_SOLVER_JACOBI: return 2;
    default: return 4;
  }
}

template <typename T>
int solve_t(lfr_ctx* ctx, const lfr_iterative_solver* s, const lfr_linop* A, const lfr_precon* P, T* x,
            const T* b, lfr_solver_stats* st) {
  Eng<T> e;
  const bool kry = s->method == LFR_SOLVER_GMRES || s->method == LFR_SOLVER_FGMRES;
  ctx->solve_bytes = -1.0;
  KTRY(e.init(ctx, A, P, nvec_for(s), kry ? s->gcrk : 0));
  int rc;
  switch (s->method) {
    case LFR_SOLVER_CG: rc = cg(e, s, x, b, st); break;
    case LFR_SOLVER_BICGSTAB: rc = bicgstab(e, s, x, b, st); break;
    case LFR_SOLVER_GMRES: rc = gmres(e, s, x, b, st); break;
    case LFR_SOLVER_FGMRES: rc = fgmres(e, s, x, b, st); break;
    case LFR_SOLVER_GCR: rc = gcr(e, s, x, b, st); break;
    case LFR_SOLVER_PRECONDITION_ONLY: rc = precondition_only(e, s, x, b, st); break;
    case LFR_SOLVER_JACOBI: rc = jacobi(e, s, x, b, st); break;
    default: rc = chebyshev(e, s, x, b, st);
  }
  ctx->solve_bytes = e.bytes_known ? e.bytes : -1.0;
  return rc;
}

// field reductions through the engine's kernels on the context workspace
int fields_dot(lfr_ctx* ctx, const lfr_vector_layout* L, const void* x, const void* y, int f0, int f1, bool take_sqrt,
               double* out) {
  const size_t need = align_up((size_t)LFR_MAX_FIELDS * kRedMaxBlocks * sizeof(double)) + 256;
  KTRY(lfr_work_reserve(ctx, need));
  double* part = (double*)ctx->work;
  double* res = (double*)((char*)ctx->work + need - 256);
  FieldNb nb{};
  nb.nfield = f1 - f0;
  for (int f = f0; f < f1; ++f) {
    const int64_t o = layout_off(*L, f), nf = L->undf[f];
    nb.nb[f - f0] = lfr::red_blocks(nf);
    double* pf = part + (int64_t)(f - f0) * kRedMaxBlocks;
    if (L->prec == 4)
      hipLaunchKernelGGL((k_fdot<float, 0>), dim3(nb.nb[f - f0]), dim3(kRedThreads), 0, ctx->stream, (float*)nullptr,
                         (const float*)x + o, (const float*)y + o, nf, (const double*)nullptr, cst(0.0), pf);
    else
      hipLaunchKernelGGL((k_fdot<double, 0>), dim3(nb.nb[f - f0]), dim3(kRedThreads), 0, ctx->stream,
                         (double*)nullptr, (const double*)x + o, (const double*)y + o, nf, (const double*)nullptr,
                         cst(0.0), pf);
  }
  hipLaunchKernelGGL(k_fields_final, dim3(1), dim3(kRedThreads), 0, ctx->stream, (const double*)part, nb, res, 0,
                     (int)take_sqrt);
  LFR_CHECK_LAUNCH(ctx);
  LFR_HIP(ctx, hipMemcpyAsync(ctx->host_scalars, res, sizeof(double), hipMemcpyDeviceToHost, ctx->stream));
  LFR_HIP(ctx, hipStreamSynchronize(ctx->stream));
  *out = ctx->host_scalars[0];
  return LFR_OK;
}

int check_space_field(lfr_ctx* ctx, const lfr_vector_layout& L, int f, const lfr_space* sp, const char* who) {
  LFR_REQUIRE(ctx, f >= 0 && f < L.nfield, "%s: field %d outside 0..%d", who, f, L.nfield - 1);
  LFR_REQUIRE(ctx, sp && (int64_t)sp->undf == L.undf[f], "%s: field %d has %lld values, its space has undf %d", who,
              f, (long long)L.undf[f], sp ? sp->undf : -1);
  return LFR_OK;
}

}  // namespace

extern "C" {

int lfr_vector_size(const lfr_vector_layout* L) { return L ? L->nfield : 0; }

int64_t lfr_vector_length(const lfr_vector_layout* L) { return layout_ok(L) ? layout_len(*L) : -1; }

int lfr_vector_dot(lfr_ctx* ctx, const lfr_vector_layout* L, const void* x, const void* y, double* out) {
  if (!ctx) return LFR_ERR_ARG;
  LFR_REQUIRE(ctx, layout_ok(L) && x && y && out, "lfr_vector_dot: bad arguments");
  return fields_dot(ctx, L, x, y, 0, L->nfield, false, out);
}

int lfr_vector_norm(lfr_ctx* ctx, const lfr_vector_layout* L, const void* x, double* out) {
  if (!ctx) return LFR_ERR_ARG;
  LFR_REQUIRE(ctx, layout_ok(L) && x && out, "lfr_vector_norm: bad arguments");
  return fields_dot(ctx, L, x, x, 0, L->nfield, true, out);
}

int lfr_vector_field_norm(lfr_ctx* ctx, const lfr_vector_layout* L, const void* x, int n, double* out) {
  if (!ctx) return LFR_ERR_ARG;
  LFR_REQUIRE(ctx, layout_ok(L) && x && out, "lfr_vector_field_norm: bad arguments");
  LFR_REQUIRE(ctx, n >= 1 && n <= L->nfield, "lfr_vector_field_norm: field %d outside 1..%d", n, L->nfield);
  return fields_dot(ctx, L, x, x, n - 1, n, true, out);
}

int lfr_linop_lma(lfr_ctx* ctx, const lfr_lma_op* A, lfr_linop** out) {
  if (!ctx || !out) return LFR_ERR_ARG;
  *out = nullptr;
  KTRY(lfr_check_op(ctx, A, "lfr_linop_lma"));
  auto* h = new lfr_linop();
  h->ctx = ctx;
  h->kind = 0;
  h->lma = *A;
  h->layout.prec = A->prec;
  h->layout.nfield = 1;
  h->layout.undf[0] = A->space->undf;
  *out = h;
  return LFR_OK;
}

int lfr_linop_block(lfr_ctx* ctx, const lfr_vector_layout* L, int nentry, const lfr_block_entry* ent,
                    lfr_linop** out) {
  if (!ctx || !out) return LFR_ERR_ARG;
  *out = nullptr;
  LFR_REQUIRE(ctx, layout_ok(L), "lfr_linop_block: bad vector layout");
  LFR_REQUIRE(ctx, nentry >= 1 && nentry <= 64 && ent, "lfr_linop_block: %d entries", nentry);
  std::vector<int> count(L->nfield, 0), last(L->nfield, -1);
  for (int q = 0; q < nentry; ++q) {
    const lfr_block_entry& b = ent[q];
    LFR_REQUIRE(ctx, b.matrix_d, "lfr_linop_block: entry %d has no matrix", q);
    KTRY(check_space_field(ctx, *L, b.row, b.to_space, "lfr_linop_block"));
    KTRY(check_space_field(ctx, *L, b.col, b.from_space, "lfr_linop_block"));
    if (b.transpose) {
      LFR_REQUIRE(ctx, b.plan && b.plan->kind == 1 && b.plan->space == b.to_space,
                  "lfr_linop_block: entry %d: transpose needs a kind-1 plan of its output space", q);
    } else if (b.plan) {
      LFR_REQUIRE(ctx, b.plan->kind == 0 && b.plan->space == b.to_space,
                  "lfr_linop_block: entry %d: plan must be a kind-0 plan of its output space", q);
    }
    LFR_REQUIRE(ctx, !(b.row == b.col && b.from_space != b.to_space && !b.plan && !b.transpose) ||
                         b.from_space->ncell == b.to_space->ncell,
                "lfr_linop_block: entry %d: spaces on different meshes", q);
    ++count[b.row];
    last[b.row] = q;
  }
  auto* h = new lfr_linop();
  h->ctx = ctx;
  h->kind = 1;
  h->layout = *L;
  h->blocks.assign(ent, ent + nentry);
  h->single.assign(L->nfield, -1);
  for (int f = 0; f < L->nfield; ++f)
    if (count[f] == 1) h->single[f] = last[f];
  *out = h;
  return LFR_OK;
}

int lfr_linop_callback(lfr_ctx* ctx, const lfr_vector_layout* L, lfr_apply_fn fn, void* user, lfr_linop** out) {
  if (!ctx || !out) return LFR_ERR_ARG;
  *out = nullptr;
  LFR_REQUIRE(ctx, layout_ok(L) && fn, "lfr_linop_callback: bad arguments");
  auto* h = new lfr_linop();
  h->ctx = ctx;
  h->kind = 2;
  h->layout = *L;
  h->fn = fn;
  h->user = user;
  *out = h;
  return LFR_OK;
}

int lfr_linop_destroy(lfr_linop* op) {
  delete op;
  return LFR_OK;
}

int lfr_linop_apply(lfr_ctx* ctx, const lfr_linop* op, const void* x, void* y) {
  if (!ctx) return LFR_ERR_ARG;
  LFR_REQUIRE(ctx, op && op->ctx == ctx && x && y, "lfr_linop_apply: bad arguments");
  LFR_REQUIRE(ctx, x != y, "lfr_linop_apply: y must not alias x");
  return linop_apply(ctx, op, x, y);
}

int lfr_precon_null(lfr_ctx* ctx, const lfr_vector_layout* L, lfr_precon** out) {
  if (!ctx || !out) return LFR_ERR_ARG;
  *out = nullptr;
  LFR_REQUIRE(ctx, layout_ok(L), "lfr_precon_null: bad vector layout");
  auto* h = new lfr_precon();
  h->ctx = ctx;
  h->kind = 0;
  h->layout = *L;
  *out = h;
  return LFR_OK;
}

int lfr_precon_diagonal(lfr_ctx* ctx, const lfr_vector_layout* L, const void* diag, lfr_precon** out) {
  if (!ctx || !out) return LFR_ERR_ARG;
  *out = nullptr;
  LFR_REQUIRE(ctx, layout_ok(L) && diag, "lfr_precon_diagonal: bad arguments");
  auto* h = new lfr_precon();
  h->ctx = ctx;
  h->kind = 1;
  h->layout = *L;
  h->diag = diag;
  *out = h;
  return LFR_OK;
}

int lfr_precon_columnwise(lfr_ctx* ctx, const lfr_vector_layout* L, const void* cm, int nrow, const lfr_space* sp,
                          const int32_t* ind, int fp64, lfr_precon** out) {
  if (!ctx || !out) return LFR_ERR_ARG;
  *out = nullptr;
  LFR_REQUIRE(ctx, layout_ok(L) && L->nfield == 1 && cm && sp && ind && nrow > 0,
              "lfr_precon_columnwise: bad arguments (one-field vectors only)");
  LFR_REQUIRE(ctx, (int64_t)sp->undf == L->undf[0], "lfr_precon_columnwise: vector length %lld, space undf %d",
              (long long)L->undf[0], sp->undf);
  LFR_REQUIRE(ctx, (int64_t)sp->max_first + nrow - 1 <= sp->undf,
              "lfr_precon_columnwise: nrow %d runs past the space's undf", nrow);
  auto* h = new lfr_precon();
  h->ctx = ctx;
  h->kind = 2;
  h->layout = *L;
  h->cm = cm;
  h->nrow = nrow;
  h->space = sp;
  h->ind = ind;
  h->fp64 = fp64;
  *out = h;
  return LFR_OK;
}

int lfr_precon_callback(lfr_ctx* ctx, const lfr_vector_layout* L, lfr_apply_fn fn, void* user, lfr_precon** out) {
  if (!ctx || !out) return LFR_ERR_ARG;
  *out = nullptr;
  LFR_REQUIRE(ctx, layout_ok(L) && fn, "lfr_precon_callback: bad arguments");
  auto* h = new lfr_precon();
  h->ctx = ctx;
  h->kind = 3;
  h->layout = *L;
  h->fn = fn;
  h->user = user;
  *out = h;
  return LFR_OK;
}

int lfr_precon_destroy(lfr_precon* p) {
  delete p;
  return LFR_OK;
}

int lfr_precon_apply(lfr_ctx* ctx, const lfr_precon* p, const void* x, void* y) {
  if (!ctx) return LFR_ERR_ARG;
  LFR_REQUIRE(ctx, p && p->ctx == ctx && x && y, "lfr_precon_apply: bad arguments");
  LFR_REQUIRE(ctx, x != y, "lfr_precon_apply: y must not alias x");
  return precon_apply(ctx, p, x, y);
}

int lfr_iterative_solve(lfr_ctx* ctx, const lfr_iterative_solver* s, const lfr_linop* A, const lfr_precon* P,
                        void* x, const void* b, lfr_solver_stats* st) {
  if (!ctx) return LFR_ERR_ARG;
  LFR_REQUIRE(ctx, s && A && P && x && b && st, "lfr_iterative_solve: null argument");
  LFR_REQUIRE(ctx, A->ctx == ctx && P->ctx == ctx, "lfr_iterative_solve: operator or preconditioner of another context");
  LFR_REQUIRE(ctx, layout_eq(A->layout, P->layout), "lfr_iterative_solve: operator and preconditioner layouts differ");
  LFR_REQUIRE(ctx, s->method >= LFR_SOLVER_CG && s->method <= LFR_SOLVER_CHEBYSHEV, "lfr_iterative_solve: method %d",
              s->method);
  LFR_REQUIRE(ctx, s->max_iter >= 0, "lfr_iterative_solve: max_iter %d", s->max_iter);
  const bool kry = s->method == LFR_SOLVER_GMRES || s->method == LFR_SOLVER_FGMRES || s->method == LFR_SOLVER_GCR;
  LFR_REQUIRE(ctx, !kry || (s->gcrk >= 1 && s->gcrk <= 256), "lfr_iterative_solve: gcrk %d outside 1..256", s->gcrk);
  LFR_REQUIRE(ctx, s->method != LFR_SOLVER_CHEBYSHEV || s->lmax != s->lmin,
              "lfr_iterative_solve: chebyshev needs lmax != lmin");
  LFR_REQUIRE(ctx, x != b, "lfr_iterative_solve: x must not alias b");
  st->iterations = 0;
  st->converged = 0;
  st->r_nrm0 = st->r_nrm = 0.0;
  if (A->layout.prec == 4) return solve_t<float>(ctx, s, A, P, (float*)x, (const float*)b, st);
  return solve_t<double>(ctx, s, A, P, (double*)x, (const double*)b, st);
}

int lfr_iterative_solve_bytes(lfr_ctx* ctx, double* bytes) {
  if (!ctx) return LFR_ERR_ARG;
  LFR_REQUIRE(ctx, bytes, "lfr_iterative_solve_bytes: null argument");
  *bytes = ctx->solve_bytes;
  return LFR_OK;
}

int lfr_mm_diagonal(lfr_ctx* ctx, int prec, void* diag, const void* M, const lfr_space* sp, const lfr_plan* plan) {
  if (!ctx) return LFR_ERR_ARG;
  LFR_REQUIRE(ctx, (prec == 4 || prec == 8) && diag && M && sp, "lfr_mm_diagonal: bad arguments");
  LFR_REQUIRE(ctx, !plan || plan->space == sp, "lfr_mm_diagonal: plan of another space");
  const int nl = sp->nlayers;
  const int64_t n3 = (int64_t)sp->ncell * nl;
  if (!plan) {
    const unsigned g = (unsigned)((n3 * sp->ndf + kT - 1) / kT);
    if (prec == 4)
      hipLaunchKernelGGL(k_mm_diag_dg<float>, dim3(g), dim3(kT), 0, ctx->stream, (float*)diag, (const float*)M,
                         sp->map0_d, sp->ncell, nl, sp->ndf, n3);
    else
      hipLaunchKernelGGL(k_mm_diag_dg<double>, dim3(g), dim3(kT), 0, ctx->stream, (double*)diag, (const double*)M,
                         sp->map0_d, sp->ncell, nl, sp->ndf, n3);
    LFR_CHECK_LAUNCH(ctx);
    return LFR_OK;
  }
  for (const auto& c : plan->classes) {
    if (c.ncol == 0) continue;
    LFR_REQUIRE(ctx, c.gsize <= 2, "lfr_mm_diagonal: %d entries of one cell in a column", c.gsize);
    const unsigned g = (unsigned)(((int64_t)c.ncol * c.height + kT - 1) / kT);
    if (prec == 4)
      hipLaunchKernelGGL(k_mm_diag_plan<float>, dim3(g), dim3(kT), 0, ctx->stream, (float*)diag, (const float*)M,
                         c.col_start_d, c.ent_d, c.ncol, c.height, c.ngroup, c.gsize, nl, n3, sp->ndf);
    else
      hipLaunchKernelGGL(k_mm_diag_plan<double>, dim3(g), dim3(kT), 0, ctx->stream, (double*)diag,
                         (const double*)M, c.col_start_d, c.ent_d, c.ncol, c.height, c.ngroup, c.gsize, nl, n3,
                         sp->ndf);
    LFR_CHECK_LAUNCH(ctx);
  }
  return LFR_OK;
}

}  // extern "C"
