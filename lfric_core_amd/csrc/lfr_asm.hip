// Local-matrix assembly by quadrature on the device, lowest-order W0.
//
// Mass matrix: compute_mass_matrix_w_scalar_code
// (components/science/source/kernel/fem/sci_compute_mass_matrix_kernel_w_scalar_mod.F90:150-174)
// with w_scalar = W0 and the coordinate field chi on W0, Cartesian (X, Y, Z)
// coordinates: the Jacobian of the reference-cube map
// (components/science/source/kernel/geometry/sci_coordinate_jacobian_mod.F90:330-345,
// jac(i, dir) = sum_df chi_i(df) dphi_df/dxi_dir, df in order) and its
// cofactor determinant (:430-440); mm(ik, df, df2) for df >= df2 is the sum,
// qp_v outer and qp_h inner, of ((((w_h w_v) phi_df) phi_df2) detJ), then
// mirrored.
//
// Stiffness (no reference kernel; used for the assembled diffusion operator):
// K(a, b) = sum_q w_h w_v (J^-T grad phi_a) . (J^-T grad phi_b) |detJ|, with
// J^-T from the cofactors.
//
// Quadrature: nq-point Gauss-Legendre on [0, 1] in each direction (the
// published rule: roots of P_nq by Newton's method from the Chebyshev
// guesses, weights 1/((1 - t^2) P'(t)^2)), horizontal points x outer, y inner
// (quadrature_xyoz_mod.F90:235-242).  W0 basis: trilinear, local dofs in the
// reference element's vertex order (bottom SW, SE, NE, NW, then top).
//
// One thread per (row a, cell-layer ik), ik fastest, so the 8 values a thread
// writes, M(ik, a, b) at ik + n3 (a + 8 b), are coalesced across a wavefront.
#include <cmath>
#include <vector>

#include "lfr_internal.h"

namespace {

constexpr int kMaxQ = 6;
constexpr int kMaxQp = kMaxQ * kMaxQ * kMaxQ;

// per quadrature point (v outer, h = (x outer, y inner) inner): basis[8],
// gradient[3][8], w_h, w_v
struct QpTable {
  double b[8];
  double g[3][8];
  double wh, wv;
};

void gauss_legendre01(int n, double* x, double* w) {
  for (int i = 0; i < n; ++i) {
    double t = std::cos(M_PI * (i + 0.75) / (n + 0.5));
    double dp = 1.0;
    for (int it = 0; it < 100; ++it) {
      double p0 = 1.0, p1 = t;
      for (int k = 2; k <= n; ++k) {
        const double p2 = ((2.0 * k - 1.0) * t * p1 - (k - 1.0) * p0) / k;
        p0 = p1;
        p1 = p2;
      }
      const double pn = n == 0 ? 1.0 : p1;
      const double pm = n == 1 ? 1.0 : p0;
      dp = n * (t * pn - pm) / (t * t - 1.0);
      const double dt = pn / dp;
      t -= dt;
      if (std::fabs(dt) < 1e-16) break;
    }
    // ascending points on [0, 1]
    x[n - 1 - i] = 0.5 * (t + 1.0);
    w[n - 1 - i] = 0.5 * 2.0 / ((1.0 - t * t) * dp * dp);
  }
}

__global__ void __launch_bounds__(256)
k_w0_assemble(double* __restrict__ mass, double* __restrict__ stiff, const double* __restrict__ chi1,
              const double* __restrict__ chi2, const double* __restrict__ chi3, const int32_t* __restrict__ map0,
              const QpTable* __restrict__ tab, int nqp, int ncell, int nl) {
  const int64_t n3 = (int64_t)ncell * nl;
  const int64_t t = (int64_t)blockIdx.x * 256 + threadIdx.x;
  if (t >= 8 * n3) return;
  const int a = (int)(t / n3);
  const int64_t ik = t - (int64_t)a * n3;
  const int c = (int)(ik / nl);
  const int k = (int)(ik - (int64_t)c * nl);
  double x1[8], x2[8], x3[8];
#pragma unroll
  for (int df = 0; df < 8; ++df) {
    const int64_t d = (int64_t)map0[(int64_t)c * 8 + df] + k;
    x1[df] = chi1[d];
    x2[df] = chi2[d];
    x3[df] = chi3[d];
  }
  double m[8], s[8];
#pragma unroll
  for (int b = 0; b < 8; ++b) {
    m[b] = 0.0;
    s[b] = 0.0;
  }
  for (int q = 0; q < nqp; ++q) {
    const QpTable& T = tab[q];
    double j[3][3];
#pragma unroll
    for (int i = 0; i < 3; ++i)
#pragma unroll
      for (int dir = 0; dir < 3; ++dir) j[i][dir] = 0.0;
#pragma unroll
    for (int df = 0; df < 8; ++df)
#pragma unroll
      for (int dir = 0; dir < 3; ++dir) {
        const double gd = T.g[dir][df];
        j[0][dir] = j[0][dir] + x1[df] * gd;
        j[1][dir] = j[1][dir] + x2[df] * gd;
        j[2][dir] = j[2][dir] + x3[df] * gd;
      }
    const double c00 = j[1][1] * j[2][2] - j[1][2] * j[2][1];
    const double c01 = j[1][0] * j[2][2] - j[1][2] * j[2][0];
    const double c02 = j[1][0] * j[2][1] - j[1][1] * j[2][0];
    const double dj = j[0][0] * c00 - j[0][1] * c01 + j[0][2] * c02;
    const double wq = T.wh * T.wv;
    // mass: df = max(a, b), df2 = min(a, b) as the mirrored lower triangle
#pragma unroll
    for (int b = 0; b < 8; ++b) {
      const int df = a > b ? a : b, df2 = a > b ? b : a;
      const double v = ((wq * T.b[df]) * T.b[df2]) * dj;
      m[b] = m[b] + v;
    }
    if (stiff) {
      // J^-T = cof(J) / det: physical gradient g = J^-T ghat
      double inv[3][3];   // inv[i][dir] = (J^-1)[dir][i]
      inv[0][0] = c00;
      inv[0][1] = -(j[0][1] * j[2][2] - j[0][2] * j[2][1]);
      inv[0][2] = j[0][1] * j[1][2] - j[0][2] * j[1][1];
      inv[1][0] = -c01;
      inv[1][1] = j[0][0] * j[2][2] - j[0][2] * j[2][0];
      inv[1][2] = -(j[0][0] * j[1][2] - j[0][2] * j[1][0]);
      inv[2][0] = c02;
      inv[2][1] = -(j[0][0] * j[2][1] - j[0][1] * j[2][0]);
      inv[2][2] = j[0][0] * j[1][1] - j[0][1] * j[1][0];
      const double rd = 1.0 / dj;
      double ga[3];
#pragma unroll
      for (int i = 0; i < 3; ++i)
        ga[i] = (inv[0][i] * T.g[0][a] + inv[1][i] * T.g[1][a] + inv[2][i] * T.g[2][a]) * rd;
      const double wa = wq * fabs(dj);
#pragma unroll
      for (int b = 0; b < 8; ++b) {
        double dot = 0.0;
#pragma unroll
        for (int i = 0; i < 3; ++i) {
          const double gb = (inv[0][i] * T.g[0][b] + inv[1][i] * T.g[1][b] + inv[2][i] * T.g[2][b]) * rd;
          dot = dot + ga[i] * gb;
        }
        s[b] = s[b] + wa * dot;
      }
    }
  }
#pragma unroll
  for (int b = 0; b < 8; ++b) {
    mass[ik + n3 * (a + 8 * b)] = m[b];
    if (stiff) stiff[ik + n3 * (a + 8 * b)] = s[b];
  }
}

}  // namespace

extern "C" int lfr_compute_mass_matrix_w0(lfr_ctx* ctx, int nq, const void* chi1_d, const void* chi2_d,
                                          const void* chi3_d, const lfr_space* space, void* mass_d,
                                          void* stiff_d) {
  if (!ctx) return LFR_ERR_ARG;
  LFR_REQUIRE(ctx, chi1_d && chi2_d && chi3_d && space && mass_d, "lfr_compute_mass_matrix_w0: null argument");
  LFR_REQUIRE(ctx, nq >= 1 && nq <= kMaxQ, "lfr_compute_mass_matrix_w0: nq must be 1..%d, got %d", kMaxQ, nq);
  LFR_REQUIRE(ctx, space->ndf == 8, "lfr_compute_mass_matrix_w0: lowest-order W0 only (ndf %d, expected 8)",
              space->ndf);
  // the four top dofs must follow their bottom dofs (vertex columns)
  for (int c = 0; c < space->ncell; ++c)
    for (int v = 0; v < 4; ++v)
      LFR_REQUIRE(ctx, space->map_h[(size_t)c * 8 + 4 + v] == space->map_h[(size_t)c * 8 + v] + 1,
                  "lfr_compute_mass_matrix_w0: cell %d is not a lowest-order W0 dofmap", c + 1);
  double x[kMaxQ], w[kMaxQ];
  gauss_legendre01(nq, x, w);
  static const int vtx[8][3] = {{0, 0, 0}, {1, 0, 0}, {1, 1, 0}, {0, 1, 0},
                                {0, 0, 1}, {1, 0, 1}, {1, 1, 1}, {0, 1, 1}};
  std::vector<QpTable> tab;
  for (int qv = 0; qv < nq; ++qv)
    for (int qx = 0; qx < nq; ++qx)
      for (int qy = 0; qy < nq; ++qy) {
        QpTable T;
        const double p[3] = {x[qx], x[qy], x[qv]};
        for (int df = 0; df < 8; ++df) {
          double f[3], d[3];
          for (int e = 0; e < 3; ++e) {
            f[e] = vtx[df][e] ? p[e] : 1.0 - p[e];
            d[e] = vtx[df][e] ? 1.0 : -1.0;
          }
          T.b[df] = f[0] * f[1] * f[2];
          T.g[0][df] = d[0] * f[1] * f[2];
          T.g[1][df] = f[0] * d[1] * f[2];
          T.g[2][df] = f[0] * f[1] * d[2];
        }
        T.wh = w[qx] * w[qy];
        T.wv = w[qv];
        tab.push_back(T);
      }
  const size_t tb = tab.size() * sizeof(QpTable);
  int rc = lfr_work_reserve(ctx, tb);
  if (rc) return rc;
  LFR_HIP(ctx, hipMemcpyAsync(ctx->work, tab.data(), tb, hipMemcpyHostToDevice, ctx->stream));
  const int64_t n = 8 * (int64_t)space->ncell * space->nlayers;
  hipLaunchKernelGGL(k_w0_assemble, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, ctx->stream, (double*)mass_d,
                     (double*)stiff_d, (const double*)chi1_d, (const double*)chi2_d, (const double*)chi3_d,
                     space->map0_d, (const QpTable*)ctx->work, (int)tab.size(), space->ncell, space->nlayers);
  LFR_CHECK_LAUNCH(ctx);
  // the table is host memory of this call: finish the copy before it goes
  LFR_HIP(ctx, hipStreamSynchronize(ctx->stream));
  return LFR_OK;
}

// ---------------------------------------------------------------------------
// Bilinear forms of the lowest-order de Rham complex from basis tables given
// by the caller (as the PSy layer hands them to the reference kernels):
//
//   out(ik, df, df2) = sum_{qp_v} sum_{qp_h} integrand(df, df2, qp)
//
// with, wt = w_h w_v, L the row table and R the column table,
//   kMapDj      ((wt L_df) R_df2) detJ                  scalar masses
//   kMapCov     (wt (Jt L_df).(Jt R_df2)) detJ          W1 mass, grad
//   kMapContra  (wt (J L_df).(J R_df2)) / detJ          W2 mass, curl
//   kMapRef     (wt L_df) R_df2                         div (no metric)
// Jt = transpose of the cofactor inverse
// (pointwise_coordinate_jacobian_inverse, sci_coordinate_jacobian_mod.F90:1140-1161),
// J and detJ from the chi gradients (:993-1103, Cartesian).  The statements
// follow compute_derham_matrices_code
// (components/science/source/kernel/fem/sci_compute_derham_matrices_kernel_mod.F90:241-352),
// compute_mass_matrix_w_scalar_code
// (sci_compute_mass_matrix_kernel_w_scalar_mod.F90:150-174) and
// compute_mass_matrix_w2_code (sci_compute_mass_matrix_kernel_w2_mod.F90:131-162).
// Symmetric forms (the masses) are computed for df >= df2 and mirrored
// (for the vector maps the dot product has the same bits either way:
// products commute exactly).
// ---------------------------------------------------------------------------
namespace {

enum { kMapDj = 0, kMapCov = 1, kMapContra = 2, kMapRef = 3 };

template <int NR, int NC, int CL, int CR>
struct FormQp {
  double l[CL][NR];   // row basis (component, dof) at the point
  double r[CR][NC];   // column basis
  double g[3][8];     // chi reference gradients
  double wh, wv;
};

// Kernel: a workgroup holds CPB = 256 / NR cell-layers, thread (li, a) =
// (cell-layer, row), li fastest, so output stores are coalesced.  Per
// quadrature point the row-0 thread of a cell-layer forms J, detJ and the
// map (J or Jt) into LDS; then every thread maps its share of the column
// vectors (columns a, a + NR, ...) into LDS and its own row vector into
// registers; then each thread sums its row.  The Jacobian is formed once
// per cell-layer and each mapped column vector once, not once per row (71 ms
// -> 45 ms for all ten outputs at C96L70).  LDS is double-buffered by point
// parity.
template <int NR, int NC, int CL, int CR, int MAP, bool SYM>
__global__ void __launch_bounds__(256)
k_form(double* __restrict__ out, const double* __restrict__ chi1, const double* __restrict__ chi2,
              const double* __restrict__ chi3, const int32_t* __restrict__ map0,
              const FormQp<NR, NC, CL, CR>* __restrict__ tab, int nqp, int ncell, int nl) {
  constexpr bool kVec = MAP == kMapCov || MAP == kMapContra;
  constexpr int KC = kVec ? 3 : 1;
  constexpr int CPB = 256 / NR;
  constexpr int NCV = kVec ? NC : 1;            // LDS column vectors (vector maps only)
  __shared__ double sm[2][CPB][10];             // map (9) and detJ
  __shared__ double srv[2][CPB][NCV][KC];
  const int64_t n3 = (int64_t)ncell * nl;
  const int li = threadIdx.x % CPB;
  const int a = threadIdx.x / CPB;              // row; a == NR: idle thread
  const int64_t ik = (int64_t)blockIdx.x * CPB + li;
  const bool valid = a < NR && ik < n3;
  const int64_t ikc = ik < n3 ? ik : n3 - 1;    // idle lanes work on a real cell-layer
  double x1[8], x2[8], x3[8];
  if (a == 0) {
    const int c = (int)(ikc / nl);
    const int k = (int)(ikc - (int64_t)c * nl);
#pragma unroll
    for (int df = 0; df < 8; ++df) {
      const int64_t d = (int64_t)map0[(int64_t)c * 8 + df] + k;
      x1[df] = chi1[d];
      x2[df] = chi2[d];
      x3[df] = chi3[d];
    }
  }
  double acc[NC];
#pragma unroll
  for (int b = 0; b < NC; ++b) acc[b] = 0.0;
  const int ar = a < NR ? a : 0;
  for (int q = 0; q < nqp; ++q) {
    const FormQp<NR, NC, CL, CR>& T = tab[q];
    const int pb = q & 1;
    if (MAP != kMapRef && a == 0) {   // the divergence form has no metric term
      double j[3][3];
#pragma unroll
      for (int i = 0; i < 3; ++i)
#pragma unroll
        for (int dir = 0; dir < 3; ++dir) j[i][dir] = 0.0;
#pragma unroll
      for (int df = 0; df < 8; ++df)
#pragma unroll
        for (int dir = 0; dir < 3; ++dir) {
          const double gd = T.g[dir][df];
          j[0][dir] = j[0][dir] + x1[df] * gd;
          j[1][dir] = j[1][dir] + x2[df] * gd;
          j[2][dir] = j[2][dir] + x3[df] * gd;
        }
      const double dj = j[0][0] * (j[1][1] * j[2][2] - j[1][2] * j[2][1]) -
                        j[0][1] * (j[1][0] * j[2][2] - j[1][2] * j[2][0]) +
                        j[0][2] * (j[1][0] * j[2][1] - j[1][1] * j[2][0]);
      double m[3][3];
      if (MAP == kMapContra) {
#pragma unroll
        for (int i = 0; i < 3; ++i)
#pragma unroll
          for (int d = 0; d < 3; ++d) m[i][d] = j[i][d];
      } else if (MAP == kMapCov) {
        const double idj = 1.0 / dj;
        double inv[3][3];
        inv[0][0] = (j[1][1] * j[2][2] - j[1][2] * j[2][1]) * idj;
        inv[0][1] = -(j[0][1] * j[2][2] - j[0][2] * j[2][1]) * idj;
        inv[0][2] = (j[0][1] * j[1][2] - j[0][2] * j[1][1]) * idj;
        inv[1][0] = -(j[1][0] * j[2][2] - j[1][2] * j[2][0]) * idj;
        inv[1][1] = (j[0][0] * j[2][2] - j[0][2] * j[2][0]) * idj;
        inv[1][2] = -(j[0][0] * j[1][2] - j[0][2] * j[1][0]) * idj;
        inv[2][0] = (j[1][0] * j[2][1] - j[1][1] * j[2][0]) * idj;
        inv[2][1] = -(j[0][0] * j[2][1] - j[0][1] * j[2][0]) * idj;
        inv[2][2] = (j[0][0] * j[1][1] - j[0][1] * j[1][0]) * idj;
#pragma unroll
        for (int i = 0; i < 3; ++i)
#pragma unroll
          for (int d = 0; d < 3; ++d) m[i][d] = inv[d][i];
      } else {
#pragma unroll
        for (int i = 0; i < 3; ++i)
#pragma unroll
          for (int d = 0; d < 3; ++d) m[i][d] = 0.0;
      }
#pragma unroll
      for (int i = 0; i < 3; ++i)
#pragma unroll
        for (int d = 0; d < 3; ++d) sm[pb][li][3 * i + d] = m[i][d];
      sm[pb][li][9] = dj;
    }
    __syncthreads();
    double m[3][3];
#pragma unroll
    for (int i = 0; i < 3; ++i)
#pragma unroll
      for (int d = 0; d < 3; ++d) m[i][d] = sm[pb][li][3 * i + d];
    const double dj = sm[pb][li][9];
    double lv[KC];
    if (kVec) {
      // this thread's share of the column vectors, and its row vector
      for (int b = ar; b < NC && a < NR; b += NR) {
#pragma unroll
        for (int i = 0; i < KC; ++i) {
          double sacc = 0.0;
#pragma unroll
          for (int d = 0; d < 3; ++d) sacc = sacc + m[i][d] * T.r[d][b];
          srv[pb][li][b < NCV ? b : 0][i] = sacc;
        }
      }
#pragma unroll
      for (int i = 0; i < KC; ++i) {
        double sacc = 0.0;
#pragma unroll
        for (int d = 0; d < 3; ++d) sacc = sacc + m[i][d] * T.l[d][ar];
        lv[i] = sacc;
      }
      __syncthreads();
    } else {
      lv[0] = T.l[0][ar];
    }
    const double wt = T.wh * T.wv;
#pragma unroll
    for (int b = 0; b < NC; ++b) {
      double v;
      if (kVec) {
        double dot = 0.0;
#pragma unroll
        for (int i = 0; i < 3; ++i) dot = dot + lv[i] * srv[pb][li][b < NCV ? b : 0][i < KC ? i : 0];
        v = MAP == kMapCov ? (wt * dot) * dj : (wt * dot) / dj;
      } else {
        double l0 = lv[0], r0 = T.r[0][b];
        if (SYM && b > ar) {   // mirrored entry: computed as (df, df2) = (b, a)
          l0 = T.r[0][b];
          r0 = lv[0];
        }
        v = MAP == kMapDj ? ((wt * l0) * r0) * dj : (wt * l0) * r0;
      }
      acc[b] = acc[b] + v;
    }
  }
  if (valid) {
#pragma unroll
    for (int b = 0; b < NC; ++b) out[ik + n3 * (a + (int64_t)NR * b)] = acc[b];
  }
}

// Fortran-order host table (ncomp, ndf, nqp_h, nqp_v) element
inline double tab_at(const double* t, int ncomp, int ndf, int nqp_h, int c, int df, int qh, int qv) {
  return t[c + ncomp * (df + (size_t)ndf * (qh + (size_t)nqp_h * qv))];
}

template <int NR, int NC, int CL, int CR, int MAP, bool SYM>
int form_assemble(lfr_ctx* ctx, int nqp_h, int nqp_v, const double* L, const double* R, const double* dchi,
                  const double* wh, const double* wv, const void* chi1, const void* chi2, const void* chi3,
                  const lfr_space* sp, void* out) {
  using Q = FormQp<NR, NC, CL, CR>;
  std::vector<Q> tab;
  for (int qv = 0; qv < nqp_v; ++qv)
    for (int qh = 0; qh < nqp_h; ++qh) {
      Q T;
      for (int c = 0; c < CL; ++c)
        for (int df = 0; df < NR; ++df) T.l[c][df] = tab_at(L, CL, NR, nqp_h, c, df, qh, qv);
      for (int c = 0; c < CR; ++c)
        for (int df = 0; df < NC; ++df) T.r[c][df] = tab_at(R, CR, NC, nqp_h, c, df, qh, qv);
      for (int dir = 0; dir < 3; ++dir)
        for (int df = 0; df < 8; ++df) T.g[dir][df] = tab_at(dchi, 3, 8, nqp_h, dir, df, qh, qv);
      T.wh = wh[qh];
      T.wv = wv[qv];
      tab.push_back(T);
    }
  const size_t tb = tab.size() * sizeof(Q);
  int rc = lfr_work_reserve(ctx, tb);
  if (rc) return rc;
  LFR_HIP(ctx, hipMemcpyAsync(ctx->work, tab.data(), tb, hipMemcpyHostToDevice, ctx->stream));
  const int64_t n3 = (int64_t)sp->ncell * sp->nlayers;
  constexpr int CPB = 256 / NR;
  hipLaunchKernelGGL((k_form<NR, NC, CL, CR, MAP, SYM>), dim3((unsigned)((n3 + CPB - 1) / CPB)), dim3(256),
                     0, ctx->stream, (double*)out, (const double*)chi1, (const double*)chi2, (const double*)chi3,
                     sp->map0_d, (const Q*)ctx->work, (int)tab.size(), sp->ncell, sp->nlayers);
  LFR_CHECK_LAUNCH(ctx);
  // the table lives in ctx->work and comes from this call's host memory:
  // finish before the next form reuses either
  LFR_HIP(ctx, hipStreamSynchronize(ctx->stream));
  return LFR_OK;
}

int check_chi(lfr_ctx* ctx, const lfr_space* sp, const char* who) {
  LFR_REQUIRE(ctx, sp->ndf == 8, "%s: the chi space must be lowest-order W0 (ndf %d)", who, sp->ndf);
  for (int c = 0; c < sp->ncell; ++c)
    for (int v = 0; v < 4; ++v)
      LFR_REQUIRE(ctx, sp->map_h[(size_t)c * 8 + 4 + v] == sp->map_h[(size_t)c * 8 + v] + 1,
                  "%s: cell %d is not a lowest-order W0 dofmap", who, c + 1);
  return LFR_OK;
}

}  // namespace

extern "C" int lfr_compute_mass_matrix(lfr_ctx* ctx, int kind, int ndf, int nqp_h, int nqp_v, const double* basis,
                                       const double* diff_basis_chi, const double* wqp_h, const double* wqp_v,
                                       const void* chi1_d, const void* chi2_d, const void* chi3_d,
                                       const lfr_space* chi_space, void* mass_d) {
  if (!ctx) return LFR_ERR_ARG;
  LFR_REQUIRE(ctx, basis && diff_basis_chi && wqp_h && wqp_v && chi1_d && chi2_d && chi3_d && chi_space && mass_d,
              "lfr_compute_mass_matrix: null argument");
  LFR_REQUIRE(ctx, kind == 0 || kind == 1, "lfr_compute_mass_matrix: kind must be 0 (scalar) or 1 (W2), got %d",
              kind);
  LFR_REQUIRE(ctx, nqp_h >= 1 && nqp_v >= 1 && nqp_h * nqp_v <= 4096, "lfr_compute_mass_matrix: bad point counts");
  int rc = check_chi(ctx, chi_space, "lfr_compute_mass_matrix");
  if (rc) return rc;
  const double *B = basis, *G = diff_basis_chi;
#define LFR_MASS(N, C, MAP)                                                                                   \
  return form_assemble<N, N, C, C, MAP, true>(ctx, nqp_h, nqp_v, B, B, G, wqp_h, wqp_v, chi1_d, chi2_d, chi3_d, \
                                              chi_space, mass_d)
  if (kind == 0) {
    switch (ndf) {
      case 1: LFR_MASS(1, 1, kMapDj);
      case 2: LFR_MASS(2, 1, kMapDj);
      case 8: LFR_MASS(8, 1, kMapDj);
      default: break;
    }
  } else {
    switch (ndf) {
      case 2: LFR_MASS(2, 3, kMapContra);
      case 4: LFR_MASS(4, 3, kMapContra);
      case 6: LFR_MASS(6, 3, kMapContra);
      default: break;
    }
  }
#undef LFR_MASS
  return lfr_set_error(ctx, LFR_ERR_ARG, "lfr_compute_mass_matrix: no kernel for kind %d with %d dofs per cell",
                       kind, ndf);
}

extern "C" int lfr_compute_derham_matrices(lfr_ctx* ctx, const lfr_derham_tables* t, const void* chi1_d,
                                           const void* chi2_d, const void* chi3_d, const lfr_space* chi_space,
                                           void* mm0, void* mm1, void* mm2, void* mm2b, void* mm3, void* mmt,
                                           void* grad, void* curl, void* div, void* broken_div) {
  if (!ctx) return LFR_ERR_ARG;
  LFR_REQUIRE(ctx, t && chi1_d && chi2_d && chi3_d && chi_space, "lfr_compute_derham_matrices: null argument");
  LFR_REQUIRE(ctx, t->nqp_h >= 1 && t->nqp_v >= 1 && t->nqp_h * t->nqp_v <= 4096 && t->wqp_h && t->wqp_v &&
                       t->diff_basis_chi,
              "lfr_compute_derham_matrices: bad quadrature");
  int rc = check_chi(ctx, chi_space, "lfr_compute_derham_matrices");
  if (rc) return rc;
  const int h = t->nqp_h, v = t->nqp_v;
  const double *G = t->diff_basis_chi, *wh = t->wqp_h, *wv = t->wqp_v;
  struct Need {
    void* out;
    const double* a;
    const double* b;
    const char* what;
  };
  const Need need[10] = {{mm0, t->basis_w0, t->basis_w0, "basis_w0"},
                         {mm1, t->basis_w1, t->basis_w1, "basis_w1"},
                         {mm2, t->basis_w2, t->basis_w2, "basis_w2"},
                         {mm2b, t->basis_w2b, t->basis_w2b, "basis_w2b"},
                         {mm3, t->basis_w3, t->basis_w3, "basis_w3"},
                         {mmt, t->basis_wt, t->basis_wt, "basis_wt"},
                         {grad, t->basis_w1, t->diff_basis_w0, "basis_w1 and diff_basis_w0"},
                         {curl, t->basis_w2, t->diff_basis_w1, "basis_w2 and diff_basis_w1"},
                         {div, t->basis_w3, t->diff_basis_w2, "basis_w3 and diff_basis_w2"},
                         {broken_div, t->basis_w3, t->diff_basis_w2b, "basis_w3 and diff_basis_w2b"}};
  for (const Need& n : need)
    LFR_REQUIRE(ctx, !n.out || (n.a && n.b), "lfr_compute_derham_matrices: an output needs %s", n.what);
#define LFR_FORM(OUT, NR, NC, CL, CR, MAP, SYM, L, R)                                                          \
  if (OUT) {                                                                                                  \
    rc = form_assemble<NR, NC, CL, CR, MAP, SYM>(ctx, h, v, L, R, G, wh, wv, chi1_d, chi2_d, chi3_d, chi_space, \
                                                 OUT);                                                        \
    if (rc) return rc;                                                                                        \
  }
  LFR_FORM(mm0, 8, 8, 1, 1, kMapDj, true, t->basis_w0, t->basis_w0)
  LFR_FORM(mm1, 12, 12, 3, 3, kMapCov, true, t->basis_w1, t->basis_w1)
  LFR_FORM(mm2, 6, 6, 3, 3, kMapContra, true, t->basis_w2, t->basis_w2)
  LFR_FORM(mm2b, 6, 6, 3, 3, kMapContra, true, t->basis_w2b, t->basis_w2b)
  LFR_FORM(mm3, 1, 1, 1, 1, kMapDj, true, t->basis_w3, t->basis_w3)
  LFR_FORM(mmt, 2, 2, 1, 1, kMapDj, true, t->basis_wt, t->basis_wt)
  LFR_FORM(grad, 12, 8, 3, 3, kMapCov, false, t->basis_w1, t->diff_basis_w0)
  LFR_FORM(curl, 6, 12, 3, 3, kMapContra, false, t->basis_w2, t->diff_basis_w1)
  LFR_FORM(div, 1, 6, 1, 1, kMapRef, false, t->basis_w3, t->diff_basis_w2)
  LFR_FORM(broken_div, 1, 6, 1, 1, kMapRef, false, t->basis_w3, t->diff_basis_w2b)
#undef LFR_FORM
  return LFR_OK;
}
