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
// Mass matrices from basis tables given by the caller (as the PSy layer hands
// them to the reference kernels): scalar spaces
// (compute_mass_matrix_w_scalar_code, sci_compute_mass_matrix_kernel_w_scalar_mod.F90:150-174)
//   mm(ik, df, df2) = sum (((w_h w_v) phi_df) phi_df2) detJ
// and W2 with the Piola map (compute_mass_matrix_w2_code,
// sci_compute_mass_matrix_kernel_w2_mod.F90:131-162)
//   mm(ik, df, df2) = sum ((w_h w_v) (J v_df).(J v_df2)) / detJ,
// df >= df2 computed and mirrored, qp_v outer and qp_h inner.  The chi
// space is lowest-order W0 (its reference gradients in the table).
// ---------------------------------------------------------------------------
namespace {

template <int NDF, int NC>
struct MassQp {
  double b[NC][NDF];   // basis(c, df) at the point
  double g[3][8];      // chi reference gradients
  double wh, wv;
};

template <int NDF, int NC>
__global__ void __launch_bounds__(256)
k_mass_assemble(double* __restrict__ mass, const double* __restrict__ chi1, const double* __restrict__ chi2,
                const double* __restrict__ chi3, const int32_t* __restrict__ map0,
                const MassQp<NDF, NC>* __restrict__ tab, int nqp, int ncell, int nl) {
  const int64_t n3 = (int64_t)ncell * nl;
  const int64_t t = (int64_t)blockIdx.x * 256 + threadIdx.x;
  if (t >= NDF * n3) return;
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
  double m[NDF];
#pragma unroll
  for (int b = 0; b < NDF; ++b) m[b] = 0.0;
  for (int q = 0; q < nqp; ++q) {
    const MassQp<NDF, NC>& T = tab[q];
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
    const double wq = T.wh * T.wv;
    if (NC == 1) {
#pragma unroll
      for (int b = 0; b < NDF; ++b) {
        const int df = a > b ? a : b, df2 = a > b ? b : a;
        const double v = ((wq * T.b[0][df]) * T.b[0][df2]) * dj;
        m[b] = m[b] + v;
      }
    } else {
      // J v for every local dof, components in matmul order
      double jv[NDF][3];
#pragma unroll
      for (int df = 0; df < NDF; ++df)
#pragma unroll
        for (int i = 0; i < 3; ++i) {
          double s = 0.0;
#pragma unroll
          for (int d = 0; d < 3; ++d) s = s + j[i][d] * T.b[d][df];
          jv[df][i] = s;
        }
#pragma unroll
      for (int b = 0; b < NDF; ++b) {
        const int df = a > b ? a : b, df2 = a > b ? b : a;
        double dot = 0.0;
#pragma unroll
        for (int i = 0; i < 3; ++i) dot = dot + jv[df][i] * jv[df2][i];
        const double v = (wq * dot) / dj;
        m[b] = m[b] + v;
      }
    }
  }
#pragma unroll
  for (int b = 0; b < NDF; ++b) mass[ik + n3 * (a + NDF * b)] = m[b];
}

template <int NDF, int NC>
int mass_assemble(lfr_ctx* ctx, int nqp_h, int nqp_v, const double* basis, const double* dchi, const double* wh,
                  const double* wv, const void* chi1, const void* chi2, const void* chi3, const lfr_space* sp,
                  void* mass) {
  std::vector<MassQp<NDF, NC>> tab;
  for (int qv = 0; qv < nqp_v; ++qv)
    for (int qh = 0; qh < nqp_h; ++qh) {
      MassQp<NDF, NC> T;
      for (int c = 0; c < NC; ++c)
        for (int df = 0; df < NDF; ++df) T.b[c][df] = basis[c + NC * (df + NDF * (qh + (size_t)nqp_h * qv))];
      for (int dir = 0; dir < 3; ++dir)
        for (int df = 0; df < 8; ++df) T.g[dir][df] = dchi[dir + 3 * (df + 8 * (qh + (size_t)nqp_h * qv))];
      T.wh = wh[qh];
      T.wv = wv[qv];
      tab.push_back(T);
    }
  const size_t tb = tab.size() * sizeof(MassQp<NDF, NC>);
  int rc = lfr_work_reserve(ctx, tb);
  if (rc) return rc;
  LFR_HIP(ctx, hipMemcpyAsync(ctx->work, tab.data(), tb, hipMemcpyHostToDevice, ctx->stream));
  const int64_t n = (int64_t)NDF * sp->ncell * sp->nlayers;
  hipLaunchKernelGGL((k_mass_assemble<NDF, NC>), dim3((unsigned)((n + 255) / 256)), dim3(256), 0, ctx->stream,
                     (double*)mass, (const double*)chi1, (const double*)chi2, (const double*)chi3, sp->map0_d,
                     (const MassQp<NDF, NC>*)ctx->work, (int)tab.size(), sp->ncell, sp->nlayers);
  LFR_CHECK_LAUNCH(ctx);
  LFR_HIP(ctx, hipStreamSynchronize(ctx->stream));   // the table is this call's host memory
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
  LFR_REQUIRE(ctx, chi_space->ndf == 8, "lfr_compute_mass_matrix: the chi space must be lowest-order W0");
  if (kind == 0) {
    switch (ndf) {
      case 1: return mass_assemble<1, 1>(ctx, nqp_h, nqp_v, basis, diff_basis_chi, wqp_h, wqp_v, chi1_d, chi2_d,
                                         chi3_d, chi_space, mass_d);
      case 2: return mass_assemble<2, 1>(ctx, nqp_h, nqp_v, basis, diff_basis_chi, wqp_h, wqp_v, chi1_d, chi2_d,
                                         chi3_d, chi_space, mass_d);
      case 8: return mass_assemble<8, 1>(ctx, nqp_h, nqp_v, basis, diff_basis_chi, wqp_h, wqp_v, chi1_d, chi2_d,
                                         chi3_d, chi_space, mass_d);
      default: break;
    }
  } else {
    switch (ndf) {
      case 2: return mass_assemble<2, 3>(ctx, nqp_h, nqp_v, basis, diff_basis_chi, wqp_h, wqp_v, chi1_d, chi2_d,
                                         chi3_d, chi_space, mass_d);
      case 4: return mass_assemble<4, 3>(ctx, nqp_h, nqp_v, basis, diff_basis_chi, wqp_h, wqp_v, chi1_d, chi2_d,
                                         chi3_d, chi_space, mass_d);
      case 6: return mass_assemble<6, 3>(ctx, nqp_h, nqp_v, basis, diff_basis_chi, wqp_h, wqp_v, chi1_d, chi2_d,
                                         chi3_d, chi_space, mass_d);
      default: break;
    }
  }
  return lfr_set_error(ctx, LFR_ERR_ARG, "lfr_compute_mass_matrix: no kernel for kind %d with %d dofs per cell",
                       kind, ndf);
}
