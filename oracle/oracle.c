/*
 * CPU oracle for the LFRic Core operator-apply / column-solve path.
 *
 * TEST INFRASTRUCTURE ONLY.  This file restates the reference's Fortran
 * kernels and PSy-layer loop shapes in plain C so that tests, the smoke check
 * and the CPU baseline leg of the benchmark can check the HIP library
 * against them.  Nothing in the shipped library calls it.
 *
 * Conventions kept from the reference so that arrays pass through unchanged:
 *   - dofmaps are map(ndf, ncell) column-major (1-based dof numbers of the
 *     bottom cell of each column; layer k adds k);
 *   - LMA matrices are local_stencil(ncell_3d, ndf_to, ndf_from) column-major,
 *     element (ik, i, j) at ik + ncell_3d*(i + ndf_to*j), ik = cell*nl + k;
 *   - columnwise matrices are columnwise_matrix(bw, nrow, ncell_2d).
 *   - Cell numbers passed to the per-cell kernels are 1-based like the
 *     Fortran `cell` argument.
 *
 * Floating-point: every accumulation is written as separate multiply and add
 * in the reference's statement order; build with -ffp-contract=off so the
 * compiler does not fuse them.
 *
 * Parity pins (see tests/): the per-cell kernels are compared bit-for-bit
 * against the reference Fortran compiled by flang (oracle/_ref) and against
 * the numbers in the reference's pFUnit tests; the solvers are pinned by
 * mathematical invariants (exact convergence on small SPD systems,
 * A-orthogonality, agreement with a dense direct solve).
 */
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

#define IDX3(ik, i, j, n3, nto) ((int64_t)(ik) + (int64_t)(n3) * ((int64_t)(i) + (int64_t)(nto) * (int64_t)(j)))

/* ------------------------------------------------------------------------ */
/* Per-column kernels                                                        */
/* ------------------------------------------------------------------------ */

/* matrix_vector_code: lhs += M x on one column
 * (components/science/source/kernel/algebra/matrix_vector_kernel_mod.F90:130-138). */
#define DEF_MATVEC(T, SUF)                                                              \
  void orc_matrix_vector_code_##SUF(int cell, int nlayers, T* lhs, const T* x,        \
                                    int ncell_3d, const T* matrix, int ndf1,           \
                                    const int* map1, int ndf2, const int* map2) {      \
    int64_t ik = (int64_t)(cell - 1) * nlayers;                                        \
    for (int df2 = 0; df2 < ndf2; ++df2) {                                             \
      int64_t i2 = map2[df2] - 1;                                                      \
      for (int df = 0; df < ndf1; ++df) {                                              \
        int64_t i1 = map1[df] - 1;                                                     \
        const T* m = matrix + IDX3(ik, df, df2, ncell_3d, ndf1);                       \
        /* Fortran array statement: the whole right-hand side is formed before    \
           the assignment, which matters when map1(df) columns overlap. */          \
        T tmp[512];                                                                    \
        for (int k = 0; k < nlayers; ++k) {                                            \
          T prod = m[k] * x[i2 + k];                                                   \
          tmp[k] = lhs[i1 + k] + prod;                                                 \
        }                                                                              \
        for (int k = 0; k < nlayers; ++k) lhs[i1 + k] = tmp[k];                        \
      }                                                                                \
    }                                                                                  \
  }
DEF_MATVEC(float, r32)
DEF_MATVEC(double, r64)

/* dg_matrix_vector_code: lhs = M x (zero, then accumulate)
 * (components/science/source/kernel/algebra/dg_matrix_vector_kernel_mod.F90:94-107). */
#define DEF_DGMATVEC(T, SUF)                                                            \
  void orc_dg_matrix_vector_code_##SUF(int cell, int nlayers, T* lhs, const T* x,     \
                                       int ncell_3d, const T* matrix, int ndf1,        \
                                       const int* map1, int ndf2, const int* map2) {   \
    for (int df1 = 0; df1 < ndf1; ++df1)                                               \
      for (int k = 0; k < nlayers; ++k) lhs[map1[df1] - 1 + k] = (T)0;                 \
    orc_matrix_vector_code_##SUF(cell, nlayers, lhs, x, ncell_3d, matrix, ndf1, map1, \
                                 ndf2, map2);                                          \
  }
DEF_DGMATVEC(float, r32)
DEF_DGMATVEC(double, r64)

/* transpose_matrix_vector_code: lhs += M^T x, M stored (ncell_3d, ndf2, ndf1)
 * (components/science/source/kernel/algebra/transpose_matrix_vector_kernel_mod.F90:92-101):
 * per layer, x_e gathered, lhs_e = matmul(transpose(M(ik,:,:)), x_e) summed
 * over df2 in ascending order, then lhs(map1(df)+k) += lhs_e(df). */
#define DEF_TMATVEC(T, SUF)                                                             \
  void orc_transpose_matrix_vector_code_##SUF(int cell, int nlayers, T* lhs,          \
                                              const T* x, int ncell_3d,                \
                                              const T* matrix, int ndf1,               \
                                              const int* map1, int ndf2,               \
                                              const int* map2) {                       \
    T x_e[512], lhs_e[512];                                                            \
    for (int k = 0; k < nlayers; ++k) {                                                \
      for (int df = 0; df < ndf2; ++df) x_e[df] = x[map2[df] - 1 + k];                 \
      int64_t ik = (int64_t)(cell - 1) * nlayers + k;                                  \
      for (int df = 0; df < ndf1; ++df) {                                              \
        T acc = (T)0;                                                                  \
        for (int df2 = 0; df2 < ndf2; ++df2) {                                         \
          T prod = matrix[IDX3(ik, df2, df, ncell_3d, ndf2)] * x_e[df2];               \
          acc = acc + prod;                                                            \
        }                                                                              \
        lhs_e[df] = acc;                                                               \
      }                                                                                \
      for (int df = 0; df < ndf1; ++df) lhs[map1[df] - 1 + k] += lhs_e[df];            \
    }                                                                                  \
  }
DEF_TMATVEC(float, r32)
DEF_TMATVEC(double, r64)

/* columnwise_op_appinv_kernel_code: Thomas solve of one tridiagonal column,
 * overwriting lhs (components/science/source/kernel/columnwise_operators/
 * sci_columnwise_op_appinv_kernel_mod.F90:116-143).  Bands (1,2,3) = (sub,
 * diag, super); one reciprocal of the pivot per row. */
#define DEF_APPINV(T, SUF)                                                              \
  void orc_columnwise_op_appinv_code_##SUF(int cell, T* lhs, const T* x,              \
                                           const T* cm, int nrow, int bw,              \
                                           const int* map, const int* ind) {           \
    T cp[4096], dp[4096];                                                              \
    const T* c = cm + (int64_t)(cell - 1) * bw * nrow;                                 \
    for (int i = 0; i < nrow; ++i) {                                                   \
      int64_t mu = (int64_t)map[0] + ind[i] - 1 - 1;                                   \
      const T* r = c + (int64_t)bw * i;                                                \
      if (i == 0) {                                                                    \
        T inv = (T)1 / r[1];                                                           \
        cp[i] = inv * r[2];                                                            \
        dp[i] = inv * x[mu];                                                           \
      } else {                                                                         \
        T den = r[1] - r[0] * cp[i - 1];                                               \
        T inv = (T)1 / den;                                                            \
        if (i < nrow - 1) cp[i] = inv * r[2];                                          \
        T num = x[mu] - r[0] * dp[i - 1];                                              \
        dp[i] = inv * num;                                                             \
      }                                                                                \
    }                                                                                  \
    for (int i = nrow - 1; i >= 0; --i) {                                              \
      int64_t mu = (int64_t)map[0] + ind[i] - 1 - 1;                                   \
      if (i < nrow - 1) {                                                              \
        T prod = cp[i] * dp[i + 1];                                                    \
        dp[i] = dp[i] - prod;                                                          \
      }                                                                                \
      lhs[mu] = dp[i];                                                                 \
    }                                                                                  \
  }
DEF_APPINV(float, r32)
DEF_APPINV(double, r64)

/* columnwise_op_app_kernel_code: lhs += A x for a banded column operator
 * (components/science/source/kernel/columnwise_operators/
 * sci_columnwise_op_app_kernel_mod.F90:116-125).  The band limits are formed
 * in the operator's real kind, as the reference does:
 * j_minus = ceiling((alpha*i-gamma_p)/beta), j_plus = floor((alpha*i+gamma_m)/beta). */
#define DEF_CMAAPP(T, SUF)                                                              \
  void orc_columnwise_op_app_code_##SUF(int cell, T* lhs, const T* x, const T* cm,    \
                                        int nrow, int ncol, int bw, int alpha,         \
                                        int beta, int gamma_m, int gamma_p,            \
                                        const int* map1, const int* ind_to,            \
                                        const int* map2, const int* ind_from) {        \
    const T* c = cm + (int64_t)(cell - 1) * bw * nrow;                                 \
    for (int i = 1; i <= nrow; ++i) {                                                  \
      int64_t mu_i = (int64_t)map1[0] + ind_to[i - 1] - 1 - 1;                         \
      int j_minus = (int)ceil((T)(alpha * i - gamma_p) / ((T)1 * beta));               \
      int j_plus = (int)floor((T)(alpha * i + gamma_m) / ((T)1 * beta));               \
      int lo = j_minus > 1 ? j_minus : 1;                                              \
      int hi = j_plus < ncol ? j_plus : ncol;                                          \
      for (int j = lo; j <= hi; ++j) {                                                 \
        int64_t mu_j = (int64_t)map2[0] + ind_from[j - 1] - 1 - 1;                     \
        T prod = c[(j - j_minus) + (int64_t)bw * (i - 1)] * x[mu_j];                   \
        lhs[mu_i] = lhs[mu_i] + prod;                                                  \
      }                                                                                \
    }                                                                                  \
  }
DEF_CMAAPP(float, r32)
DEF_CMAAPP(double, r64)

/* columnwise_op_asm_kernel_code: assemble one column's banded matrix from
 * the LMA local stencil (components/science/source/kernel/columnwise_operators/
 * sci_columnwise_op_asm_kernel_mod.F90:110-126).  column_banded_dofmap(ndf, nl). */
#define DEF_CMAASM(T, SUF)                                                              \
  void orc_columnwise_op_asm_code_##SUF(int cell, int nlayers, int ncell_3d,           \
                                        const T* local_stencil, T* cm, int nrow,       \
                                        int bw, int alpha, int beta, int gamma_p,      \
                                        int ndf_to, const int* cbd_to, int ndf_from,   \
                                        const int* cbd_from) {                         \
    T* c = cm + (int64_t)(cell - 1) * bw * nrow;                                       \
    for (int64_t q = 0; q < (int64_t)bw * nrow; ++q) c[q] = (T)0;                      \
    for (int k = 1; k <= nlayers; ++k) {                                               \
      int64_t ik = (int64_t)(cell - 1) * nlayers + k - 1;                              \
      for (int df1 = 0; df1 < ndf_to; ++df1) {                                         \
        int i = cbd_to[df1 + (int64_t)ndf_to * (k - 1)];                               \
        int j_minus = (int)ceil((T)(alpha * i - gamma_p) / ((T)1 * beta));             \
        for (int df2 = 0; df2 < ndf_from; ++df2) {                                     \
          int j = cbd_from[df2 + (int64_t)ndf_from * (k - 1)];                         \
          T* e = c + (j - j_minus) + (int64_t)bw * (i - 1);                            \
          *e = *e + local_stencil[IDX3(ik, df1, df2, ncell_3d, ndf_to)];               \
        }                                                                              \
      }                                                                                \
    }                                                                                  \
  }
DEF_CMAASM(float, r32)
DEF_CMAASM(double, r64)

/* enforce_bc_code (components/science/source/kernel/algebra/
 * sci_enforce_bc_kernel_mod.F90:63-91) over all cells in cell order:
 * field(map(df)) *= bv(df,1); field(map(df)+nl-1) *= bv(df,2); bv(ndf,2). */
#define DEF_ENFORCE_BC(T, SUF)                                                          \
  void orc_invoke_enforce_bc_##SUF(int ncell, int nlayers, T* field, int ndf,          \
                                   const int* dofmap, const int* bv) {                 \
    for (int c = 0; c < ncell; ++c) {                                                   \
      const int* map = dofmap + (int64_t)c * ndf;                                       \
      for (int df = 0; df < ndf; ++df) {                                                \
        int64_t d = (int64_t)map[df] - 1;                                               \
        field[d] = field[d] * (T)bv[df];                                                \
      }                                                                                 \
      for (int df = 0; df < ndf; ++df) {                                                \
        int64_t d = (int64_t)map[df] - 1 + nlayers - 1;                                 \
        field[d] = field[d] * (T)bv[df + ndf];                                          \
      }                                                                                 \
    }                                                                                   \
  }
DEF_ENFORCE_BC(float, r32)
DEF_ENFORCE_BC(double, r64)

/* tri_solve_code (components/science/source/kernel/solver/
 * sci_tri_solve_kernel_mod.f90:95-118) over all columns in cell order: Thomas
 * solve of tri_minus y(k-1) + tri_0 y(k) + tri_plus y(k+1) = x on the nl dofs
 * of one W3 column (map(1) = first dof); one reciprocal per row, multiplied
 * into the row; y must not alias an input. */
#define DEF_TRI_SOLVE(T, SUF)                                                           \
  void orc_invoke_tri_solve_##SUF(int ncell, int nl, T* y, const T* x, const T* t0,    \
                                  const T* tp, const T* tm, int ndf, const int* dofmap) { \
    T tpn[4096], xn[4096];                                                             \
    for (int c = 0; c < ncell; ++c) {                                                   \
      const int64_t ij = (int64_t)dofmap[(int64_t)c * ndf] - 1;                         \
      T denom = (T)1 / t0[ij];                                                          \
      tpn[0] = tp[ij] * denom;                                                          \
      xn[0] = x[ij] * denom;                                                            \
      for (int k = 1; k < nl; ++k) {                                                    \
        const T pm = tm[ij + k] * tpn[k - 1];                                           \
        denom = (T)1 / (t0[ij + k] - pm);                                               \
        tpn[k] = tp[ij + k] * denom;                                                    \
        const T xm = tm[ij + k] * xn[k - 1];                                            \
        xn[k] = (x[ij + k] - xm) * denom;                                               \
      }                                                                                 \
      y[ij + nl - 1] = xn[nl - 1];                                                      \
      for (int k = nl - 2; k >= 0; --k) {                                               \
        const T py = tpn[k] * y[ij + k + 1];                                            \
        y[ij + k] = xn[k] - py;                                                         \
      }                                                                                 \
    }                                                                                   \
  }
DEF_TRI_SOLVE(float, r32)
DEF_TRI_SOLVE(double, r64)

/* operator_tri_solve_code (components/science/source/kernel/solver/
 * sci_operator_tri_solve_kernel_mod.f90:65-120) over all columns in cell
 * order: the tridiagonal of the nl+1 dofs of one Wtheta/W2v column taken
 * from the 2x2 local matrices matrix(ncell_3d, 2, 2) (the diagonal of a
 * shared dof is matrix(ik,1,1) + matrix(ik-1,2,2)), then the same Thomas
 * solve. */
#define DEF_OP_TRI_SOLVE(T, SUF)                                                        \
  void orc_invoke_operator_tri_solve_##SUF(int ncell, int nl, T* y, const T* x,        \
                                           const T* m, int ndf, const int* dofmap) {   \
    T t0[4097], tp[4097], tm[4097], tpn[4097], xn[4097];                                \
    const int64_t n3 = (int64_t)ncell * nl;                                             \
    for (int c = 0; c < ncell; ++c) {                                                   \
      const int64_t ij = (int64_t)dofmap[(int64_t)c * ndf] - 1;                         \
      t0[0] = (T)0;                                                                     \
      tm[0] = (T)0;                                                                     \
      for (int k = 0; k < nl; ++k) {                                                    \
        const int64_t ik = (int64_t)c * nl + k;                                         \
        t0[k] = m[ik] + t0[k];                /* (ik,1,1) */                             \
        tp[k] = m[ik + n3 * 2];               /* (ik,1,2) */                             \
        tm[k + 1] = m[ik + n3 * 1];           /* (ik,2,1) */                             \
        t0[k + 1] = m[ik + n3 * 3];           /* (ik,2,2) */                             \
      }                                                                                 \
      tp[nl] = (T)0;                                                                    \
      T denom = (T)1 / t0[0];                                                           \
      tpn[0] = tp[0] * denom;                                                           \
      xn[0] = x[ij] * denom;                                                            \
      for (int k = 1; k <= nl; ++k) {                                                   \
        const T pm = tm[k] * tpn[k - 1];                                                \
        denom = (T)1 / (t0[k] - pm);                                                    \
        tpn[k] = tp[k] * denom;                                                         \
        const T xm = tm[k] * xn[k - 1];                                                 \
        xn[k] = (x[ij + k] - xm) * denom;                                               \
      }                                                                                 \
      y[ij + nl] = xn[nl];                                                              \
      for (int k = nl - 1; k >= 0; --k) {                                               \
        const T py = tpn[k] * y[ij + k + 1];                                            \
        y[ij + k] = xn[k] - py;                                                         \
      }                                                                                 \
    }                                                                                   \
  }
DEF_OP_TRI_SOLVE(float, r32)
DEF_OP_TRI_SOLVE(double, r64)

/* ------------------------------------------------------------------------ */
/* Invoke-level loops (what the generated PSy layer does around a kernel)   */
/* ------------------------------------------------------------------------ */

/* Coloured loop for a continuous (GH_INC) output: colour by colour, cells of
 * one colour in parallel (infrastructure/build/psyclone/psyclone_tools.py:71-119).
 * cells_in_colour(ncolour, maxc) column-major, 1-based.  mode: 0 forward
 * matrix_vector, 1 transpose_matrix_vector. */
#define DEF_INVOKE_COLOURED(T, SUF)                                                     \
  void orc_invoke_coloured_##SUF(int mode, int ncolour, const int* ncells_per_colour, \
                                 const int* cells_in_colour, int maxc, int nlayers,    \
                                 T* lhs, const T* x, int ncell_3d, const T* matrix,    \
                                 int ndf1, const int* dofmap1, int ndf2,               \
                                 const int* dofmap2, int nthreads) {                   \
    for (int col = 0; col < ncolour; ++col) {                                          \
      int nc = ncells_per_colour[col];                                                 \
      _Pragma("omp parallel for schedule(static) num_threads(nthreads > 0 ? nthreads : 1)") \
      for (int idx = 0; idx < nc; ++idx) {                                             \
        int cell = cells_in_colour[col + (int64_t)ncolour * idx];                      \
        const int* m1 = dofmap1 + (int64_t)(cell - 1) * ndf1;                          \
        const int* m2 = dofmap2 + (int64_t)(cell - 1) * ndf2;                          \
        if (mode == 0)                                                                 \
          orc_matrix_vector_code_##SUF(cell, nlayers, lhs, x, ncell_3d, matrix, ndf1,  \
                                       m1, ndf2, m2);                                  \
        else                                                                           \
          orc_transpose_matrix_vector_code_##SUF(cell, nlayers, lhs, x, ncell_3d,      \
                                                 matrix, ndf1, m1, ndf2, m2);          \
      }                                                                                \
    }                                                                                  \
  }
DEF_INVOKE_COLOURED(float, r32)
DEF_INVOKE_COLOURED(double, r64)

/* Plain cell loop for a discontinuous output (psyclone_tools.py:89-95).
 * mode: 0 dg_matrix_vector (overwrite), 1 dg_inc_matrix_vector (increment). */
#define DEF_INVOKE_DG(T, SUF)                                                           \
  void orc_invoke_dg_##SUF(int mode, int ncell, int nlayers, T* lhs, const T* x,      \
                           int ncell_3d, const T* matrix, int ndf1,                    \
                           const int* dofmap1, int ndf2, const int* dofmap2,           \
                           int nthreads) {                                             \
    _Pragma("omp parallel for schedule(static) num_threads(nthreads > 0 ? nthreads : 1)") \
    for (int c = 0; c < ncell; ++c) {                                                  \
      const int* m1 = dofmap1 + (int64_t)c * ndf1;                                     \
      const int* m2 = dofmap2 + (int64_t)c * ndf2;                                     \
      if (mode == 0)                                                                   \
        orc_dg_matrix_vector_code_##SUF(c + 1, nlayers, lhs, x, ncell_3d, matrix,      \
                                        ndf1, m1, ndf2, m2);                           \
      else                                                                             \
        orc_matrix_vector_code_##SUF(c + 1, nlayers, lhs, x, ncell_3d, matrix, ndf1,   \
                                     m1, ndf2, m2);                                    \
    }                                                                                  \
  }
DEF_INVOKE_DG(float, r32)
DEF_INVOKE_DG(double, r64)

/* CMA inverse over all columns (no colouring: the input space is
 * horizontally discontinuous, columnwise_operator_mod.f90:427-469). */
#define DEF_INVOKE_APPINV(T, SUF)                                                       \
  void orc_invoke_appinv_##SUF(int ncell_2d, T* lhs, const T* x, const T* cm,          \
                               int nrow, int bw, int ndf, const int* dofmap,           \
                               const int* ind, int nthreads) {                         \
    _Pragma("omp parallel for schedule(static) num_threads(nthreads > 0 ? nthreads : 1)") \
    for (int c = 0; c < ncell_2d; ++c)                                                 \
      orc_columnwise_op_appinv_code_##SUF(c + 1, lhs, x, cm, nrow, bw,                 \
                                          dofmap + (int64_t)c * ndf, ind);             \
  }
DEF_INVOKE_APPINV(float, r32)
DEF_INVOKE_APPINV(double, r64)

/* ------------------------------------------------------------------------ */
/* Vector algebra (PSyclone built-ins / abstract_vector_type methods)        */
/* ------------------------------------------------------------------------ */

/* X_innerproduct_Y with FP64 accumulation of real(x,r_double)*real(y,r_double):
 * static per-thread partials summed in thread order
 * (components/science/source/psy/sci_psykal_light_mod.f90:146-214). */
double orc_dot_r32(int64_t n, const float* x, const float* y, int nthreads) {
  if (nthreads < 1) nthreads = 1;
  double part[256];
  if (nthreads > 256) nthreads = 256;
  for (int t = 0; t < nthreads; ++t) part[t] = 0.0;
#pragma omp parallel num_threads(nthreads)
  {
    int t = 0, nt = 1;
#ifdef _OPENMP
    t = omp_get_thread_num();
    nt = omp_get_num_threads();
#endif
    int64_t chunk = (n + nt - 1) / nt, lo = t * chunk, hi = lo + chunk < n ? lo + chunk : n;
    double s = 0.0;
    for (int64_t i = lo; i < hi; ++i) s = s + (double)x[i] * (double)y[i];
    part[t] = s;
  }
  double s = 0.0;
  for (int t = 0; t < nthreads; ++t) s += part[t];
  return s;
}

double orc_dot_r64(int64_t n, const double* x, const double* y, int nthreads) {
  if (nthreads < 1) nthreads = 1;
  double part[256];
  if (nthreads > 256) nthreads = 256;
  for (int t = 0; t < nthreads; ++t) part[t] = 0.0;
#pragma omp parallel num_threads(nthreads)
  {
    int t = 0, nt = 1;
#ifdef _OPENMP
    t = omp_get_thread_num();
    nt = omp_get_num_threads();
#endif
    int64_t chunk = (n + nt - 1) / nt, lo = t * chunk, hi = lo + chunk < n ? lo + chunk : n;
    double s = 0.0;
    for (int64_t i = lo; i < hi; ++i) s = s + x[i] * y[i];
    part[t] = s;
  }
  double s = 0.0;
  for (int t = 0; t < nthreads; ++t) s += part[t];
  return s;
}

int orc_omp_max_threads(void) {
#ifdef _OPENMP
  return omp_get_max_threads();
#else
  return 1;
#endif
}

/* Intergrid transfers of lowest-order scalar fields (W3: ndf 1, levels
 * 0..nl-1; Wtheta: ndf 2, levels 0..nl), one coarse cell at a time:
 *   restrict_scalar_unweighted_code
 *   (components/science/source/kernel/inter_mesh/sci_restrict_scalar_unweighted_kernel_mod.F90:71-125):
 *     coarse(k) = sum over y outer, x inner of fine(child(x, y), k) * denom,
 *     denom = 1 / (nfx nfy), from zero;
 *   prolong_scalar_unweighted_code
 *   (.../sci_prolong_scalar_unweighted_kernel_mod.F90): fine(child, k) += coarse(k).
 * cell_map(nfx, nfy, ncell_coarse) holds 1-based fine cells. */
#define DEF_TRANSFER(T, SUF)                                                                     \
  void orc_restrict_scalar_##SUF(int ncell_coarse, int nl, int ndf, T* coarse, const T* fine,   \
                                 const int* map_coarse, const int* map_fine, const int* cell_map, \
                                 int nfx, int nfy) {                                              \
    const T denom = (T)1 / (T)(nfx * nfy);                                                       \
    const int top = nl - 2 + ndf;                                                                \
    for (int c = 0; c < ncell_coarse; ++c) {                                                     \
      const int* cm = cell_map + (int64_t)c * nfx * nfy;                                         \
      const int64_t cb = (int64_t)map_coarse[(int64_t)c * ndf] - 1;                              \
      for (int k = 0; k <= top; ++k) {                                                           \
        T v = (T)0;                                                                              \
        for (int y = 0; y < nfy; ++y)                                                            \
          for (int x = 0; x < nfx; ++x) {                                                        \
            const int f = cm[x + nfx * y] - 1;                                                   \
            const T p = fine[(int64_t)map_fine[(int64_t)f * ndf] - 1 + k] * denom;              \
            v = v + p;                                                                           \
          }                                                                                      \
        coarse[cb + k] = v;                                                                      \
      }                                                                                          \
    }                                                                                            \
  }                                                                                              \
  void orc_prolong_scalar_##SUF(int ncell_coarse, int nl, int ndf, T* fine, const T* coarse,    \
                                const int* map_coarse, const int* map_fine, const int* cell_map, \
                                int nfx, int nfy) {                                               \
    const int top = nl - 2 + ndf;                                                                \
    for (int c = 0; c < ncell_coarse; ++c) {                                                     \
      const int* cm = cell_map + (int64_t)c * nfx * nfy;                                         \
      const int64_t cb = (int64_t)map_coarse[(int64_t)c * ndf] - 1;                              \
      for (int y = 0; y < nfy; ++y)                                                              \
        for (int x = 0; x < nfx; ++x) {                                                          \
          const int f = cm[x + nfx * y] - 1;                                                     \
          const int64_t fb = (int64_t)map_fine[(int64_t)f * ndf] - 1;                            \
          for (int k = 0; k <= top; ++k) fine[fb + k] = fine[fb + k] + coarse[cb + k];          \
        }                                                                                        \
    }                                                                                            \
  }
DEF_TRANSFER(float, r32)
DEF_TRANSFER(double, r64)

/* face_selector_code (components/science/source/kernel/algebra/sci_face_selector_kernel_mod.F90)
 * over the cells in order, selectors starting at 1: a cell takes its E (N)
 * face, selector 2, unless an earlier cell took that face as its E or N
 * face.  dofmap is the lowest-order W2 map (ncell, 6), local dofs
 * W, S, E, N, B, T; the side-face dofs identify the faces. */
void orc_face_selector(int ncell, const int* dofmap, int undf, int* sel_ew, int* sel_ns) {
  char* taken = (char*)calloc((size_t)undf + 1, 1);
  for (int c = 0; c < ncell; ++c) {
    const int* m = dofmap + (int64_t)c * 6;
    sel_ew[c] = 1;
    sel_ns[c] = 1;
    if (!taken[m[2]]) { sel_ew[c] = 2; taken[m[2]] = 1; }
    if (!taken[m[3]]) { sel_ns[c] = 2; taken[m[3]] = 1; }
  }
  free(taken);
}

/* restrict_w2_code (components/science/source/kernel/inter_mesh/sci_restrict_w2_kernel_mod.F90)
 * over the coarse cells in order: faces 1 .. sel_ew + sel_ns (W, S, then E
 * or N; face 3 is N when only N is selected) take the sum, y outer and x
 * inner, of the fine faces on them (nl levels), then the B face takes the
 * sum over every child (nl + 1 levels: every horizontal face of the
 * column).  A face written by two cells keeps the later cell's value. */
#define DEF_RESTRICT_W2(T, SUF)                                                                   \
  void orc_restrict_w2_##SUF(int ncell_coarse, int nl, T* coarse, const T* fine, const int* map_c,  \
                             const int* map_f, const int* cell_map, int nfx, int nfy,              \
                             const int* sel_ew, const int* sel_ns) {                               \
    enum { W = 1, S = 2, E = 3, N = 4, B = 5 };                                                    \
    for (int c = 0; c < ncell_coarse; ++c) {                                                       \
      const int* cm = cell_map + (int64_t)c * nfx * nfy;                                           \
      const int* mc = map_c + (int64_t)c * 6;                                                      \
      const int nface = sel_ew[c] + sel_ns[c];                                                     \
      for (int face = 1; face <= nface; ++face) {                                                  \
        int df = face;                                                                             \
        if (face == 3 && sel_ns[c] == 2 && sel_ew[c] == 1) df = N;                                 \
        int x0 = 1, x1 = nfx, y0 = 1, y1 = nfy;                                                    \
        if (df == N) { y1 = 1; }                                                                   \
        else if (df == S) { y0 = nfy; }                                                            \
        else if (df == W) { x1 = 1; }                                                              \
        else if (df == E) { x0 = nfx; }                                                            \
        for (int k = 0; k < nl; ++k) {                                                             \
          T v = (T)0;                                                                              \
          for (int y = y0; y <= y1; ++y)                                                           \
            for (int x = x0; x <= x1; ++x) {                                                       \
              const int f = cm[(x - 1) + nfx * (y - 1)] - 1;                                       \
              v = v + fine[(int64_t)map_f[(int64_t)f * 6 + df - 1] - 1 + k];                      \
            }                                                                                      \
          coarse[(int64_t)mc[df - 1] - 1 + k] = v;                                                 \
        }                                                                                          \
      }                                                                                            \
      for (int k = 0; k <= nl; ++k) {                                                              \
        T v = (T)0;                                                                                \
        for (int y = 1; y <= nfy; ++y)                                                             \
          for (int x = 1; x <= nfx; ++x) {                                                         \
            const int f = cm[(x - 1) + nfx * (y - 1)] - 1;                                         \
            v = v + fine[(int64_t)map_f[(int64_t)f * 6 + B - 1] - 1 + k];                         \
          }                                                                                        \
        coarse[(int64_t)mc[B - 1] - 1 + k] = v;                                                    \
      }                                                                                            \
    }                                                                                              \
  }
DEF_RESTRICT_W2(float, r32)
DEF_RESTRICT_W2(double, r64)

/* weights_prolong_w2_kernel_code
 * (components/science/source/kernel/inter_mesh/sci_weights_prolong_w2_kernel_mod.F90)
 * and prolong_w2_kernel_code (.../sci_prolong_w2_kernel_mod.F90) over the
 * coarse cells in order (a fine face written by two coarse cells keeps the
 * later value).  dofmaps are the lowest-order W2 maps (ncell, 6). */
#define DEF_PROLONG_W2(T, SUF)                                                                    \
  void orc_weights_prolong_w2_##SUF(int ncell_coarse, int nl, T* w, const int* map_f,             \
                                    const int* cell_map, int nfx, int nfy) {                      \
    enum { W = 0, S = 1, E = 2, N = 3, B = 4 };                                                   \
    for (int c = 0; c < ncell_coarse; ++c) {                                                      \
      const int* cm = cell_map + (int64_t)c * nfx * nfy;                                          \
      for (int k = 0; k < nl; ++k) {                                                              \
        for (int x = 1; x <= nfx; ++x) {                                                          \
          int f = cm[(x - 1) + nfx * 0] - 1;                                                      \
          w[(int64_t)map_f[(int64_t)f * 6 + N] - 1 + k] = (T)1 / (T)nfx;                          \
          f = cm[(x - 1) + nfx * (nfy - 1)] - 1;                                                  \
          w[(int64_t)map_f[(int64_t)f * 6 + S] - 1 + k] = (T)1 / (T)nfx;                          \
        }                                                                                         \
        for (int y = 1; y <= nfy; ++y) {                                                          \
          int f = cm[0 + nfx * (y - 1)] - 1;                                                      \
          w[(int64_t)map_f[(int64_t)f * 6 + W] - 1 + k] = (T)1 / (T)nfy;                          \
          f = cm[(nfx - 1) + nfx * (y - 1)] - 1;                                                  \
          w[(int64_t)map_f[(int64_t)f * 6 + E] - 1 + k] = (T)1 / (T)nfy;                          \
        }                                                                                         \
        for (int y = 1; y <= nfy; ++y)                                                            \
          for (int x = 2; x <= nfx; ++x) {                                                        \
            const int f = cm[(x - 1) + nfx * (y - 1)] - 1;                                        \
            w[(int64_t)map_f[(int64_t)f * 6 + W] - 1 + k] = (T)(nfx + 1 - x) / (T)nfx;            \
          }                                                                                       \
        for (int x = 1; x <= nfx; ++x)                                                            \
          for (int y = 2; y <= nfy; ++y) {                                                        \
            const int f = cm[(x - 1) + nfx * (y - 1)] - 1;                                        \
            w[(int64_t)map_f[(int64_t)f * 6 + N] - 1 + k] = (T)(nfy + 1 - y) / (T)nfy;            \
          }                                                                                       \
      }                                                                                           \
      for (int k = 0; k <= nl; ++k)                                                               \
        for (int y = 1; y <= nfy; ++y)                                                            \
          for (int x = 1; x <= nfx; ++x) {                                                        \
            const int f = cm[(x - 1) + nfx * (y - 1)] - 1;                                        \
            w[(int64_t)map_f[(int64_t)f * 6 + B] - 1 + k] = (T)1 / (T)(nfx * nfy);                \
          }                                                                                       \
    }                                                                                             \
  }                                                                                               \
  void orc_prolong_w2_##SUF(int ncell_coarse, int nl, T* fine, const T* coarse, const T* w,       \
                            const int* map_c, const int* map_f, const int* cell_map, int nfx,     \
                            int nfy) {                                                            \
    enum { W = 0, S = 1, E = 2, N = 3, B = 4 };                                                   \
    for (int c = 0; c < ncell_coarse; ++c) {                                                      \
      const int* cm = cell_map + (int64_t)c * nfx * nfy;                                          \
      const int* mc = map_c + (int64_t)c * 6;                                                     \
      for (int k = 0; k < nl; ++k) {                                                              \
        for (int x = 1; x <= nfx; ++x) {                                                          \
          int f = cm[(x - 1)] - 1;                                                                \
          int64_t d = (int64_t)map_f[(int64_t)f * 6 + N] - 1 + k;                                 \
          fine[d] = w[d] * coarse[(int64_t)mc[N] - 1 + k];                                        \
          f = cm[(x - 1) + nfx * (nfy - 1)] - 1;                                                  \
          d = (int64_t)map_f[(int64_t)f * 6 + S] - 1 + k;                                         \
          fine[d] = w[d] * coarse[(int64_t)mc[S] - 1 + k];                                        \
        }                                                                                         \
        for (int y = 1; y <= nfy; ++y) {                                                          \
          int f = cm[nfx * (y - 1)] - 1;                                                          \
          int64_t d = (int64_t)map_f[(int64_t)f * 6 + W] - 1 + k;                                 \
          fine[d] = w[d] * coarse[(int64_t)mc[W] - 1 + k];                                        \
          f = cm[(nfx - 1) + nfx * (y - 1)] - 1;                                                  \
          d = (int64_t)map_f[(int64_t)f * 6 + E] - 1 + k;                                         \
          fine[d] = w[d] * coarse[(int64_t)mc[E] - 1 + k];                                        \
        }                                                                                         \
        for (int y = 1; y <= nfy; ++y) {                                                          \
          const T ew = w[(int64_t)map_f[(int64_t)(cm[nfx * (y - 1)] - 1) * 6 + W] - 1 + k];       \
          for (int x = 2; x <= nfx; ++x) {                                                        \
            const int64_t d = (int64_t)map_f[(int64_t)(cm[(x - 1) + nfx * (y - 1)] - 1) * 6 + W] - 1 + k; \
            const T iw = w[d];                                                                    \
            const T a = iw * coarse[(int64_t)mc[W] - 1 + k];                                      \
            const T b = ((T)1 - iw) * coarse[(int64_t)mc[E] - 1 + k];                             \
            fine[d] = ew * (a + b);                                                               \
          }                                                                                       \
        }                                                                                         \
        for (int x = 1; x <= nfx; ++x) {                                                          \
          const T ew = w[(int64_t)map_f[(int64_t)(cm[x - 1] - 1) * 6 + N] - 1 + k];               \
          for (int y = 2; y <= nfy; ++y) {                                                        \
            const int64_t d = (int64_t)map_f[(int64_t)(cm[(x - 1) + nfx * (y - 1)] - 1) * 6 + N] - 1 + k; \
            const T iw = w[d];                                                                    \
            const T a = iw * coarse[(int64_t)mc[N] - 1 + k];                                      \
            const T b = ((T)1 - iw) * coarse[(int64_t)mc[S] - 1 + k];                             \
            fine[d] = ew * (a + b);                                                               \
          }                                                                                       \
        }                                                                                         \
      }                                                                                           \
      for (int k = 0; k <= nl; ++k)                                                               \
        for (int y = 1; y <= nfy; ++y)                                                            \
          for (int x = 1; x <= nfx; ++x) {                                                        \
            const int64_t d = (int64_t)map_f[(int64_t)(cm[(x - 1) + nfx * (y - 1)] - 1) * 6 + B] - 1 + k; \
            fine[d] = w[d] * coarse[(int64_t)mc[B] - 1 + k];                                      \
          }                                                                                       \
    }                                                                                             \
  }
DEF_PROLONG_W2(float, r32)
DEF_PROLONG_W2(double, r64)

/* mm_diagonal_kernel_code (components/science/source/kernel/algebra/
 * sci_mm_diagonal_kernel_mod.F90): diag(map(df)+k) += M(ik, df, df), k outer,
 * df inner, cells colour by colour (ncolour > 0) or in order (ncolour = 0). */
#define ORC_MMDIAG(T, SUF)                                                                   \
  void orc_mm_diagonal_##SUF(int ncolour, const int* npc, const int* cic, int ncell, int nl,  \
                             T* diag, int64_t n3, const T* m, int ndf, const int* dm) {      \
    int64_t total = ncolour > 0 ? 0 : ncell;                                                 \
    for (int col = 0; col < ncolour; ++col) total += npc[col];                               \
    int col = 0, idx = 0;                                                                    \
    for (int64_t q = 0; q < total; ++q) {                                                    \
      int cell;                                                                              \
      if (ncolour == 0) {                                                                    \
        cell = (int)q + 1;                                                                   \
      } else {                                                                               \
        while (idx >= npc[col]) { ++col; idx = 0; }                                          \
        cell = cic[col + (int64_t)ncolour * idx];                                            \
        ++idx;                                                                               \
      }                                                                                      \
      const int* map = dm + (int64_t)(cell - 1) * ndf;                                       \
      for (int k = 0; k < nl; ++k) {                                                         \
        const int64_t ik = (int64_t)(cell - 1) * nl + k;                                     \
        for (int df = 0; df < ndf; ++df)                                                     \
          diag[map[df] - 1 + k] = diag[map[df] - 1 + k] + m[ik + n3 * (df + (int64_t)ndf * df)]; \
      }                                                                                      \
    }                                                                                        \
  }
ORC_MMDIAG(float, r32)
ORC_MMDIAG(double, r64)
