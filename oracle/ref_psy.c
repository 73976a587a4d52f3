/*
 * Loop driver for the reference Fortran kernels compiled by flang.
 *
 * TEST INFRASTRUCTURE ONLY.  oracle/Makefile compiles the reference's own
 * kernel sources (matrix_vector, dg_matrix_vector, dg_inc_matrix_vector,
 * transpose_matrix_vector and the columnwise app/appinv/asm kernels, plus the
 * constants/argument/kernel metadata modules they use) straight from the
 * reference tree into oracle/_ref/libref.so, and links this file with them.
 * This file is the caller: it walks cells the way the generated PSy layer
 * does (colour by colour with OpenMP over the cells of one colour for
 * continuous outputs, a plain cell loop for discontinuous ones), calling one
 * Fortran kernel per column.  It is used to pin the C oracle bit-for-bit and
 * as the CPU baseline.
 *
 * Fortran dummy arguments are passed by reference; explicit-shape arrays are
 * plain pointers.  Symbol names are flang's module-procedure mangling
 * _QM<module>P<procedure>.
 */
#include <stdint.h>

#define FK(mod, proc) _QM##mod##P##proc

#define DECL_LMA(mod, proc, T)                                                          \
  extern void FK(mod, proc)(const int* cell, const int* nlayers, T* lhs, const T* x,  \
                            const int* ncell_3d, const T* matrix, const int* ndf1,     \
                            const int* undf1, const int* map1, const int* ndf2,        \
                            const int* undf2, const int* map2);

DECL_LMA(matrix_vector_kernel_mod, matrix_vector_code_r_single, float)
DECL_LMA(matrix_vector_kernel_mod, matrix_vector_code_r_double, double)
DECL_LMA(dg_matrix_vector_kernel_mod, dg_matrix_vector_code_r_single, float)
DECL_LMA(dg_matrix_vector_kernel_mod, dg_matrix_vector_code_r_double, double)
DECL_LMA(dg_inc_matrix_vector_kernel_mod, dg_inc_matrix_vector_code_r_single, float)
DECL_LMA(dg_inc_matrix_vector_kernel_mod, dg_inc_matrix_vector_code_r_double, double)
DECL_LMA(transpose_matrix_vector_kernel_mod, transpose_matrix_vector_code_r_single, float)
DECL_LMA(transpose_matrix_vector_kernel_mod, transpose_matrix_vector_code_r_double, double)

/* r_solver is real32 in this build (R_SOLVER_PRECISION=32, the reference default). */
extern void FK(sci_columnwise_op_appinv_kernel_mod, columnwise_op_appinv_kernel_code)(
    const int* cell, const int* ncell_2d, float* lhs, const float* x, const float* cm,
    const int* nrow, const int* bw, const int* alpha, const int* beta, const int* gamma_m,
    const int* gamma_p, const int* ndf, const int* undf, const int* map, const int* ind);

extern void FK(sci_columnwise_op_app_kernel_mod, columnwise_op_app_kernel_code)(
    const int* cell, const int* ncell_2d, float* lhs, const float* x, const float* cm,
    const int* nrow, const int* ncol, const int* bw, const int* alpha, const int* beta,
    const int* gamma_m, const int* gamma_p, const int* ndf1, const int* undf1,
    const int* map1, const int* ind_to, const int* ndf2, const int* undf2, const int* map2,
    const int* ind_from);

extern void FK(sci_columnwise_op_asm_kernel_mod, columnwise_op_asm_kernel_code)(
    const int* cell, const int* nlayers, const int* ncell_2d, const int* ncell_3d,
    const float* local_stencil, float* cm, const int* nrow, const int* ncol, const int* bw,
    const int* alpha, const int* beta, const int* gamma_m, const int* gamma_p,
    const int* ndf_to, const int* cbd_to, const int* ndf_from, const int* cbd_from);

/* kernel ids: 0 matrix_vector, 1 dg_matrix_vector, 2 dg_inc_matrix_vector,
 * 3 transpose_matrix_vector */
#define DEF_CALL(T, SUF, S1, S2)                                                         \
  static void call_##SUF(int kern, int cell, int nl, T* lhs, const T* x, int n3,       \
                         const T* m, int ndf1, int undf1, const int* m1, int ndf2,     \
                         int undf2, const int* m2) {                                   \
    switch (kern) {                                                                     \
      case 0: FK(matrix_vector_kernel_mod, matrix_vector_code_##S1)(                   \
                  &cell, &nl, lhs, x, &n3, m, &ndf1, &undf1, m1, &ndf2, &undf2, m2);    \
        break;                                                                          \
      case 1: FK(dg_matrix_vector_kernel_mod, dg_matrix_vector_code_##S1)(             \
                  &cell, &nl, lhs, x, &n3, m, &ndf1, &undf1, m1, &ndf2, &undf2, m2);    \
        break;                                                                          \
      case 2: FK(dg_inc_matrix_vector_kernel_mod, dg_inc_matrix_vector_code_##S1)(     \
                  &cell, &nl, lhs, x, &n3, m, &ndf1, &undf1, m1, &ndf2, &undf2, m2);    \
        break;                                                                          \
      default: FK(transpose_matrix_vector_kernel_mod, transpose_matrix_vector_code_##S1)( \
                  &cell, &nl, lhs, x, &n3, m, &ndf1, &undf1, m1, &ndf2, &undf2, m2);    \
    }                                                                                   \
  }                                                                                     \
  /* one kernel call on one column (cell is 1-based) */                                \
  void ref_kernel_##SUF(int kern, int cell, int nl, T* lhs, const T* x, int n3,        \
                        const T* m, int ndf1, int undf1, const int* m1, int ndf2,      \
                        int undf2, const int* m2) {                                    \
    call_##SUF(kern, cell, nl, lhs, x, n3, m, ndf1, undf1, m1, ndf2, undf2, m2);       \
  }                                                                                     \
  /* coloured PSy loop: cells_in_colour(ncolour, maxc) column-major, 1-based */         \
  void ref_invoke_coloured_##SUF(int kern, int ncolour, const int* npc, const int* cic, \
                                 int nl, T* lhs, const T* x, int n3, const T* m,        \
                                 int ndf1, int undf1, const int* dm1, int ndf2,         \
                                 int undf2, const int* dm2, int nthreads) {             \
    for (int col = 0; col < ncolour; ++col) {                                           \
      int nc = npc[col];                                                                \
      _Pragma("omp parallel for schedule(static) num_threads(nthreads > 0 ? nthreads : 1)") \
      for (int idx = 0; idx < nc; ++idx) {                                              \
        int cell = cic[col + (int64_t)ncolour * idx];                                   \
        call_##SUF(kern, cell, nl, lhs, x, n3, m, ndf1, undf1,                          \
                   dm1 + (int64_t)(cell - 1) * ndf1, ndf2, undf2,                       \
                   dm2 + (int64_t)(cell - 1) * ndf2);                                   \
      }                                                                                 \
    }                                                                                   \
  }                                                                                     \
  /* plain PSy loop over cells (discontinuous output) */                                \
  void ref_invoke_cells_##SUF(int kern, int ncell, int nl, T* lhs, const T* x, int n3, \
                              const T* m, int ndf1, int undf1, const int* dm1,          \
                              int ndf2, int undf2, const int* dm2, int nthreads) {      \
    _Pragma("omp parallel for schedule(static) num_threads(nthreads > 0 ? nthreads : 1)") \
    for (int c = 0; c < ncell; ++c)                                                     \
      call_##SUF(kern, c + 1, nl, lhs, x, n3, m, ndf1, undf1, dm1 + (int64_t)c * ndf1,  \
                 ndf2, undf2, dm2 + (int64_t)c * ndf2);                                 \
  }

DEF_CALL(float, r32, r_single, x)
DEF_CALL(double, r64, r_double, x)

void ref_appinv_r32(int cell, int ncell_2d, float* lhs, const float* x, const float* cm,
                    int nrow, int bw, int alpha, int beta, int gamma_m, int gamma_p, int ndf,
                    int undf, const int* map, const int* ind) {
  FK(sci_columnwise_op_appinv_kernel_mod, columnwise_op_appinv_kernel_code)(
      &cell, &ncell_2d, lhs, x, cm, &nrow, &bw, &alpha, &beta, &gamma_m, &gamma_p, &ndf,
      &undf, map, ind);
}

void ref_invoke_appinv_r32(int ncell_2d, float* lhs, const float* x, const float* cm,
                           int nrow, int bw, int ndf, int undf, const int* dofmap,
                           const int* ind, int nthreads) {
  int one = 1, zero = 0;
#pragma omp parallel for schedule(static) num_threads(nthreads > 0 ? nthreads : 1)
  for (int c = 0; c < ncell_2d; ++c) {
    int cell = c + 1;
    FK(sci_columnwise_op_appinv_kernel_mod, columnwise_op_appinv_kernel_code)(
        &cell, &ncell_2d, lhs, x, cm, &nrow, &bw, &one, &one, &zero, &zero, &ndf, &undf,
        dofmap + (int64_t)c * ndf, ind);
  }
}

void ref_app_r32(int cell, int ncell_2d, float* lhs, const float* x, const float* cm,
                 int nrow, int ncol, int bw, int alpha, int beta, int gamma_m, int gamma_p,
                 int ndf1, int undf1, const int* map1, const int* ind_to, int ndf2, int undf2,
                 const int* map2, const int* ind_from) {
  FK(sci_columnwise_op_app_kernel_mod, columnwise_op_app_kernel_code)(
      &cell, &ncell_2d, lhs, x, cm, &nrow, &ncol, &bw, &alpha, &beta, &gamma_m, &gamma_p,
      &ndf1, &undf1, map1, ind_to, &ndf2, &undf2, map2, ind_from);
}

void ref_asm_r32(int cell, int nlayers, int ncell_2d, int ncell_3d, const float* ls, float* cm,
                 int nrow, int ncol, int bw, int alpha, int beta, int gamma_m, int gamma_p,
                 int ndf_to, const int* cbd_to, int ndf_from, const int* cbd_from) {
  FK(sci_columnwise_op_asm_kernel_mod, columnwise_op_asm_kernel_code)(
      &cell, &nlayers, &ncell_2d, &ncell_3d, ls, cm, &nrow, &ncol, &bw, &alpha, &beta,
      &gamma_m, &gamma_p, &ndf_to, cbd_to, &ndf_from, cbd_from);
}

extern void FK(sci_enforce_bc_kernel_mod, enforce_bc_code_r_single)(const int* nlayers, float* field,
                                                                    const int* ndf, const int* undf,
                                                                    const int* map, const int* bv);
extern void FK(sci_enforce_bc_kernel_mod, enforce_bc_code_r_double)(const int* nlayers, double* field,
                                                                    const int* ndf, const int* undf,
                                                                    const int* map, const int* bv);

/* enforce_bc over all cells in cell order (the kernel is called per column) */
void ref_invoke_enforce_bc_r32(int ncell, int nlayers, float* field, int ndf, int undf, const int* dofmap,
                               const int* bv) {
  for (int c = 0; c < ncell; ++c)
    FK(sci_enforce_bc_kernel_mod, enforce_bc_code_r_single)(&nlayers, field, &ndf, &undf,
                                                            dofmap + (int64_t)c * ndf, bv);
}
void ref_invoke_enforce_bc_r64(int ncell, int nlayers, double* field, int ndf, int undf, const int* dofmap,
                               const int* bv) {
  for (int c = 0; c < ncell; ++c)
    FK(sci_enforce_bc_kernel_mod, enforce_bc_code_r_double)(&nlayers, field, &ndf, &undf,
                                                            dofmap + (int64_t)c * ndf, bv);
}

/* operator_tri_solve_code (r_def = real64) per column in cell order; its
 * `ncell` argument is the 3D cell count (the first extent of matrix). */
extern void FK(sci_operator_tri_solve_kernel_mod, operator_tri_solve_code)(
    const int* cell, const int* nlayers, double* y, const double* x, const int* ncell,
    const double* matrix, const int* ndf, const int* undf, const int* map);

void ref_invoke_operator_tri_solve_r64(int ncell, int nlayers, double* y, const double* x,
                                       const double* matrix, int ndf, int undf, const int* dofmap) {
  int n3 = ncell * nlayers;
  for (int c = 0; c < ncell; ++c) {
    int cell = c + 1;
    FK(sci_operator_tri_solve_kernel_mod, operator_tri_solve_code)(&cell, &nlayers, y, x, &n3, matrix, &ndf,
                                                                 &undf, dofmap + (int64_t)c * ndf);
  }
}

/* Intergrid transfers of lowest-order scalar (W3 / Wtheta) fields, one call
 * per coarse cell with its slice of the intergrid map
 * cell_map(ncell_fine_per_coarse_x, ncell_fine_per_coarse_y, ncell_coarse). */
#define REF_TRANSFER(T, SUF)                                                                            \
  extern void FK(sci_restrict_scalar_unweighted_kernel_mod, restrict_scalar_unweighted_code_##SUF)(     \
      const int* nlayers, const int* cell_map, const int* nfx, const int* nfy, const int* ncell_fine,   \
      T* coarse, const T* fine, const int* undf_coarse, const int* map_coarse, const int* ndf,          \
      const int* undf_fine, const int* map_fine);                                                        \
  extern void FK(sci_prolong_scalar_unweighted_kernel_mod, prolong_scalar_unweighted_code_##SUF)(       \
      const int* nlayers, const int* cell_map, const int* nfx, const int* nfy, const int* ncell_fine,   \
      T* fine, const T* coarse, const int* ndf, const int* undf_fine, const int* map_fine,              \
      const int* undf_coarse, const int* map_coarse);                                                    \
  void ref_invoke_restrict_scalar_##SUF(int ncell_coarse, int ncell_fine, int nlayers, int ndf, T* coarse, \
                                        const T* fine, int undf_coarse, const int* map_coarse,          \
                                        int undf_fine, const int* map_fine, const int* cell_map, int nfx, \
                                        int nfy) {                                                        \
    for (int c = 0; c < ncell_coarse; ++c)                                                              \
      FK(sci_restrict_scalar_unweighted_kernel_mod, restrict_scalar_unweighted_code_##SUF)(            \
          &nlayers, cell_map + (int64_t)c * nfx * nfy, &nfx, &nfy, &ncell_fine, coarse, fine,          \
          &undf_coarse, map_coarse + (int64_t)c * ndf, &ndf, &undf_fine, map_fine);                     \
  }                                                                                                     \
  void ref_invoke_prolong_scalar_##SUF(int ncell_coarse, int ncell_fine, int nlayers, int ndf, T* fine,  \
                                       const T* coarse, int undf_coarse, const int* map_coarse,          \
                                       int undf_fine, const int* map_fine, const int* cell_map, int nfx,  \
                                       int nfy) {                                                         \
    for (int c = 0; c < ncell_coarse; ++c)                                                              \
      FK(sci_prolong_scalar_unweighted_kernel_mod, prolong_scalar_unweighted_code_##SUF)(              \
          &nlayers, cell_map + (int64_t)c * nfx * nfy, &nfx, &nfy, &ncell_fine, fine, coarse, &ndf,    \
          &undf_fine, map_fine, &undf_coarse, map_coarse + (int64_t)c * ndf);                           \
  }
REF_TRANSFER(float, r_single)
REF_TRANSFER(double, r_double)

/* mm_diagonal_kernel_code over all cells: colour by colour for a continuous
 * space (ncolour > 0, cells_in_colour(ncolour, maxc)), a plain cell loop for
 * a discontinuous one (ncolour = 0). */
#define DEF_MMDIAG(T, SUF, S1)                                                              \
  extern void FK(sci_mm_diagonal_kernel_mod, mm_diagonal_kernel_code_##S1)(                \
      const int* cell, const int* nlayers, T* diag, const int* ncell_3d, const T* m,         \
      const int* ndf, const int* undf, const int* map);                                     \
  void ref_invoke_mm_diagonal_##SUF(int ncolour, const int* npc, const int* cic, int ncell, \
                                    int nl, T* diag, int n3, const T* m, int ndf, int undf, \
                                    const int* dm) {                                        \
    if (ncolour == 0) {                                                                     \
      for (int c = 0; c < ncell; ++c) {                                                     \
        int cell = c + 1;                                                                   \
        FK(sci_mm_diagonal_kernel_mod, mm_diagonal_kernel_code_##S1)(                      \
            &cell, &nl, diag, &n3, m, &ndf, &undf, dm + (int64_t)c * ndf);                   \
      }                                                                                     \
      return;                                                                               \
    }                                                                                       \
    for (int col = 0; col < ncolour; ++col)                                                 \
      for (int idx = 0; idx < npc[col]; ++idx) {                                            \
        int cell = cic[col + (int64_t)ncolour * idx];                                       \
        FK(sci_mm_diagonal_kernel_mod, mm_diagonal_kernel_code_##S1)(                      \
            &cell, &nl, diag, &n3, m, &ndf, &undf, dm + (int64_t)(cell - 1) * ndf);          \
      }                                                                                     \
  }

DEF_MMDIAG(float, r32, r_single)
DEF_MMDIAG(double, r64, r_double)

/* ---- CPU baseline of the solver path ------------------------------------
 * conjugate_gradient_type%cg_solve (components/science/source/solver/
 * sci_iterative_solver_mod.f90:495-592) with the null preconditioner, as the
 * reference runs it on the CPU: each vector method is its own parallel loop
 * over the dofs (the PSyclone built-ins: setval_c, setval_X, inc_X_plus_bY,
 * inc_aX_plus_Y, inc_a_times_X), the inner products take per-thread FP64
 * partials over a static schedule summed in thread order
 * (components/science/source/psy/sci_psykal_light_mod.f90:121-134), and the
 * operator apply is setval_c(y, 0) followed by the flang-built
 * matrix_vector_code in the coloured PSy loop.  Returns the iterations done;
 * r_nrm receives the last residual norm (and r_nrm0 the first). */
#include <math.h>
#include <omp.h>
#include <stdlib.h>

static double psy_dot(const double* x, const double* y, int n, int nt) {
  double part[256];
  int nth = nt > 256 ? 256 : (nt > 0 ? nt : 1);
  for (int t = 0; t < nth; ++t) part[t] = 0.0;
#pragma omp parallel num_threads(nth)
  {
    const int t = omp_get_thread_num(), T = omp_get_num_threads();
    const int lo = (int)((int64_t)n * t / T), hi = (int)((int64_t)n * (t + 1) / T);
    double s = 0.0;
    for (int i = lo; i < hi; ++i) s += x[i] * y[i];
    part[t] = s;
  }
  double tot = 0.0;
  for (int t = 0; t < nth; ++t) tot += part[t];
  return tot;
}

static void psy_apply(int ncolour, const int* npc, const int* cic, int nl, double* y, const double* x, int n3,
                      const double* m, int ndf, int undf, const int* dm, int nt) {
#pragma omp parallel for schedule(static) num_threads(nt > 0 ? nt : 1)
  for (int i = 0; i < undf; ++i) y[i] = 0.0;
  ref_invoke_coloured_r64(0, ncolour, npc, cic, nl, y, x, n3, m, ndf, undf, dm, ndf, undf, dm, nt);
}

int ref_cg_r64(int ncolour, const int* npc, const int* cic, int nl, int n3, const double* m, int ndf,
               int undf, const int* dm, double* x, const double* b, double r_tol, int max_iter, int monitor,
               int nt, double* r_nrm0_out, double* r_nrm_out) {
  double* r = (double*)malloc(sizeof(double) * undf);
  double* p = (double*)malloc(sizeof(double) * undf);
  double* z = (double*)malloc(sizeof(double) * undf);
  const int nth = nt > 0 ? nt : 1;
  psy_apply(ncolour, npc, cic, nl, r, x, n3, m, ndf, undf, dm, nth);            /* r = A x */
#pragma omp parallel for schedule(static) num_threads(nth)
  for (int i = 0; i < undf; ++i) r[i] = -1.0 * r[i];                             /* r = -r */
#pragma omp parallel for schedule(static) num_threads(nth)
  for (int i = 0; i < undf; ++i) r[i] = r[i] + 1.0 * b[i];                       /* r += b */
  const double r0 = sqrt(psy_dot(r, r, undf, nth));
  double rn = r0;
  int it = 0;
  if (r0 >= 2.220446049250313e-16) {
#pragma omp parallel for schedule(static) num_threads(nth)
    for (int i = 0; i < undf; ++i) z[i] = r[i];                                  /* z = P^{-1} r (copy) */
    double rz = psy_dot(r, z, undf, nth);
#pragma omp parallel for schedule(static) num_threads(nth)
    for (int i = 0; i < undf; ++i) p[i] = z[i];
    for (it = 1; it <= max_iter; ++it) {
      psy_apply(ncolour, npc, cic, nl, z, p, n3, m, ndf, undf, dm, nth);        /* z = A p */
      const double alpha = rz / psy_dot(p, z, undf, nth);
#pragma omp parallel for schedule(static) num_threads(nth)
      for (int i = 0; i < undf; ++i) x[i] = x[i] + alpha * p[i];
#pragma omp parallel for schedule(static) num_threads(nth)
      for (int i = 0; i < undf; ++i) r[i] = r[i] + (-alpha) * z[i];
      if (monitor) {
        rn = sqrt(psy_dot(r, r, undf, nth));
        if (rn / r0 <= r_tol) break;
      }
#pragma omp parallel for schedule(static) num_threads(nth)
      for (int i = 0; i < undf; ++i) z[i] = r[i];
      const double rz_new = psy_dot(r, z, undf, nth);
      const double beta = rz_new / rz;
#pragma omp parallel for schedule(static) num_threads(nth)
      for (int i = 0; i < undf; ++i) p[i] = beta * p[i] + z[i];
      rz = rz_new;
    }
    if (it > max_iter) it = max_iter;
  }
  if (r_nrm0_out) *r_nrm0_out = r0;
  if (r_nrm_out) *r_nrm_out = rn;
  free(r);
  free(p);
  free(z);
  return it;
}
