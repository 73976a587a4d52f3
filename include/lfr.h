/*
 * lfr.h - C ABI of the MI355X operator-apply / column-solve library.
 *
 * One call here replaces one PSy-layer `invoke` of the reference (the loop
 * over cells or colours that the generated PSy code wraps around a kernel),
 * not one per-column kernel call: a per-cell C call would cost one GPU launch
 * per column.  Argument vocabulary follows the reference kernels:
 *
 *   - dofmaps are map(ndf, ncell) in Fortran memory order, 1-based dof
 *     numbers of the bottom cell of each column (layer k adds k);
 *   - LMA operators are local_stencil(ncell_3d, ndf_to, ndf_from) in Fortran
 *     memory order, ncell_3d = ncell * nlayers, ik = (cell-1)*nlayers + k + 1;
 *   - columnwise (CMA) operators are columnwise_matrix(bandwidth, nrow,
 *     ncell_2d) with bands (1,2,3) = (sub, diag, super) for tridiagonals;
 *   - field data are plain device arrays of undf reals; `prec` is the real
 *     kind in bytes: 4 (r_single / r_solver) or 8 (r_double / r_def).
 *
 * Every array argument named *_d is a device pointer (hipMalloc'd, e.g. by
 * lfr_malloc or by PyTorch).  Host arrays are named *_h.  All work is queued
 * on the context's stream; calls return when the work is queued unless they
 * return a host value (dots, norms, solver statistics).
 *
 * Errors: every entry point returns LFR_OK (0) or a nonzero LFR_ERR_* code
 * and records a message retrievable with lfr_last_error().  The reference
 * kernels have no error channel; its infrastructure aborts through
 * log_event(..., LOG_LEVEL_ERROR) (infrastructure/source/utilities/log_mod.F90:381-476),
 * which is what a Fortran shim over this ABI does on a nonzero return.
 *
 * A context is not re-entrant: calls on one context must not overlap.
 */
#ifndef LFR_H
#define LFR_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

enum {
  LFR_OK = 0,
  LFR_ERR_ARG = 1,           /* invalid argument / inconsistent sizes        */
  LFR_ERR_HIP = 2,           /* HIP runtime error                            */
  LFR_ERR_NOMEM = 3,         /* device allocation failed                     */
  LFR_ERR_NOT_CONVERGED = 4, /* solver hit max_iter with fail_on_non_converged */
  LFR_ERR_NAN = 5,           /* NaN detected in a solver residual            */
  LFR_ERR_SINGULAR = 6       /* zero pivot in a column solve                  */
};

typedef struct lfr_ctx lfr_ctx;
typedef struct lfr_space lfr_space;
typedef struct lfr_plan lfr_plan;
typedef struct lfr_lma_packed lfr_lma_packed;

/* ---------------------------------------------------------------------- */
/* Context, memory, streams                                                */
/* ---------------------------------------------------------------------- */

/* Create a context on `device` with its own default-priority stream. */
int lfr_init(int device, lfr_ctx** ctx);
int lfr_finalise(lfr_ctx* ctx);
const char* lfr_last_error(const lfr_ctx* ctx);
/* Queue subsequent work on an existing hipStream_t; NULL selects HIP's default
 * (null) stream.  Until the first call the context uses its own stream. */
int lfr_set_stream(lfr_ctx* ctx, void* hip_stream);
void* lfr_get_stream(const lfr_ctx* ctx);
int lfr_synchronize(lfr_ctx* ctx);
int lfr_malloc(lfr_ctx* ctx, size_t bytes, void** ptr_d);
int lfr_free(lfr_ctx* ctx, void* ptr_d);
int lfr_memcpy_h2d(lfr_ctx* ctx, void* dst_d, const void* src_h, size_t bytes);
int lfr_memcpy_d2h(lfr_ctx* ctx, void* dst_h, const void* src_d, size_t bytes);

/* ---------------------------------------------------------------------- */
/* Function spaces and gather plans                                        */
/* ---------------------------------------------------------------------- */

/* Upload the dofmap of one function space: map(ndf, ncell) host array as
 * returned by function_space_type%get_whole_dofmap()
 * (infrastructure/source/function_space/function_space_mod.F90:726-750),
 * with nlayers and undf (= last_dof_owned on one rank, :1277-1302). */
int lfr_space_create(lfr_ctx* ctx, const int32_t* dofmap_h, int ndf, int ncell, int nlayers,
                     int undf, lfr_space** space);
int lfr_space_destroy(lfr_space* space);

/* Colour-ordered inverse dofmap of a continuous output space.
 *
 * The reference writes GH_INC outputs colour by colour
 * (infrastructure/build/psyclone/psyclone_tools.py:71-119) using the cell
 * colouring of infrastructure/source/mesh/mesh_colouring_mod.F90, so a dof
 * shared by several cells receives their contributions in colour order.  The
 * plan lists, for each column of output dofs, the (cell, local dof) pairs
 * that touch it in exactly that order, so that one owner-computes gather pass
 * reproduces the coloured sum without atomics or colour barriers.
 * cells_in_colour(ncolour, maxcells) is the reference's
 * mesh%get_colour_map() array in Fortran memory order (1-based, 0-padded).
 * kind: 0 for matrix_vector order (df2 outer, df inner within a cell),
 *       1 for transpose_matrix_vector order (layer outer, df inner). */
int lfr_plan_create(lfr_ctx* ctx, const lfr_space* out_space, const int32_t* cells_in_colour_h,
                    const int32_t* ncells_per_colour_h, int ncolour, int maxcells, int kind,
                    lfr_plan** plan);
int lfr_plan_destroy(lfr_plan* plan);
/* Number of output-dof columns and of classes (for diagnostics). */
int lfr_plan_info(const lfr_plan* plan, int* ncolumns, int* nclasses, int64_t* nbytes);

/* ---------------------------------------------------------------------- */
/* LMA operator applies (one call = one invoke over all cells)             */
/* ---------------------------------------------------------------------- */

/* invoke(matrix_vector_kernel_type(lhs, x, matrix)):
 * lhs += matrix * x on a continuous output space, all cells, colour order.
 * Replaces matrix_vector_code (components/science/source/kernel/algebra/
 * matrix_vector_kernel_mod.F90:68-140) looped by the PSy layer.
 * If overwrite != 0 the call computes lhs = matrix * x, i.e. the pair
 * invoke(setval_c(lhs, 0.0), matrix_vector_kernel_type(lhs, x, matrix)) of
 * mass_matrix_operator_type%apply (components/science/source/algorithm/solver/
 * sci_mass_matrix_operator_alg_mod.x90:98-99), without reading lhs.
 * `plan` must be a kind-0 plan for `to_space`. */
int lfr_matrix_vector(lfr_ctx* ctx, int prec, void* lhs_d, const void* x_d, const void* matrix_d,
                      const lfr_space* to_space, const lfr_space* from_space,
                      const lfr_plan* plan, int overwrite);

/* invoke(dg_matrix_vector_kernel_type(lhs, x, matrix)): lhs = matrix * x on a
 * discontinuous output space (components/science/source/kernel/algebra/
 * dg_matrix_vector_kernel_mod.F90:68-154). */
int lfr_dg_matrix_vector(lfr_ctx* ctx, int prec, void* lhs_d, const void* x_d,
                         const void* matrix_d, const lfr_space* to_space,
                         const lfr_space* from_space);

/* invoke(dg_inc_matrix_vector_kernel_type(lhs, x, matrix)): lhs += matrix * x on
 * a discontinuous output space (components/science/source/kernel/algebra/
 * dg_inc_matrix_vector_kernel_mod.F90:69-144). */
int lfr_dg_inc_matrix_vector(lfr_ctx* ctx, int prec, void* lhs_d, const void* x_d,
                             const void* matrix_d, const lfr_space* to_space,
                             const lfr_space* from_space);

/* invoke(transpose_matrix_vector_kernel_type(lhs, x, matrix)): lhs += matrix^T x
 * where matrix is stored local_stencil(ncell_3d, ndf_x, ndf_lhs)
 * (components/science/source/kernel/algebra/transpose_matrix_vector_kernel_mod.F90:66-144).
 * This is also the adjoint of lfr_matrix_vector / lfr_dg_matrix_vector.
 * `plan` must be a kind-1 plan for `lhs_space`.  overwrite != 0: lhs = M^T x. */
int lfr_transpose_matrix_vector(lfr_ctx* ctx, int prec, void* lhs_d, const void* x_d,
                                const void* matrix_d, const lfr_space* lhs_space,
                                const lfr_space* x_space, const lfr_plan* plan, int overwrite);

/* lfr_matrix_vector (kind-0 plan) or lfr_transpose_matrix_vector (kind-1
 * plan of lhs' space) on a multidata field of `ndata` fields stored "ndata
 * first" (element (dof d, field r) at d*ndata + r, the layout of LFRic's
 * multidata function spaces,
 * infrastructure/source/function_space/function_space_constructor_helper_functions_mod.F90:1838-1846):
 * the same operator applied to every field, each field's result bit-identical
 * to a single-field apply.  1 <= ndata <= 64. */
int lfr_matrix_vector_ndata(lfr_ctx* ctx, int prec, void* lhs_d, const void* x_d, const void* matrix_d,
                            const lfr_space* lhs_space, const lfr_space* x_space, const lfr_plan* plan,
                            int ndata, int overwrite);

/* compute_mass_matrix_w_scalar_code for w_scalar = chi space = lowest-order
 * W0 (components/science/source/kernel/fem/sci_compute_mass_matrix_kernel_w_scalar_mod.F90:150-174)
 * over every cell-layer, with Cartesian coordinates chi1/chi2/chi3 (FP64 W0
 * fields on `space`) and the coordinate Jacobian of
 * components/science/source/kernel/geometry/sci_coordinate_jacobian_mod.F90:330-345,430-440:
 * nq-point Gauss-Legendre quadrature per direction (1 <= nq <= 6).  mass_d
 * receives local_stencil(ncell_3d, 8, 8) in FP64.  stiff_d (may be NULL)
 * receives the stiffness matrix sum_q w (J^-T grad phi_a).(J^-T grad phi_b)
 * |detJ| (no reference kernel; the diffusion term of the assembled
 * operator). */
int lfr_compute_mass_matrix_w0(lfr_ctx* ctx, int nq, const void* chi1_d, const void* chi2_d, const void* chi3_d,
                               const lfr_space* space, void* mass_d, void* stiff_d);

/* Mass matrices of lowest-order spaces from basis tables the caller
 * evaluates at the quadrature points (as the PSy layer passes them):
 *   kind 0, scalar space (W0 ndf 8, Wtheta / W2-vertical-scalar ndf 2, W3 ndf 1):
 *     compute_mass_matrix_w_scalar_code
 *     (components/science/source/kernel/fem/sci_compute_mass_matrix_kernel_w_scalar_mod.F90:150-174)
 *   kind 1, vector space with the Piola map (W2 ndf 6, W2H ndf 4, W2V ndf 2):
 *     compute_mass_matrix_w2_code
 *     (components/science/source/kernel/fem/sci_compute_mass_matrix_kernel_w2_mod.F90:131-162)
 * Host tables in Fortran order: basis(ncomp, ndf, nqp_h, nqp_v) with ncomp 1
 * (kind 0) or 3 (kind 1), diff_basis_chi(3, 8, nqp_h, nqp_v) of the
 * lowest-order W0 chi space, wqp_h(nqp_h), wqp_v(nqp_v).  chi1/2/3_d are FP64
 * fields on chi_space (lowest-order W0, Cartesian coordinates); mass_d
 * receives local_stencil(ncell_3d, ndf, ndf) in FP64.  Synchronous (the
 * tables are read from host memory during the call). */
int lfr_compute_mass_matrix(lfr_ctx* ctx, int kind, int ndf, int nqp_h, int nqp_v, const double* basis,
                            const double* diff_basis_chi, const double* wqp_h, const double* wqp_v,
                            const void* chi1_d, const void* chi2_d, const void* chi3_d,
                            const lfr_space* chi_space, void* mass_d);

/* Basis tables of the lowest-order de Rham complex at the quadrature points,
 * host FP64, Fortran order (ncomp, ndf, nqp_h, nqp_v), the func_type
 * arguments of compute_derham_matrices_code.  A table may be NULL when no
 * requested output needs it. */
typedef struct lfr_derham_tables {
  int nqp_h, nqp_v;
  const double* wqp_h;           /* (nqp_h) */
  const double* wqp_v;           /* (nqp_v) */
  const double* diff_basis_chi;  /* (3, 8, ...) lowest-order W0 gradients */
  const double* basis_w0;        /* (1, 8, ...) */
  const double* diff_basis_w0;   /* (3, 8, ...) gradient */
  const double* basis_w1;        /* (3, 12, ...) */
  const double* diff_basis_w1;   /* (3, 12, ...) curl */
  const double* basis_w2;        /* (3, 6, ...) */
  const double* diff_basis_w2;   /* (1, 6, ...) divergence */
  const double* basis_w2b;       /* (3, 6, ...) W2broken */
  const double* diff_basis_w2b;  /* (1, 6, ...) */
  const double* basis_w3;        /* (1, 1, ...) */
  const double* basis_wt;        /* (1, 2, ...) Wtheta */
} lfr_derham_tables;

/* compute_derham_matrices_code
 * (components/science/source/kernel/fem/sci_compute_derham_matrices_kernel_mod.F90:134-387)
 * over every cell-layer, lowest order, Cartesian chi (FP64 fields on the
 * lowest-order W0 chi_space): the mass matrices of W0, W1 (covariant map),
 * W2 and W2broken (Piola), W3 and Wtheta, the weak gradient (W1 x W0), curl
 * (W2 x W1), divergence (W3 x W2) and broken divergence (W3 x W2broken)
 * matrices, each local_stencil(ncell_3d, ndf_to, ndf_from) in FP64.  Any
 * output may be NULL (not computed).  Synchronous. */
int lfr_compute_derham_matrices(lfr_ctx* ctx, const lfr_derham_tables* tables, const void* chi1_d,
                                const void* chi2_d, const void* chi3_d, const lfr_space* chi_space, void* mm0,
                                void* mm1, void* mm2, void* mm2b, void* mm3, void* mmt, void* grad, void* curl,
                                void* div, void* broken_div);

/* Intergrid transfers between a coarse and a fine mesh of a multigrid
 * hierarchy: lowest-order scalar fields (W3: ndf 1, Wtheta: ndf 2) and W2
 * fluxes (ndf 6).
 * lfr_intergrid_create checks and uploads the intergrid map once: cell_map_h
 * (host, 1-based fine cells) is cell_map(nfx, nfy, ncell_coarse) in Fortran
 * order, the mesh-to-mesh connectivity of the two meshes.  Over every coarse
 * cell:
 *   lfr_restrict_scalar: restrict_scalar_unweighted_code
 *     (components/science/source/kernel/inter_mesh/sci_restrict_scalar_unweighted_kernel_mod.F90:71-125),
 *     coarse = the mean of the children (each child times 1/(nfx nfy),
 *     summed y outer, x inner); coarse_d is overwritten on the coarse columns.
 *   lfr_prolong_scalar: prolong_scalar_unweighted_code
 *     (components/science/source/kernel/inter_mesh/sci_prolong_scalar_unweighted_kernel_mod.F90),
 *     fine += coarse on every child. */
typedef struct lfr_intergrid lfr_intergrid;
int lfr_intergrid_create(lfr_ctx* ctx, const lfr_space* coarse_space, const lfr_space* fine_space,
                         const int32_t* cell_map_h, int nfx, int nfy, lfr_intergrid** intergrid);
int lfr_intergrid_destroy(lfr_intergrid* intergrid);
int lfr_restrict_scalar(lfr_ctx* ctx, int prec, void* coarse_d, const void* fine_d, const lfr_intergrid* intergrid);
int lfr_prolong_scalar(lfr_ctx* ctx, int prec, void* fine_d, const void* coarse_d, const lfr_intergrid* intergrid);
/* restrict_w2_code
 * (components/science/source/kernel/inter_mesh/sci_restrict_w2_kernel_mod.F90)
 * for an intergrid map between lowest-order W2 spaces: every coarse side
 * face receives the sum of the fine faces on it, every B face column the sum
 * over the children; the faces a coarse cell writes follow the face
 * selectors of face_selector_code
 * (components/science/source/kernel/algebra/sci_face_selector_kernel_mod.F90,
 * computed when the handle is made), and a face written by two cells keeps
 * the later cell's value, as the reference's cell loop leaves it. */
int lfr_restrict_w2(lfr_ctx* ctx, int prec, void* coarse_d, const void* fine_d, const lfr_intergrid* intergrid);
/* prolong_w2_kernel_code
 * (components/science/source/kernel/inter_mesh/sci_prolong_w2_kernel_mod.F90)
 * with the weights weights_prolong_w2_kernel_code
 * (.../sci_weights_prolong_w2_kernel_mod.F90) computes, for square children
 * (nfx = nfy): every fine face of the coarse cells is overwritten. */
int lfr_prolong_w2(lfr_ctx* ctx, int prec, void* fine_d, const void* coarse_d, const lfr_intergrid* intergrid);

/* Packed operator for repeated applies of one LMA operator (the
 * operator_type whose local_stencil, INF/operator/operator_mod.t90:31-122,
 * is assembled once and applied many times, as the mass matrices in the
 * solvers are).  Copies, in one device pass, the matrix rows that each
 * output column of `plan` reads into one contiguous block per column, in
 * the plan's summation order, so an apply streams the matrix front to back.
 * kind-0 plan: the operator of lfr_matrix_vector (x_space = from_space);
 * kind-1 plan: the operator of lfr_transpose_matrix_vector (x_space = the
 * space of x).  The handle does not reference matrix_d afterwards; it holds
 * about the matrix's bytes (plus zero padding where a column entry has no
 * layer, 1.4 % for W2 lowest order).  Every class of the plan must have a
 * record kernel (ndf_in 1, 6 or 8); otherwise LFR_ERR_ARG. */
int lfr_lma_pack(lfr_ctx* ctx, int prec, const void* matrix_d, const lfr_space* x_space,
                 const lfr_plan* plan, lfr_lma_packed** out);
int lfr_lma_packed_destroy(lfr_lma_packed* pk);
/* Device bytes held by a packed operator. */
int lfr_lma_packed_info(const lfr_lma_packed* pk, long long* nbytes);
/* lfr_matrix_vector (kind-0 pack) or lfr_transpose_matrix_vector (kind-1
 * pack) with a packed operator: bit-identical results; lhs on the plan's
 * space, x on x_space, overwrite != 0: lhs = A x. */
int lfr_matrix_vector_packed(lfr_ctx* ctx, const lfr_lma_packed* pk, void* lhs_d, const void* x_d,
                             int overwrite);


/* Gradient of an LMA apply with respect to the local matrices:
 * Mbar(ik,i,j) (+)= alpha * ybar(map_to(i)+k) * x(map_from(j)+k), for the
 * operator local_stencil(ncell_3d, ndf_to, ndf_from) of
 * lfr_matrix_vector/lfr_dg_matrix_vector.  No counterpart in the reference
 * (it has no adjoint with respect to operator coefficients); used by the
 * autograd front end. */
int lfr_lma_outer(lfr_ctx* ctx, int prec, void* Mbar_d, const void* ybar_d, const void* x_d,
                  const lfr_space* to_space, const lfr_space* from_space, double alpha,
                  int accumulate);

/* ---------------------------------------------------------------------- */
/* Vector algebra (abstract_vector_type methods / PSyclone built-ins)      */
/* ---------------------------------------------------------------------- */
/* Each call is one pass over n = undf dofs (1..last_dof_owned on one rank).
 * Scalars are r_def (double) and are cast to the field kind before use, as
 * r_solver_field_vector_type does (components/science/source/algorithm/
 * solver/sci_r_solver_field_vector_mod.x90:189-195).  Built-in semantics are
 * those named in sci_field_vector_mod.x90:132-283. */
int lfr_vec_setval(lfr_ctx* ctx, int prec, int64_t n, void* x_d, double c);            /* setval_c: x = c      */
int lfr_vec_copy(lfr_ctx* ctx, int prec, int64_t n, void* y_d, const void* x_d);       /* setval_X: y = x      */
int lfr_vec_axpy(lfr_ctx* ctx, int prec, int64_t n, double a, const void* x_d, void* y_d);  /* inc_X_plus_bY: y += a x */
int lfr_vec_aypx(lfr_ctx* ctx, int prec, int64_t n, double a, const void* x_d, void* y_d);  /* inc_aX_plus_Y: y = a y + x */
int lfr_vec_axpby(lfr_ctx* ctx, int prec, int64_t n, double a, const void* x_d, double b,
                  const void* y_d, void* z_d);                                          /* aX_plus_bY: z = a x + b y */
int lfr_vec_scale(lfr_ctx* ctx, int prec, int64_t n, double a, void* x_d);             /* inc_a_times_X: x = a x */
int lfr_vec_divide(lfr_ctx* ctx, int prec, int64_t n, void* z_d, const void* x_d,
                   const void* y_d);                                                    /* X_divideby_Y: z = x / y */
int lfr_vec_multiply(lfr_ctx* ctx, int prec, int64_t n, void* z_d, const void* x_d,
                     const void* y_d);                                                  /* X_times_Y: z = x * y */
/* Copy between kinds (copy_32_to_64 / copy_64_to_32 of
 * components/science/source/psy/sci_psykal_builtin_light_mod.f90). */
int lfr_vec_convert(lfr_ctx* ctx, int prec_out, void* y_d, int prec_in, const void* x_d, int64_t n);
/* X_innerproduct_Y: sum of real(x,r_double)*real(y,r_double), FP64
 * accumulation (components/science/source/psy/sci_psykal_light_mod.f90:146-214).
 * Deterministic: fixed grid for a given n, fixed-order partial combine.
 * lfr_vec_dot returns the value to the host; lfr_vec_dot_dev leaves it in
 * device memory (no host synchronisation). */
int lfr_vec_dot(lfr_ctx* ctx, int prec, int64_t n, const void* x_d, const void* y_d, double* out_h);
int lfr_vec_dot_dev(lfr_ctx* ctx, int prec, int64_t n, const void* x_d, const void* y_d, double* out_d);
/* lfr_vec_dot_dev without any allocation: no counterpart in the reference;
 * for callers that cannot allocate device memory (code under HIP graph
 * capture, tight serving loops).  Same grid, order and bits as
 * lfr_vec_dot_dev.  The partial sums go to a fixed scratch array of
 * 1024 doubles (8 KiB) per GPU, held until the process ends and shared by
 * every call on that GPU: calls that overlap on one GPU (two streams, or two
 * contexts on one device) race.  Any n > 0 fits (the reduction grid never
 * exceeds 1024 blocks); the result is left in out_d. */
int lfr_vec_dot_dev_noalloc(lfr_ctx* ctx, int prec, int64_t n, const void* x_d, const void* y_d,
                            double* out_d);

/* ---------------------------------------------------------------------- */
/* Solvers                                                                 */
/* ---------------------------------------------------------------------- */

/* A square LMA operator y = M x (mass_matrix_operator_type%apply without
 * boundary mask: components/science/source/algorithm/solver/
 * sci_mass_matrix_operator_alg_mod.x90:72-119).  plan: kind-0 plan of space. */
typedef struct lfr_lma_op {
  int prec;                 /* 4 or 8: kind of matrix and vectors          */
  const void* matrix_d;     /* local_stencil(ncell_3d, ndf, ndf)            */
  const lfr_space* space;
  const lfr_plan* plan;
  int apply_bc;             /* 1: enforce_bc after every apply (the space's
                               boundary values must be set), as
                               mass_matrix_operator_type%apply with apply_bc */
  const lfr_lma_packed* packed; /* NULL, or a kind-0 pack of matrix_d with
                               `plan`: applies then read the packed copy     */
} lfr_lma_op;

/* Solver settings as in the reference constructors
 * (components/science/source/solver/sci_iterative_solver_mod.f90:465-490). */
typedef struct lfr_solver_params {
  double r_tol;             /* stop when ||r||/||r0|| <= r_tol ...            */
  double a_tol;             /* ... or ||r|| <= a_tol (only if monitoring)     */
  int max_iter;
  int monitor_convergence;  /* 0: run exactly max_iter iterations             */
  int fail_on_non_converged;/* return LFR_ERR_NOT_CONVERGED when not converged */
  int preconditioner;       /* 0 null (y = x), 1 diagonal (y = x / diag)      */
  const void* diag_d;       /* diagonal for preconditioner 1                  */
  int gcrk;                 /* Krylov space size (GMRES)                      */
  int x_is_zero;            /* 1: ignore x on entry and start from x = 0      */
} lfr_solver_params;

typedef struct lfr_solver_stats {
  int iterations;
  int converged;
  double r_nrm0;            /* ||b - A x0||                                   */
  double r_nrm;             /* final ||r|| as the solver last computed it      */
} lfr_solver_stats;

/* conjugate_gradient_type%cg_solve (sci_iterative_solver_mod.f90:495-592):
 * solves A x = b; x holds the initial guess on entry.  FP64 scalars; for
 * prec 4 vectors the scalars are cast to FP32 before each update and inner
 * products accumulate in FP64. */
int lfr_cg(lfr_ctx* ctx, const lfr_lma_op* A, const lfr_solver_params* params, void* x_d,
           const void* b_d, lfr_solver_stats* stats);

/* gmres_type%gmres_solve (sci_iterative_solver_mod.f90:830-1003):
 * left-preconditioned restarted GMRES(gcrk) with modified Gram-Schmidt;
 * max_iter counts restarts; convergence is ||b - A x|| / max(||r0||, a_tol)
 * < r_tol, tested once per restart when monitoring. */
int lfr_gmres(lfr_ctx* ctx, const lfr_lma_op* A, const lfr_solver_params* params, void* x_d,
              const void* b_d, lfr_solver_stats* stats);

typedef struct lfr_refine_stats {
  int outer;                /* FP64 residual corrections applied              */
  int inner_total;          /* FP32 CG iterations over all corrections        */
  int converged;
  double b_nrm;
  double r_nrm;             /* final FP64 ||b - A x||                        */
} lfr_refine_stats;

/* Iterative refinement: FP64 residual r = b - A64 x, FP32 CG on A32 d = r
 * (inner settings; started from 0), x += d, until ||r|| <= r_tol ||b||.
 * A32 is the same operator stored in FP32 (r_solver).  No counterpart
 * routine in the reference; it combines cg_solve with the r_def / r_solver
 * split of components/science/source/algorithm/solver/
 * sci_r_solver_field_vector_mod.x90. */
int lfr_refined_cg(lfr_ctx* ctx, const lfr_lma_op* A32, const lfr_lma_op* A64, double r_tol,
                   int max_outer, const lfr_solver_params* inner, void* x64_d, const void* b64_d,
                   lfr_refine_stats* stats);

/* ---------------------------------------------------------------------- */
/* Boundary conditions                                                     */
/* ---------------------------------------------------------------------- */

/* Attach boundary_value(ndf, 2) (function_space_type%get_boundary_dofs(),
 * infrastructure/source/function_space/function_space_mod.F90:784-794;
 * Fortran order, 0/1 flags) to a space. */
int lfr_space_set_boundary(lfr_space* space, const int32_t* boundary_value_h);

/* invoke(enforce_bc_kernel_type(field)) over all cells: field(map(df)) *=
 * boundary_value(df,1), field(map(df)+nlayers-1) *= boundary_value(df,2)
 * (components/science/source/kernel/algebra/sci_enforce_bc_kernel_mod.F90:63-91). */
int lfr_enforce_bc(lfr_ctx* ctx, int prec, void* field_d, const lfr_space* space);

/* ---------------------------------------------------------------------- */
/* Column-wise (CMA) operators                                             */
/* ---------------------------------------------------------------------- */

/* invoke(columnwise_op_appinv_kernel_type(lhs, x, cma_op)) over all columns:
 * Thomas solve of each tridiagonal column, OVERWRITING
 * lhs(map(1) + indirection_dofmap(i) - 1) (components/science/source/kernel/
 * columnwise_operators/sci_columnwise_op_appinv_kernel_mod.F90:77-145; the
 * kernel's doc says lhs += A^{-1} x but the code assigns).  cm is
 * columnwise_matrix(bandwidth=3, nrow, ncell_2d); indirection_d is the
 * device copy of indirection_dofmap(nrow); space gives map(1) of each column,
 * and map(1) + nrow - 1 must not pass undf for any column.  The indirection
 * is not read on the host: entries outside 1..nrow are clamped into the
 * column (lfr_cma_create validates its own copy).
 * fp64_pivots != 0 runs the FP32 recurrence in FP64 (pivots, c', d') and
 * rounds the solution to FP32 (for ill-conditioned columns). */
int lfr_columnwise_op_appinv(lfr_ctx* ctx, int prec, void* lhs_d, const void* x_d,
                             const void* cm_d, int bandwidth, int nrow, int ncell_2d,
                             const lfr_space* space, const int32_t* indirection_d,
                             int fp64_pivots);

/* invoke(tri_solve_kernel_type(y, x, tri_0, tri_plus, tri_minus)) over every
 * column (components/science/source/kernel/solver/sci_tri_solve_kernel_mod.f90:95-118):
 * Thomas solve of tri_minus(k) y(k-1) + tri_0(k) y(k) + tri_plus(k) y(k+1) = x(k)
 * on the nlayers dofs of each column of a lowest-order W3 space (ndf 1); y is
 * overwritten and must not alias an input.  Uses undf values of context
 * workspace. */
int lfr_tri_solve(lfr_ctx* ctx, int prec, void* y_d, const void* x_d, const void* tri_0_d,
                  const void* tri_plus_d, const void* tri_minus_d, const lfr_space* space);

/* invoke(operator_tri_solve_kernel_type(y, x, matrix)) over every column
 * (components/science/source/kernel/solver/sci_operator_tri_solve_kernel_mod.f90:65-120):
 * the tridiagonal of the nlayers+1 dofs of each column of a lowest-order
 * Wtheta or W2v space (ndf 2, top dof = bottom dof + 1) assembled from the
 * 2x2 local matrices local_stencil(ncell_3d, 2, 2), then the Thomas solve;
 * y is overwritten.  The reference kernel is r_def; prec 4 runs the same
 * steps in FP32. */
int lfr_operator_tri_solve(lfr_ctx* ctx, int prec, void* y_d, const void* x_d, const void* matrix_d,
                           const lfr_space* space);

/* ---------------------------------------------------------------------- */
/* Multigrid with vertical line relaxation (W3, lowest order)              */
/* ---------------------------------------------------------------------- */
/* V-cycle for A = T + sigma_l * L_h on a hierarchy of nested cubed spheres:
 * T is a column tridiagonal (the CMA operator of columnwise_op_appinv), L_h
 * the 4-neighbour horizontal graph Laplacian.  Smoother: damped vertical line
 * Jacobi (one Thomas solve per column on the residual), restriction by cell
 * average (components/science/source/kernel/inter_mesh/
 * sci_restrict_scalar_unweighted_kernel_mod.F90), prolongation by injection
 * (.../sci_prolong_scalar_unweighted_kernel_mod.F90).  The whole V-cycle over
 * nrhs right-hand sides is captured once into a hipGraph and replayed; the
 * smoother's workgroup size is autotuned per level at creation.  No single
 * counterpart in the reference (LFRic's multigrid preconditioner lives in
 * lfric_apps); the components are the reference kernels named above.
 *   ncell[l]       cells of level l (0 = finest)
 *   cell_next_h[l] (ncell[l], 4) 0-based W,S,E,N neighbours
 *   children_h[l]  (ncell[l], 4) 0-based cells of level l-1 under each cell,
 *                  y outer, x inner (ignored for l = 0)
 *   bands_h[l]     [3][nlayers][ncell[l]] sub, diag, super of T
 *   sigma[l]       horizontal coupling of level l
 * nu1/nu2/ncoarse must be even; nrhs is a power of 2 up to 32; nlayers <= 128;
 * each level has 4x the cells of the next coarser one. */
typedef struct lfr_mg lfr_mg;
int lfr_mg_create(lfr_ctx* ctx, int prec, int nlev, int nlayers, int nrhs, const int* ncell,
                  const int32_t* const* cell_next_h, const int32_t* const* children_h,
                  const double* const* bands_h, const double* sigma, double omega, int nu1,
                  int nu2, int ncoarse, lfr_mg** mg);
int lfr_mg_destroy(lfr_mg* mg);
int lfr_mg_block_sizes(const lfr_mg* mg, int* smooth_blocks);
/* x_d, b_d: nrhs fine-level W3 fields (LFRic layout, dof = cell*nl + k),
 * stored back to back.  Runs ncycles V-cycles starting from x; if res_h is
 * not NULL it receives ||b - A x|| per right-hand side before the first cycle
 * (res_h[0..nrhs)) and after every cycle (res_h[(c+1)*nrhs + r]). */
int lfr_mg_solve(lfr_mg* mg, void* x_d, const void* b_d, int ncycles, double* res_h);

/* Banded column-operator geometry (columnwise_operator_type's alpha, beta,
 * gamma_m, gamma_p, bandwidth, nrow, ncol and its column_banded_dofmap /
 * indirection_dofmap arrays, infrastructure/source/operator/
 * columnwise_operator_mod.f90:168-191,350-396), uploaded once.  banded_*_h are
 * column_banded_dofmap(ndf, nlayers) in Fortran order, ind_*_h
 * indirection_dofmap(nrow|ncol); all 1-based.  Every (row, column) entry the
 * maps produce must lie inside the band. */
typedef struct lfr_cma lfr_cma;
int lfr_cma_create(lfr_ctx* ctx, int nlayers, int nrow, int ncol, int bandwidth, int alpha,
                   int beta, int gamma_m, int gamma_p, int ndf_to, const int32_t* banded_to_h,
                   const int32_t* ind_to_h, int ndf_from, const int32_t* banded_from_h,
                   const int32_t* ind_from_h, lfr_cma** cma);
int lfr_cma_destroy(lfr_cma* cma);

/* invoke(columnwise_op_asm_kernel_type(cma_op, lma_op)) over all columns:
 * columnwise_matrix(bandwidth, nrow, ncell_2d) assembled from
 * local_stencil(ncell_3d, ndf_to, ndf_from) (components/science/source/kernel/
 * columnwise_operators/sci_columnwise_op_asm_kernel_mod.F90:72-128), bitwise in
 * the reference's summation order. */
int lfr_columnwise_op_asm(lfr_ctx* ctx, int prec, const lfr_cma* cma, const void* local_stencil_d,
                          void* cm_d, int ncell_2d);

/* invoke(columnwise_op_app_kernel_type(lhs, x, cma_op)) over all columns:
 * lhs += A x with A banded (components/science/source/kernel/
 * columnwise_operators/sci_columnwise_op_app_kernel_mod.F90:73-124); the band
 * limits are formed in the operator's real kind as there. */
int lfr_columnwise_op_app(lfr_ctx* ctx, int prec, const lfr_cma* cma, void* lhs_d, const void* x_d,
                          const void* cm_d, const lfr_space* to_space, const lfr_space* from_space);

/* ---- multi-right-hand-side LMA apply on MFMA -----------------------------
 * 16 fields applied together: y_r (+)= A x_r.  Fields are multidata fields in
 * LFRic's "ndata first" layout with ndata = 16 (element (dof d, rhs r) at
 * d*16 + r, d the single-field dof index), FP32, 16-byte aligned.  The
 * operator is the LMA local_stencil(ncell_3d, ndf_to, ndf_from) of
 * infrastructure/source/operator/operator_mod.t90 (FP32 or FP64, prec_in 4/8),
 * repacked at creation into the handle's own layout: mprec 4 keeps FP32 and
 * runs FP32 MFMA; mprec 2 stores BF16 and rounds the inputs to BF16, with
 * FP32 accumulation.  The colouring (same arrays as lfr_plan_create) orders
 * the writes of a continuous output space; it may be NULL for a
 * discontinuous one.  Semantics per field are those of matrix_vector_code
 * (components/science/source/kernel/algebra/matrix_vector_kernel_mod.F90:68-140,
 * overwrite = 0) or setval_c(y, 0) followed by it (overwrite = 1); the
 * summation order is the MFMA's, so results are not bitwise those of the
 * single-field kernels.  No counterpart in the reference (which has no
 * multi-field LMA apply).  Shapes with kernels: W2 k=1 (36x36), W0 k=1
 * (27x27), W3 k=1 (8x8), W3<-W2 k=1 (8x36), W0 k=0 (8x8), W2 k=0 (6x6). */
typedef struct lfr_multi lfr_multi;
int lfr_multi_create(lfr_ctx* ctx, int mprec, int prec_in, const void* matrix_d,
                     const lfr_space* to_space, const lfr_space* from_space,
                     const int32_t* cells_in_colour, const int32_t* ncells_per_colour,
                     int ncolour, int maxcells, lfr_multi** out);
int lfr_multi_destroy(lfr_multi* h);
int lfr_multi_info(const lfr_multi* h, int64_t* packed_bytes, int* slots, int* vertical_pairs,
                   int* launches);
int lfr_matrix_vector_multi(lfr_ctx* ctx, lfr_multi* h, void* lhs_d, const void* x_d, int nrhs,
                            int overwrite);

/* Sixteen simultaneous solves A x_r = b_r by FP32 CG on the multi-field
 * operator A32 (lfr_matrix_vector_multi) inside an FP64 refinement loop
 * whose residuals use the single-field FP64 operator A64 (same matrix).
 * x64_d, b64_d: FP64 multidata fields (undf, 16); x holds the initial guess.
 * Each inner solve starts from zero and freezes a field once its residual
 * falls by inner_r_tol; the loop ends when ||b_r - A x_r|| <= r_tol ||b_r||
 * for every field or after max_outer corrections.  stats->r_nrm is the
 * largest final relative residual, stats->b_nrm the largest ||b_r||;
 * rel_res_h (optional) receives the 16 final relative residuals.  The inner
 * iteration is conjugate_gradient_type%cg_solve
 * (components/science/source/solver/sci_iterative_solver_mod.f90:495-592) per
 * field; the batching has no counterpart in the reference.  A64 must be a
 * kind-0 FP64 operator without apply_bc.  Returns LFR_ERR_NOT_CONVERGED (with
 * stats filled) when some field has not reached r_tol after max_outer
 * corrections, as lfr_refined_cg does. */
int lfr_refined_cg_multi(lfr_ctx* ctx, lfr_multi* A32, const lfr_lma_op* A64, double r_tol,
                         int max_outer, double inner_r_tol, int inner_max_iter, void* x64_d,
                         const void* b64_d, int nrhs, lfr_refine_stats* stats, double* rel_res_h);


/* ---- differentiable K-iteration CG -------------------------------------
 * x = CG_K(A, b): K unmonitored iterations of conjugate_gradient_type%cg_solve
 * (components/science/source/solver/sci_iterative_solver_mod.f90:495-592) from
 * x = 0 with the null preconditioner, and the exact reverse mode of that map
 * (alpha_i and beta_i depend on b, so it is nonlinear in b).  The forward
 * records the iteration's scalars and the operator products q_i = A p_i, and
 * keeps r_i, p_i only at the start of every `segment` iterations (default
 * ceil(sqrt(K)), at most 64) and through the last segment; the reverse sweep
 * recomputes the iterates of each segment from its checkpoint (vector work
 * only, the same bits) and walks it backwards with one transpose apply per
 * step.  Scalars stay on the device; no host synchronisation.  No
 * counterpart in the reference (which has no adjoint of its solvers).
 *   A        square kind-0 operator without apply_bc (packed copy optional)
 *   plan_t   kind-1 plan of A's space for A^T (or packed_t, a kind-1 pack)
 *   record   0: forward only (no tape memory kept for a reverse sweep) */
typedef struct lfr_cg_tape lfr_cg_tape;
int lfr_cg_tape_create(lfr_ctx* ctx, const lfr_lma_op* A, const lfr_plan* plan_t,
                       const lfr_lma_packed* packed_t, int iters, int segment, int record,
                       lfr_cg_tape** tape);
int lfr_cg_tape_destroy(lfr_cg_tape* tape);
int lfr_cg_tape_info(const lfr_cg_tape* tape, int64_t* bytes, int* segment);
/* x = CG_K(A, b), recording the tape. */
int lfr_cg_tape_forward(lfr_ctx* ctx, lfr_cg_tape* tape, const void* b_d, void* x_d);
/* bbar = (dx_K/db)^T xbar for the last forward's b; with Mbar_d != NULL also
 * the gradient with respect to local_stencil, sum_i qbar_i (x) p_i per
 * cell-layer (accumulate != 0: added to Mbar_d). */
int lfr_cg_tape_reverse(lfr_ctx* ctx, lfr_cg_tape* tape, const void* b_d, const void* xbar_d,
                        void* bbar_d, void* Mbar_d, int accumulate);

/* ---------------------------------------------------------------------- */
/* n-field vectors, abstract operators and preconditioners, and the        */
/* iterative solvers of sci_iterative_solver_mod over them                 */
/* ---------------------------------------------------------------------- */

#define LFR_MAX_FIELDS 8

/* Device image of a field_vector_type (r_def, prec 8) or
 * r_solver_field_vector_type (prec 4) of nfield fields
 * (infrastructure/source/field/vector_mod.F90:17-190;
 * components/science/source/algorithm/solver/sci_field_vector_mod.x90:20-56):
 * the fields lie end to end in one device buffer, field f (0-based) starting
 * at element undf[0] + ... + undf[f-1].  The elementwise methods (axpy, aypx,
 * axpby, scale, set_scalar, copy) are the lfr_vec_* calls over the whole
 * buffer; the reductions below combine per-field inner products in field
 * order, as field_vector_type does. */
typedef struct lfr_vector_layout {
  int prec;                        /* 4 (r_solver) or 8 (r_def)              */
  int nfield;                      /* 1 .. LFR_MAX_FIELDS                    */
  int64_t undf[LFR_MAX_FIELDS];    /* length of each field                   */
} lfr_vector_layout;

/* vector_size: the number of fields (sci_field_vector_mod.x90:341-347). */
int lfr_vector_size(const lfr_vector_layout* layout);
/* Total number of elements (the length of the device buffer). */
int64_t lfr_vector_length(const lfr_vector_layout* layout);
/* dot: sum over fields, in field order, of X_innerproduct_Y of each field
 * (sci_field_vector_mod.x90:85-107); FP64 accumulation. */
int lfr_vector_dot(lfr_ctx* ctx, const lfr_vector_layout* layout, const void* x_d, const void* y_d,
                   double* out_h);
/* norm = sqrt(sum over fields of X_innerproduct_X) (:59-73). */
int lfr_vector_norm(lfr_ctx* ctx, const lfr_vector_layout* layout, const void* x_d, double* out_h);
/* field_norm(n): sqrt(X_innerproduct_X) of field n, 1-based (:75-83). */
int lfr_vector_field_norm(lfr_ctx* ctx, const lfr_vector_layout* layout, const void* x_d, int n,
                          double* out_h);

/* A caller's operator or preconditioner: y = A x (or y = P^{-1} x) on the
 * vectors of the given layout, queued on `hip_stream` (or completed before
 * returning).  Returns 0 on success; any other value aborts the solve with
 * LFR_ERR_ARG.  It may call any lfr_* entry point on the same context except
 * lfr_iterative_solve. */
typedef int (*lfr_apply_fn)(void* user, const void* x_d, void* y_d, void* hip_stream);

/* abstract_linear_operator_type (components/science/source/solver/
 * sci_linear_operator_mod.f90:18-35): an object with apply(x, y), y = A x. */
typedef struct lfr_linop lfr_linop;

/* mass_matrix_operator_type (components/science/source/algorithm/solver/
 * sci_mass_matrix_operator_alg_mod.x90:72-119): y = M x on one field, with
 * enforce_bc after the apply when A->apply_bc; packed copies are used when
 * A->packed is set. */
int lfr_linop_lma(lfr_ctx* ctx, const lfr_lma_op* A, lfr_linop** out);

/* One block of a block operator: y_row += A x_col. */
typedef struct lfr_block_entry {
  int row, col;                    /* 0-based fields of y and x              */
  const void* matrix_d;            /* local_stencil(ncell_3d, ndf_to, ndf_from)
                                      of kind layout->prec                   */
  const lfr_space* to_space;       /* space of field `row`                   */
  const lfr_space* from_space;     /* space of field `col`                   */
  const lfr_plan* plan;            /* continuous output: a plan of to_space,
                                      kind 0 (kind 1 with transpose); NULL:
                                      discontinuous output (dg_inc_matrix_vector) */
  int transpose;                   /* 1: y_row += A^T x_col
                                      (transpose_matrix_vector; the matrix is
                                      stored (from_space, to_space)^T, i.e.
                                      local_stencil(ncell_3d, ndf_from, ndf_to)) */
} lfr_block_entry;

/* Block operator on n-field vectors: y = 0, then every entry adds its
 * invoke in the order given (matrix_vector for a continuous output,
 * dg_inc_matrix_vector for a discontinuous one, transpose_matrix_vector with
 * transpose).  dense_operator_type (components/science/source/algorithm/
 * solver/sci_dense_operator_alg_mod.x90:147-182) is the two-entry case
 * {(0, 0, W0 mass, plan), (1, 1, W3 mass, NULL)}.  A field of y with a single
 * entry is written by that entry's overwrite form (same values, y not read). */
int lfr_linop_block(lfr_ctx* ctx, const lfr_vector_layout* layout, int nentry,
                    const lfr_block_entry* entries, lfr_linop** out);

/* An operator supplied by the caller (the op_cb of a solver call). */
int lfr_linop_callback(lfr_ctx* ctx, const lfr_vector_layout* layout, lfr_apply_fn fn, void* user,
                       lfr_linop** out);
int lfr_linop_destroy(lfr_linop* op);
/* y = A x (queued). */
int lfr_linop_apply(lfr_ctx* ctx, const lfr_linop* op, const void* x_d, void* y_d);

/* abstract_preconditioner_type (components/science/source/solver/
 * sci_preconditioner_mod.f90:19-37): apply(x, y), y = P^{-1} x. */
typedef struct lfr_precon lfr_precon;
/* null_preconditioner_type: y = x (sci_null_preconditioner_alg_mod.x90:73-95). */
int lfr_precon_null(lfr_ctx* ctx, const lfr_vector_layout* layout, lfr_precon** out);
/* diagonal_preconditioner_type: y = x / diag, elementwise over the whole
 * vector (sci_diagonal_preconditioner_alg_mod.x90:71-106); diag_d has the
 * vector's layout and must stay valid while the preconditioner is used. */
int lfr_precon_diagonal(lfr_ctx* ctx, const lfr_vector_layout* layout, const void* diag_d,
                        lfr_precon** out);
/* Column solve: y = columnwise_op_appinv(x) on a one-field vector of `space`
 * (as lfr_columnwise_op_appinv: bandwidth 3, nrow rows per column). */
int lfr_precon_columnwise(lfr_ctx* ctx, const lfr_vector_layout* layout, const void* cm_d, int nrow,
                          const lfr_space* space, const int32_t* indirection_d, int fp64_pivots,
                          lfr_precon** out);
int lfr_precon_callback(lfr_ctx* ctx, const lfr_vector_layout* layout, lfr_apply_fn fn, void* user,
                        lfr_precon** out);
int lfr_precon_destroy(lfr_precon* p);
/* y = P^{-1} x (queued). */
int lfr_precon_apply(lfr_ctx* ctx, const lfr_precon* p, const void* x_d, void* y_d);

/* The solver types of components/science/source/solver/sci_iterative_solver_mod.f90. */
enum {
  LFR_SOLVER_CG = 0,               /* conjugate_gradient_type  (:495-592)   */
  LFR_SOLVER_BICGSTAB = 1,         /* BiCGstab_type            (:645-779)   */
  LFR_SOLVER_GMRES = 2,            /* gmres_type               (:830-1003)  */
  LFR_SOLVER_FGMRES = 3,           /* fgmres_type              (:1055-1237) */
  LFR_SOLVER_GCR = 4,              /* gcr_type                 (:1288-1408) */
  LFR_SOLVER_PRECONDITION_ONLY = 5,/* precondition_only_type   (:1664-1694) */
  LFR_SOLVER_JACOBI = 6,           /* jacobi_type              (:1745-1815) */
  LFR_SOLVER_CHEBYSHEV = 7         /* chebyshev_type           (:1868-1940) */
};

/* Constructor arguments of those types. */
typedef struct lfr_iterative_solver {
  int method;                      /* LFR_SOLVER_*                           */
  double r_tol, a_tol;
  int max_iter;                    /* iterations (GMRES, FGMRES, GCR: restarts) */
  int monitor_convergence;
  int fail_on_non_converged;
  int gcrk;                        /* Krylov vectors per restart (GMRES, FGMRES, GCR) */
  double rho_relax;                /* Jacobi relaxation factor               */
  double lmin, lmax;               /* Chebyshev eigenvalue bounds            */
} lfr_iterative_solver;

/* solver%apply(x, b) with solver = the type named by s->method over lin_op A
 * and preconditioner P; x, b: vectors of A's layout, x holds the initial
 * guess (FGMRES and GCR add their correction to x; Chebyshev starts from
 * zero; precondition_only overwrites x), exactly as the reference routines.
 * Scalars are FP64 and stay on the device; they are cast to the vector kind
 * before each update.  With monitor_convergence the residual norm is read
 * back once per check.  stats->iterations counts iterations (GMRES, FGMRES:
 * restarts; GCR: inner steps); stats->r_nrm is the last residual norm the
 * method computed.  Returns LFR_ERR_NOT_CONVERGED when monitoring,
 * fail_on_non_converged and not converged, LFR_ERR_NAN on a NaN residual. */
int lfr_iterative_solve(lfr_ctx* ctx, const lfr_iterative_solver* s, const lfr_linop* A,
                        const lfr_precon* P, void* x_d, const void* b_d, lfr_solver_stats* stats);

/* Measurement aid, no counterpart in the reference: the algorithmic HBM bytes
 * of the last lfr_iterative_solve on this context (every vector pass of the
 * method counted once per array it reads or writes, each operator apply as
 * its matrices plus its input and output fields, each preconditioner apply
 * as its copy, division or column solve), for rating a solve against the
 * memory bandwidth; -1 when a caller-written operator or preconditioner took
 * part (its traffic is unknown) or no solve has run. */
int lfr_iterative_solve_bytes(lfr_ctx* ctx, double* bytes);

/* invoke(mm_diagonal_kernel_type(diag, mass_matrix)): diag(map(df)+k) +=
 * mass_matrix(ik, df, df) over all cells (components/science/source/kernel/
 * algebra/sci_mm_diagonal_kernel_mod.F90), in the colour order of `plan`
 * (any plan of `space`) for a continuous space, or plan = NULL for a
 * discontinuous one; as used by dense_operator_type%extract_diagonal. */
int lfr_mm_diagonal(lfr_ctx* ctx, int prec, void* diag_d, const void* matrix_d, const lfr_space* space,
                    const lfr_plan* plan);

/* ---------------------------------------------------------------------------
 * Autotuner (no counterpart in the reference, whose kernels have one fixed
 * loop form).  Kernels with several result-identical launch shapes (column
 * solves: columns a wave and rows a stage; packed apply: output staging and
 * cache policy; multigrid smoother: workgroup size) time every shape on the
 * first call of a problem size and keep the fastest under a key naming the
 * kernel family and the shape ("appinv f32 nrow=70 ncell=55296", ...).
 * LFR_TUNE=0 disables timing (defaults are used); LFR_TUNE_FILE names a
 * file loaded by lfr_init and written by lfr_finalise.  Choices never change
 * results, only speed.
 * ------------------------------------------------------------------------- */
/* Store `choice` for `key` (used from then on, also by graph captures). */
int lfr_tune_set(lfr_ctx* ctx, const char* key, int choice);
/* The stored choice for `key` (LFR_ERR_ARG if none). */
int lfr_tune_get(lfr_ctx* ctx, const char* key, int* choice);
/* Number of stored choices, and entry i (0-based, key order). */
int lfr_tune_count(const lfr_ctx* ctx, int* n);
int lfr_tune_entry(lfr_ctx* ctx, int i, char* key, int keylen, int* choice);
/* Forget every stored choice. */
int lfr_tune_clear(lfr_ctx* ctx);
/* Write / read the stored choices as "choice key" lines. */
int lfr_tune_save(lfr_ctx* ctx, const char* path);
int lfr_tune_load(lfr_ctx* ctx, const char* path);

#ifdef __cplusplus
}
#endif
#endif /* LFR_H */
