!-----------------------------------------------------------------------------
! iso_c_binding interface to liblfr.so (include/lfr.h) for a PSy layer.
!
! A generated invoke such as
!   invoke( setval_c(y, 0.0), matrix_vector_kernel_type(y, x, mm) )
! (components/science/source/algorithm/solver/sci_mass_matrix_operator_alg_mod.x90:98-99)
! becomes one call of lfr_matrix_vector with overwrite = 1 on device copies of
! y, x and mm%local_stencil; the dofmaps and the colour map are uploaded once
! per function space / mesh (lfr_space_create, lfr_plan_create).  On a nonzero
! return the caller reports lfr_last_error through log_event(..., LOG_LEVEL_ERROR),
! which is how LFRic infrastructure aborts on errors.
!-----------------------------------------------------------------------------
module lfr_c_binding_mod

  use, intrinsic :: iso_c_binding, only : c_int, c_int32_t, c_int64_t, c_double, &
                                          c_ptr, c_funptr, c_size_t, c_char,     &
                                          c_null_char, c_null_ptr, c_associated, &
                                          c_f_pointer

  implicit none

  private

  integer(c_int), parameter, public :: LFR_OK = 0
  integer(c_int), parameter, public :: LFR_PREC_R_SINGLE = 4
  integer(c_int), parameter, public :: LFR_PREC_R_DOUBLE = 8

  integer(c_int), parameter, public :: LFR_MAX_FIELDS = 8
  ! solver types (the lfr_iterative_solver%method values)
  integer(c_int), parameter, public :: LFR_SOLVER_CG = 0, LFR_SOLVER_BICGSTAB = 1,         &
                                       LFR_SOLVER_GMRES = 2, LFR_SOLVER_FGMRES = 3,        &
                                       LFR_SOLVER_GCR = 4, LFR_SOLVER_PRECONDITION_ONLY = 5, &
                                       LFR_SOLVER_JACOBI = 6, LFR_SOLVER_CHEBYSHEV = 7

  !> A square LMA operator for the solvers (mass_matrix_operator_type).
  type, bind(C), public :: lfr_lma_op
    integer(c_int) :: prec = 8
    type(c_ptr)    :: matrix_d = c_null_ptr
    type(c_ptr)    :: space = c_null_ptr
    type(c_ptr)    :: plan = c_null_ptr
    integer(c_int) :: apply_bc = 0
    type(c_ptr)    :: packed = c_null_ptr
  end type lfr_lma_op

  !> Constructor arguments of conjugate_gradient_type / gmres_type.
  type, bind(C), public :: lfr_solver_params
    real(c_double) :: r_tol = 1.0e-8_c_double
    real(c_double) :: a_tol = 0.0_c_double
    integer(c_int) :: max_iter = 100
    integer(c_int) :: monitor_convergence = 1
    integer(c_int) :: fail_on_non_converged = 0
    integer(c_int) :: preconditioner = 0
    type(c_ptr)    :: diag_d = c_null_ptr
    integer(c_int) :: gcrk = 0
    integer(c_int) :: x_is_zero = 0
  end type lfr_solver_params

  type, bind(C), public :: lfr_solver_stats
    integer(c_int) :: iterations = 0
    integer(c_int) :: converged = 0
    real(c_double) :: r_nrm0 = 0.0_c_double
    real(c_double) :: r_nrm = 0.0_c_double
  end type lfr_solver_stats

  type, bind(C), public :: lfr_refine_stats
    integer(c_int) :: outer = 0
    integer(c_int) :: inner_total = 0
    integer(c_int) :: converged = 0
    real(c_double) :: b_nrm = 0.0_c_double
    real(c_double) :: r_nrm = 0.0_c_double
  end type lfr_refine_stats

  !> Fields of an n-field vector laid end to end (field_vector_type).
  type, bind(C), public :: lfr_vector_layout
    integer(c_int)     :: prec = 8
    integer(c_int)     :: nfield = 1
    integer(c_int64_t) :: undf(8) = 0
  end type lfr_vector_layout

  !> One block y_row += A x_col of a block operator (0-based fields).
  type, bind(C), public :: lfr_block_entry
    integer(c_int) :: row = 0, col = 0
    type(c_ptr)    :: matrix_d = c_null_ptr
    type(c_ptr)    :: to_space = c_null_ptr
    type(c_ptr)    :: from_space = c_null_ptr
    type(c_ptr)    :: plan = c_null_ptr
    integer(c_int) :: transpose = 0
  end type lfr_block_entry

  !> Constructor arguments of the solver types of sci_iterative_solver_mod.
  type, bind(C), public :: lfr_iterative_solver
    integer(c_int) :: method = 0
    real(c_double) :: r_tol = 1.0e-8_c_double
    real(c_double) :: a_tol = 0.0_c_double
    integer(c_int) :: max_iter = 100
    integer(c_int) :: monitor_convergence = 1
    integer(c_int) :: fail_on_non_converged = 0
    integer(c_int) :: gcrk = 10
    real(c_double) :: rho_relax = 1.0_c_double
    real(c_double) :: lmin = 0.0_c_double
    real(c_double) :: lmax = 1.0_c_double
  end type lfr_iterative_solver

  public :: lfr_init, lfr_finalise, lfr_last_error, lfr_synchronize, lfr_malloc, lfr_free, &
            lfr_memcpy_h2d, lfr_memcpy_d2h, lfr_set_stream, lfr_space_create,             &
            lfr_space_destroy, lfr_space_set_boundary, lfr_plan_create, lfr_plan_destroy, &
            lfr_matrix_vector, lfr_dg_matrix_vector, lfr_dg_inc_matrix_vector,            &
            lfr_transpose_matrix_vector, lfr_lma_pack, lfr_lma_packed_destroy,           &
            lfr_matrix_vector_packed, lfr_enforce_bc, lfr_mm_diagonal,                   &
            lfr_columnwise_op_appinv, lfr_cma_create, lfr_cma_destroy,                  &
            lfr_columnwise_op_asm, lfr_columnwise_op_app, lfr_tri_solve,                &
            lfr_operator_tri_solve, lfr_vec_dot, lfr_vec_axpy, lfr_vec_aypx,             &
            lfr_vec_axpby, lfr_vec_scale, lfr_vec_copy, lfr_vec_setval, lfr_vec_divide,  &
            lfr_cg, lfr_gmres, lfr_refined_cg, lfr_vector_dot, lfr_vector_norm,         &
            lfr_vector_field_norm, lfr_linop_lma, lfr_linop_block, lfr_linop_callback,    &
            lfr_linop_destroy, lfr_linop_apply, lfr_precon_null, lfr_precon_diagonal,    &
            lfr_precon_columnwise, lfr_precon_callback, lfr_precon_destroy,            &
            lfr_precon_apply, lfr_iterative_solve, lfr_iterative_solve_bytes,          &
            lfr_apply_fn, lfr_error_message,                                           &
            lfr_tune_set, lfr_tune_get, lfr_tune_clear, lfr_tune_save, lfr_tune_load

  interface

    integer(c_int) function lfr_init(device, ctx) bind(C, name="lfr_init")
      import :: c_int, c_ptr
      integer(c_int), value :: device
      type(c_ptr), intent(out) :: ctx
    end function lfr_init

    integer(c_int) function lfr_finalise(ctx) bind(C, name="lfr_finalise")
      import :: c_int, c_ptr
      type(c_ptr), value :: ctx
    end function lfr_finalise

    type(c_ptr) function lfr_last_error(ctx) bind(C, name="lfr_last_error")
      import :: c_ptr
      type(c_ptr), value :: ctx
    end function lfr_last_error

    integer(c_int) function lfr_synchronize(ctx) bind(C, name="lfr_synchronize")
      import :: c_int, c_ptr
      type(c_ptr), value :: ctx
    end function lfr_synchronize

    integer(c_int) function lfr_malloc(ctx, bytes, ptr) bind(C, name="lfr_malloc")
      import :: c_int, c_ptr, c_size_t
      type(c_ptr), value :: ctx
      integer(c_size_t), value :: bytes
      type(c_ptr), intent(out) :: ptr
    end function lfr_malloc

    integer(c_int) function lfr_free(ctx, ptr) bind(C, name="lfr_free")
      import :: c_int, c_ptr
      type(c_ptr), value :: ctx, ptr
    end function lfr_free

    integer(c_int) function lfr_memcpy_h2d(ctx, dst, src, bytes) bind(C, name="lfr_memcpy_h2d")
      import :: c_int, c_ptr, c_size_t
      type(c_ptr), value :: ctx, dst, src
      integer(c_size_t), value :: bytes
    end function lfr_memcpy_h2d

    integer(c_int) function lfr_memcpy_d2h(ctx, dst, src, bytes) bind(C, name="lfr_memcpy_d2h")
      import :: c_int, c_ptr, c_size_t
      type(c_ptr), value :: ctx, dst, src
      integer(c_size_t), value :: bytes
    end function lfr_memcpy_d2h

    ! dofmap(ndf, ncell) from function_space_type%get_whole_dofmap()
    integer(c_int) function lfr_space_create(ctx, dofmap, ndf, ncell, nlayers, undf, space) &
        bind(C, name="lfr_space_create")
      import :: c_int, c_int32_t, c_ptr
      type(c_ptr), value :: ctx
      integer(c_int32_t), intent(in) :: dofmap(*)
      integer(c_int), value :: ndf, ncell, nlayers, undf
      type(c_ptr), intent(out) :: space
    end function lfr_space_create

    integer(c_int) function lfr_space_destroy(space) bind(C, name="lfr_space_destroy")
      import :: c_int, c_ptr
      type(c_ptr), value :: space
    end function lfr_space_destroy

    ! cells_in_colour(ncolour, maxcells) from mesh%get_colour_map()
    integer(c_int) function lfr_plan_create(ctx, space, cells_in_colour, ncells_per_colour, &
                                            ncolour, maxcells, kind, plan)            &
        bind(C, name="lfr_plan_create")
      import :: c_int, c_int32_t, c_ptr
      type(c_ptr), value :: ctx, space
      integer(c_int32_t), intent(in) :: cells_in_colour(*), ncells_per_colour(*)
      integer(c_int), value :: ncolour, maxcells, kind
      type(c_ptr), intent(out) :: plan
    end function lfr_plan_create

    integer(c_int) function lfr_plan_destroy(plan) bind(C, name="lfr_plan_destroy")
      import :: c_int, c_ptr
      type(c_ptr), value :: plan
    end function lfr_plan_destroy

    integer(c_int) function lfr_matrix_vector(ctx, prec, lhs, x, matrix, to_space, from_space, &
                                              plan, overwrite) bind(C, name="lfr_matrix_vector")
      import :: c_int, c_ptr
      type(c_ptr), value :: ctx, lhs, x, matrix, to_space, from_space, plan
      integer(c_int), value :: prec, overwrite
    end function lfr_matrix_vector

    integer(c_int) function lfr_dg_matrix_vector(ctx, prec, lhs, x, matrix, to_space, from_space) &
        bind(C, name="lfr_dg_matrix_vector")
      import :: c_int, c_ptr
      type(c_ptr), value :: ctx, lhs, x, matrix, to_space, from_space
      integer(c_int), value :: prec
    end function lfr_dg_matrix_vector

    integer(c_int) function lfr_dg_inc_matrix_vector(ctx, prec, lhs, x, matrix, to_space, &
                                                     from_space)                      &
        bind(C, name="lfr_dg_inc_matrix_vector")
      import :: c_int, c_ptr
      type(c_ptr), value :: ctx, lhs, x, matrix, to_space, from_space
      integer(c_int), value :: prec
    end function lfr_dg_inc_matrix_vector

    integer(c_int) function lfr_transpose_matrix_vector(ctx, prec, lhs, x, matrix, lhs_space, &
                                                        x_space, plan, overwrite)          &
        bind(C, name="lfr_transpose_matrix_vector")
      import :: c_int, c_ptr
      type(c_ptr), value :: ctx, lhs, x, matrix, lhs_space, x_space, plan
      integer(c_int), value :: prec, overwrite
    end function lfr_transpose_matrix_vector

    ! packed copy of an operator applied many times (e.g. a mass matrix)
    integer(c_int) function lfr_lma_pack(ctx, prec, matrix, x_space, plan, packed) &
        bind(C, name="lfr_lma_pack")
      import :: c_int, c_ptr
      type(c_ptr), value :: ctx, matrix, x_space, plan
      integer(c_int), value :: prec
      type(c_ptr), intent(out) :: packed
    end function lfr_lma_pack

    integer(c_int) function lfr_lma_packed_destroy(packed) bind(C, name="lfr_lma_packed_destroy")
      import :: c_int, c_ptr
      type(c_ptr), value :: packed
    end function lfr_lma_packed_destroy

    integer(c_int) function lfr_matrix_vector_packed(ctx, packed, lhs, x, overwrite) &
        bind(C, name="lfr_matrix_vector_packed")
      import :: c_int, c_ptr
      type(c_ptr), value :: ctx, packed, lhs, x
      integer(c_int), value :: overwrite
    end function lfr_matrix_vector_packed

    integer(c_int) function lfr_columnwise_op_appinv(ctx, prec, lhs, x, cm, bandwidth, nrow,  &
                                                     ncell_2d, space, indirection, fp64_pivots) &
        bind(C, name="lfr_columnwise_op_appinv")
      import :: c_int, c_ptr
      type(c_ptr), value :: ctx, lhs, x, cm, space, indirection
      integer(c_int), value :: prec, bandwidth, nrow, ncell_2d, fp64_pivots
    end function lfr_columnwise_op_appinv

    integer(c_int) function lfr_vec_dot(ctx, prec, n, x, y, res) bind(C, name="lfr_vec_dot")
      import :: c_int, c_int64_t, c_double, c_ptr
      type(c_ptr), value :: ctx, x, y
      integer(c_int), value :: prec
      integer(c_int64_t), value :: n
      real(c_double), intent(out) :: res
    end function lfr_vec_dot

    integer(c_int) function lfr_vec_axpy(ctx, prec, n, a, x, y) bind(C, name="lfr_vec_axpy")
      import :: c_int, c_int64_t, c_double, c_ptr
      type(c_ptr), value :: ctx, x, y
      integer(c_int), value :: prec
      integer(c_int64_t), value :: n
      real(c_double), value :: a
    end function lfr_vec_axpy

    integer(c_int) function lfr_vec_setval(ctx, prec, n, x, c) bind(C, name="lfr_vec_setval")
      import :: c_int, c_int64_t, c_double, c_ptr
      type(c_ptr), value :: ctx, x
      integer(c_int), value :: prec
      integer(c_int64_t), value :: n
      real(c_double), value :: c
    end function lfr_vec_setval

    integer(c_int) function lfr_set_stream(ctx, stream) bind(C, name="lfr_set_stream")
      import :: c_int, c_ptr
      type(c_ptr), value :: ctx, stream
    end function lfr_set_stream

    ! boundary_value(ndf, 2) of function_space_type%get_boundary_dofs()
    integer(c_int) function lfr_space_set_boundary(space, boundary_value) &
        bind(C, name="lfr_space_set_boundary")
      import :: c_int, c_int32_t, c_ptr
      type(c_ptr), value :: space
      integer(c_int32_t), intent(in) :: boundary_value(*)
    end function lfr_space_set_boundary

    integer(c_int) function lfr_enforce_bc(ctx, prec, field, space) bind(C, name="lfr_enforce_bc")
      import :: c_int, c_ptr
      type(c_ptr), value :: ctx, field, space
      integer(c_int), value :: prec
    end function lfr_enforce_bc

    ! invoke(mm_diagonal_kernel_type(diag, matrix)); plan = c_null_ptr for a
    ! discontinuous space
    integer(c_int) function lfr_mm_diagonal(ctx, prec, diag, matrix, space, plan) &
        bind(C, name="lfr_mm_diagonal")
      import :: c_int, c_ptr
      type(c_ptr), value :: ctx, diag, matrix, space, plan
      integer(c_int), value :: prec
    end function lfr_mm_diagonal

    ! columnwise_operator_type geometry (banded / indirection dofmaps, 1-based)
    integer(c_int) function lfr_cma_create(ctx, nlayers, nrow, ncol, bandwidth, alpha, beta,     &
                                           gamma_m, gamma_p, ndf_to, banded_to, ind_to,         &
                                           ndf_from, banded_from, ind_from, cma)                &
        bind(C, name="lfr_cma_create")
      import :: c_int, c_int32_t, c_ptr
      type(c_ptr), value :: ctx
      integer(c_int), value :: nlayers, nrow, ncol, bandwidth, alpha, beta, gamma_m, gamma_p, &
                               ndf_to, ndf_from
      integer(c_int32_t), intent(in) :: banded_to(*), ind_to(*), banded_from(*), ind_from(*)
      type(c_ptr), intent(out) :: cma
    end function lfr_cma_create

    integer(c_int) function lfr_cma_destroy(cma) bind(C, name="lfr_cma_destroy")
      import :: c_int, c_ptr
      type(c_ptr), value :: cma
    end function lfr_cma_destroy

    integer(c_int) function lfr_columnwise_op_asm(ctx, prec, cma, local_stencil, cm, ncell_2d) &
        bind(C, name="lfr_columnwise_op_asm")
      import :: c_int, c_ptr
      type(c_ptr), value :: ctx, cma, local_stencil, cm
      integer(c_int), value :: prec, ncell_2d
    end function lfr_columnwise_op_asm

    integer(c_int) function lfr_columnwise_op_app(ctx, prec, cma, lhs, x, cm, to_space, from_space) &
        bind(C, name="lfr_columnwise_op_app")
      import :: c_int, c_ptr
      type(c_ptr), value :: ctx, cma, lhs, x, cm, to_space, from_space
      integer(c_int), value :: prec
    end function lfr_columnwise_op_app

    integer(c_int) function lfr_tri_solve(ctx, prec, y, x, tri_0, tri_plus, tri_minus, space) &
        bind(C, name="lfr_tri_solve")
      import :: c_int, c_ptr
      type(c_ptr), value :: ctx, y, x, tri_0, tri_plus, tri_minus, space
      integer(c_int), value :: prec
    end function lfr_tri_solve

    integer(c_int) function lfr_operator_tri_solve(ctx, prec, y, x, matrix, space) &
        bind(C, name="lfr_operator_tri_solve")
      import :: c_int, c_ptr
      type(c_ptr), value :: ctx, y, x, matrix, space
      integer(c_int), value :: prec
    end function lfr_operator_tri_solve

    integer(c_int) function lfr_vec_aypx(ctx, prec, n, a, x, y) bind(C, name="lfr_vec_aypx")
      import :: c_int, c_int64_t, c_double, c_ptr
      type(c_ptr), value :: ctx, x, y
      integer(c_int), value :: prec
      integer(c_int64_t), value :: n
      real(c_double), value :: a
    end function lfr_vec_aypx

    integer(c_int) function lfr_vec_axpby(ctx, prec, n, a, x, b, y, z) bind(C, name="lfr_vec_axpby")
      import :: c_int, c_int64_t, c_double, c_ptr
      type(c_ptr), value :: ctx, x, y, z
      integer(c_int), value :: prec
      integer(c_int64_t), value :: n
      real(c_double), value :: a, b
    end function lfr_vec_axpby

    integer(c_int) function lfr_vec_scale(ctx, prec, n, a, x) bind(C, name="lfr_vec_scale")
      import :: c_int, c_int64_t, c_double, c_ptr
      type(c_ptr), value :: ctx, x
      integer(c_int), value :: prec
      integer(c_int64_t), value :: n
      real(c_double), value :: a
    end function lfr_vec_scale

    integer(c_int) function lfr_vec_copy(ctx, prec, n, y, x) bind(C, name="lfr_vec_copy")
      import :: c_int, c_int64_t, c_ptr
      type(c_ptr), value :: ctx, y, x
      integer(c_int), value :: prec
      integer(c_int64_t), value :: n
    end function lfr_vec_copy

    integer(c_int) function lfr_vec_divide(ctx, prec, n, z, x, y) bind(C, name="lfr_vec_divide")
      import :: c_int, c_int64_t, c_ptr
      type(c_ptr), value :: ctx, z, x, y
      integer(c_int), value :: prec
      integer(c_int64_t), value :: n
    end function lfr_vec_divide

    ! conjugate_gradient_type%cg_solve over an LMA operator (fused path)
    integer(c_int) function lfr_cg(ctx, op, params, x, b, stats) bind(C, name="lfr_cg")
      import :: c_int, c_ptr, lfr_lma_op, lfr_solver_params, lfr_solver_stats
      type(c_ptr), value :: ctx, x, b
      type(lfr_lma_op), intent(in) :: op
      type(lfr_solver_params), intent(in) :: params
      type(lfr_solver_stats), intent(out) :: stats
    end function lfr_cg

    integer(c_int) function lfr_gmres(ctx, op, params, x, b, stats) bind(C, name="lfr_gmres")
      import :: c_int, c_ptr, lfr_lma_op, lfr_solver_params, lfr_solver_stats
      type(c_ptr), value :: ctx, x, b
      type(lfr_lma_op), intent(in) :: op
      type(lfr_solver_params), intent(in) :: params
      type(lfr_solver_stats), intent(out) :: stats
    end function lfr_gmres

    ! FP32 (r_solver) CG corrections inside an FP64 (r_def) residual loop
    integer(c_int) function lfr_refined_cg(ctx, op32, op64, r_tol, max_outer, inner, x64, b64, stats) &
        bind(C, name="lfr_refined_cg")
      import :: c_int, c_double, c_ptr, lfr_lma_op, lfr_solver_params, lfr_refine_stats
      type(c_ptr), value :: ctx, x64, b64
      type(lfr_lma_op), intent(in) :: op32, op64
      real(c_double), value :: r_tol
      integer(c_int), value :: max_outer
      type(lfr_solver_params), intent(in) :: inner
      type(lfr_refine_stats), intent(out) :: stats
    end function lfr_refined_cg

    ! field_vector_type%dot / %norm / %field_norm(n) on the device
    integer(c_int) function lfr_vector_dot(ctx, layout, x, y, res) bind(C, name="lfr_vector_dot")
      import :: c_int, c_double, c_ptr, lfr_vector_layout
      type(c_ptr), value :: ctx, x, y
      type(lfr_vector_layout), intent(in) :: layout
      real(c_double), intent(out) :: res
    end function lfr_vector_dot

    integer(c_int) function lfr_vector_norm(ctx, layout, x, res) bind(C, name="lfr_vector_norm")
      import :: c_int, c_double, c_ptr, lfr_vector_layout
      type(c_ptr), value :: ctx, x
      type(lfr_vector_layout), intent(in) :: layout
      real(c_double), intent(out) :: res
    end function lfr_vector_norm

    integer(c_int) function lfr_vector_field_norm(ctx, layout, x, n, res) &
        bind(C, name="lfr_vector_field_norm")
      import :: c_int, c_double, c_ptr, lfr_vector_layout
      type(c_ptr), value :: ctx, x
      type(lfr_vector_layout), intent(in) :: layout
      integer(c_int), value :: n
      real(c_double), intent(out) :: res
    end function lfr_vector_field_norm

    ! abstract_linear_operator_type objects
    integer(c_int) function lfr_linop_lma(ctx, op, linop) bind(C, name="lfr_linop_lma")
      import :: c_int, c_ptr, lfr_lma_op
      type(c_ptr), value :: ctx
      type(lfr_lma_op), intent(in) :: op
      type(c_ptr), intent(out) :: linop
    end function lfr_linop_lma

    integer(c_int) function lfr_linop_block(ctx, layout, nentry, entries, linop) &
        bind(C, name="lfr_linop_block")
      import :: c_int, c_ptr, lfr_vector_layout, lfr_block_entry
      type(c_ptr), value :: ctx
      type(lfr_vector_layout), intent(in) :: layout
      integer(c_int), value :: nentry
      type(lfr_block_entry), intent(in) :: entries(*)
      type(c_ptr), intent(out) :: linop
    end function lfr_linop_block

    integer(c_int) function lfr_linop_callback(ctx, layout, fn, user, linop) &
        bind(C, name="lfr_linop_callback")
      import :: c_int, c_ptr, c_funptr, lfr_vector_layout
      type(c_ptr), value :: ctx, user
      type(lfr_vector_layout), intent(in) :: layout
      type(c_funptr), value :: fn
      type(c_ptr), intent(out) :: linop
    end function lfr_linop_callback

    integer(c_int) function lfr_linop_destroy(linop) bind(C, name="lfr_linop_destroy")
      import :: c_int, c_ptr
      type(c_ptr), value :: linop
    end function lfr_linop_destroy

    integer(c_int) function lfr_linop_apply(ctx, linop, x, y) bind(C, name="lfr_linop_apply")
      import :: c_int, c_ptr
      type(c_ptr), value :: ctx, linop, x, y
    end function lfr_linop_apply

    ! abstract_preconditioner_type objects
    integer(c_int) function lfr_precon_null(ctx, layout, prec) bind(C, name="lfr_precon_null")
      import :: c_int, c_ptr, lfr_vector_layout
      type(c_ptr), value :: ctx
      type(lfr_vector_layout), intent(in) :: layout
      type(c_ptr), intent(out) :: prec
    end function lfr_precon_null

    integer(c_int) function lfr_precon_diagonal(ctx, layout, diag, prec) bind(C, name="lfr_precon_diagonal")
      import :: c_int, c_ptr, lfr_vector_layout
      type(c_ptr), value :: ctx, diag
      type(lfr_vector_layout), intent(in) :: layout
      type(c_ptr), intent(out) :: prec
    end function lfr_precon_diagonal

    integer(c_int) function lfr_precon_columnwise(ctx, layout, cm, nrow, space, indirection, &
                                                  fp64_pivots, prec)                       &
        bind(C, name="lfr_precon_columnwise")
      import :: c_int, c_ptr, lfr_vector_layout
      type(c_ptr), value :: ctx, cm, space, indirection
      type(lfr_vector_layout), intent(in) :: layout
      integer(c_int), value :: nrow, fp64_pivots
      type(c_ptr), intent(out) :: prec
    end function lfr_precon_columnwise

    integer(c_int) function lfr_precon_callback(ctx, layout, fn, user, prec) &
        bind(C, name="lfr_precon_callback")
      import :: c_int, c_ptr, c_funptr, lfr_vector_layout
      type(c_ptr), value :: ctx, user
      type(lfr_vector_layout), intent(in) :: layout
      type(c_funptr), value :: fn
      type(c_ptr), intent(out) :: prec
    end function lfr_precon_callback

    integer(c_int) function lfr_precon_destroy(prec) bind(C, name="lfr_precon_destroy")
      import :: c_int, c_ptr
      type(c_ptr), value :: prec
    end function lfr_precon_destroy

    integer(c_int) function lfr_precon_apply(ctx, prec, x, y) bind(C, name="lfr_precon_apply")
      import :: c_int, c_ptr
      type(c_ptr), value :: ctx, prec, x, y
    end function lfr_precon_apply

    ! solver%apply(x, b) for any solver type over any operator/preconditioner
    integer(c_int) function lfr_iterative_solve(ctx, solver, linop, prec, x, b, stats) &
        bind(C, name="lfr_iterative_solve")
      import :: c_int, c_ptr, lfr_iterative_solver, lfr_solver_stats
      type(c_ptr), value :: ctx, linop, prec, x, b
      type(lfr_iterative_solver), intent(in) :: solver
      type(lfr_solver_stats), intent(out) :: stats
    end function lfr_iterative_solve

    ! measurement aid: algorithmic HBM bytes of the last iterative solve on
    ! ctx (-1 when a caller-written operator or preconditioner took part)
    integer(c_int) function lfr_iterative_solve_bytes(ctx, bytes) bind(C, name="lfr_iterative_solve_bytes")
      import :: c_int, c_ptr, c_double
      type(c_ptr), value :: ctx
      real(c_double), intent(out) :: bytes
    end function lfr_iterative_solve_bytes

    ! autotuner: force / read a launch-shape choice, forget all, save / load
    ! them (keys and paths are NUL-terminated: trim(s)//c_null_char)
    integer(c_int) function lfr_tune_set(ctx, key, choice) bind(C, name="lfr_tune_set")
      import :: c_int, c_ptr, c_char
      type(c_ptr), value :: ctx
      character(kind=c_char), dimension(*), intent(in) :: key
      integer(c_int), value :: choice
    end function lfr_tune_set

    integer(c_int) function lfr_tune_get(ctx, key, choice) bind(C, name="lfr_tune_get")
      import :: c_int, c_ptr, c_char
      type(c_ptr), value :: ctx
      character(kind=c_char), dimension(*), intent(in) :: key
      integer(c_int), intent(out) :: choice
    end function lfr_tune_get

    integer(c_int) function lfr_tune_clear(ctx) bind(C, name="lfr_tune_clear")
      import :: c_int, c_ptr
      type(c_ptr), value :: ctx
    end function lfr_tune_clear

    integer(c_int) function lfr_tune_save(ctx, path) bind(C, name="lfr_tune_save")
      import :: c_int, c_ptr, c_char
      type(c_ptr), value :: ctx
      character(kind=c_char), dimension(*), intent(in) :: path
    end function lfr_tune_save

    integer(c_int) function lfr_tune_load(ctx, path) bind(C, name="lfr_tune_load")
      import :: c_int, c_ptr, c_char
      type(c_ptr), value :: ctx
      character(kind=c_char), dimension(*), intent(in) :: path
    end function lfr_tune_load

  end interface

  !> Interface of a caller-written operator or preconditioner (c_funloc of a
  !> bind(C) function with this interface goes to lfr_linop_callback /
  !> lfr_precon_callback): y = A x on device buffers, queued on `stream`;
  !> returns 0 on success.
  abstract interface
    integer(c_int) function lfr_apply_fn(user, x, y, stream) bind(C)
      import :: c_int, c_ptr
      type(c_ptr), value :: user, x, y, stream
    end function lfr_apply_fn
  end interface

contains

  !> Text of the last error recorded on a context (for log_event).
  function lfr_error_message(ctx) result(msg)
    type(c_ptr), intent(in) :: ctx
    character(len=1024) :: msg
    character(kind=c_char), pointer :: chars(:)
    type(c_ptr) :: p
    integer :: i
    msg = ' '
    p = lfr_last_error(ctx)
    if (.not. c_associated(p)) return
    call c_f_pointer(p, chars, [len(msg)])
    do i = 1, len(msg)
      if (chars(i) == c_null_char) exit
      msg(i:i) = chars(i)
    end do
  end function lfr_error_message

end module lfr_c_binding_mod
