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
                                          c_ptr, c_size_t, c_char, c_null_char,  &
                                          c_associated, c_f_pointer

  implicit none

  private

  integer(c_int), parameter, public :: LFR_OK = 0
  integer(c_int), parameter, public :: LFR_PREC_R_SINGLE = 4
  integer(c_int), parameter, public :: LFR_PREC_R_DOUBLE = 8

  public :: lfr_init, lfr_finalise, lfr_last_error, lfr_synchronize, lfr_malloc, lfr_free, &
            lfr_memcpy_h2d, lfr_memcpy_d2h, lfr_space_create, lfr_space_destroy,          &
            lfr_plan_create, lfr_plan_destroy, lfr_matrix_vector, lfr_dg_matrix_vector,   &
            lfr_dg_inc_matrix_vector, lfr_transpose_matrix_vector,                      &
            lfr_columnwise_op_appinv, lfr_vec_dot, lfr_vec_axpy, lfr_vec_setval,         &
            lfr_error_message

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
