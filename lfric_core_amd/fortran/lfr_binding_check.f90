!-----------------------------------------------------------------------------
! Link check of the Fortran binding: create a context through lfr_init.  On a
! host without a GPU the library reports an error code instead of a context;
! with a GPU it creates one, uploads a one-cell dofmap and releases both.
!-----------------------------------------------------------------------------
program lfr_binding_check
  use, intrinsic :: iso_c_binding, only : c_int, c_int32_t, c_ptr, c_associated
  use lfr_c_binding_mod
  implicit none
  type(c_ptr) :: ctx, space
  integer(c_int) :: rc
  integer(c_int32_t) :: dofmap(6) = [1, 4, 7, 10, 13, 14]

  rc = lfr_init(0_c_int, ctx)
  if (rc /= LFR_OK) then
    write(*, '(A, I0)') 'lfr_init returned ', rc
    stop
  end if
  rc = lfr_space_create(ctx, dofmap, 6_c_int, 1_c_int, 3_c_int, 17_c_int, space)
  write(*, '(A, I0)') 'lfr_space_create returned ', rc
  if (rc == LFR_OK) rc = lfr_space_destroy(space)
  rc = lfr_finalise(ctx)
  write(*, '(A, I0)') 'lfr_finalise returned ', rc
end program lfr_binding_check
