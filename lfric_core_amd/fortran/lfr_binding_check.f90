!-----------------------------------------------------------------------------
! Fortran driver of the C ABI through lfr_c_binding_mod.
!
!   lfr_binding_check                      create a context and report the status
!                                          (on a host without a GPU the library
!                                          returns an error code instead)
!   lfr_binding_check problem.bin outdir   run a W2 mass problem from Fortran:
!
! problem.bin (stream access, native endianness): six int32 (ndf, ncell,
! nlayers, undf, ncolour, maxcells), then dofmap(ndf, ncell) int32,
! cells_in_colour(ncolour, maxcells) int32, ncells_per_colour(ncolour)
! int32, local_stencil(ncell*nlayers, ndf, ndf) real64, x(undf) real64 and
! b(undf) real64.  The program writes, as raw real64 arrays in outdir:
!   y_ref.bin    invoke(setval_c(y, 0), matrix_vector_kernel_type(y, x, mm))
!                on the reference layout (lfr_matrix_vector, overwrite)
!   y_packed.bin the same apply through the packed copy (lfr_matrix_vector_packed)
!   x_cg.bin     25 unmonitored CG iterations from x = 0 on M x = b (lfr_cg)
!   x_bicg.bin   BiCGstab to r_tol 1e-10 from x = 0 through the generic solver
!                entry (lfr_linop_lma, lfr_precon_diagonal of mm_diagonal,
!                lfr_iterative_solve)
! and prints the solver statistics.
!-----------------------------------------------------------------------------
program lfr_binding_check
  use, intrinsic :: iso_c_binding
  use lfr_c_binding_mod
  implicit none
  type(c_ptr) :: ctx, space, plan, packed, linop, prec
  integer(c_int) :: rc
  integer(c_int32_t) :: hdr(6)
  integer(c_int32_t), allocatable, target :: dofmap(:), cic(:), npc(:)
  real(c_double), allocatable, target :: mm(:), x(:), b(:), y(:)
  type(c_ptr) :: mm_d, x_d, b_d, y_d, diag_d
  integer :: ndf, ncell, nl, undf, ncolour, maxc, n3, u, nargs
  integer(c_size_t) :: vbytes, mbytes
  character(len=4096) :: problem, outdir
  type(lfr_lma_op) :: op
  type(lfr_solver_params) :: params
  type(lfr_solver_stats) :: stats
  type(lfr_vector_layout) :: layout
  type(lfr_iterative_solver) :: solver

  nargs = command_argument_count()
  rc = lfr_init(0_c_int, ctx)
  write(*, '(A, I0)') 'lfr_init returned ', rc
  if (rc /= LFR_OK) stop
  if (nargs < 2) then
    rc = lfr_finalise(ctx)
    write(*, '(A, I0)') 'lfr_finalise returned ', rc
    stop
  end if
  call get_command_argument(1, problem)
  call get_command_argument(2, outdir)

  open(newunit=u, file=trim(problem), access='stream', form='unformatted', status='old')
  read(u) hdr
  ndf = hdr(1); ncell = hdr(2); nl = hdr(3); undf = hdr(4); ncolour = hdr(5); maxc = hdr(6)
  n3 = ncell * nl
  allocate(dofmap(ndf * ncell), cic(ncolour * maxc), npc(ncolour))
  allocate(mm(n3 * ndf * ndf), x(undf), b(undf), y(undf))
  read(u) dofmap, cic, npc, mm, x, b
  close(u)

  rc = lfr_space_create(ctx, dofmap, int(ndf, c_int), int(ncell, c_int), int(nl, c_int), &
                        int(undf, c_int), space)
  call check(rc, 'lfr_space_create')
  rc = lfr_plan_create(ctx, space, cic, npc, int(ncolour, c_int), int(maxc, c_int), 0_c_int, plan)
  call check(rc, 'lfr_plan_create')

  vbytes = int(undf, c_size_t) * 8_c_size_t
  mbytes = int(size(mm), c_size_t) * 8_c_size_t
  call check(lfr_malloc(ctx, mbytes, mm_d), 'lfr_malloc')
  call check(lfr_malloc(ctx, vbytes, x_d), 'lfr_malloc')
  call check(lfr_malloc(ctx, vbytes, b_d), 'lfr_malloc')
  call check(lfr_malloc(ctx, vbytes, y_d), 'lfr_malloc')
  call check(lfr_malloc(ctx, vbytes, diag_d), 'lfr_malloc')
  call check(lfr_memcpy_h2d(ctx, mm_d, c_loc(mm), mbytes), 'lfr_memcpy_h2d')
  call check(lfr_memcpy_h2d(ctx, x_d, c_loc(x), vbytes), 'lfr_memcpy_h2d')
  call check(lfr_memcpy_h2d(ctx, b_d, c_loc(b), vbytes), 'lfr_memcpy_h2d')

  ! invoke( setval_c(y, 0.0), matrix_vector_kernel_type(y, x, mm) )
  rc = lfr_matrix_vector(ctx, LFR_PREC_R_DOUBLE, y_d, x_d, mm_d, space, space, plan, 1_c_int)
  call check(rc, 'lfr_matrix_vector')
  call dump(y_d, 'y_ref.bin')

  ! the same apply through a packed copy
  rc = lfr_lma_pack(ctx, LFR_PREC_R_DOUBLE, mm_d, space, plan, packed)
  call check(rc, 'lfr_lma_pack')
  rc = lfr_vec_setval(ctx, LFR_PREC_R_DOUBLE, int(undf, c_int64_t), y_d, 7.0_c_double)
  rc = lfr_matrix_vector_packed(ctx, packed, y_d, x_d, 1_c_int)
  call check(rc, 'lfr_matrix_vector_packed')
  call dump(y_d, 'y_packed.bin')

  ! conjugate_gradient_type%cg_solve, 25 iterations from x = 0, packed operator
  op%prec = LFR_PREC_R_DOUBLE
  op%matrix_d = mm_d
  op%space = space
  op%plan = plan
  op%apply_bc = 0
  op%packed = packed
  params%max_iter = 25
  params%monitor_convergence = 0
  params%x_is_zero = 1
  rc = lfr_cg(ctx, op, params, y_d, b_d, stats)
  call check(rc, 'lfr_cg')
  write(*, '(A, I0, A, ES24.16)') 'cg iterations ', stats%iterations, ' r_nrm0 ', stats%r_nrm0
  call dump(y_d, 'x_cg.bin')

  ! BiCGstab with the diagonal preconditioner through the generic entry
  rc = lfr_vec_setval(ctx, LFR_PREC_R_DOUBLE, int(undf, c_int64_t), diag_d, 0.0_c_double)
  call check(lfr_mm_diagonal(ctx, LFR_PREC_R_DOUBLE, diag_d, mm_d, space, plan), 'lfr_mm_diagonal')
  call check(lfr_linop_lma(ctx, op, linop), 'lfr_linop_lma')
  layout%prec = LFR_PREC_R_DOUBLE
  layout%nfield = 1
  layout%undf = 0
  layout%undf(1) = int(undf, c_int64_t)
  call check(lfr_precon_diagonal(ctx, layout, diag_d, prec), 'lfr_precon_diagonal')
  solver%method = LFR_SOLVER_BICGSTAB
  solver%r_tol = 1.0e-10_c_double
  solver%max_iter = 200
  rc = lfr_vec_setval(ctx, LFR_PREC_R_DOUBLE, int(undf, c_int64_t), y_d, 0.0_c_double)
  rc = lfr_iterative_solve(ctx, solver, linop, prec, y_d, b_d, stats)
  call check(rc, 'lfr_iterative_solve')
  write(*, '(A, I0, A, I0)') 'bicgstab iterations ', stats%iterations, ' converged ', stats%converged
  call dump(y_d, 'x_bicg.bin')

  rc = lfr_precon_destroy(prec)
  rc = lfr_linop_destroy(linop)
  rc = lfr_lma_packed_destroy(packed)
  rc = lfr_free(ctx, mm_d)
  rc = lfr_free(ctx, x_d)
  rc = lfr_free(ctx, b_d)
  rc = lfr_free(ctx, y_d)
  rc = lfr_free(ctx, diag_d)
  rc = lfr_plan_destroy(plan)
  rc = lfr_space_destroy(space)
  rc = lfr_finalise(ctx)
  write(*, '(A, I0)') 'lfr_finalise returned ', rc

contains

  subroutine check(status, what)
    integer(c_int), intent(in) :: status
    character(len=*), intent(in) :: what
    if (status /= LFR_OK) then
      write(*, '(A, A, I0, A, A)') what, ' failed: ', status, ' ', trim(lfr_error_message(ctx))
      error stop 1
    end if
  end subroutine check

  subroutine dump(field_d, name)
    type(c_ptr), intent(in) :: field_d
    character(len=*), intent(in) :: name
    integer :: v
    call check(lfr_memcpy_d2h(ctx, c_loc(y), field_d, vbytes), 'lfr_memcpy_d2h')
    open(newunit=v, file=trim(outdir)//'/'//name, access='stream', form='unformatted', &
         status='replace')
    write(v) y
    close(v)
  end subroutine dump

end program lfr_binding_check
