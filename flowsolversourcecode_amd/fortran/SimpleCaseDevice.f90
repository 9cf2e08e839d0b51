! A Fortran host of the device-resident SIMPLE iteration through the C ABI
! (modules fsamd and fsamd_case).  It reads an OpenFOAM polyMesh (the input
! of the reference's Readit), makes every boundary patch a no-slip wall, the
! named one moving at (1, 0, 0) as the lid of a cavity, and runs niter outer
! iterations (MainProgram.for:257-348) with rho = 1, mu = 0.01 and the
! reference's settings urfMomentum 0.7, urfPressure 0.3, ASSolver 'ilu' for
! momentum (rrF 0.1, 20 iterations) and p' (rrF 0.01, 200 iterations), no
! multigrid.  It writes u, v, w, p (cells) and mdot (interior faces) as raw
! doubles to <out>.fields and the restart file of the last iteration in the
! reference's writeSolution layout to <out>.restart.  With a fifth argument
! the run starts from that restart file (readSolution, as the reference does
! with LReadOldSolution, MainProgram.for:171-177).
!
!   simple_case_device.x <polyMesh directory> <moving patch> <niter> <out> [<restart in>]
!
! Exit status 0 on success, 1 with the library's message on any failure.
program simple_case_device
  use iso_c_binding
  use fsamd, only: FS_ILU, FS_MG_OFF, fs_error_message
  use fsamd_case
  implicit none

  character(len=512) :: dir, lid, arg, out, restart_in
  integer :: niter, k, j, nb, u
  integer(c_int) :: counts(5), dims(4), rc
  type(c_ptr) :: pm, cs
  real(c_double), allocatable :: x(:), y(:), z(:), bu(:), bv(:), bw(:), fld(:)
  integer(c_int), allocatable :: face_ptr(:), face_nodes(:), owner(:), neighbour(:)
  integer(c_int), allocatable :: patch_start(:), patch_nfaces(:)
  character(kind=c_char), allocatable :: names(:)
  character(len=64) :: pname
  type(fs_simple_stats) :: st
  character(len=4), parameter :: cell_fields(4) = ['u   ', 'v   ', 'w   ', 'p   ']

  if (command_argument_count() /= 4 .and. command_argument_count() /= 5) then
    print '(a)', 'usage: simple_case_device.x <polyMesh directory> <moving patch> <niter> <out> [<restart in>]'
    stop 1
  end if
  call get_command_argument(1, dir)
  call get_command_argument(2, lid)
  call get_command_argument(3, arg)
  read (arg, *) niter
  call get_command_argument(4, out)
  restart_in = ' '
  if (command_argument_count() == 5) call get_command_argument(5, restart_in)

  call check(fs_polymesh_read(fs_cstr(dir), counts, pm), 'fs_polymesh_read')
  allocate (x(counts(1)), y(counts(1)), z(counts(1)), face_ptr(counts(2) + 1), face_nodes(counts(3)), &
            owner(counts(2)), neighbour(max(counts(4), 1)), patch_start(counts(5)), &
            patch_nfaces(counts(5)), names(64 * counts(5)))
  call check(fs_polymesh_get(pm, x, y, z, face_ptr, face_nodes, owner, neighbour, patch_start, &
                             patch_nfaces, names, 64), 'fs_polymesh_get')
  rc = fs_polymesh_destroy(pm)

  ! wall velocity of every boundary face: the named patch moves
  nb = counts(2) - counts(4)
  allocate (bu(nb), bv(nb), bw(nb))
  bu = 0.0d0
  bv = 0.0d0
  bw = 0.0d0
  do k = 1, counts(5)
    pname = ' '
    do j = 1, 64
      if (names(64 * (k - 1) + j) == c_null_char) exit
      pname(j:j) = names(64 * (k - 1) + j)
    end do
    if (trim(pname) == trim(lid)) &
      bu(patch_start(k) - counts(4) + 1:patch_start(k) - counts(4) + patch_nfaces(k)) = 1.0d0
  end do

  call check(fs_case_create(counts(1), x, y, z, counts(2), face_ptr, face_nodes, owner, counts(4), &
                            neighbour, bu, bv, bw, 1.0d0, 0.01d0, cs), 'fs_case_create')
  call check(fs_case_configure(cs, 0.7d0, 0.3d0, FS_ILU, 0.1d0, 20, FS_MG_OFF, FS_ILU, 0.01d0, 200, &
                               FS_MG_OFF), 'fs_case_configure')
  call check(fs_case_dims(cs, dims), 'fs_case_dims')
  if (len_trim(restart_in) > 0) &
    call check(fs_case_read_solution(cs, fs_cstr(restart_in), counts(5), patch_nfaces), 'fs_case_read_solution')
  if (niter > 1) call check(fs_case_iterate(cs, niter - 1, st), 'fs_case_iterate')
  ! the face gradients of the last iteration go into the restart file
  call check(fs_case_set_restart_capture(cs, 1), 'fs_case_set_restart_capture')
  call check(fs_case_iterate(cs, 1, st), 'fs_case_iterate')
  call check(fs_case_write_solution(cs, fs_cstr(trim(out)//'.restart'), counts(5), patch_nfaces), &
             'fs_case_write_solution')

  open (newunit=u, file=trim(out)//'.fields', access='stream', form='unformatted', status='replace')
  allocate (fld(dims(1)))
  do k = 1, 4
    call check(fs_case_get(cs, fs_cstr(trim(cell_fields(k))), fld, int(dims(1), c_long_long)), 'fs_case_get')
    write (u) fld
  end do
  deallocate (fld)
  allocate (fld(dims(4)))
  call check(fs_case_get(cs, fs_cstr('mdot'), fld, int(dims(4), c_long_long)), 'fs_case_get')
  write (u) fld
  close (u)
  rc = fs_case_destroy(cs)
  print '(a,i0,a,es12.5)', 'iterations ', niter, '  mass imbalance ', st%mass_l1

contains

  subroutine check(code, what)
    integer(c_int), intent(in) :: code
    character(len=*), intent(in) :: what
    if (code < 0) then
      print '(a,a,a)', trim(what), ' failed: ', trim(fs_error_message())
      stop 1
    end if
  end subroutine

end program simple_case_device
