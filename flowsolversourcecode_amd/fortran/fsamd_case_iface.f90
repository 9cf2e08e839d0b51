! ISO_C_BINDING interface to the device-resident SIMPLE iteration of
! include/fsamd_simple.h: an OpenFOAM polyMesh reader (the input of the
! reference's Readit, GridAndGeometry/ReadFoamMesh.f90:31), the case (the
! reference's ProcessGeometry plus SolveMomentumX/Y/Z and SolveContinuity per
! outer iteration, MainProgram.for:257-348) and its restart file
! (writeSolution / readSolution, WriteSolutionToContinue.for:3).
! Integers travel as int32 by value where the C side takes them by value;
! names and paths as NUL-terminated character arrays (fs_cstr).
module fsamd_case
  use iso_c_binding
  implicit none

  integer(c_int), parameter :: FS_BC_WALL = 0, FS_BC_SYMMETRY = 1
  integer(c_int), parameter :: FS_ALGORITHM_SIMPLE = 0, FS_ALGORITHM_SIMPLEC = 1

  type, bind(c) :: fs_simple_stats
    integer(c_int) :: iterations, momentum_iterations, pc_iterations, pc_unconverged, pc_levels
    real(c_double) :: seconds, momentum_seconds, pc_seconds, pc_initial_l1, pc_final_l1
    real(c_double) :: res_momentum(3), mass_l1
  end type

  interface
    integer(c_int) function fs_polymesh_read(dir, counts, mesh) bind(c)
      import :: c_int, c_char, c_ptr
      character(kind=c_char), intent(in) :: dir(*)
      integer(c_int), intent(out) :: counts(5)
      type(c_ptr), intent(out) :: mesh
    end function

    integer(c_int) function fs_polymesh_get(mesh, x, y, z, face_ptr, face_nodes, owner, neighbour, &
                                            patch_start, patch_nfaces, patch_names, name_len) bind(c)
      import :: c_int, c_double, c_char, c_ptr
      type(c_ptr), value :: mesh
      real(c_double), intent(out) :: x(*), y(*), z(*)
      integer(c_int), intent(out) :: face_ptr(*), face_nodes(*), owner(*), neighbour(*)
      integer(c_int), intent(out) :: patch_start(*), patch_nfaces(*)
      character(kind=c_char), intent(out) :: patch_names(*)
      integer(c_int), value :: name_len
    end function

    integer(c_int) function fs_polymesh_destroy(mesh) bind(c)
      import :: c_int, c_ptr
      type(c_ptr), value :: mesh
    end function

    integer(c_int) function fs_case_create(npoints, x, y, z, nfaces, face_ptr, face_nodes, owner, nif, &
                                           neighbour, wall_u, wall_v, wall_w, rho, mu, cs) bind(c)
      import :: c_int, c_double, c_ptr
      integer(c_int), value :: npoints, nfaces, nif
      real(c_double), intent(in) :: x(*), y(*), z(*), wall_u(*), wall_v(*), wall_w(*)
      integer(c_int), intent(in) :: face_ptr(*), face_nodes(*), owner(*), neighbour(*)
      real(c_double), value :: rho, mu
      type(c_ptr), intent(out) :: cs
    end function

    integer(c_int) function fs_case_destroy(cs) bind(c)
      import :: c_int, c_ptr
      type(c_ptr), value :: cs
    end function

    integer(c_int) function fs_case_dims(cs, dims) bind(c)
      import :: c_int, c_ptr
      type(c_ptr), value :: cs
      integer(c_int), intent(out) :: dims(4)
    end function

    integer(c_int) function fs_case_set_patch_types(cs, npatches, patch_nfaces, patch_type) bind(c)
      import :: c_int, c_ptr
      type(c_ptr), value :: cs
      integer(c_int), value :: npatches
      integer(c_int), intent(in) :: patch_nfaces(*), patch_type(*)
    end function

    integer(c_int) function fs_case_set_algorithm(cs, algorithm) bind(c)
      import :: c_int, c_ptr
      type(c_ptr), value :: cs
      integer(c_int), value :: algorithm
    end function

    integer(c_int) function fs_case_configure(cs, urf_u, urf_p, msolver, mrrf, mitmax, mlmg, psolver, &
                                              prrf, pitmax, plmg) bind(c)
      import :: c_int, c_double, c_ptr
      type(c_ptr), value :: cs
      real(c_double), value :: urf_u, urf_p, mrrf, prrf
      integer(c_int), value :: msolver, mitmax, mlmg, psolver, pitmax, plmg
    end function

    integer(c_int) function fs_case_iterate(cs, niter, stats) bind(c)
      import :: c_int, c_ptr, fs_simple_stats
      type(c_ptr), value :: cs
      integer(c_int), value :: niter
      type(fs_simple_stats), intent(out) :: stats
    end function

    integer(c_int) function fs_case_get(cs, name, out, count) bind(c)
      import :: c_int, c_double, c_char, c_ptr, c_long_long
      type(c_ptr), value :: cs
      character(kind=c_char), intent(in) :: name(*)
      real(c_double), intent(out) :: out(*)
      integer(c_long_long), value :: count
    end function

    integer(c_int) function fs_case_set(cs, name, in, count) bind(c)
      import :: c_int, c_double, c_char, c_ptr, c_long_long
      type(c_ptr), value :: cs
      character(kind=c_char), intent(in) :: name(*)
      real(c_double), intent(in) :: in(*)
      integer(c_long_long), value :: count
    end function

    integer(c_int) function fs_case_set_restart_capture(cs, on) bind(c)
      import :: c_int, c_ptr
      type(c_ptr), value :: cs
      integer(c_int), value :: on
    end function

    integer(c_int) function fs_case_write_solution(cs, path, npatches, patch_nfaces) bind(c)
      import :: c_int, c_char, c_ptr
      type(c_ptr), value :: cs
      character(kind=c_char), intent(in) :: path(*)
      integer(c_int), value :: npatches
      integer(c_int), intent(in) :: patch_nfaces(*)
    end function

    integer(c_int) function fs_case_read_solution(cs, path, npatches, patch_nfaces) bind(c)
      import :: c_int, c_char, c_ptr
      type(c_ptr), value :: cs
      character(kind=c_char), intent(in) :: path(*)
      integer(c_int), value :: npatches
      integer(c_int), intent(in) :: patch_nfaces(*)
    end function
  end interface

contains

  ! a Fortran string as the NUL-terminated character array C expects
  function fs_cstr(s) result(c)
    character(len=*), intent(in) :: s
    character(kind=c_char) :: c(len_trim(s) + 1)
    integer :: k
    do k = 1, len_trim(s)
      c(k) = s(k:k)
    end do
    c(len_trim(s) + 1) = c_null_char
  end function

end module fsamd_case
