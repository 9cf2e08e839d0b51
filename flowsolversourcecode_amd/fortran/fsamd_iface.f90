! ISO_C_BINDING interface to the C ABI of include/fsamd.h.
! Default INTEGER is int32 and LOGICAL is passed as int32, as the C side expects.
module fsamd
  use iso_c_binding
  implicit none

  integer(c_int), parameter :: FS_SOR = 0, FS_ILU = 1, FS_PBCG = 2, FS_DIRECT = 3, &
                               FS_GMRES = 4, FS_BICGS = 5
  integer(c_int), parameter :: FS_MG_OFF = 0, FS_MG_SOLVER = 1, FS_MG_PRECOND = 2
  integer(c_int), parameter :: FS_OPT_MG_PRESWEEP = 1, FS_OPT_MG_POSTSWEEP = 2, &
                               FS_OPT_MG_MAXCYCLES = 3, FS_OPT_MG_MAXLEVELS = 4, &
                               FS_OPT_MG_MINPARENTS = 5, FS_OPT_MG_REAGGLOMERATE = 6, &
                               FS_OPT_MG_FP32 = 7

  type, bind(c) :: fs_solve_stats
    integer(c_int) :: iterations, levels
    real(c_double) :: initial_l1, final_l1, seconds
  end type

  interface
    integer(c_int) function fs_mesh_create(n_cells, width, elem_nbr, n_nbr, n_ifaces, owner, &
                                           neighbour, i_own_nbr, i_nbr_own, mesh) bind(c)
      import :: c_int, c_ptr
      integer(c_int), value :: n_cells, width, n_ifaces
      integer(c_int), intent(in) :: elem_nbr(*), n_nbr(*)
      type(c_ptr), value :: owner, neighbour, i_own_nbr, i_nbr_own
      type(c_ptr), intent(out) :: mesh
    end function

    integer(c_int) function fs_mesh_destroy(mesh) bind(c)
      import :: c_int, c_ptr
      type(c_ptr), value :: mesh
    end function

    integer(c_int) function fs_set_option(mesh, key, value) bind(c)
      import :: c_int, c_ptr, c_double
      type(c_ptr), value :: mesh
      integer(c_int), value :: key
      real(c_double), value :: value
    end function

    integer(c_int) function fs_solve_equation(mesh, nf, fit, rrf, itmax, solver, lmultigrid, &
                                              ac, anb, bc, dphi, res4, stats) bind(c)
      import :: c_int, c_ptr, c_double, fs_solve_stats
      type(c_ptr), value :: mesh
      integer(c_int), value :: nf, itmax, solver, lmultigrid
      real(c_double), value :: rrf
      real(c_double), intent(inout) :: fit(*)
      real(c_double), intent(in) :: ac(*), anb(*), bc(*)
      real(c_double), intent(out) :: dphi(*), res4(4)
      type(fs_solve_stats), intent(out) :: stats
    end function

    type(c_ptr) function fs_last_error() bind(c)
      import :: c_ptr
    end function
  end interface

contains

  ! the reference's solver strings (User0.F90:288-301) as enum codes
  integer(c_int) function fs_solver_code(solver)
    character(len=*), intent(in) :: solver
    select case (trim(solver))
    case ('sor');    fs_solver_code = FS_SOR
    case ('ilu');    fs_solver_code = FS_ILU
    case ('pbcg');   fs_solver_code = FS_PBCG
    case ('direct'); fs_solver_code = FS_DIRECT
    case ('gmres');  fs_solver_code = FS_GMRES
    case ('bicgs');  fs_solver_code = FS_BICGS
    case default;    fs_solver_code = -1
    end select
  end function

  function fs_error_message() result(msg)
    character(len=256) :: msg
    character(kind=c_char), pointer :: p(:)
    integer :: i
    call c_f_pointer(fs_last_error(), p, [256])
    msg = ' '
    do i = 1, 256
      if (p(i) == c_null_char) exit
      msg(i:i) = p(i)
    end do
  end function
end module fsamd
