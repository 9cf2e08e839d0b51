! Drop-in replacement of the reference's
!   SUBROUTINE SolveEquation(Variable,FiT,rrF,itmax,solver,LMultigrid)
! (SolveEquations/AlgebraicSolvers/SolveAlegbraicEquations.for:4).
! Link this file in place of SolveAlegbraicEquations.for.  It reads the same
! module globals (ac, anb, bc, dphi: Variables2; ElementNeighbor,
! NumberofElementNeighbors: Geometry3; ResorAbs...: Residuals1; nIter:
! MultiGrid2; MG settings: User0), runs the solve on the MI355X through
! fs_solve_equation, and leaves FiT, dphi and the residual norms exactly where
! the reference leaves them.  It keeps the reference's branches
! (SolveAlegbraicEquations.for:45-180): multigrid from nIterStartApplyingMG
! on (only the MG variable at that iteration for 'algebraic'), the plain
! solvers before it.  The device runs the plain 'sor', 'ilu', 'pbcg', 'gmres'
! and 'bicgs' solves and the algebraic multigrid with V cycles (with the
! reference's pre/post sweep counts); 'direct', the geometric agglomerations,
! F and W cycles, and any call the device refuses (no GPU, invalid input) run
! the reference's own host code through the same subroutines SolveEquation
! calls, as do p' solves with resistance outlets (UpdateSourceResistance).  With the environment variable FSAMD_SOLVE_LOG naming a file, every
! call appends one line to it: 'device' or 'host', the variable, the solver
! and the device's iteration count.
subroutine SolveEquation(Variable, FiT, rrF, itmax, solver, LMultigrid)
  use iso_c_binding
  use fsamd
  use user0, only: nIterStartApplyingMG, MGVariable, MGType, MultiGridCycleType, MultiGridpreSweep, &
                   MultiGridpostSweep, MaxMultiGridCycles, maxNumberofCoarseLevels, &
                   minNumberofParents, reAgglomorate
  use MultiGrid2, only: nIter
  use Residuals1
  use Variables2, only: ac, anb, bc, dphi
  use DirectSolver1, only: b
  use BoundaryConditions2, only: IoutletSpecifiedResistance
  use Geometry1, only: NumberOfElements
  use Geometry3, only: ElementNeighbor, NumberofElementNeighbors
  implicit none
  character*10 Variable
  character*6 solver
  logical LMultigrid
  integer itmax
  double precision rrF
  double precision, dimension(:) :: FiT
  interface
    subroutine CalculateResiduals(NF, FiT)
      integer NF
      double precision, dimension(:) :: FiT
    end subroutine
    subroutine UpdateSourceResistance(iter, FiT)
      integer iter
      double precision, dimension(:) :: FiT
    end subroutine
    subroutine AlgebraicMultigrid(NF, rrF, itmax, solver, IResiduals, FResiduals, Variable1)
      integer NF, itmax
      double precision rrF, IResiduals, FResiduals
      character*6 solver
      character*35 Variable1
    end subroutine
  end interface
  type(c_ptr), save :: mesh = c_null_ptr
  logical, save :: device_ok = .true.
  character*35 Variable1
  integer :: NF, iter, rc, code, mg, w
  logical :: mg_applies, on_device
  double precision :: res4(4), IRes, FRes
  type(fs_solve_stats) :: st

  call GetNFfromVariable(Variable, NF, Variable1)
  ! the multigrid branches (SolveAlegbraicEquations.for:48-100): the algebraic
  ! agglomeration from the iteration after nIterStartApplyingMG, and at that
  ! iteration for the MG variable only; the geometric ones from it on
  mg_applies = .false.
  if (LMultigrid) then
    if (MGType == 'algebraic') then
      if (nIter == nIterStartApplyingMG) then
        if (Variable1 /= MGVariable) then
          dphi = 0.
          call CalculateResiduals(NF, FiT)
          call log_call('host', 0)
          return
        end if
        mg_applies = .true.
      else if (nIter > nIterStartApplyingMG) then
        mg_applies = .true.
      end if
    else if (MGType == 'geometricelement' .or. MGType == 'geometricnode') then
      mg_applies = nIter >= nIterStartApplyingMG
    end if
  end if

  code = fs_solver_code(solver)
  on_device = code >= 0 .and. solver /= 'direct'
  if (mg_applies) on_device = on_device .and. MGType == 'algebraic' .and. MultiGridCycleType == 'vcycle'
  ! outlets with a specified resistance change the p' source between sweeps
  ! (UpdateSourceResistance): that loop stays on the host
  if (.not. mg_applies .and. NF == 4 .and. IoutletSpecifiedResistance > 0 .and. &
      (code == FS_SOR .or. code == FS_ILU .or. code == FS_PBCG)) on_device = .false.
  mg = FS_MG_OFF
  if (mg_applies) then
    if (code == FS_SOR .or. code == FS_ILU) then
      mg = FS_MG_SOLVER
    else
      mg = FS_MG_PRECOND
    end if
  end if
  if (device_ok .and. on_device) then
    if (.not. c_associated(mesh)) then
      w = size(ElementNeighbor, 2)
      rc = fs_mesh_create(NumberOfElements, w, ElementNeighbor, NumberofElementNeighbors, 0, &
                          c_null_ptr, c_null_ptr, c_null_ptr, c_null_ptr, mesh)
      if (rc < 0) then
        device_ok = .false.
        print *, 'SolveEquation: device unavailable, using the host solvers: ', &
                 trim(fs_error_message())
      else
        rc = fs_set_option(mesh, FS_OPT_MG_MAXCYCLES, dble(MaxMultiGridCycles))
        rc = fs_set_option(mesh, FS_OPT_MG_MAXLEVELS, dble(maxNumberofCoarseLevels))
        rc = fs_set_option(mesh, FS_OPT_MG_MINPARENTS, dble(minNumberofParents))
        rc = fs_set_option(mesh, FS_OPT_MG_REAGGLOMERATE, dble(reAgglomorate))
        rc = fs_set_option(mesh, FS_OPT_MG_PRESWEEP, dble(MultiGridpreSweep))
        rc = fs_set_option(mesh, FS_OPT_MG_POSTSWEEP, dble(MultiGridpostSweep))
      end if
    end if
    if (device_ok) then
      rc = fs_solve_equation(mesh, NF, FiT, rrF, itmax, code, mg, ac, anb, bc, dphi, res4, st)
      if (rc >= 0) then
        ResorAbs(NF) = res4(1)
        ResorMax(NF) = res4(2)
        ResorRMS(NF) = res4(3)
        ResorScaled(NF) = res4(4)
        call log_call('device', st%iterations)
        return
      end if
    end if
  end if

  ! the reference's own host code (SolveAlegbraicEquations.for:45-180)
  dphi = 0.
  call CalculateResiduals(NF, FiT)
  IRes = ResorAbs(NF)
  FRes = IRes
  if (mg_applies) then
    call AlgebraicMultigrid(NF, rrF, itmax, solver, IRes, FRes, Variable1)
  else if (solver == 'sor' .or. solver == 'ilu' .or. solver == 'pbcg') then
    iter = 0
    if (solver == 'ilu') call ILUFactorization
    if (solver == 'pbcg') then
      call allocateForPBCG
      call sprsin
      call ILUFactorization
    end if
    do while (FRes > rrF * IRes .and. iter < itmax)
      if (NF == 4 .and. iter /= 0) call UpdateSourceResistance(iter, FiT)
      if (solver == 'sor') call SolveEquationsUsingSOR
      if (solver == 'ilu') call SolveEquationsUsingILU
      if (solver == 'pbcg') call SolveEquationsUsingPBCG(iter)
      call CheckConvergence(NF, FRes)
      iter = iter + 1
    end do
  else if (solver == 'direct') then
    call SetupMatrixforDirectSolver(NumberOfElements)
    call ludcmpF77
    call lubksbF77
    dphi = b
    call deallocateMatrixStorage
  else if (solver == 'gmres' .or. solver == 'bicgs') then
    call SolveAlgebraicSystemInParallel(solver)
  end if
  FiT = FiT + dphi
  call log_call('host', 0)

contains

  subroutine log_call(path, its)
    character(len=*), intent(in) :: path
    integer, intent(in) :: its
    character(len=1024) :: fname
    integer :: ln, st2, u
    call get_environment_variable('FSAMD_SOLVE_LOG', fname, ln, st2)
    if (st2 /= 0 .or. ln == 0) return
    open(newunit=u, file=fname(1:ln), status='unknown', position='append', action='write')
    write(u, '(a,1x,a,1x,a,1x,i0)') trim(path), trim(Variable), trim(solver), its
    close(u)
  end subroutine log_call
end subroutine SolveEquation
