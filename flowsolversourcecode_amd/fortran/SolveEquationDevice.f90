! Drop-in replacement of the reference's
!   SUBROUTINE SolveEquation(Variable,FiT,rrF,itmax,solver,LMultigrid)
! (SolveEquations/AlgebraicSolvers/SolveAlegbraicEquations.for:4).
! Link this file in place of SolveAlegbraicEquations.for.  It reads the same
! module globals (ac, anb, bc, dphi: Variables2; ElementNeighbor,
! NumberofElementNeighbors: Geometry3; ResorAbs...: Residuals1; nIter:
! MultiGrid2; MG settings: User0), runs the solve on the MI355X through
! fs_solve_equation, and leaves FiT, dphi and the residual norms exactly where
! the reference leaves them.  A device failure (negative return code, e.g.
! the 'ilu' and 'direct' solvers or no GPU) falls back to the reference's own
! host solvers, called here through the same subroutines SolveEquation uses.
subroutine SolveEquation(Variable, FiT, rrF, itmax, solver, LMultigrid)
  use iso_c_binding
  use fsamd
  use user0, only: nIterStartApplyingMG, MGVariable, MGType, MultiGridpreSweep, &
                   MultiGridpostSweep, MaxMultiGridCycles, maxNumberofCoarseLevels, &
                   minNumberofParents, reAgglomorate
  use MultiGrid2, only: nIter
  use Residuals1
  use Variables2, only: ac, anb, bc, dphi
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
  end interface
  type(c_ptr), save :: mesh = c_null_ptr
  logical, save :: device_ok = .true.
  character*35 Variable1
  integer :: NF, iter, rc, code, mg, w
  double precision :: res4(4), IRes, FRes
  type(fs_solve_stats) :: st

  call GetNFfromVariable(Variable, NF, Variable1)
  ! at nIterStartApplyingMG the reference solves only the MG variable
  ! (SolveAlegbraicEquations.for:50-65)
  if (nIter == nIterStartApplyingMG .and. LMultigrid .and. MGType == 'algebraic' .and. &
      Variable1 /= MGVariable) then
    dphi = 0.
    call CalculateResiduals(NF, FiT)
    return
  end if

  code = fs_solver_code(solver)
  mg = FS_MG_OFF
  if (LMultigrid) then
    if (code == FS_SOR .or. code == FS_ILU) then
      mg = FS_MG_SOLVER
    else
      mg = FS_MG_PRECOND
    end if
  end if
  rc = -1
  if (device_ok .and. code >= 0) then
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
      end if
    end if
    if (device_ok) then
      rc = fs_solve_equation(mesh, NF, FiT, rrF, itmax, code, mg, ac, anb, bc, dphi, res4, st)
      if (rc >= 0) then
        ResorAbs(NF) = res4(1)
        ResorMax(NF) = res4(2)
        ResorRMS(NF) = res4(3)
        ResorScaled(NF) = res4(4)
        return
      end if
    end if
  end if

  ! host fallback: the reference's non-multigrid branch (:106-180)
  dphi = 0.
  call CalculateResiduals(NF, FiT)
  IRes = ResorAbs(NF)
  FRes = IRes
  iter = 0
  if (solver == 'ilu') call ILUFactorization
  if (solver == 'pbcg') then
    call allocateForPBCG
    call sprsin
    call ILUFactorization
  end if
  do while (FRes > rrF * IRes .and. iter < itmax)
    if (solver == 'sor') call SolveEquationsUsingSOR
    if (solver == 'ilu') call SolveEquationsUsingILU
    if (solver == 'pbcg') call SolveEquationsUsingPBCG(iter)
    if (solver /= 'sor' .and. solver /= 'ilu' .and. solver /= 'pbcg') call SolveEquationsUsingSOR
    call CheckConvergence(NF, FRes)
    iter = iter + 1
  end do
  FiT = FiT + dphi
end subroutine SolveEquation
