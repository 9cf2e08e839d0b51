! Test-infrastructure driver for the reference's linear-solve path.
!
! Loads one ELL system (the reference's ac/anb/bc/ElementNeighbor layout) from a
! little-endian stream file, installs it in the reference's module globals and
! calls the reference's own routines:
!   mode 0: SolveEquation('pressc',...)    (SolveAlegbraicEquations.for:4)
!   mode 1: per-iteration history of the stationary / PBCG solvers, replaying the
!           loop of SolveAlegbraicEquations.for:111-153 with CheckConvergence
!           after every iteration.
! Output: dphi, FiT, ResorAbs/Max/RMS/Scaled(4), the residual history and, when the
! algebraic multigrid ran, its hierarchy (MultiGrid1 arrays, flattened).
! This program links against objects compiled from the reference sources; it is a
! checker only and never part of the product.
program ref_solve_driver
  use User0
  use MultiGrid1
  use MultiGrid2
  use Geometry1, only: NumberOfElements, MaximumNumberofElementFaces
  use Geometry3, only: ElementNeighbor, NumberofElementNeighbors, NElementFaces
  use Variables2, only: ac, anb, bc, dphi, dc, rc
  use Residuals1
  implicit none
  interface
    SUBROUTINE SolveEquation(Variable,FiT,rrF,itmax,solver,LMultigrid)
      character*10 Variable
      character*6 solver
      logical LMultigrid
      integer itmax
      double precision rrF
      double precision, dimension(:) :: FiT
    end SUBROUTINE SolveEquation
    SUBROUTINE CalculateResiduals(NF,FiT)
      integer NF
      double precision, dimension (:) :: FiT
    end SUBROUTINE CalculateResiduals
  end interface
  character(len=1024) :: fin, fout
  character(len=10) :: variable
  character(len=6) :: solver
  integer :: n, w, mode, isolver, itmax, lmg, icycle, npre, npost, maxcyc
  integer :: maxlev, minpar, i, j, iter, nhist, ncalls, icall, l, ntot, nntot
  double precision :: rrf, fres, ires
  double precision, allocatable :: fit(:), hist(:), bc0(:)
  logical :: lmultigrid

  call get_command_argument(1, fin)
  call get_command_argument(2, fout)
  open(unit=21, file=trim(fin), access='stream', form='unformatted', status='old')
  read(21) n, w, mode, isolver, itmax, lmg, icycle, npre, npost, maxcyc, maxlev, minpar, ncalls
  read(21) rrf
  NumberOfElements = n
  ! SolveAlgebraicSystemInParallel sizes its row lists by it
  MaximumNumberofElementFaces = w
  NElementFaces = w
  allocate(ElementNeighbor(n, w), NumberofElementNeighbors(n))
  allocate(ac(n), anb(n, w), bc(n), dphi(n), dc(n), rc(n), fit(n), bc0(n))
  read(21) ElementNeighbor
  read(21) NumberofElementNeighbors
  read(21) ac
  read(21) anb
  read(21) bc
  read(21) fit
  close(21)
  bc0 = bc

  allocate(ResorAbs(40), ResorMax(40), ResorRMS(40), ResorScaled(40))
  ResorAbs = 0.d0; ResorMax = 0.d0; ResorRMS = 0.d0; ResorScaled = 0.d0
  ! MultiGrid1 arrays must exist before ResetMultigridAllocation deallocates them.
  allocate(acMG(1), bcMG(1), anbMG(1), maxanb(1), Residuals(1), dphiMG(1))
  allocate(NumberOfElementsMG(1), NElementFacesMG(1), NBChildrenMaxMG(1))
  allocate(ijbeginE(1), ijbeginN(1), ijbeginNOld(1), ElementNeighborMG(1))
  allocate(NumberofElementNeighborsMG(1), Parents(1), Children(1))
  allocate(NumberOfChildren(1), NeighborOfParent(1), NumberOfNeighborOfParent(1))
  allocate(ElementParent(1))

  select case (isolver)
  case (0); solver = 'sor'
  case (1); solver = 'ilu'
  case (2); solver = 'pbcg'
  case (3); solver = 'direct'
  case (4); solver = 'gmres'
  case (5); solver = 'bicgs'
  case default; solver = 'sor'
  end select
  select case (icycle)
  case (0); MultiGridCycleType = 'vcycle'
  case (1); MultiGridCycleType = 'fcycle'
  case default; MultiGridCycleType = 'wcycle'
  end select
  MultiGridpreSweep = npre
  MultiGridpostSweep = npost
  MaxMultiGridCycles = maxcyc
  maxNumberofCoarseLevels = maxlev
  minNumberofParents = minpar
  MGType = 'algebraic'
  MGVariable = 'pressurecorrection'
  LTestMultiGrid = .false.
  LPrintMultiGridResiduals = .false.
  nIterStartApplyingMG = 1
  iLevelMax = 0
  reAgglomorate = 1000000
  lmultigrid = (lmg .ne. 0)
  variable = 'pressc'

  nhist = 0
  allocate(hist(max(itmax, 1) + 1))
  hist = 0.d0
  open(unit=22, file=trim(fout), access='stream', form='unformatted', status='replace')
  do icall = 1, ncalls
    nIter = icall
    nIterMG = icall
    if (icall > 1) then
      ! later calls reuse the hierarchy: perturb nothing, only the right-hand side
      bc = bc0 * (1.d0 + 0.25d0 * (icall - 1))
    end if
    if (mode == 0) then
      call SolveEquation(variable, fit, rrf, itmax, solver, lmultigrid)
    else
      dphi = 0.d0
      call CalculateResiduals(4, fit)
      ires = ResorAbs(4)
      fres = ires
      iter = 0
      if (isolver == 1) call ILUFactorization
      if (isolver == 2) then
        call allocateForPBCG
        call sprsin
        call ILUFactorization
      end if
      do while (fres > rrf * ires .and. iter < itmax)
        if (isolver == 0) call SolveEquationsUsingSOR
        if (isolver == 1) call SolveEquationsUsingILU
        if (isolver == 2) call SolveEquationsUsingPBCG(iter)
        call CheckConvergence(4, fres)
        iter = iter + 1
        hist(iter) = fres
      end do
      nhist = iter
      fit = fit + dphi
    end if
    write(22) n, nhist
    write(22) dphi
    write(22) fit
    write(22) ResorAbs(4), ResorMax(4), ResorRMS(4), ResorScaled(4)
    write(22) hist(1:max(nhist,1))
    ! the reference's hierarchy, when its AlgebraicMultigrid built one (a
    ! drop-in that solved on the device leaves iLevelMax at 0)
    if (mode == 0 .and. lmultigrid .and. iLevelMax > 0) then
      write(22) iLevelMax
      write(22) NumberOfElementsMG(1:iLevelMax), ijbeginE(1:iLevelMax), ijbeginN(1:iLevelMax), &
                NElementFacesMG(1:iLevelMax)
      ntot = sum(NumberOfElementsMG(1:iLevelMax))
      write(22) Parents(1:ntot), NumberOfChildren(1:ntot), ElementParent(1:ntot), &
                NumberofElementNeighborsMG(1:ntot)
      write(22) acMG(1:ntot)
      nntot = size(ElementNeighborMG)
      write(22) nntot, size(Children), size(anbMG)
      write(22) ElementNeighborMG
      write(22) Children
      write(22) anbMG
    else
      write(22) 0
    end if
  end do
  close(22)
end program ref_solve_driver
