! Test-infrastructure driver for the reference's geometry and SIMPLE iteration.
!
! Loads a polyMesh (points, face node lists, owner/neighbour, patches) with the
! cell topology of the reference's Readit (NumbOfElementFaces, NGlobalEFaces,
! ElementNeighbor, iOwnerNeighbor/iNeighborOwner, element node lists) from a
! little-endian stream file, installs them in the reference's module globals in
! place of Readit (ReadFoamMesh.f90:254-619) and calls the reference's own
! ProcessGeometry (ProcessGeometryRoutine.for:3).
!   mode 1: writes the geometry.
!   mode 2: then runs the reference's start-up and nouter SIMPLE outer
!           iterations of MainProgram.for:128-163,257-348 for a steady, laminar,
!           incompressible, constant-property flow bounded by no-slip walls and
!           symmetry planes (allocateBCStorage, SetBoundaryConditions,
!           SetBCPatches, CalculateFaceDensity, CalculateInitialmdot,
!           SolveMomentumX/Y/Z, SolveContinuity) and writes u, v, w, p, mdot.
!           The storage the reference's allocateStorage would allocate for
!           such a flow is allocated here and zeroed as it does
!           (AllocateStorageRoutine.for:36-147,208-259,586-700,1617-2193).
! This program links against objects compiled from the reference sources; it is a
! checker only and never part of the product.
program ref_simple_driver
  use User0
  use MultiGrid2, only: nIter, nIterMG
  use BoundaryConditions1
  use Variables1
  use Variables2
  use Variables3
  use Variables4
  use PhysicalProperties1
  use Residuals1
  use MultiGrid1
  use Geometry1, only: NumberOfNodes, NumberOfElements, NumberOfBCSets, NumbOfElementNodes, &
                       ListOfElementNodes, NumbOfElementFaces, x, y, z
  use Geometry3, only: NIFaces, NFacesTotal, NIFaceOwner, NIFaceNeighbor, NBFaces, NBFaceOwner, &
                       NGlobalEFaces, ElementNeighbor, NumberofElementNeighbors, NElementFaces, &
                       GlobalFaceNumberOfNodes, NIFaceNodes, NBFaceNodes, BCSet, BCRecord, &
                       iOwnerNeighbor, iNeighborOwner, NBFacesMax, NBFaceNeighbor
  use Geometry4
  use BoundaryConditions2, only: PeriodicPair, xTranslation, yTranslation, zTranslation, &
                                 LTranslationalPeriodicity, LRotationalPeriodicity, LPeriodicImplicit, &
                                 LcalculateBeta, periodicBeta, periodicMdot, relaxBeta, BetaIterations, &
                                 nIterCorrectBeta
  implicit none
  character(len=1024) :: fin, fout
  integer :: npoints, nfaces, nif, ncells, npatches, w, lenw, maxfn, mode, idecomp
  integer :: i, j, k, f, q, nbfmax, nb, it
  integer :: msolver, mitmax, psolver, pitmax, nouter, hrfw, hrscheme, limmom, limcont, hrstart, plmg, isimplec
  integer :: gmm, gmi, gcm, gci, gim, gic, iper, bcalc, biters, bevery
  integer, allocatable :: perpair(:)
  double precision, allocatable :: pertrans(:)
  double precision :: beta0, bmdot0, brelax
  character(len=16), parameter :: gradinterp(0:4) = (/ 'average         ', 'upwind          ', &
      'downwind        ', 'averagecorrected', 'harmonic        ' /)
  double precision :: rho, mu, urfu, urfp, mrrf, prrf, bleed
  integer, allocatable :: fnn(:), fnodes(:,:), owner(:), neigh(:), psize(:), ptype(:)
  double precision, allocatable :: wallu(:), wallv(:), wallw(:)
  character(len=6), parameter :: solvers(0:5) = (/ 'sor   ', 'ilu   ', 'pbcg  ', 'direct', 'gmres ', 'bicgs ' /)
  character(len=4), parameter :: frameworks(0:2) = (/ 'none', 'tvd ', 'nvf ' /)
  character(len=20), parameter :: schemes(0:17) = (/ 'upwind   ', 'downwind ', 'fromm    ', 'sou      ', &
      'cd       ', 'quick    ', 'superbee ', 'minmod   ', 'osher    ', 'vanleer  ', 'muscl    ', 'smart    ', &
      'hquick   ', 'umist    ', 'charm    ', 'boundedcd', 'stoic    ', 'gamma    ' /)

  call get_command_argument(1, fin)
  call get_command_argument(2, fout)
  open(unit=21, file=trim(fin), access='stream', form='unformatted', status='old')
  read(21) npoints, nfaces, nif, ncells, npatches, w, lenw, maxfn, mode, idecomp
  if (mode == 2) then
    read(21) msolver, mitmax, psolver, pitmax, nouter, hrfw, hrscheme, limmom, limcont, hrstart, plmg, isimplec, &
             gmm, gmi, gcm, gci, gim, gic, iper
    read(21) rho, mu, urfu, urfp, mrrf, prrf, bleed
    if (iper /= 0) then
      ! translational periodic pairs (the case file's block of
      ! ProblemDescriptionRoutine.for:1339-1371): partner patch (1-based, 0
      ! none) and x/y/zTranslation per patch, the bulk-flux controls
      allocate(perpair(npatches), pertrans(3 * npatches))
      read(21) perpair
      read(21) pertrans
      read(21) beta0, bmdot0, brelax
      read(21) bcalc, biters, bevery
    end if
  end if
  allocate(fnn(nfaces), fnodes(nfaces, maxfn), owner(nfaces), neigh(nif), psize(npatches), ptype(npatches))
  NumberOfNodes = npoints
  NumberOfElements = ncells
  NumberOfBCSets = npatches
  NIFaces = nif
  NFacesTotal = nfaces
  NElementFaces = w
  allocate(x(npoints), y(npoints), z(npoints))
  read(21) x, y, z
  read(21) fnn
  read(21) fnodes
  read(21) owner
  read(21) neigh
  read(21) psize
  read(21) ptype
  allocate(NumbOfElementFaces(ncells), NGlobalEFaces(ncells, w), ElementNeighbor(ncells, w))
  allocate(NumberofElementNeighbors(ncells), iOwnerNeighbor(nif), iNeighborOwner(nif))
  allocate(NumbOfElementNodes(ncells), ListOfElementNodes(ncells, lenw))
  read(21) NumbOfElementFaces
  read(21) NGlobalEFaces
  read(21) ElementNeighbor
  read(21) iOwnerNeighbor
  read(21) iNeighborOwner
  read(21) NumbOfElementNodes
  read(21) ListOfElementNodes
  if (mode == 2) then
    allocate(wallu(nfaces - nif), wallv(nfaces - nif), wallw(nfaces - nif))
    read(21) wallu, wallv, wallw
  end if
  close(21)
  NumberofElementNeighbors = NumbOfElementFaces

  ! Readit's face arrays: interior faces first, then boundary faces patch by patch
  nbfmax = maxval(psize)
  allocate(NIFaceOwner(nif), NIFaceNeighbor(nif), NIFaceNodes(nif, maxfn))
  allocate(NBFaces(npatches), NBFaceOwner(npatches, nbfmax), NBFaceNodes(npatches, nbfmax, maxfn))
  allocate(NBFaceNeighbor(npatches, nbfmax))
  NBFaceNeighbor = -1
  allocate(GlobalFaceNumberOfNodes(nfaces), BCSet(nfaces - nif), BCRecord(nfaces - nif))
  GlobalFaceNumberOfNodes = fnn
  do f = 1, nif
    NIFaceOwner(f) = owner(f)
    NIFaceNeighbor(f) = neigh(f)
    NIFaceNodes(f, :) = fnodes(f, :)
  end do
  NBFaces = psize
  f = nif
  do q = 1, npatches
    do j = 1, psize(q)
      f = f + 1
      NBFaceOwner(q, j) = owner(f)
      NBFaceNodes(q, j, :) = fnodes(f, :)
      BCSet(f - nif) = q
      BCRecord(f - nif) = j
    end do
  end do

  MethodDecomposeS = idecomp
  open(unit=11, file=trim(fout)//'.log', status='replace')
  call ProcessGeometry
  close(11)

  open(unit=22, file=trim(fout), access='stream', form='unformatted', status='replace')
  nb = nfaces - nif
  call put_faces(FaceAreax, BFaceAreax)
  call put_faces(FaceAreay, BFaceAreay)
  call put_faces(FaceAreaz, BFaceAreaz)
  call put_faces(FaceArea, BFaceArea)
  call put_faces(FaceCentroidx, BFaceCentroidx)
  call put_faces(FaceCentroidy, BFaceCentroidy)
  call put_faces(FaceCentroidz, BFaceCentroidz)
  call put_faces(FaceAreanx, BFaceAreanx)
  call put_faces(FaceAreany, BFaceAreany)
  call put_faces(FaceAreanz, BFaceAreanz)
  call put_faces(DistanceCF, BDistanceCF)
  call put_faces(DistanceCFx, BDistanceCFx)
  call put_faces(DistanceCFy, BDistanceCFy)
  call put_faces(DistanceCFz, BDistanceCFz)
  call put_faces(DistanceCFux, BDistanceCFux)
  call put_faces(DistanceCFuy, BDistanceCFuy)
  call put_faces(DistanceCFuz, BDistanceCFuz)
  call put_faces(FaceEx, BFaceEx)
  call put_faces(FaceEy, BFaceEy)
  call put_faces(FaceEz, BFaceEz)
  call put_faces(FaceE, BFaceE)
  call put_faces(gDiff, BgDiff)
  call put_faces(FaceTx, BFaceTx)
  call put_faces(FaceTy, BFaceTy)
  call put_faces(FaceTz, BFaceTz)
  call put_faces(FaceT, BFaceT)
  write(22) GFactorCF
  write(22) Volume, xc, yc, zc
  if (mode == 2) then
    call run_simple
    write(22) uVelocity, vVelocity, wVelocity, Pressure, mdot
    if (iper /= 0) then
      write(22) periodicBeta
      do q = 1, npatches
        do j = 1, psize(q)
          write(22) Bmdot(q, j)
        end do
      end do
    end if
    ! the reference's restart file (WriteSolutionToContinue.for:3) of the
    ! final state, on unit 4 as MainProgram.for:466 opens it
    open(unit=4, file=trim(fout)//'.restart', form='unformatted', status='replace')
    call writeSolution
    close(4)
  end if
  close(22)

contains

  subroutine run_simple
    ! the case: steady, laminar, incompressible, constant properties
    LSolveMomentum = .true.
    LSolveContinuity = .true.
    LSolveEnergy = .false.
    LTurbulentFlow = .false.
    LUnsteady = .false.
    Lcompressible = .false.
    LBuoyancy = .false.
    LFreeSurfaceFlow = .false.
    LFalseTransientMomentum = .false.
    Linviscid = .false.
    LfixPressure = .false.
    LanisotropicDiffusion = .false.
    LConvectScalar = .false.
    NumberOfScalarsToSolve = 0
    NumberOfrFieldsToSolve = 0
    LRelaxMomentum = .true.
    GradientInterpolationSchemeMomentum = gradinterp(gim)
    GradientInterpolationSchemeContinuity = gradinterp(gic)
    MethodCalcGradientMomentum = gmm
    nIterGradientMomentum = gmi
    MethodCalcGradientContinuity = gcm
    nIterGradientContinuity = gci
    if (isimplec /= 0) then
      Algorithm = 'simplec'
    else
      Algorithm = 'simple'
    end if
    urfMomentum = urfu
    urfPressure = urfp
    ConstantDensity = rho
    ConstantViscosity = mu
    ASSolverMomentum = solvers(msolver)
    rrFMomentum = mrrf
    ASIterMomentum = mitmax
    LMultigridMomentum = .false.
    ASSolverContinuity = solvers(psolver)
    rrFContinuity = prrf
    ASIterContinuity = pitmax
    LMultigridContinuity = plmg /= 0
    ! the algebraic agglomeration (the geometric ones read node connectivity
    ! the mesh reader builds and this driver does not install)
    MGType = 'algebraic'
    MGVariable = 'pressurecorrection'
    LTestMultiGrid = .false.
    LPrintMultiGridResiduals = .false.
    nIterStartApplyingMG = 1
    HRFrameworkMomentum = frameworks(hrfw)
    ConvectionSchemeMomentum = schemes(hrscheme)
    BleedMomentum = bleed
    nIterStartApplyingHR = hrstart
    LimitGradientMomentum = limmom > 0
    LimitGradientMomentumMethod = max(limmom, 1)
    LimitGradientContinuity = limcont > 0
    LimitGradientContinuityMethod = max(limcont, 1)
    NBFacesMax = nbfmax

    call allocateBCStorage
    do q = 1, npatches
      if (ptype(q) == 1) then
        BoundaryType(q) = 'symmetry'
      else if (ptype(q) == 2) then
        BoundaryType(q) = 'periodic'
      else
        BoundaryType(q) = 'wall'
      end if
      wallTypeM(q) = 'noslip'
      wallTypeC(q) = 'none'
      wallTypeE(q) = 'none'
      wallTypeL(q) = 'none'
      inletTypeM(q) = 'none'
      inletTypeC(q) = 'none'
      inletTypeE(q) = 'none'
      outletTypeM(q) = 'none'
      outletTypeC(q) = 'none'
      outletTypeE(q) = 'none'
    end do
    call SetBoundaryConditions
    call SetBCPatches
    LRotationalPeriodicity = .false.
    LTranslationalPeriodicity = iper /= 0
    if (iper /= 0) then
      LPeriodicImplicit = .true.
      do q = 1, npatches
        PeriodicPair(q) = perpair(q)
        xTranslation(q) = pertrans(3 * q - 2)
        yTranslation(q) = pertrans(3 * q - 1)
        zTranslation(q) = pertrans(3 * q)
      end do
      LcalculateBeta = bcalc /= 0
      periodicBeta = beta0
      periodicMdot = bmdot0
      relaxBeta = brelax
      BetaIterations = biters
      nIterCorrectBeta = bevery
      ! MainProgram.for:66
      call RotationMatrixCoefficients
    end if
    call allocate_flow_storage
    f = 0
    do q = 1, npatches
      do j = 1, psize(q)
        f = f + 1
        BuVelocity(q, j) = wallu(f)
        BvVelocity(q, j) = wallv(f)
        BwVelocity(q, j) = wallw(f)
      end do
    end do
    nIter = 0
    call SetReferencePressure
    call CalculatePhysicalProperties
    call UpdateBoundaryPressure
    call UpdateSymmetry
    call UpdateSlipWallBoundary
    call CalculateFaceDensity
    call CalculateInitialmdot
    do it = 1, nouter
      nIter = it
      nIterMG = it
      call CalculatePhysicalProperties
      call SolveMomentumX
      call SolveMomentumY
      call SolveMomentumZ
      call SolveContinuity
    end do
  end subroutine run_simple

  ! What allocateStorage holds for this flow, zeroed as it zeroes it; the
  ! densities are ConstantDensity.
  subroutine allocate_flow_storage
    integer :: n, nb2
    n = ncells
    nb2 = nbfmax
    allocate(uVelocity(n), vVelocity(n), wVelocity(n))
    allocate(BuVelocity(npatches, nb2), BvVelocity(npatches, nb2), BwVelocity(npatches, nb2))
    allocate(uVelGradx(n), uVelGrady(n), uVelGradz(n), vVelGradx(n), vVelGrady(n), vVelGradz(n))
    allocate(wVelGradx(n), wVelGrady(n), wVelGradz(n))
    allocate(BuVelGradx(npatches, nb2), BuVelGrady(npatches, nb2), BuVelGradz(npatches, nb2))
    allocate(BvVelGradx(npatches, nb2), BvVelGrady(npatches, nb2), BvVelGradz(npatches, nb2))
    allocate(BwVelGradx(npatches, nb2), BwVelGrady(npatches, nb2), BwVelGradz(npatches, nb2))
    allocate(uVelGradfx(nif), uVelGradfy(nif), uVelGradfz(nif), vVelGradfx(nif), vVelGradfy(nif))
    allocate(vVelGradfz(nif), wVelGradfx(nif), wVelGradfy(nif), wVelGradfz(nif))
    allocate(xVeldirection(npatches, nb2), yVeldirection(npatches, nb2), zVeldirection(npatches, nb2))
    allocate(Viscosity(n), BViscosity(npatches, nb2))
    allocate(ScMomentumx(n), SbMomentumx(n), ScMomentumy(n), SbMomentumy(n), ScMomentumz(n), SbMomentumz(n))
    allocate(mdot(nif), Bmdot(npatches, nb2), effdiv(n))
    allocate(Du1Velocity(n), Du2Velocity(n), uVelocityStar(n), Dv1Velocity(n), Dv2Velocity(n))
    allocate(vVelocityStar(n), Dw1Velocity(n), Dw2Velocity(n), wVelocityStar(n))
    allocate(MachNumber(n), BMachNumber(npatches, nb2))
    allocate(PressureC(n), BPressureC(npatches, nb2), Pressure(n), BPressure(npatches, nb2))
    allocate(PressGradx(n), PressGrady(n), PressGradz(n))
    allocate(BPressGradx(npatches, nb2), BPressGrady(npatches, nb2), BPressGradz(npatches, nb2))
    allocate(PressCGradx(n), PressCGrady(n), PressCGradz(n))
    allocate(BPressCGradx(npatches, nb2), BPressCGrady(npatches, nb2), BPressCGradz(npatches, nb2))
    allocate(PressGradfx(nif), PressGradfy(nif), PressGradfz(nif))
    allocate(PressCGradfx(nif), PressCGradfy(nif), PressCGradfz(nif))
    allocate(Density(n), Densityf(nif), BDensity(npatches, nb2))
    allocate(DensGradx(n), DensGrady(n), DensGradz(n))
    allocate(BDensGradx(npatches, nb2), BDensGrady(npatches, nb2), BDensGradz(npatches, nb2))
    allocate(dfidxTstar(n), dfidyTstar(n), dfidzTstar(n))
    allocate(GamaFace(nif), BGamaFace(npatches, nb2))
    allocate(FluxCf(nfaces), FluxFf(nfaces), FluxVf(nfaces), FluxTf(nfaces))
    allocate(ac(n), anb(n, w), bc(n), bcOriginal(n), dphi(n), dc(n), rc(n))
    allocate(FluxCE(n), FluxVE(n), FluxTE(n), Sc(n), Sb(n))
    allocate(ResorAbs(18), ResorMax(18), ResorRMS(18), ResorScaled(18))
    allocate(eDiffCoefficient(n), BeDiffCoefficient(npatches, nb2))
    ! the multigrid arrays exist before the first agglomeration, which
    ! deallocates and sizes them (ResetAllocationForMultiGrid.for:4)
    allocate(acMG(1), bcMG(1), anbMG(1), maxanb(1), Residuals(1), dphiMG(1))
    allocate(NumberOfElementsMG(1), NElementFacesMG(1), NBChildrenMaxMG(1), ijbeginE(1), ijbeginN(1))
    allocate(ijbeginNOld(1), ElementNeighborMG(1), NumberofElementNeighborsMG(1), Parents(1))
    allocate(Children(1), NumberOfChildren(1), NeighborOfParent(1), NumberOfNeighborOfParent(1))
    allocate(ElementParent(1))
    uVelocity = 0.; vVelocity = 0.; wVelocity = 0.
    BuVelocity = 0.; BvVelocity = 0.; BwVelocity = 0.
    uVelGradx = 0.; uVelGrady = 0.; uVelGradz = 0.; vVelGradx = 0.; vVelGrady = 0.; vVelGradz = 0.
    wVelGradx = 0.; wVelGrady = 0.; wVelGradz = 0.
    BuVelGradx = 0.; BuVelGrady = 0.; BuVelGradz = 0.; BvVelGradx = 0.; BvVelGrady = 0.; BvVelGradz = 0.
    BwVelGradx = 0.; BwVelGrady = 0.; BwVelGradz = 0.
    uVelGradfx = 0.; uVelGradfy = 0.; uVelGradfz = 0.; vVelGradfx = 0.; vVelGradfy = 0.; vVelGradfz = 0.
    wVelGradfx = 0.; wVelGradfy = 0.; wVelGradfz = 0.
    xVeldirection = 0.; yVeldirection = 0.; zVeldirection = 0.
    Viscosity = 0.; BViscosity = 0.
    ScMomentumx = 0.; SbMomentumx = 0.; ScMomentumy = 0.; SbMomentumy = 0.; ScMomentumz = 0.; SbMomentumz = 0.
    mdot = 0.; Bmdot = 0.; effdiv = 0.
    Du1Velocity = 0.; Du2Velocity = 0.; uVelocityStar = 0.; Dv1Velocity = 0.; Dv2Velocity = 0.
    vVelocityStar = 0.; Dw1Velocity = 0.; Dw2Velocity = 0.; wVelocityStar = 0.
    MachNumber = 0.; BMachNumber = 0.
    PressureC = 0.; BPressureC = 0.; Pressure = 0.; BPressure = 0.
    PressGradx = 0.; PressGrady = 0.; PressGradz = 0.; BPressGradx = 0.; BPressGrady = 0.; BPressGradz = 0.
    PressCGradx = 0.; PressCGrady = 0.; PressCGradz = 0.
    BPressCGradx = 0.; BPressCGrady = 0.; BPressCGradz = 0.
    PressGradfx = 0.; PressGradfy = 0.; PressGradfz = 0.; PressCGradfx = 0.; PressCGradfy = 0.; PressCGradfz = 0.
    Density = ConstantDensity; BDensity = ConstantDensity; Densityf = 0.
    DensGradx = 0.; DensGrady = 0.; DensGradz = 0.; BDensGradx = 0.; BDensGrady = 0.; BDensGradz = 0.
    dfidxTstar = 0.; dfidyTstar = 0.; dfidzTstar = 0.
    GamaFace = 0.; BGamaFace = 0.
    FluxCf = 0.; FluxFf = 0.; FluxVf = 0.; FluxTf = 0.
    ac = 0.; anb = 0.; bc = 0.; bcOriginal = 0.; dphi = 0.; dc = 0.; rc = 0.
    FluxCE = 0.; FluxVE = 0.; FluxTE = 0.; Sc = 0.; Sb = 0.
    ResorAbs = 0.; ResorMax = 0.; ResorRMS = 0.; ResorScaled = 0.
    eDiffCoefficient = 0.; BeDiffCoefficient = 0.
  end subroutine allocate_flow_storage

  ! one face quantity over all faces: interior values, then each patch's
  subroutine put_faces(a, b)
    double precision, intent(in) :: a(:), b(:,:)
    integer :: qq, jj
    write(22) a(1:nif)
    do qq = 1, npatches
      do jj = 1, psize(qq)
        write(22) b(qq, jj)
      end do
    end do
  end subroutine put_faces
end program ref_simple_driver
