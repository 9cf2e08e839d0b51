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
!           With iturb = 1 the flow is turbulent with the k-omega SST model
!           (ModelNumber 14, low-Reynolds-number walls, WallBCModel 'menter')
!           from the k, omega, mu_t and wall distance the input gives, and
!           every iteration ends with CalculateEffectsOfTurbulence
!           (SolveSelectedTurbulenceModelEquations.for:4,150); k, omega and
!           mu_t are written after the flow fields.
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
  use Turbulence1
  use WallDistance1, only: WallDistance, BWallDistance, MethodCalculateNormalDistance, WallResiduals
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
  integer :: gmm, gmi, gcm, gci, gim, gic, iper, bcalc, biters, bevery, iturb, hasu0
  double precision, allocatable :: u0in(:), v0in(:), w0in(:)
  integer :: tint(15)
  double precision :: tdbl(6)
  double precision, allocatable :: tk0(:), tw0(:), tmut0(:), twd(:), tbk0(:), tbw0(:), tbmut0(:)
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
             gmm, gmi, gcm, gci, gim, gic, iper, iturb
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
    if (iturb /= 0) then
      ! k and omega: solver, itmax (each), HR framework and scheme (each),
      ! MethodCalcGradient and nIterGradient (each), gradient limiter method
      ! (each, 0 off), the wall distance (0 the input's, 1 or 2 the
      ! reference's MethodCalculateNormalDistance); urfTKE, urfTOmega,
      ! urfTViscosity, rrFTKE, rrFTOmega, MaximumViscosityRatio
      read(21) tint
      read(21) tdbl
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
    if (iturb /= 0) then
      allocate(tk0(ncells), tw0(ncells), tmut0(ncells), twd(ncells))
      allocate(tbk0(nfaces - nif), tbw0(nfaces - nif), tbmut0(nfaces - nif))
      read(21) tk0, tw0, tmut0, twd
      read(21) tbk0, tbw0, tbmut0
    end if
    ! an initial velocity field (hasu0 = 1) instead of rest
    read(21) hasu0
    if (hasu0 /= 0) then
      allocate(u0in(ncells), v0in(ncells), w0in(ncells))
      read(21) u0in, v0in, w0in
    end if
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
    if (iturb /= 0) then
      write(22) TurbulentKE, TurbulentOmega, TurbulentViscosity
      call put_bfaces(BTurbulentKE)
      call put_bfaces(BTurbulentOmega)
      call put_bfaces(BTurbulentViscosity)
      write(22) F1factor, WallDistance
    end if
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
    LTurbulentFlow = iturb /= 0
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
    if (iturb /= 0) call set_turbulence_case
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
    if (hasu0 /= 0) then
      uVelocity = u0in; vVelocity = v0in; wVelocity = w0in
    end if
    if (iturb /= 0) call allocate_turbulence_storage
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
      if (iturb /= 0) call CalculateEffectsOfTurbulence
    end do
  end subroutine run_simple

  ! The k-omega SST case: the model switches of the case file
  ! (ProblemDescriptionRoutine.for) and the model constants
  ! (SetTurbulenceModelCoefficients, SetModelCoefficients.for:266-300).
  subroutine set_turbulence_case
    character(len=6), parameter :: tsolvers(0:5) = (/ 'sor   ', 'ilu   ', 'pbcg  ', 'direct', 'gmres ', 'bicgs ' /)
    TurbulenceModel = 'komegasst'
    WallTreatment = 'lowreynoldsnumber'
    WallBCModel = 'menter'
    LRough = .false.
    LRotation = .false.
    LKOmegaSSTRotationCurvatureCorrection = .false.
    LCompressibilityTemperatureCorrection = .false.
    LBuoyancyCorrection = .false.
    LCompressibilityCorrection = .false.
    LimitTurbulenceProduction = .true.
    LSolveTurbulenceKineticEnergy = .true.
    LSolveTurbulenceSpecificDissipationRate = .true.
    LSolveTurbulenceDissipationRate = .false.
    ASSolverTKE = tsolvers(tint(1))
    ASIterTKE = tint(2)
    ASSolverTOmega = tsolvers(tint(3))
    ASIterTOmega = tint(4)
    HRFrameworkTKE = frameworks(tint(5))
    ConvectionSchemeTKE = schemes(tint(6))
    HRFrameworkTOmega = frameworks(tint(7))
    ConvectionSchemeTOmega = schemes(tint(8))
    MethodCalcGradientTKE = tint(9)
    nIterGradientTKE = tint(10)
    MethodCalcGradientTOmega = tint(11)
    nIterGradientTOmega = tint(12)
    LimitGradientTKE = tint(13) > 0
    LimitGradientTKEMethod = max(tint(13), 1)
    LimitGradientTOmega = tint(14) > 0
    LimitGradientTOmegaMethod = max(tint(14), 1)
    BleedTKE = 0.
    BleedTOmega = 0.
    LMultigridTKE = .false.
    LMultigridTOmega = .false.
    LRelaxTKE = .true.
    LRelaxTOmega = .true.
    LFalseTransientTKE = .false.
    LFalseTransientTOmega = .false.
    urfTKE = tdbl(1)
    urfTOmega = tdbl(2)
    urfTViscosity = tdbl(3)
    rrFTKE = tdbl(4)
    rrFTOmega = tdbl(5)
    MaximumViscosityRatio = tdbl(6)
    GradientInterpolationSchemeTKE = 'average'
    GradientInterpolationSchemeTOmega = 'average'
    InterpolationSchemeGamaTKE = 'average'
    InterpolationSchemeGamaTOmega = 'average'
    call SetTurbulenceModelCoefficients
  end subroutine set_turbulence_case

  ! What allocateStorage holds for the k-omega SST flow
  ! (AllocateStorageRoutine.for:793-846,1239-1262,1312-1397), zeroed as it
  ! zeroes it (:2195-2243,2618-2631), with fr1Coefficient 1; then the
  ! initial k, omega, mu_t and wall distance of the input.
  subroutine allocate_turbulence_storage
    integer :: n, nb2, qq, jj, ff
    n = ncells
    nb2 = nbfmax
    allocate(rhok(n), Brhok(npatches, nb2), drhokdx(n), drhokdy(n), drhokdz(n))
    allocate(Bdrhokdx(npatches, nb2), Bdrhokdy(npatches, nb2), Bdrhokdz(npatches, nb2))
    allocate(rhoTED(n), BrhoTED(npatches, nb2), ReT(n), BReT(npatches, nb2))
    allocate(S11(n), S12(n), S13(n), S22(n), S23(n), S33(n))
    allocate(BS11(npatches, nb2), BS12(npatches, nb2), BS13(npatches, nb2), BS22(npatches, nb2))
    allocate(BS23(npatches, nb2), BS33(npatches, nb2), StrainRate(n), BStrainRate(npatches, nb2))
    allocate(W11(n), W12(n), W13(n), W22(n), W23(n), W33(n))
    allocate(BW11(npatches, nb2), BW12(npatches, nb2), BW13(npatches, nb2), BW22(npatches, nb2))
    allocate(BW23(npatches, nb2), BW33(npatches, nb2), Vorticity(n), BVorticity(npatches, nb2))
    allocate(Tau11(n), Tau12(n), Tau13(n), Tau22(n), Tau23(n), Tau33(n))
    allocate(TurbulentKE(n), BTurbulentKE(npatches, nb2), TKEGradx(n), TKEGrady(n), TKEGradz(n))
    allocate(BTKEGradx(npatches, nb2), BTKEGrady(npatches, nb2), BTKEGradz(npatches, nb2))
    allocate(TKEGradfx(nif), TKEGradfy(nif), TKEGradfz(nif), ScTKE(n), SbTKE(n))
    allocate(TurbulenceProduction(n), BTurbulenceProduction(npatches, nb2))
    allocate(TurbulentOmega(n), BTurbulentOmega(npatches, nb2), TOmegaGradx(n), TOmegaGrady(n))
    allocate(TOmegaGradz(n), BTOmegaGradx(npatches, nb2), BTOmegaGrady(npatches, nb2))
    allocate(BTOmegaGradz(npatches, nb2), TOmegaGradfx(nif), TOmegaGradfy(nif), TOmegaGradfz(nif))
    allocate(ScTOmega(n), SbTOmega(n))
    allocate(F1factor(n), BF1factor(npatches, nb2), fr1Coefficient(n), F2factor(n))
    allocate(BF2factor(npatches, nb2), F3factor(n), BF3factor(npatches, nb2), F4factor(n))
    allocate(BF4factor(npatches, nb2))
    allocate(TurbulentViscosity(n), BTurbulentViscosity(npatches, nb2))
    rhok = 0.; Brhok = 0.; drhokdx = 0.; drhokdy = 0.; drhokdz = 0.
    Bdrhokdx = 0.; Bdrhokdy = 0.; Bdrhokdz = 0.; rhoTED = 0.; BrhoTED = 0.; ReT = 0.; BReT = 0.
    S11 = 0.; S12 = 0.; S13 = 0.; S22 = 0.; S23 = 0.; S33 = 0.
    BS11 = 0.; BS12 = 0.; BS13 = 0.; BS22 = 0.; BS23 = 0.; BS33 = 0.; StrainRate = 0.; BStrainRate = 0.
    W11 = 0.; W12 = 0.; W13 = 0.; W22 = 0.; W23 = 0.; W33 = 0.
    BW11 = 0.; BW12 = 0.; BW13 = 0.; BW22 = 0.; BW23 = 0.; BW33 = 0.; Vorticity = 0.; BVorticity = 0.
    Tau11 = 0.; Tau12 = 0.; Tau13 = 0.; Tau22 = 0.; Tau23 = 0.; Tau33 = 0.
    TKEGradx = 0.; TKEGrady = 0.; TKEGradz = 0.; BTKEGradx = 0.; BTKEGrady = 0.; BTKEGradz = 0.
    TKEGradfx = 0.; TKEGradfy = 0.; TKEGradfz = 0.; ScTKE = 0.; SbTKE = 0.
    TurbulenceProduction = 0.; BTurbulenceProduction = 0.
    TOmegaGradx = 0.; TOmegaGrady = 0.; TOmegaGradz = 0.; BTOmegaGradx = 0.; BTOmegaGrady = 0.
    BTOmegaGradz = 0.; TOmegaGradfx = 0.; TOmegaGradfy = 0.; TOmegaGradfz = 0.; ScTOmega = 0.; SbTOmega = 0.
    F1factor = 0.; BF1factor = 0.; fr1Coefficient = 1.; F2factor = 0.; BF2factor = 0.
    F3factor = 0.; BF3factor = 0.; F4factor = 0.; BF4factor = 0.
    BTurbulentKE = 0.; BTurbulentOmega = 0.; BTurbulentViscosity = 0.
    TurbulentKE = tk0
    TurbulentOmega = tw0
    TurbulentViscosity = tmut0
    ! the wall distance and the nearest wall face of every cell (iTau, jTau)
    ! from the reference's CalculateNormalDistanceToWall
    ! (CalculateDistanceToWall.for:4), which writes them to unit 15
    ! (3: method 2 run to an outer residual of 1e-13 instead of WallResiduals
    ! = 1e-10, WallDistance1.F90:12)
    MethodCalculateNormalDistance = min(max(tint(15), 1), 2)
    if (tint(15) == 3) WallResiduals = 1.d-13
    open(unit=15, status='scratch')
    call CalculateNormalDistanceToWall
    close(15)
    if (tint(15) == 0) WallDistance = twd
    ff = 0
    do qq = 1, npatches
      do jj = 1, psize(qq)
        ff = ff + 1
        BTurbulentKE(qq, jj) = tbk0(ff)
        BTurbulentOmega(qq, jj) = tbw0(ff)
        BTurbulentViscosity(qq, jj) = tbmut0(ff)
      end do
    end do
  end subroutine allocate_turbulence_storage

  ! a boundary-face array in boundary-face order (patch by patch)
  subroutine put_bfaces(b)
    double precision, intent(in) :: b(:,:)
    integer :: qq, jj
    do qq = 1, npatches
      do jj = 1, psize(qq)
        write(22) b(qq, jj)
      end do
    end do
  end subroutine put_bfaces

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
