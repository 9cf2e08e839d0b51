! Test-infrastructure driver for the reference's gradient limiter and
! high-resolution convection correction.
!
! Loads one mesh slice (topology, centroids, distances, interpolation factors)
! and one field with its gradient from a little-endian stream file, installs
! them in the reference's module globals and calls the reference's own
! routines:
!   mode 1: LimitGradientVariable(..., method)   (LimitGradients.for:3)
!   mode 2: HRSchemesCorrections('velx', scheme, bleed, framework, ...)
!           (AssembleHRSchemeCorrection.for:4) on a given FluxTf
! Output: mode 1 the limited cell and boundary gradients, mode 2 FluxTf.
! This program links against objects compiled from the reference sources; it is a
! checker only and never part of the product.
program ref_conv_driver
  use Geometry1, only: NumberOfElements, NumberOfBCSets
  use Geometry3, only: NIFaces, NIFaceOwner, NIFaceNeighbor, NBFaces, NBFaceOwner, &
                       NGlobalEFaces, ElementNeighbor, NumberofElementNeighbors, NElementFaces
  use Geometry4, only: xc, yc, zc, FaceCentroidx, FaceCentroidy, FaceCentroidz, &
                       BFaceCentroidx, BFaceCentroidy, BFaceCentroidz, GFactorCF, &
                       DistanceCFx, DistanceCFy, DistanceCFz
  use Variables1, only: mdot
  use Variables2, only: FluxTf
  implicit none
  interface
    SUBROUTINE LimitGradientVariable(FiT,dfidxT,dfidyT,dfidzT, &
                                     BFiT,BdfidxT,BdfidyT,BdfidzT,LimitGradientMethod)
      integer :: LimitGradientMethod
      double precision, dimension(:) :: FiT, dfidxT, dfidyT, dfidzT
      double precision, dimension(:,:) :: BFiT, BdfidxT, BdfidyT, BdfidzT
    end SUBROUTINE LimitGradientVariable
    SUBROUTINE HRSchemesCorrections(Variable,ConvectionScheme, &
                                    Bleed,HRFramework,FiT,dfidxT,dfidyT,dfidzT)
      character*20 ConvectionScheme
      character*10 Variable
      character*4 HRFramework
      double precision :: Bleed
      double precision, dimension(:) :: FiT, dfidxT, dfidyT, dfidzT
    end SUBROUTINE HRSchemesCorrections
  end interface
  character(len=1024) :: fin, fout
  character(len=20) :: scheme
  character(len=4) :: framework
  character(len=10) :: variable
  integer :: n, w, nif, nbc, nbfmax, mode, method
  double precision :: bleed
  double precision, allocatable :: fit(:), gx(:), gy(:), gz(:)
  double precision, allocatable :: bfit(:,:), bgx(:,:), bgy(:,:), bgz(:,:)

  call get_command_argument(1, fin)
  call get_command_argument(2, fout)
  open(unit=21, file=trim(fin), access='stream', form='unformatted', status='old')
  read(21) n, w, nif, nbc, nbfmax, mode, method
  read(21) scheme, framework
  read(21) bleed
  NumberOfElements = n
  NElementFaces = w
  NIFaces = nif
  NumberOfBCSets = nbc
  allocate(NBFaces(nbc), NBFaceOwner(nbc, nbfmax), NumberofElementNeighbors(n))
  allocate(ElementNeighbor(n, w), NGlobalEFaces(n, w), NIFaceOwner(nif), NIFaceNeighbor(nif))
  allocate(xc(n), yc(n), zc(n), FaceCentroidx(nif), FaceCentroidy(nif), FaceCentroidz(nif))
  allocate(BFaceCentroidx(nbc, nbfmax), BFaceCentroidy(nbc, nbfmax), BFaceCentroidz(nbc, nbfmax))
  allocate(GFactorCF(nif), DistanceCFx(nif), DistanceCFy(nif), DistanceCFz(nif), mdot(nif), FluxTf(nif))
  allocate(fit(n), gx(n), gy(n), gz(n))
  allocate(bfit(nbc, nbfmax), bgx(nbc, nbfmax), bgy(nbc, nbfmax), bgz(nbc, nbfmax))
  read(21) NBFaces
  read(21) NBFaceOwner
  read(21) NumberofElementNeighbors
  read(21) ElementNeighbor
  read(21) NGlobalEFaces
  read(21) NIFaceOwner
  read(21) NIFaceNeighbor
  read(21) xc, yc, zc
  read(21) FaceCentroidx, FaceCentroidy, FaceCentroidz
  read(21) BFaceCentroidx, BFaceCentroidy, BFaceCentroidz
  read(21) GFactorCF, DistanceCFx, DistanceCFy, DistanceCFz
  read(21) mdot, FluxTf
  read(21) fit, gx, gy, gz
  read(21) bfit, bgx, bgy, bgz
  close(21)

  open(unit=22, file=trim(fout), access='stream', form='unformatted', status='replace')
  if (mode == 1) then
    call LimitGradientVariable(fit, gx, gy, gz, bfit, bgx, bgy, bgz, method)
    write(22) gx, gy, gz, bgx, bgy, bgz
  else
    variable = 'velx'
    call HRSchemesCorrections(variable, scheme, bleed, framework, fit, gx, gy, gz)
    write(22) FluxTf
  end if
  close(22)
end program ref_conv_driver
