! A Fortran host of the GPU path through the FIRTEZ_GPU module: the calls a
! reference build makes in place of FORWARD3D (forward.f90:43-306),
! INVERT3D (invert3d.f90:57-178), GET_PERTURBATION and HESSIAN
! (lm.f90:137-204, 30-39), and the coupled inversion between
! INIT_COUPLED_INVERSION_* and END_COUPLED_INVERSION_* (coupled_inversion.f90).
! Inputs come from the stream file named first on the command line, outputs
! go to the second; the test that writes the input compares the outputs with
! the same entry points called from Python.
PROGRAM FZ_CALLER
  USE ISO_C_BINDING
  USE FIRTEZ_GPU
  IMPLICIT NONE
  TYPE(C_PTR) :: CTX
  TYPE(FZ_LINE_T), ALLOCATABLE :: LINES(:)
  TYPE(FZ_REGION_T), ALLOCATABLE :: REGS(:)
  TYPE(FZ_INVERT_CONTROL_T) :: CTL
  INTEGER(C_INT32_T) :: HDR(8)
  INTEGER :: NL, NR, NZ, NX, NY, NP, NF, NUMW, IERR, ITS, ITSC, U, P, DTYPE, NPSY, NPSX
  REAL(C_DOUBLE), ALLOCATABLE :: Z(:), SYN3D(:,:,:), OBS3D(:,:,:), SYNI(:,:,:), CHI2(:,:)
  REAL(C_DOUBLE), ALLOCATABLE :: IJAC(:,:), DIFF(:), PEN(:), HSSR(:,:), DELTA(:), JAC(:,:), HSS(:,:)
  REAL(C_DOUBLE), ALLOCATABLE :: PSF(:,:), SYNC(:,:,:), CHIC(:,:)
  REAL(C_DOUBLE) :: LAM, TOL
  REAL(C_FLOAT), ALLOCATABLE :: CUBE(:,:,:,:), INV(:,:,:,:), DSYN(:,:,:,:), INVC(:,:,:,:)
  INTEGER(C_INT32_T), ALLOCATABLE :: DONE(:,:)
  CHARACTER(LEN=512) :: FIN, FOUT

  CALL GET_COMMAND_ARGUMENT(1, FIN)
  CALL GET_COMMAND_ARGUMENT(2, FOUT)
  OPEN(NEWUNIT=U, FILE=TRIM(FIN), ACCESS='STREAM', FORM='UNFORMATTED', STATUS='OLD')
  READ(U) HDR
  NL = HDR(1); NR = HDR(2); NZ = HDR(3); NX = HDR(4); NY = HDR(5); NP = HDR(6); NF = HDR(7); DTYPE = HDR(8)
  ALLOCATE(LINES(NL), REGS(NR), Z(NZ), CUBE(NZ, NY, NX, 12), INV(NZ, NY, NX, 12), DONE(NY, NX))
  READ(U) LINES, REGS, Z, CUBE, INV, CTL
  ALLOCATE(IJAC(NF, NP), DIFF(NF), PEN(NP), HSSR(NP, NP), JAC(NP, NF), HSS(NP, NP), DELTA(NP))
  READ(U) IJAC, DIFF, PEN, HSSR, LAM, TOL, JAC
  READ(U) NPSY, NPSX
  ALLOCATE(PSF(NPSY, NPSX))
  READ(U) PSF
  CLOSE(U)
  INVC = INV

  IF (FZ_CREATE(CTX, 0) /= 0) STOP 'fz_create'
  IERR = FZ_CONFIGURE(CTX, LINES, NL, REGS, NR, NZ, Z)
  IF (IERR /= 0) STOP 'fz_configure'
  NUMW = FZ_NUMW(CTX)
  ALLOCATE(SYN3D(4 * NUMW, NY, NX), DSYN(4 * NUMW, 5 * NZ, NY, NX), SYNI(4 * NUMW, NY, NX), CHI2(NY, NX))
  ALLOCATE(OBS3D(4 * NUMW, NY, NX))
  BLOCK
    REAL(C_FLOAT), ALLOCATABLE :: GAIN(:,:)
    ALLOCATE(GAIN(4, NUMW))
    IERR = FZ_GAIN_HSRA(CTX, GAIN)
    IF (IERR /= 0) STOP 'fz_gain_hsra'
  END BLOCK
  DONE = 0

  ! FORWARD3D with spherical response functions
  CALL FZ_FORWARD3D(CTX, NX, NY, NZ, CUBE(:,:,:,1), CUBE(:,:,:,2), CUBE(:,:,:,3), CUBE(:,:,:,4), CUBE(:,:,:,5), &
       CUBE(:,:,:,6), CUBE(:,:,:,7), CUBE(:,:,:,8), CUBE(:,:,:,9), CUBE(:,:,:,10), CUBE(:,:,:,11), &
       CUBE(:,:,:,12), DONE, FZ_RF_SPHERICAL, SYN3D, IERR, DSYN)
  IF (IERR /= 0) THEN
    PRINT *, TRIM(FZ_ERROR_MESSAGE(CTX))
    STOP 'FZ_FORWARD3D'
  END IF

  ! INVERT3D of the synthesis just made, from the perturbed initial model
  OBS3D = SYN3D
  CALL FZ_INVERT3D(CTX, NX, NY, NZ, INV(:,:,:,1), INV(:,:,:,2), INV(:,:,:,3), INV(:,:,:,4), INV(:,:,:,5), &
       INV(:,:,:,6), INV(:,:,:,7), INV(:,:,:,8), INV(:,:,:,9), INV(:,:,:,10), INV(:,:,:,11), INV(:,:,:,12), &
       OBS3D, CTL, SYNI, CHI2, ITS, IERR)
  IF (IERR /= 0) THEN
    PRINT *, TRIM(FZ_ERROR_MESSAGE(CTX))
    STOP 'FZ_INVERT3D'
  END IF

  ! LM seams of lm.f90
  CALL FZ_GET_PERTURBATION(CTX, NP, NF, IJAC, DIFF, LAM, DELTA, PEN, HSSR, DTYPE, IERR, SVDTOL=TOL)
  IF (IERR /= 0) STOP 'FZ_GET_PERTURBATION'
  CALL FZ_HESSIAN(CTX, NP, NF, JAC, HSS, IERR)
  IF (IERR /= 0) STOP 'FZ_HESSIAN'
  CALL FZ_GET_PERTURBATION(CTX, NP, NF, IJAC, DIFF, LAM, DELTA, PEN, HSSR, 3, IERR)
  IF (IERR == 0) STOP 'DTYPE 3 accepted'
  CALL FZ_GET_PERTURBATION(CTX, NP, NF, IJAC, DIFF, LAM, DELTA, PEN, HSSR, DTYPE, IERR, SVDTOL=TOL)

  ! the coupled inversion of the same synthesis through the PSF, from the
  ! same initial model; after END_COUPLED_INVERSION it must be refused
  ALLOCATE(SYNC(4 * NUMW, NY, NX), CHIC(NY, NX))
  CALL FZ_INIT_COUPLED_INVERSION(CTX, NPSY, NPSX, PSF, IERR)
  IF (IERR /= 0) STOP 'FZ_INIT_COUPLED_INVERSION'
  CALL FZ_COUPLED_INVERT3D(CTX, NX, NY, NZ, INVC(:,:,:,1), INVC(:,:,:,2), INVC(:,:,:,3), INVC(:,:,:,4), &
       INVC(:,:,:,5), INVC(:,:,:,6), INVC(:,:,:,7), INVC(:,:,:,8), INVC(:,:,:,9), INVC(:,:,:,10), INVC(:,:,:,11), &
       INVC(:,:,:,12), OBS3D, CTL, SYNC, CHIC, ITSC, IERR)
  IF (IERR /= 0) THEN
    PRINT *, TRIM(FZ_ERROR_MESSAGE(CTX))
    STOP 'FZ_COUPLED_INVERT3D'
  END IF
  CALL FZ_END_COUPLED_INVERSION(CTX, IERR)
  IF (IERR /= 0) STOP 'FZ_END_COUPLED_INVERSION'
  BLOCK
    REAL(C_FLOAT), ALLOCATABLE :: T(:,:,:,:)
    REAL(C_DOUBLE), ALLOCATABLE :: S(:,:,:), C(:,:)
    INTEGER :: IT2
    T = INV
    ALLOCATE(S(4 * NUMW, NY, NX), C(NY, NX))
    CALL FZ_COUPLED_INVERT3D(CTX, NX, NY, NZ, T(:,:,:,1), T(:,:,:,2), T(:,:,:,3), T(:,:,:,4), T(:,:,:,5), &
         T(:,:,:,6), T(:,:,:,7), T(:,:,:,8), T(:,:,:,9), T(:,:,:,10), T(:,:,:,11), T(:,:,:,12), OBS3D, CTL, &
         S, C, IT2, IERR)
    IF (IERR == 0) STOP 'coupled inversion accepted without a PSF'
  END BLOCK

  OPEN(NEWUNIT=U, FILE=TRIM(FOUT), ACCESS='STREAM', FORM='UNFORMATTED', STATUS='REPLACE')
  WRITE(U) NUMW, ITS
  WRITE(U) CUBE, SYN3D, DSYN, INV, SYNI, CHI2, DELTA, HSS
  WRITE(U) ITSC
  WRITE(U) INVC, SYNC, CHIC
  CLOSE(U)
  CALL FZ_DESTROY(CTX)
  PRINT *, 'FZ_CALLER OK', NUMW, ITS
END PROGRAM FZ_CALLER
