!===============================================================================!
! swref: a test driver around the UNCHANGED reference modules of sw-mpdec
! (fv-solver-sw/src/*.f90, compiled where they lie by oracle/Makefile).
!
! This program is test infrastructure.  It replaces only the reference's
! main program and parameter module (src/main.f90, src/parameters.f90): it
! sets the module globals that InitializeParameters would set, installs an
! initial state read from a binary file instead of FillInitialConditions,
! and then drives the reference's own public entry points:
!   FVTimeDerivative  (src/finitevolume.f90:402)   -> mode 1
!   TimeStep          (src/equation.f90:791)       -> modes 2 and 3
!   TimeDiscretization(src/timediscretization.f90:24), one CFL step per call
!                     (tEnd = dt_min, nOutputFiles = 1)  -> mode 2
!   jacobi            (src/jacobi_iteration.f90:52) -> mode 4
!   mode 5: like mode 2, but only the wall time of the step loop (SYSTEM_CLOCK)
!           and the final state are written (CPU baseline timing)
!   mode 6: the initial data (InitializeWBVariables' well-balanced averages in
!           WELLBALANCED builds, then FillInitialConditions) for
!           InitialCondition = nsteps; no state is read
!   mode 7: PostProcessing of the given state for InitialCondition = nsteps at
!           tEnd = t (its ErrorL1_nx_ny.dat in the working directory)
!   mode 8: nsteps steps as mode 2, then V(1:3,1:nx,1:ny), the primitives the
!           output writer prints (those of the last FVTD input)
! With WELLBALANCED builds the well-balanced reference state is installed the
! way InitializeWBVariables does it (src/finitevolume.f90:373-392).
!
! Builds without -DPATANKAR (the explicit schemes) have no mode 4 and write
! no production/destruction arrays in mode 1.
!
! Usage: swref <input.bin> <output.bin>
! Input (stream, little endian):
!   int32 mode, nx, nsteps, bathy_flag, has_wb, ny (0: ny = nx),
!         timescheme (0: 5, mPDeC5; src/parameters.f90:144 numbering),
!         reconstruction_fix (0: 4 = Reconstruction, as src/parameters.f90:142)
!   real64 x0, x1, gravity, cfl    (the domain is [x0,x1]^2; dy = dx*nx/ny)
!   real64 y0, y1                  [modes 6, 7: the domain is [x0,x1] x [y0,y1]]
!   real64 t                       [mode 7: tEnd]
!   real64 U(3,nx,ny)              [mode 1-3]   or  Mass(5nx*ny), Diag(nx*ny), rhs(nx*ny) [mode 4]
!   real64 Uwb(3,nx,ny)            [if has_wb]
!===============================================================================!
PROGRAM swref
USE MOD_FiniteVolume2D_vars
USE MOD_FiniteVolume2D,     ONLY: InitializeFiniteVolume
USE MOD_FiniteVolume2D,     ONLY: FVTimeDerivative
USE MOD_FiniteVolume2D,     ONLY: FillInitialConditions
USE MOD_PostProcessing,     ONLY: PostProcessing
#ifdef WELLBALANCED
USE MOD_FiniteVolume2D,     ONLY: InitializeWBVariables
#endif
USE MOD_Mesh,               ONLY: BuildMesh
USE MOD_Equation,           ONLY: TimeStep
USE MOD_TimeDiscretization, ONLY: TimeDiscretization
#ifdef PATANKAR
USE MOD_JacobiIteration,    ONLY: jacobi
#endif
IMPLICIT NONE
INTEGER            :: mode, nx, ny, nsteps, bflag, haswb, k, nrow, scheme, recfix
INTEGER(KIND=8)    :: c0, c1, crate
REAL               :: x0, x1, y0, y1, grav, cflin, dtmin, tpost
REAL, ALLOCATABLE  :: Uic(:,:,:), Uwb(:,:,:), Mass(:), Diag(:), rhs(:), sol(:)
CHARACTER(LEN=1024):: fin, fout

CALL get_command_argument(1, fin)
CALL get_command_argument(2, fout)

OPEN(10, FILE=TRIM(fin), ACCESS='stream', FORM='unformatted', STATUS='old')
READ(10) mode, nx, nsteps, bflag, haswb, ny, scheme, recfix
IF (ny .LE. 0) ny = nx
IF (scheme .LE. 0) scheme = 5
IF (recfix .LE. 0) recfix = 4
READ(10) x0, x1, grav, cflin
y0 = x0
y1 = x1
tpost = 0.
IF (mode .EQ. 6 .OR. mode .EQ. 7) READ(10) y0, y1
IF (mode .EQ. 7) READ(10) tpost

#ifdef WELLBALANCED
IF (haswb .NE. 1) THEN
  WRITE(*,*) "swref: this build is WELLBALANCED, input must carry a WB state"
  STOP 2
END IF
#else
IF (haswb .NE. 0) THEN
  WRITE(*,*) "swref: this build is not WELLBALANCED"
  STOP 2
END IF
#endif

! Module globals as InitializeParameters (src/parameters.f90:24-171) sets them.
nElemsX = nx
nElemsY = ny
MESH_X0 = (/x0, y0/)
MESH_X1 = (/x1, y1/)
Gravity = grav
CFL     = cflin
BoundaryConditionsType = (/1,1,1,1/)
BathymetryFlag    = bflag
Reconstruction    = 4
ReconstructionFix = recfix   ! src/parameters.f90:142 sets it to Reconstruction
timescheme        = scheme
WhichOutput      = 0   ! no snapshot files
nOutputFiles      = 1
InitialCondition  = 4   ! only consulted by WriteSolutionToDisk's error field

CALL InitializeFiniteVolume()
CALL BuildMesh()

OPEN(11, FILE=TRIM(fout), ACCESS='stream', FORM='unformatted', STATUS='replace')

#ifdef PATANKAR
IF (mode .EQ. 4) THEN
  nrow = nx*ny
  ALLOCATE(Mass(5*nrow), Diag(nrow), rhs(nrow), sol(nrow))
  READ(10) Mass, Diag, rhs
  sol = 0.
  CALL jacobi(RowStart, ColumnsVector, Mass, Diag, rhs, sol)
  WRITE(11) sol
  CLOSE(11)
  CLOSE(10)
  STOP
END IF
#endif

IF (mode .EQ. 6) THEN
  ! The reference's own initial data (src/main.f90:22-26) for
  ! InitialCondition = nsteps on [x0,x1]^2 with BathymetryFlag = bathy_flag:
  ! the well-balanced cell averages of InitializeWBVariables (WELLBALANCED
  ! builds; U is 0 before it, as after InitializeFiniteVolume), then the cell
  ! averages of FillInitialConditions.
  CLOSE(10)
  InitialCondition = nsteps
#ifdef WELLBALANCED
  U = 0.
  CALL InitializeWBVariables()
  WRITE(11) U(1:3,1:nx,1:ny)
#endif
  CALL FillInitialConditions()
  WRITE(11) U(1:3,1:nx,1:ny)
  CLOSE(11)
  STOP
END IF

ALLOCATE(Uic(3,nx,ny))
READ(10) Uic
IF (haswb .EQ. 1) THEN
  ALLOCATE(Uwb(3,nx,ny))
  READ(10) Uwb
END IF
CLOSE(10)

#ifdef WELLBALANCED
U = 0.
U(1:3,1:nx,1:ny) = Uwb
CALL FVTimeDerivative(0.)
UtWB = Ut
FXWB = FX
FYWB = FY
SWB  = S
#endif

U = 0.
U(1:3,1:nx,1:ny) = Uic

SELECT CASE (mode)
  CASE (1)
    CALL FVTimeDerivative(0.)
#ifdef PATANKAR
    WRITE(11) Ut, FX, FY, S, ProductionSparse, DestructionSparse
#else
    WRITE(11) Ut, FX, FY, S
#endif
    WRITE(11) WeightsGPBnd, MESH_DX, MeshGP
  CASE (2)
    DO k = 1, nsteps
      dtmin = TimeStep()
      WRITE(11) dtmin, LambdaMaxX, LambdaMaxY
      ! A non-finite dt would make TimeDiscretization loop forever
      ! (t never reaches tEnd); stop after recording it.
      IF (.NOT. (dtmin .LT. HUGE(1.0))) EXIT
      tEnd = dtmin
      CALL TimeDiscretization()
      WRITE(11) U(1:3,1:nx,1:ny)
    END DO
  CASE (3)
    dtmin = TimeStep()
    WRITE(11) dtmin, LambdaMaxX, LambdaMaxY
  CASE (8)
    ! like mode 2, then the module array V that WriteSolutionToDisk prints
    ! (src/output.f90:79): the primitives of the last FVTD input
    DO k = 1, nsteps
      dtmin = TimeStep()
      IF (.NOT. (dtmin .LT. HUGE(1.0))) EXIT
      tEnd = dtmin
      CALL TimeDiscretization()
    END DO
    WRITE(11) V(1:3,1:nx,1:ny)
  CASE (7)
    ! PostProcessing (src/postprocessing.f90:24-64) of the given state at
    ! tEnd = tpost for InitialCondition = nsteps: it writes ErrorL1_nx_ny.dat
    ! into the working directory.
    InitialCondition = nsteps
    tEnd = tpost
    CALL PostProcessing()
  CASE (5)
    CALL SYSTEM_CLOCK(c0, crate)
    DO k = 1, nsteps
      dtmin = TimeStep()
      IF (.NOT. (dtmin .LT. HUGE(1.0))) EXIT
      tEnd = dtmin
      CALL TimeDiscretization()
    END DO
    CALL SYSTEM_CLOCK(c1)
    WRITE(11) REAL(c1 - c0) / REAL(crate), REAL(k - 1)
    WRITE(11) U(1:3,1:nx,1:ny)
  CASE DEFAULT
    WRITE(*,*) "swref: unknown mode"
    STOP 2
END SELECT
CLOSE(11)

END PROGRAM swref
