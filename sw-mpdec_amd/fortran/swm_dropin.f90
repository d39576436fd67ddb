!===============================================================================!
! swm_dropin.f90 -- the module a maintainer adds to sw-mpdec so that its own
! program runs the mPDeC5 / WENO5 time step on the GPU through libswm
! (include/swm.h, bound by swm_bridge.f90).  The reference's routines call it
! from four places (the edits oracle/dropin_patch.py applies to a copy of the
! reference's sources):
!
!   TimeDiscretization (src/timediscretization.f90:57-95)
!     dt_min = TimeStep()              -> swm_dropin_timestep()
!     CASE(5) TimeDiscretizationByMPDeC5(t) -> swm_dropin_step(t)
!     before WriteSolutionToDisk(t)    -> swm_dropin_output()  (U and the writer's V)
!   InitializeWBVariables (src/finitevolume.f90:339-396), at its end
!                                      -> swm_dropin_init_wb()
!   FinalizeFiniteVolume (src/finitevolume.f90:768), at its start
!                                      -> swm_dropin_finalize()
!
! Each call site keeps the reference's own path behind swm_dropin_on(): the
! environment variable SWM_DROPIN=0 selects it, anything else (or unset) the
! GPU.  The time loop, its dt clipping, the output and PostProcessing stay in
! the reference's Fortran.  The handle is created at the first call, after
! BuildMesh has set MESH_DX, WeightsGPBnd and MeshGP; the bathymetry tables
! are the reference's own Bathymetry_X/_Y at MeshGP, so the device source term
! uses the host's libm values.  With SWM_DROPIN_DUMP set, swm_dropin_finalize
! writes the final U(1:3,1:nElemsX,1:nElemsY) to that file (raw float64), in
! either mode, so two runs can be compared bit for bit.
!===============================================================================!
MODULE MOD_SwmDropin
USE ISO_C_BINDING
USE swm_bridge
IMPLICIT NONE
PRIVATE
PUBLIC :: swm_dropin_on, swm_dropin_timestep, swm_dropin_step, swm_dropin_output
PUBLIC :: swm_dropin_init_wb, swm_dropin_finalize

TYPE(C_PTR), SAVE :: swmh = C_NULL_PTR
LOGICAL, SAVE     :: decided = .FALSE., gpu = .TRUE., uploaded = .FALSE.

CONTAINS

! The switch of every call site: .FALSE. when SWM_DROPIN=0.
LOGICAL FUNCTION swm_dropin_on()
  CHARACTER(LEN=8) :: val
  INTEGER          :: n, st
  IF (.NOT. decided) THEN
    CALL GET_ENVIRONMENT_VARIABLE("SWM_DROPIN", val, n, st)
    gpu = .NOT. (st == 0 .AND. n > 0 .AND. val(1:1) == '0')
    decided = .TRUE.
  END IF
  swm_dropin_on = gpu
END FUNCTION swm_dropin_on

! The handle, created on first use: swm_create replaces the allocations of
! InitializeFiniteVolume, and swm_set_bathy_grad takes Bathymetry_X/_Y at the
! 4x4 Gauss points of every cell, as SourceFunc evaluates them
! (src/equation.f90:398-413).
SUBROUTINE ensure_handle()
  USE MOD_FiniteVolume2D_vars, ONLY: nElemsX, nElemsY, MESH_DX, Gravity, CFL
  USE MOD_FiniteVolume2D_vars, ONLY: WeightsGPBnd, MeshGP, nGPs, BathymetryFlag
  USE MOD_Equation,            ONLY: Bathymetry_X, Bathymetry_Y
  REAL(C_DOUBLE), ALLOCATABLE :: bx(:,:,:,:), by(:,:,:,:)
  INTEGER(C_INT) :: wb
  INTEGER        :: ii, jj, xg, yg
  IF (C_ASSOCIATED(swmh)) RETURN
  wb = 0
#ifdef WELLBALANCED
  wb = 1
#endif
  CALL swm_check(C_NULL_PTR, swm_create(INT(nElemsX, C_INT), INT(nElemsY, C_INT), &
                 MESH_DX(1), MESH_DX(2), Gravity, CFL, wb, WeightsGPBnd, swmh), "swm_create")
  ! the writer prints V of each step's last FVTD input (src/output.f90:78)
  CALL swm_check(swmh, swm_keep_output_state(swmh, 1_C_INT), "swm_keep_output_state")
  IF (BathymetryFlag > 0) THEN
    ALLOCATE(bx(nElemsX, nElemsY, nGPs, nGPs), by(nElemsX, nElemsY, nGPs, nGPs))
    DO yg = 1, nGPs
      DO xg = 1, nGPs
        DO jj = 1, nElemsY
          DO ii = 1, nElemsX
            bx(ii, jj, xg, yg) = Bathymetry_X(MeshGP(:, ii, jj, xg, yg))
            by(ii, jj, xg, yg) = Bathymetry_Y(MeshGP(:, ii, jj, xg, yg))
          END DO
        END DO
      END DO
    END DO
    CALL swm_check(swmh, swm_set_bathy_grad(swmh, bx, by), "swm_set_bathy_grad")
    DEALLOCATE(bx, by)
  END IF
END SUBROUTINE ensure_handle

! InitializeWBVariables: U holds the well-balanced state; its FVTD is taken on
! the device (FXWB, FYWB, SWB).
SUBROUTINE swm_dropin_init_wb()
  USE MOD_FiniteVolume2D_vars, ONLY: U
  CALL ensure_handle()
  CALL swm_check(swmh, swm_init_wb(swmh, U, 3_C_INT), "swm_init_wb")
END SUBROUTINE swm_dropin_init_wb

! TimeStep() (src/equation.f90:791-837) of the device state; the first call
! hands the initial state U (FillInitialConditions) to the device.
REAL FUNCTION swm_dropin_timestep()
  USE MOD_FiniteVolume2D_vars, ONLY: U, LambdaMaxX, LambdaMaxY
  REAL(C_DOUBLE) :: dt_min, lx, ly
  CALL ensure_handle()
  IF (.NOT. uploaded) THEN
    CALL swm_check(swmh, swm_upload(swmh, U, 3_C_INT), "swm_upload")
    uploaded = .TRUE.
  END IF
  CALL swm_check(swmh, swm_timestep(swmh, dt_min, lx, ly), "swm_timestep")
  LambdaMaxX = lx
  LambdaMaxY = ly
  swm_dropin_timestep = dt_min
END FUNCTION swm_dropin_timestep

! TimeDiscretizationByMPDeC5(t) (src/timediscretization.f90:458-566) with the
! module's dt; a non-converged Jacobi solve is reported and the run goes on,
! as src/jacobi_iteration.f90:83-85 does.
SUBROUTINE swm_dropin_step(t)
  USE MOD_FiniteVolume2D_vars, ONLY: dt
  REAL, INTENT(IN) :: t
  INTEGER(C_INT)   :: rc, iters(13)
  rc = swm_step_mpdec5(swmh, t, dt, iters)
  CALL swm_check(swmh, rc, "swm_step_mpdec5")
  IF (rc > 0) PRINT*, "Jacobi not converged in ", rc, " solves"
END SUBROUTINE swm_dropin_step

! Before WriteSolutionToDisk: the state U and the writer's V from the device.
SUBROUTINE swm_dropin_output()
  USE MOD_FiniteVolume2D_vars, ONLY: U, V, nElemsX, nElemsY
  REAL(C_DOUBLE), ALLOCATABLE :: W(:,:,:)
  ALLOCATE(W(3, nElemsX, nElemsY))
  CALL swm_check(swmh, swm_download(swmh, W), "swm_download")
  U(1:3, 1:nElemsX, 1:nElemsY) = W
  CALL swm_check(swmh, swm_download_output_primitive(swmh, W), "swm_download_output_primitive")
  V(1:3, 1:nElemsX, 1:nElemsY) = W
  DEALLOCATE(W)
END SUBROUTINE swm_dropin_output

! FinalizeFiniteVolume: optional dump of the final U, then the handle goes.
SUBROUTINE swm_dropin_finalize()
  USE MOD_FiniteVolume2D_vars, ONLY: U, nElemsX, nElemsY
  CHARACTER(LEN=1024) :: path
  INTEGER             :: n, st, unit
  CALL GET_ENVIRONMENT_VARIABLE("SWM_DROPIN_DUMP", path, n, st)
  IF (st == 0 .AND. n > 0) THEN
    OPEN(NEWUNIT=unit, FILE=TRIM(path), ACCESS='STREAM', FORM='UNFORMATTED', &
         STATUS='REPLACE')
    WRITE(unit) U(1:3, 1:nElemsX, 1:nElemsY)
    CLOSE(unit)
  END IF
  IF (C_ASSOCIATED(swmh)) CALL swm_destroy(swmh)
  swmh = C_NULL_PTR
END SUBROUTINE swm_dropin_finalize

END MODULE MOD_SwmDropin
!===============================================================================!
