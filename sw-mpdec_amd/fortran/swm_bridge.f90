!===============================================================================!
! swm_bridge: ISO_C_BINDING interfaces to libswm (include/swm.h), the MI355X
! implementation of the PATANKAR-build mPDeC5 / WENO5 time step of sw-mpdec.
!
! Built with the reference's -fdefault-real-8, so REAL is C double; default
! INTEGER is C int.  Array arguments are passed by reference in the
! reference's own layouts (U(1:3,-2:nx+3,-2:ny+3) with ghosts = 3, or
! U(1:3,1:nx,1:ny) with ghosts = 0).
!===============================================================================!
MODULE swm_bridge
USE ISO_C_BINDING
IMPLICIT NONE
PRIVATE

INTEGER(C_INT), PARAMETER, PUBLIC :: SWM_OK = 0

PUBLIC :: swm_create, swm_destroy, swm_set_bathy_grad, swm_init_wb
PUBLIC :: swm_upload, swm_download, swm_timestep, swm_step_mpdec5, swm_advance
PUBLIC :: swm_step_mpeuler, swm_step_mpdec2, swm_set_timescheme, swm_set_hp_association
PUBLIC :: swm_set_reconstruction_fix, swm_fill_initial_data, swm_l1_error, swm_download_primitive
PUBLIC :: swm_keep_output_state, swm_download_output_primitive
PUBLIC :: swm_step_forward_euler, swm_step_ssprk4, swm_step_rk65, swm_step_dec5
PUBLIC :: swm_error_message, swm_check

INTERFACE
  ! replaces the allocations of InitializeFiniteVolume (src/finitevolume.f90:44)
  INTEGER(C_INT) FUNCTION swm_create(nx, ny, dx, dy, g, cfl, wb, wgp, h) &
      BIND(C, NAME='swm_create')
    IMPORT :: C_INT, C_DOUBLE, C_PTR
    INTEGER(C_INT), VALUE :: nx, ny, wb
    REAL(C_DOUBLE), VALUE :: dx, dy, g, cfl
    REAL(C_DOUBLE), INTENT(IN) :: wgp(4)
    TYPE(C_PTR), INTENT(OUT) :: h
  END FUNCTION

  SUBROUTINE swm_destroy(h) BIND(C, NAME='swm_destroy')
    IMPORT :: C_PTR
    TYPE(C_PTR), VALUE :: h
  END SUBROUTINE

  ! Bathymetry_X/_Y at MeshGP, arrays (1:nx,1:ny,1:4,1:4) (src/equation.f90:451-509)
  INTEGER(C_INT) FUNCTION swm_set_bathy_grad(h, bx, by) BIND(C, NAME='swm_set_bathy_grad')
    IMPORT :: C_INT, C_DOUBLE, C_PTR
    TYPE(C_PTR), VALUE :: h
    REAL(C_DOUBLE), INTENT(IN) :: bx(*), by(*)
  END FUNCTION

  ! InitializeWBVariables (src/finitevolume.f90:339-396)
  INTEGER(C_INT) FUNCTION swm_init_wb(h, Uwb, ghosts) BIND(C, NAME='swm_init_wb')
    IMPORT :: C_INT, C_DOUBLE, C_PTR
    TYPE(C_PTR), VALUE :: h
    REAL(C_DOUBLE), INTENT(IN) :: Uwb(*)
    INTEGER(C_INT), VALUE :: ghosts
  END FUNCTION

  INTEGER(C_INT) FUNCTION swm_upload(h, U, ghosts) BIND(C, NAME='swm_upload')
    IMPORT :: C_INT, C_DOUBLE, C_PTR
    TYPE(C_PTR), VALUE :: h
    REAL(C_DOUBLE), INTENT(IN) :: U(*)
    INTEGER(C_INT), VALUE :: ghosts
  END FUNCTION

  INTEGER(C_INT) FUNCTION swm_download(h, U) BIND(C, NAME='swm_download')
    IMPORT :: C_INT, C_DOUBLE, C_PTR
    TYPE(C_PTR), VALUE :: h
    REAL(C_DOUBLE), INTENT(OUT) :: U(*)
  END FUNCTION

  ! TimeStep (src/equation.f90:791-837)
  INTEGER(C_INT) FUNCTION swm_timestep(h, dt_min, lmx, lmy) BIND(C, NAME='swm_timestep')
    IMPORT :: C_INT, C_DOUBLE, C_PTR
    TYPE(C_PTR), VALUE :: h
    REAL(C_DOUBLE), INTENT(OUT) :: dt_min, lmx, lmy
  END FUNCTION

  ! TimeDiscretizationByMPDeC5 (src/timediscretization.f90:458-566)
  INTEGER(C_INT) FUNCTION swm_step_mpdec5(h, t, dt, iters) BIND(C, NAME='swm_step_mpdec5')
    IMPORT :: C_INT, C_DOUBLE, C_PTR
    TYPE(C_PTR), VALUE :: h
    REAL(C_DOUBLE), VALUE :: t, dt
    INTEGER(C_INT), INTENT(OUT) :: iters(13)
  END FUNCTION

  ! TimeDiscretizationByMPEuler (src/timediscretization.f90:671-764)
  INTEGER(C_INT) FUNCTION swm_step_mpeuler(h, t, dt, iters) BIND(C, NAME='swm_step_mpeuler')
    IMPORT :: C_INT, C_DOUBLE, C_PTR
    TYPE(C_PTR), VALUE :: h
    REAL(C_DOUBLE), VALUE :: t, dt
    INTEGER(C_INT), INTENT(OUT) :: iters(1)
  END FUNCTION

  ! TimeDiscretizationByMPDeC2 (src/timediscretization.f90:571-664)
  INTEGER(C_INT) FUNCTION swm_step_mpdec2(h, t, dt, iters) BIND(C, NAME='swm_step_mpdec2')
    IMPORT :: C_INT, C_DOUBLE, C_PTR
    TYPE(C_PTR), VALUE :: h
    REAL(C_DOUBLE), VALUE :: t, dt
    INTEGER(C_INT), INTENT(OUT) :: iters(2)
  END FUNCTION

  ! the explicit schemes as the build without PATANKAR runs them:
  ! TimeDiscretizationByForwardEuler/BySSPRK4/ByRK65/ByDeC5
  ! (src/timediscretization.f90:198-232, 104-192, 238-356, 362-451)
  INTEGER(C_INT) FUNCTION swm_step_forward_euler(h, t, dt) BIND(C, NAME='swm_step_forward_euler')
    IMPORT :: C_INT, C_DOUBLE, C_PTR
    TYPE(C_PTR), VALUE :: h
    REAL(C_DOUBLE), VALUE :: t, dt
  END FUNCTION
  INTEGER(C_INT) FUNCTION swm_step_ssprk4(h, t, dt) BIND(C, NAME='swm_step_ssprk4')
    IMPORT :: C_INT, C_DOUBLE, C_PTR
    TYPE(C_PTR), VALUE :: h
    REAL(C_DOUBLE), VALUE :: t, dt
  END FUNCTION
  INTEGER(C_INT) FUNCTION swm_step_rk65(h, t, dt) BIND(C, NAME='swm_step_rk65')
    IMPORT :: C_INT, C_DOUBLE, C_PTR
    TYPE(C_PTR), VALUE :: h
    REAL(C_DOUBLE), VALUE :: t, dt
  END FUNCTION
  INTEGER(C_INT) FUNCTION swm_step_dec5(h, t, dt) BIND(C, NAME='swm_step_dec5')
    IMPORT :: C_INT, C_DOUBLE, C_PTR
    TYPE(C_PTR), VALUE :: h
    REAL(C_DOUBLE), VALUE :: t, dt
  END FUNCTION

  ! timescheme (src/parameters.f90:144) of swm_advance: 1..7
  INTEGER(C_INT) FUNCTION swm_set_timescheme(h, timescheme) BIND(C, NAME='swm_set_timescheme')
    IMPORT :: C_INT, C_PTR
    TYPE(C_PTR), VALUE :: h
    INTEGER(C_INT), VALUE :: timescheme
  END FUNCTION

  ! 0: the reference's depths for timescheme 6 and 7; 1: by node
  INTEGER(C_INT) FUNCTION swm_set_hp_association(h, mode) BIND(C, NAME='swm_set_hp_association')
    IMPORT :: C_INT, C_PTR
    TYPE(C_PTR), VALUE :: h
    INTEGER(C_INT), VALUE :: mode
  END FUNCTION

  ! FillInitialConditions, InitializeWBVariables and the bathymetry tables
  ! (src/main.f90:22-26) on the device: InitialCondition, BathymetryFlag and
  ! MESH_X0, MESH_X1; replaces swm_set_bathy_grad + swm_init_wb + swm_upload
  INTEGER(C_INT) FUNCTION swm_fill_initial_data(h, ic, bflag, x0, x1) &
      BIND(C, NAME='swm_fill_initial_data')
    IMPORT :: C_INT, C_DOUBLE, C_PTR
    TYPE(C_PTR), VALUE :: h
    INTEGER(C_INT), VALUE :: ic, bflag
    REAL(C_DOUBLE), INTENT(IN) :: x0(2), x1(2)
  END FUNCTION

  ! PostProcessing (src/postprocessing.f90:24-64): L1 error against
  ! ExactFunction(ic, t) on the device
  INTEGER(C_INT) FUNCTION swm_l1_error(h, ic, bflag, t, x0, x1, err) &
      BIND(C, NAME='swm_l1_error')
    IMPORT :: C_INT, C_DOUBLE, C_PTR
    TYPE(C_PTR), VALUE :: h
    INTEGER(C_INT), VALUE :: ic, bflag
    REAL(C_DOUBLE), VALUE :: t
    REAL(C_DOUBLE), INTENT(IN) :: x0(2), x1(2)
    REAL(C_DOUBLE), INTENT(OUT) :: err(3)
  END FUNCTION

  ! the state WriteSolutionToDisk prints (src/output.f90:79): the input of
  ! each step's last FVTD, kept from now on; then its primitives V(1:3,nx,ny)
  INTEGER(C_INT) FUNCTION swm_keep_output_state(h, on) BIND(C, NAME='swm_keep_output_state')
    IMPORT :: C_INT, C_PTR
    TYPE(C_PTR), VALUE :: h
    INTEGER(C_INT), VALUE :: on
  END FUNCTION
  INTEGER(C_INT) FUNCTION swm_download_output_primitive(h, V) &
      BIND(C, NAME='swm_download_output_primitive')
    IMPORT :: C_INT, C_DOUBLE, C_PTR
    TYPE(C_PTR), VALUE :: h
    REAL(C_DOUBLE) :: V(*)
  END FUNCTION

  ! ConsToPrim of the state, V(1:3,nx,ny), for the output writer
  INTEGER(C_INT) FUNCTION swm_download_primitive(h, V) BIND(C, NAME='swm_download_primitive')
    IMPORT :: C_INT, C_DOUBLE, C_PTR
    TYPE(C_PTR), VALUE :: h
    REAL(C_DOUBLE) :: V(*)
  END FUNCTION

  ! ReconstructionFix (src/parameters.f90:142): 4 as the reference, 1 first
  ! order or 2 MUSCL on the cells the shock indicator flags; before swm_init_wb
  INTEGER(C_INT) FUNCTION swm_set_reconstruction_fix(h, recfix) &
      BIND(C, NAME='swm_set_reconstruction_fix')
    IMPORT :: C_INT, C_PTR
    TYPE(C_PTR), VALUE :: h
    INTEGER(C_INT), VALUE :: recfix
  END FUNCTION

  ! the time loop of TimeDiscretization to the next output time
  ! (src/timediscretization.f90:57-95), dt clipped to t_analyze and t_end
  INTEGER(C_INT) FUNCTION swm_advance(h, t, t_analyze, t_end, max_steps, steps_done, dts) &
      BIND(C, NAME='swm_advance')
    IMPORT :: C_INT, C_DOUBLE, C_PTR
    TYPE(C_PTR), VALUE :: h
    REAL(C_DOUBLE), INTENT(INOUT) :: t
    REAL(C_DOUBLE), VALUE :: t_analyze, t_end
    INTEGER(C_INT), VALUE :: max_steps
    INTEGER(C_INT), INTENT(OUT) :: steps_done
    TYPE(C_PTR), VALUE :: dts
  END FUNCTION

  FUNCTION swm_last_error_c(h) BIND(C, NAME='swm_last_error')
    IMPORT :: C_PTR
    TYPE(C_PTR), VALUE :: h
    TYPE(C_PTR) :: swm_last_error_c
  END FUNCTION
END INTERFACE

CONTAINS

! The handle's last error message as a Fortran string.
FUNCTION swm_error_message(h) RESULT(msg)
  TYPE(C_PTR), INTENT(IN) :: h
  CHARACTER(LEN=512) :: msg
  CHARACTER(KIND=C_CHAR), POINTER :: c(:)
  TYPE(C_PTR) :: p
  INTEGER :: k
  msg = ''
  p = swm_last_error_c(h)
  IF (.NOT. C_ASSOCIATED(p)) RETURN
  CALL C_F_POINTER(p, c, (/512/))
  DO k = 1, 512
    IF (c(k) == C_NULL_CHAR) EXIT
    msg(k:k) = c(k)
  END DO
END FUNCTION

! The reference's error convention for hard errors: print and STOP
! (src/timediscretization.f90:78-80).
SUBROUTINE swm_check(h, rc, what)
  TYPE(C_PTR), INTENT(IN) :: h
  INTEGER(C_INT), INTENT(IN) :: rc
  CHARACTER(LEN=*), INTENT(IN) :: what
  IF (rc < 0) THEN
    WRITE(*,*) "swm: ", TRIM(what), " failed: ", TRIM(swm_error_message(h))
    STOP 1
  END IF
END SUBROUTINE

END MODULE swm_bridge
!===============================================================================!
