!===============================================================================!
! swm_fdriver: a Fortran caller of libswm through swm_bridge, shaped like the
! reference's TimeDiscretization loop (src/timediscretization.f90:57-95) with
! timescheme = 1..7.
!   loop = 0: nsteps times dt = TimeStep(), one step of the scheme selected as
!             the reference's SELECT CASE(timescheme) does (:61-81).
!   loop = 1: the reference's output loop with tEnd = (nsteps - 1/2) * dt_min(0)
!             and 2 output intervals, each interval one swm_advance call (dt
!             clipped to tAnalyze and tEnd on the device).
!
! Reads a raw input file
!   int32  nx, ny, nsteps, wellbalanced, has_bathy, loop, timescheme (0: 5),
!          hp_association (swm_set_hp_association),
!          reconstruction_fix (0: 4, the reference's; swm_set_reconstruction_fix),
!          initial_condition (0: the data below; 1..6: made on the device by
!          swm_fill_initial_data, with no data below), bathymetry_flag
!   real64 dx, dy, gravity, cfl, wgp_bnd(4), mesh_x0(2), mesh_x1(2)
!   real64 U(3,nx,ny) [, Uwb(3,nx,ny) if wellbalanced] [, dbdx, dbdy(nx,ny,4,4)]
!          (initial_condition 0 only)
! and writes the output file
!   int32  nsteps_taken, nonconverged_total
!   real64 dt(nsteps_taken), U(3,nx,ny)
!   real64 L1error(3) at the final time [initial_condition > 0: PostProcessing,
!          src/postprocessing.f90:24-64, by swm_l1_error], then V(3,nx,ny), the
!          primitives WriteSolutionToDisk prints (swm_download_output_primitive)
! Usage: swm_fdriver <input> <output>
!===============================================================================!
PROGRAM swm_fdriver
USE ISO_C_BINDING
USE swm_bridge
IMPLICIT NONE
INTEGER(C_INT) :: hdr(11), nx, ny, nsteps, wb, hasb, rc, k, nonconv, loop, ndone, nout, nrec
INTEGER(C_INT) :: timescheme, hpa, recfix, ic, bflag
INTEGER(C_INT) :: iters(13)
REAL(C_DOUBLE) :: par(12), dx, dy, g, cfl, wgp(4), t, dt, lmx, lmy, x0(2), x1(2), l1(3)
REAL(C_DOUBLE) :: tend, tanalyze, dtanalyze
REAL(C_DOUBLE), ALLOCATABLE, TARGET :: chunk(:)
REAL(C_DOUBLE), ALLOCATABLE :: U(:,:,:), Uwb(:,:,:), bx(:,:,:,:), by(:,:,:,:), dts(:), V(:,:,:)
TYPE(C_PTR) :: h
CHARACTER(LEN=1024) :: fin, fout
INTEGER :: uin, uout

CALL GET_COMMAND_ARGUMENT(1, fin)
CALL GET_COMMAND_ARGUMENT(2, fout)
OPEN(NEWUNIT=uin, FILE=TRIM(fin), ACCESS='STREAM', FORM='UNFORMATTED', STATUS='OLD')
READ(uin) hdr
nx = hdr(1); ny = hdr(2); nsteps = hdr(3); wb = hdr(4); hasb = hdr(5); loop = hdr(6)
timescheme = hdr(7); hpa = hdr(8); recfix = hdr(9); ic = hdr(10); bflag = hdr(11)
IF (timescheme .EQ. 0) timescheme = 5
IF (recfix .EQ. 0) recfix = 4
READ(uin) par
dx = par(1); dy = par(2); g = par(3); cfl = par(4); wgp = par(5:8)
x0 = par(9:10); x1 = par(11:12)
ALLOCATE(U(3,nx,ny), dts(MAX(nsteps,1)))
IF (ic .EQ. 0) THEN
  READ(uin) U
  IF (wb .NE. 0) THEN
    ALLOCATE(Uwb(3,nx,ny))
    READ(uin) Uwb
  END IF
  IF (hasb .NE. 0) THEN
    ALLOCATE(bx(nx,ny,4,4), by(nx,ny,4,4))
    READ(uin) bx
    READ(uin) by
  END IF
END IF
CLOSE(uin)

rc = swm_create(nx, ny, dx, dy, g, cfl, wb, wgp, h)
IF (rc .NE. SWM_OK) THEN
  WRITE(*,*) "swm_create failed: ", rc
  STOP 1
END IF
CALL swm_check(h, swm_set_reconstruction_fix(h, recfix), "swm_set_reconstruction_fix")
IF (ic .GT. 0) THEN
  ! FillInitialConditions / InitializeWBVariables / the bathymetry of
  ! src/main.f90:22-26, on the device
  CALL swm_check(h, swm_fill_initial_data(h, ic, bflag, x0, x1), "swm_fill_initial_data")
  CALL swm_check(h, swm_keep_output_state(h, 1_C_INT), "swm_keep_output_state")
ELSE
  IF (hasb .NE. 0) THEN
    CALL swm_check(h, swm_set_bathy_grad(h, bx, by), "swm_set_bathy_grad")
  END IF
  IF (wb .NE. 0) THEN
    CALL swm_check(h, swm_init_wb(h, Uwb, 0_C_INT), "swm_init_wb")
  END IF
  CALL swm_check(h, swm_upload(h, U, 0_C_INT), "swm_upload")
END IF
CALL swm_check(h, swm_set_timescheme(h, timescheme), "swm_set_timescheme")
CALL swm_check(h, swm_set_hp_association(h, hpa), "swm_set_hp_association")

t = 0.0
nonconv = 0
nrec = 0
IF (loop .EQ. 0) THEN
  DO k = 1, nsteps
    CALL swm_check(h, swm_timestep(h, dt, lmx, lmy), "swm_timestep")
    IF (.NOT. (dt .LT. HUGE(dt))) THEN
      WRITE(*,*) "swm_fdriver: non-finite time step at step ", k
      STOP 2
    END IF
    SELECT CASE(timescheme)
      CASE(1)
        rc = swm_step_forward_euler(h, t, dt)
      CASE(2)
        rc = swm_step_ssprk4(h, t, dt)
      CASE(3)
        rc = swm_step_rk65(h, t, dt)
      CASE(4)
        rc = swm_step_dec5(h, t, dt)
      CASE(5)
        rc = swm_step_mpdec5(h, t, dt, iters)
      CASE(6)
        rc = swm_step_mpeuler(h, t, dt, iters(1:1))
      CASE(7)
        rc = swm_step_mpdec2(h, t, dt, iters(1:2))
      CASE DEFAULT
        WRITE(*,*) "Time discretization not implemented"
        STOP 1
    END SELECT
    CALL swm_check(h, rc, "swm_step")
    nonconv = nonconv + rc
    dts(k) = dt
    t = t + dt
  END DO
  nrec = nsteps
ELSE
  ! src/timediscretization.f90:46-95 with nOutputFiles = 2
  CALL swm_check(h, swm_timestep(h, dt, lmx, lmy), "swm_timestep")
  tend = (REAL(nsteps) - 0.5) * dt
  nout = 2
  dtanalyze = tend / REAL(nout)
  tanalyze = t + dtanalyze
  DEALLOCATE(dts)
  ALLOCATE(dts(4 * nsteps + 8), chunk(4 * nsteps + 8))
  DO
    rc = swm_advance(h, t, tanalyze, tend, 4 * nsteps + 8, ndone, C_LOC(chunk))
    CALL swm_check(h, rc, "swm_advance")
    nonconv = nonconv + rc
    dts(nrec + 1:nrec + ndone) = chunk(1:ndone)
    nrec = nrec + ndone
    IF (ABS(tanalyze - t) .LT. 1.0E-10) THEN
      ! WriteSolutionToDisk(t) would download here
      tanalyze = tanalyze + dtanalyze
      IF (tanalyze .GT. tend) tanalyze = tend
    END IF
    IF (ABS(t - tend) .LT. 1.0E-10) EXIT
  END DO
END IF
CALL swm_check(h, swm_download(h, U), "swm_download")
IF (ic .GT. 0) THEN
  ! PostProcessing at the final time (src/main.f90:28)
  CALL swm_check(h, swm_l1_error(h, ic, bflag, t, x0, x1, l1), "swm_l1_error")
  ! the writer's V after the last step (src/output.f90:79)
  ALLOCATE(V(3,nx,ny))
  CALL swm_check(h, swm_download_output_primitive(h, V), "swm_download_output_primitive")
END IF
CALL swm_destroy(h)

OPEN(NEWUNIT=uout, FILE=TRIM(fout), ACCESS='STREAM', FORM='UNFORMATTED', STATUS='REPLACE')
WRITE(uout) nrec, nonconv
WRITE(uout) dts(1:nrec)
WRITE(uout) U
IF (ic .GT. 0) WRITE(uout) l1, V
CLOSE(uout)
END PROGRAM swm_fdriver
!===============================================================================!
