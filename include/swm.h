/*
 * swm.h -- C ABI of the MI355X (gfx950) implementation of sw-mpdec's
 * PATANKAR-build time step: timescheme=5 (mPDeC5) with Reconstruction=4
 * (WENO5 at 4-point Gauss nodes), periodic boundaries, FP64 -- and, on the
 * same kernels, the reference's other time schemes (timescheme 1-4 of the
 * build without PATANKAR, 6 and 7 of the PATANKAR build), its
 * ReconstructionFix 1 and 2, and the steps before and after the time loop:
 * initial data, PostProcessing's L1 error and the writer's primitives.
 *
 * The reference (accdavlo/sw-mpdec, Fortran) has no FFI; its hot path sits
 * behind module procedures that share module-global state.  Each entry point
 * below names the reference code it replaces (fv-solver-sw/src/...).  A
 * Fortran caller binds them with ISO_C_BINDING (REAL = c_double under
 * -fdefault-real-8, INTEGER = c_int).
 *
 * Array layouts are the reference's, column-major, first index fastest:
 *   U          (1:3, 1:nx, 1:ny)                     ghosts = 0
 *              (1:3, -2:nx+3, -2:ny+3), the module array U itself, ghosts = 3
 *   Ut, S      (1:3, 1:nx, 1:ny)
 *   FX         (1:3, 0:nx, 1:ny)     FY (1:3, 1:nx, 0:ny)
 *   Prod/Dest  (5*nx*ny)   CSR slots S, W, self, E, N of each row
 *   dbdx/dbdy  (1:nx, 1:ny, 1:4, 1:4) = Bathymetry_X/_Y at MeshGP(:,i,j,xg,yg)
 *
 * All functions return SWM_OK (0) or a negative SWM_E* code; the message is
 * available from swm_last_error().  Calls on one handle must come from one
 * host thread at a time.  The library owns device memory and one HIP stream
 * per handle (a second one once swm_download_primitive_async is used); no
 * host pointer is kept after a call returns, except the destination of
 * swm_download_primitive_async, until swm_output_wait.
 */
#ifndef SWM_H
#define SWM_H

#ifdef __cplusplus
extern "C" {
#endif

#define SWM_OK 0
#define SWM_E_ARG (-1)     /* invalid argument */
#define SWM_E_HIP (-2)     /* HIP runtime failure (message names the call) */
#define SWM_E_STATE (-3)   /* call out of order (e.g. WB state missing) */
#define SWM_E_NOMEM (-4)   /* device memory exhausted */
#define SWM_E_BUDGET (-5)  /* the job does not fit the engine's memory budget now */

typedef struct swm_solver swm_solver_t;

/* Replaces InitializeFiniteVolume's allocations (src/finitevolume.f90:44-271)
 * and the geometry of BuildMesh (src/mesh.f90:28-146) for one grid.
 * wgp_bnd = WeightsGPBnd(1:4) as computed by BuildMesh.  nx, ny >= 3.
 * wellbalanced = 1 for the -DWELLBALANCED build. */
int swm_create(int nx, int ny, double dx, double dy, double gravity, double cfl,
               int wellbalanced, const double wgp_bnd[4], swm_solver_t **out);

/* Frees the handle and its device memory (FinalizeFiniteVolume,
 * src/finitevolume.f90:768-877). */
void swm_destroy(swm_solver_t *h);

/* Last error message of the handle (or of the failed swm_create when h is
 * NULL; that one is per host thread). */
const char *swm_last_error(const swm_solver_t *h);

/* Device bytes a handle of this size allocates (for admission control). */
long long swm_footprint_bytes(int nx, int ny, int wellbalanced, int bathymetry);

/* Bathymetry gradient tables used by SourceFunc (src/equation.f90:398-413,
 * Bathymetry_X/_Y :451-509), evaluated by the caller at the 4x4 Gauss points.
 * NULL, NULL selects BathymetryFlag = 0 (S = 0, src/equation.f90:320-324). */
int swm_set_bathy_grad(swm_solver_t *h, const double *dbdx, const double *dbdy);

/* InitializeWBVariables (src/finitevolume.f90:339-396): FVTimeDerivative of
 * the well-balanced state Uwb on the device, kept as FXWB/FYWB/SWB. */
int swm_init_wb(swm_solver_t *h, const double *Uwb, int ghosts);

/* The reference's initial data made on the device instead of by the host
 * (src/main.f90:22-26): the state of FillInitialConditions
 * (src/finitevolume.f90:281-333), i.e. the 4x4-Gauss cell averages of
 * ExactFunction(initial_condition, t = 0) (src/equation.f90:70-230,
 * initial_condition 1..6); the Bathymetry_X/_Y tables of bathymetry_flag
 * (src/equation.f90:419-509; 0: flat bottom); and, for a wellbalanced
 * handle, the averages of ExactFunctionWB (initial conditions 2 and 3 only,
 * else SWM_E_ARG) with their FVTD, as swm_init_wb.  One call replaces
 * swm_set_bathy_grad + swm_init_wb + swm_upload with host-evaluated data.
 * The domain is MESH_X0 = x0, MESH_X1 = x1 (src/parameters.f90), which must
 * give the handle's dx and dy exactly (|x1-x0|/n, src/mesh.f90:39-41), else
 * SWM_E_ARG.  Gravity is the handle's.  Parity: sqrt, + and * are IEEE,
 * the powers are expanded as the reference's compiler does, and the dam
 * breaks (4, 5) are bitwise the reference's; exp, sin, cos and atan are the
 * device library's, so the other initial data agree with the reference's to
 * a few ulps. */
int swm_fill_initial_data(swm_solver_t *h, int initial_condition, int bathymetry_flag,
                          const double x0[2], const double x1[2]);

/* Initial data of the benchmark workloads made on the device (no counterpart
 * in the reference, whose own initial conditions are swm_fill_initial_data's):
 * on the periodic unit square [0,1]^2 (the handle's dx, dy must be 1/nx, 1/ny
 * exactly), the 4x4-Gauss cell averages of
 *   SWM_FAMILY_VORTEX      h = 1 - eps^2/(2g) exp(1 - r^2),
 *                          (u, v) = eps exp((1 - r^2)/2) (-y', x'),
 *                          x' = (x - 1/2)/0.1, y' = (y - 1/2)/0.1, r^2 = x'^2 + y'^2;
 *   SWM_FAMILY_LAKE_BUMPS  lake at rest h + b = 1, u = v = 0, over
 *                          b = sum_k A_k G(x; cx_k, s_k) G(y; cy_k, s_k), plus
 *                          amp G(x; px, ps) G(y; py, ps) on h, where
 *                          G(x; c, s) = sum over a = -1, 0, 1 of exp(-(x-c-a)^2/s^2)
 *                          (periodic images), with the bathymetry tables of b's
 *                          exact gradient and, for a well-balanced handle, the
 *                          well-balanced state h = 1 - b and its FVTD;
 *   SWM_FAMILY_DAM_BREAK   h = h_in where (x-1/2)^2 + (y-1/2)^2 < radius^2, else
 *                          h_out; u = v = 0.
 * exp is a fixed method made of IEEE operations only (error < 0.85 ulp;
 * swm_debug_exp), so these data are bitwise the same as the host problem
 * builder's on any host.  One call replaces swm_set_bathy_grad + swm_init_wb +
 * swm_upload with host-built arrays of the same family. */
#define SWM_FAMILY_VORTEX 1
#define SWM_FAMILY_LAKE_BUMPS 2
#define SWM_FAMILY_DAM_BREAK 3
#define SWM_FAMILY_MAX_BUMPS 8
typedef struct {
  int kind;                  /* SWM_FAMILY_* */
  int nbumps;                /* lake: bumps used, 0..SWM_FAMILY_MAX_BUMPS */
  double eps;                /* vortex */
  double A[SWM_FAMILY_MAX_BUMPS], cx[SWM_FAMILY_MAX_BUMPS], cy[SWM_FAMILY_MAX_BUMPS],
      s[SWM_FAMILY_MAX_BUMPS];  /* lake: bumps */
  double amp, px, py, ps;    /* lake: surface perturbation */
  double radius, h_in, h_out;  /* dam break */
} swm_family;
int swm_fill_family(swm_solver_t *h, const swm_family *family);

/* PostProcessing (src/postprocessing.f90:24-64) on the device: the L1 error
 * of the current state against the 4x4-Gauss cell averages of
 * ExactFunction(initial_condition, t), err[v] = (sum over cells of
 * |U(v) - Uexact(v)|) * dx * dy, summed in the reference's cell order, on the
 * domain of swm_fill_initial_data.  Bitwise the reference's for the dam
 * breaks (4, 5); for the others ExactFunction's exp/sin/cos/atan are the
 * device library's (a few ulps). */
int swm_l1_error(swm_solver_t *h, int initial_condition, int bathymetry_flag, double t,
                 const double x0[2], const double x1[2], double err[3]);

/* The primitive variables (h, u, v) = ConsToPrim(U) (src/equation.f90:
 * 912-945) of the current state in the (1:3, nx, ny) layout, for output.
 * The reference's writer (WriteSolutionToDisk, src/output.f90:29-190) prints
 * the module array V, i.e. the primitives of the last FVTD input (a DeC node
 * of the step just taken), not of the state: this gives those of the state. */
int swm_download_primitive(swm_solver_t *h, double *V);

/* The state the reference's output writer prints.  WriteSolutionToDisk
 * (src/output.f90:79) writes the module array V, which the last FVTD of the
 * step filled from its input: for mPDeC5 the fourth DeC node after the fourth
 * correction (src/timediscretization.f90:548-556), for mPDeC2 the second node
 * after the first, for mPEuler U^n, for the explicit schemes the last stage
 * state -- not the new state.  With swm_keep_output_state(h, 1) every
 * following step keeps a copy of that input (3 n doubles of device memory);
 * swm_download_output_primitive then gives its primitives in the (1:3, nx, ny)
 * layout, bitwise the reference's V.  SWM_E_STATE when no step has run since
 * the state was kept. */
int swm_keep_output_state(swm_solver_t *h, int on);
int swm_download_output_primitive(swm_solver_t *h, double *V);

/* The same, asynchronous: takes a device snapshot of the primitives in
 * stream order and copies it to V on a second stream, returning at once, so
 * that the writer's copy overlaps the steps that follow (steps called after
 * this do not change what is copied).  V must stay valid until
 * swm_output_wait; pinned memory (hipHostMalloc) makes the copy truly
 * asynchronous.  A second call first lets the previous copy finish on the
 * device. */
int swm_download_primitive_async(swm_solver_t *h, double *V);
int swm_output_wait(swm_solver_t *h);

/* Copy the state U into / out of the device (the module array U). */
int swm_upload(swm_solver_t *h, const double *U, int ghosts);
int swm_download(swm_solver_t *h, double *U);

/* TimeStep (src/equation.f90:791-837): CFL time step of the device state,
 * and LambdaMaxX/Y. */
int swm_timestep(swm_solver_t *h, double *dt_min, double *lambda_max_x, double *lambda_max_y);

/* swm_timestep without any device allocation.  No counterpart in the
 * reference: it is for callers that cannot allocate device memory (code under
 * HIP graph capture, serving loops that forbid hipMalloc).  U is a DEVICE
 * array U(1:3, 1-g:nx+g, 1-g:ny+g) with g = ghosts (0 or 3, as swm_upload);
 * out is a DEVICE array that receives {dt_min, LambdaMaxX, LambdaMaxY},
 * bitwise swm_timestep's values for the same state.  The work is enqueued on
 * `stream` (a hipStream_t; NULL = the null stream) and the call returns
 * without waiting.  The per-block partial maxima live in a fixed device array
 * of the library (1.5 MiB per device, held until the process ends), so grids
 * above SWM_TIMESTEP_NOALLOC_MAX_CELLS cells are rejected with SWM_E_ARG.
 * All calls on one device share that array: calls that may overlap in time
 * on one device (several host threads or unordered streams) race. */
#define SWM_TIMESTEP_NOALLOC_MAX_CELLS (1 << 24)
int swm_timestep_noalloc(int nx, int ny, double dx, double dy, double gravity, double cfl,
                         const double *U, int ghosts, double *out, void *stream);

/* TimeDiscretizationByMPDeC5 (src/timediscretization.f90:458-566) with its
 * 13 FVTimeDerivative evaluations, 13 MODIFIED_PATANKAR_MATRIX_INVERSION
 * assemblies (:771-924) and Jacobi solves (src/jacobi_iteration.f90:52-87).
 * t is accepted for the signature's sake (periodic BCs and static bathymetry
 * make the step independent of it).  jacobi_iters (may be NULL) receives the
 * 13 sweep counts in execution order.  Returns the number (>= 0) of solves
 * that stopped at the iteration limit with the residual above tolerance --
 * the reference prints "Jacobi not converged" and continues -- or an error. */
int swm_step_mpdec5(swm_solver_t *h, double t, double dt, int jacobi_iters[13]);

/* The two other modified-Patankar schemes of the PATANKAR build, on the same
 * kernels with MSteps = 2 and thetaDeC = reshape((/0, 1/2, 0, 1/2/), (/2,2/)):
 * TimeDiscretizationByMPEuler (src/timediscretization.f90:671-764; KCorr = 1:
 * one FVTimeDerivative, one solve) and TimeDiscretizationByMPDeC2 (:571-664;
 * KCorr = 2: two of each).  Return values as swm_step_mpdec5; jacobi_iters
 * receives 1 and 2 counts.  The depths they use follow swm_set_hp_association. */
int swm_step_mpeuler(swm_solver_t *h, double t, double dt, int jacobi_iters[1]);
int swm_step_mpdec2(swm_solver_t *h, double t, double dt, int jacobi_iters[2]);

/* The explicit schemes, as the reference built without -DPATANKAR runs them:
 * there UpdateTimeDerivative forms Ut(1:3) from the face fluxes
 * (src/finitevolume.f90:529-533) instead of production/destruction terms, and
 * the step combines FVTimeDerivative stages with the source's weights, in its
 * evaluation order:
 *   TimeDiscretizationByForwardEuler (src/timediscretization.f90:198-232),
 *   TimeDiscretizationBySSPRK4 (:104-192, 5 stages),
 *   TimeDiscretizationByRK65 (:238-356, 6 stages),
 *   TimeDiscretizationByDeC5 (:362-451, 13 FVTimeDerivative evaluations).
 * No Jacobi solves; returns SWM_OK or an error.  (The PATANKAR build also
 * accepts these timescheme numbers, but it never writes Ut(1), so there the
 * depth stays frozen; that degenerate variant is not offered.) */
int swm_step_forward_euler(swm_solver_t *h, double t, double dt);
int swm_step_ssprk4(swm_solver_t *h, double t, double dt);
int swm_step_rk65(swm_solver_t *h, double t, double dt);
int swm_step_dec5(swm_solver_t *h, double t, double dt);

/* The scheme swm_run_steps and swm_advance step with, by the reference's
 * timescheme number (src/parameters.f90:144; selected at
 * src/timediscretization.f90:61-81): 1 ForwardEuler, 2 SSPRK4, 3 RK65,
 * 4 DeC5 (the explicit schemes, as above), 5 mPDeC5 (the default),
 * 6 mPEuler, 7 mPDeC2.  Other numbers: SWM_E_ARG. */
int swm_set_timescheme(swm_solver_t *h, int timescheme);

/* Which depths the two-node schemes put into the Patankar weights and the
 * right-hand side (no effect on mPDeC5).  MODIFIED_PATANKAR_MATRIX_INVERSION
 * declares hp(MSteps, nx, ny) (src/timediscretization.f90:783) but is passed
 * Up(:,1,:,:), whose first extent is 4 (src/finitevolume.f90:161); with
 * MSteps = 2 the dummy is sequence associated with that section, so
 * hp(m, cell) reads Up at node ((m-1) + 2*cell) mod 4 of cell
 * ((m-1) + 2*cell) / 4 (cells numbered from 0, x fastest), and nodes 3..4 are
 * never written and hold 0.  The reference's mPEuler and mPDeC2 therefore
 * drop the mass of every odd cell from the right-hand side (about half the
 * mass in the first step) and reach NaN states within a few steps.
 *   SWM_HP_AS_REFERENCE (default): exactly that, bitwise as the reference;
 *   SWM_HP_BY_NODE: hp(m, cell) = Up(m, cell), the same rule as mPDeC5's.
 * The by-node variant has no reference output to match; it conserves mass
 * and keeps depths non-negative. */
#define SWM_HP_AS_REFERENCE 0
#define SWM_HP_BY_NODE 1
int swm_set_hp_association(swm_solver_t *h, int mode);

/* ReconstructionFix (src/finitevolume_vars.f90:94), which the reference sets
 * equal to Reconstruction = 4 (src/parameters.f90:142).  When it differs,
 * every FVTD runs the Jameson shock indicator on the depths
 * (src/shocksindicator.f90:29-129) and the flagged cells take
 * ReconstructionFix's reconstruction instead of WENO5
 * (src/reconstruction.f90:190-342; FVTimeDerivative, src/finitevolume.f90:425-433):
 *   4 (default): WENO5 everywhere; the indicator cannot change a result;
 *   1: first order, the cell's primitives at every Gauss point;
 *   2: MUSCL with the minmod limiter.
 * The indicator writes each cell's verdict over Ind(:, ii-1:ii+1, jj) in loop
 * order, so the flag of cell ii is the verdict of cell min(ii+1, n): the
 * library reproduces that.  3 (WENO3) is rejected with SWM_E_ARG:
 * WENO3_SecondSweep defines W(1:2) only, so with 4 Gauss points the
 * reference reads undefined values.  The well-balanced reference state is an
 * FVTD under this setting: changing it on a handle discards that state, and
 * swm_init_wb must be called again before stepping. */
int swm_set_reconstruction_fix(swm_solver_t *h, int recfix);

/* Several steps without returning to the caller: before each step dt is the
 * TimeStep value (a step-count stop instead of the tEnd stop of
 * src/timediscretization.f90:57-95).  dts (may be NULL) receives nsteps dt
 * values; jacobi_iters (may be NULL) nsteps rows of the scheme's solve count
 * (13, 1 or 2; none for the explicit schemes).  Stops early with SWM_E_STATE
 * if dt is not finite. */
int swm_run_steps(swm_solver_t *h, int nsteps, double *dts, int *jacobi_iters);

/* The reference's time loop on the device (src/timediscretization.f90:57-95)
 * up to the next output time: each step's dt is
 * MIN(MIN(dt_min, t_analyze - t), MIN(dt_min, t_end - t)) with dt_min the
 * TimeStep value, then t = t + dt; returns when |t_analyze - t| < 1e-10 or
 * |t - t_end| < 1e-10 or after max_steps steps.  *t is updated in place (the
 * same double arithmetic as the reference); steps_done and dts (max_steps
 * values) may be NULL.  Returns the number of non-converged Jacobi solves, or
 * SWM_E_STATE if a dt is not finite (the state is then that before the step). */
int swm_advance(swm_solver_t *h, double *t, double t_analyze, double t_end, int max_steps,
                int *steps_done, double *dts);

/* One FVTimeDerivative of the device state (src/finitevolume.f90:402-440)
 * with its module outputs in the reference layouts, for parity tests.  Any
 * output may be NULL.  Ut(1,:,:) is 0 (unused with PATANKAR). */
int swm_fvtd(swm_solver_t *h, double *Ut, double *FX, double *FY, double *S, double *Prod,
             double *Dest);

/* Test overrides of the Jacobi stop rule (maxIter = 1000, tolerance = 1e-16,
 * src/jacobi_iteration.f90:63-64). */
int swm_set_jacobi_limits(swm_solver_t *h, int max_iter, double tol);

/* Stage positivity check (no counterpart in the reference): while on, the
 * device folds the minimum depth of every DeC stage solution (the result of
 * every Patankar solve: 13 per mPDeC5 step; the explicit schemes solve
 * nothing and record nothing) into a per-handle record; NaN if any depth was NaN.
 * Enabling resets the record.  swm_stage_min_h returns SWM_E_STATE when no
 * stage has run since. */
int swm_track_stage_min(swm_solver_t *h, int on);
int swm_stage_min_h(swm_solver_t *h, double *hmin);

/* swm_engine_set_execution / swm_engine_execution_stats /
 * swm_engine_set_tile_shape / swm_engine_tile_shape for a plain handle. */
int swm_set_execution(swm_solver_t *h, int use_graph, int margin);
int swm_execution_stats(const swm_solver_t *h, long long stats[6]);
int swm_status_reads(const swm_solver_t *h, long long *n);
int swm_set_tile_shape(swm_solver_t *h, int shape);
int swm_tile_shape(const swm_solver_t *h, int dims[4]);

/* Number of Jacobi stop decisions so far that fell inside the rounding window
 * of the parallel residual sum and were settled by the exact sequential sum. */
long long swm_ambiguous_decisions(const swm_solver_t *h);

/* Measurement (no counterpart in the reference, which has no timers): HIP
 * events around every kernel launch of the handle (swm_profile_enable), or
 * around the launches of the categories whose bit is set in mask
 * (swm_profile_categories; bit c for category c, 0 turns recording off),
 * summed per category.  While any category is recorded, steps run their
 * launches one by one (no HIP-graph replay).  Categories --
 * 0 fvtd_x, 1 fvtd_y, 2 fvtd_update, 3 momentum (the explicit schemes'
 * stage combinations; the modified-Patankar schemes' momentum update runs
 * inside the assembly kernel), 4 patankar_assemble (with that momentum
 * update, and the two-node schemes' depth view), 5 jacobi_sweep, 6 timestep,
 * 7 other.  swm_profile_read waits for the
 * stream, returns milliseconds and launch counts, and clears the record. */
#define SWM_PROF_NCAT 8
int swm_profile_enable(swm_solver_t *h, int on);
int swm_profile_categories(swm_solver_t *h, unsigned mask);
int swm_profile_read(swm_solver_t *h, double ms[SWM_PROF_NCAT], long long launches[SWM_PROF_NCAT]);
const char *swm_profile_name(int category);

/* Wait for the handle's stream. */
int swm_sync(swm_solver_t *h);

/* ------------------------------------------------------------------------
 * Batched and served solves (no counterpart in the reference, which runs one
 * grid per process): an engine holds any number of independent grids on one
 * GPU and advances every running grid by one step of its scheme per
 * swm_engine_step, each with its own CFL time step, in shared kernel
 * launches (grids of one scheme share launches).  Grids can be submitted and released between steps (continuous
 * batching); admission is checked against a device-memory budget.  Each
 * grid's results are bitwise those of the plain solver above.
 * ------------------------------------------------------------------------ */
typedef struct swm_engine swm_engine_t;

typedef struct {
  int nx, ny;
  double dx, dy, gravity, cfl;
  int wellbalanced;          /* 1: Uwb required */
  double wgp_bnd[4];         /* WeightsGPBnd(1:4) */
  const double *U0;          /* (1:3, 1:nx, 1:ny) */
  const double *Uwb;         /* (1:3, 1:nx, 1:ny) or NULL */
  const double *dbdx;        /* (1:nx, 1:ny, 1:4, 1:4) or NULL (flat bottom) */
  const double *dbdy;
  int nsteps;                /* CFL steps to run; 0: none (DONE at once); < 0: until released */
  int timescheme;            /* 0 or 5: mPDeC5; 1..4, 6, 7 as swm_set_timescheme */
  int hp_association;        /* SWM_HP_AS_REFERENCE or SWM_HP_BY_NODE (swm_set_hp_association) */
  int reconstruction_fix;    /* 0 or 4: WENO5 everywhere; 1, 2 as swm_set_reconstruction_fix */
  /* With U0 == NULL and initial_condition > 0 the job's initial data are made
   * on the device as swm_fill_initial_data does (Uwb, dbdx, dbdy unused). */
  int initial_condition;     /* 1..6 */
  int bathymetry_flag;       /* 0..2 */
  double mesh_x0[2], mesh_x1[2];
  /* With U0 == NULL and family != NULL the job's initial data are made on the
   * device as swm_fill_family does (Uwb, dbdx, dbdy, initial_condition unused). */
  const swm_family *family;
} swm_job_desc;

#define SWM_JOB_RUNNING 0
#define SWM_JOB_DONE 1
#define SWM_JOB_FAILED 2     /* TimeStep returned a non-finite dt */

typedef struct {
  int state, steps_done, steps_left;
  double t;                  /* simulated time reached */
  long long jacobi_sweeps;   /* total over all solves */
  int nonconverged;          /* solves stopped at the iteration limit */
  long long bytes;           /* device memory held */
} swm_job_info;

/* budget_bytes <= 0: no limit beyond the device's memory. */
int swm_engine_create(long long budget_bytes, swm_engine_t **out);
void swm_engine_destroy(swm_engine_t *e);
const char *swm_engine_last_error(const swm_engine_t *e);
long long swm_job_footprint_bytes(int nx, int ny, int wellbalanced, int bathymetry);
long long swm_engine_bytes_in_use(const swm_engine_t *e);
/* Copies the inputs to the device (and builds the well-balanced state);
 * returns SWM_E_BUDGET without side effects when the job does not fit. */
int swm_engine_submit(swm_engine_t *e, const swm_job_desc *job, int *job_id);
/* One step of every running job; *n_running (may be NULL) receives the
 * number still running afterwards. */
int swm_engine_step(swm_engine_t *e, int *n_running);
/* Up to max_steps steps of every running job, with one host read of the
 * device status per batch of up to 16 steps instead of one per step: the
 * device time loop commits, records and stops each job on its own (its step
 * count, a non-finite dt).  Results are those of max_steps calls of
 * swm_engine_step.  *steps_run (may be NULL) receives the steps run. */
int swm_engine_run(swm_engine_t *e, int max_steps, int *n_running, int *steps_run);
int swm_engine_job_info(swm_engine_t *e, int job_id, swm_job_info *info);
/* Current state of a job, (1:3, 1:nx, 1:ny). */
int swm_engine_fetch(swm_engine_t *e, int job_id, double *U);
/* swm_l1_error of a job's current state: a served request's error against
 * its initial condition's exact solution without fetching the state. */
int swm_engine_l1_error(swm_engine_t *e, int job_id, int initial_condition, int bathymetry_flag,
                        double t, const double x0[2], const double x1[2], double err[3]);
/* swm_keep_output_state / swm_download_output_primitive for a job: the state
 * the reference's writer prints, per served request. */
int swm_engine_keep_output_state(swm_engine_t *e, int job_id, int on);
int swm_engine_fetch_output_primitive(swm_engine_t *e, int job_id, double *V);
/* Frees the job's device memory; its id may be reused. */
int swm_engine_release(swm_engine_t *e, int job_id);
int swm_engine_set_jacobi_limits(swm_engine_t *e, int max_iter, double tol);
/* Step execution (no counterpart in the reference; results are bitwise the
 * same in every mode).  Once every job has run a step, a step launches each
 * correction's predicted Jacobi sweeps (last step's counts + margin) without
 * reading back, replayed from a captured HIP graph when use_graph is 1, and
 * checks once before committing; a step whose solves needed more sweeps is
 * redone with read-backs.  -1 lets the built-in autotuner time the
 * candidates (face-kernel tile shape, then graph 0/1 x margin 1/2) on the
 * first speculative steps and keep the fastest; margin = -2 turns
 * speculation off (every step reads back).
 * stats: speculative steps, redone steps, graph launches, graph captures,
 * graph and margin in use. */
int swm_engine_set_execution(swm_engine_t *e, int use_graph, int margin);
int swm_engine_execution_stats(const swm_engine_t *e, long long stats[6]);
/* Times the host has waited for the device status so far (no counterpart in
 * the reference): one per batch of steps of the device time loop, plus one
 * per Jacobi check while a step runs in polling mode. */
int swm_engine_status_reads(const swm_engine_t *e, long long *n);
/* Tile shape of the face kernels (no counterpart in the reference; results
 * are bitwise the same with every shape): 0 = 16x16 threads for both face
 * directions, 1 = 32x16 for x faces and 16x32 for y faces (fewer redundant
 * halo sweeps, more idle threads on small grids), -1 = chosen by the
 * autotuner (the default).  swm_engine_tile_shape returns the shape in use
 * (the tuned one once tuning has finished, 0 before) and, if dims is not
 * NULL, its x-kernel and y-kernel thread extents. */
int swm_engine_set_tile_shape(swm_engine_t *e, int shape);
int swm_engine_tile_shape(const swm_engine_t *e, int dims[4]);

/* The stage positivity check of swm_track_stage_min, per job of an engine. */
int swm_engine_track_stage_min(swm_engine_t *e, int on);
int swm_engine_stage_min_h(swm_engine_t *e, int job_id, double *hmin);
int swm_engine_profile_enable(swm_engine_t *e, int on);
int swm_engine_profile_categories(swm_engine_t *e, unsigned mask);
int swm_engine_profile_read(swm_engine_t *e, double ms[SWM_PROF_NCAT],
                            long long launches[SWM_PROF_NCAT]);

/* Self-check of the device's FP64 '/' and sqrt: q = a/b, r = sqrt(a). */
int swm_debug_divsqrt(int n, const double *a, const double *b, double *q, double *r);

/* Self-check of the division the WENO weights use when all operands lie in
 * [2^-400, 2^400]: q = a/b through one refined reciprocal of b shared by
 * every quotient with that denominator (no range check here). */
int swm_debug_shared_div(int n, const double *a, const double *b, double *q);

/* Self-check of the exponential of swm_fill_family: y = exp(x) by its fixed
 * method, for comparison with the host problem builder's. */
int swm_debug_exp(int n, const double *x, double *y);

/* Measurement (no counterpart in the reference): what a step boundary costs,
 * in microseconds per round of a chain in which every one of nblk blocks reads
 * its neighbour's value of the previous round -- [0] one kernel launch per
 * round, [1] the same launches replayed from a HIP graph, [2] one cooperative
 * launch with a grid-wide barrier per round.  SWM_E_STATE if a mode does not
 * carry the chain to every block; SWM_E_ARG if nblk blocks cannot all be
 * resident at once or rounds is outside 1..100000. */
int swm_debug_sync_cost(int nblk, int rounds, double us_per_round[3]);

/* Self-check of the WENO5 arithmetic as the kernels evaluate it
 * (WENO5_FirstSweep, src/reconstruction.f90:673-755, and WENO5_SecondSweep,
 * :999-1136) on n five-point stencils q[5i..5i+4]: first[2i..2i+1] = (wm, wp),
 * second[4i..4i+3] = the 4 Gauss-point values. */
int swm_debug_weno5(int n, const double *q, double *first, double *second);

#ifdef __cplusplus
}
#endif
#endif /* SWM_H */
