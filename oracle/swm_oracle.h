/*
 * swm_oracle.h -- CPU restatement of the sw-mpdec WENO5 time steps (mPDeC5 and
 * the other schemes of the reference's two builds), used ONLY as a test oracle
 * (by tests/, __graft_entry__.smoke() and the cpu_baseline leg of bench.py).
 * The product path never links it.
 *
 * Arrays use the reference's Fortran layouts (column-major, first index
 * fastest) so that dumps from the reference build (oracle/_ref) compare
 * byte-for-byte:
 *   U, Ut, S        (3, nx, ny)            interior only on the API
 *   FX              (3, 0:nx, 1:ny)
 *   FY              (3, 1:nx, 0:ny)
 *   Prod/Dest       (5*nx*ny)              slots S,W,self,E,N per cell
 *   bathymetry      dbdx/dbdy planes [(yg*4+xg)][j][i] (i fastest), the
 *                   values of Bathymetry_X/_Y at MeshGP(:,i,j,xg,yg)
 */
#ifndef SWM_ORACLE_H
#define SWM_ORACLE_H

#ifdef __cplusplus
extern "C" {
#endif

typedef struct orc_state orc_t;

/* nx == ny is not required.  wbnd: WeightsGPBnd(1:4) (src/mesh.f90:127).
 * wellbalanced: the -DWELLBALANCED variant.  Returns NULL on bad input. */
orc_t *orc_create(int nx, int ny, double dx, double dy, double gravity,
                  double cfl, int wellbalanced, const double wbnd[4]);
void orc_destroy(orc_t *o);

/* BathymetryFlag > 0 switches the source term on (src/equation.f90:324);
 * the tables hold d b/dx, d b/dy at the 4x4 Gauss points of every cell. */
void orc_set_bathy(orc_t *o, const double *dbdx, const double *dbdy);

/* Well-balanced reference state: FVTD(Uwb) stored as FXWB/FYWB/SWB
 * (src/finitevolume.f90:373-392). */
void orc_init_wb(orc_t *o, const double *Uwb);

void orc_set_U(orc_t *o, const double *U);
void orc_get_U(const orc_t *o, double *U);
/* The module array V (3, nx, ny) as the last FVTD left it: the primitives
 * that the reference's output writer prints (src/output.f90:79). */
void orc_get_V(const orc_t *o, double *V);

/* One FVTimeDerivative on the current U (src/finitevolume.f90:402-440).
 * Any output pointer may be NULL. */
void orc_fvtd(orc_t *o, double *Ut, double *FX, double *FY, double *S,
              double *Prod, double *Dest);

/* TimeStep (src/equation.f90:791-837): returns dt, stores LambdaMaxX/Y. */
double orc_timestep(orc_t *o, double *lmx, double *lmy);

/* One mPDeC5 step with the given dt (src/timediscretization.f90:458-566).
 * jacobi_iters[13] (may be NULL) receives the sweep count of each solve in
 * execution order.  Returns the number of solves that ended with the residual
 * above the tolerance ("Jacobi not converged", src/jacobi_iteration.f90:82-84). */
int orc_step_mpdec5(orc_t *o, double dt, int *jacobi_iters);

/* One step of a modified-Patankar scheme, by the reference's timescheme
 * number: 5 mPDeC5, 6 mPEuler (src/timediscretization.f90:671-764), 7 mPDeC2
 * (:571-664).  jacobi_iters receives 13, 1 and 2 counts respectively.
 * Returns the non-converged count, or -1 for another number. */
int orc_step_patankar(orc_t *o, int timescheme, double dt, int *jacobi_iters);

/* One step of an explicit scheme of the build without PATANKAR
 * (src/timediscretization.f90): 1 ForwardEuler (:198-232), 2 SSPRK4
 * (:104-192), 3 RK65 (:238-356), 4 DeC5 (:362-451).  Every FVTD of the oracle
 * also forms Ut(1) as that build does (src/finitevolume.f90:529-533).
 * Returns 0, or -1 for another number. */
int orc_step_explicit(orc_t *o, int timescheme, double dt);

/* on = 1: the two-node schemes read hp(m,i,j) = Up(m,1,i,j) (the evident
 * intent) instead of the reference's sequence association of hp with
 * Up(:,1,:,:) (src/timediscretization.f90:783 vs src/finitevolume.f90:161).
 * No reference output exists for on = 1; it is pinned by invariants only. */
void orc_set_hp_by_node(orc_t *o, int on);

/* ReconstructionFix (src/finitevolume_vars.f90:94): 4 (the reference's
 * setting, = Reconstruction, src/parameters.f90:142), or 1 (first order) or
 * 2 (MUSCL) for the cells the Jameson indicator flags.  Returns 0, or -1 for
 * another value. */
int orc_set_reconstruction_fix(orc_t *o, int recfix);

/* The reference's initial data for InitialCondition ic (1..6) with
 * BathymetryFlag bathy_flag (0..2) on [x0[0],x1[0]] x [x0[1],x1[1]], nx x ny
 * cells (src/main.f90:22-26): U, the cell averages of ExactFunction at the
 * 4x4 Gauss points (FillInitialConditions, src/finitevolume.f90:281-333);
 * Uwb, those of ExactFunctionWB (ic 2 and 3; InitializeWBVariables); and
 * Bathymetry_X/_Y at every Gauss point in the bathymetry-table layout.
 * Any output may be NULL.  Returns 0, or -1 for an unknown ic (or Uwb
 * requested for an ic without a well-balanced state). */
int orc_initial_data(int ic, int bathy_flag, int nx, int ny, const double x0[2],
                     const double x1[2], double gravity, double *U, double *Uwb, double *dbdx,
                     double *dbdy);

/* PostProcessing (src/postprocessing.f90:24-64): the L1 error of U (3, nx, ny)
 * against the 4x4-Gauss cell averages of ExactFunction(ic, t) on the domain
 * of orc_initial_data, err[v] = sum |U - Uexact| * dx * dy.  Returns 0, or -1
 * for an unknown ic. */
int orc_l1_error(int ic, int bathy_flag, int nx, int ny, const double x0[2], const double x1[2],
                 double gravity, double t, const double *U, double err[3]);

/* Jacobi on the reference CSR pattern (src/jacobi_iteration.f90:52-87).
 * Mass (5n), diag (n), rhs (n) -> sol (n).  Returns the sweep count. */
int orc_jacobi(const orc_t *o, const double *Mass, const double *diag,
               const double *rhs, double *sol);

/* Overrides of the Jacobi stop rule (defaults 1000 and 1e-16). */
void orc_set_jacobi_limits(orc_t *o, int max_iter, double tol);

/* thetaDeC(m, r) of src/timediscretization.f90:491-506, out[4*(m-1)+(r-1)]. */
void orc_theta(double out[16]);

/* PATANKAR_DIV (src/timediscretization.f90:913-924). */
double orc_patankar_div(double prod, double h);
/* WENO5_FirstSweep / WENO5_SecondSweep on one 5-point stencil. */
void orc_weno5_first(const double q[5], double *wm, double *wp);
void orc_weno5_second(const double q[5], double w[4]);

#ifdef __cplusplus
}
#endif
#endif
