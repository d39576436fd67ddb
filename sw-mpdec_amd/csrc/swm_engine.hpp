// swm_engine.hpp -- device data structures of the solve engine and the kernel
// launch interface shared by swm_kernels.hip and swm_engine.cpp.
//
// The engine advances any number of independent periodic grids ("jobs") in
// lockstep, one step of its modified-Patankar scheme (mPDeC5, mPEuler or
// mPDeC2) per engine step, jobs of one scheme together.  Every kernel launch covers all
// active jobs: a block looks up (job, tile) in a tile table built on the host
// whenever the job set changes.  A single plain solve is a one-job engine.
//
// Device layout of one job (FP64, structure of arrays, x fastest):
//   cell field    [j][i]         j = 0..ny-1, i = 0..nx-1   (n = nx*ny)
//   x-face field  [j][f]         f = 0..nx: face between Fortran cells f and
//                                f+1; f = 0 and f = nx are the two values of the
//                                periodic seam (src/finitevolume.f90:670)
//   y-face field  [g][i]         g = 0..ny
//   bathymetry    [yg*4+xg][j][i]
#pragma once

#include <hip/hip_runtime.h>

#include <string>

namespace swm {

// Sets the message swm_last_error(NULL) returns on the calling thread.
void set_thread_error(const std::string& msg);

constexpr int kTile = 16;             // k_fvtd_update tiles: 16x16 cells, 256 threads

// Tile shapes of the face kernels: k_fvtd_x runs xw x xh threads per tile and
// produces xw-1 faces per row; k_fvtd_y runs yw x yh threads and yh-1 faces
// per column.  Wider tiles recompute fewer halo sweeps; narrower ones waste
// fewer threads on small grids.  The engine's autotuner picks one.
struct FaceShape {
  int xw, xh, yw, yh;
};
constexpr int kNumShapes = 2;
constexpr FaceShape kShapes[kNumShapes] = {{16, 16, 16, 16}, {32, 16, 16, 32}};
constexpr int kCellBlock = 256;       // cells per block of the pointwise kernels
constexpr int kMaxSys = 3;            // Patankar systems solved together (nodes 2..4)
constexpr int kSlots = 3;             // concurrent FVTD evaluations per job
constexpr int kStates = 6;            // state / derivative slots of the explicit schemes

// Jacobi state of one system, written by the decide kernel.
struct JacState {
  int k;          // sweeps done
  int active;     // 1 while the reference loop would continue
  int cur;        // 0/1: which ping-pong buffer holds the latest iterate
  int zero;       // 1 when no sweep ran (the reference returns sol = 0)
  int ambiguous;  // stop decisions settled by the exact sequential sum
  int nonconv;    // stopped at max_iter with the residual above tolerance
  int pad[2];
  double residual;
};

// Steps of one job that one host synchronisation can cover: the device time
// loop (k_batch_begin, k_gate) runs up to kMaxBatch steps between reads of
// the status, recording each committed step.
constexpr int kMaxBatch = 16;

// Why a job's batch stopped (JobStatus::halt); while halt != 0 every gated
// kernel of the job returns at once and k_commit commits nothing.
enum HaltCode {
  kHaltNone = 0,
  kHaltBatch = 1,    // the batch's step count was reached
  kHaltCount = 2,    // the job's step count (steps_left) was reached
  kHaltReached = 3,  // t reached t_analyze or t_end (src/timediscretization.f90:85-94)
  kHaltOverrun = 4,  // a speculative step needed more Jacobi sweeps than launched
  kHaltDtBad = 5,    // TimeStep gave a non-finite dt (the step is not committed)
  kHaltMasked = 6,   // not part of the redo of another job's step
};

// One committed step of a batch.
struct StepRec {
  double dt;
  int iters[13];  // sweeps of each solve, in execution order
  int nonconv;    // solves stopped at the iteration limit
  int ambiguous;  // stop decisions settled by the exact sequential sum
  int pad;
};

// Per-job status the host reads back after every Jacobi batch.
struct JobStatus {
  JacState st[kMaxSys];
  double dt, lmx, lmy;
  int dt_bad;     // dt not finite
  int hmin_set;   // 1 once hmin holds a value
  double hmin;    // minimum depth over every DeC stage solution of the accepted
                  // steps since the last reset (NaN if any was NaN); kept only
                  // while tracking is on
  int hmin_step_set;  // the current attempt's minimum, merged into hmin by the
  double hmin_step;   // gate: a redone attempt's stages never reach hmin
  // the current step's solves, recorded after each correction's Jacobi phase
  int iters[13];  // sweeps of each solve, in execution order
  int nonconv;    // solves stopped at the iteration limit
  int ambiguous;  // stop decisions settled by the exact sequential sum
  int overrun;    // a solve was still running when its phase ended (the
                  // step ran with too few sweeps launched and must be redone)
  double dt_min;  // TimeStep's value before the time loop's clipping
  // the device time loop (k_batch_begin sets, k_gate advances)
  int halt;        // HaltCode
  int commit;      // 1: k_commit commits the step just run
  int batch_done;  // steps committed in this batch (records rec[0..batch_done))
  int batch_max;   // steps this batch may commit
  int loop;        // 1: dt = MIN(MIN(dt_min, t_an - t), MIN(dt_min, t_end - t))
                   // (src/timediscretization.f90:58-59) and t advances here
  int steps_left;  // step-count stop, counted down by k_gate: > 0 steps to go,
                   // 0 all run (k_batch_begin keeps the job halted), < 0 none
  double t, t_an, t_end;
  StepRec rec[kMaxBatch];
};

struct JobDev {
  // geometry and physics
  int nx, ny, n;
  int wb, bathy;
  int recfix;     // ReconstructionFix on flagged cells: 0 none (= WENO5), 1 first order, 2 MUSCL
  double* vout;   // 3n: h, hu, hv of the last FVTD input of a step (the writer's state), or null
  int nblk;       // cell blocks (kCellBlock cells) of this job
  double dx, dy, g, cfl;
  double wbnd[4];
  double wgp[4][4];
  // state U^n
  double *h, *hu, *hv;
  // DeC nodes m = 1..3 (index 0 unused): current depth, momenta, Jacobi ping-pong
  double* hres[4];
  double *hun[4], *hvn[4];
  double* xj[4][2];
  // per DeC node r = 0..3: FVTD outputs
  double *FU2[4], *FU3[4];
  double *qx[4], *qy[4];
  // FVTD face-flux scratch per concurrent slot
  double* FX[kSlots][3];
  double* FY[kSlots][3];
  // Patankar matrices per system: S, W, E, N off-diagonals and diagonal
  double* M[kMaxSys][5];
  // well-balanced reference (FXWB, FYWB, SWB) and bathymetry tables
  double *FXWB[3], *FYWB[3];
  double *SWB2, *SWB3;
  double *bx, *by;
  // State slots s = 0..5 (h, hu, hv), read by the FVTD: 0 is U^n, 1..3 the
  // DeC nodes (hres, hun, hvn); 4..5 alias Patankar matrices M, which the
  // explicit schemes that use them never assemble.
  double *sh[kStates], *shu[kStates], *shv[kStates];
  // Time-derivative slots r = 0..5 of the explicit schemes (Ut(1:3)): d2, d3
  // are FU2, FU3 for r < 4; d1 aliases the face-flux arrays qx (r < 4);
  // r = 4..5 alias M as well.
  double *d1[kStates], *d2[kStates], *d3[kStates];
  // reductions
  double* part;     // [kMaxSys + 1][nblk] Jacobi partial sums
  double* tspart;   // [nblk][3] TimeStep partial maxima
  JobStatus* status;  // this job's status record
};

struct TileRef {
  int job;
  int tile;
};

// Arguments shared by every launch of one engine step.  gated = 1 in the
// steps of a batch: a job whose JobStatus::halt is set is skipped.
struct EngineArgs {
  const JobDev* jobs;
  int max_iter;
  int gated;
  double tol;
};

// k_batch_begin's settings for every job of a batch (redo = 1: only jobs
// halted by an overrun run, the others are masked).
struct BatchArgs {
  int max_steps;
  int redo;
  int loop;
  double t, t_an, t_end;
};

// Which DeC nodes an FVTD launch evaluates (blockIdx.y picks one).
struct NodeList {
  int count;
  int m[kSlots];      // state slot evaluated (0 = U^n, 1..3 DeC nodes, 4..5 explicit)
  int out[kSlots];    // output slot: FU2/FU3/qx/qy[out], or d1..d3[out] with ut1
  int slot[kSlots];   // scratch slot used
  int wb_init;        // 1: building the well-balanced state (outputs to FXWB...)
  int debug_S;        // 1: also store S(2:3) to the slot's FX[.][0..1] scratch tail
  int ut1;            // 1: the build without PATANKAR: Ut(1:3) to d1..d3, no face fluxes
};

// One stage combination of an explicit scheme, per cell and variable, in the
// reference's evaluation order (every sum left to right):
//   mode 0: out = t_0 + t_1 + ...,   t_i = c_i*X_i (state) or (c_i*D_i)*dt (derivative)
//   mode 1: out = X_base + (c_0*D_0 + c_1*D_1 + ...)*dt
//   mode 2: out = X_base; out = out + (dt*c_i)*D_i for each term
// Up to kMaxOut outputs share the sources and differ in their weights c[o]
// (the DeC nodes of one correction).
constexpr int kMaxTerms = 8;
constexpr int kMaxOut = 3;
struct CombOp {
  int mode, base, nterm, nout;
  int out[kMaxOut];
  int deriv[kMaxTerms];
  int src[kMaxTerms];
  double c[kMaxOut][kMaxTerms];
};

// Patankar/Jacobi launch parameters (the same for every job of a step).
struct SysList {
  int count;
  int m[kMaxSys];       // DeC node of each system
  int src[4];           // buffer index r' holding node r's FU/q (0 at kk = 1, else r)
  int hp_is_un;         // 1: the denominators use h^n (first correction)
  int track_min;        // 1: fold the new stage depths into JobStatus::hmin
  int nr;               // thetaDeC columns summed (MSteps): 4 or 2
  int theta;            // THETA block: 0 mPDeC5, 1 the two-node schemes
  int hp_view;          // 1: denominators from hres[3], rhs from hres[2] (k_hp_view)
};

// The time schemes, by the reference's timescheme number: the modified-
// Patankar ones of the PATANKAR build and the explicit ones of the build
// without it (which compute Ut(1) instead of production/destruction terms).
struct SchemeDef {
  int timescheme;  // src/parameters.f90:144 numbering
  int msteps;      // DeC nodes (Patankar)
  int kcorr;       // corrections (Patankar)
  int theta;       // SysList::theta (Patankar)
  int nsolve;      // Jacobi solves per step (0: explicit)
  int final_state; // state slot holding U^{n+1} before the commit
};
constexpr int kNumSchemes = 7;
constexpr SchemeDef kSchemes[kNumSchemes] = {
    {5, 4, 5, 0, 13, 3}, {6, 2, 1, 1, 1, 1}, {7, 2, 2, 1, 2, 1},
    {1, 0, 0, 0, 0, 1},  {2, 0, 0, 0, 0, 1}, {3, 0, 0, 0, 0, 1}, {4, 4, 5, 0, 0, 3}};
inline bool is_explicit(const SchemeDef& d) { return d.timescheme < 5; }
inline const SchemeDef* find_scheme(int timescheme) {
  for (const SchemeDef& d : kSchemes)
    if (d.timescheme == timescheme) return &d;
  return nullptr;
}

// The reference's initial data (swm_initial.hip): InitialCondition,
// BathymetryFlag, MESH_X0, MESH_SX, Mesh_DX, Gravity, PI and the 4-point
// Gauss rule of BuildMesh (quadNodes1D, quadWeights1D; src/mesh.f90:127-132).
struct InitialSetup {
  int ic, bflag;
  double x0[2], sx[2], dx[2], g, pi, t;  // t: time of ExactFunction
  double node[4], w[4];
};

// The benchmark families' parameters (swm_family) with the mesh of the
// periodic unit square and the 4-point Gauss rule.
struct FamilySetup {
  int kind, nbumps;
  double dx, dy, g;
  double node[4], w[4];
  double eps;
  double A[8], cx[8], cy[8], s[8];
  double amp, px, py, ps;
  double radius, h_in, h_out;
};
void launch_family_data(const FamilySetup& f, int nx, int ny, double* h, double* hu, double* hv,
                        double* hwb, double* huwb, double* hvwb, double* bx, double* by,
                        hipStream_t s);
void launch_det_exp(int n, const double* x, double* y, hipStream_t s);

// h/hu/hv: the state; hwb/huwb/hvwb (may be null): the well-balanced
// state; bx/by (may be null): the bathymetry tables.
void launch_initial_data(const InitialSetup& c, int nx, int ny, double* h, double* hu, double* hv,
                         double* hwb, double* huwb, double* hvwb, double* bx, double* by,
                         hipStream_t s);
// PostProcessing's L1 error of the state against ExactFunction(c.ic, c.t)
// into err[3] (device); scratch: 3*nx*ny doubles.
void launch_l1_error(const InitialSetup& c, int nx, int ny, const double* h, const double* hu,
                     const double* hv, double* scratch, double* err, hipStream_t s);
// ConsToPrim of n cells into (1:3, n) order.
void launch_primitives(size_t n, const double* h, const double* hu, const double* hv, double* V,
                       hipStream_t s);

// ---- launches (swm_kernels.hip) -------------------------------------------
// xt / yt: face tiles of shape kShapes[shape]; ct: k_fvtd_update tiles.
void launch_fvtd(const EngineArgs& E, int shape, const TileRef* xt, int nxt, const TileRef* yt,
                 int nyt, const TileRef* ct, int nct, const NodeList& L, hipStream_t s);
void launch_assemble(const EngineArgs& E, const TileRef* cb, int ncb, const int* jl, int njobs,
                     const SysList& S, hipStream_t s);
void launch_sweep(const EngineArgs& E, const TileRef* cb, int ncb, const int* jl, int njobs,
                  const SysList& S, hipStream_t s);
void launch_finalize(const EngineArgs& E, const TileRef* cb, int ncb, const int* jl, int njobs,
                     const SysList& S, hipStream_t s);
void launch_timestep(const EngineArgs& E, const TileRef* cb, int ncb, const int* jl, int njobs,
                     hipStream_t s);
void launch_commit(const EngineArgs& E, const TileRef* cb, int ncb, int final_node,
                   hipStream_t s);
// k_gate: the commit decision and record of each job's step
void launch_gate(const EngineArgs& E, const int* jl, int njobs, int nsolve, hipStream_t s);
void launch_batch_begin(const EngineArgs& E, const int* jl, int njobs, const BatchArgs& B,
                        hipStream_t s);
// sets JobStatus::steps_left of one job (its step-count stop)
void launch_set_steps_left(JobStatus* st, int steps_left, hipStream_t s);
void launch_hp_view(const EngineArgs& E, const TileRef* cb, int ncb, int from_un, hipStream_t s);
void launch_combine(const EngineArgs& E, const TileRef* cb, int ncb, const CombOp& op,
                    hipStream_t s);
void launch_solve_done(const EngineArgs& E, const int* jl, int njobs, const SysList& S,
                       int first_solve, hipStream_t s);
void launch_copy_wb(const EngineArgs& E, const TileRef* cb, int ncb, int slot, hipStream_t s);
// node m's state into vout, for the jobs that keep the writer's state
void launch_keep_state(const EngineArgs& E, const TileRef* cb, int ncb, int m, hipStream_t s);
void launch_aos_to_soa(int nx, int ny, int ghost, const double* aos, double* h, double* hu,
                       double* hv, hipStream_t s);
void launch_soa_to_aos(int nx, int ny, const double* h, const double* hu, const double* hv,
                       double* aos, hipStream_t s);
void launch_divsqrt(int n, const double* a, const double* b, double* q, double* r,
                    hipStream_t s);
void launch_shared_div(int n, const double* a, const double* b, double* q, hipStream_t s);
void launch_weno5(int n, const double* q, double* first, double* second, hipStream_t s);

}  // namespace swm
