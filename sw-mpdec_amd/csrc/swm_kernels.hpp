// swm_kernels.hpp -- kernel parameter blocks and launch wrappers shared by the
// kernels (swm_kernels.hip) and the host orchestration (swm_api.cpp).
//
// Device layout of one solve (all FP64, structure of arrays, x fastest):
//   cell fields  [j][i]          j = 0..ny-1, i = 0..nx-1      (n = nx*ny)
//   x-face field [j][f]          f = 0..nx   (face f sits between Fortran
//                                 cells f and f+1; f = 0 and f = nx are the two
//                                 values of the periodic seam)
//   y-face field [g][i]          g = 0..ny
//   bathymetry   [yg*4+xg][j][i] 16 planes of d b/dx (and d b/dy)
#pragma once

#include <hip/hip_runtime.h>

namespace swm {

constexpr int kTile = 16;           // FVTD tiles are 16x16 cells, 256 threads
constexpr int kFaceTile = kTile - 1; // faces produced per tile row/column
constexpr int kRedBlock = 256;      // reduction kernels' block size
constexpr int kMaxSys = 3;          // Patankar systems solved together (nodes 2..4)

struct GridGeom {
  int nx, ny;
  double dx, dy, g;
  double wbnd[4];
  double wgp[4][4];
};

// Input/outputs of one FVTimeDerivative evaluation.
struct FvtdIO {
  const double* h;   // state (cell fields)
  const double* hu;
  const double* hv;
  double* FX[3];     // x-face fluxes (scratch, or FXWB when building the WB state)
  double* FY[3];
  double* Ut2;       // out: Ut(2), Ut(3)  (cell fields)
  double* Ut3;
  double* qx;        // out: (FX(1) - FXWB(1)) / dx per x face
  double* qy;        // out: (FY(1) - FYWB(1)) / dy per y face
  double* S2;        // out (optional): S(2), S(3) after WB subtraction, or SWB
  double* S3;
};

struct WbArrays {
  const double* FX[3];  // FXWB (x-face fields) or nullptr
  const double* FY[3];
  const double* S2;     // SWB(2), SWB(3) (cell fields)
  const double* S3;
};

struct Bathy {
  const double* bx;  // 16 planes, nullptr when BathymetryFlag == 0
  const double* by;
};

// Patankar system for one DeC node m (0-based 1..3)
struct PatSys {
  const double* hp;      // previous-correction depth of node m (denominators)
  double* MS;            // off-diagonals S, W, E, N and diagonal
  double* MW;
  double* ME;
  double* MN;
  double* D;
  double* x[2];          // Jacobi ping-pong
  double cth[4];         // dt*theta(m, r), r = 1..4
  int th_pos[4];         // theta(m, r) > 0
};

struct PatNodes {
  const double* qx[4];   // per DeC node r: x-face mass flux / dx
  const double* qy[4];
};

// Device-side state of the Jacobi iteration of one system.
struct JacState {
  int k;          // sweeps done
  int active;     // 1 while the reference loop would continue
  int cur;        // buffer holding the latest iterate (x[cur])
  int zero;       // 1 when no sweep ran (the reference returns sol = 0)
  int ambiguous;  // number of stop decisions that needed the exact sequential sum
  int nonconv;    // stopped at max_iter with the residual still above tol
  int pad[2];
  double residual;
};

struct JacCtl {
  JacState* st;        // [kMaxSys]
  double* partial;     // [kMaxSys][nblocks] (+ [nblocks] for sum |b| in slot kMaxSys)
  unsigned int* counter;
  int nsys;
  int max_iter;
  double tol;
};

struct TsCtl {
  double* partial;       // [nblocks][4]: max a, max b, nan flag
  unsigned int* counter;
  double* out;           // [3]: dt, LambdaMaxX, LambdaMaxY
  double cfl;
};

// ---- launch wrappers (swm_kernels.hip) ----------------------------------
void launch_fvtd_x(const GridGeom& G, const FvtdIO& io, hipStream_t s);
void launch_fvtd_y(const GridGeom& G, const FvtdIO& io, hipStream_t s);
void launch_fvtd_update(const GridGeom& G, const FvtdIO& io, const WbArrays& wb,
                        const Bathy& b, bool wb_init, hipStream_t s);
void launch_momentum(int n, const double* hun, const double* hvn, const double* const FU2[4],
                     const double* const FU3[4], const double cth[4], double* hu, double* hv,
                     hipStream_t s);
void launch_patankar_assemble(const GridGeom& G, const PatNodes& nodes, const PatSys* sys,
                              int nsys, const double* hn, const JacCtl& ctl, hipStream_t s);
void launch_jacobi_sweep(const GridGeom& G, const PatSys* sys, int nsys, const double* hn,
                         const JacCtl& ctl, hipStream_t s);
int jacobi_nblocks(const GridGeom& G);
void launch_jacobi_zero_fill(int n, const PatSys* sys, int nsys, const JacCtl& ctl,
                             hipStream_t s);
void launch_timestep(const GridGeom& G, const double* h, const double* hu, const double* hv,
                     const TsCtl& ctl, hipStream_t s);
int timestep_nblocks(const GridGeom& G);
void launch_aos_to_soa(int nx, int ny, int ghost, const double* aos, double* h, double* hu,
                       double* hv, hipStream_t s);
void launch_soa_to_aos(int nx, int ny, const double* h, const double* hu, const double* hv,
                       double* aos, hipStream_t s);
void launch_divsqrt(int n, const double* a, const double* b, double* q, double* r,
                    hipStream_t s);

}  // namespace swm
