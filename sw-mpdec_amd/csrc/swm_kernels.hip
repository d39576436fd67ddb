// swm_kernels.hip -- FP64 HIP kernels (gfx950) of the mPDeC5 / WENO5
// PATANKAR time step of sw-mpdec.
//
// One FVTimeDerivative (src/finitevolume.f90:402-440) is three launches:
//   k_fvtd_x       x faces:  halo tile -> ConsToPrim -> WENO5 first sweep along
//                  x (shared through LDS) -> second sweep along y at 4 Gauss
//                  points -> positivity limiter on h -> Rusanov -> Gauss sum
//   k_fvtd_y       the same for y faces
//   k_fvtd_update  source term (h-only WENO5 point values x -d b) and
//                  UpdateTimeDerivative: Ut(2:3), face mass fluxes / dx
// The Patankar matrix assembly fuses the first Jacobi sweep; every Jacobi
// sweep writes per-block sums of |x_new - x_old| and a one-block kernel takes
// the stop decision of src/jacobi_iteration.f90:71 on the device.  (No grid
// fences: on a multi-XCD device an agent-scope fence writes back the L2, so
// the kernel boundary is the cheap way to publish the partial sums.)
#include <hip/hip_runtime.h>

#include "swm_kernels.hpp"
#include "swm_numerics.hpp"

namespace swm {

// 0-based index of the Fortran cell c (1..n, or a periodic ghost) on an axis
// of n cells.
__device__ __forceinline__ int wrap1(int c, int n) {
  int m = (c - 1) % n;
  return m < 0 ? m + n : m;
}

// ------------------------------------------------------------------------
// x faces
// ------------------------------------------------------------------------
// Block (bx, by): Fortran cells c = 15*bx + tx (tx = 0..15, c in 0..nx+1;
// c = 0 is the ghost of nx and c = nx+1 the ghost of 1), rows
// j = 16*by + 1 + ty.  Thread (tx, ty) reconstructs cell (c, j) and, for
// tx < 15, computes the face f = c between cells c and c+1 (f = 0..nx).
// The positivity limiter runs for c = 0..nx only (src/finitevolume.f90:670),
// so the face f = nx uses the unlimited left value of cell nx+1 while f = 0
// uses the limited one of cell 1 -- the two values of the periodic seam.
__global__ __launch_bounds__(256) void k_fvtd_x(GridGeom G, FvtdIO io) {
  __shared__ double smem[3 * 400 + 2 * 3 * 320];
  double* sV = smem;               // [3][20 rows][20 cols]
  double* sQM = smem + 1200;       // [3][20 rows][16 cols]
  double* sQP = sQM + 960;
  const int nx = G.nx, ny = G.ny;
  const int tid = threadIdx.x, tx = tid & 15, ty = tid >> 4;
  const int c0 = blockIdx.x * kFaceTile;
  const int j0 = blockIdx.y * kTile + 1;

  for (int t = tid; t < 400; t += 256) {
    int lx = t % 20, ly = t / 20;
    size_t k = (size_t)wrap1(j0 - 2 + ly, ny) * nx + wrap1(c0 - 2 + lx, nx);
    double h = io.h[k], u, v;
    cons_to_prim(h, io.hu[k], io.hv[k], u, v);
    sV[t] = h;
    sV[400 + t] = u;
    sV[800 + t] = v;
  }
  __syncthreads();
  for (int t = tid; t < 960; t += 256) {
    int var = t / 320, r = t % 320, ly = r >> 4, lx = r & 15;
    const double* q = sV + var * 400 + ly * 20 + lx;
    double wm, wp;
    weno5_first(q[0], q[1], q[2], q[3], q[4], wm, wp);
    sQM[var * 320 + r] = wm;
    sQP[var * 320 + r] = wp;
  }
  __syncthreads();
  double WM[3][4], WP[3][4];
#pragma unroll
  for (int var = 0; var < 3; ++var) {
    const double* qm = sQM + var * 320 + ty * 16 + tx;
    weno5_second(qm[0], qm[16], qm[32], qm[48], qm[64], WM[var]);
    const double* qp = sQP + var * 320 + ty * 16 + tx;
    weno5_second(qp[0], qp[16], qp[32], qp[48], qp[64], WP[var]);
  }
  const double vh = sV[(ty + 2) * 20 + tx + 2];
  const int c = c0 + tx;
  if (c <= nx) positivity_limit(vh, G.wbnd, WM[0], WP[0]);
  __syncthreads();
  double* sWM = smem;  // [12][256]
#pragma unroll
  for (int a = 0; a < 12; ++a) sWM[a * 256 + tid] = WM[a >> 2][a & 3];
  __syncthreads();
  const int j = j0 + ty;
  if (tx < kFaceTile && c <= nx && j <= ny) {
    double acc1 = 0.0, acc2 = 0.0, acc3 = 0.0;
#pragma unroll
    for (int g = 0; g < 4; ++g) {
      double F1, F2, F3;
      rusanov<1, 0, 0, 1>(WP[0][g], WP[1][g], WP[2][g], sWM[(0 + g) * 256 + tid + 1],
                          sWM[(4 + g) * 256 + tid + 1], sWM[(8 + g) * 256 + tid + 1], G.g, F1,
                          F2, F3);
      acc1 = acc1 + G.wbnd[g] * F1;
      acc2 = acc2 + G.wbnd[g] * F2;
      acc3 = acc3 + G.wbnd[g] * F3;
    }
    size_t o = (size_t)(j - 1) * (nx + 1) + c;
    io.FX[0][o] = acc1;
    io.FX[1][o] = acc2;
    io.FX[2][o] = acc3;
  }
}

// ------------------------------------------------------------------------
// y faces: the transpose of k_fvtd_x.  Block (bx, by): columns
// i = 16*bx + 1 + tx, Fortran rows r = 15*by + ty (0..ny+1); face g = r.
// ------------------------------------------------------------------------
__global__ __launch_bounds__(256) void k_fvtd_y(GridGeom G, FvtdIO io) {
  __shared__ double smem[3 * 400 + 2 * 3 * 320];
  double* sV = smem;          // [3][20 rows][20 cols]
  double* sQM = smem + 1200;  // [3][16 rows][20 cols]
  double* sQP = sQM + 960;
  const int nx = G.nx, ny = G.ny;
  const int tid = threadIdx.x, tx = tid & 15, ty = tid >> 4;
  const int i0 = blockIdx.x * kTile + 1;
  const int r0 = blockIdx.y * kFaceTile;

  for (int t = tid; t < 400; t += 256) {
    int lx = t % 20, ly = t / 20;
    size_t k = (size_t)wrap1(r0 - 2 + ly, ny) * nx + wrap1(i0 - 2 + lx, nx);
    double h = io.h[k], u, v;
    cons_to_prim(h, io.hu[k], io.hv[k], u, v);
    sV[t] = h;
    sV[400 + t] = u;
    sV[800 + t] = v;
  }
  __syncthreads();
  for (int t = tid; t < 960; t += 256) {
    int var = t / 320, r = t % 320, ly = r / 20, lx = r % 20;
    const double* q = sV + var * 400 + ly * 20 + lx;
    double wm, wp;
    weno5_first(q[0], q[20], q[40], q[60], q[80], wm, wp);
    sQM[var * 320 + r] = wm;
    sQP[var * 320 + r] = wp;
  }
  __syncthreads();
  double WM[3][4], WP[3][4];
#pragma unroll
  for (int var = 0; var < 3; ++var) {
    const double* qm = sQM + var * 320 + ty * 20 + tx;
    weno5_second(qm[0], qm[1], qm[2], qm[3], qm[4], WM[var]);
    const double* qp = sQP + var * 320 + ty * 20 + tx;
    weno5_second(qp[0], qp[1], qp[2], qp[3], qp[4], WP[var]);
  }
  const double vh = sV[(ty + 2) * 20 + tx + 2];
  const int r = r0 + ty;
  if (r <= ny) positivity_limit(vh, G.wbnd, WM[0], WP[0]);
  __syncthreads();
  double* sWM = smem;
#pragma unroll
  for (int a = 0; a < 12; ++a) sWM[a * 256 + tid] = WM[a >> 2][a & 3];
  __syncthreads();
  const int i = i0 + tx;
  if (ty < kFaceTile && r <= ny && i <= nx) {
    double acc1 = 0.0, acc2 = 0.0, acc3 = 0.0;
#pragma unroll
    for (int g = 0; g < 4; ++g) {
      double F1, F2, F3;
      rusanov<0, 1, -1, 0>(WP[0][g], WP[1][g], WP[2][g], sWM[(0 + g) * 256 + tid + 16],
                           sWM[(4 + g) * 256 + tid + 16], sWM[(8 + g) * 256 + tid + 16], G.g,
                           F1, F2, F3);
      acc1 = acc1 + G.wbnd[g] * F1;
      acc2 = acc2 + G.wbnd[g] * F2;
      acc3 = acc3 + G.wbnd[g] * F3;
    }
    size_t o = (size_t)r * nx + (i - 1);
    io.FY[0][o] = acc1;
    io.FY[1][o] = acc2;
    io.FY[2][o] = acc3;
  }
}

// ------------------------------------------------------------------------
// Source term + UpdateTimeDerivative.  Block: 16x16 cells (1-based
// i = 16*bx + 1 + tx, j = 16*by + 1 + ty).
// SourceTerms (src/equation.f90:352-388): h point values at the 4x4 Gauss
// points from a WENO5 second sweep along y (columns i-2..i+2) followed by one
// along x per y Gauss point; SW = -g*h*db; S = sum_{x-GP} sum_{y-GP} W*SW.
// ------------------------------------------------------------------------
template <bool BATHY>
__global__ __launch_bounds__(256) void k_fvtd_update(GridGeom G, FvtdIO io, WbArrays wb,
                                                     Bathy B, int wb_init) {
  const int nx = G.nx, ny = G.ny;
  const int tid = threadIdx.x, tx = tid & 15, ty = tid >> 4;
  const int i0 = blockIdx.x * kTile + 1, j0 = blockIdx.y * kTile + 1;
  const int i = i0 + tx, j = j0 + ty;
  double S2 = 0.0, S3 = 0.0;
  if constexpr (BATHY) {
    __shared__ double sH[400];            // [20 rows][20 cols]
    __shared__ double sVt[4 * 16 * 20];   // [yg][16 rows][20 cols]
    for (int t = tid; t < 400; t += 256) {
      int lx = t % 20, ly = t / 20;
      sH[t] = io.h[(size_t)wrap1(j0 - 2 + ly, ny) * nx + wrap1(i0 - 2 + lx, nx)];
    }
    __syncthreads();
    for (int t = tid; t < 320; t += 256) {
      int ly = t / 20, lx = t % 20;
      const double* q = sH + ly * 20 + lx;
      double w[4];
      weno5_second(q[0], q[20], q[40], q[60], q[80], w);
#pragma unroll
      for (int yg = 0; yg < 4; ++yg) sVt[yg * 320 + t] = w[yg];
    }
    __syncthreads();
    if (i <= nx && j <= ny) {
      double hgp[4][4];  // [yg][xg]
#pragma unroll
      for (int yg = 0; yg < 4; ++yg) {
        const double* q = sVt + yg * 320 + ty * 20 + tx;
        weno5_second(q[0], q[1], q[2], q[3], q[4], hgp[yg]);
      }
      const size_t n = (size_t)nx * ny;
      const size_t k = (size_t)(j - 1) * nx + (i - 1);
#pragma unroll
      for (int xg = 0; xg < 4; ++xg)
#pragma unroll
        for (int yg = 0; yg < 4; ++yg) {
          double h = hgp[yg][xg];
          double bx = B.bx[(size_t)(yg * 4 + xg) * n + k];
          double by = B.by[(size_t)(yg * 4 + xg) * n + k];
          double sw2 = -G.g * h * bx;
          double sw3 = -G.g * h * by;
          S2 = S2 + G.wgp[xg][yg] * sw2;
          S3 = S3 + G.wgp[xg][yg] * sw3;
        }
    }
  }
  if (i > nx || j > ny) return;
  const size_t k = (size_t)(j - 1) * nx + (i - 1);
  if (wb_init) {  // building the well-balanced reference: keep S as SWB
    io.S2[k] = S2;
    io.S3[k] = S3;
    return;
  }
  const size_t fe = (size_t)(j - 1) * (nx + 1) + i, fw = fe - 1;     // x faces i, i-1
  const size_t gn = (size_t)j * nx + (i - 1), gs = gn - nx;           // y faces j, j-1
  const bool use_wb = wb.FX[0] != nullptr;
  if (use_wb) {
    S2 = S2 - wb.S2[k];
    S3 = S3 - wb.S3[k];
  }
  if (io.S2) {
    io.S2[k] = S2;
    io.S3[k] = S3;
  }
  double fxe[3], fxw[3], fyn[3], fys[3];
#pragma unroll
  for (int v = 0; v < 3; ++v) {
    fxe[v] = io.FX[v][fe];
    fxw[v] = io.FX[v][fw];
    fyn[v] = io.FY[v][gn];
    fys[v] = io.FY[v][gs];
    if (use_wb) {  // FX = FX - FXWB; FY = FY - FYWB (src/finitevolume.f90:485-488)
      fxe[v] = fxe[v] - wb.FX[v][fe];
      fxw[v] = fxw[v] - wb.FX[v][fw];
      fyn[v] = fyn[v] - wb.FY[v][gn];
      fys[v] = fys[v] - wb.FY[v][gs];
    }
  }
  // Ut(2:3) = S - (FX(i)-FX(i-1))/dx - (FY(j)-FY(j-1))/dy  (:525-527)
  io.Ut2[k] = S2 - (fxe[1] - fxw[1]) / G.dx - (fyn[1] - fys[1]) / G.dy;
  io.Ut3[k] = S3 - (fxe[2] - fxw[2]) / G.dx - (fyn[2] - fys[2]) / G.dy;
  // mass flux per face / dx: the production/destruction entries of
  // src/finitevolume.f90:496-523 are max(+-q, 0) and -min(+-q, 0) of these.
  io.qx[fe] = fxe[0] / G.dx;
  io.qy[gn] = fyn[0] / G.dy;
  if (i == 1) io.qx[fw] = fxw[0] / G.dx;
  if (j == 1) io.qy[gs] = fys[0] / G.dy;
}

// ------------------------------------------------------------------------
// DeC momentum update, src/timediscretization.f90:531-543:
//   Ua(m,2:3) = Ua(1,2:3); DO r=1,4: Ua(m) = Ua(m) + (dt*theta(m,r))*FUp(r)
// ------------------------------------------------------------------------
struct MomArgs {
  const double* hun;
  const double* hvn;
  const double* F2[4];
  const double* F3[4];
  double c[4];
  double* hu;
  double* hv;
  int n;
};

__global__ __launch_bounds__(256) void k_momentum(MomArgs a) {
  int k = blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= a.n) return;
  double u = a.hun[k], v = a.hvn[k];
#pragma unroll
  for (int r = 0; r < 4; ++r) {
    u = u + a.c[r] * a.F2[r][k];
    v = v + a.c[r] * a.F3[r][k];
  }
  a.hu[k] = u;
  a.hv[k] = v;
}

// ------------------------------------------------------------------------
// Reductions and the Jacobi stop decision
// ------------------------------------------------------------------------
__device__ __forceinline__ double wave_sum(double v) {
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, 64);
  return v;
}

// Sum over a 256-thread block; result valid in thread 0.  red: >= 4 doubles.
__device__ __forceinline__ double block_sum(double v, double* red) {
  v = wave_sum(v);
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  __syncthreads();
  if (lane == 0) red[w] = v;
  __syncthreads();
  double r = 0.0;
  if (threadIdx.x == 0) r = ((red[0] + red[1]) + red[2]) + red[3];
  return r;
}

// Decide `tol < SUM(d)/n` (src/jacobi_iteration.f90:71,77) where the
// reference sums the d_i >= 0 sequentially in row order and S is the same sum
// taken in parallel.  Both roundings differ from the exact sum by at most
// (n-1)u*sum(d), so whenever tol lies outside [S - B, S + B]/n with
// B = 8nuS the decision is that of the sequential sum.  Inside the window the
// sequential sum is recomputed exactly (by the caller).  Returns 1 continue,
// 0 stop, -1 ambiguous.
__device__ __forceinline__ int certified_continue(double S, int n, double tol) {
  if (isnan(S)) return 0;
  if (isinf(S)) return 1;
  const double u = 1.1102230246251565e-16;  // 2^-53
  double B = 8.0 * (double)n * u * S;
  double lo = (S - B) / (double)n, hi = (S + B) / (double)n;
  if (tol < lo) return 1;
  if (!(tol < hi)) return 0;
  return -1;
}

// Exact replay of the reference's sequential residual for system s.
// mode 0: sum |b_i|; mode 1: sum |x_new - x_old|.
__device__ double sequential_residual(int n, const double* a, const double* b, int mode) {
  double r = 0.0;
  for (int i = 0; i < n; ++i) r = r + (mode == 0 ? fabs(0.0 - a[i]) : fabs(a[i] - b[i]));
  return r / (double)n;
}

struct PatArgs {
  GridGeom G;
  PatNodes nodes;
  PatSys sys[kMaxSys];
  const double* hn;
  JacCtl ctl;
};

__device__ __forceinline__ void patankar_slot(const PatSys& S, const double q[4], bool east,
                                              double hK, double hL, double& M, double& D) {
  // production / destruction of this slot at the 4 DeC nodes, from the face
  // mass flux q = F/dx: west/south slots take max(F,0)/dx, -min(F,0)/dx;
  // east/north slots the same of -F (src/finitevolume.f90:496-523)
  M = 0.;
#pragma unroll
  for (int r = 0; r < 4; ++r) {
    double P, De;
    if (!east) {
      P = (q[r] > 0.) ? q[r] : 0.;
      De = (q[r] < 0.) ? -q[r] : -0.;
    } else {
      P = (q[r] < 0.) ? -q[r] : 0.;
      De = (q[r] > 0.) ? q[r] : -0.;
    }
    const double c = S.cth[r];
    if (S.th_pos[r]) {
      M = M - c * (patankar_div(P, hL));
      D = D + c * (patankar_div(De, hK));
    } else {
      D = D - c * (patankar_div(P, hK));
      M = M + c * (patankar_div(De, hL));
    }
  }
}

// One Patankar system: MODIFIED_PATANKAR_MATRIX_INVERSION's assembly
// (src/timediscretization.f90:795-887) for cell k, fused with the first
// Jacobi sweep from x^0 = rhs (src/jacobi_iteration.f90:67-76).  Returns
// |x^1 - rhs| for the residual.
template <int S_>
__device__ __forceinline__ double assemble_cell(const PatArgs& A, int k, int kS, int kW, int kE,
                                                int kN, const double* qS, const double* qW,
                                                const double* qE, const double* qN, double b,
                                                double bS, double bW, double bE, double bN) {
  const PatSys& S = A.sys[S_];
  const double hK = S.hp[k];
  double D = 1.;
  double mS, mW, mE, mN;
  patankar_slot(S, qS, false, hK, S.hp[kS], mS, D);
  patankar_slot(S, qW, false, hK, S.hp[kW], mW, D);
  patankar_slot(S, qE, true, hK, S.hp[kE], mE, D);
  patankar_slot(S, qN, true, hK, S.hp[kN], mN, D);
  S.MS[k] = mS;
  S.MW[k] = mW;
  S.ME[k] = mE;
  S.MN[k] = mN;
  S.D[k] = D;
  // sparse_matmul in CSR order S, W, self (value 0), E, N, then
  // sol = (rhs - prod)/diag
  double p = 0.;
  p = p + mS * bS;
  p = p + mW * bW;
  p = p + 0. * b;
  p = p + mE * bE;
  p = p + mN * bN;
  double x1 = (b - p) / D;
  S.x[1][k] = x1;
  return fabs(x1 - b);
}

// Assembly of up to 3 nodes' systems; per-block partial sums of |rhs| and of
// |x^1 - rhs| go to ctl.partial, the stop decision is k_jacobi_decide's.
template <int NSYS>
__global__ __launch_bounds__(256) void k_patankar_assemble(PatArgs A) {
  __shared__ double red[4];
  const int nx = A.G.nx, ny = A.G.ny, n = nx * ny;
  const int k = blockIdx.x * blockDim.x + threadIdx.x;
  double d0 = 0., d1 = 0., d2 = 0., db = 0.;
  if (k < n) {
    const int i = k % nx, j = k / nx;
    const int kS = (j == 0 ? ny - 1 : j - 1) * nx + i;
    const int kN = (j == ny - 1 ? 0 : j + 1) * nx + i;
    const int kW = j * nx + (i == 0 ? nx - 1 : i - 1);
    const int kE = j * nx + (i == nx - 1 ? 0 : i + 1);
    const size_t fW = (size_t)j * (nx + 1) + i, fE = fW + 1;
    const size_t gS = (size_t)j * nx + i, gN = gS + nx;
    double qS[4], qW[4], qE[4], qN[4];
#pragma unroll
    for (int r = 0; r < 4; ++r) {
      qS[r] = A.nodes.qy[r][gS];
      qW[r] = A.nodes.qx[r][fW];
      qE[r] = A.nodes.qx[r][fE];
      qN[r] = A.nodes.qy[r][gN];
    }
    const double b = A.hn[k];
    const double bS = A.hn[kS], bW = A.hn[kW], bE = A.hn[kE], bN = A.hn[kN];
    db = fabs(0.0 - b);
    d0 = assemble_cell<0>(A, k, kS, kW, kE, kN, qS, qW, qE, qN, b, bS, bW, bE, bN);
    if constexpr (NSYS > 1)
      d1 = assemble_cell<1>(A, k, kS, kW, kE, kN, qS, qW, qE, qN, b, bS, bW, bE, bN);
    if constexpr (NSYS > 2)
      d2 = assemble_cell<2>(A, k, kS, kW, kE, kN, qS, qW, qE, qN, b, bS, bW, bE, bN);
  }
  const int nb = gridDim.x;
  double t = block_sum(d0, red);
  if (threadIdx.x == 0) A.ctl.partial[blockIdx.x] = t;
  if constexpr (NSYS > 1) {
    t = block_sum(d1, red);
    if (threadIdx.x == 0) A.ctl.partial[(size_t)nb + blockIdx.x] = t;
  }
  if constexpr (NSYS > 2) {
    t = block_sum(d2, red);
    if (threadIdx.x == 0) A.ctl.partial[2 * (size_t)nb + blockIdx.x] = t;
  }
  t = block_sum(db, red);
  if (threadIdx.x == 0) A.ctl.partial[(size_t)kMaxSys * nb + blockIdx.x] = t;
}

struct JacArgs {
  int nx, ny;
  PatSys sys[kMaxSys];
  const double* hn;
  JacCtl ctl;
};

// One Jacobi sweep of every still-active system (src/jacobi_iteration.f90:72-80);
// per-block partial sums of |x_new - x_old|.
template <int NSYS>
__global__ __launch_bounds__(256) void k_jacobi_sweep(JacArgs A) {
  __shared__ double red[4];
  int act[NSYS], cur[NSYS];
  bool any = false;
#pragma unroll
  for (int s = 0; s < NSYS; ++s) {
    act[s] = A.ctl.st[s].active;
    cur[s] = A.ctl.st[s].cur;
    any = any || act[s];
  }
  if (!any) return;  // every system has stopped: this launch is a no-op
  const int nx = A.nx, ny = A.ny, n = nx * ny;
  const int k = blockIdx.x * blockDim.x + threadIdx.x;
  double dsum[NSYS];
#pragma unroll
  for (int s = 0; s < NSYS; ++s) dsum[s] = 0.;
  if (k < n) {
    const int i = k % nx, j = k / nx;
    const int kS = (j == 0 ? ny - 1 : j - 1) * nx + i;
    const int kN = (j == ny - 1 ? 0 : j + 1) * nx + i;
    const int kW = j * nx + (i == 0 ? nx - 1 : i - 1);
    const int kE = j * nx + (i == nx - 1 ? 0 : i + 1);
    const double b = A.hn[k];
#pragma unroll
    for (int s = 0; s < NSYS; ++s) {
      if (!act[s]) continue;
      const PatSys& S = A.sys[s];
      const double* xo = S.x[cur[s]];
      double* xn = S.x[1 - cur[s]];
      double p = 0.;
      p = p + S.MS[k] * xo[kS];
      p = p + S.MW[k] * xo[kW];
      p = p + 0. * xo[k];
      p = p + S.ME[k] * xo[kE];
      p = p + S.MN[k] * xo[kN];
      double xv = (b - p) / S.D[k];
      xn[k] = xv;
      dsum[s] = fabs(xv - xo[k]);
    }
  }
  const int nb = gridDim.x;
#pragma unroll
  for (int s = 0; s < NSYS; ++s) {
    if (!act[s]) continue;
    double t = block_sum(dsum[s], red);
    if (threadIdx.x == 0) A.ctl.partial[(size_t)s * nb + blockIdx.x] = t;
  }
}

// Sum of the nb per-block partials of slot s, in block order (one block).
__device__ double sum_partials(const double* part, int nb, double* red) {
  double t = 0.0;
  for (int b = threadIdx.x; b < nb; b += blockDim.x) t += part[b];
  t = block_sum(t, red);
  __shared__ double bc;
  if (threadIdx.x == 0) bc = t;
  __syncthreads();
  t = bc;
  __syncthreads();
  return t;
}

// Stop decision after the assembly (FIRST) or after a sweep: one block.
template <bool FIRST>
__global__ __launch_bounds__(256) void k_jacobi_decide(JacArgs A, int nb) {
  __shared__ double red[4];
  const int n = A.nx * A.ny;
  double tot[kMaxSys + 1];
  int act[kMaxSys], cur[kMaxSys];
  for (int s = 0; s < A.ctl.nsys; ++s) {
    act[s] = FIRST ? 1 : A.ctl.st[s].active;
    cur[s] = FIRST ? 0 : A.ctl.st[s].cur;
  }
  for (int s = 0; s < A.ctl.nsys; ++s)
    tot[s] = act[s] ? sum_partials(A.ctl.partial + (size_t)s * nb, nb, red) : 0.0;
  if (FIRST) tot[kMaxSys] = sum_partials(A.ctl.partial + (size_t)kMaxSys * nb, nb, red);
  if (threadIdx.x != 0) return;
  if (FIRST) {
    // initial residual SUM(ABS(0 - rhs))/n, the same for every system
    int c0 = certified_continue(tot[kMaxSys], n, A.ctl.tol);
    int amb0 = 0;
    if (c0 < 0) {
      double r = sequential_residual(n, A.hn, nullptr, 0);
      c0 = (A.ctl.tol < r) ? 1 : 0;
      amb0 = 1;
    }
    for (int s = 0; s < A.ctl.nsys; ++s) {
      JacState& st = A.ctl.st[s];
      st.ambiguous = amb0;
      st.cur = 1;
      if (!c0 || A.ctl.max_iter <= 0) {
        st.k = 0;
        st.active = 0;
        st.zero = 1;
        st.nonconv = 0;
        st.residual = tot[kMaxSys] / (double)n;
        continue;
      }
      st.k = 1;
      st.zero = 0;
      int c1 = certified_continue(tot[s], n, A.ctl.tol);
      if (c1 < 0) {
        double r = sequential_residual(n, A.sys[s].x[1], A.hn, 1);
        c1 = (A.ctl.tol < r) ? 1 : 0;
        st.ambiguous += 1;
      }
      st.residual = tot[s] / (double)n;
      st.active = (c1 && st.k < A.ctl.max_iter) ? 1 : 0;
      st.nonconv = (c1 && !st.active) ? 1 : 0;
    }
  } else {
    for (int s = 0; s < A.ctl.nsys; ++s) {
      if (!act[s]) continue;
      JacState& st = A.ctl.st[s];
      st.k += 1;
      st.cur = 1 - cur[s];
      int c = certified_continue(tot[s], n, A.ctl.tol);
      if (c < 0) {
        double r = sequential_residual(n, A.sys[s].x[st.cur], A.sys[s].x[cur[s]], 1);
        c = (A.ctl.tol < r) ? 1 : 0;
        st.ambiguous += 1;
      }
      st.residual = tot[s] / (double)n;
      st.active = (c && st.k < A.ctl.max_iter) ? 1 : 0;
      st.nonconv = (c && !st.active) ? 1 : 0;
    }
  }
}

// When no sweep ran the reference returns sol = 0 (src/jacobi_iteration.f90:68).
struct ZeroArgs {
  double* x[kMaxSys];
  JacState* st;
  int nsys, n;
};
__global__ void k_jacobi_zero_fill(ZeroArgs a) {
  int k = blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= a.n) return;
  for (int s = 0; s < a.nsys; ++s)
    if (a.st[s].zero) a.x[s][k] = 0.;
}

// ------------------------------------------------------------------------
// TimeStep (src/equation.f90:791-837).  Without NaN the sequential MAX/MIN
// fold equals dt = CFL*MIN(MIN(HUGE, dx/LX), dy/LY) with LX, LY the plain
// maxima of |u + sqrt(g h)|, |v + sqrt(g h)| and 0; with a NaN anywhere the
// fold is order dependent, so k_timestep_final replays it sequentially.
// ------------------------------------------------------------------------
struct TsArgs {
  GridGeom G;
  const double* h;
  const double* hu;
  const double* hv;
  TsCtl ctl;
};

__device__ __forceinline__ void wave_speeds(const GridGeom& G, double h, double hu, double hv,
                                            double& a, double& b) {
  double u, v;
  cons_to_prim(h, hu, hv, u, v);
  double fx = u + sqrt(G.g * h);
  double fy = v + sqrt(G.g * h);
  a = fabs(fx);
  b = fabs(fy);
}

__global__ __launch_bounds__(256) void k_timestep(TsArgs A) {
  __shared__ double sa[256], sb[256];
  __shared__ int snan[256];
  const int n = A.G.nx * A.G.ny;
  double ma = 0.0, mb = 0.0;
  int isn = 0;
  for (int k = blockIdx.x * blockDim.x + threadIdx.x; k < n; k += gridDim.x * blockDim.x) {
    double a, b;
    wave_speeds(A.G, A.h[k], A.hu[k], A.hv[k], a, b);
    if (isnan(a) || isnan(b)) isn = 1;
    if (a > ma) ma = a;
    if (b > mb) mb = b;
  }
  sa[threadIdx.x] = ma;
  sb[threadIdx.x] = mb;
  snan[threadIdx.x] = isn;
  __syncthreads();
  for (int off = 128; off > 0; off >>= 1) {
    if (threadIdx.x < off) {
      sa[threadIdx.x] = max_r(sa[threadIdx.x], sa[threadIdx.x + off]);
      sb[threadIdx.x] = max_r(sb[threadIdx.x], sb[threadIdx.x + off]);
      snan[threadIdx.x] |= snan[threadIdx.x + off];
    }
    __syncthreads();
  }
  if (threadIdx.x == 0) {
    A.ctl.partial[4 * blockIdx.x + 0] = sa[0];
    A.ctl.partial[4 * blockIdx.x + 1] = sb[0];
    A.ctl.partial[4 * blockIdx.x + 2] = snan[0] ? 1.0 : 0.0;
  }
}

__global__ void k_timestep_final(TsArgs A, int nb) {
  if (threadIdx.x != 0) return;
  const int n = A.G.nx * A.G.ny;
  double LX = 0.0, LY = 0.0;
  int anynan = 0;
  for (int b = 0; b < nb; ++b) {
    LX = max_r(LX, A.ctl.partial[4 * b + 0]);
    LY = max_r(LY, A.ctl.partial[4 * b + 1]);
    anynan |= (A.ctl.partial[4 * b + 2] != 0.0);
  }
  double ts;
  if (!anynan) {
    ts = min_r(min_r(kHuge, A.G.dx / LX), A.G.dy / LY);
  } else {
    // the reference's fold, cell by cell in its loop order
    LX = 0.0;
    LY = 0.0;
    ts = kHuge;
    for (int k = 0; k < n; ++k) {
      double a, b;
      wave_speeds(A.G, A.h[k], A.hu[k], A.hv[k], a, b);
      LX = max_r(LX, a);
      LY = max_r(LY, b);
      ts = min_r(min_r(ts, A.G.dx / LX), A.G.dy / LY);
    }
  }
  ts = A.ctl.cfl * ts;
  if (ts < kMinTimestep) ts = kMinTimestep;
  A.ctl.out[0] = ts;
  A.ctl.out[1] = LX;
  A.ctl.out[2] = LY;
}

// ------------------------------------------------------------------------
// Layout conversion between the reference's U(1:3, ...) and device SoA.
// ghost = 3: the source is U(1:3, -2:nx+3, -2:ny+3) including ghosts.
// ------------------------------------------------------------------------
__global__ void k_aos_to_soa(int nx, int ny, int ghost, const double* aos, double* h, double* hu,
                             double* hv) {
  int k = blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= nx * ny) return;
  int i = k % nx, j = k / nx;
  size_t src = 3 * ((size_t)(j + ghost) * (nx + 2 * ghost) + (i + ghost));
  h[k] = aos[src];
  hu[k] = aos[src + 1];
  hv[k] = aos[src + 2];
}

__global__ void k_soa_to_aos(int nx, int ny, const double* h, const double* hu, const double* hv,
                             double* aos) {
  int k = blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= nx * ny) return;
  aos[3 * (size_t)k] = h[k];
  aos[3 * (size_t)k + 1] = hu[k];
  aos[3 * (size_t)k + 2] = hv[k];
}

__global__ void k_divsqrt(int n, const double* a, const double* b, double* q, double* r) {
  int k = blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= n) return;
  q[k] = a[k] / b[k];
  r[k] = sqrt(a[k]);
}

// ------------------------------------------------------------------------
// launch wrappers
// ------------------------------------------------------------------------
static inline unsigned cdiv(long a, long b) { return (unsigned)((a + b - 1) / b); }

void launch_fvtd_x(const GridGeom& G, const FvtdIO& io, hipStream_t s) {
  dim3 grid(cdiv(G.nx + 1, kFaceTile), cdiv(G.ny, kTile));
  hipLaunchKernelGGL(k_fvtd_x, grid, dim3(256), 0, s, G, io);
}

void launch_fvtd_y(const GridGeom& G, const FvtdIO& io, hipStream_t s) {
  dim3 grid(cdiv(G.nx, kTile), cdiv(G.ny + 1, kFaceTile));
  hipLaunchKernelGGL(k_fvtd_y, grid, dim3(256), 0, s, G, io);
}

void launch_fvtd_update(const GridGeom& G, const FvtdIO& io, const WbArrays& wb, const Bathy& b,
                        bool wb_init, hipStream_t s) {
  dim3 grid(cdiv(G.nx, kTile), cdiv(G.ny, kTile));
  if (b.bx)
    hipLaunchKernelGGL(k_fvtd_update<true>, grid, dim3(256), 0, s, G, io, wb, b, wb_init ? 1 : 0);
  else
    hipLaunchKernelGGL(k_fvtd_update<false>, grid, dim3(256), 0, s, G, io, wb, b,
                       wb_init ? 1 : 0);
}

void launch_momentum(int n, const double* hun, const double* hvn, const double* const FU2[4],
                     const double* const FU3[4], const double cth[4], double* hu, double* hv,
                     hipStream_t s) {
  MomArgs a;
  a.hun = hun;
  a.hvn = hvn;
  for (int r = 0; r < 4; ++r) {
    a.F2[r] = FU2[r];
    a.F3[r] = FU3[r];
    a.c[r] = cth[r];
  }
  a.hu = hu;
  a.hv = hv;
  a.n = n;
  hipLaunchKernelGGL(k_momentum, dim3(cdiv(n, 256)), dim3(256), 0, s, a);
}

int jacobi_nblocks(const GridGeom& G) { return (int)cdiv((long)G.nx * G.ny, 256); }

void launch_patankar_assemble(const GridGeom& G, const PatNodes& nodes, const PatSys* sys,
                              int nsys, const double* hn, const JacCtl& ctl, hipStream_t s) {
  PatArgs A;
  A.G = G;
  A.nodes = nodes;
  for (int i = 0; i < nsys; ++i) A.sys[i] = sys[i];
  A.hn = hn;
  A.ctl = ctl;
  A.ctl.nsys = nsys;
  const int nb = jacobi_nblocks(G);
  if (nsys == 3)
    hipLaunchKernelGGL(k_patankar_assemble<3>, dim3(nb), dim3(256), 0, s, A);
  else if (nsys == 2)
    hipLaunchKernelGGL(k_patankar_assemble<2>, dim3(nb), dim3(256), 0, s, A);
  else
    hipLaunchKernelGGL(k_patankar_assemble<1>, dim3(nb), dim3(256), 0, s, A);
  JacArgs J;
  J.nx = G.nx;
  J.ny = G.ny;
  for (int i = 0; i < nsys; ++i) J.sys[i] = sys[i];
  J.hn = hn;
  J.ctl = A.ctl;
  hipLaunchKernelGGL(k_jacobi_decide<true>, dim3(1), dim3(256), 0, s, J, nb);
}

void launch_jacobi_sweep(const GridGeom& G, const PatSys* sys, int nsys, const double* hn,
                         const JacCtl& ctl, hipStream_t s) {
  JacArgs A;
  A.nx = G.nx;
  A.ny = G.ny;
  for (int i = 0; i < nsys; ++i) A.sys[i] = sys[i];
  A.hn = hn;
  A.ctl = ctl;
  A.ctl.nsys = nsys;
  const int nb = jacobi_nblocks(G);
  if (nsys == 3)
    hipLaunchKernelGGL(k_jacobi_sweep<3>, dim3(nb), dim3(256), 0, s, A);
  else if (nsys == 2)
    hipLaunchKernelGGL(k_jacobi_sweep<2>, dim3(nb), dim3(256), 0, s, A);
  else
    hipLaunchKernelGGL(k_jacobi_sweep<1>, dim3(nb), dim3(256), 0, s, A);
  hipLaunchKernelGGL(k_jacobi_decide<false>, dim3(1), dim3(256), 0, s, A, nb);
}

void launch_jacobi_zero_fill(int n, const PatSys* sys, int nsys, const JacCtl& ctl,
                             hipStream_t s) {
  ZeroArgs a;
  for (int i = 0; i < nsys; ++i) a.x[i] = sys[i].x[1];
  a.st = ctl.st;
  a.nsys = nsys;
  a.n = n;
  hipLaunchKernelGGL(k_jacobi_zero_fill, dim3(cdiv(n, 256)), dim3(256), 0, s, a);
}

int timestep_nblocks(const GridGeom& G) {
  long n = (long)G.nx * G.ny;
  long b = cdiv(n, 256 * 8);
  if (b > 2048) b = 2048;
  if (b < 1) b = 1;
  return (int)b;
}

void launch_timestep(const GridGeom& G, const double* h, const double* hu, const double* hv,
                     const TsCtl& ctl, hipStream_t s) {
  TsArgs A;
  A.G = G;
  A.h = h;
  A.hu = hu;
  A.hv = hv;
  A.ctl = ctl;
  const int nb = timestep_nblocks(G);
  hipLaunchKernelGGL(k_timestep, dim3(nb), dim3(256), 0, s, A);
  hipLaunchKernelGGL(k_timestep_final, dim3(1), dim3(64), 0, s, A, nb);
}

void launch_aos_to_soa(int nx, int ny, int ghost, const double* aos, double* h, double* hu,
                       double* hv, hipStream_t s) {
  hipLaunchKernelGGL(k_aos_to_soa, dim3(cdiv((long)nx * ny, 256)), dim3(256), 0, s, nx, ny,
                     ghost, aos, h, hu, hv);
}

void launch_soa_to_aos(int nx, int ny, const double* h, const double* hu, const double* hv,
                       double* aos, hipStream_t s) {
  hipLaunchKernelGGL(k_soa_to_aos, dim3(cdiv((long)nx * ny, 256)), dim3(256), 0, s, nx, ny, h,
                     hu, hv, aos);
}

void launch_divsqrt(int n, const double* a, const double* b, double* q, double* r,
                    hipStream_t s) {
  hipLaunchKernelGGL(k_divsqrt, dim3(cdiv(n, 256)), dim3(256), 0, s, n, a, b, q, r);
}

}  // namespace swm
