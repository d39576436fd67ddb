// swm_kernels.hip -- FP64 HIP kernels (gfx950) of the WENO5 time steps of
// sw-mpdec -- mPDeC5 and the other modified-Patankar schemes of the PATANKAR
// build, and the explicit schemes of the build without it -- over all active
// jobs of an engine.
//
// One FVTimeDerivative (src/finitevolume.f90:402-440) is three launches, each
// covering every job and up to three DeC nodes (blockIdx.y in the face
// kernels, XCD-grouped block numbers in k_fvtd_update):
//   k_fvtd_x       x faces: halo tile -> ConsToPrim -> WENO5 first sweep along
//                  x (shared through LDS) -> second sweep along y at the 4 Gauss
//                  points -> positivity limiter on h -> Rusanov -> Gauss sum
//   k_fvtd_y       the same for y faces
//   k_fvtd_update  source term (h-only WENO5 point values times -g db) and
//                  UpdateTimeDerivative: Ut(2:3) and face mass fluxes / dx
// The Patankar assembly fuses the first Jacobi sweep and carries the DeC
// momentum update of the same nodes (disjoint arrays).  Every Jacobi sweep
// writes per-block sums of |x_new - x_old|; a one-block-per-job kernel takes
// the stop decision of src/jacobi_iteration.f90:71 on the device.  No grid-wide
// fences are used: on a multi-XCD device an agent-scope fence writes back the
// L2, so the kernel boundary is the cheap way to publish partial sums.
#include <hip/hip_runtime.h>

#include "swm_engine.hpp"
#include "swm_numerics.hpp"
#include "swm_timestep.hpp"

namespace swm {

// thetaDeC: THETA[b][m][r] = thetaDeC(m+1, r+1) of block b
//   b = 0  mPDeC5  (src/timediscretization.f90:491-506)
//   b = 1  mPEuler and mPDeC2, reshape((/0, 1/2, 0, 1/2/), (/2,2/))
//          (src/timediscretization.f90:604,704)
#define SWM_THETA_TABLE                                                                         \
  {{{0.0000000000000000000, 0.0000000000000000000, 0.0000000000000000000, 0.0000000000000000000},  \
    {0.1103005664791649049, 0.1896994335208351257, -0.0339073642291439076, 0.0103005664791649201}, \
    {0.0730327668541684294, 0.4505740308958107176, 0.2269672331458314485, -0.0269672331458315692}, \
    {0.0833333333333333287, 0.4166666666666666852, 0.4166666666666666852, 0.0833333333333333287}}, \
   {{0.0, 0.0, 0.0, 0.0}, {0.5, 0.5, 0.0, 0.0}, {0.0, 0.0, 0.0, 0.0}, {0.0, 0.0, 0.0, 0.0}}}
__constant__ double THETA[2][4][4] = SWM_THETA_TABLE;

// Sign pattern of THETA[b][m][.]: bit r set when thetaDeC(m+1, r+1) > 0.  The
// assembly branches on it, uniformly for a whole launch.
constexpr double kThetaHost[2][4][4] = SWM_THETA_TABLE;
constexpr int theta_pos(int b, int m) {
  int p = 0;
  for (int r = 0; r < 4; ++r)
    if (kThetaHost[b][m][r] > 0.) p |= 1 << r;
  return p;
}
__constant__ int THETA_POS[2][4] = {{theta_pos(0, 0), theta_pos(0, 1), theta_pos(0, 2), theta_pos(0, 3)},
                                    {theta_pos(1, 0), theta_pos(1, 1), theta_pos(1, 2), theta_pos(1, 3)}};
static_assert(theta_pos(0, 1) == 0xB && theta_pos(0, 2) == 0x7 && theta_pos(0, 3) == 0xF &&
                  theta_pos(1, 1) == 0x3,
              "thetaDeC sign pattern");

// Global-address-space view of a device array.  The arrays' pointers are read
// from the job table in memory, so the compiler cannot infer their address
// space and would use flat loads and stores, which also occupy the LDS wait
// counter; through G() they become global loads and stores.
typedef __attribute__((address_space(1))) double gdouble;
__device__ __forceinline__ const gdouble* G(const double* p) { return (const gdouble*)p; }
__device__ __forceinline__ gdouble* G(double* p) { return (gdouble*)p; }

// 0-based index of Fortran cell c (1..n or a periodic ghost) on an axis of n.
// The tiles' halos keep c within [-2, n + 17], so one add or subtract of n
// suffices for n >= 17; smaller axes take the modulo.
__device__ __forceinline__ int wrap1(int c, int n) {
  int m = c - 1;
  if (m < 0)
    m += n;
  else if (m >= n)
    m -= n;
  if ((unsigned)m >= (unsigned)n) {
    m = (c - 1) % n;
    if (m < 0) m += n;
  }
  return m;
}

// Row-major cell k = k0 + t of a grid nx wide, with (i0, j0) the cell of k0
// (computed once per block): i = k % nx, j = k / nx without a per-thread
// integer division.
__device__ __forceinline__ void cell_ij(int i0, int j0, int t, int nx, int& i, int& j) {
  i = i0 + t;
  j = j0;
  while (i >= nx) {
    i -= nx;
    ++j;
  }
}

__device__ __forceinline__ int cdiv_d(int a, int b) { return (a + b - 1) / b; }

// In the steps of a batch (E.gated) a job whose batch has stopped
// (JobStatus::halt) is skipped by every kernel; the test is uniform per block.
__device__ __forceinline__ bool skipped(const EngineArgs& E, const JobDev& J) {
  return E.gated && J.status->halt != 0;
}

// State slot m of a job: m = 0 is U^n, 1..3 the DeC nodes.
__device__ __forceinline__ void node_state(const JobDev& J, int m, const gdouble*& h,
                                           const gdouble*& hu, const gdouble*& hv) {
  h = G(J.sh[m]);
  hu = G(J.shu[m]);
  hv = G(J.shv[m]);
}

// ------------------------------------------------------------------------
// Face kernels, templated on the tile shape TW x TH (threads in x, y).
//
// x faces.  Tile (bx, by) of a job: Fortran cells c = (TW-1)*bx + tx
// (tx = 0..TW-1, c in 0..nx+1; c = 0 is the ghost of nx, c = nx+1 the ghost
// of 1), rows j = TH*by + 1 + ty.  Thread (tx, ty) reconstructs cell (c, j)
// and, for tx < TW-1, computes face f = c between cells c and c+1
// (f = 0..nx).  The positivity limiter runs for c = 0..nx only
// (src/finitevolume.f90:670), so face f = nx uses the unlimited left value
// of cell nx+1 while f = 0 uses the limited one of cell 1: the two values of
// the periodic seam.  Every face is computed by exactly one thread with the
// same arithmetic whatever the shape, so the results do not depend on it.
// ------------------------------------------------------------------------
template <int TW, int TH>
struct FaceTile {
  static constexpr int NT = TW * TH;               // threads
  static constexpr int VW = TW + 4, VH = TH + 4;   // halo tile of primitives
  static constexpr int NV = VW * VH;
  // first-sweep results per variable: VH x TW (x faces) or TH x VW (y faces)
  static constexpr int NQ = (TW * VH > TH * VW) ? TW * VH : TH * VW;
  // doubles of LDS: primitives + both first sweeps, reused for the WM exchange
  static constexpr int LDS = (3 * NV + 6 * NQ > 12 * NT) ? 3 * NV + 6 * NQ : 12 * NT;
};

template <int TW, int TH>
__global__ __launch_bounds__(TW * TH) void k_fvtd_x(EngineArgs E, const TileRef* tiles,
                                                    NodeList L) {
  using FT = FaceTile<TW, TH>;
  constexpr int NT = FT::NT, VW = FT::VW, VH = FT::VH, NV = FT::NV;
  constexpr int NQ = VH * TW;  // first sweeps per variable: VH rows x TW cells
  __shared__ double smem[FT::LDS];
  double* sV = smem;            // [3][VH rows][VW cols]
  double* sQM = smem + 3 * NV;  // [3][VH rows][TW cols]
  double* sQP = sQM + 3 * NQ;
  const TileRef T = tiles[blockIdx.x];
  const JobDev& J = E.jobs[T.job];
  if (skipped(E, J)) return;
  const int m = L.m[blockIdx.y], slot = L.slot[blockIdx.y];
  const int nx = J.nx, ny = J.ny;
  const int ntx = cdiv_d(nx + 1, TW - 1);
  const int tid = threadIdx.x, tx = tid % TW, ty = tid / TW;
  const int c0 = (T.tile % ntx) * (TW - 1);
  const int j0 = (T.tile / ntx) * TH + 1;
  const gdouble *h, *hu, *hv;
  node_state(J, m, h, hu, hv);

  for (int t = tid; t < NV; t += NT) {
    int lx = t % VW, ly = t / VW;
    size_t k = (size_t)wrap1(j0 - 2 + ly, ny) * nx + wrap1(c0 - 2 + lx, nx);
    double hh = h[k], u, v;
    cons_to_prim(hh, hu[k], hv[k], u, v);
    sV[t] = hh;
    sV[NV + t] = u;
    sV[2 * NV + t] = v;
  }
  __syncthreads();
  for (int t = tid; t < 3 * NQ; t += NT) {
    int var = t / NQ, r = t % NQ, ly = r / TW, lx = r % TW;
    const double* q = sV + var * NV + ly * VW + lx;
    double wm, wp;
    weno5_first(q[0], q[1], q[2], q[3], q[4], wm, wp);
    sQM[var * NQ + r] = wm;
    sQP[var * NQ + r] = wp;
  }
  __syncthreads();
  double WM[3][4], WP[3][4];
#pragma unroll
  for (int var = 0; var < 3; ++var) {
    const double* qm = sQM + var * NQ + ty * TW + tx;
    weno5_second(qm[0], qm[TW], qm[2 * TW], qm[3 * TW], qm[4 * TW], WM[var]);
    const double* qp = sQP + var * NQ + ty * TW + tx;
    weno5_second(qp[0], qp[TW], qp[2 * TW], qp[3 * TW], qp[4 * TW], WP[var]);
  }
  const int c = c0 + tx;
  if (J.recfix) {
    // ReconstructionFix 1 or 2 on the cells the shock indicator flags.
    // ShocksIndicatorX writes each cell's verdict to Ind(1, ii-1:ii+1, jj) in
    // loop order (src/shocksindicator.f90:50-55), so the last writer wins:
    // Ind(1, c, j) = Shock(min(c + 1, nx), j), from the depths at columns
    // min(c + 1, nx) - 1 .. + 1, all inside the halo tile.
    const int lc = min(c + 1, nx) - c0 + 2;
    const double* hrow = sV + (ty + 2) * VW;
    if (jameson_shock(hrow[lc - 1], hrow[lc], hrow[lc + 1])) {
#pragma unroll
      for (int var = 0; var < 3; ++var) {
        const double* q = sV + var * NV + (ty + 2) * VW + tx + 2;
        reconstruction_fix(J.recfix, q[-1], q[0], q[1], J.dx, WM[var], WP[var]);
      }
    }
  }
  const double vh = sV[(ty + 2) * VW + tx + 2];
  if (c <= nx) positivity_limit(vh, J.wbnd, WM[0], WP[0]);
  __syncthreads();
  double* sWM = smem;  // [12][NT]
#pragma unroll
  for (int a = 0; a < 12; ++a) sWM[a * NT + tid] = WM[a >> 2][a & 3];
  __syncthreads();
  const int j = j0 + ty;
  if (tx < TW - 1 && c <= nx && j <= ny) {
    double acc1 = 0.0, acc2 = 0.0, acc3 = 0.0;
#pragma unroll
    for (int g = 0; g < 4; ++g) {
      double F1, F2, F3;
      rusanov<1, 0, 0, 1>(WP[0][g], WP[1][g], WP[2][g], sWM[(0 + g) * NT + tid + 1],
                          sWM[(4 + g) * NT + tid + 1], sWM[(8 + g) * NT + tid + 1], J.g, F1,
                          F2, F3);
      acc1 = acc1 + J.wbnd[g] * F1;
      acc2 = acc2 + J.wbnd[g] * F2;
      acc3 = acc3 + J.wbnd[g] * F3;
    }
    size_t o = (size_t)(j - 1) * (nx + 1) + c;
    if (J.wb && !L.wb_init) {  // FX = FX - FXWB (src/finitevolume.f90:485-486)
      acc1 = acc1 - G(J.FXWB[0])[o];
      acc2 = acc2 - G(J.FXWB[1])[o];
      acc3 = acc3 - G(J.FXWB[2])[o];
    }
    G(J.FX[slot][0])[o] = acc1;
    G(J.FX[slot][1])[o] = acc2;
    G(J.FX[slot][2])[o] = acc3;
  }
}

// ------------------------------------------------------------------------
// y faces: the transpose of k_fvtd_x.  Tile (bx, by): columns
// i = TW*bx + 1 + tx, Fortran rows r = (TH-1)*by + ty (0..ny+1); face g = r.
// ------------------------------------------------------------------------
template <int TW, int TH>
__global__ __launch_bounds__(TW * TH) void k_fvtd_y(EngineArgs E, const TileRef* tiles,
                                                    NodeList L) {
  using FT = FaceTile<TW, TH>;
  constexpr int NT = FT::NT, VW = FT::VW, NV = FT::NV;
  constexpr int NQ = TH * VW;  // first sweeps per variable: TH rows x VW columns
  __shared__ double smem[FT::LDS];
  double* sV = smem;            // [3][VH rows][VW cols]
  double* sQM = smem + 3 * NV;  // [3][TH rows][VW cols]
  double* sQP = sQM + 3 * NQ;
  const TileRef T = tiles[blockIdx.x];
  const JobDev& J = E.jobs[T.job];
  if (skipped(E, J)) return;
  const int m = L.m[blockIdx.y], slot = L.slot[blockIdx.y];
  const int nx = J.nx, ny = J.ny;
  const int ntx = cdiv_d(nx, TW);
  const int tid = threadIdx.x, tx = tid % TW, ty = tid / TW;
  const int i0 = (T.tile % ntx) * TW + 1;
  const int r0 = (T.tile / ntx) * (TH - 1);
  const gdouble *h, *hu, *hv;
  node_state(J, m, h, hu, hv);

  for (int t = tid; t < NV; t += NT) {
    int lx = t % VW, ly = t / VW;
    size_t k = (size_t)wrap1(r0 - 2 + ly, ny) * nx + wrap1(i0 - 2 + lx, nx);
    double hh = h[k], u, v;
    cons_to_prim(hh, hu[k], hv[k], u, v);
    sV[t] = hh;
    sV[NV + t] = u;
    sV[2 * NV + t] = v;
  }
  __syncthreads();
  for (int t = tid; t < 3 * NQ; t += NT) {
    int var = t / NQ, r = t % NQ, ly = r / VW, lx = r % VW;
    const double* q = sV + var * NV + ly * VW + lx;
    double wm, wp;
    weno5_first(q[0], q[VW], q[2 * VW], q[3 * VW], q[4 * VW], wm, wp);
    sQM[var * NQ + r] = wm;
    sQP[var * NQ + r] = wp;
  }
  __syncthreads();
  double WM[3][4], WP[3][4];
#pragma unroll
  for (int var = 0; var < 3; ++var) {
    const double* qm = sQM + var * NQ + ty * VW + tx;
    weno5_second(qm[0], qm[1], qm[2], qm[3], qm[4], WM[var]);
    const double* qp = sQP + var * NQ + ty * VW + tx;
    weno5_second(qp[0], qp[1], qp[2], qp[3], qp[4], WP[var]);
  }
  const int r = r0 + ty;
  if (J.recfix) {
    // Ind(2, i, r) = Shock(i, min(r + 1, ny)) (src/shocksindicator.f90:84-89),
    // as in k_fvtd_x
    const int lr = min(r + 1, ny) - r0 + 2;
    const double* hcol = sV + tx + 2;
    if (jameson_shock(hcol[(lr - 1) * VW], hcol[lr * VW], hcol[(lr + 1) * VW])) {
#pragma unroll
      for (int var = 0; var < 3; ++var) {
        const double* q = sV + var * NV + (ty + 2) * VW + tx + 2;
        reconstruction_fix(J.recfix, q[-VW], q[0], q[VW], J.dy, WM[var], WP[var]);
      }
    }
  }
  const double vh = sV[(ty + 2) * VW + tx + 2];
  if (r <= ny) positivity_limit(vh, J.wbnd, WM[0], WP[0]);
  __syncthreads();
  double* sWM = smem;  // [12][NT]
#pragma unroll
  for (int a = 0; a < 12; ++a) sWM[a * NT + tid] = WM[a >> 2][a & 3];
  __syncthreads();
  const int i = i0 + tx;
  if (ty < TH - 1 && r <= ny && i <= nx) {
    double acc1 = 0.0, acc2 = 0.0, acc3 = 0.0;
#pragma unroll
    for (int g = 0; g < 4; ++g) {
      double F1, F2, F3;
      rusanov<0, 1, -1, 0>(WP[0][g], WP[1][g], WP[2][g], sWM[(0 + g) * NT + tid + TW],
                           sWM[(4 + g) * NT + tid + TW], sWM[(8 + g) * NT + tid + TW], J.g,
                           F1, F2, F3);
      acc1 = acc1 + J.wbnd[g] * F1;
      acc2 = acc2 + J.wbnd[g] * F2;
      acc3 = acc3 + J.wbnd[g] * F3;
    }
    size_t o = (size_t)r * nx + (i - 1);
    if (J.wb && !L.wb_init) {  // FY = FY - FYWB (src/finitevolume.f90:487-488)
      acc1 = acc1 - G(J.FYWB[0])[o];
      acc2 = acc2 - G(J.FYWB[1])[o];
      acc3 = acc3 - G(J.FYWB[2])[o];
    }
    G(J.FY[slot][0])[o] = acc1;
    G(J.FY[slot][1])[o] = acc2;
    G(J.FY[slot][2])[o] = acc3;
  }
}

// ------------------------------------------------------------------------
// Source term + UpdateTimeDerivative.  Tile: 16x16 cells (1-based
// i = 16*bx + 1 + tx, j = 16*by + 1 + ty).  SourceTerms
// (src/equation.f90:352-388): h point values at the 4x4 Gauss points from a
// WENO5 second sweep along y (columns i-2..i+2) followed by one along x per
// y Gauss point; SW = -g*h*db; S = sum over x-GP (outer), y-GP (inner) of W*SW.
// ------------------------------------------------------------------------
// Blocks are numbered so that the L.count nodes of one tile share an XCD:
// workgroups are dealt round-robin to the 8 XCDs, so block
// ((t / 8) * count + node) * 8 + t % 8 runs tile t on XCD t % 8 for every
// node, close together in time, and the tile's bathymetry and well-balanced
// tables are read from HBM once and from that XCD's L2 for the other nodes.
__global__ __launch_bounds__(256) void k_fvtd_update(EngineArgs E, const TileRef* tiles,
                                                     int ntiles, NodeList L) {
  __shared__ double sH[400];           // [20 rows][20 cols]
  __shared__ double sVt[4 * 16 * 20];  // [yg][16 rows][20 cols]
  const int grp = blockIdx.x / (8 * L.count), rem = blockIdx.x % (8 * L.count);
  const int node = rem / 8, tile = grp * 8 + (rem % 8);
  if (tile >= ntiles) return;
  const TileRef T = tiles[tile];
  const JobDev& J = E.jobs[T.job];
  if (skipped(E, J)) return;
  const int nx = J.nx, ny = J.ny;
  const int ntx = cdiv_d(nx, kTile);
  const int tid = threadIdx.x, tx = tid & 15, ty = tid >> 4;
  const int i0 = (T.tile % ntx) * kTile + 1, j0 = (T.tile / ntx) * kTile + 1;
  const int i = i0 + tx, j = j0 + ty;
  const bool in = i <= nx && j <= ny;
  const size_t n = (size_t)J.n;
  const size_t k = in ? (size_t)(j - 1) * nx + (i - 1) : 0;
  {
    const int m = L.m[node], slot = L.slot[node];
    const gdouble *h, *hu, *hv;
    node_state(J, m, h, hu, hv);
    double S2 = 0.0, S3 = 0.0;
    if (J.bathy) {  // uniform per block
      for (int t = tid; t < 400; t += 256) {
        int lx = t % 20, ly = t / 20;
        sH[t] = h[(size_t)wrap1(j0 - 2 + ly, ny) * nx + wrap1(i0 - 2 + lx, nx)];
      }
      __syncthreads();
      for (int t = tid; t < 320; t += 256) {
        int ly = t / 20, lx = t % 20;
        const double* qh = sH + ly * 20 + lx;
        double w[4];
        weno5_second(qh[0], qh[20], qh[40], qh[60], qh[80], w);
#pragma unroll
        for (int yg = 0; yg < 4; ++yg) sVt[yg * 320 + t] = w[yg];
      }
      __syncthreads();
      if (in) {
        double hgp[4][4];  // [yg][xg]
#pragma unroll
        for (int yg = 0; yg < 4; ++yg) {
          const double* qv = sVt + yg * 320 + ty * 20 + tx;
          weno5_second(qv[0], qv[1], qv[2], qv[3], qv[4], hgp[yg]);
        }
#pragma unroll
        for (int xg = 0; xg < 4; ++xg)
#pragma unroll
          for (int yg = 0; yg < 4; ++yg) {
            double hh = hgp[yg][xg];
            double bx = G(J.bx)[(size_t)(yg * 4 + xg) * n + k];
            double by = G(J.by)[(size_t)(yg * 4 + xg) * n + k];
            double sw2 = -J.g * hh * bx;
            double sw3 = -J.g * hh * by;
            S2 = S2 + J.wgp[xg][yg] * sw2;
            S3 = S3 + J.wgp[xg][yg] * sw3;
          }
      }
    }
    if (!in) return;
    if (L.wb_init) {  // building the well-balanced reference: keep S as SWB
      G(J.SWB2)[k] = S2;
      G(J.SWB3)[k] = S3;
      return;
    }
    const size_t fe = (size_t)(j - 1) * (nx + 1) + i, fw = fe - 1;  // x faces i, i-1
    const size_t gn = (size_t)j * nx + (i - 1), gs = gn - nx;        // y faces j, j-1
    if (J.wb) {
      S2 = S2 - G(J.SWB2)[k];
      S3 = S3 - G(J.SWB3)[k];
    }
    if (L.debug_S) {  // parity tests read S(2:3) from the first matrix scratch
      G(J.M[0][0])[k] = S2;
      G(J.M[0][1])[k] = S3;
    }
    // the face fluxes, with FXWB / FYWB already subtracted by the face kernels
    double fxe[3], fxw[3], fyn[3], fys[3];
#pragma unroll
    for (int v = 0; v < 3; ++v) {
      fxe[v] = G(J.FX[slot][v])[fe];
      fxw[v] = G(J.FX[slot][v])[fw];
      fyn[v] = G(J.FY[slot][v])[gn];
      fys[v] = G(J.FY[slot][v])[gs];
    }
    // the divisions by dx and dy share one reciprocal each (uniform per job)
    const double dx = J.dx, dy = J.dy;
    const bool shared = recip_range(dx) & recip_range(dy);
    const double rdx = recip_refined(dx), rdy = recip_refined(dy);
    auto over_dx = [&](double a) { return shared ? div_pos(a, dx, rdx) : a / dx; };
    auto over_dy = [&](double a) { return shared ? div_pos(a, dy, rdy) : a / dy; };
    // Ut(2:3) = S - (FX(i)-FX(i-1))/dx - (FY(j)-FY(j-1))/dy  (:525-527)
    const int o = L.out[node];
    const double ut2 = S2 - over_dx(fxe[1] - fxw[1]) - over_dy(fyn[1] - fys[1]);
    const double ut3 = S3 - over_dx(fxe[2] - fxw[2]) - over_dy(fyn[2] - fys[2]);
    if (L.ut1) {
      // the build without PATANKAR: Ut(1:3) (:529-533), with S(1) = 0 exactly
      // (SourceFunc's first component is 0., src/equation.f90:408)
      G(J.d1[o])[k] = 0.0 - over_dx(fxe[0] - fxw[0]) - over_dy(fyn[0] - fys[0]);
      G(J.d2[o])[k] = ut2;
      G(J.d3[o])[k] = ut3;
      return;
    }
    G(J.FU2[o])[k] = ut2;
    G(J.FU3[o])[k] = ut3;
    // mass flux per face / dx: the production/destruction entries of
    // src/finitevolume.f90:496-523 are max(+-q, 0) and -min(+-q, 0) of these.
    G(J.qx[o])[fe] = over_dx(fxe[0]);
    G(J.qy[o])[gn] = over_dy(fyn[0]);
    if (i == 1) G(J.qx[o])[fw] = over_dx(fxw[0]);
    if (j == 1) G(J.qy[o])[gs] = over_dy(fys[0]);
  }
}

// ------------------------------------------------------------------------
// DeC momentum update, src/timediscretization.f90:531-543 (and :617-621,
// :717-721): Ua(m,2:3) = Ua(1,2:3); DO r=1,R: Ua(m) = Ua(m) + (dt*theta(m,r))*FUp(r)
// with R = MSteps.
// ------------------------------------------------------------------------
// One cell, every system's node m.  It runs inside the assembly kernel (the
// two read and write disjoint arrays), so that its loads overlap the
// assembly's latency-bound work.
template <int R>
__device__ __forceinline__ void momentum_cell(const JobDev& J, const SysList& S, int k,
                                              double dt) {
  double fu2[R], fu3[R];
#pragma unroll
  for (int r = 0; r < R; ++r) {
    fu2[r] = G(J.FU2[S.src[r]])[k];
    fu3[r] = G(J.FU3[S.src[r]])[k];
  }
  const double u0 = G(J.hu)[k], v0 = G(J.hv)[k];
  for (int s = 0; s < S.count; ++s) {
    const int m = S.m[s];
    double u = u0, v = v0;
#pragma unroll
    for (int r = 0; r < R; ++r) {
      const double c = dt * THETA[S.theta][m][r];
      u = u + c * fu2[r];
      v = v + c * fu3[r];
    }
    G(J.hun[m])[k] = u;
    G(J.hvn[m])[k] = v;
  }
}

// ------------------------------------------------------------------------
// Reductions
// ------------------------------------------------------------------------
__device__ __forceinline__ double wave_sum(double v) {
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, 64);
  return v;
}

// Deterministic sum over a 256-thread block, returned to every thread.
// red: 4 doubles of shared memory.
__device__ __forceinline__ double block_sum_all(double v, double* red) {
  v = wave_sum(v);
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  __syncthreads();
  if (lane == 0) red[w] = v;
  __syncthreads();
  double r = ((red[0] + red[1]) + red[2]) + red[3];
  __syncthreads();
  return r;
}

// Sum of part[0..nb) in a fixed order by one 256-thread block, with 8
// independent loads in flight per thread.
__device__ double block_sum_array(const double* part, int nb, double* red) {
  double a[8] = {0., 0., 0., 0., 0., 0., 0., 0.};
  int b = threadIdx.x;
  for (; b + 7 * 256 < nb; b += 8 * 256) {
#pragma unroll
    for (int u = 0; u < 8; ++u) a[u] += part[b + u * 256];
  }
  for (int u = 0; b < nb; b += 256, ++u) a[u & 7] += part[b];
  double t = ((a[0] + a[1]) + (a[2] + a[3])) + ((a[4] + a[5]) + (a[6] + a[7]));
  return block_sum_all(t, red);
}

// Decide `tol < SUM(d)/n` (src/jacobi_iteration.f90:71,77) where the
// reference sums the d_i >= 0 sequentially in row order and S is the same sum
// taken in parallel.  Both roundings differ from the exact sum by at most
// (n-1)u*sum(d), so whenever tol lies outside [S - B, S + B]/n with
// B = 8nuS the decision is that of the sequential sum; inside the window the
// caller recomputes the sequential sum exactly.  1 continue, 0 stop, -1 open.
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

// Exact replay of the reference's sequential residual: mode 0 sums
// |0 - a_i|, mode 1 sums |a_i - b_i|; divided by n.
__device__ double sequential_residual(int n, const double* a, const double* b, int mode) {
  double r = 0.0;
  for (int i = 0; i < n; ++i) r = r + (mode == 0 ? fabs(0.0 - a[i]) : fabs(a[i] - b[i]));
  return r / (double)n;
}

// ------------------------------------------------------------------------
// Patankar assembly (MODIFIED_PATANKAR_MATRIX_INVERSION,
// src/timediscretization.f90:795-887) fused with the first Jacobi sweep from
// x^0 = rhs (src/jacobi_iteration.f90:67-76).
// ------------------------------------------------------------------------
// PATANKAR_DIV of a zero numerator, without the division: the reference's
// 2*h*z/(h*h + MAX(h*h, 1e-6)) for z = +0 or -0 is +0 when h < MIN_DEPTH, z
// itself when h is finite (the quotient keeps the sign of z), and NaN when h
// is +Inf or NaN (2*h*z is then NaN) -- exactly what this returns.
__device__ __forceinline__ double patankar_div_zero(double z, double h) {
  return (h < kMinDepth) ? 0. : z * (h * 0. + 1.);
}

// The parts of PATANKAR_DIV (src/timediscretization.f90:913-924) that depend
// on the depth h alone: 2*h*prod/(h*h + MAX(h*h, MIN_DEPTH)) is (2h)*prod / den
// evaluated left to right, 0 when h < MIN_DEPTH; and its value for a +0 and a
// -0 numerator.
struct PdDepth {
  double twoh, den, zp, zm;
  bool dry;
};
__device__ __forceinline__ PdDepth pd_depth(double h) {
  PdDepth d;
  d.dry = h < kMinDepth;
  d.twoh = 2. * h;
  d.den = h * h + max_r(h * h, kMinDepth);
  d.zp = patankar_div_zero(0., h);
  d.zm = patankar_div_zero(-0., h);
  return d;
}

// PATANKAR_DIV(a, h) with h the neighbour's depth (useL) or the row's own.
// WET: no depth of the row is below MIN_DEPTH, so no quotient becomes 0.
template <bool WET>
__device__ __forceinline__ double pd_quot(bool useL, double a, const PdDepth& L,
                                          const PdDepth& K) {
  const double q = ((useL ? L.twoh : K.twoh) * a) / (useL ? L.den : K.den);
  return (!WET && (useL ? L.dry : K.dry)) ? 0. : q;
}

// One stencil slot of a row at the R DeC nodes (src/timediscretization.f90:
// 800-884): with P and De the slot's production and destruction,
//   theta > 0: M = M - c*PD(P, hL), D = D + c*PD(De, hK);
//   otherwise: D = D - c*PD(P, hK), M = M + c*PD(De, hL).
// s is the slot's oriented face mass flux: q = F/dx for the west and south
// slots, -q for the east and north ones, so that P = s > 0 ? s : +0 and
// De = s < 0 ? -s : -0 (src/finitevolume.f90:496-523).  At most one of P and
// De is nonzero, and the nonzero one is |s|, so each node needs one division;
// the zero one takes PdDepth's zp or zm, the value PATANKAR_DIV gives it.
// The sign of theta (bit r of pos) is the same for the whole launch, so the
// branch on it is uniform.
template <int R, bool WET>
__device__ __forceinline__ void patankar_slot(int pos, const double c[R], const double s[R],
                                              const PdDepth& L, const PdDepth& K, double& M,
                                              double& D) {
  M = 0.;
#pragma unroll
  for (int r = 0; r < R; ++r) {
    const double sr = s[r];
    const bool gt = sr > 0., lt = sr < 0.;
    if ((pos >> r) & 1) {  // L takes P, K takes De
      const double dv = pd_quot<WET>(gt, fabs(sr), L, K);
      const double tL = c[r] * (gt ? dv : L.zp);
      const double tK = c[r] * (lt ? dv : K.zm);
      M = M - tL;
      D = D + tK;
    } else {  // L takes De, K takes P
      const double dv = pd_quot<WET>(lt, fabs(sr), L, K);
      const double tL = c[r] * (lt ? dv : L.zm);
      const double tK = c[r] * (gt ? dv : K.zp);
      M = M + tL;
      D = D - tK;
    }
  }
}

// The four stencil slots of one row-system (sparse_matmul in CSR order S, W,
// self (value 0), E, N): writes the slots' matrix entries, accumulates the
// diagonal into D and returns the product with the neighbours' rhs values.
template <int R, bool WET>
__device__ __forceinline__ double patankar_row(const JobDev& J, int s, size_t k, const double* sq,
                                               int pos, const double c[R], const PdDepth& pK,
                                               double hS, double hW, double hE, double hN,
                                               double bS, double bW, double bE, double bN,
                                               double b, double& D) {
  double p = 0.;
#pragma unroll 1
  for (int sl = 0; sl < 4; ++sl) {
    // the slot's neighbour values, selected from the ones loaded by the caller
    const double hL = sl == 0 ? hS : sl == 1 ? hW : sl == 2 ? hE : hN;
    const double bL = sl == 0 ? bS : sl == 1 ? bW : sl == 2 ? bE : bN;
    double q[R];
#pragma unroll
    for (int r = 0; r < R; ++r) q[r] = sq[(sl * R + r) * 256];
    double M;
    patankar_slot<R, WET>(pos, c, q, pd_depth(hL), pK, M, D);
    G(J.M[s][sl])[k] = M;
    p = p + M * bL;
    if (sl == 1) p = p + 0. * b;
  }
  return p;
}

// The denominators and right-hand side of the systems: hp(m,.) is h^n in the
// first correction and node m's depth after it; the rhs is h^n.  With
// S.hp_view both come from the buffers k_hp_view fills.
__device__ __forceinline__ const gdouble* system_hp(const JobDev& J, const SysList& S, int m) {
  return G(S.hp_view ? J.hres[3] : (S.hp_is_un ? J.h : J.hres[m]));
}
__device__ __forceinline__ const gdouble* system_rhs(const JobDev& J, const SysList& S) {
  return G(S.hp_view ? J.hres[2] : J.h);
}

template <int R>
__global__ __launch_bounds__(256) void k_assemble(EngineArgs E, const TileRef* blocks,
                                                  SysList S) {
  __shared__ double red[4];
  __shared__ double sq[4 * R * 256];  // [slot*R + node][thread]: face mass fluxes
  const TileRef T = blocks[blockIdx.x];
  const JobDev& J = E.jobs[T.job];
  if (skipped(E, J)) return;
  const int nx = J.nx, ny = J.ny, n = J.n;
  const int tid = threadIdx.x;
  const int k = T.tile * kCellBlock + tid;
  const bool in = k < n;
  const double dt = J.status->dt;
  int kS = 0, kW = 0, kE = 0, kN = 0;  // neighbours
  double b = 0., bS = 0., bW = 0., bE = 0., bN = 0.;
  if (in) {
    const int k0 = T.tile * kCellBlock;
    int i, j;
    cell_ij(k0 % nx, k0 / nx, tid, nx, i, j);
    kS = (j == 0 ? ny - 1 : j - 1) * nx + i;
    kW = j * nx + (i == 0 ? nx - 1 : i - 1);
    kE = j * nx + (i == nx - 1 ? 0 : i + 1);
    kN = (j == ny - 1 ? 0 : j + 1) * nx + i;
    const size_t fW = (size_t)j * (nx + 1) + i, fE = fW + 1;
    const size_t gS = (size_t)j * nx + i, gN = gS + nx;
#pragma unroll
    for (int r = 0; r < R; ++r) {
      sq[(0 * R + r) * 256 + tid] = G(J.qy[S.src[r]])[gS];
      sq[(1 * R + r) * 256 + tid] = G(J.qx[S.src[r]])[fW];
      sq[(2 * R + r) * 256 + tid] = -G(J.qx[S.src[r]])[fE];  // east and north slots:
      sq[(3 * R + r) * 256 + tid] = -G(J.qy[S.src[r]])[gN];  // the flux oriented into the cell
    }
    const gdouble* rhs = system_rhs(J, S);
    b = rhs[k];
    bS = rhs[kS];
    bW = rhs[kW];
    bE = rhs[kE];
    bN = rhs[kN];
    momentum_cell<R>(J, S, k, dt);
  }
  // One system at a time, and within it one stencil slot at a time (S, W, E,
  // N: the CSR order), so that only one slot's divisions are live.  Each
  // thread reads back only its own column of sq, so no barrier is needed.
#pragma unroll 1
  for (int s = 0; s < S.count; ++s) {
    double d = 0.;
    if (in) {
      const int m = S.m[s];
      const gdouble* hp = system_hp(J, S, m);
      const double hK = hp[k], hS = hp[kS], hW = hp[kW], hE = hp[kE], hN = hp[kN];
      const PdDepth pK = pd_depth(hK);
      const int pos = THETA_POS[S.theta][m];
      double c[R];
#pragma unroll
      for (int r = 0; r < R; ++r) c[r] = dt * THETA[S.theta][m][r];
      // sol = (rhs - prod)/diag; rows of a wave with no depth below
      // MIN_DEPTH take the branch without the dry-depth selects
      double D = 1.;
      const bool dry = pK.dry | (hS < kMinDepth) | (hW < kMinDepth) | (hE < kMinDepth) |
                       (hN < kMinDepth);
      const double p =
          __any(dry) ? patankar_row<R, false>(J, s, k, sq + tid, pos, c, pK, hS, hW, hE, hN, bS,
                                              bW, bE, bN, b, D)
                     : patankar_row<R, true>(J, s, k, sq + tid, pos, c, pK, hS, hW, hE, hN, bS,
                                             bW, bE, bN, b, D);
      G(J.M[s][4])[k] = D;
      double x1 = (b - p) / D;
      G(J.xj[m][1])[k] = x1;
      d = fabs(x1 - b);
    }
    double t = block_sum_all(d, red);
    if (tid == 0) J.part[(size_t)s * J.nblk + T.tile] = t;
  }
  double t = block_sum_all(in ? fabs(0.0 - b) : 0., red);
  if (tid == 0) J.part[(size_t)kMaxSys * J.nblk + T.tile] = t;
}

// One Jacobi sweep of every still-active system (src/jacobi_iteration.f90:72-80).
__global__ __launch_bounds__(256) void k_sweep(EngineArgs E, const TileRef* blocks, SysList S) {
  __shared__ double red[4];
  const TileRef T = blocks[blockIdx.x];
  const JobDev& J = E.jobs[T.job];
  if (skipped(E, J)) return;
  int act[kMaxSys] = {0, 0, 0}, cur[kMaxSys] = {0, 0, 0};
  bool any = false;
  for (int s = 0; s < S.count; ++s) {
    act[s] = J.status->st[s].active;
    cur[s] = J.status->st[s].cur;
    any = any || act[s];
  }
  if (!any) return;  // every system of this job has stopped
  const int nx = J.nx, ny = J.ny, n = J.n;
  const int k = T.tile * kCellBlock + threadIdx.x;
  double d[kMaxSys] = {0., 0., 0.};
  if (k < n) {
    const int k0 = T.tile * kCellBlock;
    int i, j;
    cell_ij(k0 % nx, k0 / nx, threadIdx.x, nx, i, j);
    const int kS = (j == 0 ? ny - 1 : j - 1) * nx + i;
    const int kN = (j == ny - 1 ? 0 : j + 1) * nx + i;
    const int kW = j * nx + (i == 0 ? nx - 1 : i - 1);
    const int kE = j * nx + (i == nx - 1 ? 0 : i + 1);
    const double b = system_rhs(J, S)[k];
    for (int s = 0; s < S.count; ++s) {
      if (!act[s]) continue;
      const int m = S.m[s];
      const gdouble* xo = G(J.xj[m][cur[s]]);
      gdouble* xn = G(J.xj[m][1 - cur[s]]);
      double p = 0.;
      p = p + G(J.M[s][0])[k] * xo[kS];
      p = p + G(J.M[s][1])[k] * xo[kW];
      p = p + 0. * xo[k];
      p = p + G(J.M[s][2])[k] * xo[kE];
      p = p + G(J.M[s][3])[k] * xo[kN];
      double xv = (b - p) / G(J.M[s][4])[k];
      xn[k] = xv;
      d[s] = fabs(xv - xo[k]);
    }
  }
  for (int s = 0; s < S.count; ++s) {
    if (!act[s]) continue;
    double t = block_sum_all(d[s], red);
    if (threadIdx.x == 0) J.part[(size_t)s * J.nblk + T.tile] = t;
  }
}

// Stop decision, one block per job: after the assembly (FIRST) or a sweep.
template <bool FIRST>
__global__ __launch_bounds__(256) void k_decide(EngineArgs E, const int* jobs, SysList S) {
  __shared__ double red[4];
  const JobDev& J = E.jobs[jobs[blockIdx.x]];
  JobStatus* st = J.status;
  if (skipped(E, J)) return;
  const int n = J.n, nb = J.nblk;
  double tot[kMaxSys + 1] = {0., 0., 0., 0.};
  int act[kMaxSys] = {0, 0, 0}, cur[kMaxSys] = {0, 0, 0};
  for (int s = 0; s < S.count; ++s) {
    act[s] = FIRST ? 1 : st->st[s].active;
    cur[s] = FIRST ? 0 : st->st[s].cur;
  }
  for (int s = 0; s < S.count; ++s)
    if (act[s]) tot[s] = block_sum_array(J.part + (size_t)s * nb, nb, red);
  if (FIRST) tot[kMaxSys] = block_sum_array(J.part + (size_t)kMaxSys * nb, nb, red);
  if (threadIdx.x != 0) return;
  const double* rhs = S.hp_view ? J.hres[2] : J.h;
  if (FIRST) {
    // initial residual SUM(ABS(0 - rhs))/n, common to all systems
    int c0 = certified_continue(tot[kMaxSys], n, E.tol);
    int amb0 = 0;
    if (c0 < 0) {
      double r = sequential_residual(n, rhs, nullptr, 0);
      c0 = (E.tol < r) ? 1 : 0;
      amb0 = 1;
    }
    for (int s = 0; s < S.count; ++s) {
      JacState& q = st->st[s];
      q.ambiguous = amb0;
      q.cur = 1;
      if (!c0 || E.max_iter <= 0) {
        q.k = 0;
        q.active = 0;
        q.zero = 1;
        // no sweep: "not converged" when the initial residual is above the
        // tolerance (src/jacobi_iteration.f90:82-84)
        q.nonconv = c0 ? 1 : 0;
        q.residual = tot[kMaxSys] / (double)n;
        continue;
      }
      q.k = 1;
      q.zero = 0;
      int c1 = certified_continue(tot[s], n, E.tol);
      if (c1 < 0) {
        double r = sequential_residual(n, J.xj[S.m[s]][1], rhs, 1);
        c1 = (E.tol < r) ? 1 : 0;
        q.ambiguous += 1;
      }
      q.residual = tot[s] / (double)n;
      q.active = (c1 && q.k < E.max_iter) ? 1 : 0;
      q.nonconv = (c1 && !q.active) ? 1 : 0;
    }
  } else {
    for (int s = 0; s < S.count; ++s) {
      if (!act[s]) continue;
      JacState& q = st->st[s];
      q.k += 1;
      q.cur = 1 - cur[s];
      int c = certified_continue(tot[s], n, E.tol);
      if (c < 0) {
        const int m = S.m[s];
        double r = sequential_residual(n, J.xj[m][q.cur], J.xj[m][cur[s]], 1);
        c = (E.tol < r) ? 1 : 0;
        q.ambiguous += 1;
      }
      q.residual = tot[s] / (double)n;
      q.active = (c && q.k < E.max_iter) ? 1 : 0;
      q.nonconv = (c && !q.active) ? 1 : 0;
    }
  }
}

// min that propagates NaN (the stage check must see a NaN depth)
__device__ __forceinline__ double nan_min(double a, double b) {
  return (a != a || b != b) ? __builtin_nan("") : (a < b ? a : b);
}

__device__ __forceinline__ double block_min_all(double v, double* red) {
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) v = nan_min(v, __shfl_xor(v, off, 64));
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  __syncthreads();
  if (lane == 0) red[w] = v;
  __syncthreads();
  double r = nan_min(nan_min(nan_min(red[0], red[1]), red[2]), red[3]);
  __syncthreads();
  return r;
}

// Result of each solve into the node's depth: x[cur], or 0 when no sweep ran
// (src/jacobi_iteration.f90:68, src/timediscretization.f90:898-903).  With
// stage tracking on, each block also leaves the minimum of the depths it
// wrote in the (by now unused) Jacobi partial-sum slot 0.
__global__ __launch_bounds__(256) void k_finalize(EngineArgs E, const TileRef* blocks,
                                                  SysList S) {
  __shared__ double red[4];
  const TileRef T = blocks[blockIdx.x];
  const JobDev& J = E.jobs[T.job];
  if (skipped(E, J)) return;
  const int k = T.tile * kCellBlock + threadIdx.x;
  double mn = __builtin_inf();
  if (k < J.n) {
    for (int s = 0; s < S.count; ++s) {
      const int m = S.m[s];
      const JacState& q = J.status->st[s];
      const double v = q.zero ? 0. : G(J.xj[m][q.cur])[k];
      G(J.hres[m])[k] = v;
      mn = nan_min(mn, v);
    }
  }
  if (!S.track_min) return;  // uniform
  mn = block_min_all(mn, red);
  if (threadIdx.x == 0) J.part[T.tile] = mn;
}

// One block per job: fold the finalize blocks' minima into the step's
// minimum JobStatus::hmin_step (k_commit merges it into hmin once the step is
// accepted).
__global__ __launch_bounds__(256) void k_stage_min(EngineArgs E, const int* jobs) {
  __shared__ double red[4];
  const JobDev& J = E.jobs[jobs[blockIdx.x]];
  if (skipped(E, J)) return;
  double mn = __builtin_inf();
  for (int t = threadIdx.x; t < J.nblk; t += blockDim.x) mn = nan_min(mn, J.part[t]);
  mn = block_min_all(mn, red);
  if (threadIdx.x == 0) {
    JobStatus* st = J.status;
    st->hmin_step = st->hmin_step_set ? nan_min(st->hmin_step, mn) : mn;
    st->hmin_step_set = 1;
  }
}

// After a correction's Jacobi phase, one thread per job records the solves
// (sweep counts, non-converged and ambiguous stops) and whether any system was
// still running, i.e. needed more sweeps than were launched.
__global__ void k_solve_done(EngineArgs E, const int* jobs, SysList S, int first_solve) {
  if (threadIdx.x != 0) return;
  JobStatus* st = E.jobs[jobs[blockIdx.x]].status;
  if (E.gated && st->halt) return;
  if (first_solve == 0) {
    st->nonconv = 0;
    st->ambiguous = 0;
    st->overrun = 0;
    st->hmin_step_set = 0;  // this attempt's stages start a new minimum
  }
  for (int s = 0; s < S.count; ++s) {
    const JacState& q = st->st[s];
    st->iters[first_solve + s] = q.k;
    st->nonconv += q.nonconv;
    st->ambiguous += q.ambiguous;
    st->overrun |= q.active;
  }
}

// U^{n+1} = Ua(MSteps) (src/timediscretization.f90:564,662,762), node
// mf = MSteps - 1 here, for the jobs whose step k_gate accepted.
__global__ __launch_bounds__(256) void k_commit(EngineArgs E, const TileRef* blocks, int mf) {
  const TileRef T = blocks[blockIdx.x];
  const JobDev& J = E.jobs[T.job];
  const int k = T.tile * kCellBlock + threadIdx.x;
  if (!J.status->commit || k >= J.n) return;
  G(J.h)[k] = G(J.hres[mf])[k];
  G(J.hu)[k] = G(J.hun[mf])[k];
  G(J.hv)[k] = G(J.hvn[mf])[k];
}

// ------------------------------------------------------------------------
// The device time loop: up to kMaxBatch steps between host reads.
//
// k_batch_begin starts a batch (one thread per job).  After each step k_gate
// (one thread per job) decides whether the step is committed: not when the
// job's batch has already stopped, when a speculative step overran its
// launched Jacobi sweeps (U^n is intact and the host redoes the step), or when
// dt is not finite.  A committed step is recorded and advances the loop: t =
// t + dt and the output/end test of src/timediscretization.f90:83-94, the
// job's step count and the batch's.  Whatever stops the batch sets halt, and
// every later kernel of the batch skips the job.
// ------------------------------------------------------------------------
__global__ void k_batch_begin(EngineArgs E, const int* jobs, BatchArgs B) {
  if (threadIdx.x != 0) return;
  JobStatus* st = E.jobs[jobs[blockIdx.x]].status;
  if (B.redo && st->halt != kHaltOverrun) {  // keeps its own stop
    if (!st->halt) st->halt = kHaltMasked;
    st->commit = 0;
    return;
  }
  if (st->steps_left == 0) {  // the job has run all its steps
    st->halt = kHaltCount;
    st->commit = 0;
    return;
  }
  st->halt = kHaltNone;
  st->commit = 0;
  st->overrun = 0;
  st->batch_done = 0;
  st->batch_max = B.max_steps;
  if (B.redo) return;  // the redone step keeps the loop's t
  st->loop = B.loop;
  if (B.loop) {
    st->t = B.t;
    st->t_an = B.t_an;
    st->t_end = B.t_end;
  }
}

__global__ void k_gate(EngineArgs E, const int* jobs, int nsolve) {
  if (threadIdx.x != 0) return;
  JobStatus* st = E.jobs[jobs[blockIdx.x]].status;
  st->commit = 0;
  if (st->halt) return;
  if (st->overrun) {
    st->halt = kHaltOverrun;
    return;
  }
  if (st->dt_bad) {
    st->halt = kHaltDtBad;
    return;
  }
  st->commit = 1;
  StepRec& r = st->rec[st->batch_done];
  r.dt = st->dt;
  for (int i = 0; i < 13; ++i) r.iters[i] = i < nsolve ? st->iters[i] : 0;
  r.nonconv = nsolve > 0 ? st->nonconv : 0;
  r.ambiguous = nsolve > 0 ? st->ambiguous : 0;
  st->batch_done += 1;
  if (st->hmin_step_set) {  // the accepted step's stage minimum joins the record
    st->hmin = st->hmin_set ? nan_min(st->hmin, st->hmin_step) : st->hmin_step;
    st->hmin_set = 1;
    st->hmin_step_set = 0;
  }
  if (st->loop) {
    st->t = st->t + st->dt;
    if (fabs(st->t_an - st->t) < 1.0E-10 || fabs(st->t - st->t_end) < 1.0E-10)
      st->halt = kHaltReached;
  }
  if (st->steps_left > 0 && --st->steps_left == 0 && !st->halt) st->halt = kHaltCount;  // < 0: none
  if (!st->halt && st->batch_done >= st->batch_max) st->halt = kHaltBatch;
}

__global__ void k_set_steps_left(JobStatus* st, int steps_left) {
  if (threadIdx.x == 0 && blockIdx.x == 0) st->steps_left = steps_left;
}

// The reference's view of the depths in the two-node schemes.
// MODIFIED_PATANKAR_MATRIX_INVERSION declares hp(MSteps, nx, ny)
// (src/timediscretization.f90:783) and receives Up(:,1,1:nx,1:ny), whose
// first extent is 4 (src/finitevolume.f90:161), so hp is sequence associated:
// hp(m, cell) is element L = (m-1) + 2*cell of Up(1:4, 1, cells) in array
// element order, i.e. node L % 4 of cell L / 4 (cell = (i-1) + nx*(j-1), our
// k).  Nodes 3 and 4 are never written by these schemes and hold 0
// (src/finitevolume.f90:207-208).  So, with m = 2 the node solved:
//   hp(2, k) = k even ? Up(2)[k/2] : 0       -> hres[3]
//   rhs(k)  = hp(1, k) = k even ? h^n[k/2] : 0 -> hres[2]
// where Up(2) is h^n in the first correction (from_un) and node 1's depth
// after it.  hres[2] and hres[3] are free: only node 1 is solved.
__global__ __launch_bounds__(256) void k_hp_view(EngineArgs E, const TileRef* blocks,
                                                 int from_un) {
  const TileRef T = blocks[blockIdx.x];
  const JobDev& J = E.jobs[T.job];
  if (skipped(E, J)) return;
  const int k = T.tile * kCellBlock + threadIdx.x;
  if (k >= J.n) return;
  const gdouble* up2 = G(from_un ? J.h : J.hres[1]);
  const bool odd = k & 1;
  const int c = k >> 1;
  G(J.hres[3])[k] = odd ? 0. : up2[c];
  G(J.hres[2])[k] = odd ? 0. : G(J.h)[c];
}

// A stage combination of an explicit scheme (CombOp, mode 0/1/2), for the
// three variables of every cell.  Reads only slots other than op.out.
__device__ __forceinline__ const gdouble* state_var(const JobDev& J, int v, int s) {
  return G(v == 0 ? J.sh[s] : v == 1 ? J.shu[s] : J.shv[s]);
}
__device__ __forceinline__ const gdouble* deriv_var(const JobDev& J, int v, int r) {
  return G(v == 0 ? J.d1[r] : v == 1 ? J.d2[r] : J.d3[r]);
}

// Every source value is loaded once per variable (fully unrolled over
// kMaxTerms, guarded by nterm, so the values stay in registers) and serves
// each of the op.nout outputs, which differ only in their weights.
__global__ __launch_bounds__(256) void k_combine(EngineArgs E, const TileRef* blocks, CombOp op) {
  const TileRef T = blocks[blockIdx.x];
  const JobDev& J = E.jobs[T.job];
  if (skipped(E, J)) return;
  const int k = T.tile * kCellBlock + threadIdx.x;
  if (k >= J.n) return;
  const double dt = J.status->dt;
#pragma unroll
  for (int v = 0; v < 3; ++v) {
    double x[kMaxTerms];
#pragma unroll
    for (int t = 0; t < kMaxTerms; ++t)
      if (t < op.nterm)
        x[t] = op.deriv[t] ? deriv_var(J, v, op.src[t])[k] : state_var(J, v, op.src[t])[k];
    const double base = op.mode == 0 ? 0. : state_var(J, v, op.base)[k];
#pragma unroll
    for (int o = 0; o < kMaxOut; ++o) {
      if (o >= op.nout) break;
      const double* c = op.c[o];
      double acc = 0.;
      if (op.mode == 0) {
#pragma unroll
        for (int t = 0; t < kMaxTerms; ++t)
          if (t < op.nterm) {
            const double term = op.deriv[t] ? (c[t] * x[t]) * dt : c[t] * x[t];
            acc = t == 0 ? term : acc + term;
          }
      } else if (op.mode == 1) {
        double sum = 0.;
#pragma unroll
        for (int t = 0; t < kMaxTerms; ++t)
          if (t < op.nterm) {
            const double term = c[t] * x[t];
            sum = t == 0 ? term : sum + term;
          }
        acc = base + sum * dt;
      } else {
        acc = base;
#pragma unroll
        for (int t = 0; t < kMaxTerms; ++t)
          if (t < op.nterm) acc = acc + (dt * c[t]) * x[t];
      }
      const int out = op.out[o];
      G(v == 0 ? J.sh[out] : v == 1 ? J.shu[out] : J.shv[out])[k] = acc;
    }
  }
}

// FXWB/FYWB from the slot scratch after the well-balanced FVTD.  The job's
// cell blocks cover n indices; the loop also covers the nx (ny) extra faces.
__global__ __launch_bounds__(256) void k_copy_wb(EngineArgs E, const TileRef* blocks, int slot) {
  const TileRef T = blocks[blockIdx.x];
  const JobDev& J = E.jobs[T.job];
  const int nfx = (J.nx + 1) * J.ny, nfy = J.nx * (J.ny + 1);
  for (int k = T.tile * kCellBlock + threadIdx.x; k < nfx || k < nfy;
       k += J.nblk * kCellBlock) {
    for (int v = 0; v < 3; ++v) {
      if (k < nfx) G(J.FXWB[v])[k] = G(J.FX[slot][v])[k];
      if (k < nfy) G(J.FYWB[v])[k] = G(J.FY[slot][v])[k];
    }
  }
}

// ------------------------------------------------------------------------
// TimeStep (src/equation.f90:791-837): per-block maxima of the wave speeds,
// then one block per job folds them (swm_timestep.hpp) and applies the time
// loop's clipping.
// ------------------------------------------------------------------------
__global__ __launch_bounds__(256) void k_timestep(EngineArgs E, const TileRef* blocks) {
  const TileRef T = blocks[blockIdx.x];
  const JobDev& J = E.jobs[T.job];
  if (skipped(E, J)) return;
  const int k = T.tile * kCellBlock + threadIdx.x;
  const bool in = k < J.n;
  timestep_partial(in, J.g, in ? G(J.h)[k] : 0., in ? G(J.hu)[k] : 0., in ? G(J.hv)[k] : 0.,
                   J.tspart + 3 * (size_t)T.tile);
}

__global__ __launch_bounds__(256) void k_timestep_final(EngineArgs E, const int* jobs) {
  const JobDev& J = E.jobs[jobs[blockIdx.x]];
  if (skipped(E, J)) return;
  const gdouble *h = G(J.h), *hu = G(J.hu), *hv = G(J.hv);
  double ts, LX, LY;
  timestep_final(
      J.tspart, J.nblk, J.n,
      [&](int k, double& a, double& b, double& c) {
        a = h[k];
        b = hu[k];
        c = hv[k];
      },
      J.g, J.dx, J.dy, J.cfl, ts, LX, LY);
  if (threadIdx.x != 0) return;
  JobStatus* st = J.status;
  st->dt_min = ts;
  // the time loop's dt = MIN(MIN(dt_min, tAnalyze-t), MIN(dt_min, tEnd-t))
  // (src/timediscretization.f90:58-59) when the device runs the loop
  const double dt =
      st->loop ? min_r(min_r(ts, st->t_an - st->t), min_r(ts, st->t_end - st->t)) : ts;
  st->dt = dt;
  st->lmx = LX;
  st->lmy = LY;
  st->dt_bad = isfinite(dt) ? 0 : 1;
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

// a/b through the shared-reciprocal sequence (recip_refined + div_by),
// without the range check, for comparison with '/'.
__global__ void k_shared_div(int n, const double* a, const double* b, double* q) {
  int k = blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= n) return;
  q[k] = div_by(a[k], b[k], recip_refined(b[k]));
}

// WENO5 first sweep (out: wm, wp) and second sweep (out: 4 Gauss-point
// values) of n five-point stencils q[5*i .. 5*i+4], as the kernels compute them.
__global__ void k_weno5(int n, const double* q, double* first, double* second) {
  int k = blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= n) return;
  const double* s = q + 5 * (size_t)k;
  double wm, wp, w[4];
  weno5_first(s[0], s[1], s[2], s[3], s[4], wm, wp);
  weno5_second(s[0], s[1], s[2], s[3], s[4], w);
  first[2 * (size_t)k] = wm;
  first[2 * (size_t)k + 1] = wp;
  for (int g = 0; g < 4; ++g) second[4 * (size_t)k + g] = w[g];
}

// ------------------------------------------------------------------------
// The state the reference's output writer prints: WriteSolutionToDisk
// (src/output.f90:79) writes the module array V, which the last FVTD of the
// step filled from its input -- for mPDeC5 the fourth DeC node after the
// fourth correction (src/timediscretization.f90:548-556), not U^{n+1}.  A copy
// of that node's state for the jobs that asked for it (vout != null).
// ------------------------------------------------------------------------
__global__ __launch_bounds__(256) void k_keep_state(EngineArgs E, const TileRef* blocks, int m) {
  const TileRef T = blocks[blockIdx.x];
  const JobDev& J = E.jobs[T.job];
  if (skipped(E, J)) return;
  if (!J.vout) return;
  const int k = T.tile * kCellBlock + threadIdx.x;
  if (k >= J.n) return;
  const gdouble *h, *hu, *hv;
  node_state(J, m, h, hu, hv);
  gdouble* o = G(J.vout);
  o[k] = h[k];
  o[J.n + k] = hu[k];
  o[2 * (size_t)J.n + k] = hv[k];
}

// ------------------------------------------------------------------------
// launch wrappers
// ------------------------------------------------------------------------
static inline unsigned cdiv(long a, long b) { return (unsigned)((a + b - 1) / b); }

template <int XW, int XH, int YW, int YH>
static void launch_faces(const EngineArgs& E, const TileRef* xt, int nxt, const TileRef* yt,
                         int nyt, const NodeList& L, hipStream_t s) {
  if (nxt)
    hipLaunchKernelGGL((k_fvtd_x<XW, XH>), dim3(nxt, L.count), dim3(XW * XH), 0, s, E, xt, L);
  if (nyt)
    hipLaunchKernelGGL((k_fvtd_y<YW, YH>), dim3(nyt, L.count), dim3(YW * YH), 0, s, E, yt, L);
}

void launch_fvtd(const EngineArgs& E, int shape, const TileRef* xt, int nxt, const TileRef* yt,
                 int nyt, const TileRef* ct, int nct, const NodeList& L, hipStream_t s) {
  static_assert(kNumShapes == 2 && kShapes[0].xw == 16 && kShapes[0].xh == 16 &&
                    kShapes[0].yw == 16 && kShapes[0].yh == 16 && kShapes[1].xw == 32 &&
                    kShapes[1].xh == 16 && kShapes[1].yw == 16 && kShapes[1].yh == 32,
                "the instantiations below follow kShapes");
  if (shape == 1)
    launch_faces<32, 16, 16, 32>(E, xt, nxt, yt, nyt, L, s);
  else
    launch_faces<16, 16, 16, 16>(E, xt, nxt, yt, nyt, L, s);
  if (nct)
    hipLaunchKernelGGL(k_fvtd_update, dim3(cdiv(nct, 8) * 8 * L.count), dim3(256), 0, s, E, ct,
                       nct, L);
}

void launch_keep_state(const EngineArgs& E, const TileRef* cb, int ncb, int m, hipStream_t s) {
  if (ncb) hipLaunchKernelGGL(k_keep_state, dim3(ncb), dim3(256), 0, s, E, cb, m);
}

void launch_assemble(const EngineArgs& E, const TileRef* cb, int ncb, const int* jl, int njobs,
                     const SysList& S, hipStream_t s) {
  if (!ncb) return;
  if (S.nr == 2)
    hipLaunchKernelGGL(k_assemble<2>, dim3(ncb), dim3(256), 0, s, E, cb, S);
  else
    hipLaunchKernelGGL(k_assemble<4>, dim3(ncb), dim3(256), 0, s, E, cb, S);
  hipLaunchKernelGGL(k_decide<true>, dim3(njobs), dim3(256), 0, s, E, jl, S);
}

void launch_hp_view(const EngineArgs& E, const TileRef* cb, int ncb, int from_un, hipStream_t s) {
  if (ncb) hipLaunchKernelGGL(k_hp_view, dim3(ncb), dim3(256), 0, s, E, cb, from_un);
}

void launch_combine(const EngineArgs& E, const TileRef* cb, int ncb, const CombOp& op,
                    hipStream_t s) {
  if (ncb) hipLaunchKernelGGL(k_combine, dim3(ncb), dim3(256), 0, s, E, cb, op);
}

void launch_sweep(const EngineArgs& E, const TileRef* cb, int ncb, const int* jl, int njobs,
                  const SysList& S, hipStream_t s) {
  if (!ncb) return;
  hipLaunchKernelGGL(k_sweep, dim3(ncb), dim3(256), 0, s, E, cb, S);
  hipLaunchKernelGGL(k_decide<false>, dim3(njobs), dim3(256), 0, s, E, jl, S);
}

void launch_finalize(const EngineArgs& E, const TileRef* cb, int ncb, const int* jl, int njobs,
                     const SysList& S, hipStream_t s) {
  if (!ncb) return;
  hipLaunchKernelGGL(k_finalize, dim3(ncb), dim3(256), 0, s, E, cb, S);
  if (S.track_min) hipLaunchKernelGGL(k_stage_min, dim3(njobs), dim3(256), 0, s, E, jl);
}

void launch_timestep(const EngineArgs& E, const TileRef* cb, int ncb, const int* jl, int njobs,
                     hipStream_t s) {
  if (!ncb) return;
  hipLaunchKernelGGL(k_timestep, dim3(ncb), dim3(256), 0, s, E, cb);
  hipLaunchKernelGGL(k_timestep_final, dim3(njobs), dim3(256), 0, s, E, jl);
}

void launch_commit(const EngineArgs& E, const TileRef* cb, int ncb, int final_node,
                   hipStream_t s) {
  if (ncb) hipLaunchKernelGGL(k_commit, dim3(ncb), dim3(256), 0, s, E, cb, final_node);
}

void launch_gate(const EngineArgs& E, const int* jl, int njobs, int nsolve, hipStream_t s) {
  if (njobs) hipLaunchKernelGGL(k_gate, dim3(njobs), dim3(64), 0, s, E, jl, nsolve);
}

void launch_batch_begin(const EngineArgs& E, const int* jl, int njobs, const BatchArgs& B,
                        hipStream_t s) {
  if (njobs) hipLaunchKernelGGL(k_batch_begin, dim3(njobs), dim3(64), 0, s, E, jl, B);
}

void launch_set_steps_left(JobStatus* st, int steps_left, hipStream_t s) {
  hipLaunchKernelGGL(k_set_steps_left, dim3(1), dim3(64), 0, s, st, steps_left);
}

void launch_solve_done(const EngineArgs& E, const int* jl, int njobs, const SysList& S,
                       int first_solve, hipStream_t s) {
  if (njobs) hipLaunchKernelGGL(k_solve_done, dim3(njobs), dim3(64), 0, s, E, jl, S, first_solve);
}

void launch_copy_wb(const EngineArgs& E, const TileRef* cb, int ncb, int slot, hipStream_t s) {
  if (ncb) hipLaunchKernelGGL(k_copy_wb, dim3(ncb), dim3(256), 0, s, E, cb, slot);
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

void launch_weno5(int n, const double* q, double* first, double* second, hipStream_t s) {
  hipLaunchKernelGGL(k_weno5, dim3(cdiv(n, 256)), dim3(256), 0, s, n, q, first, second);
}

void launch_shared_div(int n, const double* a, const double* b, double* q, hipStream_t s) {
  hipLaunchKernelGGL(k_shared_div, dim3(cdiv(n, 256)), dim3(256), 0, s, n, a, b, q);
}

}  // namespace swm
