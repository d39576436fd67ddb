// swm_timestep.hpp -- TimeStep (src/equation.f90:791-837) as two block-level
// device routines, shared by the engine's k_timestep / k_timestep_final and by
// the allocation-free swm_timestep_noalloc.
//
// TimeStep folds dt = MIN(dt, MIN(dx/LX, dy/LY)) cell by cell, where LX, LY
// are running maxima of |u + sqrt(g h)| and |v + sqrt(g h)| (WaveSpeeds2D,
// :895-905) starting at 0.  Without NaN the fold equals
// CFL * MIN(MIN(HUGE, dx/LX), dy/LY) with LX, LY the plain maxima (MAX and MIN
// are monotone and dx/L decreases with L), which any reduction order gives
// bitwise.  With a NaN speed the fold depends on the order, so the final
// routine replays it sequentially in the reference's loop order.
#pragma once

#include <hip/hip_runtime.h>

#include "swm_numerics.hpp"

namespace swm {

// WaveSpeeds2D of one cell: |u + sqrt(g h)|, |v + sqrt(g h)|.
__device__ __forceinline__ void wave_speeds(double g, double h, double hu, double hv, double& a,
                                            double& b) {
  double u, v;
  cons_to_prim(h, hu, hv, u, v);
  double fx = u + sqrt(g * h);
  double fy = v + sqrt(g * h);
  a = fabs(fx);
  b = fabs(fy);
}

__device__ __forceinline__ double wave_max(double v) {
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) v = max_r(v, __shfl_xor(v, off, 64));
  return v;
}

// Block of 256 threads, one cell (or none: in = false) per thread: thread 0
// writes the block's (max a, max b, NaN seen) to p[0..2].
__device__ __forceinline__ void timestep_partial(bool in, double g, double h, double hu, double hv,
                                                 double* p) {
  __shared__ double sa[4], sb[4], sn[4];
  double a = 0.0, b = 0.0, nan = 0.0;
  if (in) {
    wave_speeds(g, h, hu, hv, a, b);
    if (isnan(a) || isnan(b)) {
      nan = 1.0;
      a = 0.0;
      b = 0.0;
    }
  }
  a = wave_max(a);
  b = wave_max(b);
  nan = wave_max(nan);
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  if (lane == 0) {
    sa[w] = a;
    sb[w] = b;
    sn[w] = nan;
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    p[0] = max_r(max_r(max_r(sa[0], sa[1]), sa[2]), sa[3]);
    p[1] = max_r(max_r(max_r(sb[0], sb[1]), sb[2]), sb[3]);
    p[2] = max_r(max_r(max_r(sn[0], sn[1]), sn[2]), sn[3]);
  }
}

// One block of 256 threads over the nb partial triples part[3*t..3*t+2];
// returns TimeStep's value (after the CFL factor and the MIN_TIMESTEP floor)
// and the maxima LambdaMaxX/Y in thread 0.  cell(k, h, hu, hv) loads cell k
// (0-based, x fastest) for the sequential replay when a speed was NaN.
template <class Cell>
__device__ __forceinline__ void timestep_final(const double* part, int nb, int n, Cell cell,
                                               double g, double dx, double dy, double cfl,
                                               double& ts_out, double& lx_out, double& ly_out) {
  __shared__ double sa[4], sb[4], sn[4];
  double a = 0.0, b = 0.0, nan = 0.0;
  for (int t = threadIdx.x; t < nb; t += blockDim.x) {
    a = max_r(a, part[3 * (size_t)t]);
    b = max_r(b, part[3 * (size_t)t + 1]);
    nan = max_r(nan, part[3 * (size_t)t + 2]);
  }
  a = wave_max(a);
  b = wave_max(b);
  nan = wave_max(nan);
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  if (lane == 0) {
    sa[w] = a;
    sb[w] = b;
    sn[w] = nan;
  }
  __syncthreads();
  if (threadIdx.x != 0) return;
  double LX = 0.0, LY = 0.0;
  for (int q = 0; q < 4; ++q) {
    LX = max_r(LX, sa[q]);
    LY = max_r(LY, sb[q]);
  }
  const bool anynan = (sn[0] + sn[1] + sn[2] + sn[3]) != 0.0;
  double ts;
  if (!anynan) {
    ts = min_r(min_r(kHuge, dx / LX), dy / LY);
  } else {
    // the reference's fold, cell by cell in its loop order
    LX = 0.0;
    LY = 0.0;
    ts = kHuge;
    for (int k = 0; k < n; ++k) {
      double h, hu, hv, aa, bb;
      cell(k, h, hu, hv);
      wave_speeds(g, h, hu, hv, aa, bb);
      LX = max_r(LX, aa);
      LY = max_r(LY, bb);
      ts = min_r(min_r(ts, dx / LX), dy / LY);
    }
  }
  ts = cfl * ts;
  if (ts < kMinTimestep) ts = kMinTimestep;
  ts_out = ts;
  lx_out = LX;
  ly_out = LY;
}

}  // namespace swm
