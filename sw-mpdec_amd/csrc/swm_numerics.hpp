// swm_numerics.hpp -- per-point FP64 arithmetic of the mPDeC5 / WENO5 PATANKAR
// step, shared by every kernel of the library.
//
// Each function evaluates the reference expression in the reference's order
// (Fortran: left to right, parentheses honoured) and the whole library is
// built with -ffp-contract=off, so no multiply-add is fused; '/' and sqrt are
// the IEEE correctly rounded FP64 operations.  Together that makes the device
// results bitwise equal to a contraction-free build of the reference.
// MAX/MIN follow the reference compiler's lowering, which fixes NaN
// behaviour: MAX(a,b) = (a > b) ? a : b, MIN(a,b) = (a < b) ? a : b.
#pragma once

#include <hip/hip_runtime.h>

namespace swm {

constexpr double kWenoEps = 1.0E-24;     // src/finitevolume_vars.f90:97
constexpr double kMinDepth = 1.0E-6;     // src/finitevolume_vars.f90:104
constexpr double kMinTimestep = 1.0E-30; // src/finitevolume_vars.f90:106
constexpr double kWLobatto = 1. / 12.;   // src/finitevolume_vars.f90:96
constexpr double kHuge = 1.7976931348623157e308;  // HUGE(1.0), real(8)

__device__ __forceinline__ double max_r(double a, double b) { return (a > b) ? a : b; }
__device__ __forceinline__ double min_r(double a, double b) { return (a < b) ? a : b; }

// ------------------------------------------------------------------------
// Division with a reciprocal shared between quotients of one denominator.
//
// gfx950 evaluates the IEEE quotient a/b as
//   sb = div_scale(b), sa = div_scale(a), r = rcp(sb),
//   r = fma(r, fma(-sb, r, 1), r)  (twice),  q = sa*r,
//   div_fixup(div_fmas(fma(-sb, q, sa), r, q), b, a).
// When a and b are normal, a/b is normal and their exponents are far from
// the ranges where the hardware rescales (|a|, |b| in [2^-500, 2^500] with
// |a/b| >= 2^-900 is far inside them), div_scale returns its operand,
// div_fmas is a plain fma and div_fixup returns its value, so recip_refined
// once per denominator followed by div_by per numerator yields the same bits
// as '/', with 5 + 3 instructions instead of 11 per quotient.  Callers check
// the range first and use '/' otherwise.
// ------------------------------------------------------------------------
#ifndef SWM_SHARED_RECIP
#define SWM_SHARED_RECIP 1
#endif
constexpr double kRecipLo = 0x1p-400;
constexpr double kRecipHi = 0x1p+400;

// 1 when kRecipLo <= b <= kRecipHi (0 for NaN); combined with '&' so that the
// three checks of a reconstruction need no branches.
__device__ __forceinline__ int recip_range(double b) {
  return (int)(b >= kRecipLo) & (int)(b <= kRecipHi);
}

__device__ __forceinline__ double recip_refined(double b) {
  double r = __builtin_amdgcn_rcp(b);
  double e = fma(-b, r, 1.0);
  r = fma(r, e, r);
  e = fma(-b, r, 1.0);
  r = fma(r, e, r);
  return r;
}

__device__ __forceinline__ double div_by(double a, double b, double rb) {
  double q = a * rb;
  double e = fma(-b, q, a);
  return fma(e, rb, q);
}

// a/b for a positive b in [2^-400, 2^400] with rb = recip_refined(b): the
// same bits as '/'.  A zero numerator needs no division (the quotient is a
// itself); numerators beyond the range of div_by (tiny, huge, NaN, Inf) take
// the IEEE quotient.
__device__ __forceinline__ double div_pos(double a, double b, double rb) {
#if SWM_SHARED_RECIP
  const double m = fabs(a);
  if ((int)(m >= kRecipLo) & (int)(m <= kRecipHi)) return div_by(a, b, rb);
  if (a == 0.) return a;
#endif
  return a / b;
}

// ConsToPrim, src/equation.f90:912-945:  ht = h + MIN_DEPTH/h; u = hu/ht.
__device__ __forceinline__ void cons_to_prim(double h, double hu, double hv, double& u,
                                             double& v) {
  double ht = h + kMinDepth / h;
  u = hu / ht;
  v = hv / ht;
}

// Smoothness indicators shared by both WENO5 sweeps,
// src/reconstruction.f90:699-704 and :1024-1029, and (eps + beta)^2.
__device__ __forceinline__ void weno5_den(double qm2, double qm1, double q0, double qp1,
                                          double qp2, double& d1, double& d2, double& d3) {
  double b1 = (1.0 / 3.0) * (4.0 * qm2 * qm2 - 19.0 * qm2 * qm1 + 25.0 * qm1 * qm1 +
                             11.0 * qm2 * q0 - 31.0 * qm1 * q0 + 10.0 * q0 * q0);
  double b2 = (1.0 / 3.0) * (4.0 * qm1 * qm1 - 13.0 * qm1 * q0 + 13.0 * q0 * q0 +
                             5.0 * qm1 * qp1 - 13.0 * q0 * qp1 + 4.0 * qp1 * qp1);
  double b3 = (1.0 / 3.0) * (10.0 * q0 * q0 - 31.0 * q0 * qp1 + 25.0 * qp1 * qp1 +
                             11.0 * q0 * qp2 - 19.0 * qp1 * qp2 + 4.0 * qp2 * qp2);
  d1 = (kWenoEps + b1) * (kWenoEps + b1);
  d2 = (kWenoEps + b2) * (kWenoEps + b2);
  d3 = (kWenoEps + b3) * (kWenoEps + b3);
}

// WENO5_FirstSweep, src/reconstruction.f90:673-755: cell average line values
// at the left (wm) and right (wp) faces.
__device__ __forceinline__ void weno5_first(double qm2, double qm1, double q0, double qp1,
                                            double qp2, double& wm, double& wp) {
  double d1, d2, d3;
  weno5_den(qm2, qm1, q0, qp1, qp2, d1, d2, d3);
#if SWM_SHARED_RECIP
  // With the (eps+beta)^2 in [2^-400, 2^400] the alphas lie in
  // [2^-404, 2^400], their sum in [2^-404, 2^402] and the weights above 2^-806.
  if (recip_range(d1) & recip_range(d2) & recip_range(d3)) {
    const double r1 = recip_refined(d1), r2 = recip_refined(d2), r3 = recip_refined(d3);
    {
      double a1 = div_by(3.0 / 10.0, d1, r1);
      double a2 = div_by(3.0 / 5.0, d2, r2);
      double a3 = div_by(1.0 / 10.0, d3, r3);
      double s = a1 + a2 + a3;
      double rs = recip_refined(s);
      double o1 = div_by(a1, s, rs), o2 = div_by(a2, s, rs), o3 = div_by(a3, s, rs);
      double w1 = (1.0 / 6.0) * (-qm2 + 5.0 * qm1 + 2.0 * q0);
      double w2 = (1.0 / 6.0) * (2.0 * qm1 + 5.0 * q0 - qp1);
      double w3 = (1.0 / 6.0) * (11.0 * q0 - 7.0 * qp1 + 2.0 * qp2);
      wm = o1 * w1 + o2 * w2 + o3 * w3;
    }
    {
      double a1 = div_by(1.0 / 10.0, d1, r1);
      double a2 = div_by(3.0 / 5.0, d2, r2);
      double a3 = div_by(3.0 / 10.0, d3, r3);
      double s = a1 + a2 + a3;
      double rs = recip_refined(s);
      double o1 = div_by(a1, s, rs), o2 = div_by(a2, s, rs), o3 = div_by(a3, s, rs);
      double w1 = (1.0 / 6.0) * (2.0 * qm2 - 7.0 * qm1 + 11.0 * q0);
      double w2 = (1.0 / 6.0) * (-qm1 + 5.0 * q0 + 2.0 * qp1);
      double w3 = (1.0 / 6.0) * (2.0 * q0 + 5.0 * qp1 - qp2);
      wp = o1 * w1 + o2 * w2 + o3 * w3;
    }
    return;
  }
#endif
  {
    double a1 = (3.0 / 10.0) / d1;
    double a2 = (3.0 / 5.0) / d2;
    double a3 = (1.0 / 10.0) / d3;
    double s = a1 + a2 + a3;
    double o1 = a1 / s, o2 = a2 / s, o3 = a3 / s;
    double w1 = (1.0 / 6.0) * (-qm2 + 5.0 * qm1 + 2.0 * q0);
    double w2 = (1.0 / 6.0) * (2.0 * qm1 + 5.0 * q0 - qp1);
    double w3 = (1.0 / 6.0) * (11.0 * q0 - 7.0 * qp1 + 2.0 * qp2);
    wm = o1 * w1 + o2 * w2 + o3 * w3;
  }
  {
    double a1 = (1.0 / 10.0) / d1;
    double a2 = (3.0 / 5.0) / d2;
    double a3 = (3.0 / 10.0) / d3;
    double s = a1 + a2 + a3;
    double o1 = a1 / s, o2 = a2 / s, o3 = a3 / s;
    double w1 = (1.0 / 6.0) * (2.0 * qm2 - 7.0 * qm1 + 11.0 * q0);
    double w2 = (1.0 / 6.0) * (-qm1 + 5.0 * q0 + 2.0 * qp1);
    double w3 = (1.0 / 6.0) * (2.0 * q0 + 5.0 * qp1 - qp2);
    wp = o1 * w1 + o2 * w2 + o3 * w3;
  }
}

// Coefficients of one Gauss point of WENO5_SecondSweep,
// src/reconstruction.f90:1036-1133.  The literals are the reference's decimal
// strings; signed literals are folded into the coefficients (a - c*q ==
// a + (-c)*q exactly).  c13, c14, c20, c24, c30, c31 are the reference's
// 0.0*Q terms.
struct Weno2Coef {
  double g1, g2, g3;
  double c10, c11, c12, c13, c14, c20, c21, c22, c23, c24, c30, c31, c32, c33, c34;
};

template <int G>
__device__ constexpr Weno2Coef weno2_coef() {
  if constexpr (G == 0)
    return {0.2658420974778319, 0.6112504900322486, 0.1229074124899195,
            -0.1642562761719537, 0.7590807081409336, 0.4051755680310201, 0.0000000000000000, 0.0000000000000000,
            +0.0000000000000000, 0.2663118796250726, 0.8979443965468811, -0.1642562761719537, 0.0000000000000000,
            +0.0000000000000000, 0.0000000000000000, 1.6968800354220990, -0.9631919150471715, 0.2663118796250726};
  else if constexpr (G == 1)
    return {0.1281641584355059, 0.5219691498503563, 0.3498666917141379,
            -0.1122135388132497, 0.3944175994189276, 0.7177959393943222, 0.0000000000000000, 0.0000000000000000,
            +0.0000000000000000, 0.0577769829791784, 1.0544365558340714, -0.1122135388132497, 0.0000000000000000,
            +0.0000000000000000, 0.0000000000000000, 1.2277675047716066, -0.2855444877507849, 0.0577769829791784};
  else if constexpr (G == 2)
    return {0.3498666917141379, 0.5219691498503563, 0.1281641584355059,
            +0.0577769829791784, -0.2855444877507849, 1.2277675047716066, 0.0000000000000000, 0.0000000000000000,
            +0.0000000000000000, -0.1122135388132497, 1.0544365558340714, 0.0577769829791784, 0.0000000000000000,
            +0.0000000000000000, 0.0000000000000000, 0.7177959393943222, 0.3944175994189276, -0.1122135388132497};
  else
    return {0.1229074124899195, 0.6112504900322486, 0.2658420974778319,
            +0.2663118796250726, -0.9631919150471715, 1.6968800354220990, 0.0000000000000000, 0.0000000000000000,
            +0.0000000000000000, -0.1642562761719537, 0.8979443965468811, 0.2663118796250726, 0.0000000000000000,
            +0.0000000000000000, 0.0000000000000000, 0.4051755680310201, 0.7590807081409336, -0.1642562761719537};
}

// The three candidate values, every term kept: the 0.0*Q terms propagate
// NaN/Inf and fix the sign of zero exactly as the reference does (the
// compiler shares the four distinct 0.0*q products between the Gauss points).
template <int G>
__device__ __forceinline__ void weno2_cand(double qm2, double qm1, double q0, double qp1,
                                           double qp2, double& w1, double& w2, double& w3) {
  constexpr Weno2Coef C = weno2_coef<G>();
  w1 = C.c10 * qm2 + C.c11 * qm1 + C.c12 * q0 + C.c13 * qp1 + C.c14 * qp2;
  w2 = C.c20 * qm2 + C.c21 * qm1 + C.c22 * q0 + C.c23 * qp1 + C.c24 * qp2;
  w3 = C.c30 * qm2 + C.c31 * qm1 + C.c32 * q0 + C.c33 * qp1 + C.c34 * qp2;
}

// omega_k = alpha_k / sum(alpha), alpha_k = gamma_k / (eps + beta_k)^2
// (src/reconstruction.f90:1030-1035), through shared reciprocals (SHARED) or '/'.
template <int G, bool SHARED>
__device__ __forceinline__ void weno2_weights(double d1, double d2, double d3, double r1,
                                              double r2, double r3, double& o1, double& o2,
                                              double& o3) {
  constexpr Weno2Coef C = weno2_coef<G>();
  if constexpr (SHARED) {
    double a1 = div_by(C.g1, d1, r1);
    double a2 = div_by(C.g2, d2, r2);
    double a3 = div_by(C.g3, d3, r3);
    double s = a1 + a2 + a3;
    double rs = recip_refined(s);
    o1 = div_by(a1, s, rs);
    o2 = div_by(a2, s, rs);
    o3 = div_by(a3, s, rs);
  } else {
    double a1 = C.g1 / d1;
    double a2 = C.g2 / d2;
    double a3 = C.g3 / d3;
    double s = a1 + a2 + a3;
    o1 = a1 / s;
    o2 = a2 / s;
    o3 = a3 / s;
  }
}

template <int G, bool SHARED>
__device__ __forceinline__ double weno5_second_gp(double qm2, double qm1, double q0, double qp1,
                                                  double qp2, double d1, double d2, double d3,
                                                  double r1, double r2, double r3) {
  double o1, o2, o3, w1, w2, w3;
  weno2_weights<G, SHARED>(d1, d2, d3, r1, r2, r3, o1, o2, o3);
  weno2_cand<G>(qm2, qm1, q0, qp1, qp2, w1, w2, w3);
  return (o1 * w1 + o2 * w2 + o3 * w3);
}

// WENO5_SecondSweep, src/reconstruction.f90:999-1136: point values at the
// four Gauss nodes of the cell.
__device__ __forceinline__ void weno5_second(double qm2, double qm1, double q0, double qp1,
                                             double qp2, double w[4]) {
  double d1, d2, d3;
  weno5_den(qm2, qm1, q0, qp1, qp2, d1, d2, d3);
#if SWM_SHARED_RECIP
  // the gammas lie in [0.12, 0.62]: same ranges as in weno5_first
  if (recip_range(d1) & recip_range(d2) & recip_range(d3)) {
    const double r1 = recip_refined(d1), r2 = recip_refined(d2), r3 = recip_refined(d3);
    w[0] = weno5_second_gp<0, true>(qm2, qm1, q0, qp1, qp2, d1, d2, d3, r1, r2, r3);
    w[1] = weno5_second_gp<1, true>(qm2, qm1, q0, qp1, qp2, d1, d2, d3, r1, r2, r3);
    w[2] = weno5_second_gp<2, true>(qm2, qm1, q0, qp1, qp2, d1, d2, d3, r1, r2, r3);
    w[3] = weno5_second_gp<3, true>(qm2, qm1, q0, qp1, qp2, d1, d2, d3, r1, r2, r3);
    return;
  }
#endif
  w[0] = weno5_second_gp<0, false>(qm2, qm1, q0, qp1, qp2, d1, d2, d3, 0., 0., 0.);
  w[1] = weno5_second_gp<1, false>(qm2, qm1, q0, qp1, qp2, d1, d2, d3, 0., 0., 0.);
  w[2] = weno5_second_gp<2, false>(qm2, qm1, q0, qp1, qp2, d1, d2, d3, 0., 0., 0.);
  w[3] = weno5_second_gp<3, false>(qm2, qm1, q0, qp1, qp2, d1, d2, d3, 0., 0., 0.);
}

#ifndef SWM_LIMITER_SKIP
#define SWM_LIMITER_SKIP 1
#endif

// PositivityLimiterX/Y body, src/finitevolume.f90:671-691: theta-scaling of
// the h values at the Gauss points of one cell towards the cell mean vh.
__device__ __forceinline__ void positivity_limit(double vh, const double wbnd[4], double wm[4],
                                                 double wp[4]) {
  double alpha = 0., beta = 0.;
#pragma unroll
  for (int g = 0; g < 4; ++g) {
    alpha = alpha + wbnd[g] * wm[g];
    beta = beta + wbnd[g] * wp[g];
  }
  double csi = (vh - kWLobatto * alpha - kWLobatto * beta) / (1. - 2. * kWLobatto);
#pragma unroll
  for (int g = 0; g < 4; ++g) {
    double mmin = min_r(min_r(csi, wm[g]), wp[g]);
    const double num = vh - kMinDepth, den = vh - mmin;
    double theta;
    if (vh == mmin)
      theta = 1.;
#if SWM_LIMITER_SKIP
    // |num| >= |den| with num finite makes the rounded |num/den| >= 1, so
    // MIN(1, |num/den|) is 1 without the division (the usual case: the
    // Gauss values lie close to the mean)
    else if ((int)(fabs(num) >= fabs(den)) & (int)(fabs(num) <= kHuge))
      theta = 1.;
#endif
    else
      theta = min_r(1., fabs(num / den));
    wp[g] = vh + theta * (wp[g] - vh);
    wm[g] = vh + theta * (wm[g] - vh);
  }
}

// EvaluateFlux1D, src/equation.f90:1014-1016
__device__ __forceinline__ void flux_1d(double h, double vx, double vy, double g, double& f1,
                                        double& f2, double& f3) {
  f1 = h * vx;
  f2 = h * (vx * vx) + (0.5 * g) * (h * h);
  f3 = (h * vx) * vy;
}

// RiemannSolver + RiemannSolverByRusanov at one Gauss point,
// src/equation.f90:1048-1069 and :1096-1103.  (n1,n2),(t1,t2) are the face's
// NormVect/TangVect (src/mesh.f90:74-110); they are compile-time constants
// here, and the products with 0 and 1 are kept as in the reference.
template <int N1, int N2, int T1, int T2>
__device__ __forceinline__ void rusanov(double hl, double ul, double vl, double hr, double ur,
                                        double vr, double g, double& F1, double& F2, double& F3) {
  const double n1 = N1, n2 = N2, t1 = T1, t2 = T2;
  double pl1 = hl, pl2 = n1 * ul + n2 * vl, pl3 = t1 * ul + t2 * vl;
  double pr1 = hr, pr2 = n1 * ur + n2 * vr, pr3 = t1 * ur + t2 * vr;
  double cl2 = pl1 * pl2, cl3 = pl1 * pl3;  // PrimToCons
  double cr2 = pr1 * pr2, cr3 = pr1 * pr3;
  double fl1, fl2, fl3, fr1, fr2, fr3;
  flux_1d(pl1, pl2, pl3, g, fl1, fl2, fl3);
  flux_1d(pr1, pr2, pr3, g, fr1, fr2, fr3);
  double fastl = pl2 + sqrt(g * pl1);
  double fastr = pr2 + sqrt(g * pr1);
  double lmax = max_r(fabs(fastl), fabs(fastr));
  double f1 = 0.5 * ((fl1 + fr1) - lmax * (pr1 - pl1));
  double f2 = 0.5 * ((fl2 + fr2) - lmax * (cr2 - cl2));
  double f3 = 0.5 * ((fl3 + fr3) - lmax * (cr3 - cl3));
  F1 = f1;
  F2 = n1 * f2 + t1 * f3;
  F3 = n2 * f2 + t2 * f3;
}

// JamesonIndicator, src/shocksindicator.f90:93-129, on the depths of three
// neighbouring cells.
__device__ __forceinline__ bool jameson_shock(double hm, double h0, double hp) {
  double ind = fabs(hm + hp - 2.0 * h0);
  ind = ind / (fabs(hm) + fabs(hp) + fabs(2.0 * h0));
  return ind > 0.5E-03;
}

// MINMOD, src/reconstruction.f90:387-402; SIGN(1.0, x) is copysign(1.0, x).
__device__ __forceinline__ double minmod(double x, double y) {
  return 0.5 * (copysign(1.0, x) + copysign(1.0, y)) * min_r(fabs(x), fabs(y));
}

// ReconstructionFixX/Y (src/reconstruction.f90:219-242, :298-321) for one
// variable of a flagged cell, from the cell's value q0 and its neighbours
// along the sweep direction: recfix 1 puts q0 at every Gauss point, 2 the
// MUSCL values (src/reconstruction.f90:348-381) with d = MESH_DX(dir).
__device__ __forceinline__ void reconstruction_fix(int recfix, double qm1, double q0, double qp1,
                                                   double d, double wm[4], double wp[4]) {
  double m = q0, p = q0;
  if (recfix == 2) {
    double sp = (qp1 - q0) / d;
    double sm = (q0 - qm1) / d;
    double slope = minmod(sm, sp);
    m = q0 - 0.5 * slope * d;
    p = q0 + 0.5 * slope * d;
  }
#pragma unroll
  for (int g = 0; g < 4; ++g) {
    wm[g] = m;
    wp[g] = p;
  }
}

// PATANKAR_DIV, src/timediscretization.f90:913-924
__device__ __forceinline__ double patankar_div(double prod, double h) {
  if (h < kMinDepth) return 0.;
  return 2. * h * prod / (h * h + max_r(h * h, kMinDepth));
}

}  // namespace swm
