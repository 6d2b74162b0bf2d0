// WENO building blocks shared by the single-species and mixture flux kernels (device code, FP64).
#pragma once

#include <hip/hip_runtime.h>

namespace corefl {
namespace wenodev {

// 1/d from the hardware reciprocal refined by two Newton steps (relative error at the rounding level),
// cheaper than an IEEE division; d > 0 and normal wherever it is used.
__device__ __forceinline__ double rcp(double d) {
  double r = __builtin_amdgcn_rcp(d);
  r = fma(fma(-d, r, 1.0), r, r);
  r = fma(fma(-d, r, 1.0), r, r);
  return r;
}

// The hardware reciprocal refined by one Newton step: relative error below 2.3e-15 (the hardware value is good to
// about 5e-8 and one step squares that).  Used where it scales only the nonlinear part of a WENO weight.
__device__ __forceinline__ double rcp1(double d) {
  const double r = __builtin_amdgcn_rcp(d);
  return fma(fma(-d, r, 1.0), r, r);
}

// ---------------------------------------------------------------- WENO halves
// Upwind-biased value at the right face of v[2] from v[0..4] (the fPlus part of WENO.cu:3059-3114).
// The downwind-biased value is the same function of the mirrored stencil.
__device__ __forceinline__ double weno5_half(double v0, double v1, double v2, double v3, double v4, double eps) {
  constexpr double o6 = 1.0 / 6, t12 = 13.0 / 12;
  const double p0 = o6 * (2 * v2 + 5 * v3 - v4);
  const double p1 = o6 * (-v1 + 5 * v2 + 2 * v3);
  const double p2 = o6 * (2 * v0 - 7 * v1 + 11 * v2);
  const double e0 = v2 + v4 - 2 * v3, f0 = 3 * v2 - 4 * v3 + v4;
  const double e1 = v1 + v3 - 2 * v2, f1 = v1 - v3;
  const double e2 = v0 + v2 - 2 * v1, f2 = v0 - 4 * v1 + 3 * v2;
  const double b0 = t12 * e0 * e0 + 0.25 * f0 * f0;
  const double b1 = t12 * e1 * e1 + 0.25 * f1 * f1;
  const double b2 = t12 * e2 * e2 + 0.25 * f2 * f2;
  // 1/d0 and 1/d1 from one reciprocal of their product (each d >= eps^2, so the product stays normal)
  const double d0 = (eps + b0) * (eps + b0), d1 = (eps + b1) * (eps + b1);
  const double r01 = rcp(d0 * d1);
  const double a0 = 0.3 * (d1 * r01);
  const double a1 = 0.6 * (d0 * r01);
  const double a2 = 0.1 * rcp((eps + b2) * (eps + b2));
  return (a0 * p0 + a1 * p1 + a2 * p2) * rcp(a0 + a1 + a2);
}

// ---------------------------------------------------------------- WENO7 (the fPlus part of WENO.cu:3116-3222)
// The smoothness indicator of candidate window k (v[k..k+3]) is beta_k = s1^2 + 13/12 s2^2 + 1043/960 s3^2
// + 1/12 s1 s3, with s1 a first-derivative estimate, s2 a second and s3 the third difference of the window.
// Here every beta is carried times 36, with S1 = 6 s1:
//   36 beta = S1 (S1 + 0.5 s3) + 39 s2^2 + 39.1125 s3^2,
// and eps times 36 with it: the Z weights c_k (1 + tau / (eps + beta_k)^2), tau = (b0 + 3 b1 - 3 b2 - b3)^2,
// do not change under a common scale of eps and the betas.  The third difference of a window read mirrored is
// the negative of the one read forwards, so the two halves of an interface share three of them.

// third difference of v[0..3]: -v0 + 3 v1 - 3 v2 + v3
__device__ __forceinline__ double w7_third(double v0, double v1, double v2, double v3) {
  return fma(3.0, v1 - v2, v3 - v0);
}

// 36 beta of one window from S1, s2, its third difference t (sign as read) and tt = 39.1125 t^2
__device__ __forceinline__ double w7_beta36(double S1, double s2, double t, double tt) {
  return fma(S1, fma(0.5, t, S1), fma(39.0 * s2, s2, tt));
}

// Upwind-biased value at the right face of v[3] from v[0..6], given the four windows' third differences
// t[k] (of v[k..k+3], read forwards) and tt[k] = 39.1125 t[k]^2.
__device__ __forceinline__ double weno7_half_t(const double *v, const double *t, const double *tt, double eps36) {
  const double s2m = fma(-2.0, v[3], v[2] + v[4]);
  const double b0 = w7_beta36(fma(-2.0, v[0], fma(9.0, v[1], fma(-18.0, v[2], 11.0 * v[3]))),
                              fma(4.0, v[1], fma(-5.0, v[2], fma(2.0, v[3], -v[0]))), t[0], tt[0]);
  const double b1 = w7_beta36(fma(2.0, v[4], fma(3.0, v[3], fma(-6.0, v[2], v[1]))), s2m, t[1], tt[1]);
  const double b2 = w7_beta36(fma(6.0, v[4], fma(-3.0, v[3], fma(-2.0, v[2], -v[5]))), s2m, t[2], tt[2]);
  const double b3 = w7_beta36(fma(2.0, v[6], fma(-9.0, v[5], fma(18.0, v[4], -11.0 * v[3]))),
                              fma(2.0, v[3], fma(-5.0, v[4], fma(4.0, v[5], -v[6]))), t[3], tt[3]);
  double tau = fma(3.0, b1 - b2, b0 - b3);
  tau *= tau;
  constexpr double c0 = 1.0 / 35, c1 = 12.0 / 35, c2 = 18.0 / 35, c3 = 4.0 / 35;
  // the four 1/d_k from two reciprocals of pair products (each d >= eps^2, so a product stays normal)
  const double d0 = (eps36 + b0) * (eps36 + b0), d1 = (eps36 + b1) * (eps36 + b1);
  const double d2 = (eps36 + b2) * (eps36 + b2), d3 = (eps36 + b3) * (eps36 + b3);
  const double r01 = tau * rcp1(d0 * d1), r23 = tau * rcp1(d2 * d3);
  const double a0 = fma(c0 * d1, r01, c0);
  const double a1 = fma(c1 * d0, r01, c1);
  const double a2 = fma(c2 * d3, r23, c2);
  const double a3 = fma(c3 * d2, r23, c3);
  const double q0 = -3 * v[0] + 13 * v[1] - 23 * v[2] + 25 * v[3];
  const double q1 = v[1] - 5 * v[2] + 13 * v[3] + 3 * v[4];
  const double q2 = -v[2] + 7 * v[3] + 7 * v[4] - v[5];
  const double q3 = 3 * v[3] + 13 * v[4] - 5 * v[5] + v[6];
  return (1.0 / 12) * (a0 * q0 + a1 * q1 + a2 * q2 + a3 * q3) * rcp(a0 + a1 + a2 + a3);
}

// Upwind-biased value at the right face of v[3] from v[0..6] (Z-type weights with tau7 = (b0 + 3 b1 - 3 b2 - b3)^2).
__device__ __forceinline__ double weno7_half(const double *v, double eps) {
  double t[4], tt[4];
#pragma unroll
  for (int k = 0; k < 4; ++k) {
    t[k] = w7_third(v[k], v[k + 1], v[k + 2], v[k + 3]);
    tt[k] = (39.1125 * t[k]) * t[k];
  }
  return weno7_half_t(v, t, tt, 36.0 * eps);
}

// Both halves at the interface between w[3] and w[4] of the eight values w[0..7]: lo is the value upwind from
// the left (v = w[0..6]), hi the one upwind from the right (v = w[7], w[6], .., w[1]).  The right half's
// window k is w[4-k..7-k] read backwards, whose third difference is minus the forward one of that window.
__device__ __forceinline__ void weno7_pair(const double *w, double eps, double &lo, double &hi) {
  double T[5], TT[5];
#pragma unroll
  for (int j = 0; j < 5; ++j) {
    T[j] = w7_third(w[j], w[j + 1], w[j + 2], w[j + 3]);
    TT[j] = (39.1125 * T[j]) * T[j];
  }
  const double eps36 = 36.0 * eps;
  lo = weno7_half_t(w, T, TT, eps36);
  const double m[7] = {w[7], w[6], w[5], w[4], w[3], w[2], w[1]};
  const double tm[4] = {-T[4], -T[3], -T[2], -T[1]}, ttm[4] = {TT[4], TT[3], TT[2], TT[1]};
  hi = weno7_half_t(m, tm, ttm, eps36);
}

// The linear upwind halves the hybrid scheme uses where the stencil is smooth: the WENO5 candidates with the
// optimal weights (UD5, WENO.cu:3097-3107) and the WENO7 ones (UD7, WENO.cu:3205-3220), as the reference writes them.
__device__ __forceinline__ double ud5_half(double v0, double v1, double v2, double v3, double v4) {
  constexpr double o6 = 1.0 / 6;
  const double p0 = o6 * (2 * v2 + 5 * v3 - v4);
  const double p1 = o6 * (-v1 + 5 * v2 + 2 * v3);
  const double p2 = o6 * (2 * v0 - 7 * v1 + 11 * v2);
  return 0.3 * p0 + 0.6 * p1 + 0.1 * p2;
}

__device__ __forceinline__ double ud7_half(const double *v) {
  constexpr double c0 = 1.0 / 35, c1 = 12.0 / 35, c2 = 18.0 / 35, c3 = 4.0 / 35;
  const double q0 = -3 * v[0] + 13 * v[1] - 23 * v[2] + 25 * v[3];
  const double q1 = v[1] - 5 * v[2] + 13 * v[3] + 3 * v[4];
  const double q2 = -v[2] + 7 * v[3] + 7 * v[4] - v[5];
  const double q3 = 3 * v[3] + 13 * v[4] - 5 * v[5] + v[6];
  return (1.0 / 12) * (c0 * q0 + c1 * q1 + c2 * q2 + c3 * q3);
}

// The hybrid scheme's switch for the interface between node id and id + st (WENO.cu:838-849): a shock unless
// the threshold is on and no node id - (g-1) st .. id + g st has a sensor above it.
__device__ __forceinline__ bool stencil_shock(const double *sensor, double threshold, long id, long st, int g) {
  if (!(threshold > 1e-10)) return true;
  bool shock = false;
  for (int ii = -g + 1; ii <= g; ++ii) shock = shock || sensor[id + ii * st] > threshold;
  return shock;
}

} // namespace wenodev
} // namespace corefl
