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

// Upwind-biased value at the right face of v[3] from v[0..6] (the fPlus part of WENO.cu:3116-3222,
// Z-type weights with tau7 = (b0 + 3 b1 - 3 b2 - b3)^2).
__device__ __forceinline__ double weno7_half(const double *v, double eps) {
  constexpr double o6 = 1.0 / 6, d12 = 13.0 / 12.0, d13 = 1043.0 / 960, d14 = 1.0 / 12;
  double s1 = o6 * (-2 * v[0] + 9 * v[1] - 18 * v[2] + 11 * v[3]);
  double s2 = -v[0] + 4 * v[1] - 5 * v[2] + 2 * v[3];
  double s3 = -v[0] + 3 * v[1] - 3 * v[2] + v[3];
  const double b0 = s1 * s1 + d12 * s2 * s2 + d13 * s3 * s3 + d14 * s1 * s3;
  s1 = o6 * (v[1] - 6 * v[2] + 3 * v[3] + 2 * v[4]);
  s2 = v[2] - 2 * v[3] + v[4];
  s3 = -v[1] + 3 * v[2] - 3 * v[3] + v[4];
  const double b1 = s1 * s1 + d12 * s2 * s2 + d13 * s3 * s3 + d14 * s1 * s3;
  s1 = o6 * (-2 * v[2] - 3 * v[3] + 6 * v[4] - v[5]);
  s3 = -v[2] + 3 * v[3] - 3 * v[4] + v[5];
  const double b2 = s1 * s1 + d12 * s2 * s2 + d13 * s3 * s3 + d14 * s1 * s3;
  s1 = o6 * (-11 * v[3] + 18 * v[4] - 9 * v[5] + 2 * v[6]);
  s2 = 2 * v[3] - 5 * v[4] + 4 * v[5] - v[6];
  s3 = -v[3] + 3 * v[4] - 3 * v[5] + v[6];
  const double b3 = s1 * s1 + d12 * s2 * s2 + d13 * s3 * s3 + d14 * s1 * s3;
  double tau = b0 + 3 * b1 - 3 * b2 - b3;
  tau *= tau;
  constexpr double c0 = 1.0 / 35, c1 = 12.0 / 35, c2 = 18.0 / 35, c3 = 4.0 / 35;
  // the four 1/d_k from two reciprocals of pair products (each d >= eps^2, so a product stays normal)
  const double d0 = (eps + b0) * (eps + b0), d1 = (eps + b1) * (eps + b1);
  const double d2 = (eps + b2) * (eps + b2), d3 = (eps + b3) * (eps + b3);
  const double r01 = rcp(d0 * d1), r23 = rcp(d2 * d3);
  const double a0 = c0 + c0 * tau * (d1 * r01);
  const double a1 = c1 + c1 * tau * (d0 * r01);
  const double a2 = c2 + c2 * tau * (d3 * r23);
  const double a3 = c3 + c3 * tau * (d2 * r23);
  const double q0 = -3 * v[0] + 13 * v[1] - 23 * v[2] + 25 * v[3];
  const double q1 = v[1] - 5 * v[2] + 13 * v[3] + 3 * v[4];
  const double q2 = -v[2] + 7 * v[3] + 7 * v[4] - v[5];
  const double q3 = 3 * v[3] + 13 * v[4] - 5 * v[5] + v[6];
  return (1.0 / 12) * (a0 * q0 + a1 * q1 + a2 * q2 + a3 * q3) * rcp(a0 + a1 + a2 + a3);
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
