// Collocated central first derivatives along one index of the DZone layout, shared by the single-species and
// mixture viscous kernels.
#pragma once

#include <hip/hip_runtime.h>

namespace corefl {
namespace stencil {

// Collocated first derivative along one index (ViscousScheme.cu:129-307 for ORDER 8; ORDER 4 uses the
// reference's fourth-order rows: the i==2 row as the interior stencil and the i==0/1 rows as closures).
template<int ORDER>
__device__ __forceinline__ double dd(const double *f, long id, long st, int i, int n, int lb, int rb) {
#define FF(o) f[id + (o) * st]
  if (lb && i < ORDER / 2) {
    if (i == 0) return -25.0 / 12 * FF(0) + 4.0 * FF(1) - 3 * FF(2) + 4.0 / 3 * FF(3) - 0.25 * FF(4);
    if (i == 1) return -0.25 * FF(-1) - 5.0 / 6 * FF(0) + 1.5 * FF(1) - 0.5 * FF(2) + 1.0 / 12 * FF(3);
    if (ORDER == 8) {
      if (i == 2) return -1.0 / 12 * (FF(2) - FF(-2)) + 2.0 / 3 * (FF(1) - FF(-1));
      return 0.75 * (FF(1) - FF(-1)) - 0.15 * (FF(2) - FF(-2)) + 1.0 / 60 * (FF(3) - FF(-3));
    }
  }
  if (rb && i >= n - ORDER / 2) {
    if (i == n - 1) return 25.0 / 12 * FF(0) - 4.0 * FF(-1) + 3 * FF(-2) - 4.0 / 3 * FF(-3) + 0.25 * FF(-4);
    if (i == n - 2) return 0.25 * FF(1) + 5.0 / 6 * FF(0) - 1.5 * FF(-1) + 0.5 * FF(-2) - 1.0 / 12 * FF(-3);
    if (ORDER == 8) {
      if (i == n - 3) return -1.0 / 12 * (FF(2) - FF(-2)) + 2.0 / 3 * (FF(1) - FF(-1));
      return 0.75 * (FF(1) - FF(-1)) - 0.15 * (FF(2) - FF(-2)) + 1.0 / 60 * (FF(3) - FF(-3));
    }
  }
  if constexpr (ORDER == 8)
    return 0.8 * (FF(1) - FF(-1)) + -0.2 * (FF(2) - FF(-2)) + 4.0 / 105 * (FF(3) - FF(-3)) +
           -1.0 / 280 * (FF(4) - FF(-4));
  else
    return -1.0 / 12 * (FF(2) - FF(-2)) + 2.0 / 3 * (FF(1) - FF(-1));
#undef FF
}

// Interior row of dd<ORDER> on a window w[0 .. ORDER] of consecutive values centred on w[ORDER / 2].
template<int ORDER>
__device__ __forceinline__ double dd_mid(const double *w) {
  constexpr int H = ORDER / 2;
  if constexpr (ORDER == 8)
    return 0.8 * (w[H + 1] - w[H - 1]) + -0.2 * (w[H + 2] - w[H - 2]) + 4.0 / 105 * (w[H + 3] - w[H - 3]) +
           -1.0 / 280 * (w[H + 4] - w[H - 4]);
  else
    return -1.0 / 12 * (w[H + 2] - w[H - 2]) + 2.0 / 3 * (w[H + 1] - w[H - 1]);
}

// Whether dd<ORDER> at index i of n takes a one-sided boundary row.
template<int ORDER>
__device__ __forceinline__ bool dd_closure(int i, int n, int lb, int rb) {
  return (lb && i < ORDER / 2) || (rb && i >= n - ORDER / 2);
}

} // namespace stencil
} // namespace corefl
