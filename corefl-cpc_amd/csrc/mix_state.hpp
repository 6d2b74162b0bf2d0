// Conservative variables and acoustic speed of one mixture node from its primitive state, shared by the mixture
// flow, boundary and repair kernels: compute_cv_from_bv_1_point<Mixture> + compute_total_energy
// (FieldOperation.cuh:16-84) and the acoustic speed of update_physical_properties<Mixture>
// (FieldOperation.cuh:86-120).
#pragma once

#include <hip/hip_runtime.h>

#include <cmath>

#include "chem_device.hpp"
#include "flow.hpp"
#include "flow_mix.hpp"

namespace corefl {
namespace mixdev {

constexpr int NS = 9;

__device__ inline void mix_cons(const DevBlock &b, const MixArrays &x, long id) {
  const long sz = b.sz;
  const ChemMech &m = *x.mech;
  const double rho = b.bv[id], u = b.bv[id + sz], v = b.bv[id + 2 * sz], w = b.bv[id + 3 * sz];
  const double p = b.bv[id + 4 * sz], T = b.bv[id + 5 * sz];
  double h[NS], cp[NS];
  chemdev::h_cp<NS>(m, T, h, cp);
  double E = 0.5 * (u * u + v * v + w * w), imw = 0, cpt = 0;
#pragma unroll
  for (int l = 0; l < NS; ++l) {
    const double y = x.sv[id + l * sz];
    E += h[l] * y;
    imw += y * m.imw[l];
    cpt += y * cp[l];
    b.cv[id + (5 + l) * sz] = rho * y;
  }
  b.cv[id] = rho;
  b.cv[id + sz] = rho * u;
  b.cv[id + 2 * sz] = rho * v;
  b.cv[id + 3 * sz] = rho * w;
  b.cv[id + 4 * sz] = E * rho - p;
  const double R = chemdev::kRu * imw;
  x.snd[id] = sqrt(cpt / (cpt - R) * R * T);
}

} // namespace mixdev
} // namespace corefl
