// Launchers of the physical boundary faces beyond outflow and periodic, and of the sponge layer's running mean
// (bc_kernels.hip).  MixArrays with ns = 0 selects single-species air.
#pragma once

#include <hip/hip_runtime.h>

#include "flow.hpp"
#include "flow_mix.hpp"

namespace corefl {

// apply_inflow (BoundCond.cu:2190-2705): boundary node and ghost layers of an inflow face; counts the application
// when the face draws noise.
hipError_t launch_inflow(const DevBlock &b, const MixArrays &x, const DevInflow &f, hipStream_t s);
// apply_symmetry (BoundCond.cu:2002-2063).
hipError_t launch_symmetry(const DevBlock &b, const MixArrays &x, const DevFace &f, hipStream_t s);
// Running mean of cv over the interior (nv components into mean[l * n + c]) with the step count in *iter;
// count: increment *iter afterwards (once per step, after the last block).
hipError_t launch_sponge_mean(const DevBlock &b, int nv, double *mean, unsigned long long *iter, bool count,
                              hipStream_t s);

} // namespace corefl
