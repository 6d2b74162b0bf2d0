// Step monitor and unphysical-value repair on the device (monitor_kernels.hip): the residual of
// compute_residual (Residual.cu:8-121) as a device-side reduction with a NaN flag, and limit_flow
// (TimeAdvanceFunc.cuh:117-220).
#pragma once

#include <hip/hip_runtime.h>

#include "flow.hpp"
#include "flow_mix.hpp"

namespace corefl {

// Device-side result of the last monitored step.
struct MonitorState {
  double res[4];   // RMS change of rho, |V|, p, T over the monitored step, divided by res_scale
  double scale[4]; // res_scale: the first monitored step's RMS values, at least 1e-20 (Residual.cu:70-78)
  double err_max;  // the largest of res (Residual.cu:92-98)
  int have_scale;
  int nan;   // a residual is NaN (Residual.cu:100-105)
  int count; // monitored steps so far
};

// limit_flow bounds of rho, u, v, w, p (DParameter.cu:332-349).
struct FlowLimits {
  double ll[5], ul[5];
};

// store_last_step (TimeAdvanceFunc.cu:6-19): rho, |V|, p, T of the interior nodes into last (4 * n).
hipError_t launch_store_last(const DevBlock &b, double *last, hipStream_t s);
// compute_square_of_dbv + the block sums of reduction_of_dv_squared (TimeAdvanceFunc.cu:84-100,
// Residual.cu:123-170): per-workgroup sums of the squared changes against last, 4 values per workgroup
// into part[4 * w + l], w < residual_partial_count(b).
int residual_partial_count(const DevBlock &b);
hipError_t launch_residual_partial(const DevBlock &b, const double *last, double *part, hipStream_t s);
// Sum of all blocks' partials, RMS over n_total nodes, the scale of the first call, normalisation, NaN flag.
hipError_t launch_residual_finish(const double *part, long n_part, double n_total, MonitorState *st, hipStream_t s);
// limit_flow: flag every interior node whose rho, u, v, w or p is NaN or outside the bounds, then set each flagged
// node to the mean of its unflagged neighbours in the 3x3x3 cube inside the block (or its clipped values if it has
// none), T from the equation of state and cv from bv.  x.ns = 0: single-species air.
hipError_t launch_limit_flow(const DevBlock &b, const MixArrays &x, const FlowLimits &lim, unsigned char *flag,
                             hipStream_t s);

} // namespace corefl
