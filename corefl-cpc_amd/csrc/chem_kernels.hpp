// Launchers of the batched chemistry kernels (chem_kernels.hip).
#pragma once

#include <hip/hip_runtime.h>

#include "chem.hpp"

namespace corefl {

// A batch of n constant-volume cells in device memory, structure of arrays.
struct ChemBatch {
  int n;
  double *rho;   // density, kg/m^3
  double *e;     // specific internal energy, J/kg (held constant over a sub-step)
  double *T;     // temperature, K (in: initial guess; out: final)
  double *rhoY;  // partial densities, [l * n + cell]
  double *h;     // per-cell step size carried between sub-steps (<= 0: start from dt)
  int *steps;    // accepted steps of the last sub-step (-1: attempt limit reached)
  int *rejects;  // rejected attempts of the last sub-step
  const double *dtc; // per-cell sub-step length (nullptr: the launch's dt for every cell)
  const int *list;   // integrate only cells list[0 .. *n_list) (nullptr: all n cells)
  const int *n_list; // device-side count, at most n
  const double *dt_dev; // when dtc is null: sub-step length dt_scale * dt_dev[0] (device-side step size)
  double dt_scale;
};

// Sub-step length of cell c: per-cell, device-side, or the launch's value.
__device__ __forceinline__ double cell_dt(const ChemBatch &b, long c, double dt) {
  return b.dtc ? b.dtc[c] : (b.dt_dev ? b.dt_scale * b.dt_dev[0] : dt);
}

// Launch shape picked by the autotuner.
struct ChemLaunch {
  int queue;   // 1: persistent work queue, 0: one thread per cell
  int threads; // workgroup size (multiple of 64)
  int blocks;  // persistent workgroups (queue only)
};

// e = sum_k Y_k h_k(T) - R T from T and rho Y (ConstVolumeReactor.cpp:60-68).
hipError_t launch_chem_energy(const ChemMech *mech_dev, int ns, const ChemBatch &b, hipStream_t s);
// One chemistry sub-step of length dt for every cell.  counter: device int used by the queue launch
// (cleared by the launch itself).
hipError_t launch_chem_integrate(const ChemMech *mech_dev, int ns, const ChemBatch &b, double dt, double rtol,
                                 double atol, int max_attempts, const ChemLaunch &cfg, int *counter,
                                 hipStream_t s);
hipError_t launch_chem_omega(const ChemMech *mech_dev, int ns, const ChemBatch &b, double *omega, hipStream_t s);
hipError_t launch_chem_temperature(const ChemMech *mech_dev, int ns, const ChemBatch &b, hipStream_t s);

} // namespace corefl
