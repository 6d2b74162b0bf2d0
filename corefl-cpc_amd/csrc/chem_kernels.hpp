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
  // Hand-over from the one-lane work queue to the cooperative kernel (launch mode 3): the cells that used up
  // their attempt budget, and their integration state (rhoY, T, h, steps and rejects stay in the arrays
  // above).
  int *defer;    // deferred cells (capacity n)
  int *n_defer;  // device-side count
  double *tcur;  // time reached within the sub-step
  double *elast; // error of the last accepted step (step-size controller memory)
};

// Sub-step length of cell c: per-cell, device-side, or the launch's value.
__device__ __forceinline__ double cell_dt(const ChemBatch &b, long c, double dt) {
  return b.dtc ? b.dtc[c] : (b.dt_dev ? b.dt_scale * b.dt_dev[0] : dt);
}

// Launch shape picked by the autotuner.
struct ChemLaunch {
  int queue;   // 0: one lane per cell; 1: persistent work queue, one lane per cell; 2: persistent work queue,
               // 16 lanes per cell (cooperative); 3: mode 1 for at most `budget` attempts per cell, the
               // unfinished cells then continued by mode 2
  int threads; // workgroup size (multiple of 64)
  int blocks;  // persistent workgroups (modes 1-3)
  int budget;  // attempt budget of mode 3
};

// e = sum_k Y_k h_k(T) - R T from T and rho Y (ConstVolumeReactor.cpp:60-68).
hipError_t launch_chem_energy(const ChemMech *mech_dev, int ns, const ChemBatch &b, hipStream_t s);
// One chemistry sub-step of length dt for every cell.  counter: device int used by the queue launches
// (cleared by the launch itself).  Every mode gives the same result to rounding: mode 2 distributes the same
// operations over lanes in the same order.
hipError_t launch_chem_integrate(const ChemMech *mech_dev, int ns, const ChemBatch &b, double dt, double rtol,
                                 double atol, int max_attempts, const ChemLaunch &cfg, int *counter,
                                 hipStream_t s);
hipError_t launch_chem_omega(const ChemMech *mech_dev, int ns, const ChemBatch &b, double *omega, hipStream_t s);
hipError_t launch_chem_temperature(const ChemMech *mech_dev, int ns, const ChemBatch &b, hipStream_t s);
// Cooperative integration (chem_coop.hip): work items of b (its list or all cells), fresh or resumed from the
// hand-over state.
hipError_t launch_chem_coop(const ChemMech *mech_dev, int ns, const ChemBatch &b, double dt, double rtol,
                            double atol, int max_attempts, int threads, int blocks, bool resume, int *counter,
                            hipStream_t s);

} // namespace corefl
