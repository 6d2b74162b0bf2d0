// The chemistry context behind the corefl_chem_* C ABI, shared by the chemistry and surrogate entry points.
#pragma once

#include <hip/hip_runtime.h>

#include <stdexcept>
#include <string>

#include "../../include/corefl_c.h"
#include "chem.hpp"
#include "chem_kernels.hpp"

#define COREFL_HIPCHK(x)                                                                                   \
  do {                                                                                                     \
    hipError_t e_ = (x);                                                                                   \
    if (e_ != hipSuccess) throw std::runtime_error(std::string(#x) + ": " + hipGetErrorString(e_));       \
  } while (0)

struct corefl_chem {
  corefl::ChemMech mech{};
  corefl::ChemMech *mech_dev = nullptr;
  hipStream_t stream = nullptr;
  int *counter = nullptr;
  int cap = 0;
  corefl::ChemBatch b{};
  double *save = nullptr; // autotuner copy of T, rhoY, h
  double *dtc = nullptr;  // per-cell sub-step lengths (capacity cap)
  bool have_dtc = false;
  int *list = nullptr;       // cell list of the surrogate's fallback cells (capacity cap)
  unsigned short *sfeat = nullptr; // surrogate work arrays (capacity cap): BF16 features [cap][16],
  float *sout = nullptr;           // network outputs [cap][10],
  unsigned char *sflag = nullptr;  // input-range flags [cap]
  int *list_count = nullptr; // its device-side length
  double *proj_inv = nullptr; // the surrogate projection's held-species inverse table (ElementMatrix::inv_table)
  corefl::ChemLaunch cfg{1, 64, 0};
  bool tuned = false;
  ~corefl_chem();
};

namespace corefl {

void set_last_error(const std::string &msg);
// Device tables, stream and counters on first use (needs a GPU).
void chem_ensure_device(corefl_chem *c);
// Batch arrays for n cells (grown when needed; contents undefined after growth).
void chem_ensure_batch(corefl_chem *c, long long n);
hipStream_t chem_stream(corefl_chem *c, void *stream);
void chem_check_args(corefl_chem *c, double dt, double rtol, double atol, int max_attempts);
// Surrogate sub-step of batch b (the chemistry context's cells, possibly with dt_dev / dtc) with the
// exact ROK4E fallback for gated cells, enqueued on st (implemented with the surrogate kernels).
void surrogate_substep(corefl_chem *c, corefl_surrogate *s, const ChemBatch &b, double rtol, double atol,
                       int max_attempts, hipStream_t st);
// Exact sub-step of every cell, or of list[0 .. *n_list) when list is given.
void chem_integrate(corefl_chem *c, const ChemLaunch &cfg, double dt, double rtol, double atol, int max_attempts,
                    hipStream_t s, const int *list, const int *n_list);

} // namespace corefl
