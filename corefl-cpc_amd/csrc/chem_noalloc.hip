// Allocation-free chemistry sub-step: the batched ROK4E sub-step of corefl_chem_integrate on cell arrays the
// caller owns in device memory, with the integrator's working arrays (carried step size, step counters,
// hand-over list and state, work-queue counters) in one fixed device array of this library instead of
// memory allocated per batch.  For callers that cannot allocate: code under HIP graph capture and serving
// loops.  The array is shared by every call on a GPU, so calls that overlap on one GPU race.
#include <hip/hip_runtime.h>

#include <mutex>
#include <stdexcept>
#include <string>

#include "../../include/corefl_c.h"
#include "chem_ctx.hpp"
#include "chem_kernels.hpp"

namespace {

constexpr long long kMaxCells = 1LL << 20;
constexpr size_t kDoubles = sizeof(double) * (size_t)kMaxCells;
constexpr size_t kInts = sizeof(int) * (size_t)kMaxCells;
// Layout (byte offsets): h, tcur, elast (double per cell); steps, rejects, defer (int per cell); then the
// work-queue counter and the hand-over count.  36 MiB + 256 B.
constexpr size_t kOffH = 0, kOffTcur = kDoubles, kOffElast = 2 * kDoubles, kOffSteps = 3 * kDoubles,
                 kOffRejects = kOffSteps + kInts, kOffDefer = kOffRejects + kInts, kOffCounter = kOffDefer + kInts,
                 kOffNDefer = kOffCounter + 128, kScratchBytes = kOffCounter + 256;

static __device__ __align__(256) unsigned char scratch[kScratchBytes];

__global__ void k_clear_doubles(double *p, long n) {
  const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) p[i] = 0.0;
}

constexpr int kMaxDevices = 64;
std::mutex g_mu;
unsigned char *g_ptr[kMaxDevices] = {};

// Device address of `scratch` on the current device, looked up once per device.  The host-side name of the
// array is only a stand-in for the symbol: every kernel argument below uses this address.
unsigned char *scratch_ptr() {
  int dev = 0;
  COREFL_HIPCHK(hipGetDevice(&dev));
  if (dev < 0 || dev >= kMaxDevices) throw std::runtime_error("device index out of range");
  std::lock_guard<std::mutex> lock(g_mu);
  if (!g_ptr[dev]) {
    void *p = nullptr;
    const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess || !p)
      throw std::runtime_error(std::string("hipGetSymbolAddress(scratch): ") + hipGetErrorString(e));
    g_ptr[dev] = static_cast<unsigned char *>(p);
  }
  return g_ptr[dev];
}

} // namespace

extern "C" int corefl_chem_integrate_noalloc(corefl_chem *c, long long n, const double *rho, const double *e,
                                             double *T, double *rhoY, double dt, double rtol, double atol,
                                             int max_attempts, void *stream) {
  try {
    if (!c) throw std::runtime_error("corefl_chem_integrate_noalloc: null context");
    if (!c->mech_dev)
      throw std::runtime_error("corefl_chem_integrate_noalloc: the context has no device tables yet (they are "
                               "created by its first corefl_chem_set_cells or corefl_chem_integrate)");
    if (n < 0 || n > kMaxCells)
      throw std::runtime_error("corefl_chem_integrate_noalloc: n = " + std::to_string(n) + " outside [0, " +
                               std::to_string(kMaxCells) + "] (the fixed working array's capacity)");
    if (n > 0 && (!rho || !e || !T || !rhoY)) throw std::runtime_error("corefl_chem_integrate_noalloc: null array");
    if (!(dt > 0) || !(rtol > 0) || !(atol > 0) || max_attempts < 1)
      throw std::runtime_error("corefl_chem_integrate_noalloc: dt, rtol, atol > 0 and max_attempts >= 1 required");
    if (n == 0) return 0;
    hipStream_t s = corefl::chem_stream(c, stream);
    unsigned char *base = scratch_ptr();
    corefl::ChemBatch b{};
    b.n = (int)n;
    b.rho = const_cast<double *>(rho);
    b.e = const_cast<double *>(e);
    b.T = T;
    b.rhoY = rhoY;
    b.h = reinterpret_cast<double *>(base + kOffH);
    b.tcur = reinterpret_cast<double *>(base + kOffTcur);
    b.elast = reinterpret_cast<double *>(base + kOffElast);
    b.steps = reinterpret_cast<int *>(base + kOffSteps);
    b.rejects = reinterpret_cast<int *>(base + kOffRejects);
    b.defer = reinterpret_cast<int *>(base + kOffDefer);
    b.n_defer = reinterpret_cast<int *>(base + kOffNDefer);
    int *counter = reinterpret_cast<int *>(base + kOffCounter);
    // every cell starts from the full sub-step (a zero carried step size), as after corefl_chem_set_cells;
    // the launches clear the queue counter and the hand-over count themselves
    k_clear_doubles<<<(unsigned)((n + 255) / 256), 256, 0, s>>>(b.h, (long)n);
    COREFL_HIPCHK(hipGetLastError());
    COREFL_HIPCHK(corefl::launch_chem_integrate(c->mech_dev, c->mech.ns, b, dt, rtol, atol, max_attempts, c->cfg,
                                                counter, s));
    return 0;
  } catch (const std::exception &ex) {
    corefl::set_last_error(ex.what());
    return -1;
  }
}
