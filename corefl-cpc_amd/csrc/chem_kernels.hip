// Batched finite-rate chemistry of constant-volume cells on CDNA4: NASA-7 thermo, Arrhenius / third-body /
// Lindemann / Troe kinetics with Gibbs-energy reverse rates, the analytic d(omega)/d(rho Y) Jacobian, a
// partial-pivot LU, and the adaptive 4-stage Rosenbrock ROK4E with embedded error control.  FP64.
//
// Two launch shapes integrate a batch of cells over one chemistry sub-step:
//   * static: one thread per cell;
//   * queue: persistent waves that fetch cells from a global counter as lanes finish, one ballot and one
//     atomic per wave per fetch round, so a stiff cell holds only its own lane, not a whole fixed tile.
// Every cell stops after max_attempts step attempts (flagged), so each wave's loop ends.
#include "chem_kernels.hpp"

#include <cmath>

#include "chem_device.hpp"
#include "rok4e.hpp"

namespace corefl {

namespace {

using namespace chemdev;
using namespace rok4e;

using chemdev::lu;
using chemdev::lu_solve;

// Integration state of one cell.
template<int NS> struct Cell {
  double y[NS]; // rho Y
  double rho, E, T, time, h, err_last, dt;
  int steps, rejects, attempts;
};

// One ROK4E attempt (ConstVolumeReactor.cpp:173-476 with adaptive_time_step = 1).
template<int NS>
__device__ __forceinline__ void rok4e_attempt(const ChemMech &m, Cell<NS> &s, double rtol, double atol) {
  const double dt = s.dt;
  double w[NS], J[NS * NS], k1[NS], k2[NS], k3[NS], k4[NS], ys[NS], yk[NS];
  int piv[NS];
  source<NS, true>(m, s.T, s.y, w, J);
  // The step ends the interval when it would leave less than 5% of itself (no sliver step, whose size
  // would otherwise become the proposal carried to the next sub-step).
  const double rem = dt - s.time;
  const double hs = s.h * kStretch >= rem ? rem : s.h;
#pragma unroll
  for (int q = 0; q < NS * NS; ++q) J[q] = -hs * G * J[q];
#pragma unroll
  for (int l = 0; l < NS; ++l) J[l * NS + l] += 1.0;
  lu<NS>(J, piv);
#pragma unroll
  for (int l = 0; l < NS; ++l) k1[l] = w[l];
  lu_solve<NS>(J, piv, k1);
  const double irho = 1.0 / s.rho; // products by 1/rho stand for the reference's divisions
#pragma unroll
  for (int l = 0; l < NS; ++l) {
    ys[l] = s.y[l] + hs * A21 * k1[l];
    yk[l] = ys[l] * irho;
  }
  double Ts = update_t<NS>(m, s.T, yk, s.E);
  source<NS, false>(m, Ts, ys, w, nullptr);
#pragma unroll
  for (int l = 0; l < NS; ++l) k2[l] = w[l] + G21 * k1[l];
  lu_solve<NS>(J, piv, k2);
#pragma unroll
  for (int l = 0; l < NS; ++l) {
    k2[l] -= G21 * k1[l];
    ys[l] = s.y[l] + hs * (A31 * k1[l] + A32 * k2[l]);
    yk[l] = ys[l] * irho;
  }
  Ts = update_t<NS>(m, Ts, yk, s.E);
  source<NS, false>(m, Ts, ys, w, nullptr);
#pragma unroll
  for (int l = 0; l < NS; ++l) k3[l] = w[l] + G31 * k1[l] + G32 * k2[l];
  lu_solve<NS>(J, piv, k3);
#pragma unroll
  for (int l = 0; l < NS; ++l) {
    k3[l] -= G31 * k1[l] + G32 * k2[l];
    k4[l] = w[l] + G41 * k1[l] + G42 * k2[l] + G43 * k3[l];
  }
  lu_solve<NS>(J, piv, k4);
  double err = 0, ynew[NS];
#pragma unroll
  for (int l = 0; l < NS; ++l) {
    k4[l] -= G41 * k1[l] + G42 * k2[l] + G43 * k3[l];
    ynew[l] = s.y[l] + hs * (B1 * k1[l] + B2 * k2[l] + B4 * k4[l]);
    const double dy = hs * (E1 * k1[l] + E2 * k2[l] + E3 * k3[l] + E4 * k4[l]) / (rtol * ynew[l] + atol);
    err += dy * dy;
  }
  err = sqrt(err / NS);
  double hstar = hs;
  if (s.steps > 0) hstar *= fmin(5.0, fmax(0.2, 0.8 * pow(s.err_last, 0.1) * pow(err + 1e-20, -0.175)));
  else if (!(err <= 1)) hstar *= 0.2; // a NaN first-step error also cuts the step (the reference would repeat it)
  // a step cut short by the end of the interval does not shrink the proposal
  if (err <= 1 && hs < s.h) hstar = fmax(hstar, s.h);
  s.h = hstar;
  ++s.attempts;
  if (err <= 1) {
#pragma unroll
    for (int l = 0; l < NS; ++l) {
      s.y[l] = ynew[l];
      yk[l] = ynew[l] * irho;
    }
    s.time += hs;
    s.err_last = fmax(err, 1e-30); // an exactly zero error would stall the PI controller (0^0.1 = 0)
    s.T = update_t<NS>(m, s.T, yk, s.E);
    ++s.steps;
  } else {
    ++s.rejects;
  }
}

template<int NS> __device__ __forceinline__ void load_cell(const ChemBatch &b, int c, double dt, Cell<NS> &s) {
#pragma unroll
  for (int l = 0; l < NS; ++l) s.y[l] = b.rhoY[(long)l * b.n + c];
  dt = cell_dt(b, c, dt);
  s.dt = dt;
  s.rho = b.rho[c];
  s.E = b.e[c];
  s.T = b.T[c];
  s.h = b.h[c] > 0 ? fmin(b.h[c], dt) : dt;
  s.time = 0;
  s.err_last = 0;
  s.steps = s.rejects = s.attempts = 0;
}

template<int NS> __device__ __forceinline__ void store_cell(const ChemBatch &b, int c, const Cell<NS> &s, bool ok) {
#pragma unroll
  for (int l = 0; l < NS; ++l) b.rhoY[(long)l * b.n + c] = s.y[l];
  b.T[c] = s.T;
  b.h[c] = s.h;
  b.steps[c] = ok ? s.steps : -1;
  b.rejects[c] = s.rejects;
}

// Work item w of a launch: cell w, or cell list[w] of a device-side list of *n_list cells.
__device__ __forceinline__ int work_count(const ChemBatch &b) { return b.list ? *b.n_list : b.n; }
__device__ __forceinline__ int work_cell(const ChemBatch &b, int w) { return b.list ? b.list[w] : w; }

template<int NS>
__global__ __launch_bounds__(256) void k_chem_static(const ChemMech *__restrict__ mp, ChemBatch b, double dt,
                                                     double rtol, double atol, int max_attempts) {
  const int w = blockIdx.x * blockDim.x + threadIdx.x;
  if (w >= work_count(b)) return;
  const int c = work_cell(b, w);
  const ChemMech &m = *mp;
  Cell<NS> s;
  load_cell<NS>(b, c, dt, s);
  while (s.time < s.dt && s.attempts < max_attempts) rok4e_attempt<NS>(m, s, rtol, atol);
  store_cell<NS>(b, c, s, s.time >= s.dt);
}

// Integration state of a cell that used up its attempt budget, for the cooperative kernel to resume.
template<int NS> __device__ __forceinline__ void store_partial(const ChemBatch &b, int c, const Cell<NS> &s) {
#pragma unroll
  for (int l = 0; l < NS; ++l) b.rhoY[(long)l * b.n + c] = s.y[l];
  b.T[c] = s.T;
  b.h[c] = s.h;
  b.steps[c] = s.steps;
  b.rejects[c] = s.rejects;
  b.tcur[c] = s.time;
  b.elast[c] = s.err_last;
}

// budget > 0: a cell still unfinished after `budget` attempts is handed over (appended to b.defer) instead
// of holding its lane for the rest of its chain.
template<int NS>
__global__ __launch_bounds__(256) void k_chem_queue(const ChemMech *__restrict__ mp, ChemBatch b, double dt,
                                                    double rtol, double atol, int max_attempts, int *counter,
                                                    int budget) {
  const ChemMech &m = *mp;
  const int lane = __lane_id();
  const int n_work = work_count(b);
  Cell<NS> s;
  int cell = -1;
  bool exhausted = false;
  while (true) {
    const bool need = cell < 0 && !exhausted;
    const unsigned long long mask = __ballot(need);
    if (mask) {
      const int leader = __ffsll((long long)mask) - 1;
      int base = 0;
      if (lane == leader) base = atomicAdd(counter, __popcll(mask));
      base = __shfl(base, leader, 64);
      if (need) {
        const int got = base + __popcll(mask & ((1ull << lane) - 1));
        if (got < n_work) {
          cell = work_cell(b, got);
          load_cell<NS>(b, cell, dt, s);
        } else {
          exhausted = true;
        }
      }
    }
    if (__ballot(cell >= 0) == 0) break;
    bool defer = false;
    if (cell >= 0) {
      rok4e_attempt<NS>(m, s, rtol, atol);
      if (s.time >= s.dt || s.attempts >= max_attempts) {
        store_cell<NS>(b, cell, s, s.time >= s.dt);
        cell = -1;
      } else {
        defer = budget > 0 && s.attempts >= budget;
      }
    }
    const unsigned long long dm = __ballot(defer);
    if (dm) {
      const int leader = __ffsll((long long)dm) - 1;
      int base = 0;
      if (lane == leader) base = atomicAdd(b.n_defer, __popcll(dm));
      base = __shfl(base, leader, 64);
      if (defer) {
        b.defer[base + __popcll(dm & ((1ull << lane) - 1))] = cell;
        store_partial<NS>(b, cell, s);
        cell = -1;
      }
    }
  }
}

template<int NS>
__global__ __launch_bounds__(256) void k_chem_omega(const ChemMech *__restrict__ mp, ChemBatch b, double *omega) {
  const int c = blockIdx.x * blockDim.x + threadIdx.x;
  if (c >= b.n) return;
  double y[NS], w[NS];
#pragma unroll
  for (int l = 0; l < NS; ++l) y[l] = b.rhoY[(long)l * b.n + c];
  source<NS, false>(*mp, b.T[c], y, w, nullptr);
#pragma unroll
  for (int l = 0; l < NS; ++l) omega[(long)l * b.n + c] = w[l];
}

template<int NS>
__global__ __launch_bounds__(256) void k_chem_temperature(const ChemMech *__restrict__ mp, ChemBatch b) {
  const int c = blockIdx.x * blockDim.x + threadIdx.x;
  if (c >= b.n) return;
  double yk[NS];
  const double irho = 1.0 / b.rho[c];
#pragma unroll
  for (int l = 0; l < NS; ++l) yk[l] = b.rhoY[(long)l * b.n + c] * irho;
  b.T[c] = update_t<NS>(*mp, b.T[c], yk, b.e[c]);
}

template<int NS> __global__ __launch_bounds__(256) void k_chem_energy(const ChemMech *__restrict__ mp, ChemBatch b) {
  const int c = blockIdx.x * blockDim.x + threadIdx.x;
  if (c >= b.n) return;
  const ChemMech &m = *mp;
  double h[NS], cp[NS];
  const double t = b.T[c], irho = 1.0 / b.rho[c];
  h_cp<NS>(m, t, h, cp);
  double e = 0, imw = 0;
#pragma unroll
  for (int l = 0; l < NS; ++l) {
    const double y = b.rhoY[(long)l * b.n + c] * irho;
    e += h[l] * y;
    imw += y * m.imw[l];
  }
  b.e[c] = e - kRu * imw * t;
}

__global__ void k_zero_int(int *p) { *p = 0; }

} // namespace

hipError_t launch_chem_energy(const ChemMech *mech_dev, int ns, const ChemBatch &b, hipStream_t s) {
  if (ns != 9) return hipErrorNotSupported;
  if (b.n == 0) return hipSuccess;
  k_chem_energy<9><<<(b.n + 255) / 256, 256, 0, s>>>(mech_dev, b);
  return hipGetLastError();
}

hipError_t launch_chem_integrate(const ChemMech *mech_dev, int ns, const ChemBatch &b, double dt, double rtol,
                                 double atol, int max_attempts, const ChemLaunch &cfg, int *counter,
                                 hipStream_t s) {
  if (ns != 9) return hipErrorNotSupported;
  if (b.n == 0) return hipSuccess;
  if (cfg.queue == 2) return launch_chem_coop(mech_dev, ns, b, dt, rtol, atol, max_attempts, cfg.threads,
                                              cfg.blocks, false, counter, s);
  if (cfg.queue == 3) {
    if (cfg.budget < 1) return hipErrorInvalidValue;
    k_zero_int<<<1, 1, 0, s>>>(counter);
    k_zero_int<<<1, 1, 0, s>>>(b.n_defer);
    k_chem_queue<9><<<cfg.blocks, cfg.threads, 0, s>>>(mech_dev, b, dt, rtol, atol, max_attempts, counter,
                                                       cfg.budget);
    ChemBatch d = b;
    d.list = b.defer;
    d.n_list = b.n_defer;
    return launch_chem_coop(mech_dev, ns, d, dt, rtol, atol, max_attempts, 256, cfg.blocks, true, counter, s);
  }
  if (cfg.queue) {
    k_zero_int<<<1, 1, 0, s>>>(counter);
    k_chem_queue<9><<<cfg.blocks, cfg.threads, 0, s>>>(mech_dev, b, dt, rtol, atol, max_attempts, counter, 0);
  } else {
    k_chem_static<9><<<(b.n + cfg.threads - 1) / cfg.threads, cfg.threads, 0, s>>>(mech_dev, b, dt, rtol, atol,
                                                                                  max_attempts);
  }
  return hipGetLastError();
}

hipError_t launch_chem_omega(const ChemMech *mech_dev, int ns, const ChemBatch &b, double *omega, hipStream_t s) {
  if (ns != 9) return hipErrorNotSupported;
  if (b.n == 0) return hipSuccess;
  k_chem_omega<9><<<(b.n + 255) / 256, 256, 0, s>>>(mech_dev, b, omega);
  return hipGetLastError();
}

hipError_t launch_chem_temperature(const ChemMech *mech_dev, int ns, const ChemBatch &b, hipStream_t s) {
  if (ns != 9) return hipErrorNotSupported;
  if (b.n == 0) return hipSuccess;
  k_chem_temperature<9><<<(b.n + 255) / 256, 256, 0, s>>>(mech_dev, b);
  return hipGetLastError();
}

} // namespace corefl
