// Cooperative batched chemistry: one cell per 16-lane row of a wavefront (four cells per wave), for the cells
// whose long chains of ROK4E attempts set the latency of a sub-step.
//
// The one-lane kernels keep a cell's whole state (the 9 x 9 iteration matrix, four stage vectors, the
// mechanism's working arrays) in one lane: about 500 registers, one wave per SIMD, and every instruction of
// an attempt issued one after another.  Here lane a of a row owns species a (its partial density, stage
// values and row a of the matrix) and reactions a, a + 16, ...: the reaction rates of a source evaluation
// are computed side by side, each species row accumulates its production terms and Jacobian row, and the LU
// factorisation and the solves run row-parallel.  Values are exchanged within the row by DPP broadcasts
// (row_newbcast, a constant source lane), by ds_bpermute for the data-dependent pivot row, and through a
// small per-row LDS area for the rates and concentrations that the reaction lanes read at per-lane indices.
//
// Every floating-point operation of the one-lane attempt (chem_kernels.hip) is performed here with the same
// operands in the same order: the right-looking row-parallel elimination applies to each matrix entry the
// same updates, in the same order, as the one-lane Doolittle form, and every sum over species is formed in
// ascending species order from broadcast copies.  The two kernels therefore agree to the last bit where the
// compiler contracts the same products into FMAs.
#include <hip/hip_runtime.h>

#include <cmath>
#include <type_traits>

#include "chem_device.hpp"
#include "chem_kernels.hpp"
#include "rok4e.hpp"

namespace corefl {

namespace {

using namespace chemdev;
using namespace rok4e;

constexpr int RW = 16; // lanes per cell (one DPP row)

// value of lane K of this lane's 16-lane row (DPP row_newbcast)
template<int K> __device__ __forceinline__ double row_bcast_c(double x) {
  int2 v = __builtin_bit_cast(int2, x);
  v.x = __builtin_amdgcn_update_dpp(0, v.x, 0x150 + K, 0xF, 0xF, false);
  v.y = __builtin_amdgcn_update_dpp(0, v.y, 0x150 + K, 0xF, 0xF, false);
  return __builtin_bit_cast(double, v);
}

// f(std::integral_constant<int, I>) for I = B .. E-1, unrolled at compile time: every DPP control operand
// is then a constant (a loop of broadcasts would be unrolled only after the compiler's size estimate of it)
template<int B, int E, class F> __device__ __forceinline__ void static_for(F &&f) {
  if constexpr (B < E) {
    f(std::integral_constant<int, B>{});
    static_for<B + 1, E>(f);
  }
}

// all[l] = lane l's x, l < NS
template<int NS> __device__ __forceinline__ void row_gather(double x, double *all) {
  static_for<0, NS>([&](auto L) { all[decltype(L)::value] = row_bcast_c<decltype(L)::value>(x); });
}

// Orders the LDS traffic of one wavefront: stores before this point are seen by loads after it.
__device__ __forceinline__ void wave_sync() {
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

// Species thermo tables in LDS, [field][species], read at per-lane species indices.
struct SpeciesLds {
  double lo[7][kMaxSpec], hi[7][kMaxSpec];
  double t_low[kMaxSpec], t_mid[kMaxSpec], t_high[kMaxSpec];
  double ln_t_low[kMaxSpec], ln_t_high[kMaxSpec], inv_t_low[kMaxSpec], inv_t_high[kMaxSpec];
  double mw[kMaxSpec], imw[kMaxSpec];
};

// Per-row exchange area: concentrations and g/RT written by the species lanes, rates of progress written by
// the reaction lanes.
struct RowLds {
  double c[kMaxSpec], g[kMaxSpec];
  double q1[kMaxReac], q2[kMaxReac];
};

// h_a and cp_a of species a (the one-lane h_cp, one species).
__device__ __forceinline__ void own_h_cp(const SpeciesLds &S, int a, double t, double &h, double &cp) {
  const double R = kRu * S.imw[a];
  const bool below = t < S.t_low[a], above = t > S.t_high[a];
  if (below || above) {
    const double tt = below ? S.t_low[a] : S.t_high[a];
    double c[5];
#pragma unroll
    for (int q = 0; q < 5; ++q) c[q] = below ? S.lo[q][a] : S.hi[q][a];
    const double c5 = below ? S.lo[5][a] : S.hi[5][a];
    const double tt2 = tt * tt, tt3 = tt2 * tt, tt4 = tt3 * tt, tt5 = tt4 * tt;
    const double cc = c[0] + c[1] * tt + c[2] * tt2 + c[3] * tt3 + c[4] * tt4;
    const double hh = c[0] * tt + 0.5 * c[1] * tt2 + c[2] * tt3 * kThird + 0.25 * c[3] * tt4 + 0.2 * c[4] * tt5 + c5;
    h = (hh + cc * (t - tt)) * R;
    cp = cc * R;
  } else {
    const bool use_lo = t < S.t_mid[a];
    double c[6];
#pragma unroll
    for (int q = 0; q < 6; ++q) c[q] = use_lo ? S.lo[q][a] : S.hi[q][a];
    const double t2 = t * t, t3 = t2 * t, t4 = t3 * t, t5 = t4 * t;
    h = (c[0] * t + 0.5 * c[1] * t2 + c[2] * t3 * kThird + 0.25 * c[3] * t4 + 0.2 * c[4] * t5 + c[5]) * R;
    cp = (c[0] + c[1] * t + c[2] * t2 + c[3] * t3 + c[4] * t4) * R;
  }
}

// g_a / (R T) of species a (the one-lane gibbs, one species).
__device__ __forceinline__ double own_gibbs(const SpeciesLds &S, int a, double t, double lnT, double iT) {
  const bool below = t < S.t_low[a], above = t > S.t_high[a];
  const double tt = below ? S.t_low[a] : (above ? S.t_high[a] : t);
  const bool use_lo = below || (!above && t < S.t_mid[a]);
  double c[7];
#pragma unroll
  for (int q = 0; q < 7; ++q) c[q] = use_lo ? S.lo[q][a] : S.hi[q][a];
  const double tt2 = tt * tt, tt3 = tt2 * tt, tt4 = tt3 * tt;
  const double lt = below ? S.ln_t_low[a] : (above ? S.ln_t_high[a] : lnT);
  const double itt = below ? S.inv_t_low[a] : (above ? S.inv_t_high[a] : iT);
  return c[0] * (1.0 - lt) - 0.5 * c[1] * tt - c[2] * tt2 * kSixth - c[3] * tt3 * kTwelfth - c[4] * tt4 * 0.05 +
         c[5] * itt - c[6];
}

// Newton temperature at fixed E (the one-lane update_t): yk = this lane's mass fraction (0 beyond NS).
template<int NS>
__device__ __forceinline__ double row_update_t(const ChemMech &m, const SpeciesLds &S, int a, double t, double yk,
                                               double E) {
  double y[NS];
  row_gather<NS>(yk, y);
  double imw = 0;
#pragma unroll
  for (int l = 0; l < NS; ++l) imw += y[l] * m.imw[l];
  const double R = kRu * imw;
  double err = 1e+6;
  for (int it = 0; it < 100 && err > 1e-8; ++it) {
    double h, cp;
    own_h_cp(S, a, t, h, cp);
    double hs[NS], cps[NS];
    row_gather<NS>(h, hs);
    row_gather<NS>(cp, cps);
    double cpt = 0, hh = 0;
#pragma unroll
    for (int l = 0; l < NS; ++l) {
      cpt += cps[l] * y[l];
      hh += hs[l] * y[l];
    }
    const double t1 = t - (hh - R * t - E) / (cpt - R);
    err = fabs(1 - t1 / t);
    t = t1;
  }
  return t;
}

// Rates of progress of reaction r at the row's state (the reaction body of the one-lane source).
__device__ __forceinline__ void reaction_rates(const ChemMech &m, const RowLds &X, int r, int ns, double t,
                                               double lnT, double iT, double tp, double &q1o, double &q2o) {
  const int ty = m.type[r];
  const int nfs = m.nfs[r], nbs = m.nbs[r];
  double kf = arrhenius(lnT, iT, m.A[r], m.b[r], m.Ea[r]);
  double cc = 0;
  if (ty >= 4) {
    for (int l = 0; l < ns; ++l) cc += X.c[l] * m.tb[r][l];
  }
  if (ty == 3) {
    kf += arrhenius(lnT, iT, m.A2[r], m.b2[r], m.Ea2[r]);
  } else if (ty == 4) {
    kf *= cc;
  } else if (ty > 4) {
    const double klow = arrhenius(lnT, iT, m.A2[r], m.b2[r], m.Ea2[r]), khigh = kf;
    if (khigh < 1e-25 && klow < 1e-25) {
      kf = 0;
    } else {
      const double pr = klow * cc / khigh;
      double F = 1.0;
      if (ty > 5) {
        double fc = (1 - m.troe[r][0]) * exp(-t / m.troe[r][1]) + m.troe[r][0] * exp(-t / m.troe[r][2]);
        if (ty == 7) fc += exp(-m.troe[r][3] / t);
        const double lfc = log10(fc);
        const double cq = -0.4 - 0.67 * lfc, nq = 0.75 - 1.27 * lfc;
        const double lpr = log10(pr);
        const double tmp = (lpr + cq) / (nq - 0.14 * (lpr + cq));
        F = exp(2.302585092994046 * lfc / (1.0 + tmp * tmp)); // 10^x
      }
      kf = khigh * pr / (1.0 + pr) * F;
    }
  }
  double kb = 0;
  if (ty != 0 && m.rev[r] == 1) {
    kb = arrhenius(lnT, iT, m.A2[r], m.b2[r], m.Ea2[r]);
    if (ty == 4) kb *= cc;
  }
  if (m.need_gibbs && ty != 0 && ty != 2) {
    double dg = 0;
    for (int s = 0; s < nbs; ++s) dg += X.g[m.bsp[r][s]] * m.bnu[r][s];
    for (int s = 0; s < nfs; ++s) dg -= X.g[m.fsp[r][s]] * m.fnu[r][s];
    kb = kf / (int_pow(tp, m.iorder[r]) * exp(-dg));
  }
  double q1 = 1.0, q2 = ty != 0 ? 1.0 : 0.0;
  for (int s = 0; s < nfs; ++s) q1 *= int_pow(X.c[m.fsp[r][s]], m.fnu[r][s]);
  if (ty != 0)
    for (int s = 0; s < nbs; ++s) q2 *= int_pow(X.c[m.bsp[r][s]], m.bnu[r][s]);
  q1o = q1 * kf;
  q2o = q2 * kb;
}

// Chemical source (and with JAC the Jacobian row a) of the row's cell: lane a returns omega_a; yall holds
// every species' rho Y (replicated).
template<int NS, bool JAC>
__device__ __forceinline__ double row_source(const ChemMech &m, const SpeciesLds &S, RowLds &X, int a, int gl,
                                             double t, double ya, const double *yall, double *Jrow) {
  const double lnT = log(t), iT = 1.0 / t;
  const double tp = kPatm / kRu * 1e-3 / t;
  X.c[gl] = ya * S.imw[a] * 1e-3;
  if (m.need_gibbs) X.g[gl] = own_gibbs(S, a, t, lnT, iT);
  wave_sync();
  const int nr = m.nr;
  for (int r0 = 0; r0 < nr; r0 += RW) {
    const int r = r0 + gl;
    if (r < nr) {
      double q1, q2;
      reaction_rates(m, X, r, NS, t, lnT, iT, tp, q1, q2);
      X.q1[r] = q1;
      X.q2[r] = q2;
    }
  }
  wave_sync();
  double cr = 0, de = 0;
  if constexpr (JAC) {
#pragma unroll
    for (int n = 0; n < NS; ++n) Jrow[n] = 0;
  }
  const unsigned bit = 1u << a;
  for (int r = 0; r < nr; ++r) {
    const double q1 = X.q1[r], q2 = X.q2[r];
    const int nfs = m.nfs[r], nbs = m.nbs[r];
    for (int s = 0; s < nfs; ++s) {
      if (m.fsp[r][s] == a) {
        const double nu = m.fnu[r][s];
        cr += q2 * nu;
        de += q1 * nu;
      }
    }
    for (int s = 0; s < nbs; ++s) {
      if (m.bsp[r][s] == a) {
        const double nu = m.bnu[r][s];
        cr += q1 * nu;
        de += q2 * nu;
      }
    }
    if constexpr (JAC) {
      const unsigned um = m.umask[r];
      if (m.amask[r] & bit) {
        const double dn = m.dnu[r][a];
#pragma unroll
        for (int n = 0; n < NS; ++n)
          if (um >> n & 1u) Jrow[n] += dn * (m.nf[r][n] * q1 - m.nb[r][n] * q2);
      }
    }
  }
  if constexpr (JAC) {
#pragma unroll
    for (int n = 0; n < NS; ++n) {
      const double inv = yall[n] > 1e-30 ? 1.0 / yall[n] : 0.0;
      Jrow[n] = Jrow[n] * inv * (1e+3 * S.mw[a]);
    }
  }
  wave_sync(); // the exchange area is rewritten by the next evaluation
  return cr * (1e+3 * S.mw[a]) - de * (1e+3 * S.mw[a]);
}

// Row-parallel LU with partial pivoting of the matrix whose row a is held by lane a.  Leaves L (unit, below
// the diagonal), U above it and the reciprocal pivots on it, as the one-lane lu; returns the lane whose value
// row a of a permuted right-hand side takes.
template<int NS> __device__ __forceinline__ int row_lu(double *row, int a) {
  int piv[NS];
  static_for<0, NS>([&](auto K) {
    constexpr int k = decltype(K)::value;
    double col[NS];
    static_for<k, NS>([&](auto R) { col[decltype(R)::value] = row_bcast_c<decltype(R)::value>(row[k]); });
    int ik = k;
    double amax = 0;
#pragma unroll
    for (int r = k; r < NS; ++r) {
      const double v = fabs(col[r]);
      if (r == k) {
        amax = v;
      } else if (v > amax) {
        amax = v;
        ik = r;
      }
    }
    piv[k] = ik;
    double prow[NS], krow[NS];
#pragma unroll
    for (int t = 0; t < NS; ++t) {
      prow[t] = __shfl(row[t], ik, RW);
      krow[t] = row_bcast_c<k>(row[t]);
    }
#pragma unroll
    for (int t = 0; t < NS; ++t) row[t] = a == k ? prow[t] : (a == ik ? krow[t] : row[t]);
    const double inv = 1.0 / prow[k];
    if (a > k) {
      const double l = row[k] * inv;
      row[k] = l;
#pragma unroll
      for (int t = k + 1; t < NS; ++t) row[t] -= l * prow[t];
    } else if (a == k) {
      row[k] = inv;
    }
  });
  // the sequential exchanges b[r] <-> b[piv[r]] of the solve, traced back from position a
  int pos = a;
#pragma unroll
  for (int r = NS - 1; r >= 0; --r) {
    if (pos == r) pos = piv[r];
    else if (pos == piv[r]) pos = r;
  }
  return pos;
}

// Solve with the factors of row_lu: lane a holds b_a on entry and x_a on return.
template<int NS> __device__ __forceinline__ double row_solve(const double *row, int a, int src, double b) {
  b = __shfl(b, src, RW);
  static_for<0, NS - 1>([&](auto T) {
    constexpr int t = decltype(T)::value;
    const double bt = row_bcast_c<t>(b);
    const double nb = b - row[t] * bt;
    b = a > t ? nb : b;
  });
  if (a == NS - 1) b *= row[NS - 1];
  double fin[NS];
  fin[NS - 1] = row_bcast_c<NS - 1>(b);
  static_for<0, NS - 1>([&](auto I) {
    constexpr int r = NS - 2 - decltype(I)::value;
    double v = b;
#pragma unroll
    for (int t = r + 1; t < NS; ++t) v -= row[t] * fin[t];
    v *= row[r];
    b = a == r ? v : b;
    fin[r] = row_bcast_c<r>(b);
  });
  return b;
}

// Integration state of the row's cell (replicated across the row, species values per lane).
struct RowCell {
  double y; // rho Y of this lane's species (0 beyond NS)
  double rho, E, T, time, h, err_last, dt;
  int steps, rejects, attempts;
};

// One ROK4E attempt of the row's cell: the one-lane rok4e_attempt, species-parallel.
template<int NS>
__device__ __forceinline__ void row_attempt(const ChemMech &m, const SpeciesLds &S, RowLds &X, int a, int gl,
                                            RowCell &s, double rtol, double atol) {
  const bool own = gl < NS; // lanes beyond NS carry zeros
  const double dt = s.dt;
  double yall[NS], Jrow[NS];
  row_gather<NS>(s.y, yall);
  double w = row_source<NS, true>(m, S, X, a, gl, s.T, s.y, yall, Jrow);
  const double rem = dt - s.time;
  const double hs = s.h * kStretch >= rem ? rem : s.h;
#pragma unroll
  for (int n = 0; n < NS; ++n) Jrow[n] = -hs * G * Jrow[n];
#pragma unroll
  for (int n = 0; n < NS; ++n)
    if (n == a) Jrow[n] += 1.0;
  const int src = row_lu<NS>(Jrow, a);
  const double k1 = row_solve<NS>(Jrow, a, src, w);
  const double irho = 1.0 / s.rho;
  double ys = s.y + hs * A21 * k1;
  double Ts = row_update_t<NS>(m, S, a, s.T, own ? ys * irho : 0.0, s.E);
  row_gather<NS>(ys, yall);
  w = row_source<NS, false>(m, S, X, a, gl, Ts, ys, yall, nullptr);
  double k2 = row_solve<NS>(Jrow, a, src, w + G21 * k1);
  k2 -= G21 * k1;
  ys = s.y + hs * (A31 * k1 + A32 * k2);
  Ts = row_update_t<NS>(m, S, a, Ts, own ? ys * irho : 0.0, s.E);
  row_gather<NS>(ys, yall);
  w = row_source<NS, false>(m, S, X, a, gl, Ts, ys, yall, nullptr);
  double k3 = row_solve<NS>(Jrow, a, src, w + G31 * k1 + G32 * k2);
  k3 -= G31 * k1 + G32 * k2;
  double k4 = row_solve<NS>(Jrow, a, src, w + G41 * k1 + G42 * k2 + G43 * k3);
  k4 -= G41 * k1 + G42 * k2 + G43 * k3;
  const double ynew = s.y + hs * (B1 * k1 + B2 * k2 + B4 * k4);
  const double dy = hs * (E1 * k1 + E2 * k2 + E3 * k3 + E4 * k4) / (rtol * ynew + atol);
  double dys[NS];
  row_gather<NS>(dy, dys);
  double err = 0;
#pragma unroll
  for (int l = 0; l < NS; ++l) err += dys[l] * dys[l];
  err = sqrt(err / NS);
  double hstar = hs;
  if (s.steps > 0) hstar *= fmin(5.0, fmax(0.2, 0.8 * pow(s.err_last, 0.1) * pow(err + 1e-20, -0.175)));
  else if (!(err <= 1)) hstar *= 0.2;
  if (err <= 1 && hs < s.h) hstar = fmax(hstar, s.h);
  s.h = hstar;
  ++s.attempts;
  if (err <= 1) {
    s.y = own ? ynew : 0.0;
    s.time += hs;
    s.err_last = fmax(err, 1e-30);
    s.T = row_update_t<NS>(m, S, a, s.T, s.y * irho, s.E);
    ++s.steps;
  } else {
    ++s.rejects;
  }
}

__global__ void k_zero_counter(int *p) { *p = 0; }

__device__ __forceinline__ int work_count(const ChemBatch &b) { return b.list ? *b.n_list : b.n; }
__device__ __forceinline__ int work_cell(const ChemBatch &b, int w) { return b.list ? b.list[w] : w; }

template<int NS>
__global__ __launch_bounds__(256) void k_chem_coop(const ChemMech *__restrict__ mp, ChemBatch b, double dt,
                                                   double rtol, double atol, int max_attempts, int *counter,
                                                   int resume) {
  __shared__ SpeciesLds S;
  __shared__ RowLds rows[256 / RW];
  const ChemMech &m = *mp;
  for (int i = threadIdx.x; i < kMaxSpec; i += blockDim.x) {
#pragma unroll
    for (int q = 0; q < 7; ++q) {
      S.lo[q][i] = m.lo[i][q];
      S.hi[q][i] = m.hi[i][q];
    }
    S.t_low[i] = m.t_low[i];
    S.t_mid[i] = m.t_mid[i];
    S.t_high[i] = m.t_high[i];
    S.ln_t_low[i] = m.ln_t_low[i];
    S.ln_t_high[i] = m.ln_t_high[i];
    S.inv_t_low[i] = m.inv_t_low[i];
    S.inv_t_high[i] = m.inv_t_high[i];
    S.mw[i] = m.mw[i];
    S.imw[i] = m.imw[i];
  }
  __syncthreads();
  const int lane = __lane_id();
  const int gl = lane & (RW - 1);
  const int a = gl < NS ? gl : 0; // species of this lane (lanes beyond NS mirror species 0 and carry zeros)
  RowLds &X = rows[threadIdx.x / RW];
  const int n_work = work_count(b);
  RowCell s;
  int cell = -1;
  bool exhausted = false;
  while (true) {
    // one atomic per wave per fetch round; lane 0 of each row that needs a cell takes one
    const bool need = gl == 0 && cell < 0 && !exhausted;
    const unsigned long long mask = __ballot(need);
    if (mask) {
      const int leader = __ffsll((long long)mask) - 1;
      int base = 0;
      if (lane == leader) base = atomicAdd(counter, __popcll(mask));
      base = __shfl(base, leader, 64);
      int got = need ? base + __popcll(mask & ((1ull << lane) - 1)) : 0;
      got = __shfl(got, 0, RW);
      const bool row_needs = __shfl((int)need, 0, RW) != 0;
      if (row_needs) {
        if (got < n_work) {
          cell = work_cell(b, got);
          const double dtc = cell_dt(b, cell, dt);
          s.y = gl < NS ? b.rhoY[(long)gl * b.n + cell] : 0.0;
          s.rho = b.rho[cell];
          s.E = b.e[cell];
          s.T = b.T[cell];
          s.dt = dtc;
          if (resume) {
            s.h = b.h[cell];
            s.time = b.tcur[cell];
            s.err_last = b.elast[cell];
            s.steps = b.steps[cell];
            s.rejects = b.rejects[cell];
          } else {
            s.h = b.h[cell] > 0 ? fmin(b.h[cell], dtc) : dtc;
            s.time = 0;
            s.err_last = 0;
            s.steps = s.rejects = 0;
          }
          s.attempts = s.steps + s.rejects;
        } else {
          exhausted = true;
        }
      }
    }
    if (__ballot(cell >= 0) == 0) break;
    if (cell >= 0) {
      if (s.time < s.dt && s.attempts < max_attempts) row_attempt<NS>(m, S, X, a, gl, s, rtol, atol);
      if (s.time >= s.dt || s.attempts >= max_attempts) {
        if (gl < NS) b.rhoY[(long)gl * b.n + cell] = s.y;
        if (gl == 0) {
          b.T[cell] = s.T;
          b.h[cell] = s.h;
          b.steps[cell] = s.time >= s.dt ? s.steps : -1;
          b.rejects[cell] = s.rejects;
        }
        cell = -1;
      }
    }
  }
}

} // namespace

hipError_t launch_chem_coop(const ChemMech *mech_dev, int ns, const ChemBatch &b, double dt, double rtol,
                            double atol, int max_attempts, int threads, int blocks, bool resume, int *counter,
                            hipStream_t s) {
  if (ns != 9) return hipErrorNotSupported;
  if (b.n == 0) return hipSuccess;
  if (threads != 64 && threads != 128 && threads != 256) return hipErrorInvalidValue;
  if (blocks < 1) return hipErrorInvalidValue;
  k_zero_counter<<<1, 1, 0, s>>>(counter);
  k_chem_coop<9><<<blocks, threads, 0, s>>>(mech_dev, b, dt, rtol, atol, max_attempts, counter, resume ? 1 : 0);
  return hipGetLastError();
}

} // namespace corefl
