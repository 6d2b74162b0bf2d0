// Device thermo and kinetics of one cell, shared by the chemistry and surrogate kernels: NASA-7 h, cp and
// g/RT, Arrhenius / third-body / Lindemann / Troe rates with Gibbs-energy reverse rates, the chemical source
// and its Jacobian, and the Newton temperature at fixed internal energy.  FP64.
#pragma once

#include <hip/hip_runtime.h>

#include <cmath>

#include "chem.hpp"

namespace corefl {
namespace chemdev {

constexpr double kRu = 8314.462618, kPatm = 101325.0;
// constant divisions of the polynomial fits as products (a few ulp from the divided form)
constexpr double kThird = 1.0 / 3.0, kSixth = 1.0 / 6.0, kTwelfth = 1.0 / 12.0;

// h_k and cp_k per unit mass, linear extrapolation outside [t_low, t_high] (Thermo.cu:470-520).
// h and cp of species i at t (the body of h_cp for one species).
__device__ __forceinline__ void h_cp_one(const ChemMech &m, int i, double t, double &h, double &cp) {
  const double t2 = t * t, t3 = t2 * t, t4 = t3 * t, t5 = t4 * t;
  const double R = kRu * m.imw[i];
  const bool below = t < m.t_low[i], above = t > m.t_high[i];
  if (below || above) {
    const double tt = below ? m.t_low[i] : m.t_high[i];
    const double *c = below ? m.lo[i] : m.hi[i];
    const double tt2 = tt * tt, tt3 = tt2 * tt, tt4 = tt3 * tt, tt5 = tt4 * tt;
    const double cc = c[0] + c[1] * tt + c[2] * tt2 + c[3] * tt3 + c[4] * tt4;
    const double hh = c[0] * tt + 0.5 * c[1] * tt2 + c[2] * tt3 * kThird + 0.25 * c[3] * tt4 + 0.2 * c[4] * tt5 + c[5];
    h = (hh + cc * (t - tt)) * R;
    cp = cc * R;
  } else {
    const double *c = t < m.t_mid[i] ? m.lo[i] : m.hi[i];
    h = (c[0] * t + 0.5 * c[1] * t2 + c[2] * t3 * kThird + 0.25 * c[3] * t4 + 0.2 * c[4] * t5 + c[5]) * R;
    cp = (c[0] + c[1] * t + c[2] * t2 + c[3] * t3 + c[4] * t4) * R;
  }
}

template<int NS> __device__ __forceinline__ void h_cp(const ChemMech &m, double t, double *h, double *cp) {
#pragma unroll
  for (int i = 0; i < NS; ++i) h_cp_one(m, i, t, h[i], cp[i]);
}

// g_k / (R T), clamped to the fit range (Thermo.cu:574-598).  lnT = log(t): one logarithm serves every
// species whose fit range holds t; the logarithms of the range ends are tabulated.
template<int NS>
__device__ __forceinline__ void gibbs(const ChemMech &m, double t, double lnT, double iT, double *g) {
#pragma unroll
  for (int i = 0; i < NS; ++i) {
    const bool below = t < m.t_low[i], above = t > m.t_high[i];
    const double tt = below ? m.t_low[i] : (above ? m.t_high[i] : t);
    const double *c = below ? m.lo[i] : (above ? m.hi[i] : (t < m.t_mid[i] ? m.lo[i] : m.hi[i]));
    const double tt2 = tt * tt, tt3 = tt2 * tt, tt4 = tt3 * tt;
    const double lt = below ? m.ln_t_low[i] : (above ? m.ln_t_high[i] : lnT);
    const double itt = below ? m.inv_t_low[i] : (above ? m.inv_t_high[i] : iT);
    g[i] = c[0] * (1.0 - lt) - 0.5 * c[1] * tt - c[2] * tt2 * kSixth - c[3] * tt3 * kTwelfth - c[4] * tt4 * 0.05 +
           c[5] * itt - c[6];
  }
}

// A T^b exp(-Ea/T) as A exp(b ln T - Ea/T): one exponential per rate instead of a power and an exponential
// (the same value up to rounding).
__device__ __forceinline__ double arrhenius(double lnT, double iT, double A, double b, double Ea) {
  return A * exp(b * lnT - Ea * iT);
}

// x^k for an integer k (reaction orders and stoichiometric coefficients are integers: the parser rejects
// others), by repeated products as the reference's integer branch does.
__device__ __forceinline__ double int_pow(double x, int k) {
  const int n = k < 0 ? -k : k;
  double r = 1.0;
  for (int q = 0; q < n; ++q) r *= x;
  return k < 0 ? 1.0 / r : r;
}

// Chemical source omega [kg/(m^3 s)] of one cell (FiniteRateChem.cu:1466-1605).  With JAC, also the
// Jacobian d omega / d(rho Y) with the rates of progress frozen (ConstVolumeReactor.cpp:1168-1187); columns of
// species with rho Y at or below floor are zero (1e-30 in the reactor, 1e-25 in Wu's EPI, StrangSplitting.cu:164).
// Stoichiometry is walked through each reaction's species lists (a few entries) rather than all species;
// the Jacobian keeps static indices (register-resident) and skips rows and columns through the reaction's
// species bit masks.
template<int NS, bool JAC>
__device__ __forceinline__ void source(const ChemMech &m, double t, const double *rhoY, double *w, double *J,
                                       double *wd = nullptr, double floor = 1e-30) {
  double c[NS], g[NS], cr[NS], de[NS];
#pragma unroll
  for (int l = 0; l < NS; ++l) {
    c[l] = rhoY[l] * m.imw[l] * 1e-3;
    cr[l] = 0;
    de[l] = 0;
  }
  if constexpr (JAC) {
#pragma unroll
    for (int q = 0; q < NS * NS; ++q) J[q] = 0;
  }
  // Equilibrium constants are needed unless every reversible reaction carries explicit REV parameters
  // (FiniteRateChem.cu:1519-1557); then they are applied to every reaction of type other than 0 and 2.
  const double lnT = log(t), iT = 1.0 / t;
  if (m.need_gibbs) gibbs<NS>(m, t, lnT, iT, g);
  const double tp = kPatm / kRu * 1e-3 / t;
  for (int r = 0; r < m.nr; ++r) {
    const int ty = m.type[r];
    const int nfs = m.nfs[r], nbs = m.nbs[r];
    double kf = arrhenius(lnT, iT, m.A[r], m.b[r], m.Ea[r]);
    double cc = 0;
    if (ty >= 4) {
#pragma unroll
      for (int l = 0; l < NS; ++l) cc += c[l] * m.tb[r][l];
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
      for (int s = 0; s < nbs; ++s) dg += g[m.bsp[r][s]] * m.bnu[r][s];
      for (int s = 0; s < nfs; ++s) dg -= g[m.fsp[r][s]] * m.fnu[r][s];
      kb = kf / (int_pow(tp, m.iorder[r]) * exp(-dg));
    }
    double q1 = 1.0, q2 = ty != 0 ? 1.0 : 0.0;
    for (int s = 0; s < nfs; ++s) q1 *= int_pow(c[m.fsp[r][s]], m.fnu[r][s]);
    if (ty != 0)
      for (int s = 0; s < nbs; ++s) q2 *= int_pow(c[m.bsp[r][s]], m.bnu[r][s]);
    q1 *= kf;
    q2 *= kb;
    for (int s = 0; s < nfs; ++s) {
      const int l = m.fsp[r][s];
      const double nu = m.fnu[r][s];
      cr[l] += q2 * nu;
      de[l] += q1 * nu;
    }
    for (int s = 0; s < nbs; ++s) {
      const int l = m.bsp[r][s];
      const double nu = m.bnu[r][s];
      cr[l] += q1 * nu;
      de[l] += q2 * nu;
    }
    if constexpr (JAC) {
      const unsigned um = m.umask[r], am = m.amask[r];
      double v[NS];
#pragma unroll
      for (int n = 0; n < NS; ++n) v[n] = (um >> n & 1u) ? m.nf[r][n] * q1 - m.nb[r][n] * q2 : 0.0;
#pragma unroll
      for (int a = 0; a < NS; ++a) {
        if (!(am >> a & 1u)) continue;
        const double dn = m.dnu[r][a];
#pragma unroll
        for (int n = 0; n < NS; ++n)
          if (um >> n & 1u) J[a * NS + n] += dn * v[n];
      }
    }
  }
#pragma unroll
  for (int l = 0; l < NS; ++l) {
    w[l] = cr[l] * (1e+3 * m.mw[l]) - de[l] * (1e+3 * m.mw[l]);
    if (wd) wd[l] = de[l] * (1e+3 * m.mw[l]);
  }
  if constexpr (JAC) {
#pragma unroll
    for (int n = 0; n < NS; ++n) {
      const double inv = rhoY[n] > floor ? 1.0 / rhoY[n] : 0.0;
#pragma unroll
      for (int a = 0; a < NS; ++a) J[a * NS + n] = J[a * NS + n] * inv * (1e+3 * m.mw[a]);
    }
  }
}

// Doolittle LU with partial pivoting, row-major (ConstVolumeReactor.cpp:1231-1262), and its solve.
template<int NS> __device__ __forceinline__ void lu(double *a, int *piv) {
#pragma unroll
  for (int k = 0; k < NS; ++k) {
    // first row of largest |a[r][k]|; the running maximum keeps every array index static, so the matrix
    // stays in registers
    int ik = k;
    double amax = 0;
#pragma unroll
    for (int r = k; r < NS; ++r) {
#pragma unroll
      for (int t = 0; t < k; ++t) a[r * NS + k] -= a[r * NS + t] * a[t * NS + k];
      const double v = fabs(a[r * NS + k]);
      if (r == k) {
        amax = v;
      } else if (v > amax) {
        amax = v;
        ik = r;
      }
    }
    piv[k] = ik;
    // row exchange through value selects: every index stays static and no branch is formed (a branch per
    // candidate row lets the optimiser merge them into stores through a chosen address, which sends the
    // whole matrix to scratch memory)
#pragma unroll
    for (int r = k + 1; r < NS; ++r) {
      const bool sw = r == ik;
#pragma unroll
      for (int t = 0; t < NS; ++t) {
        const double x = a[r * NS + t], y = a[k * NS + t];
        a[r * NS + t] = sw ? y : x;
        a[k * NS + t] = sw ? x : y;
      }
    }
#pragma unroll
    for (int p = k + 1; p < NS; ++p)
#pragma unroll
      for (int t = 0; t < k; ++t) a[k * NS + p] -= a[k * NS + t] * a[t * NS + p];
    const double inv = 1.0 / a[k * NS + k];
#pragma unroll
    for (int r = k + 1; r < NS; ++r) a[r * NS + k] *= inv;
    a[k * NS + k] = inv; // the solve multiplies by the reciprocal pivot
  }
}

template<int NS> __device__ __forceinline__ void lu_solve(const double *a, const int *piv, double *b) {
#pragma unroll
  for (int r = 0; r < NS; ++r) {
#pragma unroll
    for (int q = r + 1; q < NS; ++q) {
      const bool sw = piv[r] == q;
      const double x = b[q], y = b[r];
      b[q] = sw ? y : x;
      b[r] = sw ? x : y;
    }
  }
#pragma unroll
  for (int r = 1; r < NS; ++r)
#pragma unroll
    for (int t = 0; t < r; ++t) b[r] -= a[r * NS + t] * b[t];
  b[NS - 1] *= a[NS * NS - 1];
#pragma unroll
  for (int r = NS - 2; r >= 0; --r) {
#pragma unroll
    for (int t = r + 1; t < NS; ++t) b[r] -= a[r * NS + t] * b[t];
    b[r] *= a[r * NS + r];
  }
}

// Newton iteration for T at fixed specific internal energy E (ConstVolumeReactor.cpp:970-995).
template<int NS> __device__ double update_t(const ChemMech &m, double t, const double *yk, double E) {
  double imw = 0;
#pragma unroll
  for (int l = 0; l < NS; ++l) imw += yk[l] * m.imw[l];
  const double R = kRu * imw;
  double err = 1e+6;
  for (int it = 0; it < 100 && err > 1e-8; ++it) {
    // h_cp species by species, summed as they come (the same operations in the same order as summing the
    // arrays h_cp fills, without holding them)
    double cpt = 0, hh = 0;
#pragma unroll
    for (int l = 0; l < NS; ++l) {
      double h1, cp1;
      h_cp_one(m, l, t, h1, cp1);
      cpt += cp1 * yk[l];
      hh += h1 * yk[l];
    }
    const double t1 = t - (hh - R * t - E) / (cpt - R);
    err = fabs(1 - t1 / t);
    t = t1;
  }
  return t;
}

// Collision integral of binary diffusion at reduced temperature t (Transport.cu:188-190),
// 1.0548 t^-0.15504 + (t + 0.55909)^-2.1705, given lt = ln t: t^a as exp(a ln t), so that a node's pairs share
// its ln T (the powers agree with pow to a few ulps).
__device__ __forceinline__ double omega_d_log(double t, double lt) {
  return 1.0548 * exp(-0.15504 * lt) + exp(-2.1705 * log(t + 0.55909));
}

// compute_transport_property (Transport.cu:43-150) at temperature t and mass fractions y (cp: species cp per
// unit mass at t): Chapman-Enskog species viscosities mixed by Wilke's rule; species conductivities from
// Eucken's relation with the rotational-relaxation correction of linear and nonlinear molecules, mixed with
// the 1.065 factor; mixture-averaged rho D_k from the binary coefficients.
template<int NS>
__device__ inline void transport(const ChemMech &m, double t, const double *y, const double *cp, double &mu,
                                 double &lambda, double *rhoD) {
  // The reference's formulas with their powers of the reduced temperatures T k_B / eps written as exponentials
  // of ln T + ln(k_B / eps) (one logarithm per node instead of one per power), 1/(x x) for x^-2, and products
  // with tabulated or once-computed reciprocals in place of divisions: equal to the direct form to a few ulps.
  constexpr double pi = 3.14159265358979323846, Ru = 8314.462618;
  double x[NS], vs[NS];
  double imw = 0;
#pragma unroll
  for (int l = 0; l < NS; ++l) imw += y[l] * m.imw[l];
  const double mwt = 1.0 / imw;
  const double st = sqrt(t), t32 = t * st, lt = log(t), it32 = 1.0 / t32;
#pragma unroll
  for (int l = 0; l < NS; ++l) {
    x[l] = y[l] * mwt * m.imw[l];
    const double td = t * m.lj_inv[l], u = td + 0.5;
    vs[l] = m.vis_coeff[l] * st / (1.147 * exp(-0.145 * (lt + m.log_lj_inv[l])) + 1.0 / (u * u));
  }
  double isq[NS];
#pragma unroll
  for (int l = 0; l < NS; ++l) isq[l] = 1.0 / sqrt(vs[l]);
  mu = 0;
  lambda = 0;
#pragma unroll
  for (int a = 0; a < NS; ++a) {
    const double R = Ru * m.imw[a];
    double lam = 15 * 0.25 * vs[a] * R;
    const int geom = m.geometry[a];
    if (geom == 1 || geom == 2) {
      const double tr = t * m.kbe[a][a], tri = 1 / tr;
      const double q = sqrt(tri * pi) * pi;
      const double ft = 1 + 0.5 * q + (2 + 0.25 * pi * pi) * tri + q * tri;
      const double rd = m.bdc[a][a] * t32 / (omega_d_log(tr, lt + m.log_kbe[a][a]) * R * t);
      if (geom == 1) {
        const double ab = (2.5 * vs[a] - rd) / (pi * m.zrot[a] / ft * vs[a] + 10.0 / 3 * vs[a] + 2.0 * rd);
        lam += -5 * ab * vs[a] * R + rd * cp[a] + rd * (2 * ab - 2.5) * R;
      } else {
        const double ab = (2.5 * vs[a] - rd) / (pi * m.zrot[a] / ft * vs[a] + 5 * vs[a] + 2.0 * rd);
        lam += -7.5 * ab * vs[a] * R + rd * cp[a] + rd * (3 * ab - 2.5) * R;
      }
    }
    double den = 0;
    const double sa = sqrt(vs[a]);
#pragma unroll
    for (int b = 0; b < NS; ++b) {
      double phi = 1.0;
      if (a != b) {
        const double num = 1.0 + sa * isq[b] * m.wdw[a][b];
        phi = num * num * m.sqw[a][b];
      }
      den += phi * x[b];
    }
    mu += vs[a] * x[a] / den;
    lambda += lam * x[a] / (1.065 * den - 0.065 * x[a]);
  }
  constexpr double eps = 1e-12;
  double num[NS], dsum[NS];
  double sum = 0;
#pragma unroll
  for (int l = 0; l < NS; ++l) {
    num[l] = (x[l] + eps) * m.mw[l];
    sum += num[l];
    dsum[l] = 0;
  }
#pragma unroll
  for (int l = 0; l < NS; ++l)
#pragma unroll
    for (int n = l + 1; n < NS; ++n) {
      const double inv = omega_d_log(t * m.kbe[l][n], lt + m.log_kbe[l][n]) * m.inv_bdc[l][n] * it32;
      dsum[l] += (x[n] + eps) * inv;
      dsum[n] += (x[l] + eps) * inv;
    }
  const double f = 1.0 / (t * Ru);
#pragma unroll
  for (int l = 0; l < NS; ++l) rhoD[l] = (sum - num[l]) * f / dsum[l];
}

} // namespace chemdev
} // namespace corefl
