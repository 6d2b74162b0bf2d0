/*
 * oracle_chem.c -- thermo, finite-rate kinetics and the ROK4E stiff integrator of one constant-volume
 * cell (TEST INFRASTRUCTURE ONLY; see oracle.h).
 */
#include "oracle.h"

#include <math.h>
#include <string.h>

static const double kRu = 8314.462618;   /* J/(kmol K), Constants.h:17 */
static const double kPatm = 101325.0;    /* Constants.h:23 */

/* NASA-7 enthalpy and cp per unit mass with linear extrapolation outside [t_low, t_high]
 * (compute_enthalpy_and_cp, Thermo.cu:470-520 / ChemData.cpp host twin). */
void orc_chem_h_cp(const orc_mech *m, double t, double *h, double *cp) {
  const double t2 = t * t, t3 = t2 * t, t4 = t3 * t, t5 = t4 * t;
  for (int i = 0; i < m->ns; ++i) {
    const double R = kRu / m->mw[i];
    if (t < m->t_low[i] || t > m->t_high[i]) {
      const double tt = t < m->t_low[i] ? m->t_low[i] : m->t_high[i];
      const double *c = t < m->t_low[i] ? m->lo[i] : m->hi[i];
      const double tt2 = tt * tt, tt3 = tt2 * tt, tt4 = tt3 * tt, tt5 = tt4 * tt;
      double hh = c[0] * tt + 0.5 * c[1] * tt2 + c[2] * tt3 / 3 + 0.25 * c[3] * tt4 + 0.2 * c[4] * tt5 + c[5];
      const double cc = c[0] + c[1] * tt + c[2] * tt2 + c[3] * tt3 + c[4] * tt4;
      hh += cc * (t - tt);
      h[i] = hh * R;
      cp[i] = cc * R;
    } else {
      const double *c = t < m->t_mid[i] ? m->lo[i] : m->hi[i];
      h[i] = (c[0] * t + 0.5 * c[1] * t2 + c[2] * t3 / 3 + 0.25 * c[3] * t4 + 0.2 * c[4] * t5 + c[5]) * R;
      cp[i] = (c[0] + c[1] * t + c[2] * t2 + c[3] * t3 + c[4] * t4) * R;
    }
  }
}

/* g/RT per species, evaluated at the clamped temperature outside the fit range (Thermo.cu:574-598). */
static void gibbs_rt(const orc_mech *m, double t, double *g) {
  for (int i = 0; i < m->ns; ++i) {
    double tt = t;
    if (t < m->t_low[i]) tt = m->t_low[i];
    else if (t > m->t_high[i]) tt = m->t_high[i];
    const double *c = (t < m->t_low[i]) ? m->lo[i] : (t > m->t_high[i]) ? m->hi[i] : (t < m->t_mid[i] ? m->lo[i] : m->hi[i]);
    const double tt2 = tt * tt, tt3 = tt2 * tt, tt4 = tt3 * tt;
    g[i] = c[0] * (1.0 - log(tt)) - 0.5 * c[1] * tt - c[2] * tt2 / 6.0 - c[3] * tt3 / 12.0 - c[4] * tt4 * 0.05 +
           c[5] / tt - c[6];
  }
}

static double arrhenius(double t, double A, double b, double Ea) { return A * pow(t, b) * exp(-Ea / t); }

/* forward_reaction_rate / backward_reaction_rate / rate_of_progress / chemical_source
 * (FiniteRateChem.cu:1466-1605).  rhoY in kg/m^3; q1, q2 in mol/(cm^3 s); omega in kg/(m^3 s). */
void orc_chem_source(const orc_mech *m, double t, const double *rhoY, double *q1, double *q2, double *omega_d,
                     double *omega) {
  const int ns = m->ns, nr = m->nr;
  double c[ORC_MAX_SPEC], kf[ORC_MAX_REAC], kb[ORC_MAX_REAC];
  for (int l = 0; l < ns; ++l) c[l] = rhoY[l] / m->mw[l] * 1e-3;
  for (int r = 0; r < nr; ++r) {
    kf[r] = arrhenius(t, m->A[r], m->b[r], m->Ea[r]);
    const int ty = m->type[r];
    if (ty == 3) {
      kf[r] += arrhenius(t, m->A2[r], m->b2[r], m->Ea2[r]);
    } else if (ty > 3) {
      double cc = 0;
      for (int l = 0; l < ns; ++l) cc += c[l] * m->tb[r][l];
      if (ty == 4) {
        kf[r] *= cc;
      } else {
        const double klow = arrhenius(t, m->A2[r], m->b2[r], m->Ea2[r]), khigh = kf[r];
        if (khigh < 1e-25 && klow < 1e-25) {
          kf[r] = 0;
          continue;
        }
        const double pr = klow * cc / khigh;
        double F = 1.0;
        if (ty > 5) {
          double fc = (1 - m->troe[r][0]) * exp(-t / m->troe[r][1]) + m->troe[r][0] * exp(-t / m->troe[r][2]);
          if (ty == 7) fc += exp(-m->troe[r][3] / t);
          const double lfc = log10(fc);
          const double cq = -0.4 - 0.67 * lfc, nq = 0.75 - 1.27 * lfc;
          const double lpr = log10(pr);
          const double tmp = (lpr + cq) / (nq - 0.14 * (lpr + cq));
          F = pow(10, lfc / (1.0 + tmp * tmp));
        }
        kf[r] = khigh * pr / (1.0 + pr) * F;
      }
    }
  }
  int n_gibbs = nr;
  for (int r = 0; r < nr; ++r) {
    if (m->type[r] == 0) {
      kb[r] = 0;
      --n_gibbs;
    } else if (m->rev[r] == 1) {
      kb[r] = arrhenius(t, m->A2[r], m->b2[r], m->Ea2[r]);
      if (m->type[r] == 4) {
        double cc = 0;
        for (int l = 0; l < ns; ++l) cc += c[l] * m->tb[r][l];
        kb[r] *= cc;
      }
      --n_gibbs;
    }
  }
  if (n_gibbs > 0) {
    double g[ORC_MAX_SPEC];
    gibbs_rt(m, t, g);
    const double tp = kPatm / kRu * 1e-3 / t;
    for (int r = 0; r < nr; ++r) {
      if (m->type[r] != 2 && m->type[r] != 0) { /* as the reference: recomputed even after REV */
        double dg = 0;
        for (int l = 0; l < ns; ++l) dg += g[l] * (m->nb[r][l] - m->nf[r][l]);
        kb[r] = kf[r] / (pow(tp, m->order[r]) * exp(-dg));
      }
    }
  }
  for (int r = 0; r < nr; ++r) {
    q1[r] = 1.0;
    q2[r] = m->type[r] != 0 ? 1.0 : 0.0;
    for (int l = 0; l < ns; ++l) {
      q1[r] *= pow(c[l], m->nf[r][l]);
      if (m->type[r] != 0) q2[r] *= pow(c[l], m->nb[r][l]);
    }
    q1[r] *= kf[r];
    q2[r] *= kb[r];
  }
  for (int l = 0; l < ns; ++l) {
    double cr = 0, de = 0;
    for (int r = 0; r < nr; ++r) {
      cr += q2[r] * m->nf[r][l] + q1[r] * m->nb[r][l];
      de += q1[r] * m->nf[r][l] + q2[r] * m->nb[r][l];
    }
    cr *= 1e+3 * m->mw[l];
    de *= 1e+3 * m->mw[l];
    omega_d[l] = de;
    omega[l] = cr - de;
  }
}

/* d omega / d(rho Y) with q1, q2 frozen (ConstVolumeReactor.cpp:1168-1187). */
void orc_chem_jacobian(const orc_mech *m, const double *rhoY, const double *q1, const double *q2, double *jac) {
  const int ns = m->ns, nr = m->nr;
  for (int a = 0; a < ns; ++a)
    for (int n = 0; n < ns; ++n) {
      double zz = 0;
      if (rhoY[n] > 1e-30) {
        for (int r = 0; r < nr; ++r)
          zz += (m->nb[r][a] - m->nf[r][a]) * (m->nf[r][n] * q1[r] - m->nb[r][n] * q2[r]);
        zz /= rhoY[n];
      }
      jac[a * ns + n] = zz * 1e+3 * m->mw[a];
    }
}

/* Newton iteration for T at fixed internal energy E (update_t, ConstVolumeReactor.cpp:970-995). */
double orc_chem_update_t(const orc_mech *m, double t0, const double *yk, double E) {
  double imw = 0;
  for (int l = 0; l < m->ns; ++l) imw += yk[l] / m->mw[l];
  const double R = kRu * imw;
  double t = t0, err = 1e+6;
  int it = 0;
  while (err > 1e-8 && it++ < 100) {
    double h[ORC_MAX_SPEC], cp[ORC_MAX_SPEC];
    orc_chem_h_cp(m, t, h, cp);
    double cpt = 0, hh = 0;
    for (int l = 0; l < m->ns; ++l) {
      cpt += cp[l] * yk[l];
      hh += h[l] * yk[l];
    }
    const double t1 = t - (hh - R * t - E) / (cpt - R);
    err = fabs(1 - t1 / t);
    t = t1;
  }
  return t;
}

/* Doolittle LU with partial pivoting and the matching solve (ConstVolumeReactor.cpp:1231-1285). */
void orc_lu_decomp(double *a, int n, int *piv) {
  for (int k = 0; k < n; ++k) {
    int ik = k;
    for (int r = k; r < n; ++r) {
      for (int t = 0; t < k; ++t) a[r * n + k] -= a[r * n + t] * a[t * n + k];
      if (fabs(a[r * n + k]) > fabs(a[ik * n + k])) ik = r;
    }
    piv[k] = ik;
    if (ik != k)
      for (int t = 0; t < n; ++t) {
        const double s = a[ik * n + t];
        a[ik * n + t] = a[k * n + t];
        a[k * n + t] = s;
      }
    for (int p = k + 1; p < n; ++p)
      for (int t = 0; t < k; ++t) a[k * n + p] -= a[k * n + t] * a[t * n + p];
    for (int r = k + 1; r < n; ++r) a[r * n + k] /= a[k * n + k];
  }
}

void orc_lu_solve(const double *a, double *b, int n, const int *piv) {
  for (int r = 0; r < n; ++r)
    if (piv[r] != r) {
      const double s = b[piv[r]];
      b[piv[r]] = b[r];
      b[r] = s;
    }
  for (int r = 1; r < n; ++r)
    for (int t = 0; t < r; ++t) b[r] -= a[r * n + t] * b[t];
  b[n - 1] /= a[n * n - 1];
  for (int r = n - 2; r >= 0; --r) {
    for (int t = r + 1; t < n; ++t) b[r] -= a[r * n + t] * b[t];
    b[r] /= a[r * n + r];
  }
}

/* ROK4E coefficients (Define.h:28-43) */
static const double G = 0.572816062482135;
static const double A21 = 0.432364435748567, A31 = -0.51421131687617, A32 = 1.38227114461736;
#define G21 (-0.602765307997356 / G)
#define G31 (-1.389195789724843 / G)
#define G32 (1.072950969011413 / G)
#define G41 (0.992356412977094 / G)
#define G42 (-1.390032613873701 / G)
#define G43 (-0.440875890223325 / G)
static const double B1 = 0.194335256262729, B2 = 0.483167813989227, B4 = 0.322496929748044;
#define E1 (-0.217819895945721 - B1)
#define E2 (1.03130847478467 - B2)
static const double E3 = 0.186511421161047;
#define E4 (-B4)

/* Adaptive ROK4E over [0, t_end] at constant density and internal energy (ConstVolumeReactor.cpp:173-497,
 * build_arnoldi = 0, adaptive_time_step = 1).  rhoY and *T are advanced in place; *h carries the step size
 * from call to call.  Returns the number of accepted steps, or -1 when max_attempts (accepted + rejected)
 * attempts did not reach t_end; *n_rej gets the rejections.  If n_rec > 0, (time, T) after each accepted
 * step is written to rec[2*s], rec[2*s+1] for s < n_rec.
 * Two deviations, both where the reference's controller degenerates: a first-step error that is not a number
 * is treated as err > 1 (step cut to 0.2 h), where the reference leaves h unchanged (ConstVolumeReactor.cpp:
 * 421-425) and repeats the same attempt forever; and the error carried into the PI controller is floored at
 * 1e-30, because a step with an exactly zero error (frozen chemistry) makes err_last^0.1 = 0 and the reference
 * then shrinks every later accepted step by 0.2, so the sub-step never ends. */
int orc_chem_rok4e(const orc_mech *m, double rho, double E, double *rhoY, double *T, double t_end, double rtol,
                   double atol, double *h, int max_attempts, int *n_rej, double *rec, int n_rec) {
  const int ns = m->ns, nr = m->nr;
  double q1[ORC_MAX_REAC], q2[ORC_MAX_REAC], wd[ORC_MAX_SPEC], w[ORC_MAX_SPEC];
  double jac[ORC_MAX_SPEC * ORC_MAX_SPEC], lhs[ORC_MAX_SPEC * ORC_MAX_SPEC];
  double k1[ORC_MAX_SPEC], k2[ORC_MAX_SPEC], k3[ORC_MAX_SPEC], k4[ORC_MAX_SPEC], ys[ORC_MAX_SPEC],
      yk[ORC_MAX_SPEC], ynew[ORC_MAX_SPEC];
  int piv[ORC_MAX_SPEC];
  double time = 0, err_last = 0;
  int step = 0, rej = 0;
  (void)nr;
  while (time < t_end && step + rej < max_attempts) {
    orc_chem_source(m, *T, rhoY, q1, q2, wd, w);
    /* a step that would leave less than 5% of itself to t_end ends the interval (step stretching; the
     * reference takes min(h, t_end - time), whose sliver last step then becomes the carried proposal) */
    const double rem = t_end - time;
    const double hs = *h * 1.05 >= rem ? rem : *h;
    orc_chem_jacobian(m, rhoY, q1, q2, jac);
    for (int a = 0; a < ns; ++a)
      for (int n = 0; n < ns; ++n) lhs[a * ns + n] = (a == n ? 1.0 : 0.0) - hs * G * jac[a * ns + n];
    orc_lu_decomp(lhs, ns, piv);
    for (int l = 0; l < ns; ++l) k1[l] = w[l];
    orc_lu_solve(lhs, k1, ns, piv);
    for (int l = 0; l < ns; ++l) {
      ys[l] = rhoY[l] + hs * A21 * k1[l];
      yk[l] = ys[l] / rho;
    }
    double Ts = orc_chem_update_t(m, *T, yk, E);
    orc_chem_source(m, Ts, ys, q1, q2, wd, w);
    for (int l = 0; l < ns; ++l) k2[l] = w[l] + G21 * k1[l];
    orc_lu_solve(lhs, k2, ns, piv);
    for (int l = 0; l < ns; ++l) k2[l] -= G21 * k1[l];
    for (int l = 0; l < ns; ++l) {
      ys[l] = rhoY[l] + hs * (A31 * k1[l] + A32 * k2[l]);
      yk[l] = ys[l] / rho;
    }
    Ts = orc_chem_update_t(m, Ts, yk, E);
    orc_chem_source(m, Ts, ys, q1, q2, wd, w);
    for (int l = 0; l < ns; ++l) k3[l] = w[l] + G31 * k1[l] + G32 * k2[l];
    orc_lu_solve(lhs, k3, ns, piv);
    for (int l = 0; l < ns; ++l) k3[l] -= G31 * k1[l] + G32 * k2[l];
    for (int l = 0; l < ns; ++l) k4[l] = w[l] + G41 * k1[l] + G42 * k2[l] + G43 * k3[l];
    orc_lu_solve(lhs, k4, ns, piv);
    for (int l = 0; l < ns; ++l) k4[l] -= G41 * k1[l] + G42 * k2[l] + G43 * k3[l];
    double err = 0;
    for (int l = 0; l < ns; ++l) {
      ynew[l] = rhoY[l] + hs * (B1 * k1[l] + B2 * k2[l] + B4 * k4[l]);
      double dy = hs * (E1 * k1[l] + E2 * k2[l] + E3 * k3[l] + E4 * k4[l]);
      dy /= rtol * ynew[l] + atol;
      err += dy * dy;
    }
    err = sqrt(err / ns);
    double hstar = hs;
    if (step > 0)
      hstar *= fmin(5.0, fmax(0.2, 0.8 * pow(err_last, 0.1) * pow(err + 1e-20, -0.175)));
    else if (!(err <= 1))
      hstar *= 0.2;
    if (err <= 1 && hs < *h) hstar = fmax(hstar, *h); /* a step cut short by t_end keeps the proposal */
    *h = hstar;
    if (err <= 1) {
      for (int l = 0; l < ns; ++l) {
        rhoY[l] = ynew[l];
        yk[l] = ynew[l] / rho;
      }
      time += hs;
      err_last = fmax(err, 1e-30);
      *T = orc_chem_update_t(m, *T, yk, E);
      if (step < n_rec) {
        rec[2 * step] = time;
        rec[2 * step + 1] = *T;
      }
      ++step;
    } else {
      ++rej;
    }
  }
  if (n_rej) *n_rej = rej;
  return time < t_end ? -1 : step;
}
