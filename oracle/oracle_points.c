/*
 * oracle_points.c -- point functions of the oracle (TEST INFRASTRUCTURE ONLY; see oracle.h).
 * WENO weights, characteristic matrices, Riemann solvers, Sutherland, exact Riemann solution.
 */
#include "oracle.h"

#include <math.h>
#include <string.h>

static const double kGamma = 1.4;                   /* gamma_air, Constants.h:21 */
static const double kRu = 8314.462618;              /* R_u, Constants.h:17 */
static const double kMwAir = 28.9647;               /* mw_air, Constants.h:26 */

/* One upwind-biased half of WENO5 (the fPlus part of WENO.cu:3059-3114): candidate values q[0..2] at the
 * right face of v[2] from v[0..4], Jiang-Shu smoothness indicators beta[0..2] and unnormalised weights
 * alpha[0..2] (linear weights when if_shock is 0).  Returns sum(alpha q) / sum(alpha).  The downwind half
 * (fMinus) is the same function of the mirrored stencil. */
double orc_weno5_half(const double *v, double eps, int if_shock, double *beta, double *alpha) {
  const double one6th = 1.0 / 6, t12 = 13.0 / 12;
  const double q0 = one6th * (2 * v[2] + 5 * v[3] - v[4]);
  const double q1 = one6th * (-v[1] + 5 * v[2] + 2 * v[3]);
  const double q2 = one6th * (2 * v[0] - 7 * v[1] + 11 * v[2]);
  const double b0 = t12 * (v[2] + v[4] - 2 * v[3]) * (v[2] + v[4] - 2 * v[3]) +
                    0.25 * (3 * v[2] - 4 * v[3] + v[4]) * (3 * v[2] - 4 * v[3] + v[4]);
  const double b1 = t12 * (v[1] + v[3] - 2 * v[2]) * (v[1] + v[3] - 2 * v[2]) + 0.25 * (v[1] - v[3]) * (v[1] - v[3]);
  const double b2 = t12 * (v[0] + v[2] - 2 * v[1]) * (v[0] + v[2] - 2 * v[1]) +
                    0.25 * (v[0] - 4 * v[1] + 3 * v[2]) * (v[0] - 4 * v[1] + 3 * v[2]);
  double a0 = 0.3, a1 = 0.6, a2 = 0.1;
  if (if_shock) { /* Jiang-Shu weights (WENO.cu:3077-3079; the Z form there is commented out) */
    a0 = 0.3 / ((eps + b0) * (eps + b0));
    a1 = 0.6 / ((eps + b1) * (eps + b1));
    a2 = 0.1 / ((eps + b2) * (eps + b2));
  }
  if (beta) { beta[0] = b0; beta[1] = b1; beta[2] = b2; }
  if (alpha) { alpha[0] = a0; alpha[1] = a1; alpha[2] = a2; }
  if (!if_shock) return 0.3 * q0 + 0.6 * q1 + 0.1 * q2; /* the linear upwind scheme UD5 (WENO.cu:3097-3100) */
  return (a0 * q0 + a1 * q1 + a2 * q2) / (a0 + a1 + a2);
}

/* WENO.cu:3059-3114: fPlus from vp, fMinus from the mirrored vm. */
double orc_weno5(const double *vp, const double *vm, double eps, int if_shock) {
  const double m[5] = {vm[4], vm[3], vm[2], vm[1], vm[0]};
  return orc_weno5_half(vp, eps, if_shock, NULL, NULL) + orc_weno5_half(m, eps, if_shock, NULL, NULL);
}

/* smoothness indicator of one 4-point sub-stencil in the WENO.cu:3120-3140 form */
static double beta7(double s1, double s2, double s3) {
  const double d12 = 13.0 / 12.0, d13 = 1043.0 / 960, d14 = 1.0 / 12;
  return s1 * s1 + d12 * s2 * s2 + d13 * s3 * s3 + d14 * s1 * s3;
}

/* One upwind-biased half of WENO7 (the fPlus part of WENO.cu:3116-3222): candidates at the right face of v[3]
 * from v[0..6], smoothness indicators beta[0..3], Z-type weights alpha_k = d_k (1 + tau^2 / (eps + beta_k)^2)
 * with tau = beta_0 + 3 beta_1 - 3 beta_2 - beta_3 (WENO.cu:3146-3152; linear weights when if_shock is 0). */
double orc_weno7_half(const double *v, double eps, int if_shock, double *beta, double *alpha) {
  const double one6th = 1.0 / 6, one12th = 1.0 / 12;
  const double c0 = 1.0 / 35, c1 = 12.0 / 35, c2 = 18.0 / 35, c3 = 4.0 / 35;
  double s1 = one6th * (-2 * v[0] + 9 * v[1] - 18 * v[2] + 11 * v[3]);
  double s2 = -v[0] + 4 * v[1] - 5 * v[2] + 2 * v[3];
  double s3 = -v[0] + 3 * v[1] - 3 * v[2] + v[3];
  const double b0 = beta7(s1, s2, s3);
  s1 = one6th * (v[1] - 6 * v[2] + 3 * v[3] + 2 * v[4]);
  s2 = v[2] - 2 * v[3] + v[4];
  s3 = -v[1] + 3 * v[2] - 3 * v[3] + v[4];
  const double b1 = beta7(s1, s2, s3);
  s1 = one6th * (-2 * v[2] - 3 * v[3] + 6 * v[4] - v[5]);
  s3 = -v[2] + 3 * v[3] - 3 * v[4] + v[5];
  const double b2 = beta7(s1, s2, s3); /* s2 is shared with b1, as in the reference */
  s1 = one6th * (-11 * v[3] + 18 * v[4] - 9 * v[5] + 2 * v[6]);
  s2 = 2 * v[3] - 5 * v[4] + 4 * v[5] - v[6];
  s3 = -v[3] + 3 * v[4] - 3 * v[5] + v[6];
  const double b3 = beta7(s1, s2, s3);
  double a0 = c0, a1 = c1, a2 = c2, a3 = c3;
  if (if_shock) {
    double tau = b0 + 3 * b1 - 3 * b2 - b3;
    tau *= tau;
    a0 = c0 + c0 * tau / ((eps + b0) * (eps + b0));
    a1 = c1 + c1 * tau / ((eps + b1) * (eps + b1));
    a2 = c2 + c2 * tau / ((eps + b2) * (eps + b2));
    a3 = c3 + c3 * tau / ((eps + b3) * (eps + b3));
  }
  if (beta) { beta[0] = b0; beta[1] = b1; beta[2] = b2; beta[3] = b3; }
  if (alpha) { alpha[0] = a0; alpha[1] = a1; alpha[2] = a2; alpha[3] = a3; }
  const double q0 = -3 * v[0] + 13 * v[1] - 23 * v[2] + 25 * v[3];
  const double q1 = v[1] - 5 * v[2] + 13 * v[3] + 3 * v[4];
  const double q2 = -v[2] + 7 * v[3] + 7 * v[4] - v[5];
  const double q3 = 3 * v[3] + 13 * v[4] - 5 * v[5] + v[6];
  if (!if_shock) return one12th * (c0 * q0 + c1 * q1 + c2 * q2 + c3 * q3); /* UD7 (WENO.cu:3205-3212) */
  return one12th * (a0 * q0 + a1 * q1 + a2 * q2 + a3 * q3) / (a0 + a1 + a2 + a3);
}

/* WENO.cu:3116-3222: fPlus from vp, fMinus from the mirrored vm. */
double orc_weno7(const double *vp, const double *vm, double eps, int if_shock) {
  const double m[7] = {vm[6], vm[5], vm[4], vm[3], vm[2], vm[1], vm[0]};
  return orc_weno7_half(vp, eps, if_shock, NULL, NULL) + orc_weno7_half(m, eps, if_shock, NULL, NULL);
}

/* Roe-averaged interface state and the eigenvector frame (WENO.cu:22-62). */
typedef struct {
  double u, v, w, c, H, kx, ky, kz, nx, ny, nz, qx, qy, qz, Uk, Un, Uq;
} roe_frame;

static void make_frame(const double *ql, const double *qr, const double *k, roe_frame *f) {
  const double rl = ql[0], rr = qr[0];
  const double t = sqrt(rl * rr);
  const double rlc = 1 / (rl + t), rrc = 1 / (t + rr);
  f->u = rlc * ql[1] + rrc * qr[1];
  f->v = rlc * ql[2] + rrc * qr[2];
  f->w = rlc * ql[3] + rrc * qr[3];
  /* enthalpy rho*H = rho*E + p per side, Roe-averaged like the velocity */
  const double pl = (kGamma - 1) * (ql[4] - 0.5 * (ql[1] * ql[1] + ql[2] * ql[2] + ql[3] * ql[3]) / rl);
  const double pr = (kGamma - 1) * (qr[4] - 0.5 * (qr[1] * qr[1] + qr[2] * qr[2] + qr[3] * qr[3]) / rr);
  f->H = rlc * (ql[4] + pl) + rrc * (qr[4] + pr);
  const double ek = 0.5 * (f->u * f->u + f->v * f->v + f->w * f->w);
  f->c = sqrt((kGamma - 1) * (f->H - ek));
  double kx = k[0], ky = k[1], kz = k[2];
  double r = 1.0 / sqrt(kx * kx + ky * ky + kz * kz);
  kx *= r; ky *= r; kz *= r;
  double nx = 0, ny = 0, nz = 0;
  if (fabs(kx) > 0.8) {
    nx = -kz;
    nz = kx;
  } else {
    ny = kz;
    nz = -ky;
  }
  r = 1.0 / sqrt(nx * nx + ny * ny + nz * nz);
  nx *= r; ny *= r; nz *= r;
  double qx = ky * nz - kz * ny, qy = kz * nx - kx * nz, qz = kx * ny - ky * nx;
  r = 1.0 / sqrt(qx * qx + qy * qy + qz * qz);
  qx *= r; qy *= r; qz *= r;
  f->kx = kx; f->ky = ky; f->kz = kz;
  f->nx = nx; f->ny = ny; f->nz = nz;
  f->qx = qx; f->qy = qy; f->qz = qz;
  f->Uk = kx * f->u + ky * f->v + kz * f->w;
  f->Un = nx * f->u + ny * f->v + nz * f->w;
  f->Uq = qx * f->u + qy * f->v + qz * f->w;
}

static void left_matrix(const roe_frame *f, double L[5][5]) {
  const double gm1 = kGamma - 1, b = 1.0 / (f->c * f->c), c = f->c;
  const double al = 0.5 * gm1 * (f->u * f->u + f->v * f->v + f->w * f->w);
  L[0][0] = (al + f->Uk * c) * b * 0.5;
  L[0][1] = -(gm1 * f->u + f->kx * c) * b * 0.5;
  L[0][2] = -(gm1 * f->v + f->ky * c) * b * 0.5;
  L[0][3] = -(gm1 * f->w + f->kz * c) * b * 0.5;
  L[0][4] = gm1 * b * 0.5;
  L[1][0] = -f->Un; L[1][1] = f->nx; L[1][2] = f->ny; L[1][3] = f->nz; L[1][4] = 0;
  L[2][0] = -f->Uq; L[2][1] = f->qx; L[2][2] = f->qy; L[2][3] = f->qz; L[2][4] = 0;
  L[3][0] = 1 - al * b;
  L[3][1] = gm1 * f->u * b;
  L[3][2] = gm1 * f->v * b;
  L[3][3] = gm1 * f->w * b;
  L[3][4] = -gm1 * b;
  L[4][0] = (al - f->Uk * c) * b * 0.5;
  L[4][1] = -(gm1 * f->u - f->kx * c) * b * 0.5;
  L[4][2] = -(gm1 * f->v - f->ky * c) * b * 0.5;
  L[4][3] = -(gm1 * f->w - f->kz * c) * b * 0.5;
  L[4][4] = gm1 * b * 0.5;
}

/* physical = R * characteristic, written out as in WENO.cu:229-243 */
static void right_apply(const roe_frame *f, const double *w, double *q) {
  const double s = w[0] + w[3] + w[4], d = w[0] - w[4], c = f->c;
  q[0] = s;
  q[1] = f->u * s - f->kx * c * d + f->nx * w[1] + f->qx * w[2];
  q[2] = f->v * s - f->ky * c * d + f->ny * w[1] + f->qy * w[2];
  q[3] = f->w * s - f->kz * c * d + f->nz * w[1] + f->qz * w[2];
  q[4] = f->H * s - f->Uk * c * d + f->Un * w[1] + f->Uq * w[2] - c * c / (kGamma - 1) * w[3];
}

void orc_char_matrices(const double *ql, const double *qr, const double *k, double *L, double *R) {
  roe_frame f;
  make_frame(ql, qr, k, &f);
  double l5[5][5];
  left_matrix(&f, l5);
  memcpy(L, l5, sizeof(l5));
  for (int col = 0; col < 5; ++col) {
    double e[5] = {0, 0, 0, 0, 0}, q[5];
    e[col] = 1;
    right_apply(&f, e, q);
    for (int row = 0; row < 5; ++row) R[row * 5 + col] = q[row];
  }
}

/* ---------------- Riemann solvers ---------------- */

/* InviscidScheme.cu:397-469 */
static void ausm_plus(const double *pl, const double *pr, const double *ml, const double *mr, double jl,
                      double jr, double *fc) {
  const double k1 = 0.5 * (jl * ml[0] + jr * mr[0]);
  const double k2 = 0.5 * (jl * ml[1] + jr * mr[1]);
  const double k3 = 0.5 * (jl * ml[2] + jr * mr[2]);
  const double gk = sqrt(k1 * k1 + k2 * k2 + k3 * k3);
  const double ul = (k1 * pl[1] + k2 * pl[2] + k3 * pl[3]) / gk;
  const double ur = (k1 * pr[1] + k2 * pr[2] + k3 * pr[3]) / gk;
  const double c = 0.5 * (sqrt(kGamma * pl[4] / pl[0]) + sqrt(kGamma * pr[4] / pr[0]));
  const double mal = ul / c, mar = ur / c;
  const double alpha = 3 / 16.0;
  double mlp, plp, mrn, prn;
  if (fabs(mal) > 1) {
    mlp = 0.5 * (mal + fabs(mal));
    plp = mlp / mal;
  } else {
    const double a = (mal + 1) * (mal + 1) * 0.25;
    const double b = (mal * mal - 1) * (mal * mal - 1);
    mlp = a + 0.125 * b;
    plp = a * (2 - mal) + alpha * mal * b;
  }
  if (fabs(mar) > 1) {
    mrn = 0.5 * (mar - fabs(mar));
    prn = mrn / mar;
  } else {
    const double a = (mar - 1) * (mar - 1) * 0.25;
    const double b = (mar * mar - 1) * (mar * mar - 1);
    mrn = -a - 0.125 * b;
    prn = a * (2 + mar) - alpha * mar * b;
  }
  const double pc = plp * pl[4] + prn * pr[4];
  const double mh = mlp + mrn;
  const double mpos = 0.5 * (mh + fabs(mh)), mneg = 0.5 * (mh - fabs(mh));
  const double mass = c * (pl[0] * mpos + pr[0] * mneg);
  const double coeff = mass * gk;
  const double *s = mass >= 0 ? pl : pr;
  fc[0] = coeff;
  fc[1] = coeff * s[1] + pc * k1;
  fc[2] = coeff * s[2] + pc * k2;
  fc[3] = coeff * s[3] + pc * k3;
  fc[4] = coeff * (s[5] + s[4]) / s[0];
}

/* InviscidScheme.cu:471-599 (Air branch) */
static void hllc(const double *pl, const double *pr, const double *ml, const double *mr, double jl, double jr,
                 double *fc) {
  const double sl_ = sqrt(pl[0]), sr_ = sqrt(pr[0]);
  const double rlc = sl_ / (sl_ + sr_), rrc = sr_ / (sl_ + sr_);
  const double ut = pl[1] * rlc + pr[1] * rrc;
  const double vt = pl[2] * rlc + pr[2] * rrc;
  const double wt = pl[3] * rlc + pr[3] * rrc;
  const double hl = (pl[5] + pl[4]) / pl[0], hr = (pr[5] + pr[4]) / pr[0];
  const double ht = hl * rlc + hr * rrc;
  const double ekt = 0.5 * (ut * ut + vt * vt + wt * wt);
  const double ct = sqrt((kGamma - 1) * (ht - ekt));
  const double kx = 0.5 * (ml[0] + mr[0]), ky = 0.5 * (ml[1] + mr[1]), kz = 0.5 * (ml[2] + mr[2]);
  const double gk = sqrt(kx * kx + ky * ky + kz * kz);
  const double Ut = (kx * ut + ky * vt + kz * wt) / gk;
  const double ja = 0.5 * (jl + jr);
  const double Ul = kx * pl[1] + ky * pl[2] + kz * pl[3];
  const double cl = sqrt(kGamma * pl[4] / pl[0]);
  const double sl = fmin(Ul / gk - cl, Ut - ct);
  if (sl >= 0) {
    const double m = pl[0] * Ul;
    fc[0] = ja * m;
    fc[1] = ja * (m * pl[1] + pl[4] * kx);
    fc[2] = ja * (m * pl[2] + pl[4] * ky);
    fc[3] = ja * (m * pl[3] + pl[4] * kz);
    fc[4] = ja * ((pl[5] + pl[4]) * Ul);
    return;
  }
  const double Ur = kx * pr[1] + ky * pr[2] + kz * pr[3];
  const double cr = sqrt(kGamma * pr[4] / pr[0]);
  const double sr = fmax(Ur / gk + cr, Ut + ct);
  if (sr < 0) {
    const double m = pr[0] * Ur;
    fc[0] = ja * m;
    fc[1] = ja * (m * pr[1] + pr[4] * kx);
    fc[2] = ja * (m * pr[2] + pr[4] * ky);
    fc[3] = ja * (m * pr[3] + pr[4] * kz);
    fc[4] = ja * ((pr[5] + pr[4]) * Ur);
    return;
  }
  const double sm = ((pr[0] * Ur * (sr - Ur / gk) - pl[0] * Ul * (sl - Ul / gk)) / gk + pl[4] - pr[4]) /
                    (pr[0] * (sr - Ur / gk) - pl[0] * (sl - Ul / gk));
  const double pm = pl[0] * (sl - Ul / gk) * (sm - Ul / gk) + pl[4];
  if (sm >= 0) {
    const double pco = 1.0 / (sl - sm);
    const double Qc = ja * pco * sm * (sl * gk - Ul) * pl[0];
    fc[0] = Qc;
    const double dP = (sl * pm - sm * pl[4]) * pco * ja;
    fc[1] = Qc * pl[1] + dP * kx;
    fc[2] = Qc * pl[2] + dP * ky;
    fc[3] = Qc * pl[3] + dP * kz;
    fc[4] = Qc * pl[5] / pl[0] + pco * ja * (sl * pm * sm * gk - sm * pl[4] * Ul);
  } else {
    const double pco = 1.0 / (sr - sm);
    const double Qc = ja * pco * sm * (sr * gk - Ur) * pr[0];
    fc[0] = Qc;
    const double dP = (sr * pm - sm * pr[4]) * pco * ja;
    fc[1] = Qc * pr[1] + dP * kx;
    fc[2] = Qc * pr[2] + dP * ky;
    fc[3] = Qc * pr[3] + dP * kz;
    fc[4] = Qc * pr[5] / pr[0] + pco * ja * (sr * pm * sm * gk - sm * pr[4] * Ur);
  }
}

/* InviscidScheme.cu:601-765 (n_spec = 0) */
static void roe(const double *pl, const double *pr, const double *ml, const double *mr, double jl, double jr,
                double efl, double efr, double *fc) {
  /* half sum of the left and right physical fluxes, InviscidScheme.cu:601-633 */
  double JKx = jl * ml[0], JKy = jl * ml[1], JKz = jl * ml[2];
  double Uk = pl[1] * JKx + pl[2] * JKy + pl[3] * JKz;
  double co = Uk * pl[0];
  fc[0] = 0.5 * co;
  fc[1] = 0.5 * (co * pl[1] + pl[4] * JKx);
  fc[2] = 0.5 * (co * pl[2] + pl[4] * JKy);
  fc[3] = 0.5 * (co * pl[3] + pl[4] * JKz);
  fc[4] = 0.5 * Uk * (pl[4] + pl[5]);
  JKx = jr * mr[0]; JKy = jr * mr[1]; JKz = jr * mr[2];
  Uk = pr[1] * JKx + pr[2] * JKy + pr[3] * JKz;
  co = Uk * pr[0];
  fc[0] += 0.5 * co;
  fc[1] += 0.5 * (co * pr[1] + pr[4] * JKx);
  fc[2] += 0.5 * (co * pr[2] + pr[4] * JKy);
  fc[3] += 0.5 * (co * pr[3] + pr[4] * JKz);
  fc[4] += 0.5 * Uk * (pr[4] + pr[5]);

  const double dl = sqrt(pl[0]), dr = sqrt(pr[0]);
  const double inv = 1.0 / (dl + dr);
  const double u = (dl * pl[1] + dr * pr[1]) * inv;
  const double v = (dl * pl[2] + dr * pr[2]) * inv;
  const double w = (dl * pl[3] + dr * pr[3]) * inv;
  const double ek = 0.5 * (u * u + v * v + w * w);
  const double hl = (pl[5] + pl[4]) / pl[0], hr = (pr[5] + pr[4]) / pr[0];
  const double h = (dl * hl + dr * hr) * inv;
  const double g = kGamma;
  const double c = sqrt((g - 1) * (h - ek));

  double kx = 0.5 * (ml[0] + mr[0]), ky = 0.5 * (ml[1] + mr[1]), kz = 0.5 * (ml[2] + mr[2]);
  const double gk = sqrt(kx * kx + ky * ky + kz * kz);
  Uk = kx * u + ky * v + kz * w;
  double ch[3] = {Uk - gk * c, Uk, Uk + gk * c};
  const double efa = 0.5 * (efl + efr);
  for (int m = 0; m < 3; ++m) {
    ch[m] = fabs(ch[m]);
    if (ch[m] < efa) ch[m] = 0.5 * (ch[m] * ch[m] / efa + efa);
  }
  kx /= gk; ky /= gk; kz /= gk;
  Uk /= gk;
  const double ja = 0.5 * (jl + jr);
  double dq[5];
  dq[0] = ja * (pr[0] - pl[0]);
  for (int l = 1; l < 5; ++l) dq[l] = ja * (pr[0] * pr[l] - pl[0] * pl[l]);
  dq[4] = ja * (pr[5] - pl[5]);
  double c1 = (g - 1) * (ek * dq[0] - u * dq[1] - v * dq[2] - w * dq[3] + dq[4]) / (c * c);
  double c2 = (kx * dq[1] + ky * dq[2] + kz * dq[3] - Uk * dq[0]) / c;
  double c3 = dq[0] - c1;
  double L[5];
  L[0] = 0.5 * (c1 - c2);
  L[1] = kx * c3 - ((kz * v - ky * w) * dq[0] - kz * dq[2] + ky * dq[3]) / c;
  L[2] = ky * c3 - ((kx * w - kz * u) * dq[0] - kx * dq[3] + kz * dq[1]) / c;
  L[3] = kz * c3 - ((ky * u - kx * v) * dq[0] - ky * dq[1] + kx * dq[2]) / c;
  L[4] = 0.5 * (c1 + c2);
  double b[5];
  b[0] = -ch[0] * L[0];
  for (int l = 1; l < 5; ++l) b[l] = -ch[1] * L[l];
  b[4] = -ch[2] * L[4];
  const double c0 = kx * b[1] + ky * b[2] + kz * b[3];
  c1 = c0 + b[0] + b[4];
  c2 = c * (b[4] - b[0]);
  c3 = 0;
  fc[0] += 0.5 * c1;
  fc[1] += 0.5 * (u * c1 + kx * c2 - c * (kz * b[2] - ky * b[3]));
  fc[2] += 0.5 * (v * c1 + ky * c2 - c * (kx * b[3] - kz * b[1]));
  fc[3] += 0.5 * (w * c1 + kz * c2 - c * (ky * b[1] - kx * b[2]));
  fc[4] += 0.5 * (h * c1 + Uk * c2 - c * c * c0 / (g - 1) +
                  c * ((kz * v - ky * w) * b[1] + (kx * w - kz * u) * b[2] + (ky * u - kx * v) * b[3]) + c3);
}

/* InviscidScheme.cu:772-805 */
static void lax_friedrichs(const double *pl, const double *pr, const double *ml, const double *mr, double jl,
                           double jr, double *fc) {
  const double kx = 0.5 * (ml[0] + mr[0]), ky = 0.5 * (ml[1] + mr[1]), kz = 0.5 * (ml[2] + mr[2]);
  const double gk = sqrt(kx * kx + ky * ky + kz * kz);
  const double Ukl = pl[1] * kx + pl[2] * ky + pl[3] * kz;
  const double Ukr = pr[1] * kx + pr[2] * ky + pr[3] * kz;
  const double cl = sqrt(kGamma * pl[4] / pl[0]), cr = sqrt(kGamma * pr[4] / pr[0]);
  const double sp = fmax(fabs(Ukl) + cl * gk, fabs(Ukr) + cr * gk);
  const double hj = 0.5 * 0.5 * (jl + jr);
  const double rUl = pl[0] * Ukl, rUr = pr[0] * Ukr;
  fc[0] = (rUl + rUr - sp * (pr[0] - pl[0])) * hj;
  fc[1] = (rUl * pl[1] + rUr * pr[1] + kx * (pl[4] + pr[4]) - sp * (pr[1] * pr[0] - pl[1] * pl[0])) * hj;
  fc[2] = (rUl * pl[2] + rUr * pr[2] + ky * (pl[4] + pr[4]) - sp * (pr[2] * pr[0] - pl[2] * pl[0])) * hj;
  fc[3] = (rUl * pl[3] + rUr * pr[3] + kz * (pl[4] + pr[4]) - sp * (pr[3] * pr[0] - pl[3] * pl[0])) * hj;
  fc[4] = ((pl[5] + pl[4]) * Ukl + (pr[5] + pr[4]) * Ukr - sp * (pr[5] - pl[5])) * hj;
}

void orc_riemann(int scheme, const double *pl, const double *pr, const double *ml, const double *mr, double jl,
                 double jr, double efl, double efr, int n_dim_unused, double *fc) {
  (void)n_dim_unused;
  switch (scheme) {
    case 1: lax_friedrichs(pl, pr, ml, mr, jl, jr, fc); break;
    case 2: roe(pl, pr, ml, mr, jl, jr, efl, efr, fc); break;
    case 4: hllc(pl, pr, ml, mr, jl, jr, fc); break;
    case 3:
    default: ausm_plus(pl, pr, ml, mr, jl, jr, fc); break;
  }
}

/* conservative -> (rho,u,v,w,p,E) */
static void cons_to_pv(const double *q, double *pv) {
  pv[0] = q[0];
  const double ri = 1.0 / q[0];
  pv[1] = q[1] * ri;
  pv[2] = q[2] * ri;
  pv[3] = q[3] * ri;
  pv[4] = (kGamma - 1) * (q[4] - 0.5 * (q[1] * pv[1] + q[2] * pv[2] + q[3] * pv[3]));
  pv[5] = q[4];
}

/*
 * The product's WENO + Riemann pairing (reconstruction 4/5 with inviscid_scheme 1-4): the
 * reference declares reconstruction 4/5 as WENO reconstructions of conservative variables
 * with optional characteristic treatment (Parameter.cpp:211-236) but leaves the path
 * unimplemented.  Characteristic fields W = L Q (L of WENO.cu:73-124 at the Roe average of
 * the two interface nodes), one-sided WENO halves of WENO5/WENO7 (WENO.cu:3059-3222: the
 * left state is WENO(vp=W_{i-r..i+r}, vm=0), the right state WENO(vp=0, vm=W_{i-r+1..i+r+1})),
 * back-projection Q = R W, then the reference Riemann solver on the two states.
 */
int orc_interface_flux(const orc_params *p, const double (*q)[5], int n, const double *ml, const double *mr,
                       double jl, double jr, double efl, double efr, double *fc) {
  const int c = n / 2 - 1; /* index of the left node of the interface */
  double km[3] = {0.5 * (ml[0] + mr[0]), 0.5 * (ml[1] + mr[1]), 0.5 * (ml[2] + mr[2])};
  roe_frame f;
  make_frame(q[c], q[c + 1], km, &f);
  double L[5][5];
  left_matrix(&f, L);
  double W[8][5];
  for (int m = 0; m < n; ++m)
    for (int l = 0; l < 5; ++l)
      W[m][l] = L[l][0] * q[m][0] + L[l][1] * q[m][1] + L[l][2] * q[m][2] + L[l][3] * q[m][3] + L[l][4] * q[m][4];
  double wl[5], wr[5];
  const double zero[7] = {0, 0, 0, 0, 0, 0, 0};
  for (int l = 0; l < 5; ++l) {
    double s[8];
    for (int m = 0; m < n; ++m) s[m] = W[m][l];
    if (n == 8) {
      wl[l] = orc_weno7(s, zero, p->weno_eps, 1);
      wr[l] = orc_weno7(zero, s + 1, p->weno_eps, 1);
    } else {
      wl[l] = orc_weno5(s, zero, p->weno_eps, 1);
      wr[l] = orc_weno5(zero, s + 1, p->weno_eps, 1);
    }
  }
  double qL[5], qR[5], pl[6], pr[6];
  right_apply(&f, wl, qL);
  right_apply(&f, wr, qR);
  cons_to_pv(qL, pl);
  cons_to_pv(qR, pr);
  int fallback = 0;
  if (!(pl[0] > 0 && pl[4] > 0 && pr[0] > 0 && pr[4] > 0)) {
    /* positivity fallback to the first-order states of the two interface nodes */
    cons_to_pv(q[c], pl);
    cons_to_pv(q[c + 1], pr);
    fallback = 1;
  }
  orc_riemann(p->inviscid_scheme, pl, pr, ml, mr, jl, jr, efl, efr, 3, fc);
  return fallback;
}

/* minmod (Limiter.cuh:8-41; gxl::sgn(0) = 1 there, which changes nothing: min(|a|, |b|) = 0 when a or b is 0). */
static double minmod(double a, double b) {
  const double sa = a < 0 ? -1.0 : 1.0, sb = b < 0 ? -1.0 : 1.0;
  return 0.5 * (sa + sb) * fmin(fabs(a), fabs(b));
}

/* Primitive-variable interface flux for air (reconstruction 1/2/3 with inviscid_scheme 1-4): rho, u, v, w, p of the
 * four nodes pv[0..3] (the interface between pv[1] and pv[2]) reconstructed as first_order_reconstruct,
 * MUSCL_reconstruct (kappa = 1/3) or NND2_reconstruct (Reconstruction.cu:5-44) with minmod, the total energy of each
 * side from its own state (reconstruction<Air>, InviscidScheme.cu:212-217), then the Riemann solver
 * (compute_convective_term_pv_1D and Roe_compute_inviscid_flux_1D, InviscidScheme.cu:61-169, 306-395). */
void orc_interface_flux_pv(const orc_params *p, const double (*pv)[5], const double *ml, const double *mr, double jl,
                           double jr, double efl, double efr, double *fc) {
  double pl[6], pr[6];
  for (int l = 0; l < 5; ++l) {
    if (p->reconstruction == 1) {
      pl[l] = pv[1][l];
      pr[l] = pv[2][l];
      continue;
    }
    const double d0 = pv[1][l] - pv[0][l], d1 = pv[2][l] - pv[1][l], d2 = pv[3][l] - pv[2][l];
    if (p->reconstruction == 2) {
      const double kappa = 1.0 / 3.0;
      pl[l] = pv[1][l] + 0.25 * ((1 - kappa) * minmod(d0, d1) + (1 + kappa) * minmod(d1, d0));
      pr[l] = pv[2][l] - 0.25 * ((1 - kappa) * minmod(d2, d1) + (1 + kappa) * minmod(d1, d2));
    } else {
      pl[l] = pv[1][l] + 0.5 * minmod(d0, d1);
      pr[l] = pv[2][l] - 0.5 * minmod(d1, d2);
    }
  }
  const double el = 0.5 * (pl[1] * pl[1] + pl[2] * pl[2] + pl[3] * pl[3]);
  const double er = 0.5 * (pr[1] * pr[1] + pr[2] * pr[2] + pr[3] * pr[3]);
  pl[5] = el * pl[0] + pl[4] / (kGamma - 1);
  pr[5] = er * pr[0] + pr[4] / (kGamma - 1);
  orc_riemann(p->inviscid_scheme, pl, pr, ml, mr, jl, jr, efl, efr, 3, fc);
}

/* Pressure, temperature and the point quantities both LLF-split WENO schemes build their stencils
 * from (WENO.cu:890-917 component-wise, WENO.cu:1096-1136 characteristic): Uk with the unscaled
 * metric, cGradK = |grad k| sqrt(gamma R T) with T = p / (rho R) as the primitive update stores it. */
typedef struct {
  double p, Uk, cgk;
} split_point;

static split_point split_point_of(const double *q, const double *k) {
  split_point s;
  double pv[6];
  cons_to_pv(q, pv);
  const double R = kRu / kMwAir;
  const double T = pv[4] / (pv[0] * R);
  s.p = pv[4];
  s.Uk = (q[1] * k[0] + q[2] * k[1] + q[3] * k[2]) / q[0];
  s.cgk = sqrt(k[0] * k[0] + k[1] * k[1] + k[2] * k[2]) * sqrt(kGamma * R * T);
  return s;
}

/* Component-wise LLF flux splitting (schemes 51/71): WENO.cu:890-990 for the air model. Each node splits
 * its own flux with its own spectral radius |Uk| + cGradK; each conservative variable is then
 * reconstructed by WENO5/WENO7 with the epsilon scaled by v_ref^0, v_ref^2, v_ref^4. */
static void fsweno_componentwise(const orc_params *p, const double (*q)[5], const double (*met)[3],
                                 const double *jac, int n, int if_shock, double *fc) {
  double fp[8][5], fm[8][5];
  for (int m = 0; m < n; ++m) {
    const split_point s = split_point_of(q[m], met[m]);
    const double lam = fabs(s.Uk) + s.cgk;
    const double sp = 0.5 * jac[m] * (s.Uk + lam), sm = 0.5 * jac[m] * (s.Uk - lam);
    const double pk = s.p * (0.5 * jac[m]);
    fp[m][0] = sp * q[m][0];
    fm[m][0] = sm * q[m][0];
    for (int c = 0; c < 3; ++c) {
      fp[m][1 + c] = fma(sp, q[m][1 + c], pk * met[m][c]);
      fm[m][1 + c] = fma(sm, q[m][1 + c], pk * met[m][c]);
    }
    fp[m][4] = fma(sp, q[m][4], pk * s.Uk);
    fm[m][4] = fma(sm, q[m][4], pk * s.Uk);
  }
  const double e1 = p->weno_eps * p->v_ref * p->v_ref;
  const double eps[5] = {p->weno_eps, e1, e1, e1, e1 * p->v_ref * p->v_ref};
  for (int l = 0; l < 5; ++l) {
    double vp[7], vm[7];
    for (int m = 0; m < n - 1; ++m) {
      vp[m] = fp[m][l];
      vm[m] = fm[m + 1][l];
    }
    fc[l] = n == 8 ? orc_weno7(vp, vm, eps[l], if_shock) : orc_weno5(vp, vm, eps[l], if_shock);
  }
}

/* Characteristic LLF flux splitting for air (schemes 52/72): WENO.cu:1090-1145 builds the stencil,
 * weno_ch_air_x (WENO.cu:10-244) projects with the left eigenvectors of the Roe state, splits with the
 * largest spectral radius on the stencil, reconstructs each field and projects back. */
static void fsweno_characteristic(const orc_params *p, const double (*q)[5], const double (*met)[3],
                                  const double *jac, int n, int if_shock, double *fc) {
  const int c = n / 2 - 1;
  split_point s[8];
  double F[8][5]; /* F[m][0..3] = Uk, rho u Uk + p kx, rho v Uk + p ky, rho w Uk + p kz; F[m][4] = p */
  double max_lambda = 0;
  for (int m = 0; m < n; ++m) {
    s[m] = split_point_of(q[m], met[m]);
    F[m][0] = s[m].Uk;
    for (int d = 0; d < 3; ++d) F[m][1 + d] = fma(q[m][1 + d], s[m].Uk, s[m].p * met[m][d]);
    F[m][4] = s[m].p;
    const double lam = fabs(s[m].Uk) + s[m].cgk;
    if (lam > max_lambda) max_lambda = lam;
  }
  /* Roe state of the two interface nodes; the sound speed from the Roe-weighted p/rho (WENO.cu:21-30) */
  const double R = kRu / kMwAir, gm1 = kGamma - 1;
  const double t1 = sqrt(q[c][0] * q[c + 1][0]);
  const double rlc = 1 / (q[c][0] + t1), rrc = 1 / (t1 + q[c + 1][0]);
  const double um = rlc * q[c][1] + rrc * q[c + 1][1];
  const double vm = rlc * q[c][2] + rrc * q[c + 1][2];
  const double wm = rlc * q[c][3] + rrc * q[c + 1][3];
  const double Tm = (rlc * F[c][4] + rrc * F[c + 1][4]) / R;
  const double cm = sqrt(kGamma * R * Tm);
  /* frame (WENO.cu:33-64) */
  double kx = 0.5 * (met[c][0] + met[c + 1][0]), ky = 0.5 * (met[c][1] + met[c + 1][1]);
  double kz = 0.5 * (met[c][2] + met[c + 1][2]);
  double r = 1 / sqrt(kx * kx + ky * ky + kz * kz);
  kx *= r;
  ky *= r;
  kz *= r;
  const double Uk_bar = kx * um + ky * vm + kz * wm;
  double nx = 0, ny = 0, nz = 0;
  if (fabs(kx) > 0.8) {
    nx = -kz;
    nz = kx;
  } else {
    ny = kz;
    nz = -ky;
  }
  r = 1 / sqrt(nx * nx + ny * ny + nz * nz);
  nx *= r;
  ny *= r;
  nz *= r;
  double qx = ky * nz - kz * ny, qy = kz * nx - kx * nz, qz = kx * ny - ky * nx;
  r = 1 / sqrt(qx * qx + qy * qy + qz * qz);
  qx *= r;
  qy *= r;
  qz *= r;
  const double Un = nx * um + ny * vm + nz * wm, Uq = qx * um + qy * vm + qz * wm;
  const double ic2 = 1.0 / (cm * cm), al = 0.5 * gm1 * (um * um + vm * vm + wm * wm);
  const double Lm[5][5] = {
      {(al + Uk_bar * cm) * ic2 * 0.5, -(gm1 * um + kx * cm) * ic2 * 0.5, -(gm1 * vm + ky * cm) * ic2 * 0.5,
       -(gm1 * wm + kz * cm) * ic2 * 0.5, gm1 * ic2 * 0.5},
      {-Un, nx, ny, nz, 0},
      {-Uq, qx, qy, qz, 0},
      {1 - al * ic2, gm1 * um * ic2, gm1 * vm * ic2, gm1 * wm * ic2, -gm1 * ic2},
      {(al - Uk_bar * cm) * ic2 * 0.5, -(gm1 * um - kx * cm) * ic2 * 0.5, -(gm1 * vm - ky * cm) * ic2 * 0.5,
       -(gm1 * wm - kz * cm) * ic2 * 0.5, gm1 * ic2 * 0.5}};
  double fch[5];
  for (int l = 0; l < 5; ++l) {
    const double *L = Lm[l];
    double vp[7], vmn[7];
    for (int m = 0; m < n; ++m) {
      const double LF = L[0] * (q[m][0] * F[m][0]) + L[1] * F[m][1] + L[2] * F[m][2] + L[3] * F[m][3] +
                        L[4] * ((q[m][4] + F[m][4]) * F[m][0]);
      const double LQ = L[0] * q[m][0] + L[1] * q[m][1] + L[2] * q[m][2] + L[3] * q[m][3] + L[4] * q[m][4];
      if (m < n - 1) vp[m] = 0.5 * jac[m] * (LF + max_lambda * LQ);
      if (m > 0) vmn[m - 1] = 0.5 * jac[m] * (LF - max_lambda * LQ);
    }
    fch[l] = n == 8 ? orc_weno7(vp, vmn, p->weno_eps, if_shock) : orc_weno5(vp, vmn, p->weno_eps, if_shock);
  }
  /* back-projection (WENO.cu:229-243) with the Roe-averaged enthalpy */
  const double s1 = fch[0] + fch[3] + fch[4], s3 = fch[0] - fch[4];
  fc[0] = s1;
  fc[1] = um * s1 - kx * cm * s3 + nx * fch[1] + qx * fch[2];
  fc[2] = vm * s1 - ky * cm * s3 + ny * fch[1] + qy * fch[2];
  fc[3] = wm * s1 - kz * cm * s3 + nz * fch[1] + qz * fch[2];
  const double H = rlc * (q[c][4] + F[c][4]) + rrc * (q[c + 1][4] + F[c + 1][4]);
  fc[4] = H * s1 - Uk_bar * cm * s3 + Un * fch[1] + Uq * fch[2] - cm * cm / gm1 * fch[3];
}

void orc_fsweno_flux_hybrid(const orc_params *p, const double (*q)[5], const double (*met)[3], const double *jac,
                            int n, int if_shock, double *fc) {
  if (p->inviscid_scheme % 10 == 1) fsweno_componentwise(p, q, met, jac, n, if_shock, fc);
  else fsweno_characteristic(p, q, met, jac, n, if_shock, fc);
}

void orc_fsweno_flux(const orc_params *p, const double (*q)[5], const double (*met)[3], const double *jac, int n,
                     double *fc) {
  orc_fsweno_flux_hybrid(p, q, met, jac, n, 1, fc);
}

/* The shock sensors of compute_shock_sensor (FieldOperation.cu:74-179), node (i, j, k) of a block with ngg
 * ghosts.  Ducros (0): 0 where the fourth-order central divergence of the velocity is positive, else
 * div^2 / (div^2 + |curl|^2 + eps); Jameson (1): the sum over the three directions of
 * |p_+ + p_- - 2p| / (p_+ + p_- + 2p); jump (2): the largest relative forward jump of density or pressure. */
void orc_shock_sensor(int type, double eps, int mx, int my, int mz, int ngg, const double *bv, const double *met,
                      double *sensor) {
  const long d2 = mx + 2 * ngg, d1 = (long)(my + 2 * ngg) * d2, sz = d1 * (mz + 2 * ngg);
#pragma omp parallel for collapse(2)
  for (int k = 0; k < mz; ++k)
    for (int j = 0; j < my; ++j)
      for (int i = 0; i < mx; ++i) {
        const long id = (long)(k + ngg) * d1 + (long)(j + ngg) * d2 + (i + ngg);
        const long st[3] = {1, d2, d1};
        double s = 1;
        if (type == 1) {
          const double p2 = 2 * bv[id + 4 * sz];
          s = 0;
          for (int d = 0; d < 3; ++d) {
            const double pp = bv[id + st[d] + 4 * sz] + bv[id - st[d] + 4 * sz];
            s += fabs(pp - p2) / (pp + p2);
          }
        } else if (type == 2) {
          const double r = bv[id], p = bv[id + 4 * sz];
          s = 0;
          for (int d = 0; d < 3; ++d) s = fmax(s, fabs(bv[id + st[d]] - r) / r);
          for (int d = 0; d < 3; ++d) s = fmax(s, fabs(bv[id + st[d] + 4 * sz] - p) / p);
        } else {
          double g[3][3]; /* g[c][d]: d(velocity c)/d(index d), fourth-order central */
          for (int c = 0; c < 3; ++c)
            for (int d = 0; d < 3; ++d) {
              const double *f = bv + (1 + c) * sz + id;
              g[c][d] = (1.0 / 12) * f[-2 * st[d]] - (2.0 / 3) * f[-st[d]] + (2.0 / 3) * f[st[d]] -
                        (1.0 / 12) * f[2 * st[d]];
            }
          double D[3][3]; /* D[c][x]: d(velocity c)/d(x) with the metrics */
          for (int c = 0; c < 3; ++c)
            for (int x = 0; x < 3; ++x)
              D[c][x] = g[c][0] * met[id + x * sz] + g[c][1] * met[id + (3 + x) * sz] + g[c][2] * met[id + (6 + x) * sz];
          double div = D[0][0] + D[1][1] + D[2][2];
          if (div > 0) {
            s = 0;
          } else {
            div *= div;
            const double cx = D[2][1] - D[1][2], cy = D[0][2] - D[2][0], cz = D[1][0] - D[0][1];
            s = div / (div + cx * cx + cy * cy + cz * cz + eps);
          }
        }
        sensor[id] = s;
      }
}

/* Transport.cu:6-9 */
double orc_sutherland(double T) {
  return 1.7894e-5 * pow(T / 288.16, 1.5) * (288.16 + 110) / (T + 110);
}

/* ---------------- exact Riemann solver (Toro ch. 4) ---------------- */

static void pressure_function(double p, double rk, double pk, double ck, double g, double *f, double *df) {
  if (p <= pk) { /* rarefaction */
    const double pr = p / pk;
    *f = 2 * ck / (g - 1) * (pow(pr, (g - 1) / (2 * g)) - 1);
    *df = 1 / (rk * ck) * pow(pr, -(g + 1) / (2 * g));
  } else { /* shock */
    const double A = 2 / ((g + 1) * rk), B = (g - 1) / (g + 1) * pk;
    const double s = sqrt(A / (p + B));
    *f = (p - pk) * s;
    *df = s * (1 - (p - pk) / (2 * (B + p)));
  }
}

double orc_exact_riemann(double rl, double ul, double pl, double rr, double ur, double pr, double g, double x0,
                         double t, int n, const double *x, double *rho, double *u, double *p) {
  const double cl = sqrt(g * pl / rl), cr = sqrt(g * pr / rr);
  double ps = 0.5 * (pl + pr) - 0.125 * (ur - ul) * (rl + rr) * (cl + cr);
  if (ps < 1e-12) ps = 1e-12;
  for (int it = 0; it < 100; ++it) {
    double fl, dfl, fr, dfr;
    pressure_function(ps, rl, pl, cl, g, &fl, &dfl);
    pressure_function(ps, rr, pr, cr, g, &fr, &dfr);
    double pn = ps - (fl + fr + ur - ul) / (dfl + dfr);
    if (pn < 1e-12) pn = 1e-12;
    const double ch = 2 * fabs(pn - ps) / (pn + ps);
    ps = pn;
    if (ch < 1e-15) break;
  }
  double fl, dfl, fr, dfr;
  pressure_function(ps, rl, pl, cl, g, &fl, &dfl);
  pressure_function(ps, rr, pr, cr, g, &fr, &dfr);
  const double us = 0.5 * (ul + ur) + 0.5 * (fr - fl);
  const double gm = (g - 1) / (g + 1);
  for (int m = 0; m < n; ++m) {
    const double S = (x[m] - x0) / t;
    double r, uu, pp;
    if (S <= us) { /* left of contact */
      if (ps > pl) {
        const double SL = ul - cl * sqrt((g + 1) / (2 * g) * ps / pl + (g - 1) / (2 * g));
        if (S <= SL) { r = rl; uu = ul; pp = pl; }
        else { r = rl * (ps / pl + gm) / (gm * ps / pl + 1); uu = us; pp = ps; }
      } else {
        const double SHL = ul - cl, csl = cl * pow(ps / pl, (g - 1) / (2 * g)), STL = us - csl;
        if (S <= SHL) { r = rl; uu = ul; pp = pl; }
        else if (S > STL) { r = rl * pow(ps / pl, 1 / g); uu = us; pp = ps; }
        else {
          const double cc = 2 / (g + 1) + gm / cl * (ul - S);
          r = rl * pow(cc, 2 / (g - 1));
          uu = 2 / (g + 1) * (cl + (g - 1) / 2 * ul + S);
          pp = pl * pow(cc, 2 * g / (g - 1));
        }
      }
    } else {
      if (ps > pr) {
        const double SR = ur + cr * sqrt((g + 1) / (2 * g) * ps / pr + (g - 1) / (2 * g));
        if (S >= SR) { r = rr; uu = ur; pp = pr; }
        else { r = rr * (ps / pr + gm) / (gm * ps / pr + 1); uu = us; pp = ps; }
      } else {
        const double SHR = ur + cr, csr = cr * pow(ps / pr, (g - 1) / (2 * g)), STR = us + csr;
        if (S >= SHR) { r = rr; uu = ur; pp = pr; }
        else if (S <= STR) { r = rr * pow(ps / pr, 1 / g); uu = us; pp = ps; }
        else {
          const double cc = 2 / (g + 1) - gm / cr * (ur - S);
          r = rr * pow(cc, 2 / (g - 1));
          uu = 2 / (g + 1) * (-cr + (g - 1) / 2 * ur + S);
          pp = pr * pow(cc, 2 * g / (g - 1));
        }
      }
    }
    rho[m] = r; u[m] = uu; p[m] = pp;
  }
  return ps;
}

/* unused constant kept for the Air temperature relation T = p mw / (rho R_u), RK.cuh:278 */
double orc_air_temperature(double p, double rho) { return p * kMwAir / rho / kRu; }
