/*
 * oracle_mix.c -- multi-species (mixture) inviscid right-hand side, SSP-RK3 step and Strang-split
 * chemistry of one block (TEST INFRASTRUCTURE ONLY; see oracle.h).
 *
 * Primitive-variable path of the reference (compute_convective_term_pv / Roe_compute_inviscid_flux,
 * InviscidScheme.cu:28-219, 221-347): reconstruction of (rho, u, v, w, p, Y_k) by first order, MUSCL or
 * NND2 (Reconstruction.cu:5-44, minmod Limiter.cuh:8-41) or, as the new pairing, component-wise WENO5-JS /
 * WENO7-Z one-sided halves (WENO.cu:3059-3222 with a zero opposite stencil), then the reference
 * post-processing for E and gamma (InviscidScheme.cu:184-218) and the mixture Riemann solvers (Roe
 * 635-765 with the entropy fix 269-304, AUSM+ 397-469, HLLC 471-599).
 */
#include "oracle.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>

static const double kRu = 8314.462618;

#define MS ORC_MAX_SPEC
#define NV (5 + MS)

typedef struct {
  int face, dir, type;
  int s[3], e[3];
} mface_t;

struct orc_mix {
  orc_params p;
  orc_mech m;
  int ns, nv, mx, my, mz, g, dim;
  long d2, d1, sz, n;
  mface_t faces[64];
  int n_face;
  int explicit_source; /* chemical source in the residual (reaction = 1, temporal_scheme 3) */
  double *cv, *qn, *bv, *sv, *c, *efd, *dq, *jac, *met, *hstep, *tscale;
  double *cn, *wu_ts; /* Wu's splitting: frozen transport residual, reaction time scale (starts at zero) */
  double pp_dt;       /* step size seen by the positivity limiter (the global dt of the current step) */
  struct {
    double *prof, *rms;
    int sub_node, sub_ghost, fluct;
    double mw;
    unsigned long long seed, call;
  } inflow[64];       /* inflow faces (orc_mix_set_inflow) */
  double *sponge_sigma, *sponge_mean; /* sponge layer: damping rate (n), running mean of cv (nv * n) */
  unsigned long long sponge_iter;
  double *sensor; /* shock sensor (sz; ghost entries stay 0) of the hybrid flux-splitting scheme */
  int viscous;        /* viscous_order 2 or 8 with transport data set */
  int bt[6];          /* physical-boundary flags per side for the one-sided collocated differences */
  orc_tran tran;
  double *mul, *cond, *rhod, *fvis[3];
};

#define ID(b, i, j, k) ((long)((k) + (b)->g) * (b)->d1 + (long)((j) + (b)->g) * (b)->d2 + ((i) + (b)->g))
#define IDQ(b, i, j, k) (((long)(k) * (b)->my + (j)) * (b)->mx + (i))

orc_mix *orc_mix_create(const orc_params *p, const orc_mech *m, int mx, int my, int mz, int ngg, const int *faces,
                        int n_face, const double *jac, const double *metric) {
  orc_mix *b = (orc_mix *)calloc(1, sizeof(orc_mix));
  b->p = *p;
  b->m = *m;
  b->ns = m->ns;
  b->nv = 5 + m->ns;
  b->mx = mx; b->my = my; b->mz = mz; b->g = ngg;
  b->dim = mz == 1 ? 2 : 3;
  b->d2 = mx + 2 * ngg;
  b->d1 = (long)(my + 2 * ngg) * b->d2;
  b->sz = b->d1 * (mz + 2 * ngg);
  b->n = (long)mx * my * mz;
  b->n_face = n_face;
  for (int f = 0; f < n_face; ++f) {
    const int *r = faces + 7 * f;
    mface_t *F = &b->faces[f];
    F->s[0] = r[0]; F->e[0] = r[1];
    F->s[1] = r[2]; F->e[1] = r[3];
    F->s[2] = r[4]; F->e[2] = r[5];
    F->type = r[6];
    F->face = 0;
    F->dir = 1;
    for (int d = 0; d < b->dim; ++d)
      if (F->s[d] == F->e[d]) {
        F->face = d;
        if (F->s[d] == 0) F->dir = -1;
        break;
      }
    for (int d = 0; d < b->dim; ++d) {
      if (d == F->face) continue;
      if (F->e[d] < F->s[d]) { const int t = F->e[d]; F->e[d] = F->s[d]; F->s[d] = t; }
    }
    if (b->dim == 2) { F->s[2] = 0; F->e[2] = 0; }
    if (F->type != 8) b->bt[2 * F->face + (F->dir > 0 ? 1 : 0)] = F->type;
  }
  b->cv = calloc((size_t)b->nv * b->sz, sizeof(double));
  b->qn = calloc((size_t)b->nv * b->sz, sizeof(double));
  b->bv = calloc((size_t)6 * b->sz, sizeof(double));
  b->sv = calloc((size_t)b->ns * b->sz, sizeof(double));
  b->c = calloc((size_t)b->sz, sizeof(double));
  b->efd = calloc((size_t)b->sz, sizeof(double));
  b->sensor = calloc((size_t)b->sz, sizeof(double));
  b->dq = calloc((size_t)b->nv * b->n, sizeof(double));
  b->hstep = calloc((size_t)b->n, sizeof(double));
  b->cn = calloc((size_t)b->nv * b->n, sizeof(double));
  b->wu_ts = calloc((size_t)b->n, sizeof(double));
  b->tscale = malloc((size_t)b->n * sizeof(double));
  for (long q = 0; q < b->n; ++q) b->tscale[q] = 1e+6;
  b->jac = malloc((size_t)b->sz * sizeof(double));
  b->met = malloc((size_t)9 * b->sz * sizeof(double));
  memcpy(b->jac, jac, (size_t)b->sz * sizeof(double));
  memcpy(b->met, metric, (size_t)9 * b->sz * sizeof(double));
  return b;
}

void orc_mix_destroy(orc_mix *b) {
  if (!b) return;
  for (int f = 0; f < 64; ++f) {
    free(b->inflow[f].prof);
    free(b->inflow[f].rms);
  }
  free(b->sponge_sigma);
  free(b->sponge_mean);
  free(b->cv); free(b->qn); free(b->bv); free(b->sv); free(b->c); free(b->efd); free(b->dq); free(b->hstep);
  free(b->sensor);
  free(b->tscale); free(b->cn); free(b->wu_ts);
  free(b->mul); free(b->cond); free(b->rhod);
  for (int d = 0; d < 3; ++d) free(b->fvis[d]);
  free(b->jac); free(b->met);
  free(b);
}

/* compute_cv_from_bv_1_point<Mixture> + compute_total_energy (FieldOperation.cuh:16-84) */
static void cv_from_bv(orc_mix *b, long id) {
  const long s = b->sz;
  const double rho = b->bv[id], u = b->bv[id + s], v = b->bv[id + 2 * s], w = b->bv[id + 3 * s];
  b->cv[id] = rho;
  b->cv[id + s] = rho * u;
  b->cv[id + 2 * s] = rho * v;
  b->cv[id + 3 * s] = rho * w;
  for (int l = 0; l < b->ns; ++l) b->cv[id + (5 + l) * s] = rho * b->sv[id + l * s];
  double h[MS], cp[MS];
  orc_chem_h_cp(&b->m, b->bv[id + 5 * s], h, cp);
  double E = 0.5 * (u * u + v * v + w * w);
  for (int l = 0; l < b->ns; ++l) E += h[l] * b->sv[id + l * s];
  E *= rho;
  E -= b->bv[id + 4 * s];
  b->cv[id + 4 * s] = E;
}

/* update_physical_properties<Mixture> (FieldOperation.cuh:86-120): the acoustic speed */
static void properties(orc_mix *b, long id) {
  const long s = b->sz;
  double h[MS], cp[MS];
  const double T = b->bv[id + 5 * s];
  orc_chem_h_cp(&b->m, T, h, cp);
  double imw = 0, cpt = 0;
  for (int l = 0; l < b->ns; ++l) {
    const double y = b->sv[id + l * s];
    imw += y / b->m.mw[l];
    cpt += y * cp[l];
  }
  const double R = kRu * imw;
  const double gam = cpt / (cpt - R);
  b->c[id] = sqrt(gam * R * T);
}

static void all_properties(orc_mix *b) {
#pragma omp parallel for
  for (long id = 0; id < b->sz; ++id) properties(b, id);
}

static void copy_node(orc_mix *b, long dst, long src) {
  for (int l = 0; l < 6; ++l) b->bv[dst + l * b->sz] = b->bv[src + l * b->sz];
  for (int l = 0; l < b->ns; ++l) b->sv[dst + l * b->sz] = b->sv[src + l * b->sz];
}

/* apply_outflow (BoundCond.cu:2065-2188) with the species */
static void apply_outflow(orc_mix *b, const mface_t *F) {
  int d[3] = {0, 0, 0};
  d[F->face] = F->dir;
  for (int k = F->s[2]; k <= F->e[2]; ++k)
    for (int j = F->s[1]; j <= F->e[1]; ++j)
      for (int i = F->s[0]; i <= F->e[0]; ++i) {
        const long id = ID(b, i, j, k);
        for (int g = 1; g <= b->g; ++g) {
          const long ig = ID(b, i + g * d[0], j + g * d[1], k + g * d[2]);
          copy_node(b, ig, id);
          cv_from_bv(b, ig);
        }
      }
}

/* apply_periodic (BoundCond.cu:3033-3083) with the species */
static void apply_periodic(orc_mix *b, const mface_t *F) {
  int d[3] = {0, 0, 0};
  d[F->face] = F->dir;
  const int ext[3] = {b->mx, b->my, b->mz};
  for (int k = F->s[2]; k <= F->e[2]; ++k)
    for (int j = F->s[1]; j <= F->e[1]; ++j)
      for (int i = F->s[0]; i <= F->e[0]; ++i) {
        int o[3] = {i, j, k};
        o[F->face] = F->dir < 0 ? ext[F->face] - 1 : 0;
        const long id = ID(b, i, j, k), io = ID(b, o[0], o[1], o[2]);
        for (int l = 0; l < 6; ++l) {
          const double a = 0.5 * (b->bv[id + l * b->sz] + b->bv[io + l * b->sz]);
          b->bv[id + l * b->sz] = a;
          b->bv[io + l * b->sz] = a;
        }
        for (int l = 0; l < b->ns; ++l) {
          const double a = 0.5 * (b->sv[id + l * b->sz] + b->sv[io + l * b->sz]);
          b->sv[id + l * b->sz] = a;
          b->sv[io + l * b->sz] = a;
        }
        cv_from_bv(b, id);
        for (int g = 1; g <= b->g; ++g) {
          const long ig = ID(b, i + g * d[0], j + g * d[1], k + g * d[2]);
          const long is = ID(b, o[0] + g * d[0], o[1] + g * d[1], o[2] + g * d[2]);
          copy_node(b, ig, is);
          cv_from_bv(b, ig);
        }
      }
}

/* cv and acoustic speed of the face nodes (from 0 when the face nodes are included) and of their ghost layers */
static void face_cv(orc_mix *b, const mface_t *F, int from) {
  int d[3] = {0, 0, 0};
  d[F->face] = F->dir;
  for (int k = F->s[2]; k <= F->e[2]; ++k)
    for (int j = F->s[1]; j <= F->e[1]; ++j)
      for (int i = F->s[0]; i <= F->e[0]; ++i)
        for (int g = from; g <= b->g; ++g) {
          const long id = ID(b, i + g * d[0], j + g * d[1], k + g * d[2]);
          cv_from_bv(b, id);
          properties(b, id);
        }
}

/* DBoundCond::apply_boundary_conditions order (BoundCond.cu:3489-3619): outflow, inflow, symmetry, periodic */
static void apply_bcs(orc_mix *b) {
  for (int f = 0; f < b->n_face; ++f)
    if (b->faces[f].type == 6) apply_outflow(b, &b->faces[f]);
  for (int f = 0; f < b->n_face; ++f) {
    const mface_t *F = &b->faces[f];
    if (F->type != 5 || !b->inflow[f].prof) continue;
    /* the subsonic value reads the acoustic speed of the current state */
    if (b->inflow[f].sub_node) all_properties(b);
    orc_bc_inflow(b->mx, b->my, b->mz, b->g, F->s, F->e, F->face, F->dir, b->bv, b->sv, b->ns, b->c, b->met,
                  b->inflow[f].prof, b->inflow[f].rms, b->inflow[f].sub_node, b->inflow[f].sub_ghost,
                  b->inflow[f].fluct, b->inflow[f].mw, b->inflow[f].seed, b->inflow[f].call);
    face_cv(b, F, 0);
    if (b->inflow[f].fluct) ++b->inflow[f].call;
  }
  for (int f = 0; f < b->n_face; ++f) {
    const mface_t *F = &b->faces[f];
    if (F->type != 3) continue;
    orc_bc_symmetry(b->mx, b->my, b->mz, b->g, F->s, F->e, F->face, F->dir, b->bv, b->sv, b->ns, b->met);
    face_cv(b, F, 1);
  }
  for (int f = 0; f < b->n_face; ++f)
    if (b->faces[f].type == 8) apply_periodic(b, &b->faces[f]);
}

void orc_mix_set_inflow(orc_mix *b, int face, const double *prof, const double *rms, int sub_node, int sub_ghost,
                        int fluct, double mw, unsigned long long seed, unsigned long long first_call) {
  const mface_t *F = &b->faces[face];
  const long npt = (long)(F->e[0] - F->s[0] + 1) * (F->e[1] - F->s[1] + 1) * (F->e[2] - F->s[2] + 1);
  free(b->inflow[face].prof);
  free(b->inflow[face].rms);
  b->inflow[face].prof = malloc(sizeof(double) * (6 + b->ns) * (size_t)npt);
  b->inflow[face].rms = malloc(sizeof(double) * (size_t)npt);
  memcpy(b->inflow[face].prof, prof, sizeof(double) * (6 + b->ns) * (size_t)npt);
  memcpy(b->inflow[face].rms, rms, sizeof(double) * (size_t)npt);
  b->inflow[face].sub_node = sub_node;
  b->inflow[face].sub_ghost = sub_ghost;
  b->inflow[face].fluct = fluct;
  b->inflow[face].mw = mw;
  b->inflow[face].seed = seed;
  b->inflow[face].call = first_call;
}

void orc_mix_set_sponge(orc_mix *b, const double *sigma) {
  free(b->sponge_sigma);
  free(b->sponge_mean);
  b->sponge_sigma = malloc(sizeof(double) * (size_t)b->n);
  memcpy(b->sponge_sigma, sigma, sizeof(double) * (size_t)b->n);
  b->sponge_mean = calloc((size_t)b->nv * b->n, sizeof(double));
  b->sponge_iter = 0;
}

/* The sponge layer's relaxation -J sigma (U - U_mean) added to the residual, and its running mean. */
static void sponge_source(orc_mix *b) {
  if (!b->sponge_sigma) return;
  for (int k = 0; k < b->mz; ++k)
    for (int j = 0; j < b->my; ++j)
      for (int i = 0; i < b->mx; ++i) {
        const long id = ID(b, i, j, k), iq = IDQ(b, i, j, k);
        const double sg = b->sponge_sigma[iq] * b->jac[id];
        for (int l = 0; l < b->nv; ++l)
          b->dq[iq + l * b->n] -= sg * (b->cv[id + l * b->sz] - b->sponge_mean[iq + l * b->n]);
      }
}

static void sponge_mean(orc_mix *b) {
  if (!b->sponge_sigma) return;
  const double n = (double)b->sponge_iter, inv = 1.0 / (n + 1);
  for (int k = 0; k < b->mz; ++k)
    for (int j = 0; j < b->my; ++j)
      for (int i = 0; i < b->mx; ++i) {
        const long id = ID(b, i, j, k), iq = IDQ(b, i, j, k);
        for (int l = 0; l < b->nv; ++l)
          b->sponge_mean[iq + l * b->n] = (b->sponge_mean[iq + l * b->n] * n + b->cv[id + l * b->sz]) * inv;
      }
  ++b->sponge_iter;
}

void orc_mix_set_state(orc_mix *b, const double *bv, const double *sv) {
  memcpy(b->bv, bv, sizeof(double) * 6 * b->sz);
  memcpy(b->sv, sv, sizeof(double) * b->ns * b->sz);
#pragma omp parallel for
  for (long id = 0; id < b->sz; ++id) cv_from_bv(b, id);
  apply_bcs(b);
  all_properties(b);
}

void orc_mix_set_raw(orc_mix *b, const double *bv, const double *sv, const double *cv) {
  memcpy(b->bv, bv, sizeof(double) * 6 * b->sz);
  memcpy(b->sv, sv, sizeof(double) * b->ns * b->sz);
  memcpy(b->cv, cv, sizeof(double) * b->nv * b->sz);
  all_properties(b);
}

void orc_mix_get(const orc_mix *b, int which, double *out) {
  switch (which) {
    case 0: memcpy(out, b->cv, sizeof(double) * b->nv * b->sz); break;
    case 1: memcpy(out, b->bv, sizeof(double) * 6 * b->sz); break;
    case 2: memcpy(out, b->dq, sizeof(double) * b->nv * b->n); break;
    case 3: memcpy(out, b->sv, sizeof(double) * b->ns * b->sz); break;
    case 4: memcpy(out, b->c, sizeof(double) * b->sz); break;
    case 9: memcpy(out, b->sensor, sizeof(double) * b->sz); break;
    default: break;
  }
}

/* compute_entropy_fix_delta<Mixture> (InviscidScheme.cu:269-304) */
static void entropy_fix(orc_mix *b) {
  const long s = b->sz;
  const int kl = b->dim == 3 ? -1 : 0, kh = b->dim == 3 ? b->mz + 1 : 1;
#pragma omp parallel for collapse(2)
  for (int k = kl; k < kh; ++k)
    for (int j = -1; j < b->my + 1; ++j)
      for (int i = -1; i < b->mx + 1; ++i) {
        const long id = ID(b, i, j, k);
        const double *M = b->met, *bv = b->bv;
        const double u = bv[id + s], v = bv[id + 2 * s], w = bv[id + 3 * s];
        const double U = fabs(u * M[id] + v * M[id + s] + w * M[id + 2 * s]);
        const double V = fabs(u * M[id + 3 * s] + v * M[id + 4 * s] + w * M[id + 5 * s]);
        const double W = fabs(u * M[id + 6 * s] + v * M[id + 7 * s] + w * M[id + 8 * s]);
        const double kx = sqrt(M[id] * M[id] + M[id + s] * M[id + s] + M[id + 2 * s] * M[id + 2 * s]);
        const double ky = sqrt(M[id + 3 * s] * M[id + 3 * s] + M[id + 4 * s] * M[id + 4 * s] +
                               M[id + 5 * s] * M[id + 5 * s]);
        const double kz = sqrt(M[id + 6 * s] * M[id + 6 * s] + M[id + 7 * s] * M[id + 7 * s] +
                               M[id + 8 * s] * M[id + 8 * s]);
        const double c = b->c[id];
        if (b->dim == 2)
          b->efd[id] = b->p.entropy_fix_factor * (U + V + c * 0.5 * (kx + ky));
        else
          b->efd[id] = b->p.entropy_fix_factor * (U + V + W + c * (kx + ky + kz) / 3.0);
      }
}

static double minmod(double a, double c) {
  const double sa = (a > 0) - (a < 0), sc = (c > 0) - (c < 0);
  return 0.5 * (sa + sc) * fmin(fabs(a), fabs(c));
}

/* Left/right primitive states at the interface between stencil points r-1 and r of pv[2r][nv]. */
static void reconstruct(const orc_mix *b, double (*pv)[NV], int r, double *pl, double *pr) {
  const int nv = b->nv, rec = b->p.reconstruction;
  const int c = r - 1; /* left node */
  for (int l = 0; l < nv; ++l) {
    if (rec == 1) {
      pl[l] = pv[c][l];
      pr[l] = pv[c + 1][l];
    } else if (rec == 2 || rec == 3) {
      const double d0 = pv[c][l] - pv[c - 1][l], d1 = pv[c + 1][l] - pv[c][l], d2 = pv[c + 2][l] - pv[c + 1][l];
      if (rec == 2) {
        const double kappa = 1.0 / 3.0;
        pl[l] = pv[c][l] + 0.25 * ((1 - kappa) * minmod(d0, d1) + (1 + kappa) * minmod(d1, d0));
        pr[l] = pv[c + 1][l] - 0.25 * ((1 - kappa) * minmod(d2, d1) + (1 + kappa) * minmod(d1, d2));
      } else {
        pl[l] = pv[c][l] + 0.5 * minmod(d0, d1);
        pr[l] = pv[c + 1][l] - 0.5 * minmod(d1, d2);
      }
    } else {
      double s[8], zero[8] = {0};
      for (int q = 0; q < 2 * r; ++q) s[q] = pv[q][l];
      if (rec == 4) {
        pl[l] = orc_weno5(s, zero, b->p.weno_eps, 1);
        pr[l] = orc_weno5(zero, s + 1, b->p.weno_eps, 1);
      } else {
        pl[l] = orc_weno7(s, zero, b->p.weno_eps, 1);
        pr[l] = orc_weno7(zero, s + 1, b->p.weno_eps, 1);
      }
    }
  }
  if (rec >= 4 && (pl[0] <= 0 || pl[4] <= 0 || pr[0] <= 0 || pr[4] <= 0)) {
    for (int l = 0; l < nv; ++l) {
      pl[l] = pv[c][l];
      pr[l] = pv[c + 1][l];
    }
  }
}

/* reconstruction<Mixture> post-processing (InviscidScheme.cu:184-210): E at [nv], gamma at [nv+1] */
static void post(const orc_mix *b, double *pv) {
  const int ns = b->ns, nv = b->nv;
  double e = 0.5 * (pv[1] * pv[1] + pv[2] * pv[2] + pv[3] * pv[3]);
  double imw = 0;
  for (int l = 0; l < ns; ++l) imw += pv[5 + l] / b->m.mw[l];
  const double t = pv[4] / (pv[0] * kRu * imw);
  double h[MS], cp[MS];
  orc_chem_h_cp(&b->m, t, h, cp);
  double cpt = 0, cvt = 0;
  for (int l = 0; l < ns; ++l) {
    cpt += cp[l] * pv[5 + l];
    cvt += pv[5 + l] * (cp[l] - kRu / b->m.mw[l]);
    e += h[l] * pv[5 + l];
  }
  pv[nv] = pv[0] * e - pv[4];
  pv[nv + 1] = cpt / cvt;
}

/* compute_half_sum_left_right_flux (InviscidScheme.cu:601-633) + riemannSolver_Roe<Mixture> (635-765) */
static void roe(const orc_mix *b, const double *pl, const double *pr, const double *ml, const double *mr, double jl,
                double jr, double efl, double efr, double *fc) {
  const int ns = b->ns, nv = b->nv;
  double JKx = jl * ml[0], JKy = jl * ml[1], JKz = jl * ml[2];
  double Uk = pl[1] * JKx + pl[2] * JKy + pl[3] * JKz;
  double co = Uk * pl[0];
  fc[0] = 0.5 * co;
  fc[1] = 0.5 * (co * pl[1] + pl[4] * JKx);
  fc[2] = 0.5 * (co * pl[2] + pl[4] * JKy);
  fc[3] = 0.5 * (co * pl[3] + pl[4] * JKz);
  fc[4] = 0.5 * Uk * (pl[4] + pl[nv]);
  for (int l = 5; l < nv; ++l) fc[l] = 0.5 * co * pl[l];
  JKx = jr * mr[0]; JKy = jr * mr[1]; JKz = jr * mr[2];
  Uk = pr[1] * JKx + pr[2] * JKy + pr[3] * JKz;
  co = Uk * pr[0];
  fc[0] += 0.5 * co;
  fc[1] += 0.5 * (co * pr[1] + pr[4] * JKx);
  fc[2] += 0.5 * (co * pr[2] + pr[4] * JKy);
  fc[3] += 0.5 * (co * pr[3] + pr[4] * JKz);
  fc[4] += 0.5 * Uk * (pr[4] + pr[nv]);
  for (int l = 5; l < nv; ++l) fc[l] += 0.5 * co * pr[l];

  const double dl = sqrt(pl[0]), dr = sqrt(pr[0]);
  const double inv = 1.0 / (dl + dr);
  const double u = (dl * pl[1] + dr * pr[1]) * inv;
  const double v = (dl * pl[2] + dr * pr[2]) * inv;
  const double w = (dl * pl[3] + dr * pr[3]) * inv;
  const double ek = 0.5 * (u * u + v * v + w * w);
  const double hl = (pl[nv] + pl[4]) / pl[0], hr = (pr[nv] + pr[4]) / pr[0];
  const double h = (dl * hl + dr * hr) * inv;
  double svm[MS];
  for (int l = 0; l < ns; ++l) svm[l] = (dl * pl[l + 5] + dr * pr[l + 5]) * inv;
  double mw_inv = 0;
  for (int l = 0; l < ns; ++l) mw_inv += svm[l] / b->m.mw[l];
  const double tl = pl[4] / pl[0], tr = pr[4] / pr[0];
  const double t = (dl * tl + dr * tr) * inv / (kRu * mw_inv);
  double hi[MS], cpi[MS];
  orc_chem_h_cp(&b->m, t, hi, cpi);
  double cvt = 0, cpt = 0;
  for (int l = 0; l < ns; ++l) {
    cpt += cpi[l] * svm[l];
    cvt += svm[l] * (cpi[l] - kRu / b->m.mw[l]);
  }
  const double gamma = cpt / cvt;
  const double c = sqrt(gamma * kRu * mw_inv * t);
  const double mw = 1.0 / mw_inv;

  double kx = 0.5 * (ml[0] + mr[0]), ky = 0.5 * (ml[1] + mr[1]), kz = 0.5 * (ml[2] + mr[2]);
  const double gradK = sqrt(kx * kx + ky * ky + kz * kz);
  double U = kx * u + ky * v + kz * w;
  double ch[3] = {U - gradK * c, U, U + gradK * c};
  const double efa = 0.5 * (efl + efr);
  for (int q = 0; q < 3; ++q) {
    ch[q] = fabs(ch[q]);
    if (ch[q] < efa) ch[q] = 0.5 * (ch[q] * ch[q] / efa + efa);
  }
  kx /= gradK;
  ky /= gradK;
  kz /= gradK;
  U /= gradK;
  const double jave = 0.5 * (jl + jr);
  double dq[NV];
  dq[0] = jave * (pr[0] - pl[0]);
  for (int l = 1; l < nv; ++l) dq[l] = jave * (pr[0] * pr[l] - pl[0] * pl[l]);
  dq[4] = jave * (pr[nv] - pl[nv]);
  double c1 = (gamma - 1) * (ek * dq[0] - u * dq[1] - v * dq[2] - w * dq[3] + dq[4]) / (c * c);
  double c2 = (kx * dq[1] + ky * dq[2] + kz * dq[3] - U * dq[0]) / c;
  for (int l = 0; l < ns; ++l) c1 += (mw / b->m.mw[l] - hi[l] * (gamma - 1) / (c * c)) * dq[l + 5];
  double c3 = dq[0] - c1;
  double LDq[NV];
  LDq[0] = 0.5 * (c1 - c2);
  LDq[1] = kx * c3 - ((kz * v - ky * w) * dq[0] - kz * dq[2] + ky * dq[3]) / c;
  LDq[2] = ky * c3 - ((kx * w - kz * u) * dq[0] - kx * dq[3] + kz * dq[1]) / c;
  LDq[3] = kz * c3 - ((ky * u - kx * v) * dq[0] - ky * dq[1] + kx * dq[2]) / c;
  LDq[4] = 0.5 * (c1 + c2);
  for (int l = 0; l < ns; ++l) LDq[l + 5] = dq[l + 5] - svm[l] * dq[0];
  double bb[NV];
  bb[0] = -ch[0] * LDq[0];
  for (int l = 1; l < nv; ++l) bb[l] = -ch[1] * LDq[l];
  bb[4] = -ch[2] * LDq[4];
  const double c0 = kx * bb[1] + ky * bb[2] + kz * bb[3];
  c1 = c0 + bb[0] + bb[4];
  c2 = c * (bb[4] - bb[0]);
  c3 = 0;
  for (int l = 0; l < ns; ++l) c3 += (hi[l] - mw / b->m.mw[l] * c * c / (gamma - 1)) * bb[l + 5];
  fc[0] += 0.5 * c1;
  fc[1] += 0.5 * (u * c1 + kx * c2 - c * (kz * bb[2] - ky * bb[3]));
  fc[2] += 0.5 * (v * c1 + ky * c2 - c * (kx * bb[3] - kz * bb[1]));
  fc[3] += 0.5 * (w * c1 + kz * c2 - c * (ky * bb[1] - kx * bb[2]));
  fc[4] += 0.5 * (h * c1 + U * c2 - c * c * c0 / (gamma - 1) +
                  c * ((kz * v - ky * w) * bb[1] + (kx * w - kz * u) * bb[2] + (ky * u - kx * v) * bb[3]) + c3);
  for (int l = 0; l < ns; ++l) fc[5 + l] += 0.5 * (bb[l + 5] + svm[l] * c1);
}

/* riemannSolver_ausmPlus<Mixture> (InviscidScheme.cu:397-469) */
static void ausm(const orc_mix *b, const double *pl, const double *pr, const double *ml, const double *mr, double jl,
                 double jr, double *fc) {
  const int nv = b->nv;
  const double k1 = 0.5 * (jl * ml[0] + jr * mr[0]), k2 = 0.5 * (jl * ml[1] + jr * mr[1]),
               k3 = 0.5 * (jl * ml[2] + jr * mr[2]);
  const double gk = sqrt(k1 * k1 + k2 * k2 + k3 * k3);
  const double ul = (k1 * pl[1] + k2 * pl[2] + k3 * pl[3]) / gk, ur = (k1 * pr[1] + k2 * pr[2] + k3 * pr[3]) / gk;
  const double c = 0.5 * (sqrt(pl[nv + 1] * pl[4] / pl[0]) + sqrt(pr[nv + 1] * pr[4] / pr[0]));
  const double Ml = ul / c, Mr = ur / c, alpha = 3.0 / 16.0;
  double mlp, mrn, plp, prn;
  if (fabs(Ml) > 1) {
    mlp = 0.5 * (Ml + fabs(Ml));
    plp = mlp / Ml;
  } else {
    const double a = (Ml + 1) * (Ml + 1) * 0.25, q = (Ml * Ml - 1) * (Ml * Ml - 1);
    mlp = a + 0.125 * q;
    plp = a * (2 - Ml) + alpha * Ml * q;
  }
  if (fabs(Mr) > 1) {
    mrn = 0.5 * (Mr - fabs(Mr));
    prn = mrn / Mr;
  } else {
    const double a = (Mr - 1) * (Mr - 1) * 0.25, q = (Mr * Mr - 1) * (Mr * Mr - 1);
    mrn = -a - 0.125 * q;
    prn = a * (2 + Mr) - alpha * Mr * q;
  }
  const double pc = plp * pl[4] + prn * pr[4];
  const double mh = mlp + mrn;
  const double mp = 0.5 * (mh + fabs(mh)), mn = 0.5 * (mh - fabs(mh));
  const double mf = c * (pl[0] * mp + pr[0] * mn);
  const double co = mf * gk;
  const double *s = mf >= 0 ? pl : pr;
  fc[0] = co;
  fc[1] = co * s[1] + pc * k1;
  fc[2] = co * s[2] + pc * k2;
  fc[3] = co * s[3] + pc * k3;
  fc[4] = co * (s[nv] + s[4]) / s[0];
  for (int l = 5; l < nv; ++l) fc[l] = co * s[l];
}

/* riemannSolver_hllc<Mixture> (InviscidScheme.cu:471-599) */
static void hllc(const orc_mix *b, const double *pl, const double *pr, const double *ml, const double *mr, double jl,
                 double jr, double *fc) {
  const int ns = b->ns, nv = b->nv;
  const double rl = sqrt(pl[0]) / (sqrt(pl[0]) + sqrt(pr[0])), rr = sqrt(pr[0]) / (sqrt(pl[0]) + sqrt(pr[0]));
  const double ut = pl[1] * rl + pr[1] * rr, vt = pl[2] * rl + pr[2] * rr, wt = pl[3] * rl + pr[3] * rr;
  double svm[MS], mw_inv = 0;
  for (int l = 0; l < ns; ++l) {
    svm[l] = rl * pl[l + 5] + rr * pr[l + 5];
    mw_inv += svm[l] / b->m.mw[l];
  }
  const double t = (rl * pl[4] / pl[0] + rr * pr[4] / pr[0]) / (kRu * mw_inv);
  double hi[MS], cpi[MS];
  orc_chem_h_cp(&b->m, t, hi, cpi);
  double cvt = 0, cpt = 0;
  for (int l = 0; l < ns; ++l) {
    cpt += cpi[l] * svm[l];
    cvt += svm[l] * (cpi[l] - kRu / b->m.mw[l]);
  }
  const double ct = sqrt(cpt / cvt * kRu * mw_inv * t);
  const double kx = 0.5 * (ml[0] + mr[0]), ky = 0.5 * (ml[1] + mr[1]), kz = 0.5 * (ml[2] + mr[2]);
  const double gk = sqrt(kx * kx + ky * ky + kz * kz);
  const double Ub = (kx * ut + ky * vt + kz * wt) / gk;
  const double ja = 0.5 * (jl + jr);
  const double Ul = kx * pl[1] + ky * pl[2] + kz * pl[3];
  const double cl = sqrt(pl[nv + 1] * pl[4] / pl[0]);
  const double sl = fmin(Ul / gk - cl, Ub - ct);
  if (sl >= 0) {
    const double rU = pl[0] * Ul;
    fc[0] = ja * rU;
    fc[1] = ja * (rU * pl[1] + pl[4] * kx);
    fc[2] = ja * (rU * pl[2] + pl[4] * ky);
    fc[3] = ja * (rU * pl[3] + pl[4] * kz);
    fc[4] = ja * ((pl[nv] + pl[4]) * Ul);
    for (int l = 5; l < nv; ++l) fc[l] = ja * rU * pl[l];
    return;
  }
  const double Ur = kx * pr[1] + ky * pr[2] + kz * pr[3];
  const double cr = sqrt(pr[nv + 1] * pr[4] / pr[0]);
  const double sr = fmax(Ur / gk + cr, Ub + ct);
  if (sr < 0) {
    const double rU = pr[0] * Ur;
    fc[0] = ja * rU;
    fc[1] = ja * (rU * pr[1] + pr[4] * kx);
    fc[2] = ja * (rU * pr[2] + pr[4] * ky);
    fc[3] = ja * (rU * pr[3] + pr[4] * kz);
    fc[4] = ja * ((pr[nv] + pr[4]) * Ur);
    for (int l = 5; l < nv; ++l) fc[l] = ja * rU * pr[l];
    return;
  }
  const double sm = ((pr[0] * Ur * (sr - Ur / gk) - pl[0] * Ul * (sl - Ul / gk)) / gk + pl[4] - pr[4]) /
                    (pr[0] * (sr - Ur / gk) - pl[0] * (sl - Ul / gk));
  const double pm = pl[0] * (sl - Ul / gk) * (sm - Ul / gk) + pl[4];
  const int left = sm >= 0;
  const double *q = left ? pl : pr;
  const double S = left ? sl : sr, Uq = left ? Ul : Ur;
  const double pC = 1.0 / (S - sm);
  const double QC = ja * pC * sm * (S * gk - Uq) * q[0];
  fc[0] = QC;
  const double dP = (S * pm - sm * q[4]) * pC * ja;
  fc[1] = QC * q[1] + dP * kx;
  fc[2] = QC * q[2] + dP * ky;
  fc[3] = QC * q[3] + dP * kz;
  fc[4] = QC * q[nv] / q[0] + pC * ja * (S * pm * sm * gk - sm * q[4] * Uq);
  for (int l = 5; l < nv; ++l) fc[l] = QC * q[l];
}

/* LLF flux-splitting WENO with species, compute_convective_term_weno_x<Mixture> (WENO.cu:890-990 component-wise
 * 51/71; WENO.cu:1146-1470 characteristic 52/72) for the interface between stencil nodes n/2-1 and n/2.  Per node
 * m: the conservative state q (nv), the sweep's metric k, the Jacobian, pressure and acoustic speed.  The
 * characteristic variant extends the air eigenvectors with the species: the Roe state carries Roe-averaged mass
 * fractions, the sound speed comes from the mixture gamma at the Roe-weighted temperature, the rows of the
 * five flow fields gain alpha_l rho Y_l terms (alpha_l = (gamma R_l T - (gamma - 1) h_l) / c^2), and each species
 * field is rho Y_l - Y_l,Roe rho. */
static void fsweno_mix(const orc_mix *b, double (*q)[NV], double (*k)[3], const double *jac, const double *pp,
                       const double *snd, int n, int shock, double *fc) {
  const int nv = b->nv, ns = b->ns;
  const double eps = b->p.weno_eps;
  double Uk[8], lam[8];
  double lmax = 0;
  for (int m = 0; m < n; ++m) {
    Uk[m] = (q[m][1] * k[m][0] + q[m][2] * k[m][1] + q[m][3] * k[m][2]) / q[m][0];
    lam[m] = fabs(Uk[m]) + sqrt(k[m][0] * k[m][0] + k[m][1] * k[m][1] + k[m][2] * k[m][2]) * snd[m];
    if (lam[m] > lmax) lmax = lam[m];
  }
  if (b->p.inviscid_scheme % 10 == 1) {
    const double e1 = eps * b->p.v_ref * b->p.v_ref;
    for (int l = 0; l < nv; ++l) {
      double vp[7], vm[7];
      for (int m = 0; m < n; ++m) {
        const double hj = 0.5 * jac[m];
        const double sp = hj * (Uk[m] + lam[m]), sm = hj * (Uk[m] - lam[m]), pk = pp[m] * hj;
        double fp, fm;
        if (l == 0 || l >= 5) {
          fp = sp * q[m][l];
          fm = sm * q[m][l];
        } else if (l < 4) {
          fp = fma(sp, q[m][l], pk * k[m][l - 1]);
          fm = fma(sm, q[m][l], pk * k[m][l - 1]);
        } else {
          fp = fma(sp, q[m][4], pk * Uk[m]);
          fm = fma(sm, q[m][4], pk * Uk[m]);
        }
        if (m < n - 1) vp[m] = fp;
        if (m > 0) vm[m - 1] = fm;
      }
      const double el = (l >= 1 && l <= 3) ? e1 : (l == 4 ? e1 * b->p.v_ref * b->p.v_ref : eps);
      fc[l] = n == 8 ? orc_weno7(vp, vm, el, shock) : orc_weno5(vp, vm, el, shock);
    }
    return;
  }
  const int c = n / 2 - 1;
  const double t1 = sqrt(q[c][0] * q[c + 1][0]);
  const double rlc = 1 / (q[c][0] + t1), rrc = 1 / (t1 + q[c + 1][0]);
  const double um = rlc * q[c][1] + rrc * q[c + 1][1];
  const double vm = rlc * q[c][2] + rrc * q[c + 1][2];
  const double wm = rlc * q[c][3] + rrc * q[c + 1][3];
  double svm[MS], Rl[MS];
  double R = 0;
  for (int l = 0; l < ns; ++l) {
    svm[l] = rlc * q[c][5 + l] + rrc * q[c + 1][5 + l];
    Rl[l] = kRu / b->m.mw[l];
    R += svm[l] * Rl[l];
  }
  const double T = (rlc * pp[c] + rrc * pp[c + 1]) / R;
  double h[MS], cpl[MS];
  orc_chem_h_cp(&b->m, T, h, cpl);
  double cp = 0;
  for (int l = 0; l < ns; ++l) cp += svm[l] * cpl[l];
  const double gam = cp / (cp - R);
  const double cm = sqrt(gam * R * T);
  const double gm1 = gam - 1;
  double kx = 0.5 * (k[c][0] + k[c + 1][0]), ky = 0.5 * (k[c][1] + k[c + 1][1]), kz = 0.5 * (k[c][2] + k[c + 1][2]);
  double r = 1 / sqrt(kx * kx + ky * ky + kz * kz);
  kx *= r;
  ky *= r;
  kz *= r;
  const double Ub = kx * um + ky * vm + kz * wm;
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
  const double ic2 = 1.0 / (cm * cm);
  double al[MS];
  for (int l = 0; l < ns; ++l) al[l] = (gam * Rl[l] * T - gm1 * h[l]) * ic2;
  const double a = 0.5 * gm1 * (um * um + vm * vm + wm * wm);
  double sbq[8], sbf[8];
  for (int m = 0; m < n; ++m) {
    double sp = 0, sm = 0;
    for (int l = 0; l < ns; ++l) {
      sp = fma(al[l], q[m][5 + l], sp);
      sm = fma(al[l], q[m][5 + l] * Uk[m], sm);
    }
    sbq[m] = sp;
    sbf[m] = sm;
  }
  const double Lm[5][5] = {
      {(a + Ub * cm) * ic2 * 0.5, -(gm1 * um + kx * cm) * ic2 * 0.5, -(gm1 * vm + ky * cm) * ic2 * 0.5,
       -(gm1 * wm + kz * cm) * ic2 * 0.5, gm1 * ic2 * 0.5},
      {-Un, nx, ny, nz, 0},
      {-Uq, qx, qy, qz, 0},
      {1 - a * ic2, gm1 * um * ic2, gm1 * vm * ic2, gm1 * wm * ic2, -gm1 * ic2},
      {(a - Ub * cm) * ic2 * 0.5, -(gm1 * um - kx * cm) * ic2 * 0.5, -(gm1 * vm - ky * cm) * ic2 * 0.5,
       -(gm1 * wm - kz * cm) * ic2 * 0.5, gm1 * ic2 * 0.5}};
  const double tb[5] = {0.5, 0, 0, -1, 0.5};
  double fch[NV];
  for (int l = 0; l < 5; ++l) {
    const double *L = Lm[l];
    double vp[7], vmn[7];
    for (int m = 0; m < n; ++m) {
      const double F1 = fma(q[m][1], Uk[m], pp[m] * k[m][0]), F2 = fma(q[m][2], Uk[m], pp[m] * k[m][1]);
      const double F3 = fma(q[m][3], Uk[m], pp[m] * k[m][2]);
      const double LF = L[0] * (q[m][0] * Uk[m]) + L[1] * F1 + L[2] * F2 + L[3] * F3 +
                        L[4] * ((q[m][4] + pp[m]) * Uk[m]) + tb[l] * sbf[m];
      const double LQ = L[0] * q[m][0] + L[1] * q[m][1] + L[2] * q[m][2] + L[3] * q[m][3] + L[4] * q[m][4] +
                        tb[l] * sbq[m];
      if (m < n - 1) vp[m] = 0.5 * jac[m] * (LF + lmax * LQ);
      if (m > 0) vmn[m - 1] = 0.5 * jac[m] * (LF - lmax * LQ);
    }
    fch[l] = n == 8 ? orc_weno7(vp, vmn, eps, shock) : orc_weno5(vp, vmn, eps, shock);
  }
  for (int l = 0; l < ns; ++l) {
    double vp[7], vmn[7];
    for (int m = 0; m < n; ++m) {
      const double LQ = q[m][5 + l] - svm[l] * q[m][0];
      const double LF = Uk[m] * LQ;
      if (m < n - 1) vp[m] = 0.5 * jac[m] * (LF + lmax * LQ);
      if (m > 0) vmn[m - 1] = 0.5 * jac[m] * (LF - lmax * LQ);
    }
    fch[5 + l] = n == 8 ? orc_weno7(vp, vmn, eps, shock) : orc_weno5(vp, vmn, eps, shock);
  }
  const double s1 = fch[0] + fch[3] + fch[4], s3 = fch[0] - fch[4];
  fc[0] = s1;
  fc[1] = um * s1 - kx * cm * s3 + nx * fch[1] + qx * fch[2];
  fc[2] = vm * s1 - ky * cm * s3 + ny * fch[1] + qy * fch[2];
  fc[3] = wm * s1 - kz * cm * s3 + nz * fch[1] + qz * fch[2];
  const double H = rlc * (q[c][4] + pp[c]) + rrc * (q[c + 1][4] + pp[c + 1]);
  fc[4] = H * s1 - Ub * cm * s3 + Un * fch[1] + Uq * fch[2] - cm * cm / gm1 * fch[3];
  double sa = 0;
  for (int l = 0; l < ns; ++l) {
    fc[5 + l] = svm[l] * s1 + fch[5 + l];
    sa += al[l] * fch[5 + l];
  }
  fc[4] -= sa * cm * cm / gm1;
}

/* Positivity limiter of the species fluxes (WENO.cu:991-1087 component-wise, 1471-1540 characteristic): where
 * alpha J rho Y_L / 2 - dt f or alpha J rho Y_R / 2 + dt f would be negative (alpha = 1/3 in 3D, 1/2 in 2D), the
 * species flux f is blended towards a first-order split flux f1 of the two interface nodes, by the largest
 * factor that keeps both non-negative; from the first offending species on.  f1 splits with |U_k| + c|k| in
 * the component-wise schemes and with U_k + c|k| in the characteristic ones, as the reference does. */
static void positivity_limit(const orc_mix *b, double (*q)[NV], double (*k)[3], const double *jac,
                             const double *snd, int n, double *fc) {
  const int ns = b->ns, c = n / 2 - 1;
  const double dt = b->pp_dt;
  const double alpha = b->dim == 3 ? 1.0 / 3.0 : 0.5;
  const double aL = 0.5 * alpha * jac[c], aR = 0.5 * alpha * jac[c + 1];
  int start = -1;
  for (int l = 0; l < ns && start < 0; ++l) {
    const double up = aL * q[c][5 + l] - dt * fc[5 + l];
    const double um = aR * q[c + 1][5 + l] + dt * fc[5 + l];
    if (up < 0 || um < 0) start = l;
  }
  if (start < 0) return;
  double U[2], s[2];
  for (int e = 0; e < 2; ++e) {
    const double *kk = k[c + e], *qq = q[c + e];
    U[e] = (kk[0] * qq[1] + kk[1] * qq[2] + kk[2] * qq[3]) / qq[0];
    const double cg = sqrt(kk[0] * kk[0] + kk[1] * kk[1] + kk[2] * kk[2]) * snd[c + e];
    s[e] = b->p.inviscid_scheme % 10 == 1 ? fabs(U[e]) + cg : cg;
  }
  for (int l = start; l < ns; ++l) {
    const double yl = q[c][5 + l], yr = q[c + 1][5 + l];
    const double fp = 0.5 * jac[c] * (U[0] + s[0]) * yl, fm = 0.5 * jac[c + 1] * (U[1] - s[1]) * yr;
    const double f1 = 0.5 * (fp + U[0] * jac[c] * yl) + 0.5 * (fm - U[1] * jac[c + 1] * yr);
    double tp = 1, tm = 1;
    const double up = aL * yl - dt * fc[5 + l];
    if (up < 0) {
      const double up1 = aL * yl - dt * f1;
      if (fabs(up - up1) > 1e-20 * b->p.rho_ref) tp = fmin(fmax((0 - up1) / (up - up1), 0.0), 1.0);
    }
    const double um = aR * yr + dt * fc[5 + l];
    if (um < 0) {
      const double um1 = aR * yr + dt * f1;
      if (fabs(um - um1) > 1e-20 * b->p.rho_ref) tm = fmin(fmax((0 - um1) / (um - um1), 0.0), 1.0);
    }
    fc[5 + l] = fmin(tp, tm) * (fc[5 + l] - f1) + f1;
  }
}

void orc_mix_positivity_point(const orc_params *p, int ns, int dim, double dt, const double *q, const double *k,
                              const double *jac, const double *snd, int n, double *fc) {
  orc_mix b;
  memset(&b, 0, sizeof b);
  b.p = *p;
  b.ns = ns;
  b.nv = 5 + ns;
  b.dim = dim;
  b.pp_dt = dt;
  double qq[8][NV], kk[8][3];
  for (int m = 0; m < n; ++m) {
    for (int l = 0; l < 5 + ns; ++l) qq[m][l] = q[m * (5 + ns) + l];
    for (int c = 0; c < 3; ++c) kk[m][c] = k[3 * m + c];
  }
  positivity_limit(&b, qq, kk, jac, snd, n, fc);
}

static void inviscid_dir(orc_mix *b, int dir) {
  const int fs = b->p.inviscid_scheme > 50;
  const int rec = b->p.reconstruction;
  const int r = fs ? (b->p.inviscid_scheme > 70 ? 4 : 3) : (rec == 5 ? 4 : (rec == 4 ? 3 : 2)); /* half width */
  const int n = 2 * r, nv = b->nv, ns = b->ns;
  const long s = b->sz;
  const long st = dir == 0 ? 1 : (dir == 1 ? b->d2 : b->d1);
  const int ext[3] = {b->mx, b->my, b->mz};
  const int m = ext[dir];
  const int oa = dir == 0 ? 1 : 0, ob = dir == 2 ? 1 : 2;
#pragma omp parallel for collapse(2)
  for (int B = 0; B < ext[ob]; ++B)
    for (int A = 0; A < ext[oa]; ++A) {
      double *Fl = malloc(sizeof(double) * nv * (m + 1));
      for (int t = -1; t < m; ++t) {
        int idx[3];
        idx[dir] = t;
        idx[oa] = A;
        idx[ob] = B;
        const long id = ID(b, idx[0], idx[1], idx[2]);
        if (fs) {
          double q[8][NV], kk[8][3], jj[8], pp[8], cc[8];
          for (int mm = 0; mm < n; ++mm) {
            const long im = id + (mm - r + 1) * st;
            for (int l = 0; l < nv; ++l) q[mm][l] = b->cv[im + l * s];
            for (int c = 0; c < 3; ++c) kk[mm][c] = b->met[im + (3 * dir + c) * s];
            jj[mm] = b->jac[im];
            pp[mm] = b->bv[im + 4 * s];
            cc[mm] = b->c[im];
          }
          int shock = 1; /* the hybrid scheme's switch (WENO.cu:838-849) */
          if (b->p.sensor_threshold > 1e-10) {
            shock = 0;
            for (int ii = -b->g + 1; ii <= b->g; ++ii)
              if (b->sensor[id + ii * st] > b->p.sensor_threshold) shock = 1;
          }
          fsweno_mix(b, q, kk, jj, pp, cc, n, shock, Fl + (long)nv * (t + 1));
          if (b->p.positive_preserving) positivity_limit(b, q, kk, jj, cc, n, Fl + (long)nv * (t + 1));
          continue;
        }
        double pv[8][NV];
        for (int q = 0; q < n; ++q) {
          const long iq = id + (q - r + 1) * st;
          for (int l = 0; l < 5; ++l) pv[q][l] = b->bv[iq + l * s];
          for (int l = 0; l < ns; ++l) pv[q][5 + l] = b->sv[iq + l * s];
        }
        double pl[NV + 2], pr[NV + 2];
        reconstruct(b, pv, r, pl, pr);
        post(b, pl);
        post(b, pr);
        double ml[3], mr[3];
        for (int c = 0; c < 3; ++c) {
          ml[c] = b->met[id + (3 * dir + c) * s];
          mr[c] = b->met[id + st + (3 * dir + c) * s];
        }
        double *fc = Fl + (long)nv * (t + 1);
        if (b->p.inviscid_scheme == 2)
          roe(b, pl, pr, ml, mr, b->jac[id], b->jac[id + st], b->efd[id], b->efd[id + st], fc);
        else if (b->p.inviscid_scheme == 4)
          hllc(b, pl, pr, ml, mr, b->jac[id], b->jac[id + st], fc);
        else
          ausm(b, pl, pr, ml, mr, b->jac[id], b->jac[id + st], fc);
      }
      for (int t = 0; t < m; ++t) {
        int idx[3];
        idx[dir] = t;
        idx[oa] = A;
        idx[ob] = B;
        const long iq = IDQ(b, idx[0], idx[1], idx[2]);
        for (int l = 0; l < nv; ++l) b->dq[iq + l * b->n] -= Fl[(long)nv * (t + 1) + l] - Fl[(long)nv * t + l];
      }
      free(Fl);
    }
}

/* finite_rate_chemistry with chemSrcMethod 0 (FiniteRateChem.cu:1388-1464): dq += J omega, and the
 * reaction time scale min_k rho Y_k / omega_d,k */
static void explicit_source(orc_mix *b) {
  const long s = b->sz;
  const int ns = b->ns;
#pragma omp parallel for collapse(2)
  for (int k = 0; k < b->mz; ++k)
    for (int j = 0; j < b->my; ++j)
      for (int i = 0; i < b->mx; ++i) {
        const long id = ID(b, i, j, k), iq = IDQ(b, i, j, k);
        const double rho = b->bv[id];
        double ry[MS], q1[ORC_MAX_REAC], q2[ORC_MAX_REAC], wd[MS], w[MS];
        for (int l = 0; l < ns; ++l) ry[l] = rho * b->sv[id + l * s];
        orc_chem_source(&b->m, b->bv[id + 5 * s], ry, q1, q2, wd, w);
        double ts = 1e+6;
        for (int l = 0; l < ns; ++l) {
          b->dq[iq + (5 + l) * b->n] += b->jac[id] * w[l];
          if (b->sv[id + l * s] > 1e-25 && fabs(wd[l]) > 1e-25) ts = fmin(ts, fabs(ry[l] / wd[l]));
        }
        b->tscale[iq] = ts;
      }
}

void orc_mix_set_explicit_source(orc_mix *b, int on) { b->explicit_source = on; }

static double omega_d(double t) { return 1.0548 * pow(t, -0.15504) + pow(t + 0.55909, -2.1705); }

/* Pair coefficients of Species::read_tran (ChemData.cpp:779-831). */
typedef struct {
  double vc[MS], wdw[MS][MS], sqw[MS][MS], bdc[MS][MS], kbe[MS][MS], zr[MS];
} tran_coef;

static void tran_coeffs(const orc_mech *m, const orc_tran *tr, tran_coef *c) {
  const int ns = m->ns;
  const double pi = 3.14159265358979323846, kb = 1.380649e-23;
  for (int k = 0; k < ns; ++k) {
    c->vc[k] = 2.6693e-6 * sqrt(m->mw[k]) / (tr->sigma[k] * tr->sigma[k]);
    const double ti = tr->eps_kb[k] / 298;
    c->zr[k] = tr->zrot[k] * (1 + 0.5 * pi * pi * pi * sqrt(ti) + (2 + 0.25 * pi * pi) * ti + sqrt(pi * pi * pi) * sqrt(ti * ti * ti));
    for (int j = 0; j < ns; ++j) {
      c->wdw[k][j] = pow(m->mw[j] / m->mw[k], 0.25);
      c->sqw[k][j] = 1.0 / sqrt(8 * (1 + m->mw[k] / m->mw[j]));
    }
  }
  for (int j = 0; j < ns; ++j) {
    c->bdc[j][j] = 0.0188324 / (sqrt(m->mw[j] * 0.5) * tr->sigma[j] * tr->sigma[j]);
    c->kbe[j][j] = 1.0 / tr->eps_kb[j];
    for (int k = j + 1; k < ns; ++k) {
      const double W = m->mw[j] * m->mw[k] / (m->mw[j] + m->mw[k]);
      double xi = 1;
      if ((tr->dipole[j] != 0) != (tr->dipole[k] != 0)) {
        const int p = tr->dipole[j] != 0 ? j : k, n = tr->dipole[j] != 0 ? k : j;
        const double conv = 1e-18 * sqrt(1e-7) * pow(1e+8, 1.5) / sqrt(kb);
        const double mp = conv * tr->dipole[p] / sqrt(tr->eps_kb[p] * pow(tr->sigma[p], 3));
        xi = 1.0 + 0.25 * tr->polar[n] / pow(tr->sigma[n], 3) * mp * mp * sqrt(tr->eps_kb[p] / tr->eps_kb[n]);
      }
      const double sg = 0.5 * (tr->sigma[j] + tr->sigma[k]) * pow(xi, -1.0 / 6.0);
      c->bdc[j][k] = c->bdc[k][j] = 0.0188324 / (sqrt(W) * sg * sg);
      c->kbe[j][k] = c->kbe[k][j] = 1.0 / (xi * xi * sqrt(tr->eps_kb[j] * tr->eps_kb[k]));
    }
  }
}

/* Pure-species viscosity vs, conductivity lam (Eucken with the rotational-relaxation correction for linear and
 * non-linear molecules) and self-diffusion rho D_aa rd at T (Transport.cu:67-105). */
static void species_transport(const orc_mech *m, const orc_tran *tr, const tran_coef *c, double T, const double *cp,
                              int a, double *vs, double *lam, double *rd_out) {
  const double pi = 3.14159265358979323846;
  const double st = sqrt(T), t32 = T * st;
  const double td = T / tr->eps_kb[a];
  *vs = c->vc[a] * st / (1.147 * pow(td, -0.145) + pow(td + 0.5, -2));
  const double R = kRu / m->mw[a];
  const double trd = T * c->kbe[a][a];
  const double rd = c->bdc[a][a] * t32 / (omega_d(trd) * R * T);
  double l = 15 * 0.25 * *vs * R;
  if (tr->geometry[a] == 1 || tr->geometry[a] == 2) {
    const double tri = 1 / trd, q = sqrt(tri * pi) * pi;
    const double ft = 1 + 0.5 * q + (2 + 0.25 * pi * pi) * tri + q * tri;
    if (tr->geometry[a] == 1) {
      const double ab = (2.5 * *vs - rd) / (pi * c->zr[a] / ft * *vs + 10.0 / 3 * *vs + 2.0 * rd);
      l += -5 * ab * *vs * R + rd * cp[a] + rd * (2 * ab - 2.5) * R;
    } else {
      const double ab = (2.5 * *vs - rd) / (pi * c->zr[a] / ft * *vs + 5 * *vs + 2.0 * rd);
      l += -7.5 * ab * *vs * R + rd * cp[a] + rd * (3 * ab - 2.5) * R;
    }
  }
  *lam = l;
  if (rd_out) *rd_out = rd;
}

void orc_transport_species(const orc_mech *m, const orc_tran *tr, double T, int k, double *out) {
  tran_coef c;
  tran_coeffs(m, tr, &c);
  double h[MS], cp[MS];
  orc_chem_h_cp(m, T, h, cp);
  species_transport(m, tr, &c, T, cp, k, &out[0], &out[1], &out[2]);
}

void orc_transport(const orc_mech *m, const orc_tran *tr, double T, const double *Y, double *mu_out,
                   double *lambda_out, double *rhoD) {
  const int ns = m->ns;
  tran_coef c;
  tran_coeffs(m, tr, &c);
  /* properties (Transport.cu:62-150) */
  double h[MS], cp[MS], X[MS], vs[MS], lam[MS];
  orc_chem_h_cp(m, T, h, cp);
  double imw = 0;
  for (int l = 0; l < ns; ++l) imw += Y[l] / m->mw[l];
  const double mwt = 1.0 / imw, st = sqrt(T), t32 = T * st;
  for (int l = 0; l < ns; ++l) {
    X[l] = Y[l] * mwt / m->mw[l];
    species_transport(m, tr, &c, T, cp, l, &vs[l], &lam[l], NULL);
  }
  double mu = 0, lambda = 0;
  for (int a = 0; a < ns; ++a) {
    double den = 0;
    for (int b = 0; b < ns; ++b) {
      double phi = 1.0;
      if (b != a) {
        const double num = 1.0 + sqrt(vs[a] / vs[b]) * c.wdw[a][b];
        phi = num * num * c.sqw[a][b];
      }
      den += phi * X[b];
    }
    mu += vs[a] * X[a] / den;
    lambda += lam[a] * X[a] / (1.065 * den - 0.065 * X[a]);
  }
  *mu_out = mu;
  *lambda_out = lambda;
  const double eps = 1e-12;
  double num[MS], dsum[MS], sum = 0;
  for (int l = 0; l < ns; ++l) {
    num[l] = (X[l] + eps) * m->mw[l];
    sum += num[l];
    dsum[l] = 0;
  }
  for (int l = 0; l < ns; ++l)
    for (int n = l + 1; n < ns; ++n) {
      const double inv = omega_d(T * c.kbe[l][n]) / (c.bdc[l][n] * t32);
      dsum[l] += (X[n] + eps) * inv;
      dsum[n] += (X[l] + eps) * inv;
    }
  for (int l = 0; l < ns; ++l) rhoD[l] = (sum - num[l]) / (T * kRu) / dsum[l];
}

void orc_mix_set_transport(orc_mix *b, const orc_tran *t) {
  b->tran = *t;
  b->viscous = b->p.viscous_order == 2 || b->p.viscous_order == 8;
  if (!b->viscous || b->mul) return;
  b->mul = calloc((size_t)b->sz, sizeof(double));
  b->cond = calloc((size_t)b->sz, sizeof(double));
  b->rhod = calloc((size_t)b->ns * b->sz, sizeof(double));
  for (int d = 0; d < 3; ++d) b->fvis[d] = calloc((size_t)b->nv * b->sz, sizeof(double));
}

/* compute_fv/gv/hv_2nd_order<Mixture> (ViscousScheme.cu:614-778): face t+1/2 along dir stored at node t */
static void mix_viscous_faces(orc_mix *b, int dir) {
  const long s = b->sz;
  const int ns = b->ns;
  const long st[3] = {1, b->d2, b->d1};
  const int ext[3] = {b->mx, b->my, b->mz};
  int lo[3] = {0, 0, 0};
  lo[dir] = -1;
  double *F = b->fvis[dir];
#pragma omp parallel for collapse(2)
  for (int k = lo[2]; k < ext[2]; ++k)
    for (int j = lo[1]; j < ext[1]; ++j)
      for (int i = lo[0]; i < ext[0]; ++i) {
        const long id = ID(b, i, j, k), ie = id + st[dir];
        const double *M = b->met;
        double m9[9];
        for (int c = 0; c < 9; ++c) m9[c] = 0.5 * (M[id + c * s] + M[ie + c * s]);
        double gq[4 + MS][3];
        for (int q = 0; q < 4 + ns; ++q) {
          const double *f = q < 3 ? b->bv + (1 + q) * s : (q == 3 ? b->bv + 5 * s : b->sv + (q - 4) * s);
          double g[3];
          for (int a = 0; a < 3; ++a) {
            if (a == dir) g[a] = f[ie] - f[id];
            else if (a == 2 && b->dim == 2) g[a] = 0;
            else g[a] = 0.25 * (f[id + st[a]] - f[id - st[a]] + f[ie + st[a]] - f[ie - st[a]]);
          }
          for (int c = 0; c < 3; ++c) gq[q][c] = g[0] * m9[c] + g[1] * m9[3 + c] + g[2] * m9[6 + c];
        }
        const double mu = 0.5 * (b->mul[id] + b->mul[ie]);
        const double txx = mu * (4 * gq[0][0] - 2 * gq[1][1] - 2 * gq[2][2]) / 3.0;
        const double tyy = mu * (4 * gq[1][1] - 2 * gq[0][0] - 2 * gq[2][2]) / 3.0;
        const double tzz = mu * (4 * gq[2][2] - 2 * gq[0][0] - 2 * gq[1][1]) / 3.0;
        const double txy = mu * (gq[0][1] + gq[1][0]), txz = mu * (gq[0][2] + gq[2][0]), tyz = mu * (gq[1][2] + gq[2][1]);
        double kj[3];
        for (int c = 0; c < 3; ++c)
          kj[c] = 0.5 * (M[id + (3 * dir + c) * s] * b->jac[id] + M[ie + (3 * dir + c) * s] * b->jac[ie]);
        const double f1 = kj[0] * txx + kj[1] * txy + kj[2] * txz;
        const double f2 = kj[0] * txy + kj[1] * tyy + kj[2] * tyz;
        const double f3 = kj[0] * txz + kj[1] * tyz + kj[2] * tzz;
        const double um = 0.5 * (b->bv[id + s] + b->bv[ie + s]);
        const double vm = 0.5 * (b->bv[id + 2 * s] + b->bv[ie + 2 * s]);
        const double wm = 0.5 * (b->bv[id + 3 * s] + b->bv[ie + 3 * s]);
        const double lam = 0.5 * (b->cond[id] + b->cond[ie]);
        double e = um * f1 + vm * f2 + wm * f3 + lam * (kj[0] * gq[3][0] + kj[1] * gq[3][1] + kj[2] * gq[3][2]);
        double yk[MS], dk[MS], gy[MS], sgw = 0, srdy = 0, cvt = 0, imw = 0;
        for (int l = 0; l < ns; ++l) {
          yk[l] = 0.5 * (b->sv[id + l * s] + b->sv[ie + l * s]);
          dk[l] = 0.5 * (b->rhod[id + l * s] + b->rhod[ie + l * s]);
          gy[l] = kj[0] * gq[4 + l][0] + kj[1] * gq[4 + l][1] + kj[2] * gq[4 + l][2];
          cvt += dk[l] * gy[l];
          sgw += gy[l] / b->m.mw[l];
          imw += yk[l] / b->m.mw[l];
          srdy += dk[l] * yk[l];
        }
        const double mwt = 1.0 / imw;
        cvt -= mwt * srdy * sgw;
        double hh[MS], cp[MS];
        orc_chem_h_cp(&b->m, 0.5 * (b->bv[id + 5 * s] + b->bv[ie + 5 * s]), hh, cp);
        for (int l = 0; l < ns; ++l) {
          const double jd = dk[l] * (gy[l] - mwt * yk[l] * sgw) - yk[l] * cvt;
          F[id + (5 + l) * s] = jd;
          e += hh[l] * jd;
        }
        F[id + s] = f1;
        F[id + 2 * s] = f2;
        F[id + 3 * s] = f3;
        F[id + 4 * s] = e;
      }
}

static void mix_transport(orc_mix *b) {
  const long s = b->sz;
#pragma omp parallel for
  for (long id = 0; id < s; ++id) {
    const double T = b->bv[id + 5 * s];
    if (!(T > 0)) continue;
    double Y[MS], rd[MS];
    for (int l = 0; l < b->ns; ++l) Y[l] = b->sv[id + l * s];
    orc_transport(&b->m, &b->tran, T, Y, &b->mul[id], &b->cond[id], rd);
    for (int l = 0; l < b->ns; ++l) b->rhod[id + l * s] = rd[l];
  }
}

/* compute_viscous_flux_collocated<Mixture, 8> + compute_viscous_flux_collocated_scalar<8> (ViscousScheme.cu:309-566,
 * gradPInDiffusionFlux off): node fluxes of every interior node into fvis[0..2] (components 1..nv-1). */
static void mix_viscous_nodes(orc_mix *b) {
  const long s = b->sz;
  const int ns = b->ns, order = b->p.viscous_order;
#pragma omp parallel for collapse(2)
  for (int k = 0; k < b->mz; ++k)
    for (int j = 0; j < b->my; ++j)
      for (int i = 0; i < b->mx; ++i) {
        const long id = ID(b, i, j, k);
        const double *M = b->met;
        const double xx = M[id], xy = M[id + s], xz = M[id + 2 * s];
        const double ex = M[id + 3 * s], ey = M[id + 4 * s], ez = M[id + 5 * s];
        const double zx = M[id + 6 * s], zy = M[id + 7 * s], zz = M[id + 8 * s];
        double gp[4 + MS][3]; /* physical gradients of u, v, w, T, Y_l */
        for (int q = 0; q < 4 + ns; ++q) {
          const double *f = q < 3 ? b->bv + (1 + q) * s : (q == 3 ? b->bv + 5 * s : b->sv + (q - 4) * s);
          const double a = orc_deriv(f, id, 1, i, b->mx, b->bt[0], b->bt[1], order);
          const double e = orc_deriv(f, id, b->d2, j, b->my, b->bt[2], b->bt[3], order);
          const double c = b->dim == 3 ? orc_deriv(f, id, b->d1, k, b->mz, b->bt[4], b->bt[5], order) : 0.0;
          gp[q][0] = a * xx + e * ex + c * zx;
          gp[q][1] = a * xy + e * ey + c * zy;
          gp[q][2] = a * xz + e * ez + c * zz;
        }
        const double mu = b->mul[id];
        const double txx = mu * (4 * gp[0][0] - 2 * gp[1][1] - 2 * gp[2][2]) / 3.0;
        const double tyy = mu * (4 * gp[1][1] - 2 * gp[0][0] - 2 * gp[2][2]) / 3.0;
        const double tzz = mu * (4 * gp[2][2] - 2 * gp[0][0] - 2 * gp[1][1]) / 3.0;
        const double txy = mu * (gp[0][1] + gp[1][0]), txz = mu * (gp[0][2] + gp[2][0]);
        const double tyz = mu * (gp[1][2] + gp[2][1]);
        double *fv = b->fvis[0], *gv = b->fvis[1], *hv = b->fvis[2];
        fv[id + s] = xx * txx + xy * txy + xz * txz;
        fv[id + 2 * s] = xx * txy + xy * tyy + xz * tyz;
        fv[id + 3 * s] = xx * txz + xy * tyz + xz * tzz;
        gv[id + s] = ex * txx + ey * txy + ez * txz;
        gv[id + 2 * s] = ex * txy + ey * tyy + ez * tyz;
        gv[id + 3 * s] = ex * txz + ey * tyz + ez * tzz;
        hv[id + s] = zx * txx + zy * txy + zz * txz;
        hv[id + 2 * s] = zx * txy + zy * tyy + zz * tyz;
        hv[id + 3 * s] = zx * txz + zy * tyz + zz * tzz;
        const double lam = b->cond[id];
        const double u = b->bv[id + s], v = b->bv[id + 2 * s], w = b->bv[id + 3 * s];
        const double Ex = u * txx + v * txy + w * txz + lam * gp[3][0];
        const double Ey = u * txy + v * tyy + w * tyz + lam * gp[3][1];
        const double Ez = u * txz + v * tyz + w * tzz + lam * gp[3][2];
        double e4[3] = {xx * Ex + xy * Ey + xz * Ez, ex * Ex + ey * Ey + ez * Ez, zx * Ex + zy * Ey + zz * Ez};
        /* species diffusion: Term 1 (mass-fraction gradients), Term 2 (molar-mass gradient), correction velocity */
        double yk[MS], dk[MS], drv[MS][3], cvt[3] = {0, 0, 0}, sgw[3] = {0, 0, 0}, imw = 0, srdy = 0;
        for (int l = 0; l < ns; ++l) {
          yk[l] = b->sv[id + l * s];
          dk[l] = b->rhod[id + l * s];
          const double *g = gp[4 + l];
          drv[l][0] = xx * g[0] + xy * g[1] + xz * g[2];
          drv[l][1] = ex * g[0] + ey * g[1] + ez * g[2];
          drv[l][2] = zx * g[0] + zy * g[1] + zz * g[2];
          for (int a = 0; a < 3; ++a) {
            cvt[a] += dk[l] * drv[l][a];
            sgw[a] += drv[l][a] / b->m.mw[l];
          }
          imw += yk[l] / b->m.mw[l];
          srdy += dk[l] * yk[l];
        }
        const double mwt = 1.0 / imw;
        for (int a = 0; a < 3; ++a) cvt[a] -= mwt * srdy * sgw[a];
        double hh[MS], cp[MS];
        orc_chem_h_cp(&b->m, b->bv[id + 5 * s], hh, cp);
        double *F[3] = {fv, gv, hv};
        for (int l = 0; l < ns; ++l)
          for (int a = 0; a < 3; ++a) {
            const double jd = dk[l] * (drv[l][a] - mwt * yk[l] * sgw[a]) - yk[l] * cvt[a];
            F[a][id + (5 + l) * s] = jd;
            e4[a] += hh[l] * jd;
          }
        fv[id + 4 * s] = e4[0];
        gv[id + 4 * s] = e4[1];
        hv[id + 4 * s] = e4[2];
      }
}

/* periodic_exchange task 1 for the node viscous fluxes (DataCommunication.cu:628-683), face after face: average
 * the coincident face nodes, copy the ghost layers from inside the opposite face. */
static void mix_exchange_viscous(orc_mix *b) {
  const int ext[3] = {b->mx, b->my, b->mz};
  const long s = b->sz;
  for (int f = 0; f < b->n_face; ++f) {
    const mface_t *F = &b->faces[f];
    if (F->type != 8) continue;
    int d[3] = {0, 0, 0};
    d[F->face] = F->dir;
    for (int k = F->s[2]; k <= F->e[2]; ++k)
      for (int j = F->s[1]; j <= F->e[1]; ++j)
        for (int i = F->s[0]; i <= F->e[0]; ++i) {
          int o[3] = {i, j, k};
          o[F->face] = F->dir < 0 ? ext[F->face] - 1 : 0;
          const long id = ID(b, i, j, k), io = ID(b, o[0], o[1], o[2]);
          for (int l = 1; l < b->nv; ++l)
            for (int a = 0; a < 3; ++a) {
              const double av = 0.5 * (b->fvis[a][id + l * s] + b->fvis[a][io + l * s]);
              b->fvis[a][id + l * s] = av;
              b->fvis[a][io + l * s] = av;
            }
          for (int g = 1; g <= b->g; ++g) {
            const long ig = ID(b, i + g * d[0], j + g * d[1], k + g * d[2]);
            const long is = ID(b, o[0] + g * d[0], o[1] + g * d[1], o[2] + g * d[2]);
            for (int l = 1; l < b->nv; ++l)
              for (int a = 0; a < 3; ++a) b->fvis[a][ig + l * s] = b->fvis[a][is + l * s];
          }
        }
  }
}

/* compute_viscous_flux_derivative<8> (ViscousScheme.cu:568-612): dq += J (dFv/dxi + dGv/deta + dHv/dzeta). */
static void mix_viscous_collocated(orc_mix *b) {
  const long s = b->sz;
  const int order = b->p.viscous_order;
  mix_viscous_nodes(b);
  mix_exchange_viscous(b);
#pragma omp parallel for collapse(2)
  for (int k = 0; k < b->mz; ++k)
    for (int j = 0; j < b->my; ++j)
      for (int i = 0; i < b->mx; ++i) {
        const long id = ID(b, i, j, k), iq = IDQ(b, i, j, k);
        const double J = b->jac[id];
        for (int l = 1; l < b->nv; ++l) {
          double d = orc_deriv(b->fvis[0] + l * s, id, 1, i, b->mx, b->bt[0], b->bt[1], order) +
                     orc_deriv(b->fvis[1] + l * s, id, b->d2, j, b->my, b->bt[2], b->bt[3], order);
          if (b->dim == 3) d += orc_deriv(b->fvis[2] + l * s, id, b->d1, k, b->mz, b->bt[4], b->bt[5], order);
          b->dq[iq + l * b->n] += d * J;
        }
      }
}

static void mix_viscous(orc_mix *b) {
  const long s = b->sz;
  mix_transport(b);
  if (b->p.viscous_order == 8) {
    mix_viscous_collocated(b);
    return;
  }
  const long st[3] = {1, b->d2, b->d1};
  for (int dir = 0; dir < b->dim; ++dir) {
    mix_viscous_faces(b, dir);
    const double *F = b->fvis[dir];
#pragma omp parallel for collapse(2)
    for (int k = 0; k < b->mz; ++k)
      for (int j = 0; j < b->my; ++j)
        for (int i = 0; i < b->mx; ++i) {
          const long id = ID(b, i, j, k), iq = IDQ(b, i, j, k);
          for (int l = 1; l < b->nv; ++l) b->dq[iq + l * b->n] += F[id + l * s] - F[id - st[dir] + l * s];
        }
  }
}

/* compute_shock_sensor at the start of a step (RK.cuh:112-117; compute_cn, StrangSplitting.cu:19-24) */
static void shock_sensor(orc_mix *b) {
  if (b->p.inviscid_scheme > 50 && b->p.sensor_threshold > 1e-10)
    orc_shock_sensor(b->p.shock_sensor, b->p.sensor_eps, b->mx, b->my, b->mz, b->g, b->bv, b->met, b->sensor);
}

static void rhs_core(orc_mix *b);

void orc_mix_rhs(orc_mix *b) {
  shock_sensor(b);
  rhs_core(b);
}

static void rhs_core(orc_mix *b) {
  memset(b->dq, 0, sizeof(double) * b->nv * b->n);
  if (b->viscous) mix_viscous(b);
  if (b->p.inviscid_scheme == 2) entropy_fix(b);
  inviscid_dir(b, 0);
  inviscid_dir(b, 1);
  if (b->dim == 3) inviscid_dir(b, 2);
  if (b->explicit_source) explicit_source(b);
  sponge_source(b);
}

static const double rk_a[3] = {1.0, 0.75, 1.0 / 3.0};
static const double rk_b[3] = {0.0, 0.25, 2.0 / 3.0};
static const double rk_c[3] = {1.0, 0.25, 2.0 / 3.0};

/* compute_temperature_and_pressure (FieldOperation.cu:4-45) at node id from the species in sv */
static void temperature_pressure(orc_mix *b, long id) {
  const long s = b->sz;
  double R = 0;
  for (int l = 0; l < b->ns; ++l) R += b->sv[id + l * s] * kRu / b->m.mw[l];
  const double u = b->bv[id + s], v = b->bv[id + 2 * s], w = b->bv[id + 3 * s];
  const double e = b->cv[id + 4 * s] / b->bv[id] - 0.5 * (u * u + v * v + w * w);
  double err = 1, t = b->bv[id + 5 * s];
  int it = 0;
  while (err > 1e-5 && it++ < 1000) {
    double h[MS], cp[MS];
    orc_chem_h_cp(&b->m, t, h, cp);
    double cpt = 0, hh = 0;
    for (int l = 0; l < b->ns; ++l) {
      cpt += cp[l] * b->sv[id + l * s];
      hh += h[l] * b->sv[id + l * s];
    }
    const double t1 = t - (hh - R * t - e) / (cpt - R);
    err = fabs(1 - t1 / t);
    t = t1;
  }
  b->bv[id + 5 * s] = t;
  b->bv[id + 4 * s] = b->bv[id] * t * R;
}

/* update_cv_and_bv_rk<Mixture> (RK.cuh:224-280) */
static void update(orc_mix *b, double dt, int rk) {
  const long s = b->sz;
  const int nv = b->nv, ns = b->ns;
#pragma omp parallel for collapse(2)
  for (int k = 0; k < b->mz; ++k)
    for (int j = 0; j < b->my; ++j)
      for (int i = 0; i < b->mx; ++i) {
        const long id = ID(b, i, j, k), iq = IDQ(b, i, j, k);
        double *cv = b->cv, *bv = b->bv;
        const double dj = dt / b->jac[id];
        for (int l = 0; l < nv; ++l)
          cv[id + l * s] = rk_a[rk] * b->qn[id + l * s] + rk_b[rk] * cv[id + l * s] + rk_c[rk] * dj * b->dq[iq + l * b->n];
        if (b->mz == 1) cv[id + 3 * s] = 0;
        bv[id] = cv[id];
        const double di = 1.0 / cv[id];
        bv[id + s] = cv[id + s] * di;
        bv[id + 2 * s] = cv[id + 2 * s] * di;
        bv[id + 3 * s] = cv[id + 3 * s] * di;
        double sum = 0;
        for (int l = 0; l < ns; ++l) {
          if (cv[id + (5 + l) * s] < 0) cv[id + (5 + l) * s] = 0;
          else sum += cv[id + (5 + l) * s];
        }
        sum = 1.0 / sum;
        const double f = cv[id] * sum;
        for (int l = 0; l < ns; ++l) {
          b->sv[id + l * s] = cv[id + (5 + l) * s] * sum;
          cv[id + (5 + l) * s] *= f;
        }
        temperature_pressure(b, id);
      }
}

double orc_mix_cfl_dt(orc_mix *b) {
  const long s = b->sz;
  if (b->viscous) mix_transport(b); /* the properties of the current state (update_physical_properties) */
  double dtmin = 1e+6;
#pragma omp parallel for collapse(2) reduction(min : dtmin)
  for (int k = 0; k < b->mz; ++k)
    for (int j = 0; j < b->my; ++j)
      for (int i = 0; i < b->mx; ++i) {
        const long id = ID(b, i, j, k);
        const double *M = b->met, *bv = b->bv;
        const double gx = sqrt(M[id] * M[id] + M[id + s] * M[id + s] + M[id + 2 * s] * M[id + 2 * s]);
        const double ge = sqrt(M[id + 3 * s] * M[id + 3 * s] + M[id + 4 * s] * M[id + 4 * s] +
                               M[id + 5 * s] * M[id + 5 * s]);
        const double gz = sqrt(M[id + 6 * s] * M[id + 6 * s] + M[id + 7 * s] * M[id + 7 * s] +
                               M[id + 8 * s] * M[id + 8 * s]);
        const double u = bv[id + s], v = bv[id + 2 * s], w = bv[id + 3 * s];
        const double U = u * M[id] + v * M[id + s] + w * M[id + 2 * s];
        const double V = u * M[id + 3 * s] + v * M[id + 4 * s] + w * M[id + 5 * s];
        const double W = u * M[id + 6 * s] + v * M[id + 7 * s] + w * M[id + 8 * s];
        const double c = b->c[id];
        double sr = fmax(fabs(U) + c * gx, fabs(V) + c * ge);
        if (b->dim == 3) sr = fmax(sr, fabs(W) + c * gz);
        if (b->viscous) { /* viscous spectral radius, TimeAdvanceFunc.cuh:66-98 */
          double ml = fmax(gx, ge);
          if (b->dim == 3) ml = fmax(ml, gz);
          double h[MS], cpl[MS], cpt = 0, R = 0, rmax = 0;
          orc_chem_h_cp(&b->m, bv[id + 5 * s], h, cpl);
          for (int l = 0; l < b->ns; ++l) {
            cpt += cpl[l] * b->sv[id + l * s];
            R += b->sv[id + l * s] * kRu / b->m.mw[l];
            rmax = fmax(rmax, b->rhod[id + l * s]);
          }
          const double ir = 1.0 / bv[id], gam = cpt / (cpt - R);
          double dv = b->mul[id] * ir;
          dv = fmax(dv, b->cond[id] * ir * gam / cpt);
          dv = fmax(dv, rmax * ir);
          sr = fmax(sr, ml * ml * dv);
        }
        double dt = b->p.cfl / sr;
        if (b->explicit_source) { /* TimeAdvanceFunc.cuh:104-111 */
          const double rdt = b->p.cfl * b->tscale[IDQ(b, i, j, k)];
          if (rdt > 1e-20 && rdt < 1e-5) dt = fmin(dt, rdt);
        }
        if (dt < dtmin) dtmin = dt;
      }
  return dtmin;
}

void orc_mix_set_step(orc_mix *b, double dt) { b->pp_dt = dt; }

static void rk3_stages(orc_mix *b, double dt);

void orc_mix_rk3_step(orc_mix *b, double dt) {
  sponge_mean(b);
  rk3_stages(b, dt);
}

static void rk3_stages(orc_mix *b, double dt) {
  b->pp_dt = dt;
  shock_sensor(b);
  memcpy(b->qn, b->cv, sizeof(double) * b->nv * b->sz);
  for (int rk = 0; rk < 3; ++rk) {
    rhs_core(b);
    update(b, dt, rk);
    apply_bcs(b);
    all_properties(b);
  }
}

/* Exact chemistry over dt in every interior cell at constant density and internal energy (ROK4E, as
 * orc_chem_rok4e), the step size carried per cell; then species, temperature and pressure, boundary
 * conditions and properties.  Returns the number of cells that did not reach dt. */
int orc_mix_chem_substep(orc_mix *b, double dt, double rtol, double atol, int max_attempts) {
  const long s = b->sz;
  const int ns = b->ns;
  int failed = 0;
#pragma omp parallel for collapse(2) reduction(+ : failed) schedule(dynamic)
  for (int k = 0; k < b->mz; ++k)
    for (int j = 0; j < b->my; ++j)
      for (int i = 0; i < b->mx; ++i) {
        const long id = ID(b, i, j, k), iq = IDQ(b, i, j, k);
        double *cv = b->cv, *bv = b->bv;
        const double rho = cv[id];
        const double u = bv[id + s], v = bv[id + 2 * s], w = bv[id + 3 * s];
        const double e = cv[id + 4 * s] / rho - 0.5 * (u * u + v * v + w * w);
        double y[MS], T = bv[id + 5 * s];
        for (int l = 0; l < ns; ++l) y[l] = cv[id + (5 + l) * s];
        double h = b->hstep[iq] > 0 ? b->hstep[iq] : dt;
        int rej = 0;
        const int n = orc_chem_rok4e(&b->m, rho, e, y, &T, dt, rtol, atol, &h, max_attempts, &rej, 0, 0);
        if (n < 0) ++failed;
        b->hstep[iq] = h;
        double R = 0;
        for (int l = 0; l < ns; ++l) {
          cv[id + (5 + l) * s] = y[l];
          b->sv[id + l * s] = y[l] / rho;
          R += y[l] / rho * kRu / b->m.mw[l];
        }
        bv[id + 5 * s] = T;
        bv[id + 4 * s] = rho * R * T;
      }
  apply_bcs(b);
  all_properties(b);
  return failed;
}

/* Strang splitting: chemistry dt/2, one SSP-RK3 flow step dt, chemistry dt/2. */
int orc_mix_strang_step(orc_mix *b, double dt, double rtol, double atol, int max_attempts) {
  sponge_mean(b);
  int f = orc_mix_chem_substep(b, 0.5 * dt, rtol, atol, max_attempts);
  rk3_stages(b, dt);
  f += orc_mix_chem_substep(b, 0.5 * dt, rtol, atol, max_attempts);
  return f;
}

/* Species clipping and renormalisation, then temperature and pressure, of node id after cv changed
 * (the tail of wu_zeroDReaction, StrangSplitting.cu:108-132 and 196-224). */
static void wu_node_state(orc_mix *b, long id) {
  const long s = b->sz;
  double *cv = b->cv, *bv = b->bv;
  bv[id] = cv[id];
  const double di = 1.0 / cv[id];
  bv[id + s] = cv[id + s] * di;
  bv[id + 2 * s] = cv[id + 2 * s] * di;
  bv[id + 3 * s] = cv[id + 3 * s] * di;
  double sum = 0;
  for (int l = 0; l < b->ns; ++l) {
    if (cv[id + (5 + l) * s] < 0) cv[id + (5 + l) * s] = 0;
    else sum += cv[id + (5 + l) * s];
  }
  sum = 1.0 / sum;
  const double f = cv[id] * sum;
  for (int l = 0; l < b->ns; ++l) {
    b->sv[id + l * s] = cv[id + (5 + l) * s] * sum;
    cv[id + (5 + l) * s] *= f;
  }
  temperature_pressure(b, id);
}

/* wu_zeroDReaction (StrangSplitting.cu:43-225) over h at node id: the chemical source at the current state
 * and the frozen transport cn advance cv by forward Euler (chem_advance_method 0), by forward Euler of the EPI
 * source (I - h J)^-1 omega (method 1, StrangSplitting.cu:157-181: J_mn = 1e3 MW_m / (rho Y_n)
 * sum_r (nu''_rm - nu'_rm)(nu'_rn q1_r - nu''_rn q2_r), rho Y_n recovered from the concentrations, columns with
 * rho Y_n <= 1e-25 zero) or by the three SSP-RK3 stages with that same source (method 3, whose rates are not
 * re-evaluated between stages); returns the reaction time scale min_k rho Y_k / omega_d,k. */
static double wu_zero_d(orc_mix *b, long id, const double *cn, double h, int method) {
  const long s = b->sz;
  const int ns = b->ns, nv = b->nv;
  double *cv = b->cv;
  const double iJ = 1.0 / b->jac[id];
  double qn[NV], ry[MS] = {0}, q1[ORC_MAX_REAC], q2[ORC_MAX_REAC], wd[MS], w[MS];
  for (int l = 0; l < nv; ++l) qn[l] = cv[id + l * s];
  for (int l = 0; l < ns; ++l) ry[l] = cv[id + (5 + l) * s];
  orc_chem_source(&b->m, b->bv[id + 5 * s], ry, q1, q2, wd, w);
  if (method == 1) {
    const orc_mech *m = &b->m;
    double a[MS * MS];
    int piv[MS];
    for (int r = 0; r < ns; ++r)
      for (int c = 0; c < ns; ++c) {
        const double rhoY = ry[c] * (1.0 / m->mw[c]) * 1e-3 * 1e+3 / (1.0 / m->mw[c]);
        double zz = 0;
        if (rhoY > 1e-25) {
          for (int q = 0; q < m->nr; ++q) zz += (m->nb[q][r] - m->nf[q][r]) * (m->nf[q][c] * q1[q] - m->nb[q][c] * q2[q]);
          zz /= rhoY;
        }
        a[r * ns + c] = (r == c ? 1.0 : 0.0) - h * (zz * 1e+3 * m->mw[r]);
      }
    orc_lu_decomp(a, ns, piv);
    orc_lu_solve(a, w, ns, piv);
  }
  double ts = 1e+6;
  if (method == 3) {
    for (int st = 0; st < 3; ++st) {
      for (int l = 0; l < 5; ++l)
        cv[id + l * s] = rk_a[st] * qn[l] + rk_b[st] * cv[id + l * s] + rk_c[st] * h * iJ * cn[l];
      ts = 1e+6;
      for (int l = 0; l < ns; ++l) {
        const double y = cv[id + (5 + l) * s];
        if (y > 1e-25 && fabs(wd[l]) > 1e-25) ts = fmin(ts, fabs(y / wd[l]));
        cv[id + (5 + l) * s] = rk_a[st] * qn[5 + l] + rk_b[st] * y + rk_c[st] * h * (w[l] + cn[5 + l] * iJ);
      }
      wu_node_state(b, id);
    }
    return ts;
  }
  for (int l = 0; l < 5; ++l) cv[id + l * s] += cn[l] * h * iJ;
  for (int l = 0; l < ns; ++l) {
    const double y = cv[id + (5 + l) * s];
    if (y > 1e-25 && fabs(wd[l]) > 1e-25) ts = fmin(ts, fabs(y / wd[l]));
    cv[id + (5 + l) * s] = y + (w[l] + cn[5 + l] * iJ) * h;
  }
  wu_node_state(b, id);
  return ts;
}

/* Wu's splitting, one step of size dt (StrangSplitting.cu:683-861): the transport residual Cn at U^n
 * (compute_cn, 15-41); per cell n_sub + 1 zero-dimensional reaction steps with Cn frozen, n_sub from the cell's
 * reaction time scale (wu_reaction_step, 373-421); boundary conditions; then three SSP-RK3 stages over dt/2 of
 * the transport change R(U) - Cn (wu_RK_subStep and update_cv_and_bv_rk_wu, 504-634). */
void orc_mix_wu_step(orc_mix *b, double dt, int method) {
  sponge_mean(b);
  const int nv = b->nv;
  const int es = b->explicit_source;
  b->explicit_source = 0;
  b->pp_dt = dt;
  orc_mix_rhs(b);
  memcpy(b->cn, b->dq, sizeof(double) * nv * b->n);
#pragma omp parallel for collapse(2) schedule(dynamic)
  for (int k = 0; k < b->mz; ++k)
    for (int j = 0; j < b->my; ++j)
      for (int i = 0; i < b->mx; ++i) {
        const long id = ID(b, i, j, k), iq = IDQ(b, i, j, k);
        double cn[NV];
        for (int l = 0; l < nv; ++l) cn[l] = b->cn[iq + l * b->n];
        double tau = b->wu_ts[iq];
        if (tau < 1e-15) tau = 0.2 * dt;
        const double ratio = dt / tau;
        const int n_sub = ratio > 1e6 ? 1000000 : (int)ratio;
        double h = dt / (n_sub + 1);
        for (int n = 0; n < n_sub; ++n) b->wu_ts[iq] = wu_zero_d(b, id, cn, h, method);
        h = dt - h * n_sub;
        if (fabs(h) > 1e-12) b->wu_ts[iq] = wu_zero_d(b, id, cn, h, method);
      }
  apply_bcs(b);
  all_properties(b);
  memcpy(b->qn, b->cv, sizeof(double) * nv * b->sz);
  for (int rk = 0; rk < 3; ++rk) {
    rhs_core(b); /* the sensor of compute_cn stays (StrangSplitting.cu:576 does not recompute it) */
    for (long q = 0; q < (long)nv * b->n; ++q) b->dq[q] -= b->cn[q];
    update(b, 0.5 * dt, rk);
    apply_bcs(b);
    all_properties(b);
  }
  b->explicit_source = es;
}

/* Interface flux of the mixture Riemann solvers for given primitive states (rho,u,v,w,p,Y_1..Y_ns). */
void orc_mix_point_flux(const orc_params *p, const orc_mech *m, const double *pl_in, const double *pr_in,
                        const double *ml, const double *mr, double jl, double jr, double efl, double efr, double *fc) {
  static orc_mix b; /* only p, m, ns, nv are read */
#pragma omp threadprivate(b)
  b.p = *p;
  b.m = *m;
  b.ns = m->ns;
  b.nv = 5 + m->ns;
  double pl[NV + 2], pr[NV + 2];
  memcpy(pl, pl_in, sizeof(double) * b.nv);
  memcpy(pr, pr_in, sizeof(double) * b.nv);
  post(&b, pl);
  post(&b, pr);
  if (p->inviscid_scheme == 2) roe(&b, pl, pr, ml, mr, jl, jr, efl, efr, fc);
  else if (p->inviscid_scheme == 4) hllc(&b, pl, pr, ml, mr, jl, jr, fc);
  else ausm(&b, pl, pr, ml, mr, jl, jr, fc);
}
