/*
 * oracle_abi.c -- the oracle behind the library's own C ABI (TEST INFRASTRUCTURE ONLY; see oracle.h): the entry points
 * of include/corefl_c.h that a single-block, single-species flow case uses, implemented over CPU memory with the
 * oracle's block (oracle_block.c), so that a test written against the ABI runs unchanged on either implementation.
 *
 * Scope: one block of calorically perfect air (species 0); the setup keys inviscid_scheme, reconstruction,
 * viscous_order, cfl, prandtl_number, entropy_fix_factor, gridScale, positive_preserving, shockSensor_threshold,
 * shock_sensor, shockSensor_epsilon, fixed_time_step, dt, total_simulation_time, canonical_problem (taylorGreen),
 * reference_state, default_init and groups_init with init_cond_l boxes (inflow structs: density / pressure /
 * temperature, mach or velocity, direction u v w); boundary types 6 (outflow) and 8 (periodic).  Anything else is an error.  corefl_rhs runs stage -1 only
 * (the oracle advances whole SSP-RK3 steps); corefl_advance takes CFL steps one at a time.
 *
 * Rules restated here from the reference: the ghost-layer count (Parameter.cpp:197-305), the reference state and
 * v_inf (BoundCond.cpp:265-295: v_inf = |V|, or the speed of sound when |V| < 1), the WENO epsilon scale
 * 1e-40 (rho_inf v_inf)^2 (DParameter.cu:25-28, WENO.cu:8), the inflow state of a struct (BoundCond.cpp:30-262 for
 * air) and the clipped final step (total_simulation_time).  The Taylor-Green field is the textbook one: u = V0 sin x
 * cos y cos z, v = -V0 cos x sin y cos z, w = 0, p = p0 + rho0 V0^2 / 16 (cos 2x + cos 2y)(cos 2z + 2), T = T0.
 */
#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include "oracle.h"

static const double kGamma = 1.4, kRu = 8314.462618, kMwAir = 28.9647;

/* ---- setup text: typed "key = value" lines at the top level and inside "struct name { ... }" ---- */

typedef struct {
  char scope[128], key[128], val[128];
} entry;

typedef struct {
  entry *e;
  int n, cap;
} table;

static void put(table *t, const char *scope, const char *key, const char *val) {
  if (t->n == t->cap) {
    t->cap = t->cap ? 2 * t->cap : 64;
    t->e = (entry *)realloc(t->e, sizeof(entry) * t->cap);
  }
  entry *x = &t->e[t->n++];
  snprintf(x->scope, sizeof x->scope, "%s", scope);
  snprintf(x->key, sizeof x->key, "%s", key);
  snprintf(x->val, sizeof x->val, "%s", val);
}

static const char *find(const table *t, const char *scope, const char *key) {
  for (int i = t->n - 1; i >= 0; --i)
    if (!strcmp(t->e[i].scope, scope) && !strcmp(t->e[i].key, key)) return t->e[i].val;
  return NULL;
}

static void parse_setup(const char *text, table *t) {
  char scope[128] = "";
  int in_array = 0;
  const char *p = text;
  while (*p) {
    char line[512];
    int n = 0;
    while (*p && *p != '\n' && n < 511) line[n++] = *p++;
    line[n] = 0;
    if (*p == '\n') ++p;
    char *c = strstr(line, "//");
    if (c) *c = 0;
    char w[4][128];
    const int k = sscanf(line, "%127s %127s %127s %127s", w[0], w[1], w[2], w[3]);
    if (k <= 0) continue;
    if (!strcmp(w[0], "}")) {
      scope[0] = 0;
      in_array = 0;
      continue;
    }
    if (in_array) continue;
    if (!strcmp(w[0], "array")) {
      in_array = 1;
      continue;
    }
    if (!strcmp(w[0], "struct") && k >= 2) {
      snprintf(scope, sizeof scope, "%s", w[1]);
      continue;
    }
    if (k == 4 && !strcmp(w[2], "=")) put(t, scope, w[1], w[3]);
  }
}

static double get_real(const table *t, const char *scope, const char *key, double def) {
  const char *v = find(t, scope, key);
  return v ? atof(v) : def;
}

static const char *get_string(const table *t, const char *scope, const char *key, const char *def) {
  const char *v = find(t, scope, key);
  return v ? v : def;
}

/* ---- context ---- */

typedef struct {
  int mx, my, mz, ngg, n_var, n_spec;
  long long sz;
  double *cv, *bv, *qn, *metric, *jac, *dq;
} corefl_block_view;

typedef struct {
  orc_params p;
  orc_block *b;
  int mx, my, mz, ngg;
  long sz;
  int fixed_dt;
  double dt_fixed, t_end, time, dt, rho_inf, v_inf;
} corefl_ctx;

static char g_err[512];

static int fail(const char *msg) {
  snprintf(g_err, sizeof g_err, "%s", msg);
  return -1;
}

const char *corefl_last_error(void) { return g_err; }

typedef struct {
  double rho, u, v, w, p, T, c;
} flow_state;

/* The inflow state of a struct (BoundCond.cpp:30-262, air): two of density, pressure and temperature; the speed
 * from mach or velocity; the direction (u, v, w), default (1, 0, 0). */
static int inflow_state(const table *t, const char *name, flow_state *s) {
  const double R = kRu / kMwAir;
  double rho = get_real(t, name, "density", -1), p = get_real(t, name, "pressure", -1);
  double T = get_real(t, name, "temperature", -1);
  if (T < 0) {
    if (rho < 0 || p < 0) return fail("oracle: an inflow struct needs two of density, pressure and temperature");
    T = p / (rho * R);
  } else if (rho < 0) {
    if (p < 0) return fail("oracle: an inflow struct needs two of density, pressure and temperature");
    rho = p / (R * T);
  } else if (p < 0) {
    p = rho * R * T;
  }
  s->c = sqrt(kGamma * R * T);
  const double mach = get_real(t, name, "mach", -1);
  const double vel = mach >= 0 ? mach * s->c : get_real(t, name, "velocity", 0);
  s->rho = rho;
  s->p = p;
  s->T = T;
  s->u = get_real(t, name, "u", 1) * vel;
  s->v = get_real(t, name, "v", 0) * vel;
  s->w = get_real(t, name, "w", 0) * vel;
  return 0;
}

int corefl_create_mem(const char *setup_text, int n_block, const int *dims, const double *const *xyz,
                      const char *boundary_text, const char *inner_text, corefl_ctx **out) {
  if (!setup_text || !dims || !xyz || !boundary_text || !out) return fail("oracle: null argument");
  if (n_block != 1 || inner_text) return fail("oracle: one block only");
  table t = {0};
  parse_setup(setup_text, &t);
  int rc = 0;
  corefl_ctx *ctx = (corefl_ctx *)calloc(1, sizeof(corefl_ctx));
  double *x = NULL, *y = NULL, *z = NULL, *jac = NULL, *met = NULL, *bv = NULL, *gx = NULL, *gy = NULL, *gzc = NULL;
  int *faces = NULL;
  if (get_real(&t, "", "species", 0) != 0) {
    rc = fail("oracle: single-species air only");
    goto done;
  }
  orc_params *p = &ctx->p;
  p->inviscid_scheme = (int)get_real(&t, "", "inviscid_scheme", 3);
  p->reconstruction = (int)get_real(&t, "", "reconstruction", 2);
  p->viscous_order = (int)get_real(&t, "", "viscous_order", 2);
  p->prandtl = get_real(&t, "", "prandtl_number", 0.72);
  p->cfl = get_real(&t, "", "cfl", 1.0);
  p->entropy_fix_factor = get_real(&t, "", "entropy_fix_factor", 0.125);
  p->positive_preserving = (int)get_real(&t, "", "positive_preserving", 0);
  const double thr = get_real(&t, "", "shockSensor_threshold", 0);
  const int fsw = p->inviscid_scheme == 51 || p->inviscid_scheme == 52 || p->inviscid_scheme == 71 ||
                  p->inviscid_scheme == 72;
  p->sensor_threshold = fsw && thr > 1e-10 ? thr : 0;
  p->shock_sensor = (int)get_real(&t, "", "shock_sensor", 0);
  p->sensor_eps = get_real(&t, "", "shockSensor_epsilon", -1);
  /* ghost layers (Parameter.cpp:197-305): 2, 3 for WENO5, 4 for WENO7, at least 4 for viscous order 8 */
  int ngg = 2;
  if (p->reconstruction == 4 && !fsw) ngg = 3;
  if (p->reconstruction == 5 && !fsw) ngg = 4;
  if (p->inviscid_scheme == 51 || p->inviscid_scheme == 52) ngg = 3;
  if (p->inviscid_scheme == 71 || p->inviscid_scheme == 72) ngg = 4;
  if (p->viscous_order == 8 && ngg < 4) ngg = 4;
  ctx->ngg = ngg;
  ctx->fixed_dt = (int)get_real(&t, "", "fixed_time_step", 0) || !strcmp(get_string(&t, "", "fixed_time_step", "0"),
                                                                          "true");
  ctx->dt_fixed = get_real(&t, "", "dt", 0);
  ctx->t_end = get_real(&t, "", "total_simulation_time", 1e10);
  /* reference state (BoundCond.cpp:265-295) and the WENO epsilon (DParameter.cu:25-28, WENO.cu:8) */
  const char *problem = get_string(&t, "", "canonical_problem", "");
  const int tgv = !strcmp(problem, "taylorGreen");
  flow_state ref;
  double T0 = 0, p0 = 0, V0 = 0, rho0 = 0;
  if (tgv) {
    T0 = get_real(&t, "taylorGreen", "temperature", 300.0);
    p0 = get_real(&t, "taylorGreen", "pressure", 101325.0);
    const double R = kRu / kMwAir;
    rho0 = p0 / (R * T0);
    V0 = get_real(&t, "taylorGreen", "mach", 0.1) * sqrt(kGamma * R * T0);
    ctx->rho_inf = rho0;
    ctx->v_inf = V0;
  } else {
    if ((rc = inflow_state(&t, get_string(&t, "", "reference_state", ""), &ref)) != 0) goto done;
    const double speed = sqrt(ref.u * ref.u + ref.v * ref.v + ref.w * ref.w);
    ctx->rho_inf = ref.rho;
    ctx->v_inf = fabs(speed) < 1 ? ref.c : speed;
  }
  p->v_ref = ctx->v_inf;
  p->rho_ref = ctx->rho_inf;
  p->weno_eps = 1e-40 * ctx->rho_inf * ctx->v_inf * ctx->rho_inf * ctx->v_inf;
  /* grid (scaled by gridScale), metrics, faces */
  const int mx = dims[0], my = dims[1], mz = dims[2];
  ctx->mx = mx;
  ctx->my = my;
  ctx->mz = mz;
  const long n = (long)mx * my * mz;
  const long sz = (long)(mx + 2 * ngg) * (my + 2 * ngg) * (mz + 2 * ngg);
  ctx->sz = sz;
  const double scale = get_real(&t, "", "gridScale", 1.0);
  x = (double *)malloc(sizeof(double) * n);
  y = (double *)malloc(sizeof(double) * n);
  z = (double *)malloc(sizeof(double) * n);
  for (long q = 0; q < n; ++q) {
    x[q] = xyz[0][q] * scale;
    y[q] = xyz[1][q] * scale;
    z[q] = xyz[2][q] * scale;
  }
  jac = (double *)malloc(sizeof(double) * sz);
  met = (double *)malloc(sizeof(double) * 9 * sz);
  if (orc_metrics(mx, my, mz, ngg, x, y, z, jac, met) != 0) {
    rc = fail("oracle: non-positive Jacobian");
    goto done;
  }
  int n_face = 0;
  const char *bp = boundary_text;
  int used = 0;
  if (sscanf(bp, "%d%n", &n_face, &used) != 1 || n_face < 0) {
    rc = fail("oracle: bad boundary text");
    goto done;
  }
  bp += used;
  faces = (int *)malloc(sizeof(int) * 7 * (n_face > 0 ? n_face : 1));
  for (int f = 0; f < n_face; ++f)
    for (int c = 0; c < 7; ++c) {
      if (sscanf(bp, "%d%n", &faces[7 * f + c], &used) != 1) {
        rc = fail("oracle: bad boundary text");
        goto done;
      }
      bp += used;
    }
  for (int f = 0; f < n_face; ++f)
    if (faces[7 * f + 6] != 6 && faces[7 * f + 6] != 8) {
      rc = fail("oracle: boundary types 6 and 8 only");
      goto done;
    }
  ctx->b = orc_block_create(p, mx, my, mz, ngg, faces, n_face, jac, met);
  /* initial state on every node of the ghost layout, at the mirrored ghost grid's coordinates there, then the
   * boundary conditions (orc_block_set_bv applies them) */
  bv = (double *)calloc(6 * sz, sizeof(double));
  gx = (double *)malloc(sizeof(double) * sz);
  gy = (double *)malloc(sizeof(double) * sz);
  gzc = (double *)malloc(sizeof(double) * sz);
  orc_ghost_coords(mx, my, mz, ngg, x, y, z, gx, gy, gzc);
  for (long id = 0; id < sz; ++id) {
    const double xq = gx[id], yq = gy[id], zq = gzc[id];
    double s[6];
    if (tgv) {
      const double X = xq / scale, Y = yq / scale, Z = zq / scale, R = kRu / kMwAir;
      const double pr = p0 + rho0 * V0 * V0 / 16 * (cos(2 * X) + cos(2 * Y)) * (cos(2 * Z) + 2);
      s[0] = pr / (R * T0);
      s[1] = V0 * sin(X) * cos(Y) * cos(Z);
      s[2] = -V0 * cos(X) * sin(Y) * cos(Z);
      s[3] = 0;
      s[4] = pr;
      s[5] = T0;
    } else {
      /* default_init, or the first init_cond_l box (x0 x1 y0 y1 z0 z1; its state the struct "name") holding
       * the node (Initialize.cu:51-95 with groups_init) */
      const char *name = get_string(&t, "", "default_init", "");
      const int groups = (int)get_real(&t, "", "groups_init", 1);
      char sc[32];
      for (int l = 0; l < groups - 1; ++l) {
        snprintf(sc, sizeof sc, "init_cond_%d", l);
        if (xq >= get_real(&t, sc, "x0", 0) && xq <= get_real(&t, sc, "x1", 0) &&
            yq >= get_real(&t, sc, "y0", 0) && yq <= get_real(&t, sc, "y1", 0) &&
            zq >= get_real(&t, sc, "z0", 0) && zq <= get_real(&t, sc, "z1", 0)) {
          name = get_string(&t, sc, "name", sc);
          break;
        }
      }
      flow_state st;
      if ((rc = inflow_state(&t, name, &st)) != 0) goto done;
      s[0] = st.rho;
      s[1] = st.u;
      s[2] = st.v;
      s[3] = st.w;
      s[4] = st.p;
      s[5] = st.T;
    }
    for (int c = 0; c < 6; ++c) bv[id + c * sz] = s[c];
  }
  orc_block_set_bv(ctx->b, bv);
done:
  free(x);
  free(y);
  free(z);
  free(jac);
  free(met);
  free(faces);
  free(bv);
  free(gx);
  free(gy);
  free(gzc);
  free(t.e);
  if (rc != 0) {
    if (ctx->b) orc_block_destroy(ctx->b);
    free(ctx);
    *out = NULL;
    return rc;
  }
  *out = ctx;
  return 0;
}

void corefl_destroy(corefl_ctx *ctx) {
  if (!ctx) return;
  orc_block_destroy(ctx->b);
  free(ctx);
}

int corefl_n_blocks(const corefl_ctx *ctx) { return ctx ? 1 : 0; }

int corefl_block_view_get(corefl_ctx *ctx, int block, corefl_block_view *v) {
  if (!ctx || !v || block != 0) return fail("oracle: bad block");
  memset(v, 0, sizeof *v);
  v->mx = ctx->mx;
  v->my = ctx->my;
  v->mz = ctx->mz;
  v->ngg = ctx->ngg;
  v->n_var = 5;
  v->n_spec = 0;
  v->sz = ctx->sz;
  return 0;
}

/* which: 0 cv, 1 bv, 2 dq, 4 jac, 5 metric (the array codes of corefl_c.h). */
int corefl_get_array(corefl_ctx *ctx, int block, int which, double *host, long long n) {
  if (!ctx || block != 0 || !host) return fail("oracle: bad arguments");
  const long nn = (long)ctx->mx * ctx->my * ctx->mz;
  long want;
  switch (which) {
    case 0: want = 5 * ctx->sz; break;
    case 1: want = 6 * ctx->sz; break;
    case 2: want = 5 * nn; break;
    case 4: want = ctx->sz; break;
    case 5: want = 9 * ctx->sz; break;
    default: return fail("oracle: array not available");
  }
  if (n != want) return fail("oracle: wrong element count");
  orc_block_get(ctx->b, which, host);
  return 0;
}

int corefl_set_state(corefl_ctx *ctx, int block, const double *bv_host, long long n) {
  if (!ctx || block != 0 || !bv_host) return fail("oracle: bad arguments");
  if (n != 6 * ctx->sz) return fail("oracle: state must have 6*sz elements");
  orc_block_set_bv(ctx->b, bv_host);
  return 0;
}

int corefl_rhs(corefl_ctx *ctx, int stage, void *stream) {
  if (!ctx) return fail("oracle: null context");
  if (stage != -1) return fail("oracle: stage -1 (residual) only");
  orc_block_rhs(ctx->b);
  return 0;
}

int corefl_compute_dt(corefl_ctx *ctx, double *dt_out) {
  if (!ctx || !dt_out) return fail("oracle: null argument");
  double dt = orc_block_cfl_dt(ctx->b);
  if (dt > ctx->t_end - ctx->time) dt = ctx->t_end - ctx->time;
  *dt_out = ctx->dt = dt;
  return 0;
}

int corefl_set_dt(corefl_ctx *ctx, double dt) {
  if (!ctx || !(dt > 0)) return fail("oracle: dt must be positive");
  ctx->dt = dt;
  return 0;
}

int corefl_rk3_step(corefl_ctx *ctx, double *dt_inout, void *stream) {
  if (!ctx || !dt_inout) return fail("oracle: null argument");
  double dt = *dt_inout;
  if (!(dt > 0)) {
    if (ctx->fixed_dt) {
      dt = ctx->dt_fixed;
    } else {
      corefl_compute_dt(ctx, &dt);
    }
  }
  orc_block_rk3_step(ctx->b, dt);
  ctx->time += dt;
  *dt_inout = dt;
  return 0;
}

int corefl_advance(corefl_ctx *ctx, int n_steps, int use_graph, void *stream) {
  for (int s = 0; s < n_steps; ++s) {
    double dt = 0;
    if (corefl_rk3_step(ctx, &dt, NULL) != 0) return -1;
  }
  return 0;
}

int corefl_get_time(corefl_ctx *ctx, double *t) {
  if (!ctx || !t) return fail("oracle: null argument");
  *t = ctx->time;
  return 0;
}

int corefl_synchronize(corefl_ctx *ctx) { return ctx ? 0 : fail("oracle: null context"); }

int corefl_get_int(const corefl_ctx *ctx, const char *name, int *out) {
  if (!ctx || !name || !out) return fail("oracle: null argument");
  if (!strcmp(name, "ngg")) *out = ctx->ngg;
  else if (!strcmp(name, "n_var")) *out = 5;
  else if (!strcmp(name, "dimension")) *out = ctx->mz == 1 ? 2 : 3;
  else if (!strcmp(name, "inviscid_scheme")) *out = ctx->p.inviscid_scheme;
  else return fail("oracle: unknown integer parameter");
  return 0;
}

int corefl_get_real(const corefl_ctx *ctx, const char *name, double *out) {
  if (!ctx || !name || !out) return fail("oracle: null argument");
  if (!strcmp(name, "weno_eps")) *out = ctx->p.weno_eps;
  else if (!strcmp(name, "v_inf")) *out = ctx->v_inf;
  else if (!strcmp(name, "rho_inf")) *out = ctx->rho_inf;
  else if (!strcmp(name, "dt")) *out = ctx->dt;
  else return fail("oracle: unknown real parameter");
  return 0;
}
