/*
 * oracle_block.c -- mesh metrics and the per-block SSP-RK3 step of the oracle
 * (TEST INFRASTRUCTURE ONLY; see oracle.h).
 */
#include "oracle.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

static const double kGamma = 1.4;
static const double kRu = 8314.462618;
static const double kMwAir = 28.9647;

int orc_num_threads(void) {
#ifdef _OPENMP
  return omp_get_max_threads();
#else
  return 1;
#endif
}

void orc_set_threads(int n) {
#ifdef _OPENMP
  if (n > 0) omp_set_num_threads(n);
#else
  (void)n;
#endif
}

/* ---------------- mesh: ghost grid + metrics ---------------- */

typedef struct {
  int mx, my, mz, g; /* g = number of ghost layers of this array */
  long d2, d1;
  double *v;
} arr3;

static inline double *A3(arr3 *a, int i, int j, int k) {
  return a->v + (long)(k + a->g) * a->d1 + (long)(j + a->g) * a->d2 + (i + a->g);
}

static void arr3_init(arr3 *a, int mx, int my, int mz, int g) {
  a->mx = mx; a->my = my; a->mz = mz; a->g = g;
  a->d2 = mx + 2 * g;
  a->d1 = (long)(my + 2 * g) * a->d2;
  a->v = (double *)calloc((size_t)a->d1 * (mz + 2 * g), sizeof(double));
}

/* Mesh.cpp:1013-1167 (single block, no inner/parallel faces) */
static void ghost_grid(arr3 *X, arr3 *Y, arr3 *Z, int ngg, int dim) {
  const int mx = X->mx, my = X->my, mz = X->mz;
  arr3 *c[3] = {X, Y, Z};
  for (int q = 0; q < 3; ++q) {
    arr3 *B = c[q];
    for (int k = 0; k < mz; ++k)
      for (int j = 0; j < my; ++j) {
        for (int i = -1; i >= -ngg - 1; --i) *A3(B, i, j, k) = 2 * *A3(B, 0, j, k) - *A3(B, -i, j, k);
        for (int i = mx; i < mx + ngg + 1; ++i)
          *A3(B, i, j, k) = 2 * *A3(B, mx - 1, j, k) - *A3(B, 2 * (mx - 1) - i, j, k);
      }
    for (int k = 0; k < mz; ++k)
      for (int i = 0; i < mx; ++i) {
        for (int j = -1; j >= -ngg - 1; --j) *A3(B, i, j, k) = 2 * *A3(B, i, 0, k) - *A3(B, i, -j, k);
        for (int j = my; j < my + ngg + 1; ++j)
          *A3(B, i, j, k) = 2 * *A3(B, i, my - 1, k) - *A3(B, i, 2 * (my - 1) - j, k);
      }
    if (dim == 3) {
      for (int j = 0; j < my; ++j)
        for (int i = 0; i < mx; ++i) {
          for (int k = -1; k >= -ngg - 1; --k) *A3(B, i, j, k) = 2 * *A3(B, i, j, 0) - *A3(B, i, j, -k);
          for (int k = mz; k < mz + ngg + 1; ++k)
            *A3(B, i, j, k) = 2 * *A3(B, i, j, mz - 1) - *A3(B, i, j, 2 * (mz - 1) - k);
        }
    } else {
      for (int j = 0; j < my; ++j)
        for (int i = 0; i < mx; ++i) {
          for (int k = -1; k >= -ngg - 1; --k) *A3(B, i, j, k) = *A3(B, i, j, 0) + (q == 2 ? k : 0);
          for (int k = mz; k < mz + ngg + 1; ++k) *A3(B, i, j, k) = *A3(B, i, j, mz - 1) + (q == 2 ? k - mz + 1 : 0);
        }
    }
  }
  /* edges and corners by the parallelogram rule */
  for (int q = 0; q < 3; ++q) {
    arr3 *B = c[q];
    for (int i = 1; i <= ngg + 1; ++i)
      for (int j = 1; j <= ngg + 1; ++j)
        for (int k = 0; k < mz; ++k) {
          const int i1 = -i, j1 = -j, i2 = mx - 1 + i, j2 = my - 1 + j;
          *A3(B, i1, j1, k) = *A3(B, i1, 0, k) + *A3(B, 0, j1, k) - *A3(B, 0, 0, k);
          *A3(B, i2, j2, k) = *A3(B, i2, my - 1, k) + *A3(B, mx - 1, j2, k) - *A3(B, mx - 1, my - 1, k);
          *A3(B, i1, j2, k) = *A3(B, i1, my - 1, k) + *A3(B, 0, j2, k) - *A3(B, 0, my - 1, k);
          *A3(B, i2, j1, k) = *A3(B, i2, 0, k) + *A3(B, mx - 1, j1, k) - *A3(B, mx - 1, 0, k);
        }
    for (int i = 1; i <= ngg + 1; ++i)
      for (int j = 0; j < my; ++j)
        for (int k = 1; k <= ngg + 1; ++k) {
          const int i1 = -i, k1 = -k, i2 = mx - 1 + i, k2 = mz - 1 + k;
          *A3(B, i1, j, k1) = *A3(B, i1, j, 0) + *A3(B, 0, j, k1) - *A3(B, 0, j, 0);
          *A3(B, i2, j, k2) = *A3(B, i2, j, mz - 1) + *A3(B, mx - 1, j, k2) - *A3(B, mx - 1, j, mz - 1);
          *A3(B, i1, j, k2) = *A3(B, i1, j, mz - 1) + *A3(B, 0, j, k2) - *A3(B, 0, j, mz - 1);
          *A3(B, i2, j, k1) = *A3(B, i2, j, 0) + *A3(B, mx - 1, j, k1) - *A3(B, mx - 1, j, 0);
        }
    for (int i = -ngg - 1; i <= mx + ngg; ++i)
      for (int j = 1; j <= ngg + 1; ++j)
        for (int k = 1; k <= ngg + 1; ++k) {
          const int j1 = -j, k1 = -k, j2 = my - 1 + j, k2 = mz - 1 + k;
          *A3(B, i, j1, k1) = *A3(B, i, j1, 0) + *A3(B, i, 0, k1) - *A3(B, i, 0, 0);
          *A3(B, i, j2, k2) = *A3(B, i, j2, mz - 1) + *A3(B, i, my - 1, k2) - *A3(B, i, my - 1, mz - 1);
          *A3(B, i, j1, k2) = *A3(B, i, j1, mz - 1) + *A3(B, i, 0, k2) - *A3(B, i, 0, mz - 1);
          *A3(B, i, j2, k1) = *A3(B, i, j2, 0) + *A3(B, i, my - 1, k1) - *A3(B, i, my - 1, 0);
        }
  }
}

typedef struct {
  arr3 *X, *Y, *Z;
  double *jac, *met;
  long sz, d2, d1;
  int g;
} metric_ctx;

/* One node of Mesh.cpp:255-293; returns 0 if jac <= 0 (caller applies the fallback). */
static int metric_point(metric_ctx *m, int i, int j, int k) {
  arr3 *X = m->X, *Y = m->Y, *Z = m->Z;
  const double dxd1 = (*A3(X, i + 1, j, k) - *A3(X, i - 1, j, k)) * 0.5;
  const double dyd1 = (*A3(Y, i + 1, j, k) - *A3(Y, i - 1, j, k)) * 0.5;
  const double dzd1 = (*A3(Z, i + 1, j, k) - *A3(Z, i - 1, j, k)) * 0.5;
  const double dxd2 = (*A3(X, i, j + 1, k) - *A3(X, i, j - 1, k)) * 0.5;
  const double dyd2 = (*A3(Y, i, j + 1, k) - *A3(Y, i, j - 1, k)) * 0.5;
  const double dzd2 = (*A3(Z, i, j + 1, k) - *A3(Z, i, j - 1, k)) * 0.5;
  const double dxd3 = (*A3(X, i, j, k + 1) - *A3(X, i, j, k - 1)) * 0.5;
  const double dyd3 = (*A3(Y, i, j, k + 1) - *A3(Y, i, j, k - 1)) * 0.5;
  const double dzd3 = (*A3(Z, i, j, k + 1) - *A3(Z, i, j, k - 1)) * 0.5;
  const double jac = dxd1 * dyd2 * dzd3 + dxd2 * dyd3 * dzd1 + dxd3 * dyd1 * dzd2 - dxd1 * dyd3 * dzd2 -
                     dxd2 * dyd1 * dzd3 - dxd3 * dyd2 * dzd1;
  if (jac <= 0) return 0;
  const long id = (long)(k + m->g) * m->d1 + (long)(j + m->g) * m->d2 + (i + m->g);
  m->jac[id] = jac;
  double *M = m->met;
  M[id + 0 * m->sz] = (dyd2 * dzd3 - dyd3 * dzd2) / jac;
  M[id + 1 * m->sz] = (dxd3 * dzd2 - dxd2 * dzd3) / jac;
  M[id + 2 * m->sz] = (dxd2 * dyd3 - dxd3 * dyd2) / jac;
  M[id + 3 * m->sz] = (dyd3 * dzd1 - dyd1 * dzd3) / jac;
  M[id + 4 * m->sz] = (dxd1 * dzd3 - dxd3 * dzd1) / jac;
  M[id + 5 * m->sz] = (dxd3 * dyd1 - dxd1 * dyd3) / jac;
  M[id + 6 * m->sz] = (dyd1 * dzd2 - dyd2 * dzd1) / jac;
  M[id + 7 * m->sz] = (dxd2 * dzd1 - dxd1 * dzd2) / jac;
  M[id + 8 * m->sz] = (dxd1 * dyd2 - dxd2 * dyd1) / jac;
  return 1;
}

static void metric_copy(metric_ctx *m, int i, int j, int k, int si, int sj, int sk) {
  const long id = (long)(k + m->g) * m->d1 + (long)(j + m->g) * m->d2 + (i + m->g);
  const long is = (long)(sk + m->g) * m->d1 + (long)(sj + m->g) * m->d2 + (si + m->g);
  m->jac[id] = m->jac[is];
  for (int l = 0; l < 9; ++l) m->met[id + l * m->sz] = m->met[is + l * m->sz];
}

void orc_ghost_coords(int mx, int my, int mz, int ngg, const double *x, const double *y, const double *z, double *gx,
                      double *gy, double *gz) {
  const int dim = mz == 1 ? 2 : 3;
  arr3 X, Y, Z;
  arr3_init(&X, mx, my, mz, ngg + 1);
  arr3_init(&Y, mx, my, mz, ngg + 1);
  arr3_init(&Z, mx, my, mz, ngg + 1);
  for (int k = 0; k < mz; ++k)
    for (int j = 0; j < my; ++j)
      for (int i = 0; i < mx; ++i) {
        const long s = ((long)k * my + j) * mx + i;
        *A3(&X, i, j, k) = x[s];
        *A3(&Y, i, j, k) = y[s];
        *A3(&Z, i, j, k) = z[s];
      }
  ghost_grid(&X, &Y, &Z, ngg, dim);
  const long d2 = mx + 2 * ngg, d1 = (long)(my + 2 * ngg) * d2;
  for (int k = -ngg; k < mz + ngg; ++k)
    for (int j = -ngg; j < my + ngg; ++j)
      for (int i = -ngg; i < mx + ngg; ++i) {
        const long id = (long)(k + ngg) * d1 + (long)(j + ngg) * d2 + (i + ngg);
        const int kk = dim == 2 ? 0 : k;
        gx[id] = *A3(&X, i, j, kk);
        gy[id] = *A3(&Y, i, j, kk);
        gz[id] = *A3(&Z, i, j, kk);
      }
  free(X.v);
  free(Y.v);
  free(Z.v);
}

int orc_metrics(int mx, int my, int mz, int ngg, const double *x, const double *y, const double *z, double *jac,
                double *metric) {
  const int dim = mz == 1 ? 2 : 3;
  arr3 X, Y, Z;
  arr3_init(&X, mx, my, mz, ngg + 1);
  arr3_init(&Y, mx, my, mz, ngg + 1);
  arr3_init(&Z, mx, my, mz, ngg + 1);
  for (int k = 0; k < mz; ++k)
    for (int j = 0; j < my; ++j)
      for (int i = 0; i < mx; ++i) {
        const long s = ((long)k * my + j) * mx + i;
        *A3(&X, i, j, k) = x[s];
        *A3(&Y, i, j, k) = y[s];
        *A3(&Z, i, j, k) = z[s];
      }
  ghost_grid(&X, &Y, &Z, ngg, dim);
  metric_ctx m = {&X, &Y, &Z, jac, metric, 0, mx + 2 * ngg, 0, ngg};
  m.d1 = (long)(my + 2 * ngg) * m.d2;
  m.sz = m.d1 * (mz + 2 * ngg);
  int rc = 0;
  /* interior, Mesh.cpp:246-294 */
  const int kmin = dim == 2 ? 0 : 1, kmax = dim == 2 ? 1 : mz - 1;
  for (int i = 1; i < mx - 1; ++i)
    for (int j = 1; j < my - 1; ++j)
      for (int k = kmin; k < kmax; ++k)
        if (!metric_point(&m, i, j, k)) rc = -1;
  /* the six boundary blocks, Mesh.cpp:300-528, with the reference's neighbour fallback */
  for (int i = 1; i < mx - 1; ++i)
    for (int j = 1; j < my - 1; ++j) {
      for (int k = 0; k >= -ngg; --k)
        if (!metric_point(&m, i, j, k)) metric_copy(&m, i, j, k, i, j, k + 1);
    }
  for (int i = 1; i < mx - 1; ++i)
    for (int j = 1; j < my - 1; ++j)
      for (int k = mz - 1; k < mz + ngg; ++k)
        if (!metric_point(&m, i, j, k)) metric_copy(&m, i, j, k, i, j, k - 1);
  for (int i = 1; i < mx - 1; ++i)
    for (int j = 0; j >= -ngg; --j)
      for (int k = -ngg; k < mz + ngg; ++k)
        if (!metric_point(&m, i, j, k)) metric_copy(&m, i, j, k, i, j + 1, k);
  for (int i = 1; i < mx - 1; ++i)
    for (int j = my - 1; j < my + ngg; ++j)
      for (int k = -ngg; k < mz + ngg; ++k)
        if (!metric_point(&m, i, j, k)) metric_copy(&m, i, j, k, i, j - 1, k);
  for (int i = 0; i >= -ngg; --i)
    for (int j = -ngg; j < my + ngg; ++j)
      for (int k = -ngg; k < mz + ngg; ++k)
        if (!metric_point(&m, i, j, k)) metric_copy(&m, i, j, k, i + 1, j, k);
  for (int i = mx - 1; i < mx + ngg; ++i)
    for (int j = -ngg; j < my + ngg; ++j)
      for (int k = -ngg; k < mz + ngg; ++k)
        if (!metric_point(&m, i, j, k)) metric_copy(&m, i, j, k, i - 1, j, k);
  free(X.v);
  free(Y.v);
  free(Z.v);
  return rc;
}

/* ---------------- symmetric-conservative metrics (Mesh.cu:10-141, 136-250) ---------------- */

/* First derivative at index i of the degree-p Lagrange interpolant through p + 1 consecutive nodes, centred where
 * they fit in [-ng, n - 1 + ng] and shifted to one side near its ends (ggxl::Derivatives, gxl_lib/Math.cuh:6-101).
 * f(l) reads the node l along the direction. */
static double lag_deriv(const double *f0, long stride, int i, int n, int ng, int p) {
  int s = i - p / 2;
  if (i <= -ng + p / 2) s = -ng;
  else if (i >= n - 1 + ng - p / 2) s = n - 1 + ng - p;
  double D = 0;
  for (int l = s; l <= s + p; ++l) {
    double den = 1, num = 0;
    for (int a = s; a <= s + p; ++a)
      if (a != l) den *= (double)(l - a);
    for (int m = s; m <= s + p; ++m) {
      if (m == l) continue;
      double prod = 1;
      for (int a = s; a <= s + p; ++a)
        if (a != l && a != m) prod *= (double)(i - a);
      num += prod;
    }
    D += num / den * f0[(long)(l - i) * stride];
  }
  return D;
}

/* symmetric_conservative_metric = order (2, 4, 6) on a 3D block: the 18 products x_a y - y_a x of coordinates and
 * their order-`order` derivatives (compute_xyz123), the three combinations D123 (compute_D123), then the Jacobian
 * (sum of the D123 derivatives) / 6 and each metric as half a difference of product derivatives over it, both with
 * sixth-order derivatives (Mesh.cu:215-243).  Every node of the ngg-ghost layout, on the same mirrored ghost grid as
 * orc_metrics; returns -1 if a Jacobian is not positive. */
int orc_sc_metrics(int mx, int my, int mz, int ngg, const double *x, const double *y, const double *z, int order,
                   double *jac, double *metric) {
  if (mz == 1) return -2;
  arr3 X, Y, Z;
  arr3_init(&X, mx, my, mz, ngg + 1);
  arr3_init(&Y, mx, my, mz, ngg + 1);
  arr3_init(&Z, mx, my, mz, ngg + 1);
  for (int k = 0; k < mz; ++k)
    for (int j = 0; j < my; ++j)
      for (int i = 0; i < mx; ++i) {
        const long s = ((long)k * my + j) * mx + i;
        *A3(&X, i, j, k) = x[s];
        *A3(&Y, i, j, k) = y[s];
        *A3(&Z, i, j, k) = z[s];
      }
  ghost_grid(&X, &Y, &Z, ngg, 3);
  arr3 P[18], D[3];
  for (int q = 0; q < 18; ++q) arr3_init(&P[q], mx, my, mz, ngg);
  for (int q = 0; q < 3; ++q) arr3_init(&D[q], mx, my, mz, ngg);
  const int ext[3] = {mx, my, mz};
  /* derivative of array a along direction d (0 xi, 1 eta, 2 zeta) at (i, j, k), order p */
#define DER(a, d, p)                                                                                                 \
  lag_deriv(A3((a), i, j, k), (d) == 0 ? 1 : ((d) == 1 ? (a)->d2 : (a)->d1), (d) == 0 ? i : ((d) == 1 ? j : k),        \
            ext[d], ngg, (p))
#pragma omp parallel for collapse(2)
  for (int k = -ngg; k < mz + ngg; ++k)
    for (int j = -ngg; j < my + ngg; ++j)
      for (int i = -ngg; i < mx + ngg; ++i) {
        const double xv = *A3(&X, i, j, k), yv = *A3(&Y, i, j, k), zv = *A3(&Z, i, j, k);
        double dx[3], dy[3], dz[3];
        for (int d = 0; d < 3; ++d) {
          /* the coordinate arrays carry ngg + 1 layers; the stencils use ngg of them, as the reference does */
          dx[d] = DER(&X, d, order);
          dy[d] = DER(&Y, d, order);
          dz[d] = DER(&Z, d, order);
        }
        const double v[18] = {dx[0] * yv - dy[0] * xv, dy[1] * xv - dx[1] * yv, dy[0] * zv - dz[0] * yv,
                              dz[1] * yv - dy[1] * zv, dz[0] * xv - dx[0] * zv, dx[1] * zv - dz[1] * xv,
                              dx[2] * yv - dy[2] * xv, dy[0] * xv - dx[0] * yv, dy[2] * zv - dz[2] * yv,
                              dz[0] * yv - dy[0] * zv, dz[2] * xv - dx[2] * zv, dx[0] * zv - dz[0] * xv,
                              dx[1] * yv - dy[1] * xv, dy[2] * xv - dx[2] * yv, dy[1] * zv - dz[1] * yv,
                              dz[2] * yv - dy[2] * zv, dz[1] * xv - dx[1] * zv, dx[2] * zv - dz[2] * xv};
        for (int q = 0; q < 18; ++q) *A3(&P[q], i, j, k) = v[q];
      }
#pragma omp parallel for collapse(2)
  for (int k = -ngg; k < mz + ngg; ++k)
    for (int j = -ngg; j < my + ngg; ++j)
      for (int i = -ngg; i < mx + ngg; ++i) {
        const double xv = *A3(&X, i, j, k), yv = *A3(&Y, i, j, k), zv = *A3(&Z, i, j, k);
        /* D123 components 2, 1, 0 of compute_D123 */
        *A3(&D[2], i, j, k) = (DER(&P[0], 1, order) + DER(&P[1], 0, order)) * zv +
                              (DER(&P[2], 1, order) + DER(&P[3], 0, order)) * xv +
                              (DER(&P[4], 1, order) + DER(&P[5], 0, order)) * yv;
        *A3(&D[1], i, j, k) = (DER(&P[6], 0, order) + DER(&P[7], 2, order)) * zv +
                              (DER(&P[8], 0, order) + DER(&P[9], 2, order)) * xv +
                              (DER(&P[10], 0, order) + DER(&P[11], 2, order)) * yv;
        *A3(&D[0], i, j, k) = (DER(&P[12], 2, order) + DER(&P[13], 1, order)) * zv +
                              (DER(&P[14], 2, order) + DER(&P[15], 1, order)) * xv +
                              (DER(&P[16], 2, order) + DER(&P[17], 1, order)) * yv;
      }
  const long d2 = mx + 2 * ngg, d1 = (long)(my + 2 * ngg) * d2, sz = d1 * (mz + 2 * ngg);
  int rc = 0;
#pragma omp parallel for collapse(2) reduction(min : rc)
  for (int k = -ngg; k < mz + ngg; ++k)
    for (int j = -ngg; j < my + ngg; ++j)
      for (int i = -ngg; i < mx + ngg; ++i) {
        const double J = (DER(&D[2], 2, 6) + DER(&D[1], 1, 6) + DER(&D[0], 0, 6)) / 6.0;
        if (!(J > 0)) rc = -1;
        const long id = (long)(k + ngg) * d1 + (long)(j + ngg) * d2 + (i + ngg);
        jac[id] = J;
        const double h = 0.5 / J;
        metric[id + 0 * sz] = (DER(&P[14], 2, 6) - DER(&P[8], 1, 6)) * h;
        metric[id + 1 * sz] = (DER(&P[16], 2, 6) - DER(&P[10], 1, 6)) * h;
        metric[id + 2 * sz] = (DER(&P[12], 2, 6) - DER(&P[6], 1, 6)) * h;
        metric[id + 3 * sz] = (DER(&P[8], 0, 6) - DER(&P[2], 2, 6)) * h;
        metric[id + 4 * sz] = (DER(&P[10], 0, 6) - DER(&P[4], 2, 6)) * h;
        metric[id + 5 * sz] = (DER(&P[6], 0, 6) - DER(&P[0], 2, 6)) * h;
        metric[id + 6 * sz] = (DER(&P[2], 1, 6) - DER(&P[14], 0, 6)) * h;
        metric[id + 7 * sz] = (DER(&P[4], 1, 6) - DER(&P[16], 0, 6)) * h;
        metric[id + 8 * sz] = (DER(&P[0], 1, 6) - DER(&P[12], 0, 6)) * h;
      }
#undef DER
  for (int q = 0; q < 18; ++q) free(P[q].v);
  for (int q = 0; q < 3; ++q) free(D[q].v);
  free(X.v);
  free(Y.v);
  free(Z.v);
  return rc;
}

/* ---------------- block state and step ---------------- */

typedef struct {
  int face, dir, type;
  int s[3], e[3];
} face_t;

/* Inflow face data (orc_block_set_inflow) */
typedef struct {
  double *prof, *rms;
  int sub_node, sub_ghost, fluct;
  double mw;
  unsigned long long seed, call;
} inflow_t;

struct orc_block {
  orc_params p;
  int mx, my, mz, g, dim;
  long d2, d1, sz, n;
  face_t faces[64];
  inflow_t inflow[64];
  int n_face;
  int bt[6]; /* physical-boundary flags per side for the one-sided viscous closures */
  double *cv, *bv, *qn, *dq, *mul, *jac, *met, *fv, *gv, *hv, *efd;
  double *sponge_sigma, *sponge_mean; /* sponge layer: damping rate (n), running mean of cv (5 * n) */
  unsigned long long sponge_iter;
  double *sensor; /* shock sensor (sz; ghost entries stay 0) of the hybrid flux-splitting scheme */
};

#define ID(b, i, j, k) ((long)((k) + (b)->g) * (b)->d1 + (long)((j) + (b)->g) * (b)->d2 + ((i) + (b)->g))
#define IDQ(b, i, j, k) (((long)(k) * (b)->my + (j)) * (b)->mx + (i))

orc_block *orc_block_create(const orc_params *p, int mx, int my, int mz, int ngg, const int *faces, int n_face,
                            const double *jac, const double *metric) {
  orc_block *b = (orc_block *)calloc(1, sizeof(orc_block));
  b->p = *p;
  b->mx = mx; b->my = my; b->mz = mz; b->g = ngg;
  b->dim = mz == 1 ? 2 : 3;
  b->d2 = mx + 2 * ngg;
  b->d1 = (long)(my + 2 * ngg) * b->d2;
  b->sz = b->d1 * (mz + 2 * ngg);
  b->n = (long)mx * my * mz;
  /* Boundary::register_boundary (Mesh.cpp:14-32) with ngg = 0; bType flags (Mesh.cpp:857-916) */
  b->n_face = n_face;
  for (int f = 0; f < n_face; ++f) {
    const int *r = faces + 7 * f;
    face_t *F = &b->faces[f];
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
  b->cv = calloc((size_t)5 * b->sz, sizeof(double));
  b->qn = calloc((size_t)5 * b->sz, sizeof(double));
  b->bv = calloc((size_t)6 * b->sz, sizeof(double));
  b->fv = calloc((size_t)5 * b->sz, sizeof(double));
  b->gv = calloc((size_t)5 * b->sz, sizeof(double));
  b->hv = calloc((size_t)5 * b->sz, sizeof(double));
  b->mul = calloc((size_t)b->sz, sizeof(double));
  b->efd = calloc((size_t)b->sz, sizeof(double));
  b->dq = calloc((size_t)5 * b->n, sizeof(double));
  b->sensor = calloc((size_t)b->sz, sizeof(double));
  b->jac = malloc((size_t)b->sz * sizeof(double));
  b->met = malloc((size_t)9 * b->sz * sizeof(double));
  memcpy(b->jac, jac, (size_t)b->sz * sizeof(double));
  memcpy(b->met, metric, (size_t)9 * b->sz * sizeof(double));
  return b;
}

void orc_block_destroy(orc_block *b) {
  if (!b) return;
  for (int f = 0; f < 64; ++f) {
    free(b->inflow[f].prof);
    free(b->inflow[f].rms);
  }
  free(b->sponge_sigma);
  free(b->sponge_mean);
  free(b->cv); free(b->qn); free(b->bv); free(b->fv); free(b->gv); free(b->hv);
  free(b->mul); free(b->efd); free(b->dq); free(b->jac); free(b->met); free(b->sensor);
  free(b);
}

/* FieldOperation.cuh:16-64 (Air) */
static void cv_from_bv(orc_block *b, long id) {
  const long s = b->sz;
  const double rho = b->bv[id], u = b->bv[id + s], v = b->bv[id + 2 * s], w = b->bv[id + 3 * s];
  b->cv[id] = rho;
  b->cv[id + s] = rho * u;
  b->cv[id + 2 * s] = rho * v;
  b->cv[id + 3 * s] = rho * w;
  double e = 0.5 * (u * u + v * v + w * w);
  e *= rho;
  e += b->bv[id + 4 * s] / (kGamma - 1);
  b->cv[id + 4 * s] = e;
}

/* FieldOperation.cuh:86-120 (Air): mul over all nodes */
static void update_properties(orc_block *b) {
#pragma omp parallel for
  for (long id = 0; id < b->sz; ++id) b->mul[id] = orc_sutherland(b->bv[id + 5 * b->sz]);
}

/* BoundCond.cu:2065-2188 without back pressure */
static void apply_outflow(orc_block *b, const face_t *F) {
  int d[3] = {0, 0, 0};
  d[F->face] = F->dir;
  for (int k = F->s[2]; k <= F->e[2]; ++k)
    for (int j = F->s[1]; j <= F->e[1]; ++j)
      for (int i = F->s[0]; i <= F->e[0]; ++i) {
        const long id = ID(b, i, j, k);
        for (int g = 1; g <= b->g; ++g) {
          const long ig = ID(b, i + g * d[0], j + g * d[1], k + g * d[2]);
          for (int l = 0; l < 6; ++l) b->bv[ig + l * b->sz] = b->bv[id + l * b->sz];
          cv_from_bv(b, ig);
        }
      }
}

/* BoundCond.cu:3033-3083 */
static void apply_periodic(orc_block *b, const face_t *F) {
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
        cv_from_bv(b, id);
        for (int g = 1; g <= b->g; ++g) {
          const long ig = ID(b, i + g * d[0], j + g * d[1], k + g * d[2]);
          const long is = ID(b, o[0] + g * d[0], o[1] + g * d[1], o[2] + g * d[2]);
          for (int l = 0; l < 6; ++l) b->bv[ig + l * b->sz] = b->bv[is + l * b->sz];
          cv_from_bv(b, ig);
        }
      }
}

/* cv of the face nodes (from 0 when the face nodes are included) and of their ghost layers */
static void face_cv(orc_block *b, const face_t *F, int from) {
  int d[3] = {0, 0, 0};
  d[F->face] = F->dir;
  for (int k = F->s[2]; k <= F->e[2]; ++k)
    for (int j = F->s[1]; j <= F->e[1]; ++j)
      for (int i = F->s[0]; i <= F->e[0]; ++i)
        for (int g = from; g <= b->g; ++g) cv_from_bv(b, ID(b, i + g * d[0], j + g * d[1], k + g * d[2]));
}

/* DBoundCond::apply_boundary_conditions order (BoundCond.cu:3489-3619): outflow, inflow, symmetry, periodic */
static void apply_bcs(orc_block *b) {
  for (int f = 0; f < b->n_face; ++f)
    if (b->faces[f].type == 6) apply_outflow(b, &b->faces[f]);
  for (int f = 0; f < b->n_face; ++f) {
    const face_t *F = &b->faces[f];
    inflow_t *I = &b->inflow[f];
    if (F->type != 5 || !I->prof) continue;
    orc_bc_inflow(b->mx, b->my, b->mz, b->g, F->s, F->e, F->face, F->dir, b->bv, NULL, 0, NULL, b->met, I->prof,
                  I->rms, I->sub_node, I->sub_ghost, I->fluct, I->mw, I->seed, I->call);
    face_cv(b, F, 0);
    if (I->fluct) ++I->call;
  }
  for (int f = 0; f < b->n_face; ++f) {
    const face_t *F = &b->faces[f];
    if (F->type != 3) continue;
    orc_bc_symmetry(b->mx, b->my, b->mz, b->g, F->s, F->e, F->face, F->dir, b->bv, NULL, 0, b->met);
    face_cv(b, F, 1);
  }
  for (int f = 0; f < b->n_face; ++f)
    if (b->faces[f].type == 8) apply_periodic(b, &b->faces[f]);
}

void orc_block_set_inflow(orc_block *b, int face, const double *prof, const double *rms, int sub_node, int sub_ghost,
                          int fluct, double mw, unsigned long long seed, unsigned long long first_call) {
  const face_t *F = &b->faces[face];
  const long npt = (long)(F->e[0] - F->s[0] + 1) * (F->e[1] - F->s[1] + 1) * (F->e[2] - F->s[2] + 1);
  inflow_t *I = &b->inflow[face];
  free(I->prof);
  free(I->rms);
  I->prof = malloc(sizeof(double) * 6 * (size_t)npt);
  I->rms = malloc(sizeof(double) * (size_t)npt);
  memcpy(I->prof, prof, sizeof(double) * 6 * (size_t)npt);
  memcpy(I->rms, rms, sizeof(double) * (size_t)npt);
  I->sub_node = sub_node;
  I->sub_ghost = sub_ghost;
  I->fluct = fluct;
  I->mw = mw;
  I->seed = seed;
  I->call = first_call;
}

void orc_block_set_sponge(orc_block *b, const double *sigma) {
  free(b->sponge_sigma);
  free(b->sponge_mean);
  b->sponge_sigma = malloc(sizeof(double) * (size_t)b->n);
  memcpy(b->sponge_sigma, sigma, sizeof(double) * (size_t)b->n);
  b->sponge_mean = calloc((size_t)5 * b->n, sizeof(double));
  b->sponge_iter = 0;
}

/* The sponge layer's relaxation -J sigma (U - U_mean) added to the residual, and its running mean. */
static void sponge_source(orc_block *b) {
  if (!b->sponge_sigma) return;
  for (int k = 0; k < b->mz; ++k)
    for (int j = 0; j < b->my; ++j)
      for (int i = 0; i < b->mx; ++i) {
        const long id = ID(b, i, j, k), iq = IDQ(b, i, j, k);
        const double sg = b->sponge_sigma[iq] * b->jac[id];
        for (int l = 0; l < 5; ++l) b->dq[iq + l * b->n] -= sg * (b->cv[id + l * b->sz] - b->sponge_mean[iq + l * b->n]);
      }
}

static void sponge_mean(orc_block *b) {
  if (!b->sponge_sigma) return;
  const double n = (double)b->sponge_iter, inv = 1.0 / (n + 1);
  for (int k = 0; k < b->mz; ++k)
    for (int j = 0; j < b->my; ++j)
      for (int i = 0; i < b->mx; ++i) {
        const long id = ID(b, i, j, k), iq = IDQ(b, i, j, k);
        for (int l = 0; l < 5; ++l)
          b->sponge_mean[iq + l * b->n] = (b->sponge_mean[iq + l * b->n] * n + b->cv[id + l * b->sz]) * inv;
      }
  ++b->sponge_iter;
}

/* periodic_exchange task 1 for the viscous fluxes (DataCommunication.cu:628-683) */
static void exchange_viscous_flux(orc_block *b) {
  const int ext[3] = {b->mx, b->my, b->mz};
  for (int f = 0; f < b->n_face; ++f) {
    const face_t *F = &b->faces[f];
    if (F->type != 8) continue;
    int d[3] = {0, 0, 0};
    d[F->face] = F->dir;
    for (int k = F->s[2]; k <= F->e[2]; ++k)
      for (int j = F->s[1]; j <= F->e[1]; ++j)
        for (int i = F->s[0]; i <= F->e[0]; ++i) {
          int o[3] = {i, j, k};
          o[F->face] = F->dir < 0 ? ext[F->face] - 1 : 0;
          const long id = ID(b, i, j, k), io = ID(b, o[0], o[1], o[2]);
          double *A[3] = {b->fv, b->gv, b->hv};
          for (int l = 0; l < 5; ++l)
            for (int a = 0; a < 3; ++a) {
              const double av = 0.5 * (A[a][id + l * b->sz] + A[a][io + l * b->sz]);
              A[a][id + l * b->sz] = av;
              A[a][io + l * b->sz] = av;
            }
          for (int g = 1; g <= b->g; ++g) {
            const long ig = ID(b, i + g * d[0], j + g * d[1], k + g * d[2]);
            const long is = ID(b, o[0] + g * d[0], o[1] + g * d[1], o[2] + g * d[2]);
            for (int l = 0; l < 5; ++l)
              for (int a = 0; a < 3; ++a) A[a][ig + l * b->sz] = A[a][is + l * b->sz];
          }
        }
  }
}

/* d/dxi of f along a direction with stride st at index i of extent n (ViscousScheme.cu:129-177 for
 * ORDER 8; ORDER 4 uses the reference's 4th-order closures as its interior and near-boundary rows). */
double orc_deriv(const double *f, long id, long st, int i, int n, int left, int right, int order) {
#define F(o) f[id + (long)(o) * st]
  if (order == 8) {
    if (left) {
      if (i == 0) return -25.0 / 12 * F(0) + 4.0 * F(1) - 3 * F(2) + 4.0 / 3 * F(3) - 0.25 * F(4);
      if (i == 1) return -0.25 * F(-1) - 5.0 / 6 * F(0) + 1.5 * F(1) - 0.5 * F(2) + 1.0 / 12 * F(3);
      if (i == 2) return -1.0 / 12 * (F(2) - F(-2)) + 2.0 / 3 * (F(1) - F(-1));
      if (i == 3) return 0.75 * (F(1) - F(-1)) - 0.15 * (F(2) - F(-2)) + 1.0 / 60 * (F(3) - F(-3));
    }
    if (right) {
      if (i == n - 1) return 25.0 / 12 * F(0) - 4.0 * F(-1) + 3 * F(-2) - 4.0 / 3 * F(-3) + 0.25 * F(-4);
      if (i == n - 2) return 0.25 * F(1) + 5.0 / 6 * F(0) - 1.5 * F(-1) + 0.5 * F(-2) - 1.0 / 12 * F(-3);
      if (i == n - 3) return -1.0 / 12 * (F(2) - F(-2)) + 2.0 / 3 * (F(1) - F(-1));
      if (i == n - 4) return 0.75 * (F(1) - F(-1)) - 0.15 * (F(2) - F(-2)) + 1.0 / 60 * (F(3) - F(-3));
    }
    return 0.8 * (F(1) - F(-1)) + -0.2 * (F(2) - F(-2)) + 4.0 / 105 * (F(3) - F(-3)) + -1.0 / 280 * (F(4) - F(-4));
  }
  if (left) {
    if (i == 0) return -25.0 / 12 * F(0) + 4.0 * F(1) - 3 * F(2) + 4.0 / 3 * F(3) - 0.25 * F(4);
    if (i == 1) return -0.25 * F(-1) - 5.0 / 6 * F(0) + 1.5 * F(1) - 0.5 * F(2) + 1.0 / 12 * F(3);
  }
  if (right) {
    if (i == n - 1) return 25.0 / 12 * F(0) - 4.0 * F(-1) + 3 * F(-2) - 4.0 / 3 * F(-3) + 0.25 * F(-4);
    if (i == n - 2) return 0.25 * F(1) + 5.0 / 6 * F(0) - 1.5 * F(-1) + 0.5 * F(-2) - 1.0 / 12 * F(-3);
  }
  return -1.0 / 12 * (F(2) - F(-2)) + 2.0 / 3 * (F(1) - F(-1));
#undef F
}

/* compute_viscous_flux_collocated<Air,ORDER> (ViscousScheme.cu:309-408): stress and heat-flux node fluxes */
static void viscous_nodes(orc_block *b) {
  const int order = b->p.viscous_order;
  const long s = b->sz;
  const double cp = kGamma * (kRu / kMwAir) / (kGamma - 1);
#pragma omp parallel for collapse(2)
  for (int k = 0; k < b->mz; ++k)
    for (int j = 0; j < b->my; ++j)
      for (int i = 0; i < b->mx; ++i) {
        const long id = ID(b, i, j, k);
        double dv[4][3];
        const int comp[4] = {1, 2, 3, 5};
        for (int q = 0; q < 4; ++q) {
          const double *f = b->bv + comp[q] * s;
          dv[q][0] = orc_deriv(f, id, 1, i, b->mx, b->bt[0], b->bt[1], order);
          dv[q][1] = orc_deriv(f, id, b->d2, j, b->my, b->bt[2], b->bt[3], order);
          dv[q][2] = b->dim == 3 ? orc_deriv(f, id, b->d1, k, b->mz, b->bt[4], b->bt[5], order) : 0.0;
        }
        const double *M = b->met;
        const double xx = M[id], xy = M[id + s], xz = M[id + 2 * s];
        const double ex = M[id + 3 * s], ey = M[id + 4 * s], ez = M[id + 5 * s];
        const double zx = M[id + 6 * s], zy = M[id + 7 * s], zz = M[id + 8 * s];
        const double ux = dv[0][0] * xx + dv[0][1] * ex + dv[0][2] * zx;
        const double uy = dv[0][0] * xy + dv[0][1] * ey + dv[0][2] * zy;
        const double uz = dv[0][0] * xz + dv[0][1] * ez + dv[0][2] * zz;
        const double vx = dv[1][0] * xx + dv[1][1] * ex + dv[1][2] * zx;
        const double vy = dv[1][0] * xy + dv[1][1] * ey + dv[1][2] * zy;
        const double vz = dv[1][0] * xz + dv[1][1] * ez + dv[1][2] * zz;
        const double wx = dv[2][0] * xx + dv[2][1] * ex + dv[2][2] * zx;
        const double wy = dv[2][0] * xy + dv[2][1] * ey + dv[2][2] * zy;
        const double wz = dv[2][0] * xz + dv[2][1] * ez + dv[2][2] * zz;
        const double mu = b->mul[id];
        const double txx = mu * (4 * ux - 2 * vy - 2 * wz) / 3.0;
        const double tyy = mu * (4 * vy - 2 * ux - 2 * wz) / 3.0;
        const double tzz = mu * (4 * wz - 2 * ux - 2 * vy) / 3.0;
        const double txy = mu * (uy + vx), txz = mu * (uz + wx), tyz = mu * (vz + wy);
        b->fv[id + s] = xx * txx + xy * txy + xz * txz;
        b->fv[id + 2 * s] = xx * txy + xy * tyy + xz * tyz;
        b->fv[id + 3 * s] = xx * txz + xy * tyz + xz * tzz;
        b->gv[id + s] = ex * txx + ey * txy + ez * txz;
        b->gv[id + 2 * s] = ex * txy + ey * tyy + ez * tyz;
        b->gv[id + 3 * s] = ex * txz + ey * tyz + ez * tzz;
        b->hv[id + s] = zx * txx + zy * txy + zz * txz;
        b->hv[id + 2 * s] = zx * txy + zy * tyy + zz * tyz;
        b->hv[id + 3 * s] = zx * txz + zy * tyz + zz * tzz;
        const double tx = dv[3][0] * xx + dv[3][1] * ex + dv[3][2] * zx;
        const double ty = dv[3][0] * xy + dv[3][1] * ey + dv[3][2] * zy;
        const double tz = dv[3][0] * xz + dv[3][1] * ez + dv[3][2] * zz;
        const double kc = mu / b->p.prandtl * cp;
        const double u = b->bv[id + s], v = b->bv[id + 2 * s], w = b->bv[id + 3 * s];
        const double Ex = u * txx + v * txy + w * txz + kc * tx;
        const double Ey = u * txy + v * tyy + w * tyz + kc * ty;
        const double Ez = u * txz + v * tyz + w * tzz + kc * tz;
        b->fv[id + 4 * s] = xx * Ex + xy * Ey + xz * Ez;
        b->gv[id + 4 * s] = ex * Ex + ey * Ey + ez * Ez;
        b->hv[id + 4 * s] = zx * Ex + zy * Ey + zz * Ez;
      }
}

/* compute_viscous_flux_derivative<ORDER> (ViscousScheme.cu:568-612): d(fv)/dxi + d(gv)/deta + d(hv)/dzeta of
 * the node fluxes, ghosts included, into dq */
static void viscous_diff(orc_block *b) {
  const int order = b->p.viscous_order;
  const long s = b->sz;
#pragma omp parallel for collapse(2)
  for (int k = 0; k < b->mz; ++k)
    for (int j = 0; j < b->my; ++j)
      for (int i = 0; i < b->mx; ++i) {
        const long id = ID(b, i, j, k), iq = IDQ(b, i, j, k);
        const double J = b->jac[id];
        for (int l = 1; l < 5; ++l) {
          double d = orc_deriv(b->fv + l * s, id, 1, i, b->mx, b->bt[0], b->bt[1], order) +
                     orc_deriv(b->gv + l * s, id, b->d2, j, b->my, b->bt[2], b->bt[3], order);
          d += b->dim == 3 ? orc_deriv(b->hv + l * s, id, b->d1, k, b->mz, b->bt[4], b->bt[5], order) : 0.0;
          b->dq[iq + l * b->n] += d * J;
        }
      }
}

/* compute_viscous_flux with collocated node fluxes (ViscousScheme.cu:7-73) on one block: node fluxes, the
 * periodic exchange of their ghosts (exchange_value task 1, DataCommunication.cu:408-456), differences */
static void viscous(orc_block *b) {
  viscous_nodes(b);
  exchange_viscous_flux(b);
  viscous_diff(b);
}

/* Second-order face-centred viscous fluxes, compute_fv/gv/hv_2nd_order<Air> (ViscousScheme.cu:614-1010):
 * on the face between nodes t and t+1 along dir, metrics and metric*Jacobian are the two nodes' averages,
 * the derivative along dir is the difference of the two nodes, the transverse ones the average of the two
 * nodes' central differences; the viscosity is the nodes' average.  Stored at node t (face t+1/2), for
 * t = -1 .. m-1, then differenced into dq (compute_dFv_dx etc., ViscousScheme.cu:75-128).  2D blocks have
 * no transverse z differences. */
static void viscous_faces(orc_block *b, int dir, double *F) {
  const long s = b->sz;
  const double cp = kGamma * (kRu / kMwAir) / (kGamma - 1);
  const long st[3] = {1, b->d2, b->d1};
  const int ext[3] = {b->mx, b->my, b->mz};
  int lo[3] = {0, 0, 0};
  lo[dir] = -1;
#pragma omp parallel for collapse(2)
  for (int k = lo[2]; k < ext[2]; ++k)
    for (int j = lo[1]; j < ext[1]; ++j)
      for (int i = lo[0]; i < ext[0]; ++i) {
        const long id = ID(b, i, j, k), ie = id + st[dir];
        const double *M = b->met;
        double m[9];
        for (int c = 0; c < 9; ++c) m[c] = 0.5 * (M[id + c * s] + M[ie + c * s]);
        const int comp[4] = {1, 2, 3, 5};
        double g[4][3];
        for (int q = 0; q < 4; ++q) {
          const double *f = b->bv + comp[q] * s;
          for (int a = 0; a < 3; ++a) {
            if (a == dir) g[q][a] = f[ie] - f[id];
            else if (a == 2 && b->dim == 2) g[q][a] = 0;
            else g[q][a] = 0.25 * (f[id + st[a]] - f[id - st[a]] + f[ie + st[a]] - f[ie - st[a]]);
          }
        }
        double d[4][3]; /* d/dx, d/dy, d/dz of u, v, w, T */
        for (int q = 0; q < 4; ++q)
          for (int c = 0; c < 3; ++c) d[q][c] = g[q][0] * m[c] + g[q][1] * m[3 + c] + g[q][2] * m[6 + c];
        const double mu = 0.5 * (b->mul[id] + b->mul[ie]);
        const double txx = mu * (4 * d[0][0] - 2 * d[1][1] - 2 * d[2][2]) / 3.0;
        const double tyy = mu * (4 * d[1][1] - 2 * d[0][0] - 2 * d[2][2]) / 3.0;
        const double tzz = mu * (4 * d[2][2] - 2 * d[0][0] - 2 * d[1][1]) / 3.0;
        const double txy = mu * (d[0][1] + d[1][0]), txz = mu * (d[0][2] + d[2][0]), tyz = mu * (d[1][2] + d[2][1]);
        double kj[3];
        for (int c = 0; c < 3; ++c)
          kj[c] = 0.5 * (M[id + (3 * dir + c) * s] * b->jac[id] + M[ie + (3 * dir + c) * s] * b->jac[ie]);
        const double f1 = kj[0] * txx + kj[1] * txy + kj[2] * txz;
        const double f2 = kj[0] * txy + kj[1] * tyy + kj[2] * tyz;
        const double f3 = kj[0] * txz + kj[1] * tyz + kj[2] * tzz;
        const double um = 0.5 * (b->bv[id + s] + b->bv[ie + s]);
        const double vm = 0.5 * (b->bv[id + 2 * s] + b->bv[ie + 2 * s]);
        const double wm = 0.5 * (b->bv[id + 3 * s] + b->bv[ie + 3 * s]);
        const double kc = mu / b->p.prandtl * cp;
        F[id + s] = f1;
        F[id + 2 * s] = f2;
        F[id + 3 * s] = f3;
        F[id + 4 * s] = um * f1 + vm * f2 + wm * f3 + kc * (kj[0] * d[3][0] + kj[1] * d[3][1] + kj[2] * d[3][2]);
      }
}

static void viscous2(orc_block *b) {
  double *F[3] = {b->fv, b->gv, b->hv};
  const long st[3] = {1, b->d2, b->d1};
  for (int dir = 0; dir < b->dim; ++dir) {
    viscous_faces(b, dir, F[dir]);
#pragma omp parallel for collapse(2)
    for (int k = 0; k < b->mz; ++k)
      for (int j = 0; j < b->my; ++j)
        for (int i = 0; i < b->mx; ++i) {
          const long id = ID(b, i, j, k), iq = IDQ(b, i, j, k);
          for (int l = 1; l < 5; ++l) b->dq[iq + l * b->n] += F[dir][id + l * b->sz] - F[dir][id - st[dir] + l * b->sz];
        }
  }
}

/* compute_entropy_fix_delta<Air> (InviscidScheme.cu:269-304), over the interior plus one layer */
static void entropy_fix(orc_block *b) {
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
        const double c = sqrt(kGamma * bv[id + 4 * s] / bv[id]);
        if (b->dim == 2)
          b->efd[id] = b->p.entropy_fix_factor * (U + V + c * 0.5 * (kx + ky));
        else
          b->efd[id] = b->p.entropy_fix_factor * (U + V + W + c * (kx + ky + kz) / 3.0);
      }
}

/* Inviscid flux of one direction and its difference into dq (the role of the WENO/pv kernels plus
 * compute_derivative_x/y/z, WENO.cu:1544-1571). */
static void inviscid_dir(orc_block *b, int dir) {
  const int fs = b->p.inviscid_scheme >= 51; /* LLF flux-splitting WENO (WENO.cu:830-990) */
  const int pvr = !fs && b->p.reconstruction <= 3; /* primitive reconstruction (InviscidScheme.cu:61-219) */
  const int r = fs ? (b->p.inviscid_scheme >= 71 ? 4 : 3) : (pvr ? 2 : (b->p.reconstruction == 5 ? 4 : 3));
  const int n = 2 * r;
  const long s = b->sz;
  const long st = dir == 0 ? 1 : (dir == 1 ? b->d2 : b->d1);
  const int ext[3] = {b->mx, b->my, b->mz};
  const int m = ext[dir];
  /* lines orthogonal to dir */
  const int oa = dir == 0 ? 1 : 0, ob = dir == 2 ? 1 : 2;
#pragma omp parallel for collapse(2)
  for (int B = 0; B < ext[ob]; ++B)
    for (int A = 0; A < ext[oa]; ++A) {
      double *Fl = malloc(sizeof(double) * 5 * (m + 1));
      for (int t = -1; t < m; ++t) { /* interface t+1/2 */
        int idx[3];
        idx[dir] = t;
        idx[oa] = A;
        idx[ob] = B;
        const long id = ID(b, idx[0], idx[1], idx[2]);
        double q[8][5];
        for (int mm = 0; mm < n; ++mm)
          for (int l = 0; l < 5; ++l) q[mm][l] = b->cv[id + (mm - r + 1) * st + l * s];
        if (fs) {
          double met[8][3], jac[8];
          for (int mm = 0; mm < n; ++mm) {
            const long im = id + (mm - r + 1) * st;
            for (int c = 0; c < 3; ++c) met[mm][c] = b->met[im + (3 * dir + c) * s];
            jac[mm] = b->jac[im];
          }
          /* the hybrid scheme's switch (WENO.cu:838-849): nodes t-ngg+1 .. t+ngg of the sensor */
          int shock = 1;
          if (b->p.sensor_threshold > 1e-10) {
            shock = 0;
            for (int ii = -b->g + 1; ii <= b->g; ++ii)
              if (b->sensor[id + ii * st] > b->p.sensor_threshold) shock = 1;
          }
          orc_fsweno_flux_hybrid(&b->p, (const double(*)[5])q, (const double(*)[3])met, jac, n, shock,
                                 Fl + 5 * (t + 1));
          continue;
        }
        double ml[3], mr[3];
        for (int c = 0; c < 3; ++c) {
          ml[c] = b->met[id + (3 * dir + c) * s];
          mr[c] = b->met[id + st + (3 * dir + c) * s];
        }
        if (pvr) {
          double pv[4][5];
          for (int mm = 0; mm < 4; ++mm)
            for (int l = 0; l < 5; ++l) pv[mm][l] = b->bv[id + (mm - 1) * st + l * s];
          orc_interface_flux_pv(&b->p, (const double(*)[5])pv, ml, mr, b->jac[id], b->jac[id + st], b->efd[id],
                                b->efd[id + st], Fl + 5 * (t + 1));
          continue;
        }
        orc_interface_flux(&b->p, (const double(*)[5])q, n, ml, mr, b->jac[id], b->jac[id + st], b->efd[id],
                           b->efd[id + st], Fl + 5 * (t + 1));
      }
      for (int t = 0; t < m; ++t) {
        int idx[3];
        idx[dir] = t;
        idx[oa] = A;
        idx[ob] = B;
        const long iq = IDQ(b, idx[0], idx[1], idx[2]);
        for (int l = 0; l < 5; ++l) b->dq[iq + l * b->n] -= Fl[5 * (t + 1) + l] - Fl[5 * t + l];
      }
      free(Fl);
    }
}

/* compute_shock_sensor at the start of a step (RK.cuh:112-117), when the hybrid scheme is on */
static void shock_sensor(orc_block *b) {
  if (b->p.inviscid_scheme >= 51 && b->p.sensor_threshold > 1e-10)
    orc_shock_sensor(b->p.shock_sensor, b->p.sensor_eps, b->mx, b->my, b->mz, b->g, b->bv, b->met, b->sensor);
}

static void rhs_core(orc_block *b);

void orc_block_rhs(orc_block *b) {
  shock_sensor(b);
  rhs_core(b);
}

static void rhs_core(orc_block *b) {
  memset(b->dq, 0, sizeof(double) * 5 * b->n);
  if (b->p.viscous_order == 2) viscous2(b);
  else if (b->p.viscous_order > 0) viscous(b);
  if (b->p.inviscid_scheme == 2) entropy_fix(b);
  inviscid_dir(b, 0);
  inviscid_dir(b, 1);
  if (b->dim == 3) inviscid_dir(b, 2);
  sponge_source(b);
}

static const double rk_a[3] = {1.0, 0.75, 1.0 / 3.0};
static const double rk_b[3] = {0.0, 0.25, 2.0 / 3.0};
static const double rk_c[3] = {1.0, 0.25, 2.0 / 3.0};

/* update_cv_and_bv_rk<Air> (RK.cuh:224-280) */
static void update(orc_block *b, double dt, int rk) {
  const long s = b->sz;
#pragma omp parallel for collapse(2)
  for (int k = 0; k < b->mz; ++k)
    for (int j = 0; j < b->my; ++j)
      for (int i = 0; i < b->mx; ++i) {
        const long id = ID(b, i, j, k), iq = IDQ(b, i, j, k);
        const double dj = dt / b->jac[id];
        for (int l = 0; l < 5; ++l)
          b->cv[id + l * s] = rk_a[rk] * b->qn[id + l * s] + rk_b[rk] * b->cv[id + l * s] +
                              rk_c[rk] * dj * b->dq[iq + l * b->n];
        if (b->mz == 1) b->cv[id + 3 * s] = 0;
        double *bv = b->bv, *cv = b->cv;
        bv[id] = cv[id];
        const double di = 1.0 / cv[id];
        bv[id + s] = cv[id + s] * di;
        bv[id + 2 * s] = cv[id + 2 * s] * di;
        bv[id + 3 * s] = cv[id + 3 * s] * di;
        const double V2 = bv[id + s] * bv[id + s] + bv[id + 2 * s] * bv[id + 2 * s] + bv[id + 3 * s] * bv[id + 3 * s];
        bv[id + 4 * s] = (kGamma - 1) * (cv[id + 4 * s] - 0.5 * bv[id] * V2);
        bv[id + 5 * s] = bv[id + 4 * s] * kMwAir * di / kRu;
      }
}

void orc_block_set_bv(orc_block *b, const double *bv) {
  memcpy(b->bv, bv, sizeof(double) * 6 * b->sz);
#pragma omp parallel for
  for (long id = 0; id < b->sz; ++id) cv_from_bv(b, id);
  update_properties(b);
  apply_bcs(b);
  update_properties(b);
}

void orc_block_set_raw(orc_block *b, const double *bv, const double *cv) {
  memcpy(b->bv, bv, sizeof(double) * 6 * b->sz);
  memcpy(b->cv, cv, sizeof(double) * 5 * b->sz);
  update_properties(b);
}

void orc_block_get(const orc_block *b, int which, double *out) {
  switch (which) {
    case 0: memcpy(out, b->cv, sizeof(double) * 5 * b->sz); break;
    case 1: memcpy(out, b->bv, sizeof(double) * 6 * b->sz); break;
    case 2: memcpy(out, b->dq, sizeof(double) * 5 * b->n); break;
    case 3: memcpy(out, b->mul, sizeof(double) * b->sz); break;
    case 4: memcpy(out, b->jac, sizeof(double) * b->sz); break;
    case 5: memcpy(out, b->met, sizeof(double) * 9 * b->sz); break;
    case 6: memcpy(out, b->fv, sizeof(double) * 5 * b->sz); break;
    case 7: memcpy(out, b->gv, sizeof(double) * 5 * b->sz); break;
    case 8: memcpy(out, b->hv, sizeof(double) * 5 * b->sz); break;
    case 9: memcpy(out, b->sensor, sizeof(double) * b->sz); break;
    default: break;
  }
}

/* local_time_step<Air> (TimeAdvanceFunc.cuh:28-115) and the global minimum (TimeAdvanceFunc.cu:102-152) */
double orc_block_cfl_dt(orc_block *b) {
  const long s = b->sz;
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
        const double c = sqrt(kGamma * (kRu / kMwAir) * bv[id + 5 * s]);
        double sr = fmax(fabs(U) + c * gx, fabs(V) + c * ge);
        if (b->dim == 3) sr = fmax(sr, fabs(W) + c * gz);
        if (b->p.viscous_order > 0) {
          double ml = fmax(gx, ge);
          if (b->dim == 3) ml = fmax(ml, gz);
          double dv = fmax(kGamma, 4.0 / 3.0) / bv[id];
          dv *= b->mul[id] / b->p.prandtl;
          sr = fmax(sr, ml * ml * dv);
        }
        const double dt = b->p.cfl / sr;
        if (dt < dtmin) dtmin = dt;
      }
  return dtmin;
}

/* One SSP-RK3 step (RK.cuh:100-192) for a single block with only outflow/periodic faces. */
void orc_block_rk3_step(orc_block *b, double dt) {
  sponge_mean(b);
  shock_sensor(b);
  memcpy(b->qn, b->cv, sizeof(double) * 5 * b->sz);
  for (int rk = 0; rk < 3; ++rk) {
    rhs_core(b);
    update(b, dt, rk);
    apply_bcs(b);
    update_properties(b);
  }
}

/* ---- Several blocks joined by block-to-block faces ----
 * data_communication / inner_communication (DataCommunication.cu:6-103) and exchange_value task 1
 * (DataCommunication.cu:408-495) restated with a pairing of nodes found from the coordinates alone: every node of
 * block a's face layer is matched with the node of block t at the same place, the ghost layers beyond it with t's
 * nodes walked away from a, and each index direction of a with the direction of t along the same grid line.  The
 * reference derives the same pairs from the inner file's signed index ranges (InnerFace::register_boundary,
 * Mesh.cpp:38-170); matching by position checks that index algebra instead of repeating it.  A link covers the
 * whole side of block a (no corner ghosts, as the reference leaves them out). */
struct orc_link {
  int a, t;
  long n;
  long *ia, *it; /* padded node index in a, in t */
  char *layer;   /* 1 on the coincident layer */
  int vmap[3], vsign[3];
};

static double dist3(const double *x, const double *y, const double *z, long p, double px, double py, double pz) {
  const double dx = x[p] - px, dy = y[p] - py, dz = z[p] - pz;
  return sqrt(dx * dx + dy * dy + dz * dz);
}

orc_link *orc_link_find(orc_block **B, int a, const double *xa, const double *ya, const double *za, int dir, int side,
                        int t, const double *xt, const double *yt, const double *zt) {
  const orc_block *A = B[a], *T = B[t];
  const int ea[3] = {A->mx, A->my, A->mz}, et[3] = {T->mx, T->my, T->mz};
#define PA(i, j, k) ((long)(i) + (long)A->mx * ((j) + (long)A->my * (k)))
#define PT(i, j, k) ((long)(i) + (long)T->mx * ((j) + (long)T->my * (k)))
  const int u = dir == 0 ? 1 : 0, v = dir == 2 ? 1 : 2; /* the two in-face directions of a */
  const int layer = side > 0 ? ea[dir] - 1 : 0;
  const long nface = (long)ea[u] * ea[v];
  orc_link *L = calloc(1, sizeof(orc_link));
  L->a = a;
  L->t = t;
  L->ia = malloc(sizeof(long) * nface * (A->g + 1));
  L->it = malloc(sizeof(long) * nface * (A->g + 1));
  L->layer = malloc((size_t)nface * (A->g + 1));
  for (long f = 0; f < nface; ++f) {
    int n[3];
    n[dir] = layer;
    n[u] = (int)(f % ea[u]);
    n[v] = (int)(f / ea[u]);
    const long p = PA(n[0], n[1], n[2]);
    int m[3] = {n[0], n[1], n[2]};
    m[dir] -= side; /* a's interior neighbour: the spacing across the face */
    const long pin = PA(m[0], m[1], m[2]);
    const double h = dist3(xa, ya, za, pin, xa[p], ya[p], za[p]);
    long best = -1;
    double bd = 1e300;
    for (long q = 0; q < (long)T->mx * T->my * T->mz; ++q) {
      const double d = dist3(xt, yt, zt, q, xa[p], ya[p], za[p]);
      if (d < bd) { bd = d; best = q; }
    }
    if (bd > 1e-6 * h) { orc_link_destroy(L); return NULL; }
    const int Q[3] = {(int)(best % T->mx), (int)((best / T->mx) % T->my), (int)(best / ((long)T->mx * T->my))};
    /* the direction of t that continues a's grid line across the face: its first node lies nearest the
     * reflection of a's interior neighbour through the face node */
    const double rx = 2 * xa[p] - xa[pin], ry = 2 * ya[p] - ya[pin], rz = 2 * za[p] - za[pin];
    int oc = -1, os = 0;
    double od = 1e300;
    for (int c = 0; c < 3; ++c)
      for (int sg = -1; sg <= 1; sg += 2) {
        int r[3] = {Q[0], Q[1], Q[2]};
        r[c] += sg;
        if (r[c] < 0 || r[c] >= et[c]) continue;
        const double d = dist3(xt, yt, zt, PT(r[0], r[1], r[2]), rx, ry, rz);
        if (d < od) { od = d; oc = c; os = sg; }
      }
    if (oc < 0 || od > 0.5 * h) { orc_link_destroy(L); return NULL; }
    for (int g = 0; g <= A->g; ++g) {
      int r[3] = {Q[0], Q[1], Q[2]}, w[3] = {n[0], n[1], n[2]};
      r[oc] += g * os;
      w[dir] += g * side;
      if (r[oc] < 0 || r[oc] >= et[oc]) { orc_link_destroy(L); return NULL; }
      L->ia[L->n] = ID(A, w[0], w[1], w[2]);
      L->it[L->n] = ID(T, r[0], r[1], r[2]);
      L->layer[L->n] = g == 0;
      ++L->n;
    }
    if (f == nface / 2) { /* the direction map, from a node in the middle of the face */
      L->vmap[dir] = oc;
      L->vsign[dir] = os * side;
      for (int d = 0; d < 3; ++d) {
        if (d == dir) continue;
        if (A->dim == 2 && d == 2) { L->vmap[2] = 2; L->vsign[2] = 1; continue; }
        int e[3] = {n[0], n[1], n[2]};
        const int up = n[d] + 1 < ea[d];
        e[d] += up ? 1 : -1;
        const long pe = PA(e[0], e[1], e[2]);
        const double sd = up ? 1.0 : -1.0;
        const double Dx = sd * (xa[pe] - xa[p]), Dy = sd * (ya[pe] - ya[p]), Dz = sd * (za[pe] - za[p]);
        double bc = -2;
        for (int c = 0; c < 3; ++c)
          for (int sg = -1; sg <= 1; sg += 2) {
            int r[3] = {Q[0], Q[1], Q[2]};
            r[c] += sg;
            if (c == oc || r[c] < 0 || r[c] >= et[c]) continue;
            const long pr = PT(r[0], r[1], r[2]);
            const double Ex = xt[pr] - xt[best], Ey = yt[pr] - yt[best], Ez = zt[pr] - zt[best];
            const double cs = (Dx * Ex + Dy * Ey + Dz * Ez) /
                              (sqrt(Dx * Dx + Dy * Dy + Dz * Dz) * sqrt(Ex * Ex + Ey * Ey + Ez * Ez));
            if (cs > bc) { bc = cs; L->vmap[d] = c; L->vsign[d] = sg; }
          }
      }
    }
  }
#undef PA
#undef PT
  return L;
}

void orc_link_destroy(orc_link *L) {
  if (!L) return;
  free(L->ia);
  free(L->it);
  free(L->layer);
  free(L);
}

long orc_link_info(const orc_link *L, int *vmap, int *vsign, long *ia, long *it, char *layer) {
  for (int d = 0; d < 3; ++d) {
    vmap[d] = L->vmap[d];
    vsign[d] = L->vsign[d];
  }
  if (ia) memcpy(ia, L->ia, sizeof(long) * L->n);
  if (it) memcpy(it, L->it, sizeof(long) * L->n);
  if (layer) memcpy(layer, L->layer, (size_t)L->n);
  return L->n;
}

/* inner_communication<Air> (DataCommunication.cu:36-103): the coincident layer takes the two blocks' average (the
 * partner's state where a's velocity is nonzero and the partner's is zero; nothing where a's is zero), the ghost
 * layers the partner's state; cv follows bv on a's nodes. */
static void link_state(orc_block **B, const orc_link *L) {
  orc_block *A = B[L->a], *T = B[L->t];
  const long sa = A->sz, st = T->sz;
  for (long q = 0; q < L->n; ++q) {
    const long i = L->ia[q], j = L->it[q];
    if (L->layer[q]) {
      const double ua = sqrt(A->bv[i + sa] * A->bv[i + sa] + A->bv[i + 2 * sa] * A->bv[i + 2 * sa] +
                             A->bv[i + 3 * sa] * A->bv[i + 3 * sa]);
      if (ua > 1e-15) {
        const double ut = sqrt(T->bv[j + st] * T->bv[j + st] + T->bv[j + 2 * st] * T->bv[j + 2 * st] +
                               T->bv[j + 3 * st] * T->bv[j + 3 * st]);
        for (int l = 0; l < 6; ++l) {
          if (ut < 1e-15) {
            A->bv[i + l * sa] = T->bv[j + l * st];
          } else {
            const double m = 0.5 * (T->bv[j + l * st] + A->bv[i + l * sa]);
            A->bv[i + l * sa] = m;
            T->bv[j + l * st] = m;
          }
        }
      }
    } else {
      for (int l = 0; l < 6; ++l) A->bv[i + l * sa] = T->bv[j + l * st];
    }
    cv_from_bv(A, i);
  }
}

/* exchange_value task 1 (DataCommunication.cu:458-495) for the collocated viscous fluxes, each of a's index
 * directions taking the partner's flux along the matching direction with its sign (the same arrays when the two
 * blocks' directions agree, as in the reference). */
static void link_viscous(orc_block **B, const orc_link *L) {
  orc_block *A = B[L->a], *T = B[L->t];
  double *FA[3] = {A->fv, A->gv, A->hv}, *FT[3] = {T->fv, T->gv, T->hv};
  for (long q = 0; q < L->n; ++q) {
    const long i = L->ia[q], j = L->it[q];
    for (int d = 0; d < 3; ++d) {
      double *P = FT[L->vmap[d]];
      const double sg = L->vsign[d];
      for (int l = 1; l < 5; ++l) {
        if (L->layer[q]) {
          const double m = 0.5 * (FA[d][i + l * A->sz] + sg * P[j + l * T->sz]);
          FA[d][i + l * A->sz] = m;
          P[j + l * T->sz] = sg * m;
        } else {
          FA[d][i + l * A->sz] = sg * P[j + l * T->sz];
        }
      }
    }
  }
}

/* data_communication over the links in order, then the physical properties of every block (RK.cuh:174-190) */
void orc_multi_exchange(orc_block **B, int nb, orc_link **L, int nl) {
  for (int l = 0; l < nl; ++l) link_state(B, L[l]);
  for (int q = 0; q < nb; ++q) update_properties(B[q]);
}

/* The residual of every block: viscous node fluxes of all blocks, their exchange across the links and then the
 * periodic faces (exchange_value, DataCommunication.cu:408-456), differences; inviscid fluxes (RK.cuh:137-157). */
static void multi_rhs_core(orc_block **B, int nb, orc_link **L, int nl);

void orc_multi_rhs(orc_block **B, int nb, orc_link **L, int nl) {
  for (int q = 0; q < nb; ++q) shock_sensor(B[q]);
  multi_rhs_core(B, nb, L, nl);
}

static void multi_rhs_core(orc_block **B, int nb, orc_link **L, int nl) {
  int collocated = 0;
  for (int q = 0; q < nb; ++q) {
    orc_block *b = B[q];
    memset(b->dq, 0, sizeof(double) * 5 * b->n);
    if (b->p.viscous_order == 2) viscous2(b);
    else if (b->p.viscous_order > 0) {
      viscous_nodes(b);
      collocated = 1;
    }
  }
  if (collocated) {
    for (int l = 0; l < nl; ++l) link_viscous(B, L[l]);
    for (int q = 0; q < nb; ++q) exchange_viscous_flux(B[q]);
    for (int q = 0; q < nb; ++q) viscous_diff(B[q]);
  }
  for (int q = 0; q < nb; ++q) {
    orc_block *b = B[q];
    if (b->p.inviscid_scheme == 2) entropy_fix(b);
    inviscid_dir(b, 0);
    inviscid_dir(b, 1);
    if (b->dim == 3) inviscid_dir(b, 2);
    sponge_source(b);
  }
}

double orc_multi_cfl_dt(orc_block **B, int nb) {
  double dt = 1e+6;
  for (int q = 0; q < nb; ++q) dt = fmin(dt, orc_block_cfl_dt(B[q]));
  return dt;
}

/* One SSP-RK3 step of all blocks (RK.cuh:100-192): per stage every block's update and boundary conditions, then
 * the block-to-block exchange and the properties. */
void orc_multi_rk3_step(orc_block **B, int nb, orc_link **L, int nl, double dt) {
  for (int q = 0; q < nb; ++q) {
    sponge_mean(B[q]);
    shock_sensor(B[q]);
    memcpy(B[q]->qn, B[q]->cv, sizeof(double) * 5 * B[q]->sz);
  }
  for (int rk = 0; rk < 3; ++rk) {
    multi_rhs_core(B, nb, L, nl);
    for (int q = 0; q < nb; ++q) {
      update(B[q], dt, rk);
      apply_bcs(B[q]);
    }
    orc_multi_exchange(B, nb, L, nl);
  }
}
