/*
 * oracle_monitor.c -- CPU restatement of the step monitor and of limit_flow (TEST INFRASTRUCTURE ONLY; see
 * oracle.h).  Arrays use the DZone layout of oracle.h (bv: 6 components of sz nodes, ngg ghosts).
 */
#include <math.h>
#include <stdlib.h>

#include "oracle.h"

static long node(int mx, int my, int ngg, int i, int j, int k) {
  const long d2 = mx + 2 * ngg, d1 = (long)(my + 2 * ngg) * d2;
  return (long)(k + ngg) * d1 + (long)(j + ngg) * d2 + (i + ngg);
}

static long padded_size(int mx, int my, int mz, int ngg) {
  return (long)(mx + 2 * ngg) * (my + 2 * ngg) * (mz + 2 * ngg);
}

/* store_last_step, TimeAdvanceFunc.cu:6-19. */
void orc_store_last(int mx, int my, int mz, int ngg, const double *bv, double *last) {
  const long sz = padded_size(mx, my, mz, ngg), n = (long)mx * my * mz;
  for (int k = 0; k < mz; ++k)
    for (int j = 0; j < my; ++j)
      for (int i = 0; i < mx; ++i) {
        const long c = ((long)k * my + j) * mx + i, id = node(mx, my, ngg, i, j, k);
        const double u = bv[id + sz], v = bv[id + 2 * sz], w = bv[id + 3 * sz];
        last[c] = bv[id];
        last[c + n] = sqrt(u * u + v * v + w * w);
        last[c + 2 * n] = bv[id + 4 * sz];
        last[c + 3 * n] = bv[id + 5 * sz];
      }
}

/* compute_square_of_dbv (TimeAdvanceFunc.cu:84-100) summed over the interior as reduction_of_dv_squared does
 * (Residual.cu:30-46): sums[l] = sum of (new - last)^2 for rho, |V|, p, T. */
void orc_residual_sums(int mx, int my, int mz, int ngg, const double *bv, const double *last, double *sums) {
  const long sz = padded_size(mx, my, mz, ngg), n = (long)mx * my * mz;
  double s[4] = {0, 0, 0, 0};
  for (int k = 0; k < mz; ++k)
    for (int j = 0; j < my; ++j)
      for (int i = 0; i < mx; ++i) {
        const long c = ((long)k * my + j) * mx + i, id = node(mx, my, ngg, i, j, k);
        const double u = bv[id + sz], v = bv[id + 2 * sz], w = bv[id + 3 * sz];
        const double d[4] = {bv[id] - last[c], sqrt(u * u + v * v + w * w) - last[c + n],
                             bv[id + 4 * sz] - last[c + 2 * n], bv[id + 5 * sz] - last[c + 3 * n]};
        for (int l = 0; l < 4; ++l) s[l] += d[l] * d[l];
      }
  for (int l = 0; l < 4; ++l) sums[l] = s[l];
}

static int out_of_bounds(const double *bv, long id, long sz, const double *ll, const double *ul) {
  for (int l = 0; l < 5; ++l) {
    const double v = bv[id + l * sz];
    if (isnan(v) || v < ll[l] || v > ul[l]) return 1;
  }
  return 0;
}

/* limit_flow, TimeAdvanceFunc.cuh:117-220: a node with rho, u, v, w or p NaN or outside [ll, ul] takes the mean
 * of the in-bounds nodes of its 3x3x3 neighbourhood inside the block (or its own values clipped to the bounds
 * when it has none); T = p / (rho R) with R = R_u / mw_air, or the node's mixture gas constant Rgas[id] when
 * Rgas is given.  Which nodes are in bounds is decided on the state before any change.  Returns the number of
 * nodes changed. */
int orc_limit_flow(int mx, int my, int mz, int ngg, double *bv, const double *ll, const double *ul,
                   const double *Rgas) {
  const long sz = padded_size(mx, my, mz, ngg), n = (long)mx * my * mz;
  unsigned char *bad = (unsigned char *)calloc((size_t)n, 1);
  double *out = (double *)malloc(sizeof(double) * 5 * (size_t)n);
  for (int k = 0; k < mz; ++k)
    for (int j = 0; j < my; ++j)
      for (int i = 0; i < mx; ++i)
        bad[((long)k * my + j) * mx + i] = (unsigned char)out_of_bounds(bv, node(mx, my, ngg, i, j, k), sz, ll, ul);
  int count = 0;
  for (int k = 0; k < mz; ++k)
    for (int j = 0; j < my; ++j)
      for (int i = 0; i < mx; ++i) {
        const long c = ((long)k * my + j) * mx + i;
        if (!bad[c]) continue;
        double sum[5] = {0, 0, 0, 0, 0};
        int kn = 0;
        for (int ka = -1; ka < 2; ++ka)
          for (int ja = -1; ja < 2; ++ja)
            for (int ia = -1; ia < 2; ++ia) {
              const int i1 = i + ia, j1 = j + ja, k1 = k + ka;
              if (i1 < 0 || i1 >= mx || j1 < 0 || j1 >= my || k1 < 0 || k1 >= mz) continue;
              if (bad[((long)k1 * my + j1) * mx + i1]) continue;
              const long jd = node(mx, my, ngg, i1, j1, k1);
              for (int l = 0; l < 5; ++l) sum[l] += bv[jd + l * sz];
              ++kn;
            }
        const long id = node(mx, my, ngg, i, j, k);
        for (int l = 0; l < 5; ++l) {
          if (kn > 0) {
            out[5 * c + l] = sum[l] * (1.0 / kn);
          } else {
            double v = bv[id + l * sz];
            v = fmax(v, ll[l]);
            out[5 * c + l] = fmin(v, ul[l]);
          }
        }
        ++count;
      }
  for (int k = 0; k < mz; ++k)
    for (int j = 0; j < my; ++j)
      for (int i = 0; i < mx; ++i) {
        const long c = ((long)k * my + j) * mx + i;
        if (!bad[c]) continue;
        const long id = node(mx, my, ngg, i, j, k);
        for (int l = 0; l < 5; ++l) bv[id + l * sz] = out[5 * c + l];
        bv[id + 5 * sz] = Rgas ? bv[id + 4 * sz] / (bv[id] * Rgas[id])
                               : bv[id + 4 * sz] * 28.9647 / (bv[id] * 8314.462618);
      }
  free(bad);
  free(out);
  return count;
}
