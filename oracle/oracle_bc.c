/*
 * oracle_bc.c -- CPU restatement of the inflow and symmetry faces and of the sponge layer's damping rate (TEST
 * INFRASTRUCTURE ONLY; see oracle.h).  The face routines work on raw DZone arrays (bv 6*sz, sv ns*sz, metric 9*sz)
 * and leave the recomputation of cv from bv to the caller, which knows its gas model.
 */
#include <math.h>
#include <stdint.h>

#include "oracle.h"

static const double kGamma = 1.4, kRu = 8314.462618, kMwAir = 28.9647;

/* SplitMix64 finaliser (Steele, Lea and Flood, "Fast splittable pseudorandom number generators", 2014). */
static uint64_t mix64(uint64_t z) {
  z += 0x9e3779b97f4a7c15ULL;
  z = (z ^ (z >> 30)) * 0xbf58476d1ce4e5b9ULL;
  z = (z ^ (z >> 27)) * 0x94d049bb133111ebULL;
  return z ^ (z >> 31);
}

/* The inflow noise: a counter-based normal deviate from (seed, application, face node, component) by Box-Muller on
 * two uniform numbers in (0, 1).  This build's own generator: the reference draws curand_normal_double from states
 * seeded with the clock (BoundCond.cu:1448-1454), which no run can reproduce. */
double orc_noise_normal(unsigned long long seed, unsigned long long call, long node, int comp) {
  const uint64_t key = mix64(mix64(mix64(seed) + call) + (uint64_t)node);
  const double u1 = ((double)(mix64(key + 2ULL * (uint64_t)comp) >> 11) + 0.5) * 0x1p-53;
  const double u2 = ((double)(mix64(key + 2ULL * (uint64_t)comp + 1) >> 11) + 0.5) * 0x1p-53;
  return sqrt(-2.0 * log(u1)) * cos(6.283185307179586 * u2);
}

static long nid(int mx, int my, int ngg, int i, int j, int k) {
  const long d2 = mx + 2 * ngg, d1 = (long)(my + 2 * ngg) * d2;
  return (long)(k + ngg) * d1 + (long)(j + ngg) * d2 + (i + ngg);
}

/* apply_inflow (BoundCond.cu:2190-2705), inflow_type 0 (constant, with inflow_sub the characteristic subsonic value
 * of BoundCond.cu:2505-2553) and 2 (mixing layer, BoundCond.cu:2451-2503), fluctuation_type 1 (white noise on u for
 * constant inflow, on u, v, w for the mixing layer, BoundCond.cu:2479-2503, 2562-2597).  prof: per face node the
 * state without noise (rho u v w p T Y..., [q * npt + node]); snd: acoustic speed (NULL: air, from T). */
void orc_bc_inflow(int mx, int my, int mz, int ngg, const int *s, const int *e, int face, int dir, double *bv,
                   double *sv, int ns, const double *snd, const double *met, const double *prof, const double *rms,
                   int sub_node, int sub_ghost, int fluct, double mw, unsigned long long seed,
                   unsigned long long call) {
  const long sz = (long)(mx + 2 * ngg) * (my + 2 * ngg) * (mz + 2 * ngg);
  const long n0 = e[0] - s[0] + 1, n1 = e[1] - s[1] + 1, n2 = e[2] - s[2] + 1, npt = n0 * n1 * n2;
  const double Rair = kRu / kMwAir;
  int d[3] = {0, 0, 0};
  d[face] = dir;
  for (long node = 0; node < npt; ++node) {
    const int i = s[0] + (int)(node % n0), j = s[1] + (int)((node / n0) % n1), k = s[2] + (int)(node / (n0 * n1));
    const long id = nid(mx, my, ngg, i, j, k);
    double rho = prof[node], u = prof[npt + node], v = prof[2 * npt + node], w = prof[3 * npt + node];
    double p = prof[4 * npt + node], T = prof[5 * npt + node];
    if (sub_node) {
      double nx = met[id + (3 * face) * sz], ny = met[id + (3 * face + 1) * sz], nz = met[id + (3 * face + 2) * sz];
      const double gi = dir / sqrt(nx * nx + ny * ny + nz * nz);
      nx *= gi;
      ny *= gi;
      nz *= gi;
      const double c0 = snd ? snd[id] : sqrt(kGamma * Rair * bv[id + 5 * sz]);
      const double Mn = (bv[id + sz] * nx + bv[id + 2 * sz] * ny + bv[id + 3 * sz] * nz) / c0;
      if (!(Mn < -1.0)) {
        const long i1 = nid(mx, my, ngg, i - d[0], j - d[1], k - d[2]);
        const double c1 = snd ? snd[i1] : sqrt(kGamma * Rair * bv[i1 + 5 * sz]);
        const double rho1 = bv[i1], p1 = bv[i1 + 4 * sz], pin = p;
        p = 0.5 * (p1 + pin - rho1 * c1 * ((u - bv[i1 + sz]) * nx + (v - bv[i1 + 2 * sz]) * ny +
                                           (w - bv[i1 + 3 * sz]) * nz));
        rho = rho + (p - pin) / (c1 * c1);
        u = u - (p - pin) / (rho1 * c1) * nx;
        v = v - (p - pin) / (rho1 * c1) * ny;
        w = w - (p - pin) / (rho1 * c1) * nz;
        T = snd ? p * mw / (rho * kRu) : p / (rho * Rair);
      }
    } else if (fluct) {
      u += orc_noise_normal(seed, call, node, 0) * rms[node];
      if (fluct == 2) {
        v += orc_noise_normal(seed, call, node, 1) * rms[node];
        w += orc_noise_normal(seed, call, node, 2) * rms[node];
      }
    }
    const double q[6] = {rho, u, v, w, p, T};
    for (int l = 0; l < 6; ++l) bv[id + l * sz] = q[l];
    for (int l = 0; l < ns; ++l) sv[id + l * sz] = prof[(6 + l) * npt + node];
    for (int g = 1; g <= ngg; ++g) {
      const long ig = nid(mx, my, ngg, i + g * d[0], j + g * d[1], k + g * d[2]);
      for (int l = 0; l < ns; ++l) sv[ig + l * sz] = prof[(6 + l) * npt + node];
      if (sub_ghost) {
        const long ii = nid(mx, my, ngg, i - g * d[0], j - g * d[1], k - g * d[2]);
        double rho2 = 2 * rho - bv[ii], p2 = 2 * p - bv[ii + 4 * sz];
        if (p2 <= 0.1 * p) p2 = 0.1 * p;
        if (rho2 <= 0.1 * rho) rho2 = 0.1 * rho;
        bv[ig] = rho2;
        bv[ig + sz] = 2 * u - bv[ii + sz];
        bv[ig + 2 * sz] = 2 * v - bv[ii + 2 * sz];
        bv[ig + 3 * sz] = 2 * w - bv[ii + 3 * sz];
        bv[ig + 4 * sz] = p2;
        bv[ig + 5 * sz] = snd ? p2 * mw / (rho2 * kRu) : p2 / (rho2 * Rair);
      } else {
        for (int l = 0; l < 6; ++l) bv[ig + l * sz] = q[l];
      }
    }
  }
}

/* apply_symmetry (BoundCond.cu:2002-2063): the boundary node takes the first interior node's state with the normal
 * velocity removed, ghost g the mirror image of interior node g. */
void orc_bc_symmetry(int mx, int my, int mz, int ngg, const int *s, const int *e, int face, int dir, double *bv,
                     double *sv, int ns, const double *met) {
  const long sz = (long)(mx + 2 * ngg) * (my + 2 * ngg) * (mz + 2 * ngg);
  int d[3] = {0, 0, 0};
  d[face] = dir;
  for (int k = s[2]; k <= e[2]; ++k)
    for (int j = s[1]; j <= e[1]; ++j)
      for (int i = s[0]; i <= e[0]; ++i) {
        const long id = nid(mx, my, ngg, i, j, k), in = nid(mx, my, ngg, i - d[0], j - d[1], k - d[2]);
        double kx = met[id + (3 * face) * sz], ky = met[id + (3 * face + 1) * sz], kz = met[id + (3 * face + 2) * sz];
        const double km = sqrt(kx * kx + ky * ky + kz * kz);
        kx /= km;
        ky /= km;
        kz /= km;
        const double u1 = bv[in + sz], v1 = bv[in + 2 * sz], w1 = bv[in + 3 * sz];
        const double uk = kx * u1 + ky * v1 + kz * w1;
        bv[id + sz] = u1 - kx * uk;
        bv[id + 2 * sz] = v1 - ky * uk;
        bv[id + 3 * sz] = w1 - kz * uk;
        bv[id] = bv[in];
        bv[id + 4 * sz] = bv[in + 4 * sz];
        bv[id + 5 * sz] = bv[in + 5 * sz];
        for (int l = 0; l < ns; ++l) sv[id + l * sz] = sv[in + l * sz];
        for (int g = 1; g <= ngg; ++g) {
          const long ig = nid(mx, my, ngg, i + g * d[0], j + g * d[1], k + g * d[2]);
          const long ii = nid(mx, my, ngg, i - g * d[0], j - g * d[1], k - g * d[2]);
          bv[ig] = bv[ii];
          const double u = bv[ii + sz], v = bv[ii + 2 * sz], w = bv[ii + 3 * sz];
          const double un = kx * u + ky * v + kz * w;
          bv[ig + sz] = u - 2 * un * kx;
          bv[ig + 2 * sz] = v - 2 * un * ky;
          bv[ig + 3 * sz] = w - 2 * un * kz;
          bv[ig + 4 * sz] = bv[ii + 4 * sz];
          bv[ig + 5 * sz] = bv[ii + 5 * sz];
          for (int l = 0; l < ns; ++l) sv[ig + l * sz] = sv[ii + l * sz];
        }
      }
}

/* Smooth ramp of the sponge layer, sponge_function 0 (SpongeLayer.cu:164-172): 6 xi^5 - 15 xi^4 + 10 xi^3. */
double orc_sponge_ramp(double xi) {
  const double x3 = xi * xi * xi;
  return 6 * x3 * xi * xi - 15 * x3 * xi + 10 * x3;
}
