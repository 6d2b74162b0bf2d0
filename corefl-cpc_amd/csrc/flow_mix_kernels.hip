// Mixture-flow kernels (9 species) for gfx950, FP64.  One thread per interface (fluxes) or node (update,
// properties); species thermodynamics from the NASA-7 fits of the mechanism.
#include "flow_mix.hpp"

#include <cmath>

#include "chem_device.hpp"
#include "mix_state.hpp"
#include "stencil.hpp"
#include "weno.hpp"

namespace corefl {

namespace {

using wenodev::weno5_half;
using wenodev::weno7_half;

constexpr int NS = 9, NV = 5 + NS;
using chemdev::kRu;
using stencil::dd;

__device__ __forceinline__ long nid(const DevBlock &b, int i, int j, int k) {
  return (long)(k + b.g) * b.d1 + (long)(j + b.g) * b.d2 + (i + b.g);
}

using mixdev::mix_cons;

__global__ void k_mix_cv_from_bv(DevBlock b, MixArrays x) {
  const long id = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (id >= b.sz) return;
  mix_cons(b, x, id);
}

// inner_communication<Mixture> (DataCommunication.cu:36-103): as the single-species face exchange, with the
// mass fractions exchanged alongside bv and cv (and the acoustic speed) recomputed from both.
__global__ __launch_bounds__(256) void k_mix_inner_exchange(DevBlock b, MixArrays x, DevBlock t, MixArrays y,
                                                            DevInner f) {
  const long n0 = f.n_point[0], n1 = f.n_point[1], n2 = f.n_point[2];
  const long item = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (item >= n0 * n1 * n2) return;
  const int n[3] = {(int)(item % n0), (int)((item / n0) % n1), (int)(item / (n0 * n1))};
  int d[3], idx[3], it[3];
  for (int c = 0; c < 3; ++c) {
    d[c] = f.loop_dir[c] * n[c];
    idx[c] = f.start[c] + d[c];
  }
  for (int c = 0; c < 3; ++c) it[c] = f.target_start[c] + f.target_loop_dir[c] * d[f.src_tar[c]];
  const long id = nid(b, idx[0], idx[1], idx[2]), jt = nid(t, it[0], it[1], it[2]);
  const long sz = b.sz, tz = t.sz;
  bool copy = idx[f.face] != f.face_layer, average = false;
  if (!copy) {
    const double us = sqrt(b.bv[id + sz] * b.bv[id + sz] + b.bv[id + 2 * sz] * b.bv[id + 2 * sz] +
                           b.bv[id + 3 * sz] * b.bv[id + 3 * sz]);
    if (us > 1e-15) {
      const double ut = sqrt(t.bv[jt + tz] * t.bv[jt + tz] + t.bv[jt + 2 * tz] * t.bv[jt + 2 * tz] +
                             t.bv[jt + 3 * tz] * t.bv[jt + 3 * tz]);
      copy = ut < 1e-15;
      average = !copy;
    }
  }
  if (average) {
    for (int l = 0; l < 6; ++l) {
      const double a = 0.5 * (t.bv[jt + l * tz] + b.bv[id + l * sz]);
      b.bv[id + l * sz] = a;
      t.bv[jt + l * tz] = a;
    }
    for (int l = 0; l < NS; ++l) {
      const double a = 0.5 * (y.sv[jt + l * tz] + x.sv[id + l * sz]);
      x.sv[id + l * sz] = a;
      y.sv[jt + l * tz] = a;
    }
  } else if (copy) {
    for (int l = 0; l < 6; ++l) b.bv[id + l * sz] = t.bv[jt + l * tz];
    for (int l = 0; l < NS; ++l) x.sv[id + l * sz] = y.sv[jt + l * tz];
  }
  mix_cons(b, x, id);
}

// apply_outflow / apply_periodic with the species (BoundCond.cu:2065-2188, 3033-3083); work items as in
// the single-species fill: one (face node, ghost layer) pair each.
__global__ __launch_bounds__(256) void k_mix_boundary(DevBlock b, MixArrays x, DevFace f) {
  const long n0 = f.end[0] - f.start[0] + 1, n1 = f.end[1] - f.start[1] + 1, n2 = f.end[2] - f.start[2] + 1;
  const long npt = n0 * n1 * n2;
  const long item = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (item >= npt * b.g) return;
  int g;
  long pt;
  if (f.face == 0) {
    g = (int)(item % b.g) + 1;
    pt = item / b.g;
  } else {
    g = (int)(item / npt) + 1;
    pt = item % npt;
  }
  const int i = f.start[0] + (int)(pt % n0), j = f.start[1] + (int)((pt / n0) % n1),
            k = f.start[2] + (int)(pt / (n0 * n1));
  int dv[3] = {0, 0, 0};
  dv[f.face] = f.direction;
  const long id = nid(b, i, j, k), sz = b.sz;
  const long ig = nid(b, i + g * dv[0], j + g * dv[1], k + g * dv[2]);
  long src = id;
  if (f.type == 8) {
    int o[3] = {i, j, k};
    const int ext[3] = {b.mx, b.my, b.mz};
    o[f.face] = f.direction < 0 ? ext[f.face] - 1 : 0;
    if (g == 1) {
      const long io = nid(b, o[0], o[1], o[2]);
      for (int l = 0; l < 6; ++l) {
        const double a = 0.5 * (b.bv[id + l * sz] + b.bv[io + l * sz]);
        b.bv[id + l * sz] = a;
        b.bv[io + l * sz] = a;
      }
      for (int l = 0; l < NS; ++l) {
        const double a = 0.5 * (x.sv[id + l * sz] + x.sv[io + l * sz]);
        x.sv[id + l * sz] = a;
        x.sv[io + l * sz] = a;
      }
      mix_cons(b, x, id);
    }
    src = nid(b, o[0] + g * dv[0], o[1] + g * dv[1], o[2] + g * dv[2]);
  }
  for (int l = 0; l < 6; ++l) b.bv[ig + l * sz] = b.bv[src + l * sz];
  for (int l = 0; l < NS; ++l) x.sv[ig + l * sz] = x.sv[src + l * sz];
  mix_cons(b, x, ig);
}

// Node (i, j) of this thread in a launch over the nx x ny nodes of each plane: every 256-thread block takes 256
// consecutive nodes in row-major order (grid.x blocks per plane, grid.z planes), so that only the last block of a
// plane has idle lanes (a 64-wide tiling of 129 nodes leaves a third wave per row with one active lane).
__device__ __forceinline__ bool chunk_node(int nx, int ny, int &i, int &j) {
  const unsigned p = blockIdx.x * 256u + threadIdx.y * 64u + threadIdx.x;
  j = (int)(p / (unsigned)nx);
  i = (int)(p - (unsigned)j * (unsigned)nx);
  return j < ny;
}

inline dim3 chunk_grid(int nx, int ny, int nz) { return dim3((unsigned)((nx * ny + 255) / 256), 1, nz); }

// compute_entropy_fix_delta<Mixture> (InviscidScheme.cu:269-304), interior plus one layer.
__global__ __launch_bounds__(256) void k_mix_entropy_fix(DevBlock b, MixArrays x, Scheme sc) {
  int i, j;
  if (!chunk_node(b.mx + 2, b.my + 2, i, j)) return;
  i -= 1;
  j -= 1;
  const int k = b.dim == 3 ? (int)blockIdx.z - 1 : 0;
  const long id = nid(b, i, j, k), sz = b.sz;
  const double *M = b.met + id;
  const double u = b.bv[id + sz], v = b.bv[id + 2 * sz], w = b.bv[id + 3 * sz];
  const double U = fabs(u * M[0] + v * M[sz] + w * M[2 * sz]);
  const double V = fabs(u * M[3 * sz] + v * M[4 * sz] + w * M[5 * sz]);
  const double W = fabs(u * M[6 * sz] + v * M[7 * sz] + w * M[8 * sz]);
  const double gx = sqrt(M[0] * M[0] + M[sz] * M[sz] + M[2 * sz] * M[2 * sz]);
  const double ge = sqrt(M[3 * sz] * M[3 * sz] + M[4 * sz] * M[4 * sz] + M[5 * sz] * M[5 * sz]);
  const double gz = sqrt(M[6 * sz] * M[6 * sz] + M[7 * sz] * M[7 * sz] + M[8 * sz] * M[8 * sz]);
  const double c = x.snd[id];
  b.efd[id] = b.dim == 2 ? sc.entropy_fix_factor * (U + V + c * 0.5 * (gx + ge))
                         : sc.entropy_fix_factor * (U + V + W + c * (gx + ge + gz) / 3.0);
}

__device__ __forceinline__ double minmod(double a, double c) {
  const double sa = (a > 0) - (a < 0), sc = (c > 0) - (c < 0);
  return 0.5 * (sa + sc) * fmin(fabs(a), fabs(c));
}

// WENO5-JS (WENO.cu:3059-3114) of one side: the reconstruction at the right edge of v[2] from v[0..4].
__device__ __forceinline__ double weno5_left(const double *v, double eps) {
  const double b0 = 13.0 / 12 * (v[0] - 2 * v[1] + v[2]) * (v[0] - 2 * v[1] + v[2]) +
                    0.25 * (v[0] - 4 * v[1] + 3 * v[2]) * (v[0] - 4 * v[1] + 3 * v[2]);
  const double b1 = 13.0 / 12 * (v[1] - 2 * v[2] + v[3]) * (v[1] - 2 * v[2] + v[3]) +
                    0.25 * (v[1] - v[3]) * (v[1] - v[3]);
  const double b2 = 13.0 / 12 * (v[2] - 2 * v[3] + v[4]) * (v[2] - 2 * v[3] + v[4]) +
                    0.25 * (3 * v[2] - 4 * v[3] + v[4]) * (3 * v[2] - 4 * v[3] + v[4]);
  // the divisions as products with refined reciprocals (wenodev::rcp); 1/d0 and 1/d1 from one reciprocal of
  // their product (each d >= eps^2, so the product stays normal)
  const double d0 = (b0 + eps) * (b0 + eps), d1 = (b1 + eps) * (b1 + eps), d2 = (b2 + eps) * (b2 + eps);
  const double r01 = wenodev::rcp(d0 * d1);
  const double a0 = 0.1 * (d1 * r01), a1 = 0.6 * (d0 * r01), a2 = 0.3 * wenodev::rcp(d2);
  constexpr double o6 = 1.0 / 6;
  const double q0 = (2 * v[0] - 7 * v[1] + 11 * v[2]) * o6, q1 = (-v[1] + 5 * v[2] + 2 * v[3]) * o6,
               q2 = (2 * v[2] + 5 * v[3] - v[4]) * o6;
  return (a0 * q0 + a1 * q1 + a2 * q2) * wenodev::rcp(a0 + a1 + a2);
}

// Primitive reconstruction at the interface between stencil points R-1 and R of pv[2R][NV].
template<int REC, int R>
__device__ __forceinline__ void reconstruct(const double (&pv)[2 * R][NV], double eps, double *pl, double *pr) {
  constexpr int c = R - 1;
#pragma unroll
  for (int l = 0; l < NV; ++l) {
    if constexpr (REC == 1) {
      pl[l] = pv[c][l];
      pr[l] = pv[c + 1][l];
    } else if constexpr (REC == 2 || REC == 3) {
      const double d0 = pv[c][l] - pv[c - 1][l], d1 = pv[c + 1][l] - pv[c][l], d2 = pv[c + 2][l] - pv[c + 1][l];
      if constexpr (REC == 2) {
        constexpr double kappa = 1.0 / 3.0;
        pl[l] = pv[c][l] + 0.25 * ((1 - kappa) * minmod(d0, d1) + (1 + kappa) * minmod(d1, d0));
        pr[l] = pv[c + 1][l] - 0.25 * ((1 - kappa) * minmod(d2, d1) + (1 + kappa) * minmod(d1, d2));
      } else {
        pl[l] = pv[c][l] + 0.5 * minmod(d0, d1);
        pr[l] = pv[c + 1][l] - 0.5 * minmod(d1, d2);
      }
    } else {
      const double vl[5] = {pv[0][l], pv[1][l], pv[2][l], pv[3][l], pv[4][l]};
      const double vr[5] = {pv[5][l], pv[4][l], pv[3][l], pv[2][l], pv[1][l]};
      pl[l] = weno5_left(vl, eps);
      pr[l] = weno5_left(vr, eps);
    }
  }
  if constexpr (REC >= 4) {
    if (pl[0] <= 0 || pl[4] <= 0 || pr[0] <= 0 || pr[4] <= 0) {
#pragma unroll
      for (int l = 0; l < NV; ++l) {
        pl[l] = pv[c][l];
        pr[l] = pv[c + 1][l];
      }
    }
  }
}

// reconstruction<Mixture> post-processing (InviscidScheme.cu:184-210): pv[NV] = E, pv[NV+1] = gamma.
__device__ __forceinline__ void post(const ChemMech &m, double *pv) {
  double e = 0.5 * (pv[1] * pv[1] + pv[2] * pv[2] + pv[3] * pv[3]), imw = 0;
#pragma unroll
  for (int l = 0; l < NS; ++l) imw += pv[5 + l] * m.imw[l];
  const double t = pv[4] / (pv[0] * kRu * imw);
  double h[NS], cp[NS];
  chemdev::h_cp<NS>(m, t, h, cp);
  double cpt = 0, cvt = 0;
#pragma unroll
  for (int l = 0; l < NS; ++l) {
    cpt += cp[l] * pv[5 + l];
    cvt += pv[5 + l] * (cp[l] - kRu * m.imw[l]);
    e += h[l] * pv[5 + l];
  }
  pv[NV] = pv[0] * e - pv[4];
  pv[NV + 1] = cpt / cvt;
}

// compute_half_sum_left_right_flux + riemannSolver_Roe<Mixture> (InviscidScheme.cu:601-765)
__device__ void roe(const ChemMech &m, const double *pl, const double *pr, const double *ml, const double *mr,
                    double jl, double jr, double efl, double efr, double *fc) {
  double JKx = jl * ml[0], JKy = jl * ml[1], JKz = jl * ml[2];
  double Uk = pl[1] * JKx + pl[2] * JKy + pl[3] * JKz;
  double co = Uk * pl[0];
  fc[0] = 0.5 * co;
  fc[1] = 0.5 * (co * pl[1] + pl[4] * JKx);
  fc[2] = 0.5 * (co * pl[2] + pl[4] * JKy);
  fc[3] = 0.5 * (co * pl[3] + pl[4] * JKz);
  fc[4] = 0.5 * Uk * (pl[4] + pl[NV]);
#pragma unroll
  for (int l = 5; l < NV; ++l) fc[l] = 0.5 * co * pl[l];
  JKx = jr * mr[0];
  JKy = jr * mr[1];
  JKz = jr * mr[2];
  Uk = pr[1] * JKx + pr[2] * JKy + pr[3] * JKz;
  co = Uk * pr[0];
  fc[0] += 0.5 * co;
  fc[1] += 0.5 * (co * pr[1] + pr[4] * JKx);
  fc[2] += 0.5 * (co * pr[2] + pr[4] * JKy);
  fc[3] += 0.5 * (co * pr[3] + pr[4] * JKz);
  fc[4] += 0.5 * Uk * (pr[4] + pr[NV]);
#pragma unroll
  for (int l = 5; l < NV; ++l) fc[l] += 0.5 * co * pr[l];

  const double dl = sqrt(pl[0]), dr = sqrt(pr[0]);
  const double inv = 1.0 / (dl + dr);
  const double u = (dl * pl[1] + dr * pr[1]) * inv, v = (dl * pl[2] + dr * pr[2]) * inv,
               w = (dl * pl[3] + dr * pr[3]) * inv;
  const double ek = 0.5 * (u * u + v * v + w * w);
  const double h = (dl * (pl[NV] + pl[4]) / pl[0] + dr * (pr[NV] + pr[4]) / pr[0]) * inv;
  double svm[NS], mw_inv = 0;
#pragma unroll
  for (int l = 0; l < NS; ++l) {
    svm[l] = (dl * pl[l + 5] + dr * pr[l + 5]) * inv;
    mw_inv += svm[l] * m.imw[l];
  }
  const double t = (dl * pl[4] / pl[0] + dr * pr[4] / pr[0]) * inv / (kRu * mw_inv);
  double hi[NS], cpi[NS];
  chemdev::h_cp<NS>(m, t, hi, cpi);
  double cvt = 0, cpt = 0;
#pragma unroll
  for (int l = 0; l < NS; ++l) {
    cpt += cpi[l] * svm[l];
    cvt += svm[l] * (cpi[l] - kRu * m.imw[l]);
  }
  const double gamma = cpt / cvt;
  const double c = sqrt(gamma * kRu * mw_inv * t);
  const double mw = 1.0 / mw_inv;
  double kx = 0.5 * (ml[0] + mr[0]), ky = 0.5 * (ml[1] + mr[1]), kz = 0.5 * (ml[2] + mr[2]);
  const double gradK = sqrt(kx * kx + ky * ky + kz * kz);
  double U = kx * u + ky * v + kz * w;
  double ch[3] = {U - gradK * c, U, U + gradK * c};
  const double efa = 0.5 * (efl + efr);
#pragma unroll
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
#pragma unroll
  for (int l = 1; l < NV; ++l) dq[l] = jave * (pr[0] * pr[l] - pl[0] * pl[l]);
  dq[4] = jave * (pr[NV] - pl[NV]);
  const double ic2 = 1.0 / (c * c);
  double c1 = (gamma - 1) * (ek * dq[0] - u * dq[1] - v * dq[2] - w * dq[3] + dq[4]) / (c * c);
  double c2 = (kx * dq[1] + ky * dq[2] + kz * dq[3] - U * dq[0]) / c;
#pragma unroll
  for (int l = 0; l < NS; ++l) c1 += (mw * m.imw[l] - hi[l] * (gamma - 1) * ic2) * dq[l + 5];
  double c3 = dq[0] - c1;
  double L[NV];
  L[0] = 0.5 * (c1 - c2);
  L[1] = kx * c3 - ((kz * v - ky * w) * dq[0] - kz * dq[2] + ky * dq[3]) / c;
  L[2] = ky * c3 - ((kx * w - kz * u) * dq[0] - kx * dq[3] + kz * dq[1]) / c;
  L[3] = kz * c3 - ((ky * u - kx * v) * dq[0] - ky * dq[1] + kx * dq[2]) / c;
  L[4] = 0.5 * (c1 + c2);
#pragma unroll
  for (int l = 0; l < NS; ++l) L[l + 5] = dq[l + 5] - svm[l] * dq[0];
  double bb[NV];
  bb[0] = -ch[0] * L[0];
#pragma unroll
  for (int l = 1; l < NV; ++l) bb[l] = -ch[1] * L[l];
  bb[4] = -ch[2] * L[4];
  const double c0 = kx * bb[1] + ky * bb[2] + kz * bb[3];
  c1 = c0 + bb[0] + bb[4];
  c2 = c * (bb[4] - bb[0]);
  c3 = 0;
#pragma unroll
  for (int l = 0; l < NS; ++l) c3 += (hi[l] - mw * m.imw[l] * c * c / (gamma - 1)) * bb[l + 5];
  fc[0] += 0.5 * c1;
  fc[1] += 0.5 * (u * c1 + kx * c2 - c * (kz * bb[2] - ky * bb[3]));
  fc[2] += 0.5 * (v * c1 + ky * c2 - c * (kx * bb[3] - kz * bb[1]));
  fc[3] += 0.5 * (w * c1 + kz * c2 - c * (ky * bb[1] - kx * bb[2]));
  fc[4] += 0.5 * (h * c1 + U * c2 - c * c * c0 / (gamma - 1) +
                  c * ((kz * v - ky * w) * bb[1] + (kx * w - kz * u) * bb[2] + (ky * u - kx * v) * bb[3]) + c3);
#pragma unroll
  for (int l = 0; l < NS; ++l) fc[5 + l] += 0.5 * (bb[l + 5] + svm[l] * c1);
}

// riemannSolver_ausmPlus<Mixture> (InviscidScheme.cu:397-469)
__device__ void ausm(const double *pl, const double *pr, const double *ml, const double *mr, double jl, double jr,
                     double *fc) {
  const double k1 = 0.5 * (jl * ml[0] + jr * mr[0]), k2 = 0.5 * (jl * ml[1] + jr * mr[1]),
               k3 = 0.5 * (jl * ml[2] + jr * mr[2]);
  const double gk = sqrt(k1 * k1 + k2 * k2 + k3 * k3);
  const double ul = (k1 * pl[1] + k2 * pl[2] + k3 * pl[3]) / gk, ur = (k1 * pr[1] + k2 * pr[2] + k3 * pr[3]) / gk;
  const double c = 0.5 * (sqrt(pl[NV + 1] * pl[4] / pl[0]) + sqrt(pr[NV + 1] * pr[4] / pr[0]));
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
  const double mf = c * (pl[0] * 0.5 * (mh + fabs(mh)) + pr[0] * 0.5 * (mh - fabs(mh)));
  const double co = mf * gk;
  double s[NV + 1];
#pragma unroll
  for (int l = 0; l < NV + 1; ++l) s[l] = mf >= 0 ? pl[l] : pr[l];
  fc[0] = co;
  fc[1] = co * s[1] + pc * k1;
  fc[2] = co * s[2] + pc * k2;
  fc[3] = co * s[3] + pc * k3;
  fc[4] = co * (s[NV] + s[4]) / s[0];
#pragma unroll
  for (int l = 5; l < NV; ++l) fc[l] = co * s[l];
}

// riemannSolver_hllc<Mixture> (InviscidScheme.cu:471-599)
__device__ void hllc(const ChemMech &m, const double *pl, const double *pr, const double *ml, const double *mr,
                     double jl, double jr, double *fc) {
  const double sl0 = sqrt(pl[0]), sr0 = sqrt(pr[0]);
  const double rl = sl0 / (sl0 + sr0), rr = sr0 / (sl0 + sr0);
  const double ut = pl[1] * rl + pr[1] * rr, vt = pl[2] * rl + pr[2] * rr, wt = pl[3] * rl + pr[3] * rr;
  double svm[NS], mw_inv = 0;
#pragma unroll
  for (int l = 0; l < NS; ++l) {
    svm[l] = rl * pl[l + 5] + rr * pr[l + 5];
    mw_inv += svm[l] * m.imw[l];
  }
  const double t = (rl * pl[4] / pl[0] + rr * pr[4] / pr[0]) / (kRu * mw_inv);
  double hi[NS], cpi[NS];
  chemdev::h_cp<NS>(m, t, hi, cpi);
  double cvt = 0, cpt = 0;
#pragma unroll
  for (int l = 0; l < NS; ++l) {
    cpt += cpi[l] * svm[l];
    cvt += svm[l] * (cpi[l] - kRu * m.imw[l]);
  }
  const double ct = sqrt(cpt / cvt * kRu * mw_inv * t);
  const double kx = 0.5 * (ml[0] + mr[0]), ky = 0.5 * (ml[1] + mr[1]), kz = 0.5 * (ml[2] + mr[2]);
  const double gk = sqrt(kx * kx + ky * ky + kz * kz);
  const double Ub = (kx * ut + ky * vt + kz * wt) / gk;
  const double ja = 0.5 * (jl + jr);
  const double Ul = kx * pl[1] + ky * pl[2] + kz * pl[3], Ur = kx * pr[1] + ky * pr[2] + kz * pr[3];
  const double cl = sqrt(pl[NV + 1] * pl[4] / pl[0]), cr = sqrt(pr[NV + 1] * pr[4] / pr[0]);
  const double sl = fmin(Ul / gk - cl, Ub - ct), sr = fmax(Ur / gk + cr, Ub + ct);
  double q[NV + 2];
  if (sl >= 0 || sr < 0) {
    const bool lft = sl >= 0;
#pragma unroll
    for (int l = 0; l < NV + 1; ++l) q[l] = lft ? pl[l] : pr[l]; // selects keep the indexing static
    const double Uq = lft ? Ul : Ur;
    const double rU = q[0] * Uq;
    fc[0] = ja * rU;
    fc[1] = ja * (rU * q[1] + q[4] * kx);
    fc[2] = ja * (rU * q[2] + q[4] * ky);
    fc[3] = ja * (rU * q[3] + q[4] * kz);
    fc[4] = ja * ((q[NV] + q[4]) * Uq);
#pragma unroll
    for (int l = 5; l < NV; ++l) fc[l] = ja * rU * q[l];
    return;
  }
  const double sm = ((pr[0] * Ur * (sr - Ur / gk) - pl[0] * Ul * (sl - Ul / gk)) / gk + pl[4] - pr[4]) /
                    (pr[0] * (sr - Ur / gk) - pl[0] * (sl - Ul / gk));
  const double pm = pl[0] * (sl - Ul / gk) * (sm - Ul / gk) + pl[4];
  const bool left = sm >= 0;
#pragma unroll
  for (int l = 0; l < NV + 1; ++l) q[l] = left ? pl[l] : pr[l];
  const double S = left ? sl : sr, Uq = left ? Ul : Ur;
  const double pC = 1.0 / (S - sm);
  const double QC = ja * pC * sm * (S * gk - Uq) * q[0];
  fc[0] = QC;
  const double dP = (S * pm - sm * q[4]) * pC * ja;
  fc[1] = QC * q[1] + dP * kx;
  fc[2] = QC * q[2] + dP * ky;
  fc[3] = QC * q[3] + dP * kz;
  fc[4] = QC * q[NV] / q[0] + pC * ja * (S * pm * sm * gk - sm * q[4] * Uq);
#pragma unroll
  for (int l = 5; l < NV; ++l) fc[l] = QC * q[l];
}

// Interface flux F(i+1/2) along DIR, stored at node i; interfaces -1/2 .. m-1/2.
template<int DIR, int REC, int RS>
__global__ __launch_bounds__(256) void k_mix_flux(DevBlock b, MixArrays x, Scheme sc, double *F) {
  constexpr int R = REC >= 4 ? 3 : 2;
  int i, j, k = blockIdx.z;
  if (!chunk_node(b.mx + (DIR == 0), b.my + (DIR == 1), i, j)) return;
  if constexpr (DIR == 0) i -= 1;
  else if constexpr (DIR == 1) j -= 1;
  else k -= 1;
  if (i >= b.mx || j >= b.my) return;
  const long st = DIR == 0 ? 1 : (DIR == 1 ? b.d2 : b.d1);
  const long id = nid(b, i, j, k), sz = b.sz;
  double pv[2 * R][NV];
#pragma unroll
  for (int q = 0; q < 2 * R; ++q) {
    const long iq = id + (q - R + 1) * st;
#pragma unroll
    for (int l = 0; l < 5; ++l) pv[q][l] = b.bv[iq + l * sz];
#pragma unroll
    for (int l = 0; l < NS; ++l) pv[q][5 + l] = x.sv[iq + l * sz];
  }
  double pl[NV + 2], pr[NV + 2];
  reconstruct<REC, R>(pv, sc.weno_eps, pl, pr);
  const ChemMech &m = *x.mech;
  post(m, pl);
  post(m, pr);
  double ml[3], mr[3];
#pragma unroll
  for (int c = 0; c < 3; ++c) {
    ml[c] = b.met[id + (3 * DIR + c) * sz];
    mr[c] = b.met[id + st + (3 * DIR + c) * sz];
  }
  double fc[NV];
  if constexpr (RS == 2) roe(m, pl, pr, ml, mr, b.jac[id], b.jac[id + st], b.efd[id], b.efd[id + st], fc);
  else if constexpr (RS == 4) hllc(m, pl, pr, ml, mr, b.jac[id], b.jac[id + st], fc);
  else ausm(pl, pr, ml, mr, b.jac[id], b.jac[id + st], fc);
#pragma unroll
  for (int l = 0; l < NV; ++l) F[id + l * sz] = fc[l];
}

__constant__ double kRkA[3] = {1.0, 0.75, 1.0 / 3.0}; // SSPRK3 (RK.cuh:11-15)
__constant__ double kRkB[3] = {0.0, 0.25, 2.0 / 3.0};
__constant__ double kRkC[3] = {1.0, 0.25, 2.0 / 3.0};

// dq = -sum_dir (F(i+1/2) - F(i-1/2)) and update_cv_and_bv_rk<Mixture> (RK.cuh:224-280) with
// compute_temperature_and_pressure (FieldOperation.cu:4-45).  Stage 0 saves U^n into qn.
// Transport properties of every node (update_physical_properties<Mixture> with compute_transport_property).
__global__ __launch_bounds__(256) void k_mix_transport(DevBlock b, MixArrays x) {
  const long id = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (id >= b.sz) return;
  const long sz = b.sz;
  const double t = b.bv[id + 5 * sz];
  if (!(t > 0)) return; // unused corner nodes hold no state
  double y[NS], h[NS], cp[NS], rd[NS], mu, lam;
#pragma unroll
  for (int l = 0; l < NS; ++l) y[l] = x.sv[id + l * sz];
  chemdev::h_cp<NS>(*x.mech, t, h, cp);
  chemdev::transport<NS>(*x.mech, t, y, cp, mu, lam, rd);
  x.mul[id] = mu;
  x.cond[id] = lam;
#pragma unroll
  for (int l = 0; l < NS; ++l) x.rhod[id + l * sz] = rd[l];
}

// compute_viscous_flux_collocated<Mixture, ORDER> + compute_viscous_flux_collocated_scalar<ORDER>
// (ViscousScheme.cu:309-566, gradPInDiffusionFlux off): node fluxes fv/gv/hv, components 1..3 the stress
// projected on each direction's metric row, 4 the work of the stress plus conduction (node conductivity) plus
// the enthalpy carried by species diffusion, 5.. the mixture-averaged diffusion fluxes with the correction
// velocity.  Gradients by the collocated ORDER-th central difference with one-sided rows at physical faces.
template<int ORDER>
__global__ __launch_bounds__(256) void k_mix_viscous_node(DevBlock b, MixArrays x) {
  int i, j;
  const int k = blockIdx.z;
  if (!chunk_node(b.mx, b.my, i, j)) return;
  const long id = nid(b, i, j, k), sz = b.sz;
  const bool z3 = b.dim == 3;
  const double *M = b.met + id;
  const double xx = M[0], xy = M[sz], xz = M[2 * sz], ex = M[3 * sz], ey = M[4 * sz], ez = M[5 * sz];
  const double zx = M[6 * sz], zy = M[7 * sz], zz = M[8 * sz];
  // physical gradient of f (chain rule through the metric)
  auto grad = [&](const double *f, double &gx, double &gy, double &gz) {
    const double a = dd<ORDER>(f, id, 1, i, b.mx, b.bt[0], b.bt[1]);
    const double e = dd<ORDER>(f, id, b.d2, j, b.my, b.bt[2], b.bt[3]);
    const double c = z3 ? dd<ORDER>(f, id, b.d1, k, b.mz, b.bt[4], b.bt[5]) : 0.0;
    gx = a * xx + e * ex + c * zx;
    gy = a * xy + e * ey + c * zy;
    gz = a * xz + e * ez + c * zz;
  };
  double ux, uy, uz, vx, vy, vz, wx, wy, wz, tx, ty, tz;
  grad(b.bv + sz, ux, uy, uz);
  grad(b.bv + 2 * sz, vx, vy, vz);
  grad(b.bv + 3 * sz, wx, wy, wz);
  grad(b.bv + 5 * sz, tx, ty, tz);
  const double mu = x.mul[id];
  const double txx = mu * (4 * ux - 2 * vy - 2 * wz) / 3.0;
  const double tyy = mu * (4 * vy - 2 * ux - 2 * wz) / 3.0;
  const double tzz = mu * (4 * wz - 2 * ux - 2 * vy) / 3.0;
  const double txy = mu * (uy + vx), txz = mu * (uz + wx), tyz = mu * (vz + wy);
  double *fv = b.fv + id, *gv = b.gv + id, *hv = b.hv + id;
  fv[sz] = xx * txx + xy * txy + xz * txz;
  fv[2 * sz] = xx * txy + xy * tyy + xz * tyz;
  fv[3 * sz] = xx * txz + xy * tyz + xz * tzz;
  gv[sz] = ex * txx + ey * txy + ez * txz;
  gv[2 * sz] = ex * txy + ey * tyy + ez * tyz;
  gv[3 * sz] = ex * txz + ey * tyz + ez * tzz;
  hv[sz] = zx * txx + zy * txy + zz * txz;
  hv[2 * sz] = zx * txy + zy * tyy + zz * tyz;
  hv[3 * sz] = zx * txz + zy * tyz + zz * tzz;
  const double lam = x.cond[id];
  const double u = b.bv[id + sz], v = b.bv[id + 2 * sz], w = b.bv[id + 3 * sz];
  const double Ex = u * txx + v * txy + w * txz + lam * tx;
  const double Ey = u * txy + v * tyy + w * tyz + lam * ty;
  const double Ez = u * txz + v * tyz + w * tzz + lam * tz;
  double e4[3] = {xx * Ex + xy * Ey + xz * Ez, ex * Ex + ey * Ey + ez * Ez, zx * Ex + zy * Ey + zz * Ez};
  const ChemMech &m = *x.mech;
  double yk[NS], dk[NS], drv[NS][3];
  double cvt[3] = {0, 0, 0}, sgw[3] = {0, 0, 0}, imw = 0, srdy = 0;
#pragma unroll
  for (int l = 0; l < NS; ++l) {
    yk[l] = x.sv[id + l * sz];
    dk[l] = x.rhod[id + l * sz];
    double gx, gy, gz;
    grad(x.sv + l * sz, gx, gy, gz);
    drv[l][0] = xx * gx + xy * gy + xz * gz;
    drv[l][1] = ex * gx + ey * gy + ez * gz;
    drv[l][2] = zx * gx + zy * gy + zz * gz;
#pragma unroll
    for (int a = 0; a < 3; ++a) {
      cvt[a] += dk[l] * drv[l][a];
      sgw[a] += drv[l][a] * m.imw[l];
    }
    imw += yk[l] * m.imw[l];
    srdy += dk[l] * yk[l];
  }
  const double mwt = 1.0 / imw;
#pragma unroll
  for (int a = 0; a < 3; ++a) cvt[a] -= mwt * srdy * sgw[a];
  double h[NS], cp[NS];
  chemdev::h_cp<NS>(m, b.bv[id + 5 * sz], h, cp);
  double *F[3] = {fv, gv, hv};
#pragma unroll
  for (int l = 0; l < NS; ++l)
#pragma unroll
    for (int a = 0; a < 3; ++a) {
      const double jd = dk[l] * (drv[l][a] - mwt * yk[l] * sgw[a]) - yk[l] * cvt[a];
      F[a][(5 + l) * sz] = jd;
      e4[a] += h[l] * jd;
    }
  fv[4 * sz] = e4[0];
  gv[4 * sz] = e4[1];
  hv[4 * sz] = e4[2];
}

// compute_fv/gv/hv_2nd_order<Mixture> (ViscousScheme.cu:614-778 and the gv/hv analogues): the face flux of the
// air scheme with the nodes' average viscosity and conductivity, plus species diffusion with the correction
// velocity (mixture-averaged rho D_k, mass fractions and their gradients at the face) and its enthalpy flux;
// stored at node t for the face t+1/2 along DIR, t = -1 .. m-1, components 1..4 and 5..5+NS-1.
template<int DIR>
__global__ __launch_bounds__(256) void k_mix_viscous_face(DevBlock b, MixArrays x, Scheme sc) {
  int i, j, k = blockIdx.z;
  if (!chunk_node(b.mx + (DIR == 0), b.my + (DIR == 1), i, j)) return;
  if constexpr (DIR == 0) i -= 1;
  else if constexpr (DIR == 1) j -= 1;
  else k -= 1;
  if (i >= b.mx || j >= b.my) return;
  const long sz = b.sz;
  const long st[3] = {1, b.d2, b.d1};
  const long id = nid(b, i, j, k), ie = id + st[DIR];
  const double *M = b.met;
  double m[9];
#pragma unroll
  for (int c = 0; c < 9; ++c) m[c] = 0.5 * (M[id + c * sz] + M[ie + c * sz]);
  auto grad = [&](const double *f, double *gx) { // d f / d(x, y, z) at the face
    double g[3];
#pragma unroll
    for (int a = 0; a < 3; ++a) {
      if (a == DIR) g[a] = f[ie] - f[id];
      else if (a == 2 && b.dim == 2) g[a] = 0;
      else g[a] = 0.25 * (f[id + st[a]] - f[id - st[a]] + f[ie + st[a]] - f[ie - st[a]]);
    }
#pragma unroll
    for (int c = 0; c < 3; ++c) gx[c] = g[0] * m[c] + g[1] * m[3 + c] + g[2] * m[6 + c];
  };
  double du[3], dv[3], dw[3], dT[3];
  grad(b.bv + sz, du);
  grad(b.bv + 2 * sz, dv);
  grad(b.bv + 3 * sz, dw);
  grad(b.bv + 5 * sz, dT);
  const double mu = 0.5 * (x.mul[id] + x.mul[ie]);
  const double txx = mu * (4 * du[0] - 2 * dv[1] - 2 * dw[2]) / 3.0;
  const double tyy = mu * (4 * dv[1] - 2 * du[0] - 2 * dw[2]) / 3.0;
  const double tzz = mu * (4 * dw[2] - 2 * du[0] - 2 * dv[1]) / 3.0;
  const double txy = mu * (du[1] + dv[0]), txz = mu * (du[2] + dw[0]), tyz = mu * (dv[2] + dw[1]);
  double kj[3];
#pragma unroll
  for (int c = 0; c < 3; ++c)
    kj[c] = 0.5 * (M[id + (3 * DIR + c) * sz] * b.jac[id] + M[ie + (3 * DIR + c) * sz] * b.jac[ie]);
  const double f1 = kj[0] * txx + kj[1] * txy + kj[2] * txz;
  const double f2 = kj[0] * txy + kj[1] * tyy + kj[2] * tyz;
  const double f3 = kj[0] * txz + kj[1] * tyz + kj[2] * tzz;
  const double um = 0.5 * (b.bv[id + sz] + b.bv[ie + sz]);
  const double vm = 0.5 * (b.bv[id + 2 * sz] + b.bv[ie + 2 * sz]);
  const double wm = 0.5 * (b.bv[id + 3 * sz] + b.bv[ie + 3 * sz]);
  const double lam = 0.5 * (x.cond[id] + x.cond[ie]);
  double e = um * f1 + vm * f2 + wm * f3 + lam * (kj[0] * dT[0] + kj[1] * dT[1] + kj[2] * dT[2]);
  // species diffusion (ViscousScheme.cu:697-778; no pressure-gradient term)
  const ChemMech &mc = *x.mech;
  double yk[NS], dk[NS], gy[NS];
  double sgw = 0, srdy = 0, cvt = 0, imw = 0;
#pragma unroll
  for (int l = 0; l < NS; ++l) {
    yk[l] = 0.5 * (x.sv[id + l * sz] + x.sv[ie + l * sz]);
    dk[l] = 0.5 * (x.rhod[id + l * sz] + x.rhod[ie + l * sz]);
    double gY[3];
    grad(x.sv + l * sz, gY);
    gy[l] = kj[0] * gY[0] + kj[1] * gY[1] + kj[2] * gY[2];
    cvt += dk[l] * gy[l];
    sgw += gy[l] * mc.imw[l];
    imw += yk[l] * mc.imw[l];
    srdy += dk[l] * yk[l];
  }
  const double mwt = 1.0 / imw;
  cvt -= mwt * srdy * sgw;
  double hh[NS], cp[NS];
  chemdev::h_cp<NS>(mc, 0.5 * (b.bv[id + 5 * sz] + b.bv[ie + 5 * sz]), hh, cp);
  double *F = DIR == 0 ? b.fv : (DIR == 1 ? b.gv : b.hv);
#pragma unroll
  for (int l = 0; l < NS; ++l) {
    const double jd = dk[l] * (gy[l] - mwt * yk[l] * sgw) - yk[l] * cvt;
    F[id + (5 + l) * sz] = jd;
    e += hh[l] * jd;
  }
  F[id + sz] = f1;
  F[id + 2 * sz] = f2;
  F[id + 3 * sz] = f3;
  F[id + 4 * sz] = e;
}

// Species clipping and renormalisation of q, then the Newton temperature at the internal energy of q
// (update_cv_and_bv_rk<Mixture> tail, RK.cuh:224-280; compute_temperature_and_pressure,
// FieldOperation.cu:4-45); returns T, y the mass fractions, R the mixture gas constant.
__device__ __forceinline__ double mix_state(const ChemMech &m, double *q, double t0, double *y, double &R) {
  const double di = 1.0 / q[0];
  const double u = q[1] * di, v = q[2] * di, w = q[3] * di;
  double sum = 0;
#pragma unroll
  for (int l = 0; l < NS; ++l) {
    if (q[5 + l] < 0) q[5 + l] = 0;
    else sum += q[5 + l];
  }
  sum = 1.0 / sum;
  const double f = q[0] * sum;
  R = 0;
#pragma unroll
  for (int l = 0; l < NS; ++l) {
    y[l] = q[5 + l] * sum;
    q[5 + l] *= f;
    R += y[l] * kRu * m.imw[l];
  }
  // Newton iteration for T at fixed e, |1 - T1/T| < 1e-5, at most 1000 iterations
  const double e = q[4] * di - 0.5 * (u * u + v * v + w * w);
  double t = t0, err = 1;
  for (int it = 0; it < 1000 && err > 1e-5; ++it) {
    double h[NS], cp[NS];
    chemdev::h_cp<NS>(m, t, h, cp);
    double cpt = 0, hh = 0;
#pragma unroll
    for (int l = 0; l < NS; ++l) {
      cpt += cp[l] * y[l];
      hh += h[l] * y[l];
    }
    const double t1 = t - (hh - R * t - e) / (cpt - R);
    err = fabs(1 - t1 / t);
    t = t1;
  }
  return t;
}

// cn: the frozen transport residual of Wu's splitting (update_cv_and_bv_rk_wu, StrangSplitting.cu:504-562:
// the stage advances with R(U) - cn over dts * dt), or nullptr for the plain stage (dts = 1).
template<bool RHS_ONLY>
__global__ __launch_bounds__(256) void k_mix_update(DevBlock b, MixArrays x, int rk, const double *__restrict__ dtp,
                                                    const double *__restrict__ cn, double dts) {
  int i, j;
  const int k = blockIdx.z;
  if (!chunk_node(b.mx, b.my, i, j)) return;
  const long id = nid(b, i, j, k), sz = b.sz;
  double r[NV];
#pragma unroll
  for (int l = 0; l < NV; ++l) {
    double s = -(b.fx[id + l * sz] - b.fx[id - 1 + l * sz]) - (b.fy[id + l * sz] - b.fy[id - b.d2 + l * sz]);
    if (b.dim == 3) s -= b.fz[id + l * sz] - b.fz[id - b.d1 + l * sz];
    r[l] = s;
  }
  const long iq = ((long)k * b.my + j) * b.mx + i;
  if (x.visc_order == 2) { // face-centred viscous fluxes (compute_dFv_dx etc., ViscousScheme.cu:75-128)
#pragma unroll
    for (int l = 1; l < NV; ++l) {
      double v = b.fv[id + l * sz] - b.fv[id - 1 + l * sz];
      v += b.gv[id + l * sz] - b.gv[id - b.d2 + l * sz];
      if (b.dim == 3) v += b.hv[id + l * sz] - b.hv[id - b.d1 + l * sz];
      r[l] += v;
    }
  } else if (x.visc_order == 8) { // compute_viscous_flux_derivative<8> (ViscousScheme.cu:568-612)
    const double J = b.jac[id];
#pragma unroll
    for (int l = 1; l < NV; ++l) {
      double v = dd<8>(b.fv + l * sz, id, 1, i, b.mx, b.bt[0], b.bt[1]) +
                 dd<8>(b.gv + l * sz, id, b.d2, j, b.my, b.bt[2], b.bt[3]);
      if (b.dim == 3) v += dd<8>(b.hv + l * sz, id, b.d1, k, b.mz, b.bt[4], b.bt[5]);
      r[l] += v * J;
    }
  }
  if (x.explicit_source) {
    // finite_rate_chemistry with chemSrcMethod 0 (FiniteRateChem.cu:1388-1464): dq += J omega and the
    // reaction time scale min_k rho Y_k / omega_d,k
    const double rho = b.bv[id];
    double ry[NS], w[NS], wd[NS];
#pragma unroll
    for (int l = 0; l < NS; ++l) ry[l] = rho * x.sv[id + l * sz];
    chemdev::source<NS, false>(*x.mech, b.bv[id + 5 * sz], ry, w, nullptr, wd);
    const double J = b.jac[id];
    double ts = 1e+6;
#pragma unroll
    for (int l = 0; l < NS; ++l) {
      r[5 + l] += J * w[l];
      if (x.sv[id + l * sz] > 1e-25 && fabs(wd[l]) > 1e-25) ts = fmin(ts, fabs(ry[l] / wd[l]));
    }
    x.tscale[iq] = ts;
  }
  if (b.sponge_sigma) {
    // sponge layer: relaxation towards the running mean, -sigma (U - U_mean) per unit volume
    const double sg = b.sponge_sigma[iq] * b.jac[id];
#pragma unroll
    for (int l = 0; l < NV; ++l) r[l] -= sg * (b.cv[id + l * sz] - b.sponge_mean[iq + l * b.n]);
  }
  if constexpr (RHS_ONLY) {
#pragma unroll
    for (int l = 0; l < NV; ++l) b.dq[iq + l * b.n] = r[l];
    return;
  }
  if (cn) {
#pragma unroll
    for (int l = 0; l < NV; ++l) r[l] -= cn[iq + l * b.n];
  }
  const double dj = dtp[0] * dts / b.jac[id];
  double q[NV];
#pragma unroll
  for (int l = 0; l < NV; ++l) {
    const double u0 = b.cv[id + l * sz];
    if (rk == 0) {
      b.qn[id + l * sz] = u0;
      q[l] = kRkA[0] * u0 + kRkB[0] * u0 + kRkC[0] * dj * r[l];
    } else {
      q[l] = kRkA[rk] * b.qn[id + l * sz] + kRkB[rk] * u0 + kRkC[rk] * dj * r[l];
    }
  }
  if (b.mz == 1) q[3] = 0;
  const double di = 1.0 / q[0];
  const double u = q[1] * di, v = q[2] * di, w = q[3] * di;
  double y[NS], R;
  const ChemMech &m = *x.mech;
  const double t = mix_state(m, q, b.bv[id + 5 * sz], y, R);
#pragma unroll
  for (int l = 0; l < NV; ++l) b.cv[id + l * sz] = q[l];
  b.bv[id] = q[0];
  b.bv[id + sz] = u;
  b.bv[id + 2 * sz] = v;
  b.bv[id + 3 * sz] = w;
  b.bv[id + 4 * sz] = q[0] * t * R;
  b.bv[id + 5 * sz] = t;
#pragma unroll
  for (int l = 0; l < NS; ++l) x.sv[id + l * sz] = y[l];
  // acoustic speed of the new state (update_physical_properties<Mixture>, FieldOperation.cuh:86-120)
  double h[NS], cp[NS], cpt = 0;
  chemdev::h_cp<NS>(m, t, h, cp);
#pragma unroll
  for (int l = 0; l < NS; ++l) cpt += cp[l] * y[l];
  x.snd[id] = sqrt(cpt / (cpt - R) * R * t);
}

// local_time_step<Mixture> (TimeAdvanceFunc.cuh:28-115, inviscid) fused with a workgroup minimum.
__global__ __launch_bounds__(256) void k_mix_local_dt(DevBlock b, MixArrays x, Scheme sc, double *part) {
  int i, j;
  const int k = blockIdx.z;
  const bool in = chunk_node(b.mx, b.my, i, j);
  double dt = 1e+6;
  if (in) {
    const long id = nid(b, i, j, k), sz = b.sz;
    const double *M = b.met + id;
    const double gx = sqrt(M[0] * M[0] + M[sz] * M[sz] + M[2 * sz] * M[2 * sz]);
    const double ge = sqrt(M[3 * sz] * M[3 * sz] + M[4 * sz] * M[4 * sz] + M[5 * sz] * M[5 * sz]);
    const double gz = sqrt(M[6 * sz] * M[6 * sz] + M[7 * sz] * M[7 * sz] + M[8 * sz] * M[8 * sz]);
    const double u = b.bv[id + sz], v = b.bv[id + 2 * sz], w = b.bv[id + 3 * sz];
    const double c = x.snd[id];
    double s = fmax(fabs(u * M[0] + v * M[sz] + w * M[2 * sz]) + c * gx,
                    fabs(u * M[3 * sz] + v * M[4 * sz] + w * M[5 * sz]) + c * ge);
    if (b.dim == 3) s = fmax(s, fabs(u * M[6 * sz] + v * M[7 * sz] + w * M[8 * sz]) + c * gz);
    if (x.visc_order) { // viscous spectral radius (TimeAdvanceFunc.cuh:66-98)
      double ml = fmax(gx, ge);
      if (b.dim == 3) ml = fmax(ml, gz);
      double h[NS], cp[NS], cpt = 0, rmax = 0, R = 0;
      chemdev::h_cp<NS>(*x.mech, b.bv[id + 5 * sz], h, cp);
#pragma unroll
      for (int l = 0; l < NS; ++l) {
        cpt += cp[l] * x.sv[id + l * sz];
        R += x.sv[id + l * sz] * kRu * x.mech->imw[l];
        rmax = fmax(rmax, x.rhod[id + l * sz]);
      }
      const double ir = 1.0 / b.bv[id], gam = cpt / (cpt - R);
      double dv = x.mul[id] * ir;
      dv = fmax(dv, x.cond[id] * ir * gam / cpt);
      dv = fmax(dv, rmax * ir);
      s = fmax(s, ml * ml * dv);
    }
    dt = sc.cfl / s;
    if (x.explicit_source) {
      const double rdt = sc.cfl * x.tscale[((long)k * b.my + j) * b.mx + i];
      if (rdt > 1e-20 && rdt < 1e-5) dt = fmin(dt, rdt);
    }
  }
  for (int off = 32; off > 0; off >>= 1) dt = fmin(dt, __shfl_xor(dt, off, 64));
  __shared__ double wmin[4];
  if (threadIdx.x == 0) wmin[threadIdx.y] = dt;
  __syncthreads();
  if (threadIdx.x == 0 && threadIdx.y == 0) {
    const long p = ((long)blockIdx.z * gridDim.y + blockIdx.y) * gridDim.x + blockIdx.x;
    part[p] = fmin(fmin(wmin[0], wmin[1]), fmin(wmin[2], wmin[3]));
  }
}

__global__ __launch_bounds__(256) void k_mix_chem_gather(DevBlock b, MixArrays x, ChemBatch c) {
  int i, j;
  const int k = blockIdx.z;
  if (!chunk_node(b.mx, b.my, i, j)) return;
  const long id = nid(b, i, j, k), sz = b.sz;
  const long q = ((long)k * b.my + j) * b.mx + i;
  const double rho = b.cv[id];
  const double u = b.bv[id + sz], v = b.bv[id + 2 * sz], w = b.bv[id + 3 * sz];
  c.rho[q] = rho;
  c.e[q] = b.cv[id + 4 * sz] / rho - 0.5 * (u * u + v * v + w * w);
  c.T[q] = b.bv[id + 5 * sz];
#pragma unroll
  for (int l = 0; l < NS; ++l) c.rhoY[(long)l * c.n + q] = b.cv[id + (5 + l) * sz];
}

__global__ __launch_bounds__(256) void k_mix_chem_scatter(DevBlock b, MixArrays x, ChemBatch c) {
  int i, j;
  const int k = blockIdx.z;
  if (!chunk_node(b.mx, b.my, i, j)) return;
  const long id = nid(b, i, j, k), sz = b.sz;
  const long q = ((long)k * b.my + j) * b.mx + i;
  const ChemMech &m = *x.mech;
  const double rho = b.cv[id], T = c.T[q];
  double R = 0;
#pragma unroll
  for (int l = 0; l < NS; ++l) {
    const double ry = c.rhoY[(long)l * c.n + q];
    b.cv[id + (5 + l) * sz] = ry;
    x.sv[id + l * sz] = ry / rho;
    R += ry / rho * kRu * m.imw[l];
  }
  b.bv[id + 5 * sz] = T;
  b.bv[id + 4 * sz] = rho * R * T;
  double h[NS], cp[NS];
  chemdev::h_cp<NS>(m, T, h, cp);
  double cpt = 0;
#pragma unroll
  for (int l = 0; l < NS; ++l) cpt += cp[l] * x.sv[id + l * sz];
  x.snd[id] = sqrt(cpt / (cpt - R) * R * T);
}

// the grid of k_mix_local_dt: the same workgroups as the single-species time-step kernel's
dim3 node_grid(const DevBlock &b) { return chunk_grid(b.mx, b.my, b.mz); }

template<int DIR, int REC>
void dispatch_rs(const DevBlock &b, const MixArrays &x, const Scheme &sc, dim3 grid, double *F, hipStream_t s) {
  const dim3 tpb(64, 4, 1);
  switch (sc.inviscid_scheme) {
    case 2: k_mix_flux<DIR, REC, 2><<<grid, tpb, 0, s>>>(b, x, sc, F); break;
    case 4: k_mix_flux<DIR, REC, 4><<<grid, tpb, 0, s>>>(b, x, sc, F); break;
    default: k_mix_flux<DIR, REC, 3><<<grid, tpb, 0, s>>>(b, x, sc, F); break;
  }
}

// LLF flux-splitting WENO with species (WENO.cu:890-990 component-wise 51/71, WENO.cu:1146-1470
// characteristic 52/72): one thread per interface t+1/2 along DIR, stencil nodes t-R+1 .. t+R.  The spectral
// radius uses the stored acoustic speed.  The characteristic variant projects the five flow fields with the
// air eigenvectors at the mixture's Roe state plus alpha_l rho Y_l terms, and each species field as
// rho Y_l - Y_l,Roe rho; the flow fields are split and reconstructed in one pass over the stencil, the species
// one at a time, to stay within the register file.
template<int DIR, int R, bool CH>
__global__ __launch_bounds__(256) void k_mix_fsweno(DevBlock b, MixArrays x, Scheme sc, double *F) {
  int i, j, k = blockIdx.z;
  if (!chunk_node(b.mx + (DIR == 0), b.my + (DIR == 1), i, j)) return;
  if constexpr (DIR == 0) i -= 1;
  else if constexpr (DIR == 1) j -= 1;
  else k -= 1;
  if (i >= b.mx || j >= b.my) return;
  const long st = DIR == 0 ? 1 : (DIR == 1 ? b.d2 : b.d1);
  const long id = nid(b, i, j, k), sz = b.sz;
  constexpr int N = 2 * R;
  // the hybrid scheme's switch (WENO.cu:838-849): linear upwind weights where the stencil is free of shocks
  const bool shock = wenodev::stencil_shock(b.sensor, sc.sensor_threshold, id, st, b.g);
  auto weno = [shock](const double *vp, const double *vn, double eps) {
    if constexpr (R == 4) {
      return shock ? weno7_half(vp, eps) + weno7_half(vn, eps) : wenodev::ud7_half(vp) + wenodev::ud7_half(vn);
    } else {
      return shock ? weno5_half(vp[0], vp[1], vp[2], vp[3], vp[4], eps) +
                         weno5_half(vn[0], vn[1], vn[2], vn[3], vn[4], eps)
                   : wenodev::ud5_half(vp[0], vp[1], vp[2], vp[3], vp[4]) +
                         wenodev::ud5_half(vn[0], vn[1], vn[2], vn[3], vn[4]);
    }
  };
  // node quantities: U_k (unscaled metric), radius, half Jacobian
  double Uk[N], lam[N], hj[N], cg[N];
  double lmax = 0;
#pragma unroll
  for (int m = 0; m < N; ++m) {
    const long im = id + (m - R + 1) * st;
    const double k0 = b.met[im + (3 * DIR) * sz], k1 = b.met[im + (3 * DIR + 1) * sz],
                 k2 = b.met[im + (3 * DIR + 2) * sz];
    Uk[m] = (b.cv[im + sz] * k0 + b.cv[im + 2 * sz] * k1 + b.cv[im + 3 * sz] * k2) / b.cv[im];
    cg[m] = sqrt(k0 * k0 + k1 * k1 + k2 * k2) * x.snd[im];
    lam[m] = fabs(Uk[m]) + cg[m];
    lmax = fmax(lmax, lam[m]);
    hj[m] = 0.5 * b.jac[im];
  }
  double fs[NS]; // species fluxes, limited below before they are stored
  if constexpr (!CH) {
    const double e0 = sc.weno_eps, e1 = e0 * sc.v_ref * sc.v_ref, e4 = e1 * sc.v_ref * sc.v_ref;
    for (int l = 0; l < NV; ++l) {
      double vp[N - 1], vn[N - 1];
#pragma unroll
      for (int m = 0; m < N; ++m) {
        const long im = id + (m - R + 1) * st;
        const double ql = b.cv[im + l * sz];
        const double sp = hj[m] * (Uk[m] + lam[m]), sm = hj[m] * (Uk[m] - lam[m]);
        double fp = sp * ql, fm = sm * ql;
        if (l >= 1 && l <= 4) {
          const double pk = b.bv[im + 4 * sz] * hj[m];
          const double kl = l == 4 ? Uk[m] : b.met[im + (3 * DIR + l - 1) * sz];
          fp = fma(sp, ql, pk * kl);
          fm = fma(sm, ql, pk * kl);
        }
        if (m < N - 1) vp[m] = fp;
        if (m > 0) vn[N - 1 - m] = fm;
      }
      const double f = weno(vp, vn, l == 0 || l >= 5 ? e0 : (l == 4 ? e4 : e1));
      if (l < 5) F[id + l * sz] = f;
      else fs[l - 5] = f;
    }
  } else {
    const ChemMech &mc = *x.mech;
    // Roe state of the two interface nodes (WENO.cu:1146-1178)
    const long ia = id, ib = id + st;
    const double ra = b.cv[ia], rb = b.cv[ib];
    const double sab = sqrt(ra * rb);
    const double rlc = 1 / (ra + sab), rrc = 1 / (sab + rb);
    const double um = rlc * b.cv[ia + sz] + rrc * b.cv[ib + sz];
    const double vm = rlc * b.cv[ia + 2 * sz] + rrc * b.cv[ib + 2 * sz];
    const double wm = rlc * b.cv[ia + 3 * sz] + rrc * b.cv[ib + 3 * sz];
    const double pa = b.bv[ia + 4 * sz], pb = b.bv[ib + 4 * sz];
    double svm[NS], al[NS];
    double Rm = 0;
#pragma unroll
    for (int l = 0; l < NS; ++l) {
      svm[l] = rlc * b.cv[ia + (5 + l) * sz] + rrc * b.cv[ib + (5 + l) * sz];
      Rm += svm[l] * (kRu * mc.imw[l]);
    }
    const double T = (rlc * pa + rrc * pb) / Rm;
    double hh[NS], cpl[NS];
    chemdev::h_cp<NS>(mc, T, hh, cpl);
    double cp = 0;
#pragma unroll
    for (int l = 0; l < NS; ++l) cp += svm[l] * cpl[l];
    const double gam = cp / (cp - Rm);
    const double cm = sqrt(gam * Rm * T);
    const double gm1 = gam - 1;
    double kx = 0.5 * (b.met[ia + (3 * DIR) * sz] + b.met[ib + (3 * DIR) * sz]);
    double ky = 0.5 * (b.met[ia + (3 * DIR + 1) * sz] + b.met[ib + (3 * DIR + 1) * sz]);
    double kz = 0.5 * (b.met[ia + (3 * DIR + 2) * sz] + b.met[ib + (3 * DIR + 2) * sz]);
    {
      const double r = 1.0 / sqrt(kx * kx + ky * ky + kz * kz);
      kx *= r; ky *= r; kz *= r;
    }
    double nx = 0, ny = 0, nz = 0;
    if (fabs(kx) > 0.8) {
      nx = -kz; nz = kx;
    } else {
      ny = kz; nz = -ky;
    }
    {
      const double r = 1.0 / sqrt(nx * nx + ny * ny + nz * nz);
      nx *= r; ny *= r; nz *= r;
    }
    double qx = ky * nz - kz * ny, qy = kz * nx - kx * nz, qz = kx * ny - ky * nx;
    {
      const double r = 1.0 / sqrt(qx * qx + qy * qy + qz * qz);
      qx *= r; qy *= r; qz *= r;
    }
    const double Ub = kx * um + ky * vm + kz * wm;
    const double Un = nx * um + ny * vm + nz * wm, Uq = qx * um + qy * vm + qz * wm;
    const double ic2 = 1.0 / (cm * cm), hic2 = 0.5 * ic2, a = 0.5 * gm1 * (um * um + vm * vm + wm * wm);
#pragma unroll
    for (int l = 0; l < NS; ++l) al[l] = (gam * (kRu * mc.imw[l]) * T - gm1 * hh[l]) * ic2;
    // flow fields: rows 0 and 4 are A +- B, row 3 is x0 - 2A (the L of WENO.cu:1253-1291 in factored form);
    // the species terms enter rows 0/4 with +1/2 and row 3 with -1 times sum_l alpha_l x_(5+l)
    double vp[5][N - 1], vn[5][N - 1];
#pragma unroll
    for (int m = 0; m < N; ++m) {
      const long im = id + (m - R + 1) * st;
      double q[5];
#pragma unroll
      for (int l = 0; l < 5; ++l) q[l] = b.cv[im + l * sz];
      double sb = 0;
#pragma unroll
      for (int l = 0; l < NS; ++l) sb = fma(al[l], b.cv[im + (5 + l) * sz], sb);
      const double p = b.bv[im + 4 * sz];
      const double k0 = b.met[im + (3 * DIR) * sz], k1 = b.met[im + (3 * DIR + 1) * sz],
                   k2 = b.met[im + (3 * DIR + 2) * sz];
      const double u = Uk[m];
      const double fl[5] = {q[0] * u, fma(q[1], u, p * k0), fma(q[2], u, p * k1), fma(q[3], u, p * k2),
                            (q[4] + p) * u};
      double LF[5], LQ[5];
      auto project = [&](const double *v, double s5, double *w) {
        const double A = hic2 * (a * v[0] - gm1 * (um * v[1] + vm * v[2] + wm * v[3]) + gm1 * v[4]);
        const double B = hic2 * cm * (Ub * v[0] - (kx * v[1] + ky * v[2] + kz * v[3]));
        w[0] = A + B + 0.5 * s5;
        w[1] = nx * v[1] + ny * v[2] + nz * v[3] - Un * v[0];
        w[2] = qx * v[1] + qy * v[2] + qz * v[3] - Uq * v[0];
        w[3] = v[0] - 2 * A - s5;
        w[4] = A - B + 0.5 * s5;
      };
      project(fl, sb * u, LF);
      project(q, sb, LQ);
#pragma unroll
      for (int l = 0; l < 5; ++l) {
        if (m < N - 1) vp[l][m] = hj[m] * (LF[l] + lmax * LQ[l]);
        if (m > 0) vn[l][N - 1 - m] = hj[m] * (LF[l] - lmax * LQ[l]);
      }
    }
    const double eps = sc.weno_eps;
    double fch[5];
#pragma unroll
    for (int l = 0; l < 5; ++l) fch[l] = weno(vp[l], vn[l], eps);
    const double s1 = fch[0] + fch[3] + fch[4], s3 = fch[0] - fch[4];
    double sa = 0;
    for (int l = 0; l < NS; ++l) {
      double wp[N - 1], wn[N - 1];
#pragma unroll
      for (int m = 0; m < N; ++m) {
        const long im = id + (m - R + 1) * st;
        const double LQ = b.cv[im + (5 + l) * sz] - svm[l] * b.cv[im];
        const double LF = Uk[m] * LQ;
        if (m < N - 1) wp[m] = hj[m] * (LF + lmax * LQ);
        if (m > 0) wn[N - 1 - m] = hj[m] * (LF - lmax * LQ);
      }
      const double f = weno(wp, wn, eps);
      fs[l] = svm[l] * s1 + f;
      sa += al[l] * f;
    }
    F[id] = s1;
    F[id + sz] = um * s1 - kx * cm * s3 + nx * fch[1] + qx * fch[2];
    F[id + 2 * sz] = vm * s1 - ky * cm * s3 + ny * fch[1] + qy * fch[2];
    F[id + 3 * sz] = wm * s1 - kz * cm * s3 + nz * fch[1] + qz * fch[2];
    const double H = rlc * (b.cv[ia + 4 * sz] + pa) + rrc * (b.cv[ib + 4 * sz] + pb);
    F[id + 4 * sz] = H * s1 - Ub * cm * s3 + Un * fch[1] + Uq * fch[2] - cm * cm / gm1 * fch[3] - sa * cm * cm / gm1;
  }
  // positivity limiter of the species fluxes (WENO.cu:991-1087 component-wise, 1471-1540 characteristic): where
  // alpha J rho Y_L / 2 - dt f or alpha J rho Y_R / 2 + dt f would be negative (alpha = 1/3 in 3D, 1/2 in 2D),
  // f is blended towards a first-order split flux of the two interface nodes, from the first offending
  // species on; that flux splits with |U_k| + c|k| (component-wise) or U_k + c|k| (characteristic)
  if (sc.positive_preserving) {
    const double dt = x.dt[0];
    const double alpha = b.dim == 3 ? 1.0 / 3.0 : 0.5;
    const long ia = id, ib = id + st;
    const double jl = 2 * hj[R - 1], jr = 2 * hj[R];
    const double aL = 0.5 * alpha * jl, aR = 0.5 * alpha * jr;
    int start = NS;
    for (int l = NS - 1; l >= 0; --l) {
      const double up = aL * b.cv[ia + (5 + l) * sz] - dt * fs[l];
      const double um = aR * b.cv[ib + (5 + l) * sz] + dt * fs[l];
      if (up < 0 || um < 0) start = l;
    }
    const double sl = CH ? cg[R - 1] : lam[R - 1];
    const double sr = CH ? cg[R] : lam[R];
    for (int l = start; l < NS; ++l) {
      const double yl = b.cv[ia + (5 + l) * sz], yr = b.cv[ib + (5 + l) * sz];
      const double fp = 0.5 * jl * (Uk[R - 1] + sl) * yl, fm = 0.5 * jr * (Uk[R] - sr) * yr;
      const double f1 = 0.5 * (fp + Uk[R - 1] * jl * yl) + 0.5 * (fm - Uk[R] * jr * yr);
      double tp = 1, tm = 1;
      const double up = aL * yl - dt * fs[l];
      if (up < 0) {
        const double up1 = aL * yl - dt * f1;
        if (fabs(up - up1) > 1e-20 * sc.rho_ref) tp = fmin(fmax((0 - up1) / (up - up1), 0.0), 1.0);
      }
      const double um = aR * yr + dt * fs[l];
      if (um < 0) {
        const double um1 = aR * yr + dt * f1;
        if (fabs(um - um1) > 1e-20 * sc.rho_ref) tm = fmin(fmax((0 - um1) / (um - um1), 0.0), 1.0);
      }
      fs[l] = fmin(tp, tm) * (fs[l] - f1) + f1;
    }
  }
#pragma unroll
  for (int l = 0; l < NS; ++l) F[id + (5 + l) * sz] = fs[l];
}

template<int DIR>
void dispatch_fs(const DevBlock &b, const MixArrays &x, const Scheme &sc, dim3 grid, double *F, hipStream_t s) {
  const dim3 tpb(64, 4, 1);
  switch (sc.inviscid_scheme) {
    case 51: k_mix_fsweno<DIR, 3, false><<<grid, tpb, 0, s>>>(b, x, sc, F); break;
    case 52: k_mix_fsweno<DIR, 3, true><<<grid, tpb, 0, s>>>(b, x, sc, F); break;
    case 71: k_mix_fsweno<DIR, 4, false><<<grid, tpb, 0, s>>>(b, x, sc, F); break;
    default: k_mix_fsweno<DIR, 4, true><<<grid, tpb, 0, s>>>(b, x, sc, F); break;
  }
}

template<int DIR>
void dispatch_rec(const DevBlock &b, const MixArrays &x, const Scheme &sc, dim3 grid, double *F, hipStream_t s) {
  if (sc.inviscid_scheme > 50) return dispatch_fs<DIR>(b, x, sc, grid, F, s);
  switch (sc.reconstruction) {
    case 1: dispatch_rs<DIR, 1>(b, x, sc, grid, F, s); break;
    case 2: dispatch_rs<DIR, 2>(b, x, sc, grid, F, s); break;
    case 3: dispatch_rs<DIR, 3>(b, x, sc, grid, F, s); break;
    default: dispatch_rs<DIR, 4>(b, x, sc, grid, F, s); break;
  }
}

} // namespace

hipError_t launch_mix_cv_from_bv_all(const DevBlock &b, const MixArrays &x, hipStream_t s) {
  if (x.ns != NS) return hipErrorNotSupported;
  k_mix_cv_from_bv<<<(unsigned)((b.sz + 255) / 256), 256, 0, s>>>(b, x);
  return hipGetLastError();
}

hipError_t launch_mix_inner_exchange(const DevBlock &b, const MixArrays &x, const DevBlock &t, const MixArrays &y,
                                     const DevInner &f, hipStream_t s) {
  if (x.ns != NS || y.ns != NS) return hipErrorNotSupported;
  const long n = (long)f.n_point[0] * f.n_point[1] * f.n_point[2];
  k_mix_inner_exchange<<<(unsigned)((n + 255) / 256), 256, 0, s>>>(b, x, t, y, f);
  return hipGetLastError();
}

hipError_t launch_mix_boundary(const DevBlock &b, const MixArrays &x, const DevFace &f, hipStream_t s) {
  if (x.ns != NS) return hipErrorNotSupported;
  const long n = (long)(f.end[0] - f.start[0] + 1) * (f.end[1] - f.start[1] + 1) * (f.end[2] - f.start[2] + 1) * b.g;
  k_mix_boundary<<<(unsigned)((n + 255) / 256), 256, 0, s>>>(b, x, f);
  return hipGetLastError();
}

hipError_t launch_mix_entropy_fix(const DevBlock &b, const MixArrays &x, const Scheme &sc, hipStream_t s) {
  const dim3 grid = chunk_grid(b.mx + 2, b.my + 2, b.dim == 3 ? b.mz + 2 : 1);
  k_mix_entropy_fix<<<grid, dim3(64, 4, 1), 0, s>>>(b, x, sc);
  return hipGetLastError();
}

hipError_t launch_mix_inviscid_flux(const DevBlock &b, const MixArrays &x, const Scheme &sc, int dir, hipStream_t s) {
  if (x.ns != NS) return hipErrorNotSupported;
  if (dir == 0) dispatch_rec<0>(b, x, sc, chunk_grid(b.mx + 1, b.my, b.mz), b.fx, s);
  else if (dir == 1) dispatch_rec<1>(b, x, sc, chunk_grid(b.mx, b.my + 1, b.mz), b.fy, s);
  else dispatch_rec<2>(b, x, sc, chunk_grid(b.mx, b.my, b.mz + 1), b.fz, s);
  return hipGetLastError();
}

// Wu's splitting reaction sub-step (wu_reaction_step, StrangSplitting.cu:373-421; wu_zeroDReaction, 43-225):
// per interior cell, n_sub = dt / tau steps of dt / (n_sub + 1) and a last one to dt, tau the cell's reaction
// time scale from its previous sub-step (0.2 dt at the start); each advances the conservative state by the
// chemical source of its start state and the frozen transport residual cn (the dq array), by forward Euler
// (METHOD 0), by forward Euler of the EPI source (I - h J)^-1 omega with J the Jacobian of omega in rho Y at
// frozen rates of progress (METHOD 1, StrangSplitting.cu:157-181), or by three SSP-RK3 stages with that same
// source (METHOD 3), then clips and renormalises the species and solves for T.  n_sub is capped at 10^6.
template<int METHOD>
__global__ __launch_bounds__(256) void k_mix_wu_reaction(DevBlock b, MixArrays x, const double *__restrict__ dtp) {
  int i, j;
  const int k = blockIdx.z;
  if (!chunk_node(b.mx, b.my, i, j)) return;
  const long id = nid(b, i, j, k), sz = b.sz;
  const long iq = ((long)k * b.my + j) * b.mx + i;
  const ChemMech &m = *x.mech;
  const double dt = dtp[0];
  const double iJ = 1.0 / b.jac[id];
  double cn[NV], q[NV], y[NS];
#pragma unroll
  for (int l = 0; l < NV; ++l) {
    cn[l] = b.dq[iq + l * b.n];
    q[l] = b.cv[id + l * sz];
  }
  double t = b.bv[id + 5 * sz], R = 0;
  double ts = x.tscale[iq];
  const double tau = ts < 1e-15 ? 0.2 * dt : ts;
  const double ratio = dt / tau;
  const int n_sub = ratio > 1e6 ? 1000000 : (int)ratio;
  const double h0 = dt / (n_sub + 1);
  for (int n = 0; n <= n_sub; ++n) {
    const double h = n < n_sub ? h0 : dt - h0 * n_sub;
    if (n == n_sub && !(fabs(h) > 1e-12)) break;
    double w[NS], wd[NS];
    if constexpr (METHOD == 1) {
      double a[NS * NS];
      int piv[NS];
      chemdev::source<NS, true>(m, t, q + 5, w, a, wd, 1e-25);
#pragma unroll
      for (int r = 0; r < NS; ++r)
#pragma unroll
        for (int c = 0; c < NS; ++c) a[r * NS + c] = (r == c ? 1.0 : 0.0) - h * a[r * NS + c];
      chemdev::lu<NS>(a, piv);
      chemdev::lu_solve<NS>(a, piv, w);
    } else {
      chemdev::source<NS, false>(m, t, q + 5, w, nullptr, wd);
    }
    if constexpr (METHOD == 3) {
      double qn[NV];
#pragma unroll
      for (int l = 0; l < NV; ++l) qn[l] = q[l];
      for (int st = 0; st < 3; ++st) {
#pragma unroll
        for (int l = 0; l < 5; ++l) q[l] = kRkA[st] * qn[l] + kRkB[st] * q[l] + kRkC[st] * h * iJ * cn[l];
        ts = 1e+6;
#pragma unroll
        for (int l = 0; l < NS; ++l) {
          if (q[5 + l] > 1e-25 && fabs(wd[l]) > 1e-25) ts = fmin(ts, fabs(q[5 + l] / wd[l]));
          q[5 + l] = kRkA[st] * qn[5 + l] + kRkB[st] * q[5 + l] + kRkC[st] * h * (w[l] + cn[5 + l] * iJ);
        }
        t = mix_state(m, q, t, y, R);
      }
    } else {
#pragma unroll
      for (int l = 0; l < 5; ++l) q[l] += cn[l] * h * iJ;
      ts = 1e+6;
#pragma unroll
      for (int l = 0; l < NS; ++l) {
        if (q[5 + l] > 1e-25 && fabs(wd[l]) > 1e-25) ts = fmin(ts, fabs(q[5 + l] / wd[l]));
        q[5 + l] += (w[l] + cn[5 + l] * iJ) * h;
      }
      t = mix_state(m, q, t, y, R);
    }
  }
  x.tscale[iq] = ts;
#pragma unroll
  for (int l = 0; l < NV; ++l) b.cv[id + l * sz] = q[l];
  const double di = 1.0 / q[0];
  b.bv[id] = q[0];
  b.bv[id + sz] = q[1] * di;
  b.bv[id + 2 * sz] = q[2] * di;
  b.bv[id + 3 * sz] = q[3] * di;
  b.bv[id + 4 * sz] = q[0] * t * R;
  b.bv[id + 5 * sz] = t;
#pragma unroll
  for (int l = 0; l < NS; ++l) x.sv[id + l * sz] = y[l];
  double hh[NS], cp[NS], cpt = 0;
  chemdev::h_cp<NS>(m, t, hh, cp);
#pragma unroll
  for (int l = 0; l < NS; ++l) cpt += cp[l] * y[l];
  x.snd[id] = sqrt(cpt / (cpt - R) * R * t);
}

hipError_t launch_mix_transport(const DevBlock &b, const MixArrays &x, hipStream_t s) {
  if (x.ns != NS) return hipErrorNotSupported;
  k_mix_transport<<<(unsigned)((b.sz + 255) / 256), 256, 0, s>>>(b, x);
  return hipGetLastError();
}

hipError_t launch_mix_viscous(const DevBlock &b, const MixArrays &x, const Scheme &sc, hipStream_t s) {
  if (x.ns != NS) return hipErrorNotSupported;
  k_mix_transport<<<(unsigned)((b.sz + 255) / 256), 256, 0, s>>>(b, x);
  const dim3 tpb(64, 4, 1);
  if (sc.viscous_order == 8) {
    k_mix_viscous_node<8><<<chunk_grid(b.mx, b.my, b.mz), tpb, 0, s>>>(b, x);
    return hipGetLastError();
  }
  if (sc.viscous_order != 2) return hipErrorNotSupported;
  k_mix_viscous_face<0><<<chunk_grid(b.mx + 1, b.my, b.mz), tpb, 0, s>>>(b, x, sc);
  k_mix_viscous_face<1><<<chunk_grid(b.mx, b.my + 1, b.mz), tpb, 0, s>>>(b, x, sc);
  if (b.dim == 3) k_mix_viscous_face<2><<<chunk_grid(b.mx, b.my, b.mz + 1), tpb, 0, s>>>(b, x, sc);
  return hipGetLastError();
}

hipError_t launch_mix_wu_reaction(const DevBlock &b, const MixArrays &x, const double *dt_dev, int method,
                                  hipStream_t s) {
  if (x.ns != NS) return hipErrorNotSupported;
  if (method == 3) k_mix_wu_reaction<3><<<chunk_grid(b.mx, b.my, b.mz), dim3(64, 4, 1), 0, s>>>(b, x, dt_dev);
  else if (method == 1) k_mix_wu_reaction<1><<<chunk_grid(b.mx, b.my, b.mz), dim3(64, 4, 1), 0, s>>>(b, x, dt_dev);
  else k_mix_wu_reaction<0><<<chunk_grid(b.mx, b.my, b.mz), dim3(64, 4, 1), 0, s>>>(b, x, dt_dev);
  return hipGetLastError();
}

hipError_t launch_mix_wu_update(const DevBlock &b, const MixArrays &x, int rk, const double *dt_dev, hipStream_t s) {
  if (x.ns != NS) return hipErrorNotSupported;
  k_mix_update<false><<<chunk_grid(b.mx, b.my, b.mz), dim3(64, 4, 1), 0, s>>>(b, x, rk, dt_dev, b.dq, 0.5);
  return hipGetLastError();
}

hipError_t launch_mix_update(const DevBlock &b, const MixArrays &x, int rk, const double *dt_dev, bool rhs_only,
                             hipStream_t s) {
  if (x.ns != NS) return hipErrorNotSupported;
  if (rhs_only) k_mix_update<true><<<chunk_grid(b.mx, b.my, b.mz), dim3(64, 4, 1), 0, s>>>(b, x, rk, dt_dev, nullptr, 1.0);
  else k_mix_update<false><<<chunk_grid(b.mx, b.my, b.mz), dim3(64, 4, 1), 0, s>>>(b, x, rk, dt_dev, nullptr, 1.0);
  return hipGetLastError();
}

hipError_t launch_mix_local_dt(const DevBlock &b, const MixArrays &x, const Scheme &sc, double *part, hipStream_t s) {
  k_mix_local_dt<<<node_grid(b), dim3(64, 4, 1), 0, s>>>(b, x, sc, part);
  return hipGetLastError();
}

hipError_t launch_mix_chem_gather(const DevBlock &b, const MixArrays &x, const ChemBatch &c, hipStream_t s) {
  if (x.ns != NS || c.n != b.n) return hipErrorInvalidValue;
  k_mix_chem_gather<<<chunk_grid(b.mx, b.my, b.mz), dim3(64, 4, 1), 0, s>>>(b, x, c);
  return hipGetLastError();
}

hipError_t launch_mix_chem_scatter(const DevBlock &b, const MixArrays &x, const ChemBatch &c, hipStream_t s) {
  if (x.ns != NS || c.n != b.n) return hipErrorInvalidValue;
  k_mix_chem_scatter<<<chunk_grid(b.mx, b.my, b.mz), dim3(64, 4, 1), 0, s>>>(b, x, c);
  return hipGetLastError();
}

} // namespace corefl
