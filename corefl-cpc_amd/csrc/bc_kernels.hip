// Physical boundary faces beyond outflow and periodic, and the sponge layer, for gfx950 in FP64: inflow
// (constant, subsonic, and the mixing layer's tanh profile with white noise) and symmetry faces for air and
// mixtures; the sponge layer's running mean.
#include "bc.hpp"

#include <cmath>

#include "mix_state.hpp"

namespace corefl {

namespace {

constexpr double kGamma = 1.4, kRu = 8314.462618, kMwAir = 28.9647, kRair = kRu / kMwAir;
constexpr int kMaxNs = 9;

__device__ __forceinline__ long nid(const DevBlock &b, int i, int j, int k) {
  return (long)(k + b.g) * b.d1 + (long)(j + b.g) * b.d2 + (i + b.g);
}

template<bool MIX> __device__ __forceinline__ void cons(const DevBlock &b, const MixArrays &x, long id) {
  if constexpr (MIX) {
    mixdev::mix_cons(b, x, id);
  } else {
    const long sz = b.sz;
    const double rho = b.bv[id], u = b.bv[id + sz], v = b.bv[id + 2 * sz], w = b.bv[id + 3 * sz];
    b.cv[id] = rho;
    b.cv[id + sz] = rho * u;
    b.cv[id + 2 * sz] = rho * v;
    b.cv[id + 3 * sz] = rho * w;
    b.cv[id + 4 * sz] = 0.5 * (u * u + v * v + w * w) * rho + b.bv[id + 4 * sz] / (kGamma - 1);
  }
}

// apply_inflow (BoundCond.cu:2190-2705) for inflow_type 0 (constant, optionally subsonic) and 2 (mixing layer),
// fluctuation_type 0 or 1 (white noise).  One thread per face node: the boundary node, then its ghost layers.
template<bool MIX>
__global__ __launch_bounds__(256) void k_inflow(DevBlock b, MixArrays x, DevInflow f) {
  const long n0 = f.end[0] - f.start[0] + 1, n1 = f.end[1] - f.start[1] + 1, n2 = f.end[2] - f.start[2] + 1;
  const long npt = n0 * n1 * n2;
  const long node = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (node >= npt) return;
  const int i = f.start[0] + (int)(node % n0), j = f.start[1] + (int)((node / n0) % n1),
            k = f.start[2] + (int)(node / (n0 * n1));
  const int ns = MIX ? x.ns : 0;
  const long sz = b.sz, id = nid(b, i, j, k);
  const long st[3] = {1, b.d2, b.d1};
  const long dstep = f.direction * st[f.face]; // one node outwards along the normal
  double rho = f.prof[node], u = f.prof[npt + node], v = f.prof[2 * npt + node], w = f.prof[3 * npt + node];
  double p = f.prof[4 * npt + node], T = f.prof[5 * npt + node];
  double Y[kMaxNs];
  for (int l = 0; l < ns; ++l) Y[l] = f.prof[(6 + l) * npt + node];
  if (f.sub_node) {
    double nx = b.met[id + (3 * f.face) * sz], ny = b.met[id + (3 * f.face + 1) * sz],
           nz = b.met[id + (3 * f.face + 2) * sz];
    const double gi = f.direction / sqrt(nx * nx + ny * ny + nz * nz);
    nx *= gi;
    ny *= gi;
    nz *= gi;
    const double c0 = MIX ? x.snd[id] : sqrt(kGamma * kRair * b.bv[id + 5 * sz]);
    const double Mn = (b.bv[id + sz] * nx + b.bv[id + 2 * sz] * ny + b.bv[id + 3 * sz] * nz) / c0;
    if (!(Mn < -1.0)) {
      const long i1 = id - dstep;
      const double c1 = MIX ? x.snd[i1] : sqrt(kGamma * kRair * b.bv[i1 + 5 * sz]);
      const double rho1 = b.bv[i1], p1 = b.bv[i1 + 4 * sz];
      const double pin = p;
      p = 0.5 * (p1 + pin - rho1 * c1 * ((u - b.bv[i1 + sz]) * nx + (v - b.bv[i1 + 2 * sz]) * ny +
                                         (w - b.bv[i1 + 3 * sz]) * nz));
      const double dpr = (p - pin) / (rho1 * c1);
      rho = rho + (p - pin) / (c1 * c1);
      u = u - dpr * nx;
      v = v - dpr * ny;
      w = w - dpr * nz;
      T = MIX ? p * f.mw / (rho * kRu) : p / (rho * kRair);
    }
  } else if (f.fluct) {
    const unsigned long long call = *f.calls;
    const double r = f.rms[node];
    u += noise_normal(f.seed, call, node, 0) * r;
    if (f.fluct == 2) {
      v += noise_normal(f.seed, call, node, 1) * r;
      w += noise_normal(f.seed, call, node, 2) * r;
    }
  }
  b.bv[id] = rho;
  b.bv[id + sz] = u;
  b.bv[id + 2 * sz] = v;
  b.bv[id + 3 * sz] = w;
  b.bv[id + 4 * sz] = p;
  b.bv[id + 5 * sz] = T;
  for (int l = 0; l < ns; ++l) x.sv[id + l * sz] = Y[l];
  cons<MIX>(b, x, id);
  for (int g = 1; g <= b.g; ++g) {
    const long ig = id + g * dstep;
    for (int l = 0; l < ns; ++l) x.sv[ig + l * sz] = Y[l];
    if (f.sub_ghost) {
      const long ii = id - g * dstep;
      double rho2 = 2 * rho - b.bv[ii], p2 = 2 * p - b.bv[ii + 4 * sz];
      if (p2 <= 0.1 * p) p2 = 0.1 * p;
      if (rho2 <= 0.1 * rho) rho2 = 0.1 * rho;
      const double u2 = 2 * u - b.bv[ii + sz], v2 = 2 * v - b.bv[ii + 2 * sz], w2 = 2 * w - b.bv[ii + 3 * sz];
      b.bv[ig] = rho2;
      b.bv[ig + sz] = u2;
      b.bv[ig + 2 * sz] = v2;
      b.bv[ig + 3 * sz] = w2;
      b.bv[ig + 4 * sz] = p2;
      b.bv[ig + 5 * sz] = MIX ? p2 * f.mw / (rho2 * kRu) : p2 / (rho2 * kRair);
    } else {
      b.bv[ig] = rho;
      b.bv[ig + sz] = u;
      b.bv[ig + 2 * sz] = v;
      b.bv[ig + 3 * sz] = w;
      b.bv[ig + 4 * sz] = p;
      b.bv[ig + 5 * sz] = T;
    }
    cons<MIX>(b, x, ig);
  }
}

__global__ void k_count_call(unsigned long long *calls) { calls[0] += 1; }

// apply_symmetry (BoundCond.cu:2002-2063): the boundary node takes the first interior node's state with the normal
// velocity removed (its cv is left as it was, as in the reference); ghost g mirrors interior node g.
template<bool MIX>
__global__ __launch_bounds__(256) void k_symmetry(DevBlock b, MixArrays x, DevFace f) {
  const long n0 = f.end[0] - f.start[0] + 1, n1 = f.end[1] - f.start[1] + 1, n2 = f.end[2] - f.start[2] + 1;
  const long node = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (node >= n0 * n1 * n2) return;
  const int i = f.start[0] + (int)(node % n0), j = f.start[1] + (int)((node / n0) % n1),
            k = f.start[2] + (int)(node / (n0 * n1));
  const int ns = MIX ? x.ns : 0;
  const long sz = b.sz, id = nid(b, i, j, k);
  const long st[3] = {1, b.d2, b.d1};
  const long dstep = f.direction * st[f.face];
  double kx = b.met[id + (3 * f.face) * sz], ky = b.met[id + (3 * f.face + 1) * sz],
         kz = b.met[id + (3 * f.face + 2) * sz];
  const double km = sqrt(kx * kx + ky * ky + kz * kz);
  kx /= km;
  ky /= km;
  kz /= km;
  const long in = id - dstep;
  const double u1 = b.bv[in + sz], v1 = b.bv[in + 2 * sz], w1 = b.bv[in + 3 * sz];
  const double uk = kx * u1 + ky * v1 + kz * w1;
  b.bv[id + sz] = u1 - kx * uk;
  b.bv[id + 2 * sz] = v1 - ky * uk;
  b.bv[id + 3 * sz] = w1 - kz * uk;
  b.bv[id] = b.bv[in];
  b.bv[id + 4 * sz] = b.bv[in + 4 * sz];
  b.bv[id + 5 * sz] = b.bv[in + 5 * sz];
  for (int l = 0; l < ns; ++l) x.sv[id + l * sz] = x.sv[in + l * sz];
  for (int g = 1; g <= b.g; ++g) {
    const long ig = id + g * dstep, ii = id - g * dstep;
    b.bv[ig] = b.bv[ii];
    const double u = b.bv[ii + sz], v = b.bv[ii + 2 * sz], w = b.bv[ii + 3 * sz];
    const double un = kx * u + ky * v + kz * w;
    b.bv[ig + sz] = u - 2 * un * kx;
    b.bv[ig + 2 * sz] = v - 2 * un * ky;
    b.bv[ig + 3 * sz] = w - 2 * un * kz;
    b.bv[ig + 4 * sz] = b.bv[ii + 4 * sz];
    b.bv[ig + 5 * sz] = b.bv[ii + 5 * sz];
    for (int l = 0; l < ns; ++l) x.sv[ig + l * sz] = x.sv[ii + l * sz];
    cons<MIX>(b, x, ig);
  }
}

// Running mean of the conservative variables over the steps taken (the intent of update_sponge_layer_value,
// SpongeLayer.cu:28-46, whose body the reference leaves commented out): mean = (mean n + U) / (n + 1).
__global__ __launch_bounds__(256) void k_sponge_mean(DevBlock b, int nv, double *mean, const unsigned long long *iter) {
  const long c = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (c >= b.n) return;
  const int i = (int)(c % b.mx), j = (int)((c / b.mx) % b.my), k = (int)(c / ((long)b.mx * b.my));
  const long id = nid(b, i, j, k);
  const double n = (double)iter[0], inv = 1.0 / (n + 1);
  for (int l = 0; l < nv; ++l) mean[c + l * b.n] = (mean[c + l * b.n] * n + b.cv[id + l * b.sz]) * inv;
}

unsigned face_nodes(const int *start, const int *end) {
  const long n = (long)(end[0] - start[0] + 1) * (end[1] - start[1] + 1) * (end[2] - start[2] + 1);
  return (unsigned)((n + 255) / 256);
}

} // namespace

hipError_t launch_inflow(const DevBlock &b, const MixArrays &x, const DevInflow &f, hipStream_t s) {
  if (x.ns != 0 && x.ns != kMaxNs) return hipErrorNotSupported;
  if (x.ns) k_inflow<true><<<face_nodes(f.start, f.end), 256, 0, s>>>(b, x, f);
  else k_inflow<false><<<face_nodes(f.start, f.end), 256, 0, s>>>(b, x, f);
  if (f.fluct) k_count_call<<<1, 1, 0, s>>>(f.calls);
  return hipGetLastError();
}

hipError_t launch_symmetry(const DevBlock &b, const MixArrays &x, const DevFace &f, hipStream_t s) {
  if (x.ns != 0 && x.ns != kMaxNs) return hipErrorNotSupported;
  if (x.ns) k_symmetry<true><<<face_nodes(f.start, f.end), 256, 0, s>>>(b, x, f);
  else k_symmetry<false><<<face_nodes(f.start, f.end), 256, 0, s>>>(b, x, f);
  return hipGetLastError();
}

hipError_t launch_sponge_mean(const DevBlock &b, int nv, double *mean, unsigned long long *iter, bool count,
                              hipStream_t s) {
  if (b.n > 0) k_sponge_mean<<<(unsigned)((b.n + 255) / 256), 256, 0, s>>>(b, nv, mean, iter);
  if (count) k_count_call<<<1, 1, 0, s>>>(iter);
  return hipGetLastError();
}

} // namespace corefl
