// Step monitor (residual RMS with a NaN flag) and limit_flow for gfx950, FP64.  The residual is a two-level
// reduction with a fixed grid and a fixed summation order, so a monitored step gives the same bits every time.
#include "monitor.hpp"

#include <cmath>

#include "chem_device.hpp"

namespace corefl {

namespace {

constexpr double kGamma = 1.4, kRu = 8314.462618, kMwAir = 28.9647;
constexpr int kResThreads = 256, kResMaxBlocks = 1024;

__device__ __forceinline__ long nid(const DevBlock &b, int i, int j, int k) {
  return (long)(k + b.g) * b.d1 + (long)(j + b.g) * b.d2 + (i + b.g);
}

__device__ __forceinline__ long node_of(const DevBlock &b, long c) {
  const int i = (int)(c % b.mx);
  const long r = c / b.mx;
  return nid(b, i, (int)(r % b.my), (int)(r / b.my));
}

__global__ __launch_bounds__(256) void k_store_last(DevBlock b, double *last) {
  const long c = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (c >= b.n) return;
  const long id = node_of(b, c), sz = b.sz;
  const double u = b.bv[id + sz], v = b.bv[id + 2 * sz], w = b.bv[id + 3 * sz];
  last[c] = b.bv[id];
  last[c + b.n] = sqrt(u * u + v * v + w * w);
  last[c + 2 * b.n] = b.bv[id + 4 * sz];
  last[c + 3 * b.n] = b.bv[id + 5 * sz];
}

// Sum of 4 values over the workgroup (wave64 shuffles, then one LDS row per wave); result valid in thread 0.
__device__ __forceinline__ void block_sum4(double v[4]) {
  __shared__ double red[4][kResThreads / 64];
#pragma unroll
  for (int l = 0; l < 4; ++l)
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) v[l] += __shfl_down(v[l], o, 64);
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  if (lane == 0)
#pragma unroll
    for (int l = 0; l < 4; ++l) red[l][wave] = v[l];
  __syncthreads();
  if (threadIdx.x == 0)
#pragma unroll
    for (int l = 0; l < 4; ++l) {
      double s = 0;
      for (int q = 0; q < kResThreads / 64; ++q) s += red[l][q];
      v[l] = s;
    }
}

__global__ __launch_bounds__(kResThreads) void k_residual_partial(DevBlock b, const double *last, double *part) {
  double acc[4] = {0, 0, 0, 0};
  const long sz = b.sz;
  for (long c = (long)blockIdx.x * blockDim.x + threadIdx.x; c < b.n; c += (long)gridDim.x * blockDim.x) {
    const long id = node_of(b, c);
    const double u = b.bv[id + sz], v = b.bv[id + 2 * sz], w = b.bv[id + 3 * sz];
    const double d0 = b.bv[id] - last[c];
    const double d1 = sqrt(u * u + v * v + w * w) - last[c + b.n];
    const double d2 = b.bv[id + 4 * sz] - last[c + 2 * b.n];
    const double d3 = b.bv[id + 5 * sz] - last[c + 3 * b.n];
    acc[0] += d0 * d0;
    acc[1] += d1 * d1;
    acc[2] += d2 * d2;
    acc[3] += d3 * d3;
  }
  block_sum4(acc);
  if (threadIdx.x == 0)
#pragma unroll
    for (int l = 0; l < 4; ++l) part[4 * blockIdx.x + l] = acc[l];
}

__global__ __launch_bounds__(kResThreads) void k_residual_finish(const double *part, long n_part, double n_total,
                                                                 MonitorState *st) {
  double acc[4] = {0, 0, 0, 0};
  for (long w = threadIdx.x; w < n_part; w += blockDim.x)
#pragma unroll
    for (int l = 0; l < 4; ++l) acc[l] += part[4 * w + l];
  block_sum4(acc);
  if (threadIdx.x != 0) return;
  double res[4];
  for (int l = 0; l < 4; ++l) res[l] = sqrt(acc[l] / n_total);
  if (!st->have_scale) {
    for (int l = 0; l < 4; ++l) st->scale[l] = res[l] < 1e-20 ? 1e-20 : res[l];
    st->have_scale = 1;
  }
  int nan = 0;
  for (int l = 0; l < 4; ++l) {
    res[l] /= st->scale[l];
    st->res[l] = res[l];
    if (isnan(fabs(res[l]))) nan = 1;
  }
  double err_max = res[0];
  for (int l = 1; l < 4; ++l)
    if (fabs(res[l]) > err_max) err_max = res[l];
  st->err_max = err_max;
  st->nan = nan;
  st->count += 1;
}

__device__ __forceinline__ bool unphysical(const DevBlock &b, long id, const FlowLimits &L) {
  bool bad = false;
#pragma unroll
  for (int l = 0; l < 5; ++l) {
    const double v = b.bv[id + l * b.sz];
    bad |= isnan(v) || v < L.ll[l] || v > L.ul[l];
  }
  return bad;
}

__global__ __launch_bounds__(256) void k_limit_flag(DevBlock b, FlowLimits L, unsigned char *flag) {
  const long c = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (c >= b.n) return;
  flag[c] = unphysical(b, node_of(b, c), L) ? 1 : 0;
}

// The repair reads only unflagged neighbours, which no thread writes: the result does not depend on the
// order in which the flagged nodes are repaired.
template<bool MIX>
__global__ __launch_bounds__(256) void k_limit_repair(DevBlock b, MixArrays x, FlowLimits L, const unsigned char *flag) {
  const long c = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (c >= b.n || !flag[c]) return;
  const int i = (int)(c % b.mx), j = (int)((c / b.mx) % b.my), k = (int)(c / ((long)b.mx * b.my));
  const long id = nid(b, i, j, k), sz = b.sz;
  double sum[5] = {0, 0, 0, 0, 0};
  int kn = 0;
  for (int ka = -1; ka < 2; ++ka) {
    const int k1 = k + ka;
    if (k1 < 0 || k1 >= b.mz) continue;
    for (int ja = -1; ja < 2; ++ja) {
      const int j1 = j + ja;
      if (j1 < 0 || j1 >= b.my) continue;
      for (int ia = -1; ia < 2; ++ia) {
        const int i1 = i + ia;
        if (i1 < 0 || i1 >= b.mx) continue;
        if (flag[((long)k1 * b.my + j1) * b.mx + i1]) continue;
        const long jd = nid(b, i1, j1, k1);
#pragma unroll
        for (int l = 0; l < 5; ++l) sum[l] += b.bv[jd + l * sz];
        ++kn;
      }
    }
  }
  double q[5];
  if (kn > 0) {
    const double inv = 1.0 / kn;
#pragma unroll
    for (int l = 0; l < 5; ++l) q[l] = sum[l] * inv;
  } else {
#pragma unroll
    for (int l = 0; l < 5; ++l) q[l] = fmin(fmax(b.bv[id + l * sz], L.ll[l]), L.ul[l]);
  }
#pragma unroll
  for (int l = 0; l < 5; ++l) b.bv[id + l * sz] = q[l];
  const double rho = q[0], u = q[1], v = q[2], w = q[3], p = q[4];
  if constexpr (!MIX) {
    const double T = p * kMwAir / (rho * kRu);
    b.bv[id + 5 * sz] = T;
    b.cv[id] = rho;
    b.cv[id + sz] = rho * u;
    b.cv[id + 2 * sz] = rho * v;
    b.cv[id + 3 * sz] = rho * w;
    b.cv[id + 4 * sz] = 0.5 * (u * u + v * v + w * w) * rho + p / (kGamma - 1);
  } else {
    constexpr int NS = 9;
    const ChemMech &m = *x.mech;
    double imw = 0, y[NS];
#pragma unroll
    for (int l = 0; l < NS; ++l) {
      y[l] = x.sv[id + l * sz];
      imw += y[l] * m.imw[l];
    }
    const double T = p / (rho * kRu * imw);
    b.bv[id + 5 * sz] = T;
    double h[NS], cp[NS];
    chemdev::h_cp<NS>(m, T, h, cp);
    double E = 0.5 * (u * u + v * v + w * w), cpt = 0;
#pragma unroll
    for (int l = 0; l < NS; ++l) {
      E += h[l] * y[l];
      cpt += y[l] * cp[l];
      b.cv[id + (5 + l) * sz] = rho * y[l];
    }
    b.cv[id] = rho;
    b.cv[id + sz] = rho * u;
    b.cv[id + 2 * sz] = rho * v;
    b.cv[id + 3 * sz] = rho * w;
    b.cv[id + 4 * sz] = E * rho - p;
    const double R = kRu * imw;
    x.snd[id] = sqrt(cpt / (cpt - R) * R * T);
  }
}

unsigned grid_of(long n) { return (unsigned)((n + 255) / 256); }

} // namespace

hipError_t launch_store_last(const DevBlock &b, double *last, hipStream_t s) {
  if (b.n == 0) return hipSuccess;
  k_store_last<<<grid_of(b.n), 256, 0, s>>>(b, last);
  return hipGetLastError();
}

int residual_partial_count(const DevBlock &b) {
  const long g = (b.n + kResThreads - 1) / kResThreads;
  return (int)(g < kResMaxBlocks ? (g < 1 ? 1 : g) : kResMaxBlocks);
}

hipError_t launch_residual_partial(const DevBlock &b, const double *last, double *part, hipStream_t s) {
  k_residual_partial<<<residual_partial_count(b), kResThreads, 0, s>>>(b, last, part);
  return hipGetLastError();
}

hipError_t launch_residual_finish(const double *part, long n_part, double n_total, MonitorState *st, hipStream_t s) {
  k_residual_finish<<<1, kResThreads, 0, s>>>(part, n_part, n_total, st);
  return hipGetLastError();
}

hipError_t launch_limit_flow(const DevBlock &b, const MixArrays &x, const FlowLimits &lim, unsigned char *flag,
                             hipStream_t s) {
  if (b.n == 0) return hipSuccess;
  if (x.ns != 0 && x.ns != 9) return hipErrorNotSupported;
  k_limit_flag<<<grid_of(b.n), 256, 0, s>>>(b, lim, flag);
  if (x.ns) k_limit_repair<true><<<grid_of(b.n), 256, 0, s>>>(b, x, lim, flag);
  else k_limit_repair<false><<<grid_of(b.n), 256, 0, s>>>(b, x, lim, flag);
  return hipGetLastError();
}

} // namespace corefl
