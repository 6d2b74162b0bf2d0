// Symmetric-conservative metrics (symmetric_conservative_metric = 2, 4, 6; Mesh.cu:10-250) on gfx950 in FP64: the
// 18 coordinate products x_a y - y_a x with their order-P derivatives, the three combinations D123, then the
// Jacobian and the nine metrics from sixth-order derivatives of those, over every node of the ngg-ghost layout.
// A set-up step: one thread per node, three passes, the intermediate fields in device memory.
#include <hip/hip_runtime.h>

#include <string>
#include <vector>

#include "mesh.hpp"

namespace corefl {

namespace {

// First-derivative weights of the degree-P Lagrange interpolant through P + 1 consecutive nodes, for each position
// r = i - s of the evaluation node in the stencil: w[r][l] = L_l'(r) (ggxl::Derivatives, gxl_lib/Math.cuh:6-101).
// Numerators and denominators are integers, exact in double, so each weight is their correctly rounded quotient.
struct LagWeights {
  double w[7][7];
};

LagWeights lag_weights(int P) {
  LagWeights W{};
  for (int r = 0; r <= P; ++r)
    for (int l = 0; l <= P; ++l) {
      double den = 1, num = 0;
      for (int a = 0; a <= P; ++a)
        if (a != l) den *= (double)(l - a);
      for (int m = 0; m <= P; ++m) {
        if (m == l) continue;
        double prod = 1;
        for (int a = 0; a <= P; ++a)
          if (a != l && a != m) prod *= (double)(r - a);
        num += prod;
      }
      W.w[r][l] = num / den;
    }
  return W;
}

struct Grid {
  int n[3], ng;
  long pd2, pd1; // strides of the (ngg + 1)-layer coordinate arrays
  long d2, d1;   // strides of the ngg-layer node arrays
};

// d f / d(index along dir) at the node with padded index id and index c along dir: the centred stencil where it
// fits in [-ng, n - 1 + ng], else the one-sided one at that end.
template<int P>
__device__ __forceinline__ double lagd(const double *f, long id, long stride, int c, int n, int ng,
                                       const LagWeights &W) {
  int s = c - P / 2;
  if (c <= -ng + P / 2) s = -ng;
  else if (c >= n - 1 + ng - P / 2) s = n - 1 + ng - P;
  const double *w = W.w[c - s];
  const double *b = f + id + (long)(s - c) * stride;
  double D = 0;
#pragma unroll
  for (int l = 0; l <= P; ++l) D += w[l] * b[l * stride];
  return D;
}

__device__ __forceinline__ bool node_of(const Grid &g, long t, int *c) {
  const long e0 = g.n[0] + 2L * g.ng, e1 = g.n[1] + 2L * g.ng, e2 = g.n[2] + 2L * g.ng;
  if (t >= e0 * e1 * e2) return false;
  c[0] = (int)(t % e0) - g.ng;
  c[1] = (int)((t / e0) % e1) - g.ng;
  c[2] = (int)(t / (e0 * e1)) - g.ng;
  return true;
}

// compute_xyz123 (Mesh.cu:10-56)
template<int P>
__global__ __launch_bounds__(256) void k_sc_products(Grid g, const double *X, const double *Y, const double *Z,
                                                     double *Q, LagWeights W) {
  int c[3];
  const long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (!node_of(g, t, c)) return;
  const long pid = (c[2] + g.ng + 1L) * g.pd1 + (c[1] + g.ng + 1L) * g.pd2 + (c[0] + g.ng + 1L);
  const long ps[3] = {1, g.pd2, g.pd1};
  double dx[3], dy[3], dz[3];
#pragma unroll
  for (int d = 0; d < 3; ++d) {
    dx[d] = lagd<P>(X, pid, ps[d], c[d], g.n[d], g.ng, W);
    dy[d] = lagd<P>(Y, pid, ps[d], c[d], g.n[d], g.ng, W);
    dz[d] = lagd<P>(Z, pid, ps[d], c[d], g.n[d], g.ng, W);
  }
  const double x = X[pid], y = Y[pid], z = Z[pid];
  const double v[18] = {dx[0] * y - dy[0] * x, dy[1] * x - dx[1] * y, dy[0] * z - dz[0] * y, dz[1] * y - dy[1] * z,
                        dz[0] * x - dx[0] * z, dx[1] * z - dz[1] * x, dx[2] * y - dy[2] * x, dy[0] * x - dx[0] * y,
                        dy[2] * z - dz[2] * y, dz[0] * y - dy[0] * z, dz[2] * x - dx[2] * z, dx[0] * z - dz[0] * x,
                        dx[1] * y - dy[1] * x, dy[2] * x - dx[2] * y, dy[1] * z - dz[1] * y, dz[2] * y - dy[2] * z,
                        dz[1] * x - dx[1] * z, dx[2] * z - dz[2] * x};
  const long sz = g.d1 * (g.n[2] + 2L * g.ng);
#pragma unroll
  for (int q = 0; q < 18; ++q) Q[t + q * sz] = v[q];
}

// compute_D123 (Mesh.cu:58-93)
template<int P>
__global__ __launch_bounds__(256) void k_sc_d123(Grid g, const double *X, const double *Y, const double *Z,
                                                 const double *Q, double *D, LagWeights W) {
  int c[3];
  const long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (!node_of(g, t, c)) return;
  const long pid = (c[2] + g.ng + 1L) * g.pd1 + (c[1] + g.ng + 1L) * g.pd2 + (c[0] + g.ng + 1L);
  const long sz = g.d1 * (g.n[2] + 2L * g.ng);
  const long st[3] = {1, g.d2, g.d1};
  auto der = [&](int q, int d) { return lagd<P>(Q + q * sz, t, st[d], c[d], g.n[d], g.ng, W); };
  const double x = X[pid], y = Y[pid], z = Z[pid];
  D[t + 2 * sz] = (der(0, 1) + der(1, 0)) * z + (der(2, 1) + der(3, 0)) * x + (der(4, 1) + der(5, 0)) * y;
  D[t + sz] = (der(6, 0) + der(7, 2)) * z + (der(8, 0) + der(9, 2)) * x + (der(10, 0) + der(11, 2)) * y;
  D[t] = (der(12, 2) + der(13, 1)) * z + (der(14, 2) + der(15, 1)) * x + (der(16, 2) + der(17, 1)) * y;
}

// The Jacobian and metrics with sixth-order derivatives (Mesh.cu:215-243); bad[0] = 1 where J <= 0.
__global__ __launch_bounds__(256) void k_sc_metric(Grid g, const double *Q, const double *D, double *jac, double *met,
                                                   int *bad, LagWeights W) {
  int c[3];
  const long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (!node_of(g, t, c)) return;
  const long sz = g.d1 * (g.n[2] + 2L * g.ng);
  const long st[3] = {1, g.d2, g.d1};
  auto der = [&](const double *f, int d) { return lagd<6>(f, t, st[d], c[d], g.n[d], g.ng, W); };
  const double J = (der(D + 2 * sz, 2) + der(D + sz, 1) + der(D, 0)) / 6.0;
  if (!(J > 0)) bad[0] = 1;
  jac[t] = J;
  const double h = 0.5 / J;
  auto q = [&](int k) { return Q + k * sz; };
  met[t] = (der(q(14), 2) - der(q(8), 1)) * h;
  met[t + sz] = (der(q(16), 2) - der(q(10), 1)) * h;
  met[t + 2 * sz] = (der(q(12), 2) - der(q(6), 1)) * h;
  met[t + 3 * sz] = (der(q(8), 0) - der(q(2), 2)) * h;
  met[t + 4 * sz] = (der(q(10), 0) - der(q(4), 2)) * h;
  met[t + 5 * sz] = (der(q(6), 0) - der(q(0), 2)) * h;
  met[t + 6 * sz] = (der(q(2), 1) - der(q(14), 0)) * h;
  met[t + 7 * sz] = (der(q(4), 1) - der(q(16), 0)) * h;
  met[t + 8 * sz] = (der(q(0), 1) - der(q(12), 0)) * h;
}

template<int P>
void launch_first_stages(const Grid &g, const double *X, const double *Y, const double *Z, double *Q, double *D,
                         unsigned nb) {
  const LagWeights W = lag_weights(P);
  k_sc_products<P><<<nb, 256>>>(g, X, Y, Z, Q, W);
  k_sc_d123<P><<<nb, 256>>>(g, X, Y, Z, Q, D, W);
}

} // namespace

bool sc_metrics_device(HostBlock &b, int order, std::string &err) {
  if (b.dim != 3) {
    err = "symmetric_conservative_metric needs a 3D block";
    return false;
  }
  const int P = order == 4 || order == 6 ? order : 2; // other nonzero values run the second-order stage
  const int ng = b.ngg;
  if (b.mx + 2 * ng < 7 || b.my + 2 * ng < 7 || b.mz + 2 * ng < 7) {
    err = "symmetric_conservative_metric: a block needs at least 7 nodes per direction with its ghosts";
    return false;
  }
  Grid g{};
  g.n[0] = b.mx;
  g.n[1] = b.my;
  g.n[2] = b.mz;
  g.ng = ng;
  g.pd2 = b.mx + 2L * (ng + 1);
  g.pd1 = (b.my + 2L * (ng + 1)) * g.pd2;
  g.d2 = b.pad2();
  g.d1 = b.pad1();
  const long sz = b.size(), psz = (long)b.x.size();
  double *X = nullptr, *Y = nullptr, *Z = nullptr, *Q = nullptr, *D = nullptr, *jac = nullptr, *met = nullptr;
  int *bad = nullptr;
  auto check = [&](hipError_t e) {
    if (e != hipSuccess && err.empty()) err = std::string("symmetric_conservative_metric: ") + hipGetErrorString(e);
    return e == hipSuccess;
  };
  bool ok = check(hipMalloc(&X, sizeof(double) * psz)) && check(hipMalloc(&Y, sizeof(double) * psz)) &&
            check(hipMalloc(&Z, sizeof(double) * psz)) && check(hipMalloc(&Q, sizeof(double) * 18 * sz)) &&
            check(hipMalloc(&D, sizeof(double) * 3 * sz)) && check(hipMalloc(&jac, sizeof(double) * sz)) &&
            check(hipMalloc(&met, sizeof(double) * 9 * sz)) && check(hipMalloc(&bad, sizeof(int)));
  int h_bad = 0;
  if (ok) {
    ok = check(hipMemcpy(X, b.x.data(), sizeof(double) * psz, hipMemcpyHostToDevice)) &&
         check(hipMemcpy(Y, b.y.data(), sizeof(double) * psz, hipMemcpyHostToDevice)) &&
         check(hipMemcpy(Z, b.z.data(), sizeof(double) * psz, hipMemcpyHostToDevice)) &&
         check(hipMemcpy(bad, &h_bad, sizeof(int), hipMemcpyHostToDevice));
  }
  if (ok) {
    const unsigned nb = (unsigned)((sz + 255) / 256);
    if (P == 6) launch_first_stages<6>(g, X, Y, Z, Q, D, nb);
    else if (P == 4) launch_first_stages<4>(g, X, Y, Z, Q, D, nb);
    else launch_first_stages<2>(g, X, Y, Z, Q, D, nb);
    k_sc_metric<<<nb, 256>>>(g, Q, D, jac, met, bad, lag_weights(6));
    ok = check(hipGetLastError()) && check(hipDeviceSynchronize());
  }
  if (ok) {
    b.jac.assign(sz, 0.0);
    b.metric.assign(9 * sz, 0.0);
    ok = check(hipMemcpy(b.jac.data(), jac, sizeof(double) * sz, hipMemcpyDeviceToHost)) &&
         check(hipMemcpy(b.metric.data(), met, sizeof(double) * 9 * sz, hipMemcpyDeviceToHost)) &&
         check(hipMemcpy(&h_bad, bad, sizeof(int), hipMemcpyDeviceToHost));
  }
  for (void *p : {(void *)X, (void *)Y, (void *)Z, (void *)Q, (void *)D, (void *)jac, (void *)met, (void *)bad})
    if (p) (void)hipFree(p);
  if (ok && h_bad) {
    err = "negative Jacobian from the symmetric-conservative metrics";
    ok = false;
  }
  return ok;
}

} // namespace corefl
