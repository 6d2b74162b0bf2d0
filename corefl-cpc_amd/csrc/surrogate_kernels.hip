// Surrogate MLP kernels for gfx950: BF16 MFMA (v_mfma_f32_32x32x16_bf16) with FP32 accumulation.
//
// MFMA 32x32x16 operand layout (wave64): lane l supplies A[i][k] for i = l % 32, k = 8 (l / 32) + 0..7 and
// B[k][j] for j = l % 32, k = 8 (l / 32) + 0..7; it receives D[i][j] for j = l % 32 in 16 registers,
// register r holding row i = 8 (r / 4) + 4 (l / 32) + r % 4.  Both operands are therefore read as eight
// consecutive k values of one row (A) or one column (B): activations are stored [cell][k] and weights as
// W^T [n][k], so every operand fetch is one 16-byte LDS read.
#include "surrogate.hpp"

#include <cmath>

#include "chem_device.hpp"

namespace corefl {

namespace {

typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef short shortx8 __attribute__((ext_vector_type(8)));
typedef float floatx16 __attribute__((ext_vector_type(16)));

__device__ __forceinline__ floatx16 mfma(const shortx8 &a, const shortx8 &b, const floatx16 &c) {
  return __builtin_amdgcn_mfma_f32_32x32x16_bf16(__builtin_bit_cast(bf16x8, a), __builtin_bit_cast(bf16x8, b), c,
                                                 0, 0, 0);
}

// FP32 -> BF16, round to nearest even (finite inputs).
__device__ __forceinline__ unsigned short f2bf(float f) {
  unsigned u = __float_as_uint(f);
  u += 0x7FFFu + ((u >> 16) & 1u);
  return (unsigned short)(u >> 16);
}

// Hidden activations.  GELU (Hendrycks and Gimpel, "Gaussian Error Linear Units", 2016) in its tanh form:
// 0.5 z (1 + tanh(u)) = z / (1 + exp(-2u)), u = sqrt(2/pi) (z + 0.044715 z^3), with the hardware exp2 and
// reciprocal (1 ulp each; the activation is rounded to BF16 next).  Hard swish (Howard et al., "Searching for
// MobileNetV3", ICCV 2019), z clamp(z / 6 + 1/2, 0, 1): a GELU-like shape in three VALU operations without a
// transcendental (0.194 ms per 2^20 rows against the tanh GELU's 0.271 ms in the inference kernel).
constexpr float kGs = 0.7978845608028654f, kGa = 0.044715f;
// -2u log2(e) = z (c1 + c3 z^2)
constexpr float kG1 = -2.0f * 0.7978845608028654f * 1.4426950408889634f;
constexpr float kG3 = -2.0f * 0.7978845608028654f * 0.044715f * 1.4426950408889634f;
__device__ __forceinline__ float sig2u(float z) {
  const float w = z * __builtin_fmaf(kG3, z * z, kG1);
  return __builtin_amdgcn_rcpf(1.0f + __builtin_amdgcn_exp2f(w)); // exp2 -> +inf gives 0, the z -> -inf limit
}
constexpr float kActSlope = 1.0f / 6.0f;
template<int ACT> __device__ __forceinline__ float act(float z) {
  if constexpr (ACT == kActGelu) return z * sig2u(z);
  else return z * __builtin_amdgcn_fmed3f(__builtin_fmaf(z, kActSlope, 0.5f), 0.0f, 1.0f);
}
// d act / dz (hard swish: 0 below -3, 1 above 3, z / 3 + 1/2 between)
template<int ACT> __device__ __forceinline__ float act_grad(float z) {
  if constexpr (ACT == kActGelu) {
    const float s = sig2u(z);
    return s + z * s * (1.0f - s) * 2.0f * kGs * (1.0f + 3.0f * kGa * z * z);
  } else {
    return z <= -3.0f ? 0.0f : (z >= 3.0f ? 1.0f : __builtin_fmaf(z, 2.0f * kActSlope, 0.5f));
  }
}

// ------------------------------------------------------------------------------------------------ GEMM
constexpr int kTS = 40; // LDS row stride (elements) of a 32-deep k chunk, padded against bank conflicts

__global__ __launch_bounds__(256) void k_gemm(GemmArgs a) {
  __shared__ __attribute__((aligned(16))) unsigned short As[64 * kTS];
  __shared__ __attribute__((aligned(16))) unsigned short Bs[64 * kTS];
  const int t = threadIdx.x, wave = t >> 6, lane = t & 63;
  const int m0 = blockIdx.x * 64, n0 = blockIdx.y * 64;
  const int kb = blockIdx.z * a.k_chunk;
  const int ke = min(a.K, kb + a.k_chunk);
  const int wm = wave >> 1, wn = wave & 1;
  const int lr = t >> 2, lk = (t & 3) * 8; // staging: row lr of the tile, k offset lk
  floatx16 acc;
#pragma unroll
  for (int r = 0; r < 16; ++r) acc[r] = 0.0f;
  for (int k0 = kb; k0 < ke; k0 += 32) {
    {
      const int m = m0 + lr;
#pragma unroll
      for (int q = 0; q < 8; ++q) {
        const int k = k0 + lk + q;
        As[lr * kTS + lk + q] = (m < a.M && k < ke) ? a.A[m * a.sam + k * a.sak] : (unsigned short)0;
      }
      const int n = n0 + lr;
#pragma unroll
      for (int q = 0; q < 8; ++q) {
        const int k = k0 + lk + q;
        Bs[lr * kTS + lk + q] = (n < a.N && k < ke) ? a.B[k * a.sbk + n * a.sbn] : (unsigned short)0;
      }
    }
    __syncthreads();
#pragma unroll
    for (int ks = 0; ks < 32; ks += 16) {
      const int kk = ks + 8 * (lane >> 5);
      const shortx8 av = *reinterpret_cast<const shortx8 *>(&As[(wm * 32 + (lane & 31)) * kTS + kk]);
      const shortx8 bv = *reinterpret_cast<const shortx8 *>(&Bs[(wn * 32 + (lane & 31)) * kTS + kk]);
      acc = mfma(av, bv, acc);
    }
    __syncthreads();
  }
  const int n = n0 + wn * 32 + (lane & 31);
  if (n >= a.N) return;
#pragma unroll
  for (int r = 0; r < 16; ++r) {
    const int m = m0 + wm * 32 + 8 * (r >> 2) + 4 * (lane >> 5) + (r & 3);
    if (m >= a.M) continue;
    const float v = acc[r];
    if (a.ep == kEpFwdHidden) {
      const float z = v + a.bias[n];
      a.C[m * a.ldc + n] = z;
      a.Cb[m * a.ldcb + n] = f2bf(a.act == kActGelu ? act<kActGelu>(z) : act<kActHardSwish>(z));
    } else if (a.ep == kEpFwdOut) {
      a.C[m * a.ldc + n] = v + a.bias[n];
    } else if (a.ep == kEpBwdData) {
      const float z = a.zprev[m * a.ldz + n];
      a.Cb[m * a.ldcb + n] = f2bf(v * (a.act == kActGelu ? act_grad<kActGelu>(z) : act_grad<kActHardSwish>(z)));
    } else {
      atomicAdd(&a.C[m * a.ldc + n], v);
    }
  }
}

__device__ __forceinline__ unsigned long long mix64(unsigned long long x) {
  x += 0x9E3779B97F4A7C15ull;
  x = (x ^ (x >> 30)) * 0xBF58476D1CE4E5B9ull;
  x = (x ^ (x >> 27)) * 0x94D049BB133111EBull;
  return x ^ (x >> 31);
}

__global__ void k_gather(const float *X, const float *Y, long n, unsigned long long seed, long step, int B,
                         unsigned short *H0, float *Yb) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= B) return;
  const long idx = seed ? (long)(mix64(seed ^ mix64((unsigned long long)step * 0x100000001B3ull + i)) % n) : i;
#pragma unroll
  for (int f = 0; f < 16; ++f) H0[(long)i * 16 + f] = f < kMlpIn ? f2bf(X[idx * kMlpIn + f]) : (unsigned short)0;
#pragma unroll
  for (int o = 0; o < kMlpOut; ++o) Yb[(long)i * kMlpOut + o] = Y[idx * kMlpOut + o];
}

__global__ __launch_bounds__(256) void k_loss(const float *out, const float *Yb, int B, unsigned short *G,
                                              float *loss) {
  __shared__ float part[4];
  const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
  const float inv = 1.0f / (float)((long)B * kMlpOut);
  float sq = 0.0f;
  if (i < (long)B * kMlpOut) {
    const float d = out[i] - Yb[i];
    G[i] = f2bf(2.0f * d * inv);
    sq = d * d * inv;
  }
  for (int o = 32; o > 0; o >>= 1) sq += __shfl_xor(sq, o, 64);
  if ((threadIdx.x & 63) == 0) part[threadIdx.x >> 6] = sq;
  __syncthreads();
  if (threadIdx.x == 0) atomicAdd(loss, part[0] + part[1] + part[2] + part[3]);
}

__device__ __forceinline__ float bf2f(unsigned short h) { return __uint_as_float((unsigned)h << 16); }

// 256 threads: column n = t % 128, rows b0 + t / 128 + 2q of a 256-row chunk.
__global__ __launch_bounds__(256) void k_colsum(const unsigned short *G, int B, int N, long ldg, float *db) {
  const int n = threadIdx.x & 127;
  if (n >= N) return;
  const int b0 = blockIdx.x * 256 + (threadIdx.x >> 7);
  float s = 0.0f;
  for (int q = 0; q < 128; ++q) {
    const int b = b0 + 2 * q;
    if (b < B) s += bf2f(G[(long)b * ldg + n]);
  }
  atomicAdd(&db[n], s);
}

__global__ void k_adam(float *p, float *m, float *v, const float *g, long n, float lr, float b1, float b2, float eps,
                       float c1, float c2, unsigned short *wt, int K, int N, int Kp) {
  const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const float gi = g[i];
  const float mi = b1 * m[i] + (1.0f - b1) * gi;
  const float vi = b2 * v[i] + (1.0f - b2) * gi * gi;
  m[i] = mi;
  v[i] = vi;
  const float pi = p[i] - lr * (mi / c1) / (sqrtf(vi / c2) + eps);
  p[i] = pi;
  if (wt) wt[(i % N) * Kp + i / N] = f2bf(pi);
}

__global__ void k_to_bf16t(const float *W, int K, int N, int Kp, unsigned short *wt) {
  const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= (long)K * N) return;
  wt[(i % N) * Kp + i / N] = f2bf(W[i]);
}

__global__ void k_fill(float *p, long n, float v) {
  const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) p[i] = v;
}

// ---------------------------------------------------------------------------------------- fused inference
constexpr int kS0 = 24;  // LDS row stride of W0^T (16 + 8)
constexpr int kSH = 136; // LDS row stride of 128-deep rows (128 + 8)

struct FusedArgs {
  long n;
  const float *x; // raw mode
  float *y;
  ChemBatch b;    // chemistry mode
  double dt;
  const ChemMech *mech;
  SurrogateScaling sc;
  SurrogateGate gate;
  ElementMatrix em;
  int *fb_list, *fb_count;
};

// copy rows x cols BF16 from global [rows][cols] into LDS with row stride `stride` (cols multiple of 8)
__device__ __forceinline__ void load_rows(unsigned short *dst, int stride, const unsigned short *src, int rows,
                                          int cols) {
  const int per_row = cols / 8;
  for (int q = threadIdx.x; q < rows * per_row; q += blockDim.x) {
    const int r = q / per_row, c = (q % per_row) * 8;
    *reinterpret_cast<shortx8 *>(&dst[r * stride + c]) = *reinterpret_cast<const shortx8 *>(&src[r * cols + c]);
  }
}

// Mass fractions, pressure and the in-range flag of cell c; features scaled into f[0..12).
__device__ __forceinline__ bool chem_features(const FusedArgs &a, long c, float *f) {
  const ChemBatch &b = a.b;
  const ChemMech &m = *a.mech;
  const double rho = b.rho[c], T = b.T[c], dt = cell_dt(b, c, a.dt);
  const double irho = 1.0 / rho;
  double imw = 0.0;
  double y[9];
#pragma unroll
  for (int k = 0; k < 9; ++k) {
    y[k] = b.rhoY[(long)k * b.n + c] * irho;
    imw += y[k] * m.imw[k];
  }
  const double p = rho * chemdev::kRu * imw * T;
  const SurrogateScaling &s = a.sc;
  // The features are rounded to BF16 (8 significant bits) for the network: FP32 arithmetic with the
  // hardware log2 / exp2 (1 ulp) is exact to far below that.
  f[0] = ((float)T - s.in_mu[0]) / s.in_sig[0];
  f[1] = (__builtin_amdgcn_logf((float)p) * 0.6931471805599453f - s.in_mu[1]) / s.in_sig[1];
  const float ap = (float)s.y_power;
#pragma unroll
  for (int k = 0; k < 9; ++k) {
    float yf;
    if (s.y_power == 1.0) {
      yf = (float)y[k];
    } else {
      // y^a = 2^(a log2 y); y scaled by 2^64 first so that mass fractions below the FP32 normal range keep
      // their logarithm
      yf = y[k] > 0 ? __builtin_amdgcn_exp2f(ap * (__builtin_amdgcn_logf((float)(y[k] * 0x1p64)) - 64.0f)) : 0.0f;
    }
    f[2 + k] = (yf - s.in_mu[2 + k]) / s.in_sig[2 + k];
  }
  f[11] = (__builtin_amdgcn_logf((float)dt) * 0.6931471805599453f - s.in_mu[11]) / s.in_sig[11];
  const SurrogateGate &g = a.gate;
  return T >= g.t_min && T <= g.t_max && p >= g.p_min && p <= g.p_max && dt >= g.dt_min && dt <= g.dt_max;
}

// Network output q -> rate (inverse of the signed-logarithm target transform).  FP32: the network's
// outputs carry BF16-level error, and r0 (e^|v| - 1) is taken with the hardware exp2 (its absolute error
// near v = 0, r0 * 1e-7, is far below the rates that matter).
__device__ __forceinline__ double rate_of(const SurrogateScaling &s, int q, float o) {
  const float v = o * (float)s.out_sig[q] + (float)s.out_mu[q];
  if (!(s.out_r0[q] > 0)) return (double)v;
  const float m = (float)s.out_r0[q] * (__builtin_amdgcn_exp2f(fabsf(v) * 1.4426950408889634f) - 1.0f);
  return (double)(v < 0 ? -m : m);
}

// Network outputs o[10] of cell c -> projected mass fractions and energy-consistent temperature; writes the
// cell if the gate passes, else appends it to the fallback list.
// The NASA-7 data update_t reads, staged in LDS by each workgroup of k_chem_post.
struct ThermoLds {
  double lo[9][7], hi[9][7];
  double t_low[9], t_mid[9], t_high[9], R[9], imw[9];
};

// h and cp of species i at t from the LDS table: chemdev::h_cp_one's operations in its order.
__device__ __forceinline__ void h_cp_lds(const ThermoLds &th, int i, double t, double &h, double &cp) {
  using chemdev::kThird;
  const double t2 = t * t, t3 = t2 * t, t4 = t3 * t, t5 = t4 * t;
  const double R = th.R[i];
  const bool below = t < th.t_low[i], above = t > th.t_high[i];
  if (below || above) {
    const double tt = below ? th.t_low[i] : th.t_high[i];
    const double *c = below ? th.lo[i] : th.hi[i];
    const double tt2 = tt * tt, tt3 = tt2 * tt, tt4 = tt3 * tt, tt5 = tt4 * tt;
    const double cc = c[0] + c[1] * tt + c[2] * tt2 + c[3] * tt3 + c[4] * tt4;
    const double hh = c[0] * tt + 0.5 * c[1] * tt2 + c[2] * tt3 * kThird + 0.25 * c[3] * tt4 + 0.2 * c[4] * tt5 + c[5];
    h = (hh + cc * (t - tt)) * R;
    cp = cc * R;
  } else {
    const double *c = t < th.t_mid[i] ? th.lo[i] : th.hi[i];
    h = (c[0] * t + 0.5 * c[1] * t2 + c[2] * t3 * kThird + 0.25 * c[3] * t4 + 0.2 * c[4] * t5 + c[5]) * R;
    cp = (c[0] + c[1] * t + c[2] * t2 + c[3] * t3 + c[4] * t4) * R;
  }
}

__device__ void chem_finish(const FusedArgs &a, long c, const float *o, bool in_range, const ThermoLds &th) {
  // The projection's passes and the Newton iteration stay rolled up, the held-species inverses come from a
  // table and the thermo coefficients from LDS, so that the register footprint (122 VGPRs) leaves four waves
  // per SIMD to hide the memory latency (16.7M cells: 1.13 ms, against 1.71 ms with the passes unrolled, the
  // held-species system solved in registers and the coefficients held in registers at 254 VGPRs).
  const ChemBatch &b = a.b;
  const double rho = b.rho[c], T = b.T[c], dt = cell_dt(b, c, a.dt);
  const double irho = 1.0 / rho;
  double y0[9], dy[9];
#pragma unroll
  for (int k = 0; k < 9; ++k) {
    y0[k] = b.rhoY[(long)k * b.n + c] * irho;
    dy[k] = dt * rate_of(a.sc, k, o[k]);
  }
  const double t_net = T + dt * rate_of(a.sc, 9, o[9]);
  // least-squares projection onto element conservation (rows e >= ne are identity padding), species
  // driven below zero held at zero
  unsigned clamp = 0;
  bool ok = in_range;
  double yn[9];
#pragma unroll 1
  // each pass that does not end the loop holds at least one more species, so nine passes always decide
  for (int pass = 0; pass < 9 && ok; ++pass) {
    double res[4];
#pragma unroll
    for (int e = 0; e < 4; ++e) {
      double r = 0.0;
#pragma unroll
      for (int k = 0; k < 9; ++k) r += a.em.E[e][k] * ((clamp >> k) & 1u ? -y0[k] : dy[k]);
      res[e] = r;
    }
    // the inverse normal matrix over the species not held at zero, inverted on the host: the mechanism's own
    // when none is held, else the table's entry for the held set (singular: the cell falls back).  An element
    // that only held species carry is conserved only if they held none of it (a pure stream of the mixing
    // layer holds no hydrogen, or no oxygen): else the cell falls back.
    const double *Mi = clamp == 0 ? &a.em.M0inv[0][0] : a.em.inv_table + (long)clamp * kProjTableStride;
    if (clamp != 0) {
      const double tag = Mi[16];
      bool lost = !(tag > 0);
      const unsigned drop = lost ? 0u : (unsigned)tag - 1u;
#pragma unroll
      for (int e = 0; e < 4; ++e) lost |= ((drop >> e) & 1u) && fabs(res[e]) > 1e-14;
      if (lost) {
        ok = false;
        break;
      }
    }
    double x[4];
#pragma unroll
    for (int e = 0; e < 4; ++e) {
      x[e] = 0.0;
#pragma unroll
      for (int f = 0; f < 4; ++f) x[e] += Mi[4 * e + f] * res[f];
    }
#pragma unroll
    for (int e = 0; e < 4; ++e) res[e] = x[e];
    unsigned neg = 0;
#pragma unroll
    for (int k = 0; k < 9; ++k) {
      if ((clamp >> k) & 1u) {
        yn[k] = 0.0;
      } else {
        double corr = 0.0;
#pragma unroll
        for (int e = 0; e < 4; ++e) corr += a.em.E[e][k] * res[e];
        yn[k] = y0[k] + (dy[k] - corr);
        if (yn[k] < 0.0) neg |= 1u << k;
      }
    }
    if (!neg) break;
    clamp |= neg;
    if (pass == 8) ok = false;
  }
  double t_new = t_net;
  if (ok) {
    // chemdev::update_t with the coefficients read from LDS in every iteration (the same operations in the same
    // order; the empty asm statement with a memory clobber keeps the compiler from hoisting those reads, which
    // would hold every coefficient in registers)
    t_new = fmin(fmax(t_net, 200.0), 5000.0);
    double imw = 0;
#pragma unroll
    for (int l = 0; l < 9; ++l) imw += yn[l] * th.imw[l];
    const double R = chemdev::kRu * imw, E = b.e[c];
    double err = 1e+6;
#pragma unroll 1
    for (int it = 0; it < 100 && err > 1e-8; ++it) {
      asm volatile("" ::: "memory");
      double cpt = 0, hh = 0;
#pragma unroll
      for (int l = 0; l < 9; ++l) {
        double h1, cp1;
        h_cp_lds(th, l, t_new, h1, cp1);
        cpt += cp1 * yn[l];
        hh += h1 * yn[l];
      }
      const double t1 = t_new - (hh - R * t_new - E) / (cpt - R);
      err = fabs(1 - t1 / t_new);
      t_new = t1;
    }
    ok = isfinite(t_new) && fabs(t_new - t_net) <= a.gate.dT_max && fabs(t_new - T) <= a.gate.dT_step;
  }
  if (ok) {
#pragma unroll
    for (int k = 0; k < 9; ++k) b.rhoY[(long)k * b.n + c] = rho * yn[k];
    b.T[c] = t_new;
    b.steps[c] = 0;
    b.rejects[c] = 0;
  } else {
    const int slot = atomicAdd(a.fb_count, 1);
    a.fb_list[slot] = (int)c;
  }
}

// Inference MLP, one pass over the batch.  The product of each layer is formed transposed, out^T = W^T h^T:
// the weights (LDS-resident for the whole persistent launch) are the MFMA's A operand and the activations
// its B operand.  The accumulator layout of out^T (lane = cell, 16 neurons per lane) then maps onto the next
// layer's B operand with one exchange of half a neuron group between lanes l and l + 32 (v_permlane32_swap),
// so activations stay in registers from the features to the outputs.  Each layer's accumulators start from
// its bias through one more MFMA: the FP32 bias split exactly into three BF16 parts (A operand, k = 0..2)
// times ones (B operand), which takes the bias off the LDS read path and the VALU.  A wave runs kMlpTiles
// 32-cell tiles at once (each weight fragment read from LDS then feeds kMlpTiles MFMAs); twelve waves per
// workgroup, three per SIMD (one workgroup per CU: the weights fill 125 KB of LDS), overlap one wave's MFMAs
// with another's activation arithmetic.  Measured on 16.7M cells with the tanh GELU: 3.29 ms at twelve waves,
// 3.37 ms at eight, 4.07 ms with the bias added on the VALU from LDS and the half-group exchange through
// ds_bpermute; issuing each block's k-steps of the next layer right after its activation (a software
// pipeline within the wave) measured 3.49 ms.  Two tiles at four waves measured slower than one at eight.
// Computing the features inside this kernel from the chemistry batch (both lanes of a cell computing all twelve)
// and the input-range check in the post-processing kernel measured 3.80 + 1.44 ms against 0.35 + 3.29 + 1.17 ms
// with the separate feature kernel: the feature work is FP64 and VALU-heavy, and this kernel is VALU-bound.
constexpr int kMlpTiles = 1, kMlpWaves = 12;
constexpr int kBiasFrags = 4 * 4 + 1; // (layer, 32-neuron block) of the hidden layers, then the output layer

typedef unsigned uintx4 __attribute__((ext_vector_type(4)));
typedef __bf16 bf16x2 __attribute__((ext_vector_type(2)));
typedef float floatx2 __attribute__((ext_vector_type(2)));

__device__ __forceinline__ unsigned pack_bf16(float a, float b) {
  const floatx2 v = {a, b};
  return __builtin_bit_cast(unsigned, __builtin_convertvector(v, bf16x2)); // round to nearest even
}

// x = hi + mid + lo exactly, each part BF16 (round to nearest even; each remainder is exact in FP32), as the
// A-operand dwords {hi | mid << 16, lo}.
__device__ __forceinline__ uint2 split_bf16x3(float x) {
  const unsigned short hi = f2bf(x);
  const float r1 = x - __uint_as_float((unsigned)hi << 16);
  const unsigned short mid = f2bf(r1);
  const float r2 = r1 - __uint_as_float((unsigned)mid << 16);
  return make_uint2((unsigned)hi | ((unsigned)mid << 16), (unsigned)f2bf(r2));
}

// Activation of one 32-neuron block of accumulators (bias included), packed as BF16 pairs: p[q][0..1] holds
// neurons 8q + 4h + 0..3 of the block (h = lane / 32).
template<int ACT> __device__ __forceinline__ void act_pack(const floatx16 &acc, unsigned (&p)[4][2]) {
#pragma unroll
  for (int q = 0; q < 4; ++q) {
    p[q][0] = pack_bf16(act<ACT>(acc[4 * q]), act<ACT>(acc[4 * q + 1]));
    p[q][1] = pack_bf16(act<ACT>(acc[4 * q + 2]), act<ACT>(acc[4 * q + 3]));
  }
}

// B fragments of the next layer for the two k-steps 2j, 2j + 1 covered by block j (16 neurons each): lane
// (cell, h) needs group 2 par + h (neurons 16 par + 8h + 0..7 of the block), whose other half its partner lane
// (l ^ 32) holds.  v_permlane32_swap on (A = p[2 par], B = p[2 par + 1]) swaps A's upper 32 lanes with B's
// lower 32: afterwards A holds {own low half | partner's low half of group 2 par + 1} and B {partner's upper
// half of group 2 par | own upper half}, which is {A, B} = group 2 par + h in neuron order on every lane.
__device__ __forceinline__ void next_operand(const unsigned (&p)[4][2], shortx8 &b0, shortx8 &b1) {
#pragma unroll
  for (int par = 0; par < 2; ++par) {
    const auto s0 = __builtin_amdgcn_permlane32_swap(p[2 * par][0], p[2 * par + 1][0], false, false);
    const auto s1 = __builtin_amdgcn_permlane32_swap(p[2 * par][1], p[2 * par + 1][1], false, false);
    const uintx4 v = {s0[0], s1[0], s0[1], s1[1]};
    (par == 0 ? b0 : b1) = __builtin_bit_cast(shortx8, v);
  }
}

template<int ACT>
__global__ __launch_bounds__(64 * kMlpWaves) void k_mlp(MlpDev net, const unsigned short *__restrict__ feat,
                                                         float *__restrict__ out, long n) {
  __shared__ __attribute__((aligned(16))) unsigned short w0[128 * kS0];
  __shared__ __attribute__((aligned(16))) unsigned short wh[3][128 * kSH];
  __shared__ __attribute__((aligned(16))) unsigned short w4[32 * kSH];
  // bias A fragments [frag][h][row]: the three BF16 parts on the h = 0 lanes, zeros on the h = 1 lanes
  __shared__ __attribute__((aligned(16))) uint2 bfrag[kBiasFrags][2][32];
  load_rows(w0, kS0, net.wt[0], 128, 16);
  for (int l = 1; l < 4; ++l) load_rows(wh[l - 1], kSH, net.wt[l], 128, 128);
  load_rows(w4, kSH, net.wt[4], 32, 128);
  for (int q = threadIdx.x; q < kBiasFrags * 32; q += blockDim.x) {
    const int fr = q / 32, row = q % 32;
    const float b = fr < 16 ? net.bias[fr / 4][(fr % 4) * 32 + row] : (row < kMlpOut ? net.bias[4][row] : 0.0f);
    bfrag[fr][0][row] = split_bf16x3(b);
    bfrag[fr][1][row] = make_uint2(0u, 0u);
  }
  __syncthreads();
  const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
  const int cell = lane & 31, h = lane >> 5, kh = 8 * h;
  // B operand of the bias MFMA: ones at k = 0..2 (BF16 1.0 = 0x3f80), zero elsewhere
  const shortx8 ones = __builtin_bit_cast(shortx8, h == 0 ? uintx4{0x3f803f80u, 0x3f80u, 0u, 0u} : uintx4{0u, 0u, 0u, 0u});
  const floatx16 zero = {};
  constexpr long kCellsPerWave = 32 * kMlpTiles;
  const long n_units = (n + kCellsPerWave - 1) / kCellsPerWave;
  const int waves = gridDim.x * kMlpWaves;
  for (long unit = (long)blockIdx.x * kMlpWaves + wave; unit < n_units; unit += waves) {
    // the bias and weight fragments are re-read from LDS for every tile, not hoisted into registers
    asm volatile("" ::: "memory");
    const long c0 = unit * kCellsPerWave;
    floatx16 acc[kMlpTiles][4];
    shortx8 bop[kMlpTiles][8]; // B operands of the current layer, k-steps 0..7
    // layer 0: 16 features (12 used) -> 128
    {
      shortx8 f[kMlpTiles];
#pragma unroll
      for (int t = 0; t < kMlpTiles; ++t) {
        const long c = c0 + 32 * t + cell;
        f[t] = shortx8{0, 0, 0, 0, 0, 0, 0, 0};
        if (c < n) f[t] = *reinterpret_cast<const shortx8 *>(&feat[c * 16 + kh]);
      }
#pragma unroll
      for (int j = 0; j < 4; ++j) {
        const uint2 bf = bfrag[j][h][cell];
        const shortx8 ab = __builtin_bit_cast(shortx8, uintx4{bf.x, bf.y, 0u, 0u});
        const shortx8 a = *reinterpret_cast<const shortx8 *>(&w0[(j * 32 + cell) * kS0 + kh]);
#pragma unroll
        for (int t = 0; t < kMlpTiles; ++t) acc[t][j] = mfma(a, f[t], mfma(ab, ones, zero));
      }
    }
#pragma unroll
    for (int l = 0; l < 4; ++l) {
      // activation of layer l's output -> B operands of layer l + 1
#pragma unroll
      for (int t = 0; t < kMlpTiles; ++t)
#pragma unroll
        for (int j = 0; j < 4; ++j) {
          unsigned p[4][2];
          act_pack<ACT>(acc[t][j], p);
          next_operand(p, bop[t][2 * j], bop[t][2 * j + 1]);
        }
      if (l == 3) break;
      const unsigned short *w = wh[l];
#pragma unroll
      for (int j = 0; j < 4; ++j) {
        const uint2 bf = bfrag[4 * (l + 1) + j][h][cell];
        const shortx8 ab = __builtin_bit_cast(shortx8, uintx4{bf.x, bf.y, 0u, 0u});
#pragma unroll
        for (int t = 0; t < kMlpTiles; ++t) acc[t][j] = mfma(ab, ones, zero);
      }
#pragma unroll
      for (int ks = 0; ks < 8; ++ks)
#pragma unroll
        for (int j = 0; j < 4; ++j) {
          const shortx8 a = *reinterpret_cast<const shortx8 *>(&w[(j * 32 + cell) * kSH + ks * 16 + kh]);
#pragma unroll
          for (int t = 0; t < kMlpTiles; ++t) acc[t][j] = mfma(a, bop[t][ks], acc[t][j]);
        }
    }
    // output layer: 128 -> 10 (32-wide tile)
    floatx16 o[kMlpTiles];
    {
      const uint2 bf = bfrag[16][h][cell];
      const shortx8 ab = __builtin_bit_cast(shortx8, uintx4{bf.x, bf.y, 0u, 0u});
#pragma unroll
      for (int t = 0; t < kMlpTiles; ++t) o[t] = mfma(ab, ones, zero);
    }
#pragma unroll
    for (int ks = 0; ks < 8; ++ks) {
      const shortx8 a = *reinterpret_cast<const shortx8 *>(&w4[cell * kSH + ks * 16 + kh]);
#pragma unroll
      for (int t = 0; t < kMlpTiles; ++t) o[t] = mfma(a, bop[t][ks], o[t]);
    }
#pragma unroll
    for (int t = 0; t < kMlpTiles; ++t) {
      const long c = c0 + 32 * t + cell;
      if (c >= n) continue;
#pragma unroll
      for (int r = 0; r < 16; ++r) {
        const int j = 8 * (r >> 2) + 4 * h + (r & 3);
        if (j < kMlpOut) out[c * kMlpOut + j] = o[t][r];
      }
    }
  }
}

// Features of every cell of the batch as BF16 rows [16] (zero padded) and the input-range flag.
__global__ __launch_bounds__(256) void k_chem_feat(FusedArgs a, unsigned short *__restrict__ feat,
                                                   unsigned char *__restrict__ flag) {
  const long c = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (c >= a.n) return;
  float f[16];
#pragma unroll
  for (int q = 0; q < 16; ++q) f[q] = 0.0f;
  flag[c] = chem_features(a, c, f) ? 1 : 0;
  shortx8 lo, hi;
#pragma unroll
  for (int q = 0; q < 8; ++q) {
    lo[q] = (short)f2bf(f[q]);
    hi[q] = (short)f2bf(f[8 + q]);
  }
  *reinterpret_cast<shortx8 *>(&feat[c * 16]) = lo;
  *reinterpret_cast<shortx8 *>(&feat[c * 16 + 8]) = hi;
}

// Projection, temperature, gate and write-back of every cell from the network outputs.
__global__ __launch_bounds__(256) void k_chem_post(FusedArgs a, const float *__restrict__ out,
                                                   const unsigned char *__restrict__ flag) {
  __shared__ ThermoLds th;
  const ChemMech &m = *a.mech;
  for (int q = threadIdx.x; q < 9 * 7; q += blockDim.x) {
    th.lo[q / 7][q % 7] = m.lo[q / 7][q % 7];
    th.hi[q / 7][q % 7] = m.hi[q / 7][q % 7];
  }
  if (threadIdx.x < 9) {
    const int i = threadIdx.x;
    th.t_low[i] = m.t_low[i];
    th.t_mid[i] = m.t_mid[i];
    th.t_high[i] = m.t_high[i];
    th.R[i] = chemdev::kRu * m.imw[i];
    th.imw[i] = m.imw[i];
  }
  __syncthreads();
  const long c = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (c >= a.n) return;
  float o[kMlpOut];
#pragma unroll
  for (int q = 0; q < kMlpOut; ++q) o[q] = out[c * kMlpOut + q];
  chem_finish(a, c, o, flag[c] != 0, th);
}

__global__ void k_rows_to_feat(const float *__restrict__ X, long n, unsigned short *__restrict__ feat) {
  const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
#pragma unroll
  for (int f = 0; f < 16; ++f) feat[i * 16 + f] = f < kMlpIn ? f2bf(X[i * kMlpIn + f]) : (unsigned short)0;
}

__global__ void k_zero_count(int *p) { *p = 0; }

inline unsigned grid1(long n, int t) { return (unsigned)((n + t - 1) / t); }

} // namespace

hipError_t launch_gemm(const GemmArgs &a, hipStream_t s) {
  if (a.M <= 0 || a.N <= 0 || a.K <= 0) return hipSuccess;
  if (a.k_chunk <= 0 || a.k_chunk % 32) return hipErrorInvalidValue;
  const int zs = (a.K + a.k_chunk - 1) / a.k_chunk;
  if (zs > 1 && a.ep != kEpAtomic) return hipErrorInvalidValue;
  dim3 grid((a.M + 63) / 64, (a.N + 63) / 64, zs);
  k_gemm<<<grid, 256, 0, s>>>(a);
  return hipGetLastError();
}

hipError_t launch_gather(const float *X, const float *Y, long n, unsigned long long seed, long step, int B,
                         unsigned short *H0, float *Yb, hipStream_t s) {
  if (B <= 0) return hipSuccess;
  if (n <= 0) return hipErrorInvalidValue;
  k_gather<<<grid1(B, 256), 256, 0, s>>>(X, Y, n, seed, step, B, H0, Yb);
  return hipGetLastError();
}

hipError_t launch_loss(const float *out, const float *Yb, int B, unsigned short *G, float *loss, hipStream_t s) {
  if (B <= 0) return hipSuccess;
  k_loss<<<grid1((long)B * kMlpOut, 256), 256, 0, s>>>(out, Yb, B, G, loss);
  return hipGetLastError();
}

hipError_t launch_colsum(const unsigned short *G, int B, int N, long ldg, float *db, hipStream_t s) {
  if (B <= 0) return hipSuccess;
  if (N > 128) return hipErrorInvalidValue;
  k_colsum<<<grid1(B, 256), 256, 0, s>>>(G, B, N, ldg, db);
  return hipGetLastError();
}

hipError_t launch_adam(float *p, float *m, float *v, const float *g, long n, float lr, float b1, float b2, float eps,
                       float c1, float c2, unsigned short *wt, int K, int N, int Kp, hipStream_t s) {
  if (n <= 0) return hipSuccess;
  k_adam<<<grid1(n, 256), 256, 0, s>>>(p, m, v, g, n, lr, b1, b2, eps, c1, c2, wt, K, N, Kp);
  return hipGetLastError();
}

hipError_t launch_to_bf16t(const float *W, int K, int N, int Kp, unsigned short *wt, hipStream_t s) {
  k_to_bf16t<<<grid1((long)K * N, 256), 256, 0, s>>>(W, K, N, Kp, wt);
  return hipGetLastError();
}

hipError_t launch_fill(float *p, long n, float v, hipStream_t s) {
  if (n <= 0) return hipSuccess;
  k_fill<<<grid1(n, 256), 256, 0, s>>>(p, n, v);
  return hipGetLastError();
}

hipError_t launch_rows_to_feat(const float *X, long n, unsigned short *feat, hipStream_t s) {
  if (n <= 0) return hipSuccess;
  k_rows_to_feat<<<grid1(n, 256), 256, 0, s>>>(X, n, feat);
  return hipGetLastError();
}

hipError_t launch_mlp_raw(const MlpDev &net, long n, const unsigned short *feat, float *y, int blocks,
                          hipStream_t s) {
  if (n <= 0) return hipSuccess;
  const long units = (n + 32 * kMlpTiles * kMlpWaves - 1) / (32 * kMlpTiles * kMlpWaves);
  if (net.act == kActGelu) k_mlp<kActGelu><<<(unsigned)std::min<long>(units, blocks), 64 * kMlpWaves, 0, s>>>(net, feat, y, n);
  else k_mlp<kActHardSwish><<<(unsigned)std::min<long>(units, blocks), 64 * kMlpWaves, 0, s>>>(net, feat, y, n);
  return hipGetLastError();
}

hipError_t launch_mlp_chem(const MlpDev &net, const ChemMech *mech_dev, const ChemBatch &b, double dt,
                           const SurrogateScaling &sc, const SurrogateGate &gate, const ElementMatrix &em,
                           const SurrogateWork &w, int *fb_list, int *fb_count, int blocks, hipStream_t s) {
  k_zero_count<<<1, 1, 0, s>>>(fb_count);
  if (b.n <= 0) return hipGetLastError();
  FusedArgs a{};
  a.n = b.n;
  a.b = b;
  a.dt = dt;
  a.mech = mech_dev;
  a.sc = sc;
  a.gate = gate;
  a.em = em;
  a.fb_list = fb_list;
  a.fb_count = fb_count;
  const unsigned g = grid1(b.n, 256);
  k_chem_feat<<<g, 256, 0, s>>>(a, w.feat, w.flag);
  const long units = (b.n + 32 * kMlpTiles * kMlpWaves - 1) / (32 * kMlpTiles * kMlpWaves);
  if (net.act == kActGelu)
    k_mlp<kActGelu><<<(unsigned)std::min<long>(units, blocks), 64 * kMlpWaves, 0, s>>>(net, w.feat, w.out, b.n);
  else
    k_mlp<kActHardSwish><<<(unsigned)std::min<long>(units, blocks), 64 * kMlpWaves, 0, s>>>(net, w.feat, w.out, b.n);
  k_chem_post<<<g, 256, 0, s>>>(a, w.out, w.flag);
  return hipGetLastError();
}

} // namespace corefl
