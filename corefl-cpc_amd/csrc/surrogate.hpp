// Chemistry surrogate: a 12 -> 4x128 (GELU or hard swish) -> 10 MLP with BF16 weights and FP32 accumulation on the
// gfx950 matrix cores.  Training (forward, backward, Adam) runs as tiled MFMA GEMMs with fused epilogues;
// inference runs as one persistent kernel holding every weight in LDS, fused with the feature scaling, the
// element-conserving projection, the temperature recovery and the fallback gate.
#pragma once

#include <vector>

#include <hip/hip_runtime.h>

#include "chem.hpp"
#include "chem_kernels.hpp"

namespace corefl {

constexpr int kMlpLayers = 5, kMlpIn = 12, kMlpOut = 10, kMlpHidden = 128;
constexpr int kMlpK[kMlpLayers] = {12, 128, 128, 128, 128};
constexpr int kMlpN[kMlpLayers] = {128, 128, 128, 128, 10};
constexpr int kMlpKp[kMlpLayers] = {16, 128, 128, 128, 128}; // K padded to the 16-deep MFMA step
constexpr int kMlpNp[kMlpLayers] = {128, 128, 128, 128, 32}; // N padded to the 32-wide MFMA tile

// Device weights: BF16 W^T, [Np][Kp] row-major (k contiguous), zero padded; FP32 biases [N].
// Hidden activation: GELU in its tanh form (the network the specification names) or hard swish.
constexpr int kActGelu = 0, kActHardSwish = 1;

struct MlpDev {
  const unsigned short *wt[kMlpLayers];
  const float *bias[kMlpLayers];
  int act; // kActGelu or kActHardSwish
};

// Feature and target scaling.  Features (T, ln p, Y_1^a..Y_9^a, ln dt) -> (f - mu) / sig, with a = y_power
// (1: plain mass fractions; a < 1 spreads the small radical fractions).  Targets: s = out * sig + mu is the
// signed logarithm of the rate, rate = sign(s) r0 (exp|s| - 1) for (dY_k/dt, dT/dt); r0 <= 0 means
// rate = s (no transform).
struct SurrogateScaling {
  float in_mu[kMlpIn], in_sig[kMlpIn];
  double out_mu[kMlpOut], out_sig[kMlpOut], out_r0[kMlpOut];
  double y_power;
};

// A cell uses the network only inside this input box, when the network's own temperature change agrees
// with the one recovered from energy conservation to within dT_max, and when its temperature changes by at
// most dT_step over the sub-step (fast-reacting cells go to the exact integrator).
struct SurrogateGate {
  double t_min, t_max, p_min, p_max, dt_min, dt_max, dT_max, dT_step;
};

// Element conservation: E[e][k] = (atoms of element e in species k) / MW_k, at most 4 elements; rows
// e >= ne are zero.
struct ElementMatrix {
  int ne;
  double E[4][kMaxSpec];
  // inverse of the projection's normal matrix E E^T (rows e >= ne: identity) when no species is held at
  // zero, the common case
  double M0inv[4][4];
  // the same for every set of species held at zero (bit k: species k): 512 entries of 17 doubles, the 4 x 4
  // inverse of the normal matrix over the other species and 1 + m (0 where that matrix is singular), m the
  // bit set of the elements no free species carries (their rows and columns are the identity's)
  const double *inv_table;
};

constexpr int kProjTableEntries = 512, kProjTableStride = 17;
// E, M0inv and the held-species inverse table of a mechanism (host); false if it has more than 4 elements or
// E E^T is singular.
bool projection_tables(const ChemMech &mech, ElementMatrix &em, std::vector<double> &table);

// C = A B with A(m,k) = A[m*sam + k*sak], B(k,n) = B[k*sbk + n*sbn] in BF16; FP32 accumulation.
enum GemmEpilogue {
  kEpFwdHidden = 0, // z = acc + bias: C (FP32) = z, Cb (BF16) = act(z)
  kEpFwdOut = 1,    // C = acc + bias
  kEpBwdData = 2,   // Cb = acc * act'(zprev)
  kEpAtomic = 3     // C += acc (split over K by gridDim.z)
};

struct GemmArgs {
  int M, N, K;
  const unsigned short *A;
  long sam, sak;
  const unsigned short *B;
  long sbk, sbn;
  int ep;
  float *C;
  long ldc;
  unsigned short *Cb;
  long ldcb;
  const float *bias;
  const float *zprev;
  long ldz;
  int k_chunk; // K extent of one gridDim.z slice (multiple of 32); K for the non-atomic epilogues
  int act;     // hidden activation of the forward and data-gradient epilogues
};

hipError_t launch_gemm(const GemmArgs &a, hipStream_t s);
// Mini-batch assembly: rows idx_i of X [n][12] -> BF16 H0 [B][16] (zero padded), Y [n][10] -> Yb [B][10];
// idx_i = i when seed == 0, else a hash of (seed, step, i) modulo n.
hipError_t launch_gather(const float *X, const float *Y, long n, unsigned long long seed, long step, int B,
                         unsigned short *H0, float *Yb, hipStream_t s);
// G = BF16(2 (out - y) / (B * 10)); *loss += mean squared error.
hipError_t launch_loss(const float *out, const float *Yb, int B, unsigned short *G, float *loss, hipStream_t s);
// db[n] += sum_b G[b][n] (BF16 G, row stride ldg).
hipError_t launch_colsum(const unsigned short *G, int B, int N, long ldg, float *db, hipStream_t s);
// Adam on n FP32 parameters; when wt != nullptr also writes BF16 W^T (K x N weights, row stride Kp).
hipError_t launch_adam(float *p, float *m, float *v, const float *g, long n, float lr, float b1, float b2, float eps,
                       float c1, float c2, unsigned short *wt, int K, int N, int Kp, hipStream_t s);
hipError_t launch_to_bf16t(const float *W, int K, int N, int Kp, unsigned short *wt, hipStream_t s);
hipError_t launch_fill(float *p, long n, float v, hipStream_t s);

// Work arrays of a surrogate chemistry step for n cells: BF16 features [n][16], FP32 outputs [n][10] and
// the input-range flags [n].
struct SurrogateWork {
  unsigned short *feat;
  float *out;
  unsigned char *flag;
};

// FP32 feature rows X [n][12] -> BF16 rows [n][16] (zero padded).
hipError_t launch_rows_to_feat(const float *X, long n, unsigned short *feat, hipStream_t s);
// Inference.  Raw mode: BF16 feature rows feat [n][16] (as launch_gather writes them) -> outputs y [n][10].
hipError_t launch_mlp_raw(const MlpDev &net, long n, const unsigned short *feat, float *y, int blocks,
                          hipStream_t s);
// Chemistry mode: every cell of b (n cells) is advanced by dt (b.dtc per cell if set) with the network
// (features, network, projection / temperature / gate kernels); cells that fail the gate are appended to
// fb_list (count in *fb_count, zeroed by the launch) and left unchanged for the exact integrator.
hipError_t launch_mlp_chem(const MlpDev &net, const ChemMech *mech_dev, const ChemBatch &b, double dt,
                           const SurrogateScaling &sc, const SurrogateGate &gate, const ElementMatrix &em,
                           const SurrogateWork &w, int *fb_list, int *fb_count, int blocks, hipStream_t s);

} // namespace corefl
