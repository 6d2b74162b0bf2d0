// C ABI of the chemistry surrogate: parameters, scaling and gate, GPU training (forward, backward, Adam),
// raw inference, and the gated surrogate chemistry sub-step with exact-integrator fallback.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <memory>
#include <stdexcept>
#include <string>
#include <vector>

#include "../../include/corefl_c.h"
#include "chem_ctx.hpp"
#include "surrogate.hpp"

#define HIPCHK COREFL_HIPCHK

using namespace corefl;

namespace {

// flat parameter layout: W_0 (K_0 x N_0, row-major), b_0, W_1, b_1, ..., W_4, b_4
struct Offsets {
  long w[kMlpLayers], b[kMlpLayers], total;
  Offsets() {
    long o = 0;
    for (int l = 0; l < kMlpLayers; ++l) {
      w[l] = o;
      o += (long)kMlpK[l] * kMlpN[l];
      b[l] = o;
      o += kMlpN[l];
    }
    total = o;
  }
};
const Offsets kOff;

} // namespace

struct corefl_surrogate {
  hipStream_t stream = nullptr;
  int cap = 0; // mini-batch capacity
  float *p = nullptr, *m = nullptr, *v = nullptr, *g = nullptr;
  unsigned short *wt[kMlpLayers] = {};
  long adam_t = 0;
  // activations of one mini-batch
  unsigned short *H[kMlpLayers] = {}; // H[0]: [cap][16]; H[l>0]: [cap][128]
  float *Z[kMlpLayers - 1] = {};      // pre-activations of the hidden layers, [cap][128]
  float *out = nullptr, *Yb = nullptr;
  unsigned short *G4 = nullptr, *Ga = nullptr, *Gb = nullptr;
  float *loss = nullptr; // per-step losses
  long loss_cap = 0;
  // training set
  float *X = nullptr, *Y = nullptr;
  long n_data = 0;
  SurrogateScaling sc{};
  SurrogateGate gate{};
  bool have_scaling = false, have_gate = false, have_params = false;
  int cus = 0;
  double step_dt = 0; // uniform sub-step length of the current corefl_chem_surrogate_step call
  int act = kActGelu;  // hidden activation
  ~corefl_surrogate();
};

namespace {

int guarded_impl(const std::exception &e) {
  set_last_error(e.what());
  return -1;
}
template<class F> int guarded(F &&f) {
  try {
    f();
    return 0;
  } catch (const std::exception &e) {
    return guarded_impl(e);
  }
}

template<class T> void dalloc(T **p, size_t count) {
  HIPCHK(hipMalloc((void **)p, sizeof(T) * count));
}

MlpDev net_of(const corefl_surrogate *s) {
  MlpDev n{};
  for (int l = 0; l < kMlpLayers; ++l) {
    n.wt[l] = s->wt[l];
    n.bias[l] = s->p + kOff.b[l];
  }
  n.act = s->act;
  return n;
}

GemmArgs gemm(int M, int N, int K, const unsigned short *A, long sam, long sak, const unsigned short *B, long sbk,
              long sbn, int ep) {
  GemmArgs a{};
  a.M = M;
  a.N = N;
  a.K = K;
  a.A = A;
  a.sam = sam;
  a.sak = sak;
  a.B = B;
  a.sbk = sbk;
  a.sbn = sbn;
  a.ep = ep;
  a.k_chunk = ((K + 31) / 32) * 32;
  return a;
}

// Forward pass of the B rows in H[0]; outputs in s->out.
void forward(corefl_surrogate *s, int B) {
  for (int l = 0; l < kMlpLayers; ++l) {
    const long ld = l == 0 ? 16 : 128;
    GemmArgs a = gemm(B, kMlpN[l], kMlpKp[l], s->H[l], ld, 1, s->wt[l], 1, kMlpKp[l],
                      l < kMlpLayers - 1 ? kEpFwdHidden : kEpFwdOut);
    a.bias = s->p + kOff.b[l];
    a.act = s->act;
    if (l < kMlpLayers - 1) {
      a.C = s->Z[l];
      a.ldc = 128;
      a.Cb = s->H[l + 1];
      a.ldcb = 128;
    } else {
      a.C = s->out;
      a.ldc = kMlpOut;
    }
    HIPCHK(launch_gemm(a, s->stream));
  }
}

// Loss and gradients of the B rows (targets in s->Yb) into s->g; loss added to *loss_slot.
void backward(corefl_surrogate *s, int B, float *loss_slot) {
  HIPCHK(launch_fill(s->g, kOff.total, 0.0f, s->stream));
  HIPCHK(launch_loss(s->out, s->Yb, B, s->G4, loss_slot, s->stream));
  const unsigned short *G = s->G4;
  long ldg = kMlpOut;
  for (int l = kMlpLayers - 1; l >= 0; --l) {
    const long ldh = l == 0 ? 16 : 128;
    // dW_l = H_l^T G, split over the batch
    GemmArgs w = gemm(kMlpK[l], kMlpN[l], B, s->H[l], 1, ldh, G, ldg, 1, kEpAtomic);
    w.C = s->g + kOff.w[l];
    w.ldc = kMlpN[l];
    w.k_chunk = 512;
    HIPCHK(launch_gemm(w, s->stream));
    HIPCHK(launch_colsum(G, B, kMlpN[l], ldg, s->g + kOff.b[l], s->stream));
    if (l > 0) {
      unsigned short *Gn = (G == s->Ga) ? s->Gb : s->Ga;
      GemmArgs d = gemm(B, kMlpK[l], kMlpN[l], G, ldg, 1, s->wt[l], kMlpKp[l], 1, kEpBwdData);
      d.zprev = s->Z[l - 1];
      d.ldz = 128;
      d.act = s->act;
      d.Cb = Gn;
      d.ldcb = 128;
      HIPCHK(launch_gemm(d, s->stream));
      G = Gn;
      ldg = 128;
    }
  }
}

void adam(corefl_surrogate *s, float lr) {
  const float b1 = 0.9f, b2 = 0.999f, eps = 1e-8f;
  ++s->adam_t;
  const float c1 = 1.0f - std::pow(b1, (float)s->adam_t), c2 = 1.0f - std::pow(b2, (float)s->adam_t);
  for (int l = 0; l < kMlpLayers; ++l) {
    const long nw = (long)kMlpK[l] * kMlpN[l];
    HIPCHK(launch_adam(s->p + kOff.w[l], s->m + kOff.w[l], s->v + kOff.w[l], s->g + kOff.w[l], nw, lr, b1, b2, eps, c1,
                       c2, s->wt[l], kMlpK[l], kMlpN[l], kMlpKp[l], s->stream));
    HIPCHK(launch_adam(s->p + kOff.b[l], s->m + kOff.b[l], s->v + kOff.b[l], s->g + kOff.b[l], kMlpN[l], lr, b1, b2,
                       eps, c1, c2, nullptr, 0, 1, 0, s->stream));
  }
}

void upload_rows(corefl_surrogate *s, long n, const float *X, const float *Y, float **dX, float **dY) {
  dalloc(dX, (size_t)n * kMlpIn);
  dalloc(dY, (size_t)n * kMlpOut);
  HIPCHK(hipMemcpyAsync(*dX, X, sizeof(float) * n * kMlpIn, hipMemcpyHostToDevice, s->stream));
  if (Y) HIPCHK(hipMemcpyAsync(*dY, Y, sizeof(float) * n * kMlpOut, hipMemcpyHostToDevice, s->stream));
  else HIPCHK(hipMemsetAsync(*dY, 0, sizeof(float) * n * kMlpOut, s->stream));
}

void need_params(const corefl_surrogate *s) {
  if (!s->have_params) throw std::runtime_error("surrogate parameters are not set (corefl_surrogate_set_params)");
}

} // namespace

namespace corefl {

namespace {

// Inverse of the normal matrix E E^T over the species not in `held` (rows and columns e >= ne: identity) by
// Gauss-Jordan elimination with partial pivoting; false if it is singular.
// Inverse of the projection's normal matrix E E^T over the species not held at zero (rows e >= ne: identity).
// An element that no free species carries has a zero row: that row and column are replaced by the identity and
// the element's bit set in *dropped (the projection then leaves it alone, which conserves it only when the held
// species carried none of it: the kernel checks that its residual is zero).  False if the matrix is singular
// otherwise.
bool normal_inverse(const ElementMatrix &em, int ns, unsigned held, double (&inv)[4][4], unsigned *dropped) {
  unsigned drop = 0;
  for (int e = 0; e < em.ne; ++e) {
    bool carried = false;
    for (int k = 0; k < ns; ++k) carried |= !((held >> k) & 1u) && em.E[e][k] != 0.0;
    if (!carried) drop |= 1u << e;
  }
  double A[4][8] = {};
  for (int e = 0; e < 4; ++e) {
    for (int f = 0; f < 4; ++f) {
      double v = e == f ? 1.0 : 0.0;
      if (e < em.ne && f < em.ne && !((drop >> e) & 1u) && !((drop >> f) & 1u)) {
        v = 0;
        for (int k = 0; k < ns; ++k)
          if (!((held >> k) & 1u)) v += em.E[e][k] * em.E[f][k];
      }
      A[e][f] = v;
    }
    A[e][4 + e] = 1.0;
  }
  for (int k = 0; k < 4; ++k) {
    int p = k;
    for (int r = k + 1; r < 4; ++r)
      if (std::fabs(A[r][k]) > std::fabs(A[p][k])) p = r;
    if (!(std::fabs(A[p][k]) > 1e-300)) return false;
    for (int c = 0; c < 8; ++c) std::swap(A[k][c], A[p][c]);
    const double iv = 1.0 / A[k][k];
    for (int c = 0; c < 8; ++c) A[k][c] *= iv;
    for (int r = 0; r < 4; ++r) {
      if (r == k) continue;
      const double f = A[r][k];
      for (int c = 0; c < 8; ++c) A[r][c] -= f * A[k][c];
    }
  }
  for (int e = 0; e < 4; ++e)
    for (int f = 0; f < 4; ++f) inv[e][f] = A[e][4 + f];
  if (dropped) *dropped = drop;
  return true;
}

} // namespace

bool projection_tables(const ChemMech &mech, ElementMatrix &em, std::vector<double> &table) {
  if (mech.ne > 4 || mech.ns > 9) return false;
  em = ElementMatrix{};
  em.ne = mech.ne;
  for (int e = 0; e < mech.ne; ++e)
    for (int k = 0; k < mech.ns; ++k) em.E[e][k] = mech.comp[k][e] / mech.mw[k];
  unsigned drop0 = 0;
  if (!normal_inverse(em, mech.ns, 0u, em.M0inv, &drop0) || drop0 != 0) return false;
  table.assign((size_t)kProjTableEntries * kProjTableStride, 0.0);
  for (unsigned held = 0; held < (unsigned)kProjTableEntries; ++held) {
    double inv[4][4];
    unsigned drop = 0;
    double *t = table.data() + (size_t)held * kProjTableStride;
    if (!normal_inverse(em, mech.ns, held, inv, &drop)) continue;
    for (int e = 0; e < 4; ++e)
      for (int f = 0; f < 4; ++f) t[4 * e + f] = inv[e][f];
    t[16] = 1.0 + drop;
  }
  return true;
}

void surrogate_substep(corefl_chem *c, corefl_surrogate *s, const ChemBatch &b, double rtol, double atol,
                       int max_attempts, hipStream_t st) {
  need_params(s);
  if (!s->have_scaling || !s->have_gate) throw std::runtime_error("surrogate scaling and gate must be set first");
  const ChemMech &mech = c->mech;
  if (mech.ne > 4) throw std::runtime_error("surrogate projection: at most 4 elements are supported");
  ElementMatrix em{};
  std::vector<double> unused;
  if (!projection_tables(mech, em, unused)) throw std::runtime_error("surrogate projection: singular element matrix");
  em.inv_table = c->proj_inv;
  // dt of the launch: only used when the batch has neither per-cell nor device-side lengths
  const double dt = s->step_dt;
  const SurrogateWork w{c->sfeat, c->sout, c->sflag};
  HIPCHK(launch_mlp_chem(net_of(s), c->mech_dev, b, dt, s->sc, s->gate, em, w, c->list, c->list_count, s->cus, st));
  ChemBatch fb = b;
  fb.list = c->list;
  fb.n_list = c->list_count;
  // The gated cells are the stiff ones (igniting, or outside the training range) and few: the cooperative
  // kernel (lowest latency per step attempt) at two workgroups of four waves per CU.
  const ChemLaunch fallback{2, 256, 2 * s->cus, 0};
  HIPCHK(launch_chem_integrate(c->mech_dev, c->mech.ns, fb, dt, rtol, atol, max_attempts, fallback, c->counter, st));
}

} // namespace corefl

corefl_surrogate::~corefl_surrogate() {
  for (void *q : {(void *)p, (void *)m, (void *)v, (void *)g, (void *)out, (void *)Yb, (void *)G4, (void *)Ga,
                  (void *)Gb, (void *)loss, (void *)X, (void *)Y})
    if (q) (void)hipFree(q);
  for (int l = 0; l < kMlpLayers; ++l) {
    if (wt[l]) (void)hipFree(wt[l]);
    if (H[l]) (void)hipFree(H[l]);
    if (l < kMlpLayers - 1 && Z[l]) (void)hipFree(Z[l]);
  }
  if (stream) (void)hipStreamDestroy(stream);
}

extern "C" {

long long corefl_surrogate_n_params(void) { return kOff.total; }

int corefl_surrogate_create(int batch_capacity, corefl_surrogate **out) {
  return guarded([&] {
    if (!out) throw std::runtime_error("corefl_surrogate_create: null output");
    if (batch_capacity < 1 || batch_capacity > (1 << 20))
      throw std::runtime_error("corefl_surrogate_create: batch capacity must be in [1, 2^20]");
    auto s = std::make_unique<corefl_surrogate>();
    s->sc.y_power = 1.0;
    HIPCHK(hipStreamCreateWithFlags(&s->stream, hipStreamNonBlocking));
    HIPCHK(hipDeviceGetAttribute(&s->cus, hipDeviceAttributeMultiprocessorCount, 0));
    const int B = batch_capacity;
    s->cap = B;
    for (float **q : {&s->p, &s->m, &s->v, &s->g}) dalloc(q, (size_t)kOff.total);
    HIPCHK(hipMemsetAsync(s->m, 0, sizeof(float) * kOff.total, s->stream));
    HIPCHK(hipMemsetAsync(s->v, 0, sizeof(float) * kOff.total, s->stream));
    for (int l = 0; l < kMlpLayers; ++l) {
      dalloc(&s->wt[l], (size_t)kMlpNp[l] * kMlpKp[l]);
      HIPCHK(hipMemsetAsync(s->wt[l], 0, sizeof(unsigned short) * kMlpNp[l] * kMlpKp[l], s->stream));
      dalloc(&s->H[l], (size_t)B * (l == 0 ? 16 : 128));
      if (l < kMlpLayers - 1) dalloc(&s->Z[l], (size_t)B * 128);
    }
    dalloc(&s->out, (size_t)B * kMlpOut);
    dalloc(&s->Yb, (size_t)B * kMlpOut);
    dalloc(&s->G4, (size_t)B * kMlpOut);
    dalloc(&s->Ga, (size_t)B * 128);
    dalloc(&s->Gb, (size_t)B * 128);
    HIPCHK(hipStreamSynchronize(s->stream));
    *out = s.release();
  });
}

void corefl_surrogate_destroy(corefl_surrogate *s) { delete s; }

int corefl_surrogate_set_params(corefl_surrogate *s, const float *params, long long n) {
  return guarded([&] {
    if (!s || !params) throw std::runtime_error("corefl_surrogate_set_params: null argument");
    if (n != kOff.total) throw std::runtime_error("corefl_surrogate_set_params: expected " + std::to_string(kOff.total));
    HIPCHK(hipMemcpyAsync(s->p, params, sizeof(float) * n, hipMemcpyHostToDevice, s->stream));
    HIPCHK(hipMemsetAsync(s->m, 0, sizeof(float) * n, s->stream));
    HIPCHK(hipMemsetAsync(s->v, 0, sizeof(float) * n, s->stream));
    s->adam_t = 0;
    for (int l = 0; l < kMlpLayers; ++l)
      HIPCHK(launch_to_bf16t(s->p + kOff.w[l], kMlpK[l], kMlpN[l], kMlpKp[l], s->wt[l], s->stream));
    HIPCHK(hipStreamSynchronize(s->stream));
    s->have_params = true;
  });
}

int corefl_surrogate_get_params(corefl_surrogate *s, float *params, long long n) {
  return guarded([&] {
    if (!s || !params) throw std::runtime_error("corefl_surrogate_get_params: null argument");
    if (n != kOff.total) throw std::runtime_error("corefl_surrogate_get_params: expected " + std::to_string(kOff.total));
    need_params(s);
    HIPCHK(hipMemcpyAsync(params, s->p, sizeof(float) * n, hipMemcpyDeviceToHost, s->stream));
    HIPCHK(hipStreamSynchronize(s->stream));
  });
}

int corefl_surrogate_set_scaling(corefl_surrogate *s, const double *in_mu, const double *in_sig, const double *out_mu,
                                 const double *out_sig) {
  return guarded([&] {
    if (!s || !in_mu || !in_sig || !out_mu || !out_sig) throw std::runtime_error("corefl_surrogate_set_scaling: null");
    for (int q = 0; q < kMlpIn; ++q) {
      if (!(in_sig[q] > 0)) throw std::runtime_error("corefl_surrogate_set_scaling: in_sig must be positive");
      s->sc.in_mu[q] = (float)in_mu[q];
      s->sc.in_sig[q] = (float)in_sig[q];
    }
    for (int q = 0; q < kMlpOut; ++q) {
      if (!(out_sig[q] > 0)) throw std::runtime_error("corefl_surrogate_set_scaling: out_sig must be positive");
      s->sc.out_mu[q] = out_mu[q];
      s->sc.out_sig[q] = out_sig[q];
    }
    s->have_scaling = true;
  });
}

int corefl_surrogate_set_transform(corefl_surrogate *s, double y_power, const double *out_r0) {
  return guarded([&] {
    if (!s) throw std::runtime_error("corefl_surrogate_set_transform: null context");
    if (!(y_power > 0) || y_power > 1) throw std::runtime_error("corefl_surrogate_set_transform: y_power in (0, 1]");
    s->sc.y_power = y_power;
    for (int q = 0; q < kMlpOut; ++q) s->sc.out_r0[q] = out_r0 ? out_r0[q] : 0.0;
  });
}

int corefl_surrogate_set_gate(corefl_surrogate *s, const double *gate) {
  return guarded([&] {
    if (!s || !gate) throw std::runtime_error("corefl_surrogate_set_gate: null argument");
    s->gate = SurrogateGate{gate[0], gate[1], gate[2], gate[3], gate[4], gate[5], gate[6], gate[7]};
    if (!(s->gate.t_min <= s->gate.t_max && s->gate.p_min <= s->gate.p_max && s->gate.dt_min <= s->gate.dt_max &&
          s->gate.dT_max >= 0 && s->gate.dT_step >= 0))
      throw std::runtime_error("corefl_surrogate_set_gate: empty range");
    s->have_gate = true;
  });
}

int corefl_surrogate_set_activation(corefl_surrogate *s, int act) {
  return guarded([&] {
    if (!s || (act != kActGelu && act != kActHardSwish))
      throw std::runtime_error("corefl_surrogate_set_activation: act must be 0 (GELU) or 1 (hard swish)");
    s->act = act;
  });
}

int corefl_surrogate_set_data(corefl_surrogate *s, long long n, const float *X, const float *Y) {
  return guarded([&] {
    if (!s || n < 1 || !X || !Y) throw std::runtime_error("corefl_surrogate_set_data: bad arguments");
    if (s->X) (void)hipFree(s->X);
    if (s->Y) (void)hipFree(s->Y);
    s->X = s->Y = nullptr;
    s->n_data = 0;
    upload_rows(s, n, X, Y, &s->X, &s->Y);
    HIPCHK(hipStreamSynchronize(s->stream));
    s->n_data = n;
  });
}

int corefl_surrogate_train(corefl_surrogate *s, int batch, int steps, float lr, unsigned long long seed,
                           float *loss_out) {
  return guarded([&] {
    if (!s) throw std::runtime_error("corefl_surrogate_train: null context");
    need_params(s);
    if (!s->n_data) throw std::runtime_error("corefl_surrogate_train: no training data (corefl_surrogate_set_data)");
    if (batch < 1 || batch > s->cap) throw std::runtime_error("corefl_surrogate_train: batch outside [1, capacity]");
    if (steps < 0 || !(lr > 0)) throw std::runtime_error("corefl_surrogate_train: steps >= 0 and lr > 0 required");
    if (seed == 0) seed = 1;
    if (steps > s->loss_cap) {
      if (s->loss) (void)hipFree(s->loss);
      s->loss = nullptr;
      dalloc(&s->loss, (size_t)steps);
      s->loss_cap = steps;
    }
    if (steps) HIPCHK(launch_fill(s->loss, steps, 0.0f, s->stream));
    for (int st = 0; st < steps; ++st) {
      HIPCHK(launch_gather(s->X, s->Y, s->n_data, seed, s->adam_t, batch, s->H[0], s->Yb, s->stream));
      forward(s, batch);
      backward(s, batch, s->loss + st);
      adam(s, lr);
    }
    if (loss_out && steps)
      HIPCHK(hipMemcpyAsync(loss_out, s->loss, sizeof(float) * steps, hipMemcpyDeviceToHost, s->stream));
    HIPCHK(hipStreamSynchronize(s->stream));
  });
}

int corefl_surrogate_grad(corefl_surrogate *s, long long n, const float *X, const float *Y, float *loss,
                          float *grads) {
  return guarded([&] {
    if (!s || !X || !Y || !loss || !grads) throw std::runtime_error("corefl_surrogate_grad: null argument");
    need_params(s);
    if (n < 1 || n > s->cap) throw std::runtime_error("corefl_surrogate_grad: n outside [1, capacity]");
    float *dX = nullptr, *dY = nullptr, *dl = nullptr;
    try {
      upload_rows(s, n, X, Y, &dX, &dY);
      dalloc(&dl, 1);
      HIPCHK(launch_fill(dl, 1, 0.0f, s->stream));
      HIPCHK(launch_gather(dX, dY, n, 0, 0, (int)n, s->H[0], s->Yb, s->stream));
      forward(s, (int)n);
      backward(s, (int)n, dl);
      HIPCHK(hipMemcpyAsync(loss, dl, sizeof(float), hipMemcpyDeviceToHost, s->stream));
      HIPCHK(hipMemcpyAsync(grads, s->g, sizeof(float) * kOff.total, hipMemcpyDeviceToHost, s->stream));
      HIPCHK(hipStreamSynchronize(s->stream));
    } catch (...) {
      (void)hipStreamSynchronize(s->stream);
      for (float *q : {dX, dY, dl})
        if (q) (void)hipFree(q);
      throw;
    }
    for (float *q : {dX, dY, dl}) (void)hipFree(q);
  });
}

int corefl_surrogate_forward(corefl_surrogate *s, long long n, const float *X, float *out, int fused) {
  return guarded([&] {
    if (!s || !X || !out) throw std::runtime_error("corefl_surrogate_forward: null argument");
    need_params(s);
    if (n < 1 || n > (1LL << 30)) throw std::runtime_error("corefl_surrogate_forward: n outside [1, 2^30]");
    if (!fused && n > s->cap) throw std::runtime_error("corefl_surrogate_forward: n above the batch capacity");
    float *dX = nullptr, *dY = nullptr;
    try {
      upload_rows(s, n, X, nullptr, &dX, &dY);
      if (fused) {
        unsigned short *feat = nullptr;
        dalloc(&feat, (size_t)n * 16);
        const hipError_t e1 = launch_rows_to_feat(dX, n, feat, s->stream);
        const hipError_t e2 = e1 == hipSuccess ? launch_mlp_raw(net_of(s), n, feat, dY, s->cus, s->stream) : e1;
        const hipError_t e3 = hipMemcpyAsync(out, dY, sizeof(float) * n * kMlpOut, hipMemcpyDeviceToHost, s->stream);
        (void)hipStreamSynchronize(s->stream);
        (void)hipFree(feat);
        HIPCHK(e2);
        HIPCHK(e3);
      } else {
        HIPCHK(launch_gather(dX, dY, n, 0, 0, (int)n, s->H[0], s->Yb, s->stream));
        forward(s, (int)n);
        HIPCHK(hipMemcpyAsync(out, s->out, sizeof(float) * n * kMlpOut, hipMemcpyDeviceToHost, s->stream));
      }
      HIPCHK(hipStreamSynchronize(s->stream));
    } catch (...) {
      (void)hipStreamSynchronize(s->stream);
      for (float *q : {dX, dY})
        if (q) (void)hipFree(q);
      throw;
    }
    for (float *q : {dX, dY}) (void)hipFree(q);
  });
}

int corefl_chem_surrogate_step(corefl_chem *c, corefl_surrogate *s, double dt, double rtol, double atol,
                               int max_attempts, int *n_fallback, void *stream) {
  return guarded([&] {
    if (!c || !s) throw std::runtime_error("corefl_chem_surrogate_step: null context");
    need_params(s);
    if (!s->have_scaling || !s->have_gate)
      throw std::runtime_error("corefl_chem_surrogate_step: scaling and gate must be set first");
    chem_check_args(c, dt, rtol, atol, max_attempts);
    chem_ensure_device(c);
    ChemBatch b = c->b;
    b.dtc = dt > 0 ? nullptr : c->dtc;
    const hipStream_t st = chem_stream(c, stream);
    s->step_dt = dt;
    surrogate_substep(c, s, b, rtol, atol, max_attempts, st);
    if (n_fallback) {
      HIPCHK(hipMemcpyAsync(n_fallback, c->list_count, sizeof(int), hipMemcpyDeviceToHost, st));
      HIPCHK(hipStreamSynchronize(st));
    }
  });
}

} // extern "C"
