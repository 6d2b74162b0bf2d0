// C ABI implementation: setup -> mesh -> device state -> graph-captured SSP-RK3 steps.
#include "../../include/corefl_c.h"

#include <hip/hip_runtime.h>

#include <algorithm>
#include <cctype>
#include <cmath>
#include <cstring>
#include <fstream>
#include <memory>
#include <sstream>
#include <stdexcept>
#include <string>
#include <vector>

#include "bc.hpp"
#include "chem.hpp"
#include "chem_ctx.hpp"
#include "flow.hpp"
#include "flow_mix.hpp"
#include "mesh.hpp"
#include "monitor.hpp"
#include "parameter.hpp"

namespace {

thread_local std::string g_last_error;

constexpr double kGamma = 1.4, kRu = 8314.462618, kMwAir = 28.9647;

struct HipError : std::runtime_error {
  using std::runtime_error::runtime_error;
};

#define HIPCHK(x)                                                                                          \
  do {                                                                                                     \
    hipError_t e_ = (x);                                                                                   \
    if (e_ != hipSuccess) throw HipError(std::string(#x) + ": " + hipGetErrorString(e_));                 \
  } while (0)

double sutherland(double T) { return 1.7894e-5 * std::pow(T / 288.16, 1.5) * (288.16 + 110) / (T + 110); }

// Air state of an inflow-type struct (Inflow::Inflow, BoundCond.cpp:158-262, single species).
struct FlowState {
  double rho, u, v, w, p, T, c, mu;
  std::vector<double> Y; // mass fractions (mixtures only)
};

FlowState inflow_state(const corefl::Struct &s, const corefl::ChemMech *mech = nullptr) {
  using namespace corefl;
  double mach = struct_real(s, "mach", -1), p = struct_real(s, "pressure", -1);
  double T = struct_real(s, "temperature", -1), vel = struct_real(s, "velocity", 0);
  double rho = struct_real(s, "density", -1), aoa = struct_real(s, "aoa", -1000);
  double u = struct_real(s, "u", 1), v = struct_real(s, "v", 0), w = struct_real(s, "w", 0);
  const double re = struct_real(s, "reynolds", -1);
  double mw = kMwAir, gamma = kGamma;
  std::vector<double> Y;
  if (mech) {
    // species mass fractions by name (Inflow::Inflow, BoundCond.cpp:158-262)
    Y.assign(mech->ns, 0.0);
    double sum = 0, imw = 0;
    for (int k = 0; k < mech->ns; ++k) {
      Y[k] = struct_real(s, mech->species[k], 0.0);
      sum += Y[k];
      imw += Y[k] / mech->mw[k];
    }
    if (!(sum > 0)) throw std::runtime_error("inflow struct of a mixture: give the species mass fractions");
    for (auto &y : Y) y /= sum;
    mw = sum / imw;
  }
  if (T < 0) T = p * mw / (rho * kRu);
  if (mech) {
    double h[kMaxSpec], cp[kMaxSpec], cpt = 0;
    host_h_cp(*mech, T, h, cp);
    for (int k = 0; k < mech->ns; ++k) cpt += Y[k] * cp[k];
    gamma = cpt / (cpt - kRu / mw);
  }
  const double mu = sutherland(T);
  const double c = std::sqrt(gamma * kRu / mw * T);
  if (mach < 0) mach = vel / c;
  else vel = mach * c;
  if (aoa > -1000) {
    u = std::cos(aoa * M_PI / 180) * vel;
    v = std::sin(aoa * M_PI / 180) * vel;
    w = 0;
  } else {
    u *= vel;
    v *= vel;
    w *= vel;
  }
  if (rho < 0) {
    if (p < 0) {
      if (re < 0) throw std::runtime_error("inflow struct: need density, pressure or reynolds");
      rho = mu * re / vel;
      p = rho * T * kRu / mw;
    } else {
      rho = p * mw / (kRu * T);
    }
  }
  return {rho, u, v, w, p, T, c, mu, Y};
}

bool is_mixing_layer(const corefl::Parameter &P) {
  std::string cp = P.get_string("canonical_problem");
  for (auto &ch : cp) ch = (char)std::toupper((unsigned char)ch);
  return cp == "MIXINGLAYER"; // Parameter.cpp:617-625: problem_type 1
}

// The two streams of a mixing layer (get_mixing_layer_info, MixingLayer.cu:9-130): temperature, pressure and
// mass fractions of the structs upper_stream / lower_stream; their Mach numbers given, or derived from the
// convective Mach number ma_c with velocity_ratio or one given Mach number.  Velocity along x only.
struct MixingLayer {
  FlowState upper, lower;
  double delta_omega, delta_u, uc;
};

MixingLayer mixing_layer(const corefl::Parameter &P, const corefl::ChemMech *mech) {
  using namespace corefl;
  const auto &up = P.get_struct("upper_stream"), &lo = P.get_struct("lower_stream");
  auto stream = [&](const Struct &s, double &c) {
    FlowState f{};
    f.T = struct_real(s, "temperature", -1);
    f.p = struct_real(s, "pressure", -1);
    if (!(f.T > 0) || !(f.p > 0)) throw std::runtime_error("mixing layer streams need temperature and pressure");
    double imw = 1.0 / kMwAir, gamma = kGamma;
    if (mech) {
      f.Y.assign(mech->ns, 0.0);
      imw = 0;
      double h[kMaxSpec], cp[kMaxSpec], cpt = 0;
      host_h_cp(*mech, f.T, h, cp);
      for (int k = 0; k < mech->ns; ++k) {
        f.Y[k] = struct_real(s, mech->species[k], 0.0);
        imw += f.Y[k] / mech->mw[k];
        cpt += cp[k] * f.Y[k];
      }
      if (!(imw > 0)) throw std::runtime_error("mixing layer streams of a mixture need species mass fractions");
      gamma = cpt / (cpt - kRu * imw);
    }
    c = std::sqrt(gamma * kRu * imw * f.T);
    f.rho = f.p / (kRu * imw * f.T);
    f.c = c;
    return f;
  };
  double cu = 0, cl = 0;
  MixingLayer ml{stream(up, cu), stream(lo, cl), P.get_real("delta_omega"), 0, 0};
  const double ma_c = P.get_real("ma_c"), vr = P.get_real("velocity_ratio");
  const bool upper_faster = P.get_bool("upper_faster");
  double mu = struct_real(up, "mach", -1), ml_ = struct_real(lo, "mach", -1);
  if (vr > 0) {
    double slow, fast;
    if (vr > 1) {
      slow = ma_c * (cu + cl) / (vr - 1);
      fast = vr * slow;
    } else {
      fast = ma_c * (cu + cl) / (1 - vr);
      slow = fast * vr;
    }
    mu = (upper_faster ? fast : slow) / cu;
    ml_ = (upper_faster ? slow : fast) / cl;
  } else if (mu < 0 && ml_ < 0) {
    throw std::runtime_error("mixing layer: give a Mach number of one stream, or velocity_ratio with ma_c");
  } else if (mu > 0 && ml_ < 0) {
    ml_ = upper_faster ? (mu - ma_c) * cu / cl - ma_c : (ma_c + mu) * cu / cl + ma_c;
  } else if (mu < 0) {
    mu = upper_faster ? (ma_c + ml_) * cl / cu + ma_c : (ml_ - ma_c) * cl / cu - ma_c;
  }
  ml.upper.u = mu * cu;
  ml.lower.u = ml_ * cl;
  ml.upper.mu = ml.lower.mu = 0;
  // Driver.cu:286-293: convective velocity and velocity difference
  ml.uc = (ml.upper.u * cl + ml.lower.u * cu) / (cu + cl);
  ml.delta_u = std::abs(ml.upper.u - ml.lower.u);
  return ml;
}

struct Block {
  corefl::HostBlock host;
  corefl::DevBlock dev{};
  corefl::MixArrays mix{}; // species arrays (ns = 0: single-species air)
  double *hstep = nullptr; // per-cell chemistry step size carried between sub-steps
  std::vector<corefl::DevFace> faces; // in boundary-file order
  std::vector<corefl::DevInner> inner; // block-to-block faces, in inner-file order
  std::vector<int> inner_target;       // their partner blocks
  long part_offset = 0;
  int part_count = 0;
  std::vector<corefl::DevInflow> inflow; // inflow data per face (used for faces of type 5)
  double *sponge_sigma = nullptr, *sponge_mean = nullptr;
  double *last = nullptr;          // rho, |V|, p, T at the start of a monitored step (4 * n)
  unsigned char *lflag = nullptr;  // limit_flow flags (n)
  long res_offset = 0;             // this block's residual partials
  std::vector<double> coords[3];   // x, y, z of the block's nodes (i fastest), for flowfield output
};

} // namespace

struct corefl_ctx {
  corefl::Parameter param;
  corefl::Scheme scheme{};
  std::vector<Block> blocks;
  hipStream_t stream = nullptr;
  // the viscous fluxes of a stage run on `side` while the inviscid fluxes run on the caller's stream (fork and
  // join through the two events; inside a graph capture they become graph edges)
  hipStream_t side = nullptr;
  hipEvent_t ev_fork = nullptr, ev_join = nullptr;
  static constexpr int kMaxSlabs = 64;
  hipEvent_t ev_slab[kMaxSlabs] = {}; // the inviscid fluxes of z-slab q are done (air stages in slabs)
  int slab_setting = -1;              // stage_slab_planes
  double *dt_dev = nullptr, *time_dev = nullptr, *parts = nullptr;
  long n_parts = 0;
  bool fixed_dt = false;
  double t_end = 1e10;
  double host_dt = 0; // source of host-to-device copies of the step size
  double *sub_dt_dev = nullptr; // length of a caller's chemistry sub-step (corefl_chem_substep)
  double sub_dt = 0;
  hipGraphExec_t graph_exec = nullptr;
  hipGraph_t graph = nullptr;
  // step monitor (compute_residual, Residual.cu:8-121) and limit_flow (TimeAdvanceFunc.cuh:117-220)
  int output_screen = 0;            // residual on step 1 and every output_screen steps; 0: never
  bool limit_flow = false;
  corefl::FlowLimits limits{};
  corefl::MonitorState *mon = nullptr;      // device
  corefl::MonitorState *mon_host = nullptr; // pinned copy read after a monitored step
  double *res_parts = nullptr;
  long n_res_parts = 0;
  double n_nodes_total = 0;
  long step_count = 0;              // steps taken (the reference's step counter)
  unsigned long long *sponge_iter = nullptr; // steps averaged into the sponge layer's mean (device)
  int n_inflow = 0;                           // inflow faces so far (each draws its own noise stream)
  hipGraphExec_t graph_exec_mon = nullptr; // the captured monitored step
  hipGraph_t graph_mon = nullptr;
  std::vector<void *> allocations;
  // mixtures
  int ns = 0;
  bool strang = false;          // temporal_scheme 5: chemistry dt/2, SSP-RK3, chemistry dt/2
  int wu_method = -1;           // temporal_scheme 4 (Wu's splitting): chem_advance_method 0, 1 or 3
  corefl_chem *chem = nullptr;  // mechanism, device tables and the cell batch of the chemistry sub-steps
  corefl_surrogate *surrogate = nullptr; // not owned; when set, sub-steps use it with the exact fallback
  double chem_rtol = 1e-6, chem_atol = 1e-8;
  int chem_max_attempts = 100000;
  std::string input_dir;

  double *alloc(size_t n) {
    void *p = nullptr;
    HIPCHK(hipMalloc(&p, n * sizeof(double)));
    HIPCHK(hipMemset(p, 0, n * sizeof(double)));
    allocations.push_back(p);
    return static_cast<double *>(p);
  }
  ~corefl_ctx() {
    delete chem;
    drop_graph();
    if (mon_host) (void)hipHostFree(mon_host);
    for (void *p : allocations) (void)hipFree(p);
    for (hipEvent_t e : ev_slab)
      if (e) (void)hipEventDestroy(e);
    if (ev_fork) (void)hipEventDestroy(ev_fork);
    if (ev_join) (void)hipEventDestroy(ev_join);
    if (side) (void)hipStreamDestroy(side);
    if (stream) (void)hipStreamDestroy(stream);
  }

  hipStream_t pick(void *s) const { return s ? static_cast<hipStream_t>(s) : stream; }

  // Forget the captured step (its executable graph and the graph it was instantiated from).
  void drop_graph() {
    if (graph_exec) (void)hipGraphExecDestroy(graph_exec);
    if (graph) (void)hipGraphDestroy(graph);
    if (graph_exec_mon) (void)hipGraphExecDestroy(graph_exec_mon);
    if (graph_mon) (void)hipGraphDestroy(graph_mon);
    graph_exec = nullptr;
    graph = nullptr;
    graph_exec_mon = nullptr;
    graph_mon = nullptr;
  }

  // The reference stores the state at steps that are multiples of output_screen and computes the residual
  // after those steps and after step 1 (RK.cuh:92-96, 196-205).
  bool monitored(long step) const { return output_screen > 0 && (step == 1 || step % output_screen == 0); }
  bool stores_last(long step) const { return output_screen > 0 && step % output_screen == 0; }

  void limit(Block &b, hipStream_t s) {
    if (limit_flow) HIPCHK(corefl::launch_limit_flow(b.dev, b.mix, limits, b.lflag, s));
  }

  void apply_bcs(Block &b, hipStream_t s) {
    // DBoundCond::apply_boundary_conditions order (BoundCond.cu:3489-3619): outflow, inflow, symmetry, periodic.
    for (int type : {6, 5, 3, 8})
      for (size_t q = 0; q < b.faces.size(); ++q) {
        const auto &f = b.faces[q];
        if (f.type != type) continue;
        if (type == 5) HIPCHK(corefl::launch_inflow(b.dev, b.mix, b.inflow[q], s));
        else if (type == 3) HIPCHK(corefl::launch_symmetry(b.dev, b.mix, f, s));
        else if (ns) HIPCHK(corefl::launch_mix_boundary(b.dev, b.mix, f, s));
        else HIPCHK(corefl::launch_boundary(b.dev, f, s));
      }
  }

  // data_communication (DataCommunication.cu:6-34): every block-to-block face in block and file order, one
  // launch each, as the reference averages coincident nodes face after face.
  void communicate(hipStream_t s) {
    for (auto &b : blocks)
      for (size_t l = 0; l < b.inner.size(); ++l) {
        Block &t = blocks[b.inner_target[l]];
        if (ns) HIPCHK(corefl::launch_mix_inner_exchange(b.dev, b.mix, t.dev, t.mix, b.inner[l], s));
        else HIPCHK(corefl::launch_inner_exchange(b.dev, t.dev, b.inner[l], s));
      }
  }

  // A state set block by block is exchanged across the block-to-block faces before the next computation,
  // once every block holds its new state (the reference communicates once after initialising all blocks,
  // Driver.cu:160-190).
  bool exchange_pending = false;
  void flush_exchange(hipStream_t s) {
    if (!exchange_pending) return;
    communicate(s);
    exchange_pending = false;
  }

  // exchange_value task 1 (DataCommunication.cu:408-456): block-to-block faces, then periodic faces.
  void viscous_exchange(hipStream_t s) {
    if (scheme.viscous_order == 2) return; // face fluxes come from the exchanged state itself
    const int nv = 5 + ns;
    for (auto &b : blocks)
      for (size_t l = 0; l < b.inner.size(); ++l)
        HIPCHK(corefl::launch_inner_viscous_exchange(b.dev, blocks[b.inner_target[l]].dev, b.inner[l], s, nv));
    for (auto &b : blocks)
      for (const auto &f : b.faces)
        if (f.type == 8) HIPCHK(corefl::launch_viscous_exchange(b.dev, f, s, nv));
  }

  // Chemistry over scale * (*dtp) (a device scalar) in every interior cell of every block (constant density
  // and internal energy), exact or through the surrogate with its gate and exact fallback, then boundary
  // conditions and the block-to-block exchange.
  void chem_substep(double scale, hipStream_t s, const double *dtp = nullptr, int use_surrogate = -1) {
    const bool surr = use_surrogate < 0 ? surrogate != nullptr : use_surrogate != 0;
    for (auto &b : blocks) {
      corefl::chem_ensure_batch(chem, b.dev.n);
      HIPCHK(corefl::launch_mix_chem_gather(b.dev, b.mix, chem->b, s));
      corefl::ChemBatch cb = chem->b;
      cb.h = b.hstep; // this block's carried step sizes
      cb.dt_dev = dtp ? dtp : dt_dev;
      cb.dt_scale = scale;
      if (surr)
        corefl::surrogate_substep(chem, surrogate, cb, chem_rtol, chem_atol, chem_max_attempts, s);
      else
        HIPCHK(corefl::launch_chem_integrate(chem->mech_dev, ns, cb, 0.0, chem_rtol, chem_atol, chem_max_attempts,
                                             chem->cfg, chem->counter, s));
      HIPCHK(corefl::launch_mix_chem_scatter(b.dev, b.mix, chem->b, s));
      apply_bcs(b, s);
    }
    communicate(s);
  }

  // Residual of one stage for all blocks; stage < 0 writes dq only.
  void stage(int rk, hipStream_t s) {
    const bool rhs_only = rk < 0;
    // compute_shock_sensor once per step, before its first residual (RK.cuh:112-117)
    if (rk <= 0)
      for (auto &b : blocks) HIPCHK(corefl::launch_shock_sensor(b.dev, scheme, s));
    if (!ns && slab_planes() > 0) {
      air_stage_in_slabs(rk, s);
    } else {
      fluxes_then(s, [&](Block &b) {
        if (ns) HIPCHK(corefl::launch_mix_update(b.dev, b.mix, rhs_only ? 0 : rk, dt_dev, rhs_only, s));
        else HIPCHK(corefl::launch_update(b.dev, scheme, rhs_only ? 0 : rk, dt_dev, rhs_only, s));
        if (!rhs_only) {
          limit(b, s);
          apply_bcs(b, s);
        }
      });
    }
    if (!rhs_only) communicate(s);
  }

  // Slab thickness (planes) of the air stages in z-slabs (setup key stage_slab_planes; -1: two march lengths);
  // 0 when that is 0 or a block is two-dimensional or too thin.
  int slab_planes() const {
    const int P = slab_setting < 0 ? 2 * corefl::march_planes() : slab_setting;
    if (P == 0) return 0;
    for (const auto &b : blocks)
      if (b.dev.dim != 3 || b.dev.mz < 2 * P || (b.dev.mz + P - 1) / P > kMaxSlabs) return 0;
    return P;
  }

  // One air stage with the blocks cut into z-slabs of P planes: the inviscid fluxes of the slabs run on the side
  // stream, one slab after another, while s runs the viscous fluxes and then the update of each slab as soon as
  // its fluxes are there, so the FP64-bound flux kernels and the HBM-bound update run at the same time.  The
  // update of slab q overwrites its planes, which the z-fluxes of slab q + 1 still read (the WENO7 stencil of
  // the interface k + 1/2 spans planes k - 3 .. k + 4), so it waits for slab q + 1's fluxes; the fluxes of later
  // slabs read no plane of slab q.  The same kernels compute the same values as the unsliced stage.
  void air_stage_in_slabs(int rk, hipStream_t s) {
    const bool rhs_only = rk < 0;
    const int P = slab_planes();
    HIPCHK(hipEventRecord(ev_fork, s));
    HIPCHK(hipStreamWaitEvent(side, ev_fork, 0));
    if (scheme.viscous_order > 0) {
      for (auto &b : blocks) HIPCHK(corefl::launch_viscous_flux(b.dev, scheme, s));
      viscous_exchange(s);
    }
    for (auto &b : blocks) {
      const int n_slab = (b.dev.mz + P - 1) / P;
      if (scheme.inviscid_scheme == 2) HIPCHK(corefl::launch_entropy_fix(b.dev, scheme, side));
      for (int q = 0; q < n_slab; ++q) {
        for (int d = 0; d < 3; ++d) HIPCHK(corefl::launch_inviscid_flux(b.dev, scheme, d, side, q * P, (q + 1) * P));
        HIPCHK(hipEventRecord(ev_slab[q], side));
      }
      for (int q = 0; q < n_slab; ++q) {
        HIPCHK(hipStreamWaitEvent(s, ev_slab[std::min(q + 1, n_slab - 1)], 0));
        HIPCHK(corefl::launch_update(b.dev, scheme, rhs_only ? 0 : rk, dt_dev, rhs_only, s, q * P, (q + 1) * P));
      }
      if (!rhs_only) {
        limit(b, s);
        apply_bcs(b, s);
      }
    }
  }

  void enqueue_dt(hipStream_t s, int clip) {
    for (auto &b : blocks) {
      if (ns && b.mix.mul) HIPCHK(corefl::launch_mix_transport(b.dev, b.mix, s));
      if (ns) HIPCHK(corefl::launch_mix_local_dt(b.dev, b.mix, scheme, parts + b.part_offset, s));
      else HIPCHK(corefl::launch_local_dt(b.dev, scheme, parts + b.part_offset, s));
    }
    HIPCHK(corefl::launch_dt_reduce(parts, n_parts, dt_dev, time_dev, t_end, clip, s));
  }

  // The viscous fluxes of every block (and their exchange for the collocated scheme) on the side stream, the
  // inviscid fluxes of every block on s at the same time (they read the same state and write other arrays),
  // then `update` of every block on s once both are done.  The latency-bound viscous kernels fill the gaps the
  // FP64-bound inviscid flux kernels leave.
  template<class F> void fluxes_then(hipStream_t s, F &&update) {
    const bool visc = scheme.viscous_order > 0;
    if (visc) {
      HIPCHK(hipEventRecord(ev_fork, s));
      HIPCHK(hipStreamWaitEvent(side, ev_fork, 0));
      for (auto &b : blocks) {
        if (ns) HIPCHK(corefl::launch_mix_viscous(b.dev, b.mix, scheme, side));
        else HIPCHK(corefl::launch_viscous_flux(b.dev, scheme, side));
      }
      viscous_exchange(side);
      HIPCHK(hipEventRecord(ev_join, side));
    }
    for (auto &b : blocks) {
      if (ns) {
        if (scheme.inviscid_scheme == 2) HIPCHK(corefl::launch_mix_entropy_fix(b.dev, b.mix, scheme, s));
        for (int d = 0; d < b.dev.dim; ++d) HIPCHK(corefl::launch_mix_inviscid_flux(b.dev, b.mix, scheme, d, s));
      } else {
        if (scheme.inviscid_scheme == 2) HIPCHK(corefl::launch_entropy_fix(b.dev, scheme, s));
        for (int d = 0; d < b.dev.dim; ++d) HIPCHK(corefl::launch_inviscid_flux(b.dev, scheme, d, s));
      }
    }
    if (visc) HIPCHK(hipStreamWaitEvent(s, ev_join, 0));
    for (auto &b : blocks) update(b);
  }

  // wu_splitting, one step (StrangSplitting.cu:683-861): the transport residual Cn at U^n (compute_cn) kept in
  // dq; the per-cell reaction sub-step with Cn frozen (wu_reaction_subStep); then three SSP-RK3 stages over
  // dt/2 of R(U) - Cn (wu_RK_subStep).  Boundary conditions and block exchanges after each part.
  void wu_step(hipStream_t s) {
    // compute_cn computes the shock sensor (StrangSplitting.cu:19-24); the transport stages keep it
    for (auto &b : blocks) HIPCHK(corefl::launch_shock_sensor(b.dev, scheme, s));
    fluxes_then(s, [&](Block &b) { HIPCHK(corefl::launch_mix_update(b.dev, b.mix, 0, dt_dev, true, s)); });
    for (auto &b : blocks) {
      HIPCHK(corefl::launch_mix_wu_reaction(b.dev, b.mix, dt_dev, wu_method, s));
      apply_bcs(b, s);
    }
    communicate(s);
    for (int rk = 0; rk < 3; ++rk) {
      fluxes_then(s, [&](Block &b) {
        HIPCHK(corefl::launch_mix_wu_update(b.dev, b.mix, rk, dt_dev, s));
        apply_bcs(b, s);
      });
      communicate(s);
    }
  }

  // One step; store: save rho, |V|, p, T first; residual: the monitor's reduction after the step.
  void enqueue_step(hipStream_t s, bool compute_dt, bool store = false, bool residual = false) {
    if (store)
      for (auto &b : blocks) HIPCHK(corefl::launch_store_last(b.dev, b.last, s));
    if (sponge_iter) // the sponge layer's running mean at the start of the step (RK.cuh:105-110)
      for (size_t q = 0; q < blocks.size(); ++q)
        HIPCHK(corefl::launch_sponge_mean(blocks[q].dev, 5 + ns, blocks[q].sponge_mean, sponge_iter,
                                          q + 1 == blocks.size(), s));
    if (compute_dt) enqueue_dt(s, 1);
    if (wu_method >= 0) {
      wu_step(s);
    } else {
      if (strang) chem_substep(0.5, s);
      for (int rk = 0; rk < 3; ++rk) stage(rk, s);
      if (strang) chem_substep(0.5, s);
    }
    HIPCHK(corefl::launch_time_advance(time_dev, dt_dev, s));
    if (residual) {
      for (auto &b : blocks) HIPCHK(corefl::launch_residual_partial(b.dev, b.last, res_parts + 4 * b.res_offset, s));
      HIPCHK(corefl::launch_residual_finish(res_parts, n_res_parts, n_nodes_total, mon, s));
    }
  }

  // After a monitored step: wait for it and stop on a NaN residual, as compute_residual does (Residual.cu:100-118).
  void check_monitor(hipStream_t s, long step) {
    HIPCHK(hipMemcpyAsync(mon_host, mon, sizeof(corefl::MonitorState), hipMemcpyDeviceToHost, s));
    HIPCHK(hipStreamSynchronize(s));
    if (mon_host->nan) throw std::runtime_error("NaN occurred in step " + std::to_string(step) + ". Stop simulation.");
  }
};

namespace {

std::string read_file(const std::string &path, bool binary = false) {
  std::ifstream f(path, binary ? std::ios::binary : std::ios::in);
  if (!f) throw std::runtime_error("cannot open " + path);
  std::stringstream ss;
  ss << f.rdbuf();
  return ss.str();
}

// Initial primitive state on every node of one block, ghosts included (Initialize.cu:51-95 and
// initialize_bv_with_inflow in Field.cu:280-340; canonical_problem "taylorGreen" is new).
std::vector<double> initial_bv(const corefl::Parameter &P, const corefl::HostBlock &hb,
                               const corefl::ChemMech *mech = nullptr, std::vector<double> *sv = nullptr) {
  const long sz = hb.size();
  std::vector<double> bv(6 * sz, 0.0);
  if (mech) sv->assign((size_t)mech->ns * sz, 0.0);
  const int g = hb.ngg, G = g + 1;
  const long c2 = hb.mx + 2L * G, c1 = (hb.my + 2L * G) * c2;
  auto coord = [&](const std::vector<double> &a, int i, int j, int k) {
    return a[(k + G) * c1 + (j + G) * c2 + (i + G)];
  };
  auto put = [&](long id, const FlowState &s) {
    if (mech)
      for (int l = 0; l < mech->ns; ++l) (*sv)[id + l * sz] = s.Y[l];
    bv[id] = s.rho;
    bv[id + sz] = s.u;
    bv[id + 2 * sz] = s.v;
    bv[id + 3 * sz] = s.w;
    bv[id + 4 * sz] = s.p;
    bv[id + 5 * sz] = s.T;
  };
  const std::string problem = P.get_string("canonical_problem");
  if (problem == "taylorGreen") {
    if (mech) throw std::runtime_error("canonical_problem taylorGreen is single-species");
    const auto &s = P.get_struct("taylorGreen");
    const double T0 = corefl::struct_real(s, "temperature", 300.0);
    const double p0 = corefl::struct_real(s, "pressure", 101325.0);
    const double mach = corefl::struct_real(s, "mach", 0.1);
    const double R = kRu / kMwAir, rho0 = p0 / (R * T0), V0 = mach * std::sqrt(kGamma * R * T0);
    const double L = P.get_real("gridScale");
    for (int k = -g; k < hb.mz + g; ++k)
      for (int j = -g; j < hb.my + g; ++j)
        for (int i = -g; i < hb.mx + g; ++i) {
          const double X = coord(hb.x, i, j, k) / L, Y = coord(hb.y, i, j, k) / L, Z = coord(hb.z, i, j, k) / L;
          const double p = p0 + rho0 * V0 * V0 / 16 * (std::cos(2 * X) + std::cos(2 * Y)) * (std::cos(2 * Z) + 2);
          FlowState st{p / (R * T0), V0 * std::sin(X) * std::cos(Y) * std::cos(Z),
                       -V0 * std::cos(X) * std::sin(Y) * std::cos(Z), 0.0, p, T0, 0, 0};
          put(hb.idx(i, j, k), st);
        }
    return bv;
  }
  if (is_mixing_layer(P)) {
    // initialize_mixing_layer_with_info (Initialize.cu:970-1035): tanh profile of u across y = 0, every other
    // variable the upper stream's for y >= 0 and the lower stream's below, on every node including ghosts
    const MixingLayer ml = mixing_layer(P, mech);
    for (int k = -g; k < hb.mz + g; ++k)
      for (int j = -g; j < hb.my + g; ++j)
        for (int i = -g; i < hb.mx + g; ++i) {
          const double y = coord(hb.y, i, j, k);
          FlowState st = y >= 0 ? ml.upper : ml.lower;
          st.u = 0.5 * (ml.upper.u + ml.lower.u) + 0.5 * (ml.upper.u - ml.lower.u) * std::tanh(2 * y / ml.delta_omega);
          st.v = st.w = 0;
          put(hb.idx(i, j, k), st);
        }
    return bv;
  }
  const int groups = P.get_int("groups_init");
  std::vector<FlowState> states{inflow_state(P.get_struct(P.get_string("default_init")), mech)};
  std::vector<double> boxes;
  for (int l = 0; l < groups - 1; ++l) {
    const auto &pc = P.get_struct("init_cond_" + std::to_string(l));
    for (const char *key : {"x0", "x1", "y0", "y1", "z0", "z1"}) boxes.push_back(corefl::struct_real(pc, key, 0));
    const std::string name = corefl::struct_string(pc, "name", "init_cond_" + std::to_string(l));
    states.push_back(inflow_state(P.get_struct(name), mech));
  }
  for (int k = -g; k < hb.mz + g; ++k)
    for (int j = -g; j < hb.my + g; ++j)
      for (int i = -g; i < hb.mx + g; ++i) {
        const double x = coord(hb.x, i, j, k), y = coord(hb.y, i, j, k), z = coord(hb.z, i, j, k);
        int pick = 0;
        for (int l = 0; l < groups - 1; ++l) {
          const double *r = &boxes[6 * l];
          if (x >= r[0] && x <= r[1] && y >= r[2] && y <= r[3] && z >= r[4] && z <= r[5]) {
            pick = l + 1;
            break;
          }
        }
        put(hb.idx(i, j, k), states[pick]);
      }
  return bv;
}

// The struct of the boundary condition with this label and type (BoundCond.cpp registers boundary_conditions by
// label).
const corefl::Struct &bc_struct(const corefl::Parameter &P, int label, const char *type) {
  for (const auto &n : P.get_string_array("boundary_conditions")) {
    if (!P.has_struct(n)) continue;
    const auto &st = P.get_struct(n);
    if (corefl::struct_string(st, "type", "") == type && corefl::struct_int(st, "label", -1) == label) return st;
  }
  throw std::runtime_error(std::string("no ") + type + " struct with label " + std::to_string(label) +
                           " in boundary_conditions");
}

// Inflow face data (Inflow::Inflow, BoundCond.cpp:9-330, and the values apply_inflow assigns): the state of every
// face node (the mixing layer's tanh profile of the two streams, or the struct's constant state), the noise level
// per node for fluctuation_type 1, and the subsonic options.
corefl::DevInflow make_inflow(corefl_ctx *ctx, const corefl::Parameter &P, const corefl::HostBlock &h,
                              const corefl::Face &f, const corefl::ChemMech *mech) {
  using namespace corefl;
  const Struct &st = bc_struct(P, f.type, "inflow");
  DevInflow d{};
  for (int q = 0; q < 3; ++q) { d.start[q] = f.start[q]; d.end[q] = f.end[q]; }
  d.face = f.face;
  d.direction = f.direction;
  const bool ml = is_mixing_layer(P); // inflow_type 2 (BoundCond.cpp:13-16)
  if (struct_int(st, "inflow_type", 0) == 1 && !ml)
    throw std::runtime_error("profile inflow (inflow_type 1) is not on this path");
  const int sub = struct_int(st, "inflow_sub", 0), fluct = struct_int(st, "fluctuation_type", 0);
  if (fluct != 0 && fluct != 1) throw std::runtime_error("inflow fluctuation_type must be 0 or 1 (white noise)");
  const double intensity = struct_real(st, "fluctuation_intensity", 0.0);
  MixingLayer mlay{};
  FlowState cst{};
  if (ml) mlay = mixing_layer(P, mech);
  else cst = inflow_state(st, mech);
  const FlowState &own = ml ? mlay.upper : cst;
  double imw = 1.0 / kMwAir;
  if (mech) {
    imw = 0;
    for (int k = 0; k < mech->ns; ++k) imw += own.Y[k] / mech->mw[k];
  }
  d.mw = 1.0 / imw;
  d.sub_node = !ml && sub;
  d.sub_ghost = sub;
  d.fluct = fluct == 1 ? (ml ? 2 : 1) : 0;
  const int ns = mech ? mech->ns : 0;
  const long n0 = f.end[0] - f.start[0] + 1, n1 = f.end[1] - f.start[1] + 1, n2 = f.end[2] - f.start[2] + 1;
  const long npt = n0 * n1 * n2;
  std::vector<double> prof((6 + ns) * npt), rms(npt, 0.0);
  const int G = h.ngg + 1;
  const long c2 = h.mx + 2L * G, c1 = (h.my + 2L * G) * c2;
  const double vel = std::sqrt(cst.u * cst.u + cst.v * cst.v + cst.w * cst.w);
  for (long node = 0; node < npt; ++node) {
    const int i = f.start[0] + (int)(node % n0), j = f.start[1] + (int)((node / n0) % n1),
              k = f.start[2] + (int)(node / (n0 * n1));
    FlowState s0 = cst;
    if (ml) {
      const double y = h.y[(k + G) * c1 + (j + G) * c2 + (i + G)];
      s0 = y >= 0 ? mlay.upper : mlay.lower;
      s0.u = 0.5 * (mlay.upper.u + mlay.lower.u) + 0.5 * (mlay.upper.u - mlay.lower.u) * std::tanh(2 * y / mlay.delta_omega);
      s0.v = s0.w = 0;
      rms[node] = intensity * s0.u * std::exp(-y * y);
    } else {
      rms[node] = intensity * vel;
    }
    const double q[6] = {s0.rho, s0.u, s0.v, s0.w, s0.p, s0.T};
    for (int c = 0; c < 6; ++c) prof[c * npt + node] = q[c];
    for (int l = 0; l < ns; ++l) prof[(6 + l) * npt + node] = s0.Y[l];
  }
  double *pd = ctx->alloc(prof.size()), *rd = ctx->alloc(npt);
  HIPCHK(hipMemcpy(pd, prof.data(), prof.size() * sizeof(double), hipMemcpyHostToDevice));
  HIPCHK(hipMemcpy(rd, rms.data(), rms.size() * sizeof(double), hipMemcpyHostToDevice));
  d.prof = pd;
  d.rms = rd;
  d.calls = reinterpret_cast<unsigned long long *>(ctx->alloc(1));
  d.seed = (unsigned long long)P.get_int("fluctuation_seed") * 0x100000001b3ULL + (unsigned long long)ctx->n_inflow++;
  return d;
}

// Sponge layer (the reference's design, SpongeLayer.cu:48-172 and DParameter.cu:360-395, whose code is commented
// out): sigma = sum over the active sides of sigma_side S(xi), S(xi) = 6 xi^5 - 15 xi^4 + 10 xi^3, xi the
// normalised depth into the layer (clipped at 1), sigma_side = spongeCoefficient v_char / layer thickness; the
// residual gets -J sigma (U - U_mean) with U_mean the running mean of U over the steps.
void setup_sponge(corefl_ctx *ctx, const corefl::Parameter &P, Block &B, const corefl::ChemMech *mech) {
  using corefl::HostBlock;
  const HostBlock &h = B.host;
  if (P.get_int("sponge_function") != 0) throw std::runtime_error("sponge_function must be 0");
  double v_char = P.get_real("v_inf");
  if (is_mixing_layer(P)) {
    const MixingLayer ml = mixing_layer(P, mech);
    v_char = P.get_int("characteristic_velocity_ml") == 0 ? ml.uc : ml.delta_u;
  }
  const double coef = P.get_real("spongeCoefficient");
  struct Side {
    int axis, mode;
    double a, b;
  };
  std::vector<Side> sides;
  const char *axes[3] = {"X", "Y", "Z"};
  for (int d = 0; d < 3; ++d) {
    const int mode = P.get_int(std::string("sponge") + axes[d]);
    if (mode == 1 || mode == 3)
      sides.push_back({d, -1, P.get_real(std::string("sponge") + axes[d] + "MinusStart"),
                       P.get_real(std::string("sponge") + axes[d] + "MinusEnd")});
    if (mode == 2 || mode == 3)
      sides.push_back({d, 1, P.get_real(std::string("sponge") + axes[d] + "PlusStart"),
                       P.get_real(std::string("sponge") + axes[d] + "PlusEnd")});
  }
  for (const auto &sd : sides)
    if (!(std::abs(sd.b - sd.a) > 0)) throw std::runtime_error("sponge layer of zero thickness");
  const long n = (long)h.mx * h.my * h.mz;
  std::vector<double> sigma(n, 0.0);
  const int G = h.ngg + 1;
  const long c2 = h.mx + 2L * G, c1 = (h.my + 2L * G) * c2;
  for (int k = 0; k < h.mz; ++k)
    for (int j = 0; j < h.my; ++j)
      for (int i = 0; i < h.mx; ++i) {
        const long gi = (k + G) * c1 + (j + G) * c2 + (i + G);
        const double xyz[3] = {h.x[gi], h.y[gi], h.z[gi]};
        double sg = 0;
        for (const auto &sd : sides) {
          const double q = xyz[sd.axis];
          if (sd.mode < 0 ? !(q < sd.a) : !(q > sd.a)) continue;
          double xi = (q - sd.a) / (sd.b - sd.a);
          if (xi > 1) xi = 1;
          const double x3 = xi * xi * xi;
          sg += coef * v_char / std::abs(sd.b - sd.a) * (6 * x3 * xi * xi - 15 * x3 * xi + 10 * x3);
        }
        sigma[((long)k * h.my + j) * h.mx + i] = sg;
      }
  B.sponge_sigma = ctx->alloc(n);
  HIPCHK(hipMemcpy(B.sponge_sigma, sigma.data(), n * sizeof(double), hipMemcpyHostToDevice));
  B.sponge_mean = ctx->alloc((size_t)(5 + ctx->ns) * n);
  B.dev.sponge_sigma = B.sponge_sigma;
  B.dev.sponge_mean = B.sponge_mean;
  if (!ctx->sponge_iter) ctx->sponge_iter = reinterpret_cast<unsigned long long *>(ctx->alloc(1));
}

void build(corefl_ctx *ctx, std::vector<corefl::HostBlock> blocks) {
  auto &P = ctx->param;
  const int species = P.get_int("species");
  const int reaction = species ? P.get_int("reaction") : 0;
  const int temporal = P.get_int("temporal_scheme");
  const int rec = P.get_int("reconstruction");
  const int inv = P.get_int("inviscid_scheme");
  const int vo = P.get_int("viscous_order");
  if (species) {
    // Species::Species / Reaction::Reaction read input/<mechanism_file> and input/<therm_file>
    // (ChemData.cpp:10-100, 868-920)
    auto path = [&](const std::string &f) { return ctx->input_dir.empty() ? f : ctx->input_dir + "/" + f; };
    const std::string mech_text = read_file(path(P.get_string("mechanism_file")));
    std::string therm_text;
    {
      std::ifstream t(path(P.get_string("therm_file")));
      if (t) {
        std::stringstream ss;
        ss << t.rdbuf();
        therm_text = ss.str();
      }
    }
    ctx->chem = new corefl_chem();
    ctx->chem->mech = corefl::parse_mechanism(mech_text, therm_text);
    ctx->ns = ctx->chem->mech.ns;
    if (ctx->ns != 9) throw std::runtime_error("the mixture kernels are instantiated for 9 species");
    const bool fs = inv == 51 || inv == 52 || inv == 71 || inv == 72;
    if (!fs && (rec < 1 || rec > 4))
      throw std::runtime_error("with species, reconstruction must be 1 (first order), 2 (MUSCL), 3 (NND2) or 4 "
                               "(component-wise WENO5)");
    if (!fs && (inv < 2 || inv > 4))
      throw std::runtime_error("with species, inviscid_scheme must be 2 (Roe), 3 (AUSM+), 4 (HLLC) or 51/52/71/72 "
                               "(WENO5/WENO7 flux splitting)");
    if (vo != 0 && vo != 2 && vo != 8)
      throw std::runtime_error("with species, viscous_order must be 0, 2 (face-centred) or 8 (collocated)");
    if (vo > 0 && P.get_bool("gradPInDiffusionFlux"))
      throw std::runtime_error("gradPInDiffusionFlux (pressure-driven species diffusion) is not on this path");
    if (vo > 0) {
      // Species::read_tran (ChemData.cpp:90-98, 740-831)
      corefl::parse_transport(read_file(path(P.get_string("transport_file"))), ctx->chem->mech);
    }
    if (reaction && temporal != 3 && temporal != 4 && temporal != 5)
      throw std::runtime_error("reacting flow: temporal_scheme 3 (explicit source), 4 (Wu's splitting) or 5 (Strang "
                               "splitting)");
    if (reaction && temporal == 4) {
      // wu_zeroDReaction (StrangSplitting.cu:43-225): 0 forward Euler, 1 EPI, 3 SSP-RK3 stages; 2 (HCST) has an
      // empty branch for generic mechanisms (StrangSplitting.cu:155-156), so it advances as 0.  The hard-coded
      // hydrogen mechanisms (hardCodedMech > 0, StrangSplitting.cu:226-371, 479-481) are not on this path.
      const int cam = P.get_int("chem_advance_method");
      if (cam < 0 || cam > 3)
        throw std::runtime_error("with temporal_scheme 4, chem_advance_method must be 0 (explicit), 1 (EPI), 2 (HCST) "
                                 "or 3 (SSP-RK3)");
      if (P.get_int("hardCodedMech") != 0)
        throw std::runtime_error("hardCodedMech > 0 (hard-coded hydrogen mechanisms) is not supported; use the "
                                 "mechanism file");
      ctx->wu_method = cam == 2 ? 0 : cam;
    }
    // chemSrcMethod 1 and 2 make finite_rate_chemistry also store the source Jacobian (FiniteRateChem.cu:1450-1462),
    // which only the implicit solvers read; SSP-RK3 adds the source explicitly whatever its value (RK.cuh:145-156)
    if (reaction && temporal == 3 && (P.get_int("chemSrcMethod") < 0 || P.get_int("chemSrcMethod") > 2))
      throw std::runtime_error("chemSrcMethod must be 0 (explicit), 1 (EPI) or 2 (diagonal approximation)");
    if (!reaction && temporal != 3) throw std::runtime_error("temporal_scheme must be 3 (SSP-RK3)");
    ctx->strang = reaction && temporal == 5;
    P.set_int("n_spec", ctx->ns);
    P.set_int("n_var", 5 + ctx->ns);
    ctx->chem_rtol = P.get_real("chemistry_rtol");
    ctx->chem_atol = P.get_real("chemistry_atol");
    ctx->chem_max_attempts = P.get_int("chemistry_max_attempts");
    if (ctx->chem_max_attempts < 1) throw std::runtime_error("chemistry_max_attempts must be at least 1");
    corefl::chem_ensure_device(ctx->chem);
  } else {
    const bool fs = inv == 51 || inv == 52 || inv == 71 || inv == 72;
    if (!fs && (rec < 1 || rec > 5))
      throw std::runtime_error("reconstruction must be 1 (first order), 2 (MUSCL), 3 (NND2), 4 (WENO5) or 5 (WENO7)");
    if (!fs && rec <= 3 && P.get_int("limiter") != 0 && P.get_int("limiter") != 1)
      throw std::runtime_error("limiter must be 0 or 1 (minmod)");
    if (!fs && (inv < 1 || inv > 4))
      throw std::runtime_error("inviscid_scheme must be 1 (LF), 2 (Roe), 3 (AUSM+), 4 (HLLC), or 51/52/71/72 (WENO5/WENO7 "
                               "flux splitting, component-wise/characteristic)");
    if (vo != 0 && vo != 2 && vo != 4 && vo != 8)
      throw std::runtime_error("viscous_order must be 0, 2, 4 or 8 on this path");
    if (temporal != 3) throw std::runtime_error("temporal_scheme must be 3 (SSP-RK3)");
  }
  const corefl::ChemMech *mech = ctx->chem ? &ctx->chem->mech : nullptr;
  const int ngg = P.get_int("ngg");
  P.set_int("dimension", blocks.at(0).dim);
  // reference state and the WENO epsilon scale (BoundCond.cpp:276-295, DParameter.cu:25-28, WENO.cu:8)
  double rho_inf, v_inf, p_inf, c_inf;
  if (P.get_string("canonical_problem") == "taylorGreen") {
    const auto &s = P.get_struct("taylorGreen");
    const double T0 = corefl::struct_real(s, "temperature", 300.0), p0 = corefl::struct_real(s, "pressure", 101325.0);
    const double R = kRu / kMwAir;
    rho_inf = p0 / (R * T0);
    c_inf = std::sqrt(kGamma * R * T0);
    v_inf = corefl::struct_real(s, "mach", 0.1) * c_inf;
    p_inf = p0;
  } else if (is_mixing_layer(P)) {
    // Inflow::Inflow for a mixing layer (BoundCond.cpp:30-60, 265-275): the upper stream, and v_inf =
    // |velocity - v_lower| with v_lower the lower stream's (zero) v component, as written
    const MixingLayer ml = mixing_layer(P, mech);
    rho_inf = ml.upper.rho;
    p_inf = ml.upper.p;
    c_inf = ml.upper.c;
    v_inf = std::abs(std::abs(ml.upper.u) - 0.0);
  } else {
    const FlowState r = inflow_state(P.get_struct(P.get_string("reference_state")), mech);
    rho_inf = r.rho;
    p_inf = r.p;
    c_inf = r.c;
    v_inf = std::abs(std::sqrt(r.u * r.u + r.v * r.v + r.w * r.w)) < 1 ? r.c : std::sqrt(r.u * r.u + r.v * r.v + r.w * r.w);
  }
  P.set_real("rho_inf", rho_inf);
  P.set_real("v_inf", v_inf);
  P.set_real("p_inf", p_inf);
  P.set_real("speed_of_sound", c_inf);
  P.set_real("weno_eps", 1e-40 * rho_inf * v_inf * rho_inf * v_inf);
  // hybrid flux-splitting scheme (Parameter.cpp:236-255; DParameter.cu:48-50): schemes 51/52/71/72 with a
  // positive shockSensor_threshold
  const bool fsw = inv == 51 || inv == 52 || inv == 71 || inv == 72;
  const double sensor_thr = fsw && P.get_real("shockSensor_threshold") > 1e-10 ? P.get_real("shockSensor_threshold") : 0;
  if (sensor_thr > 0 && (P.get_int("shock_sensor") < 0 || P.get_int("shock_sensor") > 2))
    throw std::runtime_error("shock_sensor must be 0 (Ducros), 1 (Jameson) or 2 (density/pressure jump)");
  ctx->scheme = {inv, P.get_int("reconstruction"), vo, P.get_real("weno_eps"), P.get_real("entropy_fix_factor"),
                 P.get_real("prandtl_number"), P.get_real("cfl"), v_inf, P.get_bool("positive_preserving") ? 1 : 0,
                 rho_inf, sensor_thr, P.get_int("shock_sensor"), P.get_real("shockSensor_epsilon")};
  ctx->fixed_dt = P.get_bool("fixed_time_step");
  ctx->t_end = P.get_real("total_simulation_time");

  ctx->output_screen = P.get_int("output_screen");
  ctx->limit_flow = P.get_bool("limit_flow");
  {
    // limit_flow bounds (DParameter.cu:332-349)
    auto &L = ctx->limits;
    L.ll[0] = 1e-6 * rho_inf;
    L.ul[0] = 1e+3 * rho_inf;
    const double vref = std::abs(v_inf) < 1 ? c_inf : v_inf;
    for (int l = 1; l < 4; ++l) {
      L.ll[l] = -1000 * vref;
      L.ul[l] = 1000 * vref;
    }
    L.ll[4] = 1e-6 * p_inf;
    L.ul[4] = 1e+3 * p_inf;
  }

  HIPCHK(hipStreamCreate(&ctx->stream));
  HIPCHK(hipStreamCreate(&ctx->side));
  ctx->slab_setting = P.get_int("stage_slab_planes");
  // a slab must be thicker than the 3 planes the WENO7 z-stencil reaches below an interface, and start on a
  // march boundary of the update kernel
  if (ctx->slab_setting > 0 && (ctx->slab_setting < 8 || ctx->slab_setting % corefl::march_planes() != 0))
    throw std::runtime_error("stage_slab_planes must be -1, 0, or a multiple of " +
                             std::to_string(corefl::march_planes()) + " of at least 8");
  HIPCHK(hipEventCreateWithFlags(&ctx->ev_fork, hipEventDisableTiming));
  HIPCHK(hipEventCreateWithFlags(&ctx->ev_join, hipEventDisableTiming));
  for (hipEvent_t &e : ctx->ev_slab) HIPCHK(hipEventCreateWithFlags(&e, hipEventDisableTiming));
  ctx->dt_dev = ctx->alloc(1);
  ctx->time_dev = ctx->alloc(1);
  ctx->sub_dt_dev = ctx->alloc(1);
  // Block-to-block faces: the coincident layer on a side of the block, the in-face ranges inside it, and every
  // partner node inside the target block (the exchange kernels index both blocks without further checks).
  for (size_t bi = 0; bi < blocks.size(); ++bi) {
    const auto &hb = blocks[bi];
    const int ext[3] = {hb.mx, hb.my, hb.mz};
    for (const auto &f : hb.inner) {
      if (f.target_block < 0 || f.target_block >= (int)blocks.size() || f.target_block == (int)bi)
        throw std::runtime_error("a block-to-block face must join two different blocks");
      const auto &T = blocks[f.target_block];
      const int text[3] = {T.mx, T.my, T.mz};
      if (f.face_layer != (f.direction < 0 ? 0 : ext[f.face] - 1))
        throw std::runtime_error("block-to-block face is not on a block side");
      for (int d = 0; d < 3; ++d) {
        if (d == f.face) continue;
        const int lo = std::min(f.start[d], f.end[d]), hi = std::max(f.start[d], f.end[d]);
        if (lo < 0 || hi >= ext[d]) throw std::runtime_error("block-to-block face range outside the block");
      }
      for (int c = 0; c < 8; ++c) {
        int dd[3];
        for (int q = 0; q < 3; ++q) dd[q] = f.loop_dir[q] * ((c >> q & 1) ? f.n_point[q] - 1 : 0);
        for (int q = 0; q < 3; ++q) {
          const int t = f.target_start[q] + f.target_loop_dir[q] * dd[f.src_tar[q]];
          if (f.src_tar[q] < 0 || f.src_tar[q] > 2 || t < 0 || t >= text[q])
            throw std::runtime_error("block-to-block face maps outside its target block");
        }
      }
    }
  }
  for (auto &hb : blocks) {
    // Every kernel indexes at most ngg nodes beyond the interior; reject blocks and faces that would
    // make any of those accesses leave the allocation.
    const int ext[3] = {hb.mx, hb.my, hb.mz};
    if (hb.mx < 3 || hb.my < 3 || (hb.mz != 1 && hb.mz < 3))
      throw std::runtime_error("block extents must be >= 3 (mz = 1 for 2D blocks)");
    for (int d = 0; d < hb.dim; ++d)
      if (ext[d] < ngg + 2) throw std::runtime_error("block extent smaller than ngg + 2");
    if (vo > 2)
      for (int d = 0; d < hb.dim; ++d)
        if (ext[d] < 5) throw std::runtime_error("viscous closures need at least 5 nodes per direction");
    for (auto &f : hb.faces) {
      if (f.type != 3 && f.type != 5 && f.type != 6 && f.type != 8)
        throw std::runtime_error("boundary type " + std::to_string(f.type) + " is not supported on this path "
                                 "(3 symmetry, 5 inflow, 6 outflow, 8 periodic)");
      for (int d = 0; d < 3; ++d)
        if (f.start[d] < 0 || f.end[d] < f.start[d] || f.end[d] >= ext[d])
          throw std::runtime_error("boundary face range outside the block");
      const int nf = f.start[f.face];
      if (f.end[f.face] != nf || (nf != 0 && nf != ext[f.face] - 1))
        throw std::runtime_error("boundary face is not on a block side");
    }
    Block B;
    B.host = std::move(hb);
    auto &h = B.host;
    auto &d = B.dev;
    d.mx = h.mx; d.my = h.my; d.mz = h.mz; d.g = ngg; d.dim = h.dim;
    d.d2 = h.pad2(); d.d1 = h.pad1(); d.sz = h.size(); d.n = (long)h.mx * h.my * h.mz;
    for (int q = 0; q < 6; ++q) d.bt[q] = h.btype[q];
    const size_t sz = d.sz;
    const size_t nv = 5 + ctx->ns;
    d.cv = ctx->alloc(nv * sz);
    d.qn = ctx->alloc(nv * sz);
    d.bv = ctx->alloc(6 * sz);
    d.jac = ctx->alloc(sz);
    d.met = ctx->alloc(9 * sz);
    d.fx = ctx->alloc(nv * sz);
    d.fy = ctx->alloc(nv * sz);
    d.fz = d.dim == 3 ? ctx->alloc(nv * sz) : d.fy;
    if (ctx->ns) {
      B.mix.ns = ctx->ns;
      B.mix.sv = ctx->alloc(ctx->ns * sz);
      B.mix.snd = ctx->alloc(sz);
      B.mix.mech = ctx->chem->mech_dev;
      B.mix.dt = ctx->dt_dev;
      B.mix.explicit_source = reaction && temporal == 3;
      if (ctx->wu_method >= 0) B.mix.tscale = ctx->alloc(d.n); // reaction time scale, zero at the start
      if (B.mix.explicit_source) {
        B.mix.tscale = ctx->alloc(d.n);
        const std::vector<double> big(d.n, 1e+6);
        HIPCHK(hipMemcpy(B.mix.tscale, big.data(), d.n * sizeof(double), hipMemcpyHostToDevice));
      }
      if (ctx->strang) B.hstep = ctx->alloc(d.n);
    }
    if (vo > 0) {
      d.fv = ctx->alloc(nv * sz);
      d.gv = ctx->alloc(nv * sz);
      d.hv = ctx->alloc(nv * sz);
      if (ctx->ns) {
        B.mix.visc_order = vo;
        B.mix.mul = ctx->alloc(sz);
        B.mix.cond = ctx->alloc(sz);
        B.mix.rhod = ctx->alloc(ctx->ns * sz);
      }
    }
    if (inv == 2) d.efd = ctx->alloc(sz);
    if (ctx->scheme.sensor_threshold > 0) d.sensor = ctx->alloc(sz); // zeroed; the ghost entries stay 0
    d.dq = ctx->alloc(nv * d.n);
    HIPCHK(hipMemcpy(d.jac, h.jac.data(), sz * sizeof(double), hipMemcpyHostToDevice));
    HIPCHK(hipMemcpy(d.met, h.metric.data(), 9 * sz * sizeof(double), hipMemcpyHostToDevice));
    for (const auto &f : h.faces) {
      corefl::DevFace df{};
      for (int q = 0; q < 3; ++q) { df.start[q] = f.start[q]; df.end[q] = f.end[q]; }
      df.type = f.type; df.face = f.face; df.direction = f.direction;
      B.faces.push_back(df);
      B.inflow.push_back(f.type == 5 ? make_inflow(ctx, P, h, f, mech) : corefl::DevInflow{});
    }
    if (P.get_bool("sponge_layer")) setup_sponge(ctx, P, B, mech);
    for (const auto &f : h.inner) {
      corefl::DevInner di{};
      for (int q = 0; q < 3; ++q) {
        di.start[q] = f.start[q];
        di.loop_dir[q] = f.loop_dir[q];
        di.n_point[q] = f.n_point[q];
        di.target_start[q] = f.target_start[q];
        di.target_loop_dir[q] = f.target_loop_dir[q];
        di.src_tar[q] = f.src_tar[q];
      }
      di.face = f.face;
      di.face_layer = f.face_layer;
      for (int q = 0; q < 3; ++q) {
        di.vmap[f.src_tar[q]] = q;
        di.vsign[f.src_tar[q]] = f.target_loop_dir[q];
      }
      B.inner.push_back(di);
      B.inner_target.push_back(f.target_block);
    }
    B.part_offset = ctx->n_parts;
    B.part_count = corefl::dt_partial_count(d);
    ctx->n_parts += B.part_count;
    if (ctx->output_screen > 0) {
      B.last = ctx->alloc(4 * d.n); // zero: the reference's first residual (step 1) measures against zero
      B.res_offset = ctx->n_res_parts;
      ctx->n_res_parts += corefl::residual_partial_count(d);
      ctx->n_nodes_total += (double)d.n;
    }
    if (ctx->limit_flow) B.lflag = reinterpret_cast<unsigned char *>(ctx->alloc((d.n + 7) / 8));
    ctx->blocks.push_back(std::move(B));
  }
  ctx->parts = ctx->alloc(ctx->n_parts);
  if (ctx->output_screen > 0) {
    ctx->res_parts = ctx->alloc(4 * ctx->n_res_parts);
    ctx->mon = reinterpret_cast<corefl::MonitorState *>(ctx->alloc((sizeof(corefl::MonitorState) + 7) / 8));
    HIPCHK(hipHostMalloc(reinterpret_cast<void **>(&ctx->mon_host), sizeof(corefl::MonitorState), hipHostMallocDefault));
    std::memset(ctx->mon_host, 0, sizeof(corefl::MonitorState));
  }
  if (ctx->strang) { // the chemistry batch is allocated here, never inside a captured step
    long n_max = 0;
    for (auto &B : ctx->blocks) n_max = std::max(n_max, B.dev.n);
    corefl::chem_ensure_batch(ctx->chem, n_max);
  }
  // initial condition, then initialize_computation (Driver.cu:126-211)
  for (auto &B : ctx->blocks) {
    std::vector<double> sv;
    const auto bv = initial_bv(P, B.host, mech, &sv);
    HIPCHK(hipMemcpy(B.dev.bv, bv.data(), bv.size() * sizeof(double), hipMemcpyHostToDevice));
    if (ctx->ns) {
      HIPCHK(hipMemcpy(B.mix.sv, sv.data(), sv.size() * sizeof(double), hipMemcpyHostToDevice));
      HIPCHK(corefl::launch_mix_cv_from_bv_all(B.dev, B.mix, ctx->stream));
    } else {
      HIPCHK(corefl::launch_cv_from_bv_all(B.dev, ctx->stream));
    }
    ctx->apply_bcs(B, ctx->stream);
  }
  ctx->communicate(ctx->stream);
  ctx->host_dt = P.get_real("dt");
  HIPCHK(hipMemcpyAsync(ctx->dt_dev, &ctx->host_dt, sizeof(double), hipMemcpyHostToDevice, ctx->stream));
  HIPCHK(hipStreamSynchronize(ctx->stream));
  // keep the block nodes' coordinates for flowfield output; free the padded host copies
  for (auto &B : ctx->blocks) {
    const auto &H = B.host;
    const long G = H.ngg + 1, p2 = H.mx + 2 * G, p1 = (H.my + 2 * G) * p2;
    for (int c = 0; c < 3; ++c) {
      const auto &src = c == 0 ? H.x : (c == 1 ? H.y : H.z);
      auto &dst = B.coords[c];
      dst.resize((size_t)H.mx * H.my * H.mz);
      size_t q = 0;
      for (int k = 0; k < H.mz; ++k)
        for (int j = 0; j < H.my; ++j)
          for (int i = 0; i < H.mx; ++i) dst[q++] = src[(k + G) * p1 + (j + G) * p2 + (i + G)];
    }
    std::vector<double>().swap(B.host.x);
    std::vector<double>().swap(B.host.y);
    std::vector<double>().swap(B.host.z);
  }
}

// ---------------------------------------------------------------- flowfield files
// The reference's restart file, output/flowfield.plt: a Tecplot binary (#!TDV112) file of one ordered zone per
// block with every variable in double precision at the nodes (FieldIO.h:89-430, 709-825).  Variables: x, y, z,
// density, u, v, w, pressure, temperature, the species mass fractions under their mechanism names, mach
// (Parameter.cpp:750-858 for a laminar run); no ghost layers (the reference writes with ngg_output = 0).
struct PltWriter {
  std::ofstream f;
  void i32(int32_t v) { f.write(reinterpret_cast<const char *>(&v), 4); }
  void f32(float v) { f.write(reinterpret_cast<const char *>(&v), 4); }
  void f64(double v) { f.write(reinterpret_cast<const char *>(&v), 8); }
  // a string as one int32 per character and a terminating 0 (gxl::write_str, gxl_lib/MyString.cpp:151-163)
  void str(const std::string &t) {
    for (char c : t) i32(static_cast<int32_t>(c));
    i32(0);
  }
};

struct PltReader {
  std::ifstream f;
  std::string path;
  template<typename T> T get() {
    T v{};
    f.read(reinterpret_cast<char *>(&v), sizeof(T));
    if (!f) throw std::runtime_error("flowfield file " + path + " ends early");
    return v;
  }
  std::string str() {
    std::string t;
    for (int32_t c = get<int32_t>(); c != 0; c = get<int32_t>()) t.push_back(static_cast<char>(c));
    return t;
  }
  void skip(long long n) { f.seekg(n, std::ios::cur); }
};

std::vector<std::string> flowfield_names(const corefl_ctx *ctx) {
  std::vector<std::string> v{"x", "y", "z", "density", "u", "v", "w", "pressure", "temperature"};
  for (int l = 0; l < ctx->ns; ++l) v.emplace_back(ctx->chem->mech.species[l]);
  v.emplace_back("mach");
  return v;
}

std::string upper(std::string t) {
  for (auto &c : t) c = static_cast<char>(std::toupper(static_cast<unsigned char>(c)));
  return t;
}

void write_flowfield(corefl_ctx *ctx, const std::string &path) {
  ctx->flush_exchange(ctx->stream);
  HIPCHK(hipStreamSynchronize(ctx->stream));
  double time = 0;
  HIPCHK(hipMemcpy(&time, ctx->time_dev, sizeof(double), hipMemcpyDeviceToHost));
  const auto names = flowfield_names(ctx);
  const int nv = static_cast<int>(names.size()), ns = ctx->ns;
  PltWriter w;
  w.f.open(path, std::ios::binary | std::ios::trunc);
  if (!w.f) throw std::runtime_error("cannot write " + path);
  // header (FieldIO.h:108-218)
  w.f.write("#!TDV112", 8);
  w.i32(1);
  w.i32(0);
  w.str("Solution file");
  w.i32(nv);
  for (const auto &n : names) w.str(n);
  for (size_t b = 0; b < ctx->blocks.size(); ++b) {
    const auto &H = ctx->blocks[b].host;
    w.f32(299.0f);
    w.str("zone " + std::to_string(b));
    w.i32(-1);
    w.i32(-2);
    w.f64(time);
    w.i32(-1);
    w.i32(0);
    w.i32(0);
    w.i32(0);
    w.i32(0);
    w.i32(H.mx);
    w.i32(H.my);
    w.i32(H.mz);
    w.i32(0);
  }
  w.f32(357.0f);
  // data section (FieldIO.h:243-430)
  for (auto &B : ctx->blocks) {
    const auto &H = B.host;
    const long sz = B.dev.sz, n = (long)H.mx * H.my * H.mz;
    std::vector<double> bv(6 * sz), sv(ns * sz), snd(ns ? sz : 0);
    HIPCHK(hipMemcpy(bv.data(), B.dev.bv, 6 * sz * sizeof(double), hipMemcpyDeviceToHost));
    if (ns) {
      HIPCHK(hipMemcpy(sv.data(), B.mix.sv, ns * sz * sizeof(double), hipMemcpyDeviceToHost));
      HIPCHK(hipMemcpy(snd.data(), B.mix.snd, sz * sizeof(double), hipMemcpyDeviceToHost));
    }
    std::vector<std::vector<double>> var(nv, std::vector<double>(n));
    long q = 0;
    for (int k = 0; k < H.mz; ++k)
      for (int j = 0; j < H.my; ++j)
        for (int i = 0; i < H.mx; ++i, ++q) {
          const long id = H.idx(i, j, k);
          for (int c = 0; c < 3; ++c) var[c][q] = B.coords[c][q];
          for (int l = 0; l < 6; ++l) var[3 + l][q] = bv[l * sz + id];
          for (int l = 0; l < ns; ++l) var[9 + l][q] = sv[l * sz + id];
          const double u = bv[sz + id], v = bv[2 * sz + id], ww = bv[3 * sz + id];
          // speed of sound: the mixture's, or the perfect gas's (update_physical_properties, FieldOperation.cuh)
          const double c = ns ? snd[id] : std::sqrt(1.4 * 8314.462618 / 28.9647 * bv[5 * sz + id]);
          var[nv - 1][q] = std::sqrt(u * u + v * v + ww * ww) / c;
        }
    w.f32(299.0f);
    for (int l = 0; l < nv; ++l) w.i32(2);
    w.i32(0);
    w.i32(0);
    w.i32(-1);
    for (int l = 0; l < nv; ++l) {
      const auto mm = std::minmax_element(var[l].begin(), var[l].end());
      w.f64(*mm.first);
      w.f64(*mm.second);
    }
    for (int l = 0; l < nv; ++l) w.f.write(reinterpret_cast<const char *>(var[l].data()), n * sizeof(double));
  }
  if (!w.f) throw std::runtime_error("writing " + path + " failed");
}

// read_flowfield (Initialize.cu:97-345) for a file whose zones are the current blocks, in order or permuted
// (read_flowfield_with_same_block): density, u, v, w, pressure, temperature and the species found by name
// (identify_variable_labels, Initialize.cu:346-405) replace the current state, with the file's ghost layers where
// it has them; variables not found keep their current values; then cv, boundary conditions and block-to-block
// exchanges follow as after any state upload, and the time is the file's solution time.
void read_flowfield(corefl_ctx *ctx, const std::string &path) {
  PltReader r;
  r.path = path;
  r.f.open(path, std::ios::binary);
  if (!r.f) throw std::runtime_error("cannot open flowfield file " + path);
  char magic[8];
  r.f.read(magic, 8);
  if (!r.f || std::string(magic, 8) != "#!TDV112") throw std::runtime_error(path + " is not a #!TDV112 file");
  r.skip(8);
  r.str();
  const int nv = r.get<int32_t>();
  if (nv < 3 || nv > 10000) throw std::runtime_error(path + ": bad variable count");
  std::vector<std::string> names(nv);
  for (auto &n : names) n = r.str();
  // labels: 0..5 basic variables, 6 + l species l, -1 ignored
  std::vector<int> label(nv, -1);
  for (int v = 3; v < nv; ++v) {
    const std::string n = upper(names[v]);
    if (n == "DENSITY" || n == "ROE" || n == "RHO") label[v] = 0;
    else if (n == "U") label[v] = 1;
    else if (n == "V") label[v] = 2;
    else if (n == "W") label[v] = 3;
    else if (n == "P" || n == "PRESSURE") label[v] = 4;
    else if (n == "T" || n == "TEMPERATURE") label[v] = 5;
    else
      for (int l = 0; l < ctx->ns; ++l)
        if (n == upper(ctx->chem->mech.species[l])) label[v] = 6 + l;
  }
  std::vector<int> zx, zy, zz;
  double time = 0;
  float marker = r.get<float>();
  while (std::fabs(marker - 299.0f) < 1e-3f) {
    r.str();
    r.skip(8);
    time = r.get<double>();
    r.skip(20);
    zx.push_back(r.get<int32_t>());
    zy.push_back(r.get<int32_t>());
    zz.push_back(r.get<int32_t>());
    r.skip(4);
    marker = r.get<float>();
  }
  if (std::fabs(marker - 357.0f) > 1e-3f) throw std::runtime_error(path + ": no end-of-header marker");
  const int nz = static_cast<int>(zx.size()), nb = static_cast<int>(ctx->blocks.size());
  if (nz != nb) throw std::runtime_error(path + ": " + std::to_string(nz) + " zones for " + std::to_string(nb) +
                                         " blocks (reading onto a different mesh is not supported)");
  // zone z of the file feeds block order[z]: the same block where the sizes match in order, else the unique block
  // with the zone's node counts (the reference's same-block search)
  auto fits = [&](int z, int b) {
    const auto &H = ctx->blocks[b].host;
    const int g = (zx[z] - H.mx) / 2;
    return g >= 0 && g <= H.ngg && zx[z] == H.mx + 2 * g && zy[z] == H.my + 2 * g && zz[z] == H.mz + 2 * g;
  };
  std::vector<int> order(nz, -1);
  bool same = true;
  for (int z = 0; z < nz && same; ++z) same = fits(z, z);
  if (same) {
    for (int z = 0; z < nz; ++z) order[z] = z;
  } else {
    for (int b = 0; b < nb; ++b) {
      int hit = 0;
      for (int z = 0; z < nz; ++z)
        if (fits(z, b)) {
          ++hit;
          order[z] = b;
        }
      if (hit != 1) throw std::runtime_error(path + ": the zones do not match the blocks one to one");
    }
  }
  HIPCHK(hipStreamSynchronize(ctx->stream));
  const int ns = ctx->ns;
  for (int z = 0; z < nz; ++z) {
    auto &B = ctx->blocks[order[z]];
    const auto &H = B.host;
    const long sz = B.dev.sz, N = (long)zx[z] * zy[z] * zz[z];
    const int g = (zx[z] - H.mx) / 2;
    std::vector<double> bv(6 * sz), sv(ns * sz), buf(N);
    HIPCHK(hipMemcpy(bv.data(), B.dev.bv, 6 * sz * sizeof(double), hipMemcpyDeviceToHost));
    if (ns) HIPCHK(hipMemcpy(sv.data(), B.mix.sv, ns * sz * sizeof(double), hipMemcpyDeviceToHost));
    r.skip(4 + 4LL * nv + 12 + 16LL * nv + 3 * N * 8);
    for (int v = 3; v < nv; ++v) {
      if (label[v] < 0) {
        r.skip(N * 8);
        continue;
      }
      r.f.read(reinterpret_cast<char *>(buf.data()), N * sizeof(double));
      if (!r.f) throw std::runtime_error("flowfield file " + path + " ends early");
      double *dst = label[v] < 6 ? bv.data() + label[v] * sz : sv.data() + (label[v] - 6) * sz;
      long q = 0;
      for (int k = -g; k < H.mz + g; ++k)
        for (int j = -g; j < H.my + g; ++j)
          for (int i = -g; i < H.mx + g; ++i) dst[H.idx(i, j, k)] = buf[q++];
    }
    HIPCHK(hipMemcpy(B.dev.bv, bv.data(), 6 * sz * sizeof(double), hipMemcpyHostToDevice));
    if (ns) {
      HIPCHK(hipMemcpy(B.mix.sv, sv.data(), ns * sz * sizeof(double), hipMemcpyHostToDevice));
      HIPCHK(corefl::launch_mix_cv_from_bv_all(B.dev, B.mix, ctx->stream));
    } else {
      HIPCHK(corefl::launch_cv_from_bv_all(B.dev, ctx->stream));
    }
    ctx->apply_bcs(B, ctx->stream);
  }
  ctx->exchange_pending = true;
  ctx->flush_exchange(ctx->stream);
  HIPCHK(hipMemcpyAsync(ctx->time_dev, &time, sizeof(double), hipMemcpyHostToDevice, ctx->stream));
  HIPCHK(hipStreamSynchronize(ctx->stream));
  ctx->drop_graph();
}

int fail(const std::exception &e) {
  g_last_error = e.what();
  return -1;
}

// Boundary and block-to-block faces, ghost grid (mirrored, then the neighbours' nodes across block-to-block
// faces: Mesh.cpp:1013-1083), metrics (second-order, or symmetric-conservative with symmetric_conservative_metric).
std::vector<corefl::HostBlock> prepare_blocks(corefl::Parameter &P, std::vector<corefl::HostBlock> blocks,
                                              const std::string &boundary_text, const std::string &inner_text) {
  corefl::parse_boundaries(boundary_text, blocks, P.get_int("periodic_label"));
  if (!inner_text.empty()) corefl::parse_inner_faces(inner_text, blocks);
  for (auto &b : blocks) corefl::mirror_ghost_grid(b);
  corefl::fill_inner_ghost_grid(blocks);
  const int scm = P.get_int("symmetric_conservative_metric");
  for (auto &b : blocks) {
    std::string err;
    // the standard metrics check the interior Jacobian; the symmetric-conservative ones then replace them
    // (compute_jac_metric, Mesh.cu:136-141)
    if (!corefl::compute_metrics(b, err, false)) throw std::runtime_error(err);
    if (scm != 0 && !corefl::sc_metrics_device(b, scm, err)) throw std::runtime_error(err);
  }
  return blocks;
}

} // namespace

namespace corefl {
void set_last_error(const std::string &msg) { g_last_error = msg; }
} // namespace corefl

extern "C" {

const char *corefl_last_error(void) { return g_last_error.c_str(); }

int corefl_create(const char *input_dir, corefl_ctx **out) {
  try {
    auto ctx = std::make_unique<corefl_ctx>();
    const std::string dir(input_dir);
    ctx->input_dir = dir;
    ctx->param.parse(read_file(dir + "/setup.txt"));
    ctx->param.deduce();
    auto &P = ctx->param;
    // Mesh::read_grid (Mesh.cpp:692-788): grid%4d.dat (binary) or grid%4d.grd (ASCII)
    auto blocks = P.get_int("gridIsBinary")
                      ? corefl::read_grid_binary(dir + "/grid/grid   0.dat", P.get_int("ngg"), P.get_real("gridScale"))
                      : corefl::read_grid_ascii(dir + "/grid/grid   0.grd", P.get_int("ngg"), P.get_real("gridScale"));
    // inner%4d.txt lists the block-to-block faces; a single block needs none
    std::string inner_text;
    if (std::ifstream(dir + "/boundary_condition/inner   0.txt")) inner_text = read_file(dir + "/boundary_condition/inner   0.txt");
    blocks = prepare_blocks(P, std::move(blocks), read_file(dir + "/boundary_condition/boundary   0.txt"), inner_text);
    build(ctx.get(), std::move(blocks));
    // initial = 1: continue from the run directory's output/flowfield.plt (Initialize.cu:36-41, 97-345); the
    // input directory sits in the run directory
    if (P.get_int("initial") == 1) read_flowfield(ctx.get(), dir + "/../output/flowfield.plt");
    *out = ctx.release();
    return 0;
  } catch (const std::exception &e) {
    return fail(e);
  }
}

int corefl_create_mem(const char *setup_text, int n_block, const int *dims, const double *const *xyz,
                      const char *boundary_text, const char *inner_text, corefl_ctx **out) {
  try {
    auto ctx = std::make_unique<corefl_ctx>();
    ctx->param.parse(setup_text);
    ctx->param.deduce();
    auto &P = ctx->param;
    std::vector<corefl::HostBlock> blocks;
    for (int b = 0; b < n_block; ++b)
      blocks.push_back(corefl::make_block(dims[3 * b], dims[3 * b + 1], dims[3 * b + 2], P.get_int("ngg"),
                                          xyz[3 * b], xyz[3 * b + 1], xyz[3 * b + 2], P.get_real("gridScale")));
    blocks = prepare_blocks(P, std::move(blocks), boundary_text, inner_text ? inner_text : "");
    build(ctx.get(), std::move(blocks));
    *out = ctx.release();
    return 0;
  } catch (const std::exception &e) {
    return fail(e);
  }
}

void corefl_destroy(corefl_ctx *ctx) { delete ctx; }

int corefl_n_blocks(const corefl_ctx *ctx) { return ctx ? static_cast<int>(ctx->blocks.size()) : -1; }

int corefl_block_view_get(corefl_ctx *ctx, int block, corefl_block_view *v) {
  if (!ctx || block < 0 || block >= (int)ctx->blocks.size() || !v) {
    g_last_error = "corefl_block_view_get: bad arguments";
    return -1;
  }
  const auto &d = ctx->blocks[block].dev;
  v->mx = d.mx; v->my = d.my; v->mz = d.mz; v->ngg = d.g; v->n_var = 5 + ctx->ns; v->n_spec = ctx->ns;
  v->sz = d.sz;
  v->cv = d.cv; v->bv = d.bv; v->qn = d.qn; v->metric = d.met; v->jac = d.jac; v->dq = d.dq;
  return 0;
}

int corefl_get_array(corefl_ctx *ctx, int block, int which, double *host, long long n) {
  try {
    if (!ctx || block < 0 || block >= (int)ctx->blocks.size()) throw std::runtime_error("bad block index");
    const auto &d = ctx->blocks[block].dev;
    const auto &mx = ctx->blocks[block].mix;
    const double *src = nullptr;
    long long cnt = 0;
    const long nv = 5 + ctx->ns;
    switch (which) {
      case COREFL_CV: src = d.cv; cnt = nv * d.sz; break;
      case COREFL_BV: src = d.bv; cnt = 6 * d.sz; break;
      case COREFL_SV: src = mx.sv; cnt = ctx->ns * d.sz; break;
      case COREFL_DQ: src = d.dq; cnt = nv * d.n; break;
      case COREFL_JAC: src = d.jac; cnt = d.sz; break;
      case COREFL_METRIC: src = d.met; cnt = 9 * d.sz; break;
      case COREFL_FV: src = d.fv; cnt = 5 * d.sz; break;
      case COREFL_GV: src = d.gv; cnt = 5 * d.sz; break;
      case COREFL_HV: src = d.hv; cnt = 5 * d.sz; break;
      case COREFL_SENSOR: src = d.sensor; cnt = d.sz; break;
      default: throw std::runtime_error("unknown array selector");
    }
    if (!src) throw std::runtime_error("array not allocated for this scheme");
    if (n != cnt) throw std::runtime_error("host buffer has " + std::to_string(n) + " elements, need " + std::to_string(cnt));
    HIPCHK(hipStreamSynchronize(ctx->stream));
    HIPCHK(hipMemcpy(host, src, cnt * sizeof(double), hipMemcpyDeviceToHost));
    return 0;
  } catch (const std::exception &e) {
    return fail(e);
  }
}

int corefl_set_state(corefl_ctx *ctx, int block, const double *bv_host, long long n) {
  try {
    if (!ctx || block < 0 || block >= (int)ctx->blocks.size()) throw std::runtime_error("bad block index");
    auto &B = ctx->blocks[block];
    if (n != (6 + ctx->ns) * B.dev.sz) throw std::runtime_error("state must have (6 + n_spec)*sz elements");
    HIPCHK(hipStreamSynchronize(ctx->stream));
    HIPCHK(hipMemcpy(B.dev.bv, bv_host, 6 * B.dev.sz * sizeof(double), hipMemcpyHostToDevice));
    if (ctx->ns) {
      HIPCHK(hipMemcpy(B.mix.sv, bv_host + 6 * B.dev.sz, ctx->ns * B.dev.sz * sizeof(double), hipMemcpyHostToDevice));
      HIPCHK(corefl::launch_mix_cv_from_bv_all(B.dev, B.mix, ctx->stream));
    } else {
      HIPCHK(corefl::launch_cv_from_bv_all(B.dev, ctx->stream));
    }
    ctx->apply_bcs(B, ctx->stream);
    ctx->exchange_pending = true;
    HIPCHK(hipStreamSynchronize(ctx->stream));
    return 0;
  } catch (const std::exception &e) {
    return fail(e);
  }
}

int corefl_rhs(corefl_ctx *ctx, int stage, void *stream) {
  try {
    if (!ctx) throw std::runtime_error("null context");
    if (stage > 2) throw std::runtime_error("stage must be -1 (residual only) or an SSP-RK3 stage 0, 1, 2");
    if (stage >= 0 && ctx->wu_method >= 0)
      throw std::runtime_error("Wu's splitting advances whole steps only (corefl_rk3_step / corefl_advance)");
    hipStream_t s = ctx->pick(stream);
    ctx->flush_exchange(s);
    ctx->stage(stage < 0 ? -1 : stage, s);
    return 0;
  } catch (const std::exception &e) {
    return fail(e);
  }
}

int corefl_set_dt(corefl_ctx *ctx, double dt) {
  try {
    if (!ctx || !(dt > 0)) throw std::runtime_error("corefl_set_dt: need a context and dt > 0");
    ctx->host_dt = dt;
    HIPCHK(hipMemcpyAsync(ctx->dt_dev, &ctx->host_dt, sizeof(double), hipMemcpyHostToDevice, ctx->stream));
    HIPCHK(hipStreamSynchronize(ctx->stream));
    return 0;
  } catch (const std::exception &e) {
    return fail(e);
  }
}

int corefl_chem_substep(corefl_ctx *ctx, double dt, int mode, void *stream) {
  try {
    if (!ctx) throw std::runtime_error("null context");
    if (!ctx->chem || !ctx->strang)
      throw std::runtime_error("chemistry sub-steps need a reacting mixture with temporal_scheme 5");
    if (!(dt > 0)) throw std::runtime_error("dt must be positive");
    if (mode != 0 && mode != 1) throw std::runtime_error("mode must be 0 (exact) or 1 (surrogate with gate and fallback)");
    if (mode == 1 && !ctx->surrogate) throw std::runtime_error("mode 1 needs corefl_set_chemistry_surrogate first");
    hipStream_t s = ctx->pick(stream);
    ctx->flush_exchange(s);
    ctx->sub_dt = dt;
    HIPCHK(hipMemcpyAsync(ctx->sub_dt_dev, &ctx->sub_dt, sizeof(double), hipMemcpyHostToDevice, s));
    ctx->chem_substep(1.0, s, ctx->sub_dt_dev, mode);
    HIPCHK(hipStreamSynchronize(s));
    return 0;
  } catch (const std::exception &e) {
    return fail(e);
  }
}

int corefl_compute_dt(corefl_ctx *ctx, double *dt_out) {
  try {
    ctx->flush_exchange(ctx->stream);
    ctx->enqueue_dt(ctx->stream, 0);
    HIPCHK(hipMemcpyAsync(dt_out, ctx->dt_dev, sizeof(double), hipMemcpyDeviceToHost, ctx->stream));
    HIPCHK(hipStreamSynchronize(ctx->stream));
    return 0;
  } catch (const std::exception &e) {
    return fail(e);
  }
}

int corefl_rk3_step(corefl_ctx *ctx, double *dt_inout, void *stream) {
  try {
    hipStream_t s = ctx->pick(stream);
    ctx->flush_exchange(s);
    const bool given = dt_inout && *dt_inout > 0;
    if (given) HIPCHK(hipMemcpyAsync(ctx->dt_dev, dt_inout, sizeof(double), hipMemcpyHostToDevice, s));
    else if (ctx->fixed_dt) {
      ctx->host_dt = ctx->param.get_real("dt");
      HIPCHK(hipMemcpyAsync(ctx->dt_dev, &ctx->host_dt, sizeof(double), hipMemcpyHostToDevice, s));
    }
    const long step = ++ctx->step_count;
    const bool mon = ctx->monitored(step);
    ctx->enqueue_step(s, !given && !ctx->fixed_dt, ctx->stores_last(step), mon);
    if (dt_inout && !given) {
      HIPCHK(hipMemcpyAsync(dt_inout, ctx->dt_dev, sizeof(double), hipMemcpyDeviceToHost, s));
    }
    HIPCHK(hipStreamSynchronize(s));
    if (mon) ctx->check_monitor(s, step);
    return 0;
  } catch (const std::exception &e) {
    return fail(e);
  }
}

int corefl_advance(corefl_ctx *ctx, int n_steps, int use_graph, void *stream) {
  try {
    hipStream_t s = ctx->pick(stream);
    ctx->flush_exchange(s); // outside any capture
    if (ctx->fixed_dt) {
      ctx->host_dt = ctx->param.get_real("dt");
      HIPCHK(hipMemcpyAsync(ctx->dt_dev, &ctx->host_dt, sizeof(double), hipMemcpyHostToDevice, s));
      HIPCHK(hipStreamSynchronize(s));
    }
    // Two captured steps: the plain one and, when the monitor is on, the monitored one (state saved before,
    // residual reduced after); the monitor only ever runs on step 1 and on multiples of output_screen.
    auto capture = [&](hipGraph_t &g, hipGraphExec_t &ge, bool store, bool residual) {
      if (ge) return;
      HIPCHK(hipStreamBeginCapture(s, hipStreamCaptureModeThreadLocal));
      try {
        ctx->enqueue_step(s, !ctx->fixed_dt, store, residual);
      } catch (...) {
        hipGraph_t tmp = nullptr;
        (void)hipStreamEndCapture(s, &tmp);
        if (tmp) (void)hipGraphDestroy(tmp);
        throw;
      }
      HIPCHK(hipStreamEndCapture(s, &g));
      HIPCHK(hipGraphInstantiate(&ge, g, nullptr, nullptr, 0));
    };
    for (int n = 0; n < n_steps; ++n) {
      const long step = ++ctx->step_count;
      const bool mon = ctx->monitored(step), store = ctx->stores_last(step);
      if (!use_graph) {
        ctx->enqueue_step(s, !ctx->fixed_dt, store, mon);
      } else if (mon) {
        // step 1 measures against the zero initial store unless it is also a multiple of output_screen
        if (!store) {
          ctx->enqueue_step(s, !ctx->fixed_dt, false, true);
        } else {
          capture(ctx->graph_mon, ctx->graph_exec_mon, true, true);
          HIPCHK(hipGraphLaunch(ctx->graph_exec_mon, s));
        }
      } else {
        capture(ctx->graph, ctx->graph_exec, false, false);
        HIPCHK(hipGraphLaunch(ctx->graph_exec, s));
      }
      if (mon) ctx->check_monitor(s, step);
    }
    return 0;
  } catch (const std::exception &e) {
    return fail(e);
  }
}

int corefl_residual(corefl_ctx *ctx, double *res, int *nan_flag, int *n_monitored) {
  try {
    if (!ctx || !res) throw std::runtime_error("corefl_residual: null argument");
    if (ctx->output_screen <= 0) throw std::runtime_error("the residual monitor is off (output_screen = 0)");
    HIPCHK(hipStreamSynchronize(ctx->stream));
    HIPCHK(hipMemcpy(ctx->mon_host, ctx->mon, sizeof(corefl::MonitorState), hipMemcpyDeviceToHost));
    for (int l = 0; l < 4; ++l) res[l] = ctx->mon_host->res[l];
    if (nan_flag) *nan_flag = ctx->mon_host->nan;
    if (n_monitored) *n_monitored = ctx->mon_host->count;
    return 0;
  } catch (const std::exception &e) {
    return fail(e);
  }
}

int corefl_limit_flow(corefl_ctx *ctx, void *stream) {
  try {
    if (!ctx) throw std::runtime_error("null context");
    if (!ctx->limit_flow) throw std::runtime_error("limit_flow is off in this setup");
    hipStream_t s = ctx->pick(stream);
    ctx->flush_exchange(s);
    for (auto &b : ctx->blocks) {
      ctx->limit(b, s);
      ctx->apply_bcs(b, s);
    }
    ctx->communicate(s);
    HIPCHK(hipStreamSynchronize(s));
    return 0;
  } catch (const std::exception &e) {
    return fail(e);
  }
}

int corefl_get_time(corefl_ctx *ctx, double *t) {
  try {
    HIPCHK(hipStreamSynchronize(ctx->stream));
    HIPCHK(hipMemcpy(t, ctx->time_dev, sizeof(double), hipMemcpyDeviceToHost));
    return 0;
  } catch (const std::exception &e) {
    return fail(e);
  }
}

int corefl_write_flowfield(corefl_ctx *ctx, const char *path) {
  try {
    if (!ctx || !path) throw std::runtime_error("corefl_write_flowfield: null argument");
    write_flowfield(ctx, path);
    return 0;
  } catch (const std::exception &e) {
    return fail(e);
  }
}

int corefl_read_flowfield(corefl_ctx *ctx, const char *path) {
  try {
    if (!ctx || !path) throw std::runtime_error("corefl_read_flowfield: null argument");
    read_flowfield(ctx, path);
    return 0;
  } catch (const std::exception &e) {
    return fail(e);
  }
}

int corefl_synchronize(corefl_ctx *ctx) {
  try {
    HIPCHK(hipStreamSynchronize(ctx->stream));
    HIPCHK(hipDeviceSynchronize());
    return 0;
  } catch (const std::exception &e) {
    return fail(e);
  }
}

int corefl_profile_kernels(corefl_ctx *ctx, int reps, double *ms_out, int n_out) {
  try {
    if (reps < 1 || n_out < 1) throw std::runtime_error("reps and n_out must be positive");
    hipStream_t s = ctx->stream;
    ctx->flush_exchange(s);
    hipEvent_t e0, e1;
    HIPCHK(hipEventCreate(&e0));
    HIPCHK(hipEventCreate(&e1));
    std::vector<double> ms(8, 0.0);
    auto timed = [&](int slot, auto &&enqueue) {
      HIPCHK(hipEventRecord(e0, s));
      for (int r = 0; r < reps; ++r) enqueue();
      HIPCHK(hipEventRecord(e1, s));
      HIPCHK(hipEventSynchronize(e1));
      float t = 0;
      HIPCHK(hipEventElapsedTime(&t, e0, e1));
      ms[slot] = t / reps;
    };
    auto &sc = ctx->scheme;
    for (auto &b : ctx->blocks) {
      if (sc.viscous_order > 0) {
        timed(0, [&] { HIPCHK(corefl::launch_viscous_flux(b.dev, sc, s)); });
        timed(1, [&] { ctx->viscous_exchange(s); });
      }
      if (sc.inviscid_scheme == 2) HIPCHK(corefl::launch_entropy_fix(b.dev, sc, s));
      timed(2, [&] { HIPCHK(corefl::launch_inviscid_flux(b.dev, sc, 0, s)); });
      timed(3, [&] { HIPCHK(corefl::launch_inviscid_flux(b.dev, sc, 1, s)); });
      if (b.dev.dim == 3) timed(4, [&] { HIPCHK(corefl::launch_inviscid_flux(b.dev, sc, 2, s)); });
      timed(5, [&] { HIPCHK(corefl::launch_update(b.dev, sc, 1, ctx->dt_dev, false, s)); });
      timed(6, [&] { ctx->apply_bcs(b, s); });
      timed(7, [&] { ctx->enqueue_dt(s, 0); });
      break; // first block only
    }
    HIPCHK(hipEventDestroy(e0));
    HIPCHK(hipEventDestroy(e1));
    for (int q = 0; q < n_out && q < 8; ++q) ms_out[q] = ms[q];
    return 0;
  } catch (const std::exception &e) {
    return fail(e);
  }
}

int corefl_chemistry(corefl_ctx *ctx, int block, corefl_chem **out) {
  try {
    if (!ctx || !out || block < 0 || block >= (int)ctx->blocks.size()) throw std::runtime_error("bad arguments");
    if (!ctx->chem) throw std::runtime_error("this setup has no species (species = 0)");
    auto &b = ctx->blocks[block];
    corefl::chem_ensure_batch(ctx->chem, b.dev.n);
    HIPCHK(hipStreamSynchronize(ctx->stream));
    HIPCHK(corefl::launch_mix_chem_gather(b.dev, b.mix, ctx->chem->b, ctx->chem->stream));
    HIPCHK(hipMemsetAsync(ctx->chem->b.h, 0, sizeof(double) * b.dev.n, ctx->chem->stream));
    HIPCHK(hipStreamSynchronize(ctx->chem->stream));
    ctx->chem->have_dtc = false;
    // the caller may retune the chemistry launch: the next graph replay captures the step afresh
    ctx->drop_graph();
    *out = ctx->chem;
    return 0;
  } catch (const std::exception &e) {
    return fail(e);
  }
}

int corefl_set_chemistry_surrogate(corefl_ctx *ctx, corefl_surrogate *s) {
  try {
    if (!ctx) throw std::runtime_error("null context");
    if (s && !ctx->strang) throw std::runtime_error("the surrogate replaces split chemistry sub-steps (temporal_scheme 5)");
    ctx->surrogate = s;
    ctx->drop_graph(); // the captured step embeds the old choice
    return 0;
  } catch (const std::exception &e) {
    return fail(e);
  }
}

int corefl_get_int(const corefl_ctx *ctx, const char *name, int *out) {
  try {
    *out = ctx->param.get_int(name);
    return 0;
  } catch (const std::exception &e) {
    return fail(e);
  }
}

int corefl_get_real(const corefl_ctx *ctx, const char *name, double *out) {
  try {
    *out = ctx->param.get_real(name);
    return 0;
  } catch (const std::exception &e) {
    return fail(e);
  }
}

} // extern "C"
