// Device-side block description and the launchers of the flow kernels (flow_kernels.hip).
#pragma once

#include <hip/hip_runtime.h>

#include <cmath>

namespace corefl {

// One block's device arrays in the DZone layout (Field.h:12-110): f(i,j,k,l) = p[idx(i,j,k) + l*sz].
struct DevBlock {
  int mx, my, mz, g, dim;
  long d2, d1, sz, n; // padded strides, padded size, interior node count
  double *cv, *bv, *qn, *jac, *met; // cv/qn: 5*sz, bv: 6*sz, met: 9*sz
  double *fx, *fy, *fz;             // interface fluxes F(i+1/2) stored at node i (5*sz each)
  double *fv, *gv, *hv;             // collocated viscous fluxes at nodes (5*sz each, l = 1..4 used)
  double *efd;                      // Roe entropy-fix delta (sz)
  double *dq;                       // residual (5*n), written only by the RHS-only path
  int bt[6];                        // physical-boundary flags il, ir, jl, jr, kl, kr
  // sponge layer (null when off): damping rate sigma per interior node (n) and the running mean of cv (nv * n)
  const double *sponge_sigma, *sponge_mean;
  // shock sensor of the hybrid flux-splitting scheme (null when off; sz, ghost entries 0)
  double *sensor;
};

struct DevFace {
  int start[3], end[3];
  int type, face, direction;
};

// A block-to-block face on the device (see InnerFace in mesh.hpp).
struct DevInner {
  int start[3], loop_dir[3], n_point[3];
  int target_start[3], target_loop_dir[3], src_tar[3];
  int face, face_layer;
  // the partner's index direction vmap[s] runs along this block's direction s, with sign vsign[s]
  int vmap[3], vsign[3];
};

// An inflow face (apply_inflow, BoundCond.cu:2190-2705) with its values per face node precomputed on the host:
// the constant inflow state or the mixing layer's tanh profile of the two streams.
struct DevInflow {
  int start[3], end[3];
  int face, direction;
  int sub_node;  // constant inflow with inflow_sub: characteristic subsonic boundary value
  int sub_ghost; // inflow_sub: ghost layers extrapolated from the boundary value, with floors on rho and p
  int fluct;     // 0 none; 1 white noise on u (constant inflow); 2 white noise on u, v and w (mixing layer)
  double mw;     // molar mass of the struct's own stream (temperature of subsonic values, mixtures)
  const double *prof; // per face node (6 + ns) values rho u v w p T Y..., [q * n_face_nodes + node]
  const double *rms;  // per face node standard deviation of the noise (fluct > 0)
  unsigned long long seed;
  unsigned long long *calls; // device count of applications: every application draws fresh noise
};

struct Scheme {
  int inviscid_scheme; // 1 LF, 2 Roe, 3 AUSM+, 4 HLLC; 51/52/71/72 LLF-split WENO5/7, component-wise/characteristic
  int reconstruction;  // schemes 1-4: 1 first order, 2 MUSCL, 3 NND2 (primitive, minmod); 4 WENO5, 5 WENO7
                       // (characteristic)
  int viscous_order;   // 0, 2 (face-centred), 4, 8 (collocated)
  double weno_eps, entropy_fix_factor, prandtl, cfl;
  double v_ref;        // v_inf: the component-wise WENO epsilon is weno_eps v_ref^0, ^2, ^4 (WENO.cu:946-949)
  int positive_preserving; // species-flux positivity limiter of the flux-splitting schemes (mixtures)
  double rho_ref;          // rho_inf: the limiter's threshold scale
  // hybrid scheme (51/52/71/72 with shockSensor_threshold > 1e-10): linear upwind weights where no node of the
  // interface's stencil has a sensor above the threshold (WENO.cu:838-849); the sensor kind and its epsilon
  double sensor_threshold;
  int shock_sensor;
  double sensor_eps;
};

// Enqueue on `s`.  Each returns hipSuccess or the launch error.
hipError_t launch_cv_from_bv_all(const DevBlock &b, hipStream_t s);
hipError_t launch_boundary(const DevBlock &b, const DevFace &f, hipStream_t s);
hipError_t launch_viscous_flux(const DevBlock &b, const Scheme &sc, hipStream_t s);
// Ghost layers of the collocated viscous fluxes across a periodic face, components 1 .. nv-1.
hipError_t launch_viscous_exchange(const DevBlock &b, const DevFace &f, hipStream_t s, int nv = 5);
// inner_communication<Air> / inner_exchange task 1 (DataCommunication.cu:36-103, 458-495) for one face of b
// whose partner block is t.
hipError_t launch_inner_exchange(const DevBlock &b, const DevBlock &t, const DevInner &f, hipStream_t s);
hipError_t launch_inner_viscous_exchange(const DevBlock &b, const DevBlock &t, const DevInner &f, hipStream_t s,
                                         int nv = 5);
hipError_t launch_entropy_fix(const DevBlock &b, const Scheme &sc, hipStream_t s);
// compute_shock_sensor (FieldOperation.cu:74-179) into b.sensor at the block's nodes; a no-op when b.sensor is null.
hipError_t launch_shock_sensor(const DevBlock &b, const Scheme &sc, hipStream_t s);
// Interface fluxes along dir of the node planes [kb, ke) (the whole block by default).
hipError_t launch_inviscid_flux(const DevBlock &b, const Scheme &sc, int dir, hipStream_t s, int kb = 0,
                                int ke = 1 << 30);
// Residual + SSP-RK3 stage update of the node planes [kb, ke) (kb a multiple of march_planes()); rhs_only writes dq
// instead of updating.  dt is read from *dt_dev.
hipError_t launch_update(const DevBlock &b, const Scheme &sc, int rk, const double *dt_dev, bool rhs_only,
                         hipStream_t s, int kb = 0, int ke = 1 << 30);
// Planes one thread of the z-marching node kernels walks along k.
int march_planes();
// CFL time step: per-workgroup partial minima into part[0 .. dt_partial_count(b)), then the global
// minimum of n partials into dt_dev[0], clipped to t_end - time_dev[0] when clip != 0.
int dt_partial_count(const DevBlock &b);
hipError_t launch_local_dt(const DevBlock &b, const Scheme &sc, double *part, hipStream_t s);
hipError_t launch_dt_reduce(const double *part, long n, double *dt_dev, const double *time_dev, double t_end,
                            int clip, hipStream_t s);
hipError_t launch_time_advance(double *time_dev, const double *dt_dev, hipStream_t s);

// Normal deviate of the inflow noise (host and device): a counter-based draw from (seed, application, node,
// component) through the SplitMix64 finaliser and Box-Muller.
__host__ __device__ inline unsigned long long noise_mix(unsigned long long z) {
  z += 0x9e3779b97f4a7c15ULL;
  z = (z ^ (z >> 30)) * 0xbf58476d1ce4e5b9ULL;
  z = (z ^ (z >> 27)) * 0x94d049bb133111ebULL;
  return z ^ (z >> 31);
}
__host__ __device__ inline double noise_normal(unsigned long long seed, unsigned long long call, long node, int comp) {
  const unsigned long long key = noise_mix(noise_mix(noise_mix(seed) + call) + (unsigned long long)node);
  const double u1 = ((double)(noise_mix(key + 2ULL * comp) >> 11) + 0.5) * 0x1p-53;
  const double u2 = ((double)(noise_mix(key + 2ULL * comp + 1) >> 11) + 0.5) * 0x1p-53;
  return sqrt(-2.0 * log(u1)) * cos(6.283185307179586 * u2);
}

} // namespace corefl
