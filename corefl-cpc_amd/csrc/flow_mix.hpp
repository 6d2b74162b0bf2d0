// Multi-species (mixture) flow on one block: the reference's primitive-variable path
// (compute_convective_term_pv / Roe_compute_inviscid_flux) with mixture thermodynamics, the mixture SSP-RK3
// update with the Newton temperature, boundary fills with species, and the Strang-split chemistry
// exchange with the batched chemistry kernels.
#pragma once

#include <hip/hip_runtime.h>

#include "chem.hpp"
#include "chem_kernels.hpp"
#include "flow.hpp"

namespace corefl {

// Species arrays of a mixture block (DZone sv and acoustic_speed, Field.cu:473-570); cv, qn, fx/fy/fz
// of the DevBlock hold 5 + ns components and dq (5 + ns) * n.
struct MixArrays {
  int ns;
  double *sv;  // Y_k, ns * sz
  double *snd; // acoustic speed, sz
  const ChemMech *mech; // device tables
  int explicit_source;  // add the chemical source to the residual (reaction = 1 with temporal_scheme 3)
  double *tscale;       // reaction time scale per interior node (n), written with the explicit source
  const double *dt;     // the step size on the device (read by the positivity limiter)
  // viscous mixtures: node transport properties (viscosity, conductivity: sz; rho D_k: ns * sz)
  double *mul, *cond, *rhod;
  int visc_order; // 0 inviscid, 2 face-centred fluxes, 8 collocated node fluxes (exchanged, then differenced)
};

hipError_t launch_mix_cv_from_bv_all(const DevBlock &b, const MixArrays &x, hipStream_t s);
hipError_t launch_mix_boundary(const DevBlock &b, const MixArrays &x, const DevFace &f, hipStream_t s);
// inner_communication<Mixture> (DataCommunication.cu:36-103) for one face of b whose partner block is t.
hipError_t launch_mix_inner_exchange(const DevBlock &b, const MixArrays &x, const DevBlock &t, const MixArrays &y,
                                     const DevInner &f, hipStream_t s);
hipError_t launch_mix_entropy_fix(const DevBlock &b, const MixArrays &x, const Scheme &sc, hipStream_t s);
// reconstruction 1 first order, 2 MUSCL, 3 NND2 (minmod), 4 / 5 component-wise WENO5 / WENO7;
// inviscid_scheme 2 Roe, 3 AUSM+, 4 HLLC.
hipError_t launch_mix_inviscid_flux(const DevBlock &b, const MixArrays &x, const Scheme &sc, int dir, hipStream_t s);
hipError_t launch_mix_update(const DevBlock &b, const MixArrays &x, int rk, const double *dt_dev, bool rhs_only,
                             hipStream_t s);
// Wu's splitting (temporal_scheme 4): the reaction sub-step with the frozen transport residual held in b.dq,
// and the correction stage over dt/2 with R(U) - b.dq (StrangSplitting.cu:373-421, 504-634).
hipError_t launch_mix_wu_reaction(const DevBlock &b, const MixArrays &x, const double *dt_dev, int method,
                                  hipStream_t s);
hipError_t launch_mix_wu_update(const DevBlock &b, const MixArrays &x, int rk, const double *dt_dev, hipStream_t s);
// Viscous mixtures: transport properties on every node (compute_transport_property, Transport.cu:43-150), then
// for viscous_order 2 the face-centred viscous fluxes with species diffusion (compute_fv/gv/hv_2nd_order<Mixture>,
// ViscousScheme.cu:614-1010), for viscous_order 8 the collocated node fluxes (compute_viscous_flux_collocated +
// _scalar, ViscousScheme.cu:309-566; their ghost layers come from the viscous-flux exchange) into b.fv / b.gv /
// b.hv (nv components).
hipError_t launch_mix_viscous(const DevBlock &b, const MixArrays &x, const Scheme &sc, hipStream_t s);
hipError_t launch_mix_transport(const DevBlock &b, const MixArrays &x, hipStream_t s);
hipError_t launch_mix_local_dt(const DevBlock &b, const MixArrays &x, const Scheme &sc, double *part, hipStream_t s);
// Interior cells -> chemistry batch (rho, e, T, rho Y; n = mx*my*mz cells, i fastest) and back.
hipError_t launch_mix_chem_gather(const DevBlock &b, const MixArrays &x, const ChemBatch &c, hipStream_t s);
hipError_t launch_mix_chem_scatter(const DevBlock &b, const MixArrays &x, const ChemBatch &c, hipStream_t s);

} // namespace corefl
