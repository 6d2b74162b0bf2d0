// Finite-rate chemistry tables (DParameter's chemistry part, DParameter.cuh:13-207) and the parser of
// COREFL's chemistry inputs: CHEMKIN mechanism + NASA-7 thermo (ChemData.cpp:10-440, 868-1177).
#pragma once

#include <string>

namespace corefl {

constexpr int kMaxSpec = 16, kMaxReac = 48, kMaxElem = 8, kMaxSide = 4;

// Plain data, copied to device memory once and read uniformly by every lane.
struct ChemMech {
  int ns, nr, ne;
  double mw[kMaxSpec];  // kg/kmol from atomic weights (Element.cpp)
  double imw[kMaxSpec]; // 1 / mw
  double t_low[kMaxSpec], t_mid[kMaxSpec], t_high[kMaxSpec];
  double ln_t_low[kMaxSpec], ln_t_high[kMaxSpec];   // log of the fit-range ends (g/RT clamps to them)
  double inv_t_low[kMaxSpec], inv_t_high[kMaxSpec]; // and their reciprocals
  double lo[kMaxSpec][7], hi[kMaxSpec][7];
  // 0 irreversible, 1 reversible, 3 merged DUPLICATE pair, 4 third body, 5 Lindemann, 6/7 Troe (3/4 params)
  int type[kMaxReac], rev[kMaxReac];
  double A[kMaxReac], b[kMaxReac], Ea[kMaxReac]; // Ea in K
  double A2[kMaxReac], b2[kMaxReac], Ea2[kMaxReac];
  double troe[kMaxReac][4];
  double order[kMaxReac]; // sum of product minus reactant coefficients
  int need_gibbs;         // some reversible reaction lacks REV parameters: use equilibrium constants
  // sparse stoichiometry: distinct species on each side with integer coefficients
  int nfs[kMaxReac], nbs[kMaxReac];
  int fsp[kMaxReac][kMaxSide], bsp[kMaxReac][kMaxSide];
  int fnu[kMaxReac][kMaxSide], bnu[kMaxReac][kMaxSide];
  double nf[kMaxReac][kMaxSpec], nb[kMaxReac][kMaxSpec]; // dense stoichiometric coefficients
  double tb[kMaxReac][kMaxSpec]; // third-body efficiencies (default 1)
  // derived once the reactions are final: bit masks of the species on either side (umask) and of those whose
  // amount changes (amask), the net coefficients nb - nf, and the integer order sum(nb) - sum(nf)
  unsigned umask[kMaxReac], amask[kMaxReac];
  int iorder[kMaxReac];
  double dnu[kMaxReac][kMaxSpec];
  int comp[kMaxSpec][kMaxElem];  // element composition
  char species[kMaxSpec][16];
  char elements[kMaxElem][4];
  // mixture-averaged transport from tran.dat (Species::read_tran, ChemData.cpp:740-831); has_transport = 1
  // once read: geometry (0 atom, 1 linear, 2 nonlinear), 1 / (eps / k_B), the Chapman-Enskog viscosity
  // coefficient, Z_rot(298) F(298); Wilke factors (W_j / W_i)^(1/4) and 1 / sqrt(8 (1 + W_i / W_j)); binary
  // diffusion coefficient D_jk p sqrt(T)^-3 Omega_D and k_B / eps_jk
  int has_transport;
  int geometry[kMaxSpec];
  double lj_inv[kMaxSpec], vis_coeff[kMaxSpec], zrot[kMaxSpec];
  double wdw[kMaxSpec][kMaxSpec], sqw[kMaxSpec][kMaxSpec], bdc[kMaxSpec][kMaxSpec], kbe[kMaxSpec][kMaxSpec];
  // ln(1 / (eps / k_B)), ln(k_B / eps_jk) and 1 / bdc: the logarithms turn the collision integrals' powers of
  // T* = T k_B / eps into exponentials of ln T + ln(k_B / eps), with one ln T per node
  double log_lj_inv[kMaxSpec], log_kbe[kMaxSpec][kMaxSpec], inv_bdc[kMaxSpec][kMaxSpec];
};

// h_k and cp_k per unit mass at t, linear extrapolation outside [t_low, t_high] (Thermo.cu:470-520).
inline void host_h_cp(const ChemMech &m, double t, double *h, double *cp) {
  for (int i = 0; i < m.ns; ++i) {
    const double R = 8314.462618 / m.mw[i];
    const bool below = t < m.t_low[i], above = t > m.t_high[i];
    const double tt = below ? m.t_low[i] : (above ? m.t_high[i] : t);
    const double *c = below ? m.lo[i] : (above ? m.hi[i] : (t < m.t_mid[i] ? m.lo[i] : m.hi[i]));
    const double hh = c[0] * tt + 0.5 * c[1] * tt * tt + c[2] * tt * tt * tt / 3 + 0.25 * c[3] * tt * tt * tt * tt +
                      0.2 * c[4] * tt * tt * tt * tt * tt + c[5];
    const double cc = c[0] + c[1] * tt + c[2] * tt * tt + c[3] * tt * tt * tt + c[4] * tt * tt * tt * tt;
    h[i] = (hh + cc * (t - tt)) * R;
    cp[i] = cc * R;
  }
}

// Throws std::runtime_error with a message on malformed input or table overflow.
ChemMech parse_mechanism(const std::string &mech_text, const std::string &therm_text);
// Read tran.dat content into m's transport tables (Species::read_tran, ChemData.cpp:740-831).
void parse_transport(const std::string &tran_text, ChemMech &m);

} // namespace corefl
