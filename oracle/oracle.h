/*
 * oracle.h -- CPU restatement of COREFL's per-step right-hand side and SSP-RK3 time advance.
 *
 * TEST INFRASTRUCTURE ONLY.  This library is the parity checker for the HIP product in
 * corefl-cpc_amd/.  Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg may
 * load it, and only as the checker or as the timed CPU baseline -- never as a code path of
 * the product.
 *
 * Every routine cites the reference file:line it restates (paths relative to the
 * Liangerty/COREFL-CPC source tree).  Arithmetic is FP64 throughout, compiled with
 * -ffp-contract=off so that no fused multiply-add is introduced where the reference has none.
 *
 * Array layout (the reference's DZone/VectorField3D layout, gxl_lib/Array.cuh:233-328):
 *   f(i,j,k,l) = data[(k+g)*d1 + (j+g)*d2 + (i+g) + l*sz],  d2 = mx+2g, d1 = (my+2g)*d2,
 *   sz = (mx+2g)(my+2g)(mz+2g), g = ngg.  dq has no ghosts (Field.cu:524).
 *
 * Pinning status per function is stated next to each declaration; see the test files in tests/
 * that exercise each check.
 */
#ifndef COREFL_ORACLE_H
#define COREFL_ORACLE_H

#ifdef __cplusplus
extern "C" {
#endif

/* Scheme selection and constants the path reads from DParameter (DParameter.cuh:13-207). */
typedef struct {
  int inviscid_scheme;   /* 1 LF, 2 Roe, 3 AUSM+, 4 HLLC; 51/52/71/72 LLF-split WENO5/7-cp/ch      */
                         /* (Parameter.cpp:218-305)                                                */
  int reconstruction;    /* 4 WENO5 / 5 WENO7 characteristic reconstruction (new pairing)         */
  int viscous_order;     /* 0 inviscid, 2 face-centred, 4 (new), 8 collocated (ViscousScheme.cu:7-73) */
  double prandtl;        /* prandtl_number, default 0.72                                          */
  double cfl;            /* cfl                                                                   */
  double entropy_fix_factor; /* default 0.125 (Parameter.cpp:380)                                 */
  double weno_eps;       /* eps_weno * weno_eps_scale = 1e-40 (rho_inf v_inf)^2 (WENO.cu:8,884)   */
  double v_ref;          /* v_inf: scales the component-wise WENO epsilon (WENO.cu:946-949)       */
  int positive_preserving; /* species-flux limiter of the flux-splitting schemes (WENO.cu:991-1087)  */
  double rho_ref;        /* rho_inf: the limiter's threshold scale                                  */
  double sensor_threshold; /* shockSensor_threshold: > 1e-10 switches 51/52/71/72 to the hybrid scheme,  */
                           /* linear upwind weights where no node of the stencil exceeds it (WENO.cu:838-849) */
  int shock_sensor;        /* 0 Ducros, 1 Jameson, 2 density/pressure jump (FieldOperation.cu:158-179)     */
  double sensor_eps;       /* shockSensor_epsilon of the Ducros sensor                                    */
} orc_params;

/* ---- point functions (pure; parity pinned by tests/test_oracle_points.py) ---- */

/* compute_shock_sensor (FieldOperation.cu:74-179) at every block node from bv (6*sz) and the metrics (9*sz);
 * the ghost entries of sensor (sz) are left as they are (the reference never fills them). */
void orc_shock_sensor(int type, double eps, int mx, int my, int mz, int ngg, const double *bv, const double *met,
                      double *sensor);
/* WENO.cu:3059-3114.  Returns fPlus(vp)+fMinus(vm); if_shock=0 gives the linear upwind weights.  */
double orc_weno5(const double *vp, const double *vm, double eps, int if_shock);
/* WENO.cu:3116-3222 (Z-type tau7 weights). */
double orc_weno7(const double *vp, const double *vm, double eps, int if_shock);
/* The upwind-biased halves the two functions above sum (fPlus of v; fMinus is the same function of the
 * mirrored stencil), with their smoothness indicators beta and unnormalised weights alpha (either may be
 * NULL): WENO5 3 sub-stencils of v[0..4], WENO7 4 sub-stencils of v[0..6]. */
double orc_weno5_half(const double *v, double eps, int if_shock, double *beta, double *alpha);
double orc_weno7_half(const double *v, double eps, int if_shock, double *beta, double *alpha);

/* Left/right characteristic matrices at an interface from the Roe average of the two
 * conservative states ql, qr (5 each) and the (unnormalised) averaged metric k[3].
 * L rows follow WENO.cu:73-124 (weno_ch_air_x); R is the explicit back-projection of
 * WENO.cu:229-243 with H = c^2/(gamma-1) + |u|^2/2 so that R*L = I.  Outputs row-major 5x5.   */
void orc_char_matrices(const double *ql, const double *qr, const double *k, double *L, double *R);

/* Riemann fluxes on reconstructed primitive states pv = (rho,u,v,w,p,E).
 * scheme: 1 LF (InviscidScheme.cu:772-805), 2 Roe (601-765), 3 AUSM+ (397-469), 4 HLLC (471-599).
 * ml/mr: the 3 metric components of the sweep direction at the two nodes; jl/jr: Jacobians;
 * efl/efr: Roe entropy-fix deltas at the two nodes (InviscidScheme.cu:269-304).               */
void orc_riemann(int scheme, const double *pl, const double *pr, const double *ml, const double *mr,
                 double jl, double jr, double efl, double efr, int n_dim_unused, double *fc);

/* Interface flux of the product's characteristic WENO + Riemann path for one interface from a
 * stencil of conservative states q[m][5], m = 0..n-1 (n = 6 for WENO5, 8 for WENO7; the
 * interface lies between m = n/2-1 and n/2).  Returns 1 if the positivity fallback fired.      */
int orc_interface_flux(const orc_params *p, const double (*q)[5], int n, const double *ml,
                       const double *mr, double jl, double jr, double efl, double efr, double *fc);

/* Primitive-variable interface flux for air: params.reconstruction 1 first order, 2 MUSCL (kappa 1/3), 3 NND2 (minmod,
 * Reconstruction.cu:5-44) of rho, u, v, w, p at the nodes pv[0..3] (interface between 1 and 2), total energy per side,
 * then params.inviscid_scheme 1-4 (InviscidScheme.cu:61-219, 306-395).  Pinned by closed forms
 * (tests/test_pv_oracle_cpu.py). */
void orc_interface_flux_pv(const orc_params *p, const double (*pv)[5], const double *ml, const double *mr, double jl,
                           double jr, double efl, double efr, double *fc);

/* Interface flux of the reference's LLF flux-splitting WENO for air: params.inviscid_scheme 51/71
 * component-wise WENO5/WENO7 (WENO.cu:890-990), 52/72 characteristic (WENO.cu:10-244, 1090-1145).
 * q[m][5], met[m][3] (metric row of the sweep direction) and jac[m] for the stencil nodes
 * m = 0..n-1 (n = 6 WENO5, 8 WENO7); the interface lies between m = n/2-1 and n/2.  The shock
 * sensor is off (shockSensor_threshold = 0, its default), so the nonlinear weights always apply. */
void orc_fsweno_flux(const orc_params *p, const double (*q)[5], const double (*met)[3], const double *jac, int n,
                     double *fc);
/* The same with the hybrid scheme's switch: if_shock = 0 reconstructs with the linear upwind weights. */
void orc_fsweno_flux_hybrid(const orc_params *p, const double (*q)[5], const double (*met)[3], const double *jac,
                            int n, int if_shock, double *fc);

/* Collocated first derivative of f along a direction of stride st at index i of extent n: order 8 the central
 * stencil of d_dXi<8> (ViscousScheme.cu:129-307) with its one-sided rows where left / right mark a physical
 * face; order 4 its fourth-order rows.  Pinned by polynomial exactness (tests/test_oracle_cpu.py). */
double orc_deriv(const double *f, long id, long st, int i, int n, int left, int right, int order);

/* Sutherland's law, Transport.cu:6-9. */
double orc_sutherland(double T);

/* ---- mesh (Mesh.cpp) ---- */

/* The coordinates of every node of the ngg-ghost layout (sz each): the block's nodes and the mirrored ghost grid
 * orc_metrics builds (Mesh.cpp:1013-1167). */
void orc_ghost_coords(int mx, int my, int mz, int ngg, const double *x, const double *y, const double *z, double *gx,
                      double *gy, double *gz);
/* Mirror ghost coordinates (Mesh.cpp:1013-1167, ngg+1 layers) and 2nd-order metrics/Jacobian
 * (Mesh.cpp:244-528) for one block.  x,y,z: raw coordinates, mx*my*mz, i fastest, already
 * scaled by gridScale.  jac (sz) and metric (9*sz) are written in the ngg-ghost layout.
 * Returns 0, or -1 if an interior Jacobian is non-positive (the reference aborts).            */
int orc_metrics(int mx, int my, int mz, int ngg, const double *x, const double *y, const double *z,
                double *jac, double *metric);

/* ---- one block, full step ---- */

typedef struct orc_block orc_block;

/* faces: n_face rows of (i1,i2,j1,j2,k1,k2,type) as in boundary%4d.txt (Mesh.cpp:837-925);
 * types 6 (outflow) and 8 (periodic) are supported. */
orc_block *orc_block_create(const orc_params *p, int mx, int my, int mz, int ngg, const int *faces,
                            int n_face, const double *jac, const double *metric);
void orc_block_destroy(orc_block *b);
/* Set primitive variables bv (6*sz, ghosts included); computes cv over all nodes
 * (FieldOperation.cuh:39-64) and mul (FieldOperation.cuh:86-120), as initialize_computation does
 * after the boundary conditions (Driver.cu:126-211: properties -> BCs -> properties).          */
void orc_block_set_bv(orc_block *b, const double *bv);
/* Overwrite bv (6*sz) and cv (5*sz) on every node as given, with no boundary fill: used to restart the
 * oracle from another implementation's exact state. */
void orc_block_set_raw(orc_block *b, const double *bv, const double *cv);
/* Copy out an array: 0 cv, 1 bv, 2 dq, 3 mul, 4 jac, 5 metric, 6 fv, 7 gv, 8 hv. */
void orc_block_get(const orc_block *b, int which, double *out);
/* One RHS evaluation dq on the current state (RK.cuh:137-157 without the update). */
void orc_block_rhs(orc_block *b);
/* CFL time step over the interior (TimeAdvanceFunc.cuh:28-115 + TimeAdvanceFunc.cu:102-152). */
double orc_block_cfl_dt(orc_block *b);
/* One SSP-RK3 step with time step dt (RK.cuh:100-192 for one block). */
void orc_block_rk3_step(orc_block *b, double dt);
/* Blocks joined by block-to-block faces, paired by coordinates (x, y, z: interior nodes, i fastest).  A link
 * joins the whole side (dir 0/1/2, side -1 low or +1 high) of block a to block t; NULL when a face node has no
 * partner at its place. */
typedef struct orc_link orc_link;
orc_link *orc_link_find(orc_block **B, int a, const double *xa, const double *ya, const double *za, int dir, int side,
                        int t, const double *xt, const double *yt, const double *zt);
void orc_link_destroy(orc_link *L);
/* The link's direction map and node pairs (padded indices; layer 1 on the coincident layer); returns the count. */
long orc_link_info(const orc_link *L, int *vmap, int *vsign, long *ia, long *it, char *layer);
void orc_multi_exchange(orc_block **B, int nb, orc_link **L, int nl);
void orc_multi_rhs(orc_block **B, int nb, orc_link **L, int nl);
double orc_multi_cfl_dt(orc_block **B, int nb);
void orc_multi_rk3_step(orc_block **B, int nb, orc_link **L, int nl, double dt);
/* Symmetric-conservative metrics of order 2, 4 or 6 (Mesh.cu:10-250) on a 3D block, same layout and ghost grid as
 * orc_metrics; returns -1 on a non-positive Jacobian, -2 for a 2D block. */
int orc_sc_metrics(int mx, int my, int mz, int ngg, const double *x, const double *y, const double *z, int order,
                   double *jac, double *metric);
/* Number of OpenMP threads the oracle uses (for the CPU baseline report). */
int orc_num_threads(void);
void orc_set_threads(int n); /* OpenMP threads of the following oracle calls */

/* ---- chemistry of one constant-volume cell ---- */

#define ORC_MAX_SPEC 16
#define ORC_MAX_REAC 48

/* Mechanism tables as the reference keeps them in DParameter (DParameter.cuh:13-207): NASA-7 two-range
 * coefficients, reaction types (0 irreversible, 1 reversible, 3 duplicate pair, 4 third body,
 * 5 Lindemann, 6/7 Troe with 3/4 parameters), Arrhenius triples (Ea in K), stoichiometry, order
 * (sum of product minus reactant coefficients), third-body efficiencies. */
typedef struct {
  int ns, nr;
  double mw[ORC_MAX_SPEC];
  double t_low[ORC_MAX_SPEC], t_mid[ORC_MAX_SPEC], t_high[ORC_MAX_SPEC];
  double lo[ORC_MAX_SPEC][7], hi[ORC_MAX_SPEC][7];
  int type[ORC_MAX_REAC], rev[ORC_MAX_REAC];
  double A[ORC_MAX_REAC], b[ORC_MAX_REAC], Ea[ORC_MAX_REAC];
  double A2[ORC_MAX_REAC], b2[ORC_MAX_REAC], Ea2[ORC_MAX_REAC];
  double troe[ORC_MAX_REAC][4];
  double order[ORC_MAX_REAC];
  double nf[ORC_MAX_REAC][ORC_MAX_SPEC], nb[ORC_MAX_REAC][ORC_MAX_SPEC];
  double tb[ORC_MAX_REAC][ORC_MAX_SPEC];
} orc_mech;

/* h_k, cp_k per unit mass (Thermo.cu:470-520). */
void orc_chem_h_cp(const orc_mech *m, double t, double *h, double *cp);
/* Chemical source and rates of progress (FiniteRateChem.cu:1466-1605). */
void orc_chem_source(const orc_mech *m, double t, const double *rhoY, double *q1, double *q2, double *omega_d,
                     double *omega);
/* d omega / d(rho Y) (ConstVolumeReactor.cpp:1168-1187). */
void orc_chem_jacobian(const orc_mech *m, const double *rhoY, const double *q1, const double *q2, double *jac);
/* Doolittle LU with partial pivoting, row-major, and its solve (ConstVolumeReactor.cpp:1231-1285,
 * FiniteRateChem.cu:1769-1822). */
void orc_lu_decomp(double *a, int n, int *piv);
void orc_lu_solve(const double *a, double *b, int n, const int *piv);
/* T from internal energy by Newton (ConstVolumeReactor.cpp:970-995). */
double orc_chem_update_t(const orc_mech *m, double t0, const double *yk, double E);
/* Adaptive ROK4E at constant rho and E (ConstVolumeReactor.cpp:173-497); -1 if max_attempts attempts
 * (accepted + rejected) end before t_end. */
int orc_chem_rok4e(const orc_mech *m, double rho, double E, double *rhoY, double *T, double t_end, double rtol,
                   double atol, double *h, int max_attempts, int *n_rej, double *rec, int n_rec);

/* ---- mixture block (oracle_mix.c): primitive reconstruction + mixture Riemann solvers, RK3,
 *      Strang-split ROK4E chemistry.  params.reconstruction: 1 first order, 2 MUSCL, 3 NND2 (minmod),
 *      4 / 5 component-wise WENO5 / WENO7; params.inviscid_scheme: 2 Roe, 3 AUSM+, 4 HLLC. ---- */
typedef struct orc_mix orc_mix;
orc_mix *orc_mix_create(const orc_params *p, const orc_mech *m, int mx, int my, int mz, int ngg, const int *faces,
                        int n_face, const double *jac, const double *metric);
void orc_mix_destroy(orc_mix *b);
/* bv (6*sz) and sv (ns*sz) on every node; cv from them, boundary conditions, acoustic speed. */
void orc_mix_set_state(orc_mix *b, const double *bv, const double *sv);
void orc_mix_set_raw(orc_mix *b, const double *bv, const double *sv, const double *cv);
/* which: 0 cv ((5+ns)*sz), 1 bv (6*sz), 2 dq ((5+ns)*n), 3 sv (ns*sz), 4 acoustic speed (sz) */
void orc_mix_get(const orc_mix *b, int which, double *out);
/* Add the explicit chemical source to the residual and limit the CFL step by the reaction time scale. */
void orc_mix_set_explicit_source(orc_mix *b, int on);
void orc_mix_rhs(orc_mix *b);
double orc_mix_cfl_dt(orc_mix *b);
void orc_mix_rk3_step(orc_mix *b, double dt);
/* Step size the positivity limiter uses in the next orc_mix_rhs (the step functions set it themselves). */
void orc_mix_set_step(orc_mix *b, double dt);
int orc_mix_chem_substep(orc_mix *b, double dt, double rtol, double atol, int max_attempts);
int orc_mix_strang_step(orc_mix *b, double dt, double rtol, double atol, int max_attempts);
/* Transport data of tran.dat per species (geometry 0/1/2, eps/k_B in K, sigma in Angstrom, dipole moment in
 * Debye, polarizability in Angstrom^3, Z_rot at 298 K). */
typedef struct {
  int geometry[ORC_MAX_SPEC];
  double eps_kb[ORC_MAX_SPEC], sigma[ORC_MAX_SPEC], dipole[ORC_MAX_SPEC], polar[ORC_MAX_SPEC], zrot[ORC_MAX_SPEC];
} orc_tran;
/* Mixture viscosity, conductivity and mixture-averaged rho D_k at T and mass fractions Y
 * (compute_transport_property, Transport.cu:43-150, with the coefficients of Species::read_tran,
 * ChemData.cpp:740-831, derived here from the raw data). */
void orc_transport(const orc_mech *m, const orc_tran *t, double T, const double *Y, double *mu, double *lambda,
                   double *rhoD);
/* Pure-species values at T of the function above: out[0] viscosity, out[1] conductivity, out[2] self-diffusion
 * rho D_kk (the coefficients of Transport.cu:67-105 for species k alone). */
void orc_transport_species(const orc_mech *m, const orc_tran *t, double T, int k, double *out);
/* Viscous mixtures (viscous_order 2): the transport data the face-centred viscous fluxes use. */
void orc_mix_set_transport(orc_mix *b, const orc_tran *t);
/* Wu's splitting (temporal_scheme 4), one step: chem_advance_method 0 (forward Euler) or 3 (SSP-RK3 stages). */
void orc_mix_wu_step(orc_mix *b, double dt, int method);
/* One interface flux of the mixture Riemann solver params->inviscid_scheme from primitive states
 * (rho,u,v,w,p,Y_1..Y_ns), after the reference post-processing for E and gamma. */
void orc_mix_point_flux(const orc_params *p, const orc_mech *m, const double *pl, const double *pr,
                        const double *ml, const double *mr, double jl, double jr, double efl, double efr, double *fc);

/* The species positivity limiter of the flux-splitting schemes (WENO.cu:991-1087) on one interface: stencil
 * states q[m][5 + ns] (conservative) with metric k[m][3] along the sweep, jac[m] and acoustic speed snd[m],
 * m = 0..n-1 with the interface between n/2-1 and n/2; fc (5 + ns) is limited in place with step dt. */
void orc_mix_positivity_point(const orc_params *p, int ns, int dim, double dt, const double *q, const double *k,
                              const double *jac, const double *snd, int n, double *fc);

/* ---- inflow and symmetry faces, sponge layer (oracle_bc.c) ---- */

/* Normal deviate of the inflow noise from (seed, application, face node, component): SplitMix64 finaliser and
 * Box-Muller.  This build's own counter-based generator (the reference draws clock-seeded curand numbers). */
double orc_noise_normal(unsigned long long seed, unsigned long long call, long node, int comp);
/* apply_inflow (BoundCond.cu:2190-2705) on face nodes s..e (normal `face`, side dir): prof (6 + ns values per face
 * node: rho u v w p T Y..., [q * npt + node]) is the state without noise; sub_node: constant inflow with inflow_sub
 * (characteristic subsonic value, BoundCond.cu:2505-2553); sub_ghost: ghosts extrapolated 2 b - interior with floors;
 * fluct 1: noise on u (rms per node, BoundCond.cu:2562-2597), 2: on u, v, w (mixing layer, 2479-2503); mw: molar mass
 * of the subsonic temperatures (mixtures); snd: acoustic speed array (NULL for air).  cv is left to the caller.
 * Pinned by closed forms (tests/test_bc_oracle_cpu.py). */
void orc_bc_inflow(int mx, int my, int mz, int ngg, const int *s, const int *e, int face, int dir, double *bv,
                   double *sv, int ns, const double *snd, const double *met, const double *prof, const double *rms,
                   int sub_node, int sub_ghost, int fluct, double mw, unsigned long long seed,
                   unsigned long long call);
/* apply_symmetry (BoundCond.cu:2002-2063).  Pinned by closed forms (tests/test_bc_oracle_cpu.py). */
void orc_bc_symmetry(int mx, int my, int mz, int ngg, const int *s, const int *e, int face, int dir, double *bv,
                     double *sv, int ns, const double *met);
/* The sponge layer's ramp 6 xi^5 - 15 xi^4 + 10 xi^3 (SpongeLayer.cu:164-172). */
double orc_sponge_ramp(double xi);
/* Inflow data of face `face` of an air block (faces of type 5): prof 6 values per face node, rms per node; the
 * noise of the block's applications counts from first_call. */
void orc_block_set_inflow(orc_block *b, int face, const double *prof, const double *rms, int sub_node, int sub_ghost,
                          int fluct, double mw, unsigned long long seed, unsigned long long first_call);
/* The same for a mixture block (prof: 6 + ns values per face node). */
void orc_mix_set_inflow(orc_mix *b, int face, const double *prof, const double *rms, int sub_node, int sub_ghost,
                        int fluct, double mw, unsigned long long seed, unsigned long long first_call);
void orc_mix_set_sponge(orc_mix *b, const double *sigma);
/* Sponge layer of an air block: damping rate per interior node; the residual gets -J sigma (U - U_mean), the mean
 * of U over the steps taken, updated at the start of every step. */
void orc_block_set_sponge(orc_block *b, const double *sigma);

/* ---- step monitor and limit_flow (oracle_monitor.c) ---- */

/* store_last_step (TimeAdvanceFunc.cu:6-19): rho, |V|, p, T of the interior nodes into last (4 * mx*my*mz). */
void orc_store_last(int mx, int my, int mz, int ngg, const double *bv, double *last);
/* Sums over the interior of the squared changes of rho, |V|, p, T against last (compute_square_of_dbv +
 * reduction_of_dv_squared, TimeAdvanceFunc.cu:84-100, Residual.cu:8-60); the residual is sqrt(sum / n) over
 * all blocks' nodes, divided by the first monitored step's value.  Pinned by closed forms
 * (tests/test_monitor_oracle_cpu.py). */
void orc_residual_sums(int mx, int my, int mz, int ngg, const double *bv, const double *last, double *sums);
/* limit_flow (TimeAdvanceFunc.cuh:117-220) on bv (6*sz, in place): returns the number of nodes changed.  Rgas:
 * per-node mixture gas constant (sz values) or NULL for air.  In-bounds status is decided before any change (the
 * reference kernel reads neighbours that other threads may be repairing).  Pinned by closed forms
 * (tests/test_monitor_oracle_cpu.py). */
int orc_limit_flow(int mx, int my, int mz, int ngg, double *bv, const double *ll, const double *ul,
                   const double *Rgas);

/* ---- exact solution ---- */

/* Exact Riemann solution of the 1D Euler equations (Toro, "Riemann Solvers and Numerical
 * Methods for Fluid Dynamics", ch. 4: pressure function + Newton iteration, then sampling).
 * Pinned by example/sodShockTube/exact.dat. Returns the star pressure. */
double orc_exact_riemann(double rl, double ul, double pl, double rr, double ur, double pr,
                         double gamma, double x0, double t, int n, const double *x, double *rho,
                         double *u, double *p);

#ifdef __cplusplus
}
#endif
#endif
