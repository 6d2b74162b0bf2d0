/*
 * corefl_c.h -- C ABI of the MI355X-native COREFL right-hand side and SSP-RK3 time advance.
 *
 * The reference has no plugin or FFI layer: its per-step path is reached through internal C++
 * templates called from RK3<mix>() (src/RK.cuh:19-222).  Each entry point below names the
 * reference interface it replaces.  Conventions:
 *   - return codes are int: 0 = OK, negative = error; corefl_last_error() gives the message
 *     (the reference polls cudaGetLastError and calls MPI_Abort, src/RK.cuh:49-53, 214-218);
 *   - the library owns every device buffer; the caller owns its host buffers;
 *   - arrays use the reference's DZone layout (src/Field.h:12-110, src/gxl_lib/Array.cuh:233-328):
 *     f(i,j,k,l) = data[(k+g)*d1 + (j+g)*d2 + (i+g) + l*sz], d2 = mx+2g, d1 = (my+2g)*d2,
 *     sz = (mx+2g)(my+2g)(mz+2g), g = ngg; dq has no ghost layers (src/Field.cu:524);
 *   - one context per host thread; all work is ordered on the context's HIP stream (or the stream
 *     passed in), with no host synchronisation inside a step.
 * No torch, HIP or C++ types appear in these signatures; streams are passed as void*.
 */
#ifndef COREFL_C_H
#define COREFL_C_H

#ifdef __cplusplus
extern "C" {
#endif

typedef struct corefl_ctx corefl_ctx;
typedef struct corefl_chem corefl_chem;           /* chemistry context, declared below */
typedef struct corefl_surrogate corefl_surrogate; /* surrogate context, declared below */

/* Zero-copy view of one block's device arrays (the DZone of src/Field.h:12-110). */
typedef struct {
  int mx, my, mz, ngg, n_var, n_spec;
  long long sz;      /* padded node count */
  double *cv;        /* conservative variables, n_var*sz */
  double *bv;        /* rho, u, v, w, p, T, 6*sz */
  double *qn;        /* U^n saved at the start of a step, n_var*sz */
  double *metric;    /* xi_x xi_y xi_z eta_x ... zeta_z, 9*sz */
  double *jac;       /* Jacobian, sz */
  double *dq;        /* residual, n_var*mx*my*mz, written by corefl_rhs */
} corefl_block_view;

/* Array selectors for corefl_get_array. */
enum {
  COREFL_CV = 0,
  COREFL_BV = 1,
  COREFL_DQ = 2,
  COREFL_SV = 3,
  COREFL_JAC = 4,
  COREFL_METRIC = 5,
  COREFL_FV = 6,
  COREFL_GV = 7,
  COREFL_HV = 8,
  COREFL_SENSOR = 9 /* shock sensor of the hybrid flux-splitting scheme (DZone::shock_sensor, src/Field.h:75) */
};

/* Replaces Parameter(&argc,&argv) + Mesh(parameter) + Driver<mix>(parameter, mesh)
 * (src/Parameter.cpp:9-22, src/Mesh.cpp:619-682, src/Driver.cu:13-124): reads
 * <input_dir>/setup.txt, <input_dir>/grid/grid   0.dat, <input_dir>/boundary_condition/boundary   0.txt
 * and, when present, <input_dir>/boundary_condition/inner   0.txt (block-to-block faces; the ghost layers
 * across them are refreshed after every stage as data_communication does, src/DataCommunication.cu:6-103),
 * computes metrics, allocates device memory and sets the initial condition
 * (src/Initialize.cu:51-95) followed by initialize_computation (src/Driver.cu:126-211). */
int corefl_create(const char *input_dir, corefl_ctx **out);

/* Same as corefl_create with the inputs in memory: setup_text is the content of setup.txt;
 * dims = n_block*(mx,my,mz); xyz[3*b+0..2] = x, y, z of block b (i fastest, mx*my*mz each);
 * boundary_text is the content of boundary%4d.txt; inner_text the content of inner%4d.txt (the
 * block-to-block faces read by Mesh::read_inner_interface, src/Mesh.cpp:927-960), or NULL for none. */
int corefl_create_mem(const char *setup_text, int n_block, const int *dims, const double *const *xyz,
                      const char *boundary_text, const char *inner_text, corefl_ctx **out);

void corefl_destroy(corefl_ctx *ctx);
const char *corefl_last_error(void);

int corefl_n_blocks(const corefl_ctx *ctx);
/* Replaces direct use of Field::h_ptr / DZone (src/Field.h:164-165). */
int corefl_block_view_get(corefl_ctx *ctx, int block, corefl_block_view *view);
/* Device-to-host copy of one array (see the COREFL_* selectors); n = element count of `host`. */
int corefl_get_array(corefl_ctx *ctx, int block, int which, double *host, long long n);
/* Host-to-device primitive state (ghosts included): bv (6*sz) followed, for mixtures, by the species mass
 * fractions sv (n_spec*sz); recomputes cv on all nodes and applies the block's boundary conditions
 * (src/FieldOperation.cuh:39-64, src/BoundCond.cu:3489-3619).  The exchange across block-to-block faces
 * (src/DataCommunication.cu:6-103) runs before the next computation, once every block has been set. */
int corefl_set_state(corefl_ctx *ctx, int block, const double *bv_host, long long n);

/* Right-hand side of the current state on every block: viscous flux (src/ViscousScheme.cu:7-73) and
 * inviscid flux (src/InviscidScheme.cu:12-26).  stage = -1 writes the residual dq and leaves the state
 * unchanged.  stage = 0, 1, 2 runs that SSP-RK3 stage as the reference's RK3 loop does (src/RK.cuh:137-192):
 * the residual, then update_cv_and_bv_rk with the stage's coefficients (src/RK.cuh:224-280; stage 0 also
 * saves U^n), boundary conditions and the block-to-block exchange, with the step size held on the device
 * (corefl_compute_dt, corefl_set_dt, the setup's fixed dt or the last corefl_rk3_step).  Stages of Wu's
 * splitting are not exposed (error). */
int corefl_rhs(corefl_ctx *ctx, int stage, void *stream);

/* Step size held on the device for the following stages (the host side of DParameter::dt that
 * global_time_step writes, src/TimeAdvanceFunc.cu:102-127). */
int corefl_set_dt(corefl_ctx *ctx, double dt);

/* CFL time step: local_time_step + global_time_step (src/TimeAdvanceFunc.cuh:28-115,
 * src/TimeAdvanceFunc.cu:102-127), reduced on the device; copies the value to *dt_out. */
int corefl_compute_dt(corefl_ctx *ctx, double *dt_out);

/* One SSP-RK3 step of every block (src/RK.cuh:84-192).  *dt_inout > 0 is used as the step;
 * otherwise the CFL step is computed on the device (clipped to total_simulation_time) and returned. */
int corefl_rk3_step(corefl_ctx *ctx, double *dt_inout, void *stream);

/* n_steps SSP-RK3 steps without host synchronisation.  The step size is the setup's fixed dt when
 * fixed_time_step = 1, else the device-side CFL step.  use_graph != 0 captures one step in a
 * hipGraph on first use and replays it. */
int corefl_advance(corefl_ctx *ctx, int n_steps, int use_graph, void *stream);

/* Residual monitor (compute_residual, src/Residual.cu:8-121), on when setup.txt gives output_screen > 0: the
 * steps taken by corefl_rk3_step / corefl_advance are counted as the reference's RK3 loop counts them; before
 * every step that is a multiple of output_screen rho, |V|, p and T are saved (store_last_step,
 * src/TimeAdvanceFunc.cu:6-19), and after that step and after step 1 the RMS change of each over all interior
 * nodes is reduced on the device (src/TimeAdvanceFunc.cu:84-100, src/Residual.cu:123-170), divided by the first
 * such value (res_scale, floored at 1e-20), and a NaN flag is set when any of them is NaN.  Those steps end with
 * a stream synchronisation, as the reference's residual copy does, and a NaN residual makes corefl_rk3_step /
 * corefl_advance return an error ("NaN occurred in step N", src/Residual.cu:100-118).  This call gives the last
 * monitored step's res[0..3] (rho, |V|, p, T), its NaN flag and the number of monitored steps so far. */
int corefl_residual(corefl_ctx *ctx, double *res, int *nan_flag, int *n_monitored);

/* limit_flow (src/TimeAdvanceFunc.cuh:117-220), which the stages also run after every update when setup.txt sets
 * limit_flow = 1: every interior node whose rho, u, v, w or p is NaN or outside the bounds of
 * src/DParameter.cu:332-349 (rho and p within 1e-6..1e3 of the reference state's, velocities within 1000 v_ref)
 * is set to the mean of the in-bounds nodes of its 3x3x3 neighbourhood inside the block (its clipped values when
 * there are none), T from the equation of state and cv from bv; then boundary conditions and block exchanges.
 * Which neighbours count as in bounds is decided before any node is changed. */
int corefl_limit_flow(corefl_ctx *ctx, void *stream);

/* Simulated physical time (device accumulator) and a stream synchronisation. */
int corefl_get_time(corefl_ctx *ctx, double *t);
int corefl_synchronize(corefl_ctx *ctx);

/* The reference's restart file flowfield.plt (Tecplot binary #!TDV112, one ordered zone per block, all variables
 * double precision at the nodes): replaces FieldIO::initialize + FieldIO::print_field (src/FieldIO.h:89-430,
 * 709-825; variables x, y, z, density, u, v, w, pressure, temperature, species, mach as
 * Parameter::get_variable_names gives them for a laminar run, src/Parameter.cpp:750-858; no ghost layers).  Writes
 * the current state and simulated time to path. */
int corefl_write_flowfield(corefl_ctx *ctx, const char *path);
/* Replaces read_flowfield (src/Initialize.cu:97-345): the file's zones must be the context's blocks (in order or
 * permuted, matched by size); density, u, v, w, pressure, temperature and species found by name replace the
 * state (with the file's ghost layers where it has them), the time becomes the file's solution time, then cv,
 * boundary conditions and block-to-block exchanges follow.  With "int initial = 1" in setup.txt, corefl_create
 * reads <input_dir>/../output/flowfield.plt this way (src/Initialize.cu:36-41). */
int corefl_read_flowfield(corefl_ctx *ctx, const char *path);

/* Kernel timing on the context's stream with HIP events (no reference counterpart: the reference has
 * no tracing, src/gxl_lib/Time.cpp:4-20).  Runs `reps` launches of each kernel class of one stage on the
 * current state and writes the average milliseconds per launch into ms_out[0..n_out):
 *   0 viscous node fluxes, 1 viscous ghost exchange (all faces), 2 x flux, 3 y flux, 4 z flux,
 *   5 residual + RK update (rk = 1 coefficients), 6 boundary fills (all faces), 7 CFL step (local + reduce).
 * The state is advanced by the update launches; call it on a scratch context. */
int corefl_profile_kernels(corefl_ctx *ctx, int reps, double *ms_out, int n_out);

/* Integer / real parameter after defaults and derived values (e.g. "ngg", "n_var", "dimension",
 * "inviscid_scheme"; "weno_eps", "rho_inf", "v_inf", "dt"). */
int corefl_get_int(const corefl_ctx *ctx, const char *name, int *out);
int corefl_get_real(const corefl_ctx *ctx, const char *name, double *out);

/* Mixtures (species = 1).  Setup keys: mechanism_file, therm_file (relative to the input directory, as
 * src/ChemData.cpp:13,82 read them), reaction, temporal_scheme 3 with reaction = 1 (explicit chemical
 * source in the residual, finite_rate_chemistry with chemSrcMethod 0, src/FiniteRateChem.cu:1388-1464)
 * or temporal_scheme 5 (new: Strang splitting, chemistry dt/2 - SSP-RK3 step - chemistry dt/2, each
 * chemistry sub-step the batched ROK4E of corefl_chem_integrate; keys chemistry_rtol, chemistry_atol).
 * The borrowed chemistry context of a mixture setup, its batch filled with the interior cells of `block`
 * (the state before the next step), for corefl_chem_autotune / corefl_chem_set_launch, which the split
 * sub-steps then use (the next graph replay of corefl_advance captures the step afresh).  Owned by ctx. */
int corefl_chemistry(corefl_ctx *ctx, int block, corefl_chem **out);
/* Split chemistry sub-steps through corefl_chem_surrogate_step's path (network, projection, gate, exact
 * fallback) instead of the exact integrator; NULL restores the exact sub-steps.  s is not owned and must
 * outlive its use.  No reference counterpart. */
int corefl_set_chemistry_surrogate(corefl_ctx *ctx, corefl_surrogate *s);
/* One chemistry sub-step of length dt over every interior cell of every block at constant density and
 * internal energy, as the split step's half steps run it, then boundary conditions and block exchanges.
 * mode 0: the batched exact ROK4E; mode 1: the surrogate set by corefl_set_chemistry_surrogate (network,
 * projection, gate, exact fallback; a gate wide open sends every cell through the network).  Needs
 * temporal_scheme 5.  Replaces the reaction part of the reference's split step (wu_reaction_step,
 * src/StrangSplitting.cu:373-421, which integrates each cell with its own sub-steps). */
int corefl_chem_substep(corefl_ctx *ctx, double dt, int mode, void *stream);

/* ---------------------------------------------------------------------------------------------------
 * Finite-rate chemistry of cell batches: CHEMKIN mechanism + NASA-7 thermo, the chemical source, and the
 * stiff per-cell sub-step (4-stage Rosenbrock ROK4E with embedded error control) batched over cells.
 * The context owns its device tables, a batch of cells (structure of arrays) and a stream.
 * ------------------------------------------------------------------------------------------------- */

/* Device pointers of the current batch (n cells); rhoY is [k * n + cell]. */
typedef struct {
  int n, n_spec;
  double *rho;   /* kg/m^3 */
  double *e;     /* specific internal energy, J/kg, constant over a sub-step */
  double *T;     /* K */
  double *rhoY;  /* kg/m^3 */
  double *h;     /* step size carried between sub-steps (0: start from dt) */
  int *steps;    /* accepted steps of the last sub-step, -1 if the attempt limit was reached */
  int *rejects;  /* rejected attempts of the last sub-step */
} corefl_chem_view;

/* Replaces Species::Species + Reaction::Reaction (src/ChemData.cpp:10-100, 868-1177) and the chemistry
 * tables of DParameter::DParameter (src/DParameter.cu:6-394): reads a CHEMKIN mechanism file (ELEMENTS,
 * SPECIES, optional THERMO, REACTIONS) and, when therm_file is non-empty, a NASA-7 thermo file.
 * Activation energies are converted to K with R_c = 1.98726 cal/(mol K); DUPLICATE pairs are merged. */
int corefl_chem_create(const char *mechanism_file, const char *therm_file, corefl_chem **out);
/* Same with the two files' contents in memory (therm_text may be ""). */
int corefl_chem_create_mem(const char *mech_text, const char *therm_text, corefl_chem **out);
void corefl_chem_destroy(corefl_chem *chem);
/* Species, reaction (after DUPLICATE merging) and element counts (Species::n_spec, Reaction::n_reac). */
int corefl_chem_info(const corefl_chem *chem, int *n_spec, int *n_reac, int *n_elem);
/* Name of species (kind 0) or element (kind 1) `index` (Species::spec_name, src/ChemData.cpp:30-60). */
int corefl_chem_name(const corefl_chem *chem, int kind, int index, char *buf, int len);
/* Host copy of one parsed table, as DParameter holds it (src/DParameter.cuh:100-180): "mw", "t_low",
 * "t_mid", "t_high", "lo", "hi" (n_spec x 7), "A", "b", "Ea", "A2", "b2", "Ea2", "order", "type", "rev",
 * "troe" (n_reac x 4), "nf", "nb", "tb" (n_reac x n_spec), "comp" (n_spec x n_elem).  n must equal the
 * table's element count.  Needs no GPU. */
int corefl_chem_table(const corefl_chem *chem, const char *name, double *out, long long n);

/* Host-to-device batch of n cells (replaces the per-cell state of DZone sv/bv that
 * finite_rate_chemistry reads, src/FiniteRateChem.cu:1388-1400).  e may be NULL: then
 * e = sum_k Y_k h_k(T) - R T is computed on the device (src/ConstVolumeReactor.cpp:60-68).
 * Resets the carried step sizes and counters. */
int corefl_chem_set_cells(corefl_chem *chem, long long n, const double *rho, const double *e, const double *T,
                          const double *rhoY);
/* Per-cell sub-step lengths (n values > 0) used by corefl_chem_integrate with dt == 0; NULL clears them.
 * corefl_chem_set_cells clears them too. */
int corefl_chem_set_cell_dt(corefl_chem *chem, const double *dt_host);
/* Device-to-host copy of the batch; any pointer may be NULL. */
int corefl_chem_get_cells(corefl_chem *chem, double *T, double *rhoY, double *e, int *steps, int *rejects);
/* Zero-copy device pointers of the batch. */
int corefl_chem_view_get(corefl_chem *chem, corefl_chem_view *view);

/* Chemical source omega_k [kg/(m^3 s)] of every cell at its current T and rhoY, copied to
 * omega_host[k * n + cell] (forward_reaction_rate + backward_reaction_rate + rate_of_progress +
 * chemical_source, src/FiniteRateChem.cu:1466-1605). */
int corefl_chem_source(corefl_chem *chem, double *omega_host);

/* One chemistry sub-step of length dt (dt == 0: each cell's own length) for every cell at constant
 * density and internal energy: adaptive
 * ROK4E (src/ConstVolumeReactor.cpp:173-476 with adaptive_time_step = 1, constants src/Define.h:28-43,
 * temperature by update_t, 970-995), error norm RMS(err_k / (rtol * rhoY_k + atol)).  Each cell stops
 * after max_attempts step attempts (steps = -1 flags it).  Launch shape: the autotuned or set one.
 * Stream-ordered on `stream` (NULL: the context's stream). */
int corefl_chem_integrate(corefl_chem *chem, double dt, double rtol, double atol, int max_attempts, void *stream);

/* Launch shape of corefl_chem_integrate (no reference counterpart: the reference integrates one cell per
 * thread, src/FiniteRateChem.cu:1388).  mode 0: one thread per cell; 1: persistent waves fetching cells
 * from a device counter, `blocks` workgroups; 2: persistent waves integrating four cells at a time, 16 lanes
 * per cell (lower latency per step attempt, for stiff cells); 3: mode 1 for at most `budget` attempts per
 * cell, the unfinished cells then continued by mode 2.  All modes give the same result to rounding.
 * threads is 64, 128 or 256; budget is used by mode 3 only. */
int corefl_chem_set_launch(corefl_chem *chem, int mode, int threads, int blocks, int budget);
int corefl_chem_get_launch(const corefl_chem *chem, int *mode, int *threads, int *blocks, int *budget);
/* Times corefl_chem_integrate(dt, rtol, atol, max_attempts) on the current batch for each candidate launch
 * shape (mode 0 with 64/128/256 threads; mode 1 with 64/128/256 threads x 4/8/16 waves per CU; mode 2 at
 * 4 and 8 waves per CU; mode 3 with budgets 4 and 16), `reps` runs each from the same saved state, keeps the
 * fastest and restores the batch.  ms_out[q] (q < n_out) gets the average milliseconds of candidate q.
 * No reference counterpart. */
int corefl_chem_autotune(corefl_chem *chem, double dt, double rtol, double atol, int max_attempts, int reps,
                         double *ms_out, int n_out);
int corefl_chem_synchronize(corefl_chem *chem);

/* Allocation-free variant of corefl_chem_integrate (no counterpart in the reference; for callers that cannot
 * allocate device memory: code under HIP graph capture and tight serving loops).  One exact ROK4E sub-step of
 * length dt > 0 for n cells held in the CALLER's device arrays: rho (n), e (n, specific internal energy, kept
 * constant), T (n, in: initial guess, out: final) and rhoY (n_spec * n, [k * n + cell], in/out); each cell starts
 * from the full sub-step, as after corefl_chem_set_cells, with the context's launch shape.  The integrator's
 * working arrays live in one fixed device array of the library (36 MiB per GPU, held until the process ends), so
 * n is at most 2^20 (larger n is an error) and the call performs no allocation.  The context must already hold
 * its device tables (created by its first corefl_chem_set_cells or corefl_chem_integrate).  Every call on a GPU
 * shares that array: calls that overlap on one GPU race.  Stream-ordered on `stream` (NULL: the context's). */
int corefl_chem_integrate_noalloc(corefl_chem *chem, long long n, const double *rho, const double *e, double *T,
                                  double *rhoY, double dt, double rtol, double atol, int max_attempts, void *stream);

/* ---------------------------------------------------------------------------------------------------
 * Chemistry surrogate (no reference counterpart: a method change).  An MLP 12 -> 4 x 128 (GELU in its tanh form,
 * or hard swish z clamp(z/6 + 1/2, 0, 1)) -> 10 with BF16 weights and FP32 accumulation on the matrix cores maps the features
 * (T, ln p, Y_1..Y_9, ln dt), each scaled as (f - mu) / sig, to the scaled rates (dY_k/dt, dT/dt).  It is
 * trained on the GPU (forward, backward and Adam kernels of this library) on states labelled by the exact
 * integrator, and replaces the exact chemistry sub-step in cells that pass its gate.
 * ------------------------------------------------------------------------------------------------- */

/* Number of parameters (51968 weights + 522 biases).  Flat layout: W_0 (12 x 128, row-major), b_0 (128),
 * W_1, b_1, W_2, b_2, W_3 (128 x 128), b_1..b_3 (128), W_4 (128 x 10), b_4 (10), in that order. */
long long corefl_surrogate_n_params(void);
/* Context with mini-batch buffers for up to batch_capacity rows. */
int corefl_surrogate_create(int batch_capacity, corefl_surrogate **out);
void corefl_surrogate_destroy(corefl_surrogate *s);
/* FP32 master parameters (also resets the Adam moments and step count). */
int corefl_surrogate_set_params(corefl_surrogate *s, const float *params, long long n);
int corefl_surrogate_get_params(corefl_surrogate *s, float *params, long long n);
/* Feature scaling (12 values each) and target scaling (10 values each). */
int corefl_surrogate_set_scaling(corefl_surrogate *s, const double *in_mu, const double *in_sig, const double *out_mu,
                                 const double *out_sig);
/* Feature and target transforms: mass-fraction features Y^y_power (y_power in (0, 1], default 1), and
 * per output a signed-logarithm target s = sign(r) ln(1 + |r|/r0) of the rate r (out_r0[q] > 0; 0 or a
 * NULL array: the scaled rate itself). */
int corefl_surrogate_set_transform(corefl_surrogate *s, double y_power, const double *out_r0);
/* Gate: gate[0..8) = T_min, T_max, p_min, p_max, dt_min, dt_max (K, Pa, s), dT_max and dT_step (K): a cell
 * uses the network only inside the box, when the network's temperature change and the one recovered from
 * energy conservation differ by at most dT_max, and when its temperature changes by at most dT_step over
 * the sub-step. */
int corefl_surrogate_set_gate(corefl_surrogate *s, const double *gate);
/* Hidden activation: 0 GELU, tanh form 0.5 z (1 + tanh(sqrt(2/pi) (z + 0.044715 z^3))) (the default), 1 hard swish.
 * It defines the network the parameters belong to: set it before training. */
int corefl_surrogate_set_activation(corefl_surrogate *s, int act);
/* Training set on the device: n rows of scaled features X (n x 12) and scaled targets Y (n x 10). */
int corefl_surrogate_set_data(corefl_surrogate *s, long long n, const float *X, const float *Y);
/* `steps` Adam steps (beta 0.9 / 0.999, eps 1e-8, learning rate lr) on mini-batches of `batch` rows drawn
 * with replacement by a hash of (seed, step); mean-squared-error loss.  loss_out[steps] (may be NULL). */
int corefl_surrogate_train(corefl_surrogate *s, int batch, int steps, float lr, unsigned long long seed,
                           float *loss_out);
/* Loss and its gradient (flat layout) on the n given rows (n <= capacity); no parameter update. */
int corefl_surrogate_grad(corefl_surrogate *s, long long n, const float *X, const float *Y, float *loss, float *grads);
/* Network outputs (n x 10) of n scaled feature rows: fused = 1 runs the fused inference kernel, 0 the
 * training kernels (n <= capacity). */
int corefl_surrogate_forward(corefl_surrogate *s, long long n, const float *X, float *out, int fused);
/* Surrogate chemistry sub-step of every cell of the batch (dt == 0: per-cell lengths): network rates,
 * least-squares projection onto element conservation (which also keeps sum Y = 1) with species that would
 * turn negative held at zero, temperature from the conserved internal energy.  Cells failing the gate are
 * integrated by the exact ROK4E sub-step (as corefl_chem_integrate) in the same stream; their count goes
 * to *n_fallback when it is not NULL (this synchronises).  Accepted cells report steps = 0. */
int corefl_chem_surrogate_step(corefl_chem *chem, corefl_surrogate *s, double dt, double rtol, double atol,
                               int max_attempts, int *n_fallback, void *stream);

#ifdef __cplusplus
}
#endif
#endif
