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
  COREFL_JAC = 4,
  COREFL_METRIC = 5,
  COREFL_FV = 6,
  COREFL_GV = 7,
  COREFL_HV = 8
};

/* Replaces Parameter(&argc,&argv) + Mesh(parameter) + Driver<mix>(parameter, mesh)
 * (src/Parameter.cpp:9-22, src/Mesh.cpp:619-682, src/Driver.cu:13-124): reads
 * <input_dir>/setup.txt, <input_dir>/grid/grid   0.dat, <input_dir>/boundary_condition/boundary   0.txt,
 * computes metrics, allocates device memory and sets the initial condition
 * (src/Initialize.cu:51-95) followed by initialize_computation (src/Driver.cu:126-211). */
int corefl_create(const char *input_dir, corefl_ctx **out);

/* Same as corefl_create with the three inputs in memory: setup_text is the content of setup.txt;
 * dims = n_block*(mx,my,mz); xyz[3*b+0..2] = x, y, z of block b (i fastest, mx*my*mz each);
 * boundary_text is the content of boundary%4d.txt. */
int corefl_create_mem(const char *setup_text, int n_block, const int *dims, const double *const *xyz,
                      const char *boundary_text, corefl_ctx **out);

void corefl_destroy(corefl_ctx *ctx);
const char *corefl_last_error(void);

int corefl_n_blocks(const corefl_ctx *ctx);
/* Replaces direct use of Field::h_ptr / DZone (src/Field.h:164-165). */
int corefl_block_view_get(corefl_ctx *ctx, int block, corefl_block_view *view);
/* Device-to-host copy of one array (see the COREFL_* selectors); n = element count of `host`. */
int corefl_get_array(corefl_ctx *ctx, int block, int which, double *host, long long n);
/* Host-to-device primitive state bv (6*sz, ghosts included); recomputes cv on all nodes and
 * applies the block's boundary conditions (src/FieldOperation.cuh:39-64, src/BoundCond.cu:3489-3619). */
int corefl_set_state(corefl_ctx *ctx, int block, const double *bv_host, long long n);

/* One right-hand-side evaluation dq on every block: viscous flux (src/ViscousScheme.cu:7-73) and
 * inviscid flux (src/InviscidScheme.cu:12-26) of the current state; dq is left in device memory. */
int corefl_rhs(corefl_ctx *ctx, void *stream);

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

/* Simulated physical time (device accumulator) and a stream synchronisation. */
int corefl_get_time(corefl_ctx *ctx, double *t);
int corefl_synchronize(corefl_ctx *ctx);

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

#ifdef __cplusplus
}
#endif
#endif
