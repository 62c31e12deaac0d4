/* dmfe.h - C ABI of the MI355X (gfx950) implementation of the per-RK-step
 * renormalized DMFT update for the spherical mixed p-spin model.
 *
 * The reference has no plugin/FFI layer: its hot path is a set of C++ free
 * functions on global state (SimulationConfig config, SimulationData* sim,
 * RKData* rk; main.cpp:32-37).  Each entry point below names the reference
 * function or call site it replaces.  Conventions:
 *   - plain pointers and sizes only; every function returns a DMFE_* status
 *     (0 on success) and dmfe_last_error() describes the last failure;
 *   - host buffers are caller-owned and only borrowed during the call;
 *   - a dmfe_ctx owns its device memory and one HIP stream; a context is not
 *     thread-safe, distinct contexts are independent;
 *   - all arithmetic is FP64; stencil/time searches are bit-exact with the
 *     reference CPU path (search_utils.cpp:13-131).
 * Library: dynamite_amd/libdmfe.so (built by dynamite_amd/build.py).
 */
#ifndef DMFE_H
#define DMFE_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* status codes */
#define DMFE_OK 0
#define DMFE_ERR_ARG (-1)     /* invalid argument (reference: std::invalid_argument) */
#define DMFE_ERR_HIP (-2)     /* HIP runtime error or no device */
#define DMFE_ERR_STATE (-3)   /* call not valid in the current state */
#define DMFE_ERR_IO (-4)      /* file read/write failure */
#define DMFE_ERR_NOMEM (-5)   /* device allocation failed */

/* interpolation methods of a grid package (grid_cli.cpp:108-161) */
#define DMFE_INTERP_POLY 0
#define DMFE_INTERP_RATIONAL 1
#define DMFE_INTERP_BSPLINE 2

/* integrators (rk_data.hpp:12, runge_kutta.cpp:254-261): rk->init values */
#define DMFE_RK54 1
#define DMFE_SSPRK104 2

/* interpolation position modes (interpolation_core.cpp:22-29) */
#define DMFE_POS_BSEARCH 0  /* interpolate() with no positions: binary search of theta*t, phi2*t */
#define DMFE_POS_ISEARCH 1  /* interpolate(posOld, posOld, false): seeded interpolation search */
#define DMFE_POS_SAME 2     /* interpolate(posOld, posOld, true): reuse positions */

/* Last error message of the calling thread ("" if none). */
const char* dmfe_last_error(void);

/* Library version string. */
const char* dmfe_version(void);

/* ---------------------------------------------------------------- grid packages */

/* Generation parameters: the `grid` subcommand options (grid_cli.cpp:18-74). */
typedef struct dmfe_grid_spec {
    int64_t N;          /* -L, grid length */
    double Tmax;        /* -M */
    double alpha;       /* -a */
    double delta;       /* -D */
    int32_t spline_order; /* -s, int.dat quadrature degree */
    int32_t method;     /* -m, DMFE_INTERP_* */
    int32_t order;      /* -o, interpolation order n */
    int32_t fh_stencil; /* -f, Floater-Hormann window (rational), <=0: n+1 */
} dmfe_grid_spec;

/* Entries per target of the stencil tables: poly n+1, rational the FH window,
 * bspline N (index_vec.cpp:19 infers it as |weights|/|inds|). */
int64_t dmfe_grid_stencil_width(int64_t N, int method, int order, int fh_stencil);

/* Generates a grid package in memory.  Replaces maybe_handle_grid_cli()
 * (grid_cli.cpp:35-233).  Caller allocates: theta, integ [N]; phi1, phi2,
 * posA1, posA2, posB2 [N*N] (phi1 and the pos arrays may be NULL);
 * inds [3*N*N] (A1, A2, B2 blocks; -1 for bspline); weights [3*N*N*m]
 * (entry-major, m = dmfe_grid_stencil_width). */
int dmfe_grid_generate(const dmfe_grid_spec* spec, double* theta, double* integ, double* phi1, double* phi2,
                       double* posA1, double* posA2, double* posB2, int64_t* inds, double* weights);

/* dmfe_grid_generate with the dense B-spline weights (method bspline, the N^4 part of the
 * generator, interpolation_weights.cpp:401-456) formed on the current HIP device: the
 * inverse collocation matrix once on the host (long double), the N^2 queries of each
 * table as combinations of its rows on the device.  Other methods and all other tables
 * are identical to dmfe_grid_generate; the dense weights differ from its per-query solve
 * in the last bits.  DMFE_ERR_HIP without a device. */
int dmfe_grid_generate_device(const dmfe_grid_spec* spec, double* theta, double* integ, double* phi1, double* phi2,
                              double* posA1, double* posA2, double* posB2, int64_t* inds, double* weights);

/* theta grid in extended precision (long double, the generator's internal
 * precision; theta_grid.cpp:71-125). */
int dmfe_grid_theta_ld(int64_t N, double Tmax, double alpha, double delta, long double* theta);

/* Writes Grid_data/<L>-format text files into dir (grid_io.cpp:26-328).
 * Optional arrays may be NULL; params_text (grid_params.txt body) may be NULL. */
int dmfe_grid_write_dir(const char* dir, int64_t N, int64_t m, const double* theta, const double* integ,
                        const double* phi1, const double* phi2, const double* posA1, const double* posA2,
                        const double* posB2, const int64_t* inds, const double* weights, const char* params_text);

/* Sizes of a grid package on disk: N from theta.dat, m = |weightsA1y|/|indsA1y|. */
int dmfe_grid_read_info(const char* dir, int64_t* N, int64_t* m);

/* Reads a grid package (import(), io_input.cpp:90-110).  phi2 may be NULL; if
 * phi2.dat is absent it is rebuilt from theta (phi_grid.cpp:17-34). */
int dmfe_grid_read_dir(const char* dir, int64_t N, int64_t m, double* theta, double* integ, double* phi2,
                       int64_t* inds, double* weights);

/* ------------------------------------------------------------- step context */

/* Grid tables handed to a context (borrowed during dmfe_ctx_create only).
 * Layout = the reference's SimulationData host vectors (simulation_data.hpp:93-94). */
typedef struct dmfe_grid {
    int64_t N;                 /* config.len */
    int64_t m;                 /* stencil width */
    const double* theta;       /* h_theta [N] */
    const double* integ;       /* h_integ [N] (int.dat) */
    const double* phi2;        /* h_phi2 [N*N] */
    const int64_t* inds[3];    /* h_indsA1y, h_indsA2y, h_indsB2y [N*N]; <0 = dense row (bspline) */
    const double* weights[3];  /* h_weightsA1y, h_weightsA2y, h_weightsB2y [N*N*m] entry-major */
} dmfe_grid;

/* Model and control parameters (SimulationConfig, config.hpp:18-56). */
typedef struct dmfe_params {
    int32_t p;            /* -p */
    int32_t p2;           /* -q */
    double lambda;        /* -l */
    double T0;            /* -T, initial temperature (inf -> 1e50) */
    double Gamma;         /* -G, bath temperature */
    double delta_t_min;   /* -d */
    double delta_max;     /* -e */
    double rmax[11];      /* stability radii per method (config.hpp:33) */
    int32_t log_response; /* -R, interpolate QR in log space */
    int32_t maint_every;  /* main-loop cadence of sparsify + re-interpolation
                             (simulation_runner.cpp:161, fixed at 100000 there);
                             <= 0 means 100000 */
} dmfe_params;

/* Fills *p with the reference defaults (config.hpp:18-56). */
void dmfe_params_default(dmfe_params* p);

typedef struct dmfe_ctx dmfe_ctx;

/* Creates a context on the current HIP device: uploads the grid tables and
 * reserves a history arena of `rows_hint` rows (grown on demand).
 * Replaces init()'s table upload (initialization.cpp:177-186, copyVectorsToGPU). */
int dmfe_ctx_create(const dmfe_grid* grid, const dmfe_params* params, int64_t rows_hint, dmfe_ctx** out);
int dmfe_ctx_destroy(dmfe_ctx* ctx);

/* Quench initial condition and first interpolation: the "New simulation" branch
 * of init() (initialization.cpp:204-223) + RK choice and interpolate() (:277-283). */
int dmfe_init_quench(dmfe_ctx* ctx);

/* History access, byte-compatible with the data.bin v1 arrays
 * (io_save_binary.cpp:66-100): t1grid[T1], QK/QR/dQK/dQR [T1*N] row-major, rvec,
 * drvec [T1].  dtratio may be NULL on set: then it is rebuilt as in
 * loadSimulationStateBinary (io_input.cpp:169-172). */
int dmfe_get_T1(const dmfe_ctx* ctx, int64_t* T1);
int dmfe_set_state(dmfe_ctx* ctx, int64_t T1, const double* t1grid, const double* QK, const double* QR,
                   const double* dQK, const double* dQR, const double* rvec, const double* drvec,
                   const double* dtratio);
int dmfe_get_state(dmfe_ctx* ctx, double* t1grid, double* QK, double* QR, double* dQK, double* dQR,
                   double* rvec, double* drvec, double* dtratio);

/* Checkpoints in the reference's data.bin version-1 layout (io_save_binary.cpp:66-100):
 * dmfe_save_state writes the header (parameters, delta, delta_t, loop, energy at
 * t1grid.back()) and the history arrays; dmfe_load_state reads such a file into the
 * context (loadSimulationStateBinary, io_input.cpp:112-175, with its parameter check)
 * and sets the resumed control as init() does (initialization.cpp:225-228,277-283).
 * *energy_out (may be NULL) = the energy stored in the file. */
int dmfe_save_state(dmfe_ctx* ctx, const char* path);
int dmfe_load_state(dmfe_ctx* ctx, const char* path, double* energy_out);

/* Control scalars (config.delta_t, delta, delta_old, loop, rk->init: 1 RK54, 2 SSPRK104,
 * 3..11 SERK2 with q = 2 (method - 2)); delta_t must be finite and > 0, delta and
 * delta_old finite, loop >= 0. */
int dmfe_set_control(dmfe_ctx* ctx, double delta_t, double delta, double delta_old, int64_t loop, int method);
int dmfe_get_control(const dmfe_ctx* ctx, double* delta_t, double* delta, double* delta_old, int64_t* loop,
                     int* method);

/* interpolate() / interpolate(posOld, posOld, same) (interpolation_core.cpp:19-114):
 * computes the time positions of theta*t and phi2*t (posB1xOld, posB2xOld) and
 * the B1 prologue (Floor, diff, rInt).  The N*N interpolated fields are not
 * materialised: they are rebuilt inside the RHS kernel. */
int dmfe_interpolate(dmfe_ctx* ctx, int mode);

/* Positions left by the last interpolation: posB1 [N], posB2 [N*N]. */
int dmfe_get_positions(dmfe_ctx* ctx, double* posB1, double* posB2);
int dmfe_set_positions(dmfe_ctx* ctx, const double* posB1, const double* posB2);

/* Materialises the interpolated fields of the last interpolation (test/debug):
 * out [15*N*N + N] = QKA1 QRA1 QKA2 QRA2 QKB1 QRB1 QKB2 QRB2 SigmaKA1 SigmaRA1
 * SigmaRA2 SigmaKB1 SigmaRB1 SigmaKB2 SigmaRB2, then rInt (the sim->h_* fields
 * written by interpolate(), interpolation_core.cpp:35-113). */
int dmfe_get_fields(dmfe_ctx* ctx, double* out);

/* QKstep() and QRstep() (time_steps.cpp:27-47, 94-100) on the current
 * interpolation state: hK, hR [N]. */
int dmfe_rhs(dmfe_ctx* ctx, double* hK, double* hR);

/* No counterpart in the reference.  For callers that cannot allocate device memory
 * (code under HIP graph capture, tight serving loops): dmfe_rhs with its scratch (the
 * quadrature partials, QK_B1[j,0], rInt and the hK/hR staging) in a file-scope device
 * array of the library (1 MiB per device, held until the process ends) instead of the
 * context's buffers; it calls no hipMalloc/hipFree.  Limit: 8*N*split + 4*N <= 131072
 * doubles, else DMFE_ERR_ARG.  All calls on one device share that device's array:
 * calls that overlap on one device race.  Results equal dmfe_rhs's. */
int dmfe_rhs_noalloc(dmfe_ctx* ctx, double* hK, double* hR);

/* rstep() (time_steps.cpp:83-92) on the last history row. */
int dmfe_rstep(dmfe_ctx* ctx, double* r);

/* update() (runge_kutta.cpp:254-261): one RK54 or SSPRK104 step; *err = the
 * embedded error norm (runge_kutta.cpp:241-251). */
int dmfe_rk_step(dmfe_ctx* ctx, double* err);

/* One iteration of the CPU main loop (simulation_runner.cpp:125-304 without
 * I/O): proposed t, update(), loop++, every 1e5 loops sparsify + interpolate +
 * method switch, then the adaptive dt rule.  *t_out = proposed time written to
 * correlation.txt (:130, :331). */
int dmfe_loop_iteration(dmfe_ctx* ctx, double* t_out);

/* Runs loop iterations while t < tmax, loop < max_loop and delta_t >= delta_t_min
 * (simulation_runner.cpp:125).  If corr is non-NULL it receives up to corr_cap
 * rows (t, E, QK(t,0)) - the correlation.txt columns (:327-341).  *steps = number
 * of iterations run.  The loop runs with its step control on the device (stage
 * arguments, error norm and dt rule evaluated by kernels, 16 steps per host round
 * trip); results are bit-identical to repeated dmfe_loop_iteration calls.  Contexts
 * on dense grids, with per-launch profiling, or with tangents together with SERK2 or
 * rollbacks use the host loop.  Rows of corr
 * are iterations: after a rollback the loop counter repeats values, as correlation.txt
 * does. */
int dmfe_run(dmfe_ctx* ctx, double tmax, int64_t max_loop, double* corr, int64_t corr_cap, int64_t* steps);

/* SERK2, the reference GPU loop's stabilized second-order-Chebyshev integrator
 * (runge_kutta.cu:228-297,397-487, option use_serk2), off by default as on the CPU path:
 * with it on, the every-maint_every maintenance moves SSPRK104 one SERK2 level up each
 * time (method m >= 3 is SERK2 with q = 2 (m - 2): 2q stages of dt / (4q) in a Chebyshev
 * recursion, the error against the embedded combination sum_k b_k g_k; step cap
 * rmax[m - 1] / specRad) and a rollback steps SERK2 one level down.  Methods 3..11 can
 * also be set directly (dmfe_set_control).  Not with sensitivities; dmfe_run runs SERK2
 * steps on its device-driven loop (the stage launches form each stage's arguments). */
int dmfe_set_serk2(dmfe_ctx* ctx, int on);

/* The rollback of the reference's GPU main loop (simulation_runner.cpp:280-303,
 * rollbackState simulation_control.cpp:14-76), off by default as on the CPU path that
 * the results are checked against: when a step ends with delta > 2 delta_max, the last
 * n = min(10, T1 - 2) steps are dropped (history truncated, delta_t = the kept last
 * step, loop -= n, re-interpolation), then delta_t = 0.5 min(max(delta_t,
 * delta_t_min), rmax[method] / specRad) and the integrator restarts as SSPRK104; the
 * maintenance's dt halving then waits 1000 loops after a rollback.  dmfe_run takes
 * the rollback on its device-driven loop (the host truncates between graph replays), on
 * the host-driven loop when tangents are carried. */
int dmfe_set_rollback(dmfe_ctx* ctx, int on);

/* The QK tail fit of the reference's main loop (--tail-fit, tailFitBlendCPU
 * tail_fit.cpp:46-265, called after update() at simulation_runner.cpp:135-141), off by
 * default: after every dmfe_loop_iteration / dmfe_run step, y = (1 - QK) / (1 - theta)^2
 * on the upper half of the last row is fitted by a constant over the 16-point window of
 * least SSE, the fit parameter is smoothed across steps, and where the residual and SSE
 * gates pass, QK is blended into 1 - alpha (1 - theta)^2 from that window on.  As in the
 * reference, the next step's first RHS evaluation uses the fields interpolated before
 * the blend.  Switching it on resets the fit's state.  N <= 1024; not with sensitivities
 * or dense grids.  dmfe_rk_step (update() alone) does not apply it. */
int dmfe_set_tail_fit(dmfe_ctx* ctx, int on);
/* The interpolation of the reference's GPU path (interpolateGPU, interpolation_core.cu),
 * off by default (the CPU path is the contract): every interpolation by bisection
 * (bsearchPosSortedGPU; the interpolation search seeded by the previous positions, the
 * SSPRK104 position reuse and RK54's `same` are not used there) and the B1 time row
 * clamped per element to [1, ceil(p) - 1] (diffNfloorKernel, :23-42) instead of
 * [1, ceil(max p) - 1].  The clamp changes no value on finite histories (at integer
 * positions both forms reduce to the same history row exactly); bisection and the
 * interpolation search bracket identically but evaluate the position with different
 * formulas, so runs differ from the default in the last bits. */
int dmfe_set_gpu_variants(dmfe_ctx* ctx, int on);
/* Blends applied since the tail fit was switched on, and the smoothed fit parameter. */
int dmfe_tail_fit_stats(dmfe_ctx* ctx, int64_t* applied, double* alpha);

/* dmfe_run replays captured hipGraphs of its 16-step launch sequence (on, default) or
 * issues the same launches directly (off).  No counterpart in the reference.  Turn it
 * off for contexts that host threads drive concurrently: a stream capture in one
 * thread makes pageable copies issued meanwhile by other threads fail. */
int dmfe_set_loop_graphs(dmfe_ctx* ctx, int on);

/* Energy at time t and QK(t,0) on the last history row (simulation_runner.cpp:308-337). */
int dmfe_observables(dmfe_ctx* ctx, double t, double* energy, double* qk0);

/* sparsifyNscale(threshold) (sparsify_utils.cpp:14-97, CPU-only build semantics). */
int dmfe_sparsify(dmfe_ctx* ctx, double threshold);

/* Kernel timing with HIP events recorded around every launch on the context's
 * stream (no counterpart in the reference; used by bench.py).  enable != 0
 * starts a fresh measurement.  kind: 0 search_kernel, 1 rhs_rows_kernel,
 * 2 finish_kernel.  *total_ms = summed event time, *launches = launch count. */
int dmfe_profile(dmfe_ctx* ctx, int enable);
int dmfe_profile_read(dmfe_ctx* ctx, int kind, double* total_ms, int64_t* launches);

/* ------------------------------------------------- parameter sensitivities
 * Derivatives of the step path with respect to Gamma (the bath temperature T, -G) and
 * lambda, carried in forward (tangent-linear) mode next to the primal run: with two
 * parameters this gives the same gradient as an adjoint sweep without storing or
 * recomputing checkpoints.  The time grid (dt sequence, positions) is frozen, i.e. the
 * derivative of the discrete map of the run that is being made.  Not available in
 * log-response mode or across a sparsification point.  The reference has no
 * counterpart: these entry points serve callers that need d(E, QK(t,0))/d(Gamma, lambda). */

/* Starts K <= 2 tangent directions right after dmfe_init_quench: direction k perturbs
 * (Gamma, lambda) by (dGamma[k], dLambda[k]).  Subsequent dmfe_rk_step /
 * dmfe_loop_iteration / dmfe_run calls advance the tangents. */
int dmfe_tangent_init(dmfe_ctx* ctx, int K, const double* dGamma, const double* dLambda);
/* Directional derivatives of E(t) (energy at time t on the last row) and QK(t,0). */
int dmfe_tangent_observables(dmfe_ctx* ctx, double t, double* dE, double* dQK0);
/* Tangent history of direction k: dQK, dQR [T1*N], dr [T1] (rvec tangent); any may be NULL. */
int dmfe_tangent_state(dmfe_ctx* ctx, int k, double* dQK, double* dQR, double* dr);
/* Quench + run (as dmfe_run) carrying both directions; grad = [dE/dGamma, dE/dlambda,
 * dQK0/dGamma, dQK0/dlambda] at the final proposed time *t_final. */
int dmfe_gradients(dmfe_ctx* ctx, double tmax, int64_t max_loop, double* grad, double* t_final, int64_t* steps);
/* The adjoint entry point of the drop-in boundary (dmfe_adjoint(ctx, seeds, dT, dlambda)):
 * the vector-Jacobian product seeds . d(E, QK(t_final,0))/d(Gamma, lambda) of a quench run
 * to tmax / max_loop: *dGamma = seeds[0] dE/dGamma + seeds[1] dQK0/dGamma, *dLambda the
 * same in lambda.  With two parameters the product is formed from the forward sweep of
 * dmfe_gradients (two tangent directions through the same run, the dt sequence frozen),
 * which costs less than a reverse sweep with checkpoint recomputation; the result is the
 * same derivative of the discrete run.  t_final, steps may be NULL. */
int dmfe_adjoint(dmfe_ctx* ctx, const double* seeds, double tmax, int64_t max_loop, double* dGamma, double* dLambda,
                 double* t_final, int64_t* steps);
/* dmfe_gradients for n (grid, params) points at once (no counterpart in the reference,
 * which runs one parameter point per process): each point gets its own context, stream
 * and host thread (at most 16 at a time), so their kernels overlap on the device.
 * grads [n][4] as in dmfe_gradients; t_final [n], steps [n] may be NULL.  The results
 * are those of dmfe_gradients on a context of each point. */
int dmfe_gradients_many(int n, const dmfe_grid* grids, const dmfe_params* params, double tmax, int64_t max_loop,
                        int64_t rows_hint, double* grads, double* t_final, int64_t* steps);

/* Dense (B-spline, m = N) stencil table `map` (0 A1, 1 A2, 2 B2) applied to B <= 16
 * right-hand sides as one FP64 MFMA contraction: Y [N*N][B] = W [N*N][N] . X [N][B]
 * (row-major host buffers).  This is indexVecLN3 with m = N (index_vec.cpp:15-61)
 * for B vectors at once.  The contraction runs `reps` times between two HIP events;
 * *ms_out (may be NULL) = average device time per contraction, host copies excluded.
 * In log-response mode only map 2 is available (A1/A2 keep the stencil layout). */
int dmfe_dense_apply(dmfe_ctx* ctx, int map, int B, const double* X, double* Y, int reps, double* ms_out);

/* Measurement utility (no counterpart in the reference): `reps` back-to-back
 * launches of the RHS kernel (rhs_rows_kernel) on the current interpolation state,
 * timed with two HIP events on the context's stream; the state is not changed.
 * *ms_out = average per launch in ms. */
int dmfe_time_rhs(dmfe_ctx* ctx, int reps, double* ms_out);

/* RHS kernel shape (no counterpart in the reference, whose CUDA kernels use fixed
 * launch shapes): G lanes per B2 target entry (4 or 8 for m <= 16, 16 for m <= 32, 64
 * for any m) and `split` column chunks per row (1 .. 2*max(1, N/256)).  The B2 stencil
 * sums are reduced over G lanes and the quadratures over split chunks, so results are
 * reproducible for a fixed shape and differ in the last bits between shapes. */
typedef struct dmfe_rhs_shape {
    int32_t G;
    int32_t split;
} dmfe_rhs_shape;

int dmfe_get_rhs_shape(const dmfe_ctx* ctx, dmfe_rhs_shape* out);
/* Sets the context's shape (before dmfe_tangent_init). */
int dmfe_set_rhs_shape(dmfe_ctx* ctx, const dmfe_rhs_shape* shape);
/* Process-wide shape table per grid class (N, m, log_response): contexts and batches take
 * their shape from it when created (default G = 4/16/64 by m, split = max(1, N/256)).
 * A batch runs one G for all instances: that of its first instance's class. */
int dmfe_tuning_get(int64_t N, int m, int log_response, dmfe_rhs_shape* out);
int dmfe_tuning_set(int64_t N, int m, int log_response, const dmfe_rhs_shape* shape);
/* Autotuner: times `reps` back-to-back RHS launches for every valid (G, split) on the
 * context's current interpolated history (t > 0; best of two sweeps), keeps the fastest
 * (the default shape when it is within 2% of the fastest) on the context and, if
 * `install`, stores it in the table for the context's class.
 * tried/tried_ms (may be NULL) receive up to max_tried candidates and their average
 * launch times in ms; *n_tried = number of candidates. */
int dmfe_autotune(dmfe_ctx* ctx, int reps, int install, dmfe_rhs_shape* best, dmfe_rhs_shape* tried,
                  double* tried_ms, int max_tried, int* n_tried);

/* Measurement utility (no counterpart in the reference): `reps` launches of a
 * plain device-to-device copy of `bytes` bytes (hbm_copy_kernel), used to
 * calibrate HBM byte counters; *ms_out = average launch time in ms. */
int dmfe_copy_probe(int64_t bytes, int reps, double* ms_out);

/* Waits for all work queued on the context's stream. */
int dmfe_synchronize(dmfe_ctx* ctx);

/* ------------------------------------------------------------ batched quenches */
/* A batch holds n independent quenches (each its own grid, model, history, dt and
 * integrator) advanced together: one batch step runs one iteration of the main
 * loop (simulation_runner.cpp:125-278) for every instance that is still running,
 * with the scalar step control on the device, so that a step is a fixed sequence
 * of kernel launches, replayed from a captured hipGraph by dmfe_batch_run.
 * Instances sharing a grid (same table pointers) share its device tables.  The
 * every-1e5-loops maintenance (sparsify + interpolate, :161-233) and history
 * growth run on the host between graph replays.  The reference runs one quench
 * per process; this has no single counterpart there and replaces a loop of
 * runSimulation() calls over parameter points. */
typedef struct dmfe_batch dmfe_batch;

/* Instance status (dmfe_batch_get_control). */
#define DMFE_BATCH_RUNNING 0
#define DMFE_BATCH_DONE 1

/* grids[n], params[n]; params[i].log_response must be equal for all i.
 * rows_cap: initial history rows per instance (grown on demand). */
int dmfe_batch_create(int n, const dmfe_grid* grids, const dmfe_params* params, int64_t rows_cap,
                      dmfe_batch** out);
int dmfe_batch_destroy(dmfe_batch* b);
int dmfe_batch_size(const dmfe_batch* b, int* n);

/* dmfe_init_quench for every instance. */
int dmfe_batch_init_quench(dmfe_batch* b);

/* Loop condition of every instance: t < tmax, loop < max_loop, dt >= delta_t_min
 * (simulation_runner.cpp:125).  Default: unlimited. */
int dmfe_batch_set_limits(dmfe_batch* b, double tmax, int64_t max_loop);

/* One batch step with direct launches; *running = instances still running. */
int dmfe_batch_step(dmfe_batch* b, int* running);

/* Forward-mode sensitivities of every instance, carried inside the batch's launches:
 * K (1..2) directions (dGamma[k], dLambda[k]), the same for all instances (the batch form
 * of dmfe_tangent_init; no reference counterpart, see dmfe_gradients).  Call right after
 * dmfe_batch_init_quench (which also drops them); linear-response form only.  The tangent
 * histories take the batch's arena size: a run that would grow the arena or reach the
 * every-1e5-loops maintenance stops with DMFE_ERR_STATE. */
int dmfe_batch_tangent_init(dmfe_batch* b, int K, const double* dGamma, const double* dLambda);

/* d(E(t), QK(t,0)) along the K directions for every instance at the proposed t of its
 * last step (as dmfe_gradients; t1grid.back() before any step): dE[i * K + k],
 * dQK0[i * K + k] (the batch form of dmfe_tangent_observables). */
int dmfe_batch_tangent_observables(dmfe_batch* b, double* dE, double* dQK0);

/* Up to max_steps batch steps, stopping early when no instance runs.  use_graph:
 * replay a hipGraph of `chunk` captured steps (host checks between replays);
 * otherwise launch directly.  *steps_out = batch steps run. */
int dmfe_batch_run(dmfe_batch* b, int64_t max_steps, int use_graph, int chunk, int64_t* steps_out);

/* Per-instance control and state, as dmfe_get_control / dmfe_set_control /
 * dmfe_get_T1 / dmfe_get_state. */
int dmfe_batch_get_control(dmfe_batch* b, int i, double* delta_t, double* delta, double* delta_old, int64_t* loop,
                           int* method, int* status);
/* Same checks as dmfe_set_control; methods 1 (RK54) and 2 (SSPRK104). */
int dmfe_batch_set_control(dmfe_batch* b, int i, double delta_t, double delta, double delta_old, int64_t loop,
                           int method);
int dmfe_batch_get_T1(dmfe_batch* b, int i, int64_t* T1);
int dmfe_batch_get_state(dmfe_batch* b, int i, double* t1grid, double* QK, double* QR, double* dQK, double* dQR,
                         double* rvec, double* drvec, double* dtratio);

/* History of instance i in the dmfe_set_state layout (t1grid[T1], QK/QR/dQK/dQR
 * [T1*N], rvec, drvec [T1]; dtratio may be NULL: rebuilt as io_input.cpp:169-172),
 * followed by a fresh binary-search interpolation (interpolate() with no positions,
 * initialization.cpp:277-283); the instance is set running, its control scalars are
 * left to dmfe_batch_set_control.  Replaces, per instance, the state load of a resumed
 * run (loadSimulationStateBinary, io_input.cpp:112-175). */
int dmfe_batch_set_state(dmfe_batch* b, int i, int64_t T1, const double* t1grid, const double* QK, const double* QR,
                         const double* dQK, const double* dQR, const double* rvec, const double* drvec,
                         const double* dtratio);

/* Correlation records (the correlation.txt columns t, E, QK(t,0), plus loop) of every
 * step, kept on the device in a ring of `cap` rows per instance (row (loop-1) % cap);
 * cap <= 0 switches recording off.  dmfe_batch_records copies instance i's ring
 * (cap x 4 doubles). */
int dmfe_batch_record(dmfe_batch* b, int cap);
int dmfe_batch_records(dmfe_batch* b, int i, double* out);

/* Lane-group width G of the batch RHS kernel (one for all instances; see
 * dmfe_rhs_shape; each instance keeps the split of its class in the shape table).
 * dmfe_batch_autotune times `reps` launches of the batch RHS on the current states for
 * every valid G (best of two sweeps, after at least one batch step), keeps the fastest
 * (the default width when within 2% of it) and returns it with the candidates and their
 * average launch times (ms). */
int dmfe_batch_get_lanes(const dmfe_batch* b, int* G);
int dmfe_batch_set_lanes(dmfe_batch* b, int G);
int dmfe_batch_autotune(dmfe_batch* b, int reps, int* G_best, int32_t* tried_G, double* tried_ms, int max_tried,
                        int* n_tried);

#ifdef __cplusplus
}
#endif
#endif /* DMFE_H */
