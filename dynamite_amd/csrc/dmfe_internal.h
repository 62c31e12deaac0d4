// Internal declarations shared by the host library (dmfe_ctx.cpp, dmfe_grid.cpp)
// and the gfx950 kernels (dmfe_kernels.hip).
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include <string>
#include <vector>

#include "../../include/dmfe.h"

namespace dmfe {

int set_error(int code, const std::string& msg);

// Grid generator pieces (dmfe_grid.cpp)
void grid_theta(std::size_t len, double Tmax, double alpha, double delta, std::vector<long double>& theta);
void grid_quadrature(const std::vector<long double>& x, int order, std::vector<long double>& w);
void grid_poly_stencils(const std::vector<long double>& x, const std::vector<long double>& q, int n,
                        int64_t* starts, double* alpha);
void grid_rational_stencils(const std::vector<long double>& x, const std::vector<long double>& q, int d, int m,
                            int64_t* starts, double* alpha);
void grid_bspline_weights(const std::vector<long double>& x, const std::vector<long double>& q, int p, double* W);
// the same weights formed on the current HIP device (launch_bspline_combine)
void grid_bspline_weights_device(const std::vector<long double>& x, const std::vector<long double>& q, int p,
                                 double* W);

constexpr int kMaxN = 2048;     // LDS sizing of the row kernels
constexpr int kMaxStages = 10;  // SSPRK(10,4)
constexpr int kSums = 8;        // quadratures per RHS row
constexpr int kMaxMethod = 11;  // integrators 1 RK54, 2 SSPRK104, 3.. SERK2 levels (the rmax table)
// stages of one step at the highest SERK2 level (q = 2 (kMaxMethod - 2), 2q stages)
constexpr int kMaxStepStages = 4 * (kMaxMethod - 2);
// RHS rows with N <= kTailMaxN keep the last two history rows (V[T1-2], D[T1-1]) in LDS
// beside V[T1-1], so B2 targets in the last time interval read the block's LDS copy.
constexpr int kTailMaxN = 1024;

// Columns of RHS block (row j, chunk c of `split`): segments of kColSeg columns dealt to the
// chunks round-robin (segment s to chunk s mod split), so that the B2 entries of a row
// whose time positions lie inside the history (a contiguous column range, the costly
// gathers) are shared by all of its blocks.  Entry e of a chunk is column
// (e / kColSeg * split + c) * kColSeg + e % kColSeg; a segment is one wavefront's worth of
// contiguous columns, so lane-per-entry reads stay coalesced.
constexpr int kColSeg = 64;
__host__ __device__ inline int rhs_chunk_col(int e, int chunk, int split) {
    return ((e / kColSeg) * split + chunk) * kColSeg + e % kColSeg;
}
// entries of chunk c (columns < N)
__host__ __device__ inline int rhs_chunk_count(int N, int chunk, int split) {
    const int nseg = (N + kColSeg - 1) / kColSeg;
    if (chunk >= nseg) return 0;
    const int segs = (nseg - chunk + split - 1) / split;
    const int last = (chunk + (segs - 1) * split == nseg - 1) ? nseg * kColSeg - N : 0;  // short final segment
    return segs * kColSeg - last;
}
// the largest chunk (LDS sizing)
__host__ __device__ inline int rhs_chunk_cap(int N, int split) {
    const int nseg = (N + kColSeg - 1) / kColSeg;
    return ((nseg + split - 1) / split) * kColSeg;
}

// dynamic LDS of one RHS row block: [N] V[T1-1], ([N] V[T1-2], [N] D[T1-1] if N <= kTailMaxN),
// [cw] B2 values, [cw] Hermite abscissae, [cw] rho of the entries' next rows, [cw] time
// rows, [cw] stencil starts, 8 bytes of alignment, and for K tangent directions [K N]
// tangent last rows + [K cw] tangent B2 values
inline size_t rhs_lds_bytes(int N, int cw, int K) {
    const int rows = N <= kTailMaxN ? 3 : 1;
    return (size_t)(rows * N + cw) * 16 + (size_t)cw * (2 * 8 + 2 * 4) + 8 + (size_t)K * (N + cw) * 16;
}

// f(q) = lambda q^p + (1-lambda) q^p2 and its derivatives (math_ops.cpp:23-37)
struct Model {
    double lambda, T0, Gamma;
    int p, p2;
};

struct Tables {
    int N, m;
    long long P;
    const double* theta;
    const double* integ;
    const double* phi2;
    const int* inds[3];    // stencil starts A1, A2, B2 [P]
    const double* w[3];    // weights: A1, A2 transposed to [m][P]; B2 entry-major [P][m]
};

struct History {
    double2* V;   // (QK, QR)   [cap][N]
    double2* D;   // (dQK, dQR) [cap][N]
    double* t1;   // t1grid
    double* rho;  // delta_t_ratio
    double* r;    // rvec
    double* dr;   // drvec
};

// State of the QK tail fit (tail_fit.cpp:46-265) that the reference keeps in function-local
// statics: the chosen window, the history length at the last call, the last rollback seen,
// the smoothed fit parameter; plus the number of blends applied.
struct TailState {
    long long prev_off;        // -1: no window chosen yet
    long long prev_t1len;
    long long known_rollback;  // starts at -1000, the initial last_rollback_loop
    long long applied;
    double alpha;
    double pad[3];
};

struct Work {
    double* posB1;    // [2][N]  posB1xOld, buffer fbuf of the interpolation that wrote it
    double* posB2;    // [2][P]  posB2xOld (double-buffered as Floor / diff)
    double* saveB1;   // [3][N] SSPRK stage positions
    double* saveB2;   // [3][P]
    int* Floor;       // [2][N]  B1 time rows (double-buffered: a launch's search writes
    double* diff;     // [2][N]  one buffer while its finish block reads the other)
    double4* rpk;     // [2][N]  rInt terms of the older history rows (rint_terms), buffered as Floor
    double* rInt;     // [N]
    double* partial;  // [N][split][kSums]
    double* qkb1c0;   // [N]  QKB1[j, 0]
    double* hK;       // [kMaxStages][N]
    double* hR;       // [kMaxStages][N]
    double* scal;     // [0] dr (drstep2), [1] error norm, [2] rstep / energy scratch
    double2* fA[2];   // dense grids: A1, A2 fields (QK, QR) of the current RHS [P], or null
    double2* sE;      // [N] SERK2: the stage two steps back (g_e), or null
    double2* sF;      // [N] SERK2: the embedded sum of b_k g_k, or null
    // tail fit on (else null): the last row as the step left it, before the blend (the next
    // step's stage-0 RHS reads its fields from it, as the reference's FSAL stage uses the
    // fields interpolated before the blend), and the fit's state
    double2* tf_pre;  // [N]
    TailState* tf;
};

enum SearchMode { kSearchB = 0, kSearchI = 1, kSearchSame = 2, kSearchCopy = 3 };

struct SearchArgs {
    int mode;
    int T1;
    const double* t1;     // t1grid in HBM; entry T1-1 is taken from tl
    double tl;            // t1grid.back()
    const double* srcB1;  // kSearchCopy source
    const double* srcB2;
    double* saveB1;       // optional second destination
    double* saveB2;
    int fbuf;             // Floor/diff buffer written
    int gpu_clamp;        // B1 time row clamped per element to [1, ceil(p) - 1] (the reference GPU path)
};

struct RhsArgs {
    int T1;
    double tl;        // t1grid.back() used by the RHS
    int split;        // column chunks per row
    int log_response;
    double* fields;   // non-null: materialise the 15 fields (debug)
    int fbuf;         // Floor/diff buffer of the current interpolation
    int G;            // lanes per B2 entry (4, 8, 16 or 64; rhs_shape_valid)
    // non-null: read the last history row (V[T1-1]) from here (the tail fit's pre-blend
    // row at stage 0); only with N <= kTailMaxN, where last-interval B2 taps come from LDS
    const double2* vlast;
};

// RHS row-kernel shape: G lanes per B2 entry (the reduction tree of the B2 stencil sums
// depends on G, so results are reproducible for a fixed G); taps per lane per gather pass
inline bool rhs_shape_valid(int G, int m) {
    return (G == 4 && m <= 16) || (G == 8 && m <= 16) || (G == 16 && m <= 32) || G == 64;
}
inline int b2_taps_per_pass(int G, int m) {
    const int nt = (m + G - 1) / G;
    return nt < 1 ? 1 : (nt > 4 ? 4 : nt);
}
inline int rhs_default_G(int m) { return m <= 16 ? 4 : (m <= 32 ? 16 : 64); }
// column chunks per row: cw = ceil(N / split) entries per block
inline int rhs_default_split(int N) { return N >= 256 ? (N / 256 > 1 ? N / 256 : 1) : 1; }
inline int rhs_max_split(int N) { return 2 * rhs_default_split(N); }

// RHS shape of a grid class (N, m, log_response): the tuned entry of the process-wide
// table (dmfe_autotune / dmfe_tuning_set) or the default; contexts and batches read it
// when they are created.
struct RhsShape {
    int G, split;
};
RhsShape rhs_shape_lookup(int N, int m, int log_response);
void rhs_shape_store(int N, int m, int log_response, RhsShape s);

struct FinishArgs {
    int mode;         // 0: RHS only (into hK[stage]); 1: full RK stage
    int stage;
    int nterms;       // h_0..h_{nterms-1} combine into g_{stage+1}
    int split;
    int lr;           // current last history row (RHS state)
    int g0row;        // step start row
    int wrow;         // row written (append: lr + 1, replace: lr)
    double t_rhs;     // t of the RHS
    double t_new;     // gt[stage + 1]
    double tdiff;
    double rho_new;
    double coef[kMaxStages];   // dt * a_{stage+1, j} (or dt * b_j at the last stage)
    int last;                  // last stage: also the error norm
    double coefE[kMaxStages];  // dt * b2_j
    double gt_err;             // |gt_S - gte|
    int fbuf;                  // Floor/diff buffer of the interpolation this RHS used
    // SERK2 stages (serk 1: g += sa h; 2: g = 2 g - g_e + sa h): the base row is g0row (the
    // start row at stage 0, else the previous stage's row), sb0 / sb weight the start row /
    // this stage in the embedded sum (W.sF); stage is 0 or 1 (hK slot)
    int serk;
    double sa, sb0, sb;
};

// Forward-mode sensitivities: K <= kMaxTan directions, direction k perturbs (Gamma, lambda)
// by (dG[k], dL[k]).  Each direction carries a tangent history (V, D, r, dr; t1grid and
// delta_t_ratio are the primal's, the time grid being frozen) and its own work arrays.
constexpr int kMaxTan = 2;
struct TanArgs {
    int K;
    double dG[kMaxTan], dL[kMaxTan];
    double2* V[kMaxTan];   // [cap][N] tangent (QK, QR)
    double2* D[kMaxTan];   // [cap][N] tangent (dQK, dQR)
    double* r[kMaxTan];    // [cap]
    double* dr[kMaxTan];   // [cap]
    double* partial;       // [K][N][split][kSums] tangent quadratures
    double* qkb1c0;        // [K][N]
    double* hK;            // [K][kMaxStages][N]
    double* hR;            // [K][kMaxStages][N]
    double* scal;          // [K][8]: 0 tangent dr of the step, 2 tangent energy
};
void launch_tan_init_n(const Model& M, const TanArgs& X, int N, hipStream_t s);
void launch_tan_rhs(const Tables& T, const History& H, const Work& W, const Model& M, const RhsArgs& a,
                    const TanArgs& X, hipStream_t s);
void launch_tan_finish(const Tables& T, const History& H, const Work& W, const Model& M, const FinishArgs& f,
                       const TanArgs& X, hipStream_t s);
void launch_tan_energy(const Tables& T, const History& H, const Model& M, const TanArgs& X, int row, double t,
                       hipStream_t s);

// finish of the previous stage (fa) and/or the next interpolation's searches (sa) in one launch
void launch_stage(const Tables& T, const History& H, const Work& W, const Model& M, const SearchArgs* sa,
                  const FinishArgs* fa, hipStream_t s);
void launch_rhs_rows(const Tables& T, const History& H, const Work& W, const Model& M, const RhsArgs& a,
                     hipStream_t s);
// Y[P][ldy] (columns 0..B-1) = W[P][K] . X[K][B] in FP64 on MFMA; B <= 16
void launch_dense_apply(const double* W, const double* X, double* Y, long long P, int K, int B, int ldy,
                        hipStream_t s);
// rstep (what = 0) or energy at time t (what = 1) on history row `row`, into W.scal[2]
void launch_row_scalar(const Tables& T, const History& H, const Work& W, const Model& M, int row, double t,
                       int what, hipStream_t s);
// sparsify: per-row values val[i] for i in [2, T1-2]
void launch_sparsify_values(const Tables& T, const History& H, int T1, double* val, hipStream_t s);
// gather kept rows: dst row r <- src row keep[r]
void launch_gather_rows(const Tables& T, const History& src, const History& dst, const int* keep, int nkeep,
                        hipStream_t s);

void launch_hbm_copy(const double2* src, double2* dst, long long n, hipStream_t s);
// W[k][i] = sum_{j <= p} B[k][j] Z[start[k] + j][i], k < Q, i < N (dense B-spline weights)
void launch_bspline_combine(const double* Z, const int* start, const double* B, int p, int N, long long Q, double* W,
                            hipStream_t s);

// Device scratch of the no-allocation RHS entry point (dmfe_rhs_noalloc): kScratchDoubles
// doubles (1 MiB) in a file-scope device array, one per device.  scratch_address() gives
// its address on the current device (resolved once per device, then cached).
constexpr size_t kScratchDoubles = 131072;
int scratch_address(double** out);
// doubles of scratch dmfe_rhs_noalloc needs: partials [N][split][8], QK_B1[j,0], rInt, hK, hR [N]
inline size_t rhs_scratch_doubles(int N, int split) { return (size_t)N * split * kSums + 4 * (size_t)N; }

// ---- batched quenches with device-side step control ----------------------------------
enum CtlStatus { kRunning = 0, kDone = 1, kNeedMaint = 2, kNeedRows = 3, kNeedRollback = 4 };

// Per-instance control state (the SimulationConfig / RKData scalars of one quench),
// owned by the device while a batch runs.
struct Ctl {
    // persistent
    int T1;             // history rows
    int method;         // DMFE_RK54 / DMFE_SSPRK104
    int status;         // CtlStatus
    int fcur;           // Floor/diff buffer of the latest interpolation
    long long loop;
    double dt, delta, delta_old;
    double t_prop;      // proposed t of the last iteration (simulation_runner.cpp:130)
    int rollback;       // the GPU main loop's rollback (simulation_runner.cpp:280-303) is on
    int pad_rb;
    long long last_rollback = -1000;  // last_rollback_loop (simulation_runner.cpp:53), -1000 until one
    // limits
    double tmax;
    long long max_loop;
    long long maint_every;
    int cap;            // history capacity (rows)
    int active;         // takes part in the current step
    double dt_min, delta_max, specRad;
    double rmax[11];
    // per-step scratch, set by the begin kernel
    int T1s, S, fc0, pad;
    double gt[kMaxStepStages + 1];  // gt[n]: time of the state the RHS of stage n reads; gt[S]: the step's end
    double tprev;       // t1grid[T1s - 2]
    double gt_err;      // |gt_S - gte|
    double rho_step;    // rho of the row the step writes (set with the last stage's arguments)
    // launch arguments of this step, written by the begin kernel (RK54 and SSPRK104; the
    // stage launches of a SERK2 step form theirs from gt[] and serk_b)
    FinishArgs fin[kMaxStages];      // finish of stage m
    SearchArgs srch[kMaxStages + 1]; // searches of stage n (mode < 0: none)
    // SERK2 (method >= 3): the embedded weights b_0..b_S of the level, set by the host
    // whenever it hands the control word to the device
    double serk_b[kMaxStepStages + 1];
};

struct InstDesc {
    Tables T;
    History H;
    Work W;
    Model M;
    Ctl* ctl;
    int split;
    int gpu_variants;  // the reference GPU path's interpolation (dmfe_set_gpu_variants)
    TanArgs X;  // forward-mode tangents carried by the instance (X.K = 0: none)
};

// Butcher tableaux of the two integrators and f^(d)(q) on the host (dmfe_ctx.cpp)
struct Tableau {
    int S;
    double A[kMaxStages][kMaxStages];
    double b[kMaxStages];
    double b2[kMaxStages];
};
const Tableau& dp54();
const Tableau& ssprk104();
double host_fl(const dmfe_params& p, double q, int d);
// uploads the constant-memory tableaux of the current device once (DMFE_* status)
int batch_upload_tableaux(const Tableau& rk54, const Tableau& ssprk);

// The main loop of one quench on the device-side step control (the batch machinery with
// one instance), borrowing the caller's tables, work buffers and stream.  The history
// arena is the caller's too, but it may be replaced (growth, sparsification): H and cap
// come back updated, the old arena freed.  Control scalars are in/out.
struct DeviceLoop {
    hipStream_t stream;
    Tables T;
    Work W;
    Model M;
    dmfe_params prm;
    int split, G;
    TanArgs X;      // X.K tangent directions carried along (their arenas must hold the run)
    int use_graph;  // replay captured 16-step graphs (else the same launches directly)
    History H;
    int64_t cap;
    int64_t T1, loop;
    int method, fcur;
    double dt, delta, delta_old;
    double t_prop;  // in: the loop variable t (t1grid.back() at a fresh call); out: last proposed t
    int gpu_variants;  // the reference GPU path's interpolation variants
    int rollback;      // rollbacks on error spikes (simulation_runner.cpp:280-303)
    int64_t last_rollback;  // in/out: last_rollback_loop
    int use_serk2;     // the maintenance promotes SSPRK104 to SERK2 levels (dmfe_set_serk2)
};
// SERK2 embedded weights b_0..b_2q of level q (runge_kutta.cu:228-270), in long double
std::vector<double> serk2_coeffs(int q);
// *slot caches the loop's launch maps and captured graphs between calls (NULL at first;
// released with device_loop_close).
int device_loop_run(dmfe_batch** slot, DeviceLoop& q, double tmax, int64_t max_loop, double* corr, int64_t corr_cap,
                    int64_t* steps);
void device_loop_close(dmfe_batch* b);
void launch_batch_begin(InstDesc* D, int n, hipStream_t s);
void launch_batch_end(InstDesc* D, int n, hipStream_t s);
void launch_batch_stage(InstDesc* D, const int2* map, int nblocks, int n, int ept, hipStream_t s);
void launch_batch_rhs(InstDesc* D, const int4* map, int nblocks, int n, int log_response, int G, int maxm, size_t lds,
                      hipStream_t s);
void launch_batch_obs(InstDesc* D, int n, double* obs, int obs_cap, hipStream_t s);
// the stage and RHS launches of a batch whose instances carry K tangent directions
void launch_batch_stage_tan(InstDesc* D, const int2* map, int nblocks, int n, int ept, int K, hipStream_t s);
void launch_batch_rhs_tan(InstDesc* D, const int4* map, int nblocks, int n, int G, int K, size_t lds, hipStream_t s);
// the step launches of one context's device loop, the instance passed by value
void launch_loop_begin(const InstDesc& d, hipStream_t s);
void launch_loop_end(const InstDesc& d, hipStream_t s);
// end of a step [+ its observables into obs] + begin of the next
void launch_loop_end_begin(const InstDesc& d, double* obs, int obs_cap, hipStream_t s);
void launch_loop_stage(const InstDesc& d, int n, hipStream_t s);  // with tangents: also their finish
void launch_loop_rhs(const InstDesc& d, int n, int log_response, int G, size_t lds, hipStream_t s);
// tail fit of one context's device loop after the step's last stage launch
void launch_loop_tail_fit(const InstDesc& d, double delta_max, hipStream_t s);
// tail fit of history row `row` (t1len rows) after a host-driven step
void launch_tail_fit(const Tables& T, const History& H, const Work& W, int row, long long t1len, double delta_max,
                     long long rollback_loop, hipStream_t s);

}  // namespace dmfe
