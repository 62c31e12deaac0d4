// Host side of the step path: context (device tables, history arena, work
// buffers, one HIP stream), the RK stage sequencing of RK54 / SSPRK104
// (runge_kutta.cpp:19-261), the scalar time bookkeeping of appendAll/replaceAll
// (time_steps.cpp:49-81,127-162) and the adaptive main loop
// (simulation_runner.cpp:125-304, CPU-only branches).  All N- and N*N-sized
// arithmetic runs in the gfx950 kernels (dmfe_kernels.hip); the host only
// replays the reference's scalar arithmetic (t1grid, delta_t_ratio, dt) in the
// same FP64 order, so that time positions are bit-identical.
#include "dmfe_internal.h"

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <limits>
#include <map>
#include <atomic>
#include <mutex>
#include <string>
#include <thread>
#include <tuple>
#include <vector>

namespace dmfe {

static thread_local std::string g_err;

int set_error(int code, const std::string& msg) {
    g_err = msg;
    return code;
}

#define HIP_TRY(expr)                                                                                     \
    do {                                                                                                  \
        hipError_t e_ = (expr);                                                                           \
        if (e_ != hipSuccess)                                                                             \
            return set_error(DMFE_ERR_HIP, std::string(#expr) + ": " + hipGetErrorString(e_));            \
    } while (0)

// process-wide RHS shapes per grid class (N, m, log_response)
static std::mutex g_shape_mu;
static std::map<std::tuple<int, int, int>, RhsShape> g_shapes;

RhsShape rhs_shape_lookup(int N, int m, int log_response) {
    std::lock_guard<std::mutex> lk(g_shape_mu);
    auto it = g_shapes.find(std::make_tuple(N, m, log_response ? 1 : 0));
    if (it != g_shapes.end()) return it->second;
    return RhsShape{rhs_default_G(m), rhs_default_split(N)};
}

void rhs_shape_store(int N, int m, int log_response, RhsShape s) {
    std::lock_guard<std::mutex> lk(g_shape_mu);
    g_shapes[std::make_tuple(N, m, log_response ? 1 : 0)] = s;
}

// Dormand-Prince 5(4) and SSPRK(10,4) tableaux (runge_kutta.cpp:22-35,149-159)
const Tableau& dp54() {
    static Tableau t = [] {
        Tableau x{};
        x.S = 7;
        const double A[7][7] = {{0, 0, 0, 0, 0, 0, 0},
                                {1.0 / 5, 0, 0, 0, 0, 0, 0},
                                {3.0 / 40, 9.0 / 40, 0, 0, 0, 0, 0},
                                {44.0 / 45, -56.0 / 15, 32.0 / 9, 0, 0, 0, 0},
                                {19372.0 / 6561, -25360.0 / 2187, 64448.0 / 6561, -212.0 / 729, 0, 0, 0},
                                {9017.0 / 3168, -355.0 / 33, 46732.0 / 5247, 49.0 / 176, -5103.0 / 18656, 0, 0},
                                {35.0 / 384, 0, 500.0 / 1113, 125.0 / 192, -2187.0 / 6784, 11.0 / 84, 0}};
        const double b[7] = {35.0 / 384, 0, 500.0 / 1113, 125.0 / 192, -2187.0 / 6784, 11.0 / 84, 0};
        const double b2[7] = {5179.0 / 57600, 0, 7571.0 / 16695, 393.0 / 640, -92097.0 / 339200, 187.0 / 2100, 1.0 / 40};
        for (int i = 0; i < 7; ++i) {
            for (int j = 0; j < 7; ++j) x.A[i][j] = A[i][j];
            x.b[i] = b[i];
            x.b2[i] = b2[i];
        }
        return x;
    }();
    return t;
}

const Tableau& ssprk104() {
    static Tableau t = [] {
        Tableau x{};
        x.S = 10;
        for (int i = 0; i < 10; ++i) {
            x.b[i] = 1.0 / 10;
            for (int j = 0; j < 10; ++j) x.A[i][j] = (j < i) ? ((i <= 4) ? 1.0 / 6 : (j <= 4 ? 1.0 / 15 : 1.0 / 6)) : 0.0;
        }
        const double b2[10] = {0., 2.0 / 9, 0, 0, 5.0 / 18, 1.0 / 3, 0., 0., 0., 1.0 / 6};
        for (int i = 0; i < 10; ++i) x.b2[i] = b2[i];
        return x;
    }();
    return t;
}

double host_fl(const dmfe_params& p, double q, int d) {
    auto pw = [](double b, int e) {
        if (e == 0) return 1.0;
        if (e == 1) return b;
        if (e == 2) return b * b;
        if (e == 3) return b * b * b;
        if (e == 4) { double s = b * b; return s * s; }
        double r = 1.0, c = b;
        while (e > 0) { if (e & 1) r *= c; c *= c; e >>= 1; }
        return r;
    };
    const double l = p.lambda;
    const int a = p.p, b = p.p2;
    switch (d) {
        case 1: return l * a * pw(q, a - 1) + (1 - l) * b * pw(q, b - 1);
        case 2: return l * a * (a - 1) * pw(q, a - 2) + (1 - l) * b * (b - 1) * pw(q, b - 2);
        default: return l * pw(q, a) + (1 - l) * pw(q, b);
    }
}

}  // namespace dmfe

using namespace dmfe;

enum KernelKind { kKindSearch = 0, kKindRhs = 1, kKindFinish = 2, kKindOther = 3, kKinds = 4 };

struct dmfe_ctx {
    int device = 0;
    hipStream_t stream = nullptr;
    dmfe_params params{};
    Model model{};
    int N = 0, m = 0;
    long long P = 0;
    int split = 2;
    int G = 8;                // lanes per B2 entry of the RHS kernel
    dmfe_batch* dloop = nullptr;  // dmfe_run's device loop (launch maps, captured graphs)
    bool loop_graphs = true;      // dmfe_run replays captured graphs (dmfe_set_loop_graphs)
    bool rollback = false;        // the reference GPU loop's rollback (dmfe_set_rollback)
    bool use_serk2 = false;       // the GPU loop's SERK2 promotion at the maintenance (dmfe_set_serk2)
    int64_t last_rollback = -1000;  // last_rollback_loop (simulation_runner.cpp:53)
    // QK tail fit after every main-loop step (dmfe_set_tail_fit; tail_fit.cpp:46-265).  While
    // it is on, W.tf_pre / W.tf point at tf_pre_buf / tf_state.  tf_fresh: W.tf_pre holds the
    // row the last interpolation saw (a tail fit after the last step wrote it); otherwise the
    // next step first copies the current last row there.
    bool gpu_variants = false;  // the reference GPU path's interpolation (dmfe_set_gpu_variants)
    bool tail_fit = false;
    bool tf_fresh = false;
    double2* tf_pre_buf = nullptr;
    TailState* tf_state = nullptr;
    bool dense_mfma = false;  // dense (B-spline) grid: A1/A2 contracted on MFMA, weights entry-major
    // forward-mode sensitivities (dmfe_tangent_init): tangent histories and work
    int tanK = 0;
    History tanH[kMaxTan] = {};
    double tanG[kMaxTan] = {}, tanL[kMaxTan] = {};
    double* tanWork = nullptr;
    // tables
    double *theta = nullptr, *integ = nullptr, *phi2 = nullptr;
    int* inds[3] = {nullptr, nullptr, nullptr};
    double* w[3] = {nullptr, nullptr, nullptr};
    // history arena
    int64_t cap = 0, T1 = 0;
    History H{};
    std::vector<double> t1, rho;  // host mirrors of t1grid and delta_t_ratio
    // work
    Work W{};
    double* fields = nullptr;
    double* pinned = nullptr;  // host scalars (pinned)
    // control (SimulationConfig / RKData scalars)
    double dt = 0, delta = 0, delta_old = 0, specRad = 0;
    int64_t loop = 0;
    int method = DMFE_RK54;
    bool interp_valid = false;
    double* scratch = nullptr;  // this device's scratch array (dmfe_rhs_noalloc), resolved at creation
    std::string scratch_msg;    // why it is unavailable
    int fcur = 0;  // Floor/diff buffer of the latest interpolation
    // optional per-kernel timing with HIP events on the context stream
    bool profiling = false;
    std::vector<hipEvent_t> ev[kKinds];  // start/stop pairs
    size_t ev_used[kKinds] = {0, 0, 0, 0};

    TanArgs tan_args() const {
        TanArgs X{};
        X.K = tanK;
        for (int k = 0; k < tanK; ++k) {
            X.dG[k] = tanG[k];
            X.dL[k] = tanL[k];
            X.V[k] = tanH[k].V;
            X.D[k] = tanH[k].D;
            X.r[k] = tanH[k].r;
            X.dr[k] = tanH[k].dr;
        }
        // tanWork: partial [K][N][split][8] | qkb1c0 [K][N] | hK, hR [K][10][N] | scal [K][8]
        double* p = tanWork;
        X.partial = p; p += (size_t)tanK * N * split * kSums;
        X.qkb1c0 = p; p += (size_t)tanK * N;
        X.hK = p; p += (size_t)tanK * kMaxStages * N;
        X.hR = p; p += (size_t)tanK * kMaxStages * N;
        X.scal = p;
        return X;
    }
    size_t tan_work_doubles() const {
        return (size_t)tanK * ((size_t)N * split * kSums + N + 2 * (size_t)kMaxStages * N + 8);
    }

    Tables tables() const {
        Tables T{};
        T.N = N; T.m = m; T.P = P;
        T.theta = theta; T.integ = integ; T.phi2 = phi2;
        for (int f = 0; f < 3; ++f) { T.inds[f] = inds[f]; T.w[f] = w[f]; }
        return T;
    }
};

namespace {

void free_history(History& h) {
    hipFree(h.V); hipFree(h.D); hipFree(h.t1); hipFree(h.rho); hipFree(h.r); hipFree(h.dr);
    h = History{};
}

int alloc_history(History& h, int64_t rows, int N) {
    const size_t n = (size_t)rows;
    if (hipMalloc(&h.V, n * N * sizeof(double2)) != hipSuccess || hipMalloc(&h.D, n * N * sizeof(double2)) != hipSuccess ||
        hipMalloc(&h.t1, n * sizeof(double)) != hipSuccess || hipMalloc(&h.rho, n * sizeof(double)) != hipSuccess ||
        hipMalloc(&h.r, n * sizeof(double)) != hipSuccess || hipMalloc(&h.dr, n * sizeof(double)) != hipSuccess) {
        free_history(h);
        return set_error(DMFE_ERR_NOMEM, "history allocation of " + std::to_string(rows) + " rows failed");
    }
    return DMFE_OK;
}

// Grows the history arena to hold at least `rows` rows (copying the live rows).
int ensure_rows(dmfe_ctx* c, int64_t rows) {
    if (rows <= c->cap) return DMFE_OK;
    int64_t nc = std::max<int64_t>(c->cap, 64);
    while (nc < rows) nc *= 2;
    History h{};
    int rc = alloc_history(h, nc, c->N);
    if (rc) return rc;
    if (c->T1 > 0) {
        const size_t n = (size_t)c->T1;
        HIP_TRY(hipMemcpyAsync(h.V, c->H.V, n * c->N * sizeof(double2), hipMemcpyDeviceToDevice, c->stream));
        HIP_TRY(hipMemcpyAsync(h.D, c->H.D, n * c->N * sizeof(double2), hipMemcpyDeviceToDevice, c->stream));
        HIP_TRY(hipMemcpyAsync(h.t1, c->H.t1, n * sizeof(double), hipMemcpyDeviceToDevice, c->stream));
        HIP_TRY(hipMemcpyAsync(h.rho, c->H.rho, n * sizeof(double), hipMemcpyDeviceToDevice, c->stream));
        HIP_TRY(hipMemcpyAsync(h.r, c->H.r, n * sizeof(double), hipMemcpyDeviceToDevice, c->stream));
        HIP_TRY(hipMemcpyAsync(h.dr, c->H.dr, n * sizeof(double), hipMemcpyDeviceToDevice, c->stream));
        HIP_TRY(hipStreamSynchronize(c->stream));
    }
    for (int k = 0; k < c->tanK; ++k) {
        History th{};
        rc = alloc_history(th, nc, c->N);
        if (rc) { free_history(h); return rc; }
        const size_t n = (size_t)c->T1;
        HIP_TRY(hipMemcpyAsync(th.V, c->tanH[k].V, n * c->N * sizeof(double2), hipMemcpyDeviceToDevice, c->stream));
        HIP_TRY(hipMemcpyAsync(th.D, c->tanH[k].D, n * c->N * sizeof(double2), hipMemcpyDeviceToDevice, c->stream));
        HIP_TRY(hipMemcpyAsync(th.r, c->tanH[k].r, n * sizeof(double), hipMemcpyDeviceToDevice, c->stream));
        HIP_TRY(hipMemcpyAsync(th.dr, c->tanH[k].dr, n * sizeof(double), hipMemcpyDeviceToDevice, c->stream));
        HIP_TRY(hipStreamSynchronize(c->stream));
        free_history(c->tanH[k]);
        c->tanH[k] = th;
    }
    free_history(c->H);
    c->H = h;
    c->cap = nc;
    return DMFE_OK;
}

void tangent_free(dmfe_ctx* c) {
    for (int k = 0; k < kMaxTan; ++k) free_history(c->tanH[k]);
    if (c->tanWork) (void)hipFree(c->tanWork);
    c->tanWork = nullptr;
    c->tanK = 0;
}

void prof_begin(dmfe_ctx* c, int kind) {
    if (!c->profiling) return;
    auto& v = c->ev[kind];
    const size_t i = c->ev_used[kind];
    if (i + 2 > v.size()) {
        hipEvent_t a, b;
        hipEventCreate(&a);
        hipEventCreate(&b);
        v.push_back(a);
        v.push_back(b);
    }
    hipEventRecord(v[i], c->stream);
}

void prof_end(dmfe_ctx* c, int kind) {
    if (!c->profiling) return;
    hipEventRecord(c->ev[kind][c->ev_used[kind] + 1], c->stream);
    c->ev_used[kind] += 2;
}

int launch_check(const char* what) {
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return set_error(DMFE_ERR_HIP, std::string(what) + ": " + hipGetErrorString(e));
    return DMFE_OK;
}

SearchArgs make_search(const dmfe_ctx* c, int mode, const double* srcB1, const double* srcB2, double* saveB1,
                       double* saveB2) {
    SearchArgs a{};
    a.mode = mode;
    a.T1 = (int)c->T1;
    a.t1 = c->H.t1;
    a.tl = c->t1[c->T1 - 1];
    a.srcB1 = srcB1; a.srcB2 = srcB2; a.saveB1 = saveB1; a.saveB2 = saveB2;
    a.fbuf = 1 - c->fcur;
    if (c->gpu_variants) {
        // the reference GPU path bisects at every interpolation (interpolateGPU ignores its
        // input positions and `same`, interpolation_core.cu:65-72) and clamps the B1 row
        // per element (:23-42)
        a.mode = kSearchB;
        a.srcB1 = a.srcB2 = nullptr;
        a.gpu_clamp = 1;
    }
    return a;
}

// One stage-boundary launch: finish of the previous stage (f) and/or the searches
// of the next interpolation (sa, interpolation_core.cpp:19-84).
int run_stage(dmfe_ctx* c, const SearchArgs* sa, const FinishArgs* f) {
    const int kind = f ? kKindFinish : kKindSearch;
    FinishArgs fl{};
    if (f) { fl = *f; fl.fbuf = c->fcur; }
    prof_begin(c, kind);
    launch_stage(c->tables(), c->H, c->W, c->model, sa, f ? &fl : nullptr, c->stream);
    prof_end(c, kind);
    if (sa) { c->interp_valid = true; c->fcur = sa->fbuf; }
    return launch_check("stage_kernel");
}

int run_search(dmfe_ctx* c, int mode, const double* srcB1, const double* srcB2, double* saveB1, double* saveB2) {
    const SearchArgs a = make_search(c, mode, srcB1, srcB2, saveB1, saveB2);
    c->tf_fresh = false;  // a fresh interpolation of the current history
    return run_stage(c, &a, nullptr);
}

int run_rhs_rows(dmfe_ctx* c, double* fields, const double2* vlast = nullptr) {
    RhsArgs a{};
    a.vlast = vlast;
    a.T1 = (int)c->T1;
    a.tl = c->t1[c->T1 - 1];
    a.split = c->split;
    a.log_response = c->params.log_response;
    a.fields = fields;
    a.fbuf = c->fcur;
    a.G = c->G;
    prof_begin(c, kKindRhs);
    if (c->dense_mfma) {
        // A1/A2 of a dense grid: (P x N) . (N x 2) on MFMA against the last row (QK, QR)
        const double* last = reinterpret_cast<const double*>(c->H.V + (size_t)(c->T1 - 1) * c->N);
        for (int f = 0; f < 2; ++f)
            launch_dense_apply(c->w[f], last, reinterpret_cast<double*>(c->W.fA[f]), c->P, c->N, 2, 2, c->stream);
    }
    if (c->tanK && !fields)  // the same RHS plus the tangent quadratures
        launch_tan_rhs(c->tables(), c->H, c->W, c->model, a, c->tan_args(), c->stream);
    else
        launch_rhs_rows(c->tables(), c->H, c->W, c->model, a, c->stream);
    prof_end(c, kKindRhs);
    return launch_check("rhs_rows_kernel");
}

int copy_out(dmfe_ctx* c, double* dst, const void* src, size_t bytes) {
    HIP_TRY(hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToHost, c->stream));
    HIP_TRY(hipStreamSynchronize(c->stream));
    return DMFE_OK;
}

// update() = RK54() / SSPRK104() (runge_kutta.cpp:19-261).  Launch sequence per
// step: [search(0) if T1 == 1] rhs(0) | finish(0)+search(1) | rhs(1) | ... |
// rhs(S-1) | finish(S-1)+final search; the host replays the scalar time
// bookkeeping (gt, t1grid, delta_t_ratio) in the reference's FP64 order.
int rk_step_serk2(dmfe_ctx* c, int q, double* err_out);

// With the tail fit on, a step's stage-0 RHS (no interpolation of its own, FSAL) reads the
// last row from W.tf_pre: the row the previous interpolation saw.  When no tail fit wrote
// it after the previous step, that is the current last row.
int tail_fit_prepare(dmfe_ctx* c) {
    if (!c->W.tf_pre) return DMFE_OK;
    if (!c->tf_fresh && c->T1 >= 1)
        HIP_TRY(hipMemcpyAsync(c->W.tf_pre, c->H.V + (size_t)(c->T1 - 1) * c->N, (size_t)c->N * sizeof(double2),
                               hipMemcpyDeviceToDevice, c->stream));
    c->tf_fresh = false;
    return DMFE_OK;
}

int rk_step(dmfe_ctx* c, double* err_out) {
    if (c->method >= 3) return rk_step_serk2(c, 2 * (c->method - 2), err_out);
    const bool ss = (c->method != DMFE_RK54);
    const Tableau& tb = ss ? ssprk104() : dp54();
    const int S = tb.S;
    int rc = ensure_rows(c, c->T1 + 1);
    if (rc) return rc;
    if (!c->interp_valid && c->T1 > 1) {
        rc = run_search(c, kSearchB, nullptr, nullptr, nullptr, nullptr);
        if (rc) return rc;
    }
    rc = tail_fit_prepare(c);
    if (rc) return rc;
    const int64_t g0row = c->T1 - 1;
    const double dt = c->dt;
    double gt[kMaxStages + 1];
    double ht[kMaxStages];
    gt[0] = c->t1[c->T1 - 1];
    FinishArgs pending{};
    bool have_pending = false;
    for (int n = 0; n < S; ++n) {
        SearchArgs sa{};
        bool search = false;
        if (c->T1 == 1 || n != 0) {
            search = true;
            if (n == 0) sa = make_search(c, kSearchB, nullptr, nullptr, nullptr, nullptr);
            else if (ss && n >= 5 && n <= 7)
                sa = make_search(c, kSearchCopy, c->W.saveB1 + (size_t)(n - 5) * c->N,
                                 c->W.saveB2 + (size_t)(n - 5) * c->P, nullptr, nullptr);
            else if (ss && n >= 2 && n <= 4)
                sa = make_search(c, kSearchI, nullptr, nullptr, c->W.saveB1 + (size_t)(n - 2) * c->N,
                                 c->W.saveB2 + (size_t)(n - 2) * c->P);
            else sa = make_search(c, kSearchI, nullptr, nullptr, nullptr, nullptr);
        }
        if (search || have_pending) {
            rc = run_stage(c, search ? &sa : nullptr, have_pending ? &pending : nullptr);
            if (rc) return rc;
            have_pending = false;
        }
        rc = run_rhs_rows(c, nullptr, search ? nullptr : c->W.tf_pre);
        if (rc) return rc;
        ht[n] = 1.0;
        FinishArgs& f = pending;
        f = FinishArgs{};
        f.mode = 1;
        f.stage = n;
        f.split = c->split;
        f.lr = (int)(c->T1 - 1);
        f.g0row = (int)g0row;
        f.t_rhs = c->t1[c->T1 - 1];
        const bool last = (n == S - 1);
        f.last = last ? 1 : 0;
        if (n == 0) {
            f.nterms = 1;
            f.coef[0] = dt * tb.A[1][0];
            gt[1] = gt[0] + dt * tb.A[1][0] * ht[0];
            // appendAll (time_steps.cpp:140-149)
            c->t1.push_back(gt[1]);
            const size_t idx = c->t1.size() - 1;
            const double tdiff = c->t1[idx] - c->t1[idx - 1];
            const double r = (idx > 1) ? tdiff / (c->t1[idx - 1] - c->t1[idx - 2]) : 0.0;
            c->rho.push_back(r);
            f.wrow = (int)idx;
            f.t_new = gt[1];
            f.tdiff = tdiff;
            f.rho_new = r;
            c->T1 += 1;
        } else {
            gt[n + 1] = gt[0];
            const int nt = last ? S : n + 1;
            f.nterms = nt;
            for (int j = 0; j < nt; ++j) {
                const double co = last ? tb.b[j] : tb.A[n + 1][j];
                f.coef[j] = dt * co;
                gt[n + 1] += dt * co * ht[j];
            }
            // replaceAll (time_steps.cpp:57-67)
            const size_t L = c->t1.size();
            c->t1[L - 1] = gt[n + 1];
            const double tdiff = c->t1[L - 1] - c->t1[L - 2];
            const double r = (L > 2) ? tdiff / (c->t1[L - 2] - c->t1[L - 3]) : 0.0;
            c->rho[L - 1] = r;
            f.wrow = (int)(L - 1);
            f.t_new = gt[n + 1];
            f.tdiff = tdiff;
            f.rho_new = r;
        }
        if (last) {
            double gte = gt[0];
            for (int j = 0; j < S; ++j) {
                f.coefE[j] = dt * tb.b2[j];
                gte += dt * tb.b2[j] * ht[j];
            }
            f.gt_err = std::fabs(gt[S] - gte);
        }
        if (c->tanK) {
            // before the primal finish of this stage, which replaces the RHS's last row
            FinishArgs tf = f;
            tf.fbuf = c->fcur;
            launch_tan_finish(c->tables(), c->H, c->W, c->model, tf, c->tan_args(), c->stream);
            rc = launch_check("tan_finish_kernel");
            if (rc) return rc;
        }
        have_pending = true;
    }
    // final interpolation: RK54 reuses the positions, SSPRK104 searches (:121, :229)
    const SearchArgs fin = make_search(c, ss ? kSearchI : kSearchSame, nullptr, nullptr, nullptr, nullptr);
    rc = run_stage(c, &fin, &pending);
    if (rc) return rc;
    HIP_TRY(hipMemcpyAsync(c->pinned, c->W.scal + 1, sizeof(double), hipMemcpyDeviceToHost, c->stream));
    HIP_TRY(hipStreamSynchronize(c->stream));
    if (err_out) *err_out = c->pinned[0];
    return DMFE_OK;
}

int sparsify(dmfe_ctx* c, double threshold) {
    if (c->tanK) return set_error(DMFE_ERR_STATE, "sparsify: not available while sensitivities are carried");
    const int64_t T1 = c->T1;
    std::vector<int> keep;
    keep.reserve((size_t)T1);
    keep.push_back(0);
    bool erased = false;
    if (T1 >= 4) {
        double* dval = nullptr;
        HIP_TRY(hipMalloc(&dval, (size_t)T1 * sizeof(double)));
        launch_sparsify_values(c->tables(), c->H, (int)T1, dval, c->stream);
        std::vector<double> val((size_t)T1, 0.0);
        hipError_t e = hipMemcpyAsync(val.data(), dval, (size_t)T1 * sizeof(double), hipMemcpyDeviceToHost, c->stream);
        if (e == hipSuccess) e = hipStreamSynchronize(c->stream);
        hipFree(dval);
        if (e != hipSuccess) return set_error(DMFE_ERR_HIP, std::string("sparsify: ") + hipGetErrorString(e));
        for (int64_t i = 2; i + 1 < T1; ++i) {
            if (val[(size_t)i] < threshold) erased = true;
            else keep.push_back((int)i);
        }
    }
    keep.push_back((int)(T1 - 1));
    if (!erased) return DMFE_OK;
    const int nk = (int)keep.size();
    History h{};
    int rc = alloc_history(h, c->cap, c->N);
    if (rc) return rc;
    int* dkeep = nullptr;
    HIP_TRY(hipMalloc(&dkeep, (size_t)nk * sizeof(int)));
    HIP_TRY(hipMemcpyAsync(dkeep, keep.data(), (size_t)nk * sizeof(int), hipMemcpyHostToDevice, c->stream));
    launch_gather_rows(c->tables(), c->H, h, dkeep, nk, c->stream);
    HIP_TRY(hipStreamSynchronize(c->stream));
    hipFree(dkeep);
    free_history(c->H);
    c->H = h;
    std::vector<double> nt1(nk), nrho(nk);
    for (int r = 0; r < nk; ++r) {
        nt1[r] = c->t1[(size_t)keep[r]];
        nrho[r] = c->rho[(size_t)keep[r]];
    }
    c->t1.swap(nt1);
    c->rho.swap(nrho);
    c->T1 = nk;
    c->interp_valid = false;
    return DMFE_OK;
}

// one iteration of runSimulation's while-loop body (simulation_runner.cpp:128-278, CPU-only branches)
int loop_iteration(dmfe_ctx* c, double* t_out) {
    const int64_t every_ = c->params.maint_every > 0 ? c->params.maint_every : 100000;
    if (c->tanK && (c->loop + 1) % every_ == 0)
        return set_error(DMFE_ERR_STATE, "sensitivities: the run would cross a sparsification point (loop " +
                                             std::to_string(c->loop + 1) + ")");
    const double t = c->t1[c->T1 - 1] + c->dt;
    c->delta_old = c->delta;
    double err = 0.0;
    int rc = rk_step(c, &err);
    if (rc) return rc;
    c->delta = err;
    if (c->tail_fit) {
        // after update() and before loop++ (simulation_runner.cpp:135-141)
        launch_tail_fit(c->tables(), c->H, c->W, (int)(c->T1 - 1), (long long)c->T1, c->params.delta_max,
                        (long long)c->last_rollback, c->stream);
        rc = launch_check("tail_fit_kernel");
        if (rc) return rc;
        c->tf_fresh = true;
    }
    c->loop++;
    const double dmax = c->params.delta_max;
    const int64_t every = c->params.maint_every > 0 ? c->params.maint_every : 100000;
    if (c->loop % every == 0) {
        int64_t prev = c->T1;
        int count = 0;
        while (count < 1) {
            rc = sparsify(c, dmax);
            if (rc) return rc;
            if (c->T1 >= prev) break;
            prev = c->T1;
            count++;
        }
        rc = run_search(c, kSearchB, nullptr, nullptr, nullptr, nullptr);
        if (rc) return rc;
        // last_rollback_loop stays -1000 unless rollbacks are on (simulation_runner.cpp:53,220)
        if (c->delta < dmax / 2 && c->loop - c->last_rollback > 1000) {
            c->dt *= 0.5;
            // CPU path: SSPRK104.  The GPU loop with use_serk2: RK54 -> SSPRK104, else one
            // SERK2 level up (init_SERK2(2 (init - 1)) sets init + 1, runge_kutta.cu:274),
            // capped at the rmax table
            if (!c->use_serk2 || c->method == DMFE_RK54) c->method = DMFE_SSPRK104;
            else if (c->method < kMaxMethod) c->method += 1;
        }
    }
    const double* rmax = c->params.rmax;
    if (c->delta < dmax && c->loop > 5 && (c->delta < 1.1 * c->delta_old || c->delta_old < dmax / 1000) &&
        rmax[c->method - 1] / c->specRad > c->dt && c->rho.back() == 1.0) {
        c->dt *= 1.01;
    } else if (c->delta > 2 * dmax && c->dt > c->params.delta_t_min) {
        c->dt *= 0.9;
    }
    if (c->method == DMFE_SSPRK104 && c->delta > dmax && rmax[0] / c->specRad > c->dt) {
        c->method = DMFE_RK54;
        c->dt *= 0.5;
    }
    double t_loop = t;
    if (c->rollback && c->delta > 2 * dmax) {
        // the GPU main loop's rollback (simulation_runner.cpp:280-303): drop the last n <= 10
        // steps (rollbackState, simulation_control.cpp:14-76: truncate the history, delta_t
        // from the kept last step, loop -= n, interpolate), then halve the step and restart
        // with SSPRK104
        const int64_t n = std::min<int64_t>(10, std::max<int64_t>(0, c->T1 - 2));
        if (n > 0) {
            c->T1 -= n;
            c->t1.resize((size_t)c->T1);
            c->rho.resize((size_t)c->T1);
            c->dt = c->t1[c->T1 - 1] - c->t1[c->T1 - 2];
            c->loop -= n;
            c->interp_valid = false;
            rc = run_search(c, kSearchB, nullptr, nullptr, nullptr, nullptr);
            if (rc) return rc;
            c->last_rollback = c->loop;
            t_loop = c->t1[c->T1 - 1] + c->dt;
        }
        c->dt = 0.5 * std::min(std::max(c->dt, c->params.delta_t_min), rmax[c->method - 1] / c->specRad);
        // a SERK2 level above the first steps down, anything else restarts as SSPRK104
        c->method = c->method > 3 ? c->method - 1 : DMFE_SSPRK104;
    }
    if (t_out) *t_out = t_loop;
    return DMFE_OK;
}

int observables(dmfe_ctx* c, double t, double* E, double* qk0) {
    launch_row_scalar(c->tables(), c->H, c->W, c->model, (int)(c->T1 - 1), t, 1, c->stream);
    int rc = launch_check("row_scalar_kernel");
    if (rc) return rc;
    HIP_TRY(hipMemcpyAsync(c->pinned, c->W.scal + 2, sizeof(double), hipMemcpyDeviceToHost, c->stream));
    HIP_TRY(hipMemcpyAsync(c->pinned + 1, &c->H.V[(size_t)(c->T1 - 1) * c->N], sizeof(double), hipMemcpyDeviceToHost,
                           c->stream));
    HIP_TRY(hipStreamSynchronize(c->stream));
    if (E) *E = c->pinned[0];
    if (qk0) *qk0 = c->pinned[1];
    return DMFE_OK;
}

void free_ctx(dmfe_ctx* c) {
    if (!c) return;
    if (c->dloop) device_loop_close(c->dloop);  // waits for the stream; frees only the loop's own buffers
    c->dloop = nullptr;
    hipFree(c->theta); hipFree(c->integ); hipFree(c->phi2);
    for (int f = 0; f < 3; ++f) { hipFree(c->inds[f]); hipFree(c->w[f]); }
    free_history(c->H);
    hipFree(c->W.posB1); hipFree(c->W.posB2); hipFree(c->W.saveB1); hipFree(c->W.saveB2);
    hipFree(c->W.Floor); hipFree(c->W.diff); hipFree(c->W.rpk); hipFree(c->W.rInt); hipFree(c->W.partial);
    hipFree(c->W.qkb1c0); hipFree(c->W.hK); hipFree(c->W.hR); hipFree(c->W.scal);
    hipFree(c->W.fA[0]); hipFree(c->W.fA[1]);
    hipFree(c->W.sE); hipFree(c->W.sF);
    hipFree(c->tf_pre_buf); hipFree(c->tf_state);
    tangent_free(c);
    hipFree(c->fields);
    if (c->pinned) hipHostFree(c->pinned);
    for (auto& v : c->ev)
        for (hipEvent_t e : v) hipEventDestroy(e);
    if (c->stream) hipStreamDestroy(c->stream);
    delete c;
}

}  // namespace

// ============================== C ABI ==============================================

extern "C" const char* dmfe_last_error(void) { return g_err.c_str(); }

extern "C" const char* dmfe_version(void) { return "dmfe-mi355x 0.1.0"; }

extern "C" void dmfe_params_default(dmfe_params* p) {
    if (!p) return;
    *p = dmfe_params{};
    p->p = 3; p->p2 = 12; p->lambda = 0.3; p->T0 = 1e50; p->Gamma = 0.0;
    p->delta_t_min = 1e-5; p->delta_max = 1e-10;
    const double r[11] = {3, 13, 20, 80, 180, 320, 500, 720, 980, 1280, 1620};
    std::memcpy(p->rmax, r, sizeof r);
    p->log_response = 0;
    p->maint_every = 100000;
}

extern "C" int dmfe_ctx_create(const dmfe_grid* g, const dmfe_params* prm, int64_t rows_hint, dmfe_ctx** out) {
    if (!g || !prm || !out) return set_error(DMFE_ERR_ARG, "dmfe_ctx_create: null argument");
    *out = nullptr;
    if (g->N < 4 || g->N > kMaxN) return set_error(DMFE_ERR_ARG, "dmfe_ctx_create: N must be in [4, 2048]");
    if (g->m < 1 || g->m > g->N) return set_error(DMFE_ERR_ARG, "dmfe_ctx_create: stencil width m must be in [1, N]");
    if (!g->theta || !g->integ || !g->phi2) return set_error(DMFE_ERR_ARG, "dmfe_ctx_create: missing theta/int/phi2");
    for (int f = 0; f < 3; ++f)
        if (!g->inds[f] || !g->weights[f]) return set_error(DMFE_ERR_ARG, "dmfe_ctx_create: missing stencil table");
    const int N = (int)g->N, m = (int)g->m;
    const long long P = (long long)N * N;
    // validate stencil windows: start in [0, N-m] (negative = dense row starting at 0)
    for (int f = 0; f < 3; ++f)
        for (long long k = 0; k < P; ++k) {
            const int64_t s = g->inds[f][k] < 0 ? 0 : g->inds[f][k];
            if (s + m > N)
                return set_error(DMFE_ERR_ARG, "dmfe_ctx_create: stencil start " + std::to_string(g->inds[f][k]) +
                                                   " + width exceeds N in table " + std::to_string(f));
        }
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || ndev < 1) return set_error(DMFE_ERR_HIP, "dmfe_ctx_create: no HIP device");
    dmfe_ctx* c = new dmfe_ctx();
    hipGetDevice(&c->device);
    c->params = *prm;
    c->model.lambda = prm->lambda; c->model.T0 = prm->T0; c->model.Gamma = prm->Gamma;
    c->model.p = prm->p; c->model.p2 = prm->p2;
    c->N = N; c->m = m; c->P = P;
    {
        const RhsShape sh = rhs_shape_lookup(N, m, prm->log_response);
        c->G = sh.G;
        c->split = sh.split;
    }
    c->dense_mfma = (m == N) && !prm->log_response;
    c->specRad = 4 * std::sqrt(host_fl(*prm, 1, 2));
    auto fail = [&](int code, const std::string& msg) { free_ctx(c); return set_error(code, msg); };
    if (hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking) != hipSuccess) return fail(DMFE_ERR_HIP, "stream create failed");
    // the no-allocation entry point's scratch on this device (looked up once per device);
    // a failure only disables dmfe_rhs_noalloc, which then reports it
    if (scratch_address(&c->scratch) != DMFE_OK) c->scratch_msg = dmfe_last_error();
    const size_t bN = (size_t)N * sizeof(double), bP = (size_t)P * sizeof(double);
    bool ok = hipMalloc(&c->theta, bN) == hipSuccess && hipMalloc(&c->integ, bN) == hipSuccess &&
              hipMalloc(&c->phi2, bP) == hipSuccess;
    for (int f = 0; ok && f < 3; ++f)
        ok = hipMalloc(&c->inds[f], (size_t)P * sizeof(int)) == hipSuccess &&
             hipMalloc(&c->w[f], bP * (size_t)m) == hipSuccess;
    ok = ok && hipMalloc(&c->W.posB1, 2 * bN) == hipSuccess && hipMalloc(&c->W.posB2, 2 * bP) == hipSuccess &&
         hipMalloc(&c->W.saveB1, 3 * bN) == hipSuccess && hipMalloc(&c->W.saveB2, 3 * bP) == hipSuccess &&
         hipMalloc(&c->W.Floor, 2 * (size_t)N * sizeof(int)) == hipSuccess && hipMalloc(&c->W.diff, 2 * bN) == hipSuccess &&
         hipMalloc(&c->W.rpk, 2 * (size_t)N * sizeof(double4)) == hipSuccess &&
         hipMalloc(&c->W.rInt, bN) == hipSuccess &&
         hipMalloc(&c->W.partial, bN * (size_t)rhs_max_split(N) * kSums) == hipSuccess &&
         hipMalloc(&c->W.qkb1c0, bN) == hipSuccess && hipMalloc(&c->W.hK, bN * kMaxStages) == hipSuccess &&
         hipMalloc(&c->W.hR, bN * kMaxStages) == hipSuccess && hipMalloc(&c->W.scal, 8 * sizeof(double)) == hipSuccess &&
         hipHostMalloc((void**)&c->pinned, 8 * sizeof(double), hipHostMallocDefault) == hipSuccess;
    ok = ok && hipMalloc(&c->W.sE, (size_t)N * sizeof(double2)) == hipSuccess &&
         hipMalloc(&c->W.sF, (size_t)N * sizeof(double2)) == hipSuccess;
    if (ok && c->dense_mfma)
        ok = hipMalloc(&c->W.fA[0], (size_t)P * sizeof(double2)) == hipSuccess &&
             hipMalloc(&c->W.fA[1], (size_t)P * sizeof(double2)) == hipSuccess;
    if (!ok) return fail(DMFE_ERR_NOMEM, "dmfe_ctx_create: device allocation failed");
    // uploads: weights transposed [P][m] -> [m][P], stencil starts to int32
    std::vector<double> wt((size_t)P * m);
    std::vector<int> it((size_t)P);
    bool up = hipMemcpy(c->theta, g->theta, bN, hipMemcpyHostToDevice) == hipSuccess &&
              hipMemcpy(c->integ, g->integ, bN, hipMemcpyHostToDevice) == hipSuccess &&
              hipMemcpy(c->phi2, g->phi2, bP, hipMemcpyHostToDevice) == hipSuccess;
    for (int f = 0; up && f < 3; ++f) {
        // A1/A2 are read one target per lane: transposed [m][P] so a wavefront's taps are
        // contiguous; B2 is read one tap per lane: kept entry-major [P][m] as on disk.
        // Dense grids contract A1/A2 on MFMA from the entry-major rows.
        const bool transpose = f < 2 && !c->dense_mfma;
        for (long long k = 0; k < P; ++k) {
            it[(size_t)k] = g->inds[f][k] < 0 ? 0 : (int)g->inds[f][k];
            if (transpose)
                for (int d = 0; d < m; ++d) wt[(size_t)d * P + k] = g->weights[f][(size_t)k * m + d];
        }
        const double* src = transpose ? wt.data() : g->weights[f];
        up = hipMemcpy(c->inds[f], it.data(), (size_t)P * sizeof(int), hipMemcpyHostToDevice) == hipSuccess &&
             hipMemcpy(c->w[f], src, bP * (size_t)m, hipMemcpyHostToDevice) == hipSuccess;
    }
    // posB1xOld = 1, posB2xOld = 0 (initialization.cpp:234-235)
    std::vector<double> ones((size_t)N, 1.0), zeros((size_t)P, 0.0);
    up = up && hipMemcpy(c->W.posB1, ones.data(), bN, hipMemcpyHostToDevice) == hipSuccess &&
         hipMemcpy(c->W.posB1 + N, ones.data(), bN, hipMemcpyHostToDevice) == hipSuccess &&
         hipMemcpy(c->W.posB2, zeros.data(), bP, hipMemcpyHostToDevice) == hipSuccess &&
         hipMemcpy(c->W.posB2 + P, zeros.data(), bP, hipMemcpyHostToDevice) == hipSuccess &&
         // on the context's own stream: a null-stream memset is not ordered with it and, with
         // other host threads using the device, can land after the context's first kernels
         hipMemsetAsync(c->W.scal, 0, 8 * sizeof(double), c->stream) == hipSuccess &&
         hipStreamSynchronize(c->stream) == hipSuccess;
    if (!up) return fail(DMFE_ERR_HIP, "dmfe_ctx_create: upload failed");
    int rc = ensure_rows(c, std::max<int64_t>(rows_hint, 64));
    if (rc) { std::string msg = g_err; free_ctx(c); return set_error(rc, msg); }
    *out = c;
    return DMFE_OK;
}

extern "C" int dmfe_ctx_destroy(dmfe_ctx* c) {
    if (c) hipStreamSynchronize(c->stream);
    free_ctx(c);
    return DMFE_OK;
}

extern "C" int dmfe_init_quench(dmfe_ctx* c) {
    if (!c) return set_error(DMFE_ERR_ARG, "null context");
    tangent_free(c);
    const int N = c->N;
    int rc = ensure_rows(c, 1);
    if (rc) return rc;
    c->T1 = 1;
    c->t1.assign(1, 0.0);
    c->rho.assign(1, 0.0);
    c->specRad = 4 * std::sqrt(host_fl(c->params, 1, 2));
    c->dt = c->params.delta_t_min;
    c->loop = 0;
    c->delta = 1;
    c->delta_old = 0;
    std::vector<double2> v((size_t)N, make_double2(1.0, 1.0)), d((size_t)N, make_double2(0.0, 0.0));
    const double zero = 0.0;
    const double r0 = c->params.Gamma + host_fl(c->params, 1, 1) / c->params.T0;
    HIP_TRY(hipMemcpyAsync(c->H.V, v.data(), N * sizeof(double2), hipMemcpyHostToDevice, c->stream));
    HIP_TRY(hipMemcpyAsync(c->H.D, d.data(), N * sizeof(double2), hipMemcpyHostToDevice, c->stream));
    HIP_TRY(hipMemcpyAsync(c->H.t1, &zero, sizeof(double), hipMemcpyHostToDevice, c->stream));
    HIP_TRY(hipMemcpyAsync(c->H.rho, &zero, sizeof(double), hipMemcpyHostToDevice, c->stream));
    HIP_TRY(hipMemcpyAsync(c->H.r, &r0, sizeof(double), hipMemcpyHostToDevice, c->stream));
    // drvec[0] = rstep() on the initial row (initialization.cpp:222)
    launch_row_scalar(c->tables(), c->H, c->W, c->model, 0, 0.0, 0, c->stream);
    rc = launch_check("row_scalar_kernel");
    if (rc) return rc;
    HIP_TRY(hipMemcpyAsync(c->H.dr, c->W.scal + 2, sizeof(double), hipMemcpyDeviceToDevice, c->stream));
    c->method = (c->dt < c->params.rmax[0] / c->specRad) ? DMFE_RK54 : DMFE_SSPRK104;
    rc = run_search(c, kSearchB, nullptr, nullptr, nullptr, nullptr);
    if (rc) return rc;
    HIP_TRY(hipStreamSynchronize(c->stream));
    return DMFE_OK;
}

extern "C" int dmfe_get_T1(const dmfe_ctx* c, int64_t* T1) {
    if (!c || !T1) return set_error(DMFE_ERR_ARG, "null argument");
    *T1 = c->T1;
    return DMFE_OK;
}

extern "C" int dmfe_set_state(dmfe_ctx* c, int64_t T1, const double* t1grid, const double* QK, const double* QR,
                              const double* dQK, const double* dQR, const double* rvec, const double* drvec,
                              const double* dtratio) {
    if (!c || T1 < 1 || !t1grid || !QK || !QR || !dQK || !dQR || !rvec || !drvec)
        return set_error(DMFE_ERR_ARG, "dmfe_set_state: null argument or T1 < 1");
    for (int64_t i = 1; i < T1; ++i)
        if (!(t1grid[i] > t1grid[i - 1])) return set_error(DMFE_ERR_ARG, "dmfe_set_state: t1grid must be strictly increasing");
    tangent_free(c);  // a new history ends the tangent run (a rejected call keeps it)
    int rc = ensure_rows(c, T1 + 1);
    if (rc) return rc;
    const size_t N = (size_t)c->N, n = (size_t)T1;
    c->T1 = T1;
    c->t1.assign(t1grid, t1grid + n);
    c->rho.assign(n, 0.0);
    if (dtratio) {
        c->rho.assign(dtratio, dtratio + n);
    } else {  // io_input.cpp:169-172
        for (size_t i = 2; i < n; ++i) c->rho[i] = (c->t1[i] - c->t1[i - 1]) / (c->t1[i - 1] - c->t1[i - 2]);
    }
    std::vector<double2> v(n * N), d(n * N);
    for (size_t k = 0; k < n * N; ++k) { v[k] = make_double2(QK[k], QR[k]); d[k] = make_double2(dQK[k], dQR[k]); }
    HIP_TRY(hipMemcpyAsync(c->H.V, v.data(), n * N * sizeof(double2), hipMemcpyHostToDevice, c->stream));
    HIP_TRY(hipMemcpyAsync(c->H.D, d.data(), n * N * sizeof(double2), hipMemcpyHostToDevice, c->stream));
    HIP_TRY(hipMemcpyAsync(c->H.t1, c->t1.data(), n * sizeof(double), hipMemcpyHostToDevice, c->stream));
    HIP_TRY(hipMemcpyAsync(c->H.rho, c->rho.data(), n * sizeof(double), hipMemcpyHostToDevice, c->stream));
    HIP_TRY(hipMemcpyAsync(c->H.r, rvec, n * sizeof(double), hipMemcpyHostToDevice, c->stream));
    HIP_TRY(hipMemcpyAsync(c->H.dr, drvec, n * sizeof(double), hipMemcpyHostToDevice, c->stream));
    HIP_TRY(hipStreamSynchronize(c->stream));
    c->interp_valid = false;
    return DMFE_OK;
}

extern "C" int dmfe_get_state(dmfe_ctx* c, double* t1grid, double* QK, double* QR, double* dQK, double* dQR,
                              double* rvec, double* drvec, double* dtratio) {
    if (!c) return set_error(DMFE_ERR_ARG, "null context");
    const size_t N = (size_t)c->N, n = (size_t)c->T1;
    if (t1grid) std::memcpy(t1grid, c->t1.data(), n * sizeof(double));
    if (dtratio) std::memcpy(dtratio, c->rho.data(), n * sizeof(double));
    if (QK || QR || dQK || dQR) {
        std::vector<double2> v(n * N), d(n * N);
        HIP_TRY(hipMemcpyAsync(v.data(), c->H.V, n * N * sizeof(double2), hipMemcpyDeviceToHost, c->stream));
        HIP_TRY(hipMemcpyAsync(d.data(), c->H.D, n * N * sizeof(double2), hipMemcpyDeviceToHost, c->stream));
        HIP_TRY(hipStreamSynchronize(c->stream));
        for (size_t k = 0; k < n * N; ++k) {
            if (QK) QK[k] = v[k].x;
            if (QR) QR[k] = v[k].y;
            if (dQK) dQK[k] = d[k].x;
            if (dQR) dQR[k] = d[k].y;
        }
    }
    if (rvec) { int rc = copy_out(c, rvec, c->H.r, n * sizeof(double)); if (rc) return rc; }
    if (drvec) { int rc = copy_out(c, drvec, c->H.dr, n * sizeof(double)); if (rc) return rc; }
    return DMFE_OK;
}

extern "C" int dmfe_set_control(dmfe_ctx* c, double delta_t, double delta, double delta_old, int64_t loop, int method) {
    if (!c || method < DMFE_RK54 || method > kMaxMethod)
        return set_error(DMFE_ERR_ARG, "dmfe_set_control: bad argument (method 1 RK54, 2 SSPRK104, 3.." +
                                           std::to_string(kMaxMethod) + " SERK2(2 (method - 2)))");
    if (!(delta_t > 0.0) || !std::isfinite(delta_t) || !std::isfinite(delta) || !std::isfinite(delta_old) || loop < 0)
        return set_error(DMFE_ERR_ARG, "dmfe_set_control: delta_t must be finite and > 0, delta and delta_old finite, "
                                       "loop >= 0");
    c->dt = delta_t; c->delta = delta; c->delta_old = delta_old; c->loop = loop; c->method = method;
    return DMFE_OK;
}

extern "C" int dmfe_get_control(const dmfe_ctx* c, double* delta_t, double* delta, double* delta_old, int64_t* loop,
                                int* method) {
    if (!c) return set_error(DMFE_ERR_ARG, "null context");
    if (delta_t) *delta_t = c->dt;
    if (delta) *delta = c->delta;
    if (delta_old) *delta_old = c->delta_old;
    if (loop) *loop = c->loop;
    if (method) *method = c->method;
    return DMFE_OK;
}

extern "C" int dmfe_interpolate(dmfe_ctx* c, int mode) {
    if (!c || c->T1 < 1) return set_error(DMFE_ERR_ARG, "dmfe_interpolate: no state");
    if (mode != DMFE_POS_BSEARCH && mode != DMFE_POS_ISEARCH && mode != DMFE_POS_SAME)
        return set_error(DMFE_ERR_ARG, "dmfe_interpolate: bad mode");
    const int km = mode == DMFE_POS_BSEARCH ? kSearchB : (mode == DMFE_POS_ISEARCH ? kSearchI : kSearchSame);
    int rc = run_search(c, km, nullptr, nullptr, nullptr, nullptr);
    if (rc) return rc;
    HIP_TRY(hipStreamSynchronize(c->stream));
    return DMFE_OK;
}

extern "C" int dmfe_get_positions(dmfe_ctx* c, double* posB1, double* posB2) {
    if (!c) return set_error(DMFE_ERR_ARG, "null context");
    const double* b1 = c->W.posB1 + (size_t)c->fcur * c->N;  // the latest interpolation's buffer
    const double* b2 = c->W.posB2 + (size_t)c->fcur * c->P;
    if (posB1) { int rc = copy_out(c, posB1, b1, (size_t)c->N * sizeof(double)); if (rc) return rc; }
    if (posB2) { int rc = copy_out(c, posB2, b2, (size_t)c->P * sizeof(double)); if (rc) return rc; }
    return DMFE_OK;
}

extern "C" int dmfe_set_positions(dmfe_ctx* c, const double* posB1, const double* posB2) {
    if (!c || !posB1 || !posB2) return set_error(DMFE_ERR_ARG, "null argument");
    c->tf_fresh = false;
    HIP_TRY(hipMemcpyAsync(c->W.posB1 + (size_t)c->fcur * c->N, posB1, (size_t)c->N * sizeof(double),
                           hipMemcpyHostToDevice, c->stream));
    HIP_TRY(hipMemcpyAsync(c->W.posB2 + (size_t)c->fcur * c->P, posB2, (size_t)c->P * sizeof(double),
                           hipMemcpyHostToDevice, c->stream));
    HIP_TRY(hipStreamSynchronize(c->stream));
    return DMFE_OK;
}

extern "C" int dmfe_get_fields(dmfe_ctx* c, double* out) {
    if (!c || !out || c->T1 < 1) return set_error(DMFE_ERR_ARG, "dmfe_get_fields: bad argument");
    if (!c->fields) HIP_TRY(hipMalloc(&c->fields, (size_t)c->P * 15 * sizeof(double)));
    int rc = run_rhs_rows(c, c->fields);
    if (rc) return rc;
    FinishArgs f{};  // RHS epilogue only: also leaves rInt in the work buffer
    f.mode = 0;
    f.split = c->split;
    f.lr = (int)(c->T1 - 1);
    f.t_rhs = c->t1[c->T1 - 1];
    rc = run_stage(c, nullptr, &f);
    if (rc) return rc;
    rc = copy_out(c, out, c->fields, (size_t)c->P * 15 * sizeof(double));
    if (rc) return rc;
    return copy_out(c, out + (size_t)c->P * 15, c->W.rInt, (size_t)c->N * sizeof(double));
}

extern "C" int dmfe_rhs(dmfe_ctx* c, double* hK, double* hR) {
    if (!c || !hK || !hR || c->T1 < 1) return set_error(DMFE_ERR_ARG, "dmfe_rhs: bad argument");
    int rc = run_rhs_rows(c, nullptr);
    if (rc) return rc;
    FinishArgs f{};
    f.mode = 0;
    f.stage = 0;
    f.split = c->split;
    f.lr = (int)(c->T1 - 1);
    f.t_rhs = c->t1[c->T1 - 1];
    rc = run_stage(c, nullptr, &f);
    if (rc) return rc;
    rc = copy_out(c, hK, c->W.hK, (size_t)c->N * sizeof(double));
    if (rc) return rc;
    return copy_out(c, hR, c->W.hR, (size_t)c->N * sizeof(double));
}

// dmfe_rhs with its scratch (quadrature partials, QK_B1[j,0], rInt and the hK / hR
// staging) in the library's file-scope device array instead of the context's work
// buffers: no device memory is allocated or freed.  Same kernels and launch shape as
// dmfe_rhs, hence the same results.  Every call on a device shares that device's array,
// so calls that overlap on one device (contexts driven from several host threads) race.
// Nothing in the array has to start at zero: the RHS kernel writes every partial it sums.
extern "C" int dmfe_rhs_noalloc(dmfe_ctx* c, double* hK, double* hR) {
    if (!c || !hK || !hR || c->T1 < 1) return set_error(DMFE_ERR_ARG, "dmfe_rhs_noalloc: bad argument");
    const size_t need = rhs_scratch_doubles(c->N, c->split), avail = kScratchDoubles;
    if (need > avail)
        return set_error(DMFE_ERR_ARG, "dmfe_rhs_noalloc: N = " + std::to_string(c->N) + " with split " +
                                           std::to_string(c->split) + " needs " + std::to_string(need) +
                                           " doubles of scratch, more than the " + std::to_string(avail) +
                                           " available (8 N split + 4 N must fit)");
    if (!c->scratch) return set_error(DMFE_ERR_HIP, "dmfe_rhs_noalloc: device scratch unavailable: " + c->scratch_msg);
    Work W = c->W;
    double* p = c->scratch;
    W.partial = p; p += (size_t)c->N * c->split * kSums;
    W.qkb1c0 = p; p += c->N;
    W.rInt = p; p += c->N;
    W.hK = p; p += c->N;
    W.hR = p;
    RhsArgs a{};
    a.T1 = (int)c->T1;
    a.tl = c->t1[c->T1 - 1];
    a.split = c->split;
    a.log_response = c->params.log_response;
    a.fields = nullptr;
    a.fbuf = c->fcur;
    a.G = c->G;
    if (c->dense_mfma) {  // A1/A2 of a dense grid into the context's field buffers, as dmfe_rhs
        const double* last = reinterpret_cast<const double*>(c->H.V + (size_t)(c->T1 - 1) * c->N);
        for (int f = 0; f < 2; ++f)
            launch_dense_apply(c->w[f], last, reinterpret_cast<double*>(W.fA[f]), c->P, c->N, 2, 2, c->stream);
    }
    launch_rhs_rows(c->tables(), c->H, W, c->model, a, c->stream);
    int rc = launch_check("rhs_rows_kernel");
    if (rc) return rc;
    FinishArgs f{};
    f.mode = 0;
    f.stage = 0;
    f.split = c->split;
    f.lr = (int)(c->T1 - 1);
    f.t_rhs = c->t1[c->T1 - 1];
    f.fbuf = c->fcur;
    launch_stage(c->tables(), c->H, W, c->model, nullptr, &f, c->stream);
    rc = launch_check("stage_kernel");
    if (rc) return rc;
    rc = copy_out(c, hK, W.hK, (size_t)c->N * sizeof(double));
    if (rc) return rc;
    return copy_out(c, hR, W.hR, (size_t)c->N * sizeof(double));
}

extern "C" int dmfe_rstep(dmfe_ctx* c, double* r) {
    if (!c || !r || c->T1 < 1) return set_error(DMFE_ERR_ARG, "dmfe_rstep: bad argument");
    launch_row_scalar(c->tables(), c->H, c->W, c->model, (int)(c->T1 - 1), c->t1[c->T1 - 1], 0, c->stream);
    int rc = launch_check("row_scalar_kernel");
    if (rc) return rc;
    return copy_out(c, r, c->W.scal + 2, sizeof(double));
}

extern "C" int dmfe_rk_step(dmfe_ctx* c, double* err) {
    if (!c || c->T1 < 1) return set_error(DMFE_ERR_ARG, "dmfe_rk_step: no state");
    const int rc = rk_step(c, err);
    c->tf_fresh = false;  // update() alone: no tail fit follows
    return rc;
}

extern "C" int dmfe_loop_iteration(dmfe_ctx* c, double* t_out) {
    if (!c || c->T1 < 1) return set_error(DMFE_ERR_ARG, "dmfe_loop_iteration: no state");
    return loop_iteration(c, t_out);
}

// dmfe_run on the device-side step control: the whole main loop runs as replayed
// hipGraphs of 16 steps with the dt rule on the device; the host only services arena
// growth and the maintenance cadence between replays.  Same arithmetic as the host loop.
// With tangents the run goes in segments that fit the arenas (grown here, tangents
// included) and stops before a sparsification point, as loop_iteration does.
static int run_device_loop(dmfe_ctx* c, double tmax, int64_t max_loop, double* corr, int64_t corr_cap,
                           int64_t* steps, double* t_out = nullptr) {
    if (!c->interp_valid && c->T1 > 1) {
        int rc = run_search(c, kSearchB, nullptr, nullptr, nullptr, nullptr);
        if (rc) return rc;
    }
    if (int rc = tail_fit_prepare(c)) return rc;
    int64_t done = 0;  // iterations (a rollback moves the loop counter back)
    const int64_t every = c->params.maint_every > 0 ? c->params.maint_every : 100000;
    double t_prop = c->t1[c->T1 - 1];
    int rc = DMFE_OK;
    for (;;) {
        int64_t lim = max_loop;
        bool at_maint = false;
        if (c->tanK) {
            rc = ensure_rows(c, c->T1 + std::max<int64_t>(256, c->T1 / 2));
            if (rc) break;
            lim = std::min<int64_t>(lim, c->loop + (c->cap - c->T1 - 1));
            const int64_t next_maint = (c->loop / every + 1) * every;
            if (lim >= next_maint) {
                lim = next_maint - 1;
                at_maint = true;
            }
        }
        DeviceLoop q{};
        q.stream = c->stream;
        q.T = c->tables();
        q.W = c->W;
        q.M = c->model;
        q.prm = c->params;
        q.split = c->split;
        q.G = c->G;
        q.use_graph = c->loop_graphs ? 1 : 0;
        if (c->tanK) q.X = c->tan_args();
        q.H = c->H;
        q.cap = c->cap;
        q.T1 = c->T1;
        q.loop = c->loop;
        q.method = c->method;
        q.fcur = c->fcur;
        q.dt = c->dt;
        q.delta = c->delta;
        q.delta_old = c->delta_old;
        q.t_prop = t_prop;
        q.gpu_variants = c->gpu_variants ? 1 : 0;
        q.rollback = c->rollback ? 1 : 0;
        q.use_serk2 = c->use_serk2 ? 1 : 0;
        q.last_rollback = c->last_rollback;
        double* cseg = (corr && done < corr_cap) ? corr + 3 * done : nullptr;
        int64_t seg_iters = 0;
        rc = device_loop_run(&c->dloop, q, tmax, lim, cseg, cseg ? corr_cap - done : 0, &seg_iters);
        done += seg_iters;
        c->last_rollback = q.last_rollback;
        c->H = q.H;
        c->cap = q.cap;
        c->T1 = q.T1;
        c->loop = q.loop;
        c->method = q.method;
        c->fcur = q.fcur;
        c->dt = q.dt;
        c->delta = q.delta;
        c->delta_old = q.delta_old;
        c->interp_valid = true;
        // the loop's tail fit (or its maintenance's interpolation) left W.tf_pre current
        if (c->W.tf_pre) c->tf_fresh = true;
        t_prop = q.t_prop;
        if (rc || !c->tanK || c->loop < lim || lim >= max_loop) break;
        // a segment ended at its own limit: would the main loop go on?
        const double eps = std::numeric_limits<double>::epsilon();
        const bool more = t_prop < tmax &&
                          c->dt >= c->params.delta_t_min - eps * std::max(std::fabs(c->dt), std::fabs(c->params.delta_t_min));
        if (!more) break;
        if (at_maint && c->loop == lim) {
            rc = set_error(DMFE_ERR_STATE, "sensitivities: the run would cross a sparsification point (loop " +
                                               std::to_string(c->loop + 1) + ")");
            break;
        }
    }
    if (steps) *steps = done;
    if (t_out) *t_out = t_prop;
    // host mirrors of t1grid and delta_t_ratio
    c->t1.resize((size_t)c->T1);
    c->rho.resize((size_t)c->T1);
    HIP_TRY(hipMemcpyAsync(c->t1.data(), c->H.t1, (size_t)c->T1 * sizeof(double), hipMemcpyDeviceToHost, c->stream));
    HIP_TRY(hipMemcpyAsync(c->rho.data(), c->H.rho, (size_t)c->T1 * sizeof(double), hipMemcpyDeviceToHost, c->stream));
    HIP_TRY(hipStreamSynchronize(c->stream));
    return rc;
}

// ---- SERK2 (the reference GPU loop's stabilized integrator) --------------------------
// Coefficients (runge_kutta.cu:228-270): s = 2q stages with Chebyshev stage polynomials
// T_i(x) T_2(x)^(j-1), x = 1 + z q / s^2; b[0..s] make sum_k b_k P_k(z) match
// R(z) = T_s(w0 + w1 z) / T_s(w0) at z = 1..s+1, w0 = 1 + (19/20) / s^2,
// w1 = T_s(w0) / (s U_{s-1}(w0)); solved in long double by Gaussian elimination with
// partial pivoting.
namespace {  // the same unnamed namespace as the step functions
long double cheb_T(int n, long double x) {
    long double a = 1.0L, b = x;
    if (n == 0) return a;
    for (int k = 1; k < n; ++k) {
        const long double c2 = 2.0L * x * b - a;
        a = b;
        b = c2;
    }
    return b;
}
long double cheb_U(int n, long double x) {
    long double a = 1.0L, b = 2.0L * x;
    if (n == 0) return a;
    for (int k = 1; k < n; ++k) {
        const long double c2 = 2.0L * x * b - a;
        a = b;
        b = c2;
    }
    return b;
}

std::vector<double> serk2_solve(int q) {
    const int m = 2, s = m * q, n = s + 1;
    const long double a2 = (long double)q / ((long double)s * s);
    const long double w0 = 1.0L + (19.0L / 20.0L) / ((long double)s * s);
    const long double w1 = cheb_T(s, w0) / ((long double)s * cheb_U(s - 1, w0));
    std::vector<std::vector<long double>> A((size_t)n, std::vector<long double>((size_t)n + 1, 0.0L));
    for (int r = 0; r < n; ++r) {
        const long double z = (long double)(r + 1), x = 1.0L + a2 * z;
        auto& row = A[(size_t)r];
        row[0] = 1.0L;
        long double tm = 1.0L;
        for (int j = 1; j <= q; ++j) {
            for (int i = 1; i <= m; ++i) row[(size_t)(i + m * (j - 1))] = cheb_T(i, x) * tm;
            tm *= cheb_T(m, x);
        }
        row[(size_t)n] = cheb_T(s, w0 + w1 * z) / cheb_T(s, w0);
    }
    for (int k = 0; k < n; ++k) {
        int pr = k;
        for (int r = k + 1; r < n; ++r)
            if (std::fabs(A[(size_t)r][(size_t)k]) > std::fabs(A[(size_t)pr][(size_t)k])) pr = r;
        std::swap(A[(size_t)k], A[(size_t)pr]);
        for (int r = k + 1; r < n; ++r) {
            const long double f = A[(size_t)r][(size_t)k] / A[(size_t)k][(size_t)k];
            for (int col = k; col <= n; ++col) A[(size_t)r][(size_t)col] -= f * A[(size_t)k][(size_t)col];
        }
    }
    std::vector<double> b((size_t)n);
    for (int k = n - 1; k >= 0; --k) {
        long double v = A[(size_t)k][(size_t)n];
        for (int col = k + 1; col < n; ++col) v -= A[(size_t)k][(size_t)col] * (long double)b[(size_t)col];
        b[(size_t)k] = (double)(v / A[(size_t)k][(size_t)k]);
    }
    return b;
}

}  // namespace

// the weights of each level are solved once per process
std::vector<double> dmfe::serk2_coeffs(int q) {
    static std::mutex mu;
    static std::map<int, std::vector<double>> cache;
    std::lock_guard<std::mutex> lock(mu);
    auto it = cache.find(q);
    if (it == cache.end()) it = cache.emplace(q, serk2_solve(q)).first;
    return it->second;
}

namespace {
// One SERK2 step (SERK2GPU, runge_kutta.cu:397-487, on this path's searches): stage n
// searches the positions (stage 0 only on a fresh history), evaluates the RHS and the
// stage launch of n+1 finishes it: g += dt alpha h (even n), g = 2 g - g_e + 2 dt alpha h
// (odd n, g_e := g), alpha = 1 / (4q); the embedded sum F = b_0 g_0 + sum b_{n+1} g_{n+1}
// and the error sum |F - g_s| + |t + dt - gt_s| at the last stage.
int rk_step_serk2(dmfe_ctx* c, int q, double* err_out) {
    if (c->tanK) return set_error(DMFE_ERR_STATE, "sensitivities are not carried through SERK2 steps");
    const int S = 2 * q;
    const std::vector<double> b = serk2_coeffs(q);
    int rc = ensure_rows(c, c->T1 + 1);
    if (rc) return rc;
    if (!c->interp_valid && c->T1 > 1) {
        rc = run_search(c, kSearchB, nullptr, nullptr, nullptr, nullptr);
        if (rc) return rc;
    }
    rc = tail_fit_prepare(c);
    if (rc) return rc;
    const double dt = c->dt, alpha = 1.0 / (4 * q);
    const int64_t g0row = c->T1 - 1;
    const double t0 = c->t1[c->T1 - 1];
    double gt = t0, gte = t0;
    const double gtf = t0 + dt;
    FinishArgs pending{};
    bool have_pending = false;
    for (int n = 0; n < S; ++n) {
        SearchArgs sa{};
        bool search = false;
        if (c->T1 == 1 || n != 0) {
            search = true;
            sa = make_search(c, n == 0 ? kSearchB : kSearchI, nullptr, nullptr, nullptr, nullptr);
        }
        if (search || have_pending) {
            rc = run_stage(c, search ? &sa : nullptr, have_pending ? &pending : nullptr);
            if (rc) return rc;
            have_pending = false;
        }
        rc = run_rhs_rows(c, nullptr, search ? nullptr : c->W.tf_pre);
        if (rc) return rc;
        FinishArgs& f = pending;
        f = FinishArgs{};
        f.mode = 1;
        f.stage = n == 0 ? 0 : 1;
        f.nterms = 1;
        f.split = c->split;
        f.lr = (int)(c->T1 - 1);
        f.t_rhs = c->t1[c->T1 - 1];
        f.serk = (n % 2 == 1) ? 2 : 1;
        f.sa = (n % 2 == 1) ? 2 * dt * alpha : dt * alpha;
        f.sb0 = b[0];
        f.sb = b[(size_t)n + 1];
        f.last = (n == S - 1) ? 1 : 0;
        if (n == 0) {
            gt = t0 + dt * alpha * 1.0;
            c->t1.push_back(gt);
            const size_t idx = c->t1.size() - 1;
            const double tdiff = c->t1[idx] - c->t1[idx - 1];
            const double r = (idx > 1) ? tdiff / (c->t1[idx - 1] - c->t1[idx - 2]) : 0.0;
            c->rho.push_back(r);
            f.g0row = (int)g0row;
            f.wrow = (int)idx;
            f.t_new = gt;
            f.tdiff = tdiff;
            f.rho_new = r;
            c->T1 += 1;
        } else {
            if (n % 2 == 0) {
                gt = gt + dt * alpha * 1.0;
            } else {
                gt = 2 * gt - gte + 2 * dt * alpha * 1.0;
                gte = gt;
            }
            const size_t L = c->t1.size();
            c->t1[L - 1] = gt;
            const double tdiff = c->t1[L - 1] - c->t1[L - 2];
            const double r = (L > 2) ? tdiff / (c->t1[L - 2] - c->t1[L - 3]) : 0.0;
            c->rho[L - 1] = r;
            f.g0row = (int)(L - 1);  // the previous stage's values
            f.wrow = (int)(L - 1);
            f.t_new = gt;
            f.tdiff = tdiff;
            f.rho_new = r;
        }
        if (f.last) f.gt_err = std::fabs(gtf - gt);
        have_pending = true;
    }
    // positions for the next step, whose first stage reuses them
    const SearchArgs fin = make_search(c, kSearchI, nullptr, nullptr, nullptr, nullptr);
    rc = run_stage(c, &fin, &pending);
    if (rc) return rc;
    HIP_TRY(hipMemcpyAsync(c->pinned, c->W.scal + 1, sizeof(double), hipMemcpyDeviceToHost, c->stream));
    HIP_TRY(hipStreamSynchronize(c->stream));
    if (err_out) *err_out = c->pinned[0];
    return DMFE_OK;
}
}  // namespace

extern "C" int dmfe_set_serk2(dmfe_ctx* c, int on) {
    if (!c) return set_error(DMFE_ERR_ARG, "dmfe_set_serk2: null context");
    c->use_serk2 = on != 0;
    return DMFE_OK;
}

extern "C" int dmfe_set_rollback(dmfe_ctx* c, int on) {
    if (!c) return set_error(DMFE_ERR_ARG, "dmfe_set_rollback: null context");
    c->rollback = on != 0;
    return DMFE_OK;
}

// The QK tail fit near theta -> 1 after every main-loop step (--tail-fit; tail_fit.cpp:46-265,
// simulation_runner.cpp:135-141).  Switching it on starts the fit's state afresh (the
// reference's function-local statics at program start).
extern "C" int dmfe_set_tail_fit(dmfe_ctx* c, int on) {
    if (!c) return set_error(DMFE_ERR_ARG, "dmfe_set_tail_fit: null context");
    if (!on) {
        c->tail_fit = false;
        c->W.tf_pre = nullptr;
        c->W.tf = nullptr;
        return DMFE_OK;
    }
    if (c->N > kTailMaxN)
        return set_error(DMFE_ERR_ARG, "dmfe_set_tail_fit: N = " + std::to_string(c->N) + " exceeds " +
                                           std::to_string(kTailMaxN));
    if (c->dense_mfma) return set_error(DMFE_ERR_STATE, "dmfe_set_tail_fit: not available on dense (B-spline) grids");
    if (c->tanK) return set_error(DMFE_ERR_STATE, "dmfe_set_tail_fit: not available while sensitivities are carried");
    if (!c->tf_pre_buf && hipMalloc(&c->tf_pre_buf, (size_t)c->N * sizeof(double2)) != hipSuccess)
        return set_error(DMFE_ERR_NOMEM, "dmfe_set_tail_fit: allocation failed");
    if (!c->tf_state && hipMalloc(&c->tf_state, sizeof(TailState)) != hipSuccess)
        return set_error(DMFE_ERR_NOMEM, "dmfe_set_tail_fit: allocation failed");
    TailState st{};
    st.prev_off = -1;
    st.prev_t1len = 0;
    st.known_rollback = -1000;
    st.applied = 0;
    st.alpha = 0.0;
    HIP_TRY(hipMemcpyAsync(c->tf_state, &st, sizeof st, hipMemcpyHostToDevice, c->stream));
    HIP_TRY(hipStreamSynchronize(c->stream));
    c->tail_fit = true;
    c->tf_fresh = false;
    c->W.tf_pre = c->tf_pre_buf;
    c->W.tf = c->tf_state;
    return DMFE_OK;
}

extern "C" int dmfe_set_gpu_variants(dmfe_ctx* c, int on) {
    if (!c) return set_error(DMFE_ERR_ARG, "dmfe_set_gpu_variants: null context");
    c->gpu_variants = on != 0;
    c->interp_valid = false;  // the next step starts from a fresh interpolation
    return DMFE_OK;
}

extern "C" int dmfe_tail_fit_stats(dmfe_ctx* c, int64_t* applied, double* alpha) {
    if (!c) return set_error(DMFE_ERR_ARG, "dmfe_tail_fit_stats: null context");
    if (!c->tf_state) return set_error(DMFE_ERR_STATE, "dmfe_tail_fit_stats: the tail fit was never switched on");
    TailState st{};
    HIP_TRY(hipMemcpyAsync(&st, c->tf_state, sizeof st, hipMemcpyDeviceToHost, c->stream));
    HIP_TRY(hipStreamSynchronize(c->stream));
    if (applied) *applied = st.applied;
    if (alpha) *alpha = st.alpha;
    return DMFE_OK;
}

extern "C" int dmfe_set_loop_graphs(dmfe_ctx* c, int on) {
    if (!c) return set_error(DMFE_ERR_ARG, "dmfe_set_loop_graphs: null context");
    c->loop_graphs = on != 0;
    return DMFE_OK;
}

extern "C" int dmfe_run(dmfe_ctx* c, double tmax, int64_t max_loop, double* corr, int64_t corr_cap, int64_t* steps) {
    if (!c || c->T1 < 1) return set_error(DMFE_ERR_ARG, "dmfe_run: no state");
    // dense MFMA contexts, per-launch profiling, and tangents with SERK2 or rollbacks keep
    // the host-driven loop (which reports that SERK2 steps carry no tangents)
    if (!c->dense_mfma && !c->profiling && !(c->tanK && (c->rollback || c->use_serk2 || c->method >= 3)))
        return run_device_loop(c, tmax, max_loop, corr, corr_cap, steps);
    int64_t done = 0;
    double t = c->t1[c->T1 - 1];
    const double eps = std::numeric_limits<double>::epsilon();
    while (t < tmax && c->loop < max_loop &&
           c->dt >= c->params.delta_t_min - eps * std::max(std::fabs(c->dt), std::fabs(c->params.delta_t_min))) {
        int rc = loop_iteration(c, &t);
        if (rc) { if (steps) *steps = done; return rc; }
        if (corr && done < corr_cap) {
            double E = 0, q0 = 0;
            rc = observables(c, t, &E, &q0);
            if (rc) { if (steps) *steps = done; return rc; }
            corr[3 * done + 0] = t;
            corr[3 * done + 1] = E;
            corr[3 * done + 2] = q0;
        }
        ++done;
    }
    if (steps) *steps = done;
    return DMFE_OK;
}

extern "C" int dmfe_observables(dmfe_ctx* c, double t, double* energy, double* qk0) {
    if (!c || c->T1 < 1) return set_error(DMFE_ERR_ARG, "dmfe_observables: no state");
    return observables(c, t, energy, qk0);
}

extern "C" int dmfe_sparsify(dmfe_ctx* c, double threshold) {
    if (!c || c->T1 < 1) return set_error(DMFE_ERR_ARG, "dmfe_sparsify: no state");
    return sparsify(c, threshold);
}

extern "C" int dmfe_synchronize(dmfe_ctx* c) {
    if (!c) return set_error(DMFE_ERR_ARG, "null context");
    HIP_TRY(hipStreamSynchronize(c->stream));
    return DMFE_OK;
}

extern "C" int dmfe_profile(dmfe_ctx* c, int enable) {
    if (!c) return set_error(DMFE_ERR_ARG, "null context");
    HIP_TRY(hipStreamSynchronize(c->stream));
    c->profiling = enable != 0;
    for (int k = 0; k < kKinds; ++k) c->ev_used[k] = 0;
    return DMFE_OK;
}

extern "C" int dmfe_profile_read(dmfe_ctx* c, int kind, double* total_ms, int64_t* launches) {
    if (!c || kind < 0 || kind >= kKinds || !total_ms || !launches)
        return set_error(DMFE_ERR_ARG, "dmfe_profile_read: bad argument");
    HIP_TRY(hipStreamSynchronize(c->stream));
    double tot = 0.0;
    const size_t n = c->ev_used[kind];
    for (size_t i = 0; i + 1 < n; i += 2) {
        float ms = 0.f;
        HIP_TRY(hipEventElapsedTime(&ms, c->ev[kind][i], c->ev[kind][i + 1]));
        tot += ms;
    }
    *total_ms = tot;
    *launches = (int64_t)(n / 2);
    return DMFE_OK;
}

extern "C" int dmfe_tangent_init(dmfe_ctx* c, int K, const double* dGamma, const double* dLambda) {
    if (!c || K < 1 || K > kMaxTan || !dGamma || !dLambda)
        return set_error(DMFE_ERR_ARG, "dmfe_tangent_init: bad argument (1 <= K <= 2)");
    if (c->T1 != 1) return set_error(DMFE_ERR_STATE, "dmfe_tangent_init: call right after dmfe_init_quench");
    if (c->params.log_response)
        return set_error(DMFE_ERR_STATE, "dmfe_tangent_init: the log-response interpolation is not differentiated");
    if (c->tail_fit) return set_error(DMFE_ERR_STATE, "dmfe_tangent_init: the tail fit is not differentiated");
    tangent_free(c);
    for (int k = 0; k < K; ++k) {
        int rc = alloc_history(c->tanH[k], c->cap, c->N);
        if (rc) { tangent_free(c); return rc; }
        c->tanG[k] = dGamma[k];
        c->tanL[k] = dLambda[k];
    }
    c->tanK = K;
    if (hipMalloc(&c->tanWork, c->tan_work_doubles() * sizeof(double)) != hipSuccess) {
        tangent_free(c);
        return set_error(DMFE_ERR_NOMEM, "dmfe_tangent_init: work allocation failed");
    }
    HIP_TRY(hipMemsetAsync(c->tanWork, 0, c->tan_work_doubles() * sizeof(double), c->stream));
    launch_tan_init_n(c->model, c->tan_args(), c->N, c->stream);
    int rc = launch_check("tan_init_kernel");
    if (rc) return rc;
    HIP_TRY(hipStreamSynchronize(c->stream));
    return DMFE_OK;
}

extern "C" int dmfe_tangent_observables(dmfe_ctx* c, double t, double* dE, double* dQK0) {
    if (!c || c->tanK < 1) return set_error(DMFE_ERR_STATE, "dmfe_tangent_observables: no sensitivities carried");
    const TanArgs X = c->tan_args();
    launch_tan_energy(c->tables(), c->H, c->model, X, (int)(c->T1 - 1), t, c->stream);
    int rc = launch_check("tan_energy_kernel");
    if (rc) return rc;
    for (int k = 0; k < c->tanK; ++k) {
        double e = 0.0, q = 0.0;
        HIP_TRY(hipMemcpyAsync(&e, X.scal + k * 8 + 2, sizeof(double), hipMemcpyDeviceToHost, c->stream));
        HIP_TRY(hipMemcpyAsync(&q, &X.V[k][(size_t)(c->T1 - 1) * c->N].x, sizeof(double), hipMemcpyDeviceToHost,
                               c->stream));
        HIP_TRY(hipStreamSynchronize(c->stream));
        if (dE) dE[k] = e;
        if (dQK0) dQK0[k] = q;
    }
    return DMFE_OK;
}

extern "C" int dmfe_tangent_state(dmfe_ctx* c, int k, double* dQK, double* dQR, double* dr) {
    if (!c || k < 0 || k >= c->tanK) return set_error(DMFE_ERR_ARG, "dmfe_tangent_state: bad direction");
    const size_t n = (size_t)c->T1 * c->N;
    std::vector<double2> v(n);
    HIP_TRY(hipMemcpyAsync(v.data(), c->tanH[k].V, n * sizeof(double2), hipMemcpyDeviceToHost, c->stream));
    if (dr) HIP_TRY(hipMemcpyAsync(dr, c->tanH[k].r, (size_t)c->T1 * sizeof(double), hipMemcpyDeviceToHost, c->stream));
    HIP_TRY(hipStreamSynchronize(c->stream));
    for (size_t i = 0; i < n; ++i) {
        if (dQK) dQK[i] = v[i].x;
        if (dQR) dQR[i] = v[i].y;
    }
    return DMFE_OK;
}

extern "C" int dmfe_adjoint(dmfe_ctx* c, const double* seeds, double tmax, int64_t max_loop, double* dGamma,
                            double* dLambda, double* t_final, int64_t* steps) {
    if (!c || !seeds || !dGamma || !dLambda) return set_error(DMFE_ERR_ARG, "dmfe_adjoint: null argument");
    if (!std::isfinite(seeds[0]) || !std::isfinite(seeds[1]))
        return set_error(DMFE_ERR_ARG, "dmfe_adjoint: seeds must be finite");
    double g[4];
    const int rc = dmfe_gradients(c, tmax, max_loop, g, t_final, steps);
    if (rc) return rc;
    // seeds^T J with J = [[dE/dGamma, dE/dlambda], [dQK0/dGamma, dQK0/dlambda]]
    *dGamma = seeds[0] * g[0] + seeds[1] * g[2];
    *dLambda = seeds[0] * g[1] + seeds[1] * g[3];
    return DMFE_OK;
}

extern "C" int dmfe_gradients(dmfe_ctx* c, double tmax, int64_t max_loop, double* grad, double* t_final,
                              int64_t* steps) {
    if (!c || !grad) return set_error(DMFE_ERR_ARG, "dmfe_gradients: null argument");
    int rc = dmfe_init_quench(c);
    if (rc) return rc;
    const double g[2] = {1.0, 0.0}, l[2] = {0.0, 1.0};
    rc = dmfe_tangent_init(c, 2, g, l);
    if (rc) return rc;
    int64_t done = 0;
    double t = c->t1[c->T1 - 1];
    if (!c->dense_mfma && !c->profiling && !c->rollback && !c->use_serk2 && c->method < 3) {
        rc = run_device_loop(c, tmax, max_loop, nullptr, 0, &done, &t);
        if (rc) return rc;
    } else {
        const double eps = std::numeric_limits<double>::epsilon();
        while (t < tmax && c->loop < max_loop &&
               c->dt >= c->params.delta_t_min - eps * std::max(std::fabs(c->dt), std::fabs(c->params.delta_t_min))) {
            rc = loop_iteration(c, &t);
            if (rc) return rc;
            ++done;
        }
    }
    double dE[2], dQ[2];
    rc = dmfe_tangent_observables(c, t, dE, dQ);
    if (rc) return rc;
    grad[0] = dE[0]; grad[1] = dE[1]; grad[2] = dQ[0]; grad[3] = dQ[1];
    if (t_final) *t_final = t;
    if (steps) *steps = done;
    return DMFE_OK;
}

// n (Gamma, lambda) points at once: one context per point, each on its own stream and
// host thread (min(n, 16) workers), so the points' kernels overlap on the device.
extern "C" int dmfe_gradients_many(int n, const dmfe_grid* grids, const dmfe_params* params, double tmax,
                                   int64_t max_loop, int64_t rows_hint, double* grads, double* t_final,
                                   int64_t* steps) {
    if (n < 1 || !grids || !params || !grads) return set_error(DMFE_ERR_ARG, "dmfe_gradients_many: bad argument");
    int dev = 0;
    HIP_TRY(hipGetDevice(&dev));
    std::vector<int> rc((size_t)n, DMFE_OK);
    std::vector<std::string> msg((size_t)n);
    std::atomic<int> next{0};
    auto worker = [&]() {
        (void)hipSetDevice(dev);
        for (int i = next++; i < n; i = next++) {
            dmfe_ctx* c = nullptr;
            double tf = 0.0;
            int64_t st = 0;
            int r = dmfe_ctx_create(&grids[i], &params[i], rows_hint, &c);
            // no stream capture while other workers copy and allocate
            if (r == DMFE_OK) r = dmfe_set_loop_graphs(c, 0);
            if (r == DMFE_OK) r = dmfe_gradients(c, tmax, max_loop, grads + 4 * (size_t)i, &tf, &st);
            if (r != DMFE_OK) msg[(size_t)i] = dmfe_last_error();
            if (c) dmfe_ctx_destroy(c);
            rc[(size_t)i] = r;
            if (t_final) t_final[i] = tf;
            if (steps) steps[i] = st;
        }
    };
    const int nw = std::min(n, 16);
    std::vector<std::thread> pool;
    for (int w = 0; w < nw; ++w) pool.emplace_back(worker);
    for (auto& t : pool) t.join();
    for (int i = 0; i < n; ++i)
        if (rc[(size_t)i] != DMFE_OK)
            return set_error(rc[(size_t)i], "dmfe_gradients_many: point " + std::to_string(i) + ": " + msg[(size_t)i]);
    return DMFE_OK;
}

// data.bin version 1 (io_save_binary.cpp:66-100): int version, size_t T1, size_t N,
// int p, int p2, double lambda, T0, Gamma, delta, delta_t, int loop, double energy, then
// t1grid [T1], QK, QR, dQK, dQR [T1*N], rvec, drvec [T1]; raw, unpadded, native endian.
extern "C" int dmfe_save_state(dmfe_ctx* c, const char* path) {
    if (!c || !path || c->T1 < 1) return set_error(DMFE_ERR_ARG, "dmfe_save_state: no state or null path");
    const size_t N = (size_t)c->N, n = (size_t)c->T1;
    // energy at t1grid.back() on the last row, as the writer computes it
    double energy = 0.0, qk0 = 0.0;
    int rc = observables(c, c->t1[n - 1], &energy, &qk0);
    if (rc) return rc;
    std::vector<double> QK(n * N), QR(n * N), dQK(n * N), dQR(n * N), rv(n), drv(n);
    rc = dmfe_get_state(c, nullptr, QK.data(), QR.data(), dQK.data(), dQR.data(), rv.data(), drv.data(), nullptr);
    if (rc) return rc;
    FILE* f = std::fopen(path, "wb");
    if (!f) return set_error(DMFE_ERR_IO, std::string("dmfe_save_state: cannot open ") + path);
    const int version = 1;
    const size_t t1size = n, len = N;
    const int p = c->params.p, p2 = c->params.p2, loop = (int)c->loop;
    bool ok = std::fwrite(&version, sizeof(int), 1, f) == 1 && std::fwrite(&t1size, sizeof(size_t), 1, f) == 1 &&
              std::fwrite(&len, sizeof(size_t), 1, f) == 1 && std::fwrite(&p, sizeof(int), 1, f) == 1 &&
              std::fwrite(&p2, sizeof(int), 1, f) == 1 &&
              std::fwrite(&c->params.lambda, sizeof(double), 1, f) == 1 &&
              std::fwrite(&c->params.T0, sizeof(double), 1, f) == 1 &&
              std::fwrite(&c->params.Gamma, sizeof(double), 1, f) == 1 &&
              std::fwrite(&c->delta, sizeof(double), 1, f) == 1 && std::fwrite(&c->dt, sizeof(double), 1, f) == 1 &&
              std::fwrite(&loop, sizeof(int), 1, f) == 1 && std::fwrite(&energy, sizeof(double), 1, f) == 1 &&
              std::fwrite(c->t1.data(), sizeof(double), n, f) == n;
    for (const std::vector<double>* a : {&QK, &QR, &dQK, &dQR, &rv, &drv})
        ok = ok && std::fwrite(a->data(), sizeof(double), a->size(), f) == a->size();
    ok = (std::fclose(f) == 0) && ok;
    if (!ok) return set_error(DMFE_ERR_IO, std::string("dmfe_save_state: write failed: ") + path);
    return DMFE_OK;
}

// loadSimulationStateBinary (io_input.cpp:112-175) and the resume branch of init()
// (initialization.cpp:225-228,277-283): parameters must match (p, p2 exactly; lambda, T0,
// Gamma within 1e-10); delta, delta_t and loop come from the file, delta_old = 0, the
// integrator follows from delta_t, and the next step re-interpolates.
extern "C" int dmfe_load_state(dmfe_ctx* c, const char* path, double* energy_out) {
    if (!c || !path) return set_error(DMFE_ERR_ARG, "dmfe_load_state: null argument");
    FILE* f = std::fopen(path, "rb");
    if (!f) return set_error(DMFE_ERR_IO, std::string("dmfe_load_state: cannot open ") + path);
    int version = 0, p = 0, p2 = 0, loop = 0;
    size_t t1size = 0, len = 0;
    double lambda = 0, T0 = 0, Gamma = 0, delta = 0, dt = 0, energy = 0;
    bool ok = std::fread(&version, sizeof(int), 1, f) == 1 && std::fread(&t1size, sizeof(size_t), 1, f) == 1 &&
              std::fread(&len, sizeof(size_t), 1, f) == 1 && std::fread(&p, sizeof(int), 1, f) == 1 &&
              std::fread(&p2, sizeof(int), 1, f) == 1 && std::fread(&lambda, sizeof(double), 1, f) == 1 &&
              std::fread(&T0, sizeof(double), 1, f) == 1 && std::fread(&Gamma, sizeof(double), 1, f) == 1 &&
              std::fread(&delta, sizeof(double), 1, f) == 1 && std::fread(&dt, sizeof(double), 1, f) == 1 &&
              std::fread(&loop, sizeof(int), 1, f) == 1 && std::fread(&energy, sizeof(double), 1, f) == 1;
    if (!ok || version != 1) {
        std::fclose(f);
        return set_error(DMFE_ERR_IO, std::string("dmfe_load_state: not a version-1 data.bin: ") + path);
    }
    if (len != (size_t)c->N || t1size < 1 || t1size > ((size_t)1 << 31)) {
        std::fclose(f);
        return set_error(DMFE_ERR_STATE, "dmfe_load_state: file grid length " + std::to_string(len) + " != context N " +
                                             std::to_string(c->N));
    }
    if (p != c->params.p || p2 != c->params.p2 || std::fabs(lambda - c->params.lambda) > 1e-10 ||
        std::fabs(T0 - c->params.T0) > 1e-10 || std::fabs(Gamma - c->params.Gamma) > 1e-10) {
        std::fclose(f);
        return set_error(DMFE_ERR_STATE, "dmfe_load_state: file parameters don't match the context parameters");
    }
    const size_t n = t1size, N = len;
    std::vector<double> t1(n), QK(n * N), QR(n * N), dQK(n * N), dQR(n * N), rv(n), drv(n);
    for (std::vector<double>* a : {&t1, &QK, &QR, &dQK, &dQR, &rv, &drv})
        ok = ok && std::fread(a->data(), sizeof(double), a->size(), f) == a->size();
    std::fclose(f);
    if (!ok) return set_error(DMFE_ERR_IO, std::string("dmfe_load_state: truncated file: ") + path);
    int rc = dmfe_set_state(c, (int64_t)n, t1.data(), QK.data(), QR.data(), dQK.data(), dQR.data(), rv.data(), drv.data(),
                            nullptr);
    if (rc) return rc;
    c->delta = delta;
    c->dt = dt;
    c->delta_old = 0.0;
    c->loop = loop;
    c->method = (c->dt < c->params.rmax[0] / c->specRad) ? DMFE_RK54 : DMFE_SSPRK104;
    rc = run_search(c, kSearchB, nullptr, nullptr, nullptr, nullptr);
    if (rc) return rc;
    HIP_TRY(hipStreamSynchronize(c->stream));
    if (energy_out) *energy_out = energy;
    return DMFE_OK;
}

extern "C" int dmfe_time_rhs(dmfe_ctx* c, int reps, double* ms_out) {
    if (!c || reps < 1 || c->T1 < 1) return set_error(DMFE_ERR_ARG, "dmfe_time_rhs: bad argument");
    if (!c->interp_valid) return set_error(DMFE_ERR_STATE, "dmfe_time_rhs: no interpolation state");
    const bool was = c->profiling;
    c->profiling = false;
    hipEvent_t e0, e1;
    HIP_TRY(hipEventCreate(&e0));
    HIP_TRY(hipEventCreate(&e1));
    int rc = run_rhs_rows(c, nullptr);  // warm
    HIP_TRY(hipEventRecord(e0, c->stream));
    for (int r = 0; r < reps && rc == DMFE_OK; ++r) rc = run_rhs_rows(c, nullptr);
    HIP_TRY(hipEventRecord(e1, c->stream));
    hipError_t e = hipEventSynchronize(e1);
    float ms = 0.f;
    if (e == hipSuccess) e = hipEventElapsedTime(&ms, e0, e1);
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    c->profiling = was;
    if (rc) return rc;
    if (e != hipSuccess) return set_error(DMFE_ERR_HIP, std::string("dmfe_time_rhs: ") + hipGetErrorString(e));
    if (ms_out) *ms_out = ms / reps;
    return DMFE_OK;
}

// ---- RHS shapes and the autotuner ---------------------------------------------------------
static int check_shape(int64_t N, int m, const dmfe_rhs_shape* s, const char* who) {
    if (!s) return set_error(DMFE_ERR_ARG, std::string(who) + ": null shape");
    if (!rhs_shape_valid(s->G, m))
        return set_error(DMFE_ERR_ARG, std::string(who) + ": G must be 4 or 8 (m <= 16), 16 (m <= 32) or 64; got G=" +
                                           std::to_string(s->G) + " for m=" + std::to_string(m));
    if (s->split < 1 || s->split > rhs_max_split((int)N))
        return set_error(DMFE_ERR_ARG, std::string(who) + ": split must be in [1, " +
                                           std::to_string(rhs_max_split((int)N)) + "]");
    return DMFE_OK;
}

extern "C" int dmfe_get_rhs_shape(const dmfe_ctx* c, dmfe_rhs_shape* out) {
    if (!c || !out) return set_error(DMFE_ERR_ARG, "dmfe_get_rhs_shape: null argument");
    out->G = c->G;
    out->split = c->split;
    return DMFE_OK;
}

extern "C" int dmfe_set_rhs_shape(dmfe_ctx* c, const dmfe_rhs_shape* s) {
    if (!c) return set_error(DMFE_ERR_ARG, "dmfe_set_rhs_shape: null context");
    if (int rc = check_shape(c->N, c->m, s, "dmfe_set_rhs_shape")) return rc;
    if (c->tanK) return set_error(DMFE_ERR_STATE, "dmfe_set_rhs_shape: set the shape before dmfe_tangent_init");
    c->G = s->G;
    c->split = s->split;
    return DMFE_OK;
}

extern "C" int dmfe_tuning_get(int64_t N, int m, int log_response, dmfe_rhs_shape* out) {
    if (N < 1 || m < 1 || !out) return set_error(DMFE_ERR_ARG, "dmfe_tuning_get: bad argument");
    const RhsShape sh = rhs_shape_lookup((int)N, m, log_response);
    out->G = sh.G;
    out->split = sh.split;
    return DMFE_OK;
}

extern "C" int dmfe_tuning_set(int64_t N, int m, int log_response, const dmfe_rhs_shape* s) {
    if (N < 1 || m < 1) return set_error(DMFE_ERR_ARG, "dmfe_tuning_set: bad argument");
    if (int rc = check_shape(N, m, s, "dmfe_tuning_set")) return rc;
    rhs_shape_store((int)N, m, log_response, RhsShape{s->G, s->split});
    return DMFE_OK;
}

extern "C" int dmfe_autotune(dmfe_ctx* c, int reps, int install, dmfe_rhs_shape* best, dmfe_rhs_shape* tried,
                             double* tried_ms, int max_tried, int* n_tried) {
    if (!c || reps < 1 || max_tried < 0) return set_error(DMFE_ERR_ARG, "dmfe_autotune: bad argument");
    if (c->tanK) return set_error(DMFE_ERR_STATE, "dmfe_autotune: tune before dmfe_tangent_init");
    if (!c->interp_valid || c->T1 < 2 || !(c->t1[c->T1 - 1] > 0.0))
        return set_error(DMFE_ERR_STATE, "dmfe_autotune: needs an interpolated history with t > 0");
    std::vector<RhsShape> cand;
    for (int G : {4, 8, 16, 64})
        if (rhs_shape_valid(G, c->m))
            for (int sp : {rhs_default_split(c->N), rhs_max_split(c->N)}) cand.push_back(RhsShape{G, sp});
    const RhsShape keep{c->G, c->split};
    std::vector<double> ms(cand.size(), std::numeric_limits<double>::infinity());
    for (int round = 0; round < 2; ++round)  // best of two sweeps
        for (size_t q = 0; q < cand.size(); ++q) {
            c->G = cand[q].G;
            c->split = cand[q].split;
            double t = 0.0;
            const int rc = dmfe_time_rhs(c, reps, &t);
            if (rc) {
                c->G = keep.G;
                c->split = keep.split;
                return rc;
            }
            ms[q] = std::min(ms[q], t);
        }
    size_t qb = 0, qd = 0;
    for (size_t q = 1; q < cand.size(); ++q)
        if (ms[q] < ms[qb]) qb = q;
    // the default shape stays unless a candidate beats it by more than 2% (timing noise)
    for (size_t q = 0; q < cand.size(); ++q)
        if (cand[q].G == rhs_default_G(c->m) && cand[q].split == rhs_default_split(c->N)) qd = q;
    if (ms[qd] <= 1.02 * ms[qb]) qb = qd;
    c->G = cand[qb].G;
    c->split = cand[qb].split;
    if (install) rhs_shape_store(c->N, c->m, c->params.log_response, cand[qb]);
    if (best) {
        best->G = cand[qb].G;
        best->split = cand[qb].split;
    }
    const int nt = (int)std::min<size_t>(cand.size(), (size_t)max_tried);
    for (int q = 0; q < nt; ++q) {
        if (tried) {
            tried[q].G = cand[(size_t)q].G;
            tried[q].split = cand[(size_t)q].split;
        }
        if (tried_ms) tried_ms[q] = ms[(size_t)q];
    }
    if (n_tried) *n_tried = (int)cand.size();
    return DMFE_OK;
}

extern "C" int dmfe_dense_apply(dmfe_ctx* c, int map, int B, const double* X, double* Y, int reps, double* ms_out) {
    if (!c || !X || !Y || B < 1 || B > 16 || reps < 1 || map < 0 || map > 2)
        return set_error(DMFE_ERR_ARG, "dmfe_dense_apply: bad argument (map 0..2, 1 <= B <= 16, reps >= 1)");
    if (c->m != c->N) return set_error(DMFE_ERR_STATE, "dmfe_dense_apply: the grid is not dense (m != N)");
    if (map < 2 && !c->dense_mfma)
        return set_error(DMFE_ERR_STATE, "dmfe_dense_apply: A1/A2 tables are stencil-major in log-response mode");
    const size_t bx = (size_t)c->N * B * sizeof(double), by = (size_t)c->P * B * sizeof(double);
    double *dx = nullptr, *dy = nullptr;
    HIP_TRY(hipMalloc(&dx, bx));
    if (hipMalloc(&dy, by) != hipSuccess) {
        (void)hipFree(dx);
        return set_error(DMFE_ERR_NOMEM, "dmfe_dense_apply: output allocation failed");
    }
    hipEvent_t e0 = nullptr, e1 = nullptr;
    hipError_t e = hipMemcpyAsync(dx, X, bx, hipMemcpyHostToDevice, c->stream);
    if (e == hipSuccess) e = hipEventCreate(&e0);
    if (e == hipSuccess) e = hipEventCreate(&e1);
    if (e == hipSuccess) e = hipEventRecord(e0, c->stream);
    for (int r = 0; r < reps && e == hipSuccess; ++r) {
        launch_dense_apply(c->w[map], dx, dy, c->P, c->N, B, B, c->stream);
        e = hipGetLastError();
    }
    if (e == hipSuccess) e = hipEventRecord(e1, c->stream);
    if (e == hipSuccess) e = hipMemcpyAsync(Y, dy, by, hipMemcpyDeviceToHost, c->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(c->stream);
    float ms = 0.f;
    if (e == hipSuccess) e = hipEventElapsedTime(&ms, e0, e1);
    if (e0) (void)hipEventDestroy(e0);
    if (e1) (void)hipEventDestroy(e1);
    (void)hipFree(dx);
    (void)hipFree(dy);
    if (e != hipSuccess) return set_error(DMFE_ERR_HIP, std::string("dmfe_dense_apply: ") + hipGetErrorString(e));
    if (ms_out) *ms_out = ms / reps;
    return DMFE_OK;
}

extern "C" int dmfe_copy_probe(int64_t bytes, int reps, double* ms_out) {
    if (bytes < 16 || reps < 1) return set_error(DMFE_ERR_ARG, "dmfe_copy_probe: bad argument");
    const long long n = bytes / 16;
    double2 *a = nullptr, *b = nullptr;
    HIP_TRY(hipMalloc(&a, (size_t)n * 16));
    HIP_TRY(hipMalloc(&b, (size_t)n * 16));
    HIP_TRY(hipMemset(a, 0, (size_t)n * 16));
    hipEvent_t e0, e1;
    hipEventCreate(&e0);
    hipEventCreate(&e1);
    launch_hbm_copy(a, b, n, nullptr);
    hipEventRecord(e0, nullptr);
    for (int r = 0; r < reps; ++r) launch_hbm_copy(a, b, n, nullptr);
    hipEventRecord(e1, nullptr);
    hipError_t e = hipEventSynchronize(e1);
    float ms = 0.f;
    hipEventElapsedTime(&ms, e0, e1);
    hipEventDestroy(e0);
    hipEventDestroy(e1);
    hipFree(a);
    hipFree(b);
    if (e != hipSuccess) return set_error(DMFE_ERR_HIP, std::string("dmfe_copy_probe: ") + hipGetErrorString(e));
    if (ms_out) *ms_out = ms / reps;
    return DMFE_OK;
}
