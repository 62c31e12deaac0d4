// Batched quenches: many independent instances (own T, lambda, L, history, dt and
// integrator) advanced together, one launch per stage for all of them.  The
// per-instance scalar state of the reference's main loop (SimulationConfig /
// RKData) lives on the device (struct Ctl) and is updated by the begin / end
// kernels, so a batch step is a fixed launch sequence that is captured once into a
// hipGraph and replayed; the host only looks at the instances every few steps, to
// stop, to grow a history arena, or to run the every-1e5-loops maintenance
// (sparsify + interpolate, simulation_runner.cpp:161-233) that stays on the host.
#include "dmfe_internal.h"

#include <algorithm>
#include <cmath>
#include <cstring>
#include <limits>
#include <string>
#include <vector>

using namespace dmfe;

namespace {

#define HIP_TRY(expr)                                                                                     \
    do {                                                                                                  \
        hipError_t e_ = (expr);                                                                           \
        if (e_ != hipSuccess)                                                                             \
            return set_error(DMFE_ERR_HIP, std::string(#expr) + ": " + hipGetErrorString(e_));            \
    } while (0)

struct GridDev {
    const dmfe_grid* key = nullptr;
    int N = 0, m = 0;
    long long P = 0;
    double *theta = nullptr, *integ = nullptr, *phi2 = nullptr;
    int* inds[3] = {nullptr, nullptr, nullptr};
    double* w[3] = {nullptr, nullptr, nullptr};
    Tables tables() const {
        Tables T{};
        T.N = N; T.m = m; T.P = P;
        T.theta = theta; T.integ = integ; T.phi2 = phi2;
        for (int f = 0; f < 3; ++f) { T.inds[f] = inds[f]; T.w[f] = w[f]; }
        return T;
    }
};

bool same_grid(const dmfe_grid* a, const dmfe_grid* b) {
    if (a->N != b->N || a->m != b->m || a->theta != b->theta || a->integ != b->integ || a->phi2 != b->phi2)
        return false;
    for (int f = 0; f < 3; ++f)
        if (a->inds[f] != b->inds[f] || a->weights[f] != b->weights[f]) return false;
    return true;
}

void free_hist(History& h) {
    (void)hipFree(h.V); (void)hipFree(h.D); (void)hipFree(h.t1); (void)hipFree(h.rho); (void)hipFree(h.r);
    (void)hipFree(h.dr);
    h = History{};
}

int alloc_hist(History& h, int64_t rows, int N) {
    const size_t n = (size_t)rows;
    if (hipMalloc(&h.V, n * N * sizeof(double2)) != hipSuccess || hipMalloc(&h.D, n * N * sizeof(double2)) != hipSuccess ||
        hipMalloc(&h.t1, n * sizeof(double)) != hipSuccess || hipMalloc(&h.rho, n * sizeof(double)) != hipSuccess ||
        hipMalloc(&h.r, n * sizeof(double)) != hipSuccess || hipMalloc(&h.dr, n * sizeof(double)) != hipSuccess) {
        free_hist(h);
        return set_error(DMFE_ERR_NOMEM, "batch: history allocation of " + std::to_string(rows) + " rows failed");
    }
    return DMFE_OK;
}

struct Inst {
    int gi = 0;
    dmfe_params prm{};
    Model M{};
    History H{};
    Work W{};
    int split = 1;
    int gpu_variants = 0;  // the reference GPU path's interpolation (a context's device loop)
    int use_serk2 = 0;     // the maintenance promotes to SERK2 levels (a context's device loop)
    int64_t cap = 0;
    TanArgs X{};  // tangents (dmfe_batch_tangent_init, or a context's device loop); X.K = 0: none
    History tH[kMaxTan] = {};  // the tangent histories this batch owns (dmfe_batch_tangent_init)
    double* tWork = nullptr;   // their work arrays
};

}  // namespace

struct dmfe_batch {
    int device = 0;
    hipStream_t stream = nullptr;
    std::vector<GridDev> grids;
    std::vector<Inst> inst;
    std::vector<Ctl> ctl;  // host copy, refreshed at synchronisation points
    InstDesc* d_desc = nullptr;
    Ctl* d_ctl = nullptr;
    int2* d_smap = nullptr;
    int4* d_rmap = nullptr;
    int nb_stage = 0, nb_rhs = 0, ept = 1, G = 8, maxm = 1, log_response = 0;
    int tanK = 0;  // device loop of a context with tangents (launch shapes depend on it)
    size_t lds_rhs = 0;
    double* d_obs = nullptr;
    int obs_cap = 0;
    bool obs_on = false;
    // captured step sequences by stage count (graph_slot): 7 and 10 stages, and the SERK2
    // levels' 4, 8, ..., kMaxStepStages
    static constexpr int kSlots = 2 + kMaxStepStages / 4;
    hipGraphExec_t graph[kSlots] = {};
    int graph_steps[kSlots] = {};
    bool graph_obs[kSlots] = {};
    bool graph_s0[kSlots] = {};  // the graph holds the stage-0 search launch
    int64_t steps_run = 0;
    bool borrowed = false;  // device_loop_run: tables, instance buffers and stream belong to a context
    // device loop: the correlation record of the last rollback iteration, recomputed after
    // the rollback (t, E, QK(t,0)); rb_pending until device_loop_run stores it
    bool rb_pending = false;
    double rb_rec[3] = {0.0, 0.0, 0.0};
    InstDesc desc0{};       // device loop: instance 0's descriptor, passed by value to its kernels
};

namespace {

int n_inst(const dmfe_batch* b) { return (int)b->inst.size(); }

void drop_graphs(dmfe_batch* b);

int upload_desc(dmfe_batch* b) {
    std::vector<InstDesc> d(b->inst.size());
    for (size_t i = 0; i < b->inst.size(); ++i) {
        const Inst& x = b->inst[i];
        d[i].T = b->grids[(size_t)x.gi].tables();
        d[i].H = x.H;
        d[i].W = x.W;
        d[i].M = x.M;
        d[i].ctl = b->d_ctl + i;
        d[i].split = x.split;
        d[i].gpu_variants = x.gpu_variants;
        d[i].X = x.X;
    }
    HIP_TRY(hipMemcpyAsync(b->d_desc, d.data(), d.size() * sizeof(InstDesc), hipMemcpyHostToDevice, b->stream));
    HIP_TRY(hipStreamSynchronize(b->stream));
    if (b->borrowed && (std::memcmp(&b->desc0, &d[0], sizeof(InstDesc)) != 0)) {
        b->desc0 = d[0];
        drop_graphs(b);  // the device loop's captured launches carry the descriptor by value
    }
    return DMFE_OK;
}

int pull_ctl(dmfe_batch* b) {
    HIP_TRY(hipMemcpyAsync(b->ctl.data(), b->d_ctl, b->ctl.size() * sizeof(Ctl), hipMemcpyDeviceToHost, b->stream));
    HIP_TRY(hipStreamSynchronize(b->stream));
    return DMFE_OK;
}

int push_ctl(dmfe_batch* b) {
    // a SERK2 level's embedded weights travel with the control word
    for (Ctl& c : b->ctl)
        if (c.method >= 3) {
            const std::vector<double> w = serk2_coeffs(2 * (c.method - 2));
            for (size_t k = 0; k < w.size() && k <= (size_t)kMaxStepStages; ++k) c.serk_b[k] = w[k];
        }
    HIP_TRY(hipMemcpyAsync(b->d_ctl, b->ctl.data(), b->ctl.size() * sizeof(Ctl), hipMemcpyHostToDevice, b->stream));
    HIP_TRY(hipStreamSynchronize(b->stream));
    return DMFE_OK;
}

void drop_graphs(dmfe_batch* b) {
    for (int k = 0; k < dmfe_batch::kSlots; ++k) {
        if (b->graph[k]) (void)hipGraphExecDestroy(b->graph[k]);
        b->graph[k] = nullptr;
        b->graph_steps[k] = 0;
    }
}

// single-instance interpolation (search B) of instance i, outside a step
int interpolate_inst(dmfe_batch* b, int i) {
    Inst& x = b->inst[(size_t)i];
    Ctl& c = b->ctl[(size_t)i];
    double tl = 0.0;
    HIP_TRY(hipMemcpyAsync(&tl, x.H.t1 + (c.T1 - 1), sizeof(double), hipMemcpyDeviceToHost, b->stream));
    HIP_TRY(hipStreamSynchronize(b->stream));
    SearchArgs a{};
    a.mode = kSearchB;
    a.T1 = c.T1;
    a.t1 = x.H.t1;
    a.tl = tl;
    a.fbuf = 1 - c.fcur;
    launch_stage(b->grids[(size_t)x.gi].tables(), x.H, x.W, x.M, &a, nullptr, b->stream);
    HIP_TRY(hipGetLastError());
    c.fcur = a.fbuf;
    // a fresh interpolation: the next stage-0 RHS reads the current last row
    if (x.W.tf_pre)
        HIP_TRY(hipMemcpyAsync(x.W.tf_pre, x.H.V + (size_t)(c.T1 - 1) * b->grids[(size_t)x.gi].N,
                               (size_t)b->grids[(size_t)x.gi].N * sizeof(double2), hipMemcpyDeviceToDevice, b->stream));
    return DMFE_OK;
}

// The stage-0 launch holds only the search of the first step (T1 = 1: the bsearch of
// initialization); later steps reuse the previous step's last interpolation and there is
// no finish before stage 0, so the launch is left out once every running instance has
// two or more rows (T1 only grows between the host's looks at the instances).
bool needs_stage0(const dmfe_batch* b) {
    for (const Ctl& c : b->ctl)
        if (c.status == kRunning && c.T1 <= 1) return true;
    return false;
}

// One batch step: begin | (stage n, rhs n) for n < S | stage S | end [| observables].
// The device loop of a context (borrowed) issues step pos of `count` consecutive steps:
// a step's end, its observables and the next step's begin are one launch, so only the
// first step launches its begin and only the last its end.
int enqueue_step(dmfe_batch* b, int S, int pos = 0, int count = 1) {
    const int n = n_inst(b);
    const bool s0 = needs_stage0(b);
    if (b->borrowed) {  // one context: descriptor by value, no launch maps
        const InstDesc& d = b->desc0;
        if (pos == 0) launch_loop_begin(d, b->stream);
        for (int k = 0; k < S; ++k) {
            if (k > 0 || s0) launch_loop_stage(d, k, b->stream);
            // with tangents the next stage launch runs their finish of stage k before the
            // primal finish, as the host loop does
            launch_loop_rhs(d, k, b->log_response, b->G, b->lds_rhs, b->stream);
        }
        launch_loop_stage(d, S, b->stream);
        // the tail fit runs after update() and before the loop counter moves
        // (simulation_runner.cpp:135-141)
        if (d.W.tf) launch_loop_tail_fit(d, b->inst[0].prm.delta_max, b->stream);
        if (pos + 1 < count) {
            launch_loop_end_begin(d, b->obs_on ? b->d_obs : nullptr, b->obs_cap, b->stream);
        } else {
            launch_loop_end(d, b->stream);
            if (b->obs_on) launch_batch_obs(b->d_desc, n, b->d_obs, b->obs_cap, b->stream);
        }
        HIP_TRY(hipGetLastError());
        return DMFE_OK;
    }
    launch_batch_begin(b->d_desc, n, b->stream);
    const int K = b->tanK;  // tangent directions carried by every instance (0: none)
    for (int k = 0; k < S; ++k) {
        if (k > 0 || s0) {
            if (K) launch_batch_stage_tan(b->d_desc, b->d_smap, b->nb_stage, k, b->ept, K, b->stream);
            else launch_batch_stage(b->d_desc, b->d_smap, b->nb_stage, k, b->ept, b->stream);
        }
        if (K) launch_batch_rhs_tan(b->d_desc, b->d_rmap, b->nb_rhs, k, b->G, K, b->lds_rhs, b->stream);
        else launch_batch_rhs(b->d_desc, b->d_rmap, b->nb_rhs, k, b->log_response, b->G, b->maxm, b->lds_rhs, b->stream);
    }
    if (K) launch_batch_stage_tan(b->d_desc, b->d_smap, b->nb_stage, S, b->ept, K, b->stream);
    else launch_batch_stage(b->d_desc, b->d_smap, b->nb_stage, S, b->ept, b->stream);
    launch_batch_end(b->d_desc, n, b->stream);
    if (b->obs_on) launch_batch_obs(b->d_desc, n, b->d_obs, b->obs_cap, b->stream);
    HIP_TRY(hipGetLastError());
    return DMFE_OK;
}

int stages_of(int method) { return method >= 3 ? 4 * (method - 2) : (method == DMFE_RK54 ? 7 : 10); }

int max_stages(const dmfe_batch* b) {
    int S = 0;
    for (const Ctl& c : b->ctl)
        if (c.status == kRunning) S = std::max(S, stages_of(c.method));
    return S;
}

// graph of an S-stage step sequence: 7 -> 0, 10 -> 1, SERK2 S = 4k -> 1 + k
int graph_slot(int S) { return S == 7 ? 0 : (S == 10 ? 1 : 1 + S / 4); }

int build_graph(dmfe_batch* b, int which, int S, int steps) {
    const bool s0 = needs_stage0(b);
    if (b->graph[which] && b->graph_steps[which] == steps && b->graph_obs[which] == b->obs_on &&
        b->graph_s0[which] == s0)
        return DMFE_OK;
    if (b->graph[which]) (void)hipGraphExecDestroy(b->graph[which]);
    b->graph[which] = nullptr;
    hipGraph_t g = nullptr;
    // relaxed: other host threads (contexts driven concurrently) may allocate or copy
    // while this stream captures; the captured region itself holds kernel launches only
    HIP_TRY(hipStreamBeginCapture(b->stream, hipStreamCaptureModeRelaxed));
    int rc = DMFE_OK;
    for (int s = 0; s < steps && rc == DMFE_OK; ++s) rc = enqueue_step(b, S, s, steps);
    hipError_t e = hipStreamEndCapture(b->stream, &g);
    if (rc) { if (g) (void)hipGraphDestroy(g); return rc; }
    if (e != hipSuccess) return set_error(DMFE_ERR_HIP, std::string("graph capture: ") + hipGetErrorString(e));
    e = hipGraphInstantiate(&b->graph[which], g, nullptr, nullptr, 0);
    (void)hipGraphDestroy(g);
    if (e != hipSuccess) return set_error(DMFE_ERR_HIP, std::string("graph instantiate: ") + hipGetErrorString(e));
    b->graph_steps[which] = steps;
    b->graph_obs[which] = b->obs_on;
    b->graph_s0[which] = s0;
    return DMFE_OK;
}

// sparsifyNscale (CPU-only build semantics) on instance i; keeps the arena capacity
int sparsify_inst(dmfe_batch* b, int i, double threshold) {
    Inst& x = b->inst[(size_t)i];
    Ctl& c = b->ctl[(size_t)i];
    const int64_t T1 = c.T1;
    const GridDev& gd = b->grids[(size_t)x.gi];
    std::vector<int> keep{0};
    bool erased = false;
    if (T1 >= 4) {
        double* dval = nullptr;
        HIP_TRY(hipMalloc(&dval, (size_t)T1 * sizeof(double)));
        launch_sparsify_values(gd.tables(), x.H, (int)T1, dval, b->stream);
        std::vector<double> val((size_t)T1, 0.0);
        hipError_t e = hipMemcpyAsync(val.data(), dval, (size_t)T1 * sizeof(double), hipMemcpyDeviceToHost, b->stream);
        if (e == hipSuccess) e = hipStreamSynchronize(b->stream);
        (void)hipFree(dval);
        if (e != hipSuccess) return set_error(DMFE_ERR_HIP, std::string("batch sparsify: ") + hipGetErrorString(e));
        for (int64_t r = 2; r + 1 < T1; ++r) {
            if (val[(size_t)r] < threshold) erased = true;
            else keep.push_back((int)r);
        }
    }
    keep.push_back((int)(T1 - 1));
    if (!erased) return DMFE_OK;
    History h{};
    int rc = alloc_hist(h, x.cap, gd.N);
    if (rc) return rc;
    int* dkeep = nullptr;
    HIP_TRY(hipMalloc(&dkeep, keep.size() * sizeof(int)));
    HIP_TRY(hipMemcpyAsync(dkeep, keep.data(), keep.size() * sizeof(int), hipMemcpyHostToDevice, b->stream));
    launch_gather_rows(gd.tables(), x.H, h, dkeep, (int)keep.size(), b->stream);
    HIP_TRY(hipStreamSynchronize(b->stream));
    (void)hipFree(dkeep);
    free_hist(x.H);
    x.H = h;
    c.T1 = (int)keep.size();
    return DMFE_OK;
}

// The rollback of the GPU main loop (simulation_runner.cpp:280-303, rollbackState
// simulation_control.cpp:14-76) after an iteration whose error norm exceeded 2 delta_max:
// drop the last n = min(10, T1 - 2) rows, delta_t = the kept last step, loop -= n, a fresh
// interpolation; the iteration's correlation record is that of the truncated history at
// t1grid.back() + delta_t; then delta_t = 0.5 min(max(delta_t, dt_min), rmax / specRad)
// and the integrator restarts as SSPRK104 (SERK2 levels above the first step down one).
int rollback_inst(dmfe_batch* b, int i) {
    Inst& x = b->inst[(size_t)i];
    Ctl& c = b->ctl[(size_t)i];
    const dmfe_params& p = x.prm;
    const int N = b->grids[(size_t)x.gi].N;
    const int64_t n = std::min<int64_t>(10, std::max<int64_t>(0, (int64_t)c.T1 - 2));
    if (n > 0) {
        c.T1 -= (int)n;
        double tl[2] = {0.0, 0.0};
        HIP_TRY(hipMemcpyAsync(tl, x.H.t1 + (c.T1 - 2), 2 * sizeof(double), hipMemcpyDeviceToHost, b->stream));
        HIP_TRY(hipStreamSynchronize(b->stream));
        c.dt = tl[1] - tl[0];
        c.loop -= n;
        int rc = interpolate_inst(b, i);
        if (rc) return rc;
        c.last_rollback = c.loop;
        c.t_prop = tl[1] + c.dt;
        // the record of this iteration (simulation_runner.cpp:305-341 after the rollback)
        launch_row_scalar(b->grids[(size_t)x.gi].tables(), x.H, x.W, x.M, c.T1 - 1, c.t_prop, 1, b->stream);
        HIP_TRY(hipGetLastError());
        double e[2] = {0.0, 0.0};
        HIP_TRY(hipMemcpyAsync(e, x.W.scal + 2, sizeof(double), hipMemcpyDeviceToHost, b->stream));
        HIP_TRY(hipMemcpyAsync(e + 1, &x.H.V[(size_t)(c.T1 - 1) * N].x, sizeof(double), hipMemcpyDeviceToHost,
                               b->stream));
        HIP_TRY(hipStreamSynchronize(b->stream));
        b->rb_pending = true;
        b->rb_rec[0] = c.t_prop;
        b->rb_rec[1] = e[0];
        b->rb_rec[2] = e[1];
    }
    c.dt = 0.5 * std::min(std::max(c.dt, p.delta_t_min), p.rmax[c.method - 1] / c.specRad);
    c.method = c.method > 3 ? c.method - 1 : DMFE_SSPRK104;
    c.status = kRunning;
    return DMFE_OK;
}

// the rest of the main-loop body after loop % cadence == 0 (simulation_runner.cpp:161-278)
int maintenance(dmfe_batch* b, int i) {
    Inst& x = b->inst[(size_t)i];
    Ctl& c = b->ctl[(size_t)i];
    const double dmax = x.prm.delta_max;
    int64_t prev = c.T1;
    for (int count = 0; count < 1; ++count) {
        int rc = sparsify_inst(b, i, dmax);
        if (rc) return rc;
        if (c.T1 >= prev) break;
        prev = c.T1;
    }
    int rc = interpolate_inst(b, i);
    if (rc) return rc;
    // last_rollback_loop stays -1000 unless rollbacks are on (simulation_runner.cpp:53,220)
    if (c.delta < dmax / 2 && c.loop - c.last_rollback > 1000) {
        c.dt *= 0.5;
        // SSPRK104, or with SERK2 on (a context's device loop) one level up from SSPRK104
        // and SERK2 (init_SERK2(2 (init - 1)) sets init + 1, runge_kutta.cu:274), capped at
        // the rmax table
        if (!x.use_serk2 || c.method == DMFE_RK54) c.method = DMFE_SSPRK104;
        else if (c.method < kMaxMethod) c.method += 1;
    }
    double rho_last = 0.0;
    HIP_TRY(hipMemcpyAsync(&rho_last, x.H.rho + (c.T1 - 1), sizeof(double), hipMemcpyDeviceToHost, b->stream));
    HIP_TRY(hipStreamSynchronize(b->stream));
    const double* rmax = x.prm.rmax;
    if (c.delta < dmax && c.loop > 5 && (c.delta < 1.1 * c.delta_old || c.delta_old < dmax / 1000) &&
        rmax[c.method - 1] / c.specRad > c.dt && rho_last == 1.0) {
        c.dt *= 1.01;
    } else if (c.delta > 2 * dmax && c.dt > x.prm.delta_t_min) {
        c.dt *= 0.9;
    }
    if (c.method == DMFE_SSPRK104 && c.delta > dmax && rmax[0] / c.specRad > c.dt) {
        c.method = DMFE_RK54;
        c.dt *= 0.5;
    }
    c.status = kRunning;
    if (c.rollback && c.delta > 2 * dmax) return rollback_inst(b, i);
    return DMFE_OK;
}

int grow_inst(dmfe_batch* b, int i) {
    Inst& x = b->inst[(size_t)i];
    Ctl& c = b->ctl[(size_t)i];
    const int N = b->grids[(size_t)x.gi].N;
    const int64_t nc = std::max<int64_t>(2 * x.cap, c.T1 + 64);
    History h{};
    int rc = alloc_hist(h, nc, N);
    if (rc) return rc;
    const size_t n = (size_t)c.T1;
    HIP_TRY(hipMemcpyAsync(h.V, x.H.V, n * N * sizeof(double2), hipMemcpyDeviceToDevice, b->stream));
    HIP_TRY(hipMemcpyAsync(h.D, x.H.D, n * N * sizeof(double2), hipMemcpyDeviceToDevice, b->stream));
    HIP_TRY(hipMemcpyAsync(h.t1, x.H.t1, n * sizeof(double), hipMemcpyDeviceToDevice, b->stream));
    HIP_TRY(hipMemcpyAsync(h.rho, x.H.rho, n * sizeof(double), hipMemcpyDeviceToDevice, b->stream));
    HIP_TRY(hipMemcpyAsync(h.r, x.H.r, n * sizeof(double), hipMemcpyDeviceToDevice, b->stream));
    HIP_TRY(hipMemcpyAsync(h.dr, x.H.dr, n * sizeof(double), hipMemcpyDeviceToDevice, b->stream));
    HIP_TRY(hipStreamSynchronize(b->stream));
    free_hist(x.H);
    x.H = h;
    x.cap = nc;
    c.cap = (int)nc;
    c.status = kRunning;
    return DMFE_OK;
}

// host-side service of the instances that stopped for the host
int service(dmfe_batch* b, bool* changed) {
    *changed = false;
    for (int i = 0; i < n_inst(b); ++i) {
        Ctl& c = b->ctl[(size_t)i];
        int rc = DMFE_OK;
        if (b->inst[(size_t)i].X.K && (c.status == kNeedMaint || c.status == kNeedRows))
            return set_error(DMFE_ERR_STATE, "instance " + std::to_string(i) +
                                                 " with tangents: history arena full or maintenance point reached "
                                                 "(size the arena for the run; tangents do not cross a sparsification)");
        if (c.status == kNeedMaint) rc = maintenance(b, i);
        else if (c.status == kNeedRows) rc = grow_inst(b, i);
        else if (c.status == kNeedRollback) rc = rollback_inst(b, i);
        else continue;
        if (rc) return rc;
        *changed = true;
    }
    if (*changed) {
        int rc = upload_desc(b);
        if (rc) return rc;
        rc = push_ctl(b);
        if (rc) return rc;
        // captured graphs stay valid: kernels read the per-instance pointers from d_desc
    }
    return DMFE_OK;
}

void free_batch(dmfe_batch* b) {
    if (!b) return;
    if (b->stream) (void)hipStreamSynchronize(b->stream);
    drop_graphs(b);
    if (b->borrowed) {
        (void)hipFree(b->d_desc); (void)hipFree(b->d_ctl); (void)hipFree(b->d_smap); (void)hipFree(b->d_rmap);
        (void)hipFree(b->d_obs);
        delete b;
        return;
    }
    for (GridDev& g : b->grids) {
        (void)hipFree(g.theta); (void)hipFree(g.integ); (void)hipFree(g.phi2);
        for (int f = 0; f < 3; ++f) { (void)hipFree(g.inds[f]); (void)hipFree(g.w[f]); }
    }
    for (Inst& x : b->inst) {
        for (History& h : x.tH) free_hist(h);
        (void)hipFree(x.tWork);
        free_hist(x.H);
        Work& W = x.W;
        for (void* p : {(void*)W.posB1, (void*)W.posB2, (void*)W.saveB1, (void*)W.saveB2, (void*)W.Floor, (void*)W.diff,
                        (void*)W.rpk,
                        (void*)W.rInt, (void*)W.partial, (void*)W.qkb1c0, (void*)W.hK, (void*)W.hR, (void*)W.scal})
            (void)hipFree(p);
    }
    (void)hipFree(b->d_desc); (void)hipFree(b->d_ctl); (void)hipFree(b->d_smap); (void)hipFree(b->d_rmap);
    (void)hipFree(b->d_obs);
    if (b->stream) (void)hipStreamDestroy(b->stream);
    delete b;
}

int upload_grid(GridDev& gd, const dmfe_grid* g) {
    const int N = (int)g->N, m = (int)g->m;
    const long long P = (long long)N * N;
    gd.key = g; gd.N = N; gd.m = m; gd.P = P;
    const size_t bN = (size_t)N * sizeof(double), bP = (size_t)P * sizeof(double);
    HIP_TRY(hipMalloc(&gd.theta, bN));
    HIP_TRY(hipMalloc(&gd.integ, bN));
    HIP_TRY(hipMalloc(&gd.phi2, bP));
    HIP_TRY(hipMemcpy(gd.theta, g->theta, bN, hipMemcpyHostToDevice));
    HIP_TRY(hipMemcpy(gd.integ, g->integ, bN, hipMemcpyHostToDevice));
    HIP_TRY(hipMemcpy(gd.phi2, g->phi2, bP, hipMemcpyHostToDevice));
    std::vector<double> wt((size_t)P * m);
    std::vector<int> it((size_t)P);
    for (int f = 0; f < 3; ++f) {
        for (long long k = 0; k < P; ++k) {
            it[(size_t)k] = g->inds[f][k] < 0 ? 0 : (int)g->inds[f][k];
            if (f < 2)
                for (int d = 0; d < m; ++d) wt[(size_t)d * P + k] = g->weights[f][(size_t)k * m + d];
        }
        HIP_TRY(hipMalloc(&gd.inds[f], (size_t)P * sizeof(int)));
        HIP_TRY(hipMalloc(&gd.w[f], bP * (size_t)m));
        HIP_TRY(hipMemcpy(gd.inds[f], it.data(), (size_t)P * sizeof(int), hipMemcpyHostToDevice));
        HIP_TRY(hipMemcpy(gd.w[f], f < 2 ? wt.data() : g->weights[f], bP * (size_t)m, hipMemcpyHostToDevice));
    }
    return DMFE_OK;
}

int alloc_work(Work& W, int N, int split, hipStream_t stream) {
    const size_t bN = (size_t)N * sizeof(double), bP = (size_t)N * N * sizeof(double);
    HIP_TRY(hipMalloc(&W.posB1, 2 * bN));  // double-buffered (Work::posB1)
    HIP_TRY(hipMalloc(&W.posB2, 2 * bP));
    HIP_TRY(hipMalloc(&W.saveB1, 3 * bN));
    HIP_TRY(hipMalloc(&W.saveB2, 3 * bP));
    HIP_TRY(hipMalloc(&W.Floor, 2 * (size_t)N * sizeof(int)));
    HIP_TRY(hipMalloc(&W.diff, 2 * bN));
    HIP_TRY(hipMalloc(&W.rpk, 2 * (size_t)N * sizeof(double4)));
    HIP_TRY(hipMalloc(&W.rInt, bN));
    HIP_TRY(hipMalloc(&W.partial, bN * (size_t)split * kSums));
    HIP_TRY(hipMalloc(&W.qkb1c0, bN));
    HIP_TRY(hipMalloc(&W.hK, bN * kMaxStages));
    HIP_TRY(hipMalloc(&W.hR, bN * kMaxStages));
    HIP_TRY(hipMalloc(&W.scal, 8 * sizeof(double)));
    HIP_TRY(hipMemsetAsync(W.scal, 0, 8 * sizeof(double), stream));  // ordered with the batch's stream
    HIP_TRY(hipStreamSynchronize(stream));
    return DMFE_OK;
}

}  // namespace

// ============================== C ABI ==============================================

// ---- one context's main loop on the device-side control ---------------------------------
namespace dmfe {

namespace {
constexpr int kLoopChunk = 16;  // steps per graph replay of a context's device loop

// a borrowing one-instance batch for the context of q (tables, work, stream are q's)
int device_loop_open(const DeviceLoop& q, dmfe_batch** out) {
    const int N = q.T.N;
    dmfe_batch* b = new dmfe_batch();
    b->borrowed = true;
    b->stream = q.stream;
    (void)hipGetDevice(&b->device);
    if (int rc = batch_upload_tableaux(dp54(), ssprk104())) {
        delete b;
        return rc;
    }
    b->log_response = q.prm.log_response;
    GridDev gd;
    gd.N = N; gd.m = q.T.m; gd.P = q.T.P;
    // borrowed, never freed here (free_batch skips the tables of a borrowing batch)
    gd.theta = const_cast<double*>(q.T.theta);
    gd.integ = const_cast<double*>(q.T.integ);
    gd.phi2 = const_cast<double*>(q.T.phi2);
    for (int f = 0; f < 3; ++f) {
        gd.inds[f] = const_cast<int*>(q.T.inds[f]);
        gd.w[f] = const_cast<double*>(q.T.w[f]);
    }
    b->grids.push_back(gd);
    b->inst.resize(1);
    b->inst[0].split = q.split;
    b->G = q.G;
    b->maxm = q.T.m;
    b->tanK = q.X.K;
    b->lds_rhs = rhs_lds_bytes(N, rhs_chunk_cap(N, q.split), q.X.K);
    b->ept = (N + 255) / 256;
    std::vector<int2> smap{make_int2(0, -2), make_int2(0, -1)};
    for (int k = 0; k < N; ++k) smap.push_back(make_int2(0, k));
    std::vector<int4> rmap;
    for (int j = 0; j < N; ++j)
        for (int ch = 0; ch < q.split; ++ch) rmap.push_back(make_int4(0, j, ch, 0));
    while (rmap.size() % 8) rmap.push_back(make_int4(-1, 0, 0, 0));
    b->nb_stage = (int)smap.size();
    b->nb_rhs = (int)rmap.size();
    b->ctl.assign(1, Ctl{});
    if (hipMalloc(&b->d_desc, sizeof(InstDesc)) != hipSuccess || hipMalloc(&b->d_ctl, sizeof(Ctl)) != hipSuccess ||
        hipMalloc(&b->d_smap, smap.size() * sizeof(int2)) != hipSuccess ||
        hipMalloc(&b->d_rmap, rmap.size() * sizeof(int4)) != hipSuccess ||
        hipMalloc(&b->d_obs, (size_t)kLoopChunk * 4 * sizeof(double)) != hipSuccess) {
        free_batch(b);
        return set_error(DMFE_ERR_NOMEM, "dmfe_run: device loop allocation failed");
    }
    b->obs_cap = kLoopChunk;
    if (hipMemcpyAsync(b->d_smap, smap.data(), smap.size() * sizeof(int2), hipMemcpyHostToDevice, b->stream) != hipSuccess ||
        hipMemcpyAsync(b->d_rmap, rmap.data(), rmap.size() * sizeof(int4), hipMemcpyHostToDevice, b->stream) != hipSuccess) {
        free_batch(b);
        return set_error(DMFE_ERR_HIP, "dmfe_run: device loop map upload failed");
    }
    *out = b;
    return DMFE_OK;
}
}  // namespace

void device_loop_close(dmfe_batch* b) { free_batch(b); }

int device_loop_run(dmfe_batch** slot, DeviceLoop& q, double tmax, int64_t max_loop, double* corr, int64_t corr_cap,
                    int64_t* steps) {
    dmfe_batch* b = *slot;
    if (b && (b->G != q.G || b->inst[0].split != q.split || b->stream != q.stream || b->tanK != q.X.K)) {
        free_batch(b);  // launch shape changed: maps and captured graphs are stale
        *slot = b = nullptr;
    }
    if (!b) {
        int rc = device_loop_open(q, &b);
        if (rc) return rc;
        *slot = b;
    }
    Inst& x = b->inst[0];
    x.gi = 0; x.prm = q.prm; x.M = q.M; x.H = q.H; x.W = q.W; x.cap = q.cap; x.X = q.X;
    x.gpu_variants = q.gpu_variants;
    x.use_serk2 = q.use_serk2;
    const bool rec = corr && corr_cap > 0;
    if (rec != b->obs_on) b->obs_on = rec;  // build_graph re-captures when the recording flag changes
    Ctl& c = b->ctl[0];
    c = Ctl{};
    c.status = kRunning;
    c.cap = (int)q.cap;
    c.dt_min = q.prm.delta_t_min;
    c.delta_max = q.prm.delta_max;
    c.specRad = 4 * std::sqrt(host_fl(q.prm, 1, 2));
    std::memcpy(c.rmax, q.prm.rmax, sizeof c.rmax);
    c.tmax = tmax;
    c.max_loop = max_loop;
    c.maint_every = q.prm.maint_every > 0 ? q.prm.maint_every : 100000;
    c.T1 = (int)q.T1;
    c.method = q.method;
    c.fcur = q.fcur;
    c.loop = q.loop;
    c.dt = q.dt;
    c.delta = q.delta;
    c.delta_old = q.delta_old;
    c.t_prop = q.t_prop;
    c.active = 0;
    c.rollback = q.rollback ? 1 : 0;
    c.last_rollback = q.last_rollback;
    int rc = upload_desc(b);
    if (!rc) rc = push_ctl(b);
    // iterations run: the loop counter's advance, except that a rollback moves it back
    int64_t iters = 0;
    b->rb_pending = false;
    std::vector<double> ring(rec ? (size_t)kLoopChunk * 4 : 0);
    while (!rc) {
        const int S = max_stages(b);
        if (S == 0) break;
        const int which = graph_slot(S);
        const int64_t before = b->ctl[0].loop;
        // iterations the loop counter still allows: a last partial chunk runs as direct
        // launches, not as a replay whose surplus steps would each still launch every kernel
        const int chunk = (int)std::min<int64_t>(kLoopChunk, std::max<int64_t>(max_loop - before, 1));
        hipError_t e = hipSuccess;
        if (q.use_graph && chunk == kLoopChunk) {
            rc = build_graph(b, which, S, kLoopChunk);
            if (rc) break;
            e = hipGraphLaunch(b->graph[which], b->stream);
            if (e != hipSuccess) {
                rc = set_error(DMFE_ERR_HIP, std::string("dmfe_run: graph launch: ") + hipGetErrorString(e));
                break;
            }
        } else {
            for (int s = 0; s < chunk && rc == DMFE_OK; ++s) rc = enqueue_step(b, S, s, chunk);
            if (rc) break;
        }
        rc = pull_ctl(b);
        if (rc) break;
        if (rec) {
            e = hipMemcpyAsync(ring.data(), b->d_obs, ring.size() * sizeof(double), hipMemcpyDeviceToHost, b->stream);
            if (e == hipSuccess) e = hipStreamSynchronize(b->stream);
            if (e != hipSuccess) {
                rc = set_error(DMFE_ERR_HIP, std::string("dmfe_run: records: ") + hipGetErrorString(e));
                break;
            }
            // the ring is indexed by loop; within a replay the loop counter only grows (an
            // instance that needs a rollback stops for the rest of the replay)
            for (int64_t lp = before + 1; lp <= b->ctl[0].loop; ++lp) {
                const int64_t k = iters + (lp - before) - 1;  // row of corr: the iteration
                if (k >= corr_cap) break;
                const double* o = ring.data() + (size_t)((lp - 1) % kLoopChunk) * 4;
                corr[3 * k + 0] = o[0];
                corr[3 * k + 1] = o[1];
                corr[3 * k + 2] = o[2];
            }
        }
        iters += b->ctl[0].loop - before;
        bool changed = false;
        rc = service(b, &changed);
        if (b->rb_pending) {  // the rollback iteration's record, after the rollback
            b->rb_pending = false;
            if (rec && iters >= 1 && iters - 1 < corr_cap)
                for (int v = 0; v < 3; ++v) corr[3 * (iters - 1) + v] = b->rb_rec[v];
        }
    }
    // the caller adopts the (possibly replaced) arena and the control state
    q.H = x.H;
    q.cap = x.cap;
    const Ctl& f = b->ctl[0];
    q.T1 = f.T1; q.loop = f.loop; q.method = f.method; q.fcur = f.fcur;
    q.dt = f.dt; q.delta = f.delta; q.delta_old = f.delta_old;
    q.t_prop = f.t_prop;
    q.last_rollback = f.last_rollback;
    if (steps) *steps = iters;
    return rc;
}

}  // namespace dmfe

extern "C" int dmfe_batch_create(int n, const dmfe_grid* grids, const dmfe_params* params, int64_t rows_cap,
                                 dmfe_batch** out) {
    if (!grids || !params || !out || n < 1) return set_error(DMFE_ERR_ARG, "dmfe_batch_create: bad argument");
    *out = nullptr;
    for (int i = 0; i < n; ++i) {
        const dmfe_grid* g = grids + i;
        if (g->N < 4 || g->N > kMaxN || g->m < 1 || g->m > g->N || !g->theta || !g->integ || !g->phi2)
            return set_error(DMFE_ERR_ARG, "dmfe_batch_create: bad grid for instance " + std::to_string(i));
        for (int f = 0; f < 3; ++f)
            if (!g->inds[f] || !g->weights[f]) return set_error(DMFE_ERR_ARG, "dmfe_batch_create: missing stencil table");
        if (params[i].log_response != params[0].log_response)
            return set_error(DMFE_ERR_ARG, "dmfe_batch_create: log_response must be the same for all instances");
    }
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || ndev < 1) return set_error(DMFE_ERR_HIP, "dmfe_batch_create: no HIP device");
    dmfe_batch* b = new dmfe_batch();
    (void)hipGetDevice(&b->device);
    auto fail = [&](int rc) { std::string msg = dmfe_last_error(); free_batch(b); return set_error(rc, msg); };
    if (hipStreamCreateWithFlags(&b->stream, hipStreamNonBlocking) != hipSuccess)
        return fail(set_error(DMFE_ERR_HIP, "dmfe_batch_create: stream"));
    if (int rc = batch_upload_tableaux(dp54(), ssprk104())) return fail(rc);
    b->log_response = params[0].log_response;
    b->inst.resize((size_t)n);
    b->ctl.assign((size_t)n, Ctl{});
    const int64_t cap = std::max<int64_t>(rows_cap, 64);
    int maxm = 1;
    for (int i = 0; i < n; ++i) {
        const dmfe_grid* g = grids + i;
        int gi = -1;
        for (size_t k = 0; k < b->grids.size(); ++k)
            if (same_grid(b->grids[k].key, g)) gi = (int)k;
        if (gi < 0) {
            for (int f = 0; f < 3; ++f)
                for (long long k = 0; k < g->N * g->N; ++k) {
                    const int64_t s = g->inds[f][k] < 0 ? 0 : g->inds[f][k];
                    if (s + g->m > g->N) return fail(set_error(DMFE_ERR_ARG, "dmfe_batch_create: stencil exceeds N"));
                }
            b->grids.emplace_back();
            int rc = upload_grid(b->grids.back(), g);
            if (rc) return fail(rc);
            gi = (int)b->grids.size() - 1;
        }
        Inst& x = b->inst[(size_t)i];
        x.gi = gi;
        x.prm = params[i];
        x.M.lambda = params[i].lambda; x.M.T0 = params[i].T0; x.M.Gamma = params[i].Gamma;
        x.M.p = params[i].p; x.M.p2 = params[i].p2;
        const int N = (int)g->N;
        const RhsShape sh = rhs_shape_lookup(N, (int)g->m, params[0].log_response);
        x.split = sh.split;
        if (i == 0) b->G = sh.G;
        x.cap = cap;
        int rc = alloc_hist(x.H, cap, N);
        if (rc) return fail(rc);
        rc = alloc_work(x.W, N, rhs_max_split(N), b->stream);
        if (rc) return fail(rc);
        maxm = std::max(maxm, (int)g->m);
        b->lds_rhs = std::max(b->lds_rhs, rhs_lds_bytes(N, rhs_chunk_cap(N, x.split), 0));
        b->ept = std::max(b->ept, (N + 255) / 256);
    }
    // one lane-group width for the whole batch (its first instance's class); widened if
    // another instance's stencil needs it
    if (!rhs_shape_valid(b->G, maxm)) b->G = rhs_default_G(maxm);
    b->maxm = maxm;
    // stage map: finish blocks, theta blocks, then the phi2 rows of every instance
    std::vector<int2> smap;
    for (int i = 0; i < n; ++i) smap.push_back(make_int2(i, -2));
    for (int i = 0; i < n; ++i) smap.push_back(make_int2(i, -1));
    for (int i = 0; i < n; ++i)
        for (int k = 0; k < b->grids[(size_t)b->inst[(size_t)i].gi].N; ++k) smap.push_back(make_int2(i, k));
    // rhs map: for each grid, rows x chunks x the instances sharing that grid
    std::vector<int4> rmap;
    for (size_t gi = 0; gi < b->grids.size(); ++gi) {
        std::vector<int> members;
        for (int i = 0; i < n; ++i)
            if (b->inst[(size_t)i].gi == (int)gi) members.push_back(i);
        const int N = b->grids[gi].N;
        const int split = b->inst[(size_t)members[0]].split;
        for (int j = 0; j < N; ++j)
            for (int ch = 0; ch < split; ++ch)
                for (int i : members) rmap.push_back(make_int4(i, j, ch, 0));
    }
    while (rmap.size() % 8) rmap.push_back(make_int4(-1, 0, 0, 0));
    b->nb_stage = (int)smap.size();
    b->nb_rhs = (int)rmap.size();
    if (hipMalloc(&b->d_desc, (size_t)n * sizeof(InstDesc)) != hipSuccess ||
        hipMalloc(&b->d_ctl, (size_t)n * sizeof(Ctl)) != hipSuccess ||
        hipMalloc(&b->d_smap, smap.size() * sizeof(int2)) != hipSuccess ||
        hipMalloc(&b->d_rmap, rmap.size() * sizeof(int4)) != hipSuccess)
        return fail(set_error(DMFE_ERR_NOMEM, "dmfe_batch_create: allocation failed"));
    if (hipMemcpy(b->d_smap, smap.data(), smap.size() * sizeof(int2), hipMemcpyHostToDevice) != hipSuccess ||
        hipMemcpy(b->d_rmap, rmap.data(), rmap.size() * sizeof(int4), hipMemcpyHostToDevice) != hipSuccess)
        return fail(set_error(DMFE_ERR_HIP, "dmfe_batch_create: map upload failed"));
    for (int i = 0; i < n; ++i) {
        Ctl& c = b->ctl[(size_t)i];
        const dmfe_params& p = params[i];
        c.status = kDone;  // until dmfe_batch_init_quench
        c.cap = (int)cap;
        c.dt_min = p.delta_t_min;
        c.delta_max = p.delta_max;
        c.specRad = 4 * std::sqrt(host_fl(p, 1, 2));
        std::memcpy(c.rmax, p.rmax, sizeof c.rmax);
        c.tmax = std::numeric_limits<double>::infinity();
        c.max_loop = std::numeric_limits<long long>::max();
        c.maint_every = p.maint_every > 0 ? p.maint_every : 100000;
    }
    int rc = upload_desc(b);
    if (rc) return fail(rc);
    rc = push_ctl(b);
    if (rc) return fail(rc);
    *out = b;
    return DMFE_OK;
}

extern "C" int dmfe_batch_destroy(dmfe_batch* b) {
    free_batch(b);
    return DMFE_OK;
}

extern "C" int dmfe_batch_size(const dmfe_batch* b, int* n) {
    if (!b || !n) return set_error(DMFE_ERR_ARG, "null argument");
    *n = n_inst(b);
    return DMFE_OK;
}

// initialization.cpp:204-223, 277-283 for every instance
extern "C" int dmfe_batch_init_quench(dmfe_batch* b) {
    if (!b) return set_error(DMFE_ERR_ARG, "null batch");
    drop_graphs(b);
    if (b->tanK) {  // sensitivities start again with dmfe_batch_tangent_init
        for (Inst& x : b->inst) {
            for (History& h : x.tH) free_hist(h);
            (void)hipFree(x.tWork);
            x.tWork = nullptr;
            x.X = TanArgs{};
        }
        b->tanK = 0;
        int rc = upload_desc(b);
        if (rc) return rc;
    }
    for (int i = 0; i < n_inst(b); ++i) {
        Inst& x = b->inst[(size_t)i];
        Ctl& c = b->ctl[(size_t)i];
        const GridDev& gd = b->grids[(size_t)x.gi];
        const int N = gd.N;
        std::vector<double2> v((size_t)N, make_double2(1.0, 1.0)), d((size_t)N, make_double2(0.0, 0.0));
        const double zero = 0.0;
        const double r0 = x.prm.Gamma + host_fl(x.prm, 1, 1) / x.prm.T0;
        HIP_TRY(hipMemcpyAsync(x.H.V, v.data(), N * sizeof(double2), hipMemcpyHostToDevice, b->stream));
        HIP_TRY(hipMemcpyAsync(x.H.D, d.data(), N * sizeof(double2), hipMemcpyHostToDevice, b->stream));
        HIP_TRY(hipMemcpyAsync(x.H.t1, &zero, sizeof(double), hipMemcpyHostToDevice, b->stream));
        HIP_TRY(hipMemcpyAsync(x.H.rho, &zero, sizeof(double), hipMemcpyHostToDevice, b->stream));
        HIP_TRY(hipMemcpyAsync(x.H.r, &r0, sizeof(double), hipMemcpyHostToDevice, b->stream));
        launch_row_scalar(gd.tables(), x.H, x.W, x.M, 0, 0.0, 0, b->stream);
        HIP_TRY(hipMemcpyAsync(x.H.dr, x.W.scal + 2, sizeof(double), hipMemcpyDeviceToDevice, b->stream));
        HIP_TRY(hipStreamSynchronize(b->stream));
        c.T1 = 1;
        c.dt = x.prm.delta_t_min;
        c.loop = 0;
        c.delta = 1;
        c.delta_old = 0;
        c.t_prop = 0.0;
        c.fcur = 0;
        c.method = (c.dt < x.prm.rmax[0] / c.specRad) ? DMFE_RK54 : DMFE_SSPRK104;
        c.status = kRunning;
        c.active = 0;
        // first interpolation (initialization.cpp:283)
        int rc = interpolate_inst(b, i);
        if (rc) return rc;
    }
    HIP_TRY(hipStreamSynchronize(b->stream));
    b->steps_run = 0;
    return push_ctl(b);
}

// Forward-mode sensitivities for every instance (the batch form of dmfe_tangent_init):
// K directions (dGamma[k], dLambda[k]), the same for all instances, carried through the
// batch's stage and RHS launches.  Right after dmfe_batch_init_quench; the tangent
// histories take the batch's arena size, which must hold the whole run.
extern "C" int dmfe_batch_tangent_init(dmfe_batch* b, int K, const double* dGamma, const double* dLambda) {
    if (!b || K < 1 || K > kMaxTan || !dGamma || !dLambda)
        return set_error(DMFE_ERR_ARG, "dmfe_batch_tangent_init: bad argument (1 <= K <= 2)");
    if (b->log_response)
        return set_error(DMFE_ERR_STATE, "dmfe_batch_tangent_init: the log-response interpolation is not differentiated");
    int rc = pull_ctl(b);
    if (rc) return rc;
    for (const Ctl& c : b->ctl)
        if (c.T1 != 1 || c.status != kRunning)
            return set_error(DMFE_ERR_STATE, "dmfe_batch_tangent_init: call right after dmfe_batch_init_quench");
    drop_graphs(b);
    b->lds_rhs = 0;
    for (int i = 0; i < n_inst(b); ++i) {
        Inst& x = b->inst[(size_t)i];
        const int N = b->grids[(size_t)x.gi].N;
        for (History& h : x.tH) free_hist(h);
        (void)hipFree(x.tWork);
        x.tWork = nullptr;
        x.X = TanArgs{};
        for (int k = 0; k < K; ++k) {
            rc = alloc_hist(x.tH[k], x.cap, N);
            if (rc) return rc;
        }
        // partial [K][N][split][8] | qkb1c0 [K][N] | hK, hR [K][10][N] | scal [K][8]
        const size_t nd = (size_t)K * ((size_t)N * x.split * kSums + N + 2 * (size_t)kMaxStages * N + 8);
        if (hipMalloc(&x.tWork, nd * sizeof(double)) != hipSuccess)
            return set_error(DMFE_ERR_NOMEM, "dmfe_batch_tangent_init: work allocation failed");
        HIP_TRY(hipMemsetAsync(x.tWork, 0, nd * sizeof(double), b->stream));
        TanArgs& X = x.X;
        X.K = K;
        for (int k = 0; k < K; ++k) {
            X.dG[k] = dGamma[k];
            X.dL[k] = dLambda[k];
            X.V[k] = x.tH[k].V;
            X.D[k] = x.tH[k].D;
            X.r[k] = x.tH[k].r;
            X.dr[k] = x.tH[k].dr;
        }
        double* p = x.tWork;
        X.partial = p; p += (size_t)K * N * x.split * kSums;
        X.qkb1c0 = p; p += (size_t)K * N;
        X.hK = p; p += (size_t)K * kMaxStages * N;
        X.hR = p; p += (size_t)K * kMaxStages * N;
        X.scal = p;
        launch_tan_init_n(x.M, X, N, b->stream);
        HIP_TRY(hipGetLastError());
        b->lds_rhs = std::max(b->lds_rhs, rhs_lds_bytes(N, rhs_chunk_cap(N, x.split), K));
    }
    b->tanK = K;
    HIP_TRY(hipStreamSynchronize(b->stream));
    return upload_desc(b);
}

// d(E(t), QK(t,0)) along each of the K directions for every instance, at the proposed t
// of its last step (the time of its correlation line, simulation_runner.cpp:327-341, and
// of dmfe_gradients; t1grid.back() before the first step): dE[i * K + k], dQK0[i * K + k]
// (the batch form of dmfe_tangent_observables).
extern "C" int dmfe_batch_tangent_observables(dmfe_batch* b, double* dE, double* dQK0) {
    if (!b) return set_error(DMFE_ERR_ARG, "dmfe_batch_tangent_observables: null batch");
    if (b->tanK < 1) return set_error(DMFE_ERR_STATE, "dmfe_batch_tangent_observables: no sensitivities carried");
    int rc = pull_ctl(b);
    if (rc) return rc;
    const int K = b->tanK;
    for (int i = 0; i < n_inst(b); ++i) {
        const Inst& x = b->inst[(size_t)i];
        const GridDev& gd = b->grids[(size_t)x.gi];
        const int T1 = b->ctl[(size_t)i].T1;
        double t = b->ctl[(size_t)i].t_prop;  // the proposed t of the last step, as dmfe_gradients
        if (b->ctl[(size_t)i].loop == 0) {     // no step yet: t1grid.back()
            HIP_TRY(hipMemcpyAsync(&t, x.H.t1 + (T1 - 1), sizeof(double), hipMemcpyDeviceToHost, b->stream));
            HIP_TRY(hipStreamSynchronize(b->stream));
        }
        launch_tan_energy(gd.tables(), x.H, x.M, x.X, T1 - 1, t, b->stream);
        HIP_TRY(hipGetLastError());
        for (int k = 0; k < K; ++k) {
            double e = 0.0, q = 0.0;
            HIP_TRY(hipMemcpyAsync(&e, x.X.scal + k * 8 + 2, sizeof(double), hipMemcpyDeviceToHost, b->stream));
            HIP_TRY(hipMemcpyAsync(&q, &x.X.V[k][(size_t)(T1 - 1) * gd.N].x, sizeof(double), hipMemcpyDeviceToHost,
                                   b->stream));
            HIP_TRY(hipStreamSynchronize(b->stream));
            if (dE) dE[(size_t)i * K + k] = e;
            if (dQK0) dQK0[(size_t)i * K + k] = q;
        }
    }
    return DMFE_OK;
}

// One batch step for every running instance (direct launches, then a synchronisation).
extern "C" int dmfe_batch_step(dmfe_batch* b, int* running) {
    if (!b) return set_error(DMFE_ERR_ARG, "null batch");
    const int S = max_stages(b);
    if (S > 0) {
        int rc = enqueue_step(b, S);
        if (rc) return rc;
        ++b->steps_run;
    }
    int rc = pull_ctl(b);
    if (rc) return rc;
    bool changed = false;
    rc = service(b, &changed);
    if (rc) return rc;
    int nr = 0;
    for (const Ctl& c : b->ctl) nr += c.status == kRunning;
    if (running) *running = nr;
    return DMFE_OK;
}

extern "C" int dmfe_batch_set_limits(dmfe_batch* b, double tmax, int64_t max_loop) {
    if (!b) return set_error(DMFE_ERR_ARG, "null batch");
    int rc = pull_ctl(b);
    if (rc) return rc;
    for (Ctl& c : b->ctl) {
        c.tmax = tmax;
        c.max_loop = max_loop;
        if (c.status == kDone && c.T1 >= 1) {
            // re-test the loop condition at the next begin
            const double eps = std::numeric_limits<double>::epsilon();
            if (c.t_prop < tmax && c.loop < max_loop && c.dt >= c.dt_min - eps * std::max(std::fabs(c.dt), std::fabs(c.dt_min)))
                c.status = kRunning;
        }
    }
    return push_ctl(b);
}

// Runs batch steps until every instance stops (simulation_runner.cpp:125 per instance) or
// max_steps batch steps ran.  use_graph: replay a captured hipGraph of `chunk` steps.
extern "C" int dmfe_batch_run(dmfe_batch* b, int64_t max_steps, int use_graph, int chunk, int64_t* steps_out) {
    if (!b || max_steps < 0) return set_error(DMFE_ERR_ARG, "dmfe_batch_run: bad argument");
    chunk = std::max(1, chunk);
    int64_t done = 0;
    int rc = pull_ctl(b);
    if (rc) return rc;
    while (done < max_steps) {
        const int S = max_stages(b);
        if (S == 0) break;
        const int steps = (int)std::min<int64_t>(chunk, max_steps - done);
        if (use_graph && steps == chunk) {
            const int which = graph_slot(S);
            rc = build_graph(b, which, S, steps);
            if (rc) return rc;
            HIP_TRY(hipGraphLaunch(b->graph[which], b->stream));
        } else {
            for (int s = 0; s < steps && rc == DMFE_OK; ++s) rc = enqueue_step(b, S);
            if (rc) return rc;
        }
        done += steps;
        rc = pull_ctl(b);
        if (rc) return rc;
        bool changed = false;
        rc = service(b, &changed);
        if (rc) return rc;
    }
    b->steps_run += done;
    if (steps_out) *steps_out = done;
    return DMFE_OK;
}

extern "C" int dmfe_batch_get_lanes(const dmfe_batch* b, int* G) {
    if (!b || !G) return set_error(DMFE_ERR_ARG, "dmfe_batch_get_lanes: null argument");
    *G = b->G;
    return DMFE_OK;
}

extern "C" int dmfe_batch_set_lanes(dmfe_batch* b, int G) {
    if (!b) return set_error(DMFE_ERR_ARG, "dmfe_batch_set_lanes: null batch");
    if (!rhs_shape_valid(G, b->maxm))
        return set_error(DMFE_ERR_ARG, "dmfe_batch_set_lanes: G must be 4 or 8 (m <= 16), 16 (m <= 32) or 64; got " +
                                           std::to_string(G) + " for m = " + std::to_string(b->maxm));
    if (G != b->G) {
        b->G = G;
        drop_graphs(b);  // captured launches carry the old kernel
    }
    return DMFE_OK;
}

// Times `reps` launches of the batch RHS (stage 0 of the current step arguments, all
// instances) for every valid lane-group width G, best of two sweeps, and keeps the
// fastest.  Only the RHS scratch (quadrature partials) is written; the next step's RHS
// rewrites it before use, so the batch state is unchanged.
extern "C" int dmfe_batch_autotune(dmfe_batch* b, int reps, int* G_best, int32_t* tried_G, double* tried_ms,
                                   int max_tried, int* n_tried) {
    if (!b || reps < 1 || max_tried < 0) return set_error(DMFE_ERR_ARG, "dmfe_batch_autotune: bad argument");
    if (b->steps_run < 1) return set_error(DMFE_ERR_STATE, "dmfe_batch_autotune: run at least one batch step first");
    std::vector<int> cand;
    for (int G : {4, 8, 16, 64})
        if (rhs_shape_valid(G, b->maxm)) cand.push_back(G);
    std::vector<double> ms(cand.size(), std::numeric_limits<double>::infinity());
    hipEvent_t e0 = nullptr, e1 = nullptr;
    HIP_TRY(hipEventCreate(&e0));
    if (hipEventCreate(&e1) != hipSuccess) {
        (void)hipEventDestroy(e0);
        return set_error(DMFE_ERR_HIP, "dmfe_batch_autotune: event create failed");
    }
    hipError_t e = hipSuccess;
    for (int round = 0; round < 2 && e == hipSuccess; ++round)
        for (size_t q = 0; q < cand.size() && e == hipSuccess; ++q) {
            launch_batch_rhs(b->d_desc, b->d_rmap, b->nb_rhs, 0, b->log_response, cand[q], b->maxm, b->lds_rhs, b->stream);
            e = hipEventRecord(e0, b->stream);
            for (int r = 0; r < reps && e == hipSuccess; ++r) {
                launch_batch_rhs(b->d_desc, b->d_rmap, b->nb_rhs, 0, b->log_response, cand[q], b->maxm, b->lds_rhs,
                                 b->stream);
                e = hipGetLastError();
            }
            if (e == hipSuccess) e = hipEventRecord(e1, b->stream);
            if (e == hipSuccess) e = hipEventSynchronize(e1);
            float t = 0.f;
            if (e == hipSuccess) e = hipEventElapsedTime(&t, e0, e1);
            if (e == hipSuccess) ms[q] = std::min(ms[q], (double)t / reps);
        }
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    if (e != hipSuccess) return set_error(DMFE_ERR_HIP, std::string("dmfe_batch_autotune: ") + hipGetErrorString(e));
    size_t qb = 0, qd = 0;
    for (size_t q = 1; q < cand.size(); ++q)
        if (ms[q] < ms[qb]) qb = q;
    // the default width stays unless a candidate beats it by more than 2% (timing noise)
    for (size_t q = 0; q < cand.size(); ++q)
        if (cand[q] == rhs_default_G(b->maxm)) qd = q;
    if (ms[qd] <= 1.02 * ms[qb]) qb = qd;
    int rc = dmfe_batch_set_lanes(b, cand[qb]);
    if (rc) return rc;
    if (G_best) *G_best = cand[qb];
    const int nt = (int)std::min<size_t>(cand.size(), (size_t)max_tried);
    for (int q = 0; q < nt; ++q) {
        if (tried_G) tried_G[q] = cand[(size_t)q];
        if (tried_ms) tried_ms[q] = ms[(size_t)q];
    }
    if (n_tried) *n_tried = (int)cand.size();
    return DMFE_OK;
}

extern "C" int dmfe_batch_get_control(dmfe_batch* b, int i, double* delta_t, double* delta, double* delta_old,
                                      int64_t* loop, int* method, int* status) {
    if (!b || i < 0 || i >= n_inst(b)) return set_error(DMFE_ERR_ARG, "dmfe_batch_get_control: bad argument");
    int rc = pull_ctl(b);
    if (rc) return rc;
    const Ctl& c = b->ctl[(size_t)i];
    if (delta_t) *delta_t = c.dt;
    if (delta) *delta = c.delta;
    if (delta_old) *delta_old = c.delta_old;
    if (loop) *loop = c.loop;
    if (method) *method = c.method;
    if (status) *status = c.status;
    return DMFE_OK;
}

extern "C" int dmfe_batch_set_control(dmfe_batch* b, int i, double delta_t, double delta, double delta_old, int64_t loop,
                                      int method) {
    if (!b || i < 0 || i >= n_inst(b) || (method != DMFE_RK54 && method != DMFE_SSPRK104))
        return set_error(DMFE_ERR_ARG, "dmfe_batch_set_control: bad argument");
    if (!(delta_t > 0.0) || !std::isfinite(delta_t) || !std::isfinite(delta) || !std::isfinite(delta_old) || loop < 0)
        return set_error(DMFE_ERR_ARG, "dmfe_batch_set_control: delta_t must be finite and > 0, delta and delta_old "
                                       "finite, loop >= 0");
    int rc = pull_ctl(b);
    if (rc) return rc;
    Ctl& c = b->ctl[(size_t)i];
    if (c.T1 < 1) return set_error(DMFE_ERR_STATE, "dmfe_batch_set_control: instance has no state");
    c.dt = delta_t;
    c.delta = delta;
    c.delta_old = delta_old;
    c.loop = loop;
    c.method = method;
    return push_ctl(b);
}

extern "C" int dmfe_batch_get_T1(dmfe_batch* b, int i, int64_t* T1) {
    if (!b || !T1 || i < 0 || i >= n_inst(b)) return set_error(DMFE_ERR_ARG, "dmfe_batch_get_T1: bad argument");
    int rc = pull_ctl(b);
    if (rc) return rc;
    *T1 = b->ctl[(size_t)i].T1;
    return DMFE_OK;
}

extern "C" int dmfe_batch_get_state(dmfe_batch* b, int i, double* t1grid, double* QK, double* QR, double* dQK,
                                    double* dQR, double* rvec, double* drvec, double* dtratio) {
    if (!b || i < 0 || i >= n_inst(b)) return set_error(DMFE_ERR_ARG, "dmfe_batch_get_state: bad argument");
    int rc = pull_ctl(b);
    if (rc) return rc;
    const Inst& x = b->inst[(size_t)i];
    const size_t N = (size_t)b->grids[(size_t)x.gi].N, n = (size_t)b->ctl[(size_t)i].T1;
    auto get = [&](double* dst, const double* src) -> int {
        if (!dst) return DMFE_OK;
        HIP_TRY(hipMemcpyAsync(dst, src, n * sizeof(double), hipMemcpyDeviceToHost, b->stream));
        return DMFE_OK;
    };
    if ((rc = get(t1grid, x.H.t1)) || (rc = get(rvec, x.H.r)) || (rc = get(drvec, x.H.dr)) || (rc = get(dtratio, x.H.rho)))
        return rc;
    if (QK || QR || dQK || dQR) {
        std::vector<double2> v(n * N), d(n * N);
        HIP_TRY(hipMemcpyAsync(v.data(), x.H.V, n * N * sizeof(double2), hipMemcpyDeviceToHost, b->stream));
        HIP_TRY(hipMemcpyAsync(d.data(), x.H.D, n * N * sizeof(double2), hipMemcpyDeviceToHost, b->stream));
        HIP_TRY(hipStreamSynchronize(b->stream));
        for (size_t k = 0; k < n * N; ++k) {
            if (QK) QK[k] = v[k].x;
            if (QR) QR[k] = v[k].y;
            if (dQK) dQK[k] = d[k].x;
            if (dQR) dQR[k] = d[k].y;
        }
    }
    HIP_TRY(hipStreamSynchronize(b->stream));
    return DMFE_OK;
}

// History of instance i from host arrays (the dmfe_set_state layout), then a fresh
// binary-search interpolation on it (interpolate() with no positions, as after
// loadSimulationStateBinary, initialization.cpp:277-283).  The arena grows when T1 + 1
// rows do not fit.  The instance runs again (status RUNNING); its control scalars stay
// as they are (dmfe_batch_set_control sets them).
extern "C" int dmfe_batch_set_state(dmfe_batch* b, int i, int64_t T1, const double* t1grid, const double* QK,
                                    const double* QR, const double* dQK, const double* dQR, const double* rvec,
                                    const double* drvec, const double* dtratio) {
    if (!b || i < 0 || i >= n_inst(b) || T1 < 1 || !t1grid || !QK || !QR || !dQK || !dQR || !rvec || !drvec)
        return set_error(DMFE_ERR_ARG, "dmfe_batch_set_state: bad argument");
    if (T1 > std::numeric_limits<int>::max() / 2)
        return set_error(DMFE_ERR_ARG, "dmfe_batch_set_state: T1 too large");
    for (int64_t k = 1; k < T1; ++k)
        if (!(t1grid[k] > t1grid[k - 1]))
            return set_error(DMFE_ERR_ARG, "dmfe_batch_set_state: t1grid must be strictly increasing");
    int rc = pull_ctl(b);
    if (rc) return rc;
    Inst& x = b->inst[(size_t)i];
    Ctl& c = b->ctl[(size_t)i];
    const size_t N = (size_t)b->grids[(size_t)x.gi].N, n = (size_t)T1;
    if (T1 + 1 > x.cap) {
        History h{};
        const int64_t nc = std::max<int64_t>(2 * x.cap, T1 + 64);
        rc = alloc_hist(h, nc, (int)N);
        if (rc) return rc;
        free_hist(x.H);
        x.H = h;
        x.cap = nc;
        c.cap = (int)nc;
        rc = upload_desc(b);
        if (rc) return rc;
    }
    std::vector<double> rho(n, 0.0);
    if (dtratio) {
        rho.assign(dtratio, dtratio + n);
    } else {  // io_input.cpp:169-172
        for (size_t k = 2; k < n; ++k) rho[k] = (t1grid[k] - t1grid[k - 1]) / (t1grid[k - 1] - t1grid[k - 2]);
    }
    std::vector<double2> v(n * N), d(n * N);
    for (size_t k = 0; k < n * N; ++k) { v[k] = make_double2(QK[k], QR[k]); d[k] = make_double2(dQK[k], dQR[k]); }
    HIP_TRY(hipMemcpyAsync(x.H.V, v.data(), n * N * sizeof(double2), hipMemcpyHostToDevice, b->stream));
    HIP_TRY(hipMemcpyAsync(x.H.D, d.data(), n * N * sizeof(double2), hipMemcpyHostToDevice, b->stream));
    HIP_TRY(hipMemcpyAsync(x.H.t1, t1grid, n * sizeof(double), hipMemcpyHostToDevice, b->stream));
    HIP_TRY(hipMemcpyAsync(x.H.rho, rho.data(), n * sizeof(double), hipMemcpyHostToDevice, b->stream));
    HIP_TRY(hipMemcpyAsync(x.H.r, rvec, n * sizeof(double), hipMemcpyHostToDevice, b->stream));
    HIP_TRY(hipMemcpyAsync(x.H.dr, drvec, n * sizeof(double), hipMemcpyHostToDevice, b->stream));
    HIP_TRY(hipStreamSynchronize(b->stream));
    c.T1 = (int)T1;
    c.t_prop = t1grid[n - 1];
    c.status = kRunning;
    c.active = 0;
    rc = interpolate_inst(b, i);
    if (rc) return rc;
    HIP_TRY(hipStreamSynchronize(b->stream));
    return push_ctl(b);
}

// Per-step correlation records (t, E(t), QK(t,0), loop) of every instance, kept in a
// device ring of `cap` rows per instance; cap <= 0 switches recording off.
extern "C" int dmfe_batch_record(dmfe_batch* b, int cap) {
    if (!b) return set_error(DMFE_ERR_ARG, "null batch");
    (void)hipFree(b->d_obs);
    b->d_obs = nullptr;
    b->obs_cap = 0;
    b->obs_on = false;
    drop_graphs(b);
    if (cap <= 0) return DMFE_OK;
    HIP_TRY(hipMalloc(&b->d_obs, (size_t)n_inst(b) * cap * 4 * sizeof(double)));
    HIP_TRY(hipMemsetAsync(b->d_obs, 0, (size_t)n_inst(b) * cap * 4 * sizeof(double), b->stream));
    HIP_TRY(hipStreamSynchronize(b->stream));
    b->obs_cap = cap;
    b->obs_on = true;
    return DMFE_OK;
}

extern "C" int dmfe_batch_records(dmfe_batch* b, int i, double* out) {
    if (!b || !out || i < 0 || i >= n_inst(b) || !b->obs_on) return set_error(DMFE_ERR_ARG, "dmfe_batch_records: bad argument");
    HIP_TRY(hipMemcpyAsync(out, b->d_obs + (size_t)i * b->obs_cap * 4, (size_t)b->obs_cap * 4 * sizeof(double),
                           hipMemcpyDeviceToHost, b->stream));
    HIP_TRY(hipStreamSynchronize(b->stream));
    return DMFE_OK;
}
