// Stage, search, finish, step-control and utility launches (the RHS row launches are in
// dmfe_rhs.hip and dmfe_loop.hip); device code in dmfe_device.h.
#include "dmfe_device.h"

#include <mutex>

namespace dmfe {
namespace {

#pragma clang fp contract(off)
// [0] Dormand-Prince 5(4) (runge_kutta.cpp:149-159), [1] SSPRK(10,4) (:22-35)
__constant__ double cA[2][kMaxStages][kMaxStages];
__constant__ double cB[2][kMaxStages];
__constant__ double cB2[2][kMaxStages];

template <class C>
__device__ inline void make_finish_dev(const C& c, const InstDesc& d, int m, FinishArgs& f) {
    const int S = c.S;
    const int t = c.method == DMFE_RK54 ? 0 : 1;
    const bool last = (m == S - 1);
    const double dt = c.dt;
    f = FinishArgs{};
    f.mode = 1;
    f.stage = m;
    f.split = d.split;
    f.lr = (m == 0) ? c.T1s - 1 : c.T1s;
    f.g0row = c.T1s - 1;
    f.wrow = c.T1s;
    f.t_rhs = c.gt[m];
    f.t_new = c.gt[m + 1];
    f.nterms = (m == 0) ? 1 : (last ? S : m + 1);
#pragma unroll
    for (int j = 0; j < kMaxStages; ++j) f.coef[j] = (j < f.nterms) ? dt * (last ? cB[t][j] : cA[t][m + 1][j]) : 0.0;
    f.tdiff = c.gt[m + 1] - c.gt[0];
    f.rho_new = (c.T1s > 1) ? f.tdiff / (c.gt[0] - c.tprev) : 0.0;
    f.last = last ? 1 : 0;
    if (last) {
#pragma unroll
        for (int j = 0; j < kMaxStages; ++j) f.coefE[j] = (j < S) ? dt * cB2[t][j] : 0.0;
        f.gt_err = c.gt_err;
    }
    f.fbuf = (c.fc0 + searches_before(c, m + 1)) & 1;
}

// begin: loop condition, stage times gt[] and the error's time term (runge_kutta.cpp:167,201,207-211,234-238).
// One block per instance: lane 0 runs the scalar control, then lanes 0..S build the
// search arguments of stages 0..S and lanes 16..16+S-1 the finish arguments.
constexpr int kBeginThreads = 32;
__device__ __forceinline__ void begin_body(const InstDesc& d, int* go_lds, StepHead* sh) {
    Ctl& c = *d.ctl;
    int& go = *go_lds;
    if (threadIdx.x == 0) {
        // all control fields are loaded before the first test and the first store (one
        // memory round trip), then the two time-grid entries together (the second)
        const int status = c.status, T1 = c.T1, cap = c.cap, method = c.method, fcur = c.fcur;
        const long long loop = c.loop, max_loop = c.max_loop;
        const double dt = c.dt, dt_min = c.dt_min, t_prop = c.t_prop, tmax = c.tmax, delta = c.delta;
        go = 0;
        c.active = 0;
        const double eps = 2.220446049250313e-16;
        const double tol = eps * fmax(fabs(dt), fabs(dt_min));
        if (status != kRunning) {
        } else if (!(t_prop < tmax && loop < max_loop && dt >= dt_min - tol)) {
            c.status = kDone;
        } else if (T1 + 1 > cap) {
            c.status = kNeedRows;
        } else {
            const double* t1 = d.H.t1;
            const double g0 = t1[T1 - 1];
            const double tprev = T1 >= 2 ? t1[T1 - 2] : 0.0;
            const int t = method == DMFE_RK54 ? 0 : 1;
            const int S = method >= 3 ? 4 * (method - 2) : (t == 0 ? 7 : 10);
            StepHead& h = *sh;
            h.T1s = T1;
            h.S = S;
            h.fc0 = fcur;
            h.method = method;
            h.dt = dt;
            h.tprev = tprev;
            h.gt[0] = g0;
            if (method >= 3) {
                // SERK2 stage times (runge_kutta.cu:397-487): Euler steps of dt alpha at even
                // stages, Chebyshev steps 2 g - g_e + 2 dt alpha at odd ones
                const double alpha = 1.0 / (4 * (S / 2));
                double g = g0, ge = g0;
                for (int m = 0; m < S; ++m) {
                    if (m == 0) {
                        g = g0 + dt * alpha * 1.0;
                    } else if (m % 2 == 0) {
                        g = g + dt * alpha * 1.0;
                    } else {
                        g = 2 * g - ge + 2 * dt * alpha * 1.0;
                        ge = g;
                    }
                    h.gt[m + 1] = g;
                }
                h.gt_err = fabs((g0 + dt) - h.gt[S]);
            } else {
                h.gt[1] = g0 + dt * cA[t][1][0] * 1.0;
                for (int m = 1; m < S; ++m) {
                    const bool last = (m == S - 1);
                    double g = g0;
                    const int nt = last ? S : m + 1;
                    for (int j = 0; j < nt; ++j) g += dt * (last ? cB[t][j] : cA[t][m + 1][j]) * 1.0;
                    h.gt[m + 1] = g;
                }
                double gte = g0;
                for (int j = 0; j < S; ++j) gte += dt * cB2[t][j] * 1.0;
                h.gt_err = fabs(h.gt[S] - gte);
            }
            c.T1s = T1;
            c.S = S;
            c.fc0 = fcur;
            c.tprev = tprev;
            c.gt_err = h.gt_err;
            for (int m = 0; m <= S; ++m) c.gt[m] = h.gt[m];
            c.t_prop = g0 + dt;
            c.delta_old = delta;
            // rho of the appended row, as the last stage's finish writes it (make_finish_dev)
            c.rho_step = (T1 > 1) ? (h.gt[S] - g0) / (g0 - tprev) : 0.0;
            go = 1;
        }
    }
    __syncthreads();
    if (!go) return;
    const StepHead& h = *sh;
    const int k = threadIdx.x;
    if (h.method < 3) {  // a SERK2 step's stage launches form their own arguments
        if (k <= h.S) make_search_dev(h, d, k, c.srch[k]);
        else if (k >= 16 && k - 16 < h.S) make_finish_dev(h, d, k - 16, c.fin[k - 16]);
    }
    if (k == 0) c.active = 1;
}

__global__ __launch_bounds__(kBeginThreads) void batch_begin_kernel(InstDesc* D, int n) {
    __shared__ int go;
    __shared__ StepHead sh;
    if ((int)blockIdx.x < n) begin_body(D[blockIdx.x], &go, &sh);
}

// device loop of one context: the instance comes by value (no descriptor loads)
// tail fit after a host-driven step (row = the step's last row) and after a device-loop
// step (the row the step appended: T1s; instances not running this step are skipped)
__global__ __launch_bounds__(kStageThreads) void tail_fit_kernel(const double* theta, double2* row, double2* pre,
                                                                 TailState* st, int N, long long t1len, double delta_max,
                                                                 long long rollback_loop) {
    __shared__ double lds[3 * (kTailMaxN / 2) + 72];
    tail_fit_body(theta, row, pre, st, N, t1len, delta_max, rollback_loop, lds);
}

__global__ __launch_bounds__(kStageThreads) void loop_tail_fit_kernel(InstDesc d, double delta_max) {
    __shared__ double lds[3 * (kTailMaxN / 2) + 72];
    const Ctl& c = *d.ctl;
    const int active = c.active, T1s = c.T1s;
    const long long last_rollback = c.last_rollback;
    if (!active) return;
    tail_fit_body(d.T.theta, d.H.V + (size_t)T1s * d.T.N, d.W.tf_pre, d.W.tf, d.T.N, (long long)T1s + 1, delta_max,
                  last_rollback, lds);
}

__global__ __launch_bounds__(kBeginThreads) void loop_begin_kernel(InstDesc d) {
    __shared__ int go;
    __shared__ StepHead sh;
    begin_body(d, &go, &sh);
}

// the end of one step, its observables (obs non-null) and the begin of the next in one
// launch (consecutive steps of the device loop): lane 0 runs the end's control, the block
// the observables of the step, then the begin, which reads what the end stored
__global__ __launch_bounds__(kStageThreads) void loop_end_begin_kernel(InstDesc d, double* obs, int cap) {
    __shared__ int go;
    __shared__ StepHead sh;
    __shared__ double red[2 * (kStageThreads / kWave)];
    if (threadIdx.x == 0) end_body(d);
    if (obs) {
        __syncthreads();  // the end's control word is visible to the block
        obs_body(d, 0, obs, cap, red);
        __syncthreads();  // every lane has read it before the begin rewrites it
    }
    begin_body(d, &go, &sh);
}
#pragma clang fp contract(on)

}  // namespace

// Device scratch of dmfe_rhs_noalloc (quadrature partials, QK_B1[j,0], rInt, hK, hR):
// a file-scope array that every device holds until the process ends.  In host code
// `scratch` only stands in for the symbol; its device address is resolved once per
// device with hipGetSymbolAddress (scratch_address) and only that address is handed to
// kernels and copies.
static __device__ __align__(256) double scratch[kScratchDoubles];

int scratch_address(double** out) {
    static std::mutex mu;
    static std::vector<double*> addr;
    *out = nullptr;
    int dev = 0;
    hipError_t e = hipGetDevice(&dev);
    if (e != hipSuccess) return set_error(DMFE_ERR_HIP, std::string("scratch: hipGetDevice: ") + hipGetErrorString(e));
    std::lock_guard<std::mutex> lock(mu);
    if ((size_t)dev < addr.size() && addr[(size_t)dev]) {
        *out = addr[(size_t)dev];
        return DMFE_OK;
    }
    void* ptr = nullptr;
    e = hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch));
    if (e != hipSuccess || !ptr)
        return set_error(DMFE_ERR_HIP, std::string("scratch: hipGetSymbolAddress: ") + hipGetErrorString(e));
    if (addr.size() <= (size_t)dev) addr.resize((size_t)dev + 1, nullptr);
    addr[(size_t)dev] = static_cast<double*>(ptr);
    *out = addr[(size_t)dev];
    return DMFE_OK;
}

// Dense B-spline weights of the grid generator (grid_bspline_weights_device):
// W[k][i] = sum_{j <= p} B[k][j] Z[start[k] + j][i] for queries k < Q, i < N; one block per
// query, rows of Z read coalesced, columns outside [0, N) skipped (their basis value is 0).
__global__ __launch_bounds__(256) void bspline_combine_kernel(const double* __restrict__ Z, const int* __restrict__ start,
                                                              const double* __restrict__ B, int p, int N,
                                                              double* __restrict__ W) {
    const long long k = blockIdx.x;
    const int s0 = start[k];
    const double* b = B + k * (p + 1);
    for (int i = threadIdx.x; i < N; i += 256) {
        double acc = 0.0;
        for (int j = 0; j <= p; ++j) {
            const int c = s0 + j;
            if (c >= 0 && c < N) acc = fma(b[j], Z[(size_t)c * N + i], acc);
        }
        W[(size_t)k * N + i] = acc;
    }
}

void launch_bspline_combine(const double* Z, const int* start, const double* B, int p, int N, long long Q, double* W,
                            hipStream_t s) {
    if (Q <= 0) return;
    hipLaunchKernelGGL(bspline_combine_kernel, dim3((unsigned)Q), dim3(256), 0, s, Z, start, B, p, N, W);
}

void launch_dense_apply(const double* W, const double* X, double* Y, long long P, int K, int B, int ldy,
                        hipStream_t s) {
    const long long nb = (P + kDenseRows - 1) / kDenseRows;
    hipLaunchKernelGGL(dense_apply_kernel, dim3((unsigned)nb), dim3(256), 0, s, W, X, Y, P, K, B, ldy);
}

void launch_tan_init_n(const Model& M, const TanArgs& X, int N, hipStream_t s) {
    hipLaunchKernelGGL(tan_init_kernel, dim3(X.K), dim3(256), 0, s, M, X, N);
}


void launch_tan_finish(const Tables& T, const History& H, const Work& W, const Model& M, const FinishArgs& f,
                       const TanArgs& X, hipStream_t s) {
    const int ept = (T.N + kStageThreads - 1) / kStageThreads;
#define DMFE_TF(E, KK) \
    hipLaunchKernelGGL((tan_finish_kernel<E, KK>), dim3(1), dim3(kStageThreads), 0, s, T, H, W, M, f, X)
    if (X.K == 1) {
        if (ept <= 1) DMFE_TF(1, 1);
        else if (ept <= 2) DMFE_TF(2, 1);
        else if (ept <= 4) DMFE_TF(4, 1);
        else DMFE_TF(8, 1);
    } else {
        if (ept <= 1) DMFE_TF(1, 2);
        else if (ept <= 2) DMFE_TF(2, 2);
        else if (ept <= 4) DMFE_TF(4, 2);
        else DMFE_TF(8, 2);
    }
#undef DMFE_TF
}

void launch_tan_energy(const Tables& T, const History& H, const Model& M, const TanArgs& X, int row, double t,
                       hipStream_t s) {
    if (X.K == 1) hipLaunchKernelGGL(tan_energy_kernel<1>, dim3(1), dim3(kStageThreads), 0, s, T, H, M, X, row, t);
    else hipLaunchKernelGGL(tan_energy_kernel<2>, dim3(1), dim3(kStageThreads), 0, s, T, H, M, X, row, t);
}

void launch_hbm_copy(const double2* src, double2* dst, long long n, hipStream_t s) {
    hipLaunchKernelGGL(hbm_copy_kernel, dim3(2048), dim3(256), 0, s, src, dst, n);
}

void launch_stage(const Tables& T, const History& H, const Work& W, const Model& M, const SearchArgs* sa,
                  const FinishArgs* fa, hipStream_t s) {
    int flags = 0, nb = 0;
    SearchArgs S{};
    FinishArgs F{};
    if (fa) { flags |= kHasFinish; F = *fa; ++nb; }
    if (sa) {
        S = *sa;
        flags |= kHasTheta;
        ++nb;
        flags |= kHasRows;
        nb += T.N;
    }
    if (nb == 0) return;
    const int ept = (T.N + kStageThreads - 1) / kStageThreads;
    if (ept <= 1) hipLaunchKernelGGL(stage_kernel<1>, dim3(nb), dim3(kStageThreads), 0, s, T, H, W, M, S, F, flags);
    else if (ept <= 2) hipLaunchKernelGGL(stage_kernel<2>, dim3(nb), dim3(kStageThreads), 0, s, T, H, W, M, S, F, flags);
    else if (ept <= 4) hipLaunchKernelGGL(stage_kernel<4>, dim3(nb), dim3(kStageThreads), 0, s, T, H, W, M, S, F, flags);
    else hipLaunchKernelGGL(stage_kernel<8>, dim3(nb), dim3(kStageThreads), 0, s, T, H, W, M, S, F, flags);
}


void launch_row_scalar(const Tables& T, const History& H, const Work& W, const Model& M, int row, double t,
                       int what, hipStream_t s) {
    hipLaunchKernelGGL(row_scalar_kernel, dim3(1), dim3(kStageThreads), 0, s, T, H, W, M, row, t, what);
}

void launch_sparsify_values(const Tables& T, const History& H, int T1, double* val, hipStream_t s) {
    if (T1 < 4) return;
    hipLaunchKernelGGL(sparsify_values_kernel, dim3(T1 - 3), dim3(kRowThreads), 0, s, T, H, T1, val);
}

void launch_gather_rows(const Tables& T, const History& src, const History& dst, const int* keep, int nkeep,
                        hipStream_t s) {
    if (nkeep <= 0) return;
    hipLaunchKernelGGL(gather_rows_kernel, dim3(nkeep), dim3(kRowThreads), 0, s, T, src, dst, keep, nkeep);
}


// The tableaux live in constant memory, once per device: uploaded by the first batch or
// device loop created on a device (under a lock, so concurrent first calls from several
// host threads upload once) and never rewritten while kernels read them.
int batch_upload_tableaux(const Tableau& rk54, const Tableau& ssprk) {
    static std::mutex mu;
    static std::vector<char> done;
    int dev = 0;
    hipError_t e = hipGetDevice(&dev);
    if (e != hipSuccess) return set_error(DMFE_ERR_HIP, std::string("tableaux: hipGetDevice: ") + hipGetErrorString(e));
    std::lock_guard<std::mutex> lock(mu);
    if ((size_t)dev < done.size() && done[(size_t)dev]) return DMFE_OK;
    double A[2][kMaxStages][kMaxStages] = {}, B[2][kMaxStages] = {}, B2[2][kMaxStages] = {};
    const Tableau* t[2] = {&rk54, &ssprk};
    for (int k = 0; k < 2; ++k)
        for (int i = 0; i < kMaxStages; ++i) {
            for (int j = 0; j < kMaxStages; ++j) A[k][i][j] = t[k]->A[i][j];
            B[k][i] = t[k]->b[i];
            B2[k][i] = t[k]->b2[i];
        }
    e = hipMemcpyToSymbol(HIP_SYMBOL(cA), A, sizeof A);
    if (e == hipSuccess) e = hipMemcpyToSymbol(HIP_SYMBOL(cB), B, sizeof B);
    if (e == hipSuccess) e = hipMemcpyToSymbol(HIP_SYMBOL(cB2), B2, sizeof B2);
    if (e != hipSuccess) return set_error(DMFE_ERR_HIP, std::string("tableaux upload: ") + hipGetErrorString(e));
    if (done.size() <= (size_t)dev) done.resize((size_t)dev + 1, 0);
    done[(size_t)dev] = 1;
    return DMFE_OK;
}

void launch_batch_begin(InstDesc* D, int n, hipStream_t s) {
    hipLaunchKernelGGL(batch_begin_kernel, dim3(n), dim3(kBeginThreads), 0, s, D, n);
}

void launch_batch_end(InstDesc* D, int n, hipStream_t s) {
    hipLaunchKernelGGL(batch_end_kernel, dim3((n + 63) / 64), dim3(64), 0, s, D, n);
}

void launch_batch_stage(InstDesc* D, const int2* map, int nblocks, int n, int ept, hipStream_t s) {
    if (ept <= 1) hipLaunchKernelGGL(batch_stage_kernel<1>, dim3(nblocks), dim3(kStageThreads), 0, s, D, map, n);
    else if (ept <= 2) hipLaunchKernelGGL(batch_stage_kernel<2>, dim3(nblocks), dim3(kStageThreads), 0, s, D, map, n);
    else if (ept <= 4) hipLaunchKernelGGL(batch_stage_kernel<4>, dim3(nblocks), dim3(kStageThreads), 0, s, D, map, n);
    else hipLaunchKernelGGL(batch_stage_kernel<8>, dim3(nblocks), dim3(kStageThreads), 0, s, D, map, n);
}


// ---- device loop of one context (instance by value) ----
void launch_loop_begin(const InstDesc& d, hipStream_t s) {
    hipLaunchKernelGGL(loop_begin_kernel, dim3(1), dim3(kBeginThreads), 0, s, d);
}

void launch_loop_end(const InstDesc& d, hipStream_t s) {
    hipLaunchKernelGGL(loop_end_kernel, dim3(1), dim3(64), 0, s, d);
}

void launch_loop_end_begin(const InstDesc& d, double* obs, int obs_cap, hipStream_t s) {
    hipLaunchKernelGGL(loop_end_begin_kernel, dim3(1), dim3(kStageThreads), 0, s, d, obs, obs_cap);
}

void launch_loop_tail_fit(const InstDesc& d, double delta_max, hipStream_t s) {
    hipLaunchKernelGGL(loop_tail_fit_kernel, dim3(1), dim3(kStageThreads), 0, s, d, delta_max);
}

void launch_tail_fit(const Tables& T, const History& H, const Work& W, int row, long long t1len, double delta_max,
                     long long rollback_loop, hipStream_t s) {
    hipLaunchKernelGGL(tail_fit_kernel, dim3(1), dim3(kStageThreads), 0, s, T.theta, H.V + (size_t)row * T.N, W.tf_pre,
                       W.tf, T.N, t1len, delta_max, rollback_loop);
}

void launch_loop_stage(const InstDesc& d, int n, hipStream_t s) {
    const int N = d.T.N, ept = (N + kStageThreads - 1) / kStageThreads;
    const dim3 grid((unsigned)(N + 2));
    if (d.X.K) {  // with the tangent finish of stage n-1 in block 0
#define DMFE_LST(E, KK) hipLaunchKernelGGL((loop_stage_tan_kernel<E, KK>), grid, dim3(kStageThreads), 0, s, d, n)
#define DMFE_LSTE(KK)                       \
    if (ept <= 1) DMFE_LST(1, KK);          \
    else if (ept <= 2) DMFE_LST(2, KK);     \
    else if (ept <= 4) DMFE_LST(4, KK);     \
    else DMFE_LST(8, KK);
        if (d.X.K == 1) { DMFE_LSTE(1) }
        else { DMFE_LSTE(2) }
#undef DMFE_LSTE
#undef DMFE_LST
        return;
    }
    if (ept <= 1) hipLaunchKernelGGL(loop_stage_kernel<1>, grid, dim3(kStageThreads), 0, s, d, n);
    else if (ept <= 2) hipLaunchKernelGGL(loop_stage_kernel<2>, grid, dim3(kStageThreads), 0, s, d, n);
    else if (ept <= 4) hipLaunchKernelGGL(loop_stage_kernel<4>, grid, dim3(kStageThreads), 0, s, d, n);
    else hipLaunchKernelGGL(loop_stage_kernel<8>, grid, dim3(kStageThreads), 0, s, d, n);
}


void launch_batch_obs(InstDesc* D, int n, double* obs, int obs_cap, hipStream_t s) {
    hipLaunchKernelGGL(batch_obs_kernel, dim3(n), dim3(kStageThreads), 0, s, D, obs, obs_cap);
}

}  // namespace dmfe
