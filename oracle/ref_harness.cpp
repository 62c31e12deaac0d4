// Test-only driver for the reference's own CPU sources (compiled where they lie,
// see oracle/Makefile). It owns the three globals that the reference's main.cpp
// normally defines (main.cpp:32-37: `config`, `sim`, `rk`) and exposes a flat
// extern "C" surface so the Python tests can drive the reference CPU path on the
// same inputs as the C restatement (oracle/dmfe_oracle.c) and the HIP library.
//
// Nothing here re-implements reference arithmetic: every numerical call goes to
// the reference's own functions (interpolate, update, QKstep, ... and the grid
// generator pieces).  This file is test infrastructure only.
#include <algorithm>
#include <vector>
#include <string>
#include <cstring>
#include <cstdint>
#include <cmath>

#include "core/config.hpp"
#include "simulation/simulation_data.hpp"
#include "EOMs/rk_data.hpp"
#include "EOMs/time_steps.hpp"
#include "EOMs/runge_kutta.hpp"
#include "interpolation/interpolation_core.hpp"
#include "search/search_utils.hpp"
#include "sparsify/sparsify_utils.hpp"
#include "math/math_ops.hpp"
#include "math/math_sigma.hpp"
#include "convolution/convolution.hpp"
#include "grid/theta_grid.hpp"
#include "grid/phi_grid.hpp"
#include "grid/integration.hpp"
#include "grid/pos_grid.hpp"
#include "grid/interpolation_weights.hpp"
#include "grid/grid_io.hpp"
#include "io/io_utils.hpp"
#include "simulation/simulation_control.hpp"
#include "simulation/tail_fit.hpp"

SimulationConfig config;
SimulationData* sim = nullptr;
RKData* rk = nullptr;
// defined by simulation_runner.cpp (not compiled here); read by the tail fit
int last_rollback_loop = -1000;
static bool g_tail_fit = false;

namespace {
template <class T>
void put(std::vector<T>& v, const T* src, size_t n) { v.assign(src, src + n); }
template <class T>
void get(const std::vector<T>& v, T* dst) { std::memcpy(dst, v.data(), v.size() * sizeof(T)); }
}

extern "C" {

// Grid tables as the reference's import() leaves them (io_input.cpp:90-110).
// inds are int64 here; -1 maps to SIZE_MAX exactly like the reference parser.
void ref_setup(int64_t N, const double* theta, const double* integ, const double* phi2,
               const int64_t* indsA1, const int64_t* indsA2, const int64_t* indsB2,
               const double* wA1, const double* wA2, const double* wB2, int64_t m,
               int p, int p2, double lambda, double T0, double Gamma,
               double delta_t_min, double delta_max, int log_response) {
    delete sim; delete rk;
    sim = new SimulationData();
    rk = new RKData();
    config = SimulationConfig();
    config.len = (size_t)N;
    config.p = p; config.p2 = p2; config.lambda = lambda; config.T0 = T0; config.Gamma = Gamma;
    config.delta_t_min = delta_t_min; config.delta_max = delta_max;
    config.log_response_interp = log_response != 0;
    config.gpu = false; config.debug = false; config.save_output = false;
    const size_t P = (size_t)N * (size_t)N;
    put(sim->h_theta, theta, (size_t)N);
    put(sim->h_integ, integ, (size_t)N);
    put(sim->h_phi2, phi2, P);
    sim->h_indsA1y.resize(P); sim->h_indsA2y.resize(P); sim->h_indsB2y.resize(P);
    for (size_t k = 0; k < P; ++k) {
        sim->h_indsA1y[k] = (size_t)indsA1[k];
        sim->h_indsA2y[k] = (size_t)indsA2[k];
        sim->h_indsB2y[k] = (size_t)indsB2[k];
    }
    put(sim->h_weightsA1y, wA1, P * (size_t)m);
    put(sim->h_weightsA2y, wA2, P * (size_t)m);
    put(sim->h_weightsB2y, wB2, P * (size_t)m);
    // scratch fields (initialization.cpp:233-256)
    sim->h_posB1xOld.assign((size_t)N, 1.0);
    sim->h_posB2xOld.assign(P, 0.0);
    for (auto* v : {&sim->h_SigmaKA1int, &sim->h_SigmaRA1int, &sim->h_SigmaKB1int, &sim->h_SigmaRB1int,
                    &sim->h_SigmaKA2int, &sim->h_SigmaRA2int, &sim->h_SigmaKB2int, &sim->h_SigmaRB2int,
                    &sim->h_QKA1int, &sim->h_QRA1int, &sim->h_QKB1int, &sim->h_QRB1int,
                    &sim->h_QKA2int, &sim->h_QRA2int, &sim->h_QKB2int, &sim->h_QRB2int})
        v->assign(P, 0.0);
    sim->h_rInt.assign((size_t)N, 0.0);
    sim->h_drInt.assign((size_t)N, 0.0);
}

// The "new simulation" branch of init() (initialization.cpp:204-222, 278-283).
void ref_init_quench(void) {
    sim->h_t1grid.assign(1, 0.0);
    sim->h_delta_t_ratio.assign(1, 0.0);
    config.specRad = 4 * std::sqrt(DDflambda(1));
    config.delta_t = config.delta_t_min;
    config.loop = 0;
    config.delta = 1;
    config.delta_old = 0;
    sim->h_QKv.assign(config.len, 1.0);
    sim->h_QRv.assign(config.len, 1.0);
    sim->h_dQKv.assign(config.len, 0.0);
    sim->h_dQRv.assign(config.len, 0.0);
    sim->h_rvec.assign(1, config.Gamma + Dflambda(1) / config.T0);
    sim->h_drvec.assign(1, rstep());
    rk->init = (config.delta_t < config.rmax[0] / config.specRad) ? 1 : 2;
    interpolate();
}

void ref_set_history(int64_t T1, const double* t1grid, const double* dtratio,
                     const double* rvec, const double* drvec,
                     const double* QK, const double* QR, const double* dQK, const double* dQR) {
    const size_t n = (size_t)T1 * config.len;
    put(sim->h_t1grid, t1grid, (size_t)T1);
    put(sim->h_delta_t_ratio, dtratio, (size_t)T1);
    put(sim->h_rvec, rvec, (size_t)T1);
    put(sim->h_drvec, drvec, (size_t)T1);
    put(sim->h_QKv, QK, n); put(sim->h_QRv, QR, n);
    put(sim->h_dQKv, dQK, n); put(sim->h_dQRv, dQR, n);
}

int64_t ref_T1(void) { return (int64_t)sim->h_t1grid.size(); }

// QK(t,0) of the last history row, and the time grid alone (cheap accessors for long runs)
double ref_last_qk0(void) { return sim->h_QKv[(sim->h_t1grid.size() - 1) * config.len]; }
void ref_get_t1grid(double* out) { std::copy(sim->h_t1grid.begin(), sim->h_t1grid.end(), out); }

void ref_get_history(double* t1grid, double* dtratio, double* rvec, double* drvec,
                     double* QK, double* QR, double* dQK, double* dQR) {
    get(sim->h_t1grid, t1grid); get(sim->h_delta_t_ratio, dtratio);
    get(sim->h_rvec, rvec); get(sim->h_drvec, drvec);
    get(sim->h_QKv, QK); get(sim->h_QRv, QR); get(sim->h_dQKv, dQK); get(sim->h_dQRv, dQR);
}

void ref_set_control(double delta_t, double delta, double delta_old, int loop, int rk_init) {
    config.delta_t = delta_t; config.delta = delta; config.delta_old = delta_old;
    config.loop = loop; rk->init = rk_init;
    config.specRad = 4 * std::sqrt(DDflambda(1));
}

void ref_get_control(double* delta_t, double* delta, double* delta_old, int* loop, int* rk_init) {
    *delta_t = config.delta_t; *delta = config.delta; *delta_old = config.delta_old;
    *loop = config.loop; *rk_init = rk->init;
}

void ref_set_pos_old(const double* posB1, const double* posB2) {
    put(sim->h_posB1xOld, posB1, config.len);
    put(sim->h_posB2xOld, posB2, config.len * config.len);
}

void ref_get_pos_old(double* posB1, double* posB2) {
    get(sim->h_posB1xOld, posB1); get(sim->h_posB2xOld, posB2);
}

// mode 0: interpolate() (binary search); 1: isearch seeded by posOld; 2: reuse posOld.
void ref_interpolate(int mode) {
    if (mode == 0) { interpolate(); return; }
    std::vector<double> b1 = sim->h_posB1xOld, b2 = sim->h_posB2xOld;
    interpolate(b1, b2, mode == 2);
}

// Interpolated fields in the order QKA1 QRA1 QKA2 QRA2 QKB1 QRB1 QKB2 QRB2, then
// SigmaK/R in the order KA1 RA1 RA2 KB1 RB1 KB2 RB2 (each N*N), then rInt (N).
void ref_get_fields(double* out) {
    const std::vector<double>* f[] = {
        &sim->h_QKA1int, &sim->h_QRA1int, &sim->h_QKA2int, &sim->h_QRA2int,
        &sim->h_QKB1int, &sim->h_QRB1int, &sim->h_QKB2int, &sim->h_QRB2int,
        &sim->h_SigmaKA1int, &sim->h_SigmaRA1int, &sim->h_SigmaRA2int,
        &sim->h_SigmaKB1int, &sim->h_SigmaRB1int, &sim->h_SigmaKB2int, &sim->h_SigmaRB2int};
    size_t off = 0;
    for (auto* v : f) { std::memcpy(out + off, v->data(), v->size() * sizeof(double)); off += v->size(); }
    std::memcpy(out + off, sim->h_rInt.data(), sim->h_rInt.size() * sizeof(double));
}

void ref_rhs(double* hK, double* hR) {
    std::vector<double> a = QKstep(), b = QRstep();
    get(a, hK); get(b, hR);
}

double ref_rstep(void) { return rstep(); }

double ref_drstep2(const double* qK, const double* qR, const double* dqK, const double* dqR, double t) {
    std::vector<double> a(qK, qK + config.len), b(qR, qR + config.len),
        c(dqK, dqK + config.len), d(dqR, dqR + config.len);
    return drstep2(a, b, c, d, t);
}

double ref_update(void) { return update(); }

void ref_bsearch(int64_t nlist, const double* list, int64_t nel, const double* el, double* out) {
    std::vector<double> L(list, list + nlist), E(el, el + nel);
    get(bsearchPosSorted(L, E), out);
}

void ref_isearch(int64_t nlist, const double* list, int64_t nel, const double* el,
                 const double* inits, double* out) {
    std::vector<double> L(list, list + nlist), E(el, el + nel), I(inits, inits + nel);
    get(isearchPosSortedInit(L, E, I), out);
}

void ref_sparsify(double threshold) { sparsifyNscale(threshold); }

// The energy line written to correlation.txt by the CPU branch
// (simulation_runner.cpp:334-337) at proposed time t.
double ref_energy(double t) {
    std::vector<double> temp(config.len, 0.0);
    SigmaK(getLastLenEntries(sim->h_QKv, config.len), temp);
    double qk0 = sim->h_QKv[(sim->h_t1grid.size() - 1) * config.len];
    return -(ConvA(temp, getLastLenEntries(sim->h_QRv, config.len), t)[0] + Dflambda(qk0) / config.T0);
}

// One iteration of the CPU main loop body (simulation_runner.cpp:125-278,
// CPU-only build branches; no I/O).  Returns the proposed t used for the
// correlation line; the error norm is left in config.delta.
double ref_loop_iteration(void) {
    double t = sim->h_t1grid.back() + config.delta_t;
    config.delta_old = config.delta;
    config.delta = update();
    // --tail-fit of a CUDA build run with --gpu false (simulation_runner.cpp:135-141); the
    // CPU-only build never calls it
    if (g_tail_fit) tailFitBlendCPU();
    config.loop++;
    if (config.loop % 100000 == 0) {
        size_t prev_size = sim->h_t1grid.size();
        int count = 0;
        while (count < 1) {
            sparsifyNscale(config.delta_max);
            size_t new_size = sim->h_t1grid.size();
            if (new_size >= prev_size) break;
            prev_size = new_size;
            count++;
        }
        interpolate();
        if (config.delta < config.delta_max / 2 && config.loop - (-1000) > 1000) {
            config.delta_t *= 0.5;
            rk->init = 2;
        }
    }
    if (config.delta < config.delta_max && config.loop > 5 &&
        (config.delta < 1.1 * config.delta_old || config.delta_old < config.delta_max / 1000) &&
        config.rmax[rk->init - 1] / config.specRad > config.delta_t && sim->h_delta_t_ratio.back() == 1.0) {
        config.delta_t *= 1.01;
    } else if (config.delta > 2 * config.delta_max && config.delta_t > config.delta_t_min) {
        config.delta_t *= 0.9;
    }
    if (rk->init == 2 && config.delta > config.delta_max && config.rmax[0] / config.specRad > config.delta_t) {
        rk->init = 1;
        config.delta_t *= 0.5;
    }
    return t;
}

// The tail fit after every ref_loop_iteration (tailFitBlendCPU, tail_fit.cpp:46-265).  Its
// state lives in the reference's function-local statics: one run per loaded library.
void ref_set_tail_fit(int on) { g_tail_fit = on != 0; }

// rollbackState(n) (simulation_control.cpp:14-124, CPU-only build): truncate the history
// by n rows, delta_t from the kept last step, loop -= n, interpolate().  1 on success.
int ref_rollback_state(int n) { return rollbackState(n) ? 1 : 0; }

// ---- grid generator pieces (src/grid/*) -------------------------------------
static std::vector<long double> g_theta, g_phi1, g_phi2, g_int;

void ref_grid_theta(int64_t N, double Tmax, double alpha, double delta, double* theta_out) {
    generate_theta_grid((size_t)N, Tmax, g_theta, alpha, delta);
    generate_phi_grids(g_theta, g_phi1, g_phi2);
    for (int64_t i = 0; i < N; ++i) theta_out[i] = (double)g_theta[i];
}

void ref_grid_int(int spline_order, double* int_out) {
    compute_integration_weights(g_theta, spline_order, g_int);
    for (size_t i = 0; i < g_int.size(); ++i) int_out[i] = (double)g_int[i];
}

void ref_grid_pos(double Tmax, double alpha, double delta, double* posA1, double* posA2, double* posB2) {
    std::vector<double> a, b, c;
    generate_pos_grids(g_theta.size(), Tmax, g_theta, g_phi1, g_phi2, a, b, c, alpha, delta);
    get(a, posA1); get(b, posA2); get(c, posB2);
}

// which: 0 A1 (phi1), 1 A2 (phi2), 2 B2 (theta_i/(phi2-1e-200)).
// method: 0 poly, 1 rational (Floater-Hormann window mfh), 2 bspline.
// For bspline, inds are -1 and m = N.
void ref_grid_weights(int which, int method, int n, int mfh, int64_t* inds, double* w) {
    const size_t N = g_theta.size();
    std::vector<long double> q;
    if (which == 0) q = g_phi1;
    else if (which == 1) q = g_phi2;
    else {
        const double tiny = 1e-200;  // grid_cli.cpp:172-186
        q.reserve(N * N);
        for (size_t i = 0; i < N; ++i) {
            const long double ti = g_theta[i];
            for (size_t j = 0; j < N; ++j) {
                const long double denom = g_phi2[i * N + j] - tiny;
                q.push_back(std::fabs(denom) < 1e-300L ? ((ti > 0) ? g_theta.back() : g_theta.front()) : ti / denom);
            }
        }
    }
    if (method == 0 || method == 1) {
        auto st = method == 0 ? dmfe::grid::compute_barycentric_weights(g_theta, q, n)
                              : dmfe::grid::compute_barycentric_rational_weights(g_theta, q, n, mfh);
        for (size_t k = 0; k < st.size(); ++k) {
            inds[k] = st[k].start;
            std::memcpy(w + k * st[k].alpha.size(), st[k].alpha.data(), st[k].alpha.size() * sizeof(double));
        }
    } else {
        auto W = dmfe::grid::compute_bspline_weights(g_theta, q, n);
        for (size_t k = 0; k < W.size(); ++k) {
            inds[k] = -1;
            std::memcpy(w + k * N, W[k].w.data(), N * sizeof(double));
        }
    }
}

}  // extern "C"

// Replace the generator's theta (extended precision) so the weight / quadrature
// code of the reference can be compared on exactly the same nodes.
extern "C" void ref_grid_set_theta_ld(int64_t N, const long double* th) {
    g_theta.assign(th, th + N);
    generate_phi_grids(g_theta, g_phi1, g_phi2);
}

// ---- files written by the reference's own writers ------------------------------------
// saveSimulationStateBinary (io_save_binary.cpp:22-116): data.bin version 1 of the
// current state at `path` (and params.txt beside it, io_save_params.cpp), with the
// control scalars set by ref_set_control.
extern "C" void ref_save_binary(const char* path) {
    saveSimulationStateBinary(std::string(path), config.delta, config.delta_t);
}

// The grid package of the generator state left by ref_grid_theta / ref_grid_int (plus
// the given positions and stencil tables) under Grid_data/<subdir> of the working
// directory: write_all_grids, the three interp-metadata writers and
// write_grid_generation_params (grid_io.cpp:176-328).  inds/weights: A1, A2, B2 blocks.
extern "C" void ref_grid_write(const char* subdir, const double* posA1, const double* posA2, const double* posB2,
                               const int64_t* inds, const double* weights, int64_t m, double Tmax,
                               int spline_order, const char* interp_method, int interp_order, int fh_stencil,
                               double alpha, double delta) {
    const size_t N = g_theta.size(), P = N * N;
    std::vector<double> a(posA1, posA1 + P), b(posA2, posA2 + P), c(posB2, posB2 + P);
    write_all_grids(g_theta, g_phi1, g_phi2, g_int, a, b, c, N, subdir);
    for (int f = 0; f < 3; ++f) {
        std::vector<int> it(P);
        for (size_t k = 0; k < P; ++k) it[k] = (int)inds[f * P + k];
        std::vector<double> w(weights + f * P * (size_t)m, weights + (f + 1) * P * (size_t)m);
        if (f == 0) write_A1_interp_metadata(it, w, N, subdir);
        else if (f == 1) write_A2_interp_metadata(it, w, N, subdir);
        else write_B2_interp_metadata(it, w, N, subdir);
    }
    write_grid_generation_params(N, Tmax, spline_order, interp_method, interp_order, fh_stencil, subdir, "", alpha,
                                 delta);
}
