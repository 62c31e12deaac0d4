// dmfe_run: command-line driver over the C ABI (include/dmfe.h) for one quench on one
// MI355X.  It replaces the reference's main.cpp + runSimulation for this path: the same
// model/control options (config.cpp:28-175), Grid_data/<L> packages (io_input.cpp:90-110),
// correlation.txt rows "t<TAB>E<TAB>QK(t,0)" with 14 fixed decimals
// (simulation_runner.cpp:91,327-337), data.bin version-1 checkpoints and a params.txt with
// the reference's physical / numerical / current-state keys (io_save_params.cpp:174-239).
// The output directory is <out>/p=.._q=.._lambda=.._T0=.._G=.._L=.. (io_utils.cpp:256-283
// without the date prefix, so that a rerun finds and resumes its own checkpoint).
#include "dmfe.h"

#include <cerrno>
#include <cmath>
#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>
#include <sys/stat.h>
#include <time.h>
#include <algorithm>
#include <csignal>
#include <vector>

namespace {

// SIGINT / SIGTERM / SIGHUP end the run after the current segment: the final save runs
// and the exit status is 130 (simulation_text_ui.cpp:30-65, simulation_runner.cpp:126,467)
volatile std::sig_atomic_t g_interrupted = 0;
void on_signal(int) { g_interrupted = 1; }

// main-loop iterations per dmfe_run call: the correlation rows are written and the signal
// flag is checked between calls
constexpr int64_t kSegment = 256;

struct Options {
    int p = 3, p2 = 12;
    double lambda = 0.3, T0 = 1e50, Gamma = 0.0;
    long long max_loop = 10000;
    double tmax = 1e7, dt_min = 1e-5, delta_max = 1e-10;
    long long len = 512;
    bool save = true, log_response = false, resume = true;
    bool serk2 = false, rollback = false;  // the reference GPU loop's behaviours (off: CPU path)
    bool tail_fit = false;  // -F: the QK tail fit after every step (the CPU-only build never applies it)
    std::string out = "Results", grid_dir;
    long long save_every = 0, maint_every = 100000;
    int sweeps = -1;  // -w, recorded in params.txt (the CPU build always runs one sweep)
};

void usage() {
    std::printf(
        "dmfe_run [options]\n"
        "  -p P -q P2 -l lambda -T T0|inf -G Gamma     model (f = lambda q^p + (1-lambda) q^p2)\n"
        "  -m maxLoop -t tmax -d delta_t_min -e delta_max -L len\n"
        "  -R true|false     interpolate QR in log space\n"
        "  -s true|false     write correlation.txt / data.bin / params.txt\n"
        "  -o DIR            results root (default Results)\n"
        "  --grid DIR        grid package directory (default Grid_data/<len>)\n"
        "  --save-every N    checkpoint every N loops (default: at the end only)\n"
        "  --maint-every N   sparsify/re-interpolate cadence (default 100000)\n"
        "  --fresh           do not resume from an existing data.bin\n"
        "  -S true|false     SERK2 promotion at the maintenance (GPU-loop behaviour; default false)\n"
        "  --rollback        roll back up to 10 steps on error spikes (GPU-loop behaviour)\n"
        "  -F true|false     QK tail fit near theta -> 1 after every step (default false)\n");
}

bool parse(int argc, char** argv, Options& o) {
    for (int i = 1; i < argc; ++i) {
        const std::string a = argv[i];
        auto val = [&](const char* what) -> const char* {
            if (i + 1 >= argc) {
                std::fprintf(stderr, "missing value for %s\n", what);
                std::exit(2);
            }
            return argv[++i];
        };
        if (a == "-h" || a == "--help") { usage(); std::exit(0); }
        else if (a == "-p") o.p = (int)std::atof(val("-p"));
        else if (a == "-q") o.p2 = (int)std::atof(val("-q"));
        else if (a == "-l" || a == "--lambda") o.lambda = std::atof(val("-l"));
        else if (a == "-T" || a == "--T0") { const std::string v = val("-T"); o.T0 = (v == "inf") ? 1e50 : std::atof(v.c_str()); }
        else if (a == "-G" || a == "--Gamma") o.Gamma = std::atof(val("-G"));
        else if (a == "-m") o.max_loop = (long long)std::atof(val("-m"));
        else if (a == "-t") o.tmax = std::atof(val("-t"));
        else if (a == "-d") o.dt_min = std::atof(val("-d"));
        else if (a == "-e" || a == "--error") o.delta_max = std::atof(val("-e"));
        else if (a == "-L") o.len = (long long)std::atof(val("-L"));
        else if (a == "-R") o.log_response = std::string(val("-R")) != "false";
        else if (a == "-s") o.save = std::string(val("-s")) != "false";
        else if (a == "-o" || a == "--out-dir") o.out = val("-o");
        else if (a == "--grid") o.grid_dir = val("--grid");
        else if (a == "--save-every") o.save_every = (long long)std::atof(val("--save-every"));
        else if (a == "--maint-every") o.maint_every = (long long)std::atof(val("--maint-every"));
        else if (a == "--fresh") o.resume = false;
        else if (a == "-S" || a == "--serk2") o.serk2 = std::string(val("-S")) != "false";
        else if (a == "--rollback") o.rollback = true;
        else if (a == "-w") {
            // the sparsify sweep count: recorded, but the CPU build runs one sweep
            // (simulation_runner.cpp:187-190), as here
            o.sweeps = std::atoi(val("-w"));
        } else if (a == "-D" || a == "-g" || a == "-A" || a == "-I") {
            // accepted for command-line compatibility (config.cpp:84-162); none of them changes
            // this path: debug output, the CPU/GPU switch, asynchronous export and the
            // version check of a resumed run
            const std::string v = val(a.c_str());
            std::fprintf(stderr, "note: option %s %s has no effect here\n", a.c_str(), v.c_str());
        } else if (a == "-F" || a == "--tail-fit") {
            o.tail_fit = std::string(val("-F")) != "false";
        }
        else { std::fprintf(stderr, "unknown option %s\n", a.c_str()); usage(); return false; }
    }
    if (o.grid_dir.empty()) o.grid_dir = "Grid_data/" + std::to_string(o.len);
    return true;
}

bool mkdirs(const std::string& d) {
    std::string cur;
    for (size_t i = 0; i <= d.size(); ++i) {
        if (i == d.size() || d[i] == '/') {
            if (!cur.empty() && mkdir(cur.c_str(), 0755) != 0 && errno != EEXIST) return false;
        }
        if (i < d.size()) cur += d[i];
    }
    return true;
}

std::string param_dir(const Options& o) {
    char buf[256];
    char t0[64];
    if (o.T0 >= 1e50) std::snprintf(t0, sizeof t0, "inf");
    else std::snprintf(t0, sizeof t0, "%.5f", o.T0);
    std::snprintf(buf, sizeof buf, "p=%d_q=%d_lambda=%.2f_T0=%s_G=%.2f_L=%lld", o.p, o.p2, o.lambda, t0, o.Gamma, o.len);
    std::string root = o.out;
    if (!root.empty() && root.back() != '/') root += '/';
    return root + buf;
}

bool check(int rc, const char* what) {
    if (rc == DMFE_OK) return true;
    std::fprintf(stderr, "%s failed (%d): %s\n", what, rc, dmfe_last_error());
    return false;
}

void write_params(const std::string& dir, const Options& o, dmfe_ctx* ctx, double t_last, double energy,
                  double elapsed) {
    double dt = 0, delta = 0, delta_old = 0;
    int64_t loop = 0, T1 = 0;
    int method = 0;
    dmfe_get_control(ctx, &dt, &delta, &delta_old, &loop, &method);
    dmfe_get_T1(ctx, &T1);
    std::vector<double> qk(T1 * o.len), qr(T1 * o.len), r(T1);
    dmfe_get_state(ctx, nullptr, qk.data(), qr.data(), nullptr, nullptr, r.data(), nullptr, nullptr);
    FILE* f = std::fopen((dir + "/params.txt").c_str(), "w");
    if (!f) return;
    std::fprintf(f, "runtime seconds = %.2f\n", elapsed);
    std::fprintf(f, "loops per second = %.2f\n\n", elapsed > 0 ? loop / elapsed : 0.0);
    std::fprintf(f, "# Physical Parameters\np = %d\np2 = %d\nlambda = %.16g\nT0 = %.16g\nGamma = %.16g\n\n", o.p, o.p2,
                 o.lambda, o.T0, o.Gamma);
    std::fprintf(f, "# Numerical Parameters\nlen = %lld\ntmax = %.16g\ndelta_t_min = %.16g\ndelta_max = %.16g\n", o.len,
                 o.tmax, o.dt_min, o.delta_max);
    std::fprintf(f, "use_serk2 = %s\nlog_response_interp = %s\ntail_fit_enabled = %s\nsparsify_sweeps = %d\n",
                 o.serk2 ? "true" : "false", o.log_response ? "true" : "false", o.tail_fit ? "true" : "false", o.sweeps);
    std::fprintf(f, "maxLoop = %lld\nrmax = [3, 13, 20, 80, 180, 320, 500, 720, 980, 1280, 1620]\n\n", o.max_loop);
    // provenance of the grid package, mirrored from its grid_params.txt as the reference
    // does (io_save_params.cpp:202-226): raw values, Tmax / alpha / delta as %.16g
    std::fprintf(f, "# Grid Generation Parameters\n");
    if (FILE* gp = std::fopen((o.grid_dir + "/grid_params.txt").c_str(), "r")) {
        std::vector<std::pair<std::string, std::string>> kv;
        char line[1024];
        while (std::fgets(line, sizeof line, gp)) {
            std::string l(line);
            while (!l.empty() && (l.back() == '\n' || l.back() == '\r')) l.pop_back();
            const size_t eq = l.find('=');
            if (l.empty() || l[0] == '#' || eq == std::string::npos) continue;
            kv.emplace_back(l.substr(0, eq), l.substr(eq + 1));
        }
        std::fclose(gp);
        auto get = [&](const char* k) -> const std::string* {
            for (auto& e : kv)
                if (e.first == k) return &e.second;
            return nullptr;
        };
        const char* raw[][2] = {{"subdir", "grid_subdir"}, {"len", "grid_len"}, {"Tmax", "grid_Tmax"},
                                {"spline_order", "grid_spline_order"}, {"interp_method", "grid_interp_method"},
                                {"interp_order", "grid_interp_order"}, {"fh_stencil", "grid_fh_stencil"},
                                {"alpha", "grid_alpha"}, {"delta", "grid_delta"}};
        for (auto& r : raw) {
            const std::string* v = get(r[0]);
            if (!v) continue;
            const std::string k = r[0];
            if (k == "Tmax" || k == "alpha" || k == "delta") {
                char* end = nullptr;
                const double d = std::strtod(v->c_str(), &end);
                if (end != v->c_str()) { std::fprintf(f, "%s = %.16g\n", r[1], d); continue; }
            }
            std::fprintf(f, "%s = %s\n", r[1], v->c_str());
        }
        std::fprintf(f, "\n");
    } else {
        std::fprintf(f, "# grid_params.txt not found for len=%lld\n\n", o.len);
    }
    std::fprintf(f, "# Current Simulation State\ncurrent_time = %.16g\ncurrent_loop = %lld\ncurrent_delta = %.16g\n",
                 t_last, (long long)loop, delta);
    // io_save_params.cpp:234: RK54, SSPRK104, or SERK2(q) with q = 2 (method - 2)
    const std::string mname = method == DMFE_RK54       ? std::string("RK54")
                              : method == DMFE_SSPRK104 ? std::string("SSPRK104")
                                                        : "SERK2(" + std::to_string(2 * (method - 2)) + ")";
    std::fprintf(f, "current_delta_t = %.16g\ncurrent_method = %s\ncurrent_t1grid_size = %lld\n", dt, mname.c_str(),
                 (long long)T1);
    std::fprintf(f, "current_QK0 = %.16g\ncurrent_QR0 = %.16g\ncurrent_r = %.16g\ncurrent_energy = %.16g\n",
                 qk[(T1 - 1) * o.len], qr[(T1 - 1) * o.len], r[T1 - 1], energy);
    std::fclose(f);
}

}  // namespace

int main(int argc, char** argv) {
    Options o;
    if (!parse(argc, argv, o)) return 2;
    // grid package
    int64_t N = 0, m = 0;
    if (!check(dmfe_grid_read_info(o.grid_dir.c_str(), &N, &m), "dmfe_grid_read_info")) return 1;
    if (N != o.len) {
        std::fprintf(stderr, "grid %s has length %lld, -L is %lld\n", o.grid_dir.c_str(), (long long)N, o.len);
        return 1;
    }
    const size_t P = (size_t)N * N;
    std::vector<double> theta(N), integ(N), phi2(P), weights(3 * P * m);
    std::vector<int64_t> inds(3 * P);
    if (!check(dmfe_grid_read_dir(o.grid_dir.c_str(), N, m, theta.data(), integ.data(), phi2.data(), inds.data(),
                                  weights.data()), "dmfe_grid_read_dir"))
        return 1;
    dmfe_grid g{};
    g.N = N;
    g.m = m;
    g.theta = theta.data();
    g.integ = integ.data();
    g.phi2 = phi2.data();
    for (int f = 0; f < 3; ++f) {
        g.inds[f] = inds.data() + f * P;
        g.weights[f] = weights.data() + f * P * m;
    }
    dmfe_params prm;
    dmfe_params_default(&prm);
    prm.p = o.p; prm.p2 = o.p2; prm.lambda = o.lambda; prm.T0 = o.T0; prm.Gamma = o.Gamma;
    prm.delta_t_min = o.dt_min; prm.delta_max = o.delta_max;
    prm.log_response = o.log_response ? 1 : 0;
    prm.maint_every = (int32_t)o.maint_every;
    dmfe_ctx* ctx = nullptr;
    if (!check(dmfe_ctx_create(&g, &prm, 4096, &ctx), "dmfe_ctx_create")) return 1;
    if (!check(dmfe_set_serk2(ctx, o.serk2 ? 1 : 0), "dmfe_set_serk2") ||
        !check(dmfe_set_rollback(ctx, o.rollback ? 1 : 0), "dmfe_set_rollback") ||
        !check(dmfe_set_tail_fit(ctx, o.tail_fit ? 1 : 0), "dmfe_set_tail_fit"))
        return 1;
    const std::string dir = param_dir(o);
    const std::string bin = dir + "/data.bin";
    bool resumed = false;
    if (o.save && !mkdirs(dir)) {
        std::fprintf(stderr, "cannot create %s\n", dir.c_str());
        return 1;
    }
    if (o.resume) {
        FILE* probe = std::fopen(bin.c_str(), "rb");
        if (probe) {
            std::fclose(probe);
            double e = 0;
            if (!check(dmfe_load_state(ctx, bin.c_str(), &e), "dmfe_load_state")) return 1;
            resumed = true;
            std::printf("resumed from %s (energy %.14f)\n", bin.c_str(), e);
        }
    }
    if (!resumed && !check(dmfe_init_quench(ctx), "dmfe_init_quench")) return 1;
    FILE* corr = nullptr;
    if (o.save) {
        corr = std::fopen((dir + "/correlation.txt").c_str(), "a");
        if (!corr) {
            std::fprintf(stderr, "cannot open correlation.txt in %s\n", dir.c_str());
            return 1;
        }
    }
    std::signal(SIGINT, on_signal);
    std::signal(SIGTERM, on_signal);
    std::signal(SIGHUP, on_signal);
    struct timespec t0s;
    clock_gettime(CLOCK_MONOTONIC, &t0s);
    std::vector<double> t1;
    int64_t T1 = 0;
    dmfe_get_T1(ctx, &T1);
    t1.resize(T1);
    dmfe_get_state(ctx, t1.data(), nullptr, nullptr, nullptr, nullptr, nullptr, nullptr, nullptr);
    double t = t1.back(), E = 0.0, qk0 = 0.0;
    double dt = 0, delta = 0, delta_old = 0;
    int64_t loop = 0;
    int method = 0;
    bool aborted = false;
    // rollbacks move the loop counter back by up to 10, so a call may run up to 11 times
    // as many iterations as it advances the counter
    const int64_t seg = kSegment;
    const int64_t cap = o.rollback ? 11 * kSegment + 16 : kSegment + 16;
    std::vector<double> rec(3 * cap);
    const double eps = 2.220446049250313e-16;
    // runSimulation's while-loop (simulation_runner.cpp:125) in dmfe_run calls of at most seg
    // iterations, each ending at the next checkpoint
    for (;;) {
        if (g_interrupted) {
            aborted = true;
            break;
        }
        dmfe_get_control(ctx, &dt, &delta, &delta_old, &loop, &method);
        if (!(t < o.tmax && loop < o.max_loop && dt >= o.dt_min - eps * std::fmax(std::fabs(dt), std::fabs(o.dt_min))))
            break;
        int64_t lim = std::min<int64_t>(o.max_loop, loop + seg);
        if (o.save && o.save_every > 0) lim = std::min<int64_t>(lim, (loop / o.save_every + 1) * o.save_every);
        int64_t n = 0;
        if (!check(dmfe_run(ctx, o.tmax, lim, rec.data(), cap, &n), "dmfe_run")) return 1;
        for (int64_t k = 0; k < n && k < cap; ++k) {
            t = rec[3 * k]; E = rec[3 * k + 1]; qk0 = rec[3 * k + 2];
            if (corr) std::fprintf(corr, "%.14f\t%.14f\t%.14f\n", t, E, qk0);
        }
        dmfe_get_control(ctx, &dt, &delta, &delta_old, &loop, &method);
        if (o.save && o.save_every > 0 && n > 0 && loop % o.save_every == 0) {
            std::fflush(corr);
            if (!check(dmfe_save_state(ctx, bin.c_str()), "dmfe_save_state")) return 1;
        }
        if (n == 0) break;
    }
    struct timespec t1s;
    clock_gettime(CLOCK_MONOTONIC, &t1s);
    const double elapsed = (t1s.tv_sec - t0s.tv_sec) + 1e-9 * (t1s.tv_nsec - t0s.tv_nsec);
    dmfe_get_control(ctx, &dt, &delta, &delta_old, &loop, &method);
    std::printf("%sloop %lld  t %.14f  E %.14f  QK(t,0) %.14f  dt %.6g  %.2f loops/s\n",
                aborted ? "interrupted: " : "", (long long)loop, t, E, qk0, dt, elapsed > 0 ? loop / elapsed : 0.0);
    if (o.save) {
        std::fclose(corr);
        if (!check(dmfe_save_state(ctx, bin.c_str()), "dmfe_save_state")) return 1;
        dmfe_get_T1(ctx, &T1);
        t1.resize(T1);
        dmfe_get_state(ctx, t1.data(), nullptr, nullptr, nullptr, nullptr, nullptr, nullptr, nullptr);
        double Elast = 0, q = 0;
        dmfe_observables(ctx, t1.back(), &Elast, &q);
        write_params(dir, o, ctx, t1.back(), Elast, elapsed);
    }
    dmfe_ctx_destroy(ctx);
    return aborted ? 130 : 0;
}
