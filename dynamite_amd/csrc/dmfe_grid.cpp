// Grid-package generator and Grid_data/<L> text I/O.
//
// Produces, for a length-N theta grid, everything the step path reads:
//   theta (N), int.dat quadrature weights (N), phi1/phi2 (N*N),
//   pos grids posA1y/posA2y/posB2y (N*N, written for format parity only),
//   stencil starts + weights for the three target families A1 (phi1),
//   A2 (phi2) and B2 (theta_i / phi2_ij).
// Host code only (long double, 80-bit on x86-64); no GPU calls.
//
// Algorithms follow the reference's grid generator (src/grid/*.cpp):
//   theta map           theta_grid.cpp:37-125   (atan/exp map with Lambert W_{-1} scale)
//   Lambert W_{-1}      lambert_w.cpp:8-44       (Halley iteration)
//   phi grids           phi_grid.cpp:17-34
//   int.dat             integration.cpp:14-195   (clamped B-spline quadrature, A^T w = M)
//   pos grids           pos_grid.cpp:113-206     (oversampled inverse, local barycentric)
//   poly stencils       interpolation_weights.cpp:12-147
//   rational stencils   interpolation_weights.cpp:159-279 (local Floater-Hormann)
//   dense B-spline      interpolation_weights.cpp:292-456
//   file formats        grid_io.cpp:26-328
#include "dmfe_internal.h"

#include <algorithm>
#include <cerrno>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <limits>
#include <stdexcept>
#include <string>
#include <sys/stat.h>
#include <thread>
#include <vector>

namespace dmfe {
namespace {

typedef long double ld;
const ld kPi = 3.141592653589793238462643383279502884L;

// Lower branch of Lambert W for x in [-1/e, 0): Halley's method on w e^w = x,
// started from the asymptotic series near 0- or the branch-point series near -1/e.
ld lambert_wm1(ld x) {
    if (!(x <= 0)) throw std::domain_error("lambert_wm1: x must be <= 0");
    if (x < -1.0L / std::exp(1.0L)) throw std::domain_error("lambert_wm1: x must be >= -1/e");
    if (x == 0) return -std::numeric_limits<double>::infinity();
    ld w;
    if (x > -0.1L) {
        const ld a = std::log(-x), b = std::log(-a);
        w = a - b + (b / a);
    } else {
        ld u = x * std::exp(1.0L) + 1.0L;
        u = std::min<ld>(0, std::max<ld>(-1, u));
        const ld s = std::sqrt(-2.0L * u);
        w = -1.0L - s + (s * s) / 3.0L;
    }
    for (int it = 0; it < 80; ++it) {
        const ld e = std::exp(w);
        const ld f = w * e - x;
        const ld w1 = w + 1.0L;
        ld step = f / (e * w1 - (w1 + 1.0L) * f / (2.0L * w1));
        ld next = w - step;
        if (!std::isfinite((double)next)) {  // fall back to a Newton step
            step = f / (e * w1);
            next = w - step;
        }
        w = next;
        if (std::fabs(step) <= std::numeric_limits<double>::epsilon()) break;
    }
    return w;
}

// Optional smooth odd re-mapping of the 1-based index (alpha = 0: identity).
ld remap_index(ld x, std::size_t len, double alpha, double delta) {
    if (alpha == 0.0) return x;
    const ld N = (ld)len;
    const ld half = (N - 1.0L) / 2.0L;
    ld s = (2.0L * x - N - 1.0L) / (N - 1.0L);
    s = std::min<ld>(1.0L, std::max<ld>(-1.0L, s));
    const ld d = (ld)delta;
    const ld mag = std::pow(std::fabs(s) * std::fabs(s) * std::fabs(s) + d * d * d, 1.0L / 3.0L) - d;
    const ld g = (s < 0 ? -1.0L : (s > 0 ? 1.0L : 0.0L)) * mag;
    const ld gnorm = std::pow(1.0L + d * d * d, 1.0L / 3.0L) - d;
    ld y = (ld)alpha * (half * ((g / gnorm) + 1.0L) + 1.0L) + (1.0L - (ld)alpha) * x;
    return std::min<ld>(N, std::max<ld>(1.0L, y));
}

// theta at 0-based fractional indices.  theta(x) = [g(s_1) - g(s_x)] / D with
// s_x = (x - N/2 - 1/2) * eta, g(s) = (2/pi) atan(e^{-s}), eta = -(2/N) W_{-1}(-1/(10 Tmax)).
void theta_values(const std::vector<double>& idx, std::size_t len, double Tmax, double alpha, double delta,
                  std::vector<ld>& out) {
    const ld N = (ld)len;
    const ld eta = -(2.0L / N) * lambert_wm1(-1.0L / (10.0L * (ld)Tmax));
    const ld two_pi = 2.0L / kPi;
    const ld s_first = (1.0L - (N / 2.0L) - 1.0L / 2.0L) * eta;
    const ld g_first = two_pi * std::atan(1.0L / std::exp(s_first));
    const ld s_half = ((N / 2.0L) - 1.0L / 2.0L) * eta;
    const ld D = two_pi * std::atan(std::exp(s_half)) - two_pi * std::atan(std::exp(-s_half));
    out.resize(idx.size());
    for (std::size_t k = 0; k < idx.size(); ++k) {
        const ld x = (ld)(double)remap_index((ld)idx[k] + 1.0L, len, alpha, delta);
        const ld s = (x - (N / 2.0L) - 1.0L / 2.0L) * eta;
        const ld g = two_pi * std::atan(1.0L / std::exp(s));
        ld th = ((1.0L - g) - (1.0L - g_first)) / D;
        out[k] = std::min<ld>(1.0L, std::max<ld>(0.0L, th));
    }
}

// ---- clamped B-splines (The NURBS Book, algorithms A2.1 / A2.2) --------------------
std::vector<ld> averaged_knots(const std::vector<ld>& x, int p) {
    const int N = (int)x.size();
    if (N < p + 1) throw std::invalid_argument("B-spline: need N >= p+1");
    std::vector<ld> t(N + p + 1);
    for (int i = 0; i <= p; ++i) { t[i] = x.front(); t[N + p - i] = x.back(); }
    for (int j = p + 1; j <= N - 1; ++j) {
        ld acc = 0;
        for (int i = j - p; i <= j - 1; ++i) acc += x[i];
        t[j] = acc / (ld)p;
    }
    return t;
}

int knot_span(int nb, int p, ld u, const std::vector<ld>& t) {
    const int n = nb - 1;
    if (u >= t[n + 1]) return n;
    if (u <= t[p]) return p;
    int lo = p, hi = n + 1, mid = (lo + hi) / 2;
    while (!(u >= t[mid] && u < t[mid + 1])) {
        if (u < t[mid]) hi = mid; else lo = mid;
        mid = (lo + hi) / 2;
    }
    return mid;
}

void basis_values(int span, ld u, int p, const std::vector<ld>& t, std::vector<ld>& B) {
    B.assign(p + 1, 0);
    std::vector<ld> left(p + 1), right(p + 1);
    B[0] = 1;
    for (int j = 1; j <= p; ++j) {
        left[j] = u - t[span + 1 - j];
        right[j] = t[span + j] - u;
        ld saved = 0;
        for (int r = 0; r < j; ++r) {
            const ld den = right[r + 1] + left[j - r];
            const ld q = (den != 0) ? (B[r] / den) : 0;
            B[r] = saved + right[r + 1] * q;
            saved = left[j - r] * q;
        }
        B[j] = saved;
    }
}

// Collocation matrix A_ij = B_j(x_i), dense row-major (band |i-j| <= p is non-zero).
std::vector<ld> collocation(const std::vector<ld>& x, int p, const std::vector<ld>& t) {
    const int N = (int)x.size();
    std::vector<ld> A((std::size_t)N * N, 0), B;
    for (int i = 0; i < N; ++i) {
        const int sp = knot_span(N, p, x[i], t);
        basis_values(sp, x[i], p, t, B);
        for (int j = 0; j <= p; ++j) {
            const int col = sp - p + j;
            if (col >= 0 && col < N) A[(std::size_t)i * N + col] = B[j];
        }
    }
    return A;
}

}  // namespace

// ---- int.dat: w solves A^T w = M, M_j = (t_{j+p+1} - t_j)/(p+1) (exact basis integrals) --
void grid_quadrature(const std::vector<long double>& x, int order, std::vector<long double>& w) {
    const int N = (int)x.size();
    w.clear();
    if (N == 0) return;
    int p = std::max(1, order);
    if (p > N - 1) p = std::max(1, N - 1);
    const std::vector<ld> t = averaged_knots(x, p);
    std::vector<ld> A = collocation(x, p, t);
    // banded LU without pivoting, half-bandwidth p
    for (int k = 0; k < N; ++k) {
        const ld piv = A[(std::size_t)k * N + k];
        const int iend = std::min(N - 1, k + p);
        for (int i = k + 1; i <= iend; ++i) {
            const ld l = (A[(std::size_t)i * N + k] /= piv);
            for (int j = k + 1; j <= iend; ++j) A[(std::size_t)i * N + j] -= l * A[(std::size_t)k * N + j];
        }
    }
    std::vector<ld> y(N);
    for (int j = 0; j < N; ++j) y[j] = (t[j + p + 1] - t[j]) / (ld)(p + 1);
    // U^T z = M (forward), then L^T w = z (backward, unit diagonal)
    for (int i = 0; i < N; ++i) {
        for (int j = std::max(0, i - p); j < i; ++j) y[i] -= A[(std::size_t)j * N + i] * y[j];
        y[i] /= A[(std::size_t)i * N + i];
    }
    for (int i = N - 1; i >= 0; --i)
        for (int j = i + 1; j <= std::min(N - 1, i + p); ++j) y[i] -= A[(std::size_t)j * N + i] * y[j];
    w = y;
}

void grid_theta(std::size_t len, double Tmax, double alpha, double delta, std::vector<long double>& theta) {
    std::vector<double> idx(len);
    for (std::size_t i = 0; i < len; ++i) idx[i] = (double)i;
    theta_values(idx, len, Tmax, alpha, delta, theta);
    for (std::size_t i = 1; i < len; ++i) theta[i] = std::max(theta[i], theta[i - 1]);
}

// ---- local stencils --------------------------------------------------------------
namespace {

// Window of width m around q minimising the larger distance from q to the window ends.
// The scan starts two slots left of the centred window and its right end follows
// the best window found so far (the bound is re-read every iteration, as in
// interpolation_weights.cpp:29 and :65), so it can walk further right.
int window_start(const std::vector<ld>& x, int m, ld q) {
    const int N = (int)x.size();
    const int hi = (int)(std::lower_bound(x.begin(), x.end(), q) - x.begin());
    int s = std::min(std::max(hi - m / 2, 0), N - m);
    auto spread = [&](int st) {
        const ld l = std::fabs(q - x[st]), r = std::fabs(q - x[st + m - 1]);
        return l < r ? r : l;
    };
    ld best = spread(s);
    for (int c = std::max(0, s - 2); c <= std::min(N - m, s + 2); ++c) {
        const ld v = spread(c);
        if (v < best) { best = v; s = c; }
    }
    return s;
}

void node_weights(const ld* nodes, int m, std::vector<ld>& w) {
    w.assign(m, 1);
    for (int j = 0; j < m; ++j) {
        ld den = 1;
        for (int k = 0; k < m; ++k)
            if (k != j) den *= (nodes[j] - nodes[k]);
        w[j] = 1 / den;
    }
}

void fh_weights(const std::vector<ld>& y, int d, std::vector<ld>& w) {
    const int m = (int)y.size();
    w.assign(m, 0);
    for (int j = 0; j < m; ++j) {
        ld acc = 0;
        for (int r = std::max(0, j - (m - d - 1)); r <= std::min(d, j); ++r) {
            const int i = j - r;
            ld prod = 1;
            bool dup = false;
            for (int k = 0; k <= d; ++k) {
                if (k == r) continue;
                const ld diff = y[j] - y[i + k];
                if (diff == 0) { dup = true; break; }
                prod *= diff;
            }
            if (dup || prod == 0) continue;
            acc += (((j - r) & 1) ? -1.0L : 1.0L) / prod;
        }
        w[j] = acc;
    }
}

// Neumaier-compensated sum of v_j = wj/(q - yj); also returns sum |v_j|.
ld comp_den(const std::vector<ld>& wts, const std::vector<ld>& y, ld q, std::vector<ld>& v, ld* sumabs) {
    ld s = 0, comp = 0, sa = 0;
    for (std::size_t j = 0; j < y.size(); ++j) {
        const ld t = wts[j] / (q - y[j]);
        v[j] = t;
        sa += std::fabs(t);
        const ld n = s + t;
        comp += (std::fabs(s) >= std::fabs(t)) ? (s - n) + t : (t - n) + s;
        s = n;
    }
    if (sumabs) *sumabs = sa;
    return s + comp;
}

}  // namespace

void grid_poly_stencils(const std::vector<long double>& x, const std::vector<long double>& q, int n,
                        int64_t* starts, double* alpha) {
    const int m = n + 1;
    std::vector<ld> wn;
    std::vector<ld> v(m);
    for (std::size_t k = 0; k < q.size(); ++k) {
        const int s = window_start(x, m, q[k]);
        starts[k] = s;
        double* a = alpha + k * m;
        int hit = -1;
        for (int j = 0; j < m; ++j)
            if (q[k] == x[s + j]) { hit = j; break; }
        if (hit >= 0) {
            for (int j = 0; j < m; ++j) a[j] = (j == hit) ? 1.0 : 0.0;
            continue;
        }
        node_weights(&x[s], m, wn);
        ld den = 0;
        for (int j = 0; j < m; ++j) { v[j] = wn[j] / (q[k] - x[s + j]); den += v[j]; }
        const ld inv = 1 / den;
        for (int j = 0; j < m; ++j) a[j] = (double)(v[j] * inv);
    }
}

void grid_rational_stencils(const std::vector<long double>& x, const std::vector<long double>& q, int d, int m,
                            int64_t* starts, double* alpha) {
    d = std::max(d, 0);
    m = std::max(m, d + 1);
    std::vector<ld> y(m), wfh, wpl, v(m);
    const ld inf = std::numeric_limits<ld>::infinity();
    for (std::size_t k = 0; k < q.size(); ++k) {
        const int s = window_start(x, m, q[k]);
        starts[k] = s;
        double* a = alpha + k * m;
        const ld span = x[s + m - 1] - x[s];
        const ld tol = 64 * std::numeric_limits<ld>::epsilon() * (span > 0 ? span : 1.0L);
        int hit = -1;
        for (int j = 0; j < m; ++j)
            if (std::fabs(q[k] - x[s + j]) <= tol) { hit = j; break; }
        if (hit >= 0) {
            for (int j = 0; j < m; ++j) a[j] = (j == hit) ? 1.0 : 0.0;
            continue;
        }
        const ld c = 0.5L * (x[s] + x[s + m - 1]);
        ld sc = x[s + m - 1] - x[s];
        if (sc <= 0) sc = 1;
        for (int j = 0; j < m; ++j) y[j] = (x[s + j] - c) / sc;
        const ld qn = (q[k] - c) / sc;

        fh_weights(y, d, wfh);
        ld sabs = 0;
        const ld dfh = comp_den(wfh, y, qn, v, &sabs);
        std::vector<double> afh(m, 0.0), apl(m, 0.0);
        ld leb_fh = inf, leb_pl = inf;
        if (dfh == dfh && std::isfinite(dfh)) {
            const ld inv = 1 / dfh;
            leb_fh = 0;
            for (int j = 0; j < m; ++j) { afh[j] = (double)(v[j] * inv); leb_fh += std::fabs(afh[j]); }
        }
        node_weights(y.data(), m, wpl);
        const ld dpl = comp_den(wpl, y, qn, v, nullptr);
        if (dpl == dpl && std::isfinite(dpl)) {
            const ld inv = 1 / dpl;
            leb_pl = 0;
            for (int j = 0; j < m; ++j) { apl[j] = (double)(v[j] * inv); leb_pl += std::fabs(apl[j]); }
        }
        bool poly = !(dfh == dfh) || !std::isfinite(dfh) || (sabs > 0 && std::fabs(dfh) <= 1e-10L * sabs);
        if (!poly && leb_pl < inf && leb_fh < inf && leb_fh > 2.0L * leb_pl) poly = true;
        const std::vector<double>& src = poly ? apl : afh;
        for (int j = 0; j < m; ++j) a[j] = src[j];
    }
}

// LU factorisation with partial pivoting (long double) of A^T, A_ij = B_j(x_i) the
// collocation matrix at the averaged knots t: the system every dense B-spline weight solves.
void bspline_lu(const std::vector<long double>& x, int p, const std::vector<ld>& t, std::vector<ld>& LU,
                std::vector<int>& piv) {
    const int N = (int)x.size();
    const std::vector<ld> A = collocation(x, p, t);
    LU.assign((std::size_t)N * N, 0.0L);
    for (int i = 0; i < N; ++i)
        for (int j = 0; j < N; ++j) LU[(std::size_t)i * N + j] = A[(std::size_t)j * N + i];
    piv.resize(N);
    for (int i = 0; i < N; ++i) piv[i] = i;
    for (int k = 0; k < N; ++k) {
        int r = k;
        ld best = std::fabs(LU[(std::size_t)k * N + k]);
        for (int i = k + 1; i < N; ++i) {
            const ld v = std::fabs(LU[(std::size_t)i * N + k]);
            if (v > best) { best = v; r = i; }
        }
        if (best == 0) throw std::runtime_error("B-spline collocation matrix is singular");
        if (r != k) {
            for (int j = 0; j < N; ++j) std::swap(LU[(std::size_t)k * N + j], LU[(std::size_t)r * N + j]);
            std::swap(piv[k], piv[r]);
        }
        for (int i = k + 1; i < N; ++i) {
            LU[(std::size_t)i * N + k] /= LU[(std::size_t)k * N + k];
            const ld l = LU[(std::size_t)i * N + k];
            for (int j = k + 1; j < N; ++j) LU[(std::size_t)i * N + j] -= l * LU[(std::size_t)k * N + j];
        }
    }
}

// Dense interpolating B-spline of degree p through all N nodes: the weights of
// query q solve A^T w = e(q), with A_ij = B_j(x_i) and e(q) the p+1 non-zero
// basis values at q.  A^T is LU-factorised once (partial pivoting, long double)
// and every query is a forward/back substitution, as in
// interpolation_weights.cpp:401-456; queries are spread over host threads
// (each query's arithmetic is unchanged, so results do not depend on the split).
void grid_bspline_weights(const std::vector<long double>& x, const std::vector<long double>& q, int p,
                          double* W) {
    const int N = (int)x.size();
    const std::vector<ld> t = averaged_knots(x, p);
    std::vector<ld> LU;
    std::vector<int> piv;
    bspline_lu(x, p, t, LU, piv);
    const std::size_t Q = q.size();
    auto work = [&](std::size_t q0, std::size_t q1) {
        std::vector<ld> rhs(N), sol(N), B;
        for (std::size_t k = q0; k < q1; ++k) {
            std::fill(rhs.begin(), rhs.end(), 0.0L);
            const int sp = knot_span(N, p, q[k], t);
            basis_values(sp, q[k], p, t, B);
            for (int j = 0; j <= p; ++j) {
                const int col = sp - p + j;
                if (col >= 0 && col < N) rhs[col] = B[j];
            }
            for (int i = 0; i < N; ++i) sol[i] = rhs[piv[i]];
            for (int i = 0; i < N; ++i)
                for (int j = 0; j < i; ++j) sol[i] -= LU[(std::size_t)i * N + j] * sol[j];
            for (int i = N - 1; i >= 0; --i) {
                for (int j = i + 1; j < N; ++j) sol[i] -= LU[(std::size_t)i * N + j] * sol[j];
                sol[i] /= LU[(std::size_t)i * N + i];
            }
            for (int i = 0; i < N; ++i) W[k * N + i] = (double)sol[i];
        }
    };
    unsigned nt = std::max(1u, std::min(64u, std::thread::hardware_concurrency()));
    if (const char* e = std::getenv("DMFE_GRID_THREADS")) nt = std::max(1, std::atoi(e));
    if (Q < 256) nt = 1;
    std::vector<std::thread> th;
    for (unsigned i = 0; i < nt; ++i) th.emplace_back(work, Q * i / nt, Q * (i + 1) / nt);
    for (auto& x_ : th) x_.join();
}

// The same weights formed on the current HIP device.  w(q) = (A^T)^{-1} e(q) is the
// combination of the p+1 rows of Z = ((A^T)^{-1})^T picked by the non-zero basis values of
// q: Z is formed once on the host in long double (N solves of the LU factorisation with
// unit vectors, over host threads) and rounded to double; the host evaluates each query's
// knot span and basis values; the device forms the N^2 combinations of N entries per table
// (launch_bspline_combine) in chunks and the rows are copied into W.  The explicit inverse
// and the double-precision combination change the last bits against the per-query solve
// (tested against it and by polynomial reproduction).
void grid_bspline_weights_device(const std::vector<long double>& x, const std::vector<long double>& q, int p,
                                 double* W) {
    const int N = (int)x.size();
    const std::vector<ld> t = averaged_knots(x, p);
    std::vector<ld> LU;
    std::vector<int> piv;
    bspline_lu(x, p, t, LU, piv);
    std::vector<double> Z((std::size_t)N * N);  // row c: column c of (A^T)^{-1}
    unsigned nt = std::max(1u, std::min(64u, std::thread::hardware_concurrency()));
    auto inv = [&](int c0, int c1) {
        std::vector<ld> sol(N);
        for (int c = c0; c < c1; ++c) {
            for (int i = 0; i < N; ++i) sol[i] = (piv[i] == c) ? 1.0L : 0.0L;
            for (int i = 0; i < N; ++i)
                for (int j = 0; j < i; ++j) sol[i] -= LU[(std::size_t)i * N + j] * sol[j];
            for (int i = N - 1; i >= 0; --i) {
                for (int j = i + 1; j < N; ++j) sol[i] -= LU[(std::size_t)i * N + j] * sol[j];
                sol[i] /= LU[(std::size_t)i * N + i];
            }
            for (int i = 0; i < N; ++i) Z[(std::size_t)c * N + i] = (double)sol[i];
        }
    };
    {
        std::vector<std::thread> th;
        for (unsigned i = 0; i < nt; ++i) th.emplace_back(inv, (int)((long long)N * i / nt), (int)((long long)N * (i + 1) / nt));
        for (auto& x_ : th) x_.join();
    }
    const std::size_t Q = q.size();
    const int nb = p + 1;
    std::vector<int> start(Q);
    std::vector<double> basis(Q * (std::size_t)nb);
    auto spans = [&](std::size_t q0, std::size_t q1) {
        std::vector<ld> B;
        for (std::size_t k = q0; k < q1; ++k) {
            const int sp = knot_span(N, p, q[k], t);
            basis_values(sp, q[k], p, t, B);
            int s0 = sp - p;
            for (int j = 0; j <= p; ++j) {  // columns outside [0, N) carry no weight
                const int col = s0 + j;
                basis[k * nb + j] = (col >= 0 && col < N) ? (double)B[j] : 0.0;
            }
            start[k] = s0;
        }
    };
    {
        std::vector<std::thread> th;
        for (unsigned i = 0; i < nt; ++i) th.emplace_back(spans, Q * i / nt, Q * (i + 1) / nt);
        for (auto& x_ : th) x_.join();
    }
    double *dZ = nullptr, *dB = nullptr, *dW = nullptr;
    int* dS = nullptr;
    const std::size_t chunk = std::min<std::size_t>(Q, std::max<std::size_t>(1, (std::size_t)(256u << 20) / ((std::size_t)N * 8)));
    auto check = [&](hipError_t e, const char* what) {
        if (e != hipSuccess) {
            (void)hipFree(dZ); (void)hipFree(dB); (void)hipFree(dW); (void)hipFree(dS);
            throw std::runtime_error(std::string("B-spline weights on the device: ") + what + ": " + hipGetErrorString(e));
        }
    };
    check(hipMalloc(&dZ, Z.size() * sizeof(double)), "hipMalloc");
    check(hipMalloc(&dB, basis.size() * sizeof(double)), "hipMalloc");
    check(hipMalloc(&dS, start.size() * sizeof(int)), "hipMalloc");
    check(hipMalloc(&dW, chunk * N * sizeof(double)), "hipMalloc");
    check(hipMemcpy(dZ, Z.data(), Z.size() * sizeof(double), hipMemcpyHostToDevice), "upload");
    check(hipMemcpy(dB, basis.data(), basis.size() * sizeof(double), hipMemcpyHostToDevice), "upload");
    check(hipMemcpy(dS, start.data(), start.size() * sizeof(int), hipMemcpyHostToDevice), "upload");
    for (std::size_t q0 = 0; q0 < Q; q0 += chunk) {
        const std::size_t nq = std::min(chunk, Q - q0);
        launch_bspline_combine(dZ, dS + q0, dB + q0 * nb, p, N, (long long)nq, dW, nullptr);
        check(hipGetLastError(), "launch");
        check(hipMemcpy(W + q0 * N, dW, nq * N * sizeof(double), hipMemcpyDeviceToHost), "download");
    }
    check(hipDeviceSynchronize(), "synchronize");
    (void)hipFree(dZ); (void)hipFree(dB); (void)hipFree(dW); (void)hipFree(dS);
}

// ---- pos grids (theta value -> fractional 1-based index), written for format parity ----
// The theta map x -> theta(x) of theta_values is smooth and increasing in the continuous
// index x in [0, N-1].  Its inverse at a value v: the map is tabulated at kSub points per
// index interval, v is located in the table, and the bracketing table interval is shrunk
// by safeguarded secant steps on the map itself (regula falsi with the Illinois
// modification, all queries evaluated together per step) until the bracket is below
// kTol index units or kIter steps have run.  Values at or beyond the ends of the map give
// positions 1 and N.  The result is the position y = x + 1 with theta(x) = v to the
// accuracy of the map's evaluation.
namespace {
void invert_theta_map(std::size_t N, double Tmax, double alpha, double delta, const std::vector<double>& tab_x,
                      const std::vector<ld>& tab, const std::vector<ld>& q, double* out) {
    constexpr int kIter = 8;
    constexpr double kTol = 1e-13;
    const std::size_t Q = q.size();
    std::vector<double> a(Q), b(Q), x(Q);
    std::vector<ld> fa(Q), fb(Q);
    std::vector<signed char> side(Q, 0), done(Q, 0);
    for (std::size_t k = 0; k < Q; ++k) {
        const ld v = q[k];
        if (!(v > tab.front())) { out[k] = 1.0; done[k] = 1; continue; }
        if (!(v < tab.back())) { out[k] = (double)N; done[k] = 1; continue; }
        const std::size_t i = (std::size_t)(std::upper_bound(tab.begin(), tab.end(), v) - tab.begin());  // tab[i-1] <= v < tab[i]
        a[k] = tab_x[i - 1]; fa[k] = tab[i - 1] - v;
        b[k] = tab_x[i];     fb[k] = tab[i] - v;
        if (fa[k] == 0) { out[k] = a[k] + 1.0; done[k] = 1; }
    }
    std::vector<double> xs;
    std::vector<std::size_t> who;
    std::vector<ld> fx;
    for (int it = 0; it < kIter; ++it) {
        xs.clear();
        who.clear();
        for (std::size_t k = 0; k < Q; ++k) {
            if (done[k]) continue;
            double c = (double)(((ld)a[k] * fb[k] - (ld)b[k] * fa[k]) / (fb[k] - fa[k]));
            if (!(c > a[k] && c < b[k])) c = 0.5 * (a[k] + b[k]);
            x[k] = c;
            xs.push_back(c);
            who.push_back(k);
        }
        if (who.empty()) break;
        theta_values(xs, N, Tmax, alpha, delta, fx);
        for (std::size_t r = 0; r < who.size(); ++r) {
            const std::size_t k = who[r];
            const ld f = fx[r] - q[k];
            if (f == 0) { a[k] = b[k] = x[k]; fa[k] = fb[k] = 0; }
            else if (f < 0) {
                a[k] = x[k]; fa[k] = f;
                if (side[k] < 0) fb[k] *= 0.5L;  // Illinois: the retained end keeps drawing the secant
                side[k] = -1;
            } else {
                b[k] = x[k]; fb[k] = f;
                if (side[k] > 0) fa[k] *= 0.5L;
                side[k] = 1;
            }
            if (b[k] - a[k] <= kTol) done[k] = 2;
        }
    }
    for (std::size_t k = 0; k < Q; ++k) {
        if (done[k] == 1) continue;
        const double c = (fb[k] == fa[k]) ? a[k] : (double)(((ld)a[k] * fb[k] - (ld)b[k] * fa[k]) / (fb[k] - fa[k]));
        out[k] = std::min(std::max(c, a[k]), b[k]) + 1.0;
    }
}
}  // namespace

void grid_pos(std::size_t N, double Tmax, double alpha, double delta, const std::vector<long double>& theta,
              const std::vector<long double>& phi1, const std::vector<long double>& phi2,
              double* posA1, double* posA2, double* posB2) {
    constexpr std::size_t kSub = 8;  // table points per index interval
    const std::size_t M = (N - 1) * kSub + 1;
    std::vector<double> tab_x(M);
    for (std::size_t k = 0; k < M; ++k) tab_x[k] = (double)k / (double)kSub;
    std::vector<ld> tab;
    theta_values(tab_x, N, Tmax, alpha, delta, tab);
    for (std::size_t k = 1; k < M; ++k)  // the clamp at theta = 0 and 1 leaves flat ends: keep them ordered
        if (tab[k] < tab[k - 1]) tab[k] = tab[k - 1];
    const std::size_t P = N * N;
    std::vector<ld> qB2(P);
    const ld tiny = 1e-200L;  // the B2 query point theta_i / (phi2 - 1e-200), as in the weight tables
    for (std::size_t i = 0; i < N; ++i)
        for (std::size_t j = 0; j < N; ++j) {
            const std::size_t k = i * N + j;
            const ld den = phi2[k] - tiny;
            qB2[k] = (std::fabs(den) < 1e-300L) ? ((theta[i] > 0) ? tab.back() : tab.front()) : theta[i] / den;
        }
    invert_theta_map(N, Tmax, alpha, delta, tab_x, tab, phi1, posA1);
    invert_theta_map(N, Tmax, alpha, delta, tab_x, tab, phi2, posA2);
    invert_theta_map(N, Tmax, alpha, delta, tab_x, tab, qB2, posB2);
}

}  // namespace dmfe

// ======================= C ABI ======================================================
using dmfe::set_error;

extern "C" int64_t dmfe_grid_stencil_width(int64_t N, int method, int order, int fh_stencil) {
    const int64_t n = std::max(1, order);
    if (method == DMFE_INTERP_POLY) return n + 1;
    if (method == DMFE_INTERP_RATIONAL) return std::min<int64_t>(N, std::max<int64_t>(n + 1, fh_stencil > 0 ? fh_stencil : n + 1));
    if (method == DMFE_INTERP_BSPLINE) return N;
    return -1;
}

extern "C" int dmfe_grid_theta_ld(int64_t N, double Tmax, double alpha, double delta, long double* theta) {
    if (N < 2 || !theta) return set_error(DMFE_ERR_ARG, "dmfe_grid_theta_ld: bad argument");
    try {
        std::vector<long double> th;
        dmfe::grid_theta((std::size_t)N, Tmax, alpha, delta, th);
        std::copy(th.begin(), th.end(), theta);
        return DMFE_OK;
    } catch (const std::exception& e) {
        return set_error(DMFE_ERR_ARG, std::string("dmfe_grid_theta_ld: ") + e.what());
    }
}

namespace {
int grid_generate(const dmfe_grid_spec* spec, double* theta, double* integ, double* phi1, double* phi2, double* posA1,
                  double* posA2, double* posB2, int64_t* inds, double* weights, bool device) {
    try {
        if (!spec || spec->N < 2 || !theta || !integ || !phi2 || !inds || !weights)
            return set_error(DMFE_ERR_ARG, "dmfe_grid_generate: null or invalid argument");
        const std::size_t N = (std::size_t)spec->N, P = N * N;
        const int64_t m = dmfe_grid_stencil_width(spec->N, spec->method, spec->order, spec->fh_stencil);
        if (m < 1 || (std::size_t)m > N) return set_error(DMFE_ERR_ARG, "dmfe_grid_generate: bad method/order");
        std::vector<long double> th, w, p1(P), p2(P);
        dmfe::grid_theta(N, spec->Tmax, spec->alpha, spec->delta, th);
        for (std::size_t i = 0; i < N; ++i) {
            for (std::size_t j = 0; j < N; ++j) {
                p1[i * N + j] = th[i] * th[j];
                p2[i * N + j] = th[i] + (1.0L - th[i]) * th[j];
            }
            theta[i] = (double)th[i];
        }
        dmfe::grid_quadrature(th, spec->spline_order, w);
        for (std::size_t i = 0; i < N; ++i) integ[i] = (double)w[i];
        for (std::size_t k = 0; k < P; ++k) {
            if (phi1) phi1[k] = (double)p1[k];
            phi2[k] = (double)p2[k];
        }
        if (posA1 && posA2 && posB2) dmfe::grid_pos(N, spec->Tmax, spec->alpha, spec->delta, th, p1, p2, posA1, posA2, posB2);
        std::vector<long double> qB2;
        qB2.reserve(P);
        const double tiny = 1e-200;  // a double constant, as in grid_cli.cpp:172
        for (std::size_t i = 0; i < N; ++i)
            for (std::size_t j = 0; j < N; ++j) {
                const long double den = p2[i * N + j] - tiny;
                qB2.push_back(std::fabs(den) < 1e-300L ? ((th[i] > 0) ? th.back() : th.front()) : th[i] / den);
            }
        const std::vector<long double>* qs[3] = {&p1, &p2, &qB2};
        const int n = std::max(1, spec->order);
        for (int f = 0; f < 3; ++f) {
            int64_t* st = inds + (std::size_t)f * P;
            double* wt = weights + (std::size_t)f * P * (std::size_t)m;
            if (spec->method == DMFE_INTERP_POLY) {
                dmfe::grid_poly_stencils(th, *qs[f], n, st, wt);
            } else if (spec->method == DMFE_INTERP_RATIONAL) {
                dmfe::grid_rational_stencils(th, *qs[f], n, (int)m, st, wt);
            } else {
                for (std::size_t k = 0; k < P; ++k) st[k] = -1;
                if (device) dmfe::grid_bspline_weights_device(th, *qs[f], n, wt);
                else dmfe::grid_bspline_weights(th, *qs[f], n, wt);
            }
        }
        return DMFE_OK;
    } catch (const std::exception& e) {
        return set_error(DMFE_ERR_ARG, std::string("dmfe_grid_generate: ") + e.what());
    }
}
}  // namespace

extern "C" int dmfe_grid_generate(const dmfe_grid_spec* spec, double* theta, double* integ, double* phi1,
                                  double* phi2, double* posA1, double* posA2, double* posB2,
                                  int64_t* inds, double* weights) {
    return grid_generate(spec, theta, integ, phi1, phi2, posA1, posA2, posB2, inds, weights, false);
}

extern "C" int dmfe_grid_generate_device(const dmfe_grid_spec* spec, double* theta, double* integ, double* phi1,
                                         double* phi2, double* posA1, double* posA2, double* posB2,
                                         int64_t* inds, double* weights) {
    int n = 0;
    if (hipGetDeviceCount(&n) != hipSuccess || n < 1)
        return set_error(DMFE_ERR_HIP, "dmfe_grid_generate_device: no HIP device");
    return grid_generate(spec, theta, integ, phi1, phi2, posA1, posA2, posB2, inds, weights, true);
}

// ---- text formats (grid_io.cpp:26-328, io_input.cpp:60-110) -------------------------
namespace {

bool write_vec(const std::string& path, const double* v, std::size_t n) {
    FILE* f = std::fopen(path.c_str(), "w");
    if (!f) return false;
    for (std::size_t i = 0; i < n; ++i) std::fprintf(f, "%.17e\n", v[i]);
    return std::fclose(f) == 0;
}
bool write_mat(const std::string& path, const double* v, std::size_t N) {
    FILE* f = std::fopen(path.c_str(), "w");
    if (!f) return false;
    for (std::size_t i = 0; i < N; ++i) {
        for (std::size_t j = 0; j < N; ++j) std::fprintf(f, j ? "\t%.17e" : "%.17e", v[i * N + j]);
        std::fputc('\n', f);
    }
    return std::fclose(f) == 0;
}
bool write_imat(const std::string& path, const int64_t* v, std::size_t N) {
    FILE* f = std::fopen(path.c_str(), "w");
    if (!f) return false;
    for (std::size_t i = 0; i < N; ++i) {
        for (std::size_t j = 0; j < N; ++j) std::fprintf(f, j ? "\t%lld" : "%lld", (long long)v[i * N + j]);
        std::fputc('\n', f);
    }
    return std::fclose(f) == 0;
}

// Whitespace-separated numbers; stops at the first token that does not parse,
// like the stream extraction loop of the reference's importers.
template <class T>
bool read_numbers(const std::string& path, std::vector<T>& out) {
    FILE* f = std::fopen(path.c_str(), "rb");
    if (!f) return false;
    std::fseek(f, 0, SEEK_END);
    const long sz = std::ftell(f);
    std::fseek(f, 0, SEEK_SET);
    std::string buf((std::size_t)std::max(0L, sz), '\0');
    const std::size_t got = std::fread(&buf[0], 1, buf.size(), f);
    std::fclose(f);
    buf.resize(got);
    out.clear();
    const char* p = buf.c_str();
    const char* end = p + buf.size();
    while (p < end) {
        char* e = nullptr;
        if (std::is_integral<T>::value) {
            errno = 0;
            const long long v = std::strtoll(p, &e, 10);
            if (e == p || errno) break;
            out.push_back((T)v);
        } else {
            const double v = std::strtod(p, &e);
            if (e == p) break;
            out.push_back((T)v);
        }
        p = e;
    }
    return true;
}

}  // namespace

extern "C" int dmfe_grid_write_dir(const char* dir, int64_t N, int64_t m, const double* theta, const double* integ,
                                   const double* phi1, const double* phi2, const double* posA1, const double* posA2,
                                   const double* posB2, const int64_t* inds, const double* weights,
                                   const char* params_text) {
    if (!dir || N < 1 || m < 1) return set_error(DMFE_ERR_ARG, "dmfe_grid_write_dir: bad argument");
    const std::string d(dir);
    ::mkdir(d.c_str(), 0755);
    const std::size_t n = (std::size_t)N, P = n * n;
    bool ok = write_vec(d + "/theta.dat", theta, n) && write_vec(d + "/int.dat", integ, n);
    if (ok && phi1) ok = write_mat(d + "/phi1.dat", phi1, n);
    if (ok && phi2) ok = write_mat(d + "/phi2.dat", phi2, n);
    if (ok && posA1) ok = write_mat(d + "/posA1y.dat", posA1, n);
    if (ok && posA2) ok = write_mat(d + "/posA2y.dat", posA2, n);
    if (ok && posB2) ok = write_mat(d + "/posB2y.dat", posB2, n);
    const char* in[3] = {"/indsA1y.dat", "/indsA2y.dat", "/indsB2y.dat"};
    const char* wn[3] = {"/weightsA1y.dat", "/weightsA2y.dat", "/weightsB2y.dat"};
    for (int f = 0; ok && f < 3; ++f) {
        ok = write_imat(d + in[f], inds + (std::size_t)f * P, n) &&
             write_vec(d + wn[f], weights + (std::size_t)f * P * (std::size_t)m, P * (std::size_t)m);
    }
    if (ok && params_text) {
        FILE* f = std::fopen((d + "/grid_params.txt").c_str(), "w");
        ok = f && std::fputs(params_text, f) >= 0;
        if (f) ok = (std::fclose(f) == 0) && ok;
    }
    return ok ? DMFE_OK : set_error(DMFE_ERR_IO, "dmfe_grid_write_dir: write failed under " + d);
}

extern "C" int dmfe_grid_read_info(const char* dir, int64_t* N, int64_t* m) {
    if (!dir || !N || !m) return set_error(DMFE_ERR_ARG, "dmfe_grid_read_info: null argument");
    const std::string d(dir);
    std::vector<double> th;
    std::vector<int64_t> ia;
    if (!read_numbers(d + "/theta.dat", th) || th.empty()) return set_error(DMFE_ERR_IO, "cannot read " + d + "/theta.dat");
    if (!read_numbers(d + "/indsA1y.dat", ia) || ia.size() != th.size() * th.size())
        return set_error(DMFE_ERR_IO, "cannot read " + d + "/indsA1y.dat (N*N integers expected)");
    // m = |weights| / |inds| (index_vec.cpp:19); count lines of the weight file
    FILE* f = std::fopen((d + "/weightsA1y.dat").c_str(), "rb");
    if (!f) return set_error(DMFE_ERR_IO, "cannot read " + d + "/weightsA1y.dat");
    std::size_t count = 0;
    char buf[1 << 16];
    bool intoken = false;
    std::size_t got;
    while ((got = std::fread(buf, 1, sizeof buf, f)) > 0)
        for (std::size_t i = 0; i < got; ++i) {
            const bool ws = buf[i] == ' ' || buf[i] == '\n' || buf[i] == '\t' || buf[i] == '\r';
            if (!ws && !intoken) ++count;
            intoken = !ws;
        }
    std::fclose(f);
    *N = (int64_t)th.size();
    *m = (int64_t)(count / ia.size());
    return DMFE_OK;
}

extern "C" int dmfe_grid_read_dir(const char* dir, int64_t N, int64_t m, double* theta, double* integ, double* phi2,
                                  int64_t* inds, double* weights) {
    if (!dir || N < 1 || m < 1 || !theta || !integ || !inds || !weights)
        return set_error(DMFE_ERR_ARG, "dmfe_grid_read_dir: bad argument");
    const std::string d(dir);
    const std::size_t n = (std::size_t)N, P = n * n;
    std::vector<double> v;
    auto load = [&](const char* name, double* dst, std::size_t cnt) -> bool {
        if (!read_numbers(d + name, v) || v.size() != cnt) {
            set_error(DMFE_ERR_IO, "dmfe_grid_read_dir: " + d + name + " has " + std::to_string(v.size()) +
                                       " values, expected " + std::to_string(cnt));
            return false;
        }
        std::memcpy(dst, v.data(), cnt * sizeof(double));
        return true;
    };
    if (!load("/theta.dat", theta, n) || !load("/int.dat", integ, n)) return DMFE_ERR_IO;
    if (phi2) {
        if (!read_numbers(d + "/phi2.dat", v) || v.size() != P) {
            // phi2 is fixed by theta (phi_grid.cpp:17-34); regenerate when the file is absent
            for (std::size_t i = 0; i < n; ++i)
                for (std::size_t j = 0; j < n; ++j)
                    phi2[i * n + j] = (double)((long double)theta[i] + (1.0L - (long double)theta[i]) * (long double)theta[j]);
        } else {
            std::memcpy(phi2, v.data(), P * sizeof(double));
        }
    }
    const char* in[3] = {"/indsA1y.dat", "/indsA2y.dat", "/indsB2y.dat"};
    const char* wn[3] = {"/weightsA1y.dat", "/weightsA2y.dat", "/weightsB2y.dat"};
    std::vector<int64_t> iv;
    for (int f = 0; f < 3; ++f) {
        if (!read_numbers(d + in[f], iv) || iv.size() != P)
            return set_error(DMFE_ERR_IO, std::string("dmfe_grid_read_dir: bad ") + d + in[f]);
        std::memcpy(inds + (std::size_t)f * P, iv.data(), P * sizeof(int64_t));
        if (!load(wn[f], weights + (std::size_t)f * P * (std::size_t)m, P * (std::size_t)m)) return DMFE_ERR_IO;
    }
    return DMFE_OK;
}
