// Device code of the gfx950 kernels, shared by the translation units that launch them
// (dmfe_kernels.hip, dmfe_rhs.hip, dmfe_loop.hip); internal linkage, so every unit
// compiles the kernels it instantiates.
//
// gfx950 kernels of the per-RK-stage DMFT update (FP64).
//
//   stage_kernel     one launch per stage boundary: the finish of the previous stage
//                    (block 0) together with the time positions of theta*t (B1) and
//                    phi2*t (B2) of the next interpolation - binary search or the seeded
//                    interpolation search, bit-exact with search_utils.cpp:13-131 - and
//                    the B1 prologue (Floor, diff; interpolation_core.cpp:57-84).
//   rhs_rows_kernel  one (row j, column chunk) per workgroup: rebuilds the A1/A2/B1/B2
//                    interpolants of row j (index_vec.cpp:15-156, index_mat.cpp:18-91), the
//                    self-energies (math_sigma.cpp:7-62) and the eight int.dat quadratures
//                    of QKstep/QRstep (time_steps.cpp:27-47,94-100) without materialising
//                    any N*N field in HBM.
//   finish_block     (block 0 of stage_kernel) RHS epilogue, RK stage combination
//                    (runge_kutta.cpp), append/replace of the history row
//                    (time_steps.cpp:49-81,127-162), rstep, drstep2 and the error norm.
//
// Memory layout (HBM): history as two row-major [T1][N] arrays of double2,
// V = (QK, QR) and D = (dQK, dQR), so every gather is one 16-byte load that
// serves both fields; stencil weights transposed to [m][N*N] so a wavefront's
// 64 lanes (64 consecutive columns of one row) read 512 contiguous bytes per tap.
#pragma once

#include "dmfe_internal.h"

namespace dmfe {
namespace {

constexpr int kWave = 64;

// ---- f and derivatives (math_ops.cpp:7-37) ----------------------------------------
__device__ inline double pow_int(double b, int e) {
    if (e == 0) return 1.0;
    if (e == 1) return b;
    if (e == 2) return b * b;
    if (e == 3) return b * b * b;
    if (e == 4) { const double sq = b * b; return sq * sq; }
    double r = 1.0, c = b;
    while (e > 0) {
        if (e & 1) r *= c;
        c *= c;
        e >>= 1;
    }
    return r;
}
__device__ inline double fl1(const Model& M, double q) {
    return M.lambda * M.p * pow_int(q, M.p - 1) + (1 - M.lambda) * M.p2 * pow_int(q, M.p2 - 1);
}
__device__ inline double fl2(const Model& M, double q) {
    return M.lambda * M.p * (M.p - 1) * pow_int(q, M.p - 2) + (1 - M.lambda) * M.p2 * (M.p2 - 1) * pow_int(q, M.p2 - 2);
}
__device__ inline double fl3(const Model& M, double q) {
    return M.lambda * M.p * (M.p - 1) * (M.p - 2) * pow_int(q, M.p - 3) +
           (1 - M.lambda) * M.p2 * (M.p2 - 1) * (M.p2 - 2) * pow_int(q, M.p2 - 3);
}

// ---- reductions -------------------------------------------------------------------
__device__ inline double wave_sum(double v) {
#pragma unroll
    for (int off = kWave / 2; off > 0; off >>= 1) v += __shfl_xor(v, off, kWave);
    return v;
}
__device__ inline double wave_max(double v) {
#pragma unroll
    for (int off = kWave / 2; off > 0; off >>= 1) v = fmax(v, __shfl_xor(v, off, kWave));
    return v;
}

// Sum over aligned groups of G lanes (G in 4, 8, 16, 32, 64) with DPP lane swizzles
// (quad_perm, row_half_mirror, row_mirror) and a cross-row shuffle for G >= 32.
// Every control used here reads an active lane of the same 16-lane row, so the
// destination needs no initial value.
template <int CTRL>
__device__ inline double dpp_f64(double v) {
    const long long b = __double_as_longlong(v);
    const int lo = __builtin_amdgcn_mov_dpp((int)(b & 0xffffffff), CTRL, 0xf, 0xf, true);
    const int hi = __builtin_amdgcn_mov_dpp((int)(b >> 32), CTRL, 0xf, 0xf, true);
    return __longlong_as_double(((long long)hi << 32) | (unsigned int)lo);
}
constexpr int kDppQuadSwap1 = 0xB1;     // quad_perm [1,0,3,2]: lane i <-> i^1
constexpr int kDppQuadSwap2 = 0x4E;     // quad_perm [2,3,0,1]: lane i <-> i^2
constexpr int kDppHalfMirror = 0x141;   // row_half_mirror: lane i <-> 7-i within 8
constexpr int kDppRowMirror = 0x140;    // row_mirror: lane i <-> 15-i within 16
// row_ror:n makes lane r of a 16-lane row read lane (r - n) mod 16
constexpr int kDppRowRor8 = 0x128;      // row_ror:8: lane r <- (r+8) mod 16 within the row

// Reduce-scatter of eight per-lane values over an aligned group of 8 lanes: on return
// lane (i mod 8) holds the group total of value (i mod 8).  Partners: 7-i, i^2, i^1.
__device__ inline double scatter8(const double (&v)[8], int i8) {
    const bool h = i8 & 4, b = i8 & 2, c = i8 & 1;
    double k4[4];
#pragma unroll
    for (int q = 0; q < 4; ++q) {
        const double keep = h ? v[4 + q] : v[q], send = h ? v[q] : v[4 + q];
        k4[q] = keep + dpp_f64<kDppHalfMirror>(send);
    }
    double k2[2];
#pragma unroll
    for (int q = 0; q < 2; ++q) {
        const double keep = b ? k4[2 + q] : k4[q], send = b ? k4[q] : k4[2 + q];
        k2[q] = keep + dpp_f64<kDppQuadSwap2>(send);
    }
    const double keep = c ? k2[1] : k2[0], send = c ? k2[0] : k2[1];
    return keep + dpp_f64<kDppQuadSwap1>(send);
}
template <int G>
__device__ inline double group_sum(double v) {
    v += dpp_f64<kDppQuadSwap1>(v);
    v += dpp_f64<kDppQuadSwap2>(v);
    if (G >= 8) v += dpp_f64<kDppHalfMirror>(v);
    if (G >= 16) v += dpp_f64<kDppRowMirror>(v);
    if (G >= 32) v += __shfl_xor(v, 16, kWave);
    if (G >= 64) v += __shfl_xor(v, 32, kWave);
    return v;
}

// Sums eight values over the block: scatter8 inside groups of 8 lanes, then the
// partners r+8 (row_ror:8), lane^16 and lane^32; thread q < 8 returns total q.
// Fixed order.
template <int NT>
__device__ inline double block_sum8(const double (&v)[8], double* lds /* 8*NT/64 */) {
    constexpr int NW = NT / kWave;
    const int lane = threadIdx.x % kWave, wid = threadIdx.x / kWave;
    double t = scatter8(v, lane & 7);
    t += dpp_f64<kDppRowRor8>(t);
    t += __shfl_xor(t, 16, kWave);
    t += __shfl_xor(t, 32, kWave);
    if (lane < 8) lds[lane * NW + wid] = t;
    __syncthreads();
    double s = 0.0;
    if (threadIdx.x < 8)
        for (int w = 0; w < NW; ++w) s += lds[threadIdx.x * NW + w];
    __syncthreads();
    return s;
}

// Sums K values over the block; every thread gets the totals in out[]. Fixed order.
template <int K, int NT>
__device__ inline void block_sum(double (&v)[K], double* lds /* K*NT/64 */) {
    constexpr int NW = NT / kWave;
    const int lane = threadIdx.x % kWave, wid = threadIdx.x / kWave;
#pragma unroll
    for (int k = 0; k < K; ++k) {
        const double s = wave_sum(v[k]);
        if (lane == 0) lds[k * NW + wid] = s;
    }
    __syncthreads();
#pragma unroll
    for (int k = 0; k < K; ++k) {
        double s = 0.0;
        for (int w = 0; w < NW; ++w) s += lds[k * NW + w];
        v[k] = s;
    }
    __syncthreads();
}

// ---- searches (search_utils.cpp:13-131) --------------------------------------------
// t1grid as seen by a stage: entries [0, L-1) from HBM and the newest time (the
// value the same launch's finish block is writing) from the kernel argument.
// Indices are clamped into [0, L) so that no input can read out of bounds; on the
// reference's own (well-formed) inputs the clamp never binds.
struct ListView {
    const double* g;
    size_t L;
    double last;
    __device__ inline double operator()(size_t i) const {
        if (i >= L - 1) return last;
        return g[i];
    }
};

#pragma clang fp contract(off)
__device__ double bsearch_elem(const ListView& list, double El) {
    const size_t len = list.L;
    size_t n0 = 1, n1 = len, m = 1;
    double Lm = 0.0;
    if (list(0) < El) {
        while (n0 <= n1) {
            m = (n0 + n1) / 2;
            Lm = list(m - 1);
            if (Lm == El) { n0 = m; break; }
            if (Lm < El) n0 = m + 1;
            else n1 = m - 1;
        }
    }
    if (Lm <= El) {
        if (m == len) return (double)m;
        return m + (El - Lm) / (list(m) - Lm);
    }
    if (m > 1) return m - (El - Lm) / (list(m - 2) - Lm);
    return (double)m;
}

// One element of isearchPosSortedInit with its upper bracket n1 given.
__device__ double isearch_elem(const ListView& list, double El, double init, size_t n1) {
    const size_t length = list.L;
    size_t n0 = (size_t)floor(init);
    if (n0 < 1) n0 = 1;  // the reference reads list[-1] here; positions are >= 1 on its inputs
    size_t m = (n0 + 1 < length) ? n0 + 1 : length;
    bool even = true;
    double l0, l1, Lm = list(m - 1);
    if (Lm > El) {
        const size_t a = m - 2;  // wraps for m < 2 exactly like the size_t code
        n1 = a > 1 ? a : 1;
    }
    int guard = 0;
    while ((l0 = list(n0 - 1)) <= El && El <= (l1 = list(n1 - 1)) && n0 < n1 && guard++ < 512) {
        even = !even;
        if (even) m = (size_t)((double)n0 + round((El - l0) / (l1 - l0) * (double)(n1 - n0)));
        else m = (n0 + n1) / 2;
        Lm = list(m - 1);
        if (Lm == El) { n0 = m; n1 = m - 1; }
        else if (Lm < El) n0 = m + 1;
        else n1 = m - 1;
    }
    if (Lm <= El) {
        if (m == length) return (double)m;
        return m + (El - Lm) / (list(m) - Lm);
    }
    if (m > 1) return m - (El - Lm) / (list(m - 2) - Lm);
    return (double)m;
}
#pragma clang fp contract(on)

constexpr int kStageThreads = 256;

__device__ inline size_t bracket_of(double pos, size_t L) {
    const double cv = ceil(pos);
    return (cv >= (double)L) ? L : (cv < 1.0 ? 1 : (size_t)cv);
}

// Positions of one row of `len` elements elem[l] * tl.  The reference walks a row
// from its last element to its first and feeds ceil(result) of element l+1 into
// element l as the upper bracket n1; an element's result depends on its
// predecessor only through that bracket.  Here every lane searches its elements
// with the brackets of the previous sweep until no bracket changes: after k
// sweeps the last k elements are final, and a sweep that changes no bracket
// reproduces the sequential walk exactly.  The first sweep takes its brackets
// from the previous positions (init), which are usually already exact, so most
// rows finish in one or two sweeps.
// rInt (index_vec.cpp:138-156) in two parts.  The theta-row search, which writes the B1
// prologue (Floor, diff) of an interpolation, evaluates the terms over history rows older
// than the interpolation's newest row; the finish of the previous stage may be writing
// that row's r, dr, rho in the same launch.  The finish block that needs rInt adds the
// newest row's terms (rint_finish), so it reads no history row through Floor.  The
// operations and their order are those of rint_at:
//   kind 0: every row older -> the value itself;
//   kind 1: f + 1 is the newest row -> P - ca dr[f+1] / rho[f+1];
//   kind 2: f is the newest row (last interval) -> (P + ca r[f]) - cb dr[f].
#pragma clang fp contract(off)
__device__ inline double4 rint_terms(const History& H, int f, double d, int T1, bool live) {
    if (!live) return make_double4(H.r[0], 0.0, 0.0, 0.0);
    const double d2 = d * d, d3 = d * (d * d);
    if (f < T1 - 1) {
        const double p = (1 - 3 * d2 - 2 * d3) * H.r[f - 1] + (3 * d2 + 2 * d3) * H.r[f] - (d + 2 * d2 + d3) * H.dr[f];
        if (f + 1 < T1 - 1) return make_double4(p - (d2 + d3) * H.dr[f + 1] / H.rho[f + 1], 0.0, 0.0, 0.0);
        return make_double4(p, d2 + d3, 0.0, 1.0);
    }
    return make_double4((1 - d2) * H.r[f - 1], d2, d + d2, 2.0);
}
// rw, drw, rhow: r, dr, rho of the interpolation's newest history row
__device__ inline double rint_finish(double4 q, double rw, double drw, double rhow) {
    if (q.w == 0.0) return q.x;
    if (q.w == 1.0) return q.x - q.y * drw / rhow;
    return (q.x + q.y * rw) - q.z * drw;
}
#pragma clang fp contract(on)

template <int EPT>
__device__ __forceinline__ void search_row(const ListView& list, double tl, const double* __restrict__ elem, const double* init,
                           double* out, double* out2, int len, int mode, double* lds) {
    const int tid = threadIdx.x;
    double El[EPT], ini[EPT], res[EPT];
    size_t c[EPT];
    const size_t L = list.L;
#pragma unroll
    for (int k = 0; k < EPT; ++k) {
        const int l = tid + k * kStageThreads;
        El[k] = 0.0; ini[k] = 1.0; res[k] = 0.0; c[k] = L;
        if (l < len) {
            El[k] = (mode == kSearchB) ? elem[l] * tl : tl * elem[l];
            ini[k] = init[l];
            if (mode != kSearchB && l + 1 < len) c[k] = bracket_of(init[l + 1], L);
        }
    }
    __syncthreads();  // all inits are read before any lane overwrites them (in-place rows)
    if (mode == kSearchB) {
#pragma unroll
        for (int k = 0; k < EPT; ++k) {
            const int l = tid + k * kStageThreads;
            if (l < len) res[k] = bsearch_elem(list, El[k]);
        }
    } else {
        for (int sweep = 0; sweep <= len + 1; ++sweep) {
#pragma unroll
            for (int k = 0; k < EPT; ++k) {
                const int l = tid + k * kStageThreads;
                if (l < len) {
                    res[k] = isearch_elem(list, El[k], ini[k], c[k]);
                    lds[l] = res[k];
                }
            }
            __syncthreads();
            int changed = 0;
#pragma unroll
            for (int k = 0; k < EPT; ++k) {
                const int l = tid + k * kStageThreads;
                if (l < len) {
                    const size_t nc = (l + 1 < len) ? bracket_of(lds[l + 1], L) : L;
                    changed |= (nc != c[k]);
                    c[k] = nc;
                }
            }
            if (!__syncthreads_or(changed)) break;
        }
    }
#pragma unroll
    for (int k = 0; k < EPT; ++k) {
        const int l = tid + k * kStageThreads;
        if (l < len) {
            out[l] = res[k];
            if (out2) out2[l] = res[k];
        }
    }
}

// B2 positions of phi2 row k (interpolation_core.cpp:27-29).  Positions are double
// buffered like Floor / diff: an interpolation writes buffer a.fbuf and seeds the
// interpolation search from the other one, which holds the previous interpolation's
// positions (posB2xOld) and may still be read by an RHS of the previous stage.
template <int EPT>
__device__ __forceinline__ void search_phi2_row(const Tables& T, const Work& W, const SearchArgs& a, int k, double* lds) {
    const int N = T.N;
    double* out = W.posB2 + (size_t)a.fbuf * T.P + (size_t)k * N;
    const double* old = W.posB2 + (size_t)(1 - a.fbuf) * T.P + (size_t)k * N;
    double* out2 = a.saveB2 ? a.saveB2 + (size_t)k * N : nullptr;
    if (a.mode == kSearchSame) {  // the previous positions again (interpolation_core.cpp:25-29, same = true)
        for (int l = threadIdx.x; l < N; l += kStageThreads) {
            const double v = old[l];
            out[l] = v;
            if (out2) out2[l] = v;
        }
        return;
    }
    if (a.mode == kSearchCopy) {
        const double* src = a.srcB2 + (size_t)k * N;
        for (int l = threadIdx.x; l < N; l += kStageThreads) {
            out[l] = src[l];
            if (out2) out2[l] = src[l];
        }
        return;
    }
    const ListView list{a.t1, (size_t)a.T1, a.tl};
    search_row<EPT>(list, a.tl, T.phi2 + (size_t)k * N, old, out, out2, N, a.mode, lds);
}

// B1 positions of the theta row and the B1 prologue Floor / diff
// (interpolation_core.cpp:22-24, 57-84).
template <int EPT>
__device__ __forceinline__ void search_theta_row(const Tables& T, const History& H, const Work& W, const SearchArgs& a,
                                                 double* lds, double* red) {
    const int N = T.N;
    double* pos = W.posB1 + (size_t)a.fbuf * N;  // double buffered as the B2 positions
    const double* old = W.posB1 + (size_t)(1 - a.fbuf) * N;
    if (a.mode == kSearchCopy) {
        for (int l = threadIdx.x; l < N; l += kStageThreads) pos[l] = a.srcB1[l];
    } else if (a.mode == kSearchSame) {
        for (int l = threadIdx.x; l < N; l += kStageThreads) pos[l] = old[l];
    } else {
        const ListView list{a.t1, (size_t)a.T1, a.tl};
        search_row<EPT>(list, a.tl, T.theta, old, pos, nullptr, N, a.mode, lds);
    }
    __syncthreads();
    if (a.saveB1)
        for (int l = threadIdx.x; l < N; l += kStageThreads) a.saveB1[l] = pos[l];
    // maxCeil = ceil(max pos) - 1, at least 1 (interpolation_core.cpp:58-67)
    double mx = -INFINITY;
    for (int l = threadIdx.x; l < N; l += kStageThreads) mx = fmax(mx, pos[l]);
    mx = wave_max(mx);
    if (threadIdx.x % kWave == 0) red[threadIdx.x / kWave] = mx;
    __syncthreads();
    mx = red[0];
    for (int w = 1; w < kStageThreads / kWave; ++w) mx = fmax(mx, red[w]);
    size_t maxCeil = (size_t)ceil(mx) - 1;
    if (maxCeil < 1) maxCeil = 1;
    const size_t t1len = (size_t)a.T1;
    for (int i = threadIdx.x; i < N; i += kStageThreads) {
        size_t f = (size_t)floor(pos[i]);
        if (a.gpu_clamp) {
            // diffNfloorKernel (interpolation_core.cu:23-42): each element's own ceil
            const size_t own = (size_t)ceil(pos[i]);
            if (own - 1 < f) f = own - 1;
            if (f < 1) f = 1;
        } else if (f < 1) f = 1;
        else if (f > maxCeil) f = maxCeil;
        if (f > t1len - 1) f = t1len > 1 ? t1len - 1 : 1;  // memory guard; never binds on valid positions
        const double d = (double)f - pos[i];
        W.Floor[(size_t)a.fbuf * N + i] = (int)f;
        W.diff[(size_t)a.fbuf * N + i] = d;
        W.rpk[(size_t)a.fbuf * N + i] = rint_terms(H, (int)f, d, a.T1, a.tl > 0);
    }
}

// ---- fused RHS rows ----------------------------------------------------------------
constexpr int kRowThreads = 256;

// Consecutive workgroups are dispatched round-robin over the 8 XCDs.  Logical blocks
// (row-major (row j, chunk), or the batch's launch map) are dealt to the XCDs in stripes
// of S consecutive blocks, stripe k to XCD k mod 8: a stripe's rows, which share history
// rows and cache lines, stay in one XCD's L2, and every XCD gets stripes from the whole
// row range.  The B2 work of a row falls steeply with theta_j (small theta_j: targets
// spread over the whole history; theta_j near 1: mostly the last time interval), so a
// contiguous row range per XCD (S = nb / 8) leaves some XCDs with most of the gathers.
// S is the largest power of two <= SMAX that divides nb / 8 (the identity if 8 does not
// divide nb).  Measured at L = 512: SMAX = 32 for one instance (16 rows x 2 chunks per
// stripe), 16 for the batch map (whose consecutive blocks are the instances of one row).
template <int SMAX>
__device__ inline int xcd_stripe(int b, int nb) {
    constexpr int X = 8;
    if (nb % X) return b;
    int S = SMAX;
    while (S > 1 && nb % (X * S)) S >>= 1;
    const int x = b % X, q = b / X;
    return ((q / S) * X + x) * S + q % S;
}

// ---- forward-mode sensitivities with respect to Gamma and lambda -----------------------
// The time grid (dt sequence, positions, stencil starts, Hermite abscissae) is frozen,
// so the step map is a smooth function of the history values and of (Gamma, lambda).
// Its tangent: interpolants are linear in the history, hence their tangents are the same
// stencils / Hermite forms on the tangent history; f', f'' and f''' follow the chain rule
// in q and their explicit lambda dependence (math_ops.cpp:27-37); Gamma enters through
// 2 Gamma qR in QKstep (time_steps.cpp:43) and through rstep (:91).  Lane per target,
// written for clarity: the tangents cost about two primal RHS evaluations per stage.
__device__ inline double fl1l(const Model& M, double q) {  // derivative of f' in lambda
    return M.p * pow_int(q, M.p - 1) - M.p2 * pow_int(q, M.p2 - 1);
}
__device__ inline double fl2l(const Model& M, double q) {  // derivative of f'' in lambda
    return M.p * (M.p - 1) * pow_int(q, M.p - 2) - M.p2 * (M.p2 - 1) * pow_int(q, M.p2 - 2);
}
__device__ inline double fl3l(const Model& M, double q) {  // derivative of f''' in lambda
    return M.p * (M.p - 1) * (M.p - 2) * pow_int(q, M.p - 3) -
           M.p2 * (M.p2 - 1) * (M.p2 - 2) * pow_int(q, M.p2 - 3);
}
__device__ inline double fl4(const Model& M, double q) {  // fourth derivative of f
    return M.lambda * M.p * (M.p - 1) * (M.p - 2) * (M.p - 3) * pow_int(q, M.p - 4) +
           (1 - M.lambda) * M.p2 * (M.p2 - 1) * (M.p2 - 2) * (M.p2 - 3) * pow_int(q, M.p2 - 4);
}

// history view of tangent direction k (time arrays are the primal's)
__device__ inline History tan_history(const History& H, const TanArgs& X, int k) {
    History h = H;
    h.V = X.V[k];
    h.D = X.D[k];
    h.r = X.r[k];
    h.dr = X.dr[k];
    return h;
}

__device__ inline double rint_at(const History& H, const int* Floor, const double* diff, int i, int T1, bool live) {
    if (!live) return H.r[0];
    const size_t f = (size_t)Floor[i], t1len = (size_t)T1;
    const double d = diff[i], d2 = d * d, d3 = d * (d * d);
    if (f < t1len - 1)
        return (1 - 3 * d2 - 2 * d3) * H.r[f - 1] + (3 * d2 + 2 * d3) * H.r[f] - (d + 2 * d2 + d3) * H.dr[f] -
               (d2 + d3) * H.dr[f + 1] / H.rho[f + 1];
    return (1 - d2) * H.r[f - 1] + d2 * H.r[f] - (d + d2) * H.dr[f];
}


// taps whose weights one lane requests in one memory round trip (the stencil width of the
// default poly order 9)
constexpr int kTapBatch = 10;

#pragma clang fp contract(off)  // the compensated sum below needs unfused products
template <bool LOGR>
__device__ inline double2 stencil_last(const double* __restrict__ w, long long P, long long k, int m, int s,
                                       const double2* __restrict__ vl) {
    double a = 0.0, b = 0.0;
    if (!LOGR) {
        // the weights of up to kTapBatch taps are requested together (one memory round
        // trip per batch instead of one per few taps); accumulation order d = 0, 1, ...
        for (int d0 = 0; d0 < m; d0 += kTapBatch) {
            double wd[kTapBatch];
#pragma unroll
            for (int t = 0; t < kTapBatch; ++t) wd[t] = d0 + t < m ? w[(long long)(d0 + t) * P + k] : 0.0;
#pragma unroll
            for (int t = 0; t < kTapBatch; ++t) {
                if (d0 + t < m) {
                    const double2 v = vl[s + d0 + t];
                    a = a + wd[t] * v.x;
                    b = b + wd[t] * v.y;
                }
            }
        }
        return make_double2(a, b);
    }
    // The reference accumulates sum_d w_d log(QR) in 80-bit long double
    // (index_vec.cpp:159-172); a compensated dot product (exact FMA products and
    // two-sum accumulation, Ogita-Rump-Oishi "Dot2") reproduces that accuracy in FP64.
    double lin = 0.0, ls = 0.0, lc = 0.0;
    bool invalid = false;
    for (int d = 0; d < m; ++d) {
        const double wd = w[(long long)d * P + k];
        const double2 v = vl[s + d];
        a = a + wd * v.x;
        lin += wd * v.y;
        if (v.y > 0.0) {
            const double lv = log(v.y);
            const double p = wd * lv;
            const double pe = fma(wd, lv, -p);
            const double t = ls + p;
            const double bv = t - ls;
            lc += ((ls - (t - bv)) + (p - bv)) + pe;
            ls = t;
        } else {
            invalid = true;
        }
    }
    return make_double2(a, invalid ? lin : exp(ls + lc));
}

#pragma clang fp contract(on)

// B2 with the four slice sums (Y0, Y1, DY1, DY2) closed separately and the time Hermite
// applied to the sums by the group's first lane: the log-response form, where QR is
// interpolated in log space and is not linear in the slices.
template <bool LOGR, int G>
__device__ __forceinline__ void b2_groups_sums(const Tables& T, const History& H, const Work& W, const RhsArgs& a, int j,
                                          int chunk, int C, const int* b2ix, const int* b2s, const double* b2u,
                                          const double* b2rr, double2* b2, const double2* tail /* [3][N] or nullptr */) {
    const int N = T.N, m = T.m, T1 = a.T1;
    constexpr int GPW = kWave / G;
    const int lane = threadIdx.x % kWave, wid = threadIdx.x / kWave;
    const int g = lane / G, d0 = lane % G;
    constexpr int kWaves = kRowThreads / kWave;
    for (int e0 = wid * GPW; e0 < C; e0 += GPW * kWaves) {
        const int e = e0 + g;
        const bool ev = e < C;
        const int ee = ev ? e : 0;
        const size_t ix = (size_t)b2ix[ee];
        const int sst = b2s[ee];
        const bool b2int = ix < (size_t)(T1 - 1);
        const double2* R0 = H.V + (ix - 1) * N + sst;
        const double2* R1 = H.V + ix * N + sst;
        const double2* S1 = H.D + ix * N + sst;
        const double2* S2 = H.D + (b2int ? ix + 1 : ix) * N + sst;
        const double* wk = T.w[2] + ((long long)j * N + rhs_chunk_col(ee, chunk, a.split)) * m;
        double acc8[8];
#pragma unroll
        for (int q = 0; q < 8; ++q) acc8[q] = 0.0;
        // taps dd and dd + G of a lane are loaded together (one memory round trip per
        // pass instead of one per tap); accumulation order stays dd, dd + G, ...
        if (ev && !b2int && tail) {  // last interval: the Hermite uses Y0, Y1, DY1 only
            const double2 *L0 = tail + sst, *L1 = tail + N + sst, *L2 = tail + 2 * N + sst;
            for (int dd = d0; dd < m; dd += 2 * G) {
                const bool two = dd + G < m;
                const double wd = wk[dd], we = two ? wk[dd + G] : 0.0;
                const double2 v0 = L0[dd], v1 = L1[dd], g1 = L2[dd];
                acc8[0] += wd * v0.x; acc8[1] += wd * v0.y;
                acc8[2] += wd * v1.x; acc8[3] += wd * v1.y;
                acc8[4] += wd * g1.x; acc8[5] += wd * g1.y;
                if (two) {
                    const double2 u0 = L0[dd + G], u1 = L1[dd + G], h1 = L2[dd + G];
                    acc8[0] += we * u0.x; acc8[1] += we * u0.y;
                    acc8[2] += we * u1.x; acc8[3] += we * u1.y;
                    acc8[4] += we * h1.x; acc8[5] += we * h1.y;
                }
            }
        } else if (ev) {
            for (int dd = d0; dd < m; dd += 2 * G) {
                const bool two = dd + G < m;
                const double wd = wk[dd];
                const double2 v0 = R0[dd], v1 = R1[dd], g1 = S1[dd], g2 = S2[dd];
                double we = 0.0;
                double2 u0 = make_double2(0.0, 0.0), u1 = u0, h1 = u0, h2 = u0;
                if (two) {
                    we = wk[dd + G];
                    u0 = R0[dd + G]; u1 = R1[dd + G]; h1 = S1[dd + G]; h2 = S2[dd + G];
                }
                acc8[0] += wd * v0.x; acc8[1] += wd * v0.y;
                acc8[2] += wd * v1.x; acc8[3] += wd * v1.y;
                acc8[4] += wd * g1.x; acc8[5] += wd * g1.y;
                acc8[6] += wd * g2.x; acc8[7] += wd * g2.y;
                if (two) {
                    acc8[0] += we * u0.x; acc8[1] += we * u0.y;
                    acc8[2] += we * u1.x; acc8[3] += we * u1.y;
                    acc8[4] += we * h1.x; acc8[5] += we * h1.y;
                    acc8[6] += we * h2.x; acc8[7] += we * h2.y;
                }
            }
        }
        const double y0k = group_sum<G>(acc8[0]), y0r = group_sum<G>(acc8[1]);
        const double y1k = group_sum<G>(acc8[2]), y1r = group_sum<G>(acc8[3]);
        const double e1k = group_sum<G>(acc8[4]), e1r = group_sum<G>(acc8[5]);
        const double e2k = group_sum<G>(acc8[6]), e2r = group_sum<G>(acc8[7]);
        if (d0 == 0 && ev) {
            const double u = b2u[e], u2 = u * u, u3 = u2 * u;
            double qk, qr;
            if (b2int) {
                const double rr = b2rr[e];
                const double h0 = 1 - 3 * u2 + 2 * u3, h1 = u - 2 * u2 + u3, h2 = 3 * u2 - 2 * u3, h3 = -u2 + u3;
                qk = h0 * y0k + h1 * e1k + h2 * y1k + h3 * e2k / rr;
                if (!LOGR || !(y0r > 0.0 && y1r > 0.0))
                    qr = h0 * y0r + h1 * e1r + h2 * y1r + h3 * e2r / rr;
                else
                    qr = exp(h0 * log(y0r) + h1 * (e1r / y0r) + h2 * log(y1r) + (h3 / rr) * (e2r / y1r));
            } else {
                qk = (1 - u2) * y0k + u2 * y1k + (u - u2) * e1k;
                if (!LOGR || !(y0r > 0.0 && y1r > 0.0))
                    qr = (1 - u2) * y0r + u2 * y1r + (u - u2) * e1r;
                else
                    qr = exp((1 - u2) * log(y0r) + u2 * log(y1r) + (u - u2) * (e1r / y0r));
            }
            b2[e] = make_double2(qk, qr);
        }
    }
}

// Hermite coefficient of the DY2 slice inside the history, (u^3 - u^2) / rho[ix + 1]
// (index_mat.cpp:45-70); computed once per entry, before the gathers.
__device__ __forceinline__ double b2_c3(double u, double rr) {
    const double u2 = u * u, u3 = u2 * u;
    return (-u2 + u3) / rr;
}

// Hermite coefficients of the four B2 slices (Y0, Y1, DY1, DY2) at abscissa u
// (index_mat.cpp:45-70): cubic Hermite inside the history (c3 from b2_c3), the quadratic
// form in the last interval (where DY2 does not enter).
__device__ __forceinline__ void b2_hermite(double u, bool interior, double c3, double (&c)[4]) {
    const double u2 = u * u, u3 = u2 * u;
    if (interior) {
        c[0] = 1 - 3 * u2 + 2 * u3;
        c[1] = 3 * u2 - 2 * u3;
        c[2] = u - 2 * u2 + u3;
        c[3] = c3;
    } else {
        c[0] = 1 - u2;
        c[1] = u2;
        c[2] = u - u2;
        c[3] = 0.0;
    }
}

// one stencil tap of B2: the Hermite combination of the four slices, times the tap weight
__device__ __forceinline__ void b2_tap(const double (&c)[4], double wd, double2 v0, double2 v1, double2 g1,
                                       double2 g2, double& ak, double& ar) {
    const double pk = fma(c[3], g2.x, fma(c[2], g1.x, fma(c[1], v1.x, c[0] * v0.x)));
    const double pr = fma(c[3], g2.y, fma(c[2], g1.y, fma(c[1], v1.y, c[0] * v0.y)));
    ak = fma(wd, pk, ak);
    ar = fma(wd, pr, ar);
}

// B2 (index_mat.cpp:18-91), linear form: groups of G lanes per entry, lane d holds taps
// d, d+G, ...; gathers of one wave instruction touch 64/G entries x a few contiguous
// cache lines.  The interpolant is linear in the four slices, so every lane folds the
// entry's time Hermite into its taps and a group closes only two sums (QK, QR).  Entries
// in the last time interval (ix = T1-1: V[T1-2], V[T1-1], D[T1-1] only) read those rows
// from the block's LDS copies `tail` = {V[T1-2], V[T1-1], D[T1-1]} when given.
template <bool LOGR, int G, int NT>
__device__ __forceinline__ void b2_groups(const Tables& T, const History& H, const Work& W, const RhsArgs& a, int j,
                                          int chunk, int C, const int* b2ix, const int* b2s, const double* b2u,
                                          const double* b2rr, double2* b2, const double2* tail /* [3][N] or nullptr */) {
    if constexpr (LOGR) {
        b2_groups_sums<LOGR, G>(T, H, W, a, j, chunk, C, b2ix, b2s, b2u, b2rr, b2, tail);
        return;
    }
    const int N = T.N, m = T.m, T1 = a.T1;
    constexpr int GPW = kWave / G;
    const int lane = threadIdx.x % kWave, wid = threadIdx.x / kWave;
    const int g = lane / G, d0 = lane % G;
    constexpr int kWaves = kRowThreads / kWave;
    for (int e0 = wid * GPW; e0 < C; e0 += GPW * kWaves) {
        const int e = e0 + g;
        const bool ev = e < C;
        const int ee = ev ? e : 0;
        const size_t ix = (size_t)b2ix[ee];
        const int sst = b2s[ee];
        const bool b2int = ix < (size_t)(T1 - 1);
        const double* wk = T.w[2] + ((long long)j * N + rhs_chunk_col(ee, chunk, a.split)) * m;
        double c[4];
        b2_hermite(b2u[ee], b2int, b2rr[ee], c);  // linear form: b2rr holds c3
        double ak = 0.0, ar = 0.0;
        // the NT taps dd, dd + G, ... of a lane are loaded in one pass (one memory round
        // trip instead of one per tap); accumulation order stays dd, dd + G, ...
        if (ev && !b2int && tail) {
            const double2 *L0 = tail + sst, *L1 = tail + N + sst, *L2 = tail + 2 * N + sst;
            for (int dd = d0; dd < m; dd += NT * G) {
                double w[NT];
#pragma unroll
                for (int t = 0; t < NT; ++t) w[t] = dd + t * G < m ? wk[dd + t * G] : 0.0;
#pragma unroll
                for (int t = 0; t < NT; ++t) {
                    const int d = dd + t * G;
                    if (d < m) {
                        const double2 g1 = L2[d];
                        b2_tap(c, w[t], L0[d], L1[d], g1, g1, ak, ar);
                    }
                }
            }
        } else if (ev) {
            const double2* R0 = H.V + (ix - 1) * N + sst;
            const double2* R1 = H.V + ix * N + sst;
            const double2* S1 = H.D + ix * N + sst;
            const double2* S2 = H.D + (b2int ? ix + 1 : ix) * N + sst;
            for (int dd = d0; dd < m; dd += NT * G) {
                double w[NT];
                double2 v0[NT], v1[NT], g1[NT], g2[NT];
#pragma unroll
                for (int t = 0; t < NT; ++t) {
                    const int d = dd + t * G;
                    w[t] = 0.0;
                    v0[t] = v1[t] = g1[t] = g2[t] = make_double2(0.0, 0.0);
                    if (d < m) {
                        w[t] = wk[d];
                        v0[t] = R0[d]; v1[t] = R1[d]; g1[t] = S1[d]; g2[t] = S2[d];
                    }
                }
#pragma unroll
                for (int t = 0; t < NT; ++t)
                    if (dd + t * G < m) b2_tap(c, w[t], v0[t], v1[t], g1[t], g2[t], ak, ar);
            }
        }
        const double qk = group_sum<G>(ak), qr = group_sum<G>(ar);
        if (d0 == 0 && ev) b2[e] = make_double2(qk, qr);
    }
}

// B2 (index_mat.cpp:18-91), linear form, on the block's entry lists: slots [0, nh) hold
// the entries whose four slices are gathered from the history, slots [nh, C) those of
// the last time interval (their rows are the block's LDS copies `tail`), so every
// iteration of a wave gathers one kind only.  G lanes per entry, lane d holding taps d,
// d+G, ... (NT per pass, loaded in one memory round trip together with rho[ix+1]); the
// entry's time Hermite is folded into its taps and a group closes two sums (QK, QR).
// Per entry the arithmetic and its order are those of b2_groups.
template <int G, int NT>
__device__ __forceinline__ void b2_lists(const Tables& T, const History& H, const RhsArgs& a, int j, int chunk, int C,
                                         int nh, const int* lix, const int* ls, const double* lu, const int* le,
                                         double2* b2, const double2* tail /* [3][N] or nullptr */) {
    const int N = T.N, m = T.m, T1 = a.T1;
    constexpr int GPW = kWave / G;
    const int lane = threadIdx.x % kWave, wid = threadIdx.x / kWave;
    const int g = lane / G, d0 = lane % G;
    constexpr int kWaves = kRowThreads / kWave;
    for (int k0 = wid * GPW; k0 < nh; k0 += GPW * kWaves) {
        const int k = k0 + g;
        const bool ev = k < nh;
        const int kk = ev ? k : k0;  // a valid slot: k0 < nh
        const int e = le[kk];
        const size_t ix = (size_t)lix[kk];
        const int sst = ls[kk];
        const double u = lu[kk];
        const bool b2int = ix < (size_t)(T1 - 1);
        const double* wk = T.w[2] + ((long long)j * N + rhs_chunk_col(e, chunk, a.split)) * m;
        const double2* R0 = H.V + (ix - 1) * N + sst;
        const double2* R1 = H.V + ix * N + sst;
        const double2* S1 = H.D + ix * N + sst;
        const double2* S2 = H.D + (b2int ? ix + 1 : ix) * N + sst;
        const double rr = b2int ? H.rho[ix + 1] : 1.0;
        double ak = 0.0, ar = 0.0;
        if (ev) {
            for (int dd = d0; dd < m; dd += NT * G) {
                double w[NT];
                double2 v0[NT], v1[NT], g1[NT], g2[NT];
#pragma unroll
                for (int t = 0; t < NT; ++t) {
                    const int d = dd + t * G;
                    w[t] = 0.0;
                    v0[t] = v1[t] = g1[t] = g2[t] = make_double2(0.0, 0.0);
                    if (d < m) {
                        w[t] = wk[d];
                        v0[t] = R0[d]; v1[t] = R1[d]; g1[t] = S1[d]; g2[t] = S2[d];
                    }
                }
                double c[4];
                b2_hermite(u, b2int, b2int ? b2_c3(u, rr) : 0.0, c);
#pragma unroll
                for (int t = 0; t < NT; ++t)
                    if (dd + t * G < m) b2_tap(c, w[t], v0[t], v1[t], g1[t], g2[t], ak, ar);
            }
        }
        const double qk = group_sum<G>(ak), qr = group_sum<G>(ar);
        if (d0 == 0 && ev) b2[e] = make_double2(qk, qr);
    }
    if (!tail) return;
    // last interval: one lane per entry (its weights are contiguous, its taps in LDS), so
    // all of a block's last-interval entries need one memory round trip (the weights)
    for (int k = nh + (int)threadIdx.x; k < C; k += kRowThreads) {
        const int e = le[k];
        const int sst = ls[k];
        const double* wk = T.w[2] + ((long long)j * N + rhs_chunk_col(e, chunk, a.split)) * m;
        double c[4];
        b2_hermite(lu[k], false, 0.0, c);
        const double2 *L0 = tail + sst, *L1 = tail + N + sst, *L2 = tail + 2 * N + sst;
        double ak = 0.0, ar = 0.0;
        for (int d0 = 0; d0 < m; d0 += kTapBatch) {
            double w[kTapBatch];
#pragma unroll
            for (int t = 0; t < kTapBatch; ++t) w[t] = d0 + t < m ? wk[d0 + t] : 0.0;
#pragma unroll
            for (int t = 0; t < kTapBatch; ++t) {
                const int d = d0 + t;
                if (d < m) {
                    const double2 h1 = L2[d];
                    b2_tap(c, w[t], L0[d], L1[d], h1, h1, ak, ar);
                }
            }
        }
        b2[e] = make_double2(ak, ar);
    }
}

// Row j, the columns of chunk `chunk` (rhs_chunk_col): the eight quadratures of QKstep/QRstep.
//  phase 1 (B2, index_mat.cpp:18-91): one group of G lanes per target entry, lane d
//          holding stencil tap d, so the four history-row gathers of a wave touch a
//          few contiguous cache lines instead of one line per lane; the m-tap inner
//          products are closed by an xor-shuffle reduction inside the group.  Results
//          go to LDS.
//  phase 2 (lane per entry): A1/A2 stencils on the last row (LDS), the B1 time
//          Hermite (index_vec.cpp:63-136), the self-energies and the weighted
//          products, then one block reduction.
// B2 of the primal history and of K tangent histories in one pass (linear form): the
// weights, positions and Hermite coefficients are computed once per tap and applied to
// every history; the primal's operations are exactly those of b2_groups<false>.
// Entries of the last time interval are summed tap by tap in one lane when the plain
// kernel reads them from its LDS rows (tail_rows), with b2_lists' operation order, so
// that the primal stays bit-identical to the run without tangents.
template <int G, int K>
__device__ __forceinline__ void b2_groups_multi(const Tables& T, const History& H, const TanArgs& X, const RhsArgs& a,
                                                int j, int chunk, int C, const int* b2ix, const int* b2s,
                                                const double* b2u, const double* b2rr, double2* b2, double2* tb2,
                                                int cw, bool tail_rows) {
    const int N = T.N, m = T.m, T1 = a.T1;
    constexpr int GPW = kWave / G;
    const int lane = threadIdx.x % kWave, wid = threadIdx.x / kWave;
    const int g = lane / G, d0 = lane % G;
    constexpr int kWaves = kRowThreads / kWave;
    const double2* Vh[1 + K];
    const double2* Dh[1 + K];
    Vh[0] = H.V;
    Dh[0] = H.D;
#pragma unroll
    for (int h = 0; h < K; ++h) {
        Vh[1 + h] = X.V[h];
        Dh[1 + h] = X.D[h];
    }
    for (int e0 = wid * GPW; e0 < C; e0 += GPW * kWaves) {
        const int e = e0 + g;
        const bool ev = e < C;
        const int ee = ev ? e : 0;
        const size_t ix = (size_t)b2ix[ee];
        const int sst = b2s[ee];
        const bool b2int = ix < (size_t)(T1 - 1);
        const size_t o0 = (ix - 1) * N + sst, o1 = ix * N + sst, o2 = (b2int ? ix + 1 : ix) * N + sst;
        const double* wk = T.w[2] + ((long long)j * N + rhs_chunk_col(ee, chunk, a.split)) * m;
        double c[4];
        b2_hermite(b2u[ee], b2int, b2rr[ee], c);  // b2rr holds c3
        double ak[1 + K], ar[1 + K];
#pragma unroll
        for (int h = 0; h <= K; ++h) ak[h] = ar[h] = 0.0;
        if (ev && tail_rows && !b2int) {
            if (d0 == 0)
                for (int dd = 0; dd < m; ++dd) {
                    const double wd = wk[dd];
#pragma unroll
                    for (int h = 0; h <= K; ++h) {
                        const double2 g1 = Dh[h][o1 + dd];
                        b2_tap(c, wd, Vh[h][o0 + dd], Vh[h][o1 + dd], g1, g1, ak[h], ar[h]);
                    }
                }
        } else if (ev) {
            for (int dd = d0; dd < m; dd += G) {
                const double wd = wk[dd];
#pragma unroll
                for (int h = 0; h <= K; ++h)
                    b2_tap(c, wd, Vh[h][o0 + dd], Vh[h][o1 + dd], Dh[h][o1 + dd], Dh[h][o2 + dd], ak[h], ar[h]);
            }
        }
#pragma unroll
        for (int h = 0; h <= K; ++h) {
            const double qk = group_sum<G>(ak[h]), qr = group_sum<G>(ar[h]);
            if (d0 == 0 && ev) {
                if (h == 0) b2[e] = make_double2(qk, qr);
                else tb2[(size_t)(h - 1) * cw + e] = make_double2(qk, qr);
            }
        }
    }
}

#pragma clang fp contract(off)  // the primal part must repeat stencil_last<false> exactly
// A1 or A2 stencil on the last rows of the primal and of K tangent histories (LDS),
// weights loaded once per tap; `dense`: entry-major weights (the primal then comes
// from the MFMA contraction and only the tangents are summed here).
template <int K>
__device__ inline void stencil_last_multi(const double* __restrict__ w, long long P, long long k, int m, int s,
                                          bool dense, const double2* __restrict__ vl, const double2* __restrict__ tvl,
                                          int N, double2& prim, double2 (&tan)[K]) {
    double a = 0.0, b = 0.0;
    double ta[K], tb[K];
#pragma unroll
    for (int h = 0; h < K; ++h) ta[h] = tb[h] = 0.0;
    for (int d = 0; d < m; ++d) {
        const double wd = dense ? w[k * m + d] : w[(long long)d * P + k];
        const double2 v = vl[s + d];
        a = a + wd * v.x;
        b = b + wd * v.y;
#pragma unroll
        for (int h = 0; h < K; ++h) {
            const double2 tv = tvl[(size_t)h * N + s + d];
            ta[h] = ta[h] + wd * tv.x;
            tb[h] = tb[h] + wd * tv.y;
        }
    }
    prim = make_double2(a, b);
#pragma unroll
    for (int h = 0; h < K; ++h) tan[h] = make_double2(ta[h], tb[h]);
}
#pragma clang fp contract(on)

// K > 0: also the tangent quadratures of K forward-mode directions (linear-response
// form only): the tangent interpolants are the same stencils on the tangent histories,
// the primal part is unchanged.
template <bool LOGR, int G, int K = 0, int NT = 1>
__device__ __forceinline__ void rhs_rows_body(const Tables& T, const History& H, const Work& W, const Model& M, const RhsArgs& a,
                              int j, int chunk, double2* dyn /* [N] last row, [cw] (QKB2, QRB2), [cw] u, [cw] ix, s;
                                                                K > 0: [K N] tangent last rows, [K cw] tangent B2 */,
                              double* red, const TanArgs* X = nullptr) {
    const int N = T.N, m = T.m;
    const long long P = T.P;
    const int cw = rhs_chunk_cap(N, a.split);
    const int C = rhs_chunk_count(N, chunk, a.split);
    const int T1 = a.T1;
    const bool live = a.tl > 0;
    // LDS: rows {V[T1-2], V[T1-1], D[T1-1]} (only V[T1-1] when N > kTailMaxN), then the B2 arrays
    const bool tail_rows = N <= kTailMaxN && T1 >= 2;
    double2* rows = dyn;
    double2* vlast = tail_rows ? dyn + N : dyn;
    double2* b2 = dyn + (tail_rows ? 3 * N : N);
    double* b2u = reinterpret_cast<double*>(b2 + cw);
    // per entry inside the history: the log-response form keeps rho[ix + 1], the linear form
    // the DY2 Hermite coefficient b2_c3 (1 and 0 in the last interval)
    double* b2rr = b2u + cw;
    int* b2ix = reinterpret_cast<int*>(b2rr + cw);
    int* b2s = b2ix + cw;
    // linear form (K = 0): the entry of each list slot, in the (unused) Hermite-rho array
    int* b2e = reinterpret_cast<int*>(b2rr);
    // the B2 operands of this thread's first entry and the scalars of phase 2 are
    // requested first, so their latency overlaps the row copies below
    const int e1 = threadIdx.x;
    const long long k1 = (long long)j * N + rhs_chunk_col(e1, chunk, a.split);
    double px1 = 0.0, pmax = 0.0;
    int s1 = 0;
    const double* posB2 = W.posB2 + (size_t)a.fbuf * P;  // positions of the current interpolation
    if (live) {
        pmax = posB2[P - 1];
        if (e1 < C) {
            px1 = posB2[k1];
            s1 = T.inds[2][k1];
        }
    }
    const int F = W.Floor[(size_t)a.fbuf * N + j];
    const double dB = W.diff[(size_t)a.fbuf * N + j];
    const bool b1int = F < T1 - 1;
    const double b1rho = b1int ? H.rho[F + 1] : 1.0;
    // last rows into LDS, all loads of a thread in flight before its first store
    const double2* Vl = a.vlast ? a.vlast : H.V + (size_t)(T1 - 1) * N;
    const bool stage_tail = tail_rows && live;
    const double2* Vp = H.V + (size_t)(stage_tail ? T1 - 2 : T1 - 1) * N;
    const double2* Dl = H.D + (size_t)(T1 - 1) * N;
    for (int base = 0; base < N; base += 4 * kRowThreads) {
        double2 r0[4], r1[4], r2[4];
#pragma unroll
        for (int q = 0; q < 4; ++q) {
            const int l = base + q * kRowThreads + threadIdx.x;
            r0[q] = r1[q] = r2[q] = make_double2(0.0, 0.0);
            if (l < N) {
                r0[q] = Vl[l];
                if (stage_tail) {
                    r1[q] = Vp[l];
                    r2[q] = Dl[l];
                }
            }
        }
#pragma unroll
        for (int q = 0; q < 4; ++q) {
            const int l = base + q * kRowThreads + threadIdx.x;
            if (l < N) {
                vlast[l] = r0[q];
                if (stage_tail) {
                    rows[l] = r1[q];
                    rows[2 * N + l] = r2[q];
                }
            }
        }
    }
    // K > 0 only; b2u [cw] doubles, b2rr [cw] doubles, b2ix and b2s [cw] ints each, then
    // 16-byte alignment for the tangent rows
    double2* tvlast = reinterpret_cast<double2*>(b2u + 3 * cw + (cw & 1));
    double2* tb2 = tvlast + (size_t)K * N;
    if constexpr (K > 0) {
#pragma unroll
        for (int h = 0; h < K; ++h)
            for (int l = threadIdx.x; l < N; l += kRowThreads) tvlast[h * N + l] = X->V[h][(size_t)(T1 - 1) * N + l];
    }

    // ---- phase 1: B2 ----------------------------------------------------------------
    const double c00 = H.V[0].x, c01 = H.V[0].y;
    if constexpr (K == 0 && !LOGR) {
        if (live) {
            // 1a: per entry (lane per entry, coalesced): time row, stencil start, Hermite
            // abscissa; the entries are sorted into two lists, those whose slices are
            // gathered from the history (front of the list arrays) and those of the last
            // time interval, read from the block's LDS rows (back, in reverse order)
            size_t cap_hi = (size_t)(pmax - 0.5);
            if (cap_hi > (size_t)(T1 - 1)) cap_hi = (size_t)(T1 - 1);  // memory guard
            int* cnt = reinterpret_cast<int*>(red);                    // per-wave list counts
            const int lane = threadIdx.x % kWave, wid = threadIdx.x / kWave;
            const unsigned long long below = (1ull << lane) - 1ull;
            int nh = 0, nl = 0;
            for (int r0 = 0; r0 < C; r0 += kRowThreads) {
                const int e = r0 + threadIdx.x;
                int ix = 1, sst = 0;
                double u = 0.0;
                if (e < C) {
                    const long long k = (long long)j * N + rhs_chunk_col(e, chunk, a.split);
                    const double px = e == e1 ? px1 : posB2[k];
                    size_t x = (size_t)px;
                    if (cap_hi < x) x = cap_hi;
                    if (x < 1) x = 1;
                    ix = (int)x;
                    sst = e == e1 ? s1 : T.inds[2][k];
                    u = px - x;
                }
                const bool lastrow = tail_rows && ix >= T1 - 1;
                const bool toH = e < C && !lastrow, toL = e < C && lastrow;
                const unsigned long long mh = __ballot(toH), ml = __ballot(toL);
                if (lane == 0) {
                    cnt[wid] = __popcll(mh);
                    cnt[kRowThreads / kWave + wid] = __popcll(ml);
                }
                __syncthreads();
                int oh = nh, ol = nl;
                for (int w = 0; w < kRowThreads / kWave; ++w) {
                    const int ch = cnt[w], cl = cnt[kRowThreads / kWave + w];
                    if (w < wid) { oh += ch; ol += cl; }
                    nh += ch;
                    nl += cl;
                }
                if (toH || toL) {
                    const int slot = toH ? oh + __popcll(mh & below) : C - 1 - (ol + __popcll(ml & below));
                    b2ix[slot] = ix;
                    b2s[slot] = sst;
                    b2u[slot] = u;
                    b2e[slot] = e;
                }
                __syncthreads();  // cnt is rewritten by the next round
            }
            // 1b: G lanes per entry, lane d holds taps d, d+G, ...; gathers of one wave
            // instruction touch 64/G entries x a few contiguous cache lines
            b2_lists<G, NT>(T, H, a, j, chunk, C, nh, b2ix, b2s, b2u, b2e, b2, tail_rows ? rows : nullptr);
        }
    } else if (live) {
        // 1a: per entry (lane per entry, coalesced): time row, stencil start, Hermite abscissa
        size_t cap_hi = (size_t)(pmax - 0.5);
        if (cap_hi > (size_t)(T1 - 1)) cap_hi = (size_t)(T1 - 1);  // memory guard
        for (int e = e1; e < C; e += kRowThreads) {
            const long long k = (long long)j * N + rhs_chunk_col(e, chunk, a.split);
            const double px = e == e1 ? px1 : posB2[k];
            size_t ix = (size_t)px;
            if (cap_hi < ix) ix = cap_hi;
            if (ix < 1) ix = 1;
            b2ix[e] = (int)ix;
            b2s[e] = e == e1 ? s1 : T.inds[2][k];
            b2u[e] = px - ix;
            // requested and divided here, once per entry and overlapped with the other
            // entries' loads, so that the B2 groups below start their gathers at once
            const bool inside = ix < (size_t)(T1 - 1);
            const double rr = inside ? H.rho[ix + 1] : 1.0;
            b2rr[e] = LOGR ? rr : (inside ? b2_c3(px - ix, rr) : 0.0);
        }
        __syncthreads();
        // 1b: G lanes per entry, lane d holds taps d, d+G, ...; gathers of one wave
        // instruction touch 64/G entries x a few contiguous cache lines
        if constexpr (K == 0)
            b2_groups<LOGR, G, NT>(T, H, W, a, j, chunk, C, b2ix, b2s, b2u, b2rr, b2, tail_rows ? rows : nullptr);
        else b2_groups_multi<G, K>(T, H, *X, a, j, chunk, C, b2ix, b2s, b2u, b2rr, b2, tb2, cw, tail_rows);
    }
    __syncthreads();

    // ---- phase 2: A1, A2, B1, self-energies, quadrature products -------------------------
    const double in3 = dB * dB, in4 = in3 * dB;
    const double b1a = 1 - 3 * in3 - 2 * in4, b1b = 3 * in3 + 2 * in4, b1c = dB + 2 * in3 + in4, b1d = in3 + in4;
    const double2* VF0 = H.V + (size_t)(F - 1) * N;
    const double2* VF1 = (a.vlast && F == T1 - 1) ? a.vlast : H.V + (size_t)F * N;
    const double2* DF1 = H.D + (size_t)F * N;
    const double2* DF2 = H.D + (size_t)(b1int ? F + 1 : F) * N;

    double acc[kSums];
#pragma unroll
    for (int q = 0; q < kSums; ++q) acc[q] = 0.0;
    double tacc[kSums * (K > 0 ? K : 1)];
#pragma unroll
    for (int q = 0; q < kSums * (K > 0 ? K : 1); ++q) tacc[q] = 0.0;

    for (int e = threadIdx.x; e < C; e += kRowThreads) {
        const int i = rhs_chunk_col(e, chunk, a.split);
        const long long k = (long long)j * N + i;
        double qkA1, qrA1, qkA2, qrA2, qkB1, qrB1, qkB2, qrB2;
        double2 ta1[K > 0 ? K : 1], ta2[K > 0 ? K : 1];  // tangent A1, A2 (K > 0)
        if (!live) {
            qkA1 = qkA2 = qkB1 = qkB2 = c00;
            qrA1 = qrA2 = qrB1 = qrB2 = c01;
        } else {
            // this entry's B1 rows are requested before the stencil weights
            const double2 y0 = VF0[i], y1 = VF1[i], dy1 = DF1[i], dy2 = DF2[i];
            double2 a1, a2;
            if constexpr (K > 0) {
                const bool dense = W.fA[0] != nullptr;
                stencil_last_multi<K>(T.w[0], P, k, m, T.inds[0][k], dense, vlast, tvlast, N, a1, ta1);
                stencil_last_multi<K>(T.w[1], P, k, m, T.inds[1][k], dense, vlast, tvlast, N, a2, ta2);
                if (dense) {
                    a1 = W.fA[0][k];
                    a2 = W.fA[1][k];
                }
            } else if (!LOGR && W.fA[0]) {  // dense grid: A1/A2 contracted beforehand on MFMA
                a1 = W.fA[0][k];
                a2 = W.fA[1][k];
            } else {
                a1 = stencil_last<LOGR>(T.w[0], P, k, m, T.inds[0][k], vlast);
                a2 = stencil_last<LOGR>(T.w[1], P, k, m, T.inds[1][k], vlast);
            }
            qkA1 = a1.x; qrA1 = a1.y; qkA2 = a2.x; qrA2 = a2.y;
            if (b1int) {
                qkB1 = b1a * y0.x + b1b * y1.x - b1c * dy1.x - b1d * dy2.x / b1rho;
                if (!LOGR) {
                    qrB1 = b1a * y0.y + b1b * y1.y - b1c * dy1.y - b1d * dy2.y / b1rho;
                } else if (y0.y > 0.0 && y1.y > 0.0) {
                    qrB1 = exp(b1a * log(y0.y) + b1b * log(y1.y) - b1c * (dy1.y / y0.y) - (b1d / b1rho) * (dy2.y / y1.y));
                } else {
                    qrB1 = b1a * y0.y + b1b * y1.y - b1c * dy1.y - b1d * dy2.y / b1rho;
                }
            } else {
                qkB1 = (1 - in3) * y0.x - (dB + in3) * dy1.x + in3 * y1.x;
                if (!LOGR || !(y0.y > 0.0 && y1.y > 0.0))
                    qrB1 = (1 - in3) * y0.y - (dB + in3) * dy1.y + in3 * y1.y;
                else
                    qrB1 = exp((1 - in3) * log(y0.y) + in3 * log(y1.y) - (dB + in3) * (dy1.y / y0.y));
            }
            const double2 vb2 = b2[e];
            qkB2 = vb2.x; qrB2 = vb2.y;
        }
        // self-energies (math_sigma.cpp:7-62)
        const double sKA1 = fl1(M, qkA1), sRA1 = fl2(M, qkA1) * qrA1;
        const double sRA2 = fl2(M, qkA2) * qrA2;
        const double sKB1 = fl1(M, qkB1), sRB1 = fl2(M, qkB1) * qrB1;
        const double sKB2 = fl1(M, qkB2), sRB2 = fl2(M, qkB2) * qrB2;
        if (a.fields) {
            double* f = a.fields;
            const double v[15] = {qkA1, qrA1, qkA2, qrA2, qkB1, qrB1, qkB2, qrB2,
                                  sKA1, sRA1, sRA2, sKB1, sRB1, sKB2, sRB2};
#pragma unroll
            for (int q = 0; q < 15; ++q) f[(long long)q * P + k] = v[q];
        }
        if (i == 0) W.qkb1c0[j] = qkB1;
        if constexpr (K > 0) {
#pragma unroll
            for (int h = 0; h < K; ++h) {
                double t[8];  // tangent qkA1 qrA1 qkA2 qrA2 qkB1 qrB1 qkB2 qrB2
                if (!live) {
                    const double2 v0 = X->V[h][0];
                    t[0] = t[2] = t[4] = t[6] = v0.x;
                    t[1] = t[3] = t[5] = t[7] = v0.y;
                } else {
                    t[0] = ta1[h].x;
                    t[1] = ta1[h].y;
                    t[2] = ta2[h].x;
                    t[3] = ta2[h].y;
                    const double2* tV = X->V[h];
                    const double2* tD = X->D[h];
                    const double2 y0 = tV[(size_t)(F - 1) * N + i], y1 = tV[(size_t)F * N + i];
                    const double2 dy1 = tD[(size_t)F * N + i], dy2 = tD[(size_t)(b1int ? F + 1 : F) * N + i];
                    if (b1int) {
                        t[4] = b1a * y0.x + b1b * y1.x - b1c * dy1.x - b1d * dy2.x / b1rho;
                        t[5] = b1a * y0.y + b1b * y1.y - b1c * dy1.y - b1d * dy2.y / b1rho;
                    } else {
                        t[4] = (1 - in3) * y0.x - (dB + in3) * dy1.x + in3 * y1.x;
                        t[5] = (1 - in3) * y0.y - (dB + in3) * dy1.y + in3 * y1.y;
                    }
                    const double2 tb = tb2[(size_t)h * cw + e];
                    t[6] = tb.x;
                    t[7] = tb.y;
                }
                const double l = X->dL[h];
                const double dsKA1 = fl2(M, qkA1) * t[0] + l * fl1l(M, qkA1);
                const double dsRA1 = (fl3(M, qkA1) * t[0] + l * fl2l(M, qkA1)) * qrA1 + fl2(M, qkA1) * t[1];
                const double dsRA2 = (fl3(M, qkA2) * t[2] + l * fl2l(M, qkA2)) * qrA2 + fl2(M, qkA2) * t[3];
                const double dsKB1 = fl2(M, qkB1) * t[4] + l * fl1l(M, qkB1);
                const double dsRB1 = (fl3(M, qkB1) * t[4] + l * fl2l(M, qkB1)) * qrB1 + fl2(M, qkB1) * t[5];
                const double dsKB2 = fl2(M, qkB2) * t[6] + l * fl1l(M, qkB2);
                const double dsRB2 = (fl3(M, qkB2) * t[6] + l * fl2l(M, qkB2)) * qrB2 + fl2(M, qkB2) * t[7];
                const double wt = T.integ[i];
                double* ac = tacc + kSums * h;
                ac[0] += wt * (dsRA2 * qkB2 + sRA2 * t[6]);
                ac[1] += wt * (dsRA1 * qkB1 + sRA1 * t[4]);
                ac[2] += wt * (dsKA1 * qrB1 + sKA1 * t[5]);
                ac[3] += wt * (t[3] * sKB2 + qrA2 * dsKB2);
                ac[4] += wt * (t[1] * sKB1 + qrA1 * dsKB1);
                ac[5] += wt * (t[0] * sRB1 + qkA1 * dsRB1);
                ac[6] += wt * (dsRA2 * qrB2 + sRA2 * t[7]);
                ac[7] += wt * (t[3] * sRB2 + qrA2 * dsRB2);
                if (i == 0) X->qkb1c0[(size_t)h * N + j] = t[4];
            }
        }
        // quadratures of time_steps.cpp:36-45, 97-98 (weights int.dat)
        const double wi = T.integ[i];
        acc[0] += wi * sRA2 * qkB2;  // ConvR(SigmaRA2, QKB2)
        acc[1] += wi * sRA1 * qkB1;  // ConvA(SigmaRA1, QKB1)
        acc[2] += wi * sKA1 * qrB1;  // ConvA(SigmaKA1, QRB1)
        acc[3] += wi * qrA2 * sKB2;  // ConvR(QRA2, SigmaKB2)
        acc[4] += wi * qrA1 * sKB1;  // ConvA(QRA1, SigmaKB1)
        acc[5] += wi * qkA1 * sRB1;  // ConvA(QKA1, SigmaRB1)
        acc[6] += wi * sRA2 * qrB2;  // ConvR(SigmaRA2, QRB2)
        acc[7] += wi * qrA2 * sRB2;  // ConvR(QRA2, SigmaRB2)
    }
    const double tot = block_sum8<kRowThreads>(acc, red);
    if (threadIdx.x < kSums) W.partial[((size_t)j * a.split + chunk) * kSums + threadIdx.x] = tot;
    if constexpr (K > 0) {
#pragma unroll
        for (int h = 0; h < K; ++h) {
            double v8[kSums];
#pragma unroll
            for (int q = 0; q < kSums; ++q) v8[q] = tacc[kSums * h + q];
            const double tt = block_sum8<kRowThreads>(v8, red);
            if (threadIdx.x < kSums) X->partial[(((size_t)h * N + j) * a.split + chunk) * kSums + threadIdx.x] = tt;
        }
    }
}

template <bool LOGR, int G, int NT>
__global__ __launch_bounds__(kRowThreads) void rhs_rows_kernel(Tables T, History H, Work W, Model M, RhsArgs a) {
    extern __shared__ double2 dyn[];
    __shared__ double red[kSums * (kRowThreads / kWave)];
    const int lb = xcd_stripe<32>(blockIdx.x, gridDim.x);
    rhs_rows_body<LOGR, G, 0, NT>(T, H, W, M, a, lb / a.split, lb % a.split, dyn, red);
}

// the RHS with the tangent quadratures of K sensitivity directions (linear-response form)
template <int G, int K>
__global__ __launch_bounds__(kRowThreads) void rhs_rows_tan_kernel(Tables T, History H, Work W, Model M, RhsArgs a,
                                                                   TanArgs X) {
    extern __shared__ double2 dyn[];
    __shared__ double red[kSums * (kRowThreads / kWave)];
    const int lb = xcd_stripe<32>(blockIdx.x, gridDim.x);
    rhs_rows_body<false, G, K>(T, H, W, M, a, lb / a.split, lb % a.split, dyn, red, &X);
}

// ---- finish: RHS epilogue + RK stage -------------------------------------------------
template <int EPT, bool PRE = true>
__device__ __forceinline__ void finish_block(const Tables& T, const History& H, const Work& W, const Model& M, const FinishArgs& a,
                             double* red /* 8 * kStageThreads / 64 */) {
    const int N = T.N;
    const int n = a.stage;
    const double t = a.t_rhs;
    const double2* last = H.V + (size_t)a.lr * N;
    const double qK0 = last[0].x;
    // rInt (index_vec.cpp:138-156): the current interpolation's terms over the older rows
    // plus those of its newest row a.lr (rint_finish)
    const double4* Rq = W.rpk + (size_t)a.fbuf * N;
    const double rw = H.r[a.lr], drw = H.dr[a.lr], rhow = H.rho[a.lr];
    // operands of the stage combination (earlier stages' RHS, the step's start row) do not
    // depend on this stage's RHS: they are requested first, so that their latency
    // overlaps the RHS epilogue and the reductions below
    // (earlier stages' RHS values are held in registers for EPT <= 2 only, and not in the
    // many-instance batch, whose stage launch is throughput-bound and keeps occupancy)
    constexpr bool kHold = PRE && EPT <= 2;
    constexpr int kPre = kHold ? kMaxStages : 1;
    double2 g0v[EPT];
    double hpk[EPT][kPre], hpr[EPT][kPre];
    // drstep2 of this step (stage 0 wrote it), read before this block's stores
    const double scal0 = (a.mode != 0 && n != 0) ? W.scal[0] : 0.0;
    if (a.mode != 0) {
        const double2* g0 = H.V + (size_t)a.g0row * N;
#pragma unroll
        for (int e = 0; e < EPT; ++e) {
            const int j = threadIdx.x + e * kStageThreads;
            g0v[e] = make_double2(0.0, 0.0);
#pragma unroll
            for (int q = 0; q < kPre; ++q) hpk[e][q] = hpr[e][q] = 0.0;
            if (j >= N) continue;
            g0v[e] = g0[j];
#pragma unroll
            for (int q = 0; q < kPre; ++q) {
                if (kHold && q < n && q < a.nterms) {
                    hpk[e][q] = W.hK[(size_t)q * N + j];
                    hpr[e][q] = W.hR[(size_t)q * N + j];
                }
            }
        }
    }
    const double s1 = fl1(M, qK0) / M.T0, s2 = -rint_finish(Rq[N - 1], rw, drw, rhow), s3 = qK0 / M.T0,
                 s4 = 2 * M.Gamma;
    double hk[EPT], hr[EPT];
#pragma unroll
    for (int e = 0; e < EPT; ++e) {
        const int j = threadIdx.x + e * kStageThreads;
        hk[e] = hr[e] = 0.0;
        if (j >= N) continue;
        double C[kSums];
#pragma unroll
        for (int q = 0; q < kSums; ++q) C[q] = 0.0;
        for (int c = 0; c < a.split; ++c) {
            const double* pp = W.partial + ((size_t)j * a.split + c) * kSums;
#pragma unroll
            for (int q = 0; q < kSums; ++q) C[q] += pp[q];
        }
        const double th = T.theta[j];
        const double fA = t * th, fR = t * (1 - th);
        const double ri = rint_finish(Rq[j], rw, drw, rhow);
        W.rInt[j] = ri;
        const double qK = last[j].x, qR = last[j].y, b1 = W.qkb1c0[j];
        double d1 = b1 * s1 + qK * s2;
        d1 = d1 + C[0] * fR;
        d1 = d1 + C[1] * fA;
        d1 = d1 + C[2] * fA;
        double d2 = fl1(M, b1) * s3 + qR * s4;
        d2 = d2 + C[3] * fR;
        d2 = d2 + C[4] * fA;
        d2 = d2 + C[5] * fA;
        d2 = d2 - qK * ri;
        hk[e] = d1 + d2 * th;
        hr[e] = (qR * s2 + C[6] * fR) + (qR * ri - C[7] * fR) * th;
        W.hK[(size_t)n * N + j] = hk[e];
        W.hR[(size_t)n * N + j] = hr[e];
    }
    if (a.mode == 0) return;

    // drstep2 at stage 0 (time_steps.cpp:117-125) on the step's start row
    if (n == 0) {
        double v[6] = {0, 0, 0, 0, 0, 0};
        double dsK0 = 0.0, sK0 = 0.0;
#pragma unroll
        for (int e = 0; e < EPT; ++e) {
            const int j = threadIdx.x + e * kStageThreads;
            if (j >= N) continue;
            const double qK = last[j].x, qR = last[j].y, w = T.integ[j];
            const double sK = fl1(M, qK), sR = fl2(M, qK) * qR;
            const double dsK = fl2(M, qK) * hk[e] + 0.0 * hr[e];
            const double dsR = (fl3(M, qK) * qR) * hk[e] + fl2(M, qK) * hr[e];
            v[0] += (1.0 * w) * sR * qK;
            v[1] += (1.0 * w) * sK * qR;
            v[2] += (t * w) * dsR * qK;
            v[3] += (t * w) * dsK * qR;
            v[4] += (t * w) * sR * hk[e];
            v[5] += (t * w) * sK * hr[e];
            if (j == 0) { dsK0 = dsK; sK0 = sK; }
        }
        block_sum<6, kStageThreads>(v, red);
        if (threadIdx.x == 0) {
            const double d = v[0] + v[1] + v[2] + v[3] + v[4] + v[5] + (dsK0 * qK0 + sK0 * hk[0]) / M.T0;
            W.scal[0] = d;
            red[0] = d;  // block_sum is done with red
        }
        __syncthreads();
    }
    double dr;
    if (n == 0) dr = red[0];  // LDS and global reads kept apart (no generic-pointer load)
    else dr = scal0;
    // stage combination g_{n+1} = g_0 + sum_j h_j * (dt a_{n+1,j}) and the row write
    double2* wv = H.V + (size_t)a.wrow * N;
    double2* wd = H.D + (size_t)a.wrow * N;
    double rr[2] = {0.0, 0.0};
    double err[1] = {0.0};
    double gK0 = 0.0, sKg0 = 0.0;
#pragma unroll
    for (int e = 0; e < EPT; ++e) {
        const int j = threadIdx.x + e * kStageThreads;
        if (j >= N) continue;
        const double2 g = g0v[e];
        double gK = g.x, gR = g.y;
        double errj = 0.0;
        if (a.serk) {
            // SERK2 (runge_kutta.cu:397-487): g is the previous stage (the start row at the
            // first stage); serk 1 an Euler step, 2 the Chebyshev step against W.sE
            if (a.serk == 2) {
                const double2 ge = W.sE[j];
                gK = 2 * g.x - ge.x + hk[e] * a.sa;
                gR = 2 * g.y - ge.y + hr[e] * a.sa;
                W.sE[j] = make_double2(gK, gR);
            } else {
                gK = g.x + hk[e] * a.sa;
                gR = g.y + hr[e] * a.sa;
                if (n == 0) W.sE[j] = g;
            }
            double2 F = (n == 0) ? make_double2(g.x * a.sb0, g.y * a.sb0) : W.sF[j];
            if (a.sb != 0.0) {
                F.x = F.x + gK * a.sb;
                F.y = F.y + gR * a.sb;
            }
            W.sF[j] = F;
            if (a.last) errj = fabs(F.x - gK) + fabs(F.y - gR);
        } else {
            // static trip count: a.coef stays in registers when `a` is a local copy
#pragma clang loop unroll(full)
            for (int q = 0; q < kMaxStages; ++q) {
                if (q < a.nterms) {
                    double hq, hqr;
                    if constexpr (kHold) {
                        hq = (q == n) ? hk[e] : hpk[e][q];
                        hqr = (q == n) ? hr[e] : hpr[e][q];
                    } else {
                        hq = (q == n) ? hk[e] : W.hK[(size_t)q * N + j];
                        hqr = (q == n) ? hr[e] : W.hR[(size_t)q * N + j];
                    }
                    gK = gK + hq * a.coef[q];
                    gR = gR + hqr * a.coef[q];
                }
            }
            if (a.last) {
                double eK = g.x, eR = g.y;
#pragma clang loop unroll(full)
                for (int q = 0; q < kMaxStages; ++q) {
                    if (q <= n) {
                        double hq, hqr;
                        if constexpr (kHold) {
                            hq = (q == n) ? hk[e] : hpk[e][q];
                            hqr = (q == n) ? hr[e] : hpr[e][q];
                        } else {
                            hq = (q == n) ? hk[e] : W.hK[(size_t)q * N + j];
                            hqr = (q == n) ? hr[e] : W.hR[(size_t)q * N + j];
                        }
                        eK = eK + hq * a.coefE[q];
                        eR = eR + hqr * a.coefE[q];
                    }
                }
                errj = fabs(gK - eK) + fabs(gR - eR);
            }
        }
        const double h0k = (n == 0) ? hk[e] : (kHold ? hpk[e][0] : W.hK[j]);
        const double h0r = (n == 0) ? hr[e] : (kHold ? hpr[e][0] : W.hR[j]);
        wv[j] = make_double2(gK, gR);
        wd[j] = make_double2(a.tdiff * h0k, a.tdiff * h0r);
        // rstep on the new row (time_steps.cpp:83-92)
        const double w = T.integ[j];
        const double sK = fl1(M, gK), sR = fl2(M, gK) * gR;
        rr[0] += a.t_new * w * sR * gK;
        rr[1] += a.t_new * w * sK * gR;
        if (j == 0) { gK0 = gK; sKg0 = sK; }
        err[0] += errj;
    }
    block_sum<2, kStageThreads>(rr, red);
    if (a.last) block_sum<1, kStageThreads>(err, red);
    if (threadIdx.x == 0) {
        H.t1[a.wrow] = a.t_new;
        H.rho[a.wrow] = a.rho_new;
        H.r[a.wrow] = M.Gamma + rr[0] + rr[1] + sKg0 * gK0 / M.T0;
        H.dr[a.wrow] = a.tdiff * dr;
        if (a.last) W.scal[1] = err[0] + a.gt_err;
    }
}

// ---- QK tail fit near theta -> 1 (tail_fit.cpp:46-265, tailFitBlendCPU) ---------------
// After a main-loop step: y_j = (1 - QK_j) / (1 - theta_j)^2 on the upper half of the
// last row, the constant fit of least SSE over windows of 16 (all windows on the first
// call or a repeated call on the same history length, else the previous window and its
// neighbours), alpha smoothed with an error-dependent gain (reset by a rollback), and
// where the residual and SSE gates pass, QK -> 1 - alpha (1 - theta)^2 from the window
// on with a linear ramp of 8 or 16 points.  Sums run in index order (thread 0 for the
// cross-window ones), the reference's own plain loops; its omp simd window sums may
// differ from this order by rounding.  One block; `pre` receives the row before the blend.
#pragma clang fp contract(off)
__device__ void tail_fit_body(const double* __restrict__ theta, double2* row, double2* pre, TailState* st, int N,
                              long long t1len, double delta_max, long long rollback_loop, double* lds /* >= 2N + 96 */) {
    for (int j = threadIdx.x; j < N; j += blockDim.x) pre[j] = row[j];
    const int start = N / 2, W = 16;
    if (N < 32 || t1len == 0 || N - start < W) return;
    const int Nt = N - start;
    double* y = lds;             // [Nt]
    double* qn = lds + Nt;       // [Nt] QK of the row
    double* om = lds + 2 * Nt;   // [Nt] 1 - theta
    double* wm = lds + 3 * Nt;   // [32] window means
    double* ws = wm + 32;        // [32] window SSE
    double* bc = ws + 32;        // [8] broadcast
    for (int j = start + threadIdx.x; j < N; j += blockDim.x) {
        const double one_m = 1.0 - theta[j];
        const double denom = one_m * one_m;
        const double qk = row[j].x;
        y[j - start] = (1.0 - qk) / (denom > 0 ? denom : 1e-300);
        qn[j - start] = qk;
        om[j - start] = one_m;
    }
    const long long prev_off = st->prev_off, prev_t1len = st->prev_t1len;
    const bool full = prev_off < 0 || prev_t1len == t1len;
    // candidate windows in increasing offset: all of them, or prev - W, prev, prev + W
    int ncand = 0;
    long long cand[3] = {0, 0, 0};
    if (full) {
        ncand = (Nt - W) / W + 1;
    } else {
        if (prev_off >= W) cand[ncand++] = prev_off - W;
        cand[ncand++] = prev_off;
        if (prev_off + 2 * W <= Nt) cand[ncand++] = prev_off + W;
    }
    __syncthreads();
    if ((int)threadIdx.x < ncand) {
        const long long off = full ? (long long)threadIdx.x * W : cand[threadIdx.x];
        double sum = 0.0, err = 0.0, mean = 0.0;
        if (off + W <= Nt) {
            for (int k = 0; k < W; ++k) sum += y[off + k];
            mean = sum / (double)W;
            for (int k = 0; k < W; ++k) {
                const double d = y[off + k] - mean;
                err += d * d;
            }
        } else {
            err = __builtin_inf();  // skipped candidate (never chosen)
        }
        wm[threadIdx.x] = mean;
        ws[threadIdx.x] = err;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        long long best_off = 0;
        double best_sse = __builtin_inf(), best_mean = 0.0;
        for (int q = 0; q < ncand; ++q) {
            const long long off = full ? (long long)q * W : cand[q];
            if (off + W > Nt) continue;
            if (ws[q] < best_sse) {
                best_sse = ws[q];
                best_off = off;
                best_mean = wm[q];
            }
        }
        st->prev_off = best_off;
        st->prev_t1len = t1len;
        const double alpha_raw = best_mean;
        const int j0 = start + (int)best_off;
        const double var = best_sse / (double)(W - 1);
        const double stderr_mean = sqrt((var > 0.0 ? var : 0.0) / (double)W);
        double alpha_s = st->alpha;
        const bool rollback_detected = rollback_loop != st->known_rollback;
        if (rollback_detected) {
            st->known_rollback = rollback_loop;
            if (alpha_raw != 0.0) alpha_s = alpha_raw;
        }
        const double max_gain = 0.50, min_gain = 0.10;
        const double rel_err = fabs(alpha_raw) > 1e-12 ? stderr_mean / fabs(alpha_raw) : 1.0;
        double gain = max_gain / (1.0 + 2.0 * rel_err);
        gain = gain < max_gain ? gain : max_gain;
        gain = min_gain > gain ? min_gain : gain;
        if (!(alpha_s == alpha_s) || (alpha_raw == 0.0 && rollback_detected)) {
            if (alpha_raw != 0.0) alpha_s = alpha_raw;
            else if (!(alpha_s == alpha_s)) alpha_s = 0.0;
        } else if (alpha_raw != 0.0) {
            alpha_s = (1.0 - gain) * alpha_s + gain * alpha_raw;
        }
        st->alpha = alpha_s;
        const double alpha = alpha_s;
        const int nres = N - j0;
        double mean_abs = 0.0, mean_res = 0.0, rv = 0.0;
        for (int j = j0; j < N; ++j) {
            const double o = om[j - start];
            mean_abs += fabs((1.0 - alpha * o * o) - qn[j - start]);
        }
        mean_abs /= (double)nres;
        for (int j = j0; j < N; ++j) {
            const double o = om[j - start];
            mean_res += (1.0 - alpha * o * o) - qn[j - start];
        }
        mean_res /= (double)nres;
        for (int j = j0; j < N; ++j) {
            const double o = om[j - start];
            const double d = ((1.0 - alpha * o * o) - qn[j - start]) - mean_res;
            rv += d * d;
        }
        const double den = (double)nres - 1.0;
        rv /= (1.0 > den ? 1.0 : den);
        const double res_std = sqrt(0.0 > rv ? 0.0 : rv);
        const bool apply = mean_abs <= 3.0 * res_std && best_sse <= sqrt(delta_max) && fabs(alpha) > 1e-14;
        const double ad = fabs(alpha) > 1e-12 ? fabs(alpha) : 1e-12;
        const double alpha_change = fabs(alpha_raw - alpha) / ad;
        bc[0] = apply ? 1.0 : 0.0;
        bc[1] = alpha;
        bc[2] = (double)j0;
        bc[3] = alpha_change < 0.005 ? 8.0 : 16.0;
        if (apply) st->applied += 1;
    }
    __syncthreads();
    if (bc[0] == 0.0) return;
    const double alpha = bc[1];
    const int j0 = (int)bc[2];
    const int ramp = (int)bc[3];
    for (int j = j0 + threadIdx.x; j < N; j += blockDim.x) {
        const double o = om[j - start];
        const double qk_fit = 1.0 - alpha * o * o;
        double w = 1.0;
        if (j < j0 + ramp) w = (double)(j - j0) / (double)ramp;
        const double qk_num = qn[j - start];
        row[j].x = (1.0 - w) * qk_num + w * qk_fit;
    }
}
#pragma clang fp contract(on)

// ---- one launch per stage boundary ------------------------------------------------
// Block 0 (optional): finish of the previous stage.  Then (optional) the theta-row
// search and the N phi2-row searches of the next interpolation.  The searches read
// t1grid through ListView, whose newest entry comes from the argument, so they do
// not depend on the finish block's writes; everything that does (the RHS kernel)
// is launched afterwards.
enum StageFlags { kHasFinish = 1, kHasTheta = 2, kHasRows = 4 };

template <int EPT>
__global__ __launch_bounds__(kStageThreads) void stage_kernel(Tables T, History H, Work W, Model M, SearchArgs sa,
                                                              FinishArgs fa, int flags) {
    __shared__ double lds[EPT * kStageThreads];
    __shared__ double red[8 * (kStageThreads / kWave)];
    int b = blockIdx.x;
    if (flags & kHasFinish) {
        if (b == 0) {
            finish_block<EPT>(T, H, W, M, fa, red);
            return;
        }
        --b;
    }
    if (flags & kHasTheta) {
        if (b == 0) {
            search_theta_row<EPT>(T, H, W, sa, lds, red);
            return;
        }
        --b;
    }
    if ((flags & kHasRows) && b < T.N) search_phi2_row<EPT>(T, W, sa, b, lds);
}

// rstep (what 0) / energy (what 1) on one history row
// rstep (what 0, time_steps.cpp:83-92) or the energy (what 1, simulation_runner.cpp:308-318)
// of history row `row` at time t; the value is returned to thread 0.
__device__ double row_scalar_body(const Tables& T, const History& H, const Model& M, int row, double t, int what,
                                  double* red /* 2 * kStageThreads / 64 */) {
    const int N = T.N;
    const double2* v = H.V + (size_t)row * N;
    double s[2] = {0.0, 0.0};
    for (int j = threadIdx.x; j < N; j += kStageThreads) {
        const double qK = v[j].x, qR = v[j].y, w = T.integ[j];
        const double sK = fl1(M, qK);
        if (what == 0) {
            s[0] += t * w * (fl2(M, qK) * qR) * qK;
            s[1] += t * w * sK * qR;
        } else {
            s[0] += t * w * sK * qR;
        }
    }
    block_sum<2, kStageThreads>(s, red);
    const double qK0 = v[0].x;
    if (what == 0) return M.Gamma + s[0] + s[1] + fl1(M, qK0) * qK0 / M.T0;
    return -(s[0] + fl1(M, qK0) / M.T0);
}

__global__ __launch_bounds__(kStageThreads) void row_scalar_kernel(Tables T, History H, Work W, Model M, int row,
                                                                   double t, int what) {
    __shared__ double red[2 * (kStageThreads / kWave)];
    const double v = row_scalar_body(T, H, M, row, t, what, red);
    if (threadIdx.x == 0) W.scal[2] = v;
}

// sparsifyNscale values (sparsify_utils.cpp:19-49): one workgroup per interior row i
__global__ __launch_bounds__(kRowThreads) void sparsify_values_kernel(Tables T, History H, int T1, double* val) {
    __shared__ double red[kRowThreads / kWave];
    const int i = blockIdx.x + 2;
    if (i + 1 >= T1) return;
    const int N = T.N;
    const double tleft = H.t1[i - 2], tmid = H.t1[i];
    const double td1 = H.t1[i - 1] - tleft, td2 = tmid - tleft, td3 = H.t1[i + 1] - tmid;
    double s[1] = {0.0};
    for (int j = threadIdx.x; j < N; j += kRowThreads) {
        const double2 a = H.V[(size_t)i * N + j], b = H.V[(size_t)(i - 2) * N + j];
        const double2 d1 = H.D[(size_t)(i - 1) * N + j], d2 = H.D[(size_t)(i + 1) * N + j];
        s[0] += fabs(td2 / 12.0 * (2 * (a.x - b.x) - td2 * (d1.x / td1 + d2.x / td3)));
        s[0] += fabs(td2 / 12.0 * (2 * (a.y - b.y) - td2 * (d1.y / td1 + d2.y / td3)));
    }
    block_sum<1, kRowThreads>(s, red);
    if (threadIdx.x == 0) {
        const double Rt = H.r[i] - H.r[i - 2];
        val[i] = s[0] + fabs(td2 / 12.0 * (2 * Rt - td2 * (H.dr[i - 1] / td1 + H.dr[i + 1] / td3)));
    }
}

__global__ __launch_bounds__(kRowThreads) void gather_rows_kernel(Tables T, History src, History dst, const int* keep,
                                                                  int nkeep) {
    const int r = blockIdx.x;
    if (r >= nkeep) return;
    const int s = keep[r];
    const int N = T.N;
    for (int j = threadIdx.x; j < N; j += kRowThreads) {
        dst.V[(size_t)r * N + j] = src.V[(size_t)s * N + j];
        dst.D[(size_t)r * N + j] = src.D[(size_t)s * N + j];
    }
    if (threadIdx.x == 0) {
        dst.t1[r] = src.t1[s];
        dst.rho[r] = src.rho[s];
        dst.r[r] = src.r[s];
        dst.dr[r] = src.dr[s];
    }
}

// ---- batched quenches: device-side step control -------------------------------------
// One batch step advances every running instance by one iteration of the main loop
// (simulation_runner.cpp:125-278): begin (stage times) -> for each stage n the stage
// launch (finish of n-1 + searches of n) and the RHS launch -> final stage launch ->
// end (error norm -> dt rule).  Launch arguments are fixed; everything that varies is
// read from the per-instance Ctl, so a step can be captured once into a hipGraph.
// The scalar arithmetic repeats the reference's FP64 order without contraction.


// The per-step scalars that the begin kernel's argument builders read: thread 0 writes
// them to LDS as well as to the instance's Ctl, so the builders need no global round trip.
struct StepHead {
    int T1s, S, fc0, method;
    double dt, tprev, gt_err;
    double gt[kMaxStepStages + 1];
};

#pragma clang fp contract(off)
// C is Ctl or StepHead
template <class C>
__device__ inline int searches_before(const C& c, int k) {  // searches in stages 0..k-1 of this step
    return k <= 0 ? 0 : (c.T1s == 1 ? 1 : 0) + (k - 1);
}

template <class C>
__device__ inline bool make_search_dev(const C& c, const InstDesc& d, int n, SearchArgs& a) {
    const int S = c.S;
    const bool ss = c.method != DMFE_RK54;
    a = SearchArgs{};
    a.t1 = d.H.t1;
    a.tl = c.gt[n];
    if (n == S) {
        a.mode = ss ? kSearchI : kSearchSame;
        a.T1 = c.T1s + 1;
    } else if (n == 0) {
        if (c.T1s != 1) {
            a.mode = -1;
            return false;
        }
        a.mode = kSearchB;
        a.T1 = c.T1s;
    } else {
        a.T1 = c.T1s + 1;
        a.mode = kSearchI;
        // SSPRK104 reuses the positions of stages 2..4 at stages 5..7 (runge_kutta.cpp:60-110);
        // SERK2 searches at every stage
        const bool reuse = c.method == DMFE_SSPRK104;
        if (reuse && n >= 5 && n <= 7) {
            a.mode = kSearchCopy;
            a.srcB1 = d.W.saveB1 + (size_t)(n - 5) * d.T.N;
            a.srcB2 = d.W.saveB2 + (size_t)(n - 5) * d.T.P;
        } else if (reuse && n >= 2 && n <= 4) {
            a.saveB1 = d.W.saveB1 + (size_t)(n - 2) * d.T.N;
            a.saveB2 = d.W.saveB2 + (size_t)(n - 2) * d.T.P;
        }
    }
    a.fbuf = (c.fc0 + searches_before(c, n) + 1) & 1;
    if (d.gpu_variants) {
        // the reference GPU path bisects at every interpolation (interpolateGPU ignores its
        // input positions and `same`, interpolation_core.cu:65-72)
        a.mode = kSearchB;
        a.srcB1 = a.srcB2 = nullptr;
        a.gpu_clamp = 1;
    }
    return true;
}

// Finish of SERK2 stage m (method >= 3: q = 2 (method - 2), S = 2q stages;
// runge_kutta.cu:397-487 as the host-driven rk_step_serk2 sequences it): g += dt alpha h
// at even m, g = 2 g - g_e + 2 dt alpha h at odd m, alpha = 1 / (4q), every stage
// rewriting the row appended at stage 0; the embedded sum takes b_0 of the start row and
// b_{m+1} of each stage.  gt[] holds the stage times the begin kernel formed.
template <class C>
__device__ inline void make_finish_serk_dev(const C& c, const InstDesc& d, int m, FinishArgs& f) {
    const int S = c.S, q = S / 2;
    const double dt = c.dt, alpha = 1.0 / (4 * q);
    f = FinishArgs{};
    f.mode = 1;
    f.stage = m == 0 ? 0 : 1;
    f.nterms = 1;
    f.split = d.split;
    f.lr = (m == 0) ? c.T1s - 1 : c.T1s;
    f.g0row = (m == 0) ? c.T1s - 1 : c.T1s;  // the start row, then the previous stage's
    f.wrow = c.T1s;
    f.t_rhs = c.gt[m];
    f.t_new = c.gt[m + 1];
    f.serk = (m % 2 == 1) ? 2 : 1;
    f.sa = (m % 2 == 1) ? 2 * dt * alpha : dt * alpha;
    f.sb0 = c.serk_b[0];
    f.sb = c.serk_b[m + 1];
    f.tdiff = c.gt[m + 1] - c.gt[0];
    f.rho_new = (c.T1s > 1) ? f.tdiff / (c.gt[0] - c.tprev) : 0.0;
    f.last = (m == S - 1) ? 1 : 0;
    if (f.last) f.gt_err = c.gt_err;
    f.fbuf = (c.fc0 + searches_before(c, m + 1)) & 1;
}


// end: delta, loop, maintenance cadence and the adaptive dt rule (simulation_runner.cpp:144-278)
__device__ __forceinline__ void end_body(const InstDesc& d) {
    Ctl& c = *d.ctl;
    // every operand is loaded before the first test and the first store (one memory
    // round trip); rho of the new last row is the value its finish wrote (Ctl::rho_step)
    const int active = c.active, T1s = c.T1s, fc0 = c.fc0, S = c.S, method = c.method, rollback = c.rollback;
    const long long loop = c.loop + 1, maint_every = c.maint_every;
    const double delta = d.W.scal[1], delta_old = c.delta_old, dmax = c.delta_max, specRad = c.specRad;
    const double dt_min = c.dt_min, rho_last = c.rho_step;
    double dt = c.dt;
    double rmax[11];
#pragma unroll
    for (int q = 0; q < 11; ++q) rmax[q] = c.rmax[q];
    if (!active) return;  // stays set until the next begin: the observables launch reads it
    c.delta = delta;
    c.loop = loop;
    c.T1 = T1s + 1;
    StepHead h{};
    h.T1s = T1s;
    c.fcur = (fc0 + searches_before(h, S) + 1) & 1;
    if (maint_every > 0 && loop % maint_every == 0) {
        c.status = kNeedMaint;  // sparsify, interpolate and the dt rule run on the host
        return;
    }
    double rm = rmax[0];  // rmax[method - 1] without dynamic register indexing
#pragma unroll
    for (int q = 1; q < 11; ++q) rm = (method - 1 == q) ? rmax[q] : rm;
    if (delta < dmax && loop > 5 && (delta < 1.1 * delta_old || delta_old < dmax / 1000) && rm / specRad > dt &&
        rho_last == 1.0) {
        dt *= 1.01;
    } else if (delta > 2 * dmax && dt > dt_min) {
        dt *= 0.9;
    }
    if (method == DMFE_SSPRK104 && delta > dmax && rmax[0] / specRad > dt) {
        c.method = DMFE_RK54;
        dt *= 0.5;
    }
    c.dt = dt;
    // the rollback (simulation_runner.cpp:280-303) truncates the history and interpolates
    // afresh: the host does it between graph replays
    if (rollback && delta > 2 * dmax) c.status = kNeedRollback;
}

__global__ void batch_end_kernel(InstDesc* D, int n) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) end_body(D[i]);
}

__global__ void loop_end_kernel(InstDesc d) {
    if (threadIdx.x == 0) end_body(d);
}
#pragma clang fp contract(on)

// Stage launch n of a batch step: map[b] = (instance, role); role -2 finish of stage
// n-1, -1 theta-row search, k >= 0 phi2 row k of the interpolation of stage n.
template <int EPT, bool PRE>
__device__ __forceinline__ void stage_role_body(const InstDesc& d, int role, int n, double* lds, double* red) {
    const Ctl& c = *d.ctl;
    // the control word and this block's arguments are loaded together (one memory round
    // trip) before they are tested.  The device loop (PRE, about two blocks per CU) also
    // keeps its finish arguments in registers across the block's global stores, which
    // could otherwise alias them; the many-instance batch keeps occupancy instead.
    // SERK2 steps (method >= 3, up to kMaxStepStages stages) form the arguments here from
    // the stage times; the other integrators' were written by the begin kernel
    if (role == -2) {
        if (n < 1) return;
        if constexpr (PRE) {
            const int active = c.active, S = c.S, method = c.method;
            if (method >= 3) {
                FinishArgs fa;
                if (active && n <= S) {
                    make_finish_serk_dev(c, d, n - 1, fa);
                    finish_block<EPT, PRE>(d.T, d.H, d.W, d.M, fa, red);
                }
                return;
            }
            // (the index is clamped: a many-stage step of another instance launches more stages)
            const FinishArgs fa = c.fin[min(n - 1, kMaxStages - 1)];
            if (active && n <= S) finish_block<EPT, PRE>(d.T, d.H, d.W, d.M, fa, red);
        } else if (c.active && n <= c.S) {
            if (c.method >= 3) {
                FinishArgs fa;
                make_finish_serk_dev(c, d, n - 1, fa);
                finish_block<EPT, PRE>(d.T, d.H, d.W, d.M, fa, red);
            } else {
                finish_block<EPT, PRE>(d.T, d.H, d.W, d.M, c.fin[n - 1], red);
            }
        }
        return;
    }
    const int active = c.active, S = c.S;
    SearchArgs a;
    if (c.method >= 3) {
        if (!active || n > S) return;
        make_search_dev(c, d, n, a);
    } else {
        a = c.srch[min(n, kMaxStages)];
    }
    if (!active || n > S || a.mode < 0) return;
    if (role == -1) search_theta_row<EPT>(d.T, d.H, d.W, a, lds, red);
    else search_phi2_row<EPT>(d.T, d.W, a, role, lds);
}

template <int EPT>
__global__ __launch_bounds__(kStageThreads) void batch_stage_kernel(const InstDesc* __restrict__ D,
                                                                    const int2* __restrict__ map, int n) {
    __shared__ double lds[EPT * kStageThreads];
    __shared__ double red[8 * (kStageThreads / kWave)];
    const int2 mp = map[blockIdx.x];
    if (mp.x < 0) return;
    stage_role_body<EPT, false>(D[mp.x], mp.y, n, lds, red);
}

// device loop of one context: block 0 finish, block 1 theta row, block 2 + k phi2 row k
template <int EPT>
__global__ __launch_bounds__(kStageThreads) void loop_stage_kernel(InstDesc d, int n) {
    __shared__ double lds[EPT * kStageThreads];
    __shared__ double red[8 * (kStageThreads / kWave)];
    stage_role_body<EPT, true>(d, (int)blockIdx.x - 2, n, lds, red);
}

// RHS launch n: map[b] = (instance, row j, chunk, -); blocks of one (j, chunk) for all
// instances sharing a grid are adjacent, so their stencil weights are fetched once per XCD.
__device__ __forceinline__ RhsArgs step_rhs_args(const Ctl& c, const InstDesc& d, int n, int logr) {
    RhsArgs a{};
    a.T1 = (n == 0) ? c.T1s : c.T1s + 1;
    a.tl = c.gt[n];
    a.split = d.split;
    a.log_response = logr;
    a.fields = nullptr;
    a.fbuf = (c.fc0 + searches_before(c, n + 1)) & 1;
    a.vlast = n == 0 ? d.W.tf_pre : nullptr;  // tail fit on: the pre-blend last row
    return a;
}

template <bool LOGR, int G, int NT>
__global__ __launch_bounds__(kRowThreads) void batch_rhs_kernel(const InstDesc* __restrict__ D,
                                                                const int4* __restrict__ map, int n) {
    extern __shared__ double2 dyn[];
    __shared__ double red[kSums * (kRowThreads / kWave)];
    const int4 mp = map[xcd_stripe<16>(blockIdx.x, gridDim.x)];
    if (mp.x < 0) return;
    const InstDesc& d = D[mp.x];
    const Ctl& c = *d.ctl;
    const int active = c.active, S = c.S;
    const RhsArgs a = step_rhs_args(c, d, n, LOGR);  // loaded with the control word
    if (!active || n >= S) return;
    rhs_rows_body<LOGR, G, 0, NT>(d.T, d.H, d.W, d.M, a, mp.y, mp.z, dyn, red);
}

// device loop of one context: the instance by value, block -> (row, chunk) as in rhs_rows_kernel
template <bool LOGR, int G, int NT>
__global__ __launch_bounds__(kRowThreads) void loop_rhs_kernel(InstDesc d, int n) {
    extern __shared__ double2 dyn[];
    __shared__ double red[kSums * (kRowThreads / kWave)];
    const Ctl& c = *d.ctl;
    const int active = c.active, S = c.S;
    const RhsArgs a = step_rhs_args(c, d, n, LOGR);  // loaded with the control word
    if (!active || n >= S) return;
    const int lb = xcd_stripe<32>(blockIdx.x, gridDim.x);
    rhs_rows_body<LOGR, G, 0, NT>(d.T, d.H, d.W, d.M, a, lb / d.split, lb % d.split, dyn, red);
}

// the device loop's RHS with the tangent quadratures of the instance's K directions
template <int G, int K>
__global__ __launch_bounds__(kRowThreads) void loop_rhs_tan_kernel(InstDesc d, int n) {
    extern __shared__ double2 dyn[];
    __shared__ double red[kSums * (kRowThreads / kWave)];
    const Ctl& c = *d.ctl;
    const int active = c.active, S = c.S;
    const RhsArgs a = step_rhs_args(c, d, n, 0);  // loaded with the control word
    if (!active || n >= S) return;
    const int lb = xcd_stripe<32>(blockIdx.x, gridDim.x);
    rhs_rows_body<false, G, K>(d.T, d.H, d.W, d.M, a, lb / d.split, lb % d.split, dyn, red, &d.X);
}

// Observables of the instances that just stepped: (t_prop, E(t_prop), QK(t,0)) per row
// (simulation_runner.cpp:308-337), appended at row (loop - 1) % cap of obs[inst].
// the observables of instance `inst` of the batch (one block)
__device__ __forceinline__ void obs_body(const InstDesc& d, int inst, double* obs, int cap, double* red) {
    const Ctl& c = *d.ctl;
    if (!c.active) return;  // did not step in this batch step
    const int row = c.T1 - 1;
    const double t = c.t_prop;
    const double E = row_scalar_body(d.T, d.H, d.M, row, t, 1, red);
    if (threadIdx.x == 0) {
        double* o = obs + ((size_t)inst * cap + (size_t)((c.loop - 1) % cap)) * 4;
        o[0] = t;
        o[1] = E;
        o[2] = d.H.V[(size_t)row * d.T.N].x;
        o[3] = (double)c.loop;
    }
}

__global__ __launch_bounds__(kStageThreads) void batch_obs_kernel(const InstDesc* __restrict__ D, double* obs, int cap) {
    __shared__ double red[2 * (kStageThreads / kWave)];
    obs_body(D[blockIdx.x], blockIdx.x, obs, cap, red);
}

// ---- dense (B-spline) interpolation as an FP64 MFMA contraction --------------------
// Y = W . X with W [P][K] (one dense stencil row per target, entry-major as on disk),
// X [K][B] (B <= 16 right-hand sides, e.g. the last row's QK and QR) and Y [P][ldy]
// (index_vec.cpp:15-61 with m = N).  A block of 4 waves owns 64 target rows; K is
// streamed through LDS in chunks of 64 columns with 16-byte coalesced loads (every
// row chunk is 512 contiguous bytes), and each wave issues v_mfma_f64_16x16x4f64 on
// its 16 rows: A[i][k] sits in lane i + 16k, B[k][j] in lane j + 16k, and D[i][j]
// comes back in lane j + 16 (i % 4), register i / 4.  Columns j >= B are zero.
constexpr int kDenseRows = 64;
constexpr int kDenseKC = 32;
typedef double v4d __attribute__((ext_vector_type(4)));

__device__ inline void dense_load_chunk(const double* __restrict__ W, long long P, int K, long long r0, int k0,
                                        int tid, double2 (&v)[kDenseRows * kDenseKC / 512]) {
    constexpr int kPairs = kDenseKC / 2;  // column pairs per row chunk
    const bool full = k0 + kDenseKC <= K && r0 + kDenseRows <= P && (K & 1) == 0;
#pragma unroll
    for (int q = 0; q < kDenseRows * kDenseKC / 512; ++q) {
        const int e = tid + 256 * q;  // pair index: row e / kPairs, columns 2 (e % kPairs) + {0, 1}
        const int row = e / kPairs, c2 = (e % kPairs) * 2;
        const long long gr = r0 + row;
        if (full) {
            v[q] = *reinterpret_cast<const double2*>(W + gr * K + k0 + c2);
        } else {
            v[q].x = (gr < P && k0 + c2 < K) ? W[gr * K + k0 + c2] : 0.0;
            v[q].y = (gr < P && k0 + c2 + 1 < K) ? W[gr * K + k0 + c2 + 1] : 0.0;
        }
    }
}

__global__ __launch_bounds__(256) void dense_apply_kernel(const double* __restrict__ W, const double* __restrict__ X,
                                                          double* __restrict__ Y, long long P, int K, int B,
                                                          int ldy) {
    __shared__ double ws[kDenseRows][kDenseKC + 1];  // odd row stride: column reads spread over banks
    __shared__ double xs[kDenseKC][16];
    constexpr int kPairs = kDenseKC / 2;
    constexpr int NQ = kDenseRows * kDenseKC / 512;  // 16-byte loads per thread per chunk
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const long long r0 = (long long)blockIdx.x * kDenseRows;
    v4d acc = {0.0, 0.0, 0.0, 0.0};
    double2 pre[NQ];
    dense_load_chunk(W, P, K, r0, 0, tid, pre);
    for (int k0 = 0; k0 < K; k0 += kDenseKC) {
#pragma unroll
        for (int q = 0; q < NQ; ++q) {
            const int e = tid + 256 * q;
            ws[e / kPairs][(e % kPairs) * 2] = pre[q].x;
            ws[e / kPairs][(e % kPairs) * 2 + 1] = pre[q].y;
        }
        for (int e = tid; e < kDenseKC * 16; e += 256) {
            const int kk = e >> 4, j = e & 15;
            xs[kk][j] = (k0 + kk < K && j < B) ? X[(long long)(k0 + kk) * B + j] : 0.0;
        }
        __syncthreads();
        // the next chunk's loads are in flight while this chunk is contracted
        if (k0 + kDenseKC < K) dense_load_chunk(W, P, K, r0, k0 + kDenseKC, tid, pre);
#pragma unroll
        for (int ks = 0; ks < kDenseKC; ks += 4) {
            const double a = ws[wave * 16 + (lane & 15)][ks + (lane >> 4)];
            const double b = xs[ks + (lane >> 4)][lane & 15];
            acc = __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, acc, 0, 0, 0);
        }
        __syncthreads();
    }
    const int j = lane & 15;
#pragma unroll
    for (int r = 0; r < 4; ++r) {
        const long long gr = r0 + wave * 16 + (lane >> 4) + 4 * r;
        if (gr < P && j < B) Y[gr * ldy + j] = acc[r];
    }
}

__global__ void tan_init_kernel(Model M, TanArgs X, int N) {
    const int k = blockIdx.x;
    for (int j = threadIdx.x; j < N; j += blockDim.x) {
        X.V[k][j] = make_double2(0.0, 0.0);
        X.D[k][j] = make_double2(0.0, 0.0);
    }
    if (threadIdx.x == 0) {
        // r0 = Gamma + f'(1)/T0 and dr0 = rstep() at t = 0 (initialization.cpp:219-222)
        const double d = X.dG[k] + X.dL[k] * fl1l(M, 1.0) / M.T0;
        X.r[k][0] = d;
        X.dr[k][0] = d;
    }
}

// Tangent of finish_block for stage a.stage: tangent RHS (hK, hR), tangent drstep2 at
// stage 0, tangent stage combination and row write, tangent rstep on the new row.  It
// runs before the primal finish of the same stage (which overwrites the RHS's last row
// at stages >= 1) and recomputes the primal quantities it needs.
template <int EPT, int K>
__device__ __forceinline__ void tan_finish_body(const Tables& T, const History& H, const Work& W, const Model& M,
                                                const FinishArgs& a, const TanArgs& X,
                                                double* red /* 6 K kStageThreads / 64 */) {
    const int N = T.N, n = a.stage;
    const double t = a.t_rhs;
    const double2* last = H.V + (size_t)a.lr * N;
    const bool live = t > 0;
    const int* Fl = W.Floor + (size_t)a.fbuf * N;
    const double* Df = W.diff + (size_t)a.fbuf * N;
    const double qK0 = last[0].x;
    // tangent drstep2 of this step (stage 0 wrote it), read before this block's stores
    double tdr0[K];
#pragma unroll
    for (int h = 0; h < K; ++h) tdr0[h] = (n != 0) ? X.scal[h * 8 + 0] : 0.0;
    const double s1 = fl1(M, qK0) / M.T0, s2 = -rint_at(H, Fl, Df, N - 1, a.lr + 1, live), s3 = qK0 / M.T0,
                 s4 = 2 * M.Gamma;
    double ds1[K], ds2[K], ds3[K], ds4[K];
#pragma unroll
    for (int h = 0; h < K; ++h) {
        const History Ht = tan_history(H, X, h);
        const double tq0 = X.V[h][(size_t)a.lr * N].x;
        ds1[h] = (fl2(M, qK0) * tq0 + X.dL[h] * fl1l(M, qK0)) / M.T0;
        ds2[h] = -rint_at(Ht, Fl, Df, N - 1, a.lr + 1, live);
        ds3[h] = tq0 / M.T0;
        ds4[h] = 2 * X.dG[h];
    }
    double hk[EPT], hr[EPT], thk[EPT][K], thr[EPT][K];
#pragma unroll
    for (int e = 0; e < EPT; ++e) {
        const int j = threadIdx.x + e * kStageThreads;
        hk[e] = hr[e] = 0.0;
#pragma unroll
        for (int h = 0; h < K; ++h) thk[e][h] = thr[e][h] = 0.0;
        if (j >= N) continue;
        double C[kSums];
#pragma unroll
        for (int q = 0; q < kSums; ++q) C[q] = 0.0;
        for (int c = 0; c < a.split; ++c)
#pragma unroll
            for (int q = 0; q < kSums; ++q) C[q] += W.partial[((size_t)j * a.split + c) * kSums + q];
        const double th = T.theta[j];
        const double fA = t * th, fR = t * (1 - th);
        const double ri = rint_at(H, Fl, Df, j, a.lr + 1, live);
        const double qK = last[j].x, qR = last[j].y, b1 = W.qkb1c0[j];
        double d1 = b1 * s1 + qK * s2;
        d1 = d1 + C[0] * fR;
        d1 = d1 + C[1] * fA;
        d1 = d1 + C[2] * fA;
        double d2 = fl1(M, b1) * s3 + qR * s4;
        d2 = d2 + C[3] * fR;
        d2 = d2 + C[4] * fA;
        d2 = d2 + C[5] * fA;
        d2 = d2 - qK * ri;
        hk[e] = d1 + d2 * th;
        hr[e] = (qR * s2 + C[6] * fR) + (qR * ri - C[7] * fR) * th;
#pragma unroll
        for (int h = 0; h < K; ++h) {
            double TC[kSums];
#pragma unroll
            for (int q = 0; q < kSums; ++q) TC[q] = 0.0;
            for (int c = 0; c < a.split; ++c)
#pragma unroll
                for (int q = 0; q < kSums; ++q) TC[q] += X.partial[(((size_t)h * N + j) * a.split + c) * kSums + q];
            const History Ht = tan_history(H, X, h);
            const double tri = rint_at(Ht, Fl, Df, j, a.lr + 1, live);
            const double2 tl = X.V[h][(size_t)a.lr * N + j];
            const double tb1 = X.qkb1c0[(size_t)h * N + j];
            const double td1 = tb1 * s1 + b1 * ds1[h] + tl.x * s2 + qK * ds2[h] + TC[0] * fR + TC[1] * fA + TC[2] * fA;
            const double td2 = (fl2(M, b1) * tb1 + X.dL[h] * fl1l(M, b1)) * s3 + fl1(M, b1) * ds3[h] + tl.y * s4 +
                               qR * ds4[h] + TC[3] * fR + TC[4] * fA + TC[5] * fA - tl.x * ri - qK * tri;
            thk[e][h] = td1 + td2 * th;
            thr[e][h] = (tl.y * s2 + qR * ds2[h] + TC[6] * fR) + (tl.y * ri + qR * tri - TC[7] * fR) * th;
            X.hK[((size_t)h * kMaxStages + n) * N + j] = thk[e][h];
            X.hR[((size_t)h * kMaxStages + n) * N + j] = thr[e][h];
        }
    }
    // tangent drstep2 at stage 0 (time_steps.cpp:117-125)
    double tdr[K];
    if (n == 0) {
        double v[6 * K];
#pragma unroll
        for (int q = 0; q < 6 * K; ++q) v[q] = 0.0;
        double c0t[K];
#pragma unroll
        for (int h = 0; h < K; ++h) c0t[h] = 0.0;
#pragma unroll
        for (int e = 0; e < EPT; ++e) {
            const int j = threadIdx.x + e * kStageThreads;
            if (j >= N) continue;
            const double qK = last[j].x, qR = last[j].y, w = T.integ[j];
            const double f1 = fl1(M, qK), f2 = fl2(M, qK), f3 = fl3(M, qK);
            const double sK = f1, sR = f2 * qR;
            const double dsK = f2 * hk[e], dsR = (f3 * qR) * hk[e] + f2 * hr[e];
#pragma unroll
            for (int h = 0; h < K; ++h) {
                const double l = X.dL[h];
                const double2 tl = X.V[h][(size_t)a.lr * N + j];
                const double tf2 = f3 * tl.x + l * fl2l(M, qK);
                const double tf3 = fl4(M, qK) * tl.x + l * fl3l(M, qK);
                const double tsK = f2 * tl.x + l * fl1l(M, qK);
                const double tsR = tf2 * qR + f2 * tl.y;
                const double tdsK = tf2 * hk[e] + f2 * thk[e][h];
                const double tdsR = (tf3 * qR + f3 * tl.y) * hk[e] + (f3 * qR) * thk[e][h] + tf2 * hr[e] + f2 * thr[e][h];
                double* vv = v + 6 * h;
                vv[0] += w * (tsR * qK + sR * tl.x);
                vv[1] += w * (tsK * qR + sK * tl.y);
                vv[2] += t * w * (tdsR * qK + dsR * tl.x);
                vv[3] += t * w * (tdsK * qR + dsK * tl.y);
                vv[4] += t * w * (tsR * hk[e] + sR * thk[e][h]);
                vv[5] += t * w * (tsK * hr[e] + sK * thr[e][h]);
                if (j == 0) c0t[h] = (tdsK * qK + dsK * tl.x + tsK * hk[e] + sK * thk[e][h]) / M.T0;
            }
        }
        block_sum<6 * K, kStageThreads>(v, red);
        // c0t is valid in thread 0 (entry j = 0), which writes the result
#pragma unroll
        for (int h = 0; h < K; ++h) {
            tdr[h] = v[6 * h] + v[6 * h + 1] + v[6 * h + 2] + v[6 * h + 3] + v[6 * h + 4] + v[6 * h + 5] + c0t[h];
            if (threadIdx.x == 0) X.scal[h * 8 + 0] = tdr[h];
        }
    } else {
#pragma unroll
        for (int h = 0; h < K; ++h) tdr[h] = tdr0[h];
    }
    // tangent stage combination, row write and rstep (time_steps.cpp:57-92,140-157)
    const double2* g0 = H.V + (size_t)a.g0row * N;
    double rr[2 * K];
#pragma unroll
    for (int q = 0; q < 2 * K; ++q) rr[q] = 0.0;
    double c0r[K];
#pragma unroll
    for (int h = 0; h < K; ++h) c0r[h] = 0.0;
#pragma unroll
    for (int e = 0; e < EPT; ++e) {
        const int j = threadIdx.x + e * kStageThreads;
        if (j >= N) continue;
        double gK = g0[j].x, gR = g0[j].y;
        for (int q = 0; q < a.nterms; ++q) {
            gK = gK + ((q == n) ? hk[e] : W.hK[(size_t)q * N + j]) * a.coef[q];
            gR = gR + ((q == n) ? hr[e] : W.hR[(size_t)q * N + j]) * a.coef[q];
        }
        const double w = T.integ[j];
        const double f1 = fl1(M, gK), f2 = fl2(M, gK), f3 = fl3(M, gK);
        const double sK = f1, sR = f2 * gR;
#pragma unroll
        for (int h = 0; h < K; ++h) {
            const double2 tg0 = X.V[h][(size_t)a.g0row * N + j];
            double tgK = tg0.x, tgR = tg0.y;
            for (int q = 0; q < a.nterms; ++q) {
                tgK = tgK + ((q == n) ? thk[e][h] : X.hK[((size_t)h * kMaxStages + q) * N + j]) * a.coef[q];
                tgR = tgR + ((q == n) ? thr[e][h] : X.hR[((size_t)h * kMaxStages + q) * N + j]) * a.coef[q];
            }
            const double th0k = (n == 0) ? thk[e][h] : X.hK[((size_t)h * kMaxStages) * N + j];
            const double th0r = (n == 0) ? thr[e][h] : X.hR[((size_t)h * kMaxStages) * N + j];
            X.V[h][(size_t)a.wrow * N + j] = make_double2(tgK, tgR);
            X.D[h][(size_t)a.wrow * N + j] = make_double2(a.tdiff * th0k, a.tdiff * th0r);
            const double l = X.dL[h];
            const double tsK = f2 * tgK + l * fl1l(M, gK);
            const double tsR = (f3 * tgK + l * fl2l(M, gK)) * gR + f2 * tgR;
            rr[2 * h] += a.t_new * w * (tsR * gK + sR * tgK);
            rr[2 * h + 1] += a.t_new * w * (tsK * gR + sK * tgR);
            if (j == 0) c0r[h] = (tsK * gK + sK * tgK) / M.T0;
        }
    }
    block_sum<2 * K, kStageThreads>(rr, red);
    if (threadIdx.x == 0) {
#pragma unroll
        for (int h = 0; h < K; ++h) {
            X.r[h][a.wrow] = X.dG[h] + rr[2 * h] + rr[2 * h + 1] + c0r[h];
            X.dr[h][a.wrow] = a.tdiff * tdr[h];
        }
    }
}

template <int EPT, int K>
__global__ __launch_bounds__(kStageThreads) void tan_finish_kernel(Tables T, History H, Work W, Model M, FinishArgs a,
                                                                   TanArgs X) {
    __shared__ double red[6 * K * (kStageThreads / kWave)];
    tan_finish_body<EPT, K>(T, H, W, M, a, X, red);
}

// Stage launch n of the device loop of a context with tangents: block 0 runs the tangent
// finish of stage n-1 and then, after a barrier, the primal finish (which overwrites the
// RHS's last row that the tangent finish reads), as the host loop's two launches do; the
// other blocks are the searches of stage n.
template <int EPT, int K>
__global__ __launch_bounds__(kStageThreads) void loop_stage_tan_kernel(InstDesc d, int n) {
    __shared__ double lds[EPT * kStageThreads];
    __shared__ double red[(6 * K > 8 ? 6 * K : 8) * (kStageThreads / kWave)];
    const int role = (int)blockIdx.x - 2;
    if (role == -2) {
        if (n < 1) return;
        const Ctl& c = *d.ctl;
        const int active = c.active, S = c.S;
        const FinishArgs fa = c.fin[n - 1];
        if (!active || n > S) return;
        tan_finish_body<EPT, K>(d.T, d.H, d.W, d.M, c.fin[n - 1], d.X, red);  // (indexes its coefficients)
        __syncthreads();
        finish_block<EPT, true>(d.T, d.H, d.W, d.M, fa, red);
        return;
    }
    stage_role_body<EPT, true>(d, role, n, lds, red);
}

// Stage launch n of a batch whose instances carry K tangent directions: the finish role
// runs the tangent finish of stage n-1 before the primal finish (the order of
// loop_stage_tan_kernel and of the host loop), the search roles are those of
// batch_stage_kernel.
template <int EPT, int K>
__global__ __launch_bounds__(kStageThreads) void batch_stage_tan_kernel(const InstDesc* __restrict__ D,
                                                                        const int2* __restrict__ map, int n) {
    __shared__ double lds[EPT * kStageThreads];
    __shared__ double red[(6 * K > 8 ? 6 * K : 8) * (kStageThreads / kWave)];
    const int2 mp = map[blockIdx.x];
    if (mp.x < 0) return;
    const InstDesc& d = D[mp.x];
    if (mp.y == -2) {
        if (n < 1) return;
        const Ctl& c = *d.ctl;
        const int active = c.active, S = c.S;
        if (!active || n > S) return;
        tan_finish_body<EPT, K>(d.T, d.H, d.W, d.M, c.fin[n - 1], d.X, red);
        __syncthreads();
        finish_block<EPT, false>(d.T, d.H, d.W, d.M, c.fin[n - 1], red);
        return;
    }
    stage_role_body<EPT, false>(d, mp.y, n, lds, red);
}

// RHS launch n of a batch whose instances carry K tangent directions (linear-response
// form): the primal quadratures and the tangent quadratures of every direction in one pass,
// block map as in batch_rhs_kernel.
template <int G, int K>
__global__ __launch_bounds__(kRowThreads) void batch_rhs_tan_kernel(const InstDesc* __restrict__ D,
                                                                    const int4* __restrict__ map, int n) {
    extern __shared__ double2 dyn[];
    __shared__ double red[kSums * (kRowThreads / kWave)];
    const int4 mp = map[xcd_stripe<16>(blockIdx.x, gridDim.x)];
    if (mp.x < 0) return;
    const InstDesc& d = D[mp.x];
    const Ctl& c = *d.ctl;
    const int active = c.active, S = c.S;
    const RhsArgs a = step_rhs_args(c, d, n, 0);  // loaded with the control word
    if (!active || n >= S) return;
    rhs_rows_body<false, G, K>(d.T, d.H, d.W, d.M, a, mp.y, mp.z, dyn, red, &d.X);
}

// Tangent of the energy at time t on history row `row` (simulation_runner.cpp:308-318)
template <int K>
__global__ __launch_bounds__(kStageThreads) void tan_energy_kernel(Tables T, History H, Model M, TanArgs X, int row,
                                                                   double t) {
    __shared__ double red[K * (kStageThreads / kWave)];
    const int N = T.N;
    const double2* v = H.V + (size_t)row * N;
    double s[K];
#pragma unroll
    for (int h = 0; h < K; ++h) s[h] = 0.0;
    for (int j = threadIdx.x; j < N; j += kStageThreads) {
        const double qK = v[j].x, qR = v[j].y, w = T.integ[j];
#pragma unroll
        for (int h = 0; h < K; ++h) {
            const double2 tv = X.V[h][(size_t)row * N + j];
            s[h] += t * w * ((fl2(M, qK) * tv.x + X.dL[h] * fl1l(M, qK)) * qR + fl1(M, qK) * tv.y);
        }
    }
    block_sum<K, kStageThreads>(s, red);
    if (threadIdx.x == 0) {
        const double qK0 = v[0].x;
#pragma unroll
        for (int h = 0; h < K; ++h) {
            const double tq0 = X.V[h][(size_t)row * N].x;
            X.scal[h * 8 + 2] = -(s[h] + (fl2(M, qK0) * tq0 + X.dL[h] * fl1l(M, qK0)) / M.T0);
        }
    }
}

// Plain streaming copy, used to calibrate HBM byte counters (FETCH_SIZE / WRITE_SIZE).
__global__ __launch_bounds__(256) void hbm_copy_kernel(const double2* __restrict__ src, double2* __restrict__ dst,
                                                       long long n) {
    for (long long i = (long long)blockIdx.x * 256 + threadIdx.x; i < n; i += (long long)gridDim.x * 256)
        dst[i] = src[i];
}

}  // namespace

// (G, NT) shapes of the RHS row kernels: G lanes per B2 entry, NT taps per lane per
// gather pass; the log-response form closes its sums per slice and ignores NT.
#define DMFE_SHAPES(LOGR, G, NT, LAUNCH)                 \
    if (LOGR) {                                          \
        switch (G) {                                     \
            case 4: LAUNCH(true, 4, 1); break;           \
            case 8: LAUNCH(true, 8, 1); break;           \
            case 16: LAUNCH(true, 16, 1); break;         \
            default: LAUNCH(true, 64, 1); break;         \
        }                                                \
    } else {                                             \
        switch (G) {                                     \
            case 4:                                      \
                if (NT <= 3) LAUNCH(false, 4, 3);        \
                else LAUNCH(false, 4, 4);                \
                break;                                   \
            case 8: LAUNCH(false, 8, 2); break;          \
            case 16: LAUNCH(false, 16, 2); break;        \
            default: LAUNCH(false, 64, 1); break;        \
        }                                                \
    }

}  // namespace dmfe
