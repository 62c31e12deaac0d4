/* CPU oracle: plain-C restatement of the reference CPU path for one RK step of
 * the renormalized DMFT equations (spherical mixed p-spin).
 *
 * TEST INFRASTRUCTURE ONLY: called by tests/, __graft_entry__.smoke() and the
 * cpu_baseline leg of bench.py as the checker; never by the product.
 *
 * Single-threaded.  Floating-point expressions keep the reference's evaluation
 * order (operator precedence and left-to-right vector additions), so that with
 * -fno-fast-math -ffp-contract=off results are bit-identical to the reference
 * compiled with the same flags and OMP_NUM_THREADS=1.
 *
 * Deliberate deviation: a negative stencil start (dense B-spline tables store
 * -1) is read as 0, i.e. the intended full-row contraction.  The reference
 * parses -1 into size_t and reads one element to the left
 * (io_input.cpp:75-88 + index_vec.cpp:153, index_mat.cpp:43).
 */
#include "dmfe_oracle.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>

struct or_ctx {
    size_t N, m;
    int p, p2, log_resp;
    double lambda, T0, Gamma, dt_min, delta_max, specRad;
    double rmax[11];
    double *theta, *integ, *phi2;
    size_t *iA1, *iA2, *iB2;
    double *wA1, *wA2, *wB2;
    /* history (simulation_data.hpp:96-98) */
    size_t T1, cap;
    double *t1, *rho, *rv, *drv;
    double *QK, *QR, *dQK, *dQR;
    /* interpolation state (simulation_data.hpp:93,100-101) */
    double *posB1old, *posB2old;
    double *QKA1, *QRA1, *QKA2, *QRA2, *QKB1, *QRB1, *QKB2, *QRB2;
    double *SKA1, *SRA1, *SRA2, *SKB1, *SRB1, *SKB2, *SRB2;
    double *rInt;
    /* control (config.hpp:34-38, rk_data.hpp:12) */
    double delta_t_val, delta, delta_old;
    long loop;
    int rk_init;
    int use_serk2; /* the GPU loop's SERK2 promotion at the maintenance (config.use_serk2) */
    long maint_every; /* the maintenance cadence, 100000 in the reference (0: 100000) */
    int rollback;     /* the GPU loop's rollback (simulation_runner.cpp:280-303), off = CPU path */
    long last_rollback; /* last_rollback_loop (simulation_runner.cpp:53), -1000 until a rollback */
    /* QK tail fit near theta -> 1 (tail_fit.cpp:46-265, --tail-fit): on/off and the state the
     * reference keeps in function-local statics (prev_best_off, prev_t1len, alpha_smoothed,
     * last_known_rollback), here per context */
    /* the reference GPU path's interpolation (interpolation_core.cu:23-42,65-72): every
     * interpolation by bisection (its input positions and `same` ignored) and the B1 time
     * row clamped per element to [1, ceil(p) - 1] instead of [1, ceil(max p) - 1] */
    int gpu_variants;
    int tail_fit;
    long tf_prev_off; /* -1: SIZE_MAX sentinel, no window chosen yet */
    size_t tf_prev_t1len;
    double tf_alpha;
    long tf_known_rollback;
    long tf_applied; /* blends applied (not in the reference: a test observable) */
};

static double* dalloc(size_t n) { return (double*)calloc(n ? n : 1, sizeof(double)); }
static void tail_fit_blend(or_ctx* c);

/* ---- math_ops.cpp:7-37 ------------------------------------------------------ */
static double pow_int(double base, int e) {
    if (e == 0) return 1.0;
    if (e == 1) return base;
    if (e == 2) return base * base;
    if (e == 3) return base * base * base;
    if (e == 4) { double sq = base * base; return sq * sq; }
    double result = 1.0, cur = base;
    while (e > 0) {
        if (e & 1) result *= cur;
        cur *= cur;
        e >>= 1;
    }
    return result;
}
static double fl0(const or_ctx* c, double q) {
    return c->lambda * pow_int(q, c->p) + (1 - c->lambda) * pow_int(q, c->p2);
}
static double fl1(const or_ctx* c, double q) {
    return c->lambda * c->p * pow_int(q, c->p - 1) + (1 - c->lambda) * c->p2 * pow_int(q, c->p2 - 1);
}
static double fl2(const or_ctx* c, double q) {
    return c->lambda * c->p * (c->p - 1) * pow_int(q, c->p - 2) +
           (1 - c->lambda) * c->p2 * (c->p2 - 1) * pow_int(q, c->p2 - 2);
}
static double fl3(const or_ctx* c, double q) {
    return c->lambda * c->p * (c->p - 1) * (c->p - 2) * pow_int(q, c->p - 3) +
           (1 - c->lambda) * c->p2 * (c->p2 - 1) * (c->p2 - 2) * pow_int(q, c->p2 - 3);
}
double or_flambda(or_ctx* c, double q, int d) {
    return d == 0 ? fl0(c, q) : d == 1 ? fl1(c, q) : d == 2 ? fl2(c, q) : fl3(c, q);
}

/* ---- math_sigma.cpp:7-62 ---------------------------------------------------- */
static void sigmaK(const or_ctx* c, const double* qk, double* out, size_t n) {
    for (size_t i = 0; i < n; ++i) out[i] = fl1(c, qk[i]);
}
static void sigmaR(const or_ctx* c, const double* qk, const double* qr, double* out, size_t n) {
    for (size_t i = 0; i < n; ++i) out[i] = fl2(c, qk[i]) * qr[i];
}

/* ---- convolution.cpp:5-50 --------------------------------------------------- */
/* depth > 1: out_j = (sum_i w_i f_ji g_ji) * (t theta_j) */
static void convA_rows(const or_ctx* c, const double* f, const double* g, double t, double* out) {
    const size_t N = c->N;
    for (size_t j = 0; j < N; ++j) {
        double acc = 0.0;
        for (size_t i = 0; i < N; ++i) acc += c->integ[i] * f[j * N + i] * g[j * N + i];
        out[j] = acc * (t * c->theta[j]);
    }
}
static void convR_rows(const or_ctx* c, const double* f, const double* g, double t, double* out) {
    const size_t N = c->N;
    for (size_t j = 0; j < N; ++j) {
        double acc = 0.0;
        for (size_t i = 0; i < N; ++i) acc += c->integ[i] * f[j * N + i] * g[j * N + i];
        out[j] = acc * (t * (1 - c->theta[j]));
    }
}
/* depth == 1: t * sum_j w_j f_j g_j, serial order (the OpenMP reduction with one thread) */
static double convA_scalar(const or_ctx* c, const double* f, const double* g, double t) {
    double temp = 0.0;
    for (size_t j = 0; j < c->N; ++j) temp += t * c->integ[j] * f[j] * g[j];
    return temp;
}

/* ---- search_utils.cpp:13-61 ------------------------------------------------- */
static double bsearch_one(const double* list, size_t len, double El) {
    size_t n0 = 1, n1 = len, m = 1;
    double Lm = 0.0;
    if (list[m - 1] < El) {
        while (n0 <= n1) {
            m = (n0 + n1) / 2;
            Lm = list[m - 1];
            if (Lm == El) { n0 = m; break; }
            if (Lm < El) n0 = m + 1;
            else n1 = m - 1;
        }
    }
    if (Lm <= El) {
        if (m == len) return (double)m;
        return m + (El - Lm) / (list[m] - Lm);
    }
    if (m > 1) return m - (El - Lm) / (list[m - 2] - Lm);
    return (double)m;
}
void or_bsearch(int64_t nlist, const double* list, int64_t nel, const double* el, double* out) {
    for (int64_t l = 0; l < nel; ++l) out[l] = bsearch_one(list, (size_t)nlist, el[l]);
}

/* ---- search_utils.cpp:63-131: rows of rowlen, each walked from the end -------- */
void or_isearch(int64_t nlist, const double* list, int64_t nel, int64_t rowlen, const double* el,
                const double* inits, double* out) {
    const size_t length = (size_t)nlist, len = (size_t)rowlen;
    for (size_t k = 0; k < (size_t)nel / len; ++k) {
        double l0, l1 = 0.0, Lm, El;
        size_t n0, n1, m;
        int even;
        double temp = (double)length;
        for (size_t l = len; l-- > 0;) {
            El = list[length - 1] * el[k * len + l];
            n1 = (size_t)ceil(temp);
            if (n1 > length) n1 = length;
            n0 = (size_t)floor(inits[k * len + l]);
            m = n0 + 1 < length ? n0 + 1 : length;
            even = 1;
            if ((Lm = list[m - 1]) > El) {
                size_t a = m - 2;
                n1 = a > 1 ? a : 1;
            }
            while ((l0 = list[n0 - 1]) <= El && El <= (l1 = list[n1 - 1]) && n0 < n1) {
                even = !even;
                if (even) m = (size_t)(n0 + round((El - l0) / (l1 - l0) * (n1 - n0)));
                else m = (n0 + n1) / 2;
                Lm = list[m - 1];
                if (Lm == El) { n0 = m; n1 = m - 1; }
                else if (Lm < El) n0 = m + 1;
                else n1 = m - 1;
            }
            if (Lm <= El) {
                if (m == length) temp = (double)m;
                else temp = m + (El - Lm) / (list[m] - Lm);
            } else {
                if (m > 1) temp = m - (El - Lm) / (list[m - 2] - Lm);
                else temp = (double)m;
            }
            out[k * len + l] = temp;
        }
    }
}

/* ---- index_vec.cpp:15-61 (A1 / A2 stencils on the last history row) --------- */
static void index_ln3(or_ctx* c, const double* w, const size_t* inds, double* qk, double* qr) {
    const size_t N = c->N, P = N * N, depth = c->m;
    const double* QKs = c->QK + (c->T1 - 1) * N;
    const double* QRs = c->QR + (c->T1 - 1) * N;
    double* logqr = NULL;
    if (c->log_resp) {
        logqr = dalloc(N);
        for (size_t i = 0; i < N; ++i) { double v = QRs[i]; logqr[i] = (v > 0.0) ? log(v) : v; }
    }
    for (size_t j = 0; j < P; ++j) {
        const double* ws = w + depth * j;
        const size_t s = inds[j];
        double a = 0.0;
        for (size_t d = 0; d < depth; ++d) a = a + ws[d] * QKs[s + d];
        qk[j] = a;
        if (!c->log_resp) {
            double b = 0.0;
            for (size_t d = 0; d < depth; ++d) b = b + ws[d] * QRs[s + d];
            qr[j] = b;
        } else {
            double lin = 0.0;
            long double lsum = 0.0L;
            int invalid = 0;
            for (size_t d = 0; d < depth; ++d) {
                const double v = QRs[s + d], wd = ws[d];
                lin += wd * v;
                if (v > 0.0) lsum += (long double)wd * logqr[s + d];
                else invalid = 1;
            }
            qr[j] = invalid ? lin : exp((double)lsum);
        }
    }
    free(logqr);
}

/* ---- index_vec.cpp:63-136 (B1: cubic Hermite in time, exact theta column) ---- */
static void index_n(or_ctx* c, const double* diff, const size_t* F) {
    const size_t N = c->N, t1len = c->T1;
    const double *QK = c->QK, *QR = c->QR, *dQK = c->dQK, *dQR = c->dQR, *rho = c->rho;
    for (size_t i = 0; i < N; ++i) {
        const double d = diff[i], in3 = d * d, in4 = in3 * d;
        const size_t f = F[i];
        if (f < t1len - 1) {
            for (size_t j = 0; j < N; ++j) {
                c->QKB1[j + N * i] = (1 - 3 * in3 - 2 * in4) * QK[(f - 1) * N + j] + (3 * in3 + 2 * in4) * QK[f * N + j] -
                                     (d + 2 * in3 + in4) * dQK[f * N + j] - (in3 + in4) * dQK[(f + 1) * N + j] / rho[f + 1];
                if (!c->log_resp) {
                    c->QRB1[j + N * i] = (1 - 3 * in3 - 2 * in4) * QR[(f - 1) * N + j] + (3 * in3 + 2 * in4) * QR[f * N + j] -
                                         (d + 2 * in3 + in4) * dQR[f * N + j] - (in3 + in4) * dQR[(f + 1) * N + j] / rho[f + 1];
                } else {
                    const double qb = QR[(f - 1) * N + j], qc = QR[f * N + j];
                    const double dl = dQR[f * N + j], dr = dQR[(f + 1) * N + j];
                    if (qb > 0.0 && qc > 0.0) {
                        const double fb = log(qb), fc = log(qc), gl = dl / qb, gr = dr / qc;
                        const double c1 = 1 - 3 * in3 - 2 * in4, c2 = 3 * in3 + 2 * in4;
                        const double c3 = (d + 2 * in3 + in4), c4 = (in3 + in4) / rho[f + 1];
                        c->QRB1[j + N * i] = exp(c1 * fb + c2 * fc - c3 * gl - c4 * gr);
                    } else {
                        c->QRB1[j + N * i] = (1 - 3 * in3 - 2 * in4) * qb + (3 * in3 + 2 * in4) * qc -
                                             (d + 2 * in3 + in4) * dl - (in3 + in4) * dr / rho[f + 1];
                    }
                }
            }
        } else {
            for (size_t j = 0; j < N; ++j) {
                c->QKB1[j + N * i] = (1 - in3) * QK[(f - 1) * N + j] - (d + in3) * dQK[f * N + j] + in3 * QK[f * N + j];
                if (!c->log_resp) {
                    c->QRB1[j + N * i] = (1 - in3) * QR[(f - 1) * N + j] - (d + in3) * dQR[f * N + j] + in3 * QR[f * N + j];
                } else {
                    const double qb = QR[(f - 1) * N + j], qc = QR[f * N + j], dl = dQR[f * N + j];
                    if (qb > 0.0 && qc > 0.0) {
                        const double fb = log(qb), fc = log(qc), gl = dl / qb;
                        c->QRB1[j + N * i] = exp((1 - in3) * fb + in3 * fc - (d + in3) * gl);
                    } else {
                        c->QRB1[j + N * i] = (1 - in3) * qb - (d + in3) * dl + in3 * qc;
                    }
                }
            }
        }
    }
}

/* ---- index_vec.cpp:138-156 (rInt) -------------------------------------------- */
static void index_r2(or_ctx* c, const double* diff, const size_t* F) {
    const size_t N = c->N, t1len = c->T1;
    const double *r = c->rv, *dr = c->drv, *rho = c->rho;
    for (size_t i = 0; i < N; ++i) {
        const double d = diff[i], d2 = d * d, d3 = d * (d * d);
        const size_t f = F[i];
        if (f < t1len - 1)
            c->rInt[i] = (1 - 3 * d2 - 2 * d3) * r[f - 1] + (3 * d2 + 2 * d3) * r[f] - (d + 2 * d2 + d3) * dr[f] -
                         (d2 + d3) * dr[f + 1] / rho[f + 1];
        else
            c->rInt[i] = (1 - d2) * r[f - 1] + d2 * r[f] - (d + d2) * dr[f];
    }
}

/* ---- index_mat.cpp:18-91 (B2: theta stencil on three history rows + Hermite) - */
static double iprod(const double* w, const double* x, size_t depth) {
    double a = 0.0;
    for (size_t d = 0; d < depth; ++d) a = a + w[d] * x[d];
    return a;
}
static void index_matall(or_ctx* c, const double* posx) {
    const size_t N = c->N, P = N * N, depth = c->m, t1len = c->T1;
    const size_t cap_hi = (size_t)(posx[P - 1] - 0.5);
    for (size_t j = 0; j < P; ++j) {
        size_t ix = (size_t)posx[j];
        if (cap_hi < ix) ix = cap_hi;
        if (ix < 1) ix = 1;
        const double u = posx[j] - ix, u2 = u * u, u3 = u2 * u;
        const size_t base = (ix - 1) * N + c->iB2[j];
        const double* ws = c->wB2 + depth * j;
        if (ix < t1len - 1) {
            c->QKB2[j] = (1 - 3 * u2 + 2 * u3) * iprod(ws, c->QK + base, depth) +
                         (u - 2 * u2 + u3) * iprod(ws, c->dQK + N + base, depth) +
                         (3 * u2 - 2 * u3) * iprod(ws, c->QK + N + base, depth) +
                         (-u2 + u3) * iprod(ws, c->dQK + 2 * N + base, depth) / c->rho[ix + 1];
            if (!c->log_resp) {
                c->QRB2[j] = (1 - 3 * u2 + 2 * u3) * iprod(ws, c->QR + base, depth) +
                             (u - 2 * u2 + u3) * iprod(ws, c->dQR + N + base, depth) +
                             (3 * u2 - 2 * u3) * iprod(ws, c->QR + N + base, depth) +
                             (-u2 + u3) * iprod(ws, c->dQR + 2 * N + base, depth) / c->rho[ix + 1];
            } else {
                const double qb = iprod(ws, c->QR + base, depth), qc = iprod(ws, c->QR + N + base, depth);
                const double dl = iprod(ws, c->dQR + N + base, depth), dr = iprod(ws, c->dQR + 2 * N + base, depth);
                if (qb > 0.0 && qc > 0.0) {
                    const double fb = log(qb), fc = log(qc), gl = dl / qb, gr = dr / qc, den = c->rho[ix + 1];
                    const double c1 = 1 - 3 * u2 + 2 * u3, c2 = u - 2 * u2 + u3, c3 = 3 * u2 - 2 * u3,
                                 c4 = (-u2 + u3) / den;
                    c->QRB2[j] = exp(c1 * fb + c2 * gl + c3 * fc + c4 * gr);
                } else {
                    c->QRB2[j] = (1 - 3 * u2 + 2 * u3) * qb + (u - 2 * u2 + u3) * dl + (3 * u2 - 2 * u3) * qc +
                                 (-u2 + u3) * dr / c->rho[ix + 1];
                }
            }
        } else {
            c->QKB2[j] = (1 - u2) * iprod(ws, c->QK + base, depth) + u2 * iprod(ws, c->QK + N + base, depth) +
                         (u - u2) * iprod(ws, c->dQK + N + base, depth);
            if (!c->log_resp) {
                c->QRB2[j] = (1 - u2) * iprod(ws, c->QR + base, depth) + u2 * iprod(ws, c->QR + N + base, depth) +
                             (u - u2) * iprod(ws, c->dQR + N + base, depth);
            } else {
                const double qb = iprod(ws, c->QR + base, depth), qc = iprod(ws, c->QR + N + base, depth);
                const double dl = iprod(ws, c->dQR + N + base, depth);
                if (qb > 0.0 && qc > 0.0) {
                    const double fb = log(qb), fc = log(qc), gl = dl / qb;
                    c->QRB2[j] = exp((1 - u2) * fb + u2 * fc + (u - u2) * gl);
                } else {
                    c->QRB2[j] = (1 - u2) * qb + u2 * qc + (u - u2) * dl;
                }
            }
        }
    }
}

static void fill(double* x, size_t n, double v) { for (size_t i = 0; i < n; ++i) x[i] = v; }

/* ---- interpolation_core.cpp:19-114 --------------------------------------------
 * mode 0: no input positions (binary search of theta*t, phi2*t)
 * mode 1: isearch seeded by the previous positions
 * mode 2: reuse the given positions ("same")                                      */
static void interpolate_with(or_ctx* c, int mode, const double* inB1, const double* inB2) {
    const size_t N = c->N, P = N * N, T1 = c->T1;
    if (c->gpu_variants) mode = 0;
    const double tl = c->t1[T1 - 1];
    double* posB1 = dalloc(N);
    double* posB2 = dalloc(P);
    if (mode == 0) {
        double* e = dalloc(P);
        for (size_t i = 0; i < N; ++i) e[i] = c->theta[i] * tl;
        or_bsearch((int64_t)T1, c->t1, (int64_t)N, e, posB1);
        for (size_t i = 0; i < P; ++i) e[i] = c->phi2[i] * tl;
        or_bsearch((int64_t)T1, c->t1, (int64_t)P, e, posB2);
        free(e);
    } else if (mode == 1) {
        or_isearch((int64_t)T1, c->t1, (int64_t)N, (int64_t)N, c->theta, inB1, posB1);
        or_isearch((int64_t)T1, c->t1, (int64_t)P, (int64_t)N, c->phi2, inB2, posB2);
    } else {
        memcpy(posB1, inB1, N * sizeof(double));
        memcpy(posB2, inB2, P * sizeof(double));
    }
    memcpy(c->posB1old, posB1, N * sizeof(double));
    memcpy(c->posB2old, posB2, P * sizeof(double));

    if (tl > 0) index_ln3(c, c->wA1, c->iA1, c->QKA1, c->QRA1);
    else { fill(c->QKA1, P, c->QK[0]); fill(c->QRA1, P, c->QR[0]); }
    sigmaK(c, c->QKA1, c->SKA1, P);
    sigmaR(c, c->QKA1, c->QRA1, c->SRA1, P);

    if (tl > 0) index_ln3(c, c->wA2, c->iA2, c->QKA2, c->QRA2);
    else { fill(c->QKA2, P, c->QK[0]); fill(c->QRA2, P, c->QR[0]); }
    sigmaR(c, c->QKA2, c->QRA2, c->SRA2, P);

    double maxp = posB1[0];
    for (size_t i = 1; i < N; ++i) if (posB1[i] > maxp) maxp = posB1[i];
    size_t maxCeil = (size_t)ceil(maxp) - 1;
    if (maxCeil < 1) maxCeil = 1;
    size_t* F = (size_t*)calloc(N, sizeof(size_t));
    double* diff = dalloc(N);
    for (size_t i = 0; i < N; ++i) {
        size_t f = (size_t)floor(posB1[i]);
        if (c->gpu_variants) {
            const size_t own = (size_t)ceil(posB1[i]);  /* diffNfloorKernel: own ceil per element */
            if (own - 1 < f) f = own - 1;
            if (f < 1) f = 1;
        } else if (f < 1) f = 1;
        else if (f > maxCeil) f = maxCeil;
        F[i] = f;
        diff[i] = (double)F[i] - posB1[i];
    }
    if (tl > 0) index_n(c, diff, F);
    else { fill(c->QKB1, P, c->QK[0]); fill(c->QRB1, P, c->QR[0]); }
    sigmaK(c, c->QKB1, c->SKB1, P);
    sigmaR(c, c->QKB1, c->QRB1, c->SRB1, P);

    if (tl > 0) index_matall(c, c->posB2old);
    else { fill(c->QKB2, P, c->QK[0]); fill(c->QRB2, P, c->QR[0]); }
    sigmaK(c, c->QKB2, c->SKB2, P);
    sigmaR(c, c->QKB2, c->QRB2, c->SRB2, P);

    if (tl > 0) index_r2(c, diff, F);
    else fill(c->rInt, N, c->rv[0]);

    free(F); free(diff); free(posB1); free(posB2);
}

void or_interpolate(or_ctx* c, int mode) {
    const size_t N = c->N;
    double* b1 = dalloc(N);
    double* b2 = dalloc(N * N);
    memcpy(b1, c->posB1old, N * sizeof(double));
    memcpy(b2, c->posB2old, N * N * sizeof(double));
    interpolate_with(c, mode, b1, b2);
    free(b1); free(b2);
}

void or_get_fields(const or_ctx* c, double* out) {
    const size_t P = c->N * c->N;
    const double* f[] = {c->QKA1, c->QRA1, c->QKA2, c->QRA2, c->QKB1, c->QRB1, c->QKB2, c->QRB2,
                         c->SKA1, c->SRA1, c->SRA2, c->SKB1, c->SRB1, c->SKB2, c->SRB2};
    for (int k = 0; k < 15; ++k) memcpy(out + (size_t)k * P, f[k], P * sizeof(double));
    memcpy(out + 15 * P, c->rInt, c->N * sizeof(double));
}

/* ---- time_steps.cpp:27-47 (QKstep) and :94-100 (QRstep) ------------------------ */
void or_rhs(or_ctx* c, double* hK, double* hR) {
    const size_t N = c->N, T1 = c->T1;
    const double t = c->t1[T1 - 1];
    const double* qK = c->QK + (T1 - 1) * N;
    const double* qR = c->QR + (T1 - 1) * N;
    double* d1 = dalloc(N);
    double* d2 = dalloc(N);
    double* cv = dalloc(N);
    const double s1 = fl1(c, qK[0]) / c->T0;
    const double s2 = -c->rInt[N - 1];
    for (size_t j = 0; j < N; ++j) d1[j] = c->QKB1[j * N] * s1 + qK[j] * s2;
    convR_rows(c, c->SRA2, c->QKB2, t, cv); for (size_t j = 0; j < N; ++j) d1[j] = d1[j] + cv[j];
    convA_rows(c, c->SRA1, c->QKB1, t, cv); for (size_t j = 0; j < N; ++j) d1[j] = d1[j] + cv[j];
    convA_rows(c, c->SKA1, c->QRB1, t, cv); for (size_t j = 0; j < N; ++j) d1[j] = d1[j] + cv[j];
    const double s3 = qK[0] / c->T0;
    const double s4 = 2 * c->Gamma;
    for (size_t j = 0; j < N; ++j) d2[j] = fl1(c, c->QKB1[j * N]) * s3 + qR[j] * s4;
    convR_rows(c, c->QRA2, c->SKB2, t, cv); for (size_t j = 0; j < N; ++j) d2[j] = d2[j] + cv[j];
    convA_rows(c, c->QRA1, c->SKB1, t, cv); for (size_t j = 0; j < N; ++j) d2[j] = d2[j] + cv[j];
    convA_rows(c, c->QKA1, c->SRB1, t, cv); for (size_t j = 0; j < N; ++j) d2[j] = d2[j] + cv[j];
    for (size_t j = 0; j < N; ++j) d2[j] = d2[j] - qK[j] * c->rInt[j];
    for (size_t j = 0; j < N; ++j) hK[j] = d1[j] + d2[j] * c->theta[j];

    convR_rows(c, c->SRA2, c->QRB2, t, cv);
    for (size_t j = 0; j < N; ++j) d1[j] = qR[j] * s2 + cv[j];
    convR_rows(c, c->QRA2, c->SRB2, t, cv);
    for (size_t j = 0; j < N; ++j) d2[j] = qR[j] * c->rInt[j] - cv[j];
    for (size_t j = 0; j < N; ++j) hR[j] = d1[j] + d2[j] * c->theta[j];
    free(d1); free(d2); free(cv);
}

/* ---- time_steps.cpp:83-92 --------------------------------------------------- */
double or_rstep(or_ctx* c) {
    const size_t N = c->N, T1 = c->T1;
    const double* qK = c->QK + (T1 - 1) * N;
    const double* qR = c->QR + (T1 - 1) * N;
    const double t = c->t1[T1 - 1];
    double* sK = dalloc(N);
    double* sR = dalloc(N);
    sigmaK(c, qK, sK, N);
    sigmaR(c, qK, qR, sR, N);
    const double r = c->Gamma + convA_scalar(c, sR, qK, t) + convA_scalar(c, sK, qR, t) + sK[0] * qK[0] / c->T0;
    free(sK); free(sR);
    return r;
}

/* ---- time_steps.cpp:117-125 ------------------------------------------------- */
double or_drstep2(or_ctx* c, const double* qK, const double* qR, const double* dqK, const double* dqR, double t) {
    const size_t N = c->N;
    double* sK = dalloc(N);
    double* sR = dalloc(N);
    double* dsK = dalloc(N);
    double* dsR = dalloc(N);
    sigmaK(c, qK, sK, N);
    sigmaR(c, qK, qR, sR, N);
    for (size_t i = 0; i < N; ++i) {
        dsK[i] = fl2(c, qK[i]) * dqK[i] + 0.0 * dqR[i];
        dsR[i] = (fl3(c, qK[i]) * qR[i]) * dqK[i] + fl2(c, qK[i]) * dqR[i];
    }
    const double r = convA_scalar(c, sR, qK, 1) + convA_scalar(c, sK, qR, 1) + convA_scalar(c, dsR, qK, t) +
                     convA_scalar(c, dsK, qR, t) + convA_scalar(c, sR, dqK, t) + convA_scalar(c, sK, dqR, t) +
                     (dsK[0] * qK[0] + sK[0] * dqK[0]) / c->T0;
    free(sK); free(sR); free(dsK); free(dsR);
    return r;
}

/* ---- history storage ------------------------------------------------------------ */
static void ensure_cap(or_ctx* c, size_t rows) {
    if (rows <= c->cap) return;
    size_t nc = c->cap ? c->cap : 16;
    while (nc < rows) nc *= 2;
    const size_t N = c->N;
    c->t1 = (double*)realloc(c->t1, nc * sizeof(double));
    c->rho = (double*)realloc(c->rho, nc * sizeof(double));
    c->rv = (double*)realloc(c->rv, nc * sizeof(double));
    c->drv = (double*)realloc(c->drv, nc * sizeof(double));
    c->QK = (double*)realloc(c->QK, nc * N * sizeof(double));
    c->QR = (double*)realloc(c->QR, nc * N * sizeof(double));
    c->dQK = (double*)realloc(c->dQK, nc * N * sizeof(double));
    c->dQR = (double*)realloc(c->dQR, nc * N * sizeof(double));
    c->cap = nc;
}

/* ---- time_steps.cpp:127-162 ----------------------------------------------------- */
static void append_all(or_ctx* c, const double* qK, const double* qR, const double* dqK, const double* dqR,
                       double dr, double t) {
    const size_t N = c->N;
    ensure_cap(c, c->T1 + 1);
    c->t1[c->T1] = t;
    const size_t idx = c->T1;
    const double tdiff = c->t1[idx] - c->t1[idx - 1];
    if (idx > 1) c->rho[idx] = tdiff / (c->t1[idx - 1] - c->t1[idx - 2]);
    else c->rho[idx] = 0.0;
    for (size_t i = 0; i < N; ++i) {
        c->QK[idx * N + i] = qK[i];
        c->QR[idx * N + i] = qR[i];
        c->dQK[idx * N + i] = tdiff * dqK[i];
        c->dQR[idx * N + i] = tdiff * dqR[i];
    }
    c->T1 += 1;
    c->drv[idx] = tdiff * dr;
    c->rv[idx] = or_rstep(c);
}

/* ---- time_steps.cpp:49-81 ------------------------------------------------------- */
static void replace_all(or_ctx* c, const double* qK, const double* qR, const double* dqK, const double* dqR,
                        double dr, double t) {
    const size_t N = c->N, T1 = c->T1, last = T1 - 1;
    c->t1[last] = t;
    const double tdiff = c->t1[T1 - 1] - c->t1[T1 - 2];
    if (T1 > 2) c->rho[last] = tdiff / (c->t1[T1 - 2] - c->t1[T1 - 3]);
    else c->rho[last] = 0.0;
    for (size_t i = 0; i < N; ++i) {
        c->QK[last * N + i] = qK[i];
        c->QR[last * N + i] = qR[i];
        c->dQK[last * N + i] = tdiff * dqK[i];
        c->dQR[last * N + i] = tdiff * dqR[i];
    }
    c->drv[last] = tdiff * dr;
    c->rv[last] = or_rstep(c);
}

/* ---- runge_kutta.cpp:19-252: shared stage driver ------------------------------
 * SSPRK104 (:19-144) and RK54 (:146-252) differ in tableau, in which stages
 * reuse stored positions (SSPRK: 5,6,7 reuse stages 2,3,4) and in the final
 * interpolation (RK54: reuse; SSPRK: isearch).                                      */
static double rk_generic(or_ctx* c, size_t S, const double* A, const double* b, const double* b2, int ssprk) {
    const size_t N = c->N;
    double* gK = dalloc((S + 1) * N);
    double* gR = dalloc((S + 1) * N);
    double* gt = dalloc(S + 1);
    double* hK = dalloc(S * N);
    double* hR = dalloc(S * N);
    double* ht = dalloc(S);
    double* pv1 = dalloc(3 * N);
    double* pv2 = dalloc(3 * N * N);
    double* tmp1 = dalloc(N);
    double* tmp2 = dalloc(N * N);
    memcpy(gK, c->QK + (c->T1 - 1) * N, N * sizeof(double));
    memcpy(gR, c->QR + (c->T1 - 1) * N, N * sizeof(double));
    gt[0] = c->t1[c->T1 - 1];
    const double dt = c->delta_t_val;
    double dr = 0.0;
    for (size_t n = 0; n < S; ++n) {
        if (c->T1 == 1 || n != 0) {
            if (n == 0) {
                interpolate_with(c, 0, NULL, NULL);
            } else if (ssprk && n >= 5 && n <= 7) {
                memcpy(tmp1, pv1 + (n - 5) * N, N * sizeof(double));
                memcpy(tmp2, pv2 + (n - 5) * N * N, N * N * sizeof(double));
                interpolate_with(c, 2, tmp1, tmp2);
            } else {
                memcpy(tmp1, c->posB1old, N * sizeof(double));
                memcpy(tmp2, c->posB2old, N * N * sizeof(double));
                interpolate_with(c, 1, tmp1, tmp2);
            }
        }
        if (ssprk && n >= 2 && n <= 4) {
            memcpy(pv1 + (n - 2) * N, c->posB1old, N * sizeof(double));
            memcpy(pv2 + (n - 2) * N * N, c->posB2old, N * N * sizeof(double));
        }
        or_rhs(c, hK + n * N, hR + n * N);
        ht[n] = 1.0;
        double* gKn = gK + (n + 1) * N;
        double* gRn = gR + (n + 1) * N;
        if (n == 0) {
            const double* qK = c->QK + (c->T1 - 1) * N;
            const double* qR = c->QR + (c->T1 - 1) * N;
            dr = or_drstep2(c, qK, qR, hK, hR, c->t1[c->T1 - 1]);
            const double a = dt * A[1 * S + 0];
            for (size_t i = 0; i < N; ++i) { gKn[i] = gK[i] + hK[i] * a; gRn[i] = gR[i] + hR[i] * a; }
            gt[1] = gt[0] + dt * A[1 * S + 0] * ht[0];
            append_all(c, gKn, gRn, hK, hR, ht[0] * dr, gt[1]);
        } else {
            const int last = (n == S - 1);
            memcpy(gKn, gK, N * sizeof(double));
            memcpy(gRn, gR, N * sizeof(double));
            gt[n + 1] = gt[0];
            const size_t jmax = last ? S : n + 1;
            for (size_t j = 0; j < jmax; ++j) {
                const double coef = last ? b[j] : A[(n + 1) * S + j];
                const double a = dt * coef;
                for (size_t i = 0; i < N; ++i) {
                    gKn[i] = gKn[i] + hK[j * N + i] * a;
                    gRn[i] = gRn[i] + hR[j * N + i] * a;
                }
                gt[n + 1] = gt[n + 1] + dt * coef * ht[j];
            }
            replace_all(c, gKn, gRn, hK, hR, ht[0] * dr, gt[n + 1]);
        }
    }
    memcpy(tmp1, c->posB1old, N * sizeof(double));
    memcpy(tmp2, c->posB2old, N * N * sizeof(double));
    interpolate_with(c, ssprk ? 1 : 2, tmp1, tmp2);

    double* gKe = dalloc(N);
    double* gRe = dalloc(N);
    memcpy(gKe, gK, N * sizeof(double));
    memcpy(gRe, gR, N * sizeof(double));
    double gte = gt[0];
    for (size_t j = 0; j < S; ++j) {
        const double a = dt * b2[j];
        for (size_t i = 0; i < N; ++i) { gKe[i] = gKe[i] + hK[j * N + i] * a; gRe[i] = gRe[i] + hR[j * N + i] * a; }
        gte = gte + dt * b2[j] * ht[j];
    }
    double err = 0.0;
    for (size_t i = 0; i < N; ++i) err += fabs(gK[S * N + i] - gKe[i]);
    for (size_t i = 0; i < N; ++i) err += fabs(gR[S * N + i] - gRe[i]);
    err += fabs(gt[S] - gte);
    free(gK); free(gR); free(gt); free(hK); free(hR); free(ht); free(pv1); free(pv2);
    free(tmp1); free(tmp2); free(gKe); free(gRe);
    return err;
}

/* Dormand-Prince 5(4) (runge_kutta.cpp:149-159) */
static const double DP_A[49] = {
    0.0, 0.0, 0.0, 0.0, 0.0, 0.0, 0.0,
    1.0 / 5, 0.0, 0.0, 0.0, 0.0, 0.0, 0.0,
    3.0 / 40, 9.0 / 40, 0.0, 0.0, 0.0, 0.0, 0.0,
    44.0 / 45, -56.0 / 15, 32.0 / 9, 0.0, 0.0, 0.0, 0.0,
    19372.0 / 6561, -25360.0 / 2187, 64448.0 / 6561, -212.0 / 729, 0.0, 0.0, 0.0,
    9017.0 / 3168, -355.0 / 33, 46732.0 / 5247, 49.0 / 176, -5103.0 / 18656, 0.0, 0.0,
    35.0 / 384, 0.0, 500.0 / 1113, 125.0 / 192, -2187.0 / 6784, 11.0 / 84, 0.0};
static const double DP_B[7] = {35.0 / 384, 0.0, 500.0 / 1113, 125.0 / 192, -2187.0 / 6784, 11.0 / 84, 0.0};
static const double DP_B2[7] = {5179.0 / 57600, 0.0, 7571.0 / 16695, 393.0 / 640, -92097.0 / 339200, 187.0 / 2100, 1.0 / 40};

/* SSPRK(10,4) (runge_kutta.cpp:22-35) */
static double SS_A[100], SS_B[10];
static const double SS_B2[10] = {0., 2.0 / 9, 0, 0, 5.0 / 18, 1.0 / 3, 0., 0., 0., 1.0 / 6};
static void ss_init(void) {
    for (int i = 0; i < 10; ++i) {
        SS_B[i] = 1.0 / 10;
        for (int j = 0; j < 10; ++j) {
            double v = 0.0;
            if (j < i) v = (i <= 4) ? 1.0 / 6 : (j <= 4 ? 1.0 / 15 : 1.0 / 6);
            SS_A[i * 10 + j] = v;
        }
    }
}

/* ---- SERK2 (runge_kutta.cu:228-297 coefficients, :397-487 stages) ---------------
 * s = 2q stages of size alpha = 1/(4q): even stages an Euler step g += dt alpha h,
 * odd stages the Chebyshev recursion g = 2 g - g_e + 2 dt alpha h (g_e: the value two
 * stages back, i.e. after the previous odd stage); the stage values are
 * T_i(x) T_2(x)^(j-1), x = 1 + alpha z, on y' = (z/dt) y.  b[0..s] combine them into
 * the embedded solution that matches R(z) = T_s(w0 + w1 z) / T_s(w0) at z = 1..s+1,
 * w0 = 1 + (19/20)/s^2, w1 = T_s(w0) / T_s'(w0); the error is |sum_k b_k g_k - g_s|. */
static long double cheb_T(int n, long double x) {
    long double a = 1.0L, b = x;
    if (n == 0) return a;
    for (int k = 1; k < n; ++k) { const long double c2 = 2.0L * x * b - a; a = b; b = c2; }
    return b;
}
static long double cheb_U(int n, long double x) {
    long double a = 1.0L, b = 2.0L * x;
    if (n == 0) return a;
    for (int k = 1; k < n; ++k) { const long double c2 = 2.0L * x * b - a; a = b; b = c2; }
    return b;
}
void or_serk2_coeffs(int q, double* out) {
    const int m = 2, s = m * q, n = s + 1;
    const long double mu = 19.0L / 20.0L;
    const long double a2 = (long double)q / ((long double)s * s);
    const long double w0 = 1.0L + mu / ((long double)s * s);
    const long double w1 = cheb_T(s, w0) / ((long double)s * cheb_U(s - 1, w0));
    long double* A = (long double*)calloc((size_t)n * (n + 1), sizeof(long double)); /* [n][n | rhs] */
    for (int r = 0; r < n; ++r) {
        const long double z = (long double)(r + 1), x = 1.0L + a2 * z;
        long double* row = A + (size_t)r * (n + 1);
        row[0] = 1.0L;
        long double tm = 1.0L; /* T_m(x)^(j-1) */
        for (int j = 1; j <= q; ++j) {
            for (int i = 1; i <= m; ++i) row[i + m * (j - 1)] = cheb_T(i, x) * tm;
            tm *= cheb_T(m, x);
        }
        row[n] = cheb_T(s, w0 + w1 * z) / cheb_T(s, w0);
    }
    /* Gaussian elimination with partial pivoting */
    for (int k = 0; k < n; ++k) {
        int pr = k;
        for (int r = k + 1; r < n; ++r)
            if (fabsl(A[(size_t)r * (n + 1) + k]) > fabsl(A[(size_t)pr * (n + 1) + k])) pr = r;
        if (pr != k)
            for (int col = 0; col <= n; ++col) {
                const long double t = A[(size_t)k * (n + 1) + col];
                A[(size_t)k * (n + 1) + col] = A[(size_t)pr * (n + 1) + col];
                A[(size_t)pr * (n + 1) + col] = t;
            }
        for (int r = k + 1; r < n; ++r) {
            const long double f = A[(size_t)r * (n + 1) + k] / A[(size_t)k * (n + 1) + k];
            for (int col = k; col <= n; ++col) A[(size_t)r * (n + 1) + col] -= f * A[(size_t)k * (n + 1) + col];
        }
    }
    for (int k = n - 1; k >= 0; --k) {
        long double v = A[(size_t)k * (n + 1) + n];
        for (int col = k + 1; col < n; ++col) v -= A[(size_t)k * (n + 1) + col] * (long double)out[col];
        out[k] = (double)(v / A[(size_t)k * (n + 1) + k]);
    }
    free(A);
}

static double serk2(or_ctx* c, int q) {
    const size_t N = c->N;
    const int S = 2 * q;
    double* b = dalloc((size_t)S + 1);
    or_serk2_coeffs(q, b);
    const double alpha = 1.0 / (4 * q);
    const double dt = c->delta_t_val;
    double* gK = dalloc(N);
    double* gR = dalloc(N);
    double* eK = dalloc(N); /* g_e */
    double* eR = dalloc(N);
    double* fK = dalloc(N); /* sum_k b_k g_k */
    double* fR = dalloc(N);
    double* hK = dalloc(N);
    double* hR = dalloc(N);
    double* h0K = dalloc(N);
    double* h0R = dalloc(N);
    double* tmp1 = dalloc(N);
    double* tmp2 = dalloc(N * N);
    const double* q0K = c->QK + (c->T1 - 1) * N;
    const double* q0R = c->QR + (c->T1 - 1) * N;
    for (size_t i = 0; i < N; ++i) {
        gK[i] = eK[i] = q0K[i];
        gR[i] = eR[i] = q0R[i];
        fK[i] = q0K[i] * b[0];
        fR[i] = q0R[i] * b[0];
    }
    const double t0 = c->t1[c->T1 - 1];
    double gt = t0, gte = t0;
    const double gtf = t0 + dt;
    double dr = 0.0;
    for (int n = 0; n < S; ++n) {
        if (n == 0) {
            if (c->T1 == 1) interpolate_with(c, 0, NULL, NULL);
        } else {
            memcpy(tmp1, c->posB1old, N * sizeof(double));
            memcpy(tmp2, c->posB2old, N * N * sizeof(double));
            interpolate_with(c, 1, tmp1, tmp2);
        }
        or_rhs(c, hK, hR);
        const double a = dt * alpha;
        if (n == 0) {
            memcpy(h0K, hK, N * sizeof(double));
            memcpy(h0R, hR, N * sizeof(double));
            dr = or_drstep2(c, q0K, q0R, hK, hR, t0);
            for (size_t i = 0; i < N; ++i) { gK[i] = gK[i] + hK[i] * a; gR[i] = gR[i] + hR[i] * a; }
            gt = t0 + dt * alpha * 1.0;
        } else if (n % 2 == 0) {
            for (size_t i = 0; i < N; ++i) { gK[i] = gK[i] + hK[i] * a; gR[i] = gR[i] + hR[i] * a; }
            gt = gt + dt * alpha * 1.0;
        } else {
            const double a2 = 2 * dt * alpha;
            for (size_t i = 0; i < N; ++i) {
                gK[i] = 2 * gK[i] - eK[i] + hK[i] * a2;
                gR[i] = 2 * gR[i] - eR[i] + hR[i] * a2;
                eK[i] = gK[i];
                eR[i] = gR[i];
            }
            gt = 2 * gt - gte + 2 * dt * alpha * 1.0;
            gte = gt;
        }
        if (b[n + 1] != 0.0)
            for (size_t i = 0; i < N; ++i) { fK[i] = fK[i] + gK[i] * b[n + 1]; fR[i] = fR[i] + gR[i] * b[n + 1]; }
        if (n == 0) append_all(c, gK, gR, h0K, h0R, dr, gt);
        else replace_all(c, gK, gR, h0K, h0R, dr, gt);
    }
    /* positions for the next step (the next step's first stage reuses them) */
    memcpy(tmp1, c->posB1old, N * sizeof(double));
    memcpy(tmp2, c->posB2old, N * N * sizeof(double));
    interpolate_with(c, 1, tmp1, tmp2);
    double err = 0.0;
    for (size_t i = 0; i < N; ++i) err += fabs(fK[i] - gK[i]);
    for (size_t i = 0; i < N; ++i) err += fabs(fR[i] - gR[i]);
    err += fabs(gtf - gt);
    free(b); free(gK); free(gR); free(eK); free(eR); free(fK); free(fR); free(hK); free(hR);
    free(h0K); free(h0R); free(tmp1); free(tmp2);
    return err;
}

/* ---- runge_kutta.cpp:254-261 (runge_kutta.cu:607-615 for SERK2) ------------------ */
double or_update(or_ctx* c) {
    if (c->rk_init == 1) return rk_generic(c, 7, DP_A, DP_B, DP_B2, 0);
    if (c->rk_init >= 3) return serk2(c, 2 * (c->rk_init - 2));
    ss_init();
    return rk_generic(c, 10, SS_A, SS_B, SS_B2, 1);
}

void or_set_serk2(or_ctx* c, int on) { c->use_serk2 = on; }
void or_set_maint_every(or_ctx* c, long every) { c->maint_every = every; }
void or_set_rollback(or_ctx* c, int on) { c->rollback = on; }

/* rollbackState(n) (simulation_control.cpp:88-113): resize every history array to T1 - n
 * rows, delta_t = the kept last step, loop -= n, a fresh interpolation; 0 when n >= T1 - 1 */
int or_rollback_state(or_ctx* c, int n) {
    if (n < 0 || (size_t)n >= c->T1 - 1) return 0;
    c->T1 -= (size_t)n;
    c->delta_t_val = c->t1[c->T1 - 1] - c->t1[c->T1 - 2];
    c->loop -= n;
    interpolate_with(c, 0, NULL, NULL);
    return 1;
}

/* ---- sparsify_utils.cpp:14-97 (CPU-only build) ------------------------------------ */
void or_sparsify(or_ctx* c, double threshold) {
    const size_t N = c->N, T1 = c->T1;
    size_t* keep = (size_t*)calloc(T1 + 1, sizeof(size_t));
    size_t nk = 0;
    int erased = 0;
    keep[nk++] = 0;
    for (size_t i = 2; i + 1 < T1; ++i) {
        const double tleft = c->t1[i - 2], tmid = c->t1[i];
        const double tdiff1 = c->t1[i - 1] - tleft, tdiff2 = tmid - tleft, tdiff3 = c->t1[i + 1] - tmid;
        double val = 0.0;
        for (size_t j = 0; j < N; ++j) {
            const double df1 = c->dQK[(i - 1) * N + j], df2 = c->dQK[(i + 1) * N + j];
            const double f1 = c->QK[i * N + j] - c->QK[(i - 2) * N + j];
            val += fabs(tdiff2 / 12.0 * (2 * f1 - tdiff2 * (df1 / tdiff1 + df2 / tdiff3)));
        }
        for (size_t j = 0; j < N; ++j) {
            const double f1 = c->QR[i * N + j] - c->QR[(i - 2) * N + j];
            const double df1 = c->dQR[(i - 1) * N + j], df2 = c->dQR[(i + 1) * N + j];
            val += fabs(tdiff2 / 12.0 * (2 * f1 - tdiff2 * (df1 / tdiff1 + df2 / tdiff3)));
        }
        const double dR1 = c->drv[i - 1], dR2 = c->drv[i + 1], Rt = c->rv[i] - c->rv[i - 2];
        val += fabs(tdiff2 / 12.0 * (2 * Rt - tdiff2 * (dR1 / tdiff1 + dR2 / tdiff3)));
        if (val < threshold) erased = 1;
        else keep[nk++] = i;
    }
    keep[nk++] = T1 - 1;
    if (erased) {
        for (size_t r = 0; r < nk; ++r) {
            const size_t s = keep[r];
            /* rows only move towards the front (s >= r), so in-place compaction is safe */
            if (s != r) {
                memmove(c->QK + r * N, c->QK + s * N, N * sizeof(double));
                memmove(c->QR + r * N, c->QR + s * N, N * sizeof(double));
                memmove(c->dQK + r * N, c->dQK + s * N, N * sizeof(double));
                memmove(c->dQR + r * N, c->dQR + s * N, N * sizeof(double));
            }
            c->rv[r] = c->rv[s]; c->drv[r] = c->drv[s]; c->t1[r] = c->t1[s]; c->rho[r] = c->rho[s];
        }
        c->T1 = nk;
    }
    free(keep);
}

/* ---- simulation_runner.cpp:334-337 --------------------------------------------- */
double or_energy(or_ctx* c, double t) {
    const size_t N = c->N;
    const double* qK = c->QK + (c->T1 - 1) * N;
    const double* qR = c->QR + (c->T1 - 1) * N;
    double* tmp = dalloc(N);
    sigmaK(c, qK, tmp, N);
    const double e = -(convA_scalar(c, tmp, qR, t) + fl1(c, qK[0]) / c->T0);
    free(tmp);
    return e;
}

/* ---- simulation_runner.cpp:125-278 (CPU-only branches, no I/O) ------------------
 * Returns the proposed time t written to correlation.txt.                           */
double or_loop_iteration(or_ctx* c) {
    double t = c->t1[c->T1 - 1] + c->delta_t_val;
    c->delta_old = c->delta;
    c->delta = or_update(c);
    /* the tail fit runs after update() and before loop++ (simulation_runner.cpp:135-141) */
    if (c->tail_fit) tail_fit_blend(c);
    c->loop++;
    if (c->loop % (c->maint_every > 0 ? c->maint_every : 100000) == 0) {
        size_t prev = c->T1;
        int count = 0;
        while (count < 1) {
            or_sparsify(c, c->delta_max);
            if (c->T1 >= prev) break;
            prev = c->T1;
            count++;
        }
        interpolate_with(c, 0, NULL, NULL);
        /* last_rollback_loop stays at its initial -1000 on the CPU path */
        if (c->delta < c->delta_max / 2 && c->loop - c->last_rollback > 1000) {
            c->delta_t_val *= 0.5;
            /* CPU path: SSPRK104; GPU path with use_serk2: RK54 -> SSPRK104, else one
             * SERK2 level up (init_SERK2(2 (init - 1)) sets init + 1), capped at the
             * 11 rmax entries */
            if (!c->use_serk2 || c->rk_init == 1) c->rk_init = 2;
            else if (c->rk_init < 11) c->rk_init += 1;
        }
    }
    if (c->delta < c->delta_max && c->loop > 5 &&
        (c->delta < 1.1 * c->delta_old || c->delta_old < c->delta_max / 1000) &&
        c->rmax[c->rk_init - 1] / c->specRad > c->delta_t_val && c->rho[c->T1 - 1] == 1.0) {
        c->delta_t_val *= 1.01;
    } else if (c->delta > 2 * c->delta_max && c->delta_t_val > c->dt_min) {
        c->delta_t_val *= 0.9;
    }
    if (c->rk_init == 2 && c->delta > c->delta_max && c->rmax[0] / c->specRad > c->delta_t_val) {
        c->rk_init = 1;
        c->delta_t_val *= 0.5;
    }
    /* the GPU main loop's rollback (simulation_runner.cpp:280-303) */
    if (c->rollback && c->delta > 2 * c->delta_max) {
        const long avail = (long)c->T1 - 2;
        const long n = avail < 0 ? 0 : (avail < 10 ? avail : 10);
        if (n > 0) {
            or_rollback_state(c, (int)n);
            c->last_rollback = c->loop;
            t = c->t1[c->T1 - 1] + c->delta_t_val;
        }
        {
            const double lo = c->delta_t_val > c->dt_min ? c->delta_t_val : c->dt_min;
            const double hi = c->rmax[c->rk_init - 1] / c->specRad;
            c->delta_t_val = 0.5 * (lo < hi ? lo : hi);
        }
        /* init_SERK2(2 (init - 3)) sets init - 1 (runge_kutta.cu:274-277); else SSPRK104 */
        c->rk_init = c->rk_init > 3 ? c->rk_init - 1 : 2;
    }
    return t;
}

/* ---- tail fit (tail_fit.cpp:46-265, tailFitBlendCPU) --------------------------------
 * y_j = (1 - QK_j) / (1 - theta_j)^2 on the upper half of the last row; the constant fit
 * of least SSE over windows of 16 (a full scan on the first call or a repeated call on the
 * same history length, else the previous window and its two neighbours); alpha smoothed
 * by an error-dependent gain; the blend QK -> 1 - alpha (1 - theta)^2 from the window on,
 * ramped over 8 or 16 points, applied only when the residuals and the window SSE pass the
 * gates.  Sums run in index order (the reference's omp simd reductions may reorder the
 * 16-term window sums; the result differs from this order by rounding only). */
static void tail_fit_blend(or_ctx* c) {
    const size_t len = c->N;
    if (len < 32) return;
    const size_t t1len = c->T1;
    if (t1len == 0) return;
    const size_t base = (t1len - 1) * len;
    const size_t start = len / 2, W = 16;
    if (len - start < W) return;
    const size_t Nt = len - start;
    double* y = dalloc(Nt);
    for (size_t j = start; j < len; ++j) {
        const double one_m = 1.0 - c->theta[j];
        const double denom = one_m * one_m;
        y[j - start] = (1.0 - c->QK[base + j]) / (denom > 0 ? denom : 1e-300);
    }
    const int new_time_step = c->tf_prev_t1len != t1len;
    size_t best_off = 0;
    double best_sse = INFINITY, best_mean = 0.0;
    size_t cand[3];
    size_t ncand = 0;
    if (c->tf_prev_off < 0 || !new_time_step) {
        const size_t count = (Nt - W) / W + 1;
        for (size_t w = 0; w < count; ++w) {
            const size_t off = w * W;
            double sum = 0.0, err = 0.0;
            for (size_t k = 0; k < W; ++k) sum += y[off + k];
            const double mean = sum / (double)W;
            for (size_t k = 0; k < W; ++k) { const double d = y[off + k] - mean; err += d * d; }
            if (err < best_sse) { best_sse = err; best_off = off; best_mean = mean; }
        }
    } else {
        const size_t prev = (size_t)c->tf_prev_off;
        /* candidates sorted and unique: prev - W < prev < prev + W */
        if (prev >= W) cand[ncand++] = prev - W;
        cand[ncand++] = prev;
        if (prev + W + W <= Nt) cand[ncand++] = prev + W;
        for (size_t q = 0; q < ncand; ++q) {
            const size_t off = cand[q];
            if (off + W > Nt) continue;
            double sum = 0.0, err = 0.0;
            for (size_t k = 0; k < W; ++k) sum += y[off + k];
            const double mean = sum / (double)W;
            for (size_t k = 0; k < W; ++k) { const double d = y[off + k] - mean; err += d * d; }
            if (err < best_sse) { best_sse = err; best_off = off; best_mean = mean; }
        }
    }
    free(y);
    c->tf_prev_off = (long)best_off;
    c->tf_prev_t1len = t1len;
    const double alpha_raw = best_mean;
    const size_t j0 = start + best_off;
    const double var = best_sse / (double)(W - 1);
    const double stderr_mean = sqrt((var > 0.0 ? var : 0.0) / (double)W);
    const int rollback_detected = c->last_rollback != c->tf_known_rollback;
    if (rollback_detected) {
        c->tf_known_rollback = c->last_rollback;
        if (alpha_raw != 0.0) c->tf_alpha = alpha_raw;
    }
    const double max_gain = 0.50, min_gain = 0.10;
    const double rel_err = fabs(alpha_raw) > 1e-12 ? stderr_mean / fabs(alpha_raw) : 1.0;
    double gain = max_gain / (1.0 + 2.0 * rel_err);
    gain = gain < max_gain ? gain : max_gain;
    gain = min_gain > gain ? min_gain : gain;
    if (!(c->tf_alpha == c->tf_alpha) || (alpha_raw == 0.0 && rollback_detected)) {
        if (alpha_raw != 0.0) c->tf_alpha = alpha_raw;
        else if (!(c->tf_alpha == c->tf_alpha)) c->tf_alpha = 0.0;
    } else if (alpha_raw == 0.0) {
        /* keep the smoothed value */
    } else {
        c->tf_alpha = (1.0 - gain) * c->tf_alpha + gain * alpha_raw;
    }
    const double alpha = c->tf_alpha;
    const size_t nres = len - j0;
    double mean_abs = 0.0, mean_res = 0.0, rv = 0.0;
    for (size_t j = j0; j < len; ++j) {
        const double one_m = 1.0 - c->theta[j];
        mean_abs += fabs((1.0 - alpha * one_m * one_m) - c->QK[base + j]);
    }
    mean_abs /= (double)nres;
    for (size_t j = j0; j < len; ++j) {
        const double one_m = 1.0 - c->theta[j];
        mean_res += (1.0 - alpha * one_m * one_m) - c->QK[base + j];
    }
    mean_res /= (double)nres;
    for (size_t j = j0; j < len; ++j) {
        const double one_m = 1.0 - c->theta[j];
        const double d = ((1.0 - alpha * one_m * one_m) - c->QK[base + j]) - mean_res;
        rv += d * d;
    }
    {
        const double den = (double)nres - 1.0;
        rv /= (1.0 > den ? 1.0 : den);
    }
    const double res_std = sqrt(0.0 > rv ? 0.0 : rv);
    const int apply = mean_abs <= 3.0 * res_std && best_sse <= sqrt(c->delta_max) && fabs(alpha) > 1e-14;
    if (!apply) return;
    const double ad = fabs(alpha) > 1e-12 ? fabs(alpha) : 1e-12;
    const double alpha_change = fabs(alpha_raw - alpha) / ad;
    const size_t ramp = alpha_change < 0.005 ? 8 : 16;
    for (size_t j = j0; j < len; ++j) {
        const double one_m = 1.0 - c->theta[j];
        const double qk_fit = 1.0 - alpha * one_m * one_m;
        double w = 1.0;
        if (j < j0 + ramp) w = (double)(j - j0) / (double)ramp;
        const double qk_num = c->QK[base + j];
        c->QK[base + j] = (1.0 - w) * qk_num + w * qk_fit;
    }
    c->tf_applied++;
}

void or_set_tail_fit(or_ctx* c, int on) { c->tail_fit = on; }
void or_set_gpu_variants(or_ctx* c, int on) { c->gpu_variants = on; }
long or_tail_fit_applied(const or_ctx* c) { return c->tf_applied; }
double or_tail_fit_alpha(const or_ctx* c) { return c->tf_alpha; }

/* ---- setup -------------------------------------------------------------------- */
static size_t* copy_inds(const int64_t* src, size_t n) {
    size_t* d = (size_t*)calloc(n, sizeof(size_t));
    for (size_t i = 0; i < n; ++i) d[i] = src[i] < 0 ? 0 : (size_t)src[i];
    return d;
}
static double* copy_d(const double* src, size_t n) {
    double* d = dalloc(n);
    if (src) memcpy(d, src, n * sizeof(double));
    return d;
}

or_ctx* or_create(int64_t N_, const double* theta, const double* integ, const double* phi2,
                  const int64_t* indsA1, const int64_t* indsA2, const int64_t* indsB2,
                  const double* wA1, const double* wA2, const double* wB2, int64_t m,
                  int p, int p2, double lambda, double T0, double Gamma,
                  double delta_t_min, double delta_max, int log_response) {
    or_ctx* c = (or_ctx*)calloc(1, sizeof(or_ctx));
    const size_t N = (size_t)N_, P = N * N;
    static const double rmax[11] = {3, 13, 20, 80, 180, 320, 500, 720, 980, 1280, 1620};
    c->N = N; c->m = (size_t)m;
    c->p = p; c->p2 = p2; c->lambda = lambda; c->T0 = T0; c->Gamma = Gamma;
    c->dt_min = delta_t_min; c->delta_max = delta_max; c->log_resp = log_response;
    memcpy(c->rmax, rmax, sizeof(rmax));
    c->theta = copy_d(theta, N);
    c->integ = copy_d(integ, N);
    c->phi2 = copy_d(phi2, P);
    c->iA1 = copy_inds(indsA1, P); c->iA2 = copy_inds(indsA2, P); c->iB2 = copy_inds(indsB2, P);
    c->wA1 = copy_d(wA1, P * c->m); c->wA2 = copy_d(wA2, P * c->m); c->wB2 = copy_d(wB2, P * c->m);
    c->posB1old = dalloc(N);
    for (size_t i = 0; i < N; ++i) c->posB1old[i] = 1.0;
    c->posB2old = dalloc(P);
    double** f[] = {&c->QKA1, &c->QRA1, &c->QKA2, &c->QRA2, &c->QKB1, &c->QRB1, &c->QKB2, &c->QRB2,
                    &c->SKA1, &c->SRA1, &c->SRA2, &c->SKB1, &c->SRB1, &c->SKB2, &c->SRB2};
    for (int k = 0; k < 15; ++k) *f[k] = dalloc(P);
    c->rInt = dalloc(N);
    c->specRad = 4 * sqrt(fl2(c, 1));
    c->last_rollback = -1000;
    c->tf_prev_off = -1;
    c->tf_known_rollback = -1000;
    ensure_cap(c, 16);
    return c;
}

void or_destroy(or_ctx* c) {
    if (!c) return;
    void* ptrs[] = {c->theta, c->integ, c->phi2, c->iA1, c->iA2, c->iB2, c->wA1, c->wA2, c->wB2,
                    c->t1, c->rho, c->rv, c->drv, c->QK, c->QR, c->dQK, c->dQR, c->posB1old, c->posB2old,
                    c->QKA1, c->QRA1, c->QKA2, c->QRA2, c->QKB1, c->QRB1, c->QKB2, c->QRB2,
                    c->SKA1, c->SRA1, c->SRA2, c->SKB1, c->SRB1, c->SKB2, c->SRB2, c->rInt};
    for (size_t k = 0; k < sizeof(ptrs) / sizeof(ptrs[0]); ++k) free(ptrs[k]);
    free(c);
}

/* ---- initialization.cpp:204-223, 278-283 ----------------------------------------- */
void or_init_quench(or_ctx* c) {
    const size_t N = c->N;
    ensure_cap(c, 1);
    c->T1 = 1;
    c->t1[0] = 0.0;
    c->rho[0] = 0.0;
    c->specRad = 4 * sqrt(fl2(c, 1));
    c->delta_t_val = c->dt_min;
    c->loop = 0;
    c->delta = 1;
    c->delta_old = 0;
    for (size_t i = 0; i < N; ++i) { c->QK[i] = 1.0; c->QR[i] = 1.0; c->dQK[i] = 0.0; c->dQR[i] = 0.0; }
    c->rv[0] = c->Gamma + fl1(c, 1) / c->T0;
    c->drv[0] = or_rstep(c);
    c->rk_init = (c->delta_t_val < c->rmax[0] / c->specRad) ? 1 : 2;
    interpolate_with(c, 0, NULL, NULL);
}

int64_t or_T1(const or_ctx* c) { return (int64_t)c->T1; }

void or_set_history(or_ctx* c, int64_t T1, const double* t1grid, const double* dtratio,
                    const double* rvec, const double* drvec,
                    const double* QK, const double* QR, const double* dQK, const double* dQR) {
    const size_t n = (size_t)T1, N = c->N;
    ensure_cap(c, n);
    c->T1 = n;
    memcpy(c->t1, t1grid, n * sizeof(double));
    memcpy(c->rho, dtratio, n * sizeof(double));
    memcpy(c->rv, rvec, n * sizeof(double));
    memcpy(c->drv, drvec, n * sizeof(double));
    memcpy(c->QK, QK, n * N * sizeof(double));
    memcpy(c->QR, QR, n * N * sizeof(double));
    memcpy(c->dQK, dQK, n * N * sizeof(double));
    memcpy(c->dQR, dQR, n * N * sizeof(double));
}

void or_get_history(const or_ctx* c, double* t1grid, double* dtratio, double* rvec, double* drvec,
                    double* QK, double* QR, double* dQK, double* dQR) {
    const size_t n = c->T1, N = c->N;
    memcpy(t1grid, c->t1, n * sizeof(double));
    memcpy(dtratio, c->rho, n * sizeof(double));
    memcpy(rvec, c->rv, n * sizeof(double));
    memcpy(drvec, c->drv, n * sizeof(double));
    memcpy(QK, c->QK, n * N * sizeof(double));
    memcpy(QR, c->QR, n * N * sizeof(double));
    memcpy(dQK, c->dQK, n * N * sizeof(double));
    memcpy(dQR, c->dQR, n * N * sizeof(double));
}

void or_set_control(or_ctx* c, double delta_t, double delta, double delta_old, int loop, int rk_init) {
    c->delta_t_val = delta_t; c->delta = delta; c->delta_old = delta_old; c->loop = loop; c->rk_init = rk_init;
}
void or_get_control(const or_ctx* c, double* delta_t, double* delta, double* delta_old, int* loop, int* rk_init) {
    *delta_t = c->delta_t_val; *delta = c->delta; *delta_old = c->delta_old; *loop = (int)c->loop; *rk_init = c->rk_init;
}
void or_set_pos_old(or_ctx* c, const double* posB1, const double* posB2) {
    memcpy(c->posB1old, posB1, c->N * sizeof(double));
    memcpy(c->posB2old, posB2, c->N * c->N * sizeof(double));
}
void or_get_pos_old(const or_ctx* c, double* posB1, double* posB2) {
    memcpy(posB1, c->posB1old, c->N * sizeof(double));
    memcpy(posB2, c->posB2old, c->N * c->N * sizeof(double));
}
