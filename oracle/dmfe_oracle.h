/* CPU oracle for the per-RK-step renormalized DMFT update.
 *
 * TEST INFRASTRUCTURE ONLY.  This is a plain-C restatement of the reference's
 * CPU path, used by tests/, by __graft_entry__.smoke() and by bench.py's
 * cpu_baseline leg as the checker.  The product (dynamite_amd/) never links,
 * loads or calls it.
 *
 * Every function cites the reference file:line it restates.  Evaluation order
 * of floating-point expressions follows the reference literally, so that with
 * -fno-fast-math -ffp-contract=off and one thread the results are bit-identical
 * to the reference built with the same flags (checked in tests/test_oracle_vs_ref.py).
 *
 * Pinning (tests/test_oracle_*.py):
 *   f, f', f'', f''', Sigma          closed-form polynomials, exact rationals
 *   ConvA / ConvR                    quadrature moments of int.dat, closed-form t*theta scaling
 *   bsearch / isearch                brute-force lattice + closed-form position, and the reference build
 *   LN3 / N / MatAll / R2            polynomial reproduction in theta and in time, exact row copies
 *   QKstep / QRstep / drstep2        first-step known answer (T0 = inf, t = 0)
 *   RK54 / SSPRK104                  scalar y' = -y stability polynomial (via the h_t = 1 clock)
 *   everything                       bit-exact against the reference build (oracle/_ref)
 */
#ifndef DMFE_ORACLE_H
#define DMFE_ORACLE_H
#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct or_ctx or_ctx;

or_ctx* or_create(int64_t N, const double* theta, const double* integ, const double* phi2,
                  const int64_t* indsA1, const int64_t* indsA2, const int64_t* indsB2,
                  const double* wA1, const double* wA2, const double* wB2, int64_t m,
                  int p, int p2, double lambda, double T0, double Gamma,
                  double delta_t_min, double delta_max, int log_response);
void or_destroy(or_ctx* c);

void or_init_quench(or_ctx* c);
int64_t or_T1(const or_ctx* c);
void or_set_history(or_ctx* c, int64_t T1, const double* t1grid, const double* dtratio,
                    const double* rvec, const double* drvec,
                    const double* QK, const double* QR, const double* dQK, const double* dQR);
void or_get_history(const or_ctx* c, double* t1grid, double* dtratio, double* rvec, double* drvec,
                    double* QK, double* QR, double* dQK, double* dQR);
void or_set_control(or_ctx* c, double delta_t, double delta, double delta_old, int loop, int rk_init);
void or_get_control(const or_ctx* c, double* delta_t, double* delta, double* delta_old, int* loop, int* rk_init);
void or_set_pos_old(or_ctx* c, const double* posB1, const double* posB2);
void or_get_pos_old(const or_ctx* c, double* posB1, double* posB2);

void or_interpolate(or_ctx* c, int mode);     /* 0 bsearch, 1 isearch from old, 2 reuse old */
void or_get_fields(const or_ctx* c, double* out);
void or_rhs(or_ctx* c, double* hK, double* hR);
double or_rstep(or_ctx* c);
double or_drstep2(or_ctx* c, const double* qK, const double* qR, const double* dqK, const double* dqR, double t);
double or_update(or_ctx* c);
/* SERK2 (the reference GPU loop's stabilized integrator): b[0..2q] for q, and the
 * maintenance promotion to it (use_serk2, off = CPU path). */
void or_serk2_coeffs(int q, double* out);
void or_set_serk2(or_ctx* c, int on);
/* the every-1e5-loops maintenance cadence (simulation_runner.cpp:153), lowered in tests */
void or_set_maint_every(or_ctx* c, long every);
/* the reference GPU main loop's rollback on an error spike (simulation_runner.cpp:280-303,
 * simulation_control.cpp:14-76); off = the CPU path */
void or_set_rollback(or_ctx* c, int on);
int or_rollback_state(or_ctx* c, int n);
/* the QK tail fit after every main-loop update (tail_fit.cpp:46-265); off by default */
void or_set_tail_fit(or_ctx* c, int on);
/* the reference GPU path's interpolation variants (bisection only, per-element B1 clamp) */
void or_set_gpu_variants(or_ctx* c, int on);
long or_tail_fit_applied(const or_ctx* c);
double or_tail_fit_alpha(const or_ctx* c);
double or_loop_iteration(or_ctx* c);
double or_energy(or_ctx* c, double t);
void or_sparsify(or_ctx* c, double threshold);

void or_bsearch(int64_t nlist, const double* list, int64_t nel, const double* el, double* out);
void or_isearch(int64_t nlist, const double* list, int64_t nel, int64_t rowlen, const double* el,
                const double* inits, double* out);

double or_flambda(or_ctx* c, double q, int deriv);   /* deriv 0..3 */

#ifdef __cplusplus
}
#endif
#endif
