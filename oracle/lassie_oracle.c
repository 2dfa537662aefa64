/*
 * lassie_oracle.c -- CPU FP64 restatement of LASSIE's deterministic integrator.
 *
 * TEST INFRASTRUCTURE ONLY.  This file is the oracle: it is linked, loaded or
 * executed only by tests/, by __graft_entry__.smoke() and by bench.py's
 * cpu_baseline leg, and only as the checker.  The product (liblassie.so and
 * the LASSIE command line) never calls it.
 *
 * It restates, in plain C, the arithmetic of the reference's live code path:
 *   RHS                src/c++/header/kernels.cuh:5-63      (generate_ode, myPow)
 *   RKF45 stages       src/c++/header/kernels.cuh:68-182    (rk_Fehlberg1..6, _agg)
 *   error / step factor src/c++/header/kernels.cuh:184-199  (calculate_delta)
 *   Jacobian           src/c++/header/kernels.cuh:205-271   (createJac, createJac_)
 *   BDF1 residual      src/c++/header/kernels.cuh:289-300   (calculate_nonLinear1th)
 *   I - hJ             src/c++/header/kernels.cuh:393-402   (aggiorna_Jacobiano1)
 *   Newton update      src/c++/header/kernels.cuh:462-477   (aggiorna_Newton)
 *   NaN repair         src/c++/header/kernels.cuh:512-521   (aggiornaNaN)
 *   control loop       src/c++/header/Simulation.cuh:632-925
 *   LU (cuSOLVER getrf/getrs call sites)  Simulation.cuh:1280-1326
 *
 * The operation order is specified exactly so that the HIP path can be checked
 * bit for bit: expressions are evaluated left to right as written in the
 * reference kernels, compiled with -ffp-contract=off, and fused multiply-adds
 * appear only where fma() is written.  Where the reference leaves an order to
 * a closed library (Thrust reductions, cuSOLVER), this file fixes one:
 *   - min of the step factors: sequential "first element, then strictly
 *     smaller" semantics of thrust::min_element (Simulation.cuh:753-754);
 *   - Newton norm: sum of float-rounded squares (Simulation.cuh:58-64,
 *     883-888) over a fixed 256-lane strided layout followed by a halving
 *     tree (NORM_LANES below);
 *   - LU: right-looking kij elimination with partial pivoting (first maximum
 *     of |a|, strict >, LAPACK idamax semantics), multipliers by division,
 *     updates a_ij = fma(-l_ik, u_kj, a_ij) with k ascending; getrs applies the
 *     row interchanges in order, then forward (k ascending) and backward
 *     (k descending) substitution with fma.
 *   - the step factor 0.84*(tol/R)^(1/4) is computed as 0.84*sqrt(sqrt(tol/R))
 *     (both sqrt correctly rounded), in place of pow(., 0.25).
 * Defined behaviour where the reference is undefined: stage values of feed
 * species (never written by rk_Fehlberg1..5, kernels.cuh:74-80) are the feed
 * value; the loop stops right after the last sample is stored (the reference
 * reads t_vector[Nb_times] past the end, Simulation.cuh:689, and its output
 * is already complete at that point).
 * Extensions (all off by default, which is the reference's controller):
 *   rtol > 0 sets the per-species tolerance to atol + rtol*|w1|; rtol == 0 is
 *     the reference's absolute test (kernels.cuh:191-197);
 *   bdf_regrow: after a BDF step with no LU failure and no repair, the
 *     persistent BDF step doubles, capped at be_step (the reference only ever
 *     shrinks it, Simulation.cuh:872, 907);
 *   max_rkf_retries > 0: that many consecutive rejected RKF attempts send the
 *     next rejection to BDF (the reference can cycle on rejections forever);
 *   max_newton > 0: the loop stops (as at max_iters) once the solve has run
 *     that many modified-Newton iterations, a work bound per solve;
 *   refine: the precision form of the Newton linear solve (orc_refine_solve):
 *     FP32 factors of (float)(I - hJ), FP64 iterative refinement, FP64 LU of
 *     the step's matrix when the FP32 LU meets a zero pivot or refinement
 *     does not converge; refine_max_iter == 0 takes one FP32 solve per
 *     Newton iteration, unrefined (the Newton iteration refines).
 */
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

#include "lassie_oracle.h"

#define NORM_LANES 256

/* myPow, kernels.cuh:52-63 */
double orc_mypow(double base, int ex)
{
    double p = 1.0;
    while (ex > 0) { p = p * base; ex--; }
    return p;
}

/* generate_ode, kernels.cuh:5-50 */
void orc_rhs(const orc_model *m, const double *x, double *f)
{
    for (int s = 0; s < m->S; s++) {
        double acc = 0.0;
        for (int e = m->hptr[s]; e < m->hptr[s + 1]; e++) {
            int r = m->hrx[e];
            double sk = (double)m->hcoef[e] * m->k[r];
            int a0 = m->lptr[r], a1 = m->lptr[r + 1];
            if (a1 > a0) {
                double temp = 0.0;
                for (int j = a0; j < a1; j++) {
                    if (j == a0) temp = temp + sk * orc_mypow(x[m->lsp[j]], m->lexp[j]);
                    else temp = temp * orc_mypow(x[m->lsp[j]], m->lexp[j]);
                }
                acc = acc + temp;
            } else {
                acc = acc + sk;
            }
        }
        f[s] = acc;
    }
}

/* createJac_ , kernels.cuh:219-271, row-major J[i*S + g].  Only columns g that
 * are reactants of a reaction in row i receive a nonzero term; every other
 * term the reference adds is +0, which leaves the (never negative-zero)
 * accumulator unchanged, so the sparse walk below gives the same bits. */
void orc_jac(const orc_model *m, const double *x, double *J)
{
    const int S = m->S;
    memset(J, 0, sizeof(double) * (size_t)S * S);
    for (int i = 0; i < S; i++) {
        double *row = J + (size_t)i * S;
        for (int e = m->hptr[i]; e < m->hptr[i + 1]; e++) {
            int r = m->hrx[e];
            int c = m->hcoef[e];
            int a0 = m->lptr[r], a1 = m->lptr[r + 1];
            for (int j = a0; j < a1; j++) {
                int g = m->lsp[j];
                int ex = m->lexp[j];
                double temp = 0.0;
                if (ex == 1) {
                    temp = temp + (double)c * m->k[r];
                } else if (ex > 1) {
                    temp = temp + (((double)c * m->k[r]) * (double)ex) * orc_mypow(x[g], ex - 1);
                } else {
                    continue;
                }
                for (int jj = a0; jj < a1; jj++)
                    if (m->lsp[jj] != g) temp = temp * x[m->lsp[jj]];
                row[g] = row[g] + temp;
            }
        }
    }
}

/* aggiorna_Jacobiano1, kernels.cuh:393-402: M = I - h*J (row-major, in place) */
void orc_newton_matrix(int S, double h, double *J)
{
    for (int i = 0; i < S; i++)
        for (int g = 0; g < S; g++) {
            double id = (i == g) ? 1.0 : 0.0;
            J[(size_t)i * S + g] = id - h * J[(size_t)i * S + g];
        }
}

/* LU with partial pivoting (cusolverDnDgetrf semantics, Simulation.cuh:1304).
 * Row-major, in place.  Returns 0, or k+1 for the first exactly-zero pivot
 * (elimination stops there; the caller treats the matrix as singular,
 * Simulation.cuh:1309-1312). */
/* 1: the trailing update rounds the product and the difference separately
 * (a - l*u) instead of one fma: an equally valid FP64 LU, used to measure how
 * far two FP64 solvers' trajectories drift apart under this controller. */
static int g_lu_two_roundings = 0;
void orc_set_lu_variant(int two_roundings) { g_lu_two_roundings = two_roundings; }

int orc_lu(int n, double *A, int32_t *piv)
{
    for (int k = 0; k < n; k++) {
        int p = k;
        double bv = fabs(A[(size_t)k * n + k]);
        for (int i = k + 1; i < n; i++) {
            double v = fabs(A[(size_t)i * n + k]);
            if (v > bv) { bv = v; p = i; }
        }
        piv[k] = p;
        if (A[(size_t)p * n + k] == 0.0) return k + 1;
        if (p != k) {
            for (int j = 0; j < n; j++) {
                double t = A[(size_t)k * n + j];
                A[(size_t)k * n + j] = A[(size_t)p * n + j];
                A[(size_t)p * n + j] = t;
            }
        }
        const double ukk = A[(size_t)k * n + k];
        for (int i = k + 1; i < n; i++) A[(size_t)i * n + k] = A[(size_t)i * n + k] / ukk;
        for (int i = k + 1; i < n; i++) {
            const double l = A[(size_t)i * n + k];
            double *ri = A + (size_t)i * n;
            const double *rk = A + (size_t)k * n;
            if (g_lu_two_roundings) {
                for (int j = k + 1; j < n; j++) {
                    const double prod = l * rk[j];
                    ri[j] = ri[j] - prod;
                }
            } else {
                for (int j = k + 1; j < n; j++) ri[j] = fma(-l, rk[j], ri[j]);
            }
        }
    }
    return 0;
}

/* getrs with NRHS = 1 (Simulation.cuh:1316): b is overwritten with x. */
void orc_lu_solve(int n, const double *LU, const int32_t *piv, double *b)
{
    for (int k = 0; k < n; k++) {
        int p = piv[k];
        if (p != k) { double t = b[k]; b[k] = b[p]; b[p] = t; }
    }
    for (int k = 0; k < n; k++) {
        const double yk = b[k];
        for (int i = k + 1; i < n; i++) b[i] = fma(-LU[(size_t)i * n + k], yk, b[i]);
    }
    for (int k = n - 1; k >= 0; k--) {
        b[k] = b[k] / LU[(size_t)k * n + k];
        const double xk = b[k];
        for (int i = 0; i < k; i++) b[i] = fma(-LU[(size_t)i * n + k], xk, b[i]);
    }
}

/* Precision form of the Newton linear solve (extends linearSolverLU,
 * Simulation.cuh:1280-1326): M x = rhs with the FP32 factors LU32/piv of
 * (float)M and FP64 iterative refinement against the FP64 matrix M
 * (row-major, unfactored).  Sweep: b32 = (float)r; b32 = LU32 \ b32;
 * x = x + (double)b32; r_i = rhs_i - sum_j M_ij x_j and
 * mag_i = |rhs_i| + sum_j |M_ij||x_j|, both as fma chains over j ascending
 * from 0, then mag_i = mag_i + floor_i when floor is given.  Converged when
 * |r_i| <= tol * mag_i for every i: the Oettli-Prager componentwise backward
 * error, with the species' absolute tolerance as a floor in the integrator,
 * so that components far below it (1e-40 and less, under FP32's normal
 * range) do not hold refinement back.  x starts at 0 and r at rhs.  Returns
 * 1 on convergence within max_it sweeps, else 0; *iters counts the sweeps.
 * max_it == 0: the first sweep's x with no residual and no test (returns 1):
 * the modified-Newton iteration around the solve, whose residual is formed
 * in FP64, is then the only refinement. */
int orc_refine_solve(int n, const double *M, const float *LU32, const int32_t *piv, const double *rhs,
                     const double *floor, double *x, double *r, float *b32, double tol, int max_it,
                     int64_t *iters)
{
    for (int i = 0; i < n; i++) {
        x[i] = 0.0;
        r[i] = rhs[i];
    }
    if (max_it == 0) {
        (*iters)++;
        for (int i = 0; i < n; i++) b32[i] = (float)r[i];
        orc32_lu_solve(n, LU32, piv, b32);
        for (int i = 0; i < n; i++) x[i] = x[i] + (double)b32[i];
        return 1;
    }
    for (int it = 0; it < max_it; it++) {
        (*iters)++;
        for (int i = 0; i < n; i++) b32[i] = (float)r[i];
        orc32_lu_solve(n, LU32, piv, b32);
        for (int i = 0; i < n; i++) x[i] = x[i] + (double)b32[i];
        int bad = 0;
        for (int i = 0; i < n; i++) {
            const double *row = M + (size_t)i * n;
            double pa = 0.0, pm = 0.0;
            for (int j = 0; j < n; j++) {
                pa = fma(row[j], x[j], pa);
                pm = fma(fabs(row[j]), fabs(x[j]), pm);
            }
            r[i] = rhs[i] - pa;
            double mag = fabs(rhs[i]) + pm;
            if (floor) mag = mag + floor[i];
            if (!(fabs(r[i]) <= tol * mag)) bad = 1;
        }
        if (!bad) return 1;
    }
    return 0;
}

/* Newton norm squared: Thrust transform_reduce with a float-returning square
 * (Simulation.cuh:58-64, 883-888) in a fixed order. */
double orc_normsq(int n, const double *d)
{
    double lane[NORM_LANES];
    for (int l = 0; l < NORM_LANES; l++) {
        double s = 0.0;
        for (int i = l; i < n; i += NORM_LANES) s = s + (double)(float)(d[i] * d[i]);
        lane[l] = s;
    }
    for (int stride = NORM_LANES / 2; stride > 0; stride >>= 1)
        for (int l = 0; l < stride; l++) lane[l] = lane[l] + lane[l + stride];
    return lane[0];
}

/* thrust::min_element semantics (Simulation.cuh:753-754, 773-774). */
double orc_min_first(int n, const double *d)
{
    double m = d[0];
    for (int i = 1; i < n; i++)
        if (d[i] < m) m = d[i];
    return m;
}

/* One RKF45 attempt (Simulation.cuh:692-740): fills w1, w2, delta; returns the
 * number of species that pass calculate_delta's test. */
int orc_rkf_ws_attempt(const orc_model *m, double rtol, double dt, const double *y1, orc_rkf_ws *w)
{
    const int S = m->S;
    const double *fd = m->feed;
    orc_rhs(m, y1, w->k1);
    for (int s = 0; s < S; s++) {
        if (fd[s] != 0.0) { w->k1[s] = fd[s]; w->u[s] = fd[s]; }
        else { w->k1[s] = w->k1[s] * dt; w->u[s] = y1[s] + w->k1[s] / 4.0; }
    }
    orc_rhs(m, w->u, w->k2);
    for (int s = 0; s < S; s++) {
        if (fd[s] != 0.0) { w->k2[s] = fd[s]; w->u[s] = fd[s]; }
        else {
            w->k2[s] = w->k2[s] * dt;
            w->u[s] = y1[s] + (3.0 * w->k1[s]) / 32.0 + (9.0 * w->k2[s]) / 32.0;
        }
    }
    orc_rhs(m, w->u, w->k3);
    for (int s = 0; s < S; s++) {
        if (fd[s] != 0.0) { w->k3[s] = fd[s]; w->u[s] = fd[s]; }
        else {
            w->k3[s] = w->k3[s] * dt;
            w->u[s] = y1[s] + (1932.0 * w->k1[s]) / 2197.0 - (7200.0 * w->k2[s]) / 2197.0
                      + (7296.0 * w->k3[s]) / 2197.0;
        }
    }
    orc_rhs(m, w->u, w->k4);
    for (int s = 0; s < S; s++) {
        if (fd[s] != 0.0) { w->k4[s] = fd[s]; w->u[s] = fd[s]; }
        else {
            w->k4[s] = w->k4[s] * dt;
            w->u[s] = y1[s] + (439.0 * w->k1[s]) / 216.0 - (8.0 * w->k2[s])
                      + (3680.0 * w->k3[s]) / 513.0 - (845.0 * w->k4[s]) / 4104.0;
        }
    }
    orc_rhs(m, w->u, w->k5);
    for (int s = 0; s < S; s++) {
        if (fd[s] != 0.0) { w->k5[s] = fd[s]; w->u[s] = fd[s]; }
        else {
            w->k5[s] = w->k5[s] * dt;
            w->u[s] = y1[s] - (8.0 * w->k1[s]) / 27.0 + (2.0 * w->k2[s])
                      - (3544.0 * w->k3[s]) / 2565.0 + (1859.0 * w->k4[s]) / 4104.0
                      - (11.0 * w->k5[s]) / 40.0;
        }
    }
    orc_rhs(m, w->u, w->k6);
    int nok = 0;
    for (int s = 0; s < S; s++) {
        double w1, w2;
        if (fd[s] != 0.0) { w->k6[s] = fd[s]; w1 = fd[s]; w2 = fd[s]; }
        else {
            w->k6[s] = w->k6[s] * dt;
            w1 = y1[s] + (25.0 * w->k1[s]) / 216.0 + (1408.0 * w->k3[s]) / 2565.0
                 + (2197.0 * w->k4[s]) / 4104.0 - w->k5[s] / 5.0;
            w2 = y1[s] + (16.0 * w->k1[s]) / 135.0 + (6656.0 * w->k3[s]) / 12825.0
                 + (28561.0 * w->k4[s]) / 56430.0 - (9.0 * w->k5[s] / 50.0)
                 + (2.0 * w->k6[s]) / 55.0;
        }
        w->w1[s] = w1;
        w->w2[s] = w2;
        /* calculate_delta, kernels.cuh:184-199 */
        double R = fabs(w1 - w2) / dt;
        double tol = (rtol > 0.0) ? m->atol[s] + rtol * fabs(w1) : m->atol[s];
        w->delta[s] = 0.84 * sqrt(sqrt(tol / R));
        if (R <= tol && !isnan(R) && !isinf(R) && w1 > 0) nok++;
    }
    return nok;
}

/* Exposed single attempt for known-answer tests. */
int orc_rkf_attempt(const orc_model *m, double rtol, double dt, const double *y1,
                    double *w1, double *w2, double *delta)
{
    const int S = m->S;
    double *buf = (double *)malloc(sizeof(double) * (size_t)S * 7);
    orc_rkf_ws w = {buf, buf + S, buf + 2 * S, buf + 3 * S, buf + 4 * S, buf + 5 * S, buf + 6 * S,
                w1, w2, delta};
    int nok = orc_rkf_ws_attempt(m, rtol, dt, y1, &w);
    free(buf);
    return nok;
}

/* aggiornaNaN, kernels.cuh:512-521 */
int orc_repair(int S, double *yv, const double *y1)
{
    int n = 0;
    for (int s = 0; s < S; s++)
        if (isnan(yv[s]) || isinf(yv[s]) || yv[s] < 0) { yv[s] = y1[s]; n++; }
    return n;
}

int orc_count_bad(int S, const double *yv)
{
    int n = 0;
    for (int s = 0; s < S; s++)
        if (isnan(yv[s]) || isinf(yv[s]) || yv[s] < 0) n++;
    return n;
}

/* One BDF1 step solved by modified Newton (Simulation.cuh:780-919).  y1 is
 * replaced by the new state; returns the step size taken.  *newdt is the
 * persistent BDF step (newDtBDF). */
static double bdf_step(const orc_model *m, const orc_opts *o, double t, double t1, double *newdt,
                       double *y1, orc_bdf_ws *w, orc_stats *st, int32_t *kind)
{
    const int S = m->S;
    double dt = *newdt;
    if ((t + dt) > t1) dt = fabs(t1 - t);
    memcpy(w->yv, y1, sizeof(double) * S);
    int itN = 0;
    double errN = 1000;
    int luFail = 0;
    int info = 0;
    /* precision form: FP32 factors of (float)M until refinement fails to
     * converge (or the FP32 LU meets a zero pivot); then the FP64 LU of M
     * for the rest of the step */
    int fp64 = !o->refine;
    while (itN < o->newton_iter && errN > o->newton_tol) {
        orc_rhs(m, w->yv, w->x);
        for (int s = 0; s < S; s++) w->x[s] = -(w->yv[s] - y1[s] - dt * w->x[s]);
        if (itN == 0) {
            orc_jac(m, w->yv, w->J);
            orc_newton_matrix(S, dt, w->J);
            if (!fp64) {
                const size_t nn = (size_t)S * S;
                for (size_t e = 0; e < nn; e++) w->J32[e] = (float)w->J[e];
                if (orc32_lu(S, w->J32, w->piv) != 0) {
                    fp64 = 1;
                    st->refine_fallbacks++;
                }
            }
            if (fp64) info = orc_lu(S, w->J, w->piv);
        }
        if (info != 0) {
            orc_repair(S, w->yv, y1);
            *newdt = dt * 0.1;
            luFail = 1;
            st->lu_failures++;
            break;
        }
        if (fp64) {
            orc_lu_solve(S, w->J, w->piv, w->x);
        } else if (orc_refine_solve(S, w->J, w->J32, w->piv, w->x, m->atol, w->xr, w->r, w->b32,
                                    o->refine_tol, o->refine_max_iter, &st->refine_iters)) {
            memcpy(w->x, w->xr, sizeof(double) * S);
        } else {
            fp64 = 1;
            st->refine_fallbacks++;
            info = orc_lu(S, w->J, w->piv);
            if (info != 0) {
                orc_repair(S, w->yv, y1);
                *newdt = dt * 0.1;
                luFail = 1;
                st->lu_failures++;
                break;
            }
            orc_lu_solve(S, w->J, w->piv, w->x);
        }
        for (int s = 0; s < S; s++) {
            w->yv1[s] = w->yv[s];
            if (m->feed[s] == 0) w->yv[s] = w->yv[s] + w->x[s];
            else w->yv[s] = m->feed[s];
            w->x[s] = w->yv[s] - w->yv1[s];
        }
        errN = sqrt(orc_normsq(S, w->x));
        itN = itN + 1;
    }
    st->newton_iters += itN;
    *kind = luFail ? ORC_STEP_LUFAIL : itN;
    if (!luFail) {
        if (orc_count_bad(S, w->yv) > 0) {
            orc_repair(S, w->yv, y1);
            *newdt = dt * 0.1;
            st->nan_repairs++;
        } else if (o->bdf_regrow) {
            *newdt = fmin(*newdt * 2.0, o->be_step);
        }
    }
    memcpy(y1, w->yv, sizeof(double) * S);
    return dt;
}

/* Simulation<T>::run main loop, Simulation.cuh:632-925.
 * out[i*ncs + j] = state of species cs[j] at sample i; y_final (optional) the
 * state when the loop ends.  progress (optional,
 * capacity progress_cap) receives t/t_max at the top of every loop iteration
 * (the ProgressSimulation file, Simulation.cuh:660-672).
 * Returns 0, or 1 when max_iters is reached. */
int orc_solve(const orc_model *m, const orc_opts *o, const double *tv, int nt, const int32_t *cs,
              int ncs, double *out, double *y_final, double *progress, int64_t progress_cap,
              orc_stats *st)
{
    return orc_solve_rec(m, o, tv, nt, cs, ncs, out, y_final, progress, progress_cap, st, NULL);
}

/* Appends one step to the tape (stops recording on overflow). */
static void tape_push(orc_tape *tape, double dt, int32_t kind)
{
    if (!tape) return;
    if (tape->n >= tape->cap) { tape->overflow = 1; return; }
    tape->dt[tape->n] = dt;
    tape->kind[tape->n] = kind;
    tape->n++;
}

/* orc_solve that also records the steps it takes (tape may be NULL). */
int orc_solve_rec(const orc_model *m, const orc_opts *o, const double *tv, int nt, const int32_t *cs,
                  int ncs, double *out, double *y_final, double *progress, int64_t progress_cap,
                  orc_stats *st, orc_tape *tape)
{
    const int S = m->S;
    orc_stats local;
    if (!st) st = &local;
    memset(st, 0, sizeof(*st));
    double *y1 = (double *)malloc(sizeof(double) * S);
    double *buf = (double *)malloc(sizeof(double) * (size_t)S * 14);
    orc_rkf_ws rw = {buf, buf + S, buf + 2 * S, buf + 3 * S, buf + 4 * S, buf + 5 * S, buf + 6 * S,
                 buf + 7 * S, buf + 8 * S, buf + 9 * S};
    orc_bdf_ws bw;
    bw.J = (double *)malloc(sizeof(double) * (size_t)S * S);
    bw.x = buf + 10 * S;
    bw.yv = buf + 11 * S;
    bw.yv1 = buf + 12 * S;
    bw.piv = (int32_t *)malloc(sizeof(int32_t) * S);
    bw.J32 = NULL;
    bw.b32 = NULL;
    bw.xr = bw.r = NULL;
    if (o->refine) {
        bw.J32 = (float *)malloc(sizeof(float) * (size_t)S * S);
        bw.b32 = (float *)malloc(sizeof(float) * S);
        bw.xr = (double *)malloc(sizeof(double) * 2 * (size_t)S);
        bw.r = bw.xr + S;
    }
    memcpy(y1, m->x0, sizeof(double) * S);

    int i = 0;
    double t1 = tv[0];
    double t = 0;
    double dt = o->rkf_step;
    double newdt = o->be_step;
    const double t_max = tv[nt - 1];
    int rc = 0;
    int64_t nrej = 0;
    if (tape) { tape->n = 0; tape->overflow = 0; }
    while (i < nt) {
        if (o->max_iters > 0 && st->loop_iters >= o->max_iters) { rc = 1; break; }
        if (o->max_newton > 0 && st->newton_iters >= o->max_newton) { rc = 1; break; }
        if (progress && st->loop_iters < progress_cap) progress[st->loop_iters] = t / t_max;
        st->loop_iters++;
        if (fabs(t1 - t) < 1e-5) {
            for (int j = 0; j < ncs; j++) out[(size_t)i * ncs + j] = y1[cs[j]];
            if (tape) tape->sample_step[i] = (int32_t)tape->n;
            i++;
            st->samples = i;
            if (i == nt) break;
            t1 = tv[i];
        }
        st->rkf_attempts++;
        int nok = orc_rkf_ws_attempt(m, o->rtol, dt, y1, &rw);
        double mind = orc_min_first(S, rw.delta);
        if (nok == S) {
            nrej = 0;
            st->rkf_accepts++;
            memcpy(y1, rw.w1, sizeof(double) * S);
            tape_push(tape, dt, ORC_STEP_RKF);
            double d_temp = dt * mind;
            if ((t + d_temp) > t1) dt = fabs(t1 - t);
            else dt = d_temp;
            t = t + dt;
        } else {
            double dt_old = dt;
            dt = dt * mind;
            nrej++;
            if (dt < o->stiffness_tol || isnan(dt) || isinf(dt) || fabs(dt_old - dt) < 1e-3
                || (o->max_rkf_retries > 0 && nrej >= o->max_rkf_retries)) {
                nrej = 0;
                st->bdf_steps++;
                int32_t kind = 0;
                dt = bdf_step(m, o, t, t1, &newdt, y1, &bw, st, &kind);
                tape_push(tape, dt, kind);
                t = t + dt;
            }
        }
    }
    if (y_final) memcpy(y_final, y1, sizeof(double) * S);
    st->t_final = t;
    st->dt_final = dt;
    st->newdt_final = newdt;
    free(y1);
    free(buf);
    free(bw.J);
    free(bw.piv);
    free(bw.J32);
    free(bw.b32);
    free(bw.xr);
    return rc;
}
