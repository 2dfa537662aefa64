/*
 * lassie_oracle_f32.c -- CPU restatement of LASSIE's deterministic integrator
 * in single precision: Simulation<float>, the `-p float` mode
 * (simulator.cu:259-304 instantiates Simulation<float>).
 *
 * TEST INFRASTRUCTURE ONLY.  This file is part of the oracle: it is linked,
 * loaded or executed only by tests/ and by the benchmark's CPU leg, as the
 * checker.  The product never calls it.
 *
 * Every state vector, the time, the step sizes, the tolerances and the dense
 * Newton matrix are float.  Expressions are written as the reference writes
 * them, so C's arithmetic conversions reproduce the reference's mixed
 * precision exactly:
 *   - generate_ode (kernels.cuh:5-50): (sign * k) is a float product; myPow
 *     multiplies in float and returns double (kernels.cuh:52-63); the term
 *     is accumulated in double and added to the float output, which rounds
 *     after every term;
 *   - rk_Fehlberg1..6 / _agg (kernels.cuh:68-182): k_i = k_i * dt in float;
 *     the stage and solution combinations have double literals, so they are
 *     evaluated in double and rounded to float when stored;
 *   - calculate_delta (kernels.cuh:184-199): R = |w1 - w2| / dt in float; the
 *     factor 0.84 * (tol/R)^(1/4) is taken as 0.84 * sqrtf(sqrtf(tol/R))
 *     (correctly rounded float square roots; CUDA's powf is not restated),
 *     evaluated in double and stored as float;
 *   - createJac_ (kernels.cuh:219-271): double accumulator per term, added to
 *     the float entry; aggiorna_Jacobiano1 (kernels.cuh:393-402) in float;
 *   - calculate_nonLinear1th, aggiorna_Newton (kernels.cuh:289-300, 462-477)
 *     in float;
 *   - the Newton norm (Simulation.cuh:58-64, 883-890): float squares summed
 *     in float (the reduction order fixed as in the FP64 oracle: 256 strided
 *     lanes, then a halving tree), sqrtf;
 *   - LU and getrs (cusolverDnSgetrf/Sgetrs seat, Simulation.cuh:1328-1359):
 *     the FP64 oracle's algorithm in float with fmaf;
 *   - step control (Simulation.cuh:751-781): t, t1, dt, d_temp, dt_old and
 *     newDtBDF are float; comparisons against the literals 1e-5 and 1e-3 are
 *     made in double; newDtBDF = dt * 0.1 is evaluated in double.
 * Inputs held as double (k, x0, feed, atol, t_vector and the options) are
 * rounded to float once, as the reference's reader does when it fills T
 * vectors (Reader.h:38-134).  Outputs are the float values widened to double.
 * Extensions and defined behaviour are those of the FP64 oracle
 * (lassie_oracle.c): feed stage values, rtol (tol = atol + rtol*|w1|, formed
 * in double and rounded to float), BDF regrowth, bounded RKF retries.
 */
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

#include "lassie_oracle.h"

#define NORM_LANES 256

/* myPow<float>, kernels.cuh:52-63: float products, widened on return */
static double mypow_f(float base, int ex)
{
    float p = 1.0f;
    while (ex > 0) { p = p * base; ex--; }
    return (double)p;
}

/* generate_ode<float>, kernels.cuh:5-50 */
void orc32_rhs(const orc_model *m, const float *x, float *f)
{
    for (int s = 0; s < m->S; s++) {
        float acc = 0.0f;
        for (int e = m->hptr[s]; e < m->hptr[s + 1]; e++) {
            int r = m->hrx[e];
            float sk = (float)m->hcoef[e] * (float)m->k[r];
            int a0 = m->lptr[r], a1 = m->lptr[r + 1];
            if (a1 > a0) {
                double temp = 0.0;
                for (int j = a0; j < a1; j++) {
                    if (j == a0) temp = temp + sk * mypow_f(x[m->lsp[j]], m->lexp[j]);
                    else temp = temp * mypow_f(x[m->lsp[j]], m->lexp[j]);
                }
                acc = (float)(acc + temp);
            } else {
                acc = acc + sk;
            }
        }
        f[s] = acc;
    }
}

/* createJac_<float>, kernels.cuh:219-271, row-major.  Terms that are +0 leave
 * the entry unchanged and are skipped, as in the FP64 oracle. */
void orc32_jac(const orc_model *m, const float *x, float *J)
{
    const int S = m->S;
    memset(J, 0, sizeof(float) * (size_t)S * S);
    for (int i = 0; i < S; i++) {
        float *row = J + (size_t)i * S;
        for (int e = m->hptr[i]; e < m->hptr[i + 1]; e++) {
            int r = m->hrx[e];
            int c = m->hcoef[e];
            float kf = (float)m->k[r];
            int a0 = m->lptr[r], a1 = m->lptr[r + 1];
            for (int j = a0; j < a1; j++) {
                int g = m->lsp[j];
                int ex = m->lexp[j];
                double temp = 0.0;
                if (ex == 1) {
                    temp = temp + (float)c * kf;
                } else if (ex > 1) {
                    temp = temp + ((float)c * kf) * (float)ex * mypow_f(x[g], ex - 1);
                } else {
                    continue;
                }
                for (int jj = a0; jj < a1; jj++)
                    if (m->lsp[jj] != g) temp = temp * x[m->lsp[jj]];
                row[g] = (float)(row[g] + temp);
            }
        }
    }
}

/* aggiorna_Jacobiano1<float>: M = I - h*J, float */
void orc32_newton_matrix(int S, float h, float *J)
{
    for (int i = 0; i < S; i++)
        for (int g = 0; g < S; g++) {
            float id = (i == g) ? 1.0f : 0.0f;
            J[(size_t)i * S + g] = id - h * J[(size_t)i * S + g];
        }
}

/* getrf in float: the FP64 oracle's elimination (first maximum of |a|,
 * multipliers by division, a_ij = fmaf(-l_ik, u_kj, a_ij) with k ascending) */
int orc32_lu(int n, float *A, int32_t *piv)
{
    for (int k = 0; k < n; k++) {
        int p = k;
        float bv = fabsf(A[(size_t)k * n + k]);
        for (int i = k + 1; i < n; i++) {
            float v = fabsf(A[(size_t)i * n + k]);
            if (v > bv) { bv = v; p = i; }
        }
        piv[k] = p;
        if (A[(size_t)p * n + k] == 0.0f) return k + 1;
        if (p != k) {
            for (int j = 0; j < n; j++) {
                float t = A[(size_t)k * n + j];
                A[(size_t)k * n + j] = A[(size_t)p * n + j];
                A[(size_t)p * n + j] = t;
            }
        }
        const float ukk = A[(size_t)k * n + k];
        for (int i = k + 1; i < n; i++) A[(size_t)i * n + k] = A[(size_t)i * n + k] / ukk;
        for (int i = k + 1; i < n; i++) {
            const float l = A[(size_t)i * n + k];
            float *ri = A + (size_t)i * n;
            const float *rk = A + (size_t)k * n;
            for (int j = k + 1; j < n; j++) ri[j] = fmaf(-l, rk[j], ri[j]);
        }
    }
    return 0;
}

/* getrs in float, one right-hand side */
void orc32_lu_solve(int n, const float *LU, const int32_t *piv, float *b)
{
    for (int k = 0; k < n; k++) {
        int p = piv[k];
        if (p != k) { float t = b[k]; b[k] = b[p]; b[p] = t; }
    }
    for (int k = 0; k < n; k++) {
        const float yk = b[k];
        for (int i = k + 1; i < n; i++) b[i] = fmaf(-LU[(size_t)i * n + k], yk, b[i]);
    }
    for (int k = n - 1; k >= 0; k--) {
        b[k] = b[k] / LU[(size_t)k * n + k];
        const float xk = b[k];
        for (int i = 0; i < k; i++) b[i] = fmaf(-LU[(size_t)i * n + k], xk, b[i]);
    }
}

/* transform_reduce(square, plus<float>) in the fixed lane order */
float orc32_normsq(int n, const float *d)
{
    float lane[NORM_LANES];
    for (int l = 0; l < NORM_LANES; l++) {
        float s = 0.0f;
        for (int i = l; i < n; i += NORM_LANES) s = s + d[i] * d[i];
        lane[l] = s;
    }
    for (int stride = NORM_LANES / 2; stride > 0; stride >>= 1)
        for (int l = 0; l < stride; l++) lane[l] = lane[l] + lane[l + stride];
    return lane[0];
}

/* thrust::min_element over the factors widened to double (Simulation.cuh:753) */
static double min_first_f(int n, const float *d)
{
    double m = d[0];
    for (int i = 1; i < n; i++)
        if ((double)d[i] < m) m = d[i];
    return m;
}

typedef struct {
    float *k1, *k2, *k3, *k4, *k5, *k6, *u, *w1, *w2, *delta;
} rkf32_ws;

static void atol_f(const orc_model *m, float *a)
{
    for (int s = 0; s < m->S; s++) a[s] = (float)m->atol[s];
}

/* One RKF45 attempt in float (Simulation.cuh:692-740, kernels.cuh:68-199) */
static int rkf32_attempt(const orc_model *m, const float *fd, const float *at, double rtol, float dt,
                         const float *y1, rkf32_ws *w)
{
    const int S = m->S;
    orc32_rhs(m, y1, w->k1);
    for (int s = 0; s < S; s++) {
        if (fd[s] != 0.0) { w->k1[s] = fd[s]; w->u[s] = fd[s]; }
        else { w->k1[s] = w->k1[s] * dt; w->u[s] = (float)(y1[s] + w->k1[s] / 4.0); }
    }
    orc32_rhs(m, w->u, w->k2);
    for (int s = 0; s < S; s++) {
        if (fd[s] != 0.0) { w->k2[s] = fd[s]; w->u[s] = fd[s]; }
        else {
            w->k2[s] = w->k2[s] * dt;
            w->u[s] = (float)(y1[s] + (3.0 * w->k1[s]) / 32.0 + (9.0 * w->k2[s]) / 32.0);
        }
    }
    orc32_rhs(m, w->u, w->k3);
    for (int s = 0; s < S; s++) {
        if (fd[s] != 0.0) { w->k3[s] = fd[s]; w->u[s] = fd[s]; }
        else {
            w->k3[s] = w->k3[s] * dt;
            w->u[s] = (float)(y1[s] + (1932.0 * w->k1[s]) / 2197.0 - (7200.0 * w->k2[s]) / 2197.0
                              + (7296.0 * w->k3[s]) / 2197.0);
        }
    }
    orc32_rhs(m, w->u, w->k4);
    for (int s = 0; s < S; s++) {
        if (fd[s] != 0.0) { w->k4[s] = fd[s]; w->u[s] = fd[s]; }
        else {
            w->k4[s] = w->k4[s] * dt;
            w->u[s] = (float)(y1[s] + (439.0 * w->k1[s]) / 216.0 - (8.0 * w->k2[s])
                              + (3680.0 * w->k3[s]) / 513.0 - (845.0 * w->k4[s]) / 4104.0);
        }
    }
    orc32_rhs(m, w->u, w->k5);
    for (int s = 0; s < S; s++) {
        if (fd[s] != 0.0) { w->k5[s] = fd[s]; w->u[s] = fd[s]; }
        else {
            w->k5[s] = w->k5[s] * dt;
            w->u[s] = (float)(y1[s] - (8.0 * w->k1[s]) / 27.0 + (2.0 * w->k2[s])
                              - (3544.0 * w->k3[s]) / 2565.0 + (1859.0 * w->k4[s]) / 4104.0
                              - (11.0 * w->k5[s]) / 40.0);
        }
    }
    orc32_rhs(m, w->u, w->k6);
    int nok = 0;
    for (int s = 0; s < S; s++) {
        float w1, w2;
        if (fd[s] != 0.0) { w->k6[s] = fd[s]; w1 = fd[s]; w2 = fd[s]; }
        else {
            w->k6[s] = w->k6[s] * dt;
            w1 = (float)(y1[s] + (25.0 * w->k1[s]) / 216.0 + (1408.0 * w->k3[s]) / 2565.0
                         + (2197.0 * w->k4[s]) / 4104.0 - w->k5[s] / 5.0);
            w2 = (float)(y1[s] + (16.0 * w->k1[s]) / 135.0 + (6656.0 * w->k3[s]) / 12825.0
                         + (28561.0 * w->k4[s]) / 56430.0 - (9.0 * w->k5[s] / 50.0)
                         + (2.0 * w->k6[s]) / 55.0);
        }
        w->w1[s] = w1;
        w->w2[s] = w2;
        float R = fabsf(w1 - w2) / dt;
        float tol = (rtol > 0.0) ? (float)(at[s] + rtol * fabsf(w1)) : at[s];
        w->delta[s] = (float)(0.84 * sqrtf(sqrtf(tol / R)));
        if (R <= tol && !isnan(R) && !isinf(R) && w1 > 0) nok++;
    }
    return nok;
}

/* one attempt with the model's own feed and atol, for known-answer tests */
int orc32_rkf_attempt(const orc_model *m, double rtol, float dt, const float *y1, float *w1, float *w2,
                      float *delta)
{
    const int S = m->S;
    float *buf = (float *)malloc(sizeof(float) * (size_t)S * 9);
    float *fd = buf + 7 * S, *at = buf + 8 * S;
    for (int s = 0; s < S; s++) fd[s] = (float)m->feed[s];
    atol_f(m, at);
    rkf32_ws w = {buf, buf + S, buf + 2 * S, buf + 3 * S, buf + 4 * S, buf + 5 * S, buf + 6 * S, w1, w2, delta};
    int nok = rkf32_attempt(m, fd, at, rtol, dt, y1, &w);
    free(buf);
    return nok;
}

static int bad_f(float v) { return isnan(v) || isinf(v) || v < 0; }

/* BDF1 step by modified Newton in float (Simulation.cuh:780-919) */
static float bdf32_step(const orc_model *m, const float *fd, const orc_opts *o, float t, float t1,
                        float *newdt, float *y1, float *J, float *x, float *yv, float *yv1, int32_t *piv,
                        orc_stats *st)
{
    const int S = m->S;
    const float newton_tol = (float)o->newton_tol;
    const float be_step = (float)o->be_step;
    float dt = *newdt;
    if ((t + dt) > t1) dt = fabsf(t1 - t);
    memcpy(yv, y1, sizeof(float) * S);
    int itN = 0;
    float errN = 1000;
    int luFail = 0;
    int info = 0;
    while (itN < o->newton_iter && errN > newton_tol) {
        orc32_rhs(m, yv, x);
        for (int s = 0; s < S; s++) x[s] = -(yv[s] - y1[s] - dt * x[s]);
        if (itN == 0) {
            orc32_jac(m, yv, J);
            orc32_newton_matrix(S, dt, J);
            info = orc32_lu(S, J, piv);
        }
        if (info != 0) {
            for (int s = 0; s < S; s++)
                if (bad_f(yv[s])) yv[s] = y1[s];
            *newdt = (float)(dt * 0.1);
            luFail = 1;
            st->lu_failures++;
            break;
        }
        orc32_lu_solve(S, J, piv, x);
        for (int s = 0; s < S; s++) {
            yv1[s] = yv[s];
            if (fd[s] == 0) yv[s] = yv[s] + x[s];
            else yv[s] = fd[s];
            x[s] = yv[s] - yv1[s];
        }
        errN = sqrtf(orc32_normsq(S, x));
        itN = itN + 1;
    }
    st->newton_iters += itN;
    if (!luFail) {
        int bad = 0;
        for (int s = 0; s < S; s++) bad += bad_f(yv[s]);
        if (bad > 0) {
            for (int s = 0; s < S; s++)
                if (bad_f(yv[s])) yv[s] = y1[s];
            *newdt = (float)(dt * 0.1);
            st->nan_repairs++;
        } else if (o->bdf_regrow) {
            *newdt = fminf(*newdt * 2.0f, be_step);
        }
    }
    memcpy(y1, yv, sizeof(float) * S);
    return dt;
}

/* Simulation<float>::run main loop (Simulation.cuh:632-925).  Same arguments
 * and outputs as orc_solve; out, y_final and progress hold float values. */
int orc32_solve(const orc_model *m, const orc_opts *o, const double *tv, int nt, const int32_t *cs,
                int ncs, double *out, double *y_final, double *progress, int64_t progress_cap,
                orc_stats *st)
{
    const int S = m->S;
    orc_stats local;
    if (!st) st = &local;
    memset(st, 0, sizeof(*st));
    float *buf = (float *)malloc(sizeof(float) * (size_t)S * 18);
    float *y1 = buf, *fd = buf + S, *at = buf + 2 * S;
    rkf32_ws rw = {buf + 3 * S, buf + 4 * S, buf + 5 * S, buf + 6 * S, buf + 7 * S, buf + 8 * S,
                   buf + 9 * S, buf + 10 * S, buf + 11 * S, buf + 12 * S};
    float *x = buf + 13 * S, *yv = buf + 14 * S, *yv1 = buf + 15 * S;
    float *J = (float *)malloc(sizeof(float) * (size_t)S * S);
    int32_t *piv = (int32_t *)malloc(sizeof(int32_t) * S);
    float *tvf = (float *)malloc(sizeof(float) * (size_t)nt);
    for (int s = 0; s < S; s++) {
        y1[s] = (float)m->x0[s];
        fd[s] = (float)m->feed[s];
    }
    atol_f(m, at);
    for (int i = 0; i < nt; i++) tvf[i] = (float)tv[i];
    const float stiffness_tol = (float)o->stiffness_tol;

    int i = 0;
    float t1 = tvf[0];
    float t = 0;
    float dt = (float)o->rkf_step;
    float newdt = (float)o->be_step;
    const double t_max = tvf[nt - 1];
    int rc = 0;
    int64_t nrej = 0;
    while (i < nt) {
        if (o->max_iters > 0 && st->loop_iters >= o->max_iters) { rc = 1; break; }
        if (o->max_newton > 0 && st->newton_iters >= o->max_newton) { rc = 1; break; }
        if (progress && st->loop_iters < progress_cap) progress[st->loop_iters] = t / t_max;
        st->loop_iters++;
        if (fabsf(t1 - t) < 1e-5) {
            for (int j = 0; j < ncs; j++) out[(size_t)i * ncs + j] = y1[cs[j]];
            i++;
            st->samples = i;
            if (i == nt) break;
            t1 = tvf[i];
        }
        st->rkf_attempts++;
        int nok = rkf32_attempt(m, fd, at, o->rtol, dt, y1, &rw);
        float mind = (float)min_first_f(S, rw.delta);
        if (nok == S) {
            nrej = 0;
            st->rkf_accepts++;
            memcpy(y1, rw.w1, sizeof(float) * S);
            float d_temp = dt * mind;
            if ((t + d_temp) > t1) dt = fabsf(t1 - t);
            else dt = d_temp;
            t = t + dt;
        } else {
            float dt_old = dt;
            dt = dt * mind;
            nrej++;
            if (dt < stiffness_tol || isnan(dt) || isinf(dt) || fabsf(dt_old - dt) < 1e-3
                || (o->max_rkf_retries > 0 && nrej >= o->max_rkf_retries)) {
                nrej = 0;
                st->bdf_steps++;
                dt = bdf32_step(m, fd, o, t, t1, &newdt, y1, J, x, yv, yv1, piv, st);
                t = t + dt;
            }
        }
    }
    if (y_final)
        for (int s = 0; s < S; s++) y_final[s] = y1[s];
    st->t_final = t;
    st->dt_final = dt;
    st->newdt_final = newdt;
    free(buf);
    free(J);
    free(piv);
    free(tvf);
    return rc;
}
