/*
 * lassie_oracle_adjoint.c -- CPU FP64 restatement of the discrete adjoint of
 * LASSIE's integration loop, and of its forward (tangent) sensitivities.
 *
 * TEST INFRASTRUCTURE ONLY: loaded by tests/ as the checker; the product
 * never calls it.
 *
 * The reference has no derivatives; what is differentiated is the discrete
 * map its loop computes (Simulation.cuh:632-925), with the step sequence of
 * one recorded solve held fixed (orc_tape: step kinds, step sizes, Newton
 * iteration counts).  Parameters are theta_r = log k_r and the initial state.
 *
 *   RKF45 step  (kernels.cuh:68-182): y' = w1(y) = y + 25/216 k1 + 1408/2565 k3
 *     + 2197/4104 k4 - 1/5 k5, k_i = dt f(u_i), stages u_i as in rk_Fehlberg1..5.
 *   BDF1 step   (Simulation.cuh:780-919): modified Newton with M = I - h Jq(y),
 *     Jq the reference's analytic Jacobian (createJac_, kernels.cuh:219-271)
 *     at the step's start state, run for the recorded number of iterations:
 *       r_j = -(v_j - y - h f(v_j)),  x_j = M^-1 r_j,  v_{j+1} = v_j + x_j,
 *     v_0 = y; entries that end NaN, infinite or negative are reset to y
 *     (aggiornaNaN, kernels.cuh:512-521).  The iterations are differentiated
 *     as run (including the dependence of M on y and theta), so the adjoint is
 *     the exact derivative of the computed states, which central differences
 *     of orc_replay reproduce.
 *   Feed species (M_feed != 0) hold the feed value in every stage and
 *     iterate, so nothing flows through them.
 *
 * Mass action: rate_r(x) = k_r prod_a x_a^{e_a}, f = H rate, so
 *   d f / d theta_r = H[:, r] rate_r and (J_f^T z)_g = sum_r c_r d rate_r / d x_g
 *   with c = H^T z.  Jq differs from J_f only where a reaction has another
 *   reactant of order >= 2 (createJac_ multiplies the other reactants without
 *   their exponents); its derivatives below follow createJac_ as written.
 */
#include <math.h>
#include <stdlib.h>
#include <string.h>

#include "lassie_oracle.h"

static double rate(const orc_model *m, const double *x, int r)
{
    double p = m->k[r];
    for (int j = m->lptr[r]; j < m->lptr[r + 1]; j++) p *= orc_mypow(x[m->lsp[j]], m->lexp[j]);
    return p;
}

/* d rate_r / d x_g for the reactant entry jg of reaction r. */
static double drate(const orc_model *m, const double *x, int r, int jg)
{
    const int e = m->lexp[jg];
    double p = m->k[r] * (double)e * orc_mypow(x[m->lsp[jg]], e - 1);
    for (int j = m->lptr[r]; j < m->lptr[r + 1]; j++)
        if (j != jg) p *= orc_mypow(x[m->lsp[j]], m->lexp[j]);
    return p;
}

/* createJac_'s term of reaction r in column g = lsp[jg] (without the net
 * coefficient): k e_g x_g^(e_g-1) prod_{a != g} x_a. */
static double qterm(const orc_model *m, const double *x, int r, int jg)
{
    const int e = m->lexp[jg];
    double p = m->k[r] * (double)e * orc_mypow(x[m->lsp[jg]], e - 1);
    for (int j = m->lptr[r]; j < m->lptr[r + 1]; j++)
        if (j != jg) p *= x[m->lsp[j]];
    return p;
}

/* d qterm(r, jg) / d x_b, b = lsp[jb]. */
static double dqterm(const orc_model *m, const double *x, int r, int jg, int jb)
{
    const int e = m->lexp[jg];
    double p;
    if (jb == jg) {
        if (e < 2) return 0.0;
        p = m->k[r] * (double)e * (double)(e - 1) * orc_mypow(x[m->lsp[jg]], e - 2);
        for (int j = m->lptr[r]; j < m->lptr[r + 1]; j++)
            if (j != jg) p *= x[m->lsp[j]];
    } else {
        p = m->k[r] * (double)e * orc_mypow(x[m->lsp[jg]], e - 1);
        for (int j = m->lptr[r]; j < m->lptr[r + 1]; j++)
            if (j != jg && j != jb) p *= x[m->lsp[j]];
    }
    return p;
}

/* c = H^T z (per reaction). */
static void ht(const orc_model *m, const double *z, double *c)
{
    memset(c, 0, sizeof(double) * m->R);
    for (int s = 0; s < m->S; s++)
        for (int e = m->hptr[s]; e < m->hptr[s + 1]; e++) c[m->hrx[e]] += (double)m->hcoef[e] * z[s];
}

/* out = H w. */
static void hmul(const orc_model *m, const double *w, double *out)
{
    for (int s = 0; s < m->S; s++) {
        double a = 0.0;
        for (int e = m->hptr[s]; e < m->hptr[s + 1]; e++) a += (double)m->hcoef[e] * w[m->hrx[e]];
        out[s] = a;
    }
}

/* out += scale * J_f(x)^T z ; gth += scale * (df/dtheta)^T z.  c: R scratch. */
static void vjp_f(const orc_model *m, const double *x, const double *z, double scale, double *out,
                  double *gth, double *c)
{
    ht(m, z, c);
    for (int r = 0; r < m->R; r++) {
        if (c[r] == 0.0) continue;
        for (int j = m->lptr[r]; j < m->lptr[r + 1]; j++) out[m->lsp[j]] += scale * c[r] * drate(m, x, r, j);
        if (gth) gth[r] += scale * c[r] * rate(m, x, r);
    }
}

/* out = J_f(x) v + (df/dtheta) dth.  w: R scratch. */
static void jvp_f(const orc_model *m, const double *x, const double *v, const double *dth, double *out,
                  double *w)
{
    for (int r = 0; r < m->R; r++) {
        double a = 0.0;
        for (int j = m->lptr[r]; j < m->lptr[r + 1]; j++) a += drate(m, x, r, j) * v[m->lsp[j]];
        if (dth) a += rate(m, x, r) * dth[r];
        w[r] = a;
    }
    hmul(m, w, out);
}

/* phi(y, theta) = z^T Jq(y, theta) x.  gy += scale * d phi / d y,
 * gth += scale * d phi / d theta. */
static void vjp_jq(const orc_model *m, const double *y, const double *z, const double *x, double scale,
                   double *gy, double *gth, double *c)
{
    ht(m, z, c);
    for (int r = 0; r < m->R; r++) {
        if (c[r] == 0.0) continue;
        double th = 0.0;
        for (int jg = m->lptr[r]; jg < m->lptr[r + 1]; jg++) {
            const double xg = x[m->lsp[jg]];
            th += qterm(m, y, r, jg) * xg;
            for (int jb = m->lptr[r]; jb < m->lptr[r + 1]; jb++)
                gy[m->lsp[jb]] += scale * c[r] * xg * dqterm(m, y, r, jg, jb);
        }
        gth[r] += scale * c[r] * th;
    }
}

/* out = (d Jq along (dy, dth)) x. */
static void jvp_jq(const orc_model *m, const double *y, const double *x, const double *dy, const double *dth,
                   double *out, double *w)
{
    for (int r = 0; r < m->R; r++) {
        double a = 0.0;
        for (int jg = m->lptr[r]; jg < m->lptr[r + 1]; jg++) {
            double d = qterm(m, y, r, jg) * dth[r];
            for (int jb = m->lptr[r]; jb < m->lptr[r + 1]; jb++) d += dqterm(m, y, r, jg, jb) * dy[m->lsp[jb]];
            a += d * x[m->lsp[jg]];
        }
        w[r] = a;
    }
    hmul(m, w, out);
}

/* Solve M^T z = b with the factors of orc_lu (P M = L U): U^T w = b, L^T u = w,
 * z = P^T u.  b is overwritten with z. */
static void lu_solve_t(int n, const double *LU, const int32_t *piv, double *b)
{
    for (int k = 0; k < n; k++) {
        b[k] = b[k] / LU[(size_t)k * n + k];
        const double wk = b[k];
        for (int i = k + 1; i < n; i++) b[i] -= LU[(size_t)k * n + i] * wk;
    }
    for (int k = n - 1; k >= 0; k--) {
        const double uk = b[k];
        for (int i = 0; i < k; i++) b[i] -= LU[(size_t)k * n + i] * uk;
    }
    for (int k = n - 1; k >= 0; k--) {
        const int p = piv[k];
        if (p != k) { double t = b[k]; b[k] = b[p]; b[p] = t; }
    }
}

typedef struct {
    int S, R;
    orc_rkf_ws rw;
    double *u[6];         /* stage states u_1..u_5 (u[1..5]) */
    double *J;            /* S x S factors */
    int32_t *piv;
    double *v, *x;        /* Newton iterates v_0..v_m and corrections x_0..x_{m-1} */
    int vcap;
    double *a, *b, *c, *d, *e, *kb[6], *rs;  /* S scratch (kb: stage adjoints), rs: R scratch */
} ws_t;

static int ws_init(ws_t *w, const orc_model *m, int vcap)
{
    const int S = m->S, R = m->R;
    w->S = S;
    w->R = R;
    w->vcap = vcap;
    double *buf = (double *)calloc((size_t)S * 30 + (size_t)R, sizeof(double));
    w->J = (double *)malloc(sizeof(double) * (size_t)S * S);
    w->piv = (int32_t *)malloc(sizeof(int32_t) * S);
    w->v = (double *)malloc(sizeof(double) * (size_t)S * (vcap + 1));
    w->x = (double *)malloc(sizeof(double) * (size_t)S * (vcap + 1));
    if (!buf || !w->J || !w->piv || !w->v || !w->x) return -1;
    double *p = buf;
    w->rw.k1 = p; p += S; w->rw.k2 = p; p += S; w->rw.k3 = p; p += S; w->rw.k4 = p; p += S;
    w->rw.k5 = p; p += S; w->rw.k6 = p; p += S; w->rw.u = p; p += S; w->rw.w1 = p; p += S;
    w->rw.w2 = p; p += S; w->rw.delta = p; p += S;
    for (int i = 1; i <= 5; i++) { w->u[i] = p; p += S; }
    w->a = p; p += S; w->b = p; p += S; w->c = p; p += S; w->d = p; p += S; w->e = p; p += S;
    for (int i = 1; i <= 5; i++) { w->kb[i] = p; p += S; }
    w->rs = p;
    w->u[0] = buf;  /* keeps the allocation's base for ws_free */
    return 0;
}

static void ws_free(ws_t *w)
{
    free(w->u[0]);
    free(w->J);
    free(w->piv);
    free(w->v);
    free(w->x);
}

/* Stage states of one RKF45 attempt from y (the expressions of
 * rk_Fehlberg1..5, kernels.cuh:68-159, with the stage k's in w->rw). */
static void rkf_stages(const orc_model *m, double rtol, double dt, const double *y, ws_t *w)
{
    const int S = m->S;
    const double *fd = m->feed;
    orc_rkf_ws_attempt(m, rtol, dt, y, &w->rw);
    const double *k1 = w->rw.k1, *k2 = w->rw.k2, *k3 = w->rw.k3, *k4 = w->rw.k4;
    for (int s = 0; s < S; s++) {
        w->u[1][s] = y[s];
        if (fd[s] != 0.0) {
            w->u[2][s] = w->u[3][s] = w->u[4][s] = w->u[5][s] = fd[s];
            continue;
        }
        w->u[2][s] = y[s] + k1[s] / 4.0;
        w->u[3][s] = y[s] + (3.0 * k1[s]) / 32.0 + (9.0 * k2[s]) / 32.0;
        w->u[4][s] = y[s] + (1932.0 * k1[s]) / 2197.0 - (7200.0 * k2[s]) / 2197.0 + (7296.0 * k3[s]) / 2197.0;
        w->u[5][s] = y[s] + (439.0 * k1[s]) / 216.0 - (8.0 * k2[s]) + (3680.0 * k3[s]) / 513.0
                     - (845.0 * k4[s]) / 4104.0;
    }
}

/* Butcher coefficients a_ij (stage i uses k_j) of the stages above. */
static const double A_[6][6] = {
    {0},
    {0},
    {0, 1.0 / 4.0},
    {0, 3.0 / 32.0, 9.0 / 32.0},
    {0, 1932.0 / 2197.0, -7200.0 / 2197.0, 7296.0 / 2197.0},
    {0, 439.0 / 216.0, -8.0, 3680.0 / 513.0, -845.0 / 4104.0},
};
static const double B_[6] = {0, 25.0 / 216.0, 0.0, 1408.0 / 2565.0, 2197.0 / 4104.0, -1.0 / 5.0};

/* Newton iterations of one BDF1 step from y (v_0 = y): fills w->v, w->x and the
 * factors; returns the final iterate (repaired) in y_out. -1 if the LU fails. */
static int bdf_iterates(const orc_model *m, double h, int iters, const double *y, ws_t *w, double *y_out)
{
    const int S = m->S;
    memcpy(w->v, y, sizeof(double) * S);
    orc_jac(m, y, w->J);
    orc_newton_matrix(S, h, w->J);
    if (orc_lu(S, w->J, w->piv) != 0) return -1;
    for (int j = 0; j < iters; j++) {
        const double *vj = w->v + (size_t)j * S;
        double *vn = w->v + (size_t)(j + 1) * S;
        double *xj = w->x + (size_t)j * S;
        orc_rhs(m, vj, xj);
        for (int s = 0; s < S; s++) xj[s] = -(vj[s] - y[s] - h * xj[s]);
        orc_lu_solve(S, w->J, w->piv, xj);
        for (int s = 0; s < S; s++) vn[s] = (m->feed[s] == 0) ? vj[s] + xj[s] : m->feed[s];
    }
    const double *vm = w->v + (size_t)iters * S;
    for (int s = 0; s < S; s++) {
        const double a = vm[s];
        y_out[s] = (isnan(a) || isinf(a) || a < 0) ? y[s] : a;
    }
    return 0;
}

static int max_iters_of(const orc_tape *t)
{
    int mx = 0;
    for (int64_t n = 0; n < t->n; n++)
        if (t->kind[n] > mx) mx = t->kind[n];
    return mx;
}

/* One recorded step: y -> y_out (y_out may alias nothing). */
static int replay_step(const orc_model *m, const orc_opts *o, const orc_tape *tape, int64_t n, const double *y,
                       double *y_out, ws_t *w)
{
    const int S = m->S;
    const int kind = tape->kind[n];
    const double dt = tape->dt[n];
    if (kind == ORC_STEP_RKF) {
        orc_rkf_ws_attempt(m, o->rtol, dt, y, &w->rw);
        memcpy(y_out, w->rw.w1, sizeof(double) * S);
    } else if (kind == ORC_STEP_LUFAIL) {
        memcpy(y_out, y, sizeof(double) * S);
    } else if (bdf_iterates(m, dt, kind, y, w, y_out) != 0) {
        return -1;
    }
    return 0;
}

/* Replays the recorded steps with the model's (possibly changed) k and x0;
 * out_all[i*S + s] = state at sample i.  Returns 0, or -1 when an LU that
 * succeeded in the recorded solve fails. */
int orc_replay(const orc_model *m, const orc_opts *o, const orc_tape *tape, int nt, double *out_all)
{
    const int S = m->S;
    ws_t w;
    if (ws_init(&w, m, max_iters_of(tape)) != 0) return -2;
    double *y = (double *)malloc(sizeof(double) * S), *yn = (double *)malloc(sizeof(double) * S);
    memcpy(y, m->x0, sizeof(double) * S);
    int i = 0, rc = 0;
    for (int64_t n = 0; n <= tape->n && rc == 0; n++) {
        while (i < nt && tape->sample_step[i] == n) {
            memcpy(out_all + (size_t)i * S, y, sizeof(double) * S);
            i++;
        }
        if (n == tape->n) break;
        rc = replay_step(m, o, tape, n, y, yn, &w);
        double *t = y; y = yn; yn = t;
    }
    free(y);
    free(yn);
    ws_free(&w);
    return rc;
}

/* Forward sensitivities along (dlogk, dx0): dout_all[i*S + s] = d state / d eps
 * at sample i. */
int orc_tangent(const orc_model *m, const orc_opts *o, const orc_tape *tape, int nt, const double *dlogk,
                const double *dx0, double *dout_all)
{
    const int S = m->S;
    const double *fd = m->feed;
    ws_t w;
    if (ws_init(&w, m, max_iters_of(tape)) != 0) return -2;
    double *y = (double *)malloc(sizeof(double) * S), *yn = (double *)malloc(sizeof(double) * S);
    double *dy = (double *)malloc(sizeof(double) * S), *dyn = (double *)malloc(sizeof(double) * S);
    double *dk[6];
    for (int i = 1; i <= 5; i++) dk[i] = (double *)malloc(sizeof(double) * S);
    memcpy(y, m->x0, sizeof(double) * S);
    for (int s = 0; s < S; s++) dy[s] = dx0 ? dx0[s] : 0.0;
    int i = 0, rc = 0;
    for (int64_t n = 0; n <= tape->n && rc == 0; n++) {
        while (i < nt && tape->sample_step[i] == n) {
            memcpy(dout_all + (size_t)i * S, dy, sizeof(double) * S);
            i++;
        }
        if (n == tape->n) break;
        const int kind = tape->kind[n];
        const double dt = tape->dt[n];
        if (kind == ORC_STEP_RKF) {
            rkf_stages(m, o->rtol, dt, y, &w);
            memcpy(yn, w.rw.w1, sizeof(double) * S);
            for (int st = 1; st <= 5; st++) {
                double *du = w.a;
                for (int s = 0; s < S; s++) {
                    double a = dy[s];
                    for (int j = 1; j < st; j++) a += A_[st][j] * dk[j][s];
                    du[s] = (fd[s] != 0.0) ? 0.0 : a;
                }
                jvp_f(m, w.u[st], du, dlogk, dk[st], w.rs);
                for (int s = 0; s < S; s++) dk[st][s] = (fd[s] != 0.0) ? 0.0 : dt * dk[st][s];
            }
            for (int s = 0; s < S; s++) {
                double a = dy[s];
                for (int j = 1; j <= 5; j++) a += B_[j] * dk[j][s];
                dyn[s] = (fd[s] != 0.0) ? 0.0 : a;
            }
        } else if (kind == ORC_STEP_LUFAIL) {
            memcpy(yn, y, sizeof(double) * S);
            memcpy(dyn, dy, sizeof(double) * S);
        } else {
            if (bdf_iterates(m, dt, kind, y, &w, yn) != 0) { rc = -1; break; }
            double *dv = w.a, *dr = w.b, *t2 = w.c;
            memcpy(dv, dy, sizeof(double) * S);
            for (int j = 0; j < kind; j++) {
                const double *vj = w.v + (size_t)j * S;
                const double *xj = w.x + (size_t)j * S;
                jvp_f(m, vj, dv, dlogk, dr, w.rs);
                jvp_jq(m, y, xj, dy, dlogk, t2, w.rs);
                for (int s = 0; s < S; s++) dr[s] = -dv[s] + dy[s] + dt * dr[s] + dt * t2[s];
                orc_lu_solve(S, w.J, w.piv, dr);
                for (int s = 0; s < S; s++) dv[s] = (fd[s] == 0) ? dv[s] + dr[s] : 0.0;
            }
            const double *vm = w.v + (size_t)kind * S;
            for (int s = 0; s < S; s++) {
                const double a = vm[s];
                dyn[s] = (isnan(a) || isinf(a) || a < 0) ? dy[s] : dv[s];
            }
        }
        double *t = y; y = yn; yn = t;
        t = dy; dy = dyn; dyn = t;
    }
    free(y); free(yn); free(dy); free(dyn);
    for (int k = 1; k <= 5; k++) free(dk[k]);
    ws_free(&w);
    return rc;
}

/* Reverse pass: given gout_all[i*S + s] = dL / d(state s at sample i), returns
 * grad_logk[R] = dL / d log k and grad_x0[S] = dL / d x0 (either may be NULL).
 * Every state of the replay is kept (the checker does not checkpoint). */
int orc_adjoint(const orc_model *m, const orc_opts *o, const orc_tape *tape, int nt, const double *gout_all,
                double *grad_logk, double *grad_x0)
{
    const int S = m->S, R = m->R;
    const double *fd = m->feed;
    const int64_t N = tape->n;
    ws_t w;
    if (ws_init(&w, m, max_iters_of(tape)) != 0) return -2;
    double *states = (double *)malloc(sizeof(double) * (size_t)S * (N + 1));
    double *lam = (double *)calloc(S, sizeof(double));
    double *yb = (double *)malloc(sizeof(double) * S);
    double *gth = (double *)calloc(R, sizeof(double));
    if (!states || !lam || !yb || !gth) return -2;
    memcpy(states, m->x0, sizeof(double) * S);
    int rc = 0;
    for (int64_t n = 0; n < N && rc == 0; n++)
        rc = replay_step(m, o, tape, n, states + (size_t)n * S, states + (size_t)(n + 1) * S, &w);
    int i = nt - 1;
    for (int64_t n = N; n >= 0 && rc == 0; n--) {
        while (i >= 0 && tape->sample_step[i] == n) {
            for (int s = 0; s < S; s++) lam[s] += gout_all[(size_t)i * S + s];
            i--;
        }
        if (n == 0) break;
        const int64_t step = n - 1;
        const int kind = tape->kind[step];
        const double dt = tape->dt[step];
        const double *y = states + (size_t)step * S;
        if (kind == ORC_STEP_LUFAIL) continue;
        if (kind == ORC_STEP_RKF) {
            rkf_stages(m, o->rtol, dt, y, &w);
            for (int s = 0; s < S; s++) {
                const double l = (fd[s] != 0.0) ? 0.0 : lam[s];
                yb[s] = l;
                for (int j = 1; j <= 5; j++) w.kb[j][s] = B_[j] * l;
            }
            for (int st = 5; st >= 1; st--) {
                double *ub = w.a;
                memset(ub, 0, sizeof(double) * S);
                vjp_f(m, w.u[st], w.kb[st], dt, ub, gth, w.rs);
                for (int s = 0; s < S; s++) {
                    if (fd[s] != 0.0) continue;
                    yb[s] += ub[s];
                    for (int j = 1; j < st; j++) w.kb[j][s] += A_[st][j] * ub[s];
                }
            }
            memcpy(lam, yb, sizeof(double) * S);
        } else {
            if (bdf_iterates(m, dt, kind, y, &w, w.e) != 0) { rc = -1; break; }
            double *vb = w.a, *z = w.b;
            memset(yb, 0, sizeof(double) * S);
            memcpy(vb, lam, sizeof(double) * S);
            const double *vm = w.v + (size_t)kind * S;
            for (int s = 0; s < S; s++) {
                const double a = vm[s];
                if (isnan(a) || isinf(a) || a < 0) { yb[s] += vb[s]; vb[s] = 0.0; }
            }
            for (int j = kind - 1; j >= 0; j--) {
                const double *vj = w.v + (size_t)j * S;
                const double *xj = w.x + (size_t)j * S;
                for (int s = 0; s < S; s++) {
                    if (fd[s] != 0.0) vb[s] = 0.0;
                    z[s] = vb[s];
                }
                lu_solve_t(S, w.J, w.piv, z);
                vjp_jq(m, y, z, xj, dt, yb, gth, w.rs);
                for (int s = 0; s < S; s++) {
                    vb[s] -= z[s];
                    yb[s] += z[s];
                }
                vjp_f(m, vj, z, dt, vb, gth, w.rs);
            }
            for (int s = 0; s < S; s++) lam[s] = yb[s] + vb[s];
        }
    }
    if (grad_logk) memcpy(grad_logk, gth, sizeof(double) * R);
    if (grad_x0) memcpy(grad_x0, lam, sizeof(double) * S);
    free(states); free(lam); free(yb); free(gth);
    ws_free(&w);
    return rc;
}
