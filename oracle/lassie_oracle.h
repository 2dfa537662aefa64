/*
 * lassie_oracle.h -- types and entry points of the oracle (CPU FP64 and
 * FP32 restatements of LASSIE's deterministic integrator and of its discrete
 * adjoint).
 *
 * TEST INFRASTRUCTURE ONLY: used by tests/, __graft_entry__.smoke() and
 * bench.py's cpu_baseline leg, as the checker.  The product never calls it.
 */
#ifndef LASSIE_ORACLE_H
#define LASSIE_ORACLE_H

#include <stdint.h>

typedef struct {
    int32_t S, R;
    const int32_t *lptr, *lsp, *lexp;   /* reactants: reaction-major CSR */
    const int32_t *hptr, *hrx, *hcoef;  /* net stoichiometry: species-major CSR */
    const double *k, *x0, *feed, *atol;
} orc_model;

typedef struct {
    double rkf_step;      /* Simulation.cuh:281-288, default 1e-3 */
    double be_step;       /* Simulation.cuh:272-279, default 1e-2 */
    double newton_tol;    /* Simulation.cuh:299-306, default 1e-6 */
    double stiffness_tol; /* Simulation.cuh:308-315, default 1e-6 */
    double rtol;          /* extension; 0 = reference behaviour */
    int32_t newton_iter;  /* Simulation.cuh:290-297, default 10000 */
    int32_t bdf_regrow;   /* extension: 1 = BDF step doubles back towards be_step */
    int32_t max_rkf_retries; /* extension: >0 = this many rejections in a row force BDF */
    int32_t pad_;
    int64_t max_iters;    /* safety cap on loop iterations, 0 = none */
    double refine_tol;    /* precision form: refinement stop, |r_i| <= tol*(|rhs_i| + (|M||x|)_i + atol_i) */
    int32_t refine;       /* precision form: 1 = FP32 LU + FP64 refinement, 0 = FP64 LU */
    int32_t refine_max_iter; /* refinement sweeps before the FP64 fallback */
    int64_t max_newton;   /* extension: cap on the solve's Newton iterations, checked with
                           * max_iters at the top of the loop; 0 = none */
} orc_opts;

typedef struct {
    int64_t loop_iters, rkf_attempts, rkf_accepts, bdf_steps, newton_iters;
    int64_t lu_failures, nan_repairs, samples;
    double t_final, dt_final, newdt_final;
    int64_t refine_iters, refine_fallbacks;
} orc_stats;

/* Record of the state-changing steps of one solve (the discrete map the
 * adjoint differentiates).  Step n maps state n to state n+1:
 *   kind[n] == ORC_STEP_RKF   accepted RKF45 attempt with step dt[n];
 *   kind[n] >= 0              BDF1 step of size dt[n] solved by kind[n]
 *                             modified-Newton iterations;
 *   kind[n] == ORC_STEP_LUFAIL BDF1 step whose LU found a zero pivot (the
 *                             state is unchanged).
 * Rejected RKF attempts change only the step size and are not steps.
 * sample_step[i] = number of steps taken when sample i was stored. */
#define ORC_STEP_RKF (-1)
#define ORC_STEP_LUFAIL (-2)
typedef struct {
    int64_t cap;
    int64_t n;
    double *dt;
    int32_t *kind;
    int32_t *sample_step;
    int32_t overflow;
    int32_t pad_;
} orc_tape;

typedef struct {
    double *k1, *k2, *k3, *k4, *k5, *k6, *u, *w1, *w2, *delta;
} orc_rkf_ws;

typedef struct {
    double *J, *x, *yv, *yv1;
    int32_t *piv;
    float *J32, *b32;     /* precision form only */
    double *xr, *r;
} orc_bdf_ws;

double orc_mypow(double base, int ex);
void orc_rhs(const orc_model *m, const double *x, double *f);
void orc_jac(const orc_model *m, const double *x, double *J);
void orc_newton_matrix(int S, double h, double *J);
int orc_lu(int n, double *A, int32_t *piv);
void orc_lu_solve(int n, const double *LU, const int32_t *piv, double *b);
int orc_refine_solve(int n, const double *M, const float *LU32, const int32_t *piv, const double *rhs,
                     const double *floor, double *x, double *r, float *b32, double tol, int max_it,
                     int64_t *iters);
double orc_normsq(int n, const double *d);
double orc_min_first(int n, const double *d);
int orc_rkf_ws_attempt(const orc_model *m, double rtol, double dt, const double *y1, orc_rkf_ws *w);
int orc_repair(int S, double *yv, const double *y1);
int orc_count_bad(int S, const double *yv);
int orc_solve_rec(const orc_model *m, const orc_opts *o, const double *tv, int nt, const int32_t *cs,
                  int ncs, double *out, double *y_final, double *progress, int64_t progress_cap,
                  orc_stats *st, orc_tape *tape);

/* Single-precision loop, Simulation<float> (lassie_oracle_f32.c). */
void orc32_rhs(const orc_model *m, const float *x, float *f);
void orc32_jac(const orc_model *m, const float *x, float *J);
void orc32_newton_matrix(int S, float h, float *J);
int orc32_lu(int n, float *A, int32_t *piv);
void orc32_lu_solve(int n, const float *LU, const int32_t *piv, float *b);
float orc32_normsq(int n, const float *d);
int orc32_rkf_attempt(const orc_model *m, double rtol, float dt, const float *y1, float *w1, float *w2,
                      float *delta);
int orc32_solve(const orc_model *m, const orc_opts *o, const double *tv, int nt, const int32_t *cs,
                int ncs, double *out, double *y_final, double *progress, int64_t progress_cap,
                orc_stats *st);

/* Discrete adjoint (lassie_oracle_adjoint.c). */
int orc_replay(const orc_model *m, const orc_opts *o, const orc_tape *tape, int nt, double *out_all);
int orc_tangent(const orc_model *m, const orc_opts *o, const orc_tape *tape, int nt,
                const double *dlogk, const double *dx0, double *dout_all);
int orc_adjoint(const orc_model *m, const orc_opts *o, const orc_tape *tape, int nt,
                const double *gout_all, double *grad_logk, double *grad_x0);

#endif
