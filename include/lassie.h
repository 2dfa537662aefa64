/*
 * lassie.h -- C ABI of liblassie, an MI355X (gfx950) implementation of LASSIE's
 * deterministic mass-action integrator.
 *
 * The reference has no library interface: its boundary is the process
 * `LASSIE input_folder output_folder [-v1|-v2|-v3] [-p double|float]`
 * (src/c++/simulator.cu:29-319) plus the internal seam
 * `float Simulation<T>::run(...)` (src/c++/header/Simulation.cuh:230-234),
 * whose inputs are the parsed BioSimWare files.  This header exposes that seam
 * as plain C: a model is the parsed, sparse-compiled input
 * (Simulation.cuh:397-418), a solve is the integration loop
 * (Simulation.cuh:632-925) writing the sampled trajectory (`dinamic`,
 * Simulation.cuh:500-533, 942).
 *
 * Conventions: the caller owns every host array; the library copies what it
 * needs and owns all device memory inside a model handle.  No function calls
 * exit(); errors are status codes and lassie_last_error() (thread-local text).
 * A handle is not thread-safe.
 */
#ifndef LASSIE_H
#define LASSIE_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define LASSIE_ABI_VERSION 2

/* Status codes.  The CLI maps them to the reference's exit codes
 * (simulator.cu:61-74, 192-207; Simulation.cuh:38-44, 373-389). */
#define LASSIE_OK 0
#define LASSIE_ERR_INVALID 1   /* bad argument or inconsistent model */
#define LASSIE_ERR_HIP 2       /* HIP runtime failure (reference: exit -13) */
#define LASSIE_ERR_NOMEM 3
#define LASSIE_ERR_MAXITER 4   /* opts.max_iters loop iterations reached */
#define LASSIE_ERR_UNSUPPORTED 5
#define LASSIE_ERR_DEVICE 6    /* a device-side wait exceeded its bound */
#define LASSIE_ERR_IO 7        /* file could not be read or written */
#define LASSIE_ERR_NONFINITE 8 /* a computed result is not finite; the outputs are written (see
                                * lassie_adjoint_status) */

/* Arithmetic of the Newton linear solve.  The refined forms are
 * experimental: each Newton correction meets a componentwise backward-error
 * target of refine_tol, but trajectories are not within 1e-6 of the FP64
 * method's (the loop's step decisions are discontinuous; per-instance median
 * relative errors of about 1e-3 to 5e-2 on the stiff batched workloads).
 * BF16: the panel factorizations run in FP32 and the trailing updates on
 * the BF16 matrix cores (operands rounded to BF16, one 16x16x16 BF16 MFMA
 * per 16 x 16 tile of a 16-column panel, FP32 accumulation). */
#define LASSIE_PREC_FP64 0          /* plain method: FP64 LU (cuSOLVER getrf/getrs seat) */
#define LASSIE_PREC_FP32_REFINE 1   /* FP32 LU + FP64 iterative refinement (experimental) */
#define LASSIE_PREC_BF16_REFINE 2   /* BF16 LU updates + FP64 iterative refinement (experimental) */
/* the whole loop in single precision: `-p float`, Simulation<float>
 * (simulator.cu:259-304); lassie_solve only */
#define LASSIE_PREC_FP32_LOOP 3

typedef struct lassie_model lassie_model;

/* A compiled model (Simulation.cuh:397-418 builds the same data as short4 /
 * short2 arrays; here int32 CSR).
 *   reactants, reaction-major, species ascending (create_sparse2,
 *     Simulation.cuh:1107-1130): react_ptr[R+1], react_species[], react_exp[]
 *   net stoichiometry H = (right - left)^T, species-major, reaction ascending
 *     (createTransposte + create_sparse, Simulation.cuh:1061-1105):
 *     net_ptr[S+1], net_reaction[], net_coef[]
 *   k[R] (c_vector), x0[S] (M_0), feed[S] (M_feed, NULL = all zero),
 *   atol[S] (atol_vector, NULL = 1e-12 each, Simulation.cuh:317-325). */
typedef struct {
    int32_t n_species;
    int32_t n_reactions;
    const int32_t *react_ptr;
    const int32_t *react_species;
    const int32_t *react_exp;
    const int32_t *net_ptr;
    const int32_t *net_reaction;
    const int32_t *net_coef;
    const double *k;
    const double *x0;
    const double *feed;
    const double *atol;
} lassie_model_desc;

/* Integrator options; defaults (lassie_opts_default) are the reference's
 * (Simulation.cuh:271-354). */
typedef struct {
    double rkf_step;       /* initial RKF45 step, file rkf_step, default 1e-3 */
    double be_step;        /* BDF1 step, file be_step, default 1e-2 */
    double newton_tol;     /* file newton_tol, default 1e-6 */
    double stiffness_tol;  /* file stiffness_tol, default 1e-6 */
    double rtol;           /* extension: tol = atol + rtol*|w1| when > 0; default 0 */
    int32_t newton_iter;   /* file newton_iter, default 10000 */
    int32_t bdf_regrow;    /* extension: BDF step doubles back to be_step; default 0 */
    int32_t max_rkf_retries; /* extension: N rejections in a row force BDF; default 0 = off */
    int32_t precision;     /* LASSIE_PREC_*, default LASSIE_PREC_FP64 */
    int64_t max_iters;     /* cap on loop iterations; default 100000000 */
    double refine_tol;     /* refinement stop: |r_i| <= tol*((|M||x|)_i + |b_i| + atol_i) for all i, default 1e-14 */
    int32_t refine_max_iter; /* refinement sweeps before the FP64 fallback, default 4; 0: one FP32
                              * solve per Newton iteration, no residual and no fallback (the
                              * modified-Newton iteration, whose residual is FP64, refines) */
    int32_t reserved;      /* diagnostics: bit 1 keeps the control group's state in global memory */
    int64_t max_newton;    /* extension: cap on the Newton iterations of one solve (one instance of a
                            * batch), checked where max_iters is; 0 (default) = none.  A solve that
                            * reaches it stops like one that reaches max_iters */
} lassie_opts;

typedef struct {
    int64_t loop_iters, rkf_attempts, rkf_accepts, bdf_steps, newton_iters;
    int64_t lu_failures, nan_repairs, samples;
    double t_final, dt_final, newdt_final;
    double kernel_ms;      /* device time of the integration loop */
    int64_t lu_factorizations, refine_iters, refine_fallbacks;
    double phase_ms[4];    /* device time by phase: control+RKF45, Jacobian, LU, Newton */
    int32_t groups;        /* workgroups that cooperated on the solve */
    int32_t lds_model;     /* 1 when the control group kept model and state in LDS */
} lassie_stats;

void lassie_opts_default(lassie_opts *opts);

/* Copies the model to device memory.  Reference: Simulation.cuh:470-576. */
int lassie_model_create(const lassie_model_desc *desc, lassie_model **out);
int lassie_model_destroy(lassie_model *model);

/* Integrates from t = 0 (Simulation.cuh:632-636) and stores the state of
 * species cs[j] at every sampling time t[i] into out[i*ncs + j]
 * (saveDinamic, kernels.cuh:524-535); rows of samples the loop did not reach
 * (LASSIE_ERR_MAXITER) are left unchanged.  y_final[S] (may be NULL) receives the
 * state when the loop ends; progress (may be NULL, capacity progress_cap)
 * receives t/t_max at the top of every loop iteration (ProgressSimulation,
 * Simulation.cuh:660-672).  opts.precision LASSIE_PREC_FP64 runs Simulation<double>,
 * LASSIE_PREC_FP32_LOOP runs Simulation<float> (values widened to double on
 * output), LASSIE_PREC_FP32_REFINE runs the FP64 loop with the Newton matrix
 * factored in FP32 and every Newton solve refined in FP64 (opts.refine_tol,
 * opts.refine_max_iter; a zero FP32 pivot or a refinement that does not
 * converge switches that BDF step to the FP64 LU, counted in
 * stats->refine_fallbacks).  LASSIE_PREC_BF16_REFINE is the same form with the
 * FP32 factorization's trailing updates on the BF16 matrix cores (experimental).
 * Reference: Simulation<T>::run, Simulation.cuh:230-981; the precision form
 * extends linearSolverLU, Simulation.cuh:1280-1326. */
int lassie_solve(lassie_model *model, const lassie_opts *opts, const double *t, int32_t nt,
                 const int32_t *cs, int32_t ncs, double *out, double *y_final,
                 double *progress, int64_t progress_cap, lassie_stats *stats);

/* Batched form: integrates n_models independent models (S <= 512 each) in one
 * persistent launch, every model with its own step control, all sampled at
 * the same times t.  out receives, model after model, nt x S_i values (every
 * species); y_final (may be NULL) the S_i final states, model after model;
 * stats (may be NULL) one record per model; elapsed_ms (may be NULL) the
 * device time of the launch.  opts.precision selects the Newton linear solve
 * (LASSIE_PREC_*).  Rows of samples an instance did not reach (max_iters or
 * max_newton) hold NaN.  No counterpart in the reference, which runs one
 * model per process; each instance follows Simulation<T>::run
 * (Simulation.cuh:632-925). */
int lassie_solve_batch(lassie_model *const *models, int32_t n_models, const lassie_opts *opts,
                       const double *t, int32_t nt, double *out, double *y_final, lassie_stats *stats,
                       double *elapsed_ms);

/* lassie_solve_batch without any device allocation, for callers that cannot
 * allocate (code under HIP graph capture, tight serving loops, harnesses that
 * forbid hipMalloc).  No counterpart in the reference.  Same arguments, results
 * and output layout as lassie_solve_batch, for batches it can hold: FP64
 * (opts.precision LASSIE_PREC_FP64), every model with at most 64 species and
 * at most 2048 net-stoichiometry entries (the wave-level kernel), and the
 * instance table, sampling times, counters and every output together within
 * the library's fixed 64 MiB device buffer (one per GPU, held until the
 * process ends; about 2 + 8 nt S bytes per model plus 1 KB).  Anything else
 * returns LASSIE_ERR_UNSUPPORTED with the size it would need.  All calls on
 * one GPU share that buffer: calls that overlap on one GPU (from several
 * threads or streams) race, so callers serialise them. */
int lassie_solve_batch_static(lassie_model *const *models, int32_t n_models, const lassie_opts *opts,
                              const double *t, int32_t nt, double *out, double *y_final, lassie_stats *stats,
                              double *elapsed_ms);

/* Batched plan: the batched form prepared once and replayed (parameter sweeps,
 * repeated solves).  Creation allocates every workspace, picks the kernel
 * variant (LASSIE_BATCH_VARIANTS of them: 1, 2 or 4 row tiles of the LU
 * trailing update in flight) by timing a short trial of each variant on these
 * instances (cached per device, largest S, count and precision), and captures
 * the launch sequence as hipGraphs.
 * lassie_batch_plan_run replays it, with the output layout of
 * lassie_solve_batch; rates changed with lassie_model_set_rates are used by
 * the next run.  The first run hands instances to the kernels' work queues
 * largest S first; each run records every instance's device time, and the
 * next run hands them out longest first (list scheduling on the measured
 * costs), and the workgroup-level kernel leaves the wave-level one as many
 * CUs as balance the two kernels' measured work
 * (LASSIE_BATCH_SCHEDULE=size keeps the by-size order and every CU for the
 * workgroup-level kernel).  The schedule changes no result, only when and
 * where each instance runs.  The models must outlive the plan.  lassie_batch_plan_info
 * reports the chosen variant, its row tiles and the trial times
 * (LASSIE_BATCH_VARIANTS values in ms; zeros when the choice came from the
 * cache).  Choices persist in a
 * JSON file ($LASSIE_TUNE_CACHE, else $XDG_CACHE_HOME/lassie/tune.json, else
 * ~/.cache/lassie/tune.json; LASSIE_TUNE_CACHE=off keeps them in the process),
 * keyed by device architecture, CU count, largest S, instance count and
 * precision.  No counterpart in the reference (one model per process). */
#define LASSIE_BATCH_VARIANTS 3
typedef struct lassie_batch_plan lassie_batch_plan;
int lassie_batch_plan_create(lassie_model *const *models, int32_t n_models, const lassie_opts *opts,
                             const double *t, int32_t nt, lassie_batch_plan **plan);
int lassie_batch_plan_run(lassie_batch_plan *plan, double *out, double *y_final, lassie_stats *stats,
                          double *elapsed_ms);
int lassie_batch_plan_info(const lassie_batch_plan *plan, int32_t *variant, int32_t *tile_rows, float *trial_ms);
/* Instances finished so far by the run in progress (or by the last run), for
 * a caller that follows a long lassie_batch_plan_run from another thread:
 * every workgroup publishes its count to pinned host memory when an instance
 * ends.  Reading is safe while the run is in flight.  No counterpart in the
 * reference (it runs one model per process and reports its progress in the
 * ProgressSimulation file, Simulation.cuh:660-672). */
int lassie_batch_plan_progress(const lassie_batch_plan *plan, int64_t *done);
/* How the plan divides its instances between the batched form's two kernels
 * and their device times in the last run, each measured by events on its own
 * stream (ms[0]: the wave-level kernel, one wavefront per instance of at most
 * 64 species, FP64; ms[1]: the workgroup-level kernel).  The two kernels run
 * concurrently: the workgroup-level one first, the wave-level one on the CUs
 * it leaves free.  Any output may be NULL.  No counterpart in the reference
 * (one model per process). */
int lassie_batch_plan_timing(const lassie_batch_plan *plan, int32_t *n_wave, int32_t *n_wg, float *ms);
int lassie_batch_plan_destroy(lassie_batch_plan *plan);

/* Replaces the rate constants (c_vector) of an uploaded model, k[R].  A tape
 * recorded before the change can no longer be differentiated.  No
 * counterpart in the reference (its rates are fixed per process,
 * Simulation.cuh:397-418); used by the differentiation form. */
int lassie_model_set_rates(lassie_model *model, const double *k);

/* ---- Differentiation form --------------------------------------------------
 * Gradients of a loss on the sampled trajectories with respect to log k and
 * x0, by a discrete adjoint through the RKF45 and BDF1 steps of the loop
 * (Simulation.cuh:632-925, kernels.cuh:68-199, 289-300, 393-477, 512-521).
 * No counterpart in the reference.
 *
 * lassie_forward_batch runs the batched form (FP64, same trajectories as
 * lassie_solve_batch) and records, per model, every state-changing step (size
 * and kind) and the state before every 32nd step, in a tape of at most
 * max_steps steps per model (LASSIE_ERR_NOMEM when a model takes more).  The
 * models must stay alive, with unchanged rates, until the tape is destroyed.
 * out receives nt x S_i values per model, model after model. */
typedef struct lassie_tape lassie_tape;
int lassie_forward_batch(lassie_model *const *models, int32_t n_models, const lassie_opts *opts,
                         const double *t, int32_t nt, int64_t max_steps, double *out, lassie_stats *stats,
                         double *elapsed_ms, lassie_tape **tape);

/* Reverse pass: grad_out holds dL/d(out) in the layout of out; grad_log_k
 * receives dL/dlog k (R_i per model, model after model), grad_x0 (may be NULL)
 * dL/dx0 (S_i per model).  The step sequence is the recorded one (step sizes
 * and Newton iteration counts held fixed). */
int lassie_adjoint_batch(lassie_tape *tape, const double *grad_out, double *grad_log_k, double *grad_x0,
                         double *elapsed_ms);
/* Returns LASSIE_ERR_NONFINITE (every output written) when the gradient of
 * at least one model is not finite.  That happens when a BDF1 step's
 * modified Newton iteration ran to opts.newton_iter without converging
 * (Simulation.cuh:789, 883-890: the reference keeps whatever iterate it
 * reached): the derivative of such an iterate grows geometrically with the
 * iterations and overflows.  lassie_adjoint_status reports, per model,
 * whether its last gradient was finite (1), not finite (0) or not computed
 * yet (-1), and how many of its recorded BDF1 steps ran the full
 * opts.newton_iter iterations.  Either output may be NULL. */
int lassie_adjoint_status(const lassie_tape *tape, int32_t model, int32_t *finite, int32_t *capped_newton_steps);

/* Steps recorded for one model and the largest Newton iteration count of its
 * BDF1 steps; lassie_tape_read copies the record out: dt[n_steps], kind[n_steps]
 * (-1 accepted RKF45 step, -2 BDF1 step whose LU failed, else the BDF1 step's
 * Newton iterations), sample_step[nt] (steps taken when each sample was
 * stored).  Any output may be NULL. */
int lassie_tape_info(const lassie_tape *tape, int32_t model, int64_t *n_steps, int32_t *max_newton);
int lassie_tape_read(const lassie_tape *tape, int32_t model, double *dt, int32_t *kind, int32_t *sample_step);
int lassie_tape_destroy(lassie_tape *tape);

/* Host helper: CSR arrays from dense R x S stoichiometry (int16, row-major, as
 * Reader<T>::readMatrix returns them, Reader.h:104-124), in the entry order
 * of create_sparse2 / create_sparse (Simulation.cuh:1079-1130).  Output
 * arrays must hold R+1 / nnz(left) / S+1 / nnz(H) entries; pass NULL outputs
 * to query the two counts first. */
int lassie_csr_from_dense(int32_t n_species, int32_t n_reactions, const int16_t *left,
                          const int16_t *right, int32_t *nnz_left, int32_t *nnz_net,
                          int32_t *react_ptr, int32_t *react_species, int32_t *react_exp,
                          int32_t *net_ptr, int32_t *net_reaction, int32_t *net_coef);

/* Number of visible devices and the name of device 0 (name may be NULL). */
int lassie_device_count(int32_t *count);

const char *lassie_last_error(void);
int lassie_abi_version(void);

#ifdef __cplusplus
}
#endif
#endif
