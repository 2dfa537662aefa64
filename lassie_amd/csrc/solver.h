// Launch interface of the persistent single-model solver (host <-> device).
#pragma once
#include <hip/hip_runtime.h>

#include "device_common.h"

namespace lassie {

constexpr int kMaxDenseN = 4096;  // largest S for the dense Newton matrix
constexpr int kNumVecLds = 11;    // state vectors the control group keeps in LDS

enum { MODE_CONTINUE = 0, MODE_BDF = 1, MODE_BDF_NOOP = 2, MODE_DONE = 3 };

// Control words exchanged between workgroups (one cache line each).  mode
// and h are double-buffered by loop-iteration parity: group 0 may publish
// iteration i+1 while a slower group still reads iteration i.
struct Ctl {
    int mode[2];
    int pad0[14];
    int info;
    int pad1[15];
    int abort;
    int pad2[15];
    double h[2];
    double pad4[6];
    unsigned int ngo;    // Newton iteration stamp published by group 0
    int ncont;           // 1: solve this iteration, 0: the Newton loop ends
    int pad5[14];
    unsigned int rbad;   // precision form: number of the last refinement sweep that missed its target
    int pad6[15];
};

constexpr int kTriBlock = 64;     // rows per block of the distributed triangular solves
constexpr int kFlagStride = 16;   // ints between per-block flags (one 64-byte line each)

struct SolveArgs {
    DevModel m;
    const HRec *recs;
    const double *x0;
    double rkf_step, be_step, newton_tol, stiffness_tol, rtol;
    int newton_iter, bdf_regrow, max_rkf_retries;
    long long max_iters;
    long long max_newton;  // 0: no cap on the solve's Newton iterations
    const double *tv;
    int nt;
    const int *cs;
    int ncs;
    double *out;
    double *progress;
    long long progress_cap;
    // state vectors in global memory (S each): y1 u u2 k1..k5 w1 yv x, of the
    // state type (double, or float when fp32)
    double *gvec;
    double *M;             // dense S x S Newton matrix, of the state type
    int *piv, *perm_src, *moved_to;
    int *rows, *rowcnt;    // per panel: positions with nonzero multipliers
    int *perm_full;        // final row order of the factors: row i of L U is row perm_full[i] of M
    double *tri;           // 3 x S of the state type: permuted right-hand side, forward result, solution
    unsigned int *tflag;   // 2 x blocks x kFlagStride: forward / backward block stamps
    Ctl *ctl;
    GridSync *gs;
    long long *stats_out;  // 12 counters
    double *dstats_out;    // t, dt, newdt, phase ms x4
    int G;                 // workgroups in the cooperative launch
    int lds_model;         // 1: control group keeps model + vectors in LDS
    int fp32;              // 1: single-precision loop (state, time, M and LU in float)
    int lds_strip;         // 1: groups >= 1 keep their LU strip in dynamic LDS (FP64, one strip per group)
    int strip_ld;          // leading dimension of an LDS strip (elements)
    double *yfin;          // fp32 only: final state widened to double (S)
    // precision form (FP64 loop): the Newton matrix is factored in FP32 and
    // each Newton solve is refined in FP64 against the unfactored M
    int refine;            // 1: FP32 LU + FP64 iterative refinement
    int refine_max_iter;   // sweeps before the FP64 LU of the step's matrix
    double refine_tol;     // stop when |r_i| <= tol * (|b_i| + (|M||x|)_i) for every i
    float *M32;            // S x S, column-major like M, followed by kLuSink elements
    double *rvec;          // 4 x S: right-hand side, residual, iterate (two buffers)
};

hipError_t launch_solve(const SolveArgs &args, size_t dyn_lds, hipStream_t stream);
int solver_occupancy(int n, int fp32, int refine, size_t dyn_lds, int *per_cu);
int dense_panel_width(int n);
size_t lds_model_bytes(int S, int nnz_h);

}  // namespace lassie
