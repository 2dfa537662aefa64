// Launch interface of the persistent single-model solver (host <-> device).
#pragma once
#include <hip/hip_runtime.h>

#include "device_common.h"

namespace lassie {

constexpr int kMaxDenseN = 4096;  // largest S for the dense Newton matrix

enum { MODE_CONTINUE = 0, MODE_BDF = 1, MODE_BDF_NOOP = 2, MODE_DONE = 3 };

// Control words exchanged between workgroups (one cache line each).
struct Ctl {
    int mode;
    int pad0[15];
    int info;
    int pad1[15];
    int abort;
    int pad2[15];
    double h;
    double pad3[7];
};

struct SolveArgs {
    DevModel m;
    const double *x0;
    double rkf_step, be_step, newton_tol, stiffness_tol, rtol;
    int newton_iter, bdf_regrow, max_rkf_retries;
    long long max_iters;
    const double *tv;
    int nt;
    const int *cs;
    int ncs;
    double *out;
    double *progress;
    long long progress_cap;
    double *y1, *u, *u2, *k1, *k2, *k3, *k4, *k5, *k6, *w1, *yv, *x;
    double *M;
    int *piv, *perm_src, *moved_to;
    Ctl *ctl;
    GridSync *gs;
    long long *stats_out;  // 10 counters
    double *dstats_out;    // t, dt, newdt
    int G;
};

hipError_t launch_solve(const SolveArgs &args, hipStream_t stream);
int max_coop_groups(int n, int *out);
int dense_panel_width(int n);

}  // namespace lassie
