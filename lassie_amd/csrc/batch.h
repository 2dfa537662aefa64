// Launch interface of the batched persistent solver (host <-> device).
#pragma once
#include <hip/hip_runtime.h>

#include "device_common.h"

namespace lassie {

constexpr int kBatchMaxS = 512;  // largest instance of the batched form
constexpr int kBatchNB = 16;     // panel width of the workgroup-local LU

// One instance: device pointers of an uploaded model and where its outputs go.
struct BatchInst {
    int S, R;
    const int *lptr, *lsp, *lexp, *hptr, *hrx;
    const HRec *recs;
    const double *hk, *x0, *feed, *atol;
    long long out_off, y_off;
    long long tape_off, ckpt_off;  // step record of the differentiation form
};

constexpr int kBatchStats = 18;  // counters per instance (last 4: phase times, 100 MHz ticks)
constexpr int kCkptEvery = 32;   // steps between checkpointed states

struct BatchArgs {
    const BatchInst *inst;
    int n_inst;
    const int *order;  // instances, most expensive first
    int *next;         // work-queue head
    const double *tv;
    int nt;
    double *out, *y_final;
    long long *stats;  // kBatchStats per instance
    double *dstats;    // t, dt, newdt per instance
    double rkf_step, be_step, newton_tol, stiffness_tol, rtol;
    int newton_iter, bdf_regrow, max_rkf_retries;
    long long max_iters;
    long long max_newton;  // 0: no cap on an instance's Newton iterations
    int precision;
    double refine_tol;
    int refine_max_iter;
    int s_max;
    int nh_max;    // wave kernel: most net-stoichiometry entries of an instance (LDS sizing)
    double *ws64;  // per slot: 12 vectors + dense M (FP64)
    size_t ws64_stride;
    float *ws32;   // per slot: dense M (FP32)
    size_t ws32_stride;
    int *wsi;      // per slot: pivots, row moves, panel row lists, flags
    size_t wsi_stride;
    // step record (null: none): per instance, from tape_off, the size and kind
    // of every state-changing step (kind -1 accepted RKF45, -2 BDF1 with a
    // failed LU, else the BDF1 step's Newton iterations); the state before
    // every kCkptEvery-th step (from ckpt_off); the steps taken at each sample
    double *tape_dt;
    int *tape_kind;
    int *tape_sample;  // nt per instance
    double *tape_ckpt;
    long long tape_cap;
    // per slot: instances it has finished, in pinned host memory, so that the
    // host can follow a long launch (null: not kept)
    unsigned int *done_host;
};

// Kernel variants (row tiles of the LU trailing update in flight per wave):
// 0 -> 1, 1 -> 2, 2 -> 4.  The arithmetic is the same in every variant.
constexpr int kBatchVariants = 3;
constexpr int kBatchDefaultVariant = 2;
int batch_variant_rt(int variant);
hipError_t launch_batch(const BatchArgs &args, int slots, hipStream_t stream, int variant = kBatchDefaultVariant);
int batch_occupancy(int *per_cu, int variant = kBatchDefaultVariant);
// Resets both work-queue heads (next[0], next[1]) and fills out[0, n_out)
// with NaN (samples an instance does not reach keep that value); a kernel,
// so it may be captured in a graph.
hipError_t launch_batch_reset(int *next, hipStream_t stream, double *out = nullptr, size_t n_out = 0);

}  // namespace lassie
