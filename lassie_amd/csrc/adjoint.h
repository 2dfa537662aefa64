// Launch interface of the reverse (discrete adjoint) pass of the
// differentiation form (host <-> device).
#pragma once
#include <hip/hip_runtime.h>

#include "device_common.h"

namespace lassie {

// One instance: the uploaded model, the index maps the reverse pass needs, and
// where its step record, incoming gradient and outgoing gradients live.
struct AdjInst {
    int S, R;
    const int *lptr, *lsp, *lexp, *hptr, *hrx;
    const HRec *recs;
    const double *hk, *x0, *feed, *atol, *k;
    const int *htptr, *htsp, *htcoef;  // net stoichiometry by reaction (H^T as CSR)
    const int *ltptr, *ltent, *lrx;    // reactant entries by species; reaction of each entry
    long long tape_off, ckpt_off, n_steps;
    long long g_off;                   // into grad_out (nt x S)
    long long gk_off, gx_off;          // into grad_logk (R) and grad_x0 (S)
};

struct AdjArgs {
    const AdjInst *inst;
    int n_inst;
    const int *order;
    int *next;
    int nt;
    const double *tape_dt;
    const int *tape_kind;
    const int *tape_sample;  // nt per instance
    const double *tape_ckpt;
    const double *grad_out;
    double *grad_logk, *grad_x0;
    double rtol;
    int s_max, r_max, it_max;
    double *ws;  // per slot: segment states, factors, Newton iterates, vectors, reaction scratch
    size_t ws_stride;
    int *wsi;    // per slot: pivots, row moves, panel row lists, flags
    size_t wsi_stride;
    int *status;  // per instance: 0, or 1 when a replayed LU failed
    // per slot, a second pair of iterate/correction buffers ((it_max + 1) s_max
    // doubles each) for the segment step with the most Newton iterations, so
    // that its reverse step needs no second replay; nullptr: replay it
    double *vx2;
    size_t vx2_stride;
};

hipError_t launch_adjoint(const AdjArgs &args, int slots, hipStream_t stream);
int adjoint_occupancy(int *per_cu);

}  // namespace lassie
