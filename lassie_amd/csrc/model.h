// Library-internal helpers shared by the C ABI sources (api.cpp, plan.cpp).
#pragma once
#include <hip/hip_runtime.h>

#include <string>
#include <vector>

#include "../../include/lassie.h"
#include "batch.h"

namespace lassie {

int api_fail(int code, const std::string &msg);              // sets lassie_last_error, returns code
int model_species(const lassie_model *m);
int model_device(const lassie_model *m);
int model_nnz_net(const lassie_model *m);                    // net-stoichiometry entries (RHS records)
hipStream_t model_stream(const lassie_model *m);             // the model's stream
void model_events(const lassie_model *m, hipEvent_t *ev0, hipEvent_t *ev1);
void fill_batch_inst(const lassie_model *m, BatchInst *b);   // device pointers; offsets left to the caller
void fill_batch_stats(const long long *c, const double *d, float ms, int slots, lassie_stats *s);

// The batched form runs two kernels: the wave-level one (one wavefront per
// instance) for FP64 instances of at most kWaveMaxS species whose RHS records
// fit its LDS, and the workgroup-level one for every other instance.
constexpr int kWaveMaxNh = 2048;  // net-stoichiometry entries a wave keeps in LDS
struct BatchSplit {
    std::vector<int> order;  // workgroup-kernel instances, then wave-kernel instances, each largest S first
    int n_wg = 0, n_wave = 0;
    int smax_wg = 1, smax_wave = 1, nh_max_wave = 0;
    size_t wave_lds = 0;  // dynamic LDS of one wave
};
// allow_wave false sends every instance to the workgroup kernel (precision
// forms, the step record, and opts.reserved bit 2)
BatchSplit split_batch(lassie_model *const *models, const std::vector<BatchInst> &inst, bool allow_wave);

}  // namespace lassie
