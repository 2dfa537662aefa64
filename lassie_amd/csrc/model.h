// Library-internal helpers shared by the C ABI sources (api.cpp, plan.cpp).
#pragma once
#include <string>

#include "../../include/lassie.h"
#include "batch.h"

namespace lassie {

int api_fail(int code, const std::string &msg);              // sets lassie_last_error, returns code
int model_species(const lassie_model *m);
int model_device(const lassie_model *m);
void fill_batch_inst(const lassie_model *m, BatchInst *b);   // device pointers; offsets left to the caller
void fill_batch_stats(const long long *c, const double *d, float ms, int slots, lassie_stats *s);

}  // namespace lassie
