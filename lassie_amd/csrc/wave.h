// Launch interface of the wave-level batched solver: instances with at most
// kWaveMaxS species, one wavefront per instance (host <-> device).
#pragma once
#include <hip/hip_runtime.h>

#include "batch.h"

namespace lassie {

constexpr int kWaveMaxS = 64;  // largest instance a single wave integrates (one row per lane)

// Dynamic LDS of one wave (bytes) for instances of at most smax species and
// at most nh_max net-stoichiometry entries: the Newton matrix (smax rows of
// smax + 1 doubles), five state vectors, the model's RHS records and their
// species pointers.
size_t wave_lds_bytes(int smax, int nh_max);

// Runs args.n_inst instances (args.order lists them, largest first; every one
// has S <= kWaveMaxS, FP64, no step record) on `blocks` one-wave workgroups.
hipError_t launch_wave(const BatchArgs &args, int blocks, size_t lds_bytes, hipStream_t stream);
int wave_occupancy(size_t lds_bytes, int *per_cu);

}  // namespace lassie
