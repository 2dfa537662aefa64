// Persistent choices of the batched plan's autotuner: one JSON file, keyed by
// device and workload shape, so a later process (or run) reuses a choice
// instead of timing the trials again.
#pragma once
#include <string>

namespace lassie {

// Path of the cache: $LASSIE_TUNE_CACHE, else $XDG_CACHE_HOME/lassie/tune.json,
// else $HOME/.cache/lassie/tune.json.  Empty when LASSIE_TUNE_CACHE is "off"
// or no location is known: choices then live in this process only.
std::string tune_cache_path();

// The stored choice for key, if any.
bool tune_lookup(const std::string &key, int *choice);

// Stores (replaces) key's choice with the trial times that justified it.
// Failures to write leave the file as it was.
void tune_store(const std::string &key, int choice, const float *trial_ms, int n_trials);

}  // namespace lassie
