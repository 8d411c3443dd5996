// autotune.h -- per-context kernel-variant search and its on-disk cache.
#pragma once

#include <string>
#include <utility>
#include <vector>

#include "context.h"

namespace phx {

struct TuneResult {
    TuneChoice choice;
    double default_us = 0;   // objective (HMult+relin + one op of an 8-pair batch) at the process-wide switches
    double best_us = 0;      // the same with `choice`
    std::vector<std::pair<std::string, double>> trials;
};

// Times HMult+relin at the top level (CKKS: tensor + relinearise; BFV: the
// context's multiply technique + relinearise), plain and as an 8-pair batch
// plan, on synthetic residues resident in HBM for each candidate, applies the
// fastest (by at least 2%) to ctx and returns it.
TuneResult autotune(Context &ctx, hipStream_t s);

// cache key of a context shape: "scheme=..,N=..,Q=..,P=.."
std::string tune_cache_key(u32 scheme, u64 N, u32 size_Q, u32 size_P);
// replaces (or adds) the entry for `key` in the JSON cache at path
void tune_cache_store(const std::string &path, const std::string &key, const TuneChoice &c, double us);
// reads the entry for `key`; false when the file or the entry is absent
bool tune_cache_lookup(const std::string &path, const std::string &key, TuneChoice *out);
// the same for a context's shape
bool tune_cache_load(const std::string &path, const Context &ctx, TuneChoice *out);

}  // namespace phx
