// tune.cpp -- process-wide kernel-variant switches (read by the launchers).
//
//   bconv_mfma   1: FastBConv on the matrix cores (default), 0: VALU kernel
//   ntt_fp64     1: FP64 butterflies for primes < 2^50 (default), 0: integer only;
//                read when a context is created
// Environment defaults: PHX_BCONV_VALU=1 -> bconv_mfma=0, PHX_NTT_INT_ONLY=1 -> ntt_fp64=0.
#include <atomic>
#include <cstdlib>
#include <cstring>

#include "tune.h"

namespace phx {

namespace {
bool env_is_one(const char *name) {
    const char *e = std::getenv(name);
    return e && e[0] == '1';
}
std::atomic<int> g_bconv_mfma{env_is_one("PHX_BCONV_VALU") ? 0 : 1};
std::atomic<int> g_ntt_fp64{env_is_one("PHX_NTT_INT_ONLY") ? 0 : 1};

std::atomic<int> *slot(const char *key) {
    if (!key) return nullptr;
    if (!std::strcmp(key, "bconv_mfma")) return &g_bconv_mfma;
    if (!std::strcmp(key, "ntt_fp64")) return &g_ntt_fp64;
    return nullptr;
}
}  // namespace

bool tune_set(const char *key, int value) {
    auto *s = slot(key);
    if (!s) return false;
    s->store(value);
    return true;
}

bool tune_get(const char *key, int *value) {
    auto *s = slot(key);
    if (!s) return false;
    *value = s->load();
    return true;
}

int tune_value(const char *key) {
    int v = 0;
    tune_get(key, &v);
    return v;
}

}  // namespace phx
