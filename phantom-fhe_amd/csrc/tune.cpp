// tune.cpp -- process-wide kernel-variant switches (read by the launchers).
//
//   bconv_mfma   1: FastBConv on the matrix cores (default), 0: VALU kernel
//   ntt_fp64     1: FP64 butterflies for primes < 2^50 (default), 0: integer only;
//                read when a context is created
//   ntt_fused    G in {1, 2, 4}: one workgroup of G x 256 threads runs both passes
//                of a limb (intermediate kept on chip / in the last-level cache);
//                0: two launches per transform
//   ntt_fused_min_limbs  launches over fewer limbs than this use two passes
//   ntt_chunk_limbs  two-pass launches over more limbs than this are split into
//                slices of whole polynomials of about this many limbs (0 = never)
//   ntt_cols_threads / ntt_rows_threads  tile shape of the two passes: threads per
//                workgroup (16 registers each): columns 256 or 512, or 0 (default)
//                for 512 at N = 2^17, and at N = 2^16 in launches of at least 256
//                limb-transforms, 256 otherwise; rows 128, 256 (default) or 512
//   bconv_wgs_per_cu  FastBConv on the matrix cores: workgroups per CU of its grid
//                (1, 2, 4; 0 = as many as the LDS holds, the default)
//   bconv_fused  1: FastBConv on the matrix cores fused with the column pass of the
//                forward NTT of its outputs (default, N = 2^12 .. 2^16), 0: separate
//                launches
//   batch_branches  concurrent branches of a persistent batch plan (1..8, default
//                2): plans of at least twice as many items are split into that many
//                parts that run on forked streams, and joined plans spread their
//                parts over them; 1 runs everything on the caller's stream
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
std::atomic<int> g_ntt_fused{0};
std::atomic<int> g_ntt_fused_min{512};
std::atomic<int> g_ntt_chunk{0};
std::atomic<int> g_ntt_cols{0};
std::atomic<int> g_ntt_rows{256};
std::atomic<int> g_bconv_wgs{0};
std::atomic<int> g_batch_branches{2};
std::atomic<int> g_bconv_fused{1};

std::atomic<int> *slot(const char *key) {
    if (!key) return nullptr;
    if (!std::strcmp(key, "bconv_mfma")) return &g_bconv_mfma;
    if (!std::strcmp(key, "ntt_fp64")) return &g_ntt_fp64;
    if (!std::strcmp(key, "ntt_fused")) return &g_ntt_fused;
    if (!std::strcmp(key, "ntt_fused_min_limbs")) return &g_ntt_fused_min;
    if (!std::strcmp(key, "ntt_chunk_limbs")) return &g_ntt_chunk;
    if (!std::strcmp(key, "ntt_cols_threads")) return &g_ntt_cols;
    if (!std::strcmp(key, "ntt_rows_threads")) return &g_ntt_rows;
    if (!std::strcmp(key, "bconv_wgs_per_cu")) return &g_bconv_wgs;
    if (!std::strcmp(key, "batch_branches")) return &g_batch_branches;
    if (!std::strcmp(key, "bconv_fused")) return &g_bconv_fused;
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
