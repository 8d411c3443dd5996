// noalloc.hip -- CKKS multiply_and_relin_inplace without any allocation, for
// callers that cannot allocate device memory (code under HIP graph capture,
// tight serving loops, harnesses that forbid hipMalloc).
//
// The workspace of the operation (the tensor's c2, the modup output, the key
// inner product's cx, the scaled INTT of c2 and the moddown's delta) lives in
// one file-scope device array of kScratchBytes per GPU, held from the moment
// the library's code object is loaded on that GPU until the process ends.
// Calls on one GPU share that array: calls that overlap on one GPU (two
// streams, two host threads) race and must be serialised by the caller. No
// part of the array has to start at zero: every kernel of the pipeline writes
// the words it later reads.
#include <mutex>

#include "evaluator.h"

namespace phx {

namespace {
// 192 MiB: the workspace of one HMult+relin at N = 2^16 with 35 data limbs,
// 12 special limbs and 3 digits (187.5 MiB)
constexpr size_t kScratchBytes = size_t{192} << 20;
constexpr size_t kScratchWords = kScratchBytes / 8;
constexpr int kMaxDevices = 64;

static __device__ __align__(256) u64 scratch[kScratchWords];

std::mutex g_mu;
u64 *g_ptr[kMaxDevices] = {};

// the device address of `scratch` on the current device, looked up once per device
// (host code uses this address, never the array's name)
u64 *scratch_ptr() {
    int dev = 0;
    PHX_HIP_CHECK(hipGetDevice(&dev));
    require(dev >= 0 && dev < kMaxDevices, "device index out of range");
    std::lock_guard<std::mutex> lock(g_mu);
    if (!g_ptr[dev]) {
        void *p = nullptr;
        PHX_HIP_CHECK(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
        require(p != nullptr, "static scratch has no device address");
        g_ptr[dev] = static_cast<u64 *>(p);
    }
    return g_ptr[dev];
}
}  // namespace

void multiply_and_relin_noalloc(const Context &ctx, Ciphertext &ct1, const Ciphertext &ct2, const KSwitchKey &rlk,
                                hipStream_t s) {
    require(ctx.params().scheme == kSchemeCKKS, "the allocation-free multiply_and_relin runs CKKS ciphertexts");
    require(ct1.chain_index >= 1 && ct1.chain_index <= ctx.size_Q(), "invalid chain index");
    const size_t words = multiply_and_relin_ws_words(ctx, ct1.chain_index);
    require(words <= kScratchWords, "the allocation-free multiply_and_relin needs " + std::to_string(words * 8) +
                                        " bytes of workspace, above its static scratch of " +
                                        std::to_string(kScratchBytes) + " bytes");
    multiply_and_relin_ws(ctx, ct1, ct2, rlk, scratch_ptr(), s);
}

}  // namespace phx
