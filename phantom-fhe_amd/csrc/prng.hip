// prng.hip -- the reference's seeded uniform sampler on the GPU (see prng.h).
#include "prng.h"

#include <algorithm>
#include <cstring>

namespace phx {

namespace {
constexpr int kThreads = 256;
// a block is redrawn with probability below 2^-4 per value; past this many
// draws for one block the sampler stops (probability < 2^-4000)
constexpr u64 kMaxDraws = 1000;

struct Seed64 {
    uint8_t b[kPrngSeedBytes];
};

__global__ __launch_bounds__(kThreads) void k_sample_uniform_seeded(u64 *__restrict__ out, Seed64 seed,
                                                                    const DevModulus *__restrict__ mod, u32 N, u32 L) {
    const u64 per_limb = N / 8, blocks = per_limb * L;
    for (u64 b = (u64)blockIdx.x * kThreads + threadIdx.x; b < blocks; b += (u64)gridDim.x * kThreads) {
        const u64 q = mod[b / per_limb].value;
        const u64 max_accept = ~0ull - (~0ull % q) - 1;
        u64 r[8];
        seeded_block(seed.b, b, r);
        u64 draws = 1;
        for (int i = 0; i < 8; ++i) {
            while (r[i] > max_accept && draws < kMaxDraws) {
                seeded_block(seed.b, b + draws * (u64)N * L, r);
                ++draws;
            }
            out[b * 8 + i] = r[i] % q;
        }
    }
}
}  // namespace

void sample_uniform_seeded(const Context &ctx, const uint8_t *seed, u32 L, u64 *out, hipStream_t s) {
    require(seed != nullptr && out != nullptr, "sample_uniform_seeded: null argument");
    require(L >= 1 && L <= ctx.size_QP(), "sample_uniform_seeded: invalid limb count");
    Seed64 sd;
    std::memcpy(sd.b, seed, kPrngSeedBytes);
    const u64 blocks = ctx.N() / 8 * L;
    const u32 grid = (u32)std::min<u64>((blocks + kThreads - 1) / kThreads, 4096);
    hipLaunchKernelGGL(k_sample_uniform_seeded, dim3(grid), dim3(kThreads), 0, s, out, sd, ctx.dev_mod(),
                       (u32)ctx.N(), L);
    PHX_HIP_CHECK(hipGetLastError());
}

}  // namespace phx
