// prng.h -- the reference's seeded uniform sampler (the symmetric-ciphertext
// seed format regenerates c1 from it).
//
// Block function: the Salsa20 hash (Bernstein, "Salsa20 specification",
// sections 3-8: quarterround, rowround, columnround, doubleround x 10,
// feed-forward) of a 16-word input made of the 64-byte seed's words 0..7,
// the nonce (low, high) and the seed's words 8..13 (src/prng.cu:17-131).
// Sampler (sample_uniform_poly, src/prng.cu:174-210): block b of 8 output
// coefficients (b < L N / 8, limb b / (N/8)) takes the 8 little-endian u64 of
// the block for nonce b; a value above (2^64 - 1) - ((2^64 - 1) mod q) - 1 is
// replaced, from its position on, by the values of a fresh block for nonce
// b + r L N (r = 1, 2, ... counting the blocks drawn for b); each accepted
// value is reduced mod q.
#pragma once

#include "context.h"

namespace phx {

constexpr size_t kPrngSeedBytes = 64;

__host__ __device__ inline uint32_t salsa_rotl(uint32_t v, int c) { return (v << c) | (v >> (32 - c)); }

// quarterround on (y[a], y[b], y[c], y[d]) of the specification
__host__ __device__ inline void salsa_quarter(uint32_t *y, int a, int b, int c, int d) {
    y[b] ^= salsa_rotl(y[a] + y[d], 7);
    y[c] ^= salsa_rotl(y[b] + y[a], 9);
    y[d] ^= salsa_rotl(y[c] + y[b], 13);
    y[a] ^= salsa_rotl(y[d] + y[c], 18);
}

// Salsa20 hash: out = in + doubleround^10(in), word-wise mod 2^32
__host__ __device__ inline void salsa_hash(const uint32_t in[16], uint32_t out[16]) {
    uint32_t y[16];
    for (int i = 0; i < 16; ++i) y[i] = in[i];
    for (int r = 0; r < 10; ++r) {
        // columnround
        salsa_quarter(y, 0, 4, 8, 12);
        salsa_quarter(y, 5, 9, 13, 1);
        salsa_quarter(y, 10, 14, 2, 6);
        salsa_quarter(y, 15, 3, 7, 11);
        // rowround
        salsa_quarter(y, 0, 1, 2, 3);
        salsa_quarter(y, 5, 6, 7, 4);
        salsa_quarter(y, 10, 11, 8, 9);
        salsa_quarter(y, 15, 12, 13, 14);
    }
    for (int i = 0; i < 16; ++i) out[i] = y[i] + in[i];
}

// one 64-byte block of the sampler as 8 u64 (little-endian words)
__host__ __device__ inline void seeded_block(const uint8_t *seed, u64 nonce, u64 out[8]) {
    uint32_t in[16], h[16];
    auto word = [&](int k) {
        return (uint32_t)seed[4 * k] | ((uint32_t)seed[4 * k + 1] << 8) | ((uint32_t)seed[4 * k + 2] << 16) |
               ((uint32_t)seed[4 * k + 3] << 24);
    };
    for (int k = 0; k < 8; ++k) in[k] = word(k);
    in[8] = (uint32_t)nonce;
    in[9] = (uint32_t)(nonce >> 32);
    for (int k = 8; k < 14; ++k) in[k + 2] = word(k);
    salsa_hash(in, h);
    for (int k = 0; k < 8; ++k) out[k] = (u64)h[2 * k] | ((u64)h[2 * k + 1] << 32);
}

// out [L][N] (device) <- the sampler's L x N values for seed (host, 64 bytes);
// limb l is reduced by the context's table row l
void sample_uniform_seeded(const Context &ctx, const uint8_t *seed, u32 L, u64 *out, hipStream_t s);

}  // namespace phx
