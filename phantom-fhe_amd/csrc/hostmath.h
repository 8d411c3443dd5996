// hostmath.h -- host-side parameter maths: primes, NTT roots and tables,
// RNS constants. Built from the number theory, reproducing the values the
// reference's host library computes (src/host/numth.cu, src/host/ntt.cu,
// src/host/rns.cu, src/host/modulus.cu).
#pragma once

#include <cstdint>
#include <vector>

namespace phx {
namespace hm {

using u64 = uint64_t;
using u128 = unsigned __int128;

inline u64 mulmod(u64 a, u64 b, u64 q) { return (u64)(((u128)a * b) % q); }
inline u64 addmod(u64 a, u64 b, u64 q) { u64 r = a + b; return r >= q ? r - q : r; }
u64 powmod(u64 a, u64 e, u64 q);
u64 invmod(u64 a, u64 q);                 // q prime
inline u64 shoup(u64 w, u64 q) { return (u64)(((u128)w << 64) / q); }
bool is_prime(u64 n);
int bit_count(u64 v);
u64 reverse_bits(u64 x, int bits);

// floor(2^128 / q) as (lo, hi)
void barrett_ratio(u64 q, u64 &lo, u64 &hi);

// reference: get_primes, src/host/numth.cu:207-233
std::vector<u64> get_primes(u64 N, int bits, size_t count);
// reference: CoeffModulus::Create, src/host/modulus.cu:80-111
std::vector<u64> coeff_modulus_create(u64 N, const std::vector<int> &bits);
// reference: try_minimal_primitive_root, src/host/numth.cu:309-331
u64 minimal_primitive_root(u64 q, u64 degree);

// Per-prime NTT tables (bit-reversed powers; plain, not pre-scaled by n^-1).
struct NttTable {
    u64 q = 0;
    u64 psi = 0;
    std::vector<u64> fw, fws;   // fw[brv(i)] = psi^i
    std::vector<u64> iw, iws;   // iw[brv(i)] = psi^-i
    u64 ninv = 0;
};
NttTable make_ntt_table(u64 q, int logN);

// prod_{k != skip} base[k] mod m
u64 punctured_mod(const u64 *base, size_t n, size_t skip, u64 m);
u64 product_mod(const u64 *base, size_t n, u64 m);

}  // namespace hm
}  // namespace phx
