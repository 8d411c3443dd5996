// common.h -- shared types, error handling and gfx950 modular arithmetic.
//
// All ring elements are u64 residues. Device modular arithmetic uses the
// Shoup (precomputed-quotient) and Barrett (floor(2^128/q)) methods, written
// for the CDNA4 VALU: 64x64 products lower to v_mad_u64_u32 / v_mul_hi_u32
// chains; there is no inline asm.
#pragma once

#include <hip/hip_runtime.h>

#include <cstddef>
#include <cstdint>
#include <stdexcept>
#include <string>

namespace phx {

using u64 = uint64_t;
using u32 = uint32_t;
using u128 = unsigned __int128;

// ---------------------------------------------------------------- errors
enum Status : int {
    kOk = 0,
    kInvalidArgument = -1,   // maps to std::invalid_argument
    kLogicError = -2,        // maps to std::logic_error
    kRuntimeError = -3,      // HIP / allocation failures: std::runtime_error
    kOutOfRange = -4,
};

struct Error : std::runtime_error {
    int code;
    Error(int c, const std::string &m) : std::runtime_error(m), code(c) {}
};

[[noreturn]] inline void fail(int code, const std::string &msg) { throw Error(code, msg); }
inline void require(bool cond, const std::string &msg) {
    if (!cond) fail(kInvalidArgument, msg);
}

#define PHX_HIP_CHECK(expr)                                                                     \
    do {                                                                                        \
        hipError_t e_ = (expr);                                                                 \
        if (e_ != hipSuccess)                                                                   \
            ::phx::fail(::phx::kRuntimeError,                                                   \
                        std::string("HIP error: ") + hipGetErrorString(e_) + " at " #expr);     \
    } while (0)

// ---------------------------------------------------------------- device maths
// Modulus record as the kernels read it: value, Barrett ratio floor(2^128/q)
// (lo, hi) -- hi is floor(2^64/q) for the 64-bit Barrett step.
struct DevModulus {
    u64 value;
    u64 ratio_lo;
    u64 ratio_hi;
    u64 pad;
};

// twiddle with its Shoup companion floor(w * 2^64 / q)
struct alignas(16) TwPair {
    u64 w;
    u64 ws;
};

// A device buffer seen through the global address space. Pointers that kernels
// read from device work lists are generic to the compiler, which then emits flat
// memory instructions; a flat access also counts against the LDS wait counter, so
// every LDS wait in a kernel would wait for the flat loads and stores in flight as
// well. Every pointer the kernels dereference this way is device memory.
#define PHX_GLOBAL __attribute__((address_space(1)))
template <class T>
__device__ __forceinline__ PHX_GLOBAL T *gmem(T *p) {
    return (PHX_GLOBAL T *)p;
}

__device__ __forceinline__ u64 umulhi(u64 a, u64 b) { return __umul64hi(a, b); }

// conditional subtract: x in [0, 2m) -> [0, m)
__device__ __forceinline__ u64 csub(u64 x, u64 m) { return x >= m ? x - m : x; }

__device__ __forceinline__ u64 add_mod(u64 a, u64 b, u64 q) { return csub(a + b, q); }
__device__ __forceinline__ u64 sub_mod(u64 a, u64 b, u64 q) { return csub(a + q - b, q); }

// Shoup product x*w mod q in [0, 2q) for any 64-bit x (w < q, ws = shoup(w)).
__device__ __forceinline__ u64 mul_shoup_lazy(u64 x, u64 w, u64 ws, u64 q) {
    u64 hi = umulhi(x, ws);
    return x * w - hi * q;
}
__device__ __forceinline__ u64 mul_shoup(u64 x, u64 w, u64 ws, u64 q) {
    return csub(mul_shoup_lazy(x, w, ws, q), q);
}

// Barrett reduction of a 128-bit value by q < 2^62 with r = floor(2^128/q):
// the quotient estimate floor(x * r / 2^128) is exact or one short, so one
// conditional subtraction yields the canonical residue.
__device__ __forceinline__ u64 barrett128(u128 x, u64 q, u64 r_lo, u64 r_hi) {
    u64 xl = (u64)x, xh = (u64)(x >> 64);
    u128 t = (u128)umulhi(xl, r_lo);
    t += (u128)xl * r_hi;
    u128 t2 = (u128)xh * r_lo;
    t = (t >> 64) + ((((u128)(u64)t) + (u64)t2) >> 64) + (t2 >> 64);
    u64 est = (u64)t + xh * r_hi;
    return csub(xl - est * q, q);
}

// 64-bit Barrett: x mod q using floor(2^64/q)
__device__ __forceinline__ u64 barrett64(u64 x, u64 q, u64 r_hi) {
    u64 est = umulhi(x, r_hi);
    return csub(x - est * q, q);
}

__device__ __forceinline__ u64 mul_mod(u64 a, u64 b, const DevModulus &m) {
    return barrett128((u128)a * b, m.value, m.ratio_lo, m.ratio_hi);
}

// Forward (Cooley-Tukey) butterfly on lazy values x in [0,4q):
// x' = x + w*y, y' = x - w*y, both in [0, 4q).
__device__ __forceinline__ void ct_bfly(u64 &x, u64 &y, u64 w, u64 ws, u64 q, u64 q2) {
    u64 t = mul_shoup_lazy(y, w, ws, q);
    u64 xx = csub(x, q2);
    y = xx + q2 - t;
    x = xx + t;
}

// Inverse (Gentleman-Sande) butterfly on x, y in [0, 2q):
// x' = x + y, y' = (x - y) * w, both in [0, 2q).
__device__ __forceinline__ void gs_bfly(u64 &x, u64 &y, u64 w, u64 ws, u64 q, u64 q2) {
    u64 t = x + q2 - y;
    x = csub(x + y, q2);
    y = mul_shoup_lazy(t, w, ws, q);
}

}  // namespace phx
