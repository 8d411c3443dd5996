// ntt_arith.h -- the two butterfly arithmetics of the NTT kernels.
//
// IntArith: 64-bit Shoup/Harvey lazy butterflies on u64 values (any q < 2^61),
//   values in [0, 6q) (forward) / [0, 3q) (inverse): the Shoup quotient skips
//   the low partial product (shoup_quot_lo), see ct and gs.
//
// FpArith: exact modular arithmetic on the FP64 pipe for q < 2^50 (the
//   MI355X VALU runs FP64 FMA at full rate, 32-bit integer multiplies at a
//   quarter of it). Values are integers held exactly in doubles, signed,
//   |v| < 2^53. The product y*w is split exactly as h + l with
//   h = fl(y*w), l = fma(y, w, -h); the quotient k = rint(h / q) comes from
//   the precomputed reciprocal (a Barrett reduction) and r = fma(-k, q, h) + l
//   is exact because |h - k q| and |r| are integers below 2^53.
//   Bounds (q < 2^50): reduced values |v| <= q/2 (+ eps). mulmod(y, w) for
//   |y| <= 5.3q: |h/q - k| <= 1/2 + |y| 2^-52 <= 1.83 and |l| <= |y w| 2^-53 <=
//   0.66q, so |r| <= 2.5q, and for |y| <= 3q, |r| <= 1.63q. The forward
//   butterfly reduces x on every other stage only (ct_lazy skips it): a reducing
//   stage outputs |v| <= q/2 + 2.24q < 3q (its inputs are below 4.63q), a lazy
//   one |v| <= 3q + 1.63q = 4.63q; every value stays below 2^52.3, and the last
//   stage of a register round always reduces.
#pragma once

#include "common.h"

namespace phx {

// x in [0, 2m) -> [0, m) for m < 2^63, given nm = -m mod 2^64: t = x - m as an
// add, and m added back where t went negative. Unlike a compare and select (or a
// 64-bit subtract), this needs no lane mask, whose write-then-read the gfx950
// VALU pads with wait states.
__device__ __forceinline__ u64 csub_n(u64 x, u64 m, u64 nm) {
    const u64 t = x + nm;
    return t + (m & (u64)((long long)t >> 63));
}

// floor(a b / 2^64) without the low partial product a0 b0: with a = a1 2^32 + a0
// and b = b1 2^32 + b0, a b - a0 b0 = (a1 b1 + hi(m) + hi(n)) 2^64 + lo(n) 2^32
// for m = a0 b1 and n = a1 b0 + lo(m), and a0 b0 < 2^64 adds at most one carry,
// so the result is the exact high word or one less (three 32 x 32 -> 64
// multiplies instead of four)
__device__ __forceinline__ u64 shoup_quot_lo(u64 a, u64 b) {
    const u32 a0 = (u32)a, a1 = (u32)(a >> 32), b0 = (u32)b, b1 = (u32)(b >> 32);
    const u64 m = (u64)a0 * b1;
    const u64 n = (u64)a1 * b0 + (u32)m;
    return (u64)a1 * b1 + (m >> 32) + (n >> 32);
}

struct IntArith {
    using T = u64;
    const TwPair *tw;
    u64 q, q2, q3;
    // -q, -2q, -3q and 3q + 1 mod 2^64, read from the device tables: the
    // subtractions below are additions of them (x - y = x + 3q + 1 + ~y - 3q), and
    // the optimiser, which does not see these values, keeps them additions
    u64 nq, nq2, nq3, q3p1;

    __device__ __forceinline__ static T from_canon(u64 v) { return v; }
    __device__ __forceinline__ static T from_bits(u64 b) { return b; }
    __device__ __forceinline__ static u64 to_bits(T v) { return v; }
    // forward output [0, 6q) -> [0, q)
    __device__ __forceinline__ u64 fwd_canon(T v) const {
        return csub_n(csub_n(csub_n(v, q3, nq3), q2, nq2), q, nq);
    }

    // Harvey's lazy forward butterfly (ct_bfly in common.h) with a cheaper Shoup
    // quotient: shoup_quot_lo omits the low partial product, so it is
    // floor(y w' / 2^64) or one less, and t = y w - quot q (as y w + quot (-q)) is
    // in [0, 3q) instead of [0, 2q) for any 64-bit y. x is reduced below 3q, so
    // both outputs stay below 6q < 2^64 (q < 2^61): one 32-bit multiply fewer
    // per butterfly.
    using W = TwPair;   // a twiddle as the butterflies take it
    __device__ __forceinline__ W twiddle(u32 idx) const { return tw[idx]; }
    __device__ __forceinline__ void ct_w(T &x, T &y, const W &w) const {
        const u64 t = y * w.w + shoup_quot_lo(y, w.ws) * nq;
        const u64 xx = csub_n(x, q3, nq3);
        y = xx + q3p1 + ~t;   // xx + 3q - t
        x = xx + t;
    }
    __device__ __forceinline__ void ct(T &x, T &y, u32 idx) const { ct_w(x, y, tw[idx]); }
    // the lazy integer butterfly keeps its range by the reduction above on every stage
    __device__ __forceinline__ void ct_lazy(T &x, T &y, u32 idx) const { ct(x, y, idx); }
    __device__ __forceinline__ void ct_lazy_w(T &x, T &y, const W &w) const { ct_w(x, y, w); }
    // the inverse butterfly (gs_bfly in common.h) on values below 3q: the sum is
    // reduced by 3q, and the product of x + 3q - y takes shoup_quot_lo, which
    // leaves it in [0, 3q) (as in ct)
    __device__ __forceinline__ void gs_w(T &x, T &y, const W &w) const {
        const u64 d = x + q3p1 + ~y;   // x + 3q - y
        x = csub_n(x + y, q3, nq3);
        y = d * w.w + shoup_quot_lo(d, w.ws) * nq;
    }
    __device__ __forceinline__ void gs(T &x, T &y, u32 idx) const { gs_w(x, y, tw[idx]); }
    // top inverse stage with both outputs scaled to canonical form (mul_shoup
    // takes any 64-bit input)
    __device__ __forceinline__ void gs_last(T &x, T &y, u64 s0, u64 s0s, u64 s1, u64 s1s) const {
        const u64 sum = x + y, dif = x + q3 - y;
        x = mul_shoup(sum, s0, s0s, q);
        y = mul_shoup(dif, s1, s1s, q);
    }
    __device__ __forceinline__ static u64 inv_canon(T v) { return v; }
};

struct FpArith {
    using T = double;
    const double *tw;
    double q, qinv;

    __device__ __forceinline__ static T from_canon(u64 v) {
        // exact for v < 2^52: set the exponent of 2^52 and subtract it
        return __longlong_as_double((long long)(v | 0x4330000000000000ull)) - 0x1p52;
    }
    __device__ __forceinline__ static T from_bits(u64 b) { return __longlong_as_double((long long)b); }
    __device__ __forceinline__ static u64 to_bits(T v) { return (u64)__double_as_longlong(v); }

    __device__ __forceinline__ T red(T x) const {
#pragma clang fp contract(off)
        const T k = __builtin_rint(x * qinv);
        return __builtin_fma(-k, q, x);
    }
    __device__ __forceinline__ T mulmod(T y, T w) const {
#pragma clang fp contract(off)
        const T h = y * w;
        const T l = __builtin_fma(y, w, -h);
        const T k = __builtin_rint(h * qinv);
        return __builtin_fma(-k, q, h) + l;
    }
    // canonical residue in [0, q) of an exact integer |v| < 2^53, as u64
    __device__ __forceinline__ u64 canon(T v) const {
#pragma clang fp contract(off)
        const T k = __builtin_floor(v * qinv);
        T r = __builtin_fma(-k, q, v);
        r = r < 0.0 ? r + q : r;
        r = r >= q ? r - q : r;
        return (u64)__double_as_longlong(r + 0x1p52) & 0x000fffffffffffffull;
    }
    __device__ __forceinline__ u64 fwd_canon(T v) const { return canon(v); }

    using W = double;
    __device__ __forceinline__ W twiddle(u32 idx) const { return tw[idx]; }
    __device__ __forceinline__ void ct_w(T &x, T &y, W w) const {
#pragma clang fp contract(off)
        const T t = mulmod(y, w);
        const T xr = red(x);
        x = xr + t;
        y = xr - t;
    }
    __device__ __forceinline__ void ct(T &x, T &y, u32 idx) const { ct_w(x, y, tw[idx]); }
    // the forward butterfly without the reduction of x (see the bounds above)
    __device__ __forceinline__ void ct_lazy_w(T &x, T &y, W w) const {
#pragma clang fp contract(off)
        const T t = mulmod(y, w);
        const T xr = x;
        x = xr + t;
        y = xr - t;
    }
    __device__ __forceinline__ void ct_lazy(T &x, T &y, u32 idx) const { ct_lazy_w(x, y, tw[idx]); }
    __device__ __forceinline__ void gs_w(T &x, T &y, W w) const {
#pragma clang fp contract(off)
        const T s = x + y, d = x - y;
        x = red(s);
        y = mulmod(d, w);
    }
    __device__ __forceinline__ void gs(T &x, T &y, u32 idx) const { gs_w(x, y, tw[idx]); }
    __device__ __forceinline__ void gs_last(T &x, T &y, u64 s0, u64, u64 s1, u64) const {
#pragma clang fp contract(off)
        const T s = x + y, d = x - y;
        x = mulmod(s, from_canon(s0));
        y = mulmod(d, from_canon(s1));
    }
    __device__ __forceinline__ u64 inv_canon(T v) const { return canon(v); }
};

// Karatsuba tensor of one coefficient pair (tensor_prod_2x2_rns_poly,
// src/polymath.cu:463-498): r0 = a0 b0, r2 = a1 b1, r1 = (a0 + a1)(b0 + b1) - r0 - r2
// mod q, canonical, from canonical inputs. A row whose DevModulus carries fl(1/q)
// (q < 2^50) multiplies exactly on the FP64 pipe: operands below 2q give |h/q - k|
// <= 1/2 + 4q 2^-52 <= 1.5 and |l| <= 4q^2 2^-53 <= q/2, so each product is below
// 2q and the difference below 4.5q, all exact; other rows use 128-bit Barrett.
__device__ __forceinline__ void tensor_coeff(u64 a0, u64 a1, u64 b0, u64 b1, const DevModulus &m, u64 &r0, u64 &r1,
                                             u64 &r2) {
    if (m.pad) {
#pragma clang fp contract(off)
        const FpArith f{nullptr, (double)m.value, __longlong_as_double((long long)m.pad)};
        const double x0 = FpArith::from_canon(a0), x1 = FpArith::from_canon(a1);
        const double y0 = FpArith::from_canon(b0), y1 = FpArith::from_canon(b1);
        const double p0 = f.mulmod(x0, y0), p2 = f.mulmod(x1, y1);
        const double p1 = f.mulmod(x0 + x1, y0 + y1);
        r0 = f.canon(p0);
        r2 = f.canon(p2);
        r1 = f.canon(p1 - p0 - p2);
        return;
    }
    const u64 q = m.value;
    r0 = mul_mod(a0, b0, m);
    r2 = mul_mod(a1, b1, m);
    r1 = mul_mod(add_mod(a0, a1, q), add_mod(b0, b1, q), m);
    r1 = sub_mod(sub_mod(r1, r0, q), r2, q);
}

}  // namespace phx
