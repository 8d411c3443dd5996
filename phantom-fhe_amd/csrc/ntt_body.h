// ntt_body.h -- the pass bodies of the two-pass NTT/INTT, shared by the
// translation units that instantiate its kernels (ntt.hip: launches over a limb
// map; ntt_jobs.hip: work-list launches; ntt_fused.hip: one workgroup per limb).
//
// A body runs on NT threads x 16 registers, a tile of 16 NT coefficients: the
// column passes see TC = 16 NT / R columns of all R rows, the row passes
// NT / 16 rows of 256 coefficients. NT is the tile shape the autotuner picks
// (by default 256 for the row passes, 16 rows; the column passes take 256 or 512
// by ring degree and launch size, cols_threads below).
#pragma once

#include <algorithm>
#include <type_traits>

#include "ntt.h"
#include "tune.h"

namespace phx {
namespace {

// a store policy that loads operands of its own (Pre, prefetch, apply)
template <class S, class = void>
struct HasPrefetch : std::false_type {};
template <class S>
struct HasPrefetch<S, std::void_t<typename S::Pre>> : std::true_type {};

// LDS address helpers (u64 units). Column tiles pad each row of TC columns
// by one slot; row tiles pad every 16 coefficients by one slot so the
// 16-apart accesses of the radix-16 exchange fall in distinct banks.
template <int TC>
__device__ __forceinline__ u32 col_addr(u32 r, u32 u) { return r * (TC + 1) + u; }
__device__ __forceinline__ u32 row_addr(u32 u, u32 j) { return u * 272 + j + (j >> 4); }

__device__ __forceinline__ void limb_scale(const NttTables &T, const LimbMap &M, u32 limb, u32 row,
                                           u64 &s0, u64 &s0s, u64 &s1, u64 &s1s) {
    const auto c = gmem(M.scale ? M.scale + 4 * (u64)limb : T.inv_consts + 4 * (u64)row);
    s0 = c[0];
    s0s = c[1];
    s1 = c[2];
    s1s = c[3];
}

// the integer arithmetic of table row `row`: forward (tw) or inverse (itw) twiddles
__device__ __forceinline__ IntArith int_arith(const NttTables &T, const TwPair *tw, u32 row, u64 q) {
    const u64 *c = T.int_consts + 4 * (u64)row;
    return IntArith{tw + (u64)row * T.N, q, q << 1, 3 * q, c[0], c[1], c[2], c[3]};
}
__device__ __forceinline__ IntArith int_fwd(const NttTables &T, u32 row, u64 q) { return int_arith(T, T.tw, row, q); }
__device__ __forceinline__ IntArith int_inv(const NttTables &T, u32 row, u64 q) { return int_arith(T, T.itw, row, q); }
__device__ __forceinline__ FpArith fp_arith(const double *tw, const DevModulus &m) {
    return FpArith{tw, (double)m.value, __longlong_as_double((long long)m.pad)};
}

// ---------------------------------------------------------------------------
// Forward pass 1 body: the first SR stages (pair distances N/2 .. 256) on a
// tile of TC columns x R rows.
// ---------------------------------------------------------------------------
template <int SR, int NT, class A, class Load>
__device__ __forceinline__ void fwd_cols_body(const A &a, u64 *lds, u64 *data, u32 logN, u32 c0, u32 limb, u32 z,
                                              const DevModulus &md, const Load &ld, const u32 t) {
    using T = typename A::T;
    constexpr int R = 1 << SR;
    constexpr int TC = (NT * kNttRegs) / R;
    if constexpr (SR > 8) {
        // N = 2^17: three rounds of 4 + 4 + 1 stages over 512 rows x TC columns
        // (8 at 256 threads, 16 at 512).
        // Row bits (high to low): 4 for round A, 4 for round B, 1 for round C.
        static_assert(SR == 9, "the three-round column pass covers 512 rows");
        T v[16];
        const u32 uc = t % TC, rest = t / TC;   // rest < 32
        // round A: rows rest + 32 e, distances N/2 .. 2^13
        u32 baseA[1] = {rest * 256 + c0 + uc};
#pragma unroll
        for (int e = 0; e < 16; ++e) v[e] = A::from_canon(ld(data, baseA[0] + ((u32)e << 13), limb, z, md));
        fwd_round<4, 13, A, true>(v, baseA, logN, a);
#pragma unroll
        for (int e = 0; e < 16; ++e) lds[col_addr<TC>(((u32)e << 5) + rest, uc)] = A::to_bits(v[e]);
        __syncthreads();
        // round B: rows 32 hi + lo + 2 e (hi = rest >> 1, lo = rest & 1), distances 2^12 .. 2^9
        const u32 rb = (rest >> 1) * 32 + (rest & 1);
        u32 baseB[1] = {rb * 256 + c0 + uc};
#pragma unroll
        for (int e = 0; e < 16; ++e) v[e] = A::from_bits(lds[col_addr<TC>(rb + ((u32)e << 1), uc)]);
        fwd_round<4, 9>(v, baseB, logN, a);
        __syncthreads();
#pragma unroll
        for (int e = 0; e < 16; ++e) lds[col_addr<TC>(rb + ((u32)e << 1), uc)] = A::to_bits(v[e]);
        __syncthreads();
        // round C: row pairs (2p, 2p + 1) with p = 32 b + rest, distance 2^8
        u32 baseC[8];
#pragma unroll
        for (int b = 0; b < 8; ++b) {
            const u32 r2 = 2 * ((u32)b * 32 + rest);
            baseC[b] = r2 * 256 + c0 + uc;
            v[2 * b] = A::from_bits(lds[col_addr<TC>(r2, uc)]);
            v[2 * b + 1] = A::from_bits(lds[col_addr<TC>(r2 + 1, uc)]);
        }
        fwd_round<1, 8>(v, baseC, logN, a);
#pragma unroll
        for (int b = 0; b < 8; ++b) {
            st_mid(data + baseC[b], A::to_bits(v[2 * b]));
            st_mid(data + baseC[b] + 256, A::to_bits(v[2 * b + 1]));
        }
        return;
    } else {
    constexpr int NA = SR < 4 ? SR : 4;
    constexpr int NB = SR - NA;
    constexpr int BA = 16 >> NA;
    T v[16];
    u32 baseA[BA], uA[BA], lowA[BA];
#pragma unroll
    for (int b = 0; b < BA; ++b) {
        const u32 kap = t + NT * b;
        uA[b] = kap % TC;
        lowA[b] = kap / TC;
        baseA[b] = lowA[b] * 256 + c0 + uA[b];
    }
#pragma unroll
    for (int b = 0; b < BA; ++b)
#pragma unroll
        for (int e = 0; e < (1 << NA); ++e)
            v[b * (1 << NA) + e] = A::from_canon(ld(data, baseA[b] + ((u32)e << (8 + NB)), limb, z, md));

    fwd_round<NA, 8 + NB, A, true>(v, baseA, logN, a);

    if constexpr (NB > 0) {
#pragma unroll
        for (int b = 0; b < BA; ++b)
#pragma unroll
            for (int e = 0; e < (1 << NA); ++e)
                lds[col_addr<TC>(((u32)e << NB) + lowA[b], uA[b])] = A::to_bits(v[b * (1 << NA) + e]);
        constexpr int BB = 16 >> NB;
        u32 baseB[BB], uB[BB], hpB[BB];
#pragma unroll
        for (int b = 0; b < BB; ++b) {
            const u32 kap = t + NT * b;
            uB[b] = kap % TC;
            hpB[b] = kap / TC;
            baseB[b] = (hpB[b] << NB) * 256 + c0 + uB[b];
        }
        // a full radix-16 second round loads its twiddles while the values cross
        // the barrier
        typename A::W wB[15];
        if constexpr (NB == 4) load_round_tw<8>(a, baseB[0], logN, wB);
        __syncthreads();
#pragma unroll
        for (int b = 0; b < BB; ++b)
#pragma unroll
            for (int e = 0; e < (1 << NB); ++e)
                v[b * (1 << NB) + e] = A::from_bits(lds[col_addr<TC>((hpB[b] << NB) + e, uB[b])]);
        if constexpr (NB == 4)
            fwd_round_w(v, wB, a);
        else
            fwd_round<NB, 8>(v, baseB, logN, a);
#pragma unroll
        for (int b = 0; b < BB; ++b)
#pragma unroll
            for (int e = 0; e < (1 << NB); ++e)
                st_mid(data + baseB[b] + ((u32)e << 8), A::to_bits(v[b * (1 << NB) + e]));
    } else {
#pragma unroll
        for (int b = 0; b < BA; ++b)
#pragma unroll
            for (int e = 0; e < (1 << NA); ++e)
                st_mid(data + baseA[b] + ((u32)e << 8), A::to_bits(v[b * (1 << NA) + e]));
    }
    }
}

// Forward pass 2 core: the last 8 stages (distances 128 .. 1) on 16 rows;
// returns the canonical outputs at tile index t + 256 k in out[k]. The LDS
// region is free again after the caller's next barrier.
// FromLds: the tile's values are in LDS (row_addr layout) instead of in data.
struct NoPrefetch {
    __device__ __forceinline__ void operator()() const {}
};

// FromLds: the caller loaded round 1's twiddles (w1) across its barrier.
template <int NT, class A, bool FromLds = false, class Pre = NoPrefetch>
__device__ __forceinline__ void fwd_rows_core(const A &a, u64 *lds, const u64 *data, u32 logN, u32 r0, const u32 t,
                                              u64 (&out)[16], const Pre &pre = Pre{},
                                              const typename A::W *w1 = nullptr) {
    using T = typename A::T;
    const u32 l = t & 15, u = t >> 4;
    T v[16];
    const u32 baseA = (r0 + u) * 256 + l;
    // round 1's twiddles load ahead of the values (from L2, they arrive first; the
    // waits complete in issue order, so the round's first stages can start while
    // the values' later loads are still in flight)
    typename A::W w1g[15];
    if constexpr (!FromLds) load_round_tw<4>(a, baseA, logN, w1g);
#pragma unroll
    for (int e = 0; e < 16; ++e)
        v[e] = A::from_bits(FromLds ? lds[row_addr(u, ((u32)e << 4) + l)] : ld_mid(data + baseA + ((u32)e << 4)));
    if constexpr (FromLds)
        fwd_round_w(v, w1, a);
    else
        fwd_round_w(v, w1g, a);
#pragma unroll
    for (int e = 0; e < 16; ++e) lds[row_addr(u, ((u32)e << 4) + l)] = A::to_bits(v[e]);
    // round 2's twiddles load while the values cross the barrier
    const u32 hp = l;
    u32 baseB[1] = {(r0 + u) * 256 + (hp << 4)};
    typename A::W w2[15];
    load_round_tw<0>(a, baseB[0], logN, w2);
    __syncthreads();
#pragma unroll
    for (int e = 0; e < 16; ++e) v[e] = A::from_bits(lds[row_addr(u, (hp << 4) + e)]);
    fwd_round_w(v, w2, a);
    __syncthreads();
#pragma unroll
    for (int e = 0; e < 16; ++e) lds[row_addr(u, (hp << 4) + e)] = a.fwd_canon(v[e]);
    pre();   // the store policy's loads, in flight across the barrier
    __syncthreads();
#pragma unroll
    for (int k = 0; k < 16; ++k) {
        const u32 idx = t + NT * k;
        out[k] = lds[row_addr(idx >> 8, idx & 255)];
    }
}

// Forward pass 2 over a tile whose column pass ran only its first register
// round (the column pass fused with FastBConv, bconv_ntt.hip): the column
// pass's remaining SR - 4 stages (distances 2^(SR-5+8) .. 2^8, i.e. within
// groups of 2^(SR-4) consecutive rows) on the tile's columns, then the last 8
// stages as fwd_rows_core. Thread t holds column t % 256 of rows
// r0 + 16 (t / 256) + e (NT = 256 or 512: 16 rows per 256 threads).
template <int SR, int NT, class A, class Pre = NoPrefetch>
__device__ __forceinline__ void fwd_rows_core_rb(const A &a, u64 *lds, const u64 *data, u32 logN, u32 r0,
                                                 const u32 t, u64 (&out)[16], const Pre &pre = Pre{}) {
    static_assert(NT % 256 == 0 && SR >= 4 && SR <= 8, "column-first row pass: 256 or 512 threads, N = 2^12 .. 2^16");
    using T = typename A::T;
    constexpr int NB = SR - 4;
    if constexpr (NB == 0) {
        // N = 2^12: the column pass has only the round the fused launch ran
        fwd_rows_core<NT, A, false, Pre>(a, lds, data, logN, r0, t, out, pre);
        return;
    }
    constexpr int BB = 16 >> NB;
    const u32 cc = t & 255, rb = r0 + 16 * (t >> 8);
    T v[16];
    u32 base[BB];
#pragma unroll
    for (int b = 0; b < BB; ++b) base[b] = (rb + ((u32)b << NB)) * 256 + cc;
    // a full radix-16 column round (N = 2^16) loads its twiddles ahead of the values
    typename A::W wB[15];
    if constexpr (NB == 4) load_round_tw<8>(a, base[0], logN, wB);
#pragma unroll
    for (int e = 0; e < 16; ++e) v[e] = A::from_bits(ld_mid(data + (rb + (u32)e) * 256 + cc));
    if constexpr (NB == 4)
        fwd_round_w(v, wB, a);
    else
        fwd_round<NB, 8>(v, base, logN, a);
#pragma unroll
    for (int e = 0; e < 16; ++e) lds[row_addr(16 * (t >> 8) + (u32)e, cc)] = A::to_bits(v[e]);
    // the row pass's round 1 twiddles load while the values cross the barrier
    typename A::W w1[15];
    load_round_tw<4>(a, (r0 + (t >> 4)) * 256 + (t & 15), logN, w1);
    __syncthreads();
    fwd_rows_core<NT, A, true, Pre>(a, lds, data, logN, r0, t, out, pre, w1);
}

template <int SR, int NT, class A, class Store>
__device__ __forceinline__ void fwd_rows_body_rb(const A &a, u64 *lds, u64 *data, u32 logN, u32 r0, u32 limb, u32 z,
                                                 const DevModulus &md, const Store &st, const u32 t) {
    u64 out[16];
    if constexpr (HasPrefetch<Store>::value) {
        typename Store::Pre pf[16];
        const auto pre = [&] {
#pragma unroll
            for (int k = 0; k < 16; ++k) pf[k] = st.prefetch(r0 * 256 + t + NT * k, limb, z);
        };
        fwd_rows_core_rb<SR, NT>(a, lds, data, logN, r0, t, out, pre);
#pragma unroll
        for (int k = 0; k < 16; ++k) st.apply(data, r0 * 256 + t + NT * k, out[k], pf[k], limb, z, md);
    } else {
        fwd_rows_core_rb<SR, NT>(a, lds, data, logN, r0, t, out);
#pragma unroll
        for (int k = 0; k < 16; ++k) st(data, r0 * 256 + t + NT * k, out[k], limb, z, md);
    }
}

// Forward pass 2 body: the core, then the store policy per output (a policy
// with a prefetch has its operands loaded while the outputs wait in LDS).
template <int NT, class A, class Store>
__device__ __forceinline__ void fwd_rows_body(const A &a, u64 *lds, u64 *data, u32 logN, u32 r0, u32 limb, u32 z,
                                              const DevModulus &md, const Store &st, const u32 t) {
    u64 out[16];
    if constexpr (HasPrefetch<Store>::value) {
        typename Store::Pre pf[16];
        const auto pre = [&] {
#pragma unroll
            for (int k = 0; k < 16; ++k) pf[k] = st.prefetch(r0 * 256 + t + NT * k, limb, z);
        };
        fwd_rows_core<NT, A, false>(a, lds, data, logN, r0, t, out, pre);
#pragma unroll
        for (int k = 0; k < 16; ++k) st.apply(data, r0 * 256 + t + NT * k, out[k], pf[k], limb, z, md);
    } else {
        fwd_rows_core<NT>(a, lds, data, logN, r0, t, out);
#pragma unroll
        for (int k = 0; k < 16; ++k) st(data, r0 * 256 + t + NT * k, out[k], limb, z, md);
    }
}

// Inverse pass 1 body: the first 8 stages (distances 1 .. 128) on 16 rows.
template <int NT, class A, class Load>
__device__ __forceinline__ void inv_rows_body(const A &a, u64 *lds, u64 *data, const u64 *src, u32 logN, u32 r0,
                                              u32 limb, u32 z, const DevModulus &md, const Load &ld, const u32 t) {
    using T = typename A::T;
#pragma unroll
    for (int k = 0; k < 16; ++k) {
        const u32 idx = t + NT * k;
        lds[row_addr(idx >> 8, idx & 255)] = ld(src, r0 * 256 + idx, limb, z, md);
    }
    // each register round's twiddles load while the values cross the barrier before it
    const u32 l = t & 15, u = t >> 4;
    T v[16];
    u32 baseB[1] = {(r0 + u) * 256 + (l << 4)};
    typename A::W w1[15];
    load_round_tw<0>(a, baseB[0], logN, w1);
    __syncthreads();
#pragma unroll
    for (int e = 0; e < 16; ++e) v[e] = A::from_canon(lds[row_addr(u, (l << 4) + e)]);
    inv_round_w(v, w1, a);
    __syncthreads();
#pragma unroll
    for (int e = 0; e < 16; ++e) lds[row_addr(u, (l << 4) + e)] = A::to_bits(v[e]);
    u32 baseA[1] = {(r0 + u) * 256 + l};
    typename A::W w2[15];
    load_round_tw<4>(a, baseA[0], logN, w2);
    __syncthreads();
#pragma unroll
    for (int e = 0; e < 16; ++e) v[e] = A::from_bits(lds[row_addr(u, ((u32)e << 4) + l)]);
    inv_round_w(v, w2, a);
#pragma unroll
    for (int e = 0; e < 16; ++e) st_mid(data + baseA[0] + ((u32)e << 4), A::to_bits(v[e]));
}

// Inverse pass 2 body: the remaining SR stages (distances 256 .. N/2) on
// columns; the top stage folds n^-1 (and an optional per-limb scale).
template <int SR, int NT, class A, class Store>
__device__ __forceinline__ void inv_cols_body(const A &a, u64 *lds, u64 *data, u32 logN, u32 c0, u32 limb, u32 z,
                                              const DevModulus &md, const Store &st, u64 s0, u64 s0s, u64 s1,
                                              u64 s1s, const u32 t) {
    using T = typename A::T;
    constexpr int R = 1 << SR;
    constexpr int TC = (NT * kNttRegs) / R;
    if constexpr (SR > 8) {
        // N = 2^17: the forward column pass's rounds in reverse (C, B, then A with
        // the top stage scaled)
        static_assert(SR == 9, "the three-round column pass covers 512 rows");
        T v[16];
        const u32 uc = t % TC, rest = t / TC;
        u32 baseC[8];
#pragma unroll
        for (int b = 0; b < 8; ++b) {
            const u32 r2 = 2 * ((u32)b * 32 + rest);
            baseC[b] = r2 * 256 + c0 + uc;
            v[2 * b] = A::from_bits(ld_mid(data + baseC[b]));
            v[2 * b + 1] = A::from_bits(ld_mid(data + baseC[b] + 256));
        }
        inv_round<1, 8>(v, baseC, logN, a);
#pragma unroll
        for (int b = 0; b < 8; ++b) {
            const u32 r2 = 2 * ((u32)b * 32 + rest);
            lds[col_addr<TC>(r2, uc)] = A::to_bits(v[2 * b]);
            lds[col_addr<TC>(r2 + 1, uc)] = A::to_bits(v[2 * b + 1]);
        }
        __syncthreads();
        const u32 rb = (rest >> 1) * 32 + (rest & 1);
        u32 baseB[1] = {rb * 256 + c0 + uc};
#pragma unroll
        for (int e = 0; e < 16; ++e) v[e] = A::from_bits(lds[col_addr<TC>(rb + ((u32)e << 1), uc)]);
        inv_round<4, 9>(v, baseB, logN, a);
        __syncthreads();
#pragma unroll
        for (int e = 0; e < 16; ++e) lds[col_addr<TC>(rb + ((u32)e << 1), uc)] = A::to_bits(v[e]);
        __syncthreads();
        u32 baseA[1] = {rest * 256 + c0 + uc};
#pragma unroll
        for (int e = 0; e < 16; ++e) v[e] = A::from_bits(lds[col_addr<TC>(((u32)e << 5) + rest, uc)]);
        inv_round_last<4>(v, baseA, 13, logN, a, s0, s0s, s1, s1s);
#pragma unroll
        for (int e = 0; e < 16; ++e) st(data, baseA[0] + ((u32)e << 13), a.inv_canon(v[e]), limb, z, md);
        return;
    } else {
    constexpr int NA = SR < 4 ? SR : 4;
    constexpr int NB = SR - NA;
    constexpr int BA = 16 >> NA;
    T v[16];
    u32 baseA[BA], uA[BA], lowA[BA];
#pragma unroll
    for (int b = 0; b < BA; ++b) {
        const u32 kap = t + NT * b;
        uA[b] = kap % TC;
        lowA[b] = kap / TC;
        baseA[b] = lowA[b] * 256 + c0 + uA[b];
    }
    if constexpr (NB > 0) {
        constexpr int BB = 16 >> NB;
        u32 baseB[BB], uB[BB], hpB[BB];
#pragma unroll
        for (int b = 0; b < BB; ++b) {
            const u32 kap = t + NT * b;
            uB[b] = kap % TC;
            hpB[b] = kap / TC;
            baseB[b] = (hpB[b] << NB) * 256 + c0 + uB[b];
        }
        // a full radix-16 first round (N = 2^16) loads its twiddles ahead of the
        // values (from L2, they arrive first, and waits complete in issue order)
        typename A::W wB[15];
        if constexpr (NB == 4) load_round_tw<8>(a, baseB[0], logN, wB);
#pragma unroll
        for (int b = 0; b < BB; ++b)
#pragma unroll
            for (int e = 0; e < (1 << NB); ++e)
                v[b * (1 << NB) + e] = A::from_bits(ld_mid(data + baseB[b] + ((u32)e << 8)));
        if constexpr (NB == 4)
            inv_round_w(v, wB, a);
        else
            inv_round<NB, 8>(v, baseB, logN, a);
#pragma unroll
        for (int b = 0; b < BB; ++b)
#pragma unroll
            for (int e = 0; e < (1 << NB); ++e)
                lds[col_addr<TC>((hpB[b] << NB) + e, uB[b])] = A::to_bits(v[b * (1 << NB) + e]);
        __syncthreads();
#pragma unroll
        for (int b = 0; b < BA; ++b)
#pragma unroll
            for (int e = 0; e < (1 << NA); ++e)
                v[b * (1 << NA) + e] = A::from_bits(lds[col_addr<TC>(((u32)e << NB) + lowA[b], uA[b])]);
    } else {
#pragma unroll
        for (int b = 0; b < BA; ++b)
#pragma unroll
            for (int e = 0; e < (1 << NA); ++e)
                v[b * (1 << NA) + e] = A::from_bits(ld_mid(data + baseA[b] + ((u32)e << 8)));
    }
    inv_round_last<NA>(v, baseA, 8 + NB, logN, a, s0, s0s, s1, s1s);
#pragma unroll
    for (int b = 0; b < BA; ++b)
#pragma unroll
        for (int e = 0; e < (1 << NA); ++e)
            st(data, baseA[b] + ((u32)e << (8 + NB)), a.inv_canon(v[b * (1 << NA) + e]), limb, z, md);
    }
}


template <class F>
void dispatch_sr(u32 logN, F &&f) {
    switch (logN) {
        case 12: f(std::integral_constant<int, 4>{}); break;
        case 13: f(std::integral_constant<int, 5>{}); break;
        case 14: f(std::integral_constant<int, 6>{}); break;
        case 15: f(std::integral_constant<int, 7>{}); break;
        case 16: f(std::integral_constant<int, 8>{}); break;
        case 17: f(std::integral_constant<int, 9>{}); break;
        default: fail(kInvalidArgument, "unsupported polynomial degree for the NTT (2^11 .. 2^17)");
    }
}

// policies that carry per-entry operands of a work-list launch (kBind) are
// bound to their entry once per workgroup
template <class P, class = void>
struct BindsJob : std::false_type {};
template <class P>
struct BindsJob<P, std::void_t<decltype(P::kBind)>> : std::true_type {};

template <class P>
__device__ __forceinline__ auto bind_job(const P &p, const NttJob &J) {
    if constexpr (BindsJob<P>::value)
        return p.bind(J);
    else
        return p;
}

// tile shapes a launch may take: 16 NT <= N (one tile fits in a limb); the
// context's choice, else the process-wide switch, else 256
inline u32 pick_threads(int ctx_choice, const char *key, u32 N) {
    int nt = ctx_choice >= 0 ? ctx_choice : tune_value(key);
    if (nt != 128 && nt != 256 && nt != 512) nt = 256;
    while ((u32)nt * kNttRegs > N && nt > 128) nt >>= 1;
    return (u32)nt;
}
// Column tiles of 256 threads hold 16 columns at N = 2^16, so the pass reads and
// writes its limb in 128-byte pieces 2 KiB apart; 512 threads (32 columns, 256-byte
// pieces) move the same bytes faster (752-limb forward pair 0.340 -> 0.331 ms) but
// halve the workgroups, which costs launches over few limbs (the plain HMult+relin's
// 24..106-limb passes: -1%). At N = 2^17 a 256-thread tile holds 8 columns (64-byte
// pieces) and 512 threads are faster for any launch (384 limbs: 0.490 -> 0.378 ms).
// Choice 0 takes 512 threads at N = 2^17, and at N = 2^16 for launches of at least
// kWideColsMin limb-transforms; 256 otherwise.
constexpr u64 kWideColsMin = 256;
inline u32 cols_threads(const NttTables &T, u64 transforms) {
    int choice = T.cols_nt >= 0 ? T.cols_nt : tune_value("ntt_cols_threads");
    if (choice == 0) choice = T.logN >= 17 || (T.logN == 16 && transforms >= kWideColsMin) ? 512 : 256;
    const u32 nt = pick_threads(choice, "ntt_cols_threads", T.N);
    // column passes run on 256 or 512 threads, on 128 only where N = 2^11 needs it
    return nt == 128 && 256 * kNttRegs <= T.N ? 256 : nt;
}
inline u32 rows_threads(const NttTables &T) { return pick_threads(T.rows_nt, "ntt_rows_threads", T.N); }

// column passes: (SR, NT) for log2 N and the tile's threads. N = 2^11 (8 rows of
// 256) runs one 128-thread tile per limb: 16 registers x 128 threads = the limb
template <class F>
void dispatch_cols(u32 logN, u32 nt, F &&f) {
    if (logN == 11) {
        f(std::integral_constant<int, 3>{}, std::integral_constant<int, 128>{});
        return;
    }
    dispatch_sr(logN, [&](auto sr) {
        if (nt == 512)
            f(sr, std::integral_constant<int, 512>{});
        else
            f(sr, std::integral_constant<int, 256>{});
    });
}
// row passes after the FastBConv-fused column pass: (SR, NT) for N = 2^12 .. 2^16
// and 256 or 512 threads
template <class F>
void dispatch_rows_rb(u32 logN, u32 nt, F &&f) {
    auto go = [&](auto sr) {
        if (nt == 512)
            f(sr, std::integral_constant<int, 512>{});
        else
            f(sr, std::integral_constant<int, 256>{});
    };
    switch (logN) {
        case 12: go(std::integral_constant<int, 4>{}); break;
        case 13: go(std::integral_constant<int, 5>{}); break;
        case 14: go(std::integral_constant<int, 6>{}); break;
        case 15: go(std::integral_constant<int, 7>{}); break;
        case 16: go(std::integral_constant<int, 8>{}); break;
        default: fail(kInvalidArgument, "fused FastBConv row pass: unsupported polynomial degree");
    }
}

template <class F>
void dispatch_rows_nt(u32 nt, F &&f) {
    if (nt == 512)
        f(std::integral_constant<int, 512>{});
    else if (nt == 128)
        f(std::integral_constant<int, 128>{});
    else
        f(std::integral_constant<int, 256>{});
}

}  // namespace

// one-workgroup-per-limb forms (ntt_fused.hip) of the launches that take them
void fused_forward(u32 G, const NttTables &T, const LimbMap &M, u32 nlimbs, u32 npolys, const LoadPlain &ld,
                   const StorePlain &st, hipStream_t s);
void fused_forward(u32 G, const NttTables &T, const LimbMap &M, u32 nlimbs, u32 npolys, const LoadPlain &ld,
                   const StoreModDown &st, hipStream_t s);
void fused_forward(u32 G, const NttTables &T, const LimbMap &M, u32 nlimbs, u32 npolys, const LoadRescale &ld,
                   const StoreRescale &st, hipStream_t s);
void fused_inverse(u32 G, const NttTables &T, const LimbMap &M, u32 nlimbs, u32 npolys, hipStream_t s);

}  // namespace phx
