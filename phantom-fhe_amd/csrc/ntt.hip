// ntt.hip -- two-pass NTT/INTT kernels and their launchers.
//
// Each kernel picks its butterfly arithmetic per limb (uniform across the
// workgroup): FP64 for primes below 2^50, 64-bit Shoup otherwise. Between the
// two passes the tile is stored in the arithmetic's own form (lazy u64 or
// double bits); the first load and the last store are canonical u64.
#include "ntt.h"
#include "tune.h"

#include <type_traits>

namespace phx {

namespace {

// LDS address helpers (u64 units). Column tiles pad each row of TC columns
// by one slot; row tiles pad every 16 coefficients by one slot so the
// 16-apart accesses of the radix-16 exchange fall in distinct banks.
template <int TC>
__device__ __forceinline__ u32 col_addr(u32 r, u32 u) { return r * (TC + 1) + u; }
__device__ __forceinline__ u32 row_addr(u32 u, u32 j) { return u * 272 + j + (j >> 4); }

__device__ __forceinline__ void limb_scale(const NttTables &T, const LimbMap &M, u32 limb, u32 row,
                                           u64 &s0, u64 &s0s, u64 &s1, u64 &s1s) {
    const u64 *c = M.scale ? M.scale + 4 * (u64)limb : T.inv_consts + 4 * (u64)row;
    s0 = c[0];
    s0s = c[1];
    s1 = c[2];
    s1s = c[3];
}

__device__ __forceinline__ IntArith int_arith(const TwPair *tw, u64 q) { return IntArith{tw, q, q << 1}; }
__device__ __forceinline__ FpArith fp_arith(const double *tw, const DevModulus &m) {
    return FpArith{tw, (double)m.value, __longlong_as_double((long long)m.pad)};
}

// ---------------------------------------------------------------------------
// Forward pass 1 body: the first SR stages (pair distances N/2 .. 256) on a
// tile of TC columns x R rows.
// ---------------------------------------------------------------------------
template <int SR, class A, class Load>
__device__ __forceinline__ void fwd_cols_body(const A &a, u64 *lds, u64 *data, u32 logN, u32 c0, u32 limb, u32 z,
                                              const DevModulus &md, const Load &ld, const u32 t) {
    using T = typename A::T;
    constexpr int R = 1 << SR;
    constexpr int TC = kNttTile / R;
    if constexpr (SR > 8) {
        // N = 2^17: three rounds of 4 + 4 + 1 stages over 512 rows x 8 columns.
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
            st_stream(data + baseC[b], A::to_bits(v[2 * b]));
            st_stream(data + baseC[b] + 256, A::to_bits(v[2 * b + 1]));
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
        const u32 kap = t + kNttThreads * b;
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
        __syncthreads();
        constexpr int BB = 16 >> NB;
        u32 baseB[BB], uB[BB], hpB[BB];
#pragma unroll
        for (int b = 0; b < BB; ++b) {
            const u32 kap = t + kNttThreads * b;
            uB[b] = kap % TC;
            hpB[b] = kap / TC;
            baseB[b] = (hpB[b] << NB) * 256 + c0 + uB[b];
        }
#pragma unroll
        for (int b = 0; b < BB; ++b)
#pragma unroll
            for (int e = 0; e < (1 << NB); ++e)
                v[b * (1 << NB) + e] = A::from_bits(lds[col_addr<TC>((hpB[b] << NB) + e, uB[b])]);
        fwd_round<NB, 8>(v, baseB, logN, a);
#pragma unroll
        for (int b = 0; b < BB; ++b)
#pragma unroll
            for (int e = 0; e < (1 << NB); ++e)
                st_stream(data + baseB[b] + ((u32)e << 8), A::to_bits(v[b * (1 << NB) + e]));
    } else {
#pragma unroll
        for (int b = 0; b < BA; ++b)
#pragma unroll
            for (int e = 0; e < (1 << NA); ++e)
                st_stream(data + baseA[b] + ((u32)e << 8), A::to_bits(v[b * (1 << NA) + e]));
    }
    }
}

// Forward pass 2 core: the last 8 stages (distances 128 .. 1) on 16 rows;
// returns the canonical outputs at tile index t + 256 k in out[k]. The LDS
// region is free again after the caller's next barrier.
template <class A>
__device__ __forceinline__ void fwd_rows_core(const A &a, u64 *lds, const u64 *data, u32 logN, u32 r0, const u32 t,
                                              u64 (&out)[16]) {
    using T = typename A::T;
    const u32 l = t & 15, u = t >> 4;
    T v[16];
    u32 baseA[1] = {(r0 + u) * 256 + l};
#pragma unroll
    for (int e = 0; e < 16; ++e) v[e] = A::from_bits(ld_stream(data + baseA[0] + ((u32)e << 4)));
    fwd_round<4, 4>(v, baseA, logN, a);
#pragma unroll
    for (int e = 0; e < 16; ++e) lds[row_addr(u, ((u32)e << 4) + l)] = A::to_bits(v[e]);
    __syncthreads();
    const u32 hp = l;
    u32 baseB[1] = {(r0 + u) * 256 + (hp << 4)};
#pragma unroll
    for (int e = 0; e < 16; ++e) v[e] = A::from_bits(lds[row_addr(u, (hp << 4) + e)]);
    fwd_round<4, 0>(v, baseB, logN, a);
    __syncthreads();
#pragma unroll
    for (int e = 0; e < 16; ++e) lds[row_addr(u, (hp << 4) + e)] = a.fwd_canon(v[e]);
    __syncthreads();
#pragma unroll
    for (int k = 0; k < 16; ++k) {
        const u32 idx = t + kNttThreads * k;
        out[k] = lds[row_addr(idx >> 8, idx & 255)];
    }
}

// Forward pass 2 body: the core, then the store policy per output.
template <class A, class Store>
__device__ __forceinline__ void fwd_rows_body(const A &a, u64 *lds, u64 *data, u32 logN, u32 r0, u32 limb, u32 z,
                                              const DevModulus &md, const Store &st, const u32 t) {
    u64 out[16];
    fwd_rows_core(a, lds, data, logN, r0, t, out);
#pragma unroll
    for (int k = 0; k < 16; ++k) st(data, r0 * 256 + t + kNttThreads * k, out[k], limb, z, md);
}

// Inverse pass 1 body: the first 8 stages (distances 1 .. 128) on 16 rows.
template <class A, class Load>
__device__ __forceinline__ void inv_rows_body(const A &a, u64 *lds, u64 *data, const u64 *src, u32 logN, u32 r0,
                                              u32 limb, u32 z, const DevModulus &md, const Load &ld, const u32 t) {
    using T = typename A::T;
#pragma unroll
    for (int k = 0; k < 16; ++k) {
        const u32 idx = t + kNttThreads * k;
        lds[row_addr(idx >> 8, idx & 255)] = ld(src, r0 * 256 + idx, limb, z, md);
    }
    __syncthreads();
    const u32 l = t & 15, u = t >> 4;
    T v[16];
    u32 baseB[1] = {(r0 + u) * 256 + (l << 4)};
#pragma unroll
    for (int e = 0; e < 16; ++e) v[e] = A::from_canon(lds[row_addr(u, (l << 4) + e)]);
    inv_round<4, 0>(v, baseB, logN, a);
    __syncthreads();
#pragma unroll
    for (int e = 0; e < 16; ++e) lds[row_addr(u, (l << 4) + e)] = A::to_bits(v[e]);
    __syncthreads();
    u32 baseA[1] = {(r0 + u) * 256 + l};
#pragma unroll
    for (int e = 0; e < 16; ++e) v[e] = A::from_bits(lds[row_addr(u, ((u32)e << 4) + l)]);
    inv_round<4, 4>(v, baseA, logN, a);
#pragma unroll
    for (int e = 0; e < 16; ++e) st_stream(data + baseA[0] + ((u32)e << 4), A::to_bits(v[e]));
}

// Inverse pass 2 body: the remaining SR stages (distances 256 .. N/2) on
// columns; the top stage folds n^-1 (and an optional per-limb scale).
template <int SR, class A, class Store>
__device__ __forceinline__ void inv_cols_body(const A &a, u64 *lds, u64 *data, u32 logN, u32 c0, u32 limb, u32 z,
                                              const DevModulus &md, const Store &st, u64 s0, u64 s0s, u64 s1,
                                              u64 s1s, const u32 t) {
    using T = typename A::T;
    constexpr int R = 1 << SR;
    constexpr int TC = kNttTile / R;
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
            v[2 * b] = A::from_bits(ld_stream(data + baseC[b]));
            v[2 * b + 1] = A::from_bits(ld_stream(data + baseC[b] + 256));
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
        const u32 kap = t + kNttThreads * b;
        uA[b] = kap % TC;
        lowA[b] = kap / TC;
        baseA[b] = lowA[b] * 256 + c0 + uA[b];
    }
    if constexpr (NB > 0) {
        constexpr int BB = 16 >> NB;
        u32 baseB[BB], uB[BB], hpB[BB];
#pragma unroll
        for (int b = 0; b < BB; ++b) {
            const u32 kap = t + kNttThreads * b;
            uB[b] = kap % TC;
            hpB[b] = kap / TC;
            baseB[b] = (hpB[b] << NB) * 256 + c0 + uB[b];
        }
#pragma unroll
        for (int b = 0; b < BB; ++b)
#pragma unroll
            for (int e = 0; e < (1 << NB); ++e)
                v[b * (1 << NB) + e] = A::from_bits(ld_stream(data + baseB[b] + ((u32)e << 8)));
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
                v[b * (1 << NA) + e] = A::from_bits(ld_stream(data + baseA[b] + ((u32)e << 8)));
    }
    inv_round_last<NA>(v, baseA, 8 + NB, logN, a, s0, s0s, s1, s1s);
#pragma unroll
    for (int b = 0; b < BA; ++b)
#pragma unroll
        for (int e = 0; e < (1 << NA); ++e)
            st(data, baseA[b] + ((u32)e << (8 + NB)), a.inv_canon(v[b * (1 << NA) + e]), limb, z, md);
    }
}

}  // namespace

// ---------------------------------------------------------------------------
// kernels: resolve the limb, then run the body in the limb's arithmetic
// ---------------------------------------------------------------------------
template <int SR, class Load>
__global__ __launch_bounds__(kNttThreads) void k_fwd_cols(NttTables T, LimbMap M, Load ld) {
    constexpr int R = 1 << SR;
    constexpr int TC = kNttTile / R;
    __shared__ u64 lds[R * (TC + 1)];
    const u32 z = blockIdx.y + M.poly0;
    const u32 limb = M.limb(blockIdx.z, z);
    if (limb >= M.nlimbs) return;
    const u32 row = M.row(limb);
    u64 *data = M.data + z * M.poly_stride + (u64)limb * T.N;
    const DevModulus md = T.mod[row];
    const u32 c0 = blockIdx.x * TC;
    if (md.pad)
        fwd_cols_body<SR>(fp_arith(T.twd + (u64)row * T.N, md), lds, data, T.logN, c0, limb, z, md, ld,
                          threadIdx.x);
    else
        fwd_cols_body<SR>(int_arith(T.tw + (u64)row * T.N, md.value), lds, data, T.logN, c0, limb, z, md,
                          ld, threadIdx.x);
}

template <class Store>
__global__ __launch_bounds__(kNttThreads) void k_fwd_rows(NttTables T, LimbMap M, Store st) {
    __shared__ u64 lds[16 * 272];
    const u32 z = blockIdx.y + M.poly0;
    const u32 limb = M.limb(blockIdx.z, z);
    if (limb >= M.nlimbs) return;
    const u32 row = M.row(limb);
    u64 *data = M.data + z * M.poly_stride + (u64)limb * T.N;
    const DevModulus md = T.mod[row];
    const u32 r0 = blockIdx.x * 16;
    if (md.pad)
        fwd_rows_body(fp_arith(T.twd + (u64)row * T.N, md), lds, data, T.logN, r0, limb, z, md, st,
                      threadIdx.x);
    else
        fwd_rows_body(int_arith(T.tw + (u64)row * T.N, md.value), lds, data, T.logN, r0, limb, z, md, st,
                      threadIdx.x);
}

template <class Load>
__global__ __launch_bounds__(kNttThreads) void k_inv_rows(NttTables T, LimbMap M, Load ld) {
    __shared__ u64 lds[16 * 272];
    const u32 z = blockIdx.y + M.poly0;
    const u32 limb = M.limb(blockIdx.z, z);
    if (limb >= M.nlimbs) return;
    const u32 row = M.row(limb);
    u64 *data = M.data + z * M.poly_stride + (u64)limb * T.N;
    const u64 *src = M.src ? M.src + z * M.src_poly_stride + (u64)limb * T.N : data;
    const DevModulus md = T.mod[row];
    const u32 r0 = blockIdx.x * 16;
    if (md.pad)
        inv_rows_body(fp_arith(T.itwd + (u64)row * T.N, md), lds, data, src, T.logN, r0, limb, z, md, ld,
                      threadIdx.x);
    else
        inv_rows_body(int_arith(T.itw + (u64)row * T.N, md.value), lds, data, src, T.logN, r0, limb, z, md,
                      ld, threadIdx.x);
}

template <int SR, class Store>
__global__ __launch_bounds__(kNttThreads) void k_inv_cols(NttTables T, LimbMap M, Store st) {
    constexpr int R = 1 << SR;
    constexpr int TC = kNttTile / R;
    __shared__ u64 lds[R * (TC + 1)];
    const u32 z = blockIdx.y + M.poly0;
    const u32 limb = M.limb(blockIdx.z, z);
    if (limb >= M.nlimbs) return;
    const u32 row = M.row(limb);
    u64 *data = M.data + z * M.poly_stride + (u64)limb * T.N;
    const DevModulus md = T.mod[row];
    const u32 c0 = blockIdx.x * TC;
    u64 s0, s0s, s1, s1s;
    limb_scale(T, M, limb, row, s0, s0s, s1, s1s);
    if (md.pad)
        inv_cols_body<SR>(fp_arith(T.itwd + (u64)row * T.N, md), lds, data, T.logN, c0, limb, z, md, st,
                          s0, s0s, s1, s1s, threadIdx.x);
    else
        inv_cols_body<SR>(int_arith(T.itw + (u64)row * T.N, md.value), lds, data, T.logN, c0, limb, z, md,
                          st, s0, s0s, s1, s1s, threadIdx.x);
}

// ---------------------------------------------------------------------------
// work-list kernels: entry blockIdx.y of the table gives the limb's buffer and
// row; policies that carry per-entry operands are bound to the entry once
// ---------------------------------------------------------------------------
namespace {
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
}  // namespace

template <int SR, class Load>
__global__ __launch_bounds__(kNttThreads) void k_fwd_cols_jobs(NttTables T, const NttJob *jobs, Load ld) {
    constexpr int R = 1 << SR;
    constexpr int TC = kNttTile / R;
    __shared__ u64 lds[R * (TC + 1)];
    const NttJob J = jobs[blockIdx.y];
    const DevModulus md = T.mod[J.row];
    const auto lb = bind_job(ld, J);
    const u32 c0 = blockIdx.x * TC;
    if (md.pad)
        fwd_cols_body<SR>(fp_arith(T.twd + (u64)J.row * T.N, md), lds, J.data, T.logN, c0, 0u, 0u, md, lb,
                          threadIdx.x);
    else
        fwd_cols_body<SR>(int_arith(T.tw + (u64)J.row * T.N, md.value), lds, J.data, T.logN, c0, 0u, 0u, md, lb,
                          threadIdx.x);
}

template <class Store>
__global__ __launch_bounds__(kNttThreads) void k_fwd_rows_jobs(NttTables T, const NttJob *jobs, Store st) {
    __shared__ u64 lds[16 * 272];
    const NttJob J = jobs[blockIdx.y];
    const DevModulus md = T.mod[J.row];
    const auto sb = bind_job(st, J);
    const u32 r0 = blockIdx.x * 16;
    if (md.pad)
        fwd_rows_body(fp_arith(T.twd + (u64)J.row * T.N, md), lds, J.data, T.logN, r0, 0u, 0u, md, sb, threadIdx.x);
    else
        fwd_rows_body(int_arith(T.tw + (u64)J.row * T.N, md.value), lds, J.data, T.logN, r0, 0u, 0u, md, sb,
                      threadIdx.x);
}

__global__ __launch_bounds__(kNttThreads) void k_inv_rows_jobs(NttTables T, const NttJob *jobs) {
    __shared__ u64 lds[16 * 272];
    const NttJob J = jobs[blockIdx.y];
    const DevModulus md = T.mod[J.row];
    const u64 *src = J.aux0 ? J.aux0 : J.data;
    const u32 r0 = blockIdx.x * 16;
    if (md.pad)
        inv_rows_body(fp_arith(T.itwd + (u64)J.row * T.N, md), lds, J.data, src, T.logN, r0, 0u, 0u, md,
                      LoadPlain{}, threadIdx.x);
    else
        inv_rows_body(int_arith(T.itw + (u64)J.row * T.N, md.value), lds, J.data, src, T.logN, r0, 0u, 0u, md,
                      LoadPlain{}, threadIdx.x);
}

template <int SR>
__global__ __launch_bounds__(kNttThreads) void k_inv_cols_jobs(NttTables T, const NttJob *jobs) {
    constexpr int R = 1 << SR;
    constexpr int TC = kNttTile / R;
    __shared__ u64 lds[R * (TC + 1)];
    const NttJob J = jobs[blockIdx.y];
    const DevModulus md = T.mod[J.row];
    const u64 *c = J.scale ? J.scale : T.inv_consts + 4 * (u64)J.row;
    const u64 s0 = c[0], s0s = c[1], s1 = c[2], s1s = c[3];
    const u32 c0 = blockIdx.x * TC;
    if (md.pad)
        inv_cols_body<SR>(fp_arith(T.itwd + (u64)J.row * T.N, md), lds, J.data, T.logN, c0, 0u, 0u, md, StorePlain{},
                          s0, s0s, s1, s1s, threadIdx.x);
    else
        inv_cols_body<SR>(int_arith(T.itw + (u64)J.row * T.N, md.value), lds, J.data, T.logN, c0, 0u, 0u, md,
                          StorePlain{}, s0, s0s, s1, s1s, threadIdx.x);
}

// ---------------------------------------------------------------------------
// fused kernels: one workgroup of G groups x 256 threads runs both passes of
// one limb. The pass-1 output is written in place and read back by the same
// workgroup after a barrier, so while a few hundred limbs are in flight it is
// served from L2 / the memory-attached last-level cache instead of HBM.
// Every group runs the same number of tiles, so the bodies' barriers match.
// ---------------------------------------------------------------------------
// LDS words per group: the row tile (16 x 272) or the column tile (R x (TC + 1))
template <int SR>
constexpr u32 ntt_slot() {
    constexpr u32 cols = (1u << SR) * ((u32)kNttTile / (1u << SR) + 1);
    return cols > 16 * 272 ? cols : 16 * 272;
}

template <int SR, int G, class Load, class Store>
__global__ __launch_bounds__(kNttThreads *G) void k_fwd_fused(NttTables T, LimbMap M, Load ld, Store st) {
    constexpr int TC = kNttTile >> SR;
    constexpr u32 kSlot = ntt_slot<SR>();
    __shared__ u64 lds_all[G * kSlot];
    const u32 z = blockIdx.y + M.poly0;
    const u32 limb = M.limb(blockIdx.z, z);
    if (limb >= M.nlimbs) return;
    const u32 row = M.row(limb);
    u64 *data = M.data + z * M.poly_stride + (u64)limb * T.N;
    const DevModulus md = T.mod[row];
    const u32 g = threadIdx.x / kNttThreads, t = threadIdx.x % kNttThreads;
    u64 *lds = lds_all + g * kSlot;
    const u32 tiles = T.N / kNttTile;
    auto run = [&](const auto &a) {
        for (u32 tile = g; tile < tiles; tile += G) {
            fwd_cols_body<SR>(a, lds, data, T.logN, tile * TC, limb, z, md, ld, t);
            __syncthreads();
        }
        for (u32 tile = g; tile < tiles; tile += G) {
            fwd_rows_body(a, lds, data, T.logN, tile * 16, limb, z, md, st, t);
            __syncthreads();
        }
    };
    if (md.pad)
        run(fp_arith(T.twd + (u64)row * T.N, md));
    else
        run(int_arith(T.tw + (u64)row * T.N, md.value));
}

template <int SR, int G, class Load, class Store>
__global__ __launch_bounds__(kNttThreads *G) void k_inv_fused(NttTables T, LimbMap M, Load ld, Store st) {
    constexpr int TC = kNttTile >> SR;
    constexpr u32 kSlot = ntt_slot<SR>();
    __shared__ u64 lds_all[G * kSlot];
    const u32 z = blockIdx.y + M.poly0;
    const u32 limb = M.limb(blockIdx.z, z);
    if (limb >= M.nlimbs) return;
    const u32 row = M.row(limb);
    u64 *data = M.data + z * M.poly_stride + (u64)limb * T.N;
    const u64 *src = M.src ? M.src + z * M.src_poly_stride + (u64)limb * T.N : data;
    const DevModulus md = T.mod[row];
    u64 s0, s0s, s1, s1s;
    limb_scale(T, M, limb, row, s0, s0s, s1, s1s);
    const u32 g = threadIdx.x / kNttThreads, t = threadIdx.x % kNttThreads;
    u64 *lds = lds_all + g * kSlot;
    const u32 tiles = T.N / kNttTile;
    auto run = [&](const auto &a) {
        for (u32 tile = g; tile < tiles; tile += G) {
            inv_rows_body(a, lds, data, src, T.logN, tile * 16, limb, z, md, ld, t);
            __syncthreads();
        }
        for (u32 tile = g; tile < tiles; tile += G) {
            inv_cols_body<SR>(a, lds, data, T.logN, tile * TC, limb, z, md, st, s0, s0s, s1, s1s, t);
            __syncthreads();
        }
    };
    if (md.pad)
        run(fp_arith(T.itwd + (u64)row * T.N, md));
    else
        run(int_arith(T.itw + (u64)row * T.N, md.value));
}

// ---------------------------------------------------------------------------
// launchers
// ---------------------------------------------------------------------------
namespace {
// polynomials per launch pair: slices of about ntt_chunk_limbs limbs (0 = one launch)
u32 chunk_polys(const NttTables &T, u32 nlimbs, u32 npolys) {
    const int c = T.chunk_limbs >= 0 ? T.chunk_limbs : tune_value("ntt_chunk_limbs");
    if (c <= 0) return npolys;
    return std::max<u32>(1, (u32)c / std::max<u32>(1, nlimbs));
}

// workgroup groups of the fused form for this launch, 0 = two passes
u32 fused_groups(const NttTables &T, u32 nlimbs, u32 npolys) {
    const int g = T.fused_g >= 0 ? T.fused_g : tune_value("ntt_fused");
    if (g != 1 && g != 2 && g != 4) return 0;
    const u32 tiles = T.N / kNttTile;
    if (tiles % (u32)g) return 0;
    const int min_limbs = T.fused_min >= 0 ? T.fused_min : tune_value("ntt_fused_min_limbs");
    if ((u64)nlimbs * npolys < (u64)std::max(0, min_limbs)) return 0;
    return (u32)g;
}

template <int SR, class Load, class Store>
void launch_fused_fwd(u32 G, dim3 grid, hipStream_t s, const NttTables &T, const LimbMap &M, const Load &ld,
                      const Store &st) {
    if (G == 4)
        hipLaunchKernelGGL((k_fwd_fused<SR, 4, Load, Store>), grid, dim3(kNttThreads * 4), 0, s, T, M, ld, st);
    else if (G == 2)
        hipLaunchKernelGGL((k_fwd_fused<SR, 2, Load, Store>), grid, dim3(kNttThreads * 2), 0, s, T, M, ld, st);
    else
        hipLaunchKernelGGL((k_fwd_fused<SR, 1, Load, Store>), grid, dim3(kNttThreads), 0, s, T, M, ld, st);
}

template <int SR, class Load, class Store>
void launch_fused_inv(u32 G, dim3 grid, hipStream_t s, const NttTables &T, const LimbMap &M, const Load &ld,
                      const Store &st) {
    if (G == 4)
        hipLaunchKernelGGL((k_inv_fused<SR, 4, Load, Store>), grid, dim3(kNttThreads * 4), 0, s, T, M, ld, st);
    else if (G == 2)
        hipLaunchKernelGGL((k_inv_fused<SR, 2, Load, Store>), grid, dim3(kNttThreads * 2), 0, s, T, M, ld, st);
    else
        hipLaunchKernelGGL((k_inv_fused<SR, 1, Load, Store>), grid, dim3(kNttThreads), 0, s, T, M, ld, st);
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
        default: fail(kInvalidArgument, "unsupported polynomial degree for the NTT (2^12 .. 2^17)");
    }
}

template <class Load, class Store>
void launch_forward(const NttTables &T, const LimbMap &M, u32 nlimbs, u32 npolys, const Load &ld, const Store &st,
                    hipStream_t s) {
    if (nlimbs == 0 || npolys == 0) return;
    const u32 tiles = T.N / kNttTile;
    if (const u32 G = fused_groups(T, nlimbs, npolys)) {
        dispatch_sr(T.logN, [&](auto sr) {
            launch_fused_fwd<decltype(sr)::value>(G, dim3(1, npolys, nlimbs), s, T, M, ld, st);
        });
        PHX_HIP_CHECK(hipGetLastError());
        return;
    }
    // polynomials sharing a prime run next to each other (and, with 8 or 16
    // tiles per limb, on the same XCD), so the prime's twiddles stay in L2;
    // large launches go in slices of whole polynomials so that the pass-1
    // output of a slice is still in the last-level cache when pass 2 reads it
    const u32 step = chunk_polys(T, nlimbs, npolys);
    for (u32 p0 = 0; p0 < npolys; p0 += step) {
        LimbMap Mc = M;
        Mc.poly0 = M.poly0 + p0;
        dim3 grid(tiles, std::min(step, npolys - p0), nlimbs);
        dispatch_sr(T.logN, [&](auto sr) {
            constexpr int SR = decltype(sr)::value;
            hipLaunchKernelGGL((k_fwd_cols<SR, Load>), grid, dim3(kNttThreads), 0, s, T, Mc, ld);
        });
        hipLaunchKernelGGL((k_fwd_rows<Store>), grid, dim3(kNttThreads), 0, s, T, Mc, st);
    }
    PHX_HIP_CHECK(hipGetLastError());
}
}  // namespace

void ntt_forward(const NttTables &T, const LimbMap &M, u32 nlimbs, u32 npolys, hipStream_t s) {
    launch_forward(T, M, nlimbs, npolys, LoadPlain{}, StorePlain{}, s);
}

namespace {
// grid rows per work-list launch (the limit of gridDim.y)
constexpr u32 kMaxJobsPerLaunch = 65535;

template <class Load, class Store>
void launch_forward_jobs(const NttTables &T, const NttJob *jobs, u32 njobs, const Load &ld, const Store &st,
                         hipStream_t s) {
    const u32 tiles = T.N / kNttTile;
    for (u32 j0 = 0; j0 < njobs; j0 += kMaxJobsPerLaunch) {
        const dim3 grid(tiles, std::min(kMaxJobsPerLaunch, njobs - j0));
        dispatch_sr(T.logN, [&](auto sr) {
            constexpr int SR = decltype(sr)::value;
            hipLaunchKernelGGL((k_fwd_cols_jobs<SR, Load>), grid, dim3(kNttThreads), 0, s, T, jobs + j0, ld);
        });
        hipLaunchKernelGGL((k_fwd_rows_jobs<Store>), grid, dim3(kNttThreads), 0, s, T, jobs + j0, st);
    }
    PHX_HIP_CHECK(hipGetLastError());
}
}  // namespace

void ntt_forward_jobs(const NttTables &T, const NttJob *jobs, u32 njobs, hipStream_t s) {
    launch_forward_jobs(T, jobs, njobs, LoadPlain{}, StorePlain{}, s);
}

void ntt_forward_jobs_moddown(const NttTables &T, const NttJob *jobs, u32 njobs, hipStream_t s) {
    launch_forward_jobs(T, jobs, njobs, LoadPlain{}, StoreModDownJob{}, s);
}

void ntt_forward_jobs_rescale(const NttTables &T, const NttJob *jobs, u32 njobs, hipStream_t s) {
    launch_forward_jobs(T, jobs, njobs, LoadRescaleJob{}, StoreRescaleJob{}, s);
}

void ntt_inverse_jobs(const NttTables &T, const NttJob *jobs, u32 njobs, hipStream_t s) {
    const u32 tiles = T.N / kNttTile;
    for (u32 j0 = 0; j0 < njobs; j0 += kMaxJobsPerLaunch) {
        const dim3 grid(tiles, std::min(kMaxJobsPerLaunch, njobs - j0));
        hipLaunchKernelGGL(k_inv_rows_jobs, grid, dim3(kNttThreads), 0, s, T, jobs + j0);
        dispatch_sr(T.logN, [&](auto sr) {
            constexpr int SR = decltype(sr)::value;
            hipLaunchKernelGGL((k_inv_cols_jobs<SR>), grid, dim3(kNttThreads), 0, s, T, jobs + j0);
        });
    }
    PHX_HIP_CHECK(hipGetLastError());
}

void ntt_inverse(const NttTables &T, const LimbMap &M, u32 nlimbs, u32 npolys, hipStream_t s) {
    if (nlimbs == 0 || npolys == 0) return;
    const u32 tiles = T.N / kNttTile;
    if (const u32 G = fused_groups(T, nlimbs, npolys)) {
        dispatch_sr(T.logN, [&](auto sr) {
            launch_fused_inv<decltype(sr)::value>(G, dim3(1, npolys, nlimbs), s, T, M, LoadPlain{}, StorePlain{});
        });
        PHX_HIP_CHECK(hipGetLastError());
        return;
    }
    // launch order and slicing as in launch_forward
    const u32 step = chunk_polys(T, nlimbs, npolys);
    for (u32 p0 = 0; p0 < npolys; p0 += step) {
        LimbMap Mc = M;
        Mc.poly0 = M.poly0 + p0;
        dim3 grid(tiles, std::min(step, npolys - p0), nlimbs);
        hipLaunchKernelGGL((k_inv_rows<LoadPlain>), grid, dim3(kNttThreads), 0, s, T, Mc, LoadPlain{});
        dispatch_sr(T.logN, [&](auto sr) {
            constexpr int SR = decltype(sr)::value;
            hipLaunchKernelGGL((k_inv_cols<SR, StorePlain>), grid, dim3(kNttThreads), 0, s, T, Mc, StorePlain{});
        });
    }
    PHX_HIP_CHECK(hipGetLastError());
}

void ntt_forward_store(const NttTables &T, const LimbMap &M, u32 nlimbs, u32 npolys, const StoreModDown &st,
                       hipStream_t s) {
    launch_forward(T, M, nlimbs, npolys, LoadPlain{}, st, s);
}

void ntt_forward_store_batch(const NttTables &T, const LimbMap &M, u32 nlimbs, u32 npolys,
                             const StoreModDownBatch &st, hipStream_t s) {
    launch_forward(T, M, nlimbs, npolys, LoadPlain{}, st, s);
}

void ntt_forward_rescale(const NttTables &T, const LimbMap &M, u32 nlimbs, u32 npolys, const LoadRescale &ld,
                         const StoreRescale &st, hipStream_t s) {
    launch_forward(T, M, nlimbs, npolys, ld, st, s);
}

}  // namespace phx
