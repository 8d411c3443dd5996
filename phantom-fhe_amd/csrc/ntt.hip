// ntt.hip -- two-pass NTT/INTT kernels and their launchers.
//
// Each kernel picks its butterfly arithmetic per limb (uniform across the
// workgroup): FP64 for primes below 2^50, 64-bit Shoup otherwise. Between the
// two passes the tile is stored in the arithmetic's own form (lazy u64 or
// double bits); the first load and the last store are canonical u64.
#include "ntt.h"

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
                                              const DevModulus &md, const Load &ld) {
    using T = typename A::T;
    constexpr int R = 1 << SR;
    constexpr int TC = kNttTile / R;
    constexpr int NA = SR < 4 ? SR : 4;
    constexpr int NB = SR - NA;
    constexpr int BA = 16 >> NA;
    const u32 t = threadIdx.x;
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

    fwd_round<NA>(v, baseA, 8 + NB, logN, a);

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
        fwd_round<NB>(v, baseB, 8, logN, a);
#pragma unroll
        for (int b = 0; b < BB; ++b)
#pragma unroll
            for (int e = 0; e < (1 << NB); ++e)
                data[baseB[b] + ((u32)e << 8)] = A::to_bits(v[b * (1 << NB) + e]);
    } else {
#pragma unroll
        for (int b = 0; b < BA; ++b)
#pragma unroll
            for (int e = 0; e < (1 << NA); ++e) data[baseA[b] + ((u32)e << 8)] = A::to_bits(v[b * (1 << NA) + e]);
    }
}

// Forward pass 2 body: the last 8 stages (distances 128 .. 1) on 16 rows.
template <class A, class Store>
__device__ __forceinline__ void fwd_rows_body(const A &a, u64 *lds, u64 *data, u32 logN, u32 r0, u32 limb, u32 z,
                                              const DevModulus &md, const Store &st) {
    using T = typename A::T;
    const u32 t = threadIdx.x;
    const u32 l = t & 15, u = t >> 4;
    T v[16];
    u32 baseA[1] = {(r0 + u) * 256 + l};
#pragma unroll
    for (int e = 0; e < 16; ++e) v[e] = A::from_bits(data[baseA[0] + ((u32)e << 4)]);
    fwd_round<4>(v, baseA, 4, logN, a);
#pragma unroll
    for (int e = 0; e < 16; ++e) lds[row_addr(u, ((u32)e << 4) + l)] = A::to_bits(v[e]);
    __syncthreads();
    const u32 hp = l;
    u32 baseB[1] = {(r0 + u) * 256 + (hp << 4)};
#pragma unroll
    for (int e = 0; e < 16; ++e) v[e] = A::from_bits(lds[row_addr(u, (hp << 4) + e)]);
    fwd_round<4>(v, baseB, 0, logN, a);
    __syncthreads();
#pragma unroll
    for (int e = 0; e < 16; ++e) lds[row_addr(u, (hp << 4) + e)] = a.fwd_canon(v[e]);
    __syncthreads();
#pragma unroll
    for (int k = 0; k < 16; ++k) {
        const u32 idx = t + kNttThreads * k;
        st(data, r0 * 256 + idx, lds[row_addr(idx >> 8, idx & 255)], limb, z, md);
    }
}

// Inverse pass 1 body: the first 8 stages (distances 1 .. 128) on 16 rows.
template <class A, class Load>
__device__ __forceinline__ void inv_rows_body(const A &a, u64 *lds, u64 *data, const u64 *src, u32 logN, u32 r0,
                                              u32 limb, u32 z, const DevModulus &md, const Load &ld) {
    using T = typename A::T;
    const u32 t = threadIdx.x;
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
    inv_round<4>(v, baseB, 0, logN, a);
    __syncthreads();
#pragma unroll
    for (int e = 0; e < 16; ++e) lds[row_addr(u, (l << 4) + e)] = A::to_bits(v[e]);
    __syncthreads();
    u32 baseA[1] = {(r0 + u) * 256 + l};
#pragma unroll
    for (int e = 0; e < 16; ++e) v[e] = A::from_bits(lds[row_addr(u, ((u32)e << 4) + l)]);
    inv_round<4>(v, baseA, 4, logN, a);
#pragma unroll
    for (int e = 0; e < 16; ++e) data[baseA[0] + ((u32)e << 4)] = A::to_bits(v[e]);
}

// Inverse pass 2 body: the remaining SR stages (distances 256 .. N/2) on
// columns; the top stage folds n^-1 (and an optional per-limb scale).
template <int SR, class A, class Store>
__device__ __forceinline__ void inv_cols_body(const A &a, u64 *lds, u64 *data, u32 logN, u32 c0, u32 limb, u32 z,
                                              const DevModulus &md, const Store &st, u64 s0, u64 s0s, u64 s1,
                                              u64 s1s) {
    using T = typename A::T;
    constexpr int R = 1 << SR;
    constexpr int TC = kNttTile / R;
    constexpr int NA = SR < 4 ? SR : 4;
    constexpr int NB = SR - NA;
    constexpr int BA = 16 >> NA;
    const u32 t = threadIdx.x;
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
                v[b * (1 << NB) + e] = A::from_bits(data[baseB[b] + ((u32)e << 8)]);
        inv_round<NB>(v, baseB, 8, logN, a);
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
                v[b * (1 << NA) + e] = A::from_bits(data[baseA[b] + ((u32)e << 8)]);
    }
    inv_round_last<NA>(v, baseA, 8 + NB, logN, a, s0, s0s, s1, s1s);
#pragma unroll
    for (int b = 0; b < BA; ++b)
#pragma unroll
        for (int e = 0; e < (1 << NA); ++e)
            st(data, baseA[b] + ((u32)e << (8 + NB)), a.inv_canon(v[b * (1 << NA) + e]), limb, z, md);
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
    const u32 limb = M.limb(blockIdx.z, blockIdx.y);
    if (limb >= M.nlimbs) return;
    const u32 row = M.row(limb);
    u64 *data = M.data + blockIdx.y * M.poly_stride + (u64)limb * T.N;
    const DevModulus md = T.mod[row];
    const u32 c0 = blockIdx.x * TC;
    if (md.pad)
        fwd_cols_body<SR>(fp_arith(T.twd + (u64)row * T.N, md), lds, data, T.logN, c0, limb, blockIdx.y, md, ld);
    else
        fwd_cols_body<SR>(int_arith(T.tw + (u64)row * T.N, md.value), lds, data, T.logN, c0, limb, blockIdx.y, md,
                          ld);
}

template <class Store>
__global__ __launch_bounds__(kNttThreads) void k_fwd_rows(NttTables T, LimbMap M, Store st) {
    __shared__ u64 lds[16 * 272];
    const u32 limb = M.limb(blockIdx.z, blockIdx.y);
    if (limb >= M.nlimbs) return;
    const u32 row = M.row(limb);
    u64 *data = M.data + blockIdx.y * M.poly_stride + (u64)limb * T.N;
    const DevModulus md = T.mod[row];
    const u32 r0 = blockIdx.x * 16;
    if (md.pad)
        fwd_rows_body(fp_arith(T.twd + (u64)row * T.N, md), lds, data, T.logN, r0, limb, blockIdx.y, md, st);
    else
        fwd_rows_body(int_arith(T.tw + (u64)row * T.N, md.value), lds, data, T.logN, r0, limb, blockIdx.y, md, st);
}

template <class Load>
__global__ __launch_bounds__(kNttThreads) void k_inv_rows(NttTables T, LimbMap M, Load ld) {
    __shared__ u64 lds[16 * 272];
    const u32 limb = M.limb(blockIdx.z, blockIdx.y);
    if (limb >= M.nlimbs) return;
    const u32 row = M.row(limb);
    u64 *data = M.data + blockIdx.y * M.poly_stride + (u64)limb * T.N;
    const u64 *src = M.src ? M.src + blockIdx.y * M.src_poly_stride + (u64)limb * T.N : data;
    const DevModulus md = T.mod[row];
    const u32 r0 = blockIdx.x * 16;
    if (md.pad)
        inv_rows_body(fp_arith(T.itwd + (u64)row * T.N, md), lds, data, src, T.logN, r0, limb, blockIdx.y, md, ld);
    else
        inv_rows_body(int_arith(T.itw + (u64)row * T.N, md.value), lds, data, src, T.logN, r0, limb, blockIdx.y, md,
                      ld);
}

template <int SR, class Store>
__global__ __launch_bounds__(kNttThreads) void k_inv_cols(NttTables T, LimbMap M, Store st) {
    constexpr int R = 1 << SR;
    constexpr int TC = kNttTile / R;
    __shared__ u64 lds[R * (TC + 1)];
    const u32 limb = M.limb(blockIdx.z, blockIdx.y);
    if (limb >= M.nlimbs) return;
    const u32 row = M.row(limb);
    u64 *data = M.data + blockIdx.y * M.poly_stride + (u64)limb * T.N;
    const DevModulus md = T.mod[row];
    const u32 c0 = blockIdx.x * TC;
    u64 s0, s0s, s1, s1s;
    limb_scale(T, M, limb, row, s0, s0s, s1, s1s);
    if (md.pad)
        inv_cols_body<SR>(fp_arith(T.itwd + (u64)row * T.N, md), lds, data, T.logN, c0, limb, blockIdx.y, md, st,
                          s0, s0s, s1, s1s);
    else
        inv_cols_body<SR>(int_arith(T.itw + (u64)row * T.N, md.value), lds, data, T.logN, c0, limb, blockIdx.y, md,
                          st, s0, s0s, s1, s1s);
}

// ---------------------------------------------------------------------------
// launchers
// ---------------------------------------------------------------------------
namespace {
template <class F>
void dispatch_sr(u32 logN, F &&f) {
    switch (logN) {
        case 12: f(std::integral_constant<int, 4>{}); break;
        case 13: f(std::integral_constant<int, 5>{}); break;
        case 14: f(std::integral_constant<int, 6>{}); break;
        case 15: f(std::integral_constant<int, 7>{}); break;
        case 16: f(std::integral_constant<int, 8>{}); break;
        default: fail(kInvalidArgument, "unsupported polynomial degree for the NTT (2^12 .. 2^16)");
    }
}

template <class Load, class Store>
void launch_forward(const NttTables &T, const LimbMap &M, u32 nlimbs, u32 npolys, const Load &ld, const Store &st,
                    hipStream_t s) {
    if (nlimbs == 0 || npolys == 0) return;
    const u32 tiles = T.N / kNttTile;
    // polynomials sharing a prime run next to each other (and, with 8 or 16
    // tiles per limb, on the same XCD), so the prime's twiddles stay in L2
    dim3 grid(tiles, npolys, nlimbs);
    dispatch_sr(T.logN, [&](auto sr) {
        constexpr int SR = decltype(sr)::value;
        hipLaunchKernelGGL((k_fwd_cols<SR, Load>), grid, dim3(kNttThreads), 0, s, T, M, ld);
    });
    hipLaunchKernelGGL((k_fwd_rows<Store>), grid, dim3(kNttThreads), 0, s, T, M, st);
    PHX_HIP_CHECK(hipGetLastError());
}
}  // namespace

void ntt_forward(const NttTables &T, const LimbMap &M, u32 nlimbs, u32 npolys, hipStream_t s) {
    launch_forward(T, M, nlimbs, npolys, LoadPlain{}, StorePlain{}, s);
}

void ntt_inverse(const NttTables &T, const LimbMap &M, u32 nlimbs, u32 npolys, hipStream_t s) {
    if (nlimbs == 0 || npolys == 0) return;
    const u32 tiles = T.N / kNttTile;
    // polynomials sharing a prime run next to each other (and, with 8 or 16
    // tiles per limb, on the same XCD), so the prime's twiddles stay in L2
    dim3 grid(tiles, npolys, nlimbs);
    hipLaunchKernelGGL((k_inv_rows<LoadPlain>), grid, dim3(kNttThreads), 0, s, T, M, LoadPlain{});
    dispatch_sr(T.logN, [&](auto sr) {
        constexpr int SR = decltype(sr)::value;
        hipLaunchKernelGGL((k_inv_cols<SR, StorePlain>), grid, dim3(kNttThreads), 0, s, T, M, StorePlain{});
    });
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
