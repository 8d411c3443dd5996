// ntt_fused.hip -- the one-workgroup-per-limb form of the NTT/INTT (the
// ntt_fused tuning choice): one workgroup of G x 256 threads runs both passes of
// a limb, the pass-1 output kept in L2 / the last-level cache.
#include "ntt_body.h"

namespace phx {

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
            fwd_cols_body<SR, kNttThreads>(a, lds, data, T.logN, tile * TC, limb, z, md, ld, t);
            __syncthreads();
        }
        for (u32 tile = g; tile < tiles; tile += G) {
            fwd_rows_body<kNttThreads>(a, lds, data, T.logN, tile * 16, limb, z, md, st, t);
            __syncthreads();
        }
    };
    if (md.pad)
        run(fp_arith(T.twd + (u64)row * T.N, md));
    else
        run(int_fwd(T, row, md.value));
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
            inv_rows_body<kNttThreads>(a, lds, data, src, T.logN, tile * 16, limb, z, md, ld, t);
            __syncthreads();
        }
        for (u32 tile = g; tile < tiles; tile += G) {
            inv_cols_body<SR, kNttThreads>(a, lds, data, T.logN, tile * TC, limb, z, md, st, s0, s0s, s1, s1s, t);
            __syncthreads();
        }
    };
    if (md.pad)
        run(fp_arith(T.itwd + (u64)row * T.N, md));
    else
        run(int_inv(T, row, md.value));
}

namespace {
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

}  // namespace

void fused_forward(u32 G, const NttTables &T, const LimbMap &M, u32 nlimbs, u32 npolys, const LoadPlain &ld,
                   const StorePlain &st, hipStream_t s) {
    dispatch_sr(T.logN, [&](auto sr) {
        launch_fused_fwd<decltype(sr)::value>(G, dim3(1, npolys, nlimbs), s, T, M, ld, st);
    });
    PHX_HIP_CHECK(hipGetLastError());
}

void fused_forward(u32 G, const NttTables &T, const LimbMap &M, u32 nlimbs, u32 npolys, const LoadPlain &ld,
                   const StoreModDown &st, hipStream_t s) {
    dispatch_sr(T.logN, [&](auto sr) {
        launch_fused_fwd<decltype(sr)::value>(G, dim3(1, npolys, nlimbs), s, T, M, ld, st);
    });
    PHX_HIP_CHECK(hipGetLastError());
}

void fused_forward(u32 G, const NttTables &T, const LimbMap &M, u32 nlimbs, u32 npolys, const LoadRescale &ld,
                   const StoreRescale &st, hipStream_t s) {
    dispatch_sr(T.logN, [&](auto sr) {
        launch_fused_fwd<decltype(sr)::value>(G, dim3(1, npolys, nlimbs), s, T, M, ld, st);
    });
    PHX_HIP_CHECK(hipGetLastError());
}

void fused_inverse(u32 G, const NttTables &T, const LimbMap &M, u32 nlimbs, u32 npolys, hipStream_t s) {
    dispatch_sr(T.logN, [&](auto sr) {
        launch_fused_inv<decltype(sr)::value>(G, dim3(1, npolys, nlimbs), s, T, M, LoadPlain{}, StorePlain{});
    });
    PHX_HIP_CHECK(hipGetLastError());
}

}  // namespace phx
