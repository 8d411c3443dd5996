// ntt.hip -- two-pass NTT/INTT kernels over a limb map, and their launchers.
//
// Each kernel picks its butterfly arithmetic per limb (uniform across the
// workgroup): FP64 for primes below 2^50, 64-bit Shoup otherwise. Between the
// two passes the tile is stored in the arithmetic's own form (lazy u64 or
// double bits); the first load and the last store are canonical u64. The
// column and row passes take their tile shape (threads per workgroup) from
// the context's tuning.
#include "ntt_body.h"

namespace phx {

// ---------------------------------------------------------------------------
// kernels: resolve the limb, then run the body in the limb's arithmetic
// ---------------------------------------------------------------------------
template <int SR, int NT, class Load>
__global__ __launch_bounds__(NT) void k_fwd_cols(NttTables T, LimbMap M, Load ld) {
    constexpr int R = 1 << SR;
    constexpr int TC = NT * kNttRegs / R;
    __shared__ u64 lds[R * (TC + 1)];
    const u32 z = blockIdx.y + M.poly0;
    const u32 limb = M.limb(blockIdx.z, z);
    if (limb >= M.nlimbs) return;
    const u32 row = M.row(limb);
    u64 *data = M.data + z * M.poly_stride + (u64)limb * T.N;
    const DevModulus md = T.mod[row];
    const u32 c0 = blockIdx.x * TC;
    if (md.pad)
        fwd_cols_body<SR, NT>(fp_arith(T.twd + (u64)row * T.N, md), lds, data, T.logN, c0, limb, z, md, ld,
                              threadIdx.x);
    else
        fwd_cols_body<SR, NT>(int_fwd(T, row, md.value), lds, data, T.logN, c0, limb, z, md, ld, threadIdx.x);
}

template <int NT, class Store>
__global__ __launch_bounds__(NT) void k_fwd_rows(NttTables T, LimbMap M, Store st) {
    __shared__ u64 lds[(NT / 16) * 272];
    const u32 z = blockIdx.y + M.poly0;
    const u32 limb = M.limb(blockIdx.z, z);
    if (limb >= M.nlimbs) return;
    const u32 row = M.row(limb);
    u64 *data = M.data + z * M.poly_stride + (u64)limb * T.N;
    const DevModulus md = T.mod[row];
    const u32 r0 = blockIdx.x * (NT / 16);
    if (md.pad)
        fwd_rows_body<NT>(fp_arith(T.twd + (u64)row * T.N, md), lds, data, T.logN, r0, limb, z, md, st,
                          threadIdx.x);
    else
        fwd_rows_body<NT>(int_fwd(T, row, md.value), lds, data, T.logN, r0, limb, z, md, st, threadIdx.x);
}

// row pass after the FastBConv-fused column pass (its remaining column stages first)
template <int SR, int NT, class Store>
__global__ __launch_bounds__(NT) void k_fwd_rows_rb(NttTables T, LimbMap M, Store st) {
    __shared__ u64 lds[(NT / 16) * 272];
    const u32 z = blockIdx.y + M.poly0;
    const u32 limb = M.limb(blockIdx.z, z);
    if (limb >= M.nlimbs) return;
    const u32 row = M.row(limb);
    u64 *data = M.data + z * M.poly_stride + (u64)limb * T.N;
    const DevModulus md = T.mod[row];
    const u32 r0 = blockIdx.x * (NT / 16);
    if (md.pad)
        fwd_rows_body_rb<SR, NT>(fp_arith(T.twd + (u64)row * T.N, md), lds, data, T.logN, r0, limb, z, md, st,
                                 threadIdx.x);
    else
        fwd_rows_body_rb<SR, NT>(int_fwd(T, row, md.value), lds, data, T.logN, r0, limb, z, md, st, threadIdx.x);
}

template <int NT, class Load>
__global__ __launch_bounds__(NT) void k_inv_rows(NttTables T, LimbMap M, Load ld) {
    __shared__ u64 lds[(NT / 16) * 272];
    const u32 z = blockIdx.y + M.poly0;
    const u32 limb = M.limb(blockIdx.z, z);
    if (limb >= M.nlimbs) return;
    const u32 row = M.row(limb);
    u64 *data = M.data + z * M.poly_stride + (u64)limb * T.N;
    const u64 *src = M.src ? M.src + z * M.src_poly_stride + (u64)limb * T.N : data;
    const DevModulus md = T.mod[row];
    const u32 r0 = blockIdx.x * (NT / 16);
    if (md.pad)
        inv_rows_body<NT>(fp_arith(T.itwd + (u64)row * T.N, md), lds, data, src, T.logN, r0, limb, z, md, ld,
                          threadIdx.x);
    else
        inv_rows_body<NT>(int_inv(T, row, md.value), lds, data, src, T.logN, r0, limb, z, md, ld, threadIdx.x);
}

template <int SR, int NT, class Store>
__global__ __launch_bounds__(NT) void k_inv_cols(NttTables T, LimbMap M, Store st) {
    constexpr int R = 1 << SR;
    constexpr int TC = NT * kNttRegs / R;
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
        inv_cols_body<SR, NT>(fp_arith(T.itwd + (u64)row * T.N, md), lds, data, T.logN, c0, limb, z, md, st,
                              s0, s0s, s1, s1s, threadIdx.x);
    else
        inv_cols_body<SR, NT>(int_inv(T, row, md.value), lds, data, T.logN, c0, limb, z, md,
                              st, s0, s0s, s1, s1s, threadIdx.x);
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
    if (tiles < (u32)g || tiles % (u32)g) return 0;   // N = 2^11 has no 4096-coefficient tile
    const int min_limbs = T.fused_min >= 0 ? T.fused_min : tune_value("ntt_fused_min_limbs");
    if ((u64)nlimbs * npolys < (u64)std::max(0, min_limbs)) return 0;
    return (u32)g;
}

template <class Load, class Store>
void launch_forward(const NttTables &T, const LimbMap &M, u32 nlimbs, u32 npolys, const Load &ld, const Store &st,
                    hipStream_t s) {
    if (nlimbs == 0 || npolys == 0) return;
    if (const u32 G = fused_groups(T, nlimbs, npolys)) {
        fused_forward(G, T, M, nlimbs, npolys, ld, st, s);
        return;
    }
    const u32 ntc = cols_threads(T, (u64)nlimbs * npolys), ntr = rows_threads(T);
    // polynomials sharing a prime run next to each other (and, with 8 or 16
    // tiles per limb, on the same XCD), so the prime's twiddles stay in L2;
    // large launches go in slices of whole polynomials so that the pass-1
    // output of a slice is still in the last-level cache when pass 2 reads it
    const u32 step = chunk_polys(T, nlimbs, npolys);
    for (u32 p0 = 0; p0 < npolys; p0 += step) {
        LimbMap Mc = M;
        Mc.poly0 = M.poly0 + p0;
        const u32 np = std::min(step, npolys - p0);
        dispatch_cols(T.logN, ntc, [&](auto sr, auto nt) {
            constexpr int SR = decltype(sr)::value;
            constexpr int NT = decltype(nt)::value;
            hipLaunchKernelGGL((k_fwd_cols<SR, NT, Load>), dim3(T.N / (NT * kNttRegs), np, nlimbs), dim3(NT), 0,
                               s, T, Mc, ld);
        });
        dispatch_rows_nt(ntr, [&](auto nt) {
            constexpr int NT = decltype(nt)::value;
            hipLaunchKernelGGL((k_fwd_rows<NT, Store>), dim3(T.N / (NT * kNttRegs), np, nlimbs), dim3(NT), 0, s, T,
                               Mc, st);
        });
    }
    PHX_HIP_CHECK(hipGetLastError());
}
}  // namespace

void ntt_forward(const NttTables &T, const LimbMap &M, u32 nlimbs, u32 npolys, hipStream_t s) {
    launch_forward(T, M, nlimbs, npolys, LoadPlain{}, StorePlain{}, s);
}

namespace {
// the row pass alone, over limbs whose column pass ran fused with FastBConv
// (bconv_ntt.hip: its first register round only; the rest runs here)
template <class Store>
void launch_forward_rows(const NttTables &T, const LimbMap &M, u32 nlimbs, u32 npolys, const Store &st,
                         hipStream_t s) {
    if (nlimbs == 0 || npolys == 0) return;
    const u32 nt = rows_threads(T) == 512 ? 512 : 256;
    dispatch_rows_rb(T.logN, nt, [&](auto sr, auto ntc) {
        constexpr int SR = decltype(sr)::value;
        constexpr int NT = decltype(ntc)::value;
        hipLaunchKernelGGL((k_fwd_rows_rb<SR, NT, Store>), dim3(T.N / (NT * kNttRegs), npolys, nlimbs), dim3(NT), 0,
                           s, T, M, st);
    });
    PHX_HIP_CHECK(hipGetLastError());
}
}  // namespace

void ntt_forward_rows(const NttTables &T, const LimbMap &M, u32 nlimbs, u32 npolys, hipStream_t s) {
    launch_forward_rows(T, M, nlimbs, npolys, StorePlain{}, s);
}

void ntt_forward_rows_store(const NttTables &T, const LimbMap &M, u32 nlimbs, u32 npolys, const StoreModDown &st,
                            hipStream_t s) {
    launch_forward_rows(T, M, nlimbs, npolys, st, s);
}

namespace {
template <class Load, class Store = StorePlain>
void launch_inverse(const NttTables &T, const LimbMap &M, u32 nlimbs, u32 npolys, const Load &ld, hipStream_t s) {
    const u32 ntc = cols_threads(T, (u64)nlimbs * npolys), ntr = rows_threads(T);
    // launch order and slicing as in launch_forward
    const u32 step = chunk_polys(T, nlimbs, npolys);
    for (u32 p0 = 0; p0 < npolys; p0 += step) {
        LimbMap Mc = M;
        Mc.poly0 = M.poly0 + p0;
        const u32 np = std::min(step, npolys - p0);
        dispatch_rows_nt(ntr, [&](auto nt) {
            constexpr int NT = decltype(nt)::value;
            hipLaunchKernelGGL((k_inv_rows<NT, Load>), dim3(T.N / (NT * kNttRegs), np, nlimbs), dim3(NT), 0, s, T,
                               Mc, ld);
        });
        dispatch_cols(T.logN, ntc, [&](auto sr, auto nt) {
            constexpr int SR = decltype(sr)::value;
            constexpr int NT = decltype(nt)::value;
            hipLaunchKernelGGL((k_inv_cols<SR, NT, Store>), dim3(T.N / (NT * kNttRegs), np, nlimbs),
                               dim3(NT), 0, s, T, Mc, Store{});
        });
    }
    PHX_HIP_CHECK(hipGetLastError());
}
}  // namespace

void ntt_inverse(const NttTables &T, const LimbMap &M, u32 nlimbs, u32 npolys, hipStream_t s, bool flip) {
    if (nlimbs == 0 || npolys == 0) return;
    if (flip) {
        // byte-flipped outputs for the fused FastBConv: always the two-launch form
        launch_inverse<LoadPlain, StoreFlipped>(T, M, nlimbs, npolys, LoadPlain{}, s);
        return;
    }
    if (const u32 G = fused_groups(T, nlimbs, npolys)) {
        fused_inverse(G, T, M, nlimbs, npolys, s);
        return;
    }
    launch_inverse(T, M, nlimbs, npolys, LoadPlain{}, s);
}

void ntt_inverse_tensor(const NttTables &T, const LimbMap &M, u32 nlimbs, const LoadTensor &ld, hipStream_t s,
                        bool flip) {
    if (nlimbs == 0) return;
    if (flip)
        launch_inverse<LoadTensor, StoreFlipped>(T, M, nlimbs, 1, ld, s);
    else
        launch_inverse(T, M, nlimbs, 1, ld, s);
}

void ntt_forward_store(const NttTables &T, const LimbMap &M, u32 nlimbs, u32 npolys, const StoreModDown &st,
                       hipStream_t s) {
    launch_forward(T, M, nlimbs, npolys, LoadPlain{}, st, s);
}

void ntt_forward_rescale(const NttTables &T, const LimbMap &M, u32 nlimbs, u32 npolys, const LoadRescale &ld,
                         const StoreRescale &st, hipStream_t s) {
    launch_forward(T, M, nlimbs, npolys, ld, st, s);
}

}  // namespace phx
