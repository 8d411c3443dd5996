// ntt_jobs.hip -- work-list forms of the two-pass NTT/INTT: grid row y runs
// entry y of a device table (its limb buffer, table row and policy operands),
// so one launch covers items of different levels and separate buffers.
#include "ntt_body.h"

namespace phx {

template <int SR, int NT, class Load>
__global__ __launch_bounds__(NT) void k_fwd_cols_jobs(NttTables T, const NttJob *jobs, Load ld) {
    constexpr int R = 1 << SR;
    constexpr int TC = NT * kNttRegs / R;
    __shared__ u64 lds[R * (TC + 1)];
    const NttJob J = jobs[blockIdx.y];
    const DevModulus md = T.mod[J.row];
    const auto lb = bind_job(ld, J);
    const u32 c0 = blockIdx.x * TC;
    if (md.pad)
        fwd_cols_body<SR, NT>(fp_arith(T.twd + (u64)J.row * T.N, md), lds, J.data, T.logN, c0, 0u, 0u, md, lb,
                              threadIdx.x);
    else
        fwd_cols_body<SR, NT>(int_fwd(T, J.row, md.value), lds, J.data, T.logN, c0, 0u, 0u, md, lb, threadIdx.x);
}

template <int NT, class Store>
__global__ __launch_bounds__(NT) void k_fwd_rows_jobs(NttTables T, const NttJob *jobs, Store st) {
    __shared__ u64 lds[(NT / 16) * 272];
    const NttJob J = jobs[blockIdx.y];
    const DevModulus md = T.mod[J.row];
    const auto sb = bind_job(st, J);
    const u32 r0 = blockIdx.x * (NT / 16);
    if (md.pad)
        fwd_rows_body<NT>(fp_arith(T.twd + (u64)J.row * T.N, md), lds, J.data, T.logN, r0, 0u, 0u, md, sb,
                          threadIdx.x);
    else
        fwd_rows_body<NT>(int_fwd(T, J.row, md.value), lds, J.data, T.logN, r0, 0u, 0u, md, sb, threadIdx.x);
}

// row pass after the FastBConv-fused column pass (its remaining column stages first)
template <int SR, int NT, class Store>
__global__ __launch_bounds__(NT) void k_fwd_rows_rb_jobs(NttTables T, const NttJob *jobs, Store st) {
    __shared__ u64 lds[(NT / 16) * 272];
    const NttJob J = jobs[blockIdx.y];
    const DevModulus md = T.mod[J.row];
    const auto sb = bind_job(st, J);
    const u32 r0 = blockIdx.x * (NT / 16);
    if (md.pad)
        fwd_rows_body_rb<SR, NT>(fp_arith(T.twd + (u64)J.row * T.N, md), lds, J.data, T.logN, r0, 0u, 0u, md, sb,
                                 threadIdx.x);
    else
        fwd_rows_body_rb<SR, NT>(int_fwd(T, J.row, md.value), lds, J.data, T.logN, r0, 0u, 0u, md, sb,
                                 threadIdx.x);
}

template <int NT, class Load>
__global__ __launch_bounds__(NT) void k_inv_rows_jobs(NttTables T, const NttJob *jobs, Load ld) {
    __shared__ u64 lds[(NT / 16) * 272];
    const NttJob J = jobs[blockIdx.y];
    const DevModulus md = T.mod[J.row];
    // a plain load reads aux0 (out of place) when it is set; a bound load ignores src
    const u64 *src = !BindsJob<Load>::value && J.aux0 ? J.aux0 : J.data;
    const auto lb = bind_job(ld, J);
    const u32 r0 = blockIdx.x * (NT / 16);
    if (md.pad)
        inv_rows_body<NT>(fp_arith(T.itwd + (u64)J.row * T.N, md), lds, J.data, src, T.logN, r0, 0u, 0u, md, lb,
                          threadIdx.x);
    else
        inv_rows_body<NT>(int_inv(T, J.row, md.value), lds, J.data, src, T.logN, r0, 0u, 0u, md, lb, threadIdx.x);
}

template <int SR, int NT, class Store>
__global__ __launch_bounds__(NT) void k_inv_cols_jobs(NttTables T, const NttJob *jobs) {
    constexpr int R = 1 << SR;
    constexpr int TC = NT * kNttRegs / R;
    __shared__ u64 lds[R * (TC + 1)];
    const NttJob J = jobs[blockIdx.y];
    const DevModulus md = T.mod[J.row];
    const auto c = gmem(J.scale ? J.scale : T.inv_consts + 4 * (u64)J.row);
    const u64 s0 = c[0], s0s = c[1], s1 = c[2], s1s = c[3];
    const u32 c0 = blockIdx.x * TC;
    if (md.pad)
        inv_cols_body<SR, NT>(fp_arith(T.itwd + (u64)J.row * T.N, md), lds, J.data, T.logN, c0, 0u, 0u, md,
                              Store{}, s0, s0s, s1, s1s, threadIdx.x);
    else
        inv_cols_body<SR, NT>(int_inv(T, J.row, md.value), lds, J.data, T.logN, c0, 0u, 0u, md,
                              Store{}, s0, s0s, s1, s1s, threadIdx.x);
}

namespace {
// grid rows per work-list launch (the limit of gridDim.y)
constexpr u32 kMaxJobsPerLaunch = 65535;

template <class Load, class Store>
void launch_forward_jobs(const NttTables &T, const NttJob *jobs, u32 njobs, const Load &ld, const Store &st,
                         hipStream_t s) {
    const u32 ntc = cols_threads(T, njobs), ntr = rows_threads(T);
    for (u32 j0 = 0; j0 < njobs; j0 += kMaxJobsPerLaunch) {
        const u32 nj = std::min(kMaxJobsPerLaunch, njobs - j0);
        dispatch_cols(T.logN, ntc, [&](auto sr, auto nt) {
            constexpr int SR = decltype(sr)::value;
            constexpr int NT = decltype(nt)::value;
            hipLaunchKernelGGL((k_fwd_cols_jobs<SR, NT, Load>), dim3(T.N / (NT * kNttRegs), nj), dim3(NT), 0, s,
                               T, jobs + j0, ld);
        });
        dispatch_rows_nt(ntr, [&](auto nt) {
            constexpr int NT = decltype(nt)::value;
            hipLaunchKernelGGL((k_fwd_rows_jobs<NT, Store>), dim3(T.N / (NT * kNttRegs), nj), dim3(NT), 0, s, T,
                               jobs + j0, st);
        });
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

namespace {
// the row pass alone, over entries whose column pass ran fused with FastBConv
// (bconv_ntt.hip: its first register round only; the rest runs here)
template <class Store>
void launch_forward_rows_jobs(const NttTables &T, const NttJob *jobs, u32 njobs, const Store &st, hipStream_t s) {
    const u32 nt = rows_threads(T) == 512 ? 512 : 256;
    for (u32 j0 = 0; j0 < njobs; j0 += kMaxJobsPerLaunch) {
        const u32 nj = std::min(kMaxJobsPerLaunch, njobs - j0);
        dispatch_rows_rb(T.logN, nt, [&](auto sr, auto ntc) {
            constexpr int SR = decltype(sr)::value;
            constexpr int NT = decltype(ntc)::value;
            hipLaunchKernelGGL((k_fwd_rows_rb_jobs<SR, NT, Store>), dim3(T.N / (NT * kNttRegs), nj), dim3(NT), 0, s,
                               T, jobs + j0, st);
        });
    }
    PHX_HIP_CHECK(hipGetLastError());
}
}  // namespace

void ntt_forward_rows_jobs(const NttTables &T, const NttJob *jobs, u32 njobs, hipStream_t s) {
    launch_forward_rows_jobs(T, jobs, njobs, StorePlain{}, s);
}

void ntt_forward_rows_jobs_moddown(const NttTables &T, const NttJob *jobs, u32 njobs, hipStream_t s) {
    launch_forward_rows_jobs(T, jobs, njobs, StoreModDownJob{}, s);
}

namespace {
template <class Load, class Store = StorePlain>
void launch_inverse_jobs(const NttTables &T, const NttJob *jobs, u32 njobs, const Load &ld, hipStream_t s) {
    const u32 ntc = cols_threads(T, njobs), ntr = rows_threads(T);
    for (u32 j0 = 0; j0 < njobs; j0 += kMaxJobsPerLaunch) {
        const u32 nj = std::min(kMaxJobsPerLaunch, njobs - j0);
        dispatch_rows_nt(ntr, [&](auto nt) {
            constexpr int NT = decltype(nt)::value;
            hipLaunchKernelGGL((k_inv_rows_jobs<NT, Load>), dim3(T.N / (NT * kNttRegs), nj), dim3(NT), 0, s, T,
                               jobs + j0, ld);
        });
        dispatch_cols(T.logN, ntc, [&](auto sr, auto nt) {
            constexpr int SR = decltype(sr)::value;
            constexpr int NT = decltype(nt)::value;
            hipLaunchKernelGGL((k_inv_cols_jobs<SR, NT, Store>), dim3(T.N / (NT * kNttRegs), nj), dim3(NT), 0, s,
                               T, jobs + j0);
        });
    }
    PHX_HIP_CHECK(hipGetLastError());
}
}  // namespace

void ntt_inverse_jobs(const NttTables &T, const NttJob *jobs, u32 njobs, hipStream_t s, bool flip) {
    if (flip)
        launch_inverse_jobs<LoadPlain, StoreFlipped>(T, jobs, njobs, LoadPlain{}, s);
    else
        launch_inverse_jobs(T, jobs, njobs, LoadPlain{}, s);
}

void ntt_inverse_jobs_tensor(const NttTables &T, const NttJob *jobs, u32 njobs, hipStream_t s, bool flip) {
    if (flip)
        launch_inverse_jobs<LoadTensorJob, StoreFlipped>(T, jobs, njobs, LoadTensorJob{T.N}, s);
    else
        launch_inverse_jobs(T, jobs, njobs, LoadTensorJob{T.N}, s);
}

}  // namespace phx
