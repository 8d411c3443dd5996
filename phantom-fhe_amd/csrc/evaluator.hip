// evaluator.hip -- hot-path kernels (base conversion, key inner product,
// tensor, Galois permutation) and the orchestration of key switching,
// relinearisation, rotation and rescale; key generation and symmetric
// encryption for tests.
#include <algorithm>
#include <cstdlib>
#include <cstring>

#include "evaluator.h"
#include "ntt_arith.h"
#include "prng.h"
#include "tune.h"
#include "bfv.h"

namespace phx {

namespace {
constexpr int kThreads = 256;
constexpr int kMaxDigitLimbs = 16;   // FastBConv input limbs held in registers
constexpr int kMaxBconvInputs = 63;  // FastBConv input limbs a 128-bit sum takes

inline dim3 grid_for(u64 N, u32 y, u32 z = 1) { return dim3((u32)((N + kThreads - 1) / kThreads), y, z); }
}  // namespace

// ---------------------------------------------------------------------------
// FastBConv phase 2 (+ optional copy of the digit's own limbs), one thread per
// coefficient, one grid row per job (digit or polynomial):
//   out[pos[o]][k] = (sum_i in[in_first+i][k] * mat[o][i]) mod p_o
// 128-bit accumulation, one Barrett reduction per output.
// reference: bconv_matmul_padded_unroll2_kernel src/rns_bconv.cu:455-485,
//            bconv_matmul_unroll2_kernel 143-168, modup_copy_partQl 522-528
// ---------------------------------------------------------------------------
// the descriptor and limb sets of grid row blockIdx.y
__device__ __forceinline__ void bconv_row(const BconvArgs &A, BconvDesc &d, const u64 *&in, u64 *&out, u32 &j,
                                          u32 &bt) {
    if (A.rows) {
        const BconvRow r = A.rows[blockIdx.y];
        d = *r.desc;
        in = r.in;
        out = r.out;
        j = bt = 0;
        return;
    }
    j = blockIdx.y % A.jobs_per_batch;
    bt = blockIdx.y / A.jobs_per_batch;
    d = A.descs[j * A.desc_stride];
    in = A.in + bt * A.in_batch_stride + j * A.in_job_stride;
    out = A.out + bt * A.out_batch_stride + j * A.out_job_stride;
}

__global__ __launch_bounds__(kThreads) void k_bconv(BconvArgs A) {
    BconvDesc d;
    const u64 *inb;
    u64 *outb;
    u32 j, bt;
    bconv_row(A, d, inb, outb, j, bt);
    const u32 N = A.N;
    const u32 k = blockIdx.x * kThreads + threadIdx.x;
    if (k >= N) return;
    const auto inj = gmem(inb + (u64)d.in_first * N + k);
    const auto outj = gmem(outb + k);
    const u64 *mat = A.cbuf + d.mat_off;
    const u32 *pos = reinterpret_cast<const u32 *>(A.cbuf) + d.pos_off;
    const u32 *rows = reinterpret_cast<const u32 *>(A.cbuf) + d.row_off;
    const DevModulus *mod = A.mod;
    if (d.ni <= (u32)kMaxDigitLimbs) {
        u64 x[kMaxDigitLimbs];
#pragma unroll
        for (int i = 0; i < kMaxDigitLimbs; ++i) x[i] = (u32)i < d.ni ? inj[(u64)i * N] : 0;
        for (u32 o = 0; o < d.no; ++o) {
            const u64 *mo = mat + (u64)o * d.ni;
            u128 acc = 0;
#pragma unroll
            for (int i = 0; i < kMaxDigitLimbs; ++i)
                if ((u32)i < d.ni) acc += (u128)x[i] * mo[i];
            const DevModulus m = mod[rows[o]];
            outj[(u64)pos[o] * N] = barrett128(acc, m.value, m.ratio_lo, m.ratio_hi);
        }
    } else {
        // more input limbs than registers (special_modulus_size above 16): every
        // output reads its inputs again, from the cache; below 64 inputs of 61 bits
        // the 128-bit sum cannot overflow
        for (u32 o = 0; o < d.no; ++o) {
            const u64 *mo = mat + (u64)o * d.ni;
            u128 acc = 0;
            for (u32 i = 0; i < d.ni; ++i) acc += (u128)inj[(u64)i * N] * mo[i];
            const DevModulus m = mod[rows[o]];
            outj[(u64)pos[o] * N] = barrett128(acc, m.value, m.ratio_lo, m.ratio_hi);
        }
    }
    if (A.copy_src) {
        const auto cs = gmem(A.copy_src + bt * A.copy_batch_stride + j * A.copy_job_stride + (u64)d.in_first * N + k);
        for (u32 i = 0; i < d.ni; ++i) outj[(u64)(d.own_first + i) * N] = cs[(u64)i * N];
    }
}

// ---------------------------------------------------------------------------
// FastBConv phase 2 on the matrix cores, exact (same outputs as k_bconv).
//
// With x_i = sum_a x_ia 2^(8a) (its bytes, a < 8) and
// M'_oia = 2^(8a) M[o][i] mod p_o = sum_b d_oiab 256^b (balanced base-256
// digits d in [-128, 127], b < 8):
//   T_o = sum_(i,a) x_ia M'_oia = sum_i x_i M[o][i]  (mod p_o),  T_o < 2^15 p_o.
// The matrix cores multiply signed bytes, so the inputs enter as
// s_ia = x_ia - 128 (bit 7 flipped) and Z[o,b] = sum d_oiab s_ia is an int8
// GEMM with int32 sums (|Z| <= 128*128*128 = 2^21, exact); then
//   T_o = sum_b 256^b Z[o,b] + C_o,  C_o = 128 sum_(i,a) M'_oia,
// with C_o precomputed per output. K = (i,a) = 8i + a <= 128 -> two
// v_mfma_i32_16x16x64_i8 per 16x16 tile; rows are permuted so the 8 sums of an
// output sit in one lane's accumulators. T_o is reduced with an FP64 quotient
// estimate (off by at most one) and one correction.
// ---------------------------------------------------------------------------
typedef int v4i __attribute__((ext_vector_type(4)));
constexpr u32 kBconvTileCols = 16;   // coefficients per MFMA column tile
constexpr u32 kBconvThreads = 512;   // 8 waves share one staged copy of the operands

__global__ __launch_bounds__(kBconvThreads) void k_bconv_mfma(BconvArgs A) {
    extern __shared__ __align__(16) uint8_t smem[];
    BconvDesc d;
    const u64 *inb;
    u64 *outb;
    u32 j, bt;
    bconv_row(A, d, inb, outb, j, bt);
    const u64 *cbuf = A.cbuf;
    const DevModulus *mod = A.mod;
    const u32 N = A.N;
    const u32 rt = d.rt, ni = d.ni, no = d.no;
    const u32 tid = threadIdx.x;
    // A operand tiles staged once per workgroup: [rt][2][64 lanes] x 16 B
    uint4 *As = reinterpret_cast<uint4 *>(smem);
    const uint4 *Ag = reinterpret_cast<const uint4 *>(cbuf + d.a_off);
    for (u32 i = tid; i < rt * 128; i += kBconvThreads) As[i] = Ag[i];
    // per-output constants next to the tiles: modulus, fl(1/p), bias C_o, output offset
    struct OutConst {
        u64 p;
        double pinv;    // fl(1/p)
        u64 bias_lo;    // C_o mod 2^64
        double bias;    // fl(C_o)
        u64 off;        // output limb offset pos[o] * N
        u64 pad;
    };
    OutConst *oc = reinterpret_cast<OutConst *>(smem + (size_t)rt * 2048);
    {
        const u32 *pos = reinterpret_cast<const u32 *>(cbuf) + d.pos_off;
        const u32 *rows = reinterpret_cast<const u32 *>(cbuf) + d.row_off;
        const double *pinv = reinterpret_cast<const double *>(cbuf + d.pinv_off);
        const u64 *bias = cbuf + d.a_off + (u64)rt * 256;   // after rt x 2 KiB of operand tiles
        for (u32 o = tid; o < no; o += kBconvThreads)
            oc[o] = OutConst{mod[rows[o]].value, pinv[o], bias[2 * o], __longlong_as_double((long long)bias[2 * o + 1]),
                             (u64)pos[o] * N, 0};
    }
    __syncthreads();

    const auto inj = gmem(inb + (u64)d.in_first * N);
    const auto outj = gmem(outb);
    const u32 wave = tid >> 6, lane = tid & 63, ln = lane & 15, lg = lane >> 4;
    const u32 ntiles = N / kBconvTileCols;
    const u32 groups = rt / 2;

    // each wave walks its own column tiles (no workgroup barriers below); the
    // inputs of the next tile are loaded before the current tile is computed
    const u32 nwaves = kBconvThreads / 64;
    u32 ct = blockIdx.x * nwaves + wave;
    const u32 ct_step = gridDim.x * nwaves;
    // lane (lg, ln) feeds K = 64h + 16 lg + j: the 8 bytes of inputs 8h + 2 lg and 8h + 2 lg + 1
    u64 xn[2][2];
    auto load_x = [&](u32 tile) {
        const u32 kk = tile * kBconvTileCols + ln;
#pragma unroll
        for (int h = 0; h < 2; ++h)
#pragma unroll
            for (int r = 0; r < 2; ++r) {
                const u32 i = 8 * h + 2 * lg + r;
                xn[h][r] = i < ni ? inj[(u64)i * N + kk] : 0;
            }
    };
    if (ct < ntiles) load_x(ct);
    for (; ct < ntiles; ct += ct_step) {
        const u32 k = ct * kBconvTileCols + ln;   // this lane's coefficient
        // B operand: the input bytes as signed int8, s = byte - 128 (flip bit 7)
        v4i b[2];
#pragma unroll
        for (int h = 0; h < 2; ++h)
            b[h] = v4i{(int)((u32)xn[h][0] ^ 0x80808080u), (int)((u32)(xn[h][0] >> 32) ^ 0x80808080u),
                       (int)((u32)xn[h][1] ^ 0x80808080u), (int)((u32)(xn[h][1] >> 32) ^ 0x80808080u)};
        if (ct + ct_step < ntiles) load_x(ct + ct_step);
        // 2 tiles per group of 4 outputs; lane (lg, ln) owns output 4T + lg
        auto products = [&](u32 T, v4i(&acc)[2]) {
#pragma unroll
            for (int sgrp = 0; sgrp < 2; ++sgrp) {
                const u32 t = 2 * T + sgrp;
                const uint4 ua0 = As[(t * 2 + 0) * 64 + lane];
                const uint4 ua1 = As[(t * 2 + 1) * 64 + lane];
                const v4i a0 = {(int)ua0.x, (int)ua0.y, (int)ua0.z, (int)ua0.w};
                const v4i a1 = {(int)ua1.x, (int)ua1.y, (int)ua1.z, (int)ua1.w};
                v4i c = {0, 0, 0, 0};
                c = __builtin_amdgcn_mfma_i32_16x16x64_i8(a0, b[0], c, 0, 0, 0);
                acc[sgrp] = __builtin_amdgcn_mfma_i32_16x16x64_i8(a1, b[1], c, 0, 0, 0);
            }
        };
        // Z_b = acc[b / 4][b % 4] (signed, |Z_b| <= 2^21), b < 8, paired as
        // y_k = Z_2k + 256 Z_2k+1 (|y_k| < 2^30): T = sum_k 2^(16k) y_k + C_o is in
        // [0, 2^15 p). Its low 64 bits come exactly from 64-bit integer sums; in
        // FP64 every partial sum stays below 2^17 p (a balanced-digit expansion
        // has sum_b 256^b |d_b| < 3.02 M': the top digit is at least 1 and the
        // lower ones sum to at most 0.502 256^top), so the four roundings leave T
        // within 2^-34 p, and r = T - floor(T/p - 2^-30) p lies in [0, 2p)
        auto reduce = [&](u32 T, const v4i(&acc)[2]) {
            const u32 o = 4 * T + lg;
            if (o >= no) return;
            const OutConst c = oc[o];
            const int y0 = acc[0][0] + acc[0][1] * 256, y1 = acc[0][2] + acc[0][3] * 256;
            const int y2 = acc[1][0] + acc[1][1] * 256, y3 = acc[1][2] + acc[1][3] * 256;
            const u64 lo = c.bias_lo + (u64)((long long)y0 + (long long)y1 * (1ll << 16)) +
                           (u64)(long long)y2 * (1ull << 32) + (u64)(long long)y3 * (1ull << 48);
            double v = __builtin_fma((double)y3, 0x1p48, c.bias);
            v = __builtin_fma((double)y2, 0x1p32, v);
            v = __builtin_fma((double)y1, 0x1p16, v);
            v += (double)y0;
            const u32 kq = (u32)__builtin_fmax(__builtin_floor(__builtin_fma(v, c.pinv, -0x1p-30)), 0.0);
            u64 r = lo - (u64)kq * c.p;
            r = r >= c.p ? r - c.p : r;
            outj[c.off + k] = r;
        };
        // the next group's products are issued before this group's reduction
        v4i acc0[2], acc1[2];
        products(0, acc0);
        for (u32 T = 0; T < groups; T += 2) {
            const bool odd = T + 1 < groups;
            if (odd) products(T + 1, acc1);
            reduce(T, acc0);
            if (!odd) break;
            if (T + 2 < groups) products(T + 2, acc0);
            reduce(T + 1, acc1);
        }
    }
    if (A.copy_src) {
        const auto cs = gmem(A.copy_src + bt * A.copy_batch_stride + j * A.copy_job_stride + (u64)d.in_first * N);
        const u64 total = (u64)ni * N;
        for (u64 idx = (u64)blockIdx.x * kBconvThreads + tid; idx < total; idx += (u64)gridDim.x * kBconvThreads) {
            const u64 i = idx / N, kk = idx % N;
            outj[(d.own_first + i) * N + kk] = cs[i * N + kk];
        }
    }
}

namespace {
// FastBConv dispatch: matrix cores when every job has an MFMA form, VALU otherwise
// (the bconv_mfma switch set to 0 forces the VALU kernel).
bool bconv_force_valu(const Context &ctx) {
    const int c = ctx.tuning().bconv_mfma;
    return (c >= 0 ? c : tune_value("bconv_mfma")) == 0;
}

}  // namespace

void bconv_run(const Context &ctx, const std::vector<BconvDesc> &descs_host, const BconvArgs &args, u32 jobs,
               hipStream_t s) {
    const u64 N = ctx.N();
    u32 rt_max = 0, no_max = 0;
    bool mfma = !bconv_force_valu(ctx) && N % kBconvTileCols == 0;
    for (const auto &d : descs_host) {
        if (d.rt == 0) mfma = false;
        rt_max = std::max(rt_max, d.rt);
        no_max = std::max(no_max, d.no);
    }
    const size_t lds = (size_t)rt_max * 2048 + (size_t)no_max * 48;
    static const size_t lds_limit = [] {
        int dev = 0, v = 0;
        (void)hipGetDevice(&dev);
        (void)hipDeviceGetAttribute(&v, hipDeviceAttributeMaxSharedMemoryPerBlock, dev);
        return (size_t)v;
    }();
    if (lds > lds_limit) mfma = false;
    if (mfma) {
        // one round of co-resident workgroups (no tail round); each stages the
        // operands once and walks its job's column tiles
        static const u32 cus = [] {
            int dev = 0, v = 0;
            (void)hipGetDevice(&dev);
            (void)hipDeviceGetAttribute(&v, hipDeviceAttributeMultiprocessorCount, dev);
            return (u32)std::max(1, v);
        }();
        static const size_t lds_cu = [] {
            int dev = 0, v = 0;
            (void)hipGetDevice(&dev);
            (void)hipDeviceGetAttribute(&v, hipDeviceAttributeMaxSharedMemoryPerMultiprocessor, dev);
            return (size_t)std::max(v, 65536);
        }();
        u32 per_cu = (u32)std::max<size_t>(1, std::min<size_t>(lds_cu / lds, 2048 / kBconvThreads));
        const int cap = ctx.tuning().bconv_wgs_per_cu >= 0 ? ctx.tuning().bconv_wgs_per_cu
                                                           : tune_value("bconv_wgs_per_cu");
        if (cap > 0) per_cu = std::min<u32>(per_cu, (u32)cap);
        const u32 tiles = (u32)(N / kBconvTileCols);
        const u32 waves = kBconvThreads / 64;
        const u32 gx = std::max<u32>(1, std::min<u32>((tiles + waves - 1) / waves, cus * per_cu / jobs));
        hipLaunchKernelGGL(k_bconv_mfma, dim3(gx, jobs), dim3(kBconvThreads), lds, s, args);
    } else {
        hipLaunchKernelGGL(k_bconv, grid_for(N, jobs), dim3(kThreads), 0, s, args);
    }
    PHX_HIP_CHECK(hipGetLastError());
}

namespace {
BconvArgs bconv_args(const Context &ctx, const BconvDesc *descs_dev, u32 desc_stride, u32 jobs, const u64 *in,
                     u64 in_job_stride, u64 *out, u64 out_job_stride, const u64 *copy_src, u64 copy_job_stride) {
    BconvArgs a{};
    a.cbuf = ctx.const_buf();
    a.descs = descs_dev;
    a.desc_stride = desc_stride;
    a.jobs_per_batch = jobs;
    a.in = in;
    a.in_job_stride = in_job_stride;
    a.out = out;
    a.out_job_stride = out_job_stride;
    a.copy_src = copy_src;
    a.copy_job_stride = copy_job_stride;
    a.mod = ctx.dev_mod();
    a.N = (u32)ctx.N();
    return a;
}
void bconv_launch(const Context &ctx, const BconvDesc *descs_dev, const std::vector<BconvDesc> &descs_host,
                  u32 desc_stride, u32 jobs, const u64 *in, u64 in_job_stride, u64 *out, u64 out_job_stride,
                  const u64 *copy_src, u64 copy_job_stride, hipStream_t s) {
    bconv_run(ctx, descs_host,
              bconv_args(ctx, descs_dev, desc_stride, jobs, in, in_job_stride, out, out_job_stride, copy_src,
                         copy_job_stride),
              jobs, s);
}
}  // namespace

// key_switch_inner_prod_c2_and_evk (src/eval_key_switch.cu:14-69):
// cx0[pos] = sum_j t[j][pos] * key_j.b[row], cx1[pos] = sum_j t[j][pos] * key_j.a[row]
// With c2 given, digit j's own limbs (pos / alpha == j) are read from c2 (NTT
// form) instead of from t, which then need not hold them.
__global__ __launch_bounds__(kThreads) void k_inner_product(const u64 *__restrict__ t, const u64 *__restrict__ c2,
                                                            const u64 *__restrict__ key, u64 *__restrict__ cx,
                                                            const DevModulus *__restrict__ mod, u32 N, u32 Ql,
                                                            u32 QlP, u32 size_Q, u32 QP, u32 beta, u32 alpha) {
    const u32 pos = blockIdx.y;
    const u32 k = blockIdx.x * kThreads + threadIdx.x;
    if (k >= N) return;
    const u32 row = pos < Ql ? pos : size_Q + (pos - Ql);
    const u32 own = c2 && pos < Ql ? pos / alpha : ~0u;
    const u64 QPN = (u64)QP * N;
    const DevModulus m = mod[row];
    if (m.pad) {
        // q < 2^50: exact products on the FP64 pipe (FpArith::mulmod, |r| < q), the
        // partial sums reduced every 4 digits (|sum| < 4.5 q < 2^53), then canonical
#pragma clang fp contract(off)
        const FpArith fa{nullptr, (double)m.value, __longlong_as_double((long long)m.pad)};
        double s0 = 0.0, s1 = 0.0;
        for (u32 j = 0; j < beta; ++j) {
            if (j && (j & 3) == 0) {
                s0 = fa.red(s0);
                s1 = fa.red(s1);
            }
            const double x = FpArith::from_canon(j == own ? c2[(u64)pos * N + k] : t[((u64)j * QlP + pos) * N + k]);
            const u64 *kb = key + (u64)j * 2 * QPN + (u64)row * N + k;
            s0 += fa.mulmod(x, FpArith::from_canon(kb[0]));
            s1 += fa.mulmod(x, FpArith::from_canon(kb[QPN]));
        }
        cx[(u64)pos * N + k] = fa.canon(s0);
        cx[((u64)QlP + pos) * N + k] = fa.canon(s1);
        return;
    }
    u128 a0 = 0, a1 = 0;
    for (u32 j = 0; j < beta; ++j) {
        const u64 x = j == own ? c2[(u64)pos * N + k] : t[((u64)j * QlP + pos) * N + k];
        const u64 *kb = key + (u64)j * 2 * QPN + (u64)row * N + k;
        a0 += (u128)x * kb[0];
        a1 += (u128)x * kb[QPN];
    }
    cx[(u64)pos * N + k] = barrett128(a0, m.value, m.ratio_lo, m.ratio_hi);
    cx[((u64)QlP + pos) * N + k] = barrett128(a1, m.value, m.ratio_lo, m.ratio_hi);
}

// hoisting: the inner product of the permuted digits, accumulated:
// cx[h][pos][k] (+)= sum_j t[j][pos][perm[k]] * key_j.(b|a)[row(pos)][k]
__global__ __launch_bounds__(kThreads) void k_inner_product_perm_acc(const u64 *__restrict__ t,
                                                                     const u32 *__restrict__ perm,
                                                                     const u64 *__restrict__ key,
                                                                     u64 *__restrict__ cx,
                                                                     const DevModulus *__restrict__ mod, u32 N,
                                                                     u32 Ql, u32 QlP, u32 size_Q, u32 QP, u32 beta,
                                                                     int accumulate) {
    const u32 pos = blockIdx.y;
    const u32 k = blockIdx.x * kThreads + threadIdx.x;
    if (k >= N) return;
    const u32 row = pos < Ql ? pos : size_Q + (pos - Ql);
    const u64 QPN = (u64)QP * N;
    const u32 pk = perm[k];
    u128 a0 = 0, a1 = 0;
    for (u32 j = 0; j < beta; ++j) {
        const u64 x = t[((u64)j * QlP + pos) * N + pk];
        const u64 *kb = key + (u64)j * 2 * QPN + (u64)row * N + k;
        a0 += (u128)x * kb[0];
        a1 += (u128)x * kb[QPN];
    }
    const DevModulus m = mod[row];
    u64 r0 = barrett128(a0, m.value, m.ratio_lo, m.ratio_hi);
    u64 r1 = barrett128(a1, m.value, m.ratio_lo, m.ratio_hi);
    u64 *c0 = cx + (u64)pos * N + k, *c1 = cx + ((u64)QlP + pos) * N + k;
    if (accumulate) {
        r0 = add_mod(*c0, r0, m.value);
        r1 = add_mod(*c1, r1, m.value);
    }
    *c0 = r0;
    *c1 = r1;
}

// hoisting: acc (+)= pi(src) on L limbs
__global__ __launch_bounds__(kThreads) void k_permute_acc(const u64 *__restrict__ src, u64 *__restrict__ acc,
                                                          const u32 *__restrict__ perm,
                                                          const DevModulus *__restrict__ mod, u32 N,
                                                          int accumulate) {
    const u64 off = (u64)blockIdx.y * N;
    const u32 k = blockIdx.x * kThreads + threadIdx.x;
    if (k >= N) return;
    const u64 v = src[off + perm[k]];
    acc[off + k] = accumulate ? add_mod(acc[off + k], v, mod[blockIdx.y].value) : v;
}

// tensor_prod_2x2_rns_poly (src/polymath.cu:463-498), Karatsuba middle term
// a and b may be d0 / d1 (the in-place multiply): only d2 is declared unaliased
__global__ __launch_bounds__(kThreads) void k_tensor(const u64 *a, const u64 *b, u64 *d0, u64 *d1,
                                                     u64 *__restrict__ d2,
                                                     const DevModulus *__restrict__ mod, u32 N, u64 LN) {
    const u32 limb = blockIdx.y;
    const u32 k = blockIdx.x * kThreads + threadIdx.x;
    if (k >= N) return;
    const u64 o = (u64)limb * N + k;
    const DevModulus m = mod[limb];
    u64 r0, r1, r2;
    tensor_coeff(a[o], a[LN + o], b[o], b[LN + o], m, r0, r1, r2);
    d0[o] = r0;
    d1[o] = r1;
    d2[o] = r2;
}

// apply_galois_ntt_permutation (src/galois.cu:11-18): dst[i] = src[perm[i]]
__global__ __launch_bounds__(kThreads) void k_permute(const u64 *__restrict__ src, u64 *__restrict__ dst,
                                                      const u32 *__restrict__ perm, u32 N) {
    const u64 off = (u64)blockIdx.y * N;
    const u32 k = blockIdx.x * kThreads + threadIdx.x;
    if (k >= N) return;
    dst[off + k] = src[off + perm[k]];
}

// element-wise kernels used by key generation / encryption / decryption
__global__ __launch_bounds__(kThreads) void k_mul_elementwise(const u64 *a, const u64 *b, u64 *c,
                                                              const DevModulus *mod, u32 N, u32 row0) {
    const u32 limb = blockIdx.y;
    const u32 k = blockIdx.x * kThreads + threadIdx.x;
    if (k >= N) return;
    const u64 o = (u64)limb * N + k;
    c[o] = mul_mod(a[o], b[o], mod[row0 + limb]);
}

__device__ __forceinline__ u64 mix64(u64 x) {
    x ^= x >> 30;
    x *= 0xbf58476d1ce4e5b9ull;
    x ^= x >> 27;
    x *= 0x94d049bb133111ebull;
    x ^= x >> 31;
    return x;
}
__device__ __forceinline__ u64 rnd(u64 seed, u64 a, u64 b) {
    return mix64(seed ^ mix64(a * 0x9e3779b97f4a7c15ull + b + 0x632be59bd9b4e019ull));
}
__device__ __forceinline__ u64 uniform_mod(u64 seed, u64 idx, u64 q) {
    const int bits = 64 - __clzll(q);
    const u64 mask = bits == 64 ? ~0ull : ((1ull << bits) - 1);
    for (u64 t = 0; t < 128; ++t) {
        const u64 r = rnd(seed, idx, t) & mask;
        if (r < q) return r;
    }
    return rnd(seed, idx, 999) % q;
}
// centred binomial error, variance 21/2 (the reference's 21+21-bit CBD,
// src/prng.cu:222-244)
__device__ __forceinline__ long long cbd21(u64 seed, u64 k) {
    const u64 r = rnd(seed, k, 7);
    return (long long)__popcll(r & 0x1fffffull) - (long long)__popcll((r >> 21) & 0x1fffffull);
}

// ternary secret, identical integer across limbs
__global__ void k_sample_ternary(u64 *s, const DevModulus *mod, u32 N, u64 seed) {
    const u32 limb = blockIdx.y;
    const u32 k = blockIdx.x * kThreads + threadIdx.x;
    if (k >= N) return;
    const u64 q = mod[limb].value;
    const u64 r = rnd(seed, k, 3) % 3;
    s[(u64)limb * N + k] = r == 0 ? 0 : (r == 1 ? 1 : q - 1);
}

__global__ void k_signed_to_rns(const int8_t *c, u64 *s, const DevModulus *mod, u32 N) {
    const u32 limb = blockIdx.y;
    const u32 k = blockIdx.x * kThreads + threadIdx.x;
    if (k >= N) return;
    const u64 q = mod[limb].value;
    const int v = c[k];
    s[(u64)limb * N + k] = v >= 0 ? (u64)v : q - (u64)(-v);
}

__global__ void k_sample_error(u64 *e, const DevModulus *mod, u32 N, u32 row0, u64 seed) {
    const u32 limb = blockIdx.y;
    const u32 k = blockIdx.x * kThreads + threadIdx.x;
    if (k >= N) return;
    const u64 q = mod[row0 + limb].value;
    const long long v = cbd21(seed, k);
    e[(u64)limb * N + k] = v >= 0 ? (u64)v : q - (u64)(-v);
}

// out = pk u + e (NTT form), one limb per grid row
// (multiply_and_add_rns_poly, src/polymath.cu)
__global__ void k_pk_mul_add(const u64 *pk, const u64 *u, const u64 *e, u64 *out, const DevModulus *mod, u32 N) {
    const u32 limb = blockIdx.y;
    const u32 k = blockIdx.x * kThreads + threadIdx.x;
    if (k >= N) return;
    const u64 o = (u64)limb * N + k;
    const DevModulus m = mod[limb];
    out[o] = add_mod(mul_mod(pk[o], u[o], m), e[o], m.value);
}

// a += b, one limb per grid row
__global__ void k_add_limbs(u64 *a, const u64 *b, const DevModulus *mod, u32 N) {
    const u32 limb = blockIdx.y;
    const u32 k = blockIdx.x * kThreads + threadIdx.x;
    if (k >= N) return;
    const u64 o = (u64)limb * N + k;
    a[o] = add_mod(a[o], b[o], mod[limb].value);
}

// c0 += (Q m + [-Q m]_t) / t = [[-Q m]_t t^-1]_{q_i}
// (bfv_add_timesQ_overt_kernel, src/polymath.cu:413-436)
__global__ void k_bfv_add_plain(u64 *c0, const u64 *pt, u64 negq_t, u64 negq_t_shoup, const u64 *tinv,
                                const DevModulus *mod, u64 t, u32 N) {
    const u32 limb = blockIdx.y;
    const u32 k = blockIdx.x * kThreads + threadIdx.x;
    if (k >= N) return;
    const u64 q = mod[limb].value;
    const u64 mq = mul_shoup(pt[k], negq_t, negq_t_shoup, t);
    const u64 o = (u64)limb * N + k;
    c0[o] = add_mod(c0[o], mul_shoup(mq, tinv[2 * limb], tinv[2 * limb + 1], q), q);
}

// c1 = a (uniform, NTT form), c0 = -(a s + e) + m (+ gadget * new_key on the
// digit's limbs). Rows are data rows 0..L-1 (keys: all QP rows).
// reference: encrypt_zero_symmetric src/secretkey.cu:232-289,
//            multiply_temp_mod_and_add_rns_poly src/polymath.cu:318-338
// a_in given: a is read from it (c1 may alias it), else drawn from the counter hash
__global__ void k_rlwe(u64 *c0, u64 *c1, const u64 *s, const u64 *e_ntt, const u64 *m_ntt, const u64 *gadget_key,
                       const u64 *bigP, u32 g_lo, u32 g_hi, const DevModulus *mod, u32 N, u64 seed,
                       const u64 *a_in) {
    const u32 limb = blockIdx.y;
    const u32 k = blockIdx.x * kThreads + threadIdx.x;
    if (k >= N) return;
    const u64 o = (u64)limb * N + k;
    const DevModulus md = mod[limb];
    const u64 q = md.value;
    const u64 a = a_in ? a_in[o] : uniform_mod(seed, o, q);
    u64 v = add_mod(mul_mod(a, s[o], md), e_ntt[o], q);
    v = v ? q - v : 0;
    if (m_ntt) v = add_mod(v, m_ntt[o], q);
    if (gadget_key && limb >= g_lo && limb < g_hi)
        v = add_mod(v, mul_shoup(gadget_key[o], bigP[2 * limb], bigP[2 * limb + 1], q), q);
    c0[o] = v;
    c1[o] = a;
}

// decrypt: m = c0 + c1 s + c2 s^2 (NTT form)
__global__ void k_dot_secret(const u64 *ct, const u64 *s, const u64 *s2, u64 *out, const DevModulus *mod, u32 N,
                             u32 L, u32 size, u64 s_stride) {
    const u32 limb = blockIdx.y;
    const u32 k = blockIdx.x * kThreads + threadIdx.x;
    if (k >= N) return;
    const u64 o = (u64)limb * N + k;
    const u64 LN = (u64)L * N;
    const DevModulus md = mod[limb];
    // a size-1 ciphertext is its c0 alone (needed_sk_power 0)
    u64 v = ct[o];
    if (size > 1) v = add_mod(v, mul_mod(ct[LN + o], s[o], md), md.value);
    if (size > 2) v = add_mod(v, mul_mod(ct[2 * LN + o], s2[o], md), md.value);
    out[o] = v;
}

// 16-byte streaming copy (nontemporal): each workgroup moves one 32 KiB chunk,
// every thread issuing its eight loads before its stores (~6.2 TB/s for 1 GiB
// each way on MI355X, against ~6.0 TB/s for a grid-stride loop over four
// workgroups per CU and ~4.6 TB/s for one vector per thread)
typedef unsigned int copy_vec __attribute__((ext_vector_type(4)));
constexpr u32 kCopyVecs = 8;   // vectors per thread
__global__ __launch_bounds__(kThreads) void k_copy(const copy_vec *__restrict__ src, copy_vec *__restrict__ dst,
                                                   u64 n16) {
    const u64 first = (u64)blockIdx.x * (kThreads * kCopyVecs);
    if (first + kThreads * kCopyVecs <= n16) {
        // a whole chunk: no bounds checks, which would serialise the loads
        const copy_vec *s = src + first + threadIdx.x;
        copy_vec *d = dst + first + threadIdx.x;
        copy_vec v[kCopyVecs];
#pragma unroll
        for (u32 e = 0; e < kCopyVecs; ++e) v[e] = __builtin_nontemporal_load(s + e * kThreads);
#pragma unroll
        for (u32 e = 0; e < kCopyVecs; ++e) __builtin_nontemporal_store(v[e], d + e * kThreads);
    } else {
        for (u64 i = first + threadIdx.x; i < n16; i += kThreads)
            __builtin_nontemporal_store(__builtin_nontemporal_load(src + i), dst + i);
    }
}

void copy_u64(u64 *dst, const u64 *src, size_t n, hipStream_t s) {
    require(n % 2 == 0 && ((uintptr_t)dst % 16) == 0 && ((uintptr_t)src % 16) == 0,
            "copy_u64 needs an even count and 16-byte aligned buffers");
    const u64 n16 = n / 2;
    const u64 per_block = (u64)kThreads * kCopyVecs;
    const u64 blocks = (n16 + per_block - 1) / per_block;
    if (blocks == 0) return;
    require(blocks <= 0x7fffffffu, "copy_u64: too many words");
    hipLaunchKernelGGL(k_copy, dim3((u32)blocks), dim3(kThreads), 0, s, reinterpret_cast<const copy_vec *>(src),
                       reinterpret_cast<copy_vec *>(dst), n16);
    PHX_HIP_CHECK(hipGetLastError());
}

// ---------------------------------------------------------------------------
// host orchestration
// ---------------------------------------------------------------------------
void Ciphertext::resize(const Context &ctx, u32 chain, u32 sz, hipStream_t s) {
    const u32 L = ctx.limbs_at(chain);
    const size_t need = (size_t)sz * L * ctx.N();
    if (need > capacity || !owns) {
        u64 *nd = static_cast<u64 *>(dev_alloc(need * 8, s));
        const size_t old = (size_t)size * coeff_modulus_size * poly_degree;
        if (data && old) PHX_HIP_CHECK(hipMemcpyAsync(nd, data, std::min(old, need) * 8, hipMemcpyDeviceToDevice, s));
        if (owns) dev_free(data, s);
        data = nd;
        capacity = need;
        owns = true;
    }
    chain_index = chain;
    size = sz;
    coeff_modulus_size = L;
    poly_degree = ctx.N();
}

void Ciphertext::release(hipStream_t s) {
    if (owns) dev_free(data, s);
    data = nullptr;
    capacity = 0;
}

void tensor_2x2(const Context &ctx, u32 Ql, const u64 *a, const u64 *b, u64 *d0, u64 *d1, u64 *d2, hipStream_t s) {
    tensor_2x2_mod(ctx.dev_mod(), ctx.N(), Ql, a, b, d0, d1, d2, s);
}

void tensor_2x2_mod(const DevModulus *mod, u64 N, u32 L, const u64 *a, const u64 *b, u64 *d0, u64 *d1, u64 *d2,
                    hipStream_t s) {
    hipLaunchKernelGGL(k_tensor, grid_for(N, L), dim3(kThreads), 0, s, a, b, d0, d1, d2, mod, (u32)N, (u64)L * N);
    PHX_HIP_CHECK(hipGetLastError());
}

// DRNSTool::modup for CKKS (src/rns_bconv.cu:530-628); copy_own = false leaves
// the digits' own limbs of t_mod_up unwritten (the inner product then reads c2).
// t_cks: Ql limbs of workspace for the scaled INTT of c2.
void modup_ws(const Context &ctx, u32 chain, const u64 *c2, u64 *t_mod_up, u64 *t_cks, hipStream_t s,
              bool copy_own, const LoadTensor *tensor) {
    const LevelConsts &L = ctx.level(chain);
    const u64 N = ctx.N();
    const u32 Ql = L.size_Ql, QlP = L.size_QlP, beta = L.beta, alpha = ctx.size_P();
    require(alpha <= (u32)kMaxBconvInputs, "special_modulus_size above 63 is not supported");
    // INTT(c2) * partQlHatInv (out of place); with `tensor`, c2 is the tensor
    // product computed (and written) by the first pass's load
    // the fused FastBConv takes its inputs byte-flipped from the INTT's store
    const bool fused = !copy_own && bconv_fused_ok(ctx, L.h_modup_fbconv);
    LimbMap mi = limb_map(t_cks, 0, Ql);
    mi.src = c2;
    mi.scale = L.modup_scale;
    if (tensor)
        ntt_inverse_tensor(ctx.tables(), mi, Ql, *tensor, s, fused);
    else
        ntt_inverse(ctx.tables(), mi, Ql, 1, s, fused);
    // the complement limbs of every digit in t_mod_up
    LimbMap mf = limb_map(t_mod_up, 0, QlP, (u64)QlP * N);
    mf.split = Ql;
    mf.special_row = ctx.size_Q();
    mf.skip_alpha = alpha;
    mf.skip_cap = Ql;
    mf.nlimbs = QlP;
    const u32 last_sz = Ql - alpha * (beta - 1);
    if (fused) {
        // FastBConv of every digit to its complement fused with the NTT's column
        // pass, then the row pass
        bconv_fwd_cols_run(ctx, L.h_modup_fbconv,
                           bconv_args(ctx, L.modup_fbconv, 1u, beta, t_cks, 0, t_mod_up, (u64)QlP * N, nullptr, 0),
                           beta, s);
        ntt_forward_rows(ctx.tables(), mf, QlP - last_sz, beta, s);
        return;
    }
    // FastBConv of every digit to its complement, own limbs copied, then the
    // forward NTT of the complements
    bconv_launch(ctx, L.modup_bconv, L.h_modup_bconv, 1u, beta, t_cks, 0, t_mod_up, (u64)QlP * N,
                 copy_own ? c2 : nullptr, 0, s);
    ntt_forward(ctx.tables(), mf, QlP - last_sz, beta, s);
}

void modup(const Context &ctx, u32 chain, const u64 *c2, u64 *t_mod_up, hipStream_t s, bool copy_own) {
    Scratch t_cks((size_t)ctx.level(chain).size_Ql * ctx.N() * 8, s);
    modup_ws(ctx, chain, c2, t_mod_up, t_cks.get(), s, copy_own);
}

void inner_product(const Context &ctx, u32 chain, const u64 *t_mod_up, const KSwitchKey &key, u64 *cx,
                   hipStream_t s, const u64 *c2) {
    const LevelConsts &L = ctx.level(chain);
    require(key.dnum >= L.beta, "key-switching key has too few digits for this level");
    const u64 N = ctx.N();
    hipLaunchKernelGGL(k_inner_product, grid_for(N, L.size_QlP), dim3(kThreads), 0, s, t_mod_up, c2,
                       (const u64 *)key.data, cx, ctx.dev_mod(), (u32)N, L.size_Ql, L.size_QlP, ctx.size_Q(),
                       ctx.size_QP(), L.beta, ctx.size_P());
    PHX_HIP_CHECK(hipGetLastError());
}

// DRNSTool::moddown_from_NTT for CKKS (src/rns_bconv.cu:776-828), both halves
// in one launch per step, the tail fused into the last NTT pass. delta:
// npolys x Ql limbs of workspace for the base conversion P -> Ql.
void moddown_ws(const Context &ctx, u32 chain, u64 *cx, u64 *ct, const u64 *base0, const u64 *base1, u32 npolys,
                u64 *delta, hipStream_t s) {
    const LevelConsts &L = ctx.level(chain);
    const u64 N = ctx.N();
    const u32 Ql = L.size_Ql, QlP = L.size_QlP, P = ctx.size_P();
    // INTT of the P limbs, scaled by PHatInv (FastBConv phase 1)
    LimbMap mi = limb_map(cx + (u64)Ql * N, ctx.size_Q(), P, (u64)QlP * N);
    mi.scale = L.moddown_scale;
    const bool fused = bconv_fused_ok(ctx, {L.h_moddown_fbconv});
    ntt_inverse(ctx.tables(), mi, P, npolys, s, fused);
    if (fused)
        bconv_fwd_cols_run(ctx, {L.h_moddown_fbconv},
                           bconv_args(ctx, L.moddown_fbconv, 0u, npolys, cx, (u64)QlP * N, delta, (u64)Ql * N,
                                      nullptr, 0),
                           npolys, s);
    else
        bconv_launch(ctx, L.moddown_bconv, {L.h_moddown_bconv}, 0u, npolys, cx, (u64)QlP * N, delta, (u64)Ql * N,
                     nullptr, 0, s);
    StoreModDown st{};
    st.ct = ct;
    st.base0 = base0;
    st.base1 = base1;
    st.cx = cx;
    st.pinv = L.pinv;
    st.ct_stride = (u64)Ql * N;
    st.cx_stride = (u64)QlP * N;
    st.N = (u32)N;
    if (fused)
        ntt_forward_rows_store(ctx.tables(), limb_map(delta, 0, Ql, (u64)Ql * N), Ql, npolys, st, s);
    else
        ntt_forward_store(ctx.tables(), limb_map(delta, 0, Ql, (u64)Ql * N), Ql, npolys, st, s);
}

void moddown(const Context &ctx, u32 chain, u64 *cx, u64 *ct, const u64 *base0, const u64 *base1, u32 npolys,
             hipStream_t s) {
    Scratch delta((size_t)npolys * ctx.level(chain).size_Ql * ctx.N() * 8, s);
    moddown_ws(ctx, chain, cx, ct, base0, base1, npolys, delta.get(), s);
}

// workspace of one key switch at chain: t_mod_up (beta x QlP limbs), cx (2 x QlP),
// the scaled INTT of c2 (Ql) and delta (2 x Ql)
size_t keyswitch_ws_words(const Context &ctx, u32 chain) {
    const LevelConsts &L = ctx.level(chain);
    return (size_t)ctx.N() * ((size_t)L.beta * L.size_QlP + 2ull * L.size_QlP + 3ull * L.size_Ql);
}

// keyswitch_inplace (src/eval_key_switch.cu:95-182) in a caller's workspace
void keyswitch_ws(const Context &ctx, u32 chain, const u64 *c2, const KSwitchKey &key, u64 *ct, const u64 *base0,
                  const u64 *base1, u64 *ws, hipStream_t s, const LoadTensor *tensor) {
    const LevelConsts &L = ctx.level(chain);
    require(key.data != nullptr, "key-switching key is empty");
    require(key.dnum >= L.beta, "key-switching key has too few digits for this level");
    const u64 N = ctx.N();
    u64 *t_mod_up = ws;
    u64 *cx = t_mod_up + (u64)L.beta * L.size_QlP * N;
    u64 *t_cks = cx + 2ull * L.size_QlP * N;
    u64 *delta = t_cks + (u64)L.size_Ql * N;
    modup_ws(ctx, chain, c2, t_mod_up, t_cks, s, false, tensor);
    inner_product(ctx, chain, t_mod_up, key, cx, s, c2);
    moddown_ws(ctx, chain, cx, ct, base0, base1, 2, delta, s);
}

void keyswitch(const Context &ctx, u32 chain, const u64 *c2, const KSwitchKey &key, u64 *ct, const u64 *base0,
               const u64 *base1, hipStream_t s) {
    Scratch ws(keyswitch_ws_words(ctx, chain) * 8, s);
    keyswitch_ws(ctx, chain, c2, key, ct, base0, base1, ws.get(), s);
}

size_t multiply_and_relin_ws_words(const Context &ctx, u32 chain) {
    return (size_t)ctx.level(chain).size_Ql * ctx.N() + keyswitch_ws_words(ctx, chain);
}

// CKKS multiply_and_relin_inplace in a caller's workspace: c2 (Ql limbs) then
// the key switch's workspace
void multiply_and_relin_ws(const Context &ctx, Ciphertext &ct1, const Ciphertext &ct2, const KSwitchKey &rlk,
                           u64 *ws, hipStream_t s) {
    require(ctx.params().scheme == kSchemeCKKS, "multiply_and_relin: unsupported scheme");
    require(ct1.chain_index == ct2.chain_index, "encrypted1 and encrypted2 parameter mismatch");
    require(ct1.is_ntt_form == ct2.is_ntt_form, "NTT form mismatch");
    require(ct1.is_ntt_form, "encrypted1 and encrypted2 must be in NTT form");
    require(ct1.size == 2 && ct2.size == 2, "poly number mismatch");
    const u32 Ql = ct1.coeff_modulus_size;
    const u64 N = ctx.N(), LN = (u64)Ql * N;
    u64 *c2 = ws;
    // the tensor product is the first load of the modup INTT (d0, d1 into ct1, d2 into c2)
    const LoadTensor tl{ct1.data, ct2.data, ct1.data, c2, LN, (u32)N};
    keyswitch_ws(ctx, ct1.chain_index, c2, rlk, ct1.data, ct1.data, ct1.data + LN, ws + LN, s, &tl);
    ct1.scale = ct1.scale * ct2.scale;
}

// multiply_and_relin_inplace for CKKS (src/evaluate.cu:1061-1104)
void multiply_and_relin(const Context &ctx, Ciphertext &ct1, const Ciphertext &ct2, const KSwitchKey &rlk,
                        hipStream_t s) {
    if (ctx.params().scheme == kSchemeBFV) {
        // BEHZ multiply then relinearize (src/evaluate.cu:1087-1093); for hps and
        // hps_overq the reference's fused bfv_mul_relin_hps (818-1027) runs the same
        // two steps
        require(ctx.params().mul_tech == kMulBEHZ || ctx.params().mul_tech == kMulHPS ||
                    ctx.params().mul_tech == kMulHPSOverQ || ctx.params().mul_tech == kMulHPSOverQLeveled,
                "unsupported mul tech in BFV mul&relin");
        if (ctx.params().mul_tech == kMulHPSOverQLeveled) return bfv_mul_relin_leveled(ctx, ct1, ct2, rlk, s);
        bfv_multiply(ctx, ct1, ct2, s);
        bfv_relinearize(ctx, ct1, rlk, s);
        return;
    }
    require(ctx.params().scheme == kSchemeCKKS, "multiply_and_relin: unsupported scheme");
    Scratch ws(multiply_and_relin_ws_words(ctx, ct1.chain_index) * 8, s);
    multiply_and_relin_ws(ctx, ct1, ct2, rlk, ws.get(), s);
}

// relinearize_inplace (src/evaluate.cu:1342-1374)
void relinearize(const Context &ctx, Ciphertext &ct, const KSwitchKey &rlk, hipStream_t s) {
    if (ctx.params().scheme == kSchemeBFV) return bfv_relinearize(ctx, ct, rlk, s);
    require(ct.size == 3, "destination_size must be 3");
    require(ct.is_ntt_form, "CKKS encrypted must be in NTT form");
    const u64 LN = (u64)ct.coeff_modulus_size * ctx.N();
    keyswitch(ctx, ct.chain_index, ct.data + 2 * LN, rlk, ct.data, ct.data, ct.data + LN, s);
    ct.size = 2;
}

// rescale_to_next (src/evaluate.cu:1545-1565, 1376-1427; src/rns.cu:1160-1184)
// BFV modulus switch: c_j' = (c_j - [c_last]_{q_j}) q_last^-1 mod q_j on the
// coefficient form, one grid row per kept limb, one z slice per polynomial
// (divide_and_round_q_last_kernel, src/rns.cu:1082-1126: despite its name a floor)
__global__ void k_divide_q_last(const u64 *in, u64 *out, const u64 *qinv, const DevModulus *mod, u32 N, u32 Ql) {
    const u32 j = blockIdx.y, poly = blockIdx.z;
    const u32 k = blockIdx.x * kThreads + threadIdx.x;
    if (k >= N) return;
    const DevModulus m = mod[j];
    const u64 *src = in + (u64)poly * Ql * N;
    const u64 last = barrett64(src[(u64)(Ql - 1) * N + k], m.value, m.ratio_hi);
    const u64 d = sub_mod(src[(u64)j * N + k], last, m.value);
    out[((u64)poly * (Ql - 1) + j) * N + k] = mul_shoup(d, qinv[2 * j], qinv[2 * j + 1], m.value);
}

void mod_switch_to_next(const Context &ctx, const Ciphertext &in, Ciphertext &out, hipStream_t s) {
    require(in.chain_index >= 1, "invalid chain index");
    require(in.chain_index < ctx.size_Q(), "end of modulus switching chain reached");
    require(&in != &out, "mod_switch_to_next: in and out must differ");
    const bool bfv = ctx.params().scheme == kSchemeBFV;
    require(bfv || ctx.params().scheme == kSchemeCKKS, "unsupported scheme");
    if (bfv)
        require(!in.is_ntt_form, "BFV encrypted cannot be in NTT form");
    else
        require(in.is_ntt_form, "CKKS encrypted must be in NTT form");
    const u64 N = ctx.N();
    const u32 L = in.coeff_modulus_size, size = in.size;
    out.resize(ctx, in.chain_index + 1, size, s);
    const u32 Ln = out.coeff_modulus_size;
    if (bfv) {
        // mod_switch_scale_to_next (src/evaluate.cu:1376-1427)
        hipLaunchKernelGGL(k_divide_q_last, grid_for(N, Ln, size), dim3(kThreads), 0, s, (const u64 *)in.data,
                           out.data, ctx.level(in.chain_index).rescale_qinv, ctx.dev_mod(), (u32)N, L);
        PHX_HIP_CHECK(hipGetLastError());
    } else {
        // mod_switch_drop_to_next (src/evaluate.cu:1429-1472)
        for (u32 p = 0; p < size; ++p)
            PHX_HIP_CHECK(hipMemcpyAsync(out.data + (u64)p * Ln * N, in.data + (u64)p * L * N, (size_t)Ln * N * 8,
                                         hipMemcpyDeviceToDevice, s));
    }
    out.scale = in.scale;
    out.is_ntt_form = !bfv;
    out.correction_factor = in.correction_factor;
    out.noise_scale_deg = in.noise_scale_deg;
    out.is_asymmetric = in.is_asymmetric;
    out.has_seed = false;
}

void rescale_to_next(const Context &ctx, const Ciphertext &in, Ciphertext &out, hipStream_t s) {
    require(ctx.params().scheme == kSchemeCKKS, "unsupported scheme");
    require(in.chain_index >= 1, "invalid chain index");
    require(in.chain_index < ctx.size_Q(), "end of modulus switching chain reached");
    const LevelConsts &L = ctx.level(in.chain_index);
    const u32 Ql = L.size_Ql, size = in.size;
    const u64 N = ctx.N();
    Scratch last_buf((size_t)size * N * 8, s);
    u64 *last = last_buf.get();
    LimbMap mi = limb_map(last, Ql - 1, 1, N);
    mi.src = in.data + (u64)(Ql - 1) * N;
    mi.src_poly_stride = (u64)Ql * N;
    ntt_inverse(ctx.tables(), mi, 1, size, s);
    Ciphertext dst;
    dst.resize(ctx, in.chain_index + 1, size, s);
    try {
        LoadRescale ld{last, N};
        StoreRescale st{in.data, dst.data, L.rescale_qinv, (u64)Ql * N, (u64)(Ql - 1) * N, (u32)N};
        ntt_forward_rescale(ctx.tables(), limb_map(dst.data, 0, Ql - 1, (u64)(Ql - 1) * N), Ql - 1, size, ld, st,
                            s);
    } catch (...) {
        dst.release(s);
        throw;
    }
    dst.is_ntt_form = in.is_ntt_form;
    dst.scale = in.scale / (double)ctx.moduli()[Ql - 1];
    dst.correction_factor = in.correction_factor;
    if (&out != &in) out.release(s);
    out = dst;
}

// hoisting_inplace for CKKS (src/evaluate.cu:1670-1865): ct <- sum_i rot(ct, steps[i])
// with one modup of c1; each step's permutation is applied to c0 and, inside
// the inner product, to the modup'ed digits; one moddown of the accumulated cx.
namespace {
void hoisting_at(const Context &ctx, Ciphertext &ct, const GaloisKeys &gk, const std::vector<int> &steps,
                 hipStream_t s) {
    const bool bfv = ctx.params().scheme == kSchemeBFV;
    require(bfv || ctx.params().scheme == kSchemeCKKS, "hoisting: unsupported scheme");
    require(ct.size == 2, "ciphertext size must be 2");
    if (bfv)
        require(!ct.is_ntt_form, "BFV encrypted cannot be in NTT form");
    else
        require(ct.is_ntt_form, "CKKS encrypted must be in NTT form");
    require(!steps.empty(), "hoisting needs at least one step");
    const LevelConsts &L = ctx.level(ct.chain_index);
    const u32 Ql = L.size_Ql, QlP = L.size_QlP, beta = L.beta;
    const u64 N = ctx.N(), LN = (u64)Ql * N, LPN = (u64)QlP * N;
    std::vector<std::pair<const u32 *, const KSwitchKey *>> items;
    for (int st : steps) {
        const u32 elt = galois_elt_from_step(st, N);
        auto it = std::find(gk.elts.begin(), gk.elts.end(), elt);
        require(it != gk.elts.end(), "Galois key not present in hoisting");
        const KSwitchKey &key = gk.keys[it - gk.elts.begin()];
        require(key.dnum >= beta, "key-switching key has too few digits for this level");
        items.emplace_back(ctx.galois_perm(elt), &key);
    }
    // BFV (coefficient form): the reference permutes c0 in the coefficient domain and
    // runs its BFV modup / moddown; both equal the NTT-domain form conjugated by the
    // exact transform, so the ciphertext is transformed, hoisted and transformed back
    if (bfv) ntt_forward(ctx.tables(), limb_map(ct.data, 0, Ql, LN), Ql, 2, s);
    Scratch t_buf((size_t)beta * LPN * 8, s);
    u64 *t = t_buf.get();
    Scratch cx_buf(2 * LPN * 8, s);
    u64 *cx = cx_buf.get();
    Scratch acc0_buf(LN * 8, s);
    u64 *acc0 = acc0_buf.get();
    modup(ctx, ct.chain_index, ct.data + LN, t, s, true);
    for (size_t i = 0; i < items.size(); ++i) {
        const int acc = i > 0;
        hipLaunchKernelGGL(k_permute_acc, grid_for(N, Ql), dim3(kThreads), 0, s, (const u64 *)ct.data, acc0,
                           items[i].first, ctx.dev_mod(), (u32)N, acc);
        hipLaunchKernelGGL(k_inner_product_perm_acc, grid_for(N, QlP), dim3(kThreads), 0, s, (const u64 *)t,
                           items[i].first, (const u64 *)items[i].second->data, cx, ctx.dev_mod(), (u32)N, Ql, QlP,
                           ctx.size_Q(), ctx.size_QP(), beta, acc);
        PHX_HIP_CHECK(hipGetLastError());
    }
    // ct0 = acc0 + moddown(cx0), ct1 = moddown(cx1)
    moddown(ctx, ct.chain_index, cx, ct.data, acc0, nullptr, 2, s);
    if (bfv) ntt_inverse(ctx.tables(), limb_map(ct.data, 0, Ql, LN), Ql, 2, s);
}
}  // namespace

void hoisting(const Context &ctx, Ciphertext &ct, const GaloisKeys &gk, const std::vector<int> &steps,
              hipStream_t s) {
    // hps_overq_leveled (src/evaluate.cu:1688-1865): c0 and c1 scaled down to the
    // dropped level, hoisted there, expanded back to Q
    const u32 d = ct.size == 2 && !ct.is_ntt_form ? bfv_keyswitch_drop(ctx, ct, true) : 0;
    if (!d) return hoisting_at(ctx, ct, gk, steps, s);
    const u64 N = ctx.N(), QN0 = (u64)ct.coeff_modulus_size * N;
    Ciphertext t;
    t.resize(ctx, ct.chain_index + d, 2, s);
    t.is_ntt_form = false;
    const u64 LN = (u64)t.coeff_modulus_size * N;
    try {
        hps_scale_q_ql(ctx, d, ct.data, QN0, t.data, LN, 2, s);
        hoisting_at(ctx, t, gk, steps, s);
        hps_expand_ql_q(ctx, d, t.data, LN, ct.data, QN0, 2, false, s);
    } catch (...) {
        t.release(s);
        throw;
    }
    t.release(s);
}

// apply_galois_inplace (src/evaluate.cu:1567-1630)
void apply_galois(const Context &ctx, Ciphertext &ct, u32 elt, const GaloisKeys &gk, hipStream_t s) {
    require(ct.size == 2, "encrypted size must be 2");
    auto it = std::find(gk.elts.begin(), gk.elts.end(), elt);
    require(it != gk.elts.end(), "Galois elt not present");
    const KSwitchKey &key = gk.keys[it - gk.elts.begin()];
    const bool bfv = ctx.params().scheme == kSchemeBFV;
    if (bfv)
        require(!ct.is_ntt_form, "BFV encrypted cannot be in NTT form");
    else
        require(ct.is_ntt_form, "CKKS encrypted must be in NTT form");
    const u32 Ql = ct.coeff_modulus_size;
    const u64 N = ctx.N(), LN = (u64)Ql * N;
    const u32 *perm = ctx.galois_perm(elt);
    if (const u32 d = bfv ? bfv_keyswitch_drop(ctx, ct, true) : 0) {
        // hps_overq_leveled (src/evaluate.cu:1567-1630 with src/eval_key_switch.cu:110-180):
        // c0 <- pi(c0) + expand(ks0), c1 <- expand(ks1), ks = the key switch of pi(c1)
        // scaled down to the dropped level; pi in the coefficient domain is the NTT-domain
        // permutation conjugated by the exact transform
        const u32 Qd = Ql - d;
        const u64 DN = (u64)Qd * N;
        Scratch tmp_buf(2 * LN * 8, s);
        u64 *tmp = tmp_buf.get();
        Scratch c2_buf(DN * 8, s);
        u64 *c2 = c2_buf.get();
        Scratch ks_buf(2 * DN * 8, s);
        u64 *ks = ks_buf.get();
        ntt_forward(ctx.tables(), limb_map(ct.data, 0, Ql, LN), Ql, 2, s);
        hipLaunchKernelGGL(k_permute, grid_for(N, 2 * Ql), dim3(kThreads), 0, s, (const u64 *)ct.data, tmp, perm,
                           (u32)N);
        PHX_HIP_CHECK(hipGetLastError());
        ntt_inverse(ctx.tables(), limb_map(tmp, 0, Ql, LN), Ql, 2, s);
        hps_scale_q_ql(ctx, d, tmp + LN, LN, c2, DN, 1, s);
        ntt_forward(ctx.tables(), limb_map(c2, 0, Qd), Qd, 1, s);
        keyswitch(ctx, ct.chain_index + d, c2, key, ks, nullptr, nullptr, s);
        ntt_inverse(ctx.tables(), limb_map(ks, 0, Qd, DN), Qd, 2, s);
        PHX_HIP_CHECK(hipMemcpyAsync(ct.data, tmp, LN * 8, hipMemcpyDeviceToDevice, s));
        hps_expand_ql_q(ctx, d, ks, DN, ct.data, LN, 1, true, s);
        hps_expand_ql_q(ctx, d, ks + DN, DN, ct.data + LN, LN, 1, false, s);
        return;
    }
    // BFV (coefficient form): the reference's coefficient-domain automorphism and
    // its BFV key switch equal the NTT-domain permutation and key switch
    // conjugated by the exact transform, so the ciphertext is transformed,
    // rotated and transformed back
    if (bfv) ntt_forward(ctx.tables(), limb_map(ct.data, 0, Ql, LN), Ql, 2, s);
    Scratch tmp_buf(2 * LN * 8, s);
    u64 *tmp = tmp_buf.get();
    hipLaunchKernelGGL(k_permute, grid_for(N, 2 * Ql), dim3(kThreads), 0, s, (const u64 *)ct.data, tmp, perm,
                       (u32)N);
    PHX_HIP_CHECK(hipGetLastError());
    // (c0, c1) <- (pi(c0), 0) + keyswitch(pi(c1))
    keyswitch(ctx, ct.chain_index, tmp + LN, key, ct.data, tmp, nullptr, s);
    if (bfv) ntt_inverse(ctx.tables(), limb_map(ct.data, 0, Ql, LN), Ql, 2, s);
}

u32 galois_elt_from_step(int step, u64 N) {
    // get_elt_from_step (include/galois.cuh:16-49)
    const u64 m = 2 * N;
    if (step == 0) return (u32)(m - 1);
    const bool neg = step < 0;
    const u64 ps = (u64)(neg ? -(long long)step : step);
    require(ps < N / 2, "step count too large");
    long long e = neg ? (long long)(N / 2) - (long long)ps : (long long)ps;
    u64 g = 1;
    while (e--) g = (g * 5) & (m - 1);
    return (u32)g;
}

namespace {
// non-adjacent form of step (reference: naf(), src/evaluate.cu rotate_internal)
std::vector<int> naf(int value) {
    std::vector<int> res;
    const bool sign = value < 0;
    value = std::abs(value);
    for (int i = 0; value; ++i) {
        const int zi = (value & 1) ? 2 - (value & 3) : 0;
        value = (value - zi) >> 1;
        if (zi) res.push_back((sign ? -zi : zi) * (1 << i));
    }
    return res;
}
}  // namespace

// rotate_inplace -> rotate_internal (src/evaluate.cu:1633-1668)
void rotate(const Context &ctx, Ciphertext &ct, int step, const GaloisKeys &gk, hipStream_t s) {
    const u32 elt = galois_elt_from_step(step, ctx.N());
    if (std::find(gk.elts.begin(), gk.elts.end(), elt) != gk.elts.end()) {
        apply_galois(ctx, ct, elt, gk, s);
        return;
    }
    const auto steps = naf(step);
    require(steps.size() != 1, "Galois key not present");
    for (int st : steps)
        if ((u64)std::abs(st) != (ctx.N() >> 1)) rotate(ctx, ct, st, gk, s);
}

// ---------------------------------------------------------------------------
// keys and encryption (test-side key material; seeded, reproducible)
// ---------------------------------------------------------------------------
void secret_key_generate(const Context &ctx, SecretKey &sk, u64 seed, hipStream_t s) {
    const u64 N = ctx.N();
    const u32 QP = ctx.size_QP();
    sk.s = static_cast<u64 *>(dev_alloc((size_t)QP * N * 8, s));
    hipLaunchKernelGGL(k_sample_ternary, grid_for(N, QP), dim3(kThreads), 0, s, sk.s, ctx.dev_mod(), (u32)N, seed);
    ntt_forward(ctx.tables(), limb_map(sk.s, 0, QP), QP, 1, s);
}

void secret_key_upload(const Context &ctx, SecretKey &sk, const int8_t *coeffs, hipStream_t s) {
    const u64 N = ctx.N();
    const u32 QP = ctx.size_QP();
    sk.s = static_cast<u64 *>(dev_alloc((size_t)QP * N * 8, s));
    Scratch d_buf(N, s);
    int8_t *d = reinterpret_cast<int8_t *>(d_buf.get());
    PHX_HIP_CHECK(hipMemcpyAsync(d, coeffs, N, hipMemcpyHostToDevice, s));
    hipLaunchKernelGGL(k_signed_to_rns, grid_for(N, QP), dim3(kThreads), 0, s, (const int8_t *)d, sk.s,
                       ctx.dev_mod(), (u32)N);
    ntt_forward(ctx.tables(), limb_map(sk.s, 0, QP), QP, 1, s);
}

// generate_one_kswitch_key (src/secretkey.cu:297-341), dnum = ceil(size_Q / size_P)
void kswitch_key_generate(const Context &ctx, const SecretKey &sk, const u64 *new_key, KSwitchKey &key, u64 seed,
                          hipStream_t s) {
    const u64 N = ctx.N();
    const u32 QP = ctx.size_QP(), alpha = ctx.size_P(), dnum = ctx.dnum();
    const u64 QPN = (u64)QP * N;
    key.dnum = dnum;
    key.poly_degree = N;
    key.size_QP = QP;
    key.data = static_cast<u64 *>(dev_alloc((size_t)dnum * 2 * QPN * 8, s));
    Scratch e_buf(QPN * 8, s);
    u64 *e = e_buf.get();
    for (u32 j = 0; j < dnum; ++j) {
        const u64 sj = seed * 0x100000001b3ull + j * 0x9e3779b97f4a7c15ull + 1;
        hipLaunchKernelGGL(k_sample_error, grid_for(N, QP), dim3(kThreads), 0, s, e, ctx.dev_mod(), (u32)N, 0u,
                           sj ^ 0x5555);
        ntt_forward(ctx.tables(), limb_map(e, 0, QP), QP, 1, s);
        u64 *b = key.data + (u64)j * 2 * QPN;
        const u32 g_lo = alpha * j, g_hi = std::min(alpha * (j + 1), ctx.size_Q());
        hipLaunchKernelGGL(k_rlwe, grid_for(N, QP), dim3(kThreads), 0, s, b, b + QPN, (const u64 *)sk.s,
                           (const u64 *)e, (const u64 *)nullptr, new_key, ctx.bigP_mod_q(), g_lo, g_hi,
                           ctx.dev_mod(), (u32)N, sj, (const u64 *)nullptr);
    }
    PHX_HIP_CHECK(hipGetLastError());
}

void relin_key_generate(const Context &ctx, SecretKey &sk, KSwitchKey &key, u64 seed, hipStream_t s) {
    const u64 N = ctx.N();
    const u32 QP = ctx.size_QP();
    if (!sk.s2) {
        sk.s2 = static_cast<u64 *>(dev_alloc((size_t)QP * N * 8, s));
        hipLaunchKernelGGL(k_mul_elementwise, grid_for(N, QP), dim3(kThreads), 0, s, (const u64 *)sk.s,
                           (const u64 *)sk.s, sk.s2, ctx.dev_mod(), (u32)N, 0u);
    }
    kswitch_key_generate(ctx, sk, sk.s2, key, seed, s);
}

// create_galois_keys (src/secretkey.cu:421-461): key for pi_elt(s)
void galois_key_generate(const Context &ctx, const SecretKey &sk, u32 elt, KSwitchKey &key, u64 seed,
                         hipStream_t s) {
    const u64 N = ctx.N();
    const u32 QP = ctx.size_QP();
    Scratch rs_buf((size_t)QP * N * 8, s);
    u64 *rs = rs_buf.get();
    hipLaunchKernelGGL(k_permute, grid_for(N, QP), dim3(kThreads), 0, s, (const u64 *)sk.s, rs, ctx.galois_perm(elt),
                       (u32)N);
    kswitch_key_generate(ctx, sk, rs, key, seed, s);
}

namespace {
void encrypt_rlwe(const Context &ctx, const SecretKey &sk, const u64 *plain_dev, u32 chain, Ciphertext &ct, u64 seed,
                  const uint8_t *seed_a, hipStream_t s) {
    const u64 N = ctx.N();
    const u32 L = ctx.limbs_at(chain);
    ct.resize(ctx, chain, 2, s);
    ct.is_ntt_form = true;
    ct.has_seed = seed_a != nullptr;
    if (seed_a) {
        std::memcpy(ct.seed, seed_a, sizeof(ct.seed));
        sample_uniform_seeded(ctx, seed_a, L, ct.data + (u64)L * N, s);
    }
    Scratch m_buf((size_t)L * N * 8, s);
    u64 *m = m_buf.get();
    Scratch e_buf((size_t)L * N * 8, s);
    u64 *e = e_buf.get();
    PHX_HIP_CHECK(hipMemcpyAsync(m, plain_dev, (size_t)L * N * 8, hipMemcpyDeviceToDevice, s));
    ntt_forward(ctx.tables(), limb_map(m, 0, L), L, 1, s);
    hipLaunchKernelGGL(k_sample_error, grid_for(N, L), dim3(kThreads), 0, s, e, ctx.dev_mod(), (u32)N, 0u,
                       seed ^ 0x7777);
    ntt_forward(ctx.tables(), limb_map(e, 0, L), L, 1, s);
    hipLaunchKernelGGL(k_rlwe, grid_for(N, L), dim3(kThreads), 0, s, ct.data, ct.data + (u64)L * N,
                       (const u64 *)sk.s, (const u64 *)e, (const u64 *)m, (const u64 *)nullptr, (const u64 *)nullptr,
                       0u, 0u, ctx.dev_mod(), (u32)N, seed, seed_a ? (const u64 *)(ct.data + (u64)L * N) : nullptr);
    PHX_HIP_CHECK(hipGetLastError());
}
}  // namespace

void encrypt_symmetric(const Context &ctx, const SecretKey &sk, const u64 *plain_dev, u32 chain, Ciphertext &ct,
                       u64 seed, hipStream_t s) {
    encrypt_rlwe(ctx, sk, plain_dev, chain, ct, seed, nullptr, s);
}

void encrypt_symmetric_seeded(const Context &ctx, const SecretKey &sk, const u64 *plain_dev, u32 chain,
                              Ciphertext &ct, const uint8_t *seed_a, u64 seed_e, hipStream_t s) {
    require(seed_a != nullptr, "seed is null");
    encrypt_rlwe(ctx, sk, plain_dev, chain, ct, seed_e, seed_a, s);
}

void public_key_generate(const Context &ctx, const SecretKey &sk, const uint8_t *seed_a, u64 seed_e,
                         Ciphertext &pk, hipStream_t s) {
    require(sk.s != nullptr, "secret key is empty");
    const u64 N = ctx.N();
    const u32 QP = ctx.size_QP();
    pk.resize(ctx, 0, 2, s);
    pk.is_ntt_form = true;
    pk.is_asymmetric = false;
    pk.has_seed = seed_a != nullptr;
    u64 *c1 = pk.data + (u64)QP * N;
    if (seed_a) {
        std::memcpy(pk.seed, seed_a, sizeof(pk.seed));
        sample_uniform_seeded(ctx, seed_a, QP, c1, s);
    }
    Scratch e_buf((size_t)QP * N * 8, s);
    u64 *e = e_buf.get();
    hipLaunchKernelGGL(k_sample_error, grid_for(N, QP), dim3(kThreads), 0, s, e, ctx.dev_mod(), (u32)N, 0u,
                       seed_e ^ 0x5a5au);
    ntt_forward(ctx.tables(), limb_map(e, 0, QP), QP, 1, s);
    hipLaunchKernelGGL(k_rlwe, grid_for(N, QP), dim3(kThreads), 0, s, pk.data, c1, (const u64 *)sk.s,
                       (const u64 *)e, (const u64 *)nullptr, (const u64 *)nullptr, (const u64 *)nullptr, 0u, 0u,
                       ctx.dev_mod(), (u32)N, seed_e, seed_a ? (const u64 *)c1 : nullptr);
    PHX_HIP_CHECK(hipGetLastError());
}

void encrypt_asymmetric(const Context &ctx, const Ciphertext &pk, const u64 *plain_dev, u32 chain,
                        const int8_t *u_dev, const int8_t *const e_dev[2], u64 seed, Ciphertext &ct, hipStream_t s) {
    const bool bfv = ctx.params().scheme == kSchemeBFV;
    require(bfv || ctx.params().scheme == kSchemeCKKS, "unsupported scheme.");
    require(pk.data != nullptr && pk.chain_index == 0 && pk.size == 2 && pk.coeff_modulus_size == ctx.size_QP(),
            "PhantomPublicKey has not been generated");
    require(pk.poly_degree == ctx.N(), "the public key does not match the context (ring degree)");
    require(chain == 1, "asymmetric encryption runs at the first data level");
    require(plain_dev != nullptr, "plaintext is null");
    if (bfv) require(ctx.bfv_tinv_q() != nullptr, "plain_modulus is not set");
    const u64 N = ctx.N();
    const u32 QP = ctx.size_QP(), Ql = ctx.size_Q();
    const u64 QPN = (u64)QP * N, LN = (u64)Ql * N;
    Scratch u_buf(QPN * 8, s), e_buf(QPN * 8, s), c_buf(2 * QPN * 8, s);
    u64 *u = u_buf.get(), *e = e_buf.get(), *c = c_buf.get();
    // u ternary and e_j, shared across limbs (sample_ternary_poly / sample_error_poly)
    if (u_dev)
        hipLaunchKernelGGL(k_signed_to_rns, grid_for(N, QP), dim3(kThreads), 0, s, u_dev, u, ctx.dev_mod(), (u32)N);
    else
        hipLaunchKernelGGL(k_sample_ternary, grid_for(N, QP), dim3(kThreads), 0, s, u, ctx.dev_mod(), (u32)N,
                           seed ^ 0x7575u);
    ntt_forward(ctx.tables(), limb_map(u, 0, QP), QP, 1, s);
    for (u32 j = 0; j < 2; ++j) {
        if (e_dev && e_dev[j])
            hipLaunchKernelGGL(k_signed_to_rns, grid_for(N, QP), dim3(kThreads), 0, s, e_dev[j], e, ctx.dev_mod(),
                               (u32)N);
        else
            hipLaunchKernelGGL(k_sample_error, grid_for(N, QP), dim3(kThreads), 0, s, e, ctx.dev_mod(), (u32)N, 0u,
                               seed ^ (0x6565u + j));
        ntt_forward(ctx.tables(), limb_map(e, 0, QP), QP, 1, s);
        hipLaunchKernelGGL(k_pk_mul_add, grid_for(N, QP), dim3(kThreads), 0, s, (const u64 *)pk.data + j * QPN,
                           (const u64 *)u, (const u64 *)e, c + j * QPN, ctx.dev_mod(), (u32)N);
        PHX_HIP_CHECK(hipGetLastError());
    }
    // QP -> Q (DRNSTool::moddown, src/rns_bconv.cu:712-760: for BFV the same in the
    // coefficient domain, i.e. this NTT-domain form conjugated by the exact transform)
    ct.resize(ctx, chain, 2, s);
    moddown(ctx, chain, c, ct.data, nullptr, nullptr, 2, s);
    if (bfv) {
        ntt_inverse(ctx.tables(), limb_map(ct.data, 0, Ql, LN), Ql, 2, s);
        hipLaunchKernelGGL(k_bfv_add_plain, grid_for(N, Ql), dim3(kThreads), 0, s, ct.data, plain_dev,
                           ctx.bfv_negq_mod_t(), ctx.bfv_negq_mod_t_shoup(), ctx.bfv_tinv_q(), ctx.dev_mod(),
                           ctx.params().plain_modulus, (u32)N);
    } else {
        Scratch m_buf(LN * 8, s);
        PHX_HIP_CHECK(hipMemcpyAsync(m_buf.get(), plain_dev, LN * 8, hipMemcpyDeviceToDevice, s));
        ntt_forward(ctx.tables(), limb_map(m_buf.get(), 0, Ql), Ql, 1, s);
        hipLaunchKernelGGL(k_add_limbs, grid_for(N, Ql), dim3(kThreads), 0, s, ct.data, (const u64 *)m_buf.get(),
                           ctx.dev_mod(), (u32)N);
    }
    PHX_HIP_CHECK(hipGetLastError());
    ct.is_ntt_form = !bfv;
    ct.is_asymmetric = true;
    ct.has_seed = false;
    ct.scale = 1.0;
    ct.correction_factor = 1;
    ct.noise_scale_deg = 1;
}

// BFV scale-and-round of the phase x (canonical residues [k][N], coefficient form)
// to the plaintext, one thread per coefficient: BEHZ (behz_decrypt_scale_and_round,
// src/rns.cu:1008-1090 with perform_final_multiplication 984-1006) or HPS
// (hps_decrypt_scale_and_round_kernel_small / _small_lazy / _large / _large_lazy,
// src/rns.cu:1519-1633). The HPS forms can return t itself (the reference's llround
// of a sum just below t), kept as the reference returns it.
__global__ void k_bfv_decrypt(const u64 *phase, u64 *out, BfvDecrypt D, const DevModulus *mod, u32 N, u32 k) {
#pragma clang fp contract(off)
    const u32 c = blockIdx.x * kThreads + threadIdx.x;
    if (c >= N) return;
    const u64 *tab = D.tab;
    const u64 t = D.t;
    if (D.kind == kBfvDecBEHZ) {
        // [t gamma]_q x, then FastBConv Q -> {t, gamma}
        u128 at = 0, ag = 0;
        for (u32 i = 0; i < k; ++i) {
            const u64 q = mod[i].value;
            const u64 *r = tab + 8 * i;
            const u64 x = mul_shoup(phase[(u64)i * N + c], r[0], r[1], q);
            const u64 y = mul_shoup(x, r[2], r[3], q);
            at += (u128)y * r[4];
            ag += (u128)y * r[5];
        }
        u64 xt = barrett128(at, t, D.t_ratio_lo, D.t_ratio_hi);
        u64 xg = barrett128(ag, D.gamma, D.g_ratio_lo, D.g_ratio_hi);
        xt = mul_shoup(xt, D.neg_qinv_t, D.neg_qinv_t_shoup, t);
        xg = mul_shoup(xg, D.neg_qinv_g, D.neg_qinv_g_shoup, D.gamma);
        // centred gamma part removed, then gamma^-1 mod t
        const u64 v = xg > (D.gamma >> 1) ? add_mod(xt, (D.gamma - xg) % t, t) : sub_mod(xt, xg % t, t);
        out[c] = mul_shoup(v, D.inv_gamma_t, D.inv_gamma_t_shoup, t);
        return;
    }
    const bool lazy = D.kind == kBfvDecHpsSmallLazy || D.kind == kBfvDecHpsLargeLazy;
    const bool large = D.kind == kBfvDecHpsLargeLazy || D.kind == kBfvDecHpsLarge;
    const u32 h = D.qmsb_half;
    const double tinv = 1.0 / (double)t;
    double fs = 0.0;
    u64 is = 0;
    for (u32 i = 0; i < k; ++i) {
        const u64 x = phase[(u64)i * N + c];
        const u64 *r = tab + 8 * i;
        if (!large) {
            fs = __builtin_fma((double)x, __longlong_as_double((long long)r[2]), fs);
            is += lazy ? x * r[0] : mul_shoup(x, r[0], r[1], t);
        } else {
            const u64 hi = x >> h, lo = x & ((1ull << h) - 1);
            fs = __builtin_fma((double)lo, __longlong_as_double((long long)r[2]), fs);
            fs = __builtin_fma((double)hi, __longlong_as_double((long long)r[5]), fs);
            if (lazy) {
                is += lo * r[0];
                is += hi * r[3];
            } else {
                is += mul_shoup(lo, r[0], r[1], t);
                is += mul_shoup(hi, r[3], r[4], t);
            }
        }
    }
    fs += (double)is;
    const u64 quot = (u64)(fs * tinv);
    fs -= (double)(t * quot);
    out[c] = (u64)__builtin_llround(fs);
}

void bfv_decrypt(const Context &ctx, SecretKey &sk, const Ciphertext &ct, u64 *out_dev, hipStream_t s) {
    require(ctx.params().scheme == kSchemeBFV, "bfv_decrypt: the context is not BFV");
    require(ct.chain_index >= 1 && ct.chain_index <= ctx.size_Q(), "invalid chain index");
    const BfvDecrypt &D = ctx.bfv_decrypt_consts(ct.chain_index);
    require(D.kind != kBfvDecNone, "plain_modulus is not set");
    require(!ct.is_ntt_form, "BFV encrypted must be in coefficient form");
    const u64 N = ctx.N();
    const u32 L = ct.coeff_modulus_size;
    const u64 LN = (u64)L * N;
    // phase c0 + c1 s (+ c2 s^2), via the NTT-form product
    Scratch cp_buf((size_t)ct.size * LN * 8, s), ph_buf(LN * 8, s);
    PHX_HIP_CHECK(hipMemcpyAsync(cp_buf.get(), ct.data, (size_t)ct.size * LN * 8, hipMemcpyDeviceToDevice, s));
    ntt_forward(ctx.tables(), limb_map(cp_buf.get(), 0, L, LN), L, ct.size, s);
    Ciphertext tmp;
    tmp.chain_index = ct.chain_index;
    tmp.size = ct.size;
    tmp.poly_degree = N;
    tmp.coeff_modulus_size = L;
    tmp.is_ntt_form = true;
    tmp.data = cp_buf.get();
    tmp.owns = false;
    decrypt(ctx, sk, tmp, ph_buf.get(), s);
    hipLaunchKernelGGL(k_bfv_decrypt, grid_for(N, 1), dim3(kThreads), 0, s, (const u64 *)ph_buf.get(), out_dev, D,
                       ctx.dev_mod(), (u32)N, L);
    PHX_HIP_CHECK(hipGetLastError());
}

void decrypt(const Context &ctx, SecretKey &sk, const Ciphertext &ct, u64 *out_dev, hipStream_t s) {
    const u64 N = ctx.N();
    const u32 L = ct.coeff_modulus_size;
    require(ct.is_ntt_form, "encrypted must be in NTT form");
    require(ct.size >= 1 && ct.size <= 3, "encrypted size must be 1..3");
    if (ct.size > 2 && !sk.s2) {
        const u32 QP = ctx.size_QP();
        sk.s2 = static_cast<u64 *>(dev_alloc((size_t)QP * N * 8, s));
        hipLaunchKernelGGL(k_mul_elementwise, grid_for(N, QP), dim3(kThreads), 0, s, (const u64 *)sk.s,
                           (const u64 *)sk.s, sk.s2, ctx.dev_mod(), (u32)N, 0u);
    }
    hipLaunchKernelGGL(k_dot_secret, grid_for(N, L), dim3(kThreads), 0, s, (const u64 *)ct.data, (const u64 *)sk.s,
                       (const u64 *)sk.s2, out_dev, ctx.dev_mod(), (u32)N, L, ct.size, (u64)ctx.size_QP() * N);
    PHX_HIP_CHECK(hipGetLastError());
    ntt_inverse(ctx.tables(), limb_map(out_dev, 0, L), L, 1, s);
}

}  // namespace phx
