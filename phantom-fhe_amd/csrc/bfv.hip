// bfv.hip -- BFV multiplication with BEHZ base extension and BFV
// relinearisation (coefficient-form ciphertexts).
//
// Pipeline of bfv_multiply_behz (src/evaluate.cu:440-545), per operand
// polynomial: NTT of the q part; lift q -> Bsk (fastbconv_m_tilde and the
// Montgomery step sm_mrq, one kernel: the m_tilde residue of a coefficient
// is the low 32 bits of its 64-bit sum); NTT of the Bsk part. Then the
// tensor in both bases, the inverse NTTs with the scale by t folded into the
// last stage, and one kernel for fast_floor and the Shenoy-Kumaresan
// conversion back to q. Every per-coefficient kernel keeps all of the
// coefficient's residues in registers (|q| <= 16, |Bsk| <= 18).
//
// HPS (mul_tech hps, src/evaluate.cu:648-801): extend Q -> Q u R with the
// FP64-corrected exact conversion, NTT over Q u R, tensor, INTT, then one
// kernel for the scale-and-round to R and the conversion back to Q.
//
// Relinearisation: BFV key switching equals the CKKS one conjugated by the
// NTT (modup of c2 = modup of NTT(c2) with the digit's own limbs kept;
// moddown = INTT of the NTT-form result), so it runs as NTT(c2) -> key
// switch -> INTT, added to c0 and c1.
#include "bfv.h"

#include <algorithm>
#include <cmath>

namespace phx {

namespace {
constexpr int kThreads = 256;
constexpr u32 kMaxQ = 16, kMaxBsk = kMaxQ + 2;
inline dim3 grid2(u64 N, u32 y) { return dim3((u32)((N + kThreads - 1) / kThreads), y); }
}  // namespace

// BEHZ steps (1)-(2): y_i = [x_i m_tilde Qhat_i^-1]_{q_i};
// y_p = sum_i y_i [Qhat_i]_p, y_mt = sum_i y_i [Qhat_i]_{2^32};
// r = -y_mt Q^-1 mod 2^32 in [-2^31, 2^31); out_p = (y_p + r Q) m_tilde^-1 mod p
// (src/rns.cu:1249-1333)
__global__ __launch_bounds__(kThreads) void k_behz_lift(const u64 *__restrict__ in, u64 *__restrict__ out,
                                                        BehzConsts B, const DevModulus *__restrict__ mod, u32 N,
                                                        u64 in_poly_stride, u64 out_poly_stride) {
    const u32 k = blockIdx.x * kThreads + threadIdx.x;
    if (k >= N) return;
    const u64 *x = in + blockIdx.y * in_poly_stride + k;
    u64 *o = out + blockIdx.y * out_poly_stride + k;
    u64 y[kMaxQ];
    u64 ymt = 0;
#pragma unroll
    for (u32 i = 0; i < kMaxQ; ++i)
        if (i < B.k) {
            y[i] = mul_mod(x[(u64)i * N], B.lift_f[i], mod[i]);
            ymt += y[i] * B.qhat_mt[i];
        }
    const u64 mask = 0xffffffffull;
    const u64 r = ((ymt & mask) * B.neg_qinv_mt) & mask;
    const bool neg = r >= (1ull << 31);
    for (u32 p = 0; p < B.nb; ++p) {
        const DevModulus m = mod[B.bsk_row0 + p];
        u128 acc = 0;
#pragma unroll
        for (u32 i = 0; i < kMaxQ; ++i)
            if (i < B.k) acc += (u128)y[i] * B.qhat_bsk[(u64)p * B.k + i];
        const u64 yp = barrett128(acc, m.value, m.ratio_lo, m.ratio_hi);
        const u64 rp = neg ? r + m.value - (1ull << 32) : r;
        const u64 v = barrett128((u128)rp * B.q_mod_bsk[p] + yp, m.value, m.ratio_lo, m.ratio_hi);
        o[(u64)p * N] = mul_mod(v, B.mtinv_bsk[p], m);
    }
}

// BEHZ steps (7)-(8): fast_floor (src/rns.cu:1379-1407) then fastbconv_sk
// (1455-1517) with the centred alpha_sk correction of
// multiply_and_negated_add_rns_poly (src/polymath.cu:606-634)
__global__ __launch_bounds__(kThreads) void k_behz_floor_sk(const u64 *__restrict__ zq, const u64 *__restrict__ zb,
                                                            u64 *__restrict__ out, BehzConsts B,
                                                            const DevModulus *__restrict__ mod, u32 N, u64 zq_stride,
                                                            u64 zb_stride, u64 out_stride) {
    const u32 k = blockIdx.x * kThreads + threadIdx.x;
    if (k >= N) return;
    const u64 *xq = zq + blockIdx.y * zq_stride + k;
    const u64 *xb = zb + blockIdx.y * zb_stride + k;
    u64 *o = out + blockIdx.y * out_stride + k;
    u64 y[kMaxQ];
#pragma unroll
    for (u32 i = 0; i < kMaxQ; ++i)
        if (i < B.k) y[i] = mul_mod(xq[(u64)i * N], B.qhatinv_q[i], mod[i]);
    // w_p = (z_p - FastBConv_q->p) Q^-1 mod p
    u64 w[kMaxBsk];
#pragma unroll
    for (u32 p = 0; p < kMaxBsk; ++p)
        if (p < B.nb) {
            const DevModulus m = mod[B.bsk_row0 + p];
            u128 acc = 0;
#pragma unroll
            for (u32 i = 0; i < kMaxQ; ++i)
                if (i < B.k) acc += (u128)y[i] * B.qhat_bsk[(u64)p * B.k + i];
            const u64 c = barrett128(acc, m.value, m.ratio_lo, m.ratio_hi);
            w[p] = mul_mod(add_mod(xb[(u64)p * N], c ? m.value - c : 0, m.value), B.qinv_bsk[p], m);
        }
    // Shenoy-Kumaresan from B (with m_sk) to q
    const u32 nB = B.nb - 1;
    u64 u[kMaxBsk];
    u128 am = 0;
    const DevModulus msk = mod[B.bsk_row0 + nB];
#pragma unroll
    for (u32 j = 0; j < kMaxBsk; ++j)
        if (j < nB) {
            u[j] = mul_mod(w[j], B.bhatinv_b[j], mod[B.bsk_row0 + j]);
            am += (u128)u[j] * B.bhat_msk[j];
        }
    u64 wm = 0;
#pragma unroll
    for (u32 p = 0; p < kMaxBsk; ++p)
        if (p == nB) wm = w[p];
    const u64 bc = barrett128(am, msk.value, msk.ratio_lo, msk.ratio_hi);
    const u64 alpha = mul_mod(add_mod(bc, wm ? msk.value - wm : 0, msk.value), B.binv_msk, msk);
    const bool neg = alpha > (msk.value >> 1);
    const u64 amag = neg ? msk.value - alpha : alpha;
#pragma unroll
    for (u32 i = 0; i < kMaxQ; ++i)
        if (i < B.k) {
            const DevModulus m = mod[i];
            u128 acc = 0;
#pragma unroll
            for (u32 j = 0; j < kMaxBsk; ++j)
                if (j < nB) acc += (u128)u[j] * B.bhat_q[(u64)i * nB + j];
            const u64 v = barrett128(acc, m.value, m.ratio_lo, m.ratio_hi);
            const u64 bq = B.b_mod_q[i];
            const u64 f = neg ? bq : (bq ? m.value - bq : 0);   // -alpha*B mod q, alpha centred
            const u64 corr = mul_mod(barrett64(amag, m.value, m.ratio_hi), f, m);
            o[(u64)i * N] = add_mod(v, corr, m.value);
        }
}

// HPS (mul_tech hps). bConv_HPS (src/rns_bconv.cu:277-372): y_i = [x_i Ihat_i^-1]_{i},
// out_p = sum_i y_i [Ihat_i]_p - v [I]_p with v = llround(sum_i y_i / i_i), the
// FP64 sum taken with fused multiply-adds in input order.
// hps_extend: Q part copied, R part = bConv_HPS(Q -> R), one coefficient.
__device__ __forceinline__ void hps_extend(const u64 *__restrict__ x, u64 *__restrict__ o, const HpsConsts &H,
                                           u32 N) {
    u64 y[kMaxQ];
    double frac = 0.0;
#pragma unroll
    for (u32 i = 0; i < kMaxQ; ++i)
        if (i < H.k) {
            const u64 xi = x[(u64)i * N];
            o[(u64)i * N] = xi;
            y[i] = mul_mod(xi, H.qhatinv_q[i], H.mod_qr[i]);
            frac = __builtin_fma((double)y[i], H.qinv[i], frac);
        }
    const u32 v = (u32)__builtin_llround(frac);
    for (u32 j = 0; j < H.nr; ++j) {
        const DevModulus m = H.mod_qr[H.k + j];
        u128 acc = 0;
#pragma unroll
        for (u32 i = 0; i < kMaxQ; ++i)
            if (i < H.k) acc += (u128)y[i] * H.qhat_r[(u64)j * H.k + i];
        const u64 a = barrett128(acc, m.value, m.ratio_lo, m.ratio_hi);
        o[(u64)(H.k + j) * N] = sub_mod(a, H.alphaq_r[(u64)v * H.nr + j], m.value);
    }
}

__global__ __launch_bounds__(kThreads) void k_hps_extend(const u64 *__restrict__ in, u64 *__restrict__ out,
                                                         HpsConsts H, u32 N, u64 in_stride, u64 out_stride) {
    const u32 k = blockIdx.x * kThreads + threadIdx.x;
    if (k >= N) return;
    hps_extend(in + blockIdx.y * in_stride + k, out + blockIdx.y * out_stride + k, H, N);
}

// hps_overq, second operand: R part = bConv_BEHZ_var1(S -> R) (src/rns_bconv.cu:231-245:
// y_i = [x_i negPQHatInvModq_i]_{s_i}, z_j = [sum_i y_i s_i^-1]_{r_j}) from the ks source
// moduli (Q at the top level, all of Q for hps_overq_leveled below it), then the Q part =
// bConv_HPS(R -> Q) of z (src/evaluate.cu:745-751). One coefficient.
__device__ __forceinline__ void hps_overq_switch(const u64 *__restrict__ x, u64 *__restrict__ o, const HpsConsts &H,
                                                 u32 N) {
    u64 y[kMaxQ];
#pragma unroll
    for (u32 i = 0; i < kMaxQ; ++i)
        if (i < H.ks) y[i] = mul_mod(x[(u64)i * N], H.negrq[i], H.mod_src[i]);
    u64 w[kMaxBsk];
    double frac = 0.0;
#pragma unroll
    for (u32 j = 0; j < kMaxBsk; ++j)
        if (j < H.nr) {
            const DevModulus m = H.mod_qr[H.k + j];
            u128 acc = 0;
#pragma unroll
            for (u32 i = 0; i < kMaxQ; ++i)
                if (i < H.ks) acc += (u128)y[i] * H.qinv_r[(u64)j * H.ks + i];
            const u64 z = barrett128(acc, m.value, m.ratio_lo, m.ratio_hi);
            o[(u64)(H.k + j) * N] = z;
            w[j] = mul_mod(z, H.rhatinv_r[j], m);
            frac = __builtin_fma((double)w[j], H.rinv[j], frac);
        }
    const u32 v = (u32)__builtin_llround(frac);
#pragma unroll
    for (u32 i = 0; i < kMaxQ; ++i)
        if (i < H.k) {
            const DevModulus m = H.mod_qr[i];
            u128 acc = 0;
#pragma unroll
            for (u32 j = 0; j < kMaxBsk; ++j)
                if (j < H.nr) acc += (u128)w[j] * H.rhat_q[(u64)i * H.nr + j];
            o[(u64)i * N] = sub_mod(barrett128(acc, m.value, m.ratio_lo, m.ratio_hi), H.alphar_q[(u64)v * H.k + i],
                                    m.value);
        }
}

// second operands of hps_overq, `per_item` polynomials per item; an item whose
// square flag is set (the reference's squaring path) is extended like a first operand
__global__ __launch_bounds__(kThreads) void k_hps_overq_switch(const u64 *__restrict__ in, u64 *__restrict__ out,
                                                               HpsConsts H, u32 N, u64 in_stride, u64 out_stride,
                                                               const u64 *__restrict__ square, u32 per_item) {
    const u32 k = blockIdx.x * kThreads + threadIdx.x;
    if (k >= N) return;
    const u64 *x = in + blockIdx.y * in_stride + k;
    u64 *o = out + blockIdx.y * out_stride + k;
    if (square && square[blockIdx.y / per_item])
        hps_extend(x, o, H, N);
    else
        hps_overq_switch(x, o, H, N);
}

// scaleAndRound_HPS_QlRl_Ql (src/rns.cu:1749-1796): nu = 0.5 + sum_j x_{r_j} frac_r_j
// (fused multiply-adds in order), alpha = (u64) nu,
// out_i = [sum_j x_{r_j} D_ij + x_{q_i} D_ik]_{q_i} + [alpha]_{q_i}
__global__ __launch_bounds__(kThreads) void k_hps_overq_round(const u64 *__restrict__ in, u64 *__restrict__ out,
                                                              HpsConsts H, u32 N, u64 in_stride, u64 out_stride) {
    const u32 k = blockIdx.x * kThreads + threadIdx.x;
    if (k >= N) return;
    const u64 *x = in + blockIdx.y * in_stride + k;
    u64 *o = out + blockIdx.y * out_stride + k;
    u64 xr[kMaxBsk];
    double nu = 0.5;
#pragma unroll
    for (u32 j = 0; j < kMaxBsk; ++j)
        if (j < H.nr) {
            xr[j] = x[(u64)(H.k + j) * N];
            nu = __builtin_fma((double)xr[j], H.frac_r[j], nu);
        }
    const u64 alpha = (u64)nu;
#pragma unroll
    for (u32 i = 0; i < kMaxQ; ++i)
        if (i < H.k) {
            const DevModulus m = H.mod_qr[i];
            const u64 *dm = H.divs_mod_q + (u64)i * (H.nr + 1);
            u128 acc = (u128)x[(u64)i * N] * dm[H.nr];
#pragma unroll
            for (u32 j = 0; j < kMaxBsk; ++j)
                if (j < H.nr) acc += (u128)xr[j] * dm[j];
            o[(u64)i * N] = add_mod(barrett128(acc, m.value, m.ratio_lo, m.ratio_hi),
                                    barrett64(alpha, m.value, m.ratio_hi), m.value);
        }
}

// scaleAndRound_HPS_QR_R (src/rns.cu:1700-1746) into registers, then
// bConv_HPS R -> Q into the output ciphertext
__global__ __launch_bounds__(kThreads) void k_hps_round_convert(const u64 *__restrict__ in, u64 *__restrict__ out,
                                                                HpsConsts H, u32 N, u64 in_stride, u64 out_stride) {
    const u32 k = blockIdx.x * kThreads + threadIdx.x;
    if (k >= N) return;
    const u64 *x = in + blockIdx.y * in_stride + k;
    u64 *o = out + blockIdx.y * out_stride + k;
    u64 xq[kMaxQ];
    double nu = 0.5;
#pragma unroll
    for (u32 i = 0; i < kMaxQ; ++i)
        if (i < H.k) {
            xq[i] = x[(u64)i * N];
            nu = __builtin_fma((double)xq[i], H.frac[i], nu);
        }
    const u64 alpha = (u64)nu;
    u64 y[kMaxBsk];
    double frac = 0.0;
#pragma unroll
    for (u32 j = 0; j < kMaxBsk; ++j)
        if (j < H.nr) {
            const DevModulus m = H.mod_qr[H.k + j];
            const u64 *dm = H.divs_mod_r + (u64)j * (H.k + 1);
            u128 acc = (u128)x[(u64)(H.k + j) * N] * dm[H.k];
#pragma unroll
            for (u32 i = 0; i < kMaxQ; ++i)
                if (i < H.k) acc += (u128)xq[i] * dm[i];
            const u64 r = add_mod(barrett128(acc, m.value, m.ratio_lo, m.ratio_hi),
                                  barrett64(alpha, m.value, m.ratio_hi), m.value);
            y[j] = mul_mod(r, H.rhatinv_r[j], m);
            frac = __builtin_fma((double)y[j], H.rinv[j], frac);
        }
    const u32 v = (u32)__builtin_llround(frac);
#pragma unroll
    for (u32 i = 0; i < kMaxQ; ++i)
        if (i < H.k) {
            const DevModulus m = H.mod_qr[i];
            u128 acc = 0;
#pragma unroll
            for (u32 j = 0; j < kMaxBsk; ++j)
                if (j < H.nr) acc += (u128)y[j] * H.rhat_q[(u64)i * H.nr + j];
            const u64 a = barrett128(acc, m.value, m.ratio_lo, m.ratio_hi);
            o[(u64)i * N] = sub_mod(a, H.alphar_q[(u64)v * H.k + i], m.value);
        }
}

__global__ __launch_bounds__(kThreads) void k_add_inplace(u64 *__restrict__ a, const u64 *__restrict__ b,
                                                          const DevModulus *__restrict__ mod, u32 N, u32 L) {
    const u32 k = blockIdx.x * kThreads + threadIdx.x;
    if (k >= N) return;
    const u32 limb = blockIdx.y % L;
    const u64 o = (u64)blockIdx.y * N + k;
    a[o] = add_mod(a[o], b[o], mod[limb].value);
}

// tensor of B items at fixed strides: out[z] = (a0 b0, a0 b1 + a1 b0, a1 b1)
// (tensor_prod_2x2_rns_poly, src/polymath.cu:463-498)
__global__ __launch_bounds__(kThreads) void k_tensor_items(const u64 *__restrict__ a, const u64 *__restrict__ b,
                                                           u64 *__restrict__ out, const DevModulus *__restrict__ mod,
                                                           u32 N, u32 L, u64 a_stride, u64 b_stride, u64 out_stride) {
    const u32 k = blockIdx.x * kThreads + threadIdx.x;
    if (k >= N) return;
    const u32 limb = blockIdx.y, item = blockIdx.z;
    const u64 LN = (u64)L * N, o = (u64)limb * N + k;
    const u64 *pa = a + item * a_stride + o, *pb = b + item * b_stride + o;
    const DevModulus m = mod[limb];
    const u64 q = m.value;
    const u64 a0 = pa[0], a1 = pa[LN], b0 = pb[0], b1 = pb[LN];
    const u64 r0 = mul_mod(a0, b0, m);
    const u64 r2 = mul_mod(a1, b1, m);
    u64 r1 = mul_mod(add_mod(a0, a1, q), add_mod(b0, b1, q), m);
    r1 = sub_mod(sub_mod(r1, r0, q), r2, q);
    u64 *po = out + item * out_stride + o;
    po[0] = r0;
    po[LN] = r1;
    po[2 * LN] = r2;
}

namespace {
const BehzConsts &behz_of(const Context &ctx) {
    const BehzConsts *B = ctx.behz();
    require(B != nullptr, "the context has no BEHZ tool: scheme must be BFV with mul_tech behz");
    require(B->k <= kMaxQ && B->nb <= kMaxBsk, "BEHZ multiply supports at most 16 data primes");
    return *B;
}

// NTT, lift and NTT of the two polynomials of ct into q and Bsk buffers
void behz_mul_1(const Context &ctx, const BehzConsts &B, const Ciphertext &ct, u64 *q_out, u64 *b_out,
                hipStream_t s) {
    const u64 N = ctx.N(), QN = (u64)B.k * N, BN = (u64)B.nb * N;
    PHX_HIP_CHECK(hipMemcpyAsync(q_out, ct.data, 2 * QN * 8, hipMemcpyDeviceToDevice, s));
    ntt_forward(ctx.tables(), limb_map(q_out, 0, B.k, QN), B.k, 2, s);
    hipLaunchKernelGGL(k_behz_lift, grid2(N, 2), dim3(kThreads), 0, s, (const u64 *)ct.data, b_out, B, ctx.dev_mod(),
                       (u32)N, QN, BN);
    PHX_HIP_CHECK(hipGetLastError());
    ntt_forward(ctx.tables(), limb_map(b_out, B.bsk_row0, B.nb, BN), B.nb, 2, s);
}
}  // namespace

void bfv_multiply(const Context &ctx, Ciphertext &ct1, const Ciphertext &ct2, hipStream_t s) {
    if (ctx.params().mul_tech == kMulHPS || ctx.params().mul_tech == kMulHPSOverQ ||
        ctx.params().mul_tech == kMulHPSOverQLeveled)
        return bfv_multiply_hps(ctx, ct1, ct2, s);
    const BehzConsts &B = behz_of(ctx);
    require(ct1.chain_index == ct2.chain_index, "encrypted1 and encrypted2 parameter mismatch");
    require(!ct1.is_ntt_form && !ct2.is_ntt_form, "encrypted1 or encrypted2 cannot be in NTT form");
    require(ct1.size == 2 && ct2.size == 2, "poly number mismatch");
    require(ct1.chain_index == 1, "BEHZ multiplication runs at the top data level");
    const u64 N = ctx.N(), QN = (u64)B.k * N, BN = (u64)B.nb * N;
    const bool square = &ct1 == &ct2 || ct1.data == ct2.data;
    Scratch q1_buf(3 * QN * 8, s);
    u64 *q1 = q1_buf.get();
    Scratch b1_buf(3 * BN * 8, s);
    u64 *b1 = b1_buf.get();
    Scratch q2_buf(square ? 0 : 2 * QN * 8, s), b2_buf(square ? 0 : 2 * BN * 8, s);
    u64 *q2 = square ? q1 : q2_buf.get();
    u64 *b2 = square ? b1 : b2_buf.get();
    behz_mul_1(ctx, B, ct1, q1, b1, s);
    if (!square) behz_mul_1(ctx, B, ct2, q2, b2, s);
    // (4) tensor in both bases (tensor_prod_2x2_rns_poly, src/polymath.cu:463-498)
    tensor_2x2(ctx, B.k, q1, q2, q1, q1 + QN, q1 + 2 * QN, s);
    tensor_2x2_mod(ctx.dev_mod() + B.bsk_row0, N, B.nb, b1, b2, b1, b1 + BN, b1 + 2 * BN, s);
    // (5)-(6) inverse NTTs, times t in the last stage
    LimbMap mq = limb_map(q1, 0, B.k, QN);
    mq.scale = B.tscale_q;
    ntt_inverse(ctx.tables(), mq, B.k, 3, s);
    LimbMap mb = limb_map(b1, B.bsk_row0, B.nb, BN);
    mb.scale = B.tscale_bsk;
    ntt_inverse(ctx.tables(), mb, B.nb, 3, s);
    // (7)-(8) into the resized ciphertext
    ct1.resize(ctx, ct1.chain_index, 3, s);
    hipLaunchKernelGGL(k_behz_floor_sk, grid2(N, 3), dim3(kThreads), 0, s, (const u64 *)q1, (const u64 *)b1, ct1.data,
                       B, ctx.dev_mod(), (u32)N, QN, BN, QN);
    PHX_HIP_CHECK(hipGetLastError());
}

// ---- hps_overq_leveled (src/evaluate.cu:549-646, 680-800, 845-1020;
// src/eval_key_switch.cu:110-180; src/rns.cu:1799-1846) ----

// FindLevelsToDrop (src/evaluate.cu:549-646): the noise estimate of a BFV
// ciphertext of multiplicative depth `depth` (sigma = 3.2f, assurance 36,
// ternary key bound 1, include/host/hestdparms.h:152-153) in natural logs, the
// spare bits above it, divided by the bit size of the largest data prime.
u32 hps_levels_to_drop(const Context &ctx, u64 depth, bool key_switch, bool asymmetric) {
    require(ctx.params().mul_tech == kMulHPSOverQLeveled, "FindLevelsToDrop is only used in HPS over Q Leveled");
    const double sigma = (double)3.2f, alpha = 36.0;
    const double p = (double)ctx.params().plain_modulus;
    const u32 k = ctx.size_P();
    const u32 numPartQ = ctx.level(1).beta;
    const double Bkey = 1.0;
    const u32 n = (u32)ctx.N();
    u64 qmax = 0;
    for (u32 i = 0; i < ctx.size_Q(); ++i) qmax = std::max(qmax, ctx.moduli()[i]);
    const double dcrtBits = (double)(64 - __builtin_clzll(qmax));
    const double Berr = sigma * std::sqrt(alpha);
    auto delta = [](u32 m) -> double { return 2. * std::sqrt(m); };
    auto Vnorm = [&](u32 m) -> double {
        if (asymmetric) return (1. + delta(m) * Bkey) / 2.;
        return Berr * (1. + 2. * delta(m) * Bkey);
    };
    auto noiseKS = [&](u32 m) -> double { return k * (numPartQ * delta(m) * Berr + delta(m) * Bkey + 1.0) / 2; };
    auto C1 = [&](u32 m) -> double { return delta(m) * delta(m) * p * Bkey; };
    auto C2 = [&](u32 m) -> double { return delta(m) * delta(m) * Bkey * Bkey / 2.0 + noiseKS(m); };
    auto logqBFV = [&](u32 m) -> double {
        if (depth > 0) return std::log(4 * p) + (double)(depth - 1) * std::log(C1(m)) +
                              std::log(C1(m) * Vnorm(m) + (double)depth * C2(m));
        return std::log(p * (4 * (Vnorm(m))));
    };
    double logqPrev = 6. * std::log(10);
    double logq = logqBFV(n);
    while (std::fabs(logq - logqPrev) > std::log(1.001)) {
        logqPrev = logq;
        logq = logqBFV(n);
    }
    const double loge = logq / std::log(2) - 2 - std::log2(p);
    const double logExtra = key_switch ? std::log2(noiseKS(n)) : std::log2(delta(n));
    int32_t levels = (int32_t)std::floor((loge - (double)(2 * depth) - 16 - logExtra) / dcrtBits);
    const int32_t sizeQ = (int32_t)ctx.size_Q();
    if (levels < 0)
        levels = 0;
    else if (levels > sizeQ - 1)
        levels = sizeQ - 1;
    return (u32)levels;
}

u32 bfv_keyswitch_drop(const Context &ctx, const Ciphertext &ct, bool key_switch) {
    if (ctx.params().scheme != kSchemeBFV || ctx.params().mul_tech != kMulHPSOverQLeveled) return 0;
    const u32 d = hps_levels_to_drop(ctx, (u64)ct.noise_scale_deg - 1, key_switch, ct.is_asymmetric);
    // the level drop scales all of Q down to Q_(size_Q - d) and expands back to Q
    // (src/eval_key_switch.cu:110-150 reads size_Q limbs of c2): it is defined for
    // ciphertexts at the first data level only
    require(d == 0 || ct.chain_index == 1,
            "hps_overq_leveled key switching with dropped levels needs a ciphertext at the first data level");
    return d;
}

// ExpandCRTBasis_Ql_Q (src/rns.cu:1808-1834): out_i = [x_i (Q/Ql)]_{q_i} for i < Ql and 0
// for the dropped limbs; with add (ExpandCRTBasis_Ql_Q_add_to_ct, 1836-1846)
// out_i += [x_i (Q/Ql)]_{q_i} for i < Ql, the dropped limbs unchanged
__global__ __launch_bounds__(kThreads) void k_hps_expand(const u64 *__restrict__ in, u64 *__restrict__ out,
                                                         const u64 *__restrict__ ex, const DevModulus *__restrict__ mod,
                                                         u32 N, u32 Ql, u32 k0, u64 in_stride, u64 out_stride,
                                                         int add) {
    const u32 k = blockIdx.x * kThreads + threadIdx.x;
    if (k >= N) return;
    const u32 limb = blockIdx.y % k0, poly = blockIdx.y / k0;
    u64 *o = out + poly * out_stride + (u64)limb * N + k;
    if (limb < Ql) {
        const u64 q = mod[limb].value;
        const u64 v = mul_shoup(in[poly * in_stride + (u64)limb * N + k], ex[2 * limb], ex[2 * limb + 1], q);
        *o = add ? add_mod(*o, v, q) : v;
    } else if (!add) {
        *o = 0;
    }
}

void hps_scale_q_ql(const Context &ctx, u32 d, const u64 *in, u64 in_stride, u64 *out, u64 out_stride, u32 npolys,
                    hipStream_t s) {
    const HpsConsts &D = ctx.hps_drop(d);
    hipLaunchKernelGGL(k_hps_overq_round, grid2(ctx.N(), npolys), dim3(kThreads), 0, s, in, out, D, (u32)ctx.N(),
                       in_stride, out_stride);
    PHX_HIP_CHECK(hipGetLastError());
}

void hps_expand_ql_q(const Context &ctx, u32 d, const u64 *in, u64 in_stride, u64 *out, u64 out_stride, u32 npolys,
                     bool add, hipStream_t s) {
    const u32 k0 = ctx.size_Q();
    hipLaunchKernelGGL(k_hps_expand, grid2(ctx.N(), npolys * k0), dim3(kThreads), 0, s, in, out, ctx.hps_expand(d),
                       ctx.dev_mod(), (u32)ctx.N(), k0 - d, k0, in_stride, out_stride, (int)add);
    PHX_HIP_CHECK(hipGetLastError());
}

namespace {
// HPS multiply with d dropped levels (d > 0 only for hps_overq_leveled): ct1 <- ct1 (x) ct2
// over Ql = size_Q - d and Rl; the result's polynomials sit at Ql limbs in Q-sized slots,
// the first `expand_polys` of them expanded back to Q
void hps_multiply_core(const Context &ctx, Ciphertext &ct1, const Ciphertext &ct2, u32 d, u32 expand_polys,
                       hipStream_t s) {
    const HpsConsts *Hp = ctx.params().mul_tech == kMulHPSOverQLeveled ? &ctx.hps_level(d) : ctx.hps();
    require(Hp != nullptr, "the context has no HPS tool: scheme must be BFV with mul_tech hps");
    const HpsConsts &H = *Hp;
    require(H.k <= kMaxQ && H.nr <= kMaxBsk, "HPS multiply supports at most 16 data primes");
    require(ct1.chain_index == ct2.chain_index, "encrypted1 and encrypted2 parameter mismatch");
    require(!ct1.is_ntt_form && !ct2.is_ntt_form, "encrypted1 or encrypted2 cannot be in NTT form");
    require(ct1.size == 2 && ct2.size == 2, "dest_size must be 3 when computing BFV multiplication using HPS");
    require(ct1.chain_index == 1, "HPS multiplication runs at the top data level");
    const u32 k0 = ctx.size_Q(), Ql = H.k;
    const u64 N = ctx.N(), QN0 = (u64)k0 * N, SN = (u64)(Ql + H.nr) * N;
    const bool square = &ct1 == &ct2 || ct1.data == ct2.data;
    Scratch a_buf(3 * SN * 8, s);
    u64 *a = a_buf.get();
    Scratch b_buf(square ? 0 : 2 * SN * 8, s);
    u64 *b = square ? a : b_buf.get();
    LimbMap m = limb_map(a, 0, Ql + H.nr, SN);
    m.split = Ql;
    m.special_row = H.r_row0;
    if (d) {
        // scaleAndRound_HPS_Q_Ql, then bConv_HPS Ql -> Rl (src/evaluate.cu:709-715)
        hps_scale_q_ql(ctx, d, ct1.data, QN0, a, SN, 2, s);
        hipLaunchKernelGGL(k_hps_extend, grid2(N, 2), dim3(kThreads), 0, s, (const u64 *)a, a, H, (u32)N, SN, SN);
    } else {
        hipLaunchKernelGGL(k_hps_extend, grid2(N, 2), dim3(kThreads), 0, s, (const u64 *)ct1.data, a, H, (u32)N, QN0,
                           SN);
    }
    PHX_HIP_CHECK(hipGetLastError());
    ntt_forward(ctx.tables(), m, Ql + H.nr, 2, s);
    if (!square) {
        if (H.overq)
            hipLaunchKernelGGL(k_hps_overq_switch, grid2(N, 2), dim3(kThreads), 0, s, (const u64 *)ct2.data, b, H,
                               (u32)N, QN0, SN, (const u64 *)nullptr, 2u);
        else
            hipLaunchKernelGGL(k_hps_extend, grid2(N, 2), dim3(kThreads), 0, s, (const u64 *)ct2.data, b, H, (u32)N,
                               QN0, SN);
        PHX_HIP_CHECK(hipGetLastError());
        LimbMap mb = m;
        mb.data = b;
        ntt_forward(ctx.tables(), mb, Ql + H.nr, 2, s);
    }
    tensor_2x2_mod(H.mod_qr, N, Ql + H.nr, a, b, a, a + SN, a + 2 * SN, s);
    ntt_inverse(ctx.tables(), m, Ql + H.nr, 3, s);
    ct1.resize(ctx, ct1.chain_index, 3, s);
    if (H.overq)
        hipLaunchKernelGGL(k_hps_overq_round, grid2(N, 3), dim3(kThreads), 0, s, (const u64 *)a, ct1.data, H,
                           (u32)N, SN, QN0);
    else
        hipLaunchKernelGGL(k_hps_round_convert, grid2(N, 3), dim3(kThreads), 0, s, (const u64 *)a, ct1.data, H,
                           (u32)N, SN, QN0);
    PHX_HIP_CHECK(hipGetLastError());
    if (d && expand_polys) hps_expand_ql_q(ctx, d, ct1.data, QN0, ct1.data, QN0, expand_polys, false, s);
}

// (c0, c1) of ct (Q limbs) += the BFV key switch of c2 (Ql = size_Q - d limbs,
// coefficient form) at chain 1 + d, expanded back to Q when d > 0
void bfv_keyswitch_add(const Context &ctx, Ciphertext &ct, const u64 *c2, u32 d, const KSwitchKey &key,
                       hipStream_t s) {
    const u32 k0 = ct.coeff_modulus_size, Ql = k0 - d;
    const u64 N = ctx.N(), LN = (u64)Ql * N;
    Scratch c2n_buf(LN * 8, s);
    u64 *c2n = c2n_buf.get();
    Scratch ks_buf(2 * LN * 8, s);
    u64 *ks = ks_buf.get();
    PHX_HIP_CHECK(hipMemcpyAsync(c2n, c2, LN * 8, hipMemcpyDeviceToDevice, s));
    ntt_forward(ctx.tables(), limb_map(c2n, 0, Ql), Ql, 1, s);
    keyswitch(ctx, ct.chain_index + d, c2n, key, ks, nullptr, nullptr, s);
    ntt_inverse(ctx.tables(), limb_map(ks, 0, Ql, LN), Ql, 2, s);
    if (d) {
        hps_expand_ql_q(ctx, d, ks, LN, ct.data, (u64)k0 * N, 2, true, s);
    } else {
        hipLaunchKernelGGL(k_add_inplace, grid2(N, 2 * k0), dim3(kThreads), 0, s, ct.data, (const u64 *)ks,
                           ctx.dev_mod(), (u32)N, k0);
        PHX_HIP_CHECK(hipGetLastError());
    }
}
}  // namespace

// bfv_multiply_hps (src/evaluate.cu:648-801): mul_tech hps, hps_overq and
// hps_overq_leveled (levels dropped by the noise estimate of the operands'
// degree, all three result polynomials expanded back to Q, degree + 1)
void bfv_multiply_hps(const Context &ctx, Ciphertext &ct1, const Ciphertext &ct2, hipStream_t s) {
    const bool leveled = ctx.params().mul_tech == kMulHPSOverQLeveled;
    const u32 deg = std::max(ct1.noise_scale_deg, ct2.noise_scale_deg);
    const u32 d = leveled ? hps_levels_to_drop(ctx, (u64)deg - 1, false, ct1.is_asymmetric) : 0;
    hps_multiply_core(ctx, ct1, ct2, d, 3, s);
    if (leveled) ct1.noise_scale_deg = deg + 1;
}

// bfv_mul_relin_hps (src/evaluate.cu:819-1027) for hps_overq_leveled: c0, c1 are
// expanded back to Q, c2 is key-switched at the dropped level and the result
// expanded and added (for hps and hps_overq it equals multiply then relinearize)
void bfv_mul_relin_leveled(const Context &ctx, Ciphertext &ct1, const Ciphertext &ct2, const KSwitchKey &rlk,
                           hipStream_t s) {
    require(ctx.params().mul_tech == kMulHPSOverQLeveled, "bfv_mul_relin_leveled needs mul_tech hps_overq_leveled");
    const u32 deg = std::max(ct1.noise_scale_deg, ct2.noise_scale_deg);
    const u32 d = hps_levels_to_drop(ctx, (u64)deg - 1, false, ct1.is_asymmetric);
    hps_multiply_core(ctx, ct1, ct2, d, 2, s);
    ct1.noise_scale_deg = deg + 1;
    bfv_keyswitch_add(ctx, ct1, ct1.data + 2 * (u64)ct1.coeff_modulus_size * ctx.N(), d, rlk, s);
    ct1.size = 2;
}

// ---- batched form: B items staged contiguously ----

size_t bfv_batch_work_words(const Context &ctx) {
    const u64 N = ctx.N();
    if (ctx.params().mul_tech == kMulHPSOverQLeveled)
        fail(kInvalidArgument, "hps_overq_leveled is not supported in batch plans: its level choice depends on each "
                               "ciphertext's noise degree");
    if (ctx.params().mul_tech == kMulHPS || ctx.params().mul_tech == kMulHPSOverQ) {
        const HpsConsts *H = ctx.hps();
        require(H != nullptr, "the context has no HPS tool: scheme must be BFV with mul_tech hps");
        return (size_t)7 * (H->k + H->nr) * N;   // extended operands 2 + 2, tensor 3
    }
    const BehzConsts &B = behz_of(ctx);
    return (size_t)N * (7ull * B.nb + 3ull * B.k);   // Bsk operands 2 + 2, tensor 3 (q and Bsk)
}

void bfv_multiply_batch(const Context &ctx, u64 *in1, u64 *in2, u32 items, const u64 *square, u64 *work, u64 *out,
                        hipStream_t s) {
    require(items >= 1, "empty batch");
    const u64 N = ctx.N();
    const u32 nz = items;
    if (ctx.params().mul_tech == kMulHPS || ctx.params().mul_tech == kMulHPSOverQ) {
        const HpsConsts *Hp = ctx.hps();
        require(Hp != nullptr, "the context has no HPS tool: scheme must be BFV with mul_tech hps");
        const HpsConsts &H = *Hp;
        require(H.k <= kMaxQ && H.nr <= kMaxBsk, "HPS multiply supports at most 16 data primes");
        const u32 S = H.k + H.nr;
        const u64 QN = (u64)H.k * N, SN = (u64)S * N;
        u64 *e1 = work, *e2 = work + 2ull * nz * SN, *t = work + 4ull * nz * SN;
        LimbMap m = limb_map(e1, 0, S, SN);
        m.split = H.k;
        m.special_row = H.r_row0;
        for (int o = 0; o < 2; ++o) {
            u64 *e = o ? e2 : e1;
            if (o && H.overq)
                hipLaunchKernelGGL(k_hps_overq_switch, grid2(N, 2 * nz), dim3(kThreads), 0, s, (const u64 *)in2, e,
                                   H, (u32)N, QN, SN, square, 2u);
            else
                hipLaunchKernelGGL(k_hps_extend, grid2(N, 2 * nz), dim3(kThreads), 0, s,
                                   (const u64 *)(o ? in2 : in1), e, H, (u32)N, QN, SN);
            PHX_HIP_CHECK(hipGetLastError());
            LimbMap me = m;
            me.data = e;
            ntt_forward(ctx.tables(), me, S, 2 * nz, s);
        }
        hipLaunchKernelGGL(k_tensor_items, dim3((u32)((N + kThreads - 1) / kThreads), S, nz), dim3(kThreads), 0, s,
                           (const u64 *)e1, (const u64 *)e2, t, H.mod_qr, (u32)N, S, 2 * SN, 2 * SN, 3 * SN);
        PHX_HIP_CHECK(hipGetLastError());
        LimbMap mt = m;
        mt.data = t;
        ntt_inverse(ctx.tables(), mt, S, 3 * nz, s);
        if (H.overq)
            hipLaunchKernelGGL(k_hps_overq_round, grid2(N, 3 * nz), dim3(kThreads), 0, s, (const u64 *)t, out, H,
                               (u32)N, SN, QN);
        else
            hipLaunchKernelGGL(k_hps_round_convert, grid2(N, 3 * nz), dim3(kThreads), 0, s, (const u64 *)t, out, H,
                               (u32)N, SN, QN);
        PHX_HIP_CHECK(hipGetLastError());
        return;
    }
    const BehzConsts &B = behz_of(ctx);
    const u64 QN = (u64)B.k * N, BN = (u64)B.nb * N;
    u64 *b1 = work, *b2 = work + 2ull * nz * BN, *tb = work + 4ull * nz * BN, *tq = work + 7ull * nz * BN;
    // (1)-(3) per operand: lift to Bsk, then NTT of both bases (the q part in place)
    for (int o = 0; o < 2; ++o) {
        u64 *x = o ? in2 : in1, *bb = o ? b2 : b1;
        hipLaunchKernelGGL(k_behz_lift, grid2(N, 2 * nz), dim3(kThreads), 0, s, (const u64 *)x, bb, B,
                           ctx.dev_mod(), (u32)N, QN, BN);
        PHX_HIP_CHECK(hipGetLastError());
        ntt_forward(ctx.tables(), limb_map(x, 0, B.k, QN), B.k, 2 * nz, s);
        ntt_forward(ctx.tables(), limb_map(bb, B.bsk_row0, B.nb, BN), B.nb, 2 * nz, s);
    }
    // (4) tensor in both bases
    const u32 gx = (u32)((N + kThreads - 1) / kThreads);
    hipLaunchKernelGGL(k_tensor_items, dim3(gx, B.k, nz), dim3(kThreads), 0, s, (const u64 *)in1, (const u64 *)in2, tq,
                       ctx.dev_mod(), (u32)N, B.k, 2 * QN, 2 * QN, 3 * QN);
    PHX_HIP_CHECK(hipGetLastError());
    hipLaunchKernelGGL(k_tensor_items, dim3(gx, B.nb, nz), dim3(kThreads), 0, s, (const u64 *)b1, (const u64 *)b2, tb,
                       ctx.dev_mod() + B.bsk_row0, (u32)N, B.nb, 2 * BN, 2 * BN, 3 * BN);
    PHX_HIP_CHECK(hipGetLastError());
    // (5)-(6) inverse NTTs with t folded in
    LimbMap mq = limb_map(tq, 0, B.k, QN);
    mq.scale = B.tscale_q;
    ntt_inverse(ctx.tables(), mq, B.k, 3 * nz, s);
    LimbMap mb = limb_map(tb, B.bsk_row0, B.nb, BN);
    mb.scale = B.tscale_bsk;
    ntt_inverse(ctx.tables(), mb, B.nb, 3 * nz, s);
    // (7)-(8)
    hipLaunchKernelGGL(k_behz_floor_sk, grid2(N, 3 * nz), dim3(kThreads), 0, s, (const u64 *)tq, (const u64 *)tb, out,
                       B, ctx.dev_mod(), (u32)N, QN, BN, QN);
    PHX_HIP_CHECK(hipGetLastError());
}

void bfv_relinearize(const Context &ctx, Ciphertext &ct, const KSwitchKey &rlk, hipStream_t s) {
    require(ct.size == 3, "destination_size must be 3");
    require(!ct.is_ntt_form, "BFV encrypted cannot be in NTT form");
    const u32 L = ct.coeff_modulus_size;
    const u64 N = ctx.N(), LN = (u64)L * N;
    // hps_overq_leveled: c2 scaled down to the dropped level first (src/eval_key_switch.cu:110-150)
    const u32 d = bfv_keyswitch_drop(ctx, ct, false);
    if (d) {
        Scratch c2_buf((u64)(L - d) * N * 8, s);
        u64 *c2 = c2_buf.get();
        hps_scale_q_ql(ctx, d, ct.data + 2 * LN, LN, c2, (u64)(L - d) * N, 1, s);
        bfv_keyswitch_add(ctx, ct, c2, d, rlk, s);
    } else {
        bfv_keyswitch_add(ctx, ct, ct.data + 2 * LN, 0, rlk, s);
    }
    ct.size = 2;
}

}  // namespace phx
