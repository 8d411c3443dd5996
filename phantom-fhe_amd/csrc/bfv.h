// bfv.h -- BFV multiplication (BEHZ) and relinearisation, coefficient form.
#pragma once

#include "evaluator.h"

namespace phx {

// bfv_multiply_behz (src/evaluate.cu:440-545): ct1 <- ct1 (x) ct2, size 3
void bfv_multiply(const Context &ctx, Ciphertext &ct1, const Ciphertext &ct2, hipStream_t s);
// bfv_multiply_hps (src/evaluate.cu:648-801), mul_tech hps, hps_overq and hps_overq_leveled
void bfv_multiply_hps(const Context &ctx, Ciphertext &ct1, const Ciphertext &ct2, hipStream_t s);
// Batched multiply (the context's technique) of `items` pairs staged contiguously:
// in1, in2 = [items][2][k][N] coefficient form (overwritten), out = [items][3][k][N];
// work holds items * bfv_batch_work_words(ctx) words; square (device, [items], or
// null) marks items whose operands are one ciphertext, which hps_overq extends like
// a first operand. Bit-identical per item to bfv_multiply / bfv_multiply_hps.
size_t bfv_batch_work_words(const Context &ctx);
void bfv_multiply_batch(const Context &ctx, u64 *in1, u64 *in2, u32 items, const u64 *square, u64 *work, u64 *out,
                        hipStream_t s);
// relinearize_inplace for BFV (src/evaluate.cu:1342-1374)
void bfv_relinearize(const Context &ctx, Ciphertext &ct, const KSwitchKey &rlk, hipStream_t s);

// ---- hps_overq_leveled ----
// FindLevelsToDrop (src/evaluate.cu:549-646)
u32 hps_levels_to_drop(const Context &ctx, u64 depth, bool key_switch, bool asymmetric);
// levels a key switch of ct drops: the estimate for its noise degree under
// hps_overq_leveled, 0 otherwise (src/eval_key_switch.cu:110-122)
u32 bfv_keyswitch_drop(const Context &ctx, const Ciphertext &ct, bool key_switch);
// scaleAndRound_HPS_Q_Ql (src/rns.cu:1799-1806): npolys polynomials [Q] -> [Ql], d >= 1
void hps_scale_q_ql(const Context &ctx, u32 d, const u64 *in, u64 in_stride, u64 *out, u64 out_stride, u32 npolys,
                    hipStream_t s);
// ExpandCRTBasis_Ql_Q (add = false) and ExpandCRTBasis_Ql_Q_add_to_ct (add = true),
// src/rns.cu:1808-1846: npolys polynomials [Ql] -> [Q], d >= 1
void hps_expand_ql_q(const Context &ctx, u32 d, const u64 *in, u64 in_stride, u64 *out, u64 out_stride, u32 npolys,
                     bool add, hipStream_t s);
// bfv_mul_relin_hps (src/evaluate.cu:819-1027) for hps_overq_leveled
void bfv_mul_relin_leveled(const Context &ctx, Ciphertext &ct1, const Ciphertext &ct2, const KSwitchKey &rlk,
                           hipStream_t s);

}  // namespace phx
