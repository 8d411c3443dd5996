// bfv.h -- BFV multiplication (BEHZ) and relinearisation, coefficient form.
#pragma once

#include "evaluator.h"

namespace phx {

// bfv_multiply_behz (src/evaluate.cu:440-545): ct1 <- ct1 (x) ct2, size 3
void bfv_multiply(const Context &ctx, Ciphertext &ct1, const Ciphertext &ct2, hipStream_t s);
// bfv_multiply_hps (src/evaluate.cu:648-801), mul_tech hps and hps_overq
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

}  // namespace phx
