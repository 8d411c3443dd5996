// bfv.h -- BFV multiplication (BEHZ) and relinearisation, coefficient form.
#pragma once

#include "evaluator.h"

namespace phx {

// bfv_multiply_behz (src/evaluate.cu:440-545): ct1 <- ct1 (x) ct2, size 3
void bfv_multiply(const Context &ctx, Ciphertext &ct1, const Ciphertext &ct2, hipStream_t s);
// bfv_multiply_hps (src/evaluate.cu:648-801), mul_tech hps
void bfv_multiply_hps(const Context &ctx, Ciphertext &ct1, const Ciphertext &ct2, hipStream_t s);
// relinearize_inplace for BFV (src/evaluate.cu:1342-1374)
void bfv_relinearize(const Context &ctx, Ciphertext &ct, const KSwitchKey &rlk, hipStream_t s);

}  // namespace phx
