// ckks.h -- CKKS encoding and decoding on the host (canonical embedding).
//
// Slot j of a plaintext m(X) (N coefficients, n = N/2 slots) is the value
// m(zeta^(5^j mod 2N)) / scale with zeta = exp(i*pi/N): the reference's
// PhantomCKKSEncoder ordering (rotation group 5^j, src/ckks.cu:17-63), so that
// rotation by `step` (Galois element 5^step) moves slot j + step to slot j.
// Coefficient k < n carries the real part and k + n the imaginary part of the
// half-size complex transform (decompose_array, src/rns_base.cu:49-102).
#pragma once

#include <complex>
#include <vector>

#include "context.h"

namespace phx {

// PhantomCKKSEncoder::encode (src/ckks.cu:65-135): values (at most n) ->
// coefficient-domain residues [Ql][N] of round(scale * m) at chain_index.
void ckks_encode(const Context &ctx, const std::complex<double> *values, size_t count, double scale,
                 u32 chain_index, u64 *out);

// PhantomCKKSEncoder::decode (src/ckks.cu:347-400): coefficient-domain
// residues [Ql][N] -> n slot values, CRT-composed, centred and divided by scale.
void ckks_decode(const Context &ctx, const u64 *in, u32 chain_index, double scale, std::complex<double> *values);

}  // namespace phx
