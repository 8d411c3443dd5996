// serial.h -- byte layouts of the reference's save/load streams.
//
// PhantomCiphertext::save (include/ciphertext.h:173-189): chain_index, size,
// poly_modulus_degree, coeff_modulus_size (8-byte unsigned each), scale
// (double), correction_factor (u64), noiseScaleDeg (8-byte unsigned),
// is_ntt_form, is_asymmetric (one byte each), then the data [size][L][N] u64;
// little-endian, no padding: 58 header bytes.
// PhantomCiphertext::save_symmetric / load_symmetric (include/ciphertext.h:216-305):
// the same header, c0 only ([L][N] u64), then the 64-byte seed from which load
// regenerates c1 (the seeded sampler, NTT form; inverse-transformed for a
// coefficient-form ciphertext); only 2-polynomial symmetric ciphertexts at the
// first data level.
// PhantomRelinKey::save (include/secretkey.h:129-147): dnum (8 bytes), then
// dnum public keys, each a ciphertext (chain 0, size 2, QP limbs).
// PhantomGaloisKey::save (include/secretkey.h:194-206): the number of keys
// (8 bytes), then that many relinearisation keys.
// PhantomSecretKey::save (include/secretkey.h:347-389): sk_max_power,
// poly_modulus_degree, coeff_modulus_size (QP), 8 bytes each, then the powers
// s, s^2, ... of the secret key, [max_power][QP][N] u64 in NTT form. This
// build keeps s and s^2 (decryption of 2- and 3-polynomial ciphertexts);
// a stream with higher powers loads its first two.
#pragma once

#include "evaluator.h"

namespace phx {

constexpr size_t kCtHeaderBytes = 58;

size_t ciphertext_bytes(const Ciphertext &ct);
void ciphertext_save(const Ciphertext &ct, uint8_t *buf, size_t bytes);
// consumes one ciphertext from buf; returns the bytes read
size_t ciphertext_load(const Context &ctx, Ciphertext &ct, const uint8_t *buf, size_t bytes);

size_t ciphertext_symmetric_bytes(const Ciphertext &ct);
// requires the ciphertext's c1 to still be its seed's sampler output
void ciphertext_save_symmetric(const Context &ctx, const Ciphertext &ct, uint8_t *buf, size_t bytes);
size_t ciphertext_load_symmetric(const Context &ctx, Ciphertext &ct, const uint8_t *buf, size_t bytes);

size_t kswitch_key_bytes(const Context &ctx, const KSwitchKey &key);
void kswitch_key_save(const Context &ctx, const KSwitchKey &key, uint8_t *buf, size_t bytes);
size_t kswitch_key_load(const Context &ctx, KSwitchKey &key, const uint8_t *buf, size_t bytes);

size_t secret_key_bytes(const Context &ctx, const SecretKey &sk);
void secret_key_save(const Context &ctx, const SecretKey &sk, uint8_t *buf, size_t bytes);
size_t secret_key_load(const Context &ctx, SecretKey &sk, const uint8_t *buf, size_t bytes);

}  // namespace phx
