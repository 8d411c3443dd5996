// serial.cpp -- the reference's ciphertext / key stream layouts (serial.h).
#include "serial.h"

#include <cstring>
#include <vector>

#include "ntt.h"
#include "prng.h"

namespace phx {

namespace {
struct Header {
    u64 chain_index, size, poly_degree, coeff_modulus_size;
    double scale;
    u64 correction_factor, noise_scale_deg;
    uint8_t is_ntt_form, is_asymmetric;
};

void put_header(uint8_t *p, const Header &h) {
    std::memcpy(p + 0, &h.chain_index, 8);
    std::memcpy(p + 8, &h.size, 8);
    std::memcpy(p + 16, &h.poly_degree, 8);
    std::memcpy(p + 24, &h.coeff_modulus_size, 8);
    std::memcpy(p + 32, &h.scale, 8);
    std::memcpy(p + 40, &h.correction_factor, 8);
    std::memcpy(p + 48, &h.noise_scale_deg, 8);
    p[56] = h.is_ntt_form;
    p[57] = h.is_asymmetric;
}

Header get_header(const uint8_t *p) {
    Header h{};
    std::memcpy(&h.chain_index, p + 0, 8);
    std::memcpy(&h.size, p + 8, 8);
    std::memcpy(&h.poly_degree, p + 16, 8);
    std::memcpy(&h.coeff_modulus_size, p + 24, 8);
    std::memcpy(&h.scale, p + 32, 8);
    std::memcpy(&h.correction_factor, p + 40, 8);
    std::memcpy(&h.noise_scale_deg, p + 48, 8);
    h.is_ntt_form = p[56];
    h.is_asymmetric = p[57];
    return h;
}
}  // namespace

size_t ciphertext_bytes(const Ciphertext &ct) {
    return kCtHeaderBytes + (size_t)ct.size * ct.coeff_modulus_size * ct.poly_degree * 8;
}

void ciphertext_save(const Ciphertext &ct, uint8_t *buf, size_t bytes) {
    require(ct.data != nullptr, "ciphertext is empty");
    require(bytes >= ciphertext_bytes(ct), "buffer too small for the ciphertext");
    put_header(buf, Header{ct.chain_index, ct.size, ct.poly_degree, ct.coeff_modulus_size, ct.scale,
                           ct.correction_factor, ct.noise_scale_deg, (uint8_t)ct.is_ntt_form,
                           (uint8_t)ct.is_asymmetric});
    PHX_HIP_CHECK(hipStreamSynchronize(hipStreamPerThread));
    PHX_HIP_CHECK(hipMemcpy(buf + kCtHeaderBytes, ct.data, ciphertext_bytes(ct) - kCtHeaderBytes,
                            hipMemcpyDeviceToHost));
}

size_t ciphertext_load(const Context &ctx, Ciphertext &ct, const uint8_t *buf, size_t bytes) {
    require(bytes >= kCtHeaderBytes, "stream too short for a ciphertext header");
    const Header h = get_header(buf);
    require(h.poly_degree == ctx.N(), "ciphertext poly_modulus_degree does not match the context");
    require(h.chain_index <= ctx.size_Q(), "ciphertext chain_index is invalid for the context");
    require(h.coeff_modulus_size == ctx.limbs_at((u32)h.chain_index),
            "ciphertext coeff_modulus_size does not match its chain_index");
    require(h.size >= 1 && h.size <= 3, "ciphertext size must be 1..3");
    const size_t n = (size_t)h.size * h.coeff_modulus_size * h.poly_degree;
    require(bytes >= kCtHeaderBytes + n * 8, "stream too short for the ciphertext data");
    ct.resize(ctx, (u32)h.chain_index, (u32)h.size, hipStreamPerThread);
    ct.scale = h.scale;
    ct.correction_factor = h.correction_factor;
    ct.noise_scale_deg = (u32)h.noise_scale_deg;
    ct.is_ntt_form = h.is_ntt_form != 0;
    ct.is_asymmetric = h.is_asymmetric != 0;
    PHX_HIP_CHECK(hipMemcpyAsync(ct.data, buf + kCtHeaderBytes, n * 8, hipMemcpyHostToDevice, hipStreamPerThread));
    PHX_HIP_CHECK(hipStreamSynchronize(hipStreamPerThread));
    return kCtHeaderBytes + n * 8;
}

size_t ciphertext_symmetric_bytes(const Ciphertext &ct) {
    return kCtHeaderBytes + (size_t)ct.coeff_modulus_size * ct.poly_degree * 8 + kPrngSeedBytes;
}

namespace {
// c1 of a seeded ciphertext: the sampler's output, inverse-transformed for coefficient form
u64 *regenerate_c1(const Context &ctx, const uint8_t *seed, u32 L, bool ntt_form, hipStream_t s) {
    u64 *c1 = static_cast<u64 *>(dev_alloc((size_t)L * ctx.N() * 8, s));
    sample_uniform_seeded(ctx, seed, L, c1, s);
    if (!ntt_form) ntt_inverse(ctx.tables(), limb_map(c1, 0, L), L, 1, s);
    return c1;
}
}  // namespace

void ciphertext_save_symmetric(const Context &ctx, const Ciphertext &ct, uint8_t *buf, size_t bytes) {
    require(ct.data != nullptr, "ciphertext is empty");
    if (ct.is_asymmetric) fail(kRuntimeError, "Asymmetric ciphertext does not have seed.");
    if (ct.size != 2) fail(kRuntimeError, "This method is only for 2-polynomial ciphertext.");
    require(ct.has_seed, "the ciphertext was not encrypted from a seed");
    if (ct.coeff_modulus_size != ctx.size_Q()) fail(kRuntimeError, "Only support ciphertext without modulus switching.");
    require(bytes >= ciphertext_symmetric_bytes(ct), "buffer too small for the ciphertext");
    const size_t n = (size_t)ct.coeff_modulus_size * ct.poly_degree;
    hipStream_t s = hipStreamPerThread;
    // the seed stands for c1 only while c1 is still the sampler's output
    u64 *c1 = regenerate_c1(ctx, ct.seed, ct.coeff_modulus_size, ct.is_ntt_form, s);
    std::vector<u64> want(n), have(n);
    PHX_HIP_CHECK(hipMemcpyAsync(want.data(), c1, n * 8, hipMemcpyDeviceToHost, s));
    PHX_HIP_CHECK(hipMemcpyAsync(have.data(), ct.data + n, n * 8, hipMemcpyDeviceToHost, s));
    PHX_HIP_CHECK(hipStreamSynchronize(s));
    dev_free(c1, s);
    require(want == have, "the ciphertext's c1 no longer matches its seed (it was modified after encryption)");
    put_header(buf, Header{ct.chain_index, ct.size, ct.poly_degree, ct.coeff_modulus_size, ct.scale,
                           ct.correction_factor, ct.noise_scale_deg, (uint8_t)ct.is_ntt_form,
                           (uint8_t)ct.is_asymmetric});
    PHX_HIP_CHECK(hipMemcpy(buf + kCtHeaderBytes, ct.data, n * 8, hipMemcpyDeviceToHost));
    std::memcpy(buf + kCtHeaderBytes + n * 8, ct.seed, kPrngSeedBytes);
}

size_t ciphertext_load_symmetric(const Context &ctx, Ciphertext &ct, const uint8_t *buf, size_t bytes) {
    require(bytes >= kCtHeaderBytes, "stream too short for a ciphertext header");
    const Header h = get_header(buf);
    if (h.is_asymmetric) fail(kRuntimeError, "Asymmetric ciphertext does not have seed.");
    if (h.size != 2) fail(kRuntimeError, "This method is only for 2-polynomial ciphertext.");
    require(h.poly_degree == ctx.N(), "ciphertext poly_modulus_degree does not match the context");
    if (h.coeff_modulus_size != ctx.size_Q()) fail(kRuntimeError, "Only support ciphertext without modulus switching.");
    require(h.chain_index >= 1 && h.chain_index <= ctx.size_Q() &&
                ctx.limbs_at((u32)h.chain_index) == h.coeff_modulus_size,
            "ciphertext chain_index does not match its coeff_modulus_size");
    const size_t n = (size_t)h.coeff_modulus_size * h.poly_degree;
    require(bytes >= kCtHeaderBytes + n * 8 + kPrngSeedBytes, "stream too short for the ciphertext data");
    hipStream_t s = hipStreamPerThread;
    ct.resize(ctx, (u32)h.chain_index, 2, s);
    ct.scale = h.scale;
    ct.correction_factor = h.correction_factor;
    ct.noise_scale_deg = (u32)h.noise_scale_deg;
    ct.is_ntt_form = h.is_ntt_form != 0;
    ct.is_asymmetric = false;
    ct.has_seed = true;
    std::memcpy(ct.seed, buf + kCtHeaderBytes + n * 8, kPrngSeedBytes);
    PHX_HIP_CHECK(hipMemcpyAsync(ct.data, buf + kCtHeaderBytes, n * 8, hipMemcpyHostToDevice, s));
    u64 *c1 = regenerate_c1(ctx, ct.seed, (u32)h.coeff_modulus_size, ct.is_ntt_form, s);
    PHX_HIP_CHECK(hipMemcpyAsync(ct.data + n, c1, n * 8, hipMemcpyDeviceToDevice, s));
    dev_free(c1, s);
    PHX_HIP_CHECK(hipStreamSynchronize(s));
    return kCtHeaderBytes + n * 8 + kPrngSeedBytes;
}

size_t kswitch_key_bytes(const Context &ctx, const KSwitchKey &key) {
    return 8 + (size_t)key.dnum * (kCtHeaderBytes + (size_t)2 * ctx.size_QP() * ctx.N() * 8);
}

void kswitch_key_save(const Context &ctx, const KSwitchKey &key, uint8_t *buf, size_t bytes) {
    require(key.data != nullptr, "key is empty");
    require(bytes >= kswitch_key_bytes(ctx, key), "buffer too small for the key");
    const u64 dnum = key.dnum, per = (u64)2 * ctx.size_QP() * ctx.N();
    std::memcpy(buf, &dnum, 8);
    uint8_t *p = buf + 8;
    PHX_HIP_CHECK(hipStreamSynchronize(hipStreamPerThread));
    for (u64 j = 0; j < dnum; ++j) {
        // each digit is a public key: chain 0 (the key level), size 2, all QP limbs
        put_header(p, Header{0, 2, ctx.N(), ctx.size_QP(), 1.0, 1, 1, 1, 0});
        PHX_HIP_CHECK(hipMemcpy(p + kCtHeaderBytes, key.data + j * per, per * 8, hipMemcpyDeviceToHost));
        p += kCtHeaderBytes + per * 8;
    }
}

size_t kswitch_key_load(const Context &ctx, KSwitchKey &key, const uint8_t *buf, size_t bytes) {
    require(bytes >= 8, "stream too short for a key");
    u64 dnum = 0;
    std::memcpy(&dnum, buf, 8);
    require(dnum >= 1 && dnum <= ctx.size_Q(), "key dnum is invalid for the context");
    const u64 per = (u64)2 * ctx.size_QP() * ctx.N();
    require(bytes >= 8 + dnum * (kCtHeaderBytes + per * 8), "stream too short for the key data");
    u64 *d = static_cast<u64 *>(dev_alloc(dnum * per * 8, hipStreamPerThread));
    const uint8_t *p = buf + 8;
    for (u64 j = 0; j < dnum; ++j) {
        const Header h = get_header(p);
        if (h.poly_degree != ctx.N() || h.coeff_modulus_size != ctx.size_QP() || h.size != 2) {
            dev_free(d, hipStreamPerThread);
            fail(kInvalidArgument, "key digit does not match the context's key level");
        }
        PHX_HIP_CHECK(hipMemcpyAsync(d + j * per, p + kCtHeaderBytes, per * 8, hipMemcpyHostToDevice,
                                     hipStreamPerThread));
        p += kCtHeaderBytes + per * 8;
    }
    PHX_HIP_CHECK(hipStreamSynchronize(hipStreamPerThread));
    if (key.data) dev_free(key.data, hipStreamPerThread);
    key.data = d;
    key.dnum = (u32)dnum;
    key.poly_degree = ctx.N();
    key.size_QP = ctx.size_QP();
    return (size_t)(p - buf);
}

// ---- secret key: [max_power, N, QP] then the powers [max_power][QP][N] ----

size_t secret_key_bytes(const Context &ctx, const SecretKey &sk) {
    const u64 powers = sk.s2 ? 2 : 1;
    return 24 + powers * ctx.size_QP() * ctx.N() * 8;
}

void secret_key_save(const Context &ctx, const SecretKey &sk, uint8_t *buf, size_t bytes) {
    if (!sk.s) fail(kInvalidArgument, "PhantomSecretKey has not been generated");
    require(bytes >= secret_key_bytes(ctx, sk), "buffer too small for the secret key");
    const u64 hdr[3] = {sk.s2 ? 2ull : 1ull, ctx.N(), ctx.size_QP()};
    std::memcpy(buf, hdr, 24);
    const size_t per = (size_t)ctx.size_QP() * ctx.N() * 8;
    PHX_HIP_CHECK(hipStreamSynchronize(hipStreamPerThread));
    PHX_HIP_CHECK(hipMemcpy(buf + 24, sk.s, per, hipMemcpyDeviceToHost));
    if (sk.s2) PHX_HIP_CHECK(hipMemcpy(buf + 24 + per, sk.s2, per, hipMemcpyDeviceToHost));
}

size_t secret_key_load(const Context &ctx, SecretKey &sk, const uint8_t *buf, size_t bytes) {
    require(bytes >= 24, "stream too short for a secret key header");
    u64 hdr[3];
    std::memcpy(hdr, buf, 24);
    require(hdr[1] == ctx.N(), "secret key poly_modulus_degree does not match the context");
    require(hdr[2] == ctx.size_QP(), "secret key coeff_modulus_size does not match the context's key level");
    require(hdr[0] >= 1 && hdr[0] <= 64, "secret key max power is invalid");
    const size_t per = (size_t)ctx.size_QP() * ctx.N() * 8;
    require((bytes - 24) / per >= hdr[0], "stream too short for the secret key data");
    hipStream_t s = hipStreamPerThread;
    u64 *d = static_cast<u64 *>(dev_alloc(per, s));
    u64 *d2 = hdr[0] >= 2 ? static_cast<u64 *>(dev_alloc(per, s)) : nullptr;
    PHX_HIP_CHECK(hipMemcpyAsync(d, buf + 24, per, hipMemcpyHostToDevice, s));
    if (d2) PHX_HIP_CHECK(hipMemcpyAsync(d2, buf + 24 + per, per, hipMemcpyHostToDevice, s));
    PHX_HIP_CHECK(hipStreamSynchronize(s));
    dev_free(sk.s, s);
    dev_free(sk.s2, s);
    sk.s = d;
    sk.s2 = d2;
    return 24 + hdr[0] * per;
}

}  // namespace phx
