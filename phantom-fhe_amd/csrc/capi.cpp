// capi.cpp -- the extern "C" boundary (include/phantom_amd.h).
#include "../../include/phantom_amd.h"

#include <cstdio>
#include <cstring>
#include <memory>
#include <string>

#include "batch.h"
#include "evaluator.h"
#include "autotune.h"
#include "prng.h"
#include "tune.h"
#include "ckks.h"
#include "serve.h"
#include "bfv.h"
#include "serial.h"

using namespace phx;

// Handles own their device memory: destroying a handle (or a creation that
// throws half-way) releases it.
struct phx_context {
    Context *ctx;
    ~phx_context() { delete ctx; }
};
struct phx_ciphertext {
    Ciphertext ct;
    ~phx_ciphertext() { ct.release(hipStreamPerThread); }
};
struct phx_secret_key {
    SecretKey sk;
    ~phx_secret_key() {
        dev_free(sk.s, hipStreamPerThread);
        dev_free(sk.s2, hipStreamPerThread);
    }
};
struct phx_public_key {
    Ciphertext pk;
    ~phx_public_key() { pk.release(hipStreamPerThread); }
};
struct phx_relin_key {
    KSwitchKey key;
    ~phx_relin_key() { dev_free(key.data, hipStreamPerThread); }
};
struct phx_galois_key {
    GaloisKeys gk;
    ~phx_galois_key() {
        for (auto &k : gk.keys) dev_free(k.data, hipStreamPerThread);
    }
};
struct phx_batch {
    BatchPlan *plan;
};
struct phx_server {
    Server *srv;
};

namespace {
thread_local std::string g_last_error;

template <class F>
int guard(F &&f) {
    try {
        f();
        return PHX_OK;
    } catch (const Error &e) {
        g_last_error = e.what();
        return e.code;
    } catch (const std::invalid_argument &e) {
        g_last_error = e.what();
        return PHX_EINVAL;
    } catch (const std::logic_error &e) {
        g_last_error = e.what();
        return PHX_ELOGIC;
    } catch (const std::exception &e) {
        g_last_error = e.what();
        return PHX_ERUNTIME;
    } catch (...) {
        g_last_error = "unknown error";
        return PHX_ERUNTIME;
    }
}

inline hipStream_t S(phx_stream s) { return s ? static_cast<hipStream_t>(s) : hipStreamPerThread; }
inline void nn(const void *p, const char *what) {
    if (!p) fail(kInvalidArgument, std::string(what) + " must not be null");
}
}  // namespace

namespace {
// tuning choice <-> the C array of PHX_TUNE_AXES entries {ntt_fused, ntt_fused_min_limbs,
// ntt_chunk_limbs, bconv_mfma, ntt_cols_threads, ntt_rows_threads, bconv_wgs_per_cu, bconv_fused}
TuneChoice choice_of(const int c[PHX_TUNE_AXES]) {
    TuneChoice t;
    t.ntt_fused = c[0];
    t.ntt_fused_min_limbs = c[1];
    t.ntt_chunk_limbs = c[2];
    t.bconv_mfma = c[3];
    t.ntt_cols_threads = c[4];
    t.ntt_rows_threads = c[5];
    t.bconv_wgs_per_cu = c[6];
    t.bconv_fused = c[7];
    return t;
}
void put_choice(const TuneChoice &t, int c[PHX_TUNE_AXES]) {
    c[0] = t.ntt_fused;
    c[1] = t.ntt_fused_min_limbs;
    c[2] = t.ntt_chunk_limbs;
    c[3] = t.bconv_mfma;
    c[4] = t.ntt_cols_threads;
    c[5] = t.ntt_rows_threads;
    c[6] = t.bconv_wgs_per_cu;
    c[7] = t.bconv_fused;
}
// A ciphertext handed to an evaluator entry point must belong to the context:
// its ring degree and limb count are the context's at its chain index (the
// kernels index it by the context's N and level sizes).
void check_ct(const Context &c, const Ciphertext &ct, const char *what) {
    require(ct.data != nullptr, std::string(what) + " is empty");
    require(ct.poly_degree == c.N(), std::string(what) + ": poly_modulus_degree does not match the context");
    require(ct.chain_index >= 1 && ct.chain_index <= c.size_Q() &&
                ct.coeff_modulus_size == c.limbs_at(ct.chain_index),
            std::string(what) + ": chain index / coeff_modulus_size do not match the context");
}
// a key-switching key of this context: its ring degree and QP limb count, and a
// digit for every digit the context's key switch reads
void check_key(const Context &c, const KSwitchKey &k, const char *what) {
    require(k.data != nullptr, std::string(what) + " is empty");
    require(k.poly_degree == c.N() && k.size_QP == c.size_QP() && k.dnum >= c.dnum(),
            std::string(what) + " does not match the context (ring degree, QP limbs or dnum)");
}
// A ciphertext built inside an entry point before it replaces the caller's: its
// device buffer is released if the call fails before the hand-over.
struct TempCt {
    Ciphertext ct;
    hipStream_t s;
    explicit TempCt(hipStream_t st) : s(st) {}
    TempCt(const TempCt &) = delete;
    TempCt &operator=(const TempCt &) = delete;
    ~TempCt() { ct.release(s); }
    // moves the ciphertext into dst, releasing dst's old buffer
    void hand_over(Ciphertext &dst) {
        dst.release(s);
        dst = ct;
        ct.data = nullptr;
        ct.owns = false;
        ct.capacity = 0;
    }
};
void check_keys(const Context &c, const GaloisKeys &g) {
    for (const KSwitchKey &k : g.keys) check_key(c, k, "Galois key");
}
}  // namespace

extern "C" {

const char *phx_last_error(void) { return g_last_error.c_str(); }

int phx_coeff_modulus_create(uint64_t N, const int *bits, size_t n, uint64_t *out) {
    return guard([&] {
        nn(bits, "bit_sizes");
        nn(out, "out");
        require(N >= 2048 && N <= 131072 && (N & (N - 1)) == 0, "poly_modulus_degree is invalid");
        auto v = hm::coeff_modulus_create(N, std::vector<int>(bits, bits + n));
        std::memcpy(out, v.data(), n * 8);
    });
}

int phx_galois_elt_from_step(int step, uint64_t N, uint32_t *out) {
    return guard([&] {
        nn(out, "out");
        *out = galois_elt_from_step(step, N);
    });
}

int phx_context_create(const phx_params *p, phx_context **out) {
    return guard([&] {
        nn(p, "params");
        nn(out, "out");
        nn(p->coeff_modulus, "coeff_modulus");
        Params q;
        q.scheme = p->scheme;
        q.mul_tech = p->mul_tech;
        q.poly_degree = p->poly_modulus_degree;
        q.coeff_modulus.assign(p->coeff_modulus, p->coeff_modulus + p->coeff_modulus_size);
        q.special_modulus_size = p->special_modulus_size;
        q.plain_modulus = p->plain_modulus;
        if (p->galois_elts) q.galois_elts.assign(p->galois_elts, p->galois_elts + p->galois_elts_size);
        require(q.scheme == PHX_SCHEME_CKKS || q.scheme == PHX_SCHEME_BFV || q.scheme == PHX_SCHEME_BGV,
                "unsupported scheme");
        auto h = std::make_unique<phx_context>();
        h->ctx = new Context(q);
        *out = h.release();
    });
}

int phx_context_destroy(phx_context *ctx) {
    return guard([&] {
        delete ctx;
    });
}

int phx_context_info(const phx_context *ctx, uint64_t *N, size_t *sQ, size_t *sP, size_t *dnum) {
    return guard([&] {
        nn(ctx, "context");
        if (N) *N = ctx->ctx->N();
        if (sQ) *sQ = ctx->ctx->size_Q();
        if (sP) *sP = ctx->ctx->size_P();
        if (dnum) *dnum = ctx->ctx->dnum();
    });
}

int phx_context_limbs_at(const phx_context *ctx, size_t c, size_t *limbs) {
    return guard([&] {
        nn(ctx, "context");
        nn(limbs, "limbs");
        require(c <= ctx->ctx->size_Q(), "invalid chain index");
        *limbs = ctx->ctx->limbs_at((u32)c);
    });
}

int phx_ciphertext_create(phx_ciphertext **out) {
    return guard([&] {
        nn(out, "out");
        *out = new phx_ciphertext{};
    });
}

int phx_ciphertext_destroy(phx_ciphertext *ct) {
    return guard([&] {
        delete ct;
    });
}

int phx_ciphertext_resize(const phx_context *ctx, phx_ciphertext *ct, size_t chain, size_t size, phx_stream s) {
    return guard([&] {
        nn(ctx, "context");
        nn(ct, "ciphertext");
        require(chain <= ctx->ctx->size_Q(), "invalid chain index");
        require(size >= 1 && size <= 3, "ciphertext size must be 1..3");
        // a new ciphertext takes its scheme's form: BFV coefficient, CKKS NTT
        if (!ct->ct.data) ct->ct.is_ntt_form = ctx->ctx->params().scheme != kSchemeBFV;
        ct->ct.resize(*ctx->ctx, (u32)chain, (u32)size, S(s));
    });
}

int phx_ciphertext_set_ntt_form(phx_ciphertext *ct, int is_ntt_form) {
    return guard([&] {
        nn(ct, "ciphertext");
        ct->ct.is_ntt_form = is_ntt_form != 0;
    });
}

int phx_ciphertext_upload(phx_ciphertext *ct, const uint64_t *host, size_t n, phx_stream s) {
    return guard([&] {
        nn(ct, "ciphertext");
        nn(host, "host");
        require(n == (size_t)ct->ct.size * ct->ct.coeff_modulus_size * ct->ct.poly_degree, "size mismatch");
        PHX_HIP_CHECK(hipMemcpyAsync(ct->ct.data, host, n * 8, hipMemcpyHostToDevice, S(s)));
        PHX_HIP_CHECK(hipStreamSynchronize(S(s)));
    });
}

int phx_ciphertext_download(const phx_ciphertext *ct, uint64_t *host, size_t n, phx_stream s) {
    return guard([&] {
        nn(ct, "ciphertext");
        nn(host, "host");
        require(n == (size_t)ct->ct.size * ct->ct.coeff_modulus_size * ct->ct.poly_degree, "size mismatch");
        PHX_HIP_CHECK(hipMemcpyAsync(host, ct->ct.data, n * 8, hipMemcpyDeviceToHost, S(s)));
        PHX_HIP_CHECK(hipStreamSynchronize(S(s)));
    });
}

int phx_ciphertext_copy(const phx_context *ctx, phx_ciphertext *dst, const phx_ciphertext *src, phx_stream s) {
    return guard([&] {
        nn(ctx, "context");
        nn(dst, "dst");
        nn(src, "src");
        if (dst == src) return;
        const Ciphertext &a = src->ct;
        dst->ct.resize(*ctx->ctx, a.chain_index, a.size, S(s));
        PHX_HIP_CHECK(hipMemcpyAsync(dst->ct.data, a.data, (size_t)a.size * a.coeff_modulus_size * a.poly_degree * 8,
                                     hipMemcpyDeviceToDevice, S(s)));
        dst->ct.scale = a.scale;
        dst->ct.is_ntt_form = a.is_ntt_form;
        dst->ct.correction_factor = a.correction_factor;
        dst->ct.noise_scale_deg = a.noise_scale_deg;
        dst->ct.is_asymmetric = a.is_asymmetric;
        dst->ct.has_seed = a.has_seed;
        std::memcpy(dst->ct.seed, a.seed, sizeof(a.seed));
    });
}

int phx_ciphertext_clone(const phx_ciphertext *src, phx_ciphertext **out, phx_stream s) {
    return guard([&] {
        nn(src, "src");
        nn(out, "out");
        std::unique_ptr<phx_ciphertext> c(new phx_ciphertext{});
        const Ciphertext &a = src->ct;
        c->ct = a;
        c->ct.data = nullptr;
        c->ct.capacity = 0;
        c->ct.owns = true;
        const size_t words = (size_t)a.size * a.coeff_modulus_size * a.poly_degree;
        if (a.data && words) {
            c->ct.data = static_cast<u64 *>(dev_alloc(words * 8, S(s)));
            c->ct.capacity = words;
            PHX_HIP_CHECK(hipMemcpyAsync(c->ct.data, a.data, words * 8, hipMemcpyDeviceToDevice, S(s)));
        }
        *out = c.release();
    });
}

int phx_ciphertext_info(const phx_ciphertext *ct, size_t *chain, size_t *size, size_t *L, double *scale, int *ntt) {
    return guard([&] {
        nn(ct, "ciphertext");
        if (chain) *chain = ct->ct.chain_index;
        if (size) *size = ct->ct.size;
        if (L) *L = ct->ct.coeff_modulus_size;
        if (scale) *scale = ct->ct.scale;
        if (ntt) *ntt = ct->ct.is_ntt_form ? 1 : 0;
    });
}

int phx_ciphertext_set_scale(phx_ciphertext *ct, double scale) {
    return guard([&] {
        nn(ct, "ciphertext");
        ct->ct.scale = scale;
    });
}

int phx_ciphertext_data(const phx_ciphertext *ct, uint64_t **p) {
    return guard([&] {
        nn(ct, "ciphertext");
        nn(p, "out");
        *p = ct->ct.data;
    });
}

int phx_secret_key_create(const phx_context *ctx, uint64_t seed, phx_secret_key **out) {
    return guard([&] {
        nn(ctx, "context");
        nn(out, "out");
        auto k = std::make_unique<phx_secret_key>();
        secret_key_generate(*ctx->ctx, k->sk, seed, hipStreamPerThread);
        PHX_HIP_CHECK(hipStreamSynchronize(hipStreamPerThread));
        *out = k.release();
    });
}

int phx_secret_key_create_from_coeffs(const phx_context *ctx, const int8_t *coeffs, phx_secret_key **out) {
    return guard([&] {
        nn(ctx, "context");
        nn(coeffs, "coeffs");
        nn(out, "out");
        auto k = std::make_unique<phx_secret_key>();
        secret_key_upload(*ctx->ctx, k->sk, coeffs, hipStreamPerThread);
        PHX_HIP_CHECK(hipStreamSynchronize(hipStreamPerThread));
        *out = k.release();
    });
}

int phx_secret_key_download(const phx_context *ctx, const phx_secret_key *sk, uint64_t *host) {
    return guard([&] {
        nn(ctx, "context");
        nn(sk, "secret key");
        nn(host, "host");
        PHX_HIP_CHECK(hipStreamSynchronize(hipStreamPerThread));
        PHX_HIP_CHECK(hipMemcpy(host, sk->sk.s, (size_t)ctx->ctx->size_QP() * ctx->ctx->N() * 8, hipMemcpyDeviceToHost));
    });
}

int phx_secret_key_destroy(phx_secret_key *sk) {
    return guard([&] {
        delete sk;
    });
}

int phx_relin_key_create(const phx_context *ctx, phx_secret_key *sk, uint64_t seed, phx_relin_key **out) {
    return guard([&] {
        nn(ctx, "context");
        nn(sk, "secret key");
        nn(out, "out");
        auto k = std::make_unique<phx_relin_key>();
        relin_key_generate(*ctx->ctx, sk->sk, k->key, seed, hipStreamPerThread);
        PHX_HIP_CHECK(hipStreamSynchronize(hipStreamPerThread));
        *out = k.release();
    });
}

int phx_relin_key_create_from_host(const phx_context *ctx, const uint64_t *host, size_t dnum, phx_relin_key **out) {
    return guard([&] {
        nn(ctx, "context");
        nn(host, "host");
        nn(out, "out");
        require(dnum >= 1, "dnum must be positive");
        const size_t n = dnum * 2 * (size_t)ctx->ctx->size_QP() * ctx->ctx->N();
        auto k = std::make_unique<phx_relin_key>();
        k->key.dnum = (u32)dnum;
        k->key.poly_degree = ctx->ctx->N();
        k->key.size_QP = ctx->ctx->size_QP();
        k->key.data = static_cast<u64 *>(dev_alloc(n * 8, hipStreamPerThread));
        PHX_HIP_CHECK(hipMemcpyAsync(k->key.data, host, n * 8, hipMemcpyHostToDevice, hipStreamPerThread));
        PHX_HIP_CHECK(hipStreamSynchronize(hipStreamPerThread));
        *out = k.release();
    });
}

int phx_relin_key_download(const phx_context *ctx, const phx_relin_key *key, uint64_t *host) {
    return guard([&] {
        nn(ctx, "context");
        nn(key, "key");
        nn(host, "host");
        const size_t n = (size_t)key->key.dnum * 2 * ctx->ctx->size_QP() * ctx->ctx->N();
        PHX_HIP_CHECK(hipStreamSynchronize(hipStreamPerThread));
        PHX_HIP_CHECK(hipMemcpy(host, key->key.data, n * 8, hipMemcpyDeviceToHost));
    });
}

int phx_relin_key_destroy(phx_relin_key *key) {
    return guard([&] {
        delete key;
    });
}

int phx_galois_key_create(const phx_context *ctx, const phx_secret_key *sk, const uint32_t *elts, size_t n,
                          uint64_t seed, phx_galois_key **out) {
    return guard([&] {
        nn(ctx, "context");
        nn(sk, "secret key");
        nn(elts, "elts");
        nn(out, "out");
        auto k = std::make_unique<phx_galois_key>();
        for (size_t i = 0; i < n; ++i) {
            k->gk.elts.push_back(elts[i]);
            k->gk.keys.emplace_back();   // owned by the handle before its memory is allocated
            galois_key_generate(*ctx->ctx, sk->sk, elts[i], k->gk.keys.back(), seed + 0x1000 * (i + 1),
                                hipStreamPerThread);
        }
        PHX_HIP_CHECK(hipStreamSynchronize(hipStreamPerThread));
        *out = k.release();
    });
}

int phx_galois_key_create_from_host(const phx_context *ctx, const uint32_t *elts, size_t n, const uint64_t *host,
                                    size_t dnum, phx_galois_key **out) {
    return guard([&] {
        nn(ctx, "context");
        nn(elts, "elts");
        nn(host, "host");
        nn(out, "out");
        const size_t per = dnum * 2 * (size_t)ctx->ctx->size_QP() * ctx->ctx->N();
        auto k = std::make_unique<phx_galois_key>();
        for (size_t i = 0; i < n; ++i) {
            k->gk.elts.push_back(elts[i]);
            k->gk.keys.emplace_back();
            KSwitchKey &key = k->gk.keys.back();
            key.dnum = (u32)dnum;
            key.poly_degree = ctx->ctx->N();
            key.size_QP = ctx->ctx->size_QP();
            key.data = static_cast<u64 *>(dev_alloc(per * 8, hipStreamPerThread));
            PHX_HIP_CHECK(hipMemcpyAsync(key.data, host + i * per, per * 8, hipMemcpyHostToDevice, hipStreamPerThread));
        }
        PHX_HIP_CHECK(hipStreamSynchronize(hipStreamPerThread));
        *out = k.release();
    });
}

int phx_galois_key_destroy(phx_galois_key *key) {
    return guard([&] {
        delete key;
    });
}

int phx_encrypt_symmetric(const phx_context *ctx, const phx_secret_key *sk, const uint64_t *plain, size_t chain,
                          uint64_t seed, phx_ciphertext *out, phx_stream s) {
    return guard([&] {
        nn(ctx, "context");
        nn(sk, "secret key");
        nn(plain, "plain");
        nn(out, "out");
        const Context &c = *ctx->ctx;
        require(chain >= 1 && chain <= c.size_Q(), "invalid chain index");
        const size_t n = (size_t)c.limbs_at((u32)chain) * c.N();
        Scratch d_buf(n * 8, S(s));
        u64 *d = d_buf.get();
        PHX_HIP_CHECK(hipMemcpyAsync(d, plain, n * 8, hipMemcpyHostToDevice, S(s)));
        encrypt_symmetric(c, sk->sk, d, (u32)chain, out->ct, seed, S(s));
        PHX_HIP_CHECK(hipStreamSynchronize(S(s)));
    });
}

int phx_encrypt_symmetric_seeded(const phx_context *ctx, const phx_secret_key *sk, const uint64_t *plain,
                                 size_t chain, const uint8_t *seed_a, uint64_t seed_e, phx_ciphertext *out,
                                 phx_stream s) {
    return guard([&] {
        nn(ctx, "context");
        nn(sk, "secret key");
        nn(plain, "plain");
        nn(seed_a, "seed");
        nn(out, "out");
        const Context &c = *ctx->ctx;
        require(chain >= 1 && chain <= c.size_Q(), "invalid chain index");
        const size_t n = (size_t)c.limbs_at((u32)chain) * c.N();
        Scratch d_buf(n * 8, S(s));
        u64 *d = d_buf.get();
        PHX_HIP_CHECK(hipMemcpyAsync(d, plain, n * 8, hipMemcpyHostToDevice, S(s)));
        encrypt_symmetric_seeded(c, sk->sk, d, (u32)chain, out->ct, seed_a, seed_e, S(s));
        PHX_HIP_CHECK(hipStreamSynchronize(S(s)));
    });
}

int phx_decrypt(const phx_context *ctx, phx_secret_key *sk, const phx_ciphertext *ct, uint64_t *out, phx_stream s) {
    return guard([&] {
        nn(ctx, "context");
        nn(sk, "secret key");
        nn(ct, "ciphertext");
        nn(out, "out");
        check_ct(*ctx->ctx, ct->ct, "ciphertext");
        const Context &c = *ctx->ctx;
        const size_t n = (size_t)ct->ct.coeff_modulus_size * c.N();
        Scratch d_buf(n * 8, S(s));
        u64 *d = d_buf.get();
        decrypt(c, sk->sk, ct->ct, d, S(s));
        PHX_HIP_CHECK(hipMemcpyAsync(out, d, n * 8, hipMemcpyDeviceToHost, S(s)));
        PHX_HIP_CHECK(hipStreamSynchronize(S(s)));
    });
}

int phx_decrypt_bfv(const phx_context *ctx, phx_secret_key *sk, const phx_ciphertext *ct, uint64_t *plain,
                    phx_stream s) {
    return guard([&] {
        nn(ctx, "context");
        nn(sk, "secret key");
        nn(ct, "ciphertext");
        nn(plain, "plain");
        check_ct(*ctx->ctx, ct->ct, "ciphertext");
        const size_t N = ctx->ctx->N();
        Scratch d_buf(N * 8, S(s));
        bfv_decrypt(*ctx->ctx, sk->sk, ct->ct, d_buf.get(), S(s));
        PHX_HIP_CHECK(hipMemcpyAsync(plain, d_buf.get(), N * 8, hipMemcpyDeviceToHost, S(s)));
        PHX_HIP_CHECK(hipStreamSynchronize(S(s)));
    });
}

int phx_ciphertext_save_size(const phx_ciphertext *ct, size_t *bytes) {
    return guard([&] {
        nn(ct, "ciphertext");
        nn(bytes, "bytes");
        *bytes = ciphertext_bytes(ct->ct);
    });
}

int phx_ciphertext_save(const phx_ciphertext *ct, void *buf, size_t bytes) {
    return guard([&] {
        nn(ct, "ciphertext");
        nn(buf, "buffer");
        ciphertext_save(ct->ct, static_cast<uint8_t *>(buf), bytes);
    });
}

int phx_ciphertext_load(const phx_context *ctx, phx_ciphertext *ct, const void *buf, size_t bytes) {
    return guard([&] {
        nn(ctx, "context");
        nn(ct, "ciphertext");
        nn(buf, "buffer");
        ciphertext_load(*ctx->ctx, ct->ct, static_cast<const uint8_t *>(buf), bytes);
    });
}

int phx_ciphertext_save_symmetric_size(const phx_ciphertext *ct, size_t *bytes) {
    return guard([&] {
        nn(ct, "ciphertext");
        nn(bytes, "bytes");
        *bytes = ciphertext_symmetric_bytes(ct->ct);
    });
}

int phx_ciphertext_save_symmetric(const phx_context *ctx, const phx_ciphertext *ct, void *buf, size_t bytes) {
    return guard([&] {
        nn(ctx, "context");
        nn(ct, "ciphertext");
        nn(buf, "buffer");
        ciphertext_save_symmetric(*ctx->ctx, ct->ct, static_cast<uint8_t *>(buf), bytes);
    });
}

int phx_ciphertext_load_symmetric(const phx_context *ctx, phx_ciphertext *ct, const void *buf, size_t bytes) {
    return guard([&] {
        nn(ctx, "context");
        nn(ct, "ciphertext");
        nn(buf, "buffer");
        ciphertext_load_symmetric(*ctx->ctx, ct->ct, static_cast<const uint8_t *>(buf), bytes);
    });
}

int phx_sample_uniform_poly(const phx_context *ctx, const uint8_t *seed, size_t coeff_mod_size, uint64_t *out,
                            phx_stream s) {
    return guard([&] {
        nn(ctx, "context");
        nn(seed, "seed");
        nn(out, "out");
        sample_uniform_seeded(*ctx->ctx, seed, (u32)coeff_mod_size, out, S(s));
    });
}

int phx_public_key_create(const phx_context *ctx, const phx_secret_key *sk, const uint8_t *seed_a, uint64_t seed_e,
                          phx_public_key **out) {
    return guard([&] {
        nn(ctx, "context");
        nn(sk, "secret key");
        nn(out, "out");
        auto k = std::make_unique<phx_public_key>();
        public_key_generate(*ctx->ctx, sk->sk, seed_a, seed_e, k->pk, hipStreamPerThread);
        PHX_HIP_CHECK(hipStreamSynchronize(hipStreamPerThread));
        *out = k.release();
    });
}

int phx_public_key_destroy(phx_public_key *pk) {
    return guard([&] { delete pk; });
}

int phx_public_key_download(const phx_context *ctx, const phx_public_key *pk, uint64_t *host_2_qp_n) {
    return guard([&] {
        nn(ctx, "context");
        nn(pk, "public key");
        nn(host_2_qp_n, "out");
        PHX_HIP_CHECK(hipStreamSynchronize(hipStreamPerThread));
        PHX_HIP_CHECK(hipMemcpy(host_2_qp_n, pk->pk.data, (size_t)2 * ctx->ctx->size_QP() * ctx->ctx->N() * 8,
                                hipMemcpyDeviceToHost));
    });
}

int phx_public_key_save_size(const phx_public_key *pk, size_t *bytes) {
    return guard([&] {
        nn(pk, "public key");
        nn(bytes, "bytes");
        *bytes = ciphertext_bytes(pk->pk);
    });
}

int phx_public_key_save(const phx_public_key *pk, void *buf, size_t bytes) {
    return guard([&] {
        nn(pk, "public key");
        nn(buf, "buffer");
        if (!pk->pk.data) fail(kInvalidArgument, "PhantomPublicKey has not been generated");
        ciphertext_save(pk->pk, static_cast<uint8_t *>(buf), bytes);
    });
}

int phx_public_key_load(const phx_context *ctx, const void *buf, size_t bytes, phx_public_key **out) {
    return guard([&] {
        nn(ctx, "context");
        nn(buf, "buffer");
        nn(out, "out");
        auto k = std::make_unique<phx_public_key>();
        ciphertext_load(*ctx->ctx, k->pk, static_cast<const uint8_t *>(buf), bytes);
        require(k->pk.chain_index == 0 && k->pk.size == 2 && k->pk.is_ntt_form,
                "the stream does not hold a public key (a 2-polynomial NTT-form ciphertext at the key level)");
        *out = k.release();
    });
}

int phx_encrypt_asymmetric(const phx_context *ctx, const phx_public_key *pk, const uint64_t *plain, size_t chain,
                           uint64_t seed, phx_ciphertext *out, phx_stream s) {
    return phx_encrypt_asymmetric_with_samples(ctx, pk, plain, chain, nullptr, nullptr, nullptr, seed, out, s);
}

int phx_encrypt_asymmetric_with_samples(const phx_context *ctx, const phx_public_key *pk, const uint64_t *plain,
                                        size_t chain, const int8_t *u, const int8_t *e0, const int8_t *e1,
                                        uint64_t seed, phx_ciphertext *out, phx_stream s) {
    return guard([&] {
        nn(ctx, "context");
        nn(pk, "public key");
        nn(plain, "plain");
        nn(out, "out");
        const Context &c = *ctx->ctx;
        // checked before the plaintext is read: its length follows from the first data level
        require(chain == 1, "encrypt_asymmetric switches from the key level to the first data level (chain 1) only");
        const bool bfv = c.params().scheme == kSchemeBFV;
        const u64 N = c.N();
        const size_t words = bfv ? N : (size_t)c.size_Q() * N;
        if (bfv)
            for (size_t i = 0; i < N; ++i)
                require(plain[i] < c.params().plain_modulus, "plaintext coefficient is not below plain_modulus");
        Scratch pd(words * 8, S(s));
        PHX_HIP_CHECK(hipMemcpyAsync(pd.get(), plain, words * 8, hipMemcpyHostToDevice, S(s)));
        Scratch sm(u ? 3 * N : 0, S(s));
        const int8_t *ud = nullptr;
        const int8_t *ed[2] = {nullptr, nullptr};
        if (u) {
            require(e0 && e1, "explicit samples need u, e0 and e1");
            int8_t *b = reinterpret_cast<int8_t *>(sm.get());
            PHX_HIP_CHECK(hipMemcpyAsync(b, u, N, hipMemcpyHostToDevice, S(s)));
            PHX_HIP_CHECK(hipMemcpyAsync(b + N, e0, N, hipMemcpyHostToDevice, S(s)));
            PHX_HIP_CHECK(hipMemcpyAsync(b + 2 * N, e1, N, hipMemcpyHostToDevice, S(s)));
            ud = b;
            ed[0] = b + N;
            ed[1] = b + 2 * N;
        }
        encrypt_asymmetric(c, pk->pk, pd.get(), (u32)chain, ud, ed, seed, out->ct, S(s));
        PHX_HIP_CHECK(hipStreamSynchronize(S(s)));
    });
}

int phx_secret_key_save_size(const phx_context *ctx, const phx_secret_key *sk, size_t *bytes) {
    return guard([&] {
        nn(ctx, "context");
        nn(sk, "secret key");
        nn(bytes, "bytes");
        *bytes = secret_key_bytes(*ctx->ctx, sk->sk);
    });
}

int phx_secret_key_save(const phx_context *ctx, const phx_secret_key *sk, void *buf, size_t bytes) {
    return guard([&] {
        nn(ctx, "context");
        nn(sk, "secret key");
        nn(buf, "buffer");
        secret_key_save(*ctx->ctx, sk->sk, static_cast<uint8_t *>(buf), bytes);
    });
}

int phx_secret_key_load(const phx_context *ctx, const void *buf, size_t bytes, phx_secret_key **out) {
    return guard([&] {
        nn(ctx, "context");
        nn(buf, "buffer");
        nn(out, "out");
        auto k = std::make_unique<phx_secret_key>();
        secret_key_load(*ctx->ctx, k->sk, static_cast<const uint8_t *>(buf), bytes);
        *out = k.release();
    });
}

int phx_relin_key_save_size(const phx_context *ctx, const phx_relin_key *key, size_t *bytes) {
    return guard([&] {
        nn(ctx, "context");
        nn(key, "key");
        nn(bytes, "bytes");
        *bytes = kswitch_key_bytes(*ctx->ctx, key->key);
    });
}

int phx_relin_key_save(const phx_context *ctx, const phx_relin_key *key, void *buf, size_t bytes) {
    return guard([&] {
        nn(ctx, "context");
        nn(key, "key");
        nn(buf, "buffer");
        kswitch_key_save(*ctx->ctx, key->key, static_cast<uint8_t *>(buf), bytes);
    });
}

int phx_relin_key_load(const phx_context *ctx, const void *buf, size_t bytes, phx_relin_key **out) {
    return guard([&] {
        nn(ctx, "context");
        nn(buf, "buffer");
        nn(out, "out");
        auto k = std::make_unique<phx_relin_key>();
        kswitch_key_load(*ctx->ctx, k->key, static_cast<const uint8_t *>(buf), bytes);
        *out = k.release();
    });
}

int phx_galois_key_save_size(const phx_context *ctx, const phx_galois_key *gk, size_t *bytes) {
    return guard([&] {
        nn(ctx, "context");
        nn(gk, "Galois keys");
        nn(bytes, "bytes");
        size_t n = 8;
        for (const auto &k : gk->gk.keys) n += kswitch_key_bytes(*ctx->ctx, k);
        *bytes = n;
    });
}

int phx_galois_key_save(const phx_context *ctx, const phx_galois_key *gk, void *buf, size_t bytes) {
    return guard([&] {
        nn(ctx, "context");
        nn(gk, "Galois keys");
        nn(buf, "buffer");
        auto *p = static_cast<uint8_t *>(buf);
        const u64 cnt = gk->gk.keys.size();
        require(bytes >= 8, "buffer too small for the Galois keys");
        std::memcpy(p, &cnt, 8);
        size_t off = 8;
        for (const auto &k : gk->gk.keys) {
            const size_t kb = kswitch_key_bytes(*ctx->ctx, k);
            require(off + kb <= bytes, "buffer too small for the Galois keys");
            kswitch_key_save(*ctx->ctx, k, p + off, kb);
            off += kb;
        }
    });
}

int phx_galois_key_load(const phx_context *ctx, const uint32_t *elts, size_t n, const void *buf, size_t bytes,
                        phx_galois_key **out) {
    return guard([&] {
        nn(ctx, "context");
        nn(elts, "elts");
        nn(buf, "buffer");
        nn(out, "out");
        const auto *p = static_cast<const uint8_t *>(buf);
        require(bytes >= 8, "stream too short for Galois keys");
        u64 cnt = 0;
        std::memcpy(&cnt, p, 8);
        require(cnt == n, "the number of Galois keys in the stream differs from the number of elements");
        auto g = std::make_unique<phx_galois_key>();
        size_t off = 8;
        for (size_t i = 0; i < n; ++i) {
            require((elts[i] & 1) && elts[i] < 2 * ctx->ctx->N(), "Galois element is not valid");
            g->gk.elts.push_back(elts[i]);
            g->gk.keys.emplace_back();
            off += kswitch_key_load(*ctx->ctx, g->gk.keys.back(), p + off, bytes - off);
        }
        *out = g.release();
    });
}

int phx_ckks_encode(const phx_context *ctx, const double *re, const double *im, size_t count, double scale,
                    size_t chain_index, uint64_t *out) {
    return guard([&] {
        nn(ctx, "context");
        nn(re, "values");
        nn(out, "out");
        std::vector<std::complex<double>> v(count);
        for (size_t i = 0; i < count; ++i) v[i] = {re[i], im ? im[i] : 0.0};
        ckks_encode(*ctx->ctx, v.data(), count, scale, (u32)chain_index, out);
    });
}

int phx_ckks_decode(const phx_context *ctx, const uint64_t *in, size_t chain_index, double scale, double *re,
                    double *im) {
    return guard([&] {
        nn(ctx, "context");
        nn(in, "input");
        nn(re, "re");
        const Context &c = *ctx->ctx;
        std::vector<std::complex<double>> v(c.N() / 2);
        ckks_decode(c, in, (u32)chain_index, scale, v.data());
        for (size_t i = 0; i < v.size(); ++i) {
            re[i] = v[i].real();
            if (im) im[i] = v[i].imag();
        }
    });
}

int phx_multiply_and_relin_inplace(const phx_context *ctx, phx_ciphertext *ct1, const phx_ciphertext *ct2,
                                   const phx_relin_key *rlk, phx_stream s) {
    return guard([&] {
        nn(ctx, "context");
        nn(ct1, "encrypted1");
        nn(ct2, "encrypted2");
        nn(rlk, "relin_keys");
        check_ct(*ctx->ctx, ct1->ct, "encrypted1");
        check_ct(*ctx->ctx, ct2->ct, "encrypted2");
        check_key(*ctx->ctx, rlk->key, "relin_keys");
        multiply_and_relin(*ctx->ctx, ct1->ct, ct2->ct, rlk->key, S(s));
    });
}

int phx_multiply_and_relin_inplace_noalloc(const phx_context *ctx, phx_ciphertext *ct1, const phx_ciphertext *ct2,
                                           const phx_relin_key *rlk, phx_stream s) {
    return guard([&] {
        nn(ctx, "context");
        nn(ct1, "encrypted1");
        nn(ct2, "encrypted2");
        nn(rlk, "relin_keys");
        check_ct(*ctx->ctx, ct1->ct, "encrypted1");
        check_ct(*ctx->ctx, ct2->ct, "encrypted2");
        check_key(*ctx->ctx, rlk->key, "relin_keys");
        multiply_and_relin_noalloc(*ctx->ctx, ct1->ct, ct2->ct, rlk->key, S(s));
    });
}

int phx_multiply_inplace(const phx_context *ctx, phx_ciphertext *ct1, const phx_ciphertext *ct2, phx_stream s) {
    return guard([&] {
        nn(ctx, "context");
        nn(ct1, "encrypted1");
        nn(ct2, "encrypted2");
        check_ct(*ctx->ctx, ct1->ct, "encrypted1");
        check_ct(*ctx->ctx, ct2->ct, "encrypted2");
        Ciphertext &a = ct1->ct;
        const Ciphertext &b = ct2->ct;
        if (ctx->ctx->params().scheme == kSchemeBFV) {
            // bfv_multiply (src/evaluate.cu:1046-1048), BEHZ
            require(ctx->ctx->params().mul_tech == kMulBEHZ || ctx->ctx->params().mul_tech == kMulHPS ||
                        ctx->ctx->params().mul_tech == kMulHPSOverQ ||
                        ctx->ctx->params().mul_tech == kMulHPSOverQLeveled,
                    "mul_tech not supported for bfv_multiply");
            bfv_multiply(*ctx->ctx, a, b, S(s));
            return;
        }
        require(a.chain_index == b.chain_index, "encrypted1 and encrypted2 parameter mismatch");
        require(a.is_ntt_form && b.is_ntt_form, "encrypted1 and encrypted2 must be in NTT form");
        require(a.size == 2 && b.size == 2, "only size-2 operands are supported");
        const u64 LN = (u64)a.coeff_modulus_size * ctx->ctx->N();
        const double sb = b.scale;
        // keep ct2 readable when it aliases ct1: the tensor reads both operands
        // before writing d0/d1 at the same index.
        a.resize(*ctx->ctx, a.chain_index, 3, S(s));
        const u64 *bd = (ct1 == ct2) ? a.data : b.data;
        tensor_2x2(*ctx->ctx, a.coeff_modulus_size, a.data, bd, a.data, a.data + LN, a.data + 2 * LN, S(s));
        a.scale *= sb;
    });
}

int phx_relinearize_inplace(const phx_context *ctx, phx_ciphertext *ct, const phx_relin_key *rlk, phx_stream s) {
    return guard([&] {
        nn(ctx, "context");
        nn(ct, "encrypted");
        nn(rlk, "relin_keys");
        check_ct(*ctx->ctx, ct->ct, "encrypted");
        check_key(*ctx->ctx, rlk->key, "relin_keys");
        relinearize(*ctx->ctx, ct->ct, rlk->key, S(s));
    });
}

int phx_rescale_to_next(const phx_context *ctx, const phx_ciphertext *in, phx_ciphertext *out, phx_stream s) {
    return guard([&] {
        nn(ctx, "context");
        nn(in, "encrypted");
        nn(out, "destination");
        check_ct(*ctx->ctx, in->ct, "encrypted");
        if (in == out) {
            TempCt res(S(s));
            rescale_to_next(*ctx->ctx, in->ct, res.ct, S(s));
            res.hand_over(out->ct);
        } else {
            rescale_to_next(*ctx->ctx, in->ct, out->ct, S(s));
        }
    });
}

int phx_mod_switch_to_next(const phx_context *ctx, const phx_ciphertext *in, phx_ciphertext *out, phx_stream s) {
    return guard([&] {
        nn(ctx, "context");
        nn(in, "encrypted");
        nn(out, "destination");
        check_ct(*ctx->ctx, in->ct, "encrypted");
        if (in == out) {
            TempCt res(S(s));
            mod_switch_to_next(*ctx->ctx, in->ct, res.ct, S(s));
            res.hand_over(out->ct);
        } else {
            mod_switch_to_next(*ctx->ctx, in->ct, out->ct, S(s));
        }
    });
}

int phx_mod_switch_to_inplace(const phx_context *ctx, phx_ciphertext *ct, size_t chain_index, phx_stream s) {
    return guard([&] {
        nn(ctx, "context");
        nn(ct, "encrypted");
        check_ct(*ctx->ctx, ct->ct, "encrypted");
        require(ct->ct.chain_index <= chain_index, "cannot switch to higher level modulus");
        require(chain_index <= ctx->ctx->size_Q(), "end of modulus switching chain reached");
        if (ct->ct.chain_index == chain_index) return;
        const Context &c = *ctx->ctx;
        if (c.params().scheme == kSchemeBFV) {
            // one division by the last prime per level, as the reference's loop does
            while (ct->ct.chain_index < chain_index) {
                TempCt res(S(s));
                mod_switch_to_next(c, ct->ct, res.ct, S(s));
                res.hand_over(ct->ct);
            }
            return;
        }
        // CKKS: one copy of the kept limbs (the same limbs as repeated drops)
        require(c.params().scheme == kSchemeCKKS, "unsupported scheme");
        require(ct->ct.is_ntt_form, "CKKS encrypted must be in NTT form");
        const u64 N = c.N();
        const Ciphertext &in = ct->ct;
        TempCt tmp(S(s));
        Ciphertext &res = tmp.ct;
        res.resize(c, (u32)chain_index, in.size, S(s));
        const u32 L = in.coeff_modulus_size, Ln = res.coeff_modulus_size;
        for (u32 p = 0; p < in.size; ++p)
            PHX_HIP_CHECK(hipMemcpyAsync(res.data + (u64)p * Ln * N, in.data + (u64)p * L * N, (size_t)Ln * N * 8,
                                         hipMemcpyDeviceToDevice, S(s)));
        res.scale = in.scale;
        res.is_ntt_form = true;
        res.correction_factor = in.correction_factor;
        res.noise_scale_deg = in.noise_scale_deg;
        res.is_asymmetric = in.is_asymmetric;
        tmp.hand_over(ct->ct);
    });
}

int phx_apply_galois_inplace(const phx_context *ctx, phx_ciphertext *ct, uint32_t elt, const phx_galois_key *gk,
                             phx_stream s) {
    return guard([&] {
        nn(ctx, "context");
        nn(ct, "encrypted");
        nn(gk, "galois_keys");
        check_ct(*ctx->ctx, ct->ct, "encrypted");
        check_keys(*ctx->ctx, gk->gk);
        apply_galois(*ctx->ctx, ct->ct, elt, gk->gk, S(s));
    });
}

int phx_rotate_inplace(const phx_context *ctx, phx_ciphertext *ct, int step, const phx_galois_key *gk, phx_stream s) {
    return guard([&] {
        nn(ctx, "context");
        nn(ct, "encrypted");
        nn(gk, "galois_keys");
        check_ct(*ctx->ctx, ct->ct, "encrypted");
        check_keys(*ctx->ctx, gk->gk);
        rotate(*ctx->ctx, ct->ct, step, gk->gk, S(s));
    });
}

int phx_hoisting_inplace(const phx_context *ctx, phx_ciphertext *ct, const phx_galois_key *gk, const int *steps,
                         size_t n, phx_stream s) {
    return guard([&] {
        nn(ctx, "context");
        nn(ct, "encrypted");
        nn(gk, "Galois keys");
        nn(steps, "steps");
        check_ct(*ctx->ctx, ct->ct, "encrypted");
        check_keys(*ctx->ctx, gk->gk);
        hoisting(*ctx->ctx, ct->ct, gk->gk, std::vector<int>(steps, steps + n), S(s));
    });
}

int phx_keyswitch_inplace(const phx_context *ctx, phx_ciphertext *ct, const uint64_t *c2, const phx_relin_key *key,
                          int /*is_relin*/, phx_stream s) {
    return guard([&] {
        nn(ctx, "context");
        nn(ct, "encrypted");
        nn(c2, "c2");
        nn(key, "key");
        check_ct(*ctx->ctx, ct->ct, "encrypted");
        check_key(*ctx->ctx, key->key, "key");
        require(ct->ct.size >= 2, "ciphertext must have two polynomials");
        const u64 LN = (u64)ct->ct.coeff_modulus_size * ctx->ctx->N();
        keyswitch(*ctx->ctx, ct->ct.chain_index, c2, key->key, ct->ct.data, ct->ct.data, ct->ct.data + LN, S(s));
    });
}

int phx_batch_multiply_and_relin_create(const phx_context *ctx, phx_ciphertext *const *ct1,
                                        const phx_ciphertext *const *ct2, size_t n, const phx_relin_key *rlk,
                                        phx_batch **out) {
    return guard([&] {
        nn(ctx, "context");
        nn(ct1, "ct1");
        nn(ct2, "ct2");
        nn(rlk, "relin_keys");
        nn(out, "out");
        check_key(*ctx->ctx, rlk->key, "relin_keys");
        std::vector<Ciphertext *> a;
        std::vector<const Ciphertext *> b;
        for (size_t i = 0; i < n; ++i) {
            nn(ct1[i], "ct1[i]");
            nn(ct2[i], "ct2[i]");
            check_ct(*ctx->ctx, ct1[i]->ct, "ct1[i]");
            check_ct(*ctx->ctx, ct2[i]->ct, "ct2[i]");
            a.push_back(&ct1[i]->ct);
            b.push_back(&ct2[i]->ct);
        }
        PHX_HIP_CHECK(hipStreamSynchronize(hipStreamPerThread));
        *out = new phx_batch{BatchPlan::mul_relin(*ctx->ctx, a, b, rlk->key)};
    });
}

int phx_batch_rotate_create(const phx_context *ctx, phx_ciphertext *const *cts, const int *steps, size_t n,
                            const phx_galois_key *gk, phx_batch **out) {
    return guard([&] {
        nn(ctx, "context");
        nn(cts, "cts");
        nn(steps, "steps");
        nn(gk, "galois_keys");
        nn(out, "out");
        check_keys(*ctx->ctx, gk->gk);
        std::vector<Ciphertext *> a;
        for (size_t i = 0; i < n; ++i) {
            nn(cts[i], "cts[i]");
            check_ct(*ctx->ctx, cts[i]->ct, "cts[i]");
            a.push_back(&cts[i]->ct);
        }
        PHX_HIP_CHECK(hipStreamSynchronize(hipStreamPerThread));
        *out = new phx_batch{BatchPlan::rotate(*ctx->ctx, a, std::vector<int>(steps, steps + n), gk->gk)};
    });
}

int phx_batch_join(phx_batch *const *parts, size_t n, phx_batch **out) {
    return guard([&] {
        nn(parts, "parts");
        nn(out, "out");
        require(n >= 1, "nothing to join");
        std::vector<BatchPlan *> v;
        for (size_t i = 0; i < n; ++i) {
            nn(parts[i], "parts[i]");
            v.push_back(parts[i]->plan);
        }
        *out = new phx_batch{BatchPlan::join(v)};
        for (size_t i = 0; i < n; ++i) delete parts[i];   // the plans now belong to the joined batch
    });
}

int phx_batch_run(phx_batch *b, phx_stream s) {
    return guard([&] {
        nn(b, "batch");
        b->plan->run(S(s));
    });
}

int phx_batch_capture(phx_batch *b, phx_stream s) {
    return guard([&] {
        nn(b, "batch");
        require(s != nullptr, "graph capture needs an explicit stream");
        b->plan->capture(S(s));
    });
}

int phx_batch_replay(phx_batch *b, phx_stream s) {
    return guard([&] {
        nn(b, "batch");
        b->plan->replay(S(s));
    });
}

int phx_batch_scratch_bytes(const phx_batch *b, size_t *bytes) {
    return guard([&] {
        nn(b, "batch");
        nn(bytes, "bytes");
        *bytes = b->plan->scratch_bytes();
    });
}

int phx_batch_launches(const phx_batch *b, size_t *launches) {
    return guard([&] {
        nn(b, "batch");
        nn(launches, "launches");
        *launches = b->plan->launches();
    });
}

int phx_batch_branches(const phx_batch *b, size_t *branches) {
    return guard([&] {
        nn(b, "batch");
        nn(branches, "branches");
        *branches = b->plan->branches();
    });
}

int phx_batch_destroy(phx_batch *b) {
    return guard([&] {
        if (!b) return;
        delete b->plan;
        delete b;
    });
}

int phx_server_create(const phx_context *const *ctxs, const phx_relin_key *const *rlks,
                      const phx_galois_key *const *gks, size_t nlanes, size_t mem_budget, size_t max_active,
                      size_t depth, phx_stream s, phx_server **out) {
    return guard([&] {
        nn(ctxs, "contexts");
        nn(rlks, "relinearisation keys");
        nn(gks, "Galois keys");
        nn(out, "out");
        std::vector<Server::Lane> lanes;
        for (size_t i = 0; i < nlanes; ++i) {
            nn(ctxs[i], "context");
            nn(rlks[i], "relinearisation key");
            nn(gks[i], "Galois keys");
            lanes.push_back({ctxs[i]->ctx, &rlks[i]->key, &gks[i]->gk});
            check_key(*ctxs[i]->ctx, rlks[i]->key, "relinearisation key");
            check_keys(*ctxs[i]->ctx, gks[i]->gk);
        }
        require(s != nullptr, "the server needs an explicit stream");
        *out = new phx_server{new Server(lanes, mem_budget, (u32)max_active, (u32)depth, S(s))};
    });
}

int phx_server_now(const phx_server *srv, double *t) {
    return guard([&] {
        nn(srv, "server");
        nn(t, "t");
        *t = srv->srv->now();
    });
}

int phx_server_submit(phx_server *srv, size_t lane, const phx_ciphertext *ct, const int *steps, size_t k,
                      double arrival, uint64_t *id) {
    return guard([&] {
        nn(srv, "server");
        nn(ct, "ciphertext");
        nn(steps, "steps");
        nn(id, "id");
        *id = srv->srv->submit((u32)lane, ct->ct, std::vector<int>(steps, steps + k), arrival);
    });
}

int phx_server_step(phx_server *srv, size_t *issued) {
    return guard([&] {
        nn(srv, "server");
        const u32 n = srv->srv->step();
        if (issued) *issued = n;
    });
}

int phx_server_drain(phx_server *srv) {
    return guard([&] {
        nn(srv, "server");
        srv->srv->drain();
    });
}

int phx_server_status(const phx_server *srv, uint64_t id, int *state) {
    return guard([&] {
        nn(srv, "server");
        nn(state, "state");
        *state = srv->srv->status(id);
    });
}

int phx_server_take(phx_server *srv, uint64_t id, phx_ciphertext *out, double *latency_s) {
    return guard([&] {
        nn(srv, "server");
        nn(out, "out");
        srv->srv->take(id, out->ct, latency_s);
    });
}

int phx_server_stats_get(const phx_server *srv, phx_server_stats *out) {
    return guard([&] {
        nn(srv, "server");
        nn(out, "out");
        const ServerStats st = srv->srv->stats();
        *out = phx_server_stats{st.submitted, st.admitted,   st.completed, st.ticks,        st.ops,
                                st.p50_ms,    st.p99_ms,     st.mean_ms,   st.max_ms,       st.span_s,
                                st.req_per_s, st.ops_per_s,  st.peak_reserved, st.budget, st.peak_active};
    });
}

int phx_server_destroy(phx_server *srv) {
    return guard([&] {
        if (!srv) return;
        delete srv->srv;
        delete srv;
    });
}

int phx_ntt_forward(const phx_context *ctx, uint64_t *dev, size_t L, size_t start, phx_stream s) {
    return guard([&] {
        nn(ctx, "context");
        nn(dev, "data");
        require(start + L <= ctx->ctx->size_QP(), "limb range exceeds the modulus chain");
        ntt_forward(ctx->ctx->tables(), limb_map(dev, (u32)start, (u32)L), (u32)L, 1, S(s));
    });
}

int phx_ntt_inverse(const phx_context *ctx, uint64_t *dev, size_t L, size_t start, phx_stream s) {
    return guard([&] {
        nn(ctx, "context");
        nn(dev, "data");
        require(start + L <= ctx->ctx->size_QP(), "limb range exceeds the modulus chain");
        ntt_inverse(ctx->ctx->tables(), limb_map(dev, (u32)start, (u32)L), (u32)L, 1, S(s));
    });
}

int phx_ntt_forward_batch(const phx_context *ctx, uint64_t *dev, size_t L, size_t start, size_t npolys,
                          size_t poly_stride, phx_stream s) {
    return guard([&] {
        nn(ctx, "context");
        nn(dev, "data");
        require(start + L <= ctx->ctx->size_QP(), "limb range exceeds the modulus chain");
        require(npolys <= 65535, "too many polynomials in one launch");
        ntt_forward(ctx->ctx->tables(), limb_map(dev, (u32)start, (u32)L, poly_stride), (u32)L, (u32)npolys, S(s));
    });
}

int phx_ntt_inverse_batch(const phx_context *ctx, uint64_t *dev, size_t L, size_t start, size_t npolys,
                          size_t poly_stride, phx_stream s) {
    return guard([&] {
        nn(ctx, "context");
        nn(dev, "data");
        require(start + L <= ctx->ctx->size_QP(), "limb range exceeds the modulus chain");
        require(npolys <= 65535, "too many polynomials in one launch");
        ntt_inverse(ctx->ctx->tables(), limb_map(dev, (u32)start, (u32)L, poly_stride), (u32)L, (u32)npolys, S(s));
    });
}

int phx_modup(const phx_context *ctx, size_t chain, const uint64_t *c2, uint64_t *t, phx_stream s) {
    return guard([&] {
        nn(ctx, "context");
        nn(c2, "c2");
        nn(t, "t_mod_up");
        modup(*ctx->ctx, (u32)chain, c2, t, S(s));
    });
}

int phx_key_inner_product(const phx_context *ctx, size_t chain, const uint64_t *t, const phx_relin_key *key,
                          uint64_t *cx, phx_stream s) {
    return guard([&] {
        nn(ctx, "context");
        nn(t, "t_mod_up");
        nn(key, "key");
        nn(cx, "cx");
        inner_product(*ctx->ctx, (u32)chain, t, key->key, cx, S(s));
    });
}

int phx_moddown_from_ntt(const phx_context *ctx, size_t chain, uint64_t *cx, uint64_t *out, size_t npolys,
                         phx_stream s) {
    return guard([&] {
        nn(ctx, "context");
        nn(cx, "cx");
        nn(out, "out");
        require(npolys >= 1 && npolys <= 2, "npolys must be 1 or 2");
        moddown(*ctx->ctx, (u32)chain, cx, out, nullptr, nullptr, (u32)npolys, S(s));
    });
}

int phx_malloc(size_t bytes, void **out) {
    return guard([&] {
        nn(out, "out");
        PHX_HIP_CHECK(hipMalloc(out, bytes));
    });
}
int phx_free(void *p) {
    return guard([&] { PHX_HIP_CHECK(hipFree(p)); });
}
int phx_memcpy_htod(void *dst, const void *src, size_t bytes, phx_stream s) {
    return guard([&] {
        PHX_HIP_CHECK(hipMemcpyAsync(dst, src, bytes, hipMemcpyHostToDevice, S(s)));
        PHX_HIP_CHECK(hipStreamSynchronize(S(s)));
    });
}
int phx_memcpy_dtoh(void *dst, const void *src, size_t bytes, phx_stream s) {
    return guard([&] {
        PHX_HIP_CHECK(hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToHost, S(s)));
        PHX_HIP_CHECK(hipStreamSynchronize(S(s)));
    });
}
int phx_stream_synchronize(phx_stream s) {
    return guard([&] { PHX_HIP_CHECK(hipStreamSynchronize(S(s))); });
}
int phx_device_synchronize(void) {
    return guard([&] { PHX_HIP_CHECK(hipDeviceSynchronize()); });
}
int phx_stream_create(phx_stream *out) {
    return guard([&] {
        nn(out, "out");
        hipStream_t st;
        PHX_HIP_CHECK(hipStreamCreateWithFlags(&st, hipStreamNonBlocking));
        *out = st;
    });
}
int phx_stream_destroy(phx_stream s) {
    return guard([&] {
        if (s) PHX_HIP_CHECK(hipStreamDestroy(static_cast<hipStream_t>(s)));
    });
}
int phx_event_create(void **out) {
    return guard([&] {
        nn(out, "out");
        hipEvent_t e;
        PHX_HIP_CHECK(hipEventCreate(&e));
        *out = e;
    });
}
int phx_event_record(void *ev, phx_stream s) {
    return guard([&] { PHX_HIP_CHECK(hipEventRecord(static_cast<hipEvent_t>(ev), S(s))); });
}
int phx_event_elapsed_ms(void *a, void *b, float *ms) {
    return guard([&] {
        nn(ms, "ms");
        PHX_HIP_CHECK(hipEventSynchronize(static_cast<hipEvent_t>(b)));
        PHX_HIP_CHECK(hipEventElapsedTime(ms, static_cast<hipEvent_t>(a), static_cast<hipEvent_t>(b)));
    });
}
int phx_event_destroy(void *ev) {
    return guard([&] {
        if (ev) PHX_HIP_CHECK(hipEventDestroy(static_cast<hipEvent_t>(ev)));
    });
}
int phx_tune_set(const char *key, int value) {
    return guard([&] { require(tune_set(key, value), std::string("unknown tuning key: ") + (key ? key : "(null)")); });
}

int phx_tune_get(const char *key, int *value) {
    return guard([&] {
        nn(value, "value");
        require(tune_get(key, value), std::string("unknown tuning key: ") + (key ? key : "(null)"));
    });
}

int phx_context_get_tuning(const phx_context *ctx, int choice[PHX_TUNE_AXES]) {
    return guard([&] {
        nn(ctx, "context");
        nn(choice, "choice");
        put_choice(ctx->ctx->tuning(), choice);
    });
}

int phx_context_set_tuning(phx_context *ctx, const int choice[PHX_TUNE_AXES]) {
    return guard([&] {
        nn(ctx, "context");
        nn(choice, "choice");
        ctx->ctx->set_tuning(choice_of(choice));
    });
}

int phx_autotune(phx_context *ctx, const char *cache_path, phx_tune_report *report, phx_stream s) {
    return guard([&] {
        nn(ctx, "context");
        const TuneResult r = autotune(*ctx->ctx, S(s));
        const Context &c = *ctx->ctx;
        if (cache_path && *cache_path)
            tune_cache_store(cache_path, tune_cache_key((u32)c.params().scheme, c.N(), c.size_Q(), c.size_P()),
                             r.choice, r.best_us);
        if (report) {
            put_choice(r.choice, report->choice);
            report->default_us = r.default_us;
            report->best_us = r.best_us;
            report->trials = (uint32_t)r.trials.size();
            // every candidate with its objective, as a JSON array (truncated to the buffer)
            std::string j = "[";
            for (size_t i = 0; i < r.trials.size(); ++i) {
                char us[32];
                std::snprintf(us, sizeof us, "%.2f", r.trials[i].second);
                const std::string e = std::string(i ? ", " : "") + "{\"choice\": \"" + r.trials[i].first +
                                      "\", \"us\": " + us + "}";
                if (j.size() + e.size() + 2 > sizeof report->trials_json) break;
                j += e;
            }
            j += "]";
            std::snprintf(report->trials_json, sizeof report->trials_json, "%s", j.c_str());
        }
    });
}

int phx_tune_load(phx_context *ctx, const char *cache_path, int *found) {
    return guard([&] {
        nn(ctx, "context");
        nn(cache_path, "cache_path");
        TuneChoice t;
        const bool ok = tune_cache_load(cache_path, *ctx->ctx, &t);
        if (ok) ctx->ctx->set_tuning(t);
        if (found) *found = ok ? 1 : 0;
    });
}

int phx_tune_cache_store(const char *cache_path, int scheme, uint64_t N, size_t size_Q, size_t size_P,
                         const int choice[PHX_TUNE_AXES], double hmult_relin_us) {
    return guard([&] {
        nn(cache_path, "cache_path");
        nn(choice, "choice");
        tune_cache_store(cache_path, tune_cache_key((u32)scheme, N, (u32)size_Q, (u32)size_P), choice_of(choice),
                         hmult_relin_us);
    });
}

int phx_tune_cache_lookup(const char *cache_path, int scheme, uint64_t N, size_t size_Q, size_t size_P,
                          int choice[PHX_TUNE_AXES], int *found) {
    return guard([&] {
        nn(cache_path, "cache_path");
        nn(choice, "choice");
        TuneChoice t;
        const bool ok = tune_cache_lookup(cache_path, tune_cache_key((u32)scheme, N, (u32)size_Q, (u32)size_P), &t);
        if (ok) put_choice(t, choice);
        if (found) *found = ok ? 1 : 0;
    });
}

int phx_copy_u64(uint64_t *dst, const uint64_t *src, size_t n, phx_stream s) {
    return guard([&] {
        nn(dst, "dst");
        nn(src, "src");
        copy_u64(dst, src, n, S(s));
    });
}

int phx_device_count(int *n) {
    return guard([&] {
        nn(n, "n");
        PHX_HIP_CHECK(hipGetDeviceCount(n));
    });
}

}  // extern "C"
