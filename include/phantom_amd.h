/*
 * phantom_amd.h -- C ABI of the MI355X (gfx950) implementation of
 * PhantomFHE's RNS polynomial hot path: negacyclic NTT/INTT, RNS base
 * conversion (modup/moddown) and hybrid key switching (relinearisation,
 * Galois rotation), plus CKKS rescale.
 *
 * Each entry point names the reference interface it replaces (paths are
 * relative to the PhantomFHE source tree). All functions return PHX_OK (0)
 * or a negative status; phx_last_error() returns the thread's last message.
 * Status codes map to the reference's exception types:
 *   PHX_EINVAL   -> std::invalid_argument
 *   PHX_ELOGIC   -> std::logic_error
 *   PHX_ERUNTIME -> std::runtime_error (HIP / allocation failures)
 *
 * Device buffers hold u64 residues laid out like the reference:
 *   ciphertext [poly][limb][N], key [digit][2][QP][N], t_mod_up [digit][QlP][N].
 * phx_stream is a hipStream_t; NULL selects the per-thread default stream
 * (the reference runs on cudaStreamPerThread, src/evaluate.cu:1079).
 */
#ifndef PHANTOM_AMD_H
#define PHANTOM_AMD_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define PHX_OK 0
#define PHX_EINVAL (-1)
#define PHX_ELOGIC (-2)
#define PHX_ERUNTIME (-3)

/* scheme_type / mul_tech_type values of include/host/encryptionparams.h:14-35 */
#define PHX_SCHEME_BGV 1
#define PHX_SCHEME_BFV 2
#define PHX_SCHEME_CKKS 3
#define PHX_MUL_TECH_NONE 0
#define PHX_MUL_TECH_BEHZ 1
#define PHX_MUL_TECH_HPS 2
#define PHX_MUL_TECH_HPS_OVERQ 3
#define PHX_MUL_TECH_HPS_OVERQ_LEVELED 4

typedef struct phx_context phx_context;
typedef struct phx_ciphertext phx_ciphertext;
typedef struct phx_secret_key phx_secret_key;
typedef struct phx_public_key phx_public_key;
typedef struct phx_relin_key phx_relin_key;
typedef struct phx_galois_key phx_galois_key;
typedef void *phx_stream;

/* thread-local message of the last failed call */
const char *phx_last_error(void);

/* EncryptionParameters (include/host/encryptionparams.h:37-245) */
typedef struct {
    int scheme;
    int mul_tech;
    uint64_t poly_modulus_degree;
    const uint64_t *coeff_modulus;   /* size_Q data primes, then the special primes */
    size_t coeff_modulus_size;
    size_t special_modulus_size;
    uint64_t plain_modulus;
    const uint32_t *galois_elts;     /* may be NULL */
    size_t galois_elts_size;
} phx_params;

/* ---- parameter maths ---------------------------------------------------- */
/* CoeffModulus::Create (src/host/modulus.cu:80-111) */
int phx_coeff_modulus_create(uint64_t poly_modulus_degree, const int *bit_sizes, size_t n, uint64_t *out);
/* get_elt_from_step (include/galois.cuh:16-49) */
int phx_galois_elt_from_step(int step, uint64_t poly_modulus_degree, uint32_t *out);

/* ---- context ------------------------------------------------------------ */
/* PhantomContext::PhantomContext (src/context.cu:121-232) */
int phx_context_create(const phx_params *params, phx_context **out);
int phx_context_destroy(phx_context *ctx);
/* sizes: N, size_Q, size_P, dnum (= ceil(size_Q / size_P)) */
int phx_context_info(const phx_context *ctx, uint64_t *N, size_t *size_Q, size_t *size_P, size_t *dnum);
/* limbs of chain index c (context_data_[c], src/context.cu:145-159) */
int phx_context_limbs_at(const phx_context *ctx, size_t chain_index, size_t *limbs);

/* ---- ciphertext (include/ciphertext.h:7-308) ------------------------------ */
int phx_ciphertext_create(phx_ciphertext **out);
int phx_ciphertext_destroy(phx_ciphertext *ct);
/* PhantomCiphertext::resize(context, chain_index, size, stream) (include/ciphertext.h:44-72) */
int phx_ciphertext_resize(const phx_context *ctx, phx_ciphertext *ct, size_t chain_index, size_t size,
                          phx_stream stream);
int phx_ciphertext_upload(phx_ciphertext *ct, const uint64_t *host, size_t n_elems, phx_stream stream);
int phx_ciphertext_download(const phx_ciphertext *ct, uint64_t *host, size_t n_elems, phx_stream stream);
int phx_ciphertext_copy(const phx_context *ctx, phx_ciphertext *dst, const phx_ciphertext *src, phx_stream stream);
/* PhantomCiphertext's copy constructor (include/ciphertext.h:31, a deep copy of the
 * device buffer and every field): *out is a new handle */
int phx_ciphertext_clone(const phx_ciphertext *src, phx_ciphertext **out, phx_stream stream);
int phx_ciphertext_info(const phx_ciphertext *ct, size_t *chain_index, size_t *size, size_t *coeff_modulus_size,
                        double *scale, int *is_ntt_form);
int phx_ciphertext_set_scale(phx_ciphertext *ct, double scale);
/* PhantomCiphertext::is_ntt_form (include/ciphertext.h); a new ciphertext is in NTT
 * form for CKKS and in coefficient form for BFV */
int phx_ciphertext_set_ntt_form(phx_ciphertext *ct, int is_ntt_form);
int phx_ciphertext_data(const phx_ciphertext *ct, uint64_t **dev_ptr);

/* ---- keys (include/secretkey.h) ------------------------------------------ */
/* PhantomSecretKey(context): ternary secret from a seed, stored in NTT form */
int phx_secret_key_create(const phx_context *ctx, uint64_t seed, phx_secret_key **out);
/* secret from explicit ternary coefficients (N int8 values in {-1,0,1}) */
int phx_secret_key_create_from_coeffs(const phx_context *ctx, const int8_t *coeffs, phx_secret_key **out);
int phx_secret_key_download(const phx_context *ctx, const phx_secret_key *sk, uint64_t *host_qp_n);
int phx_secret_key_destroy(phx_secret_key *sk);
/* PhantomSecretKey::gen_relinkey (src/secretkey.cu:396-419), dnum = ceil(size_Q/size_P) */
int phx_relin_key_create(const phx_context *ctx, phx_secret_key *sk, uint64_t seed, phx_relin_key **out);
/* load a key laid out [dnum][2][QP][N] (PhantomRelinKey::load, include/secretkey.h:129-162) */
int phx_relin_key_create_from_host(const phx_context *ctx, const uint64_t *host, size_t dnum, phx_relin_key **out);
int phx_relin_key_download(const phx_context *ctx, const phx_relin_key *key, uint64_t *host);
int phx_relin_key_destroy(phx_relin_key *key);
/* PhantomSecretKey::create_galois_keys (src/secretkey.cu:421-461) */
int phx_galois_key_create(const phx_context *ctx, const phx_secret_key *sk, const uint32_t *elts, size_t n,
                          uint64_t seed, phx_galois_key **out);
/* load keys [n][dnum][2][QP][N] for the listed elements (include/secretkey.h:194-219) */
int phx_galois_key_create_from_host(const phx_context *ctx, const uint32_t *elts, size_t n, const uint64_t *host,
                                    size_t dnum, phx_galois_key **out);
int phx_galois_key_destroy(phx_galois_key *key);

/* ---- symmetric encryption / decryption (CKKS, NTT form) ------------------- */
/* PhantomSecretKey::encrypt_symmetric (src/secretkey.cu:463-531); plain = [Ql][N] coefficient residues */
int phx_encrypt_symmetric(const phx_context *ctx, const phx_secret_key *sk, const uint64_t *plain_host,
                          size_t chain_index, uint64_t seed, phx_ciphertext *out, phx_stream stream);
/* encrypt_symmetric with c1 drawn from the seeded uniform sampler (sample_uniform_poly,
 * src/prng.cu:174-210) for the 64-byte seed_a, as the reference does
 * (src/secretkey.cu:232-300, 463-500); the ciphertext keeps the seed so that
 * phx_ciphertext_save_symmetric can store it instead of c1. e comes from seed_e. */
int phx_encrypt_symmetric_seeded(const phx_context *ctx, const phx_secret_key *sk, const uint64_t *plain_host,
                                 size_t chain_index, const uint8_t *seed_a, uint64_t seed_e, phx_ciphertext *out,
                                 phx_stream stream);
/* sample_uniform_poly_wrap (src/prng.cu:206-210): out (device, [coeff_mod_size][N]) <- the
 * seeded uniform sampler's output for the 64-byte seed (host) */
int phx_sample_uniform_poly(const phx_context *ctx, const uint8_t *seed, size_t coeff_mod_size, uint64_t *out,
                            phx_stream stream);
/* PhantomSecretKey::ckks_decrypt (src/secretkey.cu:533-569) followed by the INTT: out = [Ql][N] */
int phx_decrypt(const phx_context *ctx, phx_secret_key *sk, const phx_ciphertext *ct, uint64_t *out_host,
                phx_stream stream);
/* PhantomSecretKey::decrypt of a BFV ciphertext (bfv_decrypt, src/secretkey.cu:571-635): the
 * plaintext's N coefficients (host), scale-and-round by BEHZ for mul_tech behz
 * (src/rns.cu:1008-1090), by HPS otherwise (src/rns.cu:1519-1690, its four forms chosen as the
 * reference chooses them, with the level's constants); coefficient-form ciphertexts at any data level */
int phx_decrypt_bfv(const phx_context *ctx, phx_secret_key *sk, const phx_ciphertext *ct, uint64_t *plain,
                    phx_stream stream);

/* ---- serialization: the reference's stream layouts, into / out of caller buffers.
 * Ciphertext (PhantomCiphertext::save / load, include/ciphertext.h:173-216): 58-byte
 * header (chain_index, size, poly_modulus_degree, coeff_modulus_size as u64, scale as
 * double, correction_factor u64, noiseScaleDeg u64, is_ntt_form and is_asymmetric one
 * byte each), then [size][L][N] u64. Relin key (PhantomRelinKey::save / load,
 * include/secretkey.h:129-162): dnum (u64), then dnum public-key ciphertexts. Galois
 * keys (PhantomGaloisKey::save / load, include/secretkey.h:194-219): the count (u64),
 * then that many relin keys, in the order of the elts list given to load. */
int phx_ciphertext_save_size(const phx_ciphertext *ct, size_t *bytes);
int phx_ciphertext_save(const phx_ciphertext *ct, void *buf, size_t bytes);
int phx_ciphertext_load(const phx_context *ctx, phx_ciphertext *ct, const void *buf, size_t bytes);
/* PhantomCiphertext::save_symmetric / load_symmetric (include/ciphertext.h:216-305): header, c0,
 * then the 64-byte seed instead of c1 (regenerated on load). Saving requires a ciphertext from
 * phx_encrypt_symmetric_seeded (or a symmetric load) whose c1 is still the seed's output. */
int phx_ciphertext_save_symmetric_size(const phx_ciphertext *ct, size_t *bytes);
int phx_ciphertext_save_symmetric(const phx_context *ctx, const phx_ciphertext *ct, void *buf, size_t bytes);
int phx_ciphertext_load_symmetric(const phx_context *ctx, phx_ciphertext *ct, const void *buf, size_t bytes);
/* PhantomSecretKey::gen_publickey (src/secretkey.cu:382-394): pk = (-(a s + e), a) at the key
 * level (QP limbs, NTT form); a from the seeded uniform sampler for the 64-byte seed_a (NULL: the
 * library's counter hash), e from seed_e */
int phx_public_key_create(const phx_context *ctx, const phx_secret_key *sk, const uint8_t *seed_a, uint64_t seed_e,
                          phx_public_key **out);
int phx_public_key_destroy(phx_public_key *pk);
int phx_public_key_download(const phx_context *ctx, const phx_public_key *pk, uint64_t *host_2_qp_n);
/* PhantomPublicKey::save / load (include/secretkey.h:85-100): the public key's ciphertext stream */
int phx_public_key_save_size(const phx_public_key *pk, size_t *bytes);
int phx_public_key_save(const phx_public_key *pk, void *buf, size_t bytes);
int phx_public_key_load(const phx_context *ctx, const void *buf, size_t bytes, phx_public_key **out);
/* PhantomPublicKey::encrypt_asymmetric (src/secretkey.cu:11-186) at the first data level (chain 1):
 * c_j = moddown_{QP->Q}(pk_j u + e_j), then c0 += the plaintext: CKKS plain = coefficient residues
 * [size_Q][N] (from phx_ckks_encode), added in NTT form; BFV plain = N coefficients mod t, added as
 * (Q m + [-Q m]_t) / t (multiply_add_plain_with_scaling_variant). u ternary and e_0, e_1 from `seed`, or,
 * in the _with_samples form, the given int8 coefficients (N each, host) */
int phx_encrypt_asymmetric(const phx_context *ctx, const phx_public_key *pk, const uint64_t *plain, size_t chain,
                           uint64_t seed, phx_ciphertext *out, phx_stream stream);
int phx_encrypt_asymmetric_with_samples(const phx_context *ctx, const phx_public_key *pk, const uint64_t *plain,
                                        size_t chain, const int8_t *u, const int8_t *e0, const int8_t *e1,
                                        uint64_t seed, phx_ciphertext *out, phx_stream stream);
/* PhantomSecretKey::save / load (include/secretkey.h:347-389): max power, N, QP (8 bytes each),
 * then s (and s^2 when computed) in NTT form over QP; load keeps the first two powers */
int phx_secret_key_save_size(const phx_context *ctx, const phx_secret_key *sk, size_t *bytes);
int phx_secret_key_save(const phx_context *ctx, const phx_secret_key *sk, void *buf, size_t bytes);
int phx_secret_key_load(const phx_context *ctx, const void *buf, size_t bytes, phx_secret_key **out);
int phx_relin_key_save_size(const phx_context *ctx, const phx_relin_key *key, size_t *bytes);
int phx_relin_key_save(const phx_context *ctx, const phx_relin_key *key, void *buf, size_t bytes);
int phx_relin_key_load(const phx_context *ctx, const void *buf, size_t bytes, phx_relin_key **out);
int phx_galois_key_save_size(const phx_context *ctx, const phx_galois_key *gk, size_t *bytes);
int phx_galois_key_save(const phx_context *ctx, const phx_galois_key *gk, void *buf, size_t bytes);
int phx_galois_key_load(const phx_context *ctx, const uint32_t *elts, size_t n, const void *buf, size_t bytes,
                        phx_galois_key **out);

/* ---- CKKS encoder (include/ckks.h), host side ------------------------------ */
/* PhantomCKKSEncoder::encode (src/ckks.cu:65-135): count <= N/2 complex values
 * (re[i] + i*im[i], im may be NULL) -> round(scale * m) as coefficient residues
 * [Ql][N] at chain_index, ready for phx_encrypt_symmetric */
int phx_ckks_encode(const phx_context *ctx, const double *re, const double *im, size_t count, double scale,
                    size_t chain_index, uint64_t *out_host);
/* PhantomCKKSEncoder::decode (src/ckks.cu:347-400): coefficient residues [Ql][N]
 * (as phx_decrypt returns them) -> N/2 slot values re[i] + i*im[i] */
int phx_ckks_decode(const phx_context *ctx, const uint64_t *in_host, size_t chain_index, double scale, double *re,
                    double *im);

/* ---- evaluator: the hot path (include/evaluate.cuh) ------------------------ */
/* multiply_and_relin_inplace (include/evaluate.cuh:115-116, src/evaluate.cu:1061-1104) */
int phx_multiply_and_relin_inplace(const phx_context *ctx, phx_ciphertext *ct1, const phx_ciphertext *ct2,
                                   const phx_relin_key *rlk, phx_stream stream);
/* No counterpart in the reference: multiply_and_relin_inplace for CKKS that allocates
 * nothing, for callers that cannot allocate device memory (code under HIP graph capture,
 * tight serving loops, harnesses that forbid hipMalloc). Its workspace is a static device
 * array of 192 MiB per GPU (enough for N = 2^16 with 35 data limbs, 12 special limbs and
 * 3 digits); larger shapes return PHX_EINVAL before any kernel runs. Calls on one GPU
 * share that array: overlapping calls on one GPU race. Results equal
 * phx_multiply_and_relin_inplace's bit for bit. */
int phx_multiply_and_relin_inplace_noalloc(const phx_context *ctx, phx_ciphertext *ct1, const phx_ciphertext *ct2,
                                           const phx_relin_key *rlk, phx_stream stream);
/* multiply_inplace (include/evaluate.cuh:104-105): tensor only, result size 3 */
int phx_multiply_inplace(const phx_context *ctx, phx_ciphertext *ct1, const phx_ciphertext *ct2, phx_stream stream);
/* relinearize_inplace (include/evaluate.cuh:125-126, src/evaluate.cu:1342-1374) */
int phx_relinearize_inplace(const phx_context *ctx, phx_ciphertext *ct, const phx_relin_key *rlk, phx_stream stream);
/* mod_switch_to_next / mod_switch_to_inplace (include/evaluate.cuh:145-178, src/evaluate.cu:1376-1540):
 * CKKS drops the last limb(s) (NTT form, scale kept); BFV divides by the last prime per level
 * (coefficient form, floor as the reference's divide_and_round_q_last). in == out allowed. */
int phx_mod_switch_to_next(const phx_context *ctx, const phx_ciphertext *in, phx_ciphertext *out, phx_stream stream);
int phx_mod_switch_to_inplace(const phx_context *ctx, phx_ciphertext *ct, size_t chain_index, phx_stream stream);
/* rescale_to_next (include/evaluate.cuh:137, src/evaluate.cu:1545-1565) */
int phx_rescale_to_next(const phx_context *ctx, const phx_ciphertext *in, phx_ciphertext *out, phx_stream stream);
/* apply_galois_inplace (include/evaluate.cuh:215-216, src/evaluate.cu:1567-1630) */
int phx_apply_galois_inplace(const phx_context *ctx, phx_ciphertext *ct, uint32_t galois_elt,
                             const phx_galois_key *gk, phx_stream stream);
/* rotate_inplace (include/evaluate.cuh:225-226, src/evaluate.cu:1633-1668) */
int phx_rotate_inplace(const phx_context *ctx, phx_ciphertext *ct, int step, const phx_galois_key *gk,
                       phx_stream stream);
/* hoisting_inplace (include/evaluate.cuh:237-238, src/evaluate.cu:1670-1865): ct <- the sum
 * of the rotations of ct by steps[0..n-1], one modup shared by all of them (CKKS in NTT
 * form; BFV in coefficient form) */
int phx_hoisting_inplace(const phx_context *ctx, phx_ciphertext *ct, const phx_galois_key *gk, const int *steps,
                         size_t n, phx_stream stream);
/* keyswitch_inplace (include/evaluate.cuh:29-32, src/eval_key_switch.cu:95-182):
 * ct[0..1] += keyswitch(c2) with c2 = [Ql][N] (NTT form) in device memory */
int phx_keyswitch_inplace(const phx_context *ctx, phx_ciphertext *ct, const uint64_t *c2_dev,
                          const phx_relin_key *key, int is_relin, phx_stream stream);

/* ---- batched hot path (no single reference counterpart: it runs the reference's
 * multiply_and_relin_inplace / rotate_inplace over many ciphertexts at once) -----------
 * A batch binds a fixed list of ciphertexts (their device buffers must not be resized
 * while the batch lives) to preallocated scratch and device work lists; every stage
 * of the pipeline is one launch over all items, whatever their levels. run() launches eagerly,
 * capture() records the launches into a hipGraph on an explicit (non-NULL) stream and
 * replay() launches that graph. Results are bit-identical to the one-at-a-time calls.
 * A batch of at least 2 x B items (B = the "batch_branches" switch, default 2) is split
 * into B runs of items of about equal work that run concurrently: branch 0 on the
 * caller's stream, the others on batch-owned streams forked from it and joined back
 * into it, so run(), capture() and replay() still order everything on the caller's
 * stream. */
typedef struct phx_batch phx_batch;
/* ct1[i] = relin(ct1[i] * ct2[i]); ct1[i] may equal ct2[i] (squaring).
 * CKKS: NTT form, any level. BFV: coefficient form at the top level, multiplied with
 * the context's technique (BEHZ or HPS) as in multiply_inplace + relinearize_inplace
 * (src/evaluate.cu:440-801, 1342-1374). */
int phx_batch_multiply_and_relin_create(const phx_context *ctx, phx_ciphertext *const *ct1,
                                        const phx_ciphertext *const *ct2, size_t n, const phx_relin_key *rlk,
                                        phx_batch **out);
/* cts[i] = rotate(cts[i], steps[i]); each step's Galois element must have a key in gk */
int phx_batch_rotate_create(const phx_context *ctx, phx_ciphertext *const *cts, const int *steps, size_t n,
                            const phx_galois_key *gk, phx_batch **out);
/* one batch that runs `parts` (one graph when captured); parts may belong to different
 * contexts (ring degrees), joined batches are flattened into their parts. The parts are
 * spread over B concurrent branches (largest first onto the least loaded branch); the
 * parts of one branch run in order in one scratch arena of the largest one's size.
 * Takes ownership: the parts' handles are released and must not be used again; on
 * error they are left untouched. */
int phx_batch_join(phx_batch *const *parts, size_t n, phx_batch **out);
int phx_batch_run(phx_batch *batch, phx_stream stream);
int phx_batch_capture(phx_batch *batch, phx_stream stream);
int phx_batch_replay(phx_batch *batch, phx_stream stream);
/* bytes of device scratch the batch holds */
int phx_batch_scratch_bytes(const phx_batch *batch, size_t *bytes);
/* kernel launches one run issues (the kernel nodes of the captured graph): one per
 * pipeline stage across all items and levels of the batch */
int phx_batch_launches(const phx_batch *batch, size_t *launches);
/* concurrent branches of the batch (1 when it runs on the caller's stream alone) */
int phx_batch_branches(const phx_batch *batch, size_t *branches);
int phx_batch_destroy(phx_batch *batch);

/* ---- serving loop (no single reference counterpart: a request loop around the
 * reference's evaluator calls, multiply_and_relin_inplace / rescale_to_next /
 * rotate_inplace of include/evaluate.cuh:115-116, 137, 225-226).
 * A request is a CKKS ciphertext and a chain of k links, link i being
 *   ct <- rotate(rescale(ct * ct relinearised), steps[i]).
 * Evaluation keys stay resident. Each phx_server_step is one scheduling tick:
 * it admits pending requests in arrival order while their reservation (their
 * ciphertext + per-item scratch at their level) fits mem_budget_bytes and fewer
 * than max_active are active, then issues the next operation of every active
 * request as batched launches (continuous batching). The host runs ahead of the
 * device by at most `depth` ticks. Latency = completion observed - arrival. */
typedef struct phx_server phx_server;
typedef struct {
    uint64_t submitted, admitted, completed, ticks, ops;
    double p50_ms, p99_ms, mean_ms, max_ms;
    double span_s, req_per_s, ops_per_s;
    uint64_t peak_reserved_bytes, budget_bytes, peak_active;
} phx_server_stats;
/* lane i serves ciphertexts of ctxs[i] with rlks[i] and gks[i] */
int phx_server_create(const phx_context *const *ctxs, const phx_relin_key *const *rlks,
                      const phx_galois_key *const *gks, size_t nlanes, size_t mem_budget_bytes, size_t max_active,
                      size_t depth, phx_stream stream, phx_server **out);
/* seconds since the server was created (the clock of arrival times) */
int phx_server_now(const phx_server *srv, double *t);
/* copies ct (size 2, NTT form); arrival < 0 means now */
int phx_server_submit(phx_server *srv, size_t lane, const phx_ciphertext *ct, const int *steps, size_t k,
                      double arrival, uint64_t *id);
int phx_server_step(phx_server *srv, size_t *issued_ops);
int phx_server_drain(phx_server *srv);
/* 0 pending, 1 active, 2 complete (result waiting), 3 taken or unknown */
int phx_server_status(const phx_server *srv, uint64_t id, int *state);
/* moves a complete request's result into out and frees its reservation */
int phx_server_take(phx_server *srv, uint64_t id, phx_ciphertext *out, double *latency_s);
int phx_server_stats_get(const phx_server *srv, phx_server_stats *out);
int phx_server_destroy(phx_server *srv);

/* ---- low-level RNS kernels (device pointers) ------------------------------- */
/* nwt_2d_radix8_forward_inplace (include/ntt.cuh:173-226, src/ntt/fntt_2d.cu:620-653):
 * limbs [0, L) of dev use QP table rows [start, start + L) */
int phx_ntt_forward(const phx_context *ctx, uint64_t *dev, size_t L, size_t start, phx_stream stream);
/* nwt_2d_radix8_backward_inplace (src/ntt/intt_2d.cu:724-757) */
int phx_ntt_inverse(const phx_context *ctx, uint64_t *dev, size_t L, size_t start, phx_stream stream);
/* the same transforms over npolys polynomials of L limbs each (poly_stride elements apart) in one launch;
 * the reference's benchmark drives its NTT the same way over batch x limbs (benchmark/ntt_bench.cu:84-152) */
int phx_ntt_forward_batch(const phx_context *ctx, uint64_t *dev, size_t L, size_t start, size_t npolys,
                          size_t poly_stride, phx_stream stream);
int phx_ntt_inverse_batch(const phx_context *ctx, uint64_t *dev, size_t L, size_t start, size_t npolys,
                          size_t poly_stride, phx_stream stream);
/* DRNSTool::modup (include/rns.cuh:156-163, src/rns_bconv.cu:530-628):
 * c2 [Ql][N] NTT form -> t_mod_up [beta][QlP][N] */
int phx_modup(const phx_context *ctx, size_t chain_index, const uint64_t *c2_dev, uint64_t *t_mod_up_dev,
              phx_stream stream);
/* key_switch_inner_prod (include/evaluate.cuh:25-27): t_mod_up -> cx [2][QlP][N] */
int phx_key_inner_product(const phx_context *ctx, size_t chain_index, const uint64_t *t_mod_up_dev,
                          const phx_relin_key *key, uint64_t *cx_dev, phx_stream stream);
/* DRNSTool::moddown_from_NTT (src/rns_bconv.cu:776-828) on npolys polynomials of cx [npolys][QlP][N];
 * out [npolys][Ql][N]; cx's special limbs are overwritten (scratch) */
int phx_moddown_from_ntt(const phx_context *ctx, size_t chain_index, uint64_t *cx_dev, uint64_t *out_dev,
                         size_t npolys, phx_stream stream);

/* ---- device memory and timing helpers ------------------------------------- */
int phx_malloc(size_t bytes, void **out);
int phx_free(void *p);
int phx_memcpy_htod(void *dst, const void *src, size_t bytes, phx_stream stream);
int phx_memcpy_dtoh(void *dst, const void *src, size_t bytes, phx_stream stream);
int phx_stream_synchronize(phx_stream stream);
int phx_device_synchronize(void);
int phx_stream_create(phx_stream *out);
int phx_stream_destroy(phx_stream s);
int phx_event_create(void **out);
int phx_event_record(void *ev, phx_stream stream);
int phx_event_elapsed_ms(void *start, void *stop, float *ms);
int phx_event_destroy(void *ev);
int phx_device_count(int *n);
/* kernel-variant switches (no reference counterpart; used by tests and the autotuner):
 *   "bconv_mfma" 1 = FastBConv on the matrix cores (default), 0 = VALU kernel
 *   "ntt_fp64"   1 = exact FP64 butterflies for primes < 2^50 (default), 0 = 64-bit integer
 *                only; takes effect for contexts created afterwards
 *   "ntt_fused", "ntt_fused_min_limbs", "ntt_chunk_limbs", "ntt_cols_threads",
 *   "ntt_rows_threads", "bconv_wgs_per_cu", "bconv_fused": the process-wide defaults
 *                of the per-context variants below
 * unknown keys return PHX_EINVAL */
int phx_tune_set(const char *key, int value);
int phx_tune_get(const char *key, int *value);
/* Per-context kernel variants (no reference counterpart: the reference fixes its
 * launch shapes at compile time). choice[PHX_TUNE_AXES] = {ntt_fused,
 * ntt_fused_min_limbs, ntt_chunk_limbs, bconv_mfma, ntt_cols_threads,
 * ntt_rows_threads, bconv_wgs_per_cu, bconv_fused}; -1 = follow the process-wide switch of
 * that name.
 *   ntt_fused           0 = two launches per NTT; 1, 2, 4 = one workgroup of that
 *                       many x 256 threads per limb (N / 4096 a multiple of it)
 *   ntt_fused_min_limbs launches over fewer limbs keep two launches
 *   ntt_chunk_limbs     two-launch NTTs over more limbs go in slices of about this
 *                       many limbs (0 = one launch pair)
 *   bconv_mfma          1 = FastBConv on the matrix cores, 0 = VALU kernel
 *   ntt_cols_threads    tile shape of the column passes: 256 or 512 threads x 16
 *                       coefficients (16 or 32 columns at N = 2^16); 0 = 512 at
 *                       N = 2^17, and at N = 2^16 for launches of at least 256
 *                       limb-transforms, else 256
 *   ntt_rows_threads    tile shape of the row passes: 128, 256 or 512 threads
 *                       (8, 16 or 32 rows of 256 coefficients)
 *   bconv_wgs_per_cu    matrix-core FastBConv workgroups per CU: 1, 2, 4 (0 = as many
 *                       as the LDS holds)
 *   bconv_fused         1 = the matrix-core FastBConv of modup and moddown fused with the
 *                       column pass of the forward NTT of its outputs (N = 2^12 .. 2^16,
 *                       special_modulus_size up to 16), 0 = separate launches */
#define PHX_TUNE_AXES 8
typedef struct {
    int choice[PHX_TUNE_AXES];   /* the variants applied */
    double default_us;   /* objective with every switch at -1: one HMult+relin at the top level plus
                            one op of an 8-pair batch plan, microseconds */
    double best_us;      /* the same with `choice` */
    uint32_t trials;     /* candidates timed */
    char trials_json[8192];   /* every candidate and its objective: [{"choice": "...", "us": x}, ...] */
} phx_tune_report;
int phx_context_get_tuning(const phx_context *ctx, int choice[PHX_TUNE_AXES]);
int phx_context_set_tuning(phx_context *ctx, const int choice[PHX_TUNE_AXES]);
/* Autotuner: times HMult+relin at the top level, plain and batched (synthetic
 * residues in HBM), for each candidate (NTT launch form, NTT tile shape, FastBConv
 * kind and occupancy), applies the fastest (a candidate must beat the incumbent by
 * 2%) to ctx and, when cache_path is non-empty, stores it in that JSON cache under
 * (scheme, N, size_Q, size_P). Contexts created while the environment variable
 * PHX_TUNE_CACHE names such a file apply its entry. */
int phx_autotune(phx_context *ctx, const char *cache_path, phx_tune_report *report, phx_stream stream);
/* applies the cache entry for ctx's shape if there is one (*found = 0 or 1) */
int phx_tune_load(phx_context *ctx, const char *cache_path, int *found);
/* host-only cache access (no GPU needed) */
int phx_tune_cache_store(const char *cache_path, int scheme, uint64_t N, size_t size_Q, size_t size_P,
                         const int choice[PHX_TUNE_AXES], double hmult_relin_us);
int phx_tune_cache_lookup(const char *cache_path, int scheme, uint64_t N, size_t size_Q, size_t size_P,
                          int choice[PHX_TUNE_AXES], int *found);
/* streaming copy kernel of n u64 elements (no reference counterpart): the known-size
 * copy that HBM counters and bandwidth figures are calibrated against */
int phx_copy_u64(uint64_t *dst, const uint64_t *src, size_t n, phx_stream stream);

#ifdef __cplusplus
}
#endif
#endif /* PHANTOM_AMD_H */
