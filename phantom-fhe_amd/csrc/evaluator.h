// evaluator.h -- ciphertext/key types and the hot-path operations.
#pragma once

#include <vector>

#include "context.h"
#include "ntt.h"

namespace phx {

// Ciphertext: device buffer [size][coeff_modulus_size][N] (the reference's
// PhantomCiphertext layout, include/ciphertext.h:15-25).
struct Ciphertext {
    u32 chain_index = 0;
    u32 size = 0;
    u64 poly_degree = 0;
    u32 coeff_modulus_size = 0;
    double scale = 1.0;
    u64 correction_factor = 1;
    u32 noise_scale_deg = 1;
    bool is_ntt_form = true;
    bool is_asymmetric = false;
    u64 *data = nullptr;
    size_t capacity = 0;  // elements
    bool owns = true;
    // symmetric encryption from a seed: c1 is the seeded sampler's output
    // (PhantomCiphertext::seed_, include/ciphertext.h:25)
    bool has_seed = false;
    uint8_t seed[64] = {};

    void resize(const Context &ctx, u32 chain, u32 sz, hipStream_t s);
    void release(hipStream_t s);
};

// Key-switching key: [dnum][2][QP][N] (digit j = (b_j, a_j), NTT form),
// the reference's PhantomRelinKey public_keys_ layout (src/secretkey.cu:297-341)
// with dnum = ceil(size_Q / size_P).
struct KSwitchKey {
    u32 dnum = 0;
    u64 *data = nullptr;
    u64 poly_degree = 0;   // the context's N and QP limb count at creation
    u32 size_QP = 0;
};

struct GaloisKeys {
    std::vector<u32> elts;
    std::vector<KSwitchKey> keys;
};

struct SecretKey {
    u64 *s = nullptr;    // [QP][N] NTT form
    u64 *s2 = nullptr;   // s^2, [QP][N] NTT form (lazily computed)
};

// One grid row of a work-list FastBConv launch: the descriptor of the row's
// conversion (its level's), the row's input limb set and output limb set.
struct BconvRow {
    const BconvDesc *desc;
    const u64 *in;
    u64 *out;
    const BconvDesc *fdesc;   // the same conversion's fused form (bconv_ntt.hip)
};

// FastBConv launch arguments. Grid row y = batch * jobs_per_batch + job; job j
// uses descs[j * desc_stride]; pointers advance by job and batch strides. With
// rows set, grid row y is rows[y] instead (copy_src unused).
struct BconvArgs {
    const BconvRow *rows;
    const u64 *cbuf;
    const BconvDesc *descs;
    u32 desc_stride;
    u32 jobs_per_batch;
    const u64 *in;
    u64 in_job_stride, in_batch_stride;
    u64 *out;
    u64 out_job_stride, out_batch_stride;
    const u64 *copy_src;   // modup: the digit's own limbs are copied from here (or null)
    u64 copy_job_stride, copy_batch_stride;
    const DevModulus *mod;
    u32 N;
};
// runs FastBConv over `jobs` grid rows: matrix cores when every descriptor has
// an MFMA form (and the bconv_mfma switch is on), the VALU kernel otherwise
void bconv_run(const Context &ctx, const std::vector<BconvDesc> &descs_host, const BconvArgs &args, u32 jobs,
               hipStream_t s);
// FastBConv fused with the column pass of the forward NTT of its outputs
// (bconv_ntt.hip): fdescs are the fused descriptors the launch uses (args.descs /
// args.rows[].fdesc); the outputs are left as the column pass leaves them, for the
// row pass (ntt_forward_rows*). bconv_fused_ok says whether the fused launch serves
// these conversions on this context; copy_src is not supported.
bool bconv_fused_ok(const Context &ctx, const std::vector<BconvDesc> &fdescs);
void bconv_fwd_cols_run(const Context &ctx, const std::vector<BconvDesc> &fdescs, const BconvArgs &args, u32 jobs,
                        hipStream_t s);

// ---- hot path ----
void tensor_2x2(const Context &ctx, u32 Ql, const u64 *a, const u64 *b, u64 *d0, u64 *d1, u64 *d2, hipStream_t s);
// the same over L limbs whose moduli start at mod
void tensor_2x2_mod(const DevModulus *mod, u64 N, u32 L, const u64 *a, const u64 *b, u64 *d0, u64 *d1, u64 *d2,
                    hipStream_t s);
// ct[0..1] (Ql limbs each) = base0/base1 + keyswitch(c2); base may be null (zero)
void keyswitch(const Context &ctx, u32 chain, const u64 *c2, const KSwitchKey &key, u64 *ct,
               const u64 *base0, const u64 *base1, hipStream_t s);
void multiply_and_relin(const Context &ctx, Ciphertext &ct1, const Ciphertext &ct2, const KSwitchKey &rlk,
                        hipStream_t s);
void relinearize(const Context &ctx, Ciphertext &ct, const KSwitchKey &rlk, hipStream_t s);
void rescale_to_next(const Context &ctx, const Ciphertext &in, Ciphertext &out, hipStream_t s);
// mod_switch_to_next (src/evaluate.cu:1505-1540): CKKS drops the last limb
// (mod_switch_drop_to_next, 1429-1472), BFV divides by it (mod_switch_scale_to_next,
// 1376-1427, coefficient form); in != out
void mod_switch_to_next(const Context &ctx, const Ciphertext &in, Ciphertext &out, hipStream_t s);
void apply_galois(const Context &ctx, Ciphertext &ct, u32 elt, const GaloisKeys &gk, hipStream_t s);
void rotate(const Context &ctx, Ciphertext &ct, int step, const GaloisKeys &gk, hipStream_t s);
// hoisting_inplace: ct <- sum_i rot(ct, steps[i]) with one shared modup
// (CKKS in NTT form, BFV in coefficient form)
void hoisting(const Context &ctx, Ciphertext &ct, const GaloisKeys &gk, const std::vector<int> &steps,
              hipStream_t s);
u32 galois_elt_from_step(int step, u64 N);

// the same operations in a caller's workspace (device memory of *_ws_words u64
// words at the ciphertext's chain index): nothing is allocated
size_t keyswitch_ws_words(const Context &ctx, u32 chain);
// tensor (optional): the modup INTT's first pass loads the tensor product and writes c2
void keyswitch_ws(const Context &ctx, u32 chain, const u64 *c2, const KSwitchKey &key, u64 *ct,
                  const u64 *base0, const u64 *base1, u64 *ws, hipStream_t s, const LoadTensor *tensor = nullptr);
size_t multiply_and_relin_ws_words(const Context &ctx, u32 chain);
// CKKS only
void multiply_and_relin_ws(const Context &ctx, Ciphertext &ct1, const Ciphertext &ct2, const KSwitchKey &rlk,
                           u64 *ws, hipStream_t s);
// multiply_and_relin_ws in the library's static device scratch (no allocation);
// sizes beyond it are refused
void multiply_and_relin_noalloc(const Context &ctx, Ciphertext &ct1, const Ciphertext &ct2, const KSwitchKey &rlk,
                                hipStream_t s);

// ---- pieces exposed for tests / the low-level API ----
// copy_own = false: the digits' own limbs of t_mod_up are left unwritten
void modup(const Context &ctx, u32 chain, const u64 *c2, u64 *t_mod_up, hipStream_t s, bool copy_own = true);
void modup_ws(const Context &ctx, u32 chain, const u64 *c2, u64 *t_mod_up, u64 *t_cks, hipStream_t s,
              bool copy_own, const LoadTensor *tensor = nullptr);
// c2 given: the digits' own limbs are read from c2 (NTT form) instead of t_mod_up
void inner_product(const Context &ctx, u32 chain, const u64 *t_mod_up, const KSwitchKey &key, u64 *cx,
                   hipStream_t s, const u64 *c2 = nullptr);
void moddown(const Context &ctx, u32 chain, u64 *cx, u64 *ct, const u64 *base0, const u64 *base1, u32 npolys,
             hipStream_t s);
void moddown_ws(const Context &ctx, u32 chain, u64 *cx, u64 *ct, const u64 *base0, const u64 *base1, u32 npolys,
                u64 *delta, hipStream_t s);

// streaming copy (bandwidth calibration)
void copy_u64(u64 *dst, const u64 *src, size_t n, hipStream_t s);

// ---- keys and (test-side) encryption ----
void secret_key_generate(const Context &ctx, SecretKey &sk, u64 seed, hipStream_t s);
void secret_key_upload(const Context &ctx, SecretKey &sk, const int8_t *coeffs, hipStream_t s);
void kswitch_key_generate(const Context &ctx, const SecretKey &sk, const u64 *new_key, KSwitchKey &key, u64 seed,
                          hipStream_t s);
void relin_key_generate(const Context &ctx, SecretKey &sk, KSwitchKey &key, u64 seed, hipStream_t s);
void galois_key_generate(const Context &ctx, const SecretKey &sk, u32 elt, KSwitchKey &key, u64 seed,
                         hipStream_t s);
// plain: [Ql][N] coefficient-domain residues at chain index `chain`; result NTT form
void encrypt_symmetric(const Context &ctx, const SecretKey &sk, const u64 *plain_dev, u32 chain, Ciphertext &ct,
                       u64 seed, hipStream_t s);
// the same with c1 = the seeded uniform sampler's output for seed_a (64 bytes), as
// encrypt_symmetric / encrypt_zero_symmetric do (src/secretkey.cu:232-300, 463-500);
// the ciphertext keeps the seed (save_symmetric). e comes from the counter hash (seed_e).
void encrypt_symmetric_seeded(const Context &ctx, const SecretKey &sk, const u64 *plain_dev, u32 chain,
                              Ciphertext &ct, const uint8_t *seed_a, u64 seed_e, hipStream_t s);
// BFV: the plaintext (N coefficients) of a coefficient-form ciphertext at any data
// level (PhantomSecretKey::bfv_decrypt, src/secretkey.cu:571-635)
void bfv_decrypt(const Context &ctx, SecretKey &sk, const Ciphertext &ct, u64 *out_dev, hipStream_t s);
// out: [Ql][N] coefficient-domain residues of c0 + c1 s (+ c2 s^2)
void decrypt(const Context &ctx, SecretKey &sk, const Ciphertext &ct, u64 *out_dev, hipStream_t s);

// gen_publickey (src/secretkey.cu:382-394): pk = (-(a s + e), a) at the key
// level (chain 0, QP limbs, NTT form); a from the seeded uniform sampler for
// seed_a (64 bytes) or, with seed_a null, from the counter hash; e from seed_e.
void public_key_generate(const Context &ctx, const SecretKey &sk, const uint8_t *seed_a, u64 seed_e,
                         Ciphertext &pk, hipStream_t s);
// PhantomPublicKey::encrypt_asymmetric (src/secretkey.cu:11-186) at the first
// data level: c_j = moddown(pk_j u + e_j) over QP -> Q, then c0 += the plaintext
// (CKKS: plain_dev = coefficient residues [size_Q][N], added in NTT form; BFV:
// plain_dev = N coefficients mod t, added as floor(Q m / t)). u ternary and
// e_0, e_1 come from the int8 samples u_dev / e_dev[2] (device, N each) when
// given, else from the counter hash with `seed`.
void encrypt_asymmetric(const Context &ctx, const Ciphertext &pk, const u64 *plain_dev, u32 chain,
                        const int8_t *u_dev, const int8_t *const e_dev[2], u64 seed, Ciphertext &ct, hipStream_t s);

}  // namespace phx
