/*
 * oracle.h -- CPU restatement of PhantomFHE's RNS polynomial hot path.
 *
 * TEST INFRASTRUCTURE ONLY. This library is the checker: tests/, the smoke
 * entry and the benchmark's CPU-baseline leg load it; the product library
 * never links, loads or calls it.
 *
 * Every routine restates the reference algorithm step for step, with exact
 * 128-bit integer arithmetic, and cites the reference file:line it follows.
 * All outputs are canonical residues in [0, q).
 *
 * Pinning (tests/test_oracle*.py): every routine is checked against closed
 * forms, known answers or the specification's vectors. The FP64 sums of the
 * HPS multiply (bConv_HPS's v, the scale-and-round nu of QR -> R, QlRl -> Ql
 * and Q -> Ql) are pinned bit for bit by an exact-rational model of fused
 * multiply-adds in input order (tests/test_oracle_hps_fp64.py), which holds if
 * nvcc contracts the reference's `nu += (double) x * frac` that way (its
 * default); so is the HPS scale-and-round of BFV decryption, in all four of its
 * forms (tests/test_oracle_bfv_decrypt.py, on phases at rounding ties).
 * Not pinned:
 *   - orc_hps_levels_to_drop (FindLevelsToDrop): invariants only, its value is
 *     parity unpinned;
 *   - orc_sample_uniform_seeded: the Salsa20 core is pinned by the
 *     specification's vectors, the seed/nonce layout, rejection bound and
 *     redraw nonce are the reference's own arrangement: parity unpinned.
 *
 * Layouts (identical to the reference):
 *   polynomial limb set  [limb][N]            (u64)
 *   ciphertext           [poly][limb][N]
 *   relin / Galois key   [digit][2][QP][N]    (b part, then a part; NTT form)
 *   t_mod_up             [digit][QlP][N]
 */
#ifndef PHANTOM_ORACLE_H
#define PHANTOM_ORACLE_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct orc_ctx orc_ctx;

/* thread count used by the limb-parallel loops (default 1) */
void orc_set_threads(int n);
int orc_get_threads(void);

/* ---- host parameter maths ---- */
int orc_is_prime(uint64_t v);
int orc_get_primes(uint64_t N, int bits, size_t count, uint64_t *out);
int orc_coeff_modulus_create(uint64_t N, const int *bits, size_t n, uint64_t *out);
uint64_t orc_minimal_root(uint64_t q, uint64_t degree);
uint64_t orc_shoup(uint64_t w, uint64_t q);
void orc_ntt_tables(uint64_t q, int logN, uint64_t *rp, uint64_t *rps,
                    uint64_t *irp, uint64_t *irps, uint64_t *ninv);

/* ---- context (QP primes: size_Q data primes, then size_P special primes) ---- */
orc_ctx *orc_ctx_create(int logN, const uint64_t *moduli, size_t size_QP, size_t size_P);
void orc_ctx_destroy(orc_ctx *c);
const uint64_t *orc_ctx_root_powers(const orc_ctx *c, size_t prime_idx);
const uint64_t *orc_ctx_inv_root_powers(const orc_ctx *c, size_t prime_idx);

/* ---- transforms on limbs [start, start+L) of the QP chain ---- */
void orc_ntt_forward(const orc_ctx *c, uint64_t *data, size_t L, size_t start);
void orc_ntt_inverse(const orc_ctx *c, uint64_t *data, size_t L, size_t start);

/* ---- RNS base conversion ---- */
int orc_fastbconv(size_t N, const uint64_t *ibase, size_t ni, const uint64_t *obase, size_t no,
                  const uint64_t *in, uint64_t *out);

/* ---- hybrid key switching (CKKS, NTT domain), level with size_Ql data limbs ---- */
size_t orc_beta(const orc_ctx *c, size_t size_Ql);
int orc_modup(const orc_ctx *c, size_t size_Ql, const uint64_t *c2, uint64_t *t_mod_up);
int orc_inner_product(const orc_ctx *c, size_t size_Ql, const uint64_t *t_mod_up,
                      const uint64_t *key, size_t dnum, uint64_t *cx);
int orc_moddown(const orc_ctx *c, size_t size_Ql, const uint64_t *cx_i, uint64_t *out);
int orc_keyswitch(const orc_ctx *c, size_t size_Ql, const uint64_t *c2, const uint64_t *key,
                  size_t dnum, uint64_t *out2);

/* ---- evaluator ---- */
void orc_tensor(const orc_ctx *c, size_t size_Ql, const uint64_t *ct1, const uint64_t *ct2,
                uint64_t *out3);
int orc_multiply_and_relin(const orc_ctx *c, size_t size_Ql, const uint64_t *ct1,
                           const uint64_t *ct2, const uint64_t *key, size_t dnum, uint64_t *out2);
int orc_rescale(const orc_ctx *c, size_t size_Ql, const uint64_t *ct, size_t ct_size,
                uint64_t *out);
uint32_t orc_galois_elt_from_step(int step, size_t N);
void orc_galois_perm_table(int logN, uint32_t elt, uint32_t *out);
int orc_apply_galois(const orc_ctx *c, size_t size_Ql, const uint64_t *ct, uint32_t elt,
                     const uint64_t *key, size_t dnum, uint64_t *out2);

/* hoisting_inplace (CKKS): sum of the rotations of ct by elts, one shared modup;
 * keys [n][dnum][2][QP][N] in the order of elts */
int orc_hoisting(const orc_ctx *c, size_t size_Ql, const uint64_t *ct, const uint32_t *elts, size_t n,
                 const uint64_t *keys, size_t dnum, uint64_t *out2);
/* the same for BFV (ct and out2 in coefficient form) */
int orc_hoisting_bfv(const orc_ctx *c, size_t size_Ql, const uint64_t *ct, const uint32_t *elts, size_t n,
                     const uint64_t *keys, size_t dnum, uint64_t *out2);

/* key material from explicit samples: s_ntt, s2_ntt [QP][N] (NTT), a [dnum][QP][N] (NTT),
 * e [dnum][QP][N] (coefficient domain, canonical residues) */
int orc_make_kswitch_key(const orc_ctx *c, const uint64_t *s_ntt, const uint64_t *s2_ntt,
                         const uint64_t *a, const uint64_t *e, size_t dnum, uint64_t *key);

/* ---- BFV (coefficient form): BEHZ multiply at the top data level, relinearisation ---- */
typedef struct orc_behz orc_behz;
orc_behz *orc_behz_create(const orc_ctx *c, uint64_t t);
void orc_behz_destroy(orc_behz *b);
size_t orc_behz_bsk_size(const orc_behz *b);
void orc_behz_bsk(const orc_behz *b, uint64_t *out);   /* base B, then m_sk */
/* ct1, ct2 [2][size_Q][N] -> out3 [3][size_Q][N] */
int orc_bfv_multiply(const orc_ctx *c, const orc_behz *b, const uint64_t *ct1, const uint64_t *ct2,
                     uint64_t *out3);
/* ct3 [3][size_Q][N] -> out2 [2][size_Q][N] */
int orc_bfv_relinearize(const orc_ctx *c, const uint64_t *ct3, const uint64_t *key, size_t dnum,
                        uint64_t *out2);

/* ---- BFV HPS multiply (mul_tech hps) at the top data level ---- */
typedef struct orc_hps orc_hps;
orc_hps *orc_hps_create(const orc_ctx *c, uint64_t t);
/* overq != 0: mul_tech hps_overq (R = size_Q primes, t/R scale-and-round QR -> Q) */
orc_hps *orc_hps_create2(const orc_ctx *c, uint64_t t, int overq);
/* bConv_BEHZ_var1 Q -> R of one polynomial [k][N] -> [nr][N] (over-Q tool) */
void orc_hps_modswitch(const orc_ctx *c, const orc_hps *h, const uint64_t *x, uint64_t *out);
void orc_hps_destroy(orc_hps *h);
size_t orc_hps_r_size(const orc_hps *h);
void orc_hps_r(const orc_hps *h, uint64_t *out);
/* square != 0: the reference's squaring path (one operand, extended as the first) */
int orc_bfv_multiply_hps2(const orc_ctx *c, const orc_hps *h, const uint64_t *ct1, const uint64_t *ct2, int square,
                          uint64_t *out3);
int orc_bfv_multiply_hps(const orc_ctx *c, const orc_hps *h, const uint64_t *ct1, const uint64_t *ct2,
                         uint64_t *out3);

/* The FP64 sums of the HPS path and the steps that use them, exported so that the
 * tests can check every sum against an exact-rational model of sequential fused
 * multiply-adds. orc_fma_chain: out[n] = fma(x[m-1][n], w[m-1], ... fma(x[0][n], w[0],
 * init)) over nterms terms, x [nterms][N]. orc_bconv_hps: bConv_HPS of [ni][N] -> [no][N]
 * with its sum (llround'ed to v) per coefficient. orc_hps_scale_round: the scale-and-round
 * step of one polynomial [k + nr][N] (QR -> R into [nr][N] for hps, QlRl -> Ql into [k][N]
 * for hps_overq and levels) with nu per coefficient. orc_hps_frac: the frac constants
 * (tRSHatInvModsDivsFrac [k]; tQlSlHatInvModsDivsFrac [nr] for over-Q tools). */
void orc_fma_chain(const uint64_t *x, size_t nterms, size_t N, const double *w, double init, double *out);
void orc_bconv_hps(size_t N, const uint64_t *ib, size_t ni, const uint64_t *ob, size_t no, const uint64_t *in,
                   uint64_t *out, double *sum);
void orc_hps_scale_round(const orc_ctx *c, const orc_hps *h, const uint64_t *d, uint64_t *out, double *nu);
void orc_hps_frac(const orc_hps *h, double *frac_q, double *frac_r);

/* hps_overq_leveled: the HPS tool of the level with d dropped primes (Ql = size_Q - d,
 * Rl = the first Ql primes of R, the second operand switched from all of Q);
 * FindLevelsToDrop; scaleAndRound_HPS_Q_Ql ([size_Q][N] -> [Ql][N]); ExpandCRTBasis_Ql_Q
 * ([Ql][N] -> [size_Q][N]); the multiply with d dropped levels */
orc_hps *orc_hps_create_level(const orc_ctx *c, uint64_t t, size_t d);
size_t orc_hps_levels_to_drop(const orc_ctx *c, uint64_t t, size_t depth, int key_switch, int asymmetric);
void orc_hps_scale_q_ql(const orc_ctx *c, size_t d, const uint64_t *x, uint64_t *out);
/* the same with nu (the FP64 sum of scaleAndRound_HPS_Q_Ql) per coefficient */
void orc_hps_scale_q_ql_nu(const orc_ctx *c, size_t d, const uint64_t *x, uint64_t *out, double *nu);
void orc_hps_expand(const orc_ctx *c, size_t d, const uint64_t *x, uint64_t *out);
int orc_bfv_multiply_hps_level(const orc_ctx *c, const orc_hps *h, size_t d, const uint64_t *ct1,
                               const uint64_t *ct2, int square, size_t expand_polys, uint64_t *out3);

/* the reference's seeded uniform sampler (sample_uniform_poly, src/prng.cu:174-210)
 * and its block function (the Salsa20 hash, src/prng.cu:17-131) */
void orc_salsa_quarterround(uint32_t y[4]);
void orc_salsa_hash(const uint8_t in[64], uint8_t out[64]);
void orc_sample_uniform_seeded(const orc_ctx *c, const uint8_t seed[64], size_t L, uint64_t *out);

/* BFV decryption scale-and-round of a coefficient-form phase [size_Ql][N] (BEHZ or HPS) */
int orc_bfv_decrypt(const orc_ctx *c, size_t size_Ql, uint64_t t, int behz, const uint64_t *phase, uint64_t *out);

#ifdef __cplusplus
}
#endif
#endif
