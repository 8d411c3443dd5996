// context.h -- the device-resident context: modulus chain, NTT tables and
// per-level hybrid key-switching constants.
//
// Chain semantics follow the reference's PhantomContext (src/context.cu:121-232):
// chain index 0 is the key level (all QP primes), chain index c >= 1 is a data
// level holding the first size_Q - (c - 1) data primes.
#pragma once

#include <memory>
#include <vector>

#include "common.h"
#include "hostmath.h"
#include "ntt.h"

namespace phx {

enum SchemeType : int { kSchemeNone = 0, kSchemeBGV = 1, kSchemeBFV = 2, kSchemeCKKS = 3 };
enum MulTech : int { kMulNone = 0, kMulBEHZ = 1, kMulHPS = 2, kMulHPSOverQ = 3, kMulHPSOverQLeveled = 4 };

struct Params {
    int scheme = kSchemeCKKS;
    int mul_tech = kMulNone;
    u64 poly_degree = 0;
    std::vector<u64> coeff_modulus;   // size_Q data primes then size_P special primes
    u64 special_modulus_size = 1;
    u64 plain_modulus = 0;
    std::vector<u32> galois_elts;
};

// Base conversion of one digit / one polynomial (FastBConv phase 2):
// out[out_pos[o]] = (sum_i in[i] * mat[o][i]) mod omod[o]
struct BconvDesc {
    u32 ni, no;          // input limbs, output limbs
    u32 in_first;        // first input limb (within the input limb set)
    u32 own_first;       // modup: first output limb that copies the digit's own limbs
    u64 mat_off;         // offset (u64 units) of the [no][ni] matrix in the constant buffer
    u64 pos_off;         // offset (u32 units) of out_pos[no]
    u64 row_off;         // offset (u32 units) of the table rows of the outputs
    // MFMA form (ni <= 16): int8 operand tiles of the sliced matrix, see bconv_mfma
    u64 a_off;           // offset (u64 units) of rt x 2 x 64 lanes x 16 bytes, then no x (lo, fl) bias
    u64 pinv_off;        // offset (u64 units) of fl(1/p_o) as double bits, [no]
    u32 rt;              // 16-row operand tiles: 2 * ceil(no / 4); 0 = no MFMA form
    u32 pad;
};

// MFMA operand blob of one FastBConv matrix (host side). Row 4g+i of tile
// 2T+s stands for output o = 4T+g, digit b = 4s+i (b < 8), so the 8 partial
// sums of an output end up in one lane's accumulators; K = (i, a) = i*8 + a
// indexes byte a of input i; A[row][K] = digit b of the balanced base-256
// expansion (digits in [-128, 127]) of 2^(8a) * mat[o][i] mod p_o. Lane l of tile t, half h holds bytes
// A[16t + l%16][64h + 16(l/16) + j], j < 16. After the tiles: per output the
// bias C_o that restores the unsigned input bytes, as (C_o mod 2^64, fl(C_o)).
std::vector<uint8_t> bconv_mfma_operands(const std::vector<u64> &mat, u32 ni, u32 no,
                                         const std::vector<u64> &omod, u32 &rt);

struct LevelConsts {
    u32 chain_index = 0;
    u32 size_Ql = 0, size_QlP = 0, beta = 0;
    // device pointers into the context constant buffer
    const u64 *modup_scale = nullptr;     // [Ql][4] INTT final-stage constants x partQlHatInv
    const BconvDesc *modup_bconv = nullptr;  // [beta]
    const u64 *moddown_scale = nullptr;   // [P][4] INTT final-stage constants x PHatInv
    const BconvDesc *moddown_bconv = nullptr;  // [1]
    const u64 *pinv = nullptr;            // [Ql][2] P^-1 mod q_i (Shoup)
    const u64 *rescale_qinv = nullptr;    // [Ql-1][2] q_last^-1 mod q_j (Shoup)
    std::vector<BconvDesc> h_modup_bconv; // host copies
    BconvDesc h_moddown_bconv{};
    // the same conversions for the launch fused with the forward NTT's column pass
    // (outputs grouped by butterfly arithmetic, see bconv_ntt.hip; rt = 0: none)
    const BconvDesc *modup_fbconv = nullptr;   // [beta]
    const BconvDesc *moddown_fbconv = nullptr; // [1]
    std::vector<BconvDesc> h_modup_fbconv;
    BconvDesc h_moddown_fbconv{};
};

// BEHZ constants of the top data level (BFV multiply; reference DRNSTool ctor,
// src/rns.cu:392-571): q = the size_Q data primes, Bsk = B (61-bit primes) then
// m_sk, m_tilde = 2^32. Device arrays live in the context constant buffer; the
// Bsk primes' NTT tables are table rows bsk_row0 .. bsk_row0 + nb - 1.
struct BehzConsts {
    u32 k = 0, nb = 0;       // |q|, |Bsk|
    u32 bsk_row0 = 0;
    u32 pad = 0;
    u64 t = 0;
    u64 neg_qinv_mt = 0;     // -Q^-1 mod 2^32
    u64 binv_msk = 0;        // B^-1 mod m_sk
    const u64 *lift_f;       // [k]      m_tilde * Qhat_i^-1 mod q_i
    const u64 *qhat_bsk;     // [nb][k]  Qhat_i mod p
    const u64 *qhat_mt;      // [k]      Qhat_i mod 2^32
    const u64 *q_mod_bsk;    // [nb]     Q mod p
    const u64 *mtinv_bsk;    // [nb]     m_tilde^-1 mod p
    const u64 *qhatinv_q;    // [k]      Qhat_i^-1 mod q_i
    const u64 *qinv_bsk;     // [nb]     Q^-1 mod p
    const u64 *bhatinv_b;    // [nb-1]   Bhat_j^-1 mod b_j
    const u64 *bhat_msk;     // [nb-1]   Bhat_j mod m_sk
    const u64 *bhat_q;       // [k][nb-1] Bhat_j mod q_i
    const u64 *b_mod_q;      // [k]      B mod q_i
    const u64 *tscale_q;     // [k][4]   final INTT stage constants times t
    const u64 *tscale_bsk;   // [nb][4]
};

// BFV decryption scale-and-round at one data level, the reference's
// choice of method (BEHZ for mul_tech behz; HPS otherwise, in its four forms).
enum BfvDecKind { kBfvDecNone = 0, kBfvDecBEHZ, kBfvDecHpsSmallLazy, kBfvDecHpsSmall, kBfvDecHpsLargeLazy,
                  kBfvDecHpsLarge };
struct BfvDecrypt {
    int kind = kBfvDecNone;
    u32 qmsb_half = 0;    // HPS large forms: x = x_hi 2^h + x_lo with h = qMSB / 2
    // per q_i, 8 words: BEHZ {[t gamma]_q, Shoup, Qhat^-1 mod q, Shoup, Qhat mod t, Qhat mod gamma, 0, 0};
    // HPS {[t Qhat^-1 / q]_t, Shoup, frac (FP64 bits), the same for Qhat^-1 2^h, 0, 0}
    const u64 *tab = nullptr;
    u64 t = 0, t_ratio_lo = 0, t_ratio_hi = 0;
    u64 gamma = 0, g_ratio_lo = 0, g_ratio_hi = 0;
    u64 neg_qinv_t = 0, neg_qinv_t_shoup = 0, neg_qinv_g = 0, neg_qinv_g_shoup = 0;
    u64 inv_gamma_t = 0, inv_gamma_t_shoup = 0;
};

// HPS constants of the top data level (BFV multiply with mul_tech hps;
// reference DRNSTool ctor, src/rns.cu:686-783, and the BaseConverter tables of
// src/host/rns.cu:438-497): R = size_Q + 1 primes below the smallest q, as
// table rows r_row0 ..; mod_qr is the moduli of Q then R, contiguous.
struct HpsConsts {
    u32 k = 0, nr = 0;
    u32 r_row0 = 0;
    u32 pad = 0;
    const DevModulus *mod_qr;   // [k + nr]
    const u64 *qhatinv_q;       // [k]        Qhat_i^-1 mod q_i
    const u64 *qhat_r;          // [nr][k]    Qhat_i mod r_j
    const double *qinv;         // [k]        1 / q_i
    const u64 *alphaq_r;        // [k+1][nr]  v Q mod r_j
    const double *frac;         // [k]        tRSHatInvModsDivsFrac
    const u64 *divs_mod_r;      // [nr][k+1]  tRSHatInvModsDivsModr
    const u64 *rhatinv_r;       // [nr]       Rhat_j^-1 mod r_j
    const u64 *rhat_q;          // [k][nr]    Rhat_j mod q_i
    const double *rinv;         // [nr]       1 / r_j
    const u64 *alphar_q;        // [nr+1][k]  v R mod q_i
    // over Q (mul_tech hps_overq / hps_overq_leveled; nr = k): switch of the
    // second operand from its ks source moduli (Q) to R, t/R scale-and-round QR -> Q
    u32 overq = 0;
    u32 ks = 0;                 // source moduli of the switch (k, or all of Q below the top level)
    const DevModulus *mod_src;  // [ks]
    const u64 *negrq;           // [ks]       s_i - [R Shat_i^-1]_{s_i}
    const u64 *qinv_r;          // [nr][ks]   s_i^-1 mod r_j
    const double *frac_r;       // [nr]       tQlSlHatInvModsDivsFrac
    const u64 *divs_mod_q;      // [k][nr+1]  tQlSlHatInvModsDivsModq
};

// Kernel variants chosen for one context (by the autotuner or a tuning cache);
// -1 = follow the process-wide switch of the same name.
struct TuneChoice {
    int ntt_fused = -1;          // 0: two-pass NTT; G in {1, 2, 4}: one workgroup of G x 256 per limb
    int ntt_fused_min_limbs = -1;
    int ntt_chunk_limbs = -1;    // limbs per two-pass launch pair (0 = one launch)
    int bconv_mfma = -1;         // 1: FastBConv on the matrix cores, 0: VALU
    int ntt_cols_threads = -1;   // tile shape of the column passes: 256 or 512 threads, 0 by launch size
    int ntt_rows_threads = -1;   // tile shape of the row passes: 128, 256 or 512 threads
    int bconv_wgs_per_cu = -1;   // FastBConv (matrix cores) workgroups per CU: 1, 2, 4 (0 = as LDS allows)
    int bconv_fused = -1;        // 1: FastBConv fused with the forward NTT's column pass, 0: separate
};

class Context {
public:
    explicit Context(const Params &p);
    ~Context();
    Context(const Context &) = delete;
    Context &operator=(const Context &) = delete;

    const Params &params() const { return params_; }
    u64 N() const { return N_; }
    u32 logN() const { return logN_; }
    u32 size_Q() const { return size_Q_; }
    u32 size_P() const { return size_P_; }
    u32 size_QP() const { return size_QP_; }
    u32 dnum() const { return (size_Q_ + size_P_ - 1) / size_P_; }
    u32 total_levels() const { return size_Q_ + 1; }
    u32 limbs_at(u32 chain_index) const { return chain_index == 0 ? size_QP_ : size_Q_ - (chain_index - 1); }
    const std::vector<u64> &moduli() const { return params_.coeff_modulus; }
    const NttTables &tables() const { return tables_; }
    const LevelConsts &level(u32 chain_index) const;
    const DevModulus *dev_mod() const { return tables_.mod; }
    const u64 *const_buf() const { return cbuf_; }
    const u32 *const_buf32() const { return reinterpret_cast<const u32 *>(cbuf_); }
    // P mod q_i for i < size_Q (keygen gadget factor), [size_Q][2] with Shoup
    const u64 *bigP_mod_q() const { return bigP_mod_q_; }
    // BFV plaintext scaling: [-Q]_t (and its Shoup companion mod t) and
    // t^-1 mod q_i, [size_Q][2] with Shoup (null unless BFV with t set)
    u64 bfv_negq_mod_t() const { return bfv_negq_mod_t_; }
    u64 bfv_negq_mod_t_shoup() const { return bfv_negq_mod_t_shoup_; }
    const u64 *bfv_tinv_q() const { return bfv_tinv_q_; }
    // per chain index 1 .. size_Q (kind kBfvDecNone unless BFV with t set)
    const BfvDecrypt &bfv_decrypt_consts(u32 chain) const { return bfv_dec_[chain]; }
    const hm::NttTable &host_table(u32 row) const { return host_tables_[row]; }

    // BFV with BEHZ multiplication: its constants (else null)
    const BehzConsts *behz() const { return behz_.k ? &behz_ : nullptr; }
    const std::vector<u64> &bsk_primes() const { return bsk_; }
    // BFV with HPS multiplication: its constants (else null)
    const HpsConsts *hps() const { return hps_.k ? &hps_ : nullptr; }
    const std::vector<u64> &hps_r_primes() const { return hps_r_; }
    // hps_overq_leveled with d dropped levels (d = 0 is hps()): the tables for
    // Ql = size_Q - d and Rl = the first Ql primes of R; the Q -> Ql
    // scale-and-round (k = Ql, nr = d, the dropped moduli as the "R" part); the
    // expansion factors [Ql][2] = (Q/Ql mod q_i, Shoup)
    const HpsConsts &hps_level(u32 d) const;
    const HpsConsts &hps_drop(u32 d) const;
    const u64 *hps_expand(u32 d) const;

    // Galois permutation tables (NTT domain), created on first use
    const u32 *galois_perm(u32 elt) const;
    const std::vector<u32> &galois_elts() const { return params_.galois_elts; }

    // kernel variants for this context; set_tuning applies them to later launches
    const TuneChoice &tuning() const { return tune_; }
    void set_tuning(const TuneChoice &t);

private:
    TuneChoice tune_{};
    Params params_;
    u64 N_ = 0;
    u32 logN_ = 0, size_Q_ = 0, size_P_ = 0, size_QP_ = 0;
    std::vector<hm::NttTable> host_tables_;
    NttTables tables_{};
    void *tw_buf_ = nullptr;     // forward + inverse twiddles, moduli, inverse consts
    u64 *cbuf_ = nullptr;        // per-level constants
    u64 *bigP_mod_q_ = nullptr;
    u64 bfv_negq_mod_t_ = 0, bfv_negq_mod_t_shoup_ = 0;
    const u64 *bfv_tinv_q_ = nullptr;
    std::vector<BfvDecrypt> bfv_dec_;
    std::vector<LevelConsts> levels_;
    mutable std::vector<std::pair<u32, u32 *>> galois_tables_;
    BehzConsts behz_{};
    std::vector<u64> bsk_;
    HpsConsts hps_{};
    std::vector<u64> hps_r_;
    std::vector<HpsConsts> hps_lv_, hps_drop_;
    std::vector<const u64 *> hps_expand_;
};

// stream-ordered device allocation (pooled, like the reference's cuda_auto_ptr)
void *dev_alloc(size_t bytes, hipStream_t s);
void dev_free(void *p, hipStream_t s);
void pool_init();

// Device scratch owned by one scope: allocated on stream s and released on it
// when the scope ends, error paths included (0 bytes: no allocation, null).
class Scratch {
public:
    Scratch(size_t bytes, hipStream_t s) : p_(static_cast<u64 *>(dev_alloc(bytes, s))), s_(s) {}
    ~Scratch() { dev_free(p_, s_); }
    Scratch(const Scratch &) = delete;
    Scratch &operator=(const Scratch &) = delete;
    u64 *get() const { return p_; }

private:
    u64 *p_;
    hipStream_t s_;
};

}  // namespace phx
