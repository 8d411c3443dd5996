// context.cpp -- context construction: tables, per-level constants, pool.
#include "context.h"

#include <algorithm>
#include <cstdlib>
#include <cstring>
#include <mutex>
#include <thread>

#include "autotune.h"
#include "tune.h"

namespace phx {

namespace {
std::once_flag g_pool_once;
std::mutex g_galois_mu;

// byte buffer with aligned appends; offsets are returned in bytes
struct ConstBuilder {
    std::vector<uint8_t> buf;
    size_t align(size_t a) {
        size_t off = (buf.size() + a - 1) / a * a;
        buf.resize(off);
        return off;
    }
    size_t push64(const std::vector<u64> &v) {
        size_t off = align(16);
        buf.resize(off + v.size() * 8);
        std::memcpy(buf.data() + off, v.data(), v.size() * 8);
        return off;
    }
    size_t push32(const std::vector<u32> &v) {
        size_t off = align(16);
        buf.resize(off + v.size() * 4);
        std::memcpy(buf.data() + off, v.data(), v.size() * 4);
        return off;
    }
    template <class T>
    size_t push_struct(const std::vector<T> &v) {
        size_t off = align(16);
        buf.resize(off + v.size() * sizeof(T));
        std::memcpy(buf.data() + off, v.data(), v.size() * sizeof(T));
        return off;
    }
};
// offsets (bytes into the constant buffer) of one HpsConsts
struct HpsOffs {
    size_t mod_qr, qhatinv_q, qhat_r, qinv, alphaq_r, frac, divs_mod_r, rhatinv_r, rhat_q, rinv, alphar_q;
    size_t mod_src, negrq, qinv_r, frac_r, divs_mod_q;
};

size_t push_doubles(ConstBuilder &cb, const std::vector<double> &v) {
    std::vector<u64> w(v.size());
    std::memcpy(w.data(), v.data(), v.size() * 8);
    return cb.push64(w);
}

// unsigned big integer, little-endian 64-bit words
struct BigU {
    std::vector<u64> w{1};
    void mul(u64 m) {
        u128 c = 0;
        for (auto &x : w) {
            const u128 v = (u128)x * m + c;
            x = (u64)v;
            c = v >> 64;
        }
        if (c) w.push_back((u64)c);
    }
    u64 mod(u64 m) const {
        u128 r = 0;
        for (size_t i = w.size(); i-- > 0;) r = ((r << 64) | w[i]) % m;
        return (u64)r;
    }
    void div(u64 m) {
        u128 r = 0;
        for (size_t i = w.size(); i-- > 0;) {
            const u128 cur = (r << 64) | w[i];
            w[i] = (u64)(cur / m);
            r = cur % m;
        }
    }
};

// Tables of the HPS scale-and-round from S = s[0 .. ns) to its first k moduli
// with factor F = prod(mult) (scaleAndRound_HPS_QlRl_Ql with F = t Ql, and its
// reuse Q -> Ql with F = Ql; src/rns.cu:833-890, 920-1001): theta_i =
// F [(S/s_i)^-1]_{s_i}; frac[j] = [theta_{k+j}]_{s_{k+j}} / s_{k+j};
// D[i][j] = floor(theta_{k+j} / s_{k+j}) mod s_i, D[i][ns-k] = floor(theta_i / s_i) mod s_i.
void scale_round_tables(const std::vector<u64> &sv, u32 k, const std::vector<u64> &mult, std::vector<double> &frac,
                        std::vector<u64> &D) {
    const u32 ns = (u32)sv.size(), nr = ns - k;
    frac.assign(nr, 0.0);
    D.assign((size_t)k * (nr + 1), 0);
    for (u32 i = 0; i < ns; ++i) {
        BigU th;
        for (u64 m : mult) th.mul(m);
        th.mul(hm::invmod(hm::punctured_mod(sv.data(), ns, i, sv[i]), sv[i]));
        if (i >= k) frac[i - k] = (double)th.mod(sv[i]) / (double)sv[i];
        th.div(sv[i]);
        for (u32 m = 0; m < k; ++m) {
            if (i >= k) D[(size_t)m * (nr + 1) + (i - k)] = th.mod(sv[m]);
            else if (i == m) D[(size_t)m * (nr + 1) + nr] = th.mod(sv[m]);
        }
    }
}

// HPS tables for Ql = qp[0 .. k) and Rl = rp[0 .. nr) (src/rns.cu:686-905; the
// BaseConverter tables of src/host/rns.cu:438-497); overq adds bConv_BEHZ_var1
// from src[0 .. ks) to Rl and the t/Rl scale-and-round QlRl -> Ql
HpsOffs build_hps_tables(ConstBuilder &cb, const u64 *qp, u32 k, const u64 *rp, u32 nr, u64 t, bool overq,
                         const u64 *src, u32 ks) {
    HpsOffs ho{};
    const u32 ns = k + nr;
    std::vector<u64> s(qp, qp + k);
    s.insert(s.end(), rp, rp + nr);
    std::vector<u64> qhatinv(k), qhat_r((size_t)nr * k), alphaq_r((size_t)(k + 1) * nr), rhatinv(nr),
        rhat_q((size_t)k * nr), alphar_q((size_t)(nr + 1) * k), divs((size_t)nr * (k + 1));
    std::vector<double> qinv(k), rinv(nr), frac(k);
    for (u32 i = 0; i < k; ++i) {
        qhatinv[i] = hm::invmod(hm::punctured_mod(qp, k, i, qp[i]), qp[i]);
        qinv[i] = 1.0 / (double)qp[i];
        for (u32 j = 0; j < nr; ++j) qhat_r[(size_t)j * k + i] = hm::punctured_mod(qp, k, i, rp[j]);
    }
    for (u32 v = 0; v <= k; ++v)
        for (u32 j = 0; j < nr; ++j) alphaq_r[(size_t)v * nr + j] = hm::mulmod(v, hm::product_mod(qp, k, rp[j]), rp[j]);
    for (u32 j = 0; j < nr; ++j) {
        rhatinv[j] = hm::invmod(hm::punctured_mod(rp, nr, j, rp[j]), rp[j]);
        rinv[j] = 1.0 / (double)rp[j];
        for (u32 i = 0; i < k; ++i) rhat_q[(size_t)i * nr + j] = hm::punctured_mod(rp, nr, j, qp[i]);
    }
    for (u32 v = 0; v <= nr; ++v)
        for (u32 i = 0; i < k; ++i) alphar_q[(size_t)v * k + i] = hm::mulmod(v, hm::product_mod(rp, nr, qp[i]), qp[i]);
    // HPS QR -> R: theta_i = t R [(S/s_i)^-1]_{s_i} (S = Q R): frac_i = [theta_i]_{q_i} / q_i,
    // divs = floor(theta_i / s_i) mod r_j (src/rns.cu:734-781)
    for (u32 i = 0; i < ns; ++i) {
        BigU th;
        th.mul(t);
        for (u32 j = 0; j < nr; ++j) th.mul(rp[j]);
        th.mul(hm::invmod(hm::punctured_mod(s.data(), ns, i, s[i]), s[i]));
        if (i < k) frac[i] = (double)th.mod(s[i]) / (double)s[i];
        th.div(s[i]);
        for (u32 j = 0; j < nr; ++j) {
            if (i < k) divs[(size_t)j * (k + 1) + i] = th.mod(rp[j]);
            else if (i == k + j) divs[(size_t)j * (k + 1) + k] = th.mod(rp[j]);
        }
    }
    std::vector<DevModulus> mqr(ns);
    for (u32 i = 0; i < ns; ++i) {
        u64 lo, hi;
        hm::barrett_ratio(s[i], lo, hi);
        mqr[i] = {s[i], lo, hi, 0};
    }
    ho.mod_qr = cb.push_struct(mqr);
    ho.qhatinv_q = cb.push64(qhatinv);
    ho.qhat_r = cb.push64(qhat_r);
    ho.qinv = push_doubles(cb, qinv);
    ho.alphaq_r = cb.push64(alphaq_r);
    ho.frac = push_doubles(cb, frac);
    ho.divs_mod_r = cb.push64(divs);
    ho.rhatinv_r = cb.push64(rhatinv);
    ho.rhat_q = cb.push64(rhat_q);
    ho.rinv = push_doubles(cb, rinv);
    ho.alphar_q = cb.push64(alphar_q);
    if (overq) {
        // bConv_BEHZ_var1 src -> Rl (src/host/rns.cu:466-497 with P = Rl):
        // negrq_i = src_i - [Rl (S'/src_i)^-1]_{src_i} (S' = prod src), qinv_r = src_i^-1 mod r_j
        std::vector<u64> negrq(ks), qinv_r((size_t)nr * ks);
        std::vector<DevModulus> msrc(ks);
        for (u32 i = 0; i < ks; ++i) {
            const u64 sh = hm::invmod(hm::punctured_mod(src, ks, i, src[i]), src[i]);
            negrq[i] = src[i] - hm::mulmod(hm::product_mod(rp, nr, src[i]), sh, src[i]);
            for (u32 j = 0; j < nr; ++j) qinv_r[(size_t)j * ks + i] = hm::invmod(src[i] % rp[j], rp[j]);
            u64 lo, hi;
            hm::barrett_ratio(src[i], lo, hi);
            msrc[i] = {src[i], lo, hi, 0};
        }
        std::vector<u64> mult{t};
        mult.insert(mult.end(), qp, qp + k);
        std::vector<double> frac_r;
        std::vector<u64> dq;
        scale_round_tables(s, k, mult, frac_r, dq);
        ho.mod_src = cb.push_struct(msrc);
        ho.negrq = cb.push64(negrq);
        ho.qinv_r = cb.push64(qinv_r);
        ho.frac_r = push_doubles(cb, frac_r);
        ho.divs_mod_q = cb.push64(dq);
    }
    return ho;
}

HpsConsts resolve_hps(const uint8_t *cbb, const HpsOffs &ho, u32 k, u32 nr, bool overq, u32 ks) {
    auto at = [&](size_t off) { return reinterpret_cast<const u64 *>(cbb + off); };
    auto atd = [&](size_t off) { return reinterpret_cast<const double *>(cbb + off); };
    HpsConsts h{};
    h.k = k;
    h.nr = nr;
    h.mod_qr = reinterpret_cast<const DevModulus *>(cbb + ho.mod_qr);
    h.qhatinv_q = at(ho.qhatinv_q);
    h.qhat_r = at(ho.qhat_r);
    h.qinv = atd(ho.qinv);
    h.alphaq_r = at(ho.alphaq_r);
    h.frac = atd(ho.frac);
    h.divs_mod_r = at(ho.divs_mod_r);
    h.rhatinv_r = at(ho.rhatinv_r);
    h.rhat_q = at(ho.rhat_q);
    h.rinv = atd(ho.rinv);
    h.alphar_q = at(ho.alphar_q);
    if (overq) {
        h.overq = 1;
        h.ks = ks;
        h.mod_src = reinterpret_cast<const DevModulus *>(cbb + ho.mod_src);
        h.negrq = at(ho.negrq);
        h.qinv_r = at(ho.qinv_r);
        h.frac_r = atd(ho.frac_r);
        h.divs_mod_q = at(ho.divs_mod_q);
    }
    return h;
}
}  // namespace

std::vector<uint8_t> bconv_mfma_operands(const std::vector<u64> &mat, u32 ni, u32 no, const std::vector<u64> &omod,
                                         u32 &rt) {
    // 4 outputs per group of 2 tiles; tile 2T+s, row 4g+i holds output 4T+g, digit b = 4s+i
    rt = 2 * ((no + 3) / 4);
    const size_t a_bytes = (size_t)rt * 2 * 64 * 16;
    std::vector<uint8_t> blob(a_bytes + (size_t)no * 16, 0);
    // mp[o][i][a] = 2^(8a) * mat[o][i] mod p_o (a indexes the input's bytes), written
    // in balanced base 256: mp = sum_b d_b 256^b with digits d_b in [-128, 127], b < 8
    // (eight such digits reach 2^63, above every p_o < 2^61)
    std::vector<int8_t> dg((size_t)no * ni * 8 * 8);
    std::vector<u64> mp((size_t)no * ni * 8);
    for (u32 o = 0; o < no; ++o) {
        const u64 p = omod[o];
        for (u32 i = 0; i < ni; ++i) {
            u64 v = mat[(size_t)o * ni + i] % p;
            for (u32 a = 0; a < 8; ++a) {
                const size_t e = ((size_t)o * ni + i) * 8 + a;
                mp[e] = v;
                long long w = (long long)v;
                for (u32 b = 0; b < 8; ++b) {
                    long long r = w & 255;
                    if (r >= 128) r -= 256;
                    dg[e * 8 + b] = (int8_t)r;
                    w = (w - r) >> 8;
                }
                if (w != 0) fail(kLogicError, "FastBConv constant outside the balanced digit range");
                v = hm::mulmod(v, 256, p);
            }
        }
    }
    // A[row][K]: row -> (output, digit) as above, K = i*8 + a
    auto A = [&](u32 row, u32 K) -> uint8_t {
        const u32 t = row / 16, rr = row % 16;
        const u32 o = 4 * (t / 2) + rr / 4, b = 4 * (t % 2) + rr % 4, i = K / 8, a = K % 8;
        if (o >= no || i >= ni) return 0;
        return (uint8_t)dg[(((size_t)o * ni + i) * 8 + a) * 8 + b];
    };
    for (u32 t = 0; t < rt; ++t)
        for (u32 h = 0; h < 2; ++h)
            for (u32 l = 0; l < 64; ++l)
                for (u32 j = 0; j < 16; ++j)
                    blob[(((size_t)t * 2 + h) * 64 + l) * 16 + j] = A(16 * t + l % 16, 64 * h + 16 * (l / 16) + j);
    // bias of the signed input bytes: the kernel feeds s = byte - 128, so output o
    // adds back C_o = 128 * sum_b 256^b sum_(i,a) d_b(mp[o][i][a]) = 128 sum_(i,a)
    // mp[o][i][a] (< 2^75); stored as (C_o mod 2^64, fl(C_o) as double bits)
    for (u32 o = 0; o < no; ++o) {
        hm::u128 c = 0;
        for (u32 i = 0; i < ni; ++i)
            for (u32 a = 0; a < 8; ++a) c += (hm::u128)mp[((size_t)o * ni + i) * 8 + a] * 128;
        const u64 lo = (u64)c;
        const double cd = (double)(u64)(c >> 64) * 0x1p64 + (double)lo;
        std::memcpy(blob.data() + a_bytes + (size_t)o * 16, &lo, 8);
        std::memcpy(blob.data() + a_bytes + (size_t)o * 16 + 8, &cd, 8);
    }
    return blob;
}

void pool_init() {
    std::call_once(g_pool_once, [] {
        int dev = 0;
        PHX_HIP_CHECK(hipGetDevice(&dev));
        hipMemPool_t pool;
        PHX_HIP_CHECK(hipDeviceGetDefaultMemPool(&pool, dev));
        uint64_t thr = UINT64_MAX;
        PHX_HIP_CHECK(hipMemPoolSetAttribute(pool, hipMemPoolAttrReleaseThreshold, &thr));
    });
}

void *dev_alloc(size_t bytes, hipStream_t s) {
    if (bytes == 0) return nullptr;
    void *p = nullptr;
    PHX_HIP_CHECK(hipMallocAsync(&p, bytes, s));
    return p;
}

void dev_free(void *p, hipStream_t s) {
    if (p) (void)hipFreeAsync(p, s);
}

Context::Context(const Params &p) : params_(p) {
    const size_t QP = p.coeff_modulus.size();
    require(QP >= 2, "The coefficient modulus must be a vector of at least two primes");
    require(p.special_modulus_size >= 1 && p.special_modulus_size < QP, "invalid special_modulus_size");
    N_ = p.poly_degree;
    logN_ = (u32)hm::bit_count(N_) - 1;
    require(N_ == ((u64)1 << logN_) && logN_ >= 11 && logN_ <= 17,
            "poly_modulus_degree must be a power of two in [2^11, 2^17]");
    size_QP_ = (u32)QP;
    size_P_ = (u32)p.special_modulus_size;
    size_Q_ = size_QP_ - size_P_;
    for (u64 q : p.coeff_modulus) {
        require(q > 1 && (q >> 61) == 0, "value can be at most 61-bit and cannot be 1");
        require(hm::is_prime(q), "coefficient modulus must be prime");
        require(q % (2 * N_) == 1, "coefficient modulus must be 1 mod 2N");
    }
    for (size_t i = 0; i < QP; ++i)
        for (size_t j = i + 1; j < QP; ++j) require(p.coeff_modulus[i] != p.coeff_modulus[j], "moduli must be distinct");
    pool_init();

    // ---- BEHZ bases (BFV multiplication at the top data level, src/rns.cu:392-423):
    // B has size_Q 61-bit primes, one more when 32 + bits(t) + bits(Q) >= 61 (size_Q + 1);
    // m_sk is the first prime of the same search, Bsk = B then m_sk
    if (p.scheme == kSchemeBFV && p.mul_tech == kMulBEHZ) {
        require(p.plain_modulus > 1, "plain_modulus is not set");
        std::vector<u64> Qw{1};
        for (u32 i = 0; i < size_Q_; ++i) {
            u128 carry = 0;
            for (auto &w : Qw) {
                const u128 v = (u128)w * p.coeff_modulus[i] + carry;
                w = (u64)v;
                carry = v >> 64;
            }
            if (carry) Qw.push_back((u64)carry);
        }
        const int qbits = 64 * ((int)Qw.size() - 1) + hm::bit_count(Qw.back());
        u32 size_B = size_Q_;
        if (32 + hm::bit_count(p.plain_modulus) + qbits >= 61 * (int)size_Q_ + 61) ++size_B;
        const std::vector<u64> pr = hm::get_primes(N_, 61, size_B + 1);
        bsk_.assign(pr.begin() + 1, pr.end());
        bsk_.push_back(pr[0]);
        for (u64 b : bsk_)
            for (u64 q : p.coeff_modulus) require(b != q, "coefficient modulus collides with the BEHZ base");
    }
    // ---- HPS base R (src/rns.cu:686-699): size_Q + 1 primes below the smallest q,
    // stepping down by 2N within its bit size (get_primes_below, src/host/numth.cu:235-262);
    // size_Q primes for hps_overq (src/rns.cu:792-801)
    const bool overq = p.mul_tech == kMulHPSOverQ || p.mul_tech == kMulHPSOverQLeveled;
    if (p.scheme == kSchemeBFV && (p.mul_tech == kMulHPS || overq)) {
        require(p.plain_modulus > 1, "plain_modulus is not set");
        const u64 qmin = *std::min_element(p.coeff_modulus.begin(), p.coeff_modulus.begin() + size_Q_);
        const u64 lower = 1ull << (hm::bit_count(qmin) - 1);
        const size_t nr = overq ? size_Q_ : size_Q_ + 1;
        for (u64 v = qmin - 2 * N_; hps_r_.size() < nr && v > lower; v -= 2 * N_)
            if (hm::is_prime(v)) hps_r_.push_back(v);
        require(hps_r_.size() == nr, "failed to find enough qualifying primes");
        for (u64 r : hps_r_)
            for (u64 q : p.coeff_modulus) require(r != q, "coefficient modulus collides with the HPS base");
    }
    std::vector<u64> rows_p = p.coeff_modulus;
    rows_p.insert(rows_p.end(), bsk_.begin(), bsk_.end());
    rows_p.insert(rows_p.end(), hps_r_.begin(), hps_r_.end());
    const size_t R = rows_p.size();

    // ---- NTT tables (host, one thread per prime) ----
    host_tables_.resize(R);
    {
        unsigned nt = std::max(1u, std::min<unsigned>(std::thread::hardware_concurrency(), 16));
        std::vector<std::thread> th;
        for (unsigned w = 0; w < nt; ++w)
            th.emplace_back([&, w] {
                for (size_t i = w; i < R; i += nt) host_tables_[i] = hm::make_ntt_table(rows_p[i], (int)logN_);
            });
        for (auto &t : th) t.join();
    }
    // Rows with q < 2^50 run the NTT on the FP64 pipe (exact, see ntt_arith.h)
    // unless the ntt_fp64 switch is off (then every row uses 64-bit integers).
    const bool int_only = tune_value("ntt_fp64") == 0;
    const size_t twb = R * N_ * sizeof(TwPair);
    const size_t twdb = R * N_ * sizeof(double);
    const size_t modb = R * sizeof(DevModulus);
    const size_t icb = R * 4 * sizeof(u64);
    const size_t ncb = R * 4 * sizeof(u64);
    std::vector<uint8_t> h(2 * twb + 2 * twdb + modb + icb + ncb);
    TwPair *htw = reinterpret_cast<TwPair *>(h.data());
    TwPair *hitw = reinterpret_cast<TwPair *>(h.data() + twb);
    double *htwd = reinterpret_cast<double *>(h.data() + 2 * twb);
    double *hitwd = reinterpret_cast<double *>(h.data() + 2 * twb + twdb);
    DevModulus *hmod = reinterpret_cast<DevModulus *>(h.data() + 2 * twb + 2 * twdb);
    u64 *hic = reinterpret_cast<u64 *>(h.data() + 2 * twb + 2 * twdb + modb);
    u64 *hnc = reinterpret_cast<u64 *>(h.data() + 2 * twb + 2 * twdb + modb + icb);
    for (size_t r = 0; r < R; ++r) {
        const auto &t = host_tables_[r];
        for (u64 k = 0; k < N_; ++k) {
            const u64 o = r * N_ + tw_slot(k, N_);   // device storage order (ntt_kernels.h)
            htw[o] = {t.fw[k], t.fws[k]};
            hitw[o] = {t.iw[k], t.iws[k]};
            htwd[o] = (double)t.fw[k];
            hitwd[o] = (double)t.iw[k];
        }
        u64 lo, hi;
        hm::barrett_ratio(t.q, lo, hi);
        u64 fp_bits = 0;
        if (!int_only && t.q < (1ull << 50)) {
            const double qinv = 1.0 / (double)t.q;
            std::memcpy(&fp_bits, &qinv, 8);
        }
        hmod[r] = {t.q, lo, hi, fp_bits};
        const u64 s1 = hm::mulmod(t.ninv, t.iw[1], t.q);
        hic[4 * r + 0] = t.ninv;
        hic[4 * r + 1] = hm::shoup(t.ninv, t.q);
        hic[4 * r + 2] = s1;
        hic[4 * r + 3] = hm::shoup(s1, t.q);
        hnc[4 * r + 0] = 0 - t.q;
        hnc[4 * r + 1] = 0 - 2 * t.q;
        hnc[4 * r + 2] = 0 - 3 * t.q;
        hnc[4 * r + 3] = 3 * t.q + 1;
    }
    PHX_HIP_CHECK(hipMalloc(&tw_buf_, h.size()));
    PHX_HIP_CHECK(hipMemcpy(tw_buf_, h.data(), h.size(), hipMemcpyHostToDevice));
    uint8_t *db = static_cast<uint8_t *>(tw_buf_);
    tables_.tw = reinterpret_cast<const TwPair *>(db);
    tables_.itw = reinterpret_cast<const TwPair *>(db + twb);
    tables_.twd = reinterpret_cast<const double *>(db + 2 * twb);
    tables_.itwd = reinterpret_cast<const double *>(db + 2 * twb + twdb);
    tables_.mod = reinterpret_cast<const DevModulus *>(db + 2 * twb + 2 * twdb);
    tables_.inv_consts = reinterpret_cast<const u64 *>(db + 2 * twb + 2 * twdb + modb);
    tables_.int_consts = reinterpret_cast<const u64 *>(db + 2 * twb + 2 * twdb + modb + icb);
    tables_.logN = logN_;
    tables_.N = (u32)N_;
    set_tuning(TuneChoice{});

    // ---- per-level key-switching constants ----
    // reference: DRNSTool ctor, hybrid key-switching part (src/rns.cu:97-195)
    const u64 *qp = p.coeff_modulus.data();
    const u64 *Pp = qp + size_Q_;
    const u32 alpha = size_P_;
    ConstBuilder cb;
    struct Offs {
        size_t modup_scale, modup_bconv, moddown_scale, moddown_bconv, pinv, rescale, modup_fbconv, moddown_fbconv;
    };
    std::vector<Offs> offs(size_Q_ + 1);
    levels_.resize(size_Q_ + 1);
    auto row_of = [&](u32 Ql, u32 pos) -> u32 { return pos < Ql ? pos : size_Q_ + (pos - Ql); };
    // MFMA operands of a FastBConv matrix (only when K = 9 * ni fits one 128-wide slab)
    auto add_mfma = [&](BconvDesc &d, const std::vector<u64> &mat, const std::vector<u64> &om) {
        d.rt = 0;
        if (d.ni > 16) return;
        u32 rt = 0;
        std::vector<uint8_t> blob = bconv_mfma_operands(mat, d.ni, d.no, om, rt);
        std::vector<u64> words(blob.size() / 8);
        std::memcpy(words.data(), blob.data(), blob.size());
        d.a_off = cb.push64(words) / 8;
        std::vector<u64> pinv(d.no);
        for (u32 o = 0; o < d.no; ++o) {
            const double v = 1.0 / (double)om[o];
            std::memcpy(&pinv[o], &v, 8);
        }
        d.pinv_off = cb.push64(pinv) / 8;
        d.rt = rt;
    };
    // The same conversion for the launch fused with the forward NTT's column pass
    // (bconv_ntt.hip): outputs on 64-bit integer butterflies first, then those on
    // FP64 butterflies, padded to a multiple of four with outputs that are computed
    // but not stored (pos ~0, a zero matrix row, the last output's prime). A group
    // of four outputs runs one arithmetic: the integer one when its first output
    // needs it (the FP64 outputs of such a group are stored in FP64 form).
    auto row_fp = [&](u32 row) { return !int_only && host_tables_[row].q < (1ull << 50); };
    auto add_fused = [&](const BconvDesc &d, const std::vector<u64> &mat, const std::vector<u32> &pos,
                         const std::vector<u32> &rows, const std::vector<u64> &om) {
        BconvDesc f{};
        if (d.ni > 16) return f;
        std::vector<u32> order[2];
        for (u32 o = 0; o < d.no; ++o) order[row_fp(rows[o]) ? 1 : 0].push_back(o);
        std::vector<u64> fmat, fom;
        std::vector<u32> fpos, frows;
        std::vector<u32> ord = order[0];
        ord.insert(ord.end(), order[1].begin(), order[1].end());
        const size_t padded = (ord.size() + 3) / 4 * 4;
        for (size_t k = 0; k < padded; ++k) {
            const bool real = k < ord.size();
            const u32 o = real ? ord[k] : ord.back();
            for (u32 i = 0; i < d.ni; ++i) fmat.push_back(real ? mat[(size_t)o * d.ni + i] : 0);
            fom.push_back(om[o]);
            fpos.push_back(real ? pos[o] : 0xffffffffu);
            frows.push_back(rows[o]);
        }
        f.ni = d.ni;
        f.no = (u32)fpos.size();
        f.in_first = d.in_first;
        f.own_first = d.own_first;
        f.pos_off = cb.push32(fpos) / 4;
        f.row_off = cb.push32(frows) / 4;
        add_mfma(f, fmat, fom);
        return f;
    };
    auto scale_consts = [&](u32 row, u64 scale) {
        const auto &t = host_tables_[row];
        const u64 s0 = hm::mulmod(t.ninv, scale, t.q);
        const u64 s1 = hm::mulmod(hm::mulmod(t.ninv, t.iw[1], t.q), scale, t.q);
        return std::vector<u64>{s0, hm::shoup(s0, t.q), s1, hm::shoup(s1, t.q)};
    };
    for (u32 c = 1; c <= size_Q_; ++c) {
        LevelConsts &L = levels_[c];
        const u32 Ql = size_Q_ - (c - 1), QlP = Ql + size_P_;
        const u32 beta = (Ql + alpha - 1) / alpha;
        L.chain_index = c;
        L.size_Ql = Ql;
        L.size_QlP = QlP;
        L.beta = beta;
        std::vector<u64> qlp(QlP);
        for (u32 i = 0; i < QlP; ++i) qlp[i] = qp[row_of(Ql, i)];

        // modup: INTT scale partQlHatInv, then per digit FastBConv to the complement
        std::vector<u64> msc;
        std::vector<BconvDesc> descs, fdescs;
        for (u32 j = 0; j < beta; ++j) {
            const u32 st = alpha * j, sz = (j == beta - 1) ? Ql - alpha * (beta - 1) : alpha;
            for (u32 i = st; i < st + sz; ++i) {
                const u64 hinv = hm::invmod(hm::punctured_mod(qlp.data() + st, sz, i - st, qlp[i]), qlp[i]);
                auto v = scale_consts(i, hinv);
                msc.insert(msc.end(), v.begin(), v.end());
            }
            const u32 no = QlP - sz;
            std::vector<u64> mat((size_t)no * sz);
            std::vector<u32> pos(no), rows(no);
            for (u32 o = 0; o < no; ++o) {
                pos[o] = o < st ? o : o + sz;
                rows[o] = row_of(Ql, pos[o]);
                for (u32 i = 0; i < sz; ++i) mat[(size_t)o * sz + i] = hm::punctured_mod(qlp.data() + st, sz, i, qlp[pos[o]]);
            }
            BconvDesc d{};
            d.ni = sz;
            d.no = no;
            d.in_first = st;
            d.own_first = st;
            d.mat_off = cb.push64(mat) / 8;
            d.pos_off = cb.push32(pos) / 4;
            d.row_off = cb.push32(rows) / 4;
            std::vector<u64> om(no);
            for (u32 o = 0; o < no; ++o) om[o] = qlp[pos[o]];
            add_mfma(d, mat, om);
            descs.push_back(d);
            fdescs.push_back(add_fused(d, mat, pos, rows, om));
        }
        offs[c].modup_scale = cb.push64(msc);
        offs[c].modup_bconv = cb.push_struct(descs);
        offs[c].modup_fbconv = cb.push_struct(fdescs);
        L.h_modup_bconv = descs;
        L.h_modup_fbconv = fdescs;

        // moddown: INTT of the P limbs scaled by PHatInv, FastBConv P -> Ql, P^-1
        std::vector<u64> dsc;
        for (u32 j = 0; j < size_P_; ++j) {
            const u64 hinv = hm::invmod(hm::punctured_mod(Pp, size_P_, j, Pp[j]), Pp[j]);
            auto v = scale_consts(size_Q_ + j, hinv);
            dsc.insert(dsc.end(), v.begin(), v.end());
        }
        offs[c].moddown_scale = cb.push64(dsc);
        {
            std::vector<u64> mat((size_t)Ql * size_P_);
            std::vector<u32> pos(Ql), rows(Ql);
            for (u32 i = 0; i < Ql; ++i) {
                pos[i] = i;
                rows[i] = i;
                for (u32 j = 0; j < size_P_; ++j) mat[(size_t)i * size_P_ + j] = hm::punctured_mod(Pp, size_P_, j, qp[i]);
            }
            BconvDesc d{};
            d.ni = size_P_;
            d.no = Ql;
            d.in_first = Ql;
            d.own_first = 0xffffffffu;
            d.mat_off = cb.push64(mat) / 8;
            d.pos_off = cb.push32(pos) / 4;
            d.row_off = cb.push32(rows) / 4;
            const std::vector<u64> om(qp, qp + Ql);
            add_mfma(d, mat, om);
            offs[c].moddown_bconv = cb.push_struct(std::vector<BconvDesc>{d});
            L.h_moddown_bconv = d;
            L.h_moddown_fbconv = add_fused(d, mat, pos, rows, om);
            offs[c].moddown_fbconv = cb.push_struct(std::vector<BconvDesc>{L.h_moddown_fbconv});
        }
        std::vector<u64> pinv;
        for (u32 i = 0; i < Ql; ++i) {
            const u64 v = hm::invmod(hm::product_mod(Pp, size_P_, qp[i]), qp[i]);
            pinv.push_back(v);
            pinv.push_back(hm::shoup(v, qp[i]));
        }
        offs[c].pinv = cb.push64(pinv);
        std::vector<u64> rq;
        for (u32 j = 0; j + 1 < Ql; ++j) {
            const u64 v = hm::invmod(qp[Ql - 1] % qp[j], qp[j]);
            rq.push_back(v);
            rq.push_back(hm::shoup(v, qp[j]));
        }
        if (rq.empty()) rq.push_back(0);
        offs[c].rescale = cb.push64(rq);
    }
    // key level: P mod q_i (gadget factor of the key-switching keys)
    std::vector<u64> bp;
    for (u32 i = 0; i < size_Q_; ++i) {
        const u64 v = hm::product_mod(Pp, size_P_, qp[i]);
        bp.push_back(v);
        bp.push_back(hm::shoup(v, qp[i]));
    }
    const size_t bp_off = cb.push64(bp);
    // BFV plaintext scaling (multiply_add_plain_with_scaling_variant, src/scalingvariant.cu:10-34;
    // constants src/rns.cu:296-320): [-Q]_t and t^-1 mod q_i (Shoup) at the first data level
    size_t tinv_off = 0;
    if (p.scheme == kSchemeBFV && p.plain_modulus > 1) {
        const u64 t = p.plain_modulus;
        const u64 qmodt = hm::product_mod(qp, size_Q_, t);
        bfv_negq_mod_t_ = qmodt ? t - qmodt : 0;
        bfv_negq_mod_t_shoup_ = hm::shoup(bfv_negq_mod_t_, t);
        std::vector<u64> ti;
        for (u32 i = 0; i < size_Q_; ++i) {
            const u64 v = hm::invmod(t % qp[i], qp[i]);
            ti.push_back(v);
            ti.push_back(hm::shoup(v, qp[i]));
        }
        tinv_off = cb.push64(ti);
    }
    // BFV decryption scale-and-round at the first data level (PhantomSecretKey::bfv_decrypt,
    // src/secretkey.cu:571-635): BEHZ for mul_tech behz (constants src/rns.cu:330-390), HPS
    // otherwise (constants src/rns.cu:583-673, kernel choice 1634-1690)
    std::vector<size_t> dec_off(size_Q_ + 1, 0);
    bfv_dec_.assign(size_Q_ + 1, BfvDecrypt{});
    u64 dec_qmax = 0;
    for (u32 i = 0; i < size_Q_; ++i) dec_qmax = std::max(dec_qmax, qp[i]);
    for (u32 lev = 1; p.scheme == kSchemeBFV && p.plain_modulus > 1 && lev <= size_Q_; ++lev) {
        const u32 Ql = size_Q_ - (lev - 1);
        const u64 t = p.plain_modulus;
        BfvDecrypt &D = bfv_dec_[lev];
        D.t = t;
        hm::barrett_ratio(t, D.t_ratio_lo, D.t_ratio_hi);
        std::vector<u64> tab((size_t)Ql * 8, 0);
        if (p.mul_tech == kMulBEHZ) {
            const u64 g = hm::get_primes(N_, 61, 1)[0];
            D.kind = kBfvDecBEHZ;
            D.gamma = g;
            hm::barrett_ratio(g, D.g_ratio_lo, D.g_ratio_hi);
            const u64 qt = hm::product_mod(qp, Ql, t), qg = hm::product_mod(qp, Ql, g);
            const u64 it = hm::invmod(qt, t), ig = hm::invmod(qg, g);
            D.neg_qinv_t = it ? t - it : 0;
            D.neg_qinv_g = ig ? g - ig : 0;
            D.neg_qinv_t_shoup = hm::shoup(D.neg_qinv_t, t);
            D.neg_qinv_g_shoup = hm::shoup(D.neg_qinv_g, g);
            D.inv_gamma_t = hm::invmod(g % t, t);
            D.inv_gamma_t_shoup = hm::shoup(D.inv_gamma_t, t);
            for (u32 i = 0; i < Ql; ++i) {
                const u64 q = qp[i];
                const u64 tg = hm::mulmod(t % q, g % q, q);
                const u64 hinv = hm::invmod(hm::punctured_mod(qp, Ql, i, q), q);
                u64 *r = &tab[(size_t)i * 8];
                r[0] = tg;
                r[1] = hm::shoup(tg, q);
                r[2] = hinv;
                r[3] = hm::shoup(hinv, q);
                r[4] = hm::punctured_mod(qp, Ql, i, t);
                r[5] = hm::punctured_mod(qp, Ql, i, g);
            }
        } else {
            // qMSB over all of Q, the size term over this level's primes (src/rns.cu:576-592)
            const int qmsb = hm::bit_count(dec_qmax), sizeqmsb = hm::bit_count(Ql), tmsb = hm::bit_count(t);
            const bool large = qmsb + sizeqmsb >= 52;
            D.qmsb_half = (u32)(qmsb >> 1);
            if (!large)
                D.kind = qmsb + tmsb + sizeqmsb < 52 ? kBfvDecHpsSmallLazy : kBfvDecHpsSmall;
            else
                D.kind = (int)D.qmsb_half + tmsb + sizeqmsb < 52 ? kBfvDecHpsLargeLazy : kBfvDecHpsLarge;
            auto dbits = [](double d) { u64 b; std::memcpy(&b, &d, 8); return b; };
            for (u32 i = 0; i < Ql; ++i) {
                const u64 q = qp[i];
                const u64 v = hm::invmod(hm::punctured_mod(qp, Ql, i, q), q);
                u64 *r = &tab[(size_t)i * 8];
                const u128 big = (u128)t * v;
                r[0] = (u64)((big / q) % t);
                r[1] = hm::shoup(r[0], t);
                r[2] = dbits((double)(u64)(big % q) / (double)q);
                if (large) {
                    const u64 vb = (u64)(((u128)v << D.qmsb_half) % q);
                    const u128 bigb = (u128)t * vb;
                    r[3] = (u64)((bigb / q) % t);
                    r[4] = hm::shoup(r[3], t);
                    r[5] = dbits((double)(u64)(bigb % q) / (double)q);
                }
            }
        }
        dec_off[lev] = cb.push64(tab);
    }
    // BEHZ constants (src/rns.cu:424-571)
    struct BehzOffs {
        size_t lift_f, qhat_bsk, qhat_mt, q_mod_bsk, mtinv_bsk, qhatinv_q, qinv_bsk, bhatinv_b, bhat_msk, bhat_q,
            b_mod_q, tscale_q, tscale_bsk;
    } bo{};
    if (!bsk_.empty()) {
        const u32 k = size_Q_, nb = (u32)bsk_.size(), nB = nb - 1;
        const u64 mt = 1ull << 32, t = p.plain_modulus, msk = bsk_[nB];
        auto qhat_mod = [&](u32 i, u64 m) { return hm::punctured_mod(qp, k, i, m); };
        auto bhat_mod = [&](u32 j, u64 m) { return hm::punctured_mod(bsk_.data(), nB, j, m); };
        std::vector<u64> lift_f(k), qhat_mt(k), qhatinv(k), b_mod_q(k), bhat_q((size_t)k * nB);
        for (u32 i = 0; i < k; ++i) {
            const u64 hinv = hm::invmod(qhat_mod(i, qp[i]), qp[i]);
            lift_f[i] = hm::mulmod(mt % qp[i], hinv, qp[i]);
            qhat_mt[i] = qhat_mod(i, mt);
            qhatinv[i] = hinv;
            b_mod_q[i] = hm::product_mod(bsk_.data(), nB, qp[i]);
            for (u32 j = 0; j < nB; ++j) bhat_q[(size_t)i * nB + j] = bhat_mod(j, qp[i]);
        }
        std::vector<u64> qhat_bsk((size_t)nb * k), q_mod_bsk(nb), mtinv(nb), qinv(nb);
        for (u32 o = 0; o < nb; ++o) {
            const u64 pm = bsk_[o];
            for (u32 i = 0; i < k; ++i) qhat_bsk[(size_t)o * k + i] = qhat_mod(i, pm);
            q_mod_bsk[o] = hm::product_mod(qp, k, pm);
            mtinv[o] = hm::invmod(mt % pm, pm);
            qinv[o] = hm::invmod(q_mod_bsk[o], pm);
        }
        std::vector<u64> bhatinv(nB), bhat_msk(nB);
        for (u32 j = 0; j < nB; ++j) {
            bhatinv[j] = hm::invmod(bhat_mod(j, bsk_[j]), bsk_[j]);
            bhat_msk[j] = bhat_mod(j, msk);
        }
        // -Q^-1 mod 2^32 (Q odd): Newton iteration for the inverse
        const u64 qm = hm::product_mod(qp, k, mt);
        u64 inv = 1;
        for (int it = 0; it < 6; ++it) inv = (inv * (2 - qm * inv)) & (mt - 1);
        behz_.neg_qinv_mt = (mt - inv) & (mt - 1);
        behz_.binv_msk = hm::invmod(hm::product_mod(bsk_.data(), nB, msk), msk);
        auto tscale = [&](size_t row) {
            const auto &tb = host_tables_[row];
            const u64 tt = t % tb.q;
            const u64 s0 = hm::mulmod(tb.ninv, tt, tb.q);
            const u64 s1 = hm::mulmod(hm::mulmod(tb.ninv, tb.iw[1], tb.q), tt, tb.q);
            return std::vector<u64>{s0, hm::shoup(s0, tb.q), s1, hm::shoup(s1, tb.q)};
        };
        std::vector<u64> tsq, tsb;
        for (u32 i = 0; i < k; ++i) {
            auto v = tscale(i);
            tsq.insert(tsq.end(), v.begin(), v.end());
        }
        for (u32 o = 0; o < nb; ++o) {
            auto v = tscale(QP + o);
            tsb.insert(tsb.end(), v.begin(), v.end());
        }
        bo = {cb.push64(lift_f), cb.push64(qhat_bsk), cb.push64(qhat_mt), cb.push64(q_mod_bsk), cb.push64(mtinv),
              cb.push64(qhatinv), cb.push64(qinv), cb.push64(bhatinv), cb.push64(bhat_msk), cb.push64(bhat_q),
              cb.push64(b_mod_q), cb.push64(tsq), cb.push64(tsb)};
        behz_.k = k;
        behz_.nb = nb;
        behz_.bsk_row0 = (u32)QP;
        behz_.t = t;
    }
    // HPS constants: the top level, and for hps_overq_leveled one set per number
    // of dropped levels d = 1 .. size_Q - 1 (Ql = size_Q - d, Rl = the first Ql
    // primes of R; src/rns.cu:792-1004)
    std::vector<HpsOffs> hps_offs, drop_offs;
    std::vector<size_t> expand_offs;
    if (!hps_r_.empty()) {
        const u32 k0 = size_Q_;
        const u64 t = p.plain_modulus;
        const u32 levels = p.mul_tech == kMulHPSOverQLeveled ? k0 : 1;
        for (u32 d = 0; d < levels; ++d) {
            const u32 Ql = k0 - d;
            const u32 nr = overq ? Ql : (u32)hps_r_.size();
            // switch source of the second operand: Ql at the top level, all of Q below
            hps_offs.push_back(build_hps_tables(cb, qp, Ql, hps_r_.data(), nr, t, overq, qp, d ? k0 : Ql));
            if (d) {
                std::vector<u64> sQ(qp, qp + k0), ql(qp, qp + Ql);
                std::vector<double> frac;
                std::vector<u64> D;
                scale_round_tables(sQ, Ql, ql, frac, D);
                HpsOffs o{};
                o.frac_r = push_doubles(cb, frac);
                o.divs_mod_q = cb.push64(D);
                std::vector<DevModulus> mq(k0);
                for (u32 i = 0; i < k0; ++i) {
                    u64 lo, hi;
                    hm::barrett_ratio(qp[i], lo, hi);
                    mq[i] = {qp[i], lo, hi, 0};
                }
                o.mod_qr = cb.push_struct(mq);
                drop_offs.push_back(o);
                std::vector<u64> ex(2 * (size_t)Ql);
                for (u32 i = 0; i < Ql; ++i) {
                    ex[2 * i] = hm::product_mod(qp + Ql, d, qp[i]);
                    ex[2 * i + 1] = hm::shoup(ex[2 * i], qp[i]);
                }
                expand_offs.push_back(cb.push64(ex));
            } else {
                drop_offs.push_back(HpsOffs{});
                expand_offs.push_back(0);
            }
        }
    }
    cb.align(256);
    PHX_HIP_CHECK(hipMalloc(&cbuf_, cb.buf.size()));
    PHX_HIP_CHECK(hipMemcpy(cbuf_, cb.buf.data(), cb.buf.size(), hipMemcpyHostToDevice));
    uint8_t *cbb = reinterpret_cast<uint8_t *>(cbuf_);
    for (u32 c = 1; c <= size_Q_; ++c) {
        LevelConsts &L = levels_[c];
        L.modup_scale = reinterpret_cast<const u64 *>(cbb + offs[c].modup_scale);
        L.modup_bconv = reinterpret_cast<const BconvDesc *>(cbb + offs[c].modup_bconv);
        L.moddown_scale = reinterpret_cast<const u64 *>(cbb + offs[c].moddown_scale);
        L.moddown_bconv = reinterpret_cast<const BconvDesc *>(cbb + offs[c].moddown_bconv);
        L.modup_fbconv = reinterpret_cast<const BconvDesc *>(cbb + offs[c].modup_fbconv);
        L.moddown_fbconv = reinterpret_cast<const BconvDesc *>(cbb + offs[c].moddown_fbconv);
        L.pinv = reinterpret_cast<const u64 *>(cbb + offs[c].pinv);
        L.rescale_qinv = reinterpret_cast<const u64 *>(cbb + offs[c].rescale);
    }
    bigP_mod_q_ = reinterpret_cast<u64 *>(cbb + bp_off);
    if (tinv_off) bfv_tinv_q_ = reinterpret_cast<const u64 *>(cbb + tinv_off);
    for (size_t lev = 1; lev < dec_off.size(); ++lev)
        if (dec_off[lev]) bfv_dec_[lev].tab = reinterpret_cast<const u64 *>(cbb + dec_off[lev]);
    for (size_t d = 0; d < hps_offs.size(); ++d) {
        const u32 k0 = size_Q_, Ql = k0 - (u32)d;
        HpsConsts h = resolve_hps(cbb, hps_offs[d], Ql, overq ? Ql : (u32)hps_r_.size(), overq, d ? k0 : Ql);
        h.r_row0 = (u32)(QP + bsk_.size());
        if (d == 0) hps_ = h;
        hps_lv_.push_back(h);
        HpsConsts dr{};
        if (d) {
            dr.k = Ql;
            dr.nr = (u32)d;
            dr.mod_qr = reinterpret_cast<const DevModulus *>(cbb + drop_offs[d].mod_qr);
            dr.frac_r = reinterpret_cast<const double *>(cbb + drop_offs[d].frac_r);
            dr.divs_mod_q = reinterpret_cast<const u64 *>(cbb + drop_offs[d].divs_mod_q);
            dr.overq = 1;
        }
        hps_drop_.push_back(dr);
        hps_expand_.push_back(d ? reinterpret_cast<const u64 *>(cbb + expand_offs[d]) : nullptr);
    }
    if (behz_.k) {
        auto at = [&](size_t off) { return reinterpret_cast<const u64 *>(cbb + off); };
        behz_.lift_f = at(bo.lift_f);
        behz_.qhat_bsk = at(bo.qhat_bsk);
        behz_.qhat_mt = at(bo.qhat_mt);
        behz_.q_mod_bsk = at(bo.q_mod_bsk);
        behz_.mtinv_bsk = at(bo.mtinv_bsk);
        behz_.qhatinv_q = at(bo.qhatinv_q);
        behz_.qinv_bsk = at(bo.qinv_bsk);
        behz_.bhatinv_b = at(bo.bhatinv_b);
        behz_.bhat_msk = at(bo.bhat_msk);
        behz_.bhat_q = at(bo.bhat_q);
        behz_.b_mod_q = at(bo.b_mod_q);
        behz_.tscale_q = at(bo.tscale_q);
        behz_.tscale_bsk = at(bo.tscale_bsk);
    }
    levels_[0].chain_index = 0;
    levels_[0].size_Ql = size_Q_;
    levels_[0].size_QlP = size_QP_;
    // kernel variants from a tuning cache written by the autotuner, if one is named
    if (const char *path = std::getenv("PHX_TUNE_CACHE")) {
        TuneChoice t;
        if (*path && tune_cache_load(path, *this, &t)) set_tuning(t);
    }
}

Context::~Context() {
    (void)hipDeviceSynchronize();
    for (auto &g : galois_tables_) (void)hipFree(g.second);
    if (cbuf_) (void)hipFree(cbuf_);
    if (tw_buf_) (void)hipFree(tw_buf_);
}

void Context::set_tuning(const TuneChoice &t) {
    require(t.ntt_fused == -1 || t.ntt_fused == 0 || t.ntt_fused == 1 || t.ntt_fused == 2 || t.ntt_fused == 4,
            "ntt_fused must be -1, 0, 1, 2 or 4");
    require(t.bconv_mfma >= -1 && t.bconv_mfma <= 1, "bconv_mfma must be -1, 0 or 1");
    require(t.bconv_fused >= -1 && t.bconv_fused <= 1, "bconv_fused must be -1, 0 or 1");
    require(t.ntt_chunk_limbs >= -1 && t.ntt_fused_min_limbs >= -1, "limb counts must be -1 or non-negative");
    require(t.ntt_cols_threads == -1 || t.ntt_cols_threads == 0 || t.ntt_cols_threads == 256 ||
                t.ntt_cols_threads == 512,
            "ntt_cols_threads must be -1, 0, 256 or 512");
    require(t.ntt_rows_threads == -1 || t.ntt_rows_threads == 128 || t.ntt_rows_threads == 256 ||
                t.ntt_rows_threads == 512,
            "ntt_rows_threads must be -1, 128, 256 or 512");
    require(t.bconv_wgs_per_cu >= -1 && t.bconv_wgs_per_cu <= 4 && t.bconv_wgs_per_cu != 3,
            "bconv_wgs_per_cu must be -1, 0, 1, 2 or 4");
    tune_ = t;
    tables_.fused_g = t.ntt_fused;
    tables_.fused_min = t.ntt_fused_min_limbs;
    tables_.chunk_limbs = t.ntt_chunk_limbs;
    tables_.cols_nt = t.ntt_cols_threads;
    tables_.rows_nt = t.ntt_rows_threads;
}

const HpsConsts &Context::hps_level(u32 d) const {
    require(d < hps_lv_.size(), "no HPS tables for this number of dropped levels");
    return hps_lv_[d];
}

const HpsConsts &Context::hps_drop(u32 d) const {
    require(d >= 1 && d < hps_drop_.size(), "no Q -> Ql tables for this number of dropped levels");
    return hps_drop_[d];
}

const u64 *Context::hps_expand(u32 d) const {
    require(d >= 1 && d < hps_expand_.size(), "no expansion factors for this number of dropped levels");
    return hps_expand_[d];
}

const LevelConsts &Context::level(u32 chain_index) const {
    require(chain_index >= 1 && chain_index <= size_Q_, "invalid chain index");
    return levels_[chain_index];
}

// reference: PhantomGaloisTool permutation tables (include/galois.cuh:98-113)
const u32 *Context::galois_perm(u32 elt) const {
    std::lock_guard<std::mutex> lk(g_galois_mu);
    for (auto &g : galois_tables_)
        if (g.first == elt) return g.second;
    require((elt & 1) && elt < 2 * N_, "Galois element is not valid");
    std::vector<u32> t(N_);
    for (u64 i = N_; i < 2 * N_; ++i) {
        const u64 r = hm::reverse_bits(i, (int)logN_ + 1);
        u64 idx = ((u64)elt * r) >> 1;
        idx &= (N_ - 1);
        t[i - N_] = (u32)hm::reverse_bits(idx, (int)logN_);
    }
    u32 *d = nullptr;
    PHX_HIP_CHECK(hipMalloc(&d, N_ * sizeof(u32)));
    PHX_HIP_CHECK(hipMemcpy(d, t.data(), N_ * sizeof(u32), hipMemcpyHostToDevice));
    galois_tables_.emplace_back(elt, d);
    return d;
}

}  // namespace phx
