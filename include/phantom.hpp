/*
 * phantom.hpp -- C++ facade with the reference's class names and evaluator
 * signatures, over the C ABI of phantom_amd.h (header only).
 *
 * A PhantomFHE program that uses the hot path keeps its calls:
 *   phantom::EncryptionParameters, phantom::arith::CoeffModulus::Create
 *       (include/host/encryptionparams.h:37-245, include/host/modulus.h:295)
 *   PhantomContext (include/context.cuh:133-273), PhantomCiphertext (deep copy,
 *       include/ciphertext.h:7-308), PhantomRelinKey / PhantomGaloisKey (move-only,
 *       include/secretkey.h:102-220), PhantomSecretKey (include/secretkey.h:226-389)
 *   phantom::multiply_and_relin_inplace, relinearize_inplace, rescale_to_next,
 *       mod_switch_to_next, apply_galois_inplace, rotate_inplace, hoisting_inplace,
 *       keyswitch_inplace (include/evaluate.cuh:29-32, 104-238) and the 2-D NTT
 *       entry points (include/ntt.cuh:173-226, a context in place of the DNTTTable)
 * Errors are rethrown as the reference throws them: std::invalid_argument for
 * parameter errors, std::logic_error, std::runtime_error for HIP failures, with
 * the library's message.
 *
 * Differences, each forced by the C ABI owning the device state: save / load of
 * keys take the context (the library checks the key's shape against it), and
 * PhantomCiphertext::load takes the context for the same reason. Host-buffer
 * upload / download and keys from host arrays are additions (the reference fills
 * ciphertexts through encryption or its stream format). Streams: calls without a
 * stream run on the per-thread default stream, as the reference's do on
 * cudaStreamPerThread.
 */
#ifndef PHANTOM_AMD_FACADE_HPP
#define PHANTOM_AMD_FACADE_HPP

#include <hip/hip_runtime_api.h>

#include <cstdint>
#include <istream>
#include <ostream>
#include <stdexcept>
#include <string>
#include <utility>
#include <vector>

#include "phantom_amd.h"

namespace phantom {

namespace detail {
inline void check(int rc) {
    if (rc == PHX_OK) return;
    const std::string msg = phx_last_error();
    if (rc == PHX_EINVAL) throw std::invalid_argument(msg);
    if (rc == PHX_ELOGIC) throw std::logic_error(msg);
    throw std::runtime_error(msg);
}
inline phx_stream st(const hipStream_t &s) { return static_cast<phx_stream>(s); }
}  // namespace detail

enum class scheme_type : std::uint8_t { none = 0, bgv = 1, bfv = 2, ckks = 3 };
enum class mul_tech_type : std::uint8_t { none = 0, behz = 1, hps = 2, hps_overq = 3, hps_overq_leveled = 4 };

namespace arith {
class Modulus {
public:
    Modulus(std::uint64_t value = 0) : value_(value) {}
    std::uint64_t value() const noexcept { return value_; }
    bool is_zero() const noexcept { return value_ == 0; }

private:
    std::uint64_t value_;
};

struct CoeffModulus {
    // the largest primes of the given bit sizes with q = 1 mod 2N, as the reference orders them
    static std::vector<Modulus> Create(std::size_t poly_modulus_degree, const std::vector<int> &bit_sizes) {
        std::vector<std::uint64_t> v(bit_sizes.size());
        detail::check(phx_coeff_modulus_create(poly_modulus_degree, bit_sizes.data(), bit_sizes.size(), v.data()));
        return std::vector<Modulus>(v.begin(), v.end());
    }
};
}  // namespace arith

class EncryptionParameters {
public:
    explicit EncryptionParameters(scheme_type scheme = scheme_type::none)
        : scheme_(scheme), mul_tech_(scheme == scheme_type::bfv ? mul_tech_type::hps : mul_tech_type::none) {}
    void set_mul_tech(mul_tech_type t) {
        if (scheme_ == scheme_type::bfv) mul_tech_ = t;
    }
    void set_poly_modulus_degree(std::size_t n) { poly_modulus_degree_ = n; }
    void set_coeff_modulus(const std::vector<arith::Modulus> &m) { coeff_modulus_ = m; }
    void set_special_modulus_size(std::size_t n) { special_modulus_size_ = n; }
    void set_galois_elts(const std::vector<std::uint32_t> &e) { galois_elts_ = e; }
    void set_plain_modulus(const arith::Modulus &t) { plain_modulus_ = t; }
    scheme_type scheme() const noexcept { return scheme_; }
    mul_tech_type mul_tech() const noexcept { return mul_tech_; }
    std::size_t poly_modulus_degree() const noexcept { return poly_modulus_degree_; }
    const std::vector<arith::Modulus> &coeff_modulus() const noexcept { return coeff_modulus_; }
    std::size_t special_modulus_size() const noexcept { return special_modulus_size_; }
    const std::vector<std::uint32_t> &galois_elts() const noexcept { return galois_elts_; }
    const arith::Modulus &plain_modulus() const noexcept { return plain_modulus_; }

private:
    scheme_type scheme_;
    mul_tech_type mul_tech_;
    std::size_t poly_modulus_degree_ = 0;
    std::vector<arith::Modulus> coeff_modulus_;
    std::size_t special_modulus_size_ = 1;
    std::vector<std::uint32_t> galois_elts_;
    arith::Modulus plain_modulus_;
};

}  // namespace phantom

class PhantomContext {
public:
    explicit PhantomContext(const phantom::EncryptionParameters &p) : galois_elts_(p.galois_elts()) {
        std::vector<std::uint64_t> q;
        for (const auto &m : p.coeff_modulus()) q.push_back(m.value());
        phx_params c{};
        c.scheme = static_cast<int>(p.scheme());
        c.mul_tech = static_cast<int>(p.mul_tech());
        c.poly_modulus_degree = p.poly_modulus_degree();
        c.coeff_modulus = q.data();
        c.coeff_modulus_size = q.size();
        c.special_modulus_size = p.special_modulus_size();
        c.plain_modulus = p.plain_modulus().value();
        c.galois_elts = galois_elts_.empty() ? nullptr : galois_elts_.data();
        c.galois_elts_size = galois_elts_.size();
        phantom::detail::check(phx_context_create(&c, &h_));
        phantom::detail::check(phx_context_info(h_, &poly_degree_, &size_Q_, &size_P_, &dnum_));
    }
    PhantomContext(const PhantomContext &) = delete;
    void operator=(const PhantomContext &) = delete;
    ~PhantomContext() { phx_context_destroy(h_); }

    phx_context *handle() const noexcept { return h_; }
    std::size_t poly_degree() const noexcept { return poly_degree_; }
    std::size_t size_Q() const noexcept { return size_Q_; }
    std::size_t size_P() const noexcept { return size_P_; }
    std::size_t dnum() const noexcept { return dnum_; }
    // limbs of chain index c (context_data_[c]: 0 = the key level QP, 1 = the first data level)
    std::size_t limbs_at(std::size_t chain_index) const {
        std::size_t l = 0;
        phantom::detail::check(phx_context_limbs_at(h_, chain_index, &l));
        return l;
    }
    const std::vector<std::uint32_t> &galois_elts() const noexcept { return galois_elts_; }

private:
    phx_context *h_ = nullptr;
    std::uint64_t poly_degree_ = 0;
    std::size_t size_Q_ = 0, size_P_ = 0, dnum_ = 0;
    std::vector<std::uint32_t> galois_elts_;
};

class PhantomCiphertext {
public:
    PhantomCiphertext() { phantom::detail::check(phx_ciphertext_create(&h_)); }
    PhantomCiphertext(const PhantomCiphertext &o) { phantom::detail::check(phx_ciphertext_clone(o.h_, &h_, nullptr)); }
    PhantomCiphertext &operator=(const PhantomCiphertext &o) {
        if (this != &o) {
            phx_ciphertext *n = nullptr;
            phantom::detail::check(phx_ciphertext_clone(o.h_, &n, nullptr));
            phx_ciphertext_destroy(h_);
            h_ = n;
        }
        return *this;
    }
    PhantomCiphertext(PhantomCiphertext &&o) noexcept : h_(o.h_) { o.h_ = nullptr; }
    PhantomCiphertext &operator=(PhantomCiphertext &&o) noexcept {
        std::swap(h_, o.h_);
        return *this;
    }
    ~PhantomCiphertext() {
        if (h_) phx_ciphertext_destroy(h_);
    }

    void resize(const PhantomContext &context, std::size_t chain_index, std::size_t size, const hipStream_t &stream) {
        phantom::detail::check(phx_ciphertext_resize(context.handle(), h_, chain_index, size, phantom::detail::st(stream)));
    }
    std::size_t chain_index() const { return info().chain; }
    std::size_t size() const { return info().size; }
    std::size_t coeff_modulus_size() const { return info().limbs; }
    std::size_t poly_modulus_degree() const {
        // the stream format is a 58-byte header and size x limbs x N words
        std::size_t n = 0;
        phantom::detail::check(phx_ciphertext_save_size(h_, &n));
        const Info i = info();
        return i.size != 0 && i.limbs != 0 ? (n - 58) / 8 / (i.size * i.limbs) : 0;
    }
    double scale() const { return info().scale; }
    bool is_ntt_form() const { return info().ntt != 0; }
    void set_scale(double s) { phantom::detail::check(phx_ciphertext_set_scale(h_, s)); }
    void set_ntt_form(bool v) { phantom::detail::check(phx_ciphertext_set_ntt_form(h_, v ? 1 : 0)); }
    // device buffer [size][coeff_modulus_size][N]
    std::uint64_t *data() const {
        std::uint64_t *p = nullptr;
        phantom::detail::check(phx_ciphertext_data(h_, &p));
        return p;
    }
    // the reference's stream format (include/ciphertext.h:173-216)
    void save(std::ostream &stream) const {
        std::size_t n = 0;
        phantom::detail::check(phx_ciphertext_save_size(h_, &n));
        std::string buf(n, '\0');
        phantom::detail::check(phx_ciphertext_save(h_, &buf[0], n));
        stream.write(buf.data(), (std::streamsize)n);
    }
    void load(const PhantomContext &context, std::istream &stream) {
        std::string buf((std::istreambuf_iterator<char>(stream)), std::istreambuf_iterator<char>());
        phantom::detail::check(phx_ciphertext_load(context.handle(), h_, buf.data(), buf.size()));
    }
    // additions: host residues in the device layout
    void upload(const std::vector<std::uint64_t> &host) {
        phantom::detail::check(phx_ciphertext_upload(h_, host.data(), host.size(), nullptr));
    }
    std::vector<std::uint64_t> download() const {
        const Info i = info();
        std::vector<std::uint64_t> out(i.size * i.limbs * poly_modulus_degree());
        phantom::detail::check(phx_ciphertext_download(h_, out.data(), out.size(), nullptr));
        return out;
    }
    phx_ciphertext *handle() const noexcept { return h_; }

private:
    struct Info {
        std::size_t chain = 0, size = 0, limbs = 0;
        double scale = 1.0;
        int ntt = 0;
    };
    Info info() const {
        Info i;
        phantom::detail::check(phx_ciphertext_info(h_, &i.chain, &i.size, &i.limbs, &i.scale, &i.ntt));
        return i;
    }
    phx_ciphertext *h_ = nullptr;
};

class PhantomRelinKey {
public:
    PhantomRelinKey() = default;
    PhantomRelinKey(const PhantomRelinKey &) = delete;
    PhantomRelinKey &operator=(const PhantomRelinKey &) = delete;
    PhantomRelinKey(PhantomRelinKey &&o) noexcept : h_(o.h_) { o.h_ = nullptr; }
    PhantomRelinKey &operator=(PhantomRelinKey &&o) noexcept {
        std::swap(h_, o.h_);
        return *this;
    }
    ~PhantomRelinKey() {
        if (h_) phx_relin_key_destroy(h_);
    }
    // addition: a key laid out [dnum][2][QP][N] (the reference's public_keys_ layout)
    static PhantomRelinKey from_host(const PhantomContext &context, const std::uint64_t *host, std::size_t dnum) {
        PhantomRelinKey k;
        phantom::detail::check(phx_relin_key_create_from_host(context.handle(), host, dnum, &k.h_));
        return k;
    }
    void save(const PhantomContext &context, std::ostream &stream) const {
        std::size_t n = 0;
        phantom::detail::check(phx_relin_key_save_size(context.handle(), h_, &n));
        std::string buf(n, '\0');
        phantom::detail::check(phx_relin_key_save(context.handle(), h_, &buf[0], n));
        stream.write(buf.data(), (std::streamsize)n);
    }
    void load(const PhantomContext &context, std::istream &stream) {
        std::string buf((std::istreambuf_iterator<char>(stream)), std::istreambuf_iterator<char>());
        phx_relin_key *k = nullptr;
        phantom::detail::check(phx_relin_key_load(context.handle(), buf.data(), buf.size(), &k));
        if (h_) phx_relin_key_destroy(h_);
        h_ = k;
    }
    const phx_relin_key *handle() const noexcept { return h_; }

private:
    friend class PhantomSecretKey;
    phx_relin_key *h_ = nullptr;
};

class PhantomGaloisKey {
public:
    PhantomGaloisKey() = default;
    PhantomGaloisKey(const PhantomGaloisKey &) = delete;
    PhantomGaloisKey &operator=(const PhantomGaloisKey &) = delete;
    PhantomGaloisKey(PhantomGaloisKey &&o) noexcept : h_(o.h_) { o.h_ = nullptr; }
    PhantomGaloisKey &operator=(PhantomGaloisKey &&o) noexcept {
        std::swap(h_, o.h_);
        return *this;
    }
    ~PhantomGaloisKey() {
        if (h_) phx_galois_key_destroy(h_);
    }
    // addition: keys [n][dnum][2][QP][N] for the listed Galois elements
    static PhantomGaloisKey from_host(const PhantomContext &context, const std::vector<std::uint32_t> &elts,
                                      const std::uint64_t *host, std::size_t dnum) {
        PhantomGaloisKey k;
        phantom::detail::check(
            phx_galois_key_create_from_host(context.handle(), elts.data(), elts.size(), host, dnum, &k.h_));
        return k;
    }
    const phx_galois_key *handle() const noexcept { return h_; }

private:
    friend class PhantomSecretKey;
    phx_galois_key *h_ = nullptr;
};

class PhantomSecretKey {
public:
    explicit PhantomSecretKey(const PhantomContext &context, std::uint64_t seed = 1) {
        phantom::detail::check(phx_secret_key_create(context.handle(), seed, &h_));
    }
    PhantomSecretKey(const PhantomSecretKey &) = delete;
    PhantomSecretKey &operator=(const PhantomSecretKey &) = delete;
    ~PhantomSecretKey() { phx_secret_key_destroy(h_); }
    PhantomRelinKey gen_relinkey(const PhantomContext &context, std::uint64_t seed = 2) {
        PhantomRelinKey k;
        phantom::detail::check(phx_relin_key_create(context.handle(), h_, seed, &k.h_));
        return k;
    }
    // keys for the context's Galois elements (EncryptionParameters::set_galois_elts)
    PhantomGaloisKey create_galois_keys(const PhantomContext &context, std::uint64_t seed = 3) const {
        PhantomGaloisKey k;
        const auto &e = context.galois_elts();
        phantom::detail::check(phx_galois_key_create(context.handle(), h_, e.data(), e.size(), seed, &k.h_));
        return k;
    }

private:
    phx_secret_key *h_ = nullptr;
};

namespace phantom {

// ---- the hot path (include/evaluate.cuh) ----
inline void multiply_inplace(const PhantomContext &context, PhantomCiphertext &encrypted1,
                             const PhantomCiphertext &encrypted2) {
    detail::check(phx_multiply_inplace(context.handle(), encrypted1.handle(), encrypted2.handle(), nullptr));
}
inline auto multiply(const PhantomContext &context, const PhantomCiphertext &encrypted1,
                     const PhantomCiphertext &encrypted2) {
    PhantomCiphertext destination = encrypted1;
    multiply_inplace(context, destination, encrypted2);
    return destination;
}
inline void multiply_and_relin_inplace(const PhantomContext &context, PhantomCiphertext &encrypted1,
                                       const PhantomCiphertext &encrypted2, const PhantomRelinKey &relin_keys) {
    detail::check(phx_multiply_and_relin_inplace(context.handle(), encrypted1.handle(), encrypted2.handle(),
                                                 relin_keys.handle(), nullptr));
}
inline auto multiply_and_relin(const PhantomContext &context, const PhantomCiphertext &encrypted1,
                               const PhantomCiphertext &encrypted2, const PhantomRelinKey &relin_keys) {
    PhantomCiphertext destination = encrypted1;
    multiply_and_relin_inplace(context, destination, encrypted2, relin_keys);
    return destination;
}
inline void relinearize_inplace(const PhantomContext &context, PhantomCiphertext &encrypted,
                                const PhantomRelinKey &relin_keys) {
    detail::check(phx_relinearize_inplace(context.handle(), encrypted.handle(), relin_keys.handle(), nullptr));
}
inline auto relinearize(const PhantomContext &context, const PhantomCiphertext &encrypted,
                        const PhantomRelinKey &relin_keys) {
    PhantomCiphertext destination = encrypted;
    relinearize_inplace(context, destination, relin_keys);
    return destination;
}
[[nodiscard]] inline PhantomCiphertext rescale_to_next(const PhantomContext &context,
                                                       const PhantomCiphertext &encrypted) {
    PhantomCiphertext destination;
    detail::check(phx_rescale_to_next(context.handle(), encrypted.handle(), destination.handle(), nullptr));
    return destination;
}
inline void rescale_to_next_inplace(const PhantomContext &context, PhantomCiphertext &encrypted) {
    encrypted = rescale_to_next(context, encrypted);
}
[[nodiscard]] inline PhantomCiphertext mod_switch_to_next(const PhantomContext &context,
                                                          const PhantomCiphertext &encrypted) {
    PhantomCiphertext destination;
    detail::check(phx_mod_switch_to_next(context.handle(), encrypted.handle(), destination.handle(), nullptr));
    return destination;
}
inline void mod_switch_to_next_inplace(const PhantomContext &context, PhantomCiphertext &encrypted) {
    encrypted = mod_switch_to_next(context, encrypted);
}
inline void mod_switch_to_inplace(const PhantomContext &context, PhantomCiphertext &encrypted,
                                  std::size_t chain_index) {
    detail::check(phx_mod_switch_to_inplace(context.handle(), encrypted.handle(), chain_index, nullptr));
}
inline auto mod_switch_to(const PhantomContext &context, const PhantomCiphertext &encrypted,
                          std::size_t chain_index) {
    PhantomCiphertext destination = encrypted;
    mod_switch_to_inplace(context, destination, chain_index);
    return destination;
}
inline void apply_galois_inplace(const PhantomContext &context, PhantomCiphertext &encrypted,
                                 std::size_t galois_elt, const PhantomGaloisKey &galois_keys) {
    if (galois_elt > UINT32_MAX) throw std::invalid_argument("Galois elt not present");
    detail::check(phx_apply_galois_inplace(context.handle(), encrypted.handle(), (std::uint32_t)galois_elt,
                                           galois_keys.handle(), nullptr));
}
inline auto apply_galois(const PhantomContext &context, const PhantomCiphertext &encrypted, std::size_t galois_elt,
                         const PhantomGaloisKey &galois_keys) {
    PhantomCiphertext destination = encrypted;
    apply_galois_inplace(context, destination, galois_elt, galois_keys);
    return destination;
}
inline void rotate_inplace(const PhantomContext &context, PhantomCiphertext &encrypted, int step,
                           const PhantomGaloisKey &galois_key) {
    detail::check(phx_rotate_inplace(context.handle(), encrypted.handle(), step, galois_key.handle(), nullptr));
}
inline auto rotate(const PhantomContext &context, const PhantomCiphertext &encrypted, int step,
                   const PhantomGaloisKey &galois_key) {
    PhantomCiphertext destination = encrypted;
    rotate_inplace(context, destination, step, galois_key);
    return destination;
}
inline void hoisting_inplace(const PhantomContext &context, PhantomCiphertext &ct, const PhantomGaloisKey &glk,
                             const std::vector<int> &steps) {
    detail::check(phx_hoisting_inplace(context.handle(), ct.handle(), glk.handle(), steps.data(), steps.size(),
                                       nullptr));
}
inline auto hoisting(const PhantomContext &context, const PhantomCiphertext &encrypted, const PhantomGaloisKey &glk,
                     const std::vector<int> &steps) {
    PhantomCiphertext destination = encrypted;
    hoisting_inplace(context, destination, glk, steps);
    return destination;
}
// encrypted[0..1] += keyswitch(c2), c2 = [Ql][N] device residues in NTT form
inline void keyswitch_inplace(const PhantomContext &context, PhantomCiphertext &encrypted, std::uint64_t *c2,
                              const PhantomRelinKey &relin_keys, bool is_relin, const hipStream_t &stream) {
    detail::check(phx_keyswitch_inplace(context.handle(), encrypted.handle(), c2, relin_keys.handle(),
                                        is_relin ? 1 : 0, detail::st(stream)));
}

// ---- the 2-D NTT (include/ntt.cuh:173-226): limbs [0, L) of inout use QP rows [start, start + L) ----
inline void nwt_2d_radix8_forward_inplace(std::uint64_t *inout, const PhantomContext &context, std::size_t L,
                                          std::size_t start, const hipStream_t &stream) {
    detail::check(phx_ntt_forward(context.handle(), inout, L, start, detail::st(stream)));
}
inline void nwt_2d_radix8_backward_inplace(std::uint64_t *inout, const PhantomContext &context, std::size_t L,
                                           std::size_t start, const hipStream_t &stream) {
    detail::check(phx_ntt_inverse(context.handle(), inout, L, start, detail::st(stream)));
}

// get_elt_from_step (include/galois.cuh:16-49)
inline std::uint32_t get_elt_from_step(int step, std::size_t poly_modulus_degree) {
    std::uint32_t e = 0;
    detail::check(phx_galois_elt_from_step(step, poly_modulus_degree, &e));
    return e;
}

}  // namespace phantom

#endif /* PHANTOM_AMD_FACADE_HPP */
