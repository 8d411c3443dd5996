// ckks.cpp -- CKKS canonical-embedding encoder / decoder (host side).
//
// With c_k = m_k + i*m_(k+n) (k < n) and points zeta^(1+4t) = zeta * w^t,
// w = exp(2*pi*i/n), every slot is a value of the size-n transform
//   Z_t = sum_k (c_k zeta^k) w^(t k),
// because X^n = i at every point 5^j (5^j = 1 mod 4). Slot j sits at
// t(j) = (5^j mod 2N - 1) / 4. Encoding inverts the transform.
#include "ckks.h"

#include <cmath>
#include <thread>

#include "hostmath.h"

namespace phx {

namespace {

using cplx = std::complex<double>;

// in-place radix-2 transform of size n: a_t <- sum_k a_k exp(sign*2*pi*i*t*k/n)
void fft(std::vector<cplx> &a, int sign) {
    const size_t n = a.size();
    for (size_t i = 1, j = 0; i < n; ++i) {
        size_t bit = n >> 1;
        for (; j & bit; bit >>= 1) j ^= bit;
        j ^= bit;
        if (i < j) std::swap(a[i], a[j]);
    }
    for (size_t len = 2; len <= n; len <<= 1) {
        const double ang = sign * 2.0 * M_PI / (double)len;
        std::vector<cplx> w(len / 2);
        for (size_t k = 0; k < len / 2; ++k) w[k] = std::polar(1.0, ang * (double)k);
        for (size_t i = 0; i < n; i += len)
            for (size_t k = 0; k < len / 2; ++k) {
                const cplx u = a[i + k], v = a[i + k + len / 2] * w[k];
                a[i + k] = u + v;
                a[i + k + len / 2] = u - v;
            }
    }
}

// slot j -> transform index t(j)
std::vector<u32> slot_index(u64 N) {
    const u64 n = N / 2, m = 2 * N;
    std::vector<u32> idx(n);
    u64 pos = 1;
    for (u64 j = 0; j < n; ++j) {
        idx[j] = (u32)((pos - 1) / 4);
        pos = (pos * 5) & (m - 1);
    }
    return idx;
}

// |a| mod q for an integral double a >= 0
u64 mod_of_integral(double a, u64 q) {
    if (a < 0x1p64) return (u64)a % q;
    int e = 0;
    const double f = std::frexp(a, &e);        // a = f * 2^e, f in [0.5, 1)
    const u64 mant = (u64)std::ldexp(f, 53);  // exact
    return hm::mulmod(mant % q, hm::powmod(2, (u64)(e - 53), q), q);
}

template <class F>
void parallel_rows(u64 n, F &&f) {
    const unsigned T = std::max(1u, std::min(16u, std::thread::hardware_concurrency()));
    std::vector<std::thread> th;
    for (unsigned w = 0; w < T; ++w)
        th.emplace_back([&, w] {
            for (u64 i = w; i < n; i += T) f(i);
        });
    for (auto &t : th) t.join();
}

// little-endian multi-word unsigned integers
using Big = std::vector<u64>;

Big big_mul_small(const Big &a, u64 b) {
    Big r(a.size() + 1, 0);
    hm::u128 carry = 0;
    for (size_t i = 0; i < a.size(); ++i) {
        const hm::u128 t = (hm::u128)a[i] * b + carry;
        r[i] = (u64)t;
        carry = t >> 64;
    }
    r[a.size()] = (u64)carry;
    return r;
}

int big_cmp(const u64 *a, const u64 *b, size_t n) {
    for (size_t i = n; i-- > 0;)
        if (a[i] != b[i]) return a[i] < b[i] ? -1 : 1;
    return 0;
}

void big_sub(u64 *a, const u64 *b, size_t n) {   // a -= b (a >= b)
    u64 borrow = 0;
    for (size_t i = 0; i < n; ++i) {
        const u64 bi = b[i] + borrow;
        const u64 nb = (bi < borrow) || (a[i] < bi) ? 1 : 0;
        a[i] -= bi;
        borrow = nb;
    }
}

double big_to_double(const u64 *a, size_t n) {
    double r = 0;
    for (size_t i = n; i-- > 0;) r = r * 0x1p64 + (double)a[i];
    return r;
}

}  // namespace

void ckks_encode(const Context &ctx, const cplx *values, size_t count, double scale, u32 chain_index, u64 *out) {
    const u64 N = ctx.N(), n = N / 2;
    require(chain_index >= 1 && chain_index <= ctx.size_Q(), "invalid chain index");
    const u32 L = ctx.limbs_at(chain_index);
    require(count > 0, "Input vector is empty");
    require(count <= n, "Input vector exceeds max slots");
    int total_bits = 0;
    for (u32 i = 0; i < L; ++i) total_bits += hm::bit_count(ctx.moduli()[i]);
    require(scale > 0 && (int)std::log2(scale) + 1 < total_bits, "scale out of bounds");
    const std::vector<u32> idx = slot_index(N);
    std::vector<cplx> z(n, 0.0);
    for (size_t j = 0; j < count; ++j) z[idx[j]] = values[j];
    fft(z, -1);
    std::vector<double> m(N);
    double max_coeff = 0;
    for (u64 k = 0; k < n; ++k) {
        // c_k = (1/n) * IDFT_k * zeta^-k, scaled
        const cplx c = z[k] * std::polar(scale / (double)n, -M_PI * (double)k / (double)N);
        m[k] = std::round(c.real());
        m[k + n] = std::round(c.imag());
        max_coeff = std::max(max_coeff, std::max(std::fabs(m[k]), std::fabs(m[k + n])));
    }
    const int max_bits = (int)std::ceil(std::log2(std::max(max_coeff, 1.0))) + 1;
    require(max_bits < total_bits, "encoded values are too large");
    parallel_rows(L, [&](u64 i) {
        const u64 q = ctx.moduli()[i];
        for (u64 k = 0; k < N; ++k) {
            const u64 r = mod_of_integral(std::fabs(m[k]), q);
            out[i * N + k] = (m[k] < 0 && r) ? q - r : r;
        }
    });
}

void ckks_decode(const Context &ctx, const u64 *in, u32 chain_index, double scale, cplx *values) {
    const u64 N = ctx.N(), n = N / 2;
    require(chain_index >= 1 && chain_index <= ctx.size_Q(), "invalid chain index");
    require(scale > 0, "scale out of bounds");
    const u32 L = ctx.limbs_at(chain_index);
    const std::vector<u64> &q = ctx.moduli();
    // CRT constants: Q, Q/q_i and [(Q/q_i)^-1]_{q_i}
    Big Q{1};
    for (u32 i = 0; i < L; ++i) Q = big_mul_small(Q, q[i]);
    const size_t W = Q.size();
    std::vector<Big> punct(L);
    std::vector<u64> pinv(L);
    for (u32 i = 0; i < L; ++i) {
        Big p{1};
        for (u32 k = 0; k < L; ++k)
            if (k != i) p = big_mul_small(p, q[k]);
        p.resize(W, 0);
        punct[i] = p;
        pinv[i] = hm::invmod(hm::punctured_mod(q.data(), L, i, q[i]), q[i]);
    }
    Big half = Q;   // floor(Q / 2)
    for (size_t w = 0; w < W; ++w) half[w] = (half[w] >> 1) | (w + 1 < W ? half[w + 1] << 63 : 0);
    // centred value of every coefficient, divided by scale
    std::vector<double> coef(N);
    parallel_rows(N, [&](u64 k) {
        Big S(W + 1, 0);
        double frac = 0;
        for (u32 i = 0; i < L; ++i) {
            const u64 y = hm::mulmod(in[(u64)i * N + k], pinv[i], q[i]);
            frac += (double)y / (double)q[i];
            hm::u128 carry = 0;
            for (size_t w = 0; w < W; ++w) {
                const hm::u128 t = (hm::u128)punct[i][w] * y + S[w] + carry;
                S[w] = (u64)t;
                carry = t >> 64;
            }
            S[W] += (u64)carry;
        }
        // S = x + kq * Q with kq = floor(frac) up to one; reduce into [0, Q)
        u64 kq = (u64)std::floor(frac);
        kq = kq ? kq - 1 : 0;
        Big KQ = big_mul_small(Q, kq);
        KQ.resize(W + 1, 0);
        big_sub(S.data(), KQ.data(), W + 1);
        Big Qe = Q;
        Qe.resize(W + 1, 0);
        while (big_cmp(S.data(), Qe.data(), W + 1) >= 0) big_sub(S.data(), Qe.data(), W + 1);
        double v;
        if (big_cmp(S.data(), half.data(), W) > 0 || S[W]) {
            Big t = Q;
            big_sub(t.data(), S.data(), W);
            v = -big_to_double(t.data(), W);
        } else {
            v = big_to_double(S.data(), W);
        }
        coef[k] = v / scale;
    });
    std::vector<cplx> z(n);
    for (u64 k = 0; k < n; ++k)
        z[k] = cplx(coef[k], coef[k + n]) * std::polar(1.0, M_PI * (double)k / (double)N);
    fft(z, +1);
    const std::vector<u32> idx = slot_index(N);
    for (u64 j = 0; j < n; ++j) values[j] = z[idx[j]];
}

}  // namespace phx
