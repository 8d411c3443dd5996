// hostmath.cpp -- host parameter maths (see hostmath.h).
#include "hostmath.h"

#include <map>
#include <stdexcept>

namespace phx {
namespace hm {

u64 powmod(u64 a, u64 e, u64 q) {
    u64 r = 1 % q;
    a %= q;
    while (e) {
        if (e & 1) r = mulmod(r, a, q);
        a = mulmod(a, a, q);
        e >>= 1;
    }
    return r;
}

u64 invmod(u64 a, u64 q) {
    // extended Euclid (q need not be prime, but gcd(a, q) must be 1)
    __int128 t = 0, nt = 1, r = q, nr = a % q;
    while (nr != 0) {
        __int128 quo = r / nr;
        __int128 tmp = t - quo * nt;
        t = nt;
        nt = tmp;
        tmp = r - quo * nr;
        r = nr;
        nr = tmp;
    }
    if (r != 1) throw std::invalid_argument("value is not invertible modulo q");
    if (t < 0) t += q;
    return (u64)t;
}

bool is_prime(u64 n) {
    static const u64 small[] = {2, 3, 5, 7, 11, 13, 17, 19, 23, 29, 31, 37};
    if (n < 2) return false;
    for (u64 p : small) {
        if (n == p) return true;
        if (n % p == 0) return false;
    }
    u64 d = n - 1;
    int s = 0;
    while (!(d & 1)) { d >>= 1; ++s; }
    for (u64 a : small) {
        u64 x = powmod(a, d, n);
        if (x == 1 || x == n - 1) continue;
        bool composite = true;
        for (int i = 1; i < s; ++i) {
            x = mulmod(x, x, n);
            if (x == n - 1) { composite = false; break; }
        }
        if (composite) return false;
    }
    return true;
}

int bit_count(u64 v) {
    int b = 0;
    while (v) { ++b; v >>= 1; }
    return b;
}

u64 reverse_bits(u64 x, int bits) {
    u64 r = 0;
    for (int i = 0; i < bits; ++i) { r = (r << 1) | (x & 1); x >>= 1; }
    return r;
}

void barrett_ratio(u64 q, u64 &lo, u64 &hi) {
    // floor(2^128 / q): long division of (1, 0, 0) words by q
    u128 r = 1;  // remainder after the top word
    hi = (u64)(((u128)1 << 64) / q);
    r = ((u128)1 << 64) % q;
    lo = (u64)((r << 64) / q);
}

std::vector<u64> get_primes(u64 N, int bits, size_t count) {
    std::vector<u64> out;
    const u64 step = 2 * N;
    u64 v = ((u64)1 << bits) - step + 1;
    const u64 lower = (u64)1 << (bits - 1);
    while (out.size() < count && v > lower) {
        if (is_prime(v)) out.push_back(v);
        v -= step;
    }
    if (out.size() < count) throw std::logic_error("failed to find enough qualifying primes");
    return out;
}

std::vector<u64> coeff_modulus_create(u64 N, const std::vector<int> &bits) {
    std::map<int, size_t> count;
    for (int b : bits) {
        if (b < 2 || b > 60) throw std::invalid_argument("bit_sizes is invalid");
        ++count[b];
    }
    std::map<int, std::vector<u64>> pool;
    for (auto &kv : count) pool[kv.first] = get_primes(N, kv.first, kv.second);
    std::vector<u64> out;
    for (int b : bits) {
        out.push_back(pool[b].back());
        pool[b].pop_back();
    }
    return out;
}

u64 minimal_primitive_root(u64 q, u64 degree) {
    if ((q - 1) % degree) throw std::invalid_argument("invalid modulus for the NTT");
    const u64 quot = (q - 1) / degree;
    u64 g = 0;
    for (u64 x = 2; x < q && !g; ++x) {
        u64 c = powmod(x, quot, q);
        if (powmod(c, degree >> 1, q) == q - 1) g = c;
    }
    if (!g) throw std::invalid_argument("no primitive root");
    // every primitive degree-th root is an odd power of g
    const u64 g2 = mulmod(g, g, q);
    u64 best = g, cur = g;
    for (u64 i = 0; i < degree / 2; ++i) {
        if (cur < best) best = cur;
        cur = mulmod(cur, g2, q);
    }
    return best;
}

NttTable make_ntt_table(u64 q, int logN) {
    NttTable t;
    const u64 N = (u64)1 << logN;
    t.q = q;
    t.psi = minimal_primitive_root(q, 2 * N);
    const u64 ipsi = invmod(t.psi, q);
    t.fw.assign(N, 0);
    t.iw.assign(N, 0);
    u64 p = 1, ip = 1;
    for (u64 i = 0; i < N; ++i) {
        const u64 r = reverse_bits(i, logN);
        t.fw[r] = p;
        t.iw[r] = ip;
        p = mulmod(p, t.psi, q);
        ip = mulmod(ip, ipsi, q);
    }
    t.fws.resize(N);
    t.iws.resize(N);
    for (u64 i = 0; i < N; ++i) {
        t.fws[i] = shoup(t.fw[i], q);
        t.iws[i] = shoup(t.iw[i], q);
    }
    t.ninv = invmod(N % q, q);
    return t;
}

u64 punctured_mod(const u64 *base, size_t n, size_t skip, u64 m) {
    u64 r = 1 % m;
    for (size_t k = 0; k < n; ++k)
        if (k != skip) r = mulmod(r, base[k] % m, m);
    return r;
}

u64 product_mod(const u64 *base, size_t n, u64 m) {
    u64 r = 1 % m;
    for (size_t k = 0; k < n; ++k) r = mulmod(r, base[k] % m, m);
    return r;
}

}  // namespace hm
}  // namespace phx
