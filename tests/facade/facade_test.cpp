// facade_test.cpp -- drives the hot path through the C++ facade (include/phantom.hpp)
// exactly as a PhantomFHE program calls it, on seeded residues, and writes the
// inputs and outputs as raw u64 files for tests/test_gpu_facade.py to compare
// with the oracle. Checks made here: deep-copy semantics of PhantomCiphertext,
// multiply + relinearize == multiply_and_relin, and the exception types of
// parameter errors. Prints "facade: ok" on success.
//
// usage: facade_test <output directory>
#include <cstdio>
#include <fstream>
#include <sstream>
#include <string>
#include <vector>

#include "phantom.hpp"

using u64 = std::uint64_t;

namespace {
u64 splitmix(u64 &s) {
    u64 z = (s += 0x9e3779b97f4a7c15ull);
    z = (z ^ (z >> 30)) * 0xbf58476d1ce4e5b9ull;
    z = (z ^ (z >> 27)) * 0x94d049bb133111ebull;
    return z ^ (z >> 31);
}

// count blocks of N residues, block b reduced by mods[b % mods_per_block]
std::vector<u64> residues(u64 &seed, const std::vector<u64> &mods, std::size_t blocks, std::size_t N) {
    std::vector<u64> v(blocks * N);
    for (std::size_t b = 0; b < blocks; ++b)
        for (std::size_t n = 0; n < N; ++n) v[b * N + n] = splitmix(seed) % mods[b % mods.size()];
    return v;
}

void write(const std::string &dir, const std::string &name, const std::vector<u64> &v) {
    std::ofstream f(dir + "/" + name, std::ios::binary);
    f.write(reinterpret_cast<const char *>(v.data()), (std::streamsize)(v.size() * 8));
    if (!f) throw std::runtime_error("cannot write " + name);
}

int fails = 0;
void expect(bool c, const char *what) {
    if (!c) {
        std::fprintf(stderr, "facade: FAILED %s\n", what);
        ++fails;
    }
}

template <class E, class F>
void expect_throw(F &&f, const char *what) {
    try {
        f();
    } catch (const E &) {
        return;
    } catch (const std::exception &e) {
        std::fprintf(stderr, "facade: %s threw the wrong type: %s\n", what, e.what());
        ++fails;
        return;
    }
    std::fprintf(stderr, "facade: %s did not throw\n", what);
    ++fails;
}
}  // namespace

int main(int argc, char **argv) {
    if (argc != 2) {
        std::fprintf(stderr, "usage: facade_test <output directory>\n");
        return 2;
    }
    const std::string dir = argv[1];
    try {
        const std::size_t N = 4096, P = 2;
        auto coeff = phantom::arith::CoeffModulus::Create(N, {60, 50, 50, 50, 50, 60, 60});
        std::vector<u64> mods;
        for (const auto &m : coeff) mods.push_back(m.value());
        const std::vector<int> steps = {1, 2, -1};
        std::vector<std::uint32_t> elts;
        for (int s : steps) elts.push_back(phantom::get_elt_from_step(s, N));
        phantom::EncryptionParameters parms(phantom::scheme_type::ckks);
        parms.set_poly_modulus_degree(N);
        parms.set_coeff_modulus(coeff);
        parms.set_special_modulus_size(P);
        parms.set_galois_elts(elts);
        PhantomContext context(parms);
        const std::size_t Q = context.size_Q(), QP = mods.size(), dnum = context.dnum();
        const std::vector<u64> qmods(mods.begin(), mods.begin() + (long)Q);

        u64 seed = 2024;
        const auto a = residues(seed, qmods, 2 * Q, N), b = residues(seed, qmods, 2 * Q, N);
        const auto rk_host = residues(seed, mods, dnum * 2 * QP, N);
        const auto gk_host = residues(seed, mods, elts.size() * dnum * 2 * QP, N);
        const auto c2_host = residues(seed, qmods, Q, N);
        write(dir, "moduli.bin", mods);
        write(dir, "ct1.bin", a);
        write(dir, "ct2.bin", b);
        write(dir, "rk.bin", rk_host);
        write(dir, "gk.bin", gk_host);
        write(dir, "c2.bin", c2_host);
        write(dir, "elts.bin", std::vector<u64>(elts.begin(), elts.end()));

        auto rk = PhantomRelinKey::from_host(context, rk_host.data(), dnum);
        auto gk = PhantomGaloisKey::from_host(context, elts, gk_host.data(), dnum);
        PhantomRelinKey rk_moved = std::move(rk);   // keys are move-only
        PhantomCiphertext ct1, ct2;
        ct1.resize(context, 1, 2, nullptr);
        ct2.resize(context, 1, 2, nullptr);
        ct1.upload(a);
        ct2.upload(b);
        ct1.set_scale(1099511627776.0);
        ct2.set_scale(1099511627776.0);

        // multiply_and_relin (a deep copy: ct1 unchanged)
        PhantomCiphertext mr = phantom::multiply_and_relin(context, ct1, ct2, rk_moved);
        expect(ct1.download() == a, "multiply_and_relin left its operand unchanged");
        expect(mr.size() == 2 && mr.chain_index() == 1, "multiply_and_relin result shape");
        write(dir, "mr.bin", mr.download());
        // multiply then relinearize equals the fused call
        PhantomCiphertext m3 = ct1;
        phantom::multiply_inplace(context, m3, ct2);
        expect(m3.size() == 3, "multiply_inplace gives three polynomials");
        phantom::relinearize_inplace(context, m3, rk_moved);
        expect(m3.download() == mr.download(), "multiply + relinearize == multiply_and_relin");
        // rescale then rotate by 1
        PhantomCiphertext rs = phantom::rescale_to_next(context, mr);
        expect(rs.chain_index() == 2 && rs.coeff_modulus_size() == Q - 1, "rescale_to_next drops a limb");
        write(dir, "rs.bin", rs.download());
        phantom::rotate_inplace(context, rs, 1, gk);
        write(dir, "rot.bin", rs.download());
        // apply_galois with the element of step -1
        PhantomCiphertext gal = phantom::apply_galois(context, ct1, elts[2], gk);
        write(dir, "gal.bin", gal.download());
        // hoisting: the sum of the rotations by 1 and 2
        PhantomCiphertext h = phantom::hoisting(context, ct1, gk, {1, 2});
        write(dir, "hoist.bin", h.download());
        // keyswitch_inplace of an explicit c2 (device), added to a copy of ct1
        std::uint64_t *c2_dev = nullptr;
        phantom::detail::check(phx_malloc(c2_host.size() * 8, reinterpret_cast<void **>(&c2_dev)));
        phantom::detail::check(phx_memcpy_htod(c2_dev, c2_host.data(), c2_host.size() * 8, nullptr));
        PhantomCiphertext ks = ct1;
        hipStream_t stream = nullptr;
        phantom::keyswitch_inplace(context, ks, c2_dev, rk_moved, true, stream);
        phantom::detail::check(phx_stream_synchronize(nullptr));
        write(dir, "ks.bin", ks.download());
        // the 2-D NTT of ct1's c0
        phantom::detail::check(phx_memcpy_htod(c2_dev, a.data(), Q * N * 8, nullptr));
        phantom::nwt_2d_radix8_forward_inplace(c2_dev, context, Q, 0, stream);
        std::vector<u64> ntt(Q * N);
        phantom::detail::check(phx_memcpy_dtoh(ntt.data(), c2_dev, ntt.size() * 8, nullptr));
        write(dir, "ntt.bin", ntt);
        phantom::nwt_2d_radix8_backward_inplace(c2_dev, context, Q, 0, stream);
        std::vector<u64> back(Q * N);
        phantom::detail::check(phx_memcpy_dtoh(back.data(), c2_dev, back.size() * 8, nullptr));
        expect(std::vector<u64>(a.begin(), a.begin() + (long)(Q * N)) == back, "INTT(NTT(x)) == x");
        phantom::detail::check(phx_free(c2_dev));
        // stream round trip of a ciphertext
        std::stringstream ss;
        mr.save(ss);
        PhantomCiphertext loaded;
        loaded.load(context, ss);
        expect(loaded.download() == mr.download() && loaded.scale() == mr.scale(), "save / load round trip");

        // parameter errors keep the reference's exception types
        expect_throw<std::invalid_argument>([&] { phantom::multiply_and_relin_inplace(context, mr, rs, rk_moved); },
                                            "operands at different levels");
        expect_throw<std::invalid_argument>([&] { phantom::rotate_inplace(context, mr, 8, gk); },
                                            "rotation without a key");
        expect_throw<std::invalid_argument>([&] { phantom::multiply_and_relin_inplace(context, mr, ct2, rk); },
                                            "a moved-from relinearisation key");
        expect_throw<std::invalid_argument>([&] { PhantomCiphertext e; phantom::relinearize_inplace(context, e, rk_moved); },
                                            "an empty ciphertext");
    } catch (const std::exception &e) {
        std::fprintf(stderr, "facade: exception %s\n", e.what());
        return 1;
    }
    if (fails) return 1;
    std::printf("facade: ok\n");
    return 0;
}
