"""Pins the CPU oracle against closed forms, brute force, invariants and the
golden fixtures (tests/golden/golden_small.json). CPU only.

The reference holds no golden vectors for this path (its only pins are the
NTT->INTT round trip, test/ntt_test.cu:46-68, 102-121, and decryption error
bounds, examples/3_ckks.cu:18-26), so every oracle function is checked here
against an independent statement of what it computes, chosen so that a
dropped term, a wrong sign/index or a transposed operand fails.
"""
import json
import os
import random

import numpy as np
import pytest

import bigint_ref as br
import oracle_bind as ob

HERE = os.path.dirname(os.path.abspath(__file__))


@pytest.fixture(scope="module", autouse=True)
def _build():
    ob.build()
    ob.set_threads(4)


def rand_limbs(rng, moduli, N):
    return np.stack([rng.integers(0, q, N, dtype=np.uint64) for q in moduli]).astype(np.uint64)


# ---------------------------------------------------------------- primes/roots
def test_get_primes_matches_scan():
    for N, bits, cnt in [(4096, 60, 3), (65536, 50, 8), (65536, 60, 13), (8192, 54, 4)]:
        assert ob.get_primes(N, bits, cnt) == br.primes_below(bits, N, cnt)


def test_create_ordering_n16_q35():
    N = 1 << 16
    bits = [60] + [50] * 34 + [60] * 12
    got = ob.coeff_modulus_create(N, bits)
    p60 = sorted(br.primes_below(60, N, 13))
    p50 = sorted(br.primes_below(50, N, 34))
    assert got[0] == p60[0]
    assert got[1:35] == p50
    assert got[35:] == p60[1:]
    assert len(set(got)) == len(got)


def test_create_n12_three_60bit_is_largest_ascending():
    got = ob.coeff_modulus_create(4096, [60, 60, 60])
    assert got == sorted(br.primes_below(60, 4096, 3))
    assert all(q % 8192 == 1 and q < (1 << 60) for q in got)


@pytest.mark.parametrize("N", [8, 16, 32])
def test_minimal_root_bruteforce(N):
    for q in br.primes_below(20, N, 3):
        assert ob.minimal_root(q, 2 * N) == br.minimal_root_bruteforce(q, 2 * N)


def test_ntt_tables_structure():
    logN, N = 6, 64
    q = br.primes_below(60, N, 1)[0]
    rp, rps, irp, irps, ninv = ob.ntt_tables(q, logN)
    psi = ob.minimal_root(q, 2 * N)
    for i in range(N):
        assert int(rp[br.brv(i, logN)]) == pow(psi, i, q)
        assert int(rps[i]) == (int(rp[i]) << 64) // q
    assert ninv == pow(N, -1, q)
    # inverse table: psi^-i with entry 1 pre-multiplied by n^-1 (src/host/ntt.cu:53-55)
    assert int(irp[1]) == pow(psi, -(N // 2), q) * ninv % q
    assert int(irp[2]) == pow(psi, -(N // 4), q)


# ---------------------------------------------------------------- NTT
def test_ntt_golden_vectors():
    g = json.load(open(os.path.join(HERE, "golden", "golden_small.json")))
    logN, N = g["logN"], g["N"]
    for v in g["ntt"]:
        other = [x for x in br.primes_below(60, N, 4) if x != v["q"]][0]
        ctx = ob.Ctx(logN, [v["q"], other], 1)
        assert ob.minimal_root(v["q"], 2 * N) == v["psi"]
        a = np.array(v["input"], np.uint64)
        out = ctx.ntt_forward(a, 0)
        assert [int(x) for x in out] == v["output"]
        assert np.array_equal(ctx.ntt_inverse(out, 0), a)


@pytest.mark.parametrize("bits", [60, 50, 36])
def test_ntt_closed_form_small(bits):
    logN, N = 7, 128
    qs = br.primes_below(bits, N, 2)
    ctx = ob.Ctx(logN, qs, 1)
    rng = np.random.default_rng(bits)
    a = rand_limbs(rng, qs, N)
    out = ctx.ntt_forward(a, 0)
    for i, q in enumerate(qs):
        psi = ob.minimal_root(q, 2 * N)
        assert [int(x) for x in out[i]] == br.ntt_closed(a[i], q, psi, logN)


def test_ntt_n12_three_60bit_sampled_closed_form_and_roundtrip():
    """N=2^12, the 3 largest 60-bit primes = 1 mod 2N, uniform input."""
    logN, N = 12, 4096
    qs = ob.coeff_modulus_create(N, [60, 60, 60])
    ctx = ob.Ctx(logN, qs, 1)
    rng = np.random.default_rng(1)
    a = rand_limbs(rng, qs, N)
    A = ctx.ntt_forward(a, 0)
    prng = random.Random(5)
    for i, q in enumerate(qs):
        psi = ob.minimal_root(q, 2 * N)
        for k in prng.sample(range(N), 24) + [0, N - 1]:
            assert int(A[i, k]) == br.ntt_closed_at(a[i], q, psi, logN, k)
    assert np.array_equal(ctx.ntt_inverse(A, 0), a)
    assert (A < np.array(qs, np.uint64)[:, None]).all()
    # the reference's own test input: a constant polynomial (test/ntt_test.cu:102-121)
    c = np.full((3, N), 2, np.uint64)
    assert np.array_equal(ctx.ntt_inverse(ctx.ntt_forward(c, 0), 0), c)


def test_negacyclic_convolution_theorem():
    logN, N = 8, 256
    qs = br.primes_below(50, N, 2)
    ctx = ob.Ctx(logN, qs, 1)
    rng = np.random.default_rng(7)
    a, b = rand_limbs(rng, qs, N), rand_limbs(rng, qs, N)
    A, B = ctx.ntt_forward(a, 0), ctx.ntt_forward(b, 0)
    C = np.array([[int(x) * int(y) % q for x, y in zip(A[i], B[i])] for i, q in enumerate(qs)], np.uint64)
    c = ctx.ntt_inverse(C, 0)
    for i, q in enumerate(qs):
        assert [int(x) for x in c[i]] == br.negacyclic_mul(a[i], b[i], q)


# ---------------------------------------------------------------- base conversion
def test_fastbconv_golden():
    g = json.load(open(os.path.join(HERE, "golden", "golden_small.json")))["fastbconv"]
    out = ob.fastbconv(64, g["ibase"], g["obase"], np.array(g["input"], np.uint64))
    assert out.tolist() == g["output"]


def test_fastbconv_is_x_plus_small_multiple_of_Q():
    N = 32
    ib = br.primes_below(50, N, 4)
    obase = br.primes_below(60, N, 3)
    rng = np.random.default_rng(3)
    x = rand_limbs(rng, ib, N)
    out = ob.fastbconv(N, ib, obase, x)
    for k in range(N):
        X, Q = br.crt([x[i, k] for i in range(len(ib))], ib)
        for o, p in enumerate(obase):
            es = [e for e in range(len(ib)) if (X + e * Q) % p == int(out[o, k])]
            assert es, (k, o)


# ---------------------------------------------------------------- toy key-switching chain
TOY_LOGN, TOY_N = 6, 64


def toy_ctx():
    mods = ob.coeff_modulus_create(TOY_N, [60, 50, 50, 50, 50, 60, 60])
    return ob.Ctx(TOY_LOGN, mods, 2), mods


def test_moddown_inverts_multiplication_by_P():
    ctx, mods = toy_ctx()
    Ql = 5
    qlp = mods[:Ql] + mods[5:]
    Qbig = 1
    for q in mods[:Ql]:
        Qbig *= q
    P = mods[5] * mods[6]
    rng = random.Random(11)
    xs = [rng.randrange(Qbig) for _ in range(TOY_N)]
    cx = np.array([[(P * v) % q for v in xs] for q in qlp], np.uint64)
    # NTT every limb with its own table row (specials use rows 5, 6)
    cx_ntt = np.concatenate([ctx.ntt_forward(cx[:Ql], 0), ctx.ntt_forward(cx[Ql:], 5)])
    out = ctx.moddown(Ql, cx_ntt)
    want = ctx.ntt_forward(np.array([[v % q for v in xs] for q in mods[:Ql]], np.uint64), 0)
    assert np.array_equal(out, want)


def test_modup_structure():
    ctx, mods = toy_ctx()
    Ql, alpha = 5, 2
    qlp = mods[:Ql] + mods[5:]
    rows = list(range(Ql)) + [5, 6]
    rng = np.random.default_rng(5)
    c2 = rand_limbs(rng, mods[:Ql], TOY_N)
    t = ctx.modup(Ql, c2)
    assert t.shape == (3, Ql + 2, TOY_N)  # beta = ceil(5/2) = 3, last digit partial
    x = ctx.ntt_inverse(c2, 0)
    for j in range(3):
        st, sz = alpha * j, min(alpha, Ql - alpha * j)
        assert np.array_equal(t[j, st:st + sz], c2[st:st + sz])
        coef = np.stack([ctx.ntt_inverse(t[j, pos:pos + 1], rows[pos])[0] for pos in range(Ql + 2)])
        dig = mods[st:st + sz]
        for k in range(0, TOY_N, 7):
            X, Qd = br.crt([x[i, k] for i in range(st, st + sz)], dig)
            for pos in range(Ql + 2):
                if st <= pos < st + sz:
                    continue
                p = qlp[pos]
                assert any((X + e * Qd) % p == int(coef[pos, k]) for e in range(sz))


def test_inner_product_bruteforce():
    ctx, mods = toy_ctx()
    Ql = 3
    QP = 7
    qlp_rows = [0, 1, 2, 5, 6]
    rng = np.random.default_rng(9)
    beta = ctx.beta(Ql)
    t = np.stack([rand_limbs(rng, [mods[r] for r in qlp_rows], TOY_N) for _ in range(beta)])
    key = np.stack([np.stack([rand_limbs(rng, mods, TOY_N), rand_limbs(rng, mods, TOY_N)]) for _ in range(3)])
    cx = ctx.inner_product(Ql, t, key)
    for pos, r in enumerate(qlp_rows):
        q = mods[r]
        for h in range(2):
            want = [sum(int(t[j, pos, k]) * int(key[j, h, r, k]) for j in range(beta)) % q for k in range(TOY_N)]
            assert [int(v) for v in cx[h, pos]] == want
    assert QP == len(mods)


def test_tensor_bruteforce():
    ctx, mods = toy_ctx()
    Ql = 4
    rng = np.random.default_rng(2)
    a = np.stack([rand_limbs(rng, mods[:Ql], TOY_N) for _ in range(2)])
    b = np.stack([rand_limbs(rng, mods[:Ql], TOY_N) for _ in range(2)])
    d = ctx.tensor(Ql, a, b)
    for i, q in enumerate(mods[:Ql]):
        for k in range(TOY_N):
            a0, a1, b0, b1 = (int(v) for v in (a[0, i, k], a[1, i, k], b[0, i, k], b[1, i, k]))
            assert int(d[0, i, k]) == a0 * b0 % q
            assert int(d[1, i, k]) == (a0 * b1 + a1 * b0) % q
            assert int(d[2, i, k]) == a1 * b1 % q


def test_rescale_golden_floor():
    g = json.load(open(os.path.join(HERE, "golden", "golden_small.json")))["rescale"]
    mods = g["moduli"]
    ctx = ob.Ctx(6, mods + [br.primes_below(61, 64, 1)[0]], 1)
    limbs = np.array(g["input"], np.uint64)
    ct = ctx.ntt_forward(limbs, 0)[None]
    out = ctx.rescale(3, ct)[0]
    assert ctx.ntt_inverse(out, 0).tolist() == g["output"]


def test_galois_permutation_matches_automorphism():
    ctx, mods = toy_ctx()
    rng = np.random.default_rng(4)
    a = rand_limbs(rng, mods[:2], TOY_N)
    A = ctx.ntt_forward(a, 0)
    for step in [1, 3, -1, 0]:
        g = ob.galois_elt_from_step(step, TOY_N)
        perm = ob.galois_perm_table(TOY_LOGN, g)
        for i, q in enumerate(mods[:2]):
            want = ctx.ntt_forward(np.array([br.galois_coeff(a[i], g, q)], np.uint64), i)[0]
            assert np.array_equal(A[i][perm], want)


def test_galois_elt_from_step():
    N = 1 << 16
    assert ob.galois_elt_from_step(0, N) == 2 * N - 1
    assert ob.galois_elt_from_step(1, N) == 5
    assert ob.galois_elt_from_step(2, N) == 25
    assert ob.galois_elt_from_step(-1, N) == pow(5, N // 2 - 1, 2 * N)
    assert ob.galois_elt_from_step(-1, N) * 5 % (2 * N) == 1


def _centered(v, Q):
    return v - Q if v > Q // 2 else v


def _toy_secret(rng):
    s = rng.integers(-1, 2, TOY_N)
    return s


def _rns(vals, mods):
    return np.array([[int(v) % q for v in vals] for q in mods], np.uint64)


def _toy_keys(ctx, mods, s_coef, s2_coef_fn, seed):
    rng = np.random.default_rng(seed)
    s_ntt = ctx.ntt_forward(_rns(s_coef, mods), 0)
    s2 = s2_coef_fn(s_coef)
    s2_ntt = ctx.ntt_forward(_rns(s2, mods), 0)
    dnum = 3
    a = np.stack([rand_limbs(rng, mods, TOY_N) for _ in range(dnum)])
    e = np.stack([_rns(rng.integers(-4, 5, TOY_N), mods) for _ in range(dnum)])
    return ctx.make_kswitch_key(s_ntt, s2_ntt, a, e), s_ntt


def test_keyswitch_decrypts_to_c2_s2():
    ctx, mods = toy_ctx()
    rng = np.random.default_rng(21)
    s = _toy_secret(rng)
    sq = br.negacyclic_mul(s, s, 1 << 200)
    sq = [_centered(v, 1 << 200) for v in sq]
    key, s_ntt = _toy_keys(ctx, mods, s, lambda _: sq, 5)
    for Ql in (5, 4, 2):
        c2 = rand_limbs(rng, mods[:Ql], TOY_N)
        out = ctx.keyswitch(Ql, c2, key)
        # (cx0 + cx1 s) - c2 s^2 must be small (key-switching noise)
        lhs = np.array([[(int(out[0, i, k]) + int(out[1, i, k]) * int(s_ntt[i, k])) % q for k in range(TOY_N)]
                        for i, q in enumerate(mods[:Ql])], np.uint64)
        s2n = ctx.ntt_forward(_rns(sq, mods[:Ql]), 0)
        rhs = np.array([[int(c2[i, k]) * int(s2n[i, k]) % q for k in range(TOY_N)] for i, q in enumerate(mods[:Ql])],
                       np.uint64)
        diff = ctx.ntt_inverse(np.array([[(int(x) - int(y)) % q for x, y in zip(lhs[i], rhs[i])]
                                         for i, q in enumerate(mods[:Ql])], np.uint64), 0)
        for k in range(TOY_N):
            v, Q = br.crt([diff[i, k] for i in range(Ql)], mods[:Ql])
            assert abs(_centered(v, Q)) < (1 << 20), (Ql, k)


def test_keyswitch_detects_wrong_key():
    """A plausible mistake (using s instead of s^2 for the gadget) must fail the
    decryption check above: the noise becomes large."""
    ctx, mods = toy_ctx()
    rng = np.random.default_rng(22)
    s = _toy_secret(rng)
    key, s_ntt = _toy_keys(ctx, mods, s, lambda x: list(x), 6)
    Ql = 5
    c2 = rand_limbs(rng, mods[:Ql], TOY_N)
    out = ctx.keyswitch(Ql, c2, key)
    sq = [_centered(v, 1 << 200) for v in br.negacyclic_mul(s, s, 1 << 200)]
    s2n = ctx.ntt_forward(_rns(sq, mods[:Ql]), 0)
    lhs = [[(int(out[0, i, k]) + int(out[1, i, k]) * int(s_ntt[i, k]) - int(c2[i, k]) * int(s2n[i, k])) % q
            for k in range(TOY_N)] for i, q in enumerate(mods[:Ql])]
    diff = ctx.ntt_inverse(np.array(lhs, np.uint64), 0)
    big = 0
    for k in range(TOY_N):
        v, Q = br.crt([diff[i, k] for i in range(Ql)], mods[:Ql])
        big = max(big, abs(_centered(v, Q)))
    assert big > (1 << 40)


def test_multiply_and_relin_decrypts_product():
    ctx, mods = toy_ctx()
    rng = np.random.default_rng(31)
    s = _toy_secret(rng)
    sq = [_centered(v, 1 << 200) for v in br.negacyclic_mul(s, s, 1 << 200)]
    key, s_ntt = _toy_keys(ctx, mods, s, lambda _: sq, 8)
    Ql = 5
    m1 = rng.integers(-1000, 1000, TOY_N)
    m2 = rng.integers(-1000, 1000, TOY_N)

    def enc(m):
        a = rand_limbs(rng, mods[:Ql], TOY_N)
        e = ctx.ntt_forward(_rns(rng.integers(-3, 4, TOY_N), mods[:Ql]), 0)
        mm = ctx.ntt_forward(_rns(m, mods[:Ql]), 0)
        c0 = np.array([[(-int(a[i, k]) * int(s_ntt[i, k]) + int(e[i, k]) + int(mm[i, k])) % q
                        for k in range(TOY_N)] for i, q in enumerate(mods[:Ql])], np.uint64)
        return np.stack([c0, a])

    ct1, ct2 = enc(m1), enc(m2)
    out = ctx.multiply_and_relin(Ql, ct1, ct2, key)
    dec = ctx.ntt_inverse(np.array([[(int(out[0, i, k]) + int(out[1, i, k]) * int(s_ntt[i, k])) % q
                                     for k in range(TOY_N)] for i, q in enumerate(mods[:Ql])], np.uint64), 0)
    want = br.negacyclic_mul(m1, m2, 1 << 200)
    for k in range(TOY_N):
        v, Q = br.crt([dec[i, k] for i in range(Ql)], mods[:Ql])
        assert abs(_centered(v, Q) - _centered(want[k], 1 << 200)) < (1 << 22)


def test_apply_galois_decrypts_rotated_message():
    ctx, mods = toy_ctx()
    rng = np.random.default_rng(41)
    s = _toy_secret(rng)
    g = ob.galois_elt_from_step(1, TOY_N)
    sg = br.galois_coeff([int(v) % (1 << 200) for v in s], g, 1 << 200)
    sg = [_centered(v, 1 << 200) for v in sg]
    key, s_ntt = _toy_keys(ctx, mods, s, lambda _: sg, 9)
    Ql = 4
    m = rng.integers(-1000, 1000, TOY_N)
    a = rand_limbs(rng, mods[:Ql], TOY_N)
    mm = ctx.ntt_forward(_rns(m, mods[:Ql]), 0)
    c0 = np.array([[(-int(a[i, k]) * int(s_ntt[i, k]) + int(mm[i, k])) % q for k in range(TOY_N)]
                   for i, q in enumerate(mods[:Ql])], np.uint64)
    out = ctx.apply_galois(Ql, np.stack([c0, a]), g, key)
    dec = ctx.ntt_inverse(np.array([[(int(out[0, i, k]) + int(out[1, i, k]) * int(s_ntt[i, k])) % q
                                     for k in range(TOY_N)] for i, q in enumerate(mods[:Ql])], np.uint64), 0)
    want = br.galois_coeff([int(v) % (1 << 200) for v in m], g, 1 << 200)
    for k in range(TOY_N):
        v, Q = br.crt([dec[i, k] for i in range(Ql)], mods[:Ql])
        assert abs(_centered(v, Q) - _centered(want[k], 1 << 200)) < (1 << 22)


def test_hoisting_decrypts_sum_of_rotations():
    """hoisting_inplace (src/evaluate.cu:1670-1865) returns the sum of the
    rotations: decryption must give sum_i sigma_{g_i}(m) up to key-switching
    noise, with one key per element built for sigma_{g_i}(s)."""
    ctx, mods = toy_ctx()
    rng = np.random.default_rng(43)
    s = _toy_secret(rng)
    steps = (1, 3, -2)
    elts = [ob.galois_elt_from_step(st, TOY_N) for st in steps]
    keys = []
    for i, g in enumerate(elts):
        sg = [_centered(v, 1 << 200) for v in br.galois_coeff([int(v) % (1 << 200) for v in s], g, 1 << 200)]
        key, s_ntt = _toy_keys(ctx, mods, s, lambda _, sg=sg: sg, 50 + i)
        keys.append(key)
    Ql = 4
    m = rng.integers(-1000, 1000, TOY_N)
    a = rand_limbs(rng, mods[:Ql], TOY_N)
    mm = ctx.ntt_forward(_rns(m, mods[:Ql]), 0)
    c0 = np.array([[(-int(a[i, k]) * int(s_ntt[i, k]) + int(mm[i, k])) % q for k in range(TOY_N)]
                   for i, q in enumerate(mods[:Ql])], np.uint64)
    out = ctx.hoisting(Ql, np.stack([c0, a]), elts, np.stack(keys))
    dec = ctx.ntt_inverse(np.array([[(int(out[0, i, k]) + int(out[1, i, k]) * int(s_ntt[i, k])) % q
                                     for k in range(TOY_N)] for i, q in enumerate(mods[:Ql])], np.uint64), 0)
    want = [0] * TOY_N
    for g in elts:
        rot = br.galois_coeff([int(v) % (1 << 200) for v in m], g, 1 << 200)
        want = [w + _centered(r, 1 << 200) for w, r in zip(want, rot)]
    for k in range(TOY_N):
        v, Q = br.crt([dec[i, k] for i in range(Ql)], mods[:Ql])
        assert abs(_centered(v, Q) - want[k]) < (1 << 22), k


def test_hoisting_bfv_decrypts_sum_of_rotations():
    """BFV hoisting (coefficient-form ciphertext): decryption gives
    sum_i sigma_{g_i}(m) up to key-switching noise, and the c0 part equals the
    coefficient-domain automorphism sum (a key switch of c1 = 0 is zero)."""
    ctx, mods = toy_ctx()
    rng = np.random.default_rng(47)
    s = _toy_secret(rng)
    steps = (2, -1, 5)
    elts = [ob.galois_elt_from_step(st, TOY_N) for st in steps]
    keys = []
    for i, g in enumerate(elts):
        sg = [_centered(v, 1 << 200) for v in br.galois_coeff([int(v) % (1 << 200) for v in s], g, 1 << 200)]
        key, s_ntt = _toy_keys(ctx, mods, s, lambda _, sg=sg: sg, 70 + i)
        keys.append(key)
    keys = np.stack(keys)
    Ql = 4
    m = rng.integers(-1000, 1000, TOY_N)
    a = rand_limbs(rng, mods[:Ql], TOY_N)
    a_ntt = ctx.ntt_forward(a, 0)
    as_ = ctx.ntt_inverse(np.array([[int(a_ntt[i, k]) * int(s_ntt[i, k]) % q for k in range(TOY_N)]
                                    for i, q in enumerate(mods[:Ql])], np.uint64), 0)
    mr = _rns(m, mods[:Ql])
    c0 = np.array([[(int(mr[i, k]) - int(as_[i, k])) % q for k in range(TOY_N)]
                   for i, q in enumerate(mods[:Ql])], np.uint64)
    out = ctx.hoisting(Ql, np.stack([c0, a]), elts, keys, bfv=True)
    o1 = ctx.ntt_forward(out[1], 0)
    o1s = ctx.ntt_inverse(np.array([[int(o1[i, k]) * int(s_ntt[i, k]) % q for k in range(TOY_N)]
                                    for i, q in enumerate(mods[:Ql])], np.uint64), 0)
    want = [0] * TOY_N
    for g in elts:
        rot = br.galois_coeff([int(v) % (1 << 200) for v in m], g, 1 << 200)
        want = [w + _centered(r, 1 << 200) for w, r in zip(want, rot)]
    for k in range(TOY_N):
        v, Q = br.crt([(int(out[0, i, k]) + int(o1s[i, k])) % mods[i] for i in range(Ql)], mods[:Ql])
        assert abs(_centered(v, Q) - want[k]) < (1 << 22), k
    # with c1 = 0 the output is (sum_i sigma_i(c0), 0)
    z = ctx.hoisting(Ql, np.stack([c0, np.zeros_like(a)]), elts, keys, bfv=True)
    assert not z[1].any()
    for i, q in enumerate(mods[:Ql]):
        acc = [0] * TOY_N
        for g in elts:
            rot = br.galois_coeff([int(v) for v in c0[i]], g, q)
            acc = [(x + y) % q for x, y in zip(acc, rot)]
        assert [int(v) for v in z[0, i]] == acc, i
