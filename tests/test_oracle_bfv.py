"""Pins of the oracle's BFV path (BEHZ multiply, BFV relinearisation) by
closed forms in Python integers, at small ring degree.

BEHZ multiply (src/evaluate.cu:440-545): the lift (fastbconv_m_tilde +
Montgomery sm_mrq) turns the CRT integer X in [0, Q) of each input coefficient
into X' = (y + r Q) / m_tilde with y = sum_i [m_tilde x_i Qhat_i^-1]_{q_i} Qhat_i
and r = -y Q^-1 mod m_tilde taken in [-m_tilde/2, m_tilde/2): X' = X or X - Q,
the centred representative (but for a window of width ~ size_Q Q / m_tilde
around Q/2); the tensor is the integer negacyclic product z of the X'; after the
scale by t, fast_floor returns w = (t z - S) / Q where S = sum_i
[t z * Qhat_i^-1]_{q_i} Qhat_i is the FastBConv value of [t z]_Q (so w =
floor(t z / Q) - e, 0 <= e < size_Q); Shenoy-Kumaresan then returns w mod q_i
exactly. Every output residue is checked against that w.

BFV relinearisation: with a key built from a known ternary s, the output
satisfies c0' + c1' s = c0 + c1 s + c2 s^2 + small (mod Q).
"""
import numpy as np
import pytest

import oracle_bind as ob
from bigint_ref import crt, primes_below

T = 65537


def negacyclic_int(a, b):
    N = len(a)
    c = [0] * N
    for i in range(N):
        if a[i] == 0:
            continue
        for j in range(N):
            k = i + j
            if k < N:
                c[k] += a[i] * b[j]
            else:
                c[k - N] -= a[i] * b[j]
    return c


def setup(logN, bits, P):
    N = 1 << logN
    mods = ob.coeff_modulus_create(N, bits)
    return ob.Ctx(logN, mods, P), mods, N


def test_behz_bases_closed_form():
    octx, mods, N = setup(5, [30, 30, 30, 31], 1)
    b = ob.Behz(octx, T)
    k = 3
    Q = 1
    for q in mods[:k]:
        Q *= q
    size_B = k + (1 if 32 + T.bit_length() + Q.bit_length() >= 61 * k + 61 else 0)
    p = primes_below(61, N, size_B + 1)
    assert b.bsk() == p[1:] + p[:1]          # B (descending), then m_sk = the largest


@pytest.mark.parametrize("logN,bits", [(5, [30, 30, 30, 31]), (4, [50, 50, 61]), (5, [54, 54, 54, 54, 60])])
def test_behz_multiply_matches_integer_floor(logN, bits):
    octx, mods, N = setup(logN, bits, 1)
    k = len(mods) - 1
    qs = mods[:k]
    Q = 1
    for q in qs:
        Q *= q
    beh = ob.Behz(octx, T)
    rng = np.random.default_rng(logN * 7 + k)
    ct1 = np.stack([np.stack([rng.integers(0, q, N, dtype=np.uint64) for q in qs]) for _ in range(2)])
    ct2 = np.stack([np.stack([rng.integers(0, q, N, dtype=np.uint64) for q in qs]) for _ in range(2)])
    out = beh.multiply(ct1, ct2)
    qhat = [Q // q for q in qs]
    qhinv = [pow(h % q, -1, q) for h, q in zip(qhat, qs)]
    mt = 1 << 32
    qinv_mt = pow(Q % mt, -1, mt)

    def lift(res):
        y = sum(((mt * int(x)) * hi % q) * h for x, q, hi, h in zip(res, qs, qhinv, qhat))
        r = (-y * qinv_mt) % mt
        if r >= mt // 2:
            r -= mt
        assert (y + r * Q) % mt == 0
        v = (y + r * Q) // mt
        X = crt([int(x) for x in res], qs)[0]
        assert v in (X, X - Q)
        return v

    X = [[lift(ct1[p, :, n]) for n in range(N)] for p in range(2)]
    Y = [[lift(ct2[p, :, n]) for n in range(N)] for p in range(2)]
    z0 = negacyclic_int(X[0], Y[0])
    z2 = negacyclic_int(X[1], Y[1])
    z1 = [a + b for a, b in zip(negacyclic_int(X[0], Y[1]), negacyclic_int(X[1], Y[0]))]
    for d, z in enumerate((z0, z1, z2)):
        for n in range(N):
            tz = T * z[n]
            S = sum(((tz % q) * hi % q) * h for q, hi, h in zip(qs, qhinv, qhat))
            assert (tz - S) % Q == 0
            w = (tz - S) // Q
            e = (S - tz % Q) // Q
            assert 0 <= e < k
            for i, q in enumerate(qs):
                assert int(out[d, i, n]) == w % q, (d, n, i)


def test_behz_multiply_planted_error_is_caught():
    """A wrong t (a dropped scale) must fail the closed form above."""
    octx, mods, N = setup(5, [30, 30, 30, 31], 1)
    qs = mods[:3]
    rng = np.random.default_rng(3)
    ct = np.stack([np.stack([rng.integers(0, q, N, dtype=np.uint64) for q in qs]) for _ in range(2)])
    good = ob.Behz(octx, T).multiply(ct, ct)
    bad = ob.Behz(octx, T + 2).multiply(ct, ct)
    assert not np.array_equal(good, bad)


def ternary(rng, N):
    return [int(v) for v in rng.integers(-1, 2, N)]


def residues(vals, qs):
    return np.stack([np.array([int(v) % q for v in vals], np.uint64) for q in qs])


@pytest.mark.parametrize("logN,bits,P", [(5, [30, 30, 30, 31], 1), (5, [40, 40, 40, 40, 50, 50], 2)])
def test_bfv_relinearize_decrypts_like_size3(logN, bits, P):
    octx, mods, N = setup(logN, bits, P)
    QP = len(mods)
    k = QP - P
    qs = mods[:k]
    rng = np.random.default_rng(11 * logN + P)
    s = ternary(rng, N)
    s2 = negacyclic_int(s, s)
    s_ntt = octx.ntt_forward(residues(s, mods), 0)
    s2_ntt = octx.ntt_forward(residues(s2, mods), 0)
    dnum = octx.beta(k)
    a = np.stack([np.stack([rng.integers(0, q, N, dtype=np.uint64) for q in mods]) for _ in range(dnum)])
    e = np.stack([residues(rng.integers(-8, 9, N), mods) for _ in range(dnum)])
    key = octx.make_kswitch_key(s_ntt, s2_ntt, a, e)
    ct3 = np.stack([np.stack([rng.integers(0, q, N, dtype=np.uint64) for q in qs]) for _ in range(3)])
    out = ob.bfv_relinearize(octx, ct3, key)
    Q = 1
    for q in qs:
        Q *= q

    def decrypt_int(polys):
        acc = [0] * N
        sp = [1] + [0] * (N - 1)
        for c in polys:
            cint = [crt([int(c[i, n]) for i in range(k)], qs)[0] for n in range(N)]
            term = negacyclic_int(cint, sp) if sp != [1] + [0] * (N - 1) else cint
            acc = [x + y for x, y in zip(acc, term)]
            sp = negacyclic_int(sp, s)
        return acc

    d3 = decrypt_int([ct3[0], ct3[1], ct3[2]])
    d2 = decrypt_int([out[0], out[1]])
    noise = [((x - y) % Q + Q // 2) % Q - Q // 2 for x, y in zip(d2, d3)]
    assert max(abs(v) for v in noise) < 1 << 16, max(abs(v) for v in noise)


# ---------------------------------------------------------------- HPS
# bfv_multiply_hps (src/evaluate.cu:648-801): bConv_HPS lifts each input to its
# centred representative X_c in (-Q/2, Q/2] (exact CRT with the FP64 overflow
# count), the tensor z of the X_c is exact in base Q R, scaleAndRound_HPS_QR_R
# gives round(t z / Q) mod R and bConv_HPS R -> Q returns it (centred, exact)
# mod q_i. So out_i = round(t z / Q) + d mod q_i, where d comes only from the
# FP64 sum nu = 0.5 + sum_i x_i frac_i of scaleAndRound_HPS_QR_R: its absolute
# error is at most k^2 2^(qbits - 53) (x_i < 2^qbits), so d = 0 for small
# primes (exact regime, checked exactly) and |d| <= that bound + 1 for 50- and
# 54-bit primes. Which way a sum near a rounding boundary falls (the
# reference's compiled FP64 accumulation order and contraction) is the one part
# no closed form pins; the oracle and the GPU use the same fused
# multiply-adds in input order, so they agree bit for bit.

def test_hps_r_primes_closed_form():
    octx, mods, N = setup(5, [30, 30, 30, 31], 1)
    h = ob.Hps(octx, T)
    qmin = min(mods[:3])
    want, v = [], qmin - 2 * N
    from bigint_ref import is_prime
    while len(want) < 4:
        if is_prime(v):
            want.append(v)
        v -= 2 * N
    assert h.r() == want


@pytest.mark.parametrize("logN,bits", [(5, [30, 30, 30, 31]), (4, [50, 50, 61]), (5, [54, 54, 54, 54, 60])])
def test_hps_multiply_matches_integer_round(logN, bits):
    octx, mods, N = setup(logN, bits, 1)
    k = len(mods) - 1
    qs = mods[:k]
    Q = 1
    for q in qs:
        Q *= q
    h = ob.Hps(octx, T)
    rng = np.random.default_rng(logN * 13 + k)
    ct1 = np.stack([np.stack([rng.integers(0, q, N, dtype=np.uint64) for q in qs]) for _ in range(2)])
    ct2 = np.stack([np.stack([rng.integers(0, q, N, dtype=np.uint64) for q in qs]) for _ in range(2)])
    out = h.multiply(ct1, ct2)

    def centred(res):
        x = crt([int(v) for v in res], qs)[0]
        return x - Q if 2 * x > Q else x

    X = [[centred(ct1[p, :, n]) for n in range(N)] for p in range(2)]
    Y = [[centred(ct2[p, :, n]) for n in range(N)] for p in range(2)]
    zs = (negacyclic_int(X[0], Y[0]),
          [a + b for a, b in zip(negacyclic_int(X[0], Y[1]), negacyclic_int(X[1], Y[0]))],
          negacyclic_int(X[1], Y[1]))
    qbits = max(q.bit_length() for q in qs)
    err = k * k * 2.0 ** (qbits - 53)
    bound = 0 if err < 2.0 ** -10 else int(err) + 1
    worst = 0
    for d, z in enumerate(zs):
        for n in range(N):
            num = T * z[n]
            w = (2 * num + Q) // (2 * Q)        # round half up
            got = crt([int(out[d, i, n]) for i in range(k)], qs)[0]
            diff = (got - w) % Q
            diff = diff - Q if 2 * diff > Q else diff
            worst = max(worst, abs(diff))
            assert abs(diff) <= bound, (d, n, diff, bound)
    if bound:
        assert worst <= bound


# ---- HPS over Q (mul_tech hps_overq) ----

def _centred(x, M):
    return x - M if 2 * x > M else x


def test_hps_overq_r_primes_closed_form():
    """R has size_Q primes (not size_Q + 1), the largest below min q stepping by 2N"""
    octx, mods, N = setup(5, [30, 30, 30, 31], 1)
    h = ob.Hps(octx, T, overq=True)
    qmin = min(mods[:3])
    want, v = [], qmin - 2 * N
    from bigint_ref import is_prime
    while len(want) < 3:
        if is_prime(v):
            want.append(v)
        v -= 2 * N
    assert h.r() == want


@pytest.mark.parametrize("logN,bits", [(5, [30, 30, 30, 31]), (4, [54, 54, 54, 54, 60])])
def test_hps_overq_modswitch_closed_form(logN, bits):
    """bConv_BEHZ_var1 Q -> R: Z = (R X + u) / Q with 0 <= u < k Q, for X = CRT(x) in [0, Q)"""
    octx, mods, N = setup(logN, bits, 1)
    k = len(mods) - 1
    qs = mods[:k]
    h = ob.Hps(octx, T, overq=True)
    rs = h.r()
    Q, R = 1, 1
    for q in qs:
        Q *= q
    for r in rs:
        R *= r
    rng = np.random.default_rng(logN + k)
    x = np.stack([rng.integers(0, q, N, dtype=np.uint64) for q in qs])
    x[:, 0] = 0                       # X = 0 -> Z = u / Q, u = 0
    x[:, 1] = [q - 1 for q in qs]     # X = Q - 1
    z = h.modswitch(x)
    for n in range(N):
        X = crt([int(v) for v in x[:, n]], qs)[0]
        Z = crt([int(v) for v in z[:, n]], rs)[0]
        d = Z * Q - R * X
        if d < 0:
            d += R * Q                # Z wrapped past R
        assert 0 <= d < k * Q, (n, d)
    assert crt([int(v) for v in z[:, 0]], rs)[0] == 0


@pytest.mark.parametrize("logN,bits,square", [(5, [30, 30, 30, 31], False), (4, [54, 54, 54, 54, 60], False),
                                              (5, [30, 30, 30, 31], True)])
def test_hps_overq_multiply_matches_integer_round(logN, bits, square):
    """hps_overq: the output is round(t z / R) mod Q (within the FP64 bound), where z
    is the integer negacyclic tensor of the centred X (ct1 mod Q) and Y: the centred
    modulus switch of ct2 (mod R) or, for a squaring, X again"""
    octx, mods, N = setup(logN, bits, 1)
    k = len(mods) - 1
    qs = mods[:k]
    h = ob.Hps(octx, T, overq=True)
    rs = h.r()
    Q, R = 1, 1
    for q in qs:
        Q *= q
    for r in rs:
        R *= r
    rng = np.random.default_rng(logN * 7 + k + square)
    ct1 = np.stack([np.stack([rng.integers(0, q, N, dtype=np.uint64) for q in qs]) for _ in range(2)])
    ct2 = ct1 if square else np.stack([np.stack([rng.integers(0, q, N, dtype=np.uint64) for q in qs])
                                       for _ in range(2)])
    out = h.multiply(ct1, ct2, square=square)
    X = [[_centred(crt([int(v) for v in ct1[p, :, n]], qs)[0], Q) for n in range(N)] for p in range(2)]
    if square:
        Y = X
    else:
        Y = []
        for p in range(2):
            z = h.modswitch(ct2[p])
            Y.append([_centred(crt([int(v) for v in z[:, n]], rs)[0], R) for n in range(N)])
    zs = (negacyclic_int(X[0], Y[0]),
          [a + b for a, b in zip(negacyclic_int(X[0], Y[1]), negacyclic_int(X[1], Y[0]))],
          negacyclic_int(X[1], Y[1]))
    rbits = max(r.bit_length() for r in rs)
    err = k * k * 2.0 ** (rbits - 53)
    bound = 0 if err < 2.0 ** -10 else int(err) + 1
    for d, z in enumerate(zs):
        for n in range(N):
            num = T * z[n]
            w = (2 * num + R) // (2 * R)        # round half up
            got = crt([int(out[d, i, n]) for i in range(k)], qs)[0]
            diff = _centred((got - w) % Q, Q)
            assert abs(diff) <= bound, (d, n, diff, bound)


def test_hps_overq_differs_from_hps():
    """the two techniques give different (both correctly rounded) results"""
    octx, mods, N = setup(4, [30, 30, 30, 31], 1)
    rng = np.random.default_rng(3)
    qs = mods[:3]
    ct1 = np.stack([np.stack([rng.integers(0, q, N, dtype=np.uint64) for q in qs]) for _ in range(2)])
    ct2 = np.stack([np.stack([rng.integers(0, q, N, dtype=np.uint64) for q in qs]) for _ in range(2)])
    a = ob.Hps(octx, T).multiply(ct1, ct2)
    b = ob.Hps(octx, T, overq=True).multiply(ct1, ct2)
    assert not np.array_equal(a, b)


# ---- hps_overq_leveled ----

def _prod(xs):
    v = 1
    for x in xs:
        v *= x
    return v


def test_hps_level_r_is_prefix():
    """the R of the level with d dropped primes is the first Ql primes of the top level's R"""
    octx, mods, N = setup(5, [30, 30, 30, 30, 31], 1)
    top = ob.Hps(octx, T, overq=True).r()
    for d in range(1, 4):
        assert ob.HpsLevel(octx, T, d).r() == top[:4 - d]


@pytest.mark.parametrize("logN,bits", [(5, [30, 30, 30, 30, 31]), (4, [54, 54, 54, 54, 60])])
def test_hps_scale_q_ql_closed_form(logN, bits):
    """scaleAndRound_HPS_Q_Ql: round(Ql X / Q) mod q_i (i < Ql) for X = CRT_Q(x), within
    the FP64 bound of the nu sum"""
    octx, mods, N = setup(logN, bits, 1)
    k0 = len(mods) - 1
    qs = mods[:k0]
    Q = _prod(qs)
    rng = np.random.default_rng(logN + 11)
    x = np.stack([rng.integers(0, q, N, dtype=np.uint64) for q in qs])
    x[:, 0] = 0
    x[:, 1] = [q - 1 for q in qs]
    qbits = max(q.bit_length() for q in qs)
    for d in range(1, k0):
        Ql = _prod(qs[:k0 - d])
        out = ob.hps_scale_q_ql(octx, d, x)
        err = d * d * 2.0 ** (qbits - 53)
        bound = 0 if err < 2.0 ** -10 else int(err) + 1
        for n in range(N):
            X = crt([int(v) for v in x[:, n]], qs)[0]
            w = (2 * Ql * X + Q) // (2 * Q)
            got = crt([int(out[i, n]) for i in range(k0 - d)], qs[:k0 - d])[0]
            diff = _centred((got - w) % Ql, Ql)
            assert abs(diff) <= bound, (d, n, diff)
        assert crt([int(v) for v in out[:, 0]], qs[:k0 - d])[0] == 0


def test_hps_expand_closed_form():
    """ExpandCRTBasis_Ql_Q: CRT_Q(out) = CRT_Ql(x) (Q / Ql), the dropped limbs zero"""
    octx, mods, N = setup(4, [30, 30, 30, 30, 31], 1)
    qs = mods[:4]
    Q = _prod(qs)
    rng = np.random.default_rng(5)
    for d in (1, 2, 3):
        x = np.stack([rng.integers(0, q, N, dtype=np.uint64) for q in qs[:4 - d]])
        out = ob.hps_expand(octx, d, x)
        assert not out[4 - d:].any()
        for n in range(N):
            X = crt([int(v) for v in x[:, n]], qs[:4 - d])[0]
            assert crt([int(v) for v in out[:, n]], qs)[0] == X * (Q // _prod(qs[:4 - d])) % Q


def test_hps_level_modswitch_from_all_of_q():
    """below the top level the second operand is switched from all of Q to Rl:
    Z = (Rl X + u) / Q with 0 <= u < size_Q Q"""
    octx, mods, N = setup(5, [30, 30, 30, 30, 31], 1)
    qs = mods[:4]
    Q = _prod(qs)
    rng = np.random.default_rng(9)
    x = np.stack([rng.integers(0, q, N, dtype=np.uint64) for q in qs])
    for d in (1, 2):
        h = ob.HpsLevel(octx, T, d)
        rs = h.r()
        R = _prod(rs)
        z = h.modswitch(x)
        for n in range(N):
            X = crt([int(v) for v in x[:, n]], qs)[0]
            Z = crt([int(v) for v in z[:, n]], rs)[0]
            dd = Z * Q - R * X
            if dd < 0:
                dd += R * Q
            assert 0 <= dd < 4 * Q, (d, n, dd)


@pytest.mark.parametrize("square", [False, True])
def test_hps_level_multiply_matches_integer_round(square):
    """with d dropped levels: X = the centred round(Ql x1 / Q) (or, squaring, of ct1 twice),
    Y = the centred switch of ct2 to Rl; the result's first Ql limbs are
    round(t z / Rl) (Q / Ql) mod q_i, the dropped limbs zero"""
    octx, mods, N = setup(5, [30, 30, 30, 30, 31], 1)
    k0 = 4
    qs = mods[:k0]
    Q = _prod(qs)
    rng = np.random.default_rng(17 + square)
    ct1 = np.stack([np.stack([rng.integers(0, q, N, dtype=np.uint64) for q in qs]) for _ in range(2)])
    ct2 = ct1 if square else np.stack([np.stack([rng.integers(0, q, N, dtype=np.uint64) for q in qs])
                                       for _ in range(2)])
    for d in (1, 2):
        h = ob.HpsLevel(octx, T, d)
        rs = h.r()
        R = _prod(rs)
        ql = qs[:k0 - d]
        Ql = _prod(ql)
        out = h.multiply(ct1, ct2, square=square)
        X = []
        for p in range(2):
            s = ob.hps_scale_q_ql(octx, d, ct1[p])
            X.append([_centred(crt([int(v) for v in s[:, n]], ql)[0], Ql) for n in range(N)])
        if square:
            Y = X
        else:
            Y = []
            for p in range(2):
                z = h.modswitch(ct2[p])
                Y.append([_centred(crt([int(v) for v in z[:, n]], rs)[0], R) for n in range(N)])
        zs = (negacyclic_int(X[0], Y[0]),
              [a + b for a, b in zip(negacyclic_int(X[0], Y[1]), negacyclic_int(X[1], Y[0]))],
              negacyclic_int(X[1], Y[1]))
        for p, z in enumerate(zs):
            assert not out[p, k0 - d:].any()
            for n in range(N):
                w = (2 * T * z[n] + R) // (2 * R)
                got = crt([int(out[p, i, n]) for i in range(k0)], qs)[0]
                assert got == (w % Ql) * (Q // Ql) % Q, (d, p, n)


def test_levels_to_drop_invariants():
    """FindLevelsToDrop: 0 for a fresh ciphertext, non-decreasing in the depth, at most
    size_Q - 1, and no more for a key switch or an asymmetric ciphertext (larger
    extra noise / smaller fresh noise). Its constants are the reference's own
    arrangement: beyond these invariants the value is not pinned."""
    octx, mods, N = setup(13, [54, 54, 54, 54, 60], 1)
    prev = 0
    vals = []
    for depth in range(0, 12):
        v = ob.hps_levels_to_drop(octx, T, depth)
        assert prev <= v <= 3
        assert ob.hps_levels_to_drop(octx, T, depth, key_switch=True) <= v
        assert ob.hps_levels_to_drop(octx, T, depth, asymmetric=True) <= v
        prev = v
        vals.append(v)
    assert vals[0] == 0 and vals[1] == 0
    assert vals[-1] >= 1
