"""The oracle's FP64 arithmetic on the HPS paths, pinned exactly.

Every FP64 sum of the HPS multiply (bConv_HPS's overflow count v, the
scale-and-round nu of QR -> R, QlRl -> Ql and Q -> Ql) is checked bit for bit
against tests/hps_model.py: sequential fused multiply-adds in input order,
evaluated with exact rationals and rounded once per operation (IEEE round half
to even), with the constants as exact integers rounded by the reference's own
`(double) m / (double) q`. The integer results of each step are then checked
against the CRT identity that the fixed sum implies, and the whole multiply
(hps, hps_overq, hps_overq_leveled, squaring) against the composition of those
steps. The one assumption left is how the reference is compiled: nvcc
contracts `nu += (double) x * frac` into one fused multiply-add, in the loop's
order (src/rns.cu:1708-1714, 1757-1763; src/rns_bconv.cu:170-195).

Where the exact sum lies farther from an integer than the FP64 error bound,
the result equals round(t X / D) exactly; the tests count the coefficients that
fall inside the bound (where the reference's rounding, by design, may differ
from the exact one)."""
import math
from fractions import Fraction as F

import numpy as np
import pytest

import hps_model as hm
import oracle_bind as ob

T = 65537
SHAPES = [(5, [30, 30, 30, 31]), (4, [50, 50, 61]), (5, [54, 54, 54, 54, 60])]


def setup(logN, bits):
    N = 1 << logN
    mods = ob.coeff_modulus_create(N, bits)
    return ob.Ctx(logN, mods, 1), mods, N


def negacyclic(a, b):
    N = len(a)
    c = [0] * N
    for i in range(N):
        for j in range(N):
            if i + j < N:
                c[i + j] += a[i] * b[j]
            else:
                c[i + j - N] -= a[i] * b[j]
    return c


def bits_of(x):
    return np.float64(x).view(np.uint64)


# ---- the sum itself ----

def test_fma_chain_fused_in_order_round_half_even():
    cases = [
        # fused: the exact (2^52 + 1)(1 + 2^-52) - (2^52 + 2) = 2^-52; an unfused
        # multiply rounds the product to 2^52 + 2 first and gives 0
        ([2 ** 52 + 1], [1.0 + 2.0 ** -52], -(2.0 ** 52 + 2), 2.0 ** -52),
        # u64 -> double rounds to nearest even: 2^53 + 1 -> 2^53, 2^53 + 3 -> 2^53 + 4
        ([2 ** 53 + 1], [1.0], 0.0, 2.0 ** 53),
        ([2 ** 53 + 3], [1.0], 0.0, 2.0 ** 53 + 4),
        # input order: 2^60 first absorbs each 100 (half an ulp is 128); the reverse
        # order would sum 200 first and round up to 2^60 + 256
        ([2 ** 60, 100, 100], [1.0, 1.0, 1.0], 0.0, 2.0 ** 60),
    ]
    for xs, ws, init, want in cases:
        x = np.array(xs, np.uint64).reshape(len(xs), 1)
        got = ob.fma_chain(x, ws, init)[0]
        assert got == want, (xs, ws, init, got, want)
        assert hm.chain(xs, ws, init) == want
    rng = np.random.default_rng(3)
    x = rng.integers(0, 2 ** 61, (6, 64), dtype=np.uint64)
    w = rng.random(6)
    got = ob.fma_chain(x, w, 0.5)
    for n in range(64):
        assert bits_of(got[n]) == bits_of(hm.chain([int(v) for v in x[:, n]], list(w), 0.5)), n


# ---- the constants ----

@pytest.mark.parametrize("logN,bits", SHAPES)
def test_frac_constants_exact(logN, bits):
    octx, mods, N = setup(logN, bits)
    k = len(mods) - 1
    qs = mods[:k]
    h = ob.Hps(octx, T)
    rs = h.r()
    fq, _ = h.frac()
    want, _ = hm.frac_consts(T * hm.prod(rs), qs + rs, range(k))
    assert [bits_of(v) for v in fq] == [bits_of(v) for v in want]
    ho = ob.Hps(octx, T, overq=True)
    ro = ho.r()
    _, fr = ho.frac()
    want, _ = hm.frac_consts(T * hm.prod(qs), qs + ro, range(k, 2 * k))
    assert [bits_of(v) for v in fr] == [bits_of(v) for v in want]
    if k > 1:
        hl = ob.HpsLevel(octx, T, 1)
        rl = hl.r()
        _, fr = hl.frac()
        want, _ = hm.frac_consts(T * hm.prod(qs[:k - 1]), qs[:k - 1] + rl, range(k - 1, 2 * (k - 1)))
        assert [bits_of(v) for v in fr] == [bits_of(v) for v in want]


# ---- the steps ----

def _check_scale_round(out, nu, d, qs, rs, num, side, divisor, out_mods, label):
    """every coefficient: nu bit-exact with the model, out = floor(t X / D + 1/2) + delta;
    returns (coefficients inside the error bound, coefficients with delta != 0)"""
    N = d.shape[1]
    k = len(qs)
    fp, ex = hm.frac_consts(num, list(qs) + list(rs), range(k) if side == "q" else range(k, k + len(rs)))
    inside = moved = 0
    for n in range(N):
        xq = [int(v) for v in d[:k, n]]
        xr = [int(v) for v in d[k:, n]]
        val, nu_m, nu_ex = hm.scale_round(xq, xr, qs, rs, num, side, lambda X: F(T * X, divisor))
        assert bits_of(nu[n]) == bits_of(nu_m), (label, n)
        assert [int(v) for v in out[:, n]] == [val % m for m in out_mods], (label, n)
        xs = xq if side == "q" else xr
        bound = hm.fma_error_bound(xs, fp, ex, 0.5)
        assert abs(F(nu_m) - nu_ex) <= bound, (label, n)
        dist = abs(nu_ex - round(nu_ex))
        X = hm.crt(xq + xr, list(qs) + list(rs))
        exact = math.floor(F(T * X, divisor) + F(1, 2))
        if dist > bound:
            assert val == exact, (label, n)
        else:
            inside += 1
        moved += val != exact
    return inside, moved


@pytest.mark.parametrize("logN,bits", SHAPES)
def test_scale_round_qr_r_exact(logN, bits):
    octx, mods, N = setup(logN, bits)
    k = len(mods) - 1
    qs = mods[:k]
    h = ob.Hps(octx, T)
    rs = h.r()
    rng = np.random.default_rng(logN + k)
    d = np.stack([rng.integers(0, s, N, dtype=np.uint64) for s in qs + rs])
    out, nu = h.scale_round(d)
    inside, moved = _check_scale_round(out, nu, d, qs, rs, T * hm.prod(rs), "q", hm.prod(qs), rs, "QR->R")
    print(f"QR->R {bits}: {inside}/{N} coefficients within the FP64 bound, {moved} rounded off the exact value")


@pytest.mark.parametrize("logN,bits", SHAPES)
def test_scale_round_qlrl_ql_exact(logN, bits):
    octx, mods, N = setup(logN, bits)
    k = len(mods) - 1
    qs = mods[:k]
    rng = np.random.default_rng(2 * logN + k)
    tools = [("overq", ob.Hps(octx, T, overq=True), qs)]
    if k > 1:
        tools.append(("level 1", ob.HpsLevel(octx, T, 1), qs[:k - 1]))
    for label, h, ql in tools:
        rs = h.r()
        d = np.stack([rng.integers(0, s, N, dtype=np.uint64) for s in ql + rs])
        out, nu = h.scale_round(d)
        inside, moved = _check_scale_round(out, nu, d, ql, rs, T * hm.prod(ql), "r", hm.prod(rs), ql, label)
        print(f"QlRl->Ql {label} {bits}: {inside}/{N} within the bound, {moved} rounded off the exact value")


@pytest.mark.parametrize("logN,bits", SHAPES)
def test_bconv_hps_exact(logN, bits):
    octx, mods, N = setup(logN, bits)
    k = len(mods) - 1
    qs = mods[:k]
    rs = ob.Hps(octx, T).r()
    rng = np.random.default_rng(5 * logN + k)
    for ib, obase in ((qs, rs), (rs, qs)):
        x = np.stack([rng.integers(0, q, N, dtype=np.uint64) for q in ib])
        # coefficients whose sum sits on .5 within rounding: plant exact half-way values
        # (X = (Q - 1) / 2 and (Q + 1) / 2, the centred lift's two sides)
        Q = hm.prod(ib)
        for n, X in ((0, (Q - 1) // 2), (1, (Q + 1) // 2), (2, 0), (3, Q - 1)):
            x[:, n] = [X % q for q in ib]
        out, s = ob.bconv_hps(ib, obase, x)
        for n in range(N):
            res, s_m, v, val = hm.bconv_hps(ib, obase, [int(t) for t in x[:, n]])
            assert bits_of(s[n]) == bits_of(s_m), n
            assert [int(t) for t in out[:, n]] == res, n
            X = hm.crt([int(t) for t in x[:, n]], ib)
            sigma = sum(F(int(t) * pow((Q // q) % q, -1, q) % q, q) for t, q in zip(x[:, n], ib))
            assert v in (math.floor(sigma), math.floor(sigma) + 1)
            assert val == X - (v - math.floor(sigma)) * Q


@pytest.mark.parametrize("logN,bits", SHAPES)
def test_scale_q_ql_exact(logN, bits):
    octx, mods, N = setup(logN, bits)
    k = len(mods) - 1
    if k < 2:
        pytest.skip("needs two data primes")
    qs = mods[:k]
    rng = np.random.default_rng(7 * logN + k)
    for dd in range(1, k):
        x = np.stack([rng.integers(0, q, N, dtype=np.uint64) for q in qs])
        out, nu = ob.hps_scale_q_ql_nu(octx, dd, x)
        ql, dropped = qs[:k - dd], qs[k - dd:]
        # the divisor is Q / Ql and the numerator Ql: round(Ql X / Q) = round(X / dropped)
        fp, ex = hm.frac_consts(hm.prod(ql), qs, range(k - dd, k))
        for n in range(N):
            xs = [int(v) for v in x[:, n]]
            nu_m = hm.chain(xs[k - dd:], fp, 0.5)
            assert bits_of(nu[n]) == bits_of(nu_m), (dd, n)
            nu_ex = F(1, 2) + sum(v * e for v, e in zip(xs[k - dd:], ex))
            X = hm.crt(xs, qs)
            val = math.floor(F(X, hm.prod(dropped)) + F(1, 2)) + (int(nu_m) - math.floor(nu_ex))
            assert [int(v) for v in out[:, n]] == [val % q for q in ql], (dd, n)


# ---- the whole multiply, composed from the pinned steps ----

def _extend(ib, obase, poly):
    """bConv_HPS of every coefficient of poly [len(ib)][N]: the integer it represents"""
    return [hm.bconv_hps(ib, obase, [int(t) for t in poly[:, n]])[3] for n in range(poly.shape[1])]


def _tensor(A, B):
    return (negacyclic(A[0], B[0]),
            [u + v for u, v in zip(negacyclic(A[0], B[1]), negacyclic(A[1], B[0]))],
            negacyclic(A[1], B[1]))


def _round_all(z, qs, rs, num, side, divisor):
    ss = list(qs) + list(rs)
    k = len(qs)
    out = []
    for zn in z:
        xq, xr = [zn % s for s in qs], [zn % s for s in rs]
        out.append(hm.scale_round(xq, xr, qs, rs, num, side, lambda X: F(T * X, divisor))[0])
    return out


@pytest.mark.parametrize("logN,bits", SHAPES)
@pytest.mark.parametrize("square", [False, True])
def test_hps_multiply_composed_exact(logN, bits, square):
    octx, mods, N = setup(logN, bits)
    k = len(mods) - 1
    qs = mods[:k]
    rng = np.random.default_rng(11 * logN + k + square)
    ct1 = np.stack([np.stack([rng.integers(0, q, N, dtype=np.uint64) for q in qs]) for _ in range(2)])
    ct2 = ct1 if square else np.stack([np.stack([rng.integers(0, q, N, dtype=np.uint64) for q in qs])
                                       for _ in range(2)])
    # mul_tech hps: extend both Q -> R, tensor, QR -> R, bConv_HPS R -> Q
    h = ob.Hps(octx, T)
    rs = h.r()
    A = [_extend(qs, rs, ct1[p]) for p in range(2)]
    B = A if square else [_extend(qs, rs, ct2[p]) for p in range(2)]
    got = h.multiply(ct1, ct2, square=square)
    for p, z in enumerate(_tensor(A, B)):
        w = _round_all(z, qs, rs, T * hm.prod(rs), "q", hm.prod(qs))
        tr = np.array([[wn % r for wn in w] for r in rs], np.uint64)
        want = [hm.bconv_hps(rs, qs, [int(t) for t in tr[:, n]])[0] for n in range(N)]
        assert np.array_equal(got[p], np.array(want, np.uint64).T), ("hps", p)
    # hps_overq: the second operand switched Q -> R and converted back (unless squaring)
    ho = ob.Hps(octx, T, overq=True)
    ro = ho.r()
    A = [_extend(qs, ro, ct1[p]) for p in range(2)]
    if square:
        B = A
    else:
        B = [_extend(ro, qs, ho.modswitch(ct2[p])) for p in range(2)]
    got = ho.multiply(ct1, ct2, square=square)
    for p, z in enumerate(_tensor(A, B)):
        w = _round_all(z, qs, ro, T * hm.prod(qs), "r", hm.prod(ro))
        assert np.array_equal(got[p], np.array([[wn % q for wn in w] for q in qs], np.uint64)), ("overq", p)


@pytest.mark.parametrize("logN,bits", [SHAPES[0], SHAPES[2]])
def test_hps_level_multiply_composed_exact(logN, bits):
    octx, mods, N = setup(logN, bits)
    k = len(mods) - 1
    qs = mods[:k]
    rng = np.random.default_rng(13 * logN + k)
    ct1 = np.stack([np.stack([rng.integers(0, q, N, dtype=np.uint64) for q in qs]) for _ in range(2)])
    ct2 = np.stack([np.stack([rng.integers(0, q, N, dtype=np.uint64) for q in qs]) for _ in range(2)])
    for d in range(1, k):
        h = ob.HpsLevel(octx, T, d)
        ql, rl = qs[:k - d], h.r()
        Qd = hm.prod(qs[k - d:])
        A = []
        for p in range(2):
            s = ob.hps_scale_q_ql(octx, d, ct1[p])       # pinned by test_scale_q_ql_exact
            A.append(_extend(ql, rl, s))
        B = [_extend(rl, ql, h.modswitch(ct2[p])) for p in range(2)]
        got = h.multiply(ct1, ct2, expand_polys=3)
        for p, z in enumerate(_tensor(A, B)):
            w = _round_all(z, ql, rl, T * hm.prod(ql), "r", hm.prod(rl))
            want = np.zeros((k, N), np.uint64)
            for i, q in enumerate(ql):
                want[i] = [(wn % q) * (Qd % q) % q for wn in w]
            assert np.array_equal(got[p], want), (d, p)
