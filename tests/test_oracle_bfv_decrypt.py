"""Pins of the oracle's BFV decryption scale-and-round (orc_bfv_decrypt: BEHZ,
and the four HPS forms the reference picks from qMSB, tMSB and size_Q) by the
closed form m = round(t X / Q) mod t in Python integers, X the CRT integer of
the phase:
- a valid phase X = floor(Q / t) m + e with |e| < Q / (4t) decrypts to m in
  every coefficient, for every method;
- a uniform phase agrees with round(t X / Q) mod t except within floating-point
  (HPS) or gamma (BEHZ) reach of a rounding tie, and then by exactly one.
A dropped constant, a wrong index or a transposed table fails both."""
import numpy as np
import pytest

import oracle_bind as ob
from bigint_ref import crt, primes_below

N = 1024
T = 65537

# (name, data prime bits, data primes, t bits or None for 65537, behz)
CASES = [
    ("behz", 54, 3, None, True),
    ("hps_small_lazy", 30, 3, None, False),     # qMSB + tMSB + log2 size_Q < 52
    ("hps_small", 40, 3, None, False),          # qMSB + log2 size_Q < 52 <= qMSB + tMSB + ...
    ("hps_large_lazy", 54, 4, None, False),     # qMSB/2 + tMSB + log2 size_Q < 52
    ("hps_large", 54, 4, 30, False),            # neither
]


def setup(bits, k, tbits):
    data = primes_below(bits, N, k)
    special = [p for p in primes_below(60, N, 2) if p not in data][:1]
    t = T if tbits is None else primes_below(tbits, 1, 1)[0]
    octx = ob.Ctx(N.bit_length() - 1, data + special, 1)
    Q = 1
    for q in data:
        Q *= q
    return octx, data, Q, t


def residues(X, moduli):
    return np.stack([np.array([x % q for x in X], dtype=np.uint64) for q in moduli])


@pytest.mark.parametrize("name,bits,k,tbits,behz", CASES)
def test_valid_phase_decrypts_to_m(name, bits, k, tbits, behz):
    octx, data, Q, t = setup(bits, k, tbits)
    rng = np.random.default_rng(bits + k)
    delta = Q // t
    m = [int(v) for v in rng.integers(0, t, N)]
    bound = Q // (4 * t)
    e = [int(v) for v in rng.integers(-(1 << 62), 1 << 62, N)]
    e = [v % (2 * bound) - bound for v in e]
    X = [(delta * mi + ei) % Q for mi, ei in zip(m, e)]
    got = ob.bfv_decrypt(octx, t, residues(X, data), behz)
    assert [int(g) % t for g in got] == m


@pytest.mark.parametrize("name,bits,k,tbits,behz", CASES)
def test_uniform_phase_rounds_to_nearest(name, bits, k, tbits, behz):
    octx, data, Q, t = setup(bits, k, tbits)
    rng = np.random.default_rng(7 * bits + k)
    X = [crt([int(rng.integers(0, q)) for q in data], data)[0] for _ in range(N)]
    got = ob.bfv_decrypt(octx, t, residues(X, data), behz)
    want = [((t * x + Q // 2) // Q) % t for x in X]
    diff = [(int(g) - w) % t for g, w in zip(got, want)]
    bad = [d for d in diff if d != 0]
    assert all(d in (1, t - 1) for d in bad), bad[:5]
    assert len(bad) <= N // 200


@pytest.mark.parametrize("name,bits,k,tbits,behz", [c for c in CASES if not c[4]])
def test_hps_forms_bit_exact_fp64_model(name, bits, k, tbits, behz):
    """the HPS forms' FP64 arithmetic, pinned bit for bit: the oracle equals
    tests/hps_model.py's exact-rational evaluation of the reference kernel (fused
    multiply-adds in input order, every other double operation rounded once,
    the integer sum wrapping at 2^64, llround halves away from zero) at every
    coefficient, on uniform phases and on phases a few units from a rounding tie
    of t X / Q, where the floating-point evaluation alone decides the result.
    In the non-lazy small form (a small integer sum) the products' roundings
    decide such coefficients: an unfused or a reordered evaluation gives other
    results there, so the check tells those apart. In the other forms the 64-bit
    integer sum added last is large enough that the double's resolution (2^-3 ..
    2^-19) hides those roundings: every variant gives the same results."""
    import hps_model as hm

    octx, data, Q, t = setup(bits, k, tbits)
    q_max = max(octx.moduli[:len(data)])
    rng = np.random.default_rng(11 * bits + k)
    uniform = np.stack([rng.integers(0, q, N, dtype=np.uint64) for q in data])
    m = [int(v) for v in rng.integers(0, t, N)]
    d = [int(v) for v in rng.integers(-3, 4, N)]
    X = [((2 * mi + 1) * Q) // (2 * t) + di for mi, di in zip(m, d)]
    ties = residues(X, data)
    for phase in (uniform, ties):
        got = [int(v) for v in ob.bfv_decrypt(octx, t, phase, behz)]
        rows = [[int(v) for v in row] for row in phase]
        assert got == hm.decrypt_hps(rows, data, q_max, t)
    unfused = hm.decrypt_hps(rows, data, q_max, t, fused=False)
    reordered = hm.decrypt_hps(rows, data, q_max, t, reverse=True)
    n_unf = sum(a != b for a, b in zip(got, unfused))
    n_rev = sum(a != b for a, b in zip(got, reordered))
    print(f"{name}: {2 * N} coefficients bit-exact; near ties an unfused evaluation differs at {n_unf}, "
          f"a reversed order at {n_rev} of {N}")
    if name == "hps_small":
        assert n_unf > 0 and n_rev > 0
    else:
        assert n_unf == 0 and n_rev == 0
