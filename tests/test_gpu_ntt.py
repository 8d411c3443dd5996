"""GPU parity of the NTT/INTT kernels against the oracle (bit-exact)."""
import numpy as np
import pytest

import bigint_ref as br
import oracle_bind as ob
import pyphantom as ph
from gpu_helpers import make_ctx, rand_limbs, say

pytestmark = pytest.mark.gpu


def run_ntt(ctx, data, start, inverse=False):
    buf = ph.DeviceBuffer.from_array(data)
    L = data.shape[0]
    if inverse:
        ph.ntt_inverse(ctx, buf, L, start)
    else:
        ph.ntt_forward(ctx, buf, L, start)
    ph.device_synchronize()
    return buf.to_array(data.shape)


def test_n12_three_60bit_ntt_then_intt_bit_exact():
    """N=2^12, 3 limbs of the largest 60-bit primes = 1 mod 2N, uniform input."""
    say("case: N=4096 L=3")
    N = 4096
    ctx, octx, mods = make_ctx(N, [60, 60, 60], 1)
    rng = np.random.default_rng(1)
    a = rand_limbs(rng, mods, N)
    fwd = run_ntt(ctx, a, 0)
    assert np.array_equal(fwd, octx.ntt_forward(a, 0))
    inv = run_ntt(ctx, fwd, 0, inverse=True)
    assert np.array_equal(inv, a)


@pytest.mark.parametrize("logN", [11, 12, 13, 14, 15, 16, 17])
def test_ntt_parity_all_degrees(logN):
    N = 1 << logN
    say(f"case: ntt parity N=2^{logN}")
    ctx, octx, mods = make_ctx(N, [60, 50, 50, 50, 40, 60, 60], 2)
    rng = np.random.default_rng(logN)
    a = rand_limbs(rng, mods, N)
    got = run_ntt(ctx, a, 0)
    assert np.array_equal(got, octx.ntt_forward(a, 0))
    # inverse on arbitrary canonical input, with a table offset
    b = rand_limbs(rng, mods[3:6], N)
    got_i = run_ntt(ctx, b, 3, inverse=True)
    assert np.array_equal(got_i, octx.ntt_inverse(b, 3))
    assert np.array_equal(run_ntt(ctx, got_i, 3), b)


def test_ntt_edge_values():
    """all-zero, all q-1 and the reference's constant input (test/ntt_test.cu:102-121)"""
    say("case: ntt edge values N=2^16")
    N = 1 << 16
    ctx, octx, mods = make_ctx(N, [60, 50, 49, 60], 1)
    qs = np.array(mods, np.uint64)[:, None]
    L = len(mods)
    alt = np.where(np.arange(N) % 2 == 0, qs - 1, 0).astype(np.uint64)
    half = np.broadcast_to(qs // 2, (L, N)).copy()
    for a in (np.zeros((L, N), np.uint64), np.broadcast_to(qs - 1, (L, N)).copy(), np.full((L, N), 2, np.uint64),
              alt, half):
        f = run_ntt(ctx, a, 0)
        assert np.array_equal(f, octx.ntt_forward(a, 0))
        assert np.array_equal(run_ntt(ctx, f, 0, inverse=True), a)
        assert np.array_equal(run_ntt(ctx, a, 0, inverse=True), octx.ntt_inverse(a, 0))


@pytest.fixture
def fused():
    """Switch the launchers to the one-launch form for every limb count."""
    old = ph.tune_get("ntt_fused"), ph.tune_get("ntt_fused_min_limbs")
    yield lambda g: (ph.tune_set("ntt_fused", g), ph.tune_set("ntt_fused_min_limbs", 0))
    ph.tune_set("ntt_fused", old[0])
    ph.tune_set("ntt_fused_min_limbs", old[1])


@pytest.mark.parametrize("G", [1, 2, 4])
@pytest.mark.parametrize("logN", [13, 15, 16, 17])
def test_ntt_fused_parity(fused, G, logN):
    """One workgroup per limb runs both passes: same outputs as the oracle."""
    N = 1 << logN
    say(f"case: fused ntt G={G} N=2^{logN}")
    fused(G)
    ctx, octx, mods = make_ctx(N, [60, 50, 50, 49, 40, 60, 60], 2)
    rng = np.random.default_rng(100 + logN)
    a = rand_limbs(rng, mods, N)
    got = run_ntt(ctx, a, 0)
    assert np.array_equal(got, octx.ntt_forward(a, 0))
    b = rand_limbs(rng, mods[2:6], N)
    got_i = run_ntt(ctx, b, 2, inverse=True)
    assert np.array_equal(got_i, octx.ntt_inverse(b, 2))
    assert np.array_equal(run_ntt(ctx, got_i, 2), b)


def test_ntt_max_degree_edge_values():
    """N = 2^17 (the reference's largest ring degree, include/host/defines.h:23):
    edge inputs and a round trip on 60- and 50-bit primes"""
    say("case: ntt edge values N=2^17")
    N = 1 << 17
    ctx, octx, mods = make_ctx(N, [60, 50, 60], 1)
    qs = np.array(mods, np.uint64)[:, None]
    L = len(mods)
    for a in (np.zeros((L, N), np.uint64), np.broadcast_to(qs - 1, (L, N)).copy(), np.full((L, N), 2, np.uint64)):
        f = run_ntt(ctx, a, 0)
        assert np.array_equal(f, octx.ntt_forward(a, 0))
        assert np.array_equal(run_ntt(ctx, f, 0, inverse=True), a)


@pytest.mark.parametrize("logN", [12, 16])
def test_ntt_61bit_edge_values(logN):
    """Primes just below 2^61, the largest a context takes: the integer butterflies'
    lazy ranges are widest there (forward values below 6q < 2^64, inverse below 3q).
    All q - 1, alternating q - 1 / 0 and uniform inputs, forward and inverse against
    the oracle and round trips; then multiply_and_relin with such special primes and
    every input at q - 1."""
    N = 1 << logN
    say(f"case: ntt 61-bit edge values N=2^{logN}")
    p61 = br.primes_below(61, N, 5)   # coeff_modulus_create stops at 60 bits
    mods = p61[:3] + ph.coeff_modulus_create(N, [50]) + p61[3:]
    assert all(q >> 60 == 1 for q in p61)
    parms = ph.EncryptionParameters(ph.SCHEME_CKKS)
    parms.set_poly_modulus_degree(N)
    parms.set_coeff_modulus(mods)
    parms.set_special_modulus_size(2)
    ctx = ph.context(parms)
    octx = ob.Ctx(logN, mods, 2)
    qs = np.array(mods, np.uint64)[:, None]
    L = len(mods)
    rng = np.random.default_rng(61 + logN)
    alt = np.where(np.arange(N) % 2 == 0, qs - 1, 0).astype(np.uint64)
    for a in (np.broadcast_to(qs - 1, (L, N)).copy(), alt, rand_limbs(rng, mods, N)):
        f = run_ntt(ctx, a, 0)
        assert np.array_equal(f, octx.ntt_forward(a, 0))
        assert np.array_equal(run_ntt(ctx, f, 0, inverse=True), a)
        assert np.array_equal(run_ntt(ctx, a, 0, inverse=True), octx.ntt_inverse(a, 0))
    Ql = ctx.size_Q
    key = np.stack([np.stack([np.broadcast_to(qs - 1, (L, N)).copy() for _ in range(2)]) for _ in range(ctx.dnum)])
    rk = ph.relin_key(ctx, host=key)
    a = np.stack([np.broadcast_to(qs[:Ql] - 1, (Ql, N)).copy() for _ in range(2)])
    c1 = ph.ciphertext(ctx, 1, 2)
    c1.upload(a)
    c2 = ph.ciphertext(ctx, 1, 2)
    c2.upload(a)
    ph.multiply_and_relin_inplace(ctx, c1, c2, rk)
    assert np.array_equal(c1.download(N), octx.multiply_and_relin(Ql, a, a, key))
