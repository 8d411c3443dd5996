"""CKKS end to end on the GPU: encode -> encrypt -> HMult+relin -> rescale ->
rotate -> decrypt -> decode, with the error against the same computation in
FP64 on the plaintext values reported (printed) and bounded.

The encoder is pinned independently of its own decoder: the coefficient
polynomial it returns is evaluated directly at zeta^(5^j) (zeta = exp(i pi/N))
and must give scale * z_j, the slot convention of the reference's
PhantomCKKSEncoder (rotation group 5^j, src/ckks.cu:17-63)."""
import numpy as np
import pytest

import pyphantom as ph
from gpu_helpers import say

pytestmark = pytest.mark.gpu


def make(N, bits, P, steps=()):
    mods = ph.coeff_modulus_create(N, bits)
    parms = ph.EncryptionParameters(ph.SCHEME_CKKS)
    parms.set_poly_modulus_degree(N)
    parms.set_coeff_modulus(mods)
    parms.set_special_modulus_size(P)
    ctx = ph.context(parms)
    return ctx, mods


def centred(limb0, q0):
    v = limb0.astype(object)
    return np.array([int(x) - q0 if int(x) > q0 // 2 else int(x) for x in v], dtype=np.float64)


def rand_slots(rng, n):
    return rng.uniform(-1, 1, n) + 1j * rng.uniform(-1, 1, n)


def test_encoder_matches_canonical_embedding():
    N = 1 << 12
    say("case: encoder canonical embedding N=4096")
    ctx, mods = make(N, [60, 40, 40, 60], 1)
    enc = ph.ckks_encoder(ctx)
    rng = np.random.default_rng(1)
    z = rand_slots(rng, N // 2)
    scale = 2.0 ** 30
    m = enc.encode(z, scale, 1)
    # every limb holds the same integer polynomial
    c = centred(m[0], mods[0])
    assert np.array_equal(m[1], (c.astype(np.int64) % np.int64(mods[1])).astype(np.uint64))
    zeta = np.exp(1j * np.pi / N)
    k = np.arange(N)
    for j in (0, 1, 2, 7, N // 4, N // 2 - 1):
        e = pow(5, j, 2 * N)
        val = np.sum(c * zeta ** ((e * k) % (2 * N)))
        assert abs(val / scale - z[j]) < 1e-6, j
    back = enc.decode(m, 1, scale)
    assert np.max(np.abs(back - z)) < 1e-6


def test_encoder_rejects_bad_input():
    ctx, mods = make(1 << 12, [60, 40, 60], 1)
    enc = ph.ckks_encoder(ctx)
    with pytest.raises(ValueError):
        enc.encode(np.zeros(4096), 2.0 ** 30, 1)      # more than N/2 values
    with pytest.raises(ValueError):
        enc.encode(np.ones(4), 2.0 ** 170, 1)         # scale beyond the modulus
    with pytest.raises(ValueError):
        enc.encode(np.full(4, 1e45), 2.0 ** 30, 1)    # values too large


def chain_error(ctx, mods, scale, step, seed):
    """encrypt x, y; x*y relinearised, rescaled, rotated by step; decrypt and
    compare with the FP64 computation after every stage. Returns the max
    slot errors {fresh, mul_relin, rescale, rotate} and the final chain index."""
    N = ctx.N
    enc = ph.ckks_encoder(ctx)
    rng = np.random.default_rng(seed)
    x, y = rand_slots(rng, N // 2), rand_slots(rng, N // 2)
    sk = ph.secret_key(ctx, seed=seed)
    rk = sk.gen_relinkey(ctx, seed=seed + 1)
    elt = ph.galois_elt_from_step(step, N)
    gk = sk.create_galois_keys(ctx, [elt], seed=seed + 2)
    cx = sk.encrypt_symmetric(ctx, enc.encode(x, scale, 1), 1, seed=seed + 3)
    cy = sk.encrypt_symmetric(ctx, enc.encode(y, scale, 1), 1, seed=seed + 4)
    cx.set_scale(scale)
    cy.set_scale(scale)

    def err(ct, want):
        got = enc.decode(sk.decrypt(ctx, ct), ct.chain_index(), ct.scale())
        return float(np.max(np.abs(got - want)))

    e = {"fresh": err(cx, x)}
    ph.multiply_and_relin_inplace(ctx, cx, cy, rk)
    e["mul_relin"] = err(cx, x * y)
    r = ph.rescale_to_next(ctx, cx)
    e["rescale"] = err(r, x * y)
    ph.rotate_inplace(ctx, r, step, gk)
    e["rotate"] = err(r, np.roll(x * y, -step))
    return e, r.chain_index()


# The bounds below are set by the reference algorithm itself: its moddown
# uses the approximate FastBConv (src/rns_bconv.cu:212-229), which returns
# x + e*P with 0 <= e < alpha, so every key switch adds -e0 - e1*s with e of
# mean (alpha-1)/2; that non-centred term concentrates in the low slots
# (measured: coefficient-domain mean ~ +270, std ~ 830 at N=2^16), and the floor
# rescale (src/rns.cu:1170-1182) adds a smaller one of the same kind.
def test_hmult_relin_rescale_rotate_error_small():
    say("case: ckks chain error N=2^13")
    ctx, mods = make(1 << 13, [60, 40, 40, 40, 60, 60], 2)
    e, chain = chain_error(ctx, mods, 2.0 ** 40, 3, 10)
    say("ckks error N=2^13 L=4+2 scale=2^40: " + ", ".join(f"{k} {v:.3e}" for k, v in e.items()))
    assert chain == 2
    assert e["fresh"] < 1e-8 and e["mul_relin"] < 1e-8 and e["rescale"] < 1e-5 and e["rotate"] < 1e-5


def test_n16_q35_error_report():
    """N=2^16, 35 data limbs (60 + 34 x 50 bit), 12 special, dnum 3, scale 2^50."""
    say("case: ckks chain error N=2^16")
    ctx, mods = make(1 << 16, [60] + [50] * 34 + [60] * 12, 12)
    e, chain = chain_error(ctx, mods, 2.0 ** 50, 5, 20)
    say("ckks error N=2^16 L=35+12 scale=2^50: " + ", ".join(f"{k} {v:.3e}" for k, v in e.items()))
    assert chain == 2
    assert e["fresh"] < 1e-10 and e["mul_relin"] < 1e-10 and e["rescale"] < 1e-7 and e["rotate"] < 1e-6
