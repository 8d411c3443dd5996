"""Two shapes the reference accepts at the edges of this build's kernels, bit-exact
against the oracle: ring degree N = 2^11 (the smallest of the reference's 2-D NTT,
include/ntt.cuh:134-137; one 128-thread tile per limb here), for CKKS and BFV, and
a special modulus of more than 16 primes (the FastBConv inputs no longer fit the
registers of the VALU kernel's fast path, nor the matrix-core form)."""
import numpy as np
import pytest

import oracle_bind as ob
import pyphantom as ph
from gpu_helpers import make_ctx, rand_ct, rand_key, rand_limbs, say, upload_ct

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def n11():
    ob.set_threads(8)
    N = 1 << 11
    ctx, octx, mods = make_ctx(N, [60, 40, 40, 40, 40, 60, 60], 2)
    key = rand_key(np.random.default_rng(11), mods, N, ctx.dnum)
    return ctx, octx, mods, key


def test_n11_ntt_bit_exact(n11):
    ctx, octx, mods, key = n11
    N = ctx.N
    say("case: N=2^11 NTT")
    rng = np.random.default_rng(1)
    a = rand_limbs(rng, mods, N)
    buf = ph.DeviceBuffer.from_array(a)
    ph.ntt_forward(ctx, buf, len(mods), 0)
    ph.device_synchronize()
    f = buf.to_array(a.shape)
    assert np.array_equal(f, octx.ntt_forward(a, 0))
    ph.ntt_inverse(ctx, buf, len(mods), 0)
    ph.device_synchronize()
    assert np.array_equal(buf.to_array(a.shape), a)
    b = rand_limbs(rng, mods[2:5], N)
    buf = ph.DeviceBuffer.from_array(b)
    ph.ntt_inverse(ctx, buf, 3, 2)
    ph.device_synchronize()
    assert np.array_equal(buf.to_array(b.shape), octx.ntt_inverse(b, 2))


@pytest.mark.parametrize("chain", [1, 3])
def test_n11_multiply_relin_rotate_rescale(n11, chain):
    ctx, octx, mods, key = n11
    N = ctx.N
    Ql = ctx.limbs_at(chain)
    say(f"case: N=2^11 mulrelin/rotate/rescale chain={chain}")
    rng = np.random.default_rng(20 + chain)
    a, b = rand_ct(rng, mods[:Ql], N), rand_ct(rng, mods[:Ql], N)
    rk = ph.relin_key(ctx, host=key)
    c1, c2 = upload_ct(ctx, a, chain), upload_ct(ctx, b, chain)
    ph.multiply_and_relin_inplace(ctx, c1, c2, rk)
    prod = octx.multiply_and_relin(Ql, a, b, key)
    assert np.array_equal(c1.download(N), prod)
    elt = ph.galois_elt_from_step(3, N)
    gk = ph.galois_key(ctx, [elt], host=key[None])
    ph.rotate_inplace(ctx, c1, 3, gk)
    rot = octx.apply_galois(Ql, prod, elt, key)
    assert np.array_equal(c1.download(N), rot)
    r = ph.rescale_to_next(ctx, c1)
    assert np.array_equal(r.download(N), octx.rescale(Ql, rot))


def test_n11_batch_plan_and_autotune(n11):
    """a batch plan over items of two levels (two branches), eager and replayed,
    and the autotuner, which skips the forms that need a 4096-coefficient tile"""
    ctx, octx, mods, key = n11
    N = ctx.N
    say("case: N=2^11 batch plan")
    rng = np.random.default_rng(5)
    chains = [1, 2, 1, 3]
    data = [(c, ctx.limbs_at(c), rand_ct(rng, mods[:ctx.limbs_at(c)], N),
             rand_ct(rng, mods[:ctx.limbs_at(c)], N)) for c in chains]
    want = [octx.multiply_and_relin(Ql, a, b, key) for c, Ql, a, b in data]
    rk = ph.relin_key(ctx, host=key)
    c1 = [upload_ct(ctx, a, c) for c, _, a, _ in data]
    c2 = [upload_ct(ctx, b, c) for c, _, _, b in data]
    bt = ph.batch.multiply_and_relin(ctx, c1, c2, rk)
    stream = ph.stream_create()
    bt.run(stream)
    ph.synchronize(stream)
    for i, w in enumerate(want):
        assert np.array_equal(c1[i].download(N), w), i
    bt.capture(stream)
    for rep in range(2):
        for i, (c, Ql, a, b) in enumerate(data):
            c1[i].upload(a)
        ph.device_synchronize()
        bt.replay(stream)
        ph.synchronize(stream)
        for i, w in enumerate(want):
            assert np.array_equal(c1[i].download(N), w), (rep, i)
    report = ctx.autotune()
    assert report["choice"]["ntt_fused"] in (0, -1)
    x = rand_ct(rng, mods[:ctx.size_Q], N)
    y = upload_ct(ctx, x, 1)
    ph.multiply_and_relin_inplace(ctx, y, y, rk)
    assert np.array_equal(y.download(N), octx.multiply_and_relin(ctx.size_Q, x, x, key))


@pytest.fixture(scope="module")
def wide_p():
    """N = 2^12, 25 data primes and 20 special primes: dnum 2 (digits of 20 and 5)"""
    ob.set_threads(8)
    N = 1 << 12
    ctx, octx, mods = make_ctx(N, [60] + [50] * 24 + [60] * 20, 20)
    assert ctx.size_P == 20 and ctx.dnum == 2
    key = rand_key(np.random.default_rng(12), mods, N, ctx.dnum)
    return ctx, octx, mods, key


@pytest.mark.parametrize("chain", [1, 8])
def test_special_modulus_20_multiply_relin_rotate(wide_p, chain):
    ctx, octx, mods, key = wide_p
    N = ctx.N
    Ql = ctx.limbs_at(chain)
    say(f"case: size_P=20 mulrelin/rotate chain={chain}")
    rng = np.random.default_rng(40 + chain)
    a, b = rand_ct(rng, mods[:Ql], N), rand_ct(rng, mods[:Ql], N)
    rk = ph.relin_key(ctx, host=key)
    c1, c2 = upload_ct(ctx, a, chain), upload_ct(ctx, b, chain)
    ph.multiply_and_relin_inplace(ctx, c1, c2, rk)
    prod = octx.multiply_and_relin(Ql, a, b, key)
    assert np.array_equal(c1.download(N), prod)
    elt = ph.galois_elt_from_step(-1, N)
    gk = ph.galois_key(ctx, [elt], host=key[None])
    ph.rotate_inplace(ctx, c1, -1, gk)
    assert np.array_equal(c1.download(N), octx.apply_galois(Ql, prod, elt, key))


def test_special_modulus_20_batch(wide_p):
    ctx, octx, mods, key = wide_p
    N = ctx.N
    say("case: size_P=20 batch plan")
    rng = np.random.default_rng(9)
    chains = [1, 5, 1, 12]
    data = [(c, ctx.limbs_at(c), rand_ct(rng, mods[:ctx.limbs_at(c)], N),
             rand_ct(rng, mods[:ctx.limbs_at(c)], N)) for c in chains]
    rk = ph.relin_key(ctx, host=key)
    c1 = [upload_ct(ctx, a, c) for c, _, a, _ in data]
    c2 = [upload_ct(ctx, b, c) for c, _, _, b in data]
    bt = ph.batch.multiply_and_relin(ctx, c1, c2, rk)
    bt.run()
    ph.device_synchronize()
    for i, (c, Ql, a, b) in enumerate(data):
        assert np.array_equal(c1[i].download(N), octx.multiply_and_relin(Ql, a, b, key)), i


@pytest.mark.parametrize("tech", [ph.MUL_TECH_BEHZ, ph.MUL_TECH_HPS])
def test_n11_bfv_multiply_relinearize(tech):
    """BFV at N = 2^11 (t = 65537, two 54-bit primes + one special): multiply with
    BEHZ or HPS, then relinearise, each against the oracle"""
    ob.set_threads(8)
    N = 1 << 11
    say(f"case: N=2^11 bfv tech={tech}")
    mods = ph.coeff_modulus_create(N, [54, 54, 60])
    parms = ph.EncryptionParameters(ph.SCHEME_BFV)
    parms.set_poly_modulus_degree(N)
    parms.set_coeff_modulus(mods)
    parms.set_special_modulus_size(1)
    parms.set_plain_modulus(65537)
    parms.set_mul_tech(tech)
    ctx = ph.context(parms)
    octx = ob.Ctx(11, mods, 1)
    tool = ob.Behz(octx, 65537) if tech == ph.MUL_TECH_BEHZ else ob.Hps(octx, 65537, overq=False)
    rng = np.random.default_rng(tech)
    a, b = rand_ct(rng, mods[:2], N), rand_ct(rng, mods[:2], N)
    c1 = ph.ciphertext(ctx, 1, 2)
    c1.upload(a)
    c2 = ph.ciphertext(ctx, 1, 2)
    c2.upload(b)
    ph.multiply_inplace(ctx, c1, c2)
    ph.device_synchronize()
    prod = tool.multiply(a, b)
    assert np.array_equal(c1.download(N), prod)
    key = rand_key(rng, mods, N, ctx.dnum)
    rk = ph.relin_key(ctx, host=key)
    ph.relinearize_inplace(ctx, c1, rk)
    assert np.array_equal(c1.download(N), ob.bfv_relinearize(octx, prod, key))
