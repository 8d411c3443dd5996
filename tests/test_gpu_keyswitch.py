"""GPU parity of modup / key inner product / moddown / keyswitch /
multiply_and_relin / rescale / apply_galois against the oracle, bit-exact
on every output limb, on seeded uniform inputs."""
import numpy as np
import pytest

import oracle_bind as ob
import pyphantom as ph
from gpu_helpers import make_ctx, rand_ct, rand_key, rand_limbs, say, upload_ct

pytestmark = pytest.mark.gpu

SMALL = dict(N=4096, bits=[60, 50, 50, 50, 50, 50, 60, 60], P=2)   # 6 data primes, alpha 2


@pytest.fixture(scope="module")
def small():
    ob.set_threads(8)
    ctx, octx, mods = make_ctx(SMALL["N"], SMALL["bits"], SMALL["P"])
    rng = np.random.default_rng(100)
    key = rand_key(rng, mods, SMALL["N"], ctx.dnum)
    return ctx, octx, mods, key


@pytest.mark.parametrize("chain", [1, 2, 3, 6])
def test_modup_inner_moddown_pieces(small, chain):
    ctx, octx, mods, key = small
    N = ctx.N
    Ql = ctx.limbs_at(chain)
    say(f"case: pieces N={N} chain={chain} Ql={Ql}")
    rng = np.random.default_rng(chain)
    c2 = rand_limbs(rng, mods[:Ql], N)
    beta = octx.beta(Ql)
    QlP = Ql + ctx.size_P
    c2b = ph.DeviceBuffer.from_array(c2)
    tb = ph.DeviceBuffer(beta * QlP * N)
    ph.modup(ctx, chain, c2b, tb)
    ph.device_synchronize()
    t = tb.to_array((beta, QlP, N))
    assert np.array_equal(t, octx.modup(Ql, c2))
    rk = ph.relin_key(ctx, host=key)
    cxb = ph.DeviceBuffer(2 * QlP * N)
    ph.key_inner_product(ctx, chain, tb, rk, cxb)
    ph.device_synchronize()
    cx = cxb.to_array((2, QlP, N))
    assert np.array_equal(cx, octx.inner_product(Ql, t, key))
    outb = ph.DeviceBuffer(2 * Ql * N)
    ph.moddown_from_ntt(ctx, chain, cxb, outb, 2)
    ph.device_synchronize()
    out = outb.to_array((2, Ql, N))
    assert np.array_equal(out[0], octx.moddown(Ql, cx[0]))
    assert np.array_equal(out[1], octx.moddown(Ql, cx[1]))


@pytest.mark.parametrize("chain", [1, 4, 6])
def test_multiply_and_relin_small(small, chain):
    ctx, octx, mods, key = small
    N = ctx.N
    Ql = ctx.limbs_at(chain)
    say(f"case: mulrelin N={N} chain={chain}")
    rng = np.random.default_rng(10 + chain)
    a, b = rand_ct(rng, mods[:Ql], N), rand_ct(rng, mods[:Ql], N)
    rk = ph.relin_key(ctx, host=key)
    c1, c2 = upload_ct(ctx, a, chain), upload_ct(ctx, b, chain)
    ph.multiply_and_relin_inplace(ctx, c1, c2, rk)
    got = c1.download(N)
    assert got.shape == (2, Ql, N)
    assert np.array_equal(got, octx.multiply_and_relin(Ql, a, b, key))


def test_square_aliased_operands(small):
    ctx, octx, mods, key = small
    N = ctx.N
    say("case: square (ct1 is ct2)")
    rng = np.random.default_rng(77)
    a = rand_ct(rng, mods[:6], N)
    rk = ph.relin_key(ctx, host=key)
    c1 = upload_ct(ctx, a, 1)
    ph.multiply_and_relin_inplace(ctx, c1, c1, rk)
    assert np.array_equal(c1.download(N), octx.multiply_and_relin(6, a, a, key))


@pytest.mark.parametrize("chain", [1, 3, 5])
def test_rescale_small(small, chain):
    ctx, octx, mods, key = small
    N = ctx.N
    Ql = ctx.limbs_at(chain)
    say(f"case: rescale chain={chain}")
    rng = np.random.default_rng(20 + chain)
    a = rand_ct(rng, mods[:Ql], N)
    c = upload_ct(ctx, a, chain)
    c.set_scale(2.0 ** 40)
    r = ph.rescale_to_next(ctx, c)
    assert r.chain_index() == chain + 1
    assert np.array_equal(r.download(N), octx.rescale(Ql, a))
    assert r.scale() == pytest.approx(2.0 ** 40 / mods[Ql - 1])


def test_rescale_end_of_chain_raises(small):
    ctx, octx, mods, key = small
    c = upload_ct(ctx, rand_ct(np.random.default_rng(0), mods[:1], ctx.N), 6)
    with pytest.raises(ValueError):
        ph.rescale_to_next(ctx, c)


@pytest.mark.parametrize("step", [1, -3, 0])
def test_apply_galois_small(small, step):
    ctx, octx, mods, key = small
    N = ctx.N
    say(f"case: galois step={step}")
    elt = ph.galois_elt_from_step(step, N)
    rng = np.random.default_rng(30 + step)
    Ql = ctx.limbs_at(2)
    a = rand_ct(rng, mods[:Ql], N)
    gk = ph.galois_key(ctx, [elt], host=key[None])
    c = upload_ct(ctx, a, 2)
    ph.rotate_inplace(ctx, c, step, gk)
    assert np.array_equal(c.download(N), octx.apply_galois(Ql, a, elt, key))


def test_missing_galois_key_raises(small):
    ctx, octx, mods, key = small
    gk = ph.galois_key(ctx, [ph.galois_elt_from_step(1, ctx.N)], host=key[None])
    c = upload_ct(ctx, rand_ct(np.random.default_rng(0), mods[:6], ctx.N), 1)
    with pytest.raises(ValueError):
        ph.apply_galois_inplace(ctx, c, ph.galois_elt_from_step(2, ctx.N), gk)


# ---------------------------------------------------------------- full size
BITS_N16_Q35_P12 = [60] + [50] * 34 + [60] * 12


@pytest.fixture(scope="module")
def n16_q35():
    ob.set_threads(16)
    N = 1 << 16
    ctx, octx, mods = make_ctx(N, BITS_N16_Q35_P12, 12)
    assert ctx.size_Q == 35 and ctx.dnum == 3
    rng = np.random.default_rng(2026)
    key = rand_key(rng, mods, N, 3)
    return ctx, octx, mods, key


def test_n16_q35_multiply_and_relin_bit_exact(n16_q35):
    """N=2^16, 35 data limbs + 12 special, dnum = 3 (digits 12, 12, 11)."""
    ctx, octx, mods, key = n16_q35
    N = ctx.N
    say("case: N=2^16 mulrelin top level")
    rng = np.random.default_rng(5)
    a, b = rand_ct(rng, mods[:35], N), rand_ct(rng, mods[:35], N)
    rk = ph.relin_key(ctx, host=key)
    c1, c2 = upload_ct(ctx, a, 1), upload_ct(ctx, b, 1)
    ph.multiply_and_relin_inplace(ctx, c1, c2, rk)
    assert np.array_equal(c1.download(N), octx.multiply_and_relin(35, a, b, key))


def test_n16_q35_lower_level_and_rotation(n16_q35):
    ctx, octx, mods, key = n16_q35
    N = ctx.N
    chain = 16   # 20 limbs: digits 12, 8
    Ql = ctx.limbs_at(chain)
    say(f"case: N=2^16 chain={chain} mulrelin + galois")
    rng = np.random.default_rng(6)
    a, b = rand_ct(rng, mods[:Ql], N), rand_ct(rng, mods[:Ql], N)
    rk = ph.relin_key(ctx, host=key)
    c1, c2 = upload_ct(ctx, a, chain), upload_ct(ctx, b, chain)
    ph.multiply_and_relin_inplace(ctx, c1, c2, rk)
    assert np.array_equal(c1.download(N), octx.multiply_and_relin(Ql, a, b, key))
    elt = ph.galois_elt_from_step(5, N)
    gk = ph.galois_key(ctx, [elt], host=key[None])
    c3 = upload_ct(ctx, a, chain)
    ph.rotate_inplace(ctx, c3, 5, gk)
    assert np.array_equal(c3.download(N), octx.apply_galois(Ql, a, elt, key))


def test_n16_q35_rescale(n16_q35):
    ctx, octx, mods, key = n16_q35
    N = ctx.N
    say("case: N=2^16 rescale")
    a = rand_ct(np.random.default_rng(8), mods[:35], N)
    c = upload_ct(ctx, a, 1)
    r = ph.rescale_to_next(ctx, c)
    assert np.array_equal(r.download(N), octx.rescale(35, a))


def test_reference_valid_shape_36_12():
    """size_Q % size_P == 0 (the reference keygen's own assumption, src/secretkey.cu:309)."""
    ob.set_threads(16)
    N = 1 << 16
    say("case: 36+12 mulrelin")
    ctx, octx, mods = make_ctx(N, [60] + [50] * 35 + [60] * 12, 12)
    assert ctx.dnum == 3
    rng = np.random.default_rng(36)
    key = rand_key(rng, mods, N, 3)
    a, b = rand_ct(rng, mods[:36], N), rand_ct(rng, mods[:36], N)
    rk = ph.relin_key(ctx, host=key)
    c1, c2 = upload_ct(ctx, a, 1), upload_ct(ctx, b, 1)
    ph.multiply_and_relin_inplace(ctx, c1, c2, rk)
    assert np.array_equal(c1.download(N), octx.multiply_and_relin(36, a, b, key))


# ---------------------------------------------------------------- kernel variants
@pytest.mark.parametrize("bconv_mfma,ntt_fp64", [(0, 0), (0, 1), (1, 0), (1, 1)])
def test_variants_bit_exact(bconv_mfma, ntt_fp64):
    """Both FastBConv kernels (VALU / matrix cores) and both NTT arithmetics
    (64-bit integer / exact FP64) give the oracle's outputs."""
    say(f"case: variants bconv_mfma={bconv_mfma} ntt_fp64={ntt_fp64}")
    old = (ph.tune_get("bconv_mfma"), ph.tune_get("ntt_fp64"))
    try:
        ph.tune_set("bconv_mfma", bconv_mfma)
        ph.tune_set("ntt_fp64", ntt_fp64)
        N = 1 << 14
        ctx, octx, mods = make_ctx(N, [60] + [50] * 9 + [60] * 4, 4)
        rng = np.random.default_rng(50 + 2 * bconv_mfma + ntt_fp64)
        key = rand_key(rng, mods, N, ctx.dnum)
        rk = ph.relin_key(ctx, host=key)
        for chain in (1, 3):
            Ql = ctx.limbs_at(chain)
            a, b = rand_ct(rng, mods[:Ql], N), rand_ct(rng, mods[:Ql], N)
            c1, c2 = upload_ct(ctx, a, chain), upload_ct(ctx, b, chain)
            ph.multiply_and_relin_inplace(ctx, c1, c2, rk)
            assert np.array_equal(c1.download(N), octx.multiply_and_relin(Ql, a, b, key))
    finally:
        ph.tune_set("bconv_mfma", old[0])
        ph.tune_set("ntt_fp64", old[1])


def test_bconv_extreme_inputs_mfma():
    """All inputs q-1 (largest slices everywhere) through modup and moddown."""
    say("case: bconv extreme inputs")
    N = 1 << 13
    ctx, octx, mods = make_ctx(N, [60] + [50] * 13 + [60] * 14, 14)
    Ql = 14
    qs = np.array(mods[:Ql], np.uint64)[:, None]
    c2 = np.broadcast_to(qs - 1, (Ql, N)).copy()
    beta = octx.beta(Ql)
    QlP = Ql + 14
    c2b = ph.DeviceBuffer.from_array(c2)
    tb = ph.DeviceBuffer(beta * QlP * N)
    ph.modup(ctx, 1, c2b, tb)
    ph.device_synchronize()
    assert np.array_equal(tb.to_array((beta, QlP, N)), octx.modup(Ql, c2))
    qlp = np.array(mods, np.uint64)[:, None]
    cx = np.broadcast_to(qlp - 1, (QlP, N)).copy()
    cxb = ph.DeviceBuffer.from_array(cx)
    ob = ph.DeviceBuffer(Ql * N)
    ph.moddown_from_ntt(ctx, 1, cxb, ob, 1)
    ph.device_synchronize()
    assert np.array_equal(ob.to_array((Ql, N)), octx.moddown(Ql, cx))


@pytest.mark.parametrize("G", [2, 4])
def test_fused_ntt_hot_path(G):
    """The one-launch NTT form (one workgroup per limb, both passes) under
    every fused load/store policy: modup, moddown's subtract-and-scale store,
    rescale's load/store pair, rotation, and the batch plan."""
    say(f"case: fused ntt hot path G={G}")
    old = (ph.tune_get("ntt_fused"), ph.tune_get("ntt_fused_min_limbs"))
    try:
        ph.tune_set("ntt_fused", G)
        ph.tune_set("ntt_fused_min_limbs", 0)
        N = 1 << 14
        ctx, octx, mods = make_ctx(N, [60] + [50] * 9 + [60] * 4, 4)
        rng = np.random.default_rng(60 + G)
        key = rand_key(rng, mods, N, ctx.dnum)
        rk = ph.relin_key(ctx, host=key)
        for chain in (1, 4):
            Ql = ctx.limbs_at(chain)
            a, b = rand_ct(rng, mods[:Ql], N), rand_ct(rng, mods[:Ql], N)
            c1, c2 = upload_ct(ctx, a, chain), upload_ct(ctx, b, chain)
            ph.multiply_and_relin_inplace(ctx, c1, c2, rk)
            want = octx.multiply_and_relin(Ql, a, b, key)
            assert np.array_equal(c1.download(N), want)
            c1.set_scale(2.0 ** 50)
            r = ph.rescale_to_next(ctx, c1)
            assert np.array_equal(r.download(N), octx.rescale(Ql, want))
        elt = ph.galois_elt_from_step(3, N)
        gk = ph.galois_key(ctx, [elt], host=key[None])
        a = rand_ct(rng, mods[:10], N)
        c = upload_ct(ctx, a, 1)
        ph.rotate_inplace(ctx, c, 3, gk)
        assert np.array_equal(c.download(N), octx.apply_galois(10, a, elt, key))
        xs = [rand_ct(rng, mods[:10], N) for _ in range(3)]
        ys = [rand_ct(rng, mods[:10], N) for _ in range(3)]
        cx = [upload_ct(ctx, x, 1) for x in xs]
        cy = [upload_ct(ctx, y, 1) for y in ys]
        bt = ph.batch.multiply_and_relin(ctx, cx, cy, rk)
        bt.run()
        ph.device_synchronize()
        for i in range(3):
            assert np.array_equal(cx[i].download(N), octx.multiply_and_relin(10, xs[i], ys[i], key)), i
    finally:
        ph.tune_set("ntt_fused", old[0])
        ph.tune_set("ntt_fused_min_limbs", old[1])


@pytest.mark.parametrize("N,bits,P,chain", [(1 << 12, [60, 50, 50, 50, 50, 50, 60, 60], 2, 2),
                                            (1 << 16, [60] + [50] * 34 + [60] * 12, 12, 1)])
def test_hoisting_bit_exact(N, bits, P, chain):
    """hoisting_inplace: the sum of several rotations sharing one modup."""
    ob.set_threads(16)
    say(f"case: hoisting N={N} chain={chain}")
    ctx, octx, mods = make_ctx(N, bits, P)
    rng = np.random.default_rng(N + chain)
    steps = [1, 2, -3, 7]
    elts = [ph.galois_elt_from_step(s, N) for s in steps]
    keys = np.stack([rand_key(rng, mods, N, ctx.dnum) for _ in elts])
    gk = ph.galois_key(ctx, elts, host=keys)
    Ql = ctx.limbs_at(chain)
    a = rand_ct(rng, mods[:Ql], N)
    c = upload_ct(ctx, a, chain)
    ph.hoisting_inplace(ctx, c, gk, steps)
    assert np.array_equal(c.download(N), octx.hoisting(Ql, a, elts, keys))
    with pytest.raises(ValueError):
        ph.hoisting_inplace(ctx, c, gk, [5])


@pytest.mark.parametrize("N,k", [(1 << 13, 4), (1 << 15, 16)])
def test_hoisting_bfv_bit_exact(N, k):
    """BFV hoisting_inplace on a coefficient-form ciphertext, bit-exact against
    the oracle (coefficient-domain automorphism of c0, key switch of c1)."""
    ob.set_threads(16)
    say(f"case: bfv hoisting N={N} k={k}")
    mods = ph.coeff_modulus_create(N, [54] * k + [60])
    parms = ph.EncryptionParameters(ph.SCHEME_BFV)
    parms.set_poly_modulus_degree(N)
    parms.set_coeff_modulus(mods)
    parms.set_special_modulus_size(1)
    parms.set_plain_modulus(65537)
    ctx = ph.context(parms)
    octx = ob.Ctx(N.bit_length() - 1, mods, 1)
    rng = np.random.default_rng(N + k)
    steps = [1, -2, 0]
    elts = [ph.galois_elt_from_step(s, N) for s in steps]
    keys = np.stack([rand_key(rng, mods, N, ctx.dnum) for _ in elts])
    gk = ph.galois_key(ctx, elts, host=keys)
    a = rand_ct(rng, mods[:k], N)
    c = ph.ciphertext(ctx, 1, 2)
    assert not c.is_ntt_form()
    c.upload(a)
    ph.hoisting_inplace(ctx, c, gk, steps)
    assert not c.is_ntt_form()
    assert np.array_equal(c.download(N), octx.hoisting(k, a, elts, keys, bfv=True))
    c.set_ntt_form(True)
    with pytest.raises(ValueError):
        ph.hoisting_inplace(ctx, c, gk, steps)


def test_max_degree_multiply_relin_rotate_rescale():
    """N = 2^17 end to end (multiply_and_relin, rotate, rescale) against the oracle"""
    ob.set_threads(16)
    N = 1 << 17
    say("case: N=2^17 mulrelin/rotate/rescale")
    ctx, octx, mods = make_ctx(N, [60, 50, 50, 50, 60, 60], 2)
    rng = np.random.default_rng(17)
    key = rand_key(rng, mods, N, ctx.dnum)
    rk = ph.relin_key(ctx, host=key)
    Ql = ctx.size_Q
    a, b = rand_ct(rng, mods[:Ql], N), rand_ct(rng, mods[:Ql], N)
    c1, c2 = upload_ct(ctx, a, 1), upload_ct(ctx, b, 1)
    ph.multiply_and_relin_inplace(ctx, c1, c2, rk)
    want = octx.multiply_and_relin(Ql, a, b, key)
    assert np.array_equal(c1.download(N), want)
    elt = ph.galois_elt_from_step(3, N)
    gkey = rand_key(rng, mods, N, ctx.dnum)
    gk = ph.galois_key(ctx, [elt], host=gkey[None])
    ph.rotate_inplace(ctx, c1, 3, gk)
    rot = octx.apply_galois(Ql, want, elt, gkey)
    assert np.array_equal(c1.download(N), rot)
    c1.set_scale(2.0 ** 40)
    r = ph.rescale_to_next(ctx, c1)
    assert np.array_equal(r.download(N), octx.rescale(Ql, rot))


def test_ciphertext_of_another_context_rejected(small):
    """a ciphertext whose ring degree or limb count is not the context's is refused
    on the host before any kernel reads it"""
    ctx, octx, mods, key = small
    say("case: foreign ciphertext rejected")
    other, _, omods = make_ctx(2 * SMALL["N"], SMALL["bits"], SMALL["P"])
    rng = np.random.default_rng(5)
    foreign = upload_ct(other, rand_ct(rng, omods[:6], 2 * SMALL["N"]), 1)
    mine = upload_ct(ctx, rand_ct(rng, mods[:6], SMALL["N"]), 1)
    rk = ph.relin_key(ctx, host=key)
    with pytest.raises(ValueError):
        ph.multiply_and_relin_inplace(ctx, mine, foreign, rk)
    with pytest.raises(ValueError):
        ph.relinearize_inplace(ctx, foreign, rk)
    with pytest.raises(ValueError):
        ph.rescale_to_next(ctx, foreign)
    empty = ph.ciphertext()
    with pytest.raises(ValueError):
        ph.multiply_inplace(ctx, mine, empty)
    # a key of the other context (same dnum, other ring degree)
    okey = rand_key(rng, omods, 2 * SMALL["N"], other.dnum)
    ork = ph.relin_key(other, host=okey)
    with pytest.raises(ValueError):
        ph.multiply_and_relin_inplace(ctx, mine, mine, ork)
    # the valid operand is untouched by the refused calls
    ph.multiply_and_relin_inplace(ctx, mine, mine, rk)
    assert mine.size() == 2


def test_mod_switch_drops_limbs(small):
    """mod_switch_to_next / mod_switch_to_inplace (CKKS): the first limbs of each
    polynomial, NTT form and scale kept; the switched ciphertext multiplies with
    one born at that level as the oracle does"""
    ctx, octx, mods, key = small
    N = ctx.N
    say("case: mod switch drop")
    rng = np.random.default_rng(9)
    a = rand_ct(rng, mods[:6], N)
    c = upload_ct(ctx, a, 1)
    c.set_scale(2.0 ** 40)
    d = ph.mod_switch_to_next(ctx, c)
    assert d.chain_index() == 2 and d.is_ntt_form() and d.scale() == 2.0 ** 40
    assert np.array_equal(d.download(N), a[:, :5])
    ph.mod_switch_to_next_inplace(ctx, c)
    assert np.array_equal(c.download(N), a[:, :5])
    ph.mod_switch_to_inplace(ctx, c, 4)
    assert c.chain_index() == 4 and np.array_equal(c.download(N), a[:, :3])
    b = rand_ct(rng, mods[:3], N)
    e = upload_ct(ctx, b, 4)
    rk = ph.relin_key(ctx, host=key)
    ph.multiply_and_relin_inplace(ctx, c, e, rk)
    assert np.array_equal(c.download(N), octx.multiply_and_relin(3, a[:, :3], b, key))
    with pytest.raises(ValueError):
        ph.mod_switch_to_inplace(ctx, c, 2)          # higher level
    last = upload_ct(ctx, rand_ct(rng, mods[:1], N), 6)
    with pytest.raises(ValueError):
        ph.mod_switch_to_next(ctx, last)              # end of the chain
