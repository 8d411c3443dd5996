"""BFV with mul_tech hps_overq_leveled (src/evaluate.cu:549-1027,
src/eval_key_switch.cu:95-182, src/evaluate.cu:1567-1865): the number of
dropped levels follows the noise degree carried in the ciphertext, which the
tests set through the ciphertext stream header. Every path is compared bit for
bit with the oracle: multiply (all three polynomials expanded back to Q), the
fused multiply + relinearise (c2 key-switched at the dropped level),
relinearise, apply_galois and hoisting; and a real encryption decrypts."""
import struct

import numpy as np
import pytest

import oracle_bind as ob
import pyphantom as ph
from bigint_ref import crt
from gpu_helpers import rand_ct, rand_key, say

pytestmark = pytest.mark.gpu

T = 65537
N = 1 << 13
K = 4


@pytest.fixture(scope="module")
def lv():
    ob.set_threads(8)
    mods = ph.coeff_modulus_create(N, [54] * K + [60])
    parms = ph.EncryptionParameters(ph.SCHEME_BFV)
    parms.set_poly_modulus_degree(N)
    parms.set_coeff_modulus(mods)
    parms.set_special_modulus_size(1)
    parms.set_plain_modulus(T)
    parms.set_mul_tech(ph.MUL_TECH_HPS_OVERQ_LEVELED)
    ctx = ph.context(parms)
    octx = ob.Ctx(13, mods, 1)
    return ctx, octx, mods


def with_degree(ctx, arr, deg):
    """a coefficient-form ciphertext whose noise degree is deg (stream header field)"""
    c = ph.ciphertext(ctx, 1, arr.shape[0])
    c.upload(arr)
    blob = bytearray(c.save())
    blob[48:56] = struct.pack("<Q", deg)
    return ph.ciphertext.load(ctx, bytes(blob))


def degree(ct):
    return struct.unpack("<Q", ct.save()[48:56])[0]


def add_mod(a, b, mods):
    out = a.astype(object) + b.astype(object)
    return (out % np.array(mods[:a.shape[0]], dtype=object)[:, None]).astype(np.uint64)


def bfv_keyswitch(octx, c2, key, Ql):
    """the BFV key switch of a coefficient-form c2 at Ql limbs: the CKKS key switch of
    its NTT, brought back"""
    ks = octx.keyswitch(Ql, octx.ntt_forward(c2, 0), key)
    return np.stack([octx.ntt_inverse(ks[p], 0) for p in range(2)])


def test_level_choice_spans_drops(lv):
    ctx, octx, mods = lv
    drops = [ob.hps_levels_to_drop(octx, T, D - 1) for D in (1, 4, 6, 7)]
    assert drops == [0, 1, 2, 3]


@pytest.mark.parametrize("deg", [1, 4, 6, 7])
def test_multiply_bit_exact(lv, deg):
    ctx, octx, mods = lv
    say(f"case: leveled multiply deg={deg}")
    d = ob.hps_levels_to_drop(octx, T, deg - 1)
    rng = np.random.default_rng(deg)
    a, b = rand_ct(rng, mods[:K], N), rand_ct(rng, mods[:K], N)
    c1, c2 = with_degree(ctx, a, deg), with_degree(ctx, b, deg - 1 if deg > 1 else 1)
    ph.multiply_inplace(ctx, c1, c2)
    assert c1.size() == 3 and degree(c1) == deg + 1
    want = ob.HpsLevel(octx, T, d).multiply(a, b)
    assert np.array_equal(c1.download(N), want)
    # squaring extends the one operand like the first
    c3 = with_degree(ctx, a, deg)
    ph.multiply_inplace(ctx, c3, c3)
    assert np.array_equal(c3.download(N), ob.HpsLevel(octx, T, d).multiply(a, a, square=True))


@pytest.mark.parametrize("deg", [1, 4, 6])
def test_fused_multiply_relin_bit_exact(lv, deg):
    ctx, octx, mods = lv
    say(f"case: leveled mul+relin deg={deg}")
    d = ob.hps_levels_to_drop(octx, T, deg - 1)
    Ql = K - d
    rng = np.random.default_rng(100 + deg)
    key = rand_key(rng, mods, N, ctx.dnum)
    rk = ph.relin_key(ctx, host=key)
    a, b = rand_ct(rng, mods[:K], N), rand_ct(rng, mods[:K], N)
    c1, c2 = with_degree(ctx, a, deg), with_degree(ctx, b, deg)
    ph.multiply_and_relin_inplace(ctx, c1, c2, rk)
    assert c1.size() == 2 and degree(c1) == deg + 1
    m3 = ob.HpsLevel(octx, T, d).multiply(a, b, expand_polys=2)
    ks = bfv_keyswitch(octx, m3[2, :Ql], key, Ql)
    if d:
        ks = np.stack([ob.hps_expand(octx, d, ks[p]) for p in range(2)])
    want = np.stack([add_mod(m3[p], ks[p], mods) for p in range(2)])
    assert np.array_equal(c1.download(N), want)


@pytest.mark.parametrize("deg", [2, 5])
def test_relinearize_bit_exact(lv, deg):
    """relinearize_inplace of a size-3 ciphertext of degree deg: c2 scaled down by the
    drop for depth deg - 1, key-switched there, expanded and added"""
    ctx, octx, mods = lv
    say(f"case: leveled relinearize deg={deg}")
    d = ob.hps_levels_to_drop(octx, T, deg - 1)
    Ql = K - d
    rng = np.random.default_rng(200 + deg)
    key = rand_key(rng, mods, N, ctx.dnum)
    rk = ph.relin_key(ctx, host=key)
    x = rand_ct(rng, mods[:K], N, size=3)
    c = with_degree(ctx, x, deg)
    ph.relinearize_inplace(ctx, c, rk)
    c2 = ob.hps_scale_q_ql(octx, d, x[2]) if d else x[2]
    ks = bfv_keyswitch(octx, c2, key, Ql)
    if d:
        ks = np.stack([ob.hps_expand(octx, d, ks[p]) for p in range(2)])
    want = np.stack([add_mod(x[p], ks[p], mods) for p in range(2)])
    assert np.array_equal(c.download(N), want)


def coeff_automorphism(x, g, mods):
    L, n = x.shape
    j = np.arange(n, dtype=np.int64)
    dd = (j * g) % (2 * n)
    neg = dd >= n
    dd = np.where(neg, dd - n, dd)
    out = np.zeros_like(x)
    for i in range(L):
        q = np.uint64(mods[i])
        out[i, dd] = np.where(neg & (x[i] != 0), q - x[i], x[i])
    return out


@pytest.mark.parametrize("deg", [1, 5])
def test_apply_galois_and_hoisting_bit_exact(lv, deg):
    """apply_galois: c0 <- pi(c0) + expand(ks0), c1 <- expand(ks1), ks the key switch of
    pi(c1) scaled to the drop for a key switch at depth deg - 1; hoisting: c0, c1
    scaled down, hoisted at that level, expanded back"""
    ctx, octx, mods = lv
    say(f"case: leveled galois/hoisting deg={deg}")
    d = ob.hps_levels_to_drop(octx, T, deg - 1, key_switch=True)
    Ql = K - d
    rng = np.random.default_rng(300 + deg)
    steps = [1, -2]
    elts = [ph.galois_elt_from_step(s, N) for s in steps]
    keys = np.stack([rand_key(rng, mods, N, ctx.dnum) for _ in elts])
    gk = ph.galois_key(ctx, elts, host=keys)
    a = rand_ct(rng, mods[:K], N)
    c = with_degree(ctx, a, deg)
    ph.apply_galois_inplace(ctx, c, elts[0], gk)
    assert degree(c) == deg
    p0 = coeff_automorphism(a[0], elts[0], mods)
    p1 = coeff_automorphism(a[1], elts[0], mods)
    c2 = ob.hps_scale_q_ql(octx, d, p1) if d else p1
    ks = bfv_keyswitch(octx, c2, keys[0], Ql)
    if d:
        ks = np.stack([ob.hps_expand(octx, d, ks[p]) for p in range(2)])
    want = np.stack([add_mod(p0, ks[0], mods), ks[1]])
    assert np.array_equal(c.download(N), want)
    h = with_degree(ctx, a, deg)
    ph.hoisting_inplace(ctx, h, gk, steps)
    sc = np.stack([ob.hps_scale_q_ql(octx, d, a[p]) for p in range(2)]) if d else a
    hz = octx.hoisting(Ql, sc, elts, keys, bfv=True)
    if d:
        hz = np.stack([ob.hps_expand(octx, d, hz[p]) for p in range(2)])
    assert np.array_equal(h.download(N), hz)


def test_batch_plans_reject_leveled(lv):
    ctx, octx, mods = lv
    rng = np.random.default_rng(4)
    rk = ph.relin_key(ctx, host=rand_key(rng, mods, N, ctx.dnum))
    a = with_degree(ctx, rand_ct(rng, mods[:K], N), 1)
    b = with_degree(ctx, rand_ct(rng, mods[:K], N), 1)
    with pytest.raises(ValueError):
        ph.batch.multiply_and_relin(ctx, [a], [b], rk)


def poly_mul_rns(octx, x, y, L):
    fx, fy = octx.ntt_forward(x[:L], 0), octx.ntt_forward(y[:L], 0)
    prod = np.stack([(fx[i].astype(object) * fy[i].astype(object) % octx.moduli[i]).astype(np.uint64)
                     for i in range(L)])
    return octx.ntt_inverse(prod, 0)


@pytest.mark.parametrize("deg", [1, 6])
def test_encrypt_multiply_relin_decrypt(lv, deg):
    """real encryptions carrying noise degree deg: the product decrypts to m1 m2 mod t
    whether or not levels were dropped"""
    ctx, octx, mods = lv
    say(f"case: leveled semantic deg={deg}")
    qs = mods[:K]
    Q = 1
    for q in qs:
        Q *= q
    delta = Q // T
    rng = np.random.default_rng(40 + deg)
    s = rng.integers(-1, 2, N)
    res = lambda v, ms: np.stack([np.array([int(x) % q for x in v], np.uint64) for q in ms])
    s_all = res(s, mods)

    def encrypt(m):
        a = np.stack([rng.integers(0, q, N, dtype=np.uint64) for q in qs])
        e = res(rng.integers(-4, 5, N), qs)
        dm = np.stack([np.array([(delta % q) * int(v) % q for v in m], np.uint64) for q in qs])
        as_ = poly_mul_rns(octx, a, s_all, K)
        c0 = np.stack([(dm[i].astype(object) + e[i].astype(object) - as_[i].astype(object)) % qs[i]
                       for i in range(K)]).astype(np.uint64)
        return np.stack([c0, a])

    m1, m2 = rng.integers(0, T, N), rng.integers(0, T, N)
    s_ntt = octx.ntt_forward(s_all, 0)
    s2_ntt = octx.ntt_forward(poly_mul_rns(octx, s_all, s_all, len(mods)), 0)
    ka = np.stack([np.stack([rng.integers(0, q, N, dtype=np.uint64) for q in mods]) for _ in range(ctx.dnum)])
    ke = np.stack([res(rng.integers(-4, 5, N), mods) for _ in range(ctx.dnum)])
    rk = ph.relin_key(ctx, host=octx.make_kswitch_key(s_ntt, s2_ntt, ka, ke))
    c1, c2 = with_degree(ctx, encrypt(m1), deg), with_degree(ctx, encrypt(m2), deg)
    ph.multiply_and_relin_inplace(ctx, c1, c2, rk)
    out = c1.download(N)
    v = out[0].astype(object) + poly_mul_rns(octx, out[1], s_all, K).astype(object)
    v = v % np.array(qs, dtype=object)[:, None]
    dec = [((crt([int(v[i, n]) for i in range(K)], qs)[0] * T + Q // 2) // Q) % T for n in range(N)]
    # negacyclic m1 m2 over the integers (|coefficients| < N t^2 < q0 / 2), then mod t
    pm = poly_mul_rns(ob.Ctx(13, [mods[0], mods[1]], 1),
                      np.stack([np.array([int(x) for x in m1], np.uint64)] * 2),
                      np.stack([np.array([int(x) for x in m2], np.uint64)] * 2), 1)[0]
    q0 = mods[0]
    want = [((int(x) if int(x) <= q0 // 2 else int(x) - q0) % T) for x in pm]
    assert want == dec


def test_copy_keeps_noise_degree(lv):
    """a copied ciphertext keeps its noise degree, so the non-inplace
    multiply_and_relin (copy, then in place) drops the same levels"""
    ctx, octx, mods = lv
    say("case: leveled copy keeps degree")
    rng = np.random.default_rng(77)
    key = rand_key(rng, mods, N, ctx.dnum)
    rk = ph.relin_key(ctx, host=key)
    a, b = rand_ct(rng, mods[:K], N), rand_ct(rng, mods[:K], N)
    c1, c2 = with_degree(ctx, a, 6), with_degree(ctx, b, 6)
    cp = ph.ciphertext()
    cp.copy_from(ctx, c1)
    assert degree(cp) == 6
    out = ph.multiply_and_relin(ctx, c1, c2, rk)
    assert degree(out) == 7 and degree(c1) == 6
    ph.multiply_and_relin_inplace(ctx, c1, c2, rk)
    assert np.array_equal(out.download(N), c1.download(N))
