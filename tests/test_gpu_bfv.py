"""BFV on the GPU: BEHZ multiply and relinearisation bit-exact against the
oracle (coefficient-form ciphertexts), the config of 64 ciphertexts over
N in {2^13, 2^14, 2^15} with 4/8/16 limbs of 54-bit primes and t = 65537, and
a semantic check: encrypt m1, m2 -> multiply+relin -> decrypt = m1 m2 mod
(X^N + 1, t). The batched form (one plan per ring degree, joined into one graph)
is checked item by item against the plain path and the oracle."""
import numpy as np
import pytest

import oracle_bind as ob
import pyphantom as ph
from bigint_ref import crt
from gpu_helpers import rand_ct, rand_key, say

pytestmark = pytest.mark.gpu

T = 65537
LIMBS = {1 << 13: 4, 1 << 14: 8, 1 << 15: 16}


def make_bfv(N, k, tech=ph.MUL_TECH_BEHZ):
    mods = ph.coeff_modulus_create(N, [54] * k + [60])
    parms = ph.EncryptionParameters(ph.SCHEME_BFV)
    parms.set_poly_modulus_degree(N)
    parms.set_coeff_modulus(mods)
    parms.set_special_modulus_size(1)
    parms.set_plain_modulus(T)
    parms.set_mul_tech(tech)
    ctx = ph.context(parms)
    octx = ob.Ctx(N.bit_length() - 1, mods, 1)
    tool = ob.Behz(octx, T) if tech == ph.MUL_TECH_BEHZ else ob.Hps(octx, T, overq=tech == ph.MUL_TECH_HPS_OVERQ)
    return ctx, octx, tool, mods


def upload_bfv(ctx, arr):
    ct = ph.ciphertext(ctx, 1, arr.shape[0])
    assert not ct.is_ntt_form()
    ct.upload(arr)
    return ct


@pytest.fixture(scope="module")
def small():
    ob.set_threads(8)
    return make_bfv(1 << 13, 4)


def test_bfv_multiply_bit_exact(small):
    ctx, octx, beh, mods = small
    N = ctx.N
    say("case: bfv multiply N=2^13 k=4")
    rng = np.random.default_rng(1)
    a, b = rand_ct(rng, mods[:4], N), rand_ct(rng, mods[:4], N)
    c1, c2 = upload_bfv(ctx, a), upload_bfv(ctx, b)
    ph.multiply_inplace(ctx, c1, c2)
    ph.device_synchronize()
    assert c1.size() == 3
    assert np.array_equal(c1.download(N), beh.multiply(a, b))
    # squaring (both operands the same ciphertext)
    c3 = upload_bfv(ctx, a)
    ph.multiply_inplace(ctx, c3, c3)
    assert np.array_equal(c3.download(N), beh.multiply(a, a))


def test_bfv_relinearize_bit_exact(small):
    ctx, octx, beh, mods = small
    N = ctx.N
    say("case: bfv relinearize N=2^13 k=4")
    rng = np.random.default_rng(2)
    key = rand_key(rng, mods, N, ctx.dnum)
    rk = ph.relin_key(ctx, host=key)
    x = rand_ct(rng, mods[:4], N, size=3)
    c = upload_bfv(ctx, x)
    ph.relinearize_inplace(ctx, c, rk)
    assert c.size() == 2
    assert np.array_equal(c.download(N), ob.bfv_relinearize(octx, x, key))


def test_bfv_rejects_ntt_form_and_wrong_scheme(small):
    ctx, octx, beh, mods = small
    rng = np.random.default_rng(3)
    c = upload_bfv(ctx, rand_ct(rng, mods[:4], ctx.N))
    c.set_ntt_form(True)
    with pytest.raises(ValueError):
        ph.multiply_inplace(ctx, c, c)


def test_bfv_64_ciphertexts_batch_bit_exact():
    """64 ciphertexts, N uniform in {2^13, 2^14, 2^15} (4/8/16 limbs), multiply
    + relinearise each; items of every ring degree checked against the oracle."""
    ob.set_threads(16)
    say("case: 64 bfv ciphertexts")
    rng = np.random.default_rng(44)
    ctxs = {N: make_bfv(N, k) for N, k in LIMBS.items()}
    keys = {N: rand_key(np.random.default_rng(N), v[3], N, v[0].dnum) for N, v in ctxs.items()}
    rks = {N: ph.relin_key(v[0], host=keys[N]) for N, v in ctxs.items()}
    Ns = rng.choice(list(LIMBS), 64)
    items = []
    for i, N in enumerate(Ns):
        N = int(N)
        ctx, octx, beh, mods = ctxs[N]
        k = LIMBS[N]
        a = rand_ct(np.random.default_rng(1000 + i), mods[:k], N)
        b = rand_ct(np.random.default_rng(2000 + i), mods[:k], N)
        c1, c2 = upload_bfv(ctx, a), upload_bfv(ctx, b)
        ph.multiply_and_relin_inplace(ctx, c1, c2, rks[N])
        items.append((i, N, c1))
    ph.device_synchronize()
    checked = set()
    for i, N, c in items:
        if N in checked and i not in (0, 63):
            continue
        checked.add(N)
        ctx, octx, beh, mods = ctxs[N]
        k = LIMBS[N]
        a = rand_ct(np.random.default_rng(1000 + i), mods[:k], N)
        b = rand_ct(np.random.default_rng(2000 + i), mods[:k], N)
        want = ob.bfv_relinearize(octx, beh.multiply(a, b), keys[N])
        assert np.array_equal(c.download(N), want), (i, N)
    assert checked == set(LIMBS)


def negacyclic_mod_t(a, b, N, octx_t):
    fa = octx_t.ntt_forward(np.array(a, np.uint64)[None], 0)
    fb = octx_t.ntt_forward(np.array(b, np.uint64)[None], 0)
    prod = (fa.astype(object) * fb.astype(object) % T).astype(np.uint64)
    return octx_t.ntt_inverse(prod, 0)[0]


def poly_mul_rns(octx, x, y, L):
    """negacyclic products limb by limb through the oracle NTT"""
    fx = octx.ntt_forward(x, 0)
    fy = octx.ntt_forward(y, 0)
    qs = np.array(octx.moduli[:L], dtype=object)[:, None]
    return octx.ntt_inverse(((fx.astype(object) * fy.astype(object)) % qs).astype(np.uint64), 0)


@pytest.mark.parametrize("N,k,tech", [(1 << 13, 4, ph.MUL_TECH_HPS), (1 << 15, 16, ph.MUL_TECH_HPS),
                                      (1 << 13, 4, ph.MUL_TECH_HPS_OVERQ), (1 << 15, 16, ph.MUL_TECH_HPS_OVERQ)])
def test_hps_multiply_and_relin_bit_exact(N, k, tech):
    """mul_tech hps / hps_overq: FP64-corrected base extension and scale-and-round
    (over Q: the second operand switched Q -> R, t/R rounding QR -> Q; a squaring
    extends the one operand like the first), then relinearisation, bit-exact
    against the oracle (same fused multiply-adds)."""
    ob.set_threads(16)
    say(f"case: hps multiply N={N} k={k} tech={tech}")
    ctx, octx, hps, mods = make_bfv(N, k, tech)
    rng = np.random.default_rng(N + k)
    a, b = rand_ct(rng, mods[:k], N), rand_ct(rng, mods[:k], N)
    c1, c2 = upload_bfv(ctx, a), upload_bfv(ctx, b)
    ph.multiply_inplace(ctx, c1, c2)
    want3 = hps.multiply(a, b)
    assert np.array_equal(c1.download(N), want3)
    c3 = upload_bfv(ctx, a)
    ph.multiply_inplace(ctx, c3, c3)
    assert np.array_equal(c3.download(N), hps.multiply(a, a, square=True))
    key = rand_key(rng, mods, N, ctx.dnum)
    rk = ph.relin_key(ctx, host=key)
    c4, c5 = upload_bfv(ctx, a), upload_bfv(ctx, b)
    ph.multiply_and_relin_inplace(ctx, c4, c5, rk)
    assert np.array_equal(c4.download(N), ob.bfv_relinearize(octx, want3, key))


@pytest.mark.parametrize("tech", [ph.MUL_TECH_BEHZ, ph.MUL_TECH_HPS, ph.MUL_TECH_HPS_OVERQ])
def test_bfv_encrypt_multiply_relin_decrypt(tech):
    N, k = 1 << 13, 4
    say(f"case: bfv semantic multiply tech={tech}")
    ctx, octx, beh, mods = make_bfv(N, k, tech)
    qs = mods[:k]
    Q = 1
    for q in qs:
        Q *= q
    delta = Q // T
    rng = np.random.default_rng(7)
    s = rng.integers(-1, 2, N)
    res = lambda v, ms: np.stack([np.array([int(x) % q for x in v], np.uint64) for q in ms])
    s_all = res(s, mods)

    def encrypt(m):
        a = np.stack([rng.integers(0, q, N, dtype=np.uint64) for q in qs])
        e = res(rng.integers(-4, 5, N), qs)
        dm = np.stack([np.array([(delta % q) * int(v) % q for v in m], np.uint64) for q in qs])
        as_ = poly_mul_rns(octx, a, s_all[:k], k)
        c0 = np.stack([(np.array(dm[i], dtype=object) + e[i].astype(object) - as_[i].astype(object)) % qs[i]
                       for i in range(k)]).astype(np.uint64)
        return np.stack([c0, a])

    m1, m2 = rng.integers(0, T, N), rng.integers(0, T, N)
    # relinearisation key for s^2 from explicit samples (NTT form, QP limbs)
    s_ntt = octx.ntt_forward(s_all, 0)
    s2 = poly_mul_rns(octx, s_all, s_all, len(mods))
    s2_ntt = octx.ntt_forward(s2, 0)
    dnum = ctx.dnum
    ka = np.stack([np.stack([rng.integers(0, q, N, dtype=np.uint64) for q in mods]) for _ in range(dnum)])
    ke = np.stack([res(rng.integers(-4, 5, N), mods) for _ in range(dnum)])
    rk = ph.relin_key(ctx, host=octx.make_kswitch_key(s_ntt, s2_ntt, ka, ke))
    c1, c2 = upload_bfv(ctx, encrypt(m1)), upload_bfv(ctx, encrypt(m2))
    ph.multiply_and_relin_inplace(ctx, c1, c2, rk)
    out = c1.download(N)
    v = (out[0].astype(object) + poly_mul_rns(octx, out[1], s_all[:k], k).astype(object))
    v = v % np.array(qs, dtype=object)[:, None]
    dec = []
    for n in range(N):
        x = crt([int(v[i, n]) for i in range(k)], qs)[0]
        dec.append(((x * T + Q // 2) // Q) % T)
    octx_t = ob.Ctx(N.bit_length() - 1, [T, mods[0]], 1)
    want = negacyclic_mod_t([int(x) for x in m1], [int(x) for x in m2], N, octx_t)
    assert [int(x) for x in want] == dec


def plain_mul_relin(ctx, a, b, rk, square=False):
    c1 = upload_bfv(ctx, a)
    c2 = c1 if square else upload_bfv(ctx, b)
    ph.multiply_and_relin_inplace(ctx, c1, c2, rk)
    return c1.download(ctx.N)


@pytest.mark.parametrize("tech", [ph.MUL_TECH_BEHZ, ph.MUL_TECH_HPS, ph.MUL_TECH_HPS_OVERQ])
def test_bfv_batch_bit_exact(tech):
    """batched BFV multiply + relin == plain path == oracle, squaring item included"""
    ob.set_threads(8)
    N, k = 1 << 13, 4
    say(f"case: bfv batch tech={tech}")
    ctx, octx, tool, mods = make_bfv(N, k, tech)
    rng = np.random.default_rng(50 + tech)
    key = rand_key(rng, mods, N, ctx.dnum)
    rk = ph.relin_key(ctx, host=key)
    data = [(rand_ct(rng, mods[:k], N), rand_ct(rng, mods[:k], N)) for _ in range(5)]
    c1 = [upload_bfv(ctx, a) for a, _ in data]
    c2 = [upload_bfv(ctx, b) for _, b in data]
    c2[3] = c1[3]   # squaring
    bt = ph.batch.multiply_and_relin(ctx, c1, c2, rk)
    bt.run()
    ph.device_synchronize()
    for i, (a, b) in enumerate(data):
        got = c1[i].download(N)
        assert c1[i].size() == 2 and not c1[i].is_ntt_form()
        assert np.array_equal(got, plain_mul_relin(ctx, a, b, rk, square=(i == 3))), i
        if i in (0, 3):
            prod = tool.multiply(a, a, square=True) if i == 3 else tool.multiply(a, b)
            assert np.array_equal(got, ob.bfv_relinearize(octx, prod, key)), i


def test_bfv_batch_rejects():
    ctx, octx, tool, mods = make_bfv(1 << 13, 4)
    rng = np.random.default_rng(5)
    rk = ph.relin_key(ctx, host=rand_key(rng, mods, ctx.N, ctx.dnum))
    a = upload_bfv(ctx, rand_ct(rng, mods[:4], ctx.N))
    b = upload_bfv(ctx, rand_ct(rng, mods[:4], ctx.N))
    b.set_ntt_form(True)
    with pytest.raises(ValueError):
        ph.batch.multiply_and_relin(ctx, [a], [b], rk)


def test_bfv_mixed_rings_joined_batch_graph():
    """Batched form of 64 BFV ciphertexts, N uniform in {2^13, 2^14, 2^15}
    (4/8/16 limbs), one BEHZ plan per ring degree joined into one captured graph;
    every item equals the plain path, a sample equals the oracle; replayed twice
    with eager work (re-uploads) between the replays."""
    ob.set_threads(16)
    say("case: joined bfv batch over three ring degrees")
    rng = np.random.default_rng(404)
    ctxs = {N: make_bfv(N, k) for N, k in LIMBS.items()}
    keys = {N: rand_key(np.random.default_rng(N + 1), v[3], N, v[0].dnum) for N, v in ctxs.items()}
    rks = {N: ph.relin_key(v[0], host=keys[N]) for N, v in ctxs.items()}
    Ns = [int(x) for x in rng.choice(list(LIMBS), 64)]
    inputs = {}
    for i, N in enumerate(Ns):
        mods, k = ctxs[N][3], LIMBS[N]
        inputs[i] = (rand_ct(np.random.default_rng(3000 + i), mods[:k], N),
                     rand_ct(np.random.default_rng(4000 + i), mods[:k], N))
    want = {i: plain_mul_relin(ctxs[N][0], *inputs[i], rks[N]) for i, N in enumerate(Ns)}
    parts, c1s = [], {}
    for N in LIMBS:
        idx = [i for i, n in enumerate(Ns) if n == N]
        if not idx:
            continue
        ctx = ctxs[N][0]
        c1 = [upload_bfv(ctx, inputs[i][0]) for i in idx]
        c2 = [upload_bfv(ctx, inputs[i][1]) for i in idx]
        c1s.update(zip(idx, c1))
        parts.append(ph.batch.multiply_and_relin(ctx, c1, c2, rks[N]))
    jb = ph.batch.join(parts)
    stream = ph.stream_create()
    jb.capture(stream)
    for rep in range(2):
        if rep:
            for i, c in c1s.items():
                c.upload(inputs[i][0])
            ph.device_synchronize()
        jb.replay(stream)
        ph.synchronize(stream)
        for i, N in enumerate(Ns):
            assert np.array_equal(c1s[i].download(N), want[i]), (rep, i, N)
    seen = set()
    for i, N in enumerate(Ns):
        if N in seen:
            continue
        seen.add(N)
        ctx, octx, beh, mods = ctxs[N]
        a, b = inputs[i]
        assert np.array_equal(want[i], ob.bfv_relinearize(octx, beh.multiply(a, b), keys[N])), (i, N)
    assert seen == set(LIMBS)


def coeff_automorphism(x, g, mods):
    """X^j -> X^(j g) on coefficient-form limbs [L][N] (negacyclic: a wrap past N negates)"""
    L, N = x.shape
    j = np.arange(N, dtype=np.int64)
    d = (j * g) % (2 * N)
    neg = d >= N
    d = np.where(neg, d - N, d)
    out = np.zeros_like(x)
    for i in range(L):
        q = np.uint64(mods[i])
        out[i, d] = np.where(neg & (x[i] != 0), q - x[i], x[i])
    return out


def test_bfv_apply_galois_bit_exact(small):
    """BFV apply_galois / rotate on coefficient-form ciphertexts: the reference's
    coefficient-domain automorphism of c0 and c1, then the BFV key switch of
    pi(c1), i.e. the oracle's BFV relinearisation of (pi(c0), 0, pi(c1))"""
    ctx, octx, beh, mods = small
    N, k = ctx.N, ctx.size_Q
    say("case: bfv apply_galois N=2^13 k=4")
    rng = np.random.default_rng(21)
    elts = [ph.galois_elt_from_step(s, N) for s in (1, -3)] + [2 * N - 1]
    keys = np.stack([rand_key(rng, mods, N, ctx.dnum) for _ in elts])
    gk = ph.galois_key(ctx, elts, host=keys)
    a = rand_ct(rng, mods[:k], N)
    for elt, key in zip(elts, keys):
        c = upload_bfv(ctx, a)
        ph.apply_galois_inplace(ctx, c, elt, gk)
        assert not c.is_ntt_form()
        ct3 = np.stack([coeff_automorphism(a[0], elt, mods), np.zeros_like(a[0]), coeff_automorphism(a[1], elt, mods)])
        assert np.array_equal(c.download(N), ob.bfv_relinearize(octx, ct3, key)), elt
    c = upload_bfv(ctx, a)
    ph.rotate_inplace(ctx, c, 1, gk)
    ct3 = np.stack([coeff_automorphism(a[0], elts[0], mods), np.zeros_like(a[0]), coeff_automorphism(a[1], elts[0], mods)])
    assert np.array_equal(c.download(N), ob.bfv_relinearize(octx, ct3, keys[0]))
    c.set_ntt_form(True)
    with pytest.raises(ValueError):
        ph.apply_galois_inplace(ctx, c, elts[0], gk)


def test_bfv_rotate_decrypts():
    """encrypt m, rotate with a Galois key of pi(s), decrypt: the automorphism of m mod t"""
    N, k = 1 << 13, 4
    say("case: bfv rotate semantic")
    ctx, octx, beh, mods = make_bfv(N, k, ph.MUL_TECH_HPS)
    qs = mods[:k]
    Q = 1
    for q in qs:
        Q *= q
    delta = Q // T
    rng = np.random.default_rng(8)
    s = rng.integers(-1, 2, N)
    res = lambda v, ms: np.stack([np.array([int(x) % q for x in v], np.uint64) for q in ms])
    s_all = res(s, mods)
    m = rng.integers(0, T, N)
    a = np.stack([rng.integers(0, q, N, dtype=np.uint64) for q in qs])
    e = res(rng.integers(-4, 5, N), qs)
    dm = np.stack([np.array([(delta % q) * int(v) % q for v in m], np.uint64) for q in qs])
    as_ = poly_mul_rns(octx, a, s_all[:k], k)
    c0 = np.stack([(dm[i].astype(object) + e[i].astype(object) - as_[i].astype(object)) % qs[i]
                   for i in range(k)]).astype(np.uint64)
    g = ph.galois_elt_from_step(3, N)
    s_ntt = octx.ntt_forward(s_all, 0)
    sg_ntt = octx.ntt_forward(coeff_automorphism(s_all, g, mods), 0)
    dnum = ctx.dnum
    ka = np.stack([np.stack([rng.integers(0, q, N, dtype=np.uint64) for q in mods]) for _ in range(dnum)])
    ke = np.stack([res(rng.integers(-4, 5, N), mods) for _ in range(dnum)])
    gk = ph.galois_key(ctx, [g], host=octx.make_kswitch_key(s_ntt, sg_ntt, ka, ke)[None])
    c = upload_bfv(ctx, np.stack([c0, a]))
    ph.rotate_inplace(ctx, c, 3, gk)
    out = c.download(N)
    v = (out[0].astype(object) + poly_mul_rns(octx, out[1], s_all[:k], k).astype(object))
    v = v % np.array(qs, dtype=object)[:, None]
    dec = [((crt([int(v[i, n]) for i in range(k)], qs)[0] * T + Q // 2) // Q) % T for n in range(N)]
    want = coeff_automorphism(np.array([m], np.uint64), g, [T])[0]
    assert [int(x) for x in want] == dec
