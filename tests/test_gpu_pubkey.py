"""Public keys and asymmetric encryption (PhantomSecretKey::gen_publickey,
PhantomPublicKey::encrypt_asymmetric, src/secretkey.cu:11-186, 382-394).

With explicit samples u, e_0, e_1 the ciphertext is bit-exact against the
oracle's composition of the reference's steps: c_j = moddown_{QP->Q}(pk_j NTT(u)
+ NTT(e_j)) (DRNSTool::moddown, src/rns_bconv.cu:712-756), then CKKS c_0 +=
NTT(plain) and BFV c_0 += [[-Q]_t m]_t t^-1 mod q_i (bfv_add_timesQ_overt_kernel,
src/polymath.cu:413-436) on the coefficient form. The public key itself is
checked through its defining relation pk_0 + pk_1 s = -e with one small e
across every QP limb, and each ciphertext decrypts to its plaintext."""
import struct

import numpy as np
import pytest

import oracle_bind as ob
import pyphantom as ph
from bigint_ref import crt
from gpu_helpers import make_ctx, say

pytestmark = pytest.mark.gpu

T = 65537


def residues(v, moduli):
    v = np.asarray(v).astype(object)
    return np.stack([(v % q).astype(np.uint64) for q in moduli])


def centered(x, q):
    x = np.asarray(x, dtype=object) % q
    return np.where(x > q // 2, x - q, x)


def mulmod(a, b, moduli):
    qs = np.array(moduli, dtype=object)[:, None]
    return ((a.astype(object) * b.astype(object)) % qs).astype(np.uint64)


def addmod(a, b, moduli):
    qs = np.array(moduli, dtype=object)[:, None]
    return ((a.astype(object) + b.astype(object)) % qs).astype(np.uint64)


def check_public_key(ctx, octx, mods, pk, s_ntt):
    """pk_0 + pk_1 s = -e: the same small integer polynomial in every QP limb"""
    v = octx.ntt_inverse(addmod(pk[0], mulmod(pk[1], s_ntt, mods), mods), 0)
    e0 = centered(v[0], mods[0])
    assert 0 < int(np.max(np.abs(e0.astype(np.int64)))) < 64
    for i, q in enumerate(mods):
        assert np.array_equal(centered(v[i], q), e0), i


def zero_encryption(octx, mods, size_Q, pk, u, e):
    """moddown_{QP->Q}(pk_j NTT(u) + NTT(e_j)), NTT form"""
    u_ntt = octx.ntt_forward(residues(u, mods), 0)
    out = []
    for j in range(2):
        x = addmod(mulmod(pk[j], u_ntt, mods), octx.ntt_forward(residues(e[j], mods), 0), mods)
        out.append(octx.moddown(size_Q, x))
    return np.stack(out)


def samples(rng, N):
    u = rng.integers(-1, 2, N).astype(np.int8)
    e = [rng.integers(-20, 21, N).astype(np.int8) for _ in range(2)]
    return u, e


@pytest.fixture(scope="module")
def ckks():
    ob.set_threads(8)
    return make_ctx(1 << 12, [60, 50, 50, 50, 60, 60], 2)


def test_ckks_public_key_and_asymmetric_bit_exact(ckks):
    ctx, octx, mods = ckks
    N, Q = ctx.N, ctx.size_Q
    say("case: ckks public key + asymmetric encryption N=2^12")
    sk = ph.secret_key(ctx, seed=31)
    s_ntt = sk.ntt_form()
    seed_a = bytes((7 * i + 3) % 256 for i in range(64))
    pk = sk.gen_publickey(ctx, seed_a=seed_a)
    P = pk.data()
    assert np.array_equal(P[1], ob.sample_uniform_seeded(octx, seed_a, len(mods)))
    check_public_key(ctx, octx, mods, P, s_ntt)
    rng = np.random.default_rng(11)
    m = rng.integers(-(1 << 30), 1 << 30, N)
    plain = residues(m, mods[:Q])
    u, e = samples(rng, N)
    ct = pk.encrypt_asymmetric(ctx, plain, 1, u=u, e=e)
    assert ct.chain_index() == 1 and ct.is_ntt_form() and ct.size() == 2
    want = zero_encryption(octx, mods, Q, P, u, e)
    want[0] = addmod(want[0], octx.ntt_forward(plain, 0), mods[:Q])
    assert np.array_equal(ct.download(N), want)
    # decryption: c0 + c1 s = m + (small), in coefficient form
    dec = sk.decrypt(ctx, ct)
    for i, q in enumerate(mods[:Q]):
        d = centered(dec[i].astype(object) - plain[i].astype(object), q)
        assert int(np.max(np.abs(d.astype(np.int64)))) < 4 * N


def test_ckks_asymmetric_sampled(ckks):
    """the library's own samplers: deterministic per seed, fresh per seed, decrypts"""
    ctx, octx, mods = ckks
    N, Q = ctx.N, ctx.size_Q
    say("case: ckks asymmetric encryption, sampled")
    sk = ph.secret_key(ctx, seed=32)
    pk = sk.gen_publickey(ctx)
    check_public_key(ctx, octx, mods, pk.data(), sk.ntt_form())
    rng = np.random.default_rng(12)
    plain = residues(rng.integers(-(1 << 20), 1 << 20, N), mods[:Q])
    a = pk.encrypt_asymmetric(ctx, plain, 1, seed=9).download(N)
    b = pk.encrypt_asymmetric(ctx, plain, 1, seed=9).download(N)
    c = pk.encrypt_asymmetric(ctx, plain, 1, seed=10).download(N)
    assert np.array_equal(a, b)
    assert not np.array_equal(a[1], c[1])
    ct = ph.ciphertext(ctx, 1, 2)
    ct.upload(c)
    dec = sk.decrypt(ctx, ct)
    for i, q in enumerate(mods[:Q]):
        d = centered(dec[i].astype(object) - plain[i].astype(object), q)
        assert int(np.max(np.abs(d.astype(np.int64)))) < 4 * N


def test_public_key_save_load(ckks):
    ctx, octx, mods = ckks
    N = ctx.N
    say("case: public key save/load")
    sk = ph.secret_key(ctx, seed=33)
    pk = sk.gen_publickey(ctx, seed_e=17)
    blob = pk.save()
    assert len(blob) == 58 + 2 * len(mods) * N * 8
    chain, size, n, L, scale, cf, nsd, ntt, asym = struct.unpack("<QQQQdQQ??", blob[:58])
    assert (chain, size, n, L, ntt) == (0, 2, N, len(mods), True)
    assert np.array_equal(np.frombuffer(blob[58:], np.uint64).reshape(2, len(mods), N), pk.data())
    pk2 = ph.public_key.load(ctx, blob)
    assert np.array_equal(pk2.data(), pk.data())
    rng = np.random.default_rng(13)
    plain = residues(rng.integers(-100, 100, N), mods[:ctx.size_Q])
    u, e = samples(rng, N)
    assert np.array_equal(pk.encrypt_asymmetric(ctx, plain, 1, u=u, e=e).download(N),
                          pk2.encrypt_asymmetric(ctx, plain, 1, u=u, e=e).download(N))
    # a data-level ciphertext is not a public key
    ct = sk.encrypt_symmetric(ctx, plain, 1)
    with pytest.raises((ValueError, ph.PhxError)):
        ph.public_key.load(ctx, ct.save())
    with pytest.raises((ValueError, ph.PhxError)):
        ph.public_key.load(ctx, blob[:-8])


def test_asymmetric_rejects(ckks):
    ctx, octx, mods = ckks
    N, Q = ctx.N, ctx.size_Q
    sk = ph.secret_key(ctx, seed=34)
    pk = sk.gen_publickey(ctx)
    plain = np.zeros((Q, N), np.uint64)
    # the reference switches modulus from the key level to the first data level only
    with pytest.raises((ValueError, ph.PhxError)):
        pk.encrypt_asymmetric(ctx, plain, 2)


def make_bfv(N, k):
    mods = ph.coeff_modulus_create(N, [54] * k + [60])
    parms = ph.EncryptionParameters(ph.SCHEME_BFV)
    parms.set_poly_modulus_degree(N)
    parms.set_coeff_modulus(mods)
    parms.set_special_modulus_size(1)
    parms.set_plain_modulus(T)
    return ph.context(parms), ob.Ctx(N.bit_length() - 1, mods, 1), mods


@pytest.mark.parametrize("N,k", [(1 << 12, 3), (1 << 14, 6)])
def test_bfv_asymmetric_bit_exact_and_decrypts(N, k):
    say(f"case: bfv asymmetric encryption N={N} k={k}")
    ob.set_threads(8)
    ctx, octx, mods = make_bfv(N, k)
    qs = mods[:k]
    sk = ph.secret_key(ctx, seed=35)
    s_ntt = sk.ntt_form()
    pk = sk.gen_publickey(ctx, seed_a=bytes(range(64)))
    P = pk.data()
    check_public_key(ctx, octx, mods, P, s_ntt)
    rng = np.random.default_rng(N + k)
    m = rng.integers(0, T, N).astype(np.uint64)
    u, e = samples(rng, N)
    ct = pk.encrypt_asymmetric(ctx, m, 1, u=u, e=e)
    assert ct.chain_index() == 1 and not ct.is_ntt_form()
    want = zero_encryption(octx, mods, k, P, u, e)
    want = np.stack([octx.ntt_inverse(w, 0) for w in want])
    Qv = 1
    for q in qs:
        Qv *= q
    mq = (((-Qv) % T) * m.astype(object)) % T
    add = np.stack([(mq * pow(T, -1, q) % q).astype(np.uint64) for q in qs])
    want[0] = addmod(want[0], add, qs)
    got = ct.download(N)
    assert np.array_equal(got, want)

    def bfv_decrypt(c):
        """round(t (c0 + c1 s) / Q) mod t on the coefficient form"""
        v = addmod(c[0], octx.ntt_inverse(mulmod(octx.ntt_forward(c[1], 0), s_ntt[:k], qs), 0), qs)
        return np.array([((crt([int(v[i, n]) for i in range(k)], qs)[0] * T + Qv // 2) // Qv) % T
                         for n in range(N)], dtype=np.uint64)

    assert np.array_equal(bfv_decrypt(got), m)
    # the sampled form decrypts too
    assert np.array_equal(bfv_decrypt(pk.encrypt_asymmetric(ctx, m, 1, seed=5).download(N)), m)
    bad = m.copy()
    bad[3] = T
    with pytest.raises((ValueError, ph.PhxError)):
        pk.encrypt_asymmetric(ctx, bad, 1)
