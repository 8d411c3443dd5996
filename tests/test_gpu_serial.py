"""Serialization in the reference's stream layouts (PhantomCiphertext::save,
PhantomRelinKey::save, PhantomGaloisKey::save): the header fields parse as the
reference writes them, and a loaded ciphertext / key gives the same results."""
import struct

import numpy as np
import pytest

import oracle_bind as ob
import pyphantom as ph
from gpu_helpers import make_ctx, rand_ct, rand_key, say, upload_ct

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def small():
    ob.set_threads(8)
    ctx, octx, mods = make_ctx(1 << 12, [60, 50, 50, 50, 60, 60], 2)
    return ctx, octx, mods


def test_ciphertext_stream_layout_and_roundtrip(small):
    ctx, octx, mods = small
    N = ctx.N
    say("case: ciphertext save/load")
    rng = np.random.default_rng(1)
    a = rand_ct(rng, mods[:3], N)
    c = upload_ct(ctx, a, 2)
    c.set_scale(2.0 ** 45)
    blob = c.save()
    assert len(blob) == 58 + a.size * 8
    chain, size, n, L, scale, cf, nsd, ntt, asym = struct.unpack("<QQQQdQQ??", blob[:58])
    assert (chain, size, n, L, scale, cf, nsd, ntt, asym) == (2, 2, N, 3, 2.0 ** 45, 1, 1, True, False)
    assert np.array_equal(np.frombuffer(blob[58:], np.uint64).reshape(a.shape), a)
    d = ph.ciphertext.load(ctx, blob)
    assert d.chain_index() == 2 and d.scale() == 2.0 ** 45 and d.is_ntt_form()
    assert np.array_equal(d.download(N), a)
    with pytest.raises(ValueError):
        ph.ciphertext.load(ctx, blob[:100])           # truncated
    bad = bytearray(blob)
    bad[16:24] = struct.pack("<Q", 2 * N)             # wrong ring degree
    with pytest.raises(ValueError):
        ph.ciphertext.load(ctx, bytes(bad))


def test_keys_roundtrip_same_results(small):
    ctx, octx, mods = small
    N = ctx.N
    say("case: key save/load")
    rng = np.random.default_rng(2)
    key = rand_key(rng, mods, N, ctx.dnum)
    rk = ph.relin_key(ctx, host=key)
    blob = rk.save()
    per = 58 + 2 * len(mods) * N * 8
    assert len(blob) == 8 + ctx.dnum * per
    assert struct.unpack("<Q", blob[:8])[0] == ctx.dnum
    hdr = struct.unpack("<QQQQdQQ??", blob[8:66])
    assert hdr[:4] == (0, 2, N, len(mods))
    rk2 = ph.relin_key.load(ctx, blob)
    assert np.array_equal(rk2.download(), key)
    a, b = rand_ct(rng, mods[:4], N), rand_ct(rng, mods[:4], N)
    c1, c2 = upload_ct(ctx, a, 1), upload_ct(ctx, b, 1)
    ph.multiply_and_relin_inplace(ctx, c1, c2, rk2)
    assert np.array_equal(c1.download(N), octx.multiply_and_relin(4, a, b, key))
    elts = [ph.galois_elt_from_step(s, N) for s in (1, -2)]
    keys = np.stack([rand_key(rng, mods, N, ctx.dnum) for _ in elts])
    gk = ph.galois_key(ctx, elts, host=keys)
    gblob = gk.save()
    assert struct.unpack("<Q", gblob[:8])[0] == 2
    gk2 = ph.galois_key.load(ctx, elts, gblob)
    x = rand_ct(rng, mods[:4], N)
    c = upload_ct(ctx, x, 1)
    ph.rotate_inplace(ctx, c, -2, gk2)
    assert np.array_equal(c.download(N), octx.apply_galois(4, x, elts[1], keys[1]))
    with pytest.raises(ValueError):
        ph.galois_key.load(ctx, elts[:1], gblob)       # count mismatch


# ---- seeded symmetric format (save_symmetric / load_symmetric) ----

def ntt_prime_above(lo, N):
    from bigint_ref import is_prime
    q = lo - lo % (2 * N) + 1
    while q <= lo or not is_prime(q):
        q += 2 * N
    return q


def seeded_ctx(N):
    """two data primes (one just above 2^59: the sampler redraws ~1/32 of its
    words) and one special prime"""
    q0 = ntt_prime_above(1 << 59, N)
    rest = ph.coeff_modulus_create(N, [50, 60])
    mods = [q0] + rest
    parms = ph.EncryptionParameters(ph.SCHEME_CKKS)
    parms.set_poly_modulus_degree(N)
    parms.set_coeff_modulus(mods)
    parms.set_special_modulus_size(1)
    return ph.context(parms), ob.Ctx(N.bit_length() - 1, mods, 1), mods


@pytest.mark.parametrize("logN", [12, 16])
def test_seeded_sampler_matches_oracle(logN):
    N = 1 << logN
    say(f"case: seeded sampler N=2^{logN}")
    ctx, octx, mods = seeded_ctx(N)
    seed = bytes((5 * i + logN) % 256 for i in range(64))
    for L in (1, 2, 3):
        buf = ph.DeviceBuffer(L * N)
        ph.sample_uniform_poly(ctx, seed, L, buf)
        ph.device_synchronize()
        assert np.array_equal(buf.to_array((L, N)), ob.sample_uniform_seeded(octx, seed, L)), L


def test_save_load_symmetric_roundtrip():
    N = 1 << 13
    say("case: save/load symmetric")
    ctx, octx, mods = seeded_ctx(N)
    sk = ph.secret_key(ctx, seed=21)
    rng = np.random.default_rng(3)
    Ql = ctx.size_Q
    plain = np.stack([rng.integers(0, 1000, N, dtype=np.uint64) for _ in range(Ql)])
    seed = bytes(rng.integers(0, 256, 64, dtype=np.uint8))
    ct = sk.encrypt_symmetric_seeded(ctx, plain, 1, seed)
    full = ct.download(N)
    # c1 is the sampler's output (NTT form, as the reference samples it)
    assert np.array_equal(full[1], ob.sample_uniform_seeded(octx, seed, Ql))
    # decryption (coefficient form) recovers the plaintext up to the encryption noise
    dec = sk.decrypt(ctx, ct)
    for i, q in enumerate(mods[:Ql]):
        d = (dec[i].astype(object) - plain[i].astype(object)) % q
        d = np.where(d > q // 2, d - q, d)
        assert int(np.max(np.abs(d.astype(np.int64)))) < 64
    blob = ct.save_symmetric(ctx)
    assert len(blob) == 58 + Ql * N * 8 + 64
    chain, size, n, L, scale, cf, nsd, ntt, asym = struct.unpack("<QQQQdQQ??", blob[:58])
    assert (chain, size, n, L, ntt, asym) == (1, 2, N, Ql, True, False)
    assert np.array_equal(np.frombuffer(blob[58:58 + Ql * N * 8], np.uint64).reshape(Ql, N), full[0])
    assert blob[-64:] == seed
    d = ph.ciphertext.load_symmetric(ctx, blob)
    assert d.is_ntt_form() and d.chain_index() == 1
    assert np.array_equal(d.download(N), full)
    # a loaded symmetric ciphertext can be saved again, to the same bytes
    assert d.save_symmetric(ctx) == blob


def test_load_symmetric_coefficient_form():
    """a coefficient-form blob (BFV-style): c1 is the inverse transform of the sample"""
    N = 1 << 12
    say("case: load symmetric, coefficient form")
    ctx, octx, mods = seeded_ctx(N)
    Ql = ctx.size_Q
    rng = np.random.default_rng(4)
    c0 = np.stack([rng.integers(0, q, N, dtype=np.uint64) for q in mods[:Ql]])
    seed = bytes(range(100, 164))
    blob = struct.pack("<QQQQdQQ??", 1, 2, N, Ql, 1.0, 1, 1, False, False) + c0.tobytes() + seed
    d = ph.ciphertext.load_symmetric(ctx, blob)
    assert not d.is_ntt_form()
    got = d.download(N)
    assert np.array_equal(got[0], c0)
    assert np.array_equal(got[1], octx.ntt_inverse(ob.sample_uniform_seeded(octx, seed, Ql), 0))


def test_symmetric_format_rejects():
    N = 1 << 12
    ctx, octx, mods = seeded_ctx(N)
    sk = ph.secret_key(ctx, seed=22)
    Ql = ctx.size_Q
    plain = np.zeros((Ql, N), np.uint64)
    ct = sk.encrypt_symmetric_seeded(ctx, plain, 1, bytes(64))
    blob = ct.save_symmetric(ctx)
    # a modified ciphertext no longer matches its seed
    other = sk.encrypt_symmetric_seeded(ctx, plain, 1, bytes(range(64)))
    ct.upload(other.download(N))
    with pytest.raises(ValueError):
        ct.save_symmetric(ctx)
    # an unseeded ciphertext has no seed
    plain_ct = sk.encrypt_symmetric(ctx, plain, 1)
    with pytest.raises(ValueError):
        plain_ct.save_symmetric(ctx)
    # asymmetric flag, wrong size, wrong limb count, truncated: the reference's runtime errors
    asym = bytearray(blob)
    asym[57] = 1
    with pytest.raises(ph.PhxError):
        ph.ciphertext.load_symmetric(ctx, bytes(asym))
    size3 = bytearray(blob)
    size3[8:16] = (3).to_bytes(8, "little")
    with pytest.raises(ph.PhxError):
        ph.ciphertext.load_symmetric(ctx, bytes(size3))
    fewer = bytearray(blob)
    fewer[24:32] = (Ql - 1).to_bytes(8, "little")
    with pytest.raises(ph.PhxError):
        ph.ciphertext.load_symmetric(ctx, bytes(fewer))
    with pytest.raises(ValueError):
        ph.ciphertext.load_symmetric(ctx, blob[:-1])


def test_secret_key_roundtrip(small):
    """PhantomSecretKey::save layout: max power, N, QP, then s (and s^2 once a
    3-polynomial ciphertext was decrypted) in NTT form; a loaded key decrypts
    like the original"""
    ctx, octx, mods = small
    N = ctx.N
    say("case: secret key save/load")
    sk = ph.secret_key(ctx, seed=31)
    blob = sk.save()
    assert len(blob) == 24 + ctx.size_QP * N * 8
    assert struct.unpack("<QQQ", blob[:24]) == (1, N, ctx.size_QP)
    assert np.array_equal(np.frombuffer(blob[24:], np.uint64).reshape(ctx.size_QP, N), sk.ntt_form())
    sk2 = ph.secret_key.load(ctx, blob)
    assert np.array_equal(sk2.ntt_form(), sk.ntt_form())
    rng = np.random.default_rng(5)
    a = rand_ct(rng, mods[:ctx.size_Q], N, size=3)
    c = upload_ct(ctx, a, 1)
    want = sk.decrypt(ctx, c)            # computes s^2
    assert np.array_equal(sk2.decrypt(ctx, c), want)
    blob2 = sk.save()
    assert len(blob2) == 24 + 2 * ctx.size_QP * N * 8 and struct.unpack("<Q", blob2[:8])[0] == 2
    s = sk.ntt_form().astype(object)
    s2 = np.array([[(int(x) * int(x)) % mods[i] for x in s[i]] for i in range(ctx.size_QP)], dtype=object)
    assert np.array_equal(np.frombuffer(blob2[24 + ctx.size_QP * N * 8:], np.uint64).reshape(ctx.size_QP, N),
                          s2.astype(np.uint64))
    sk3 = ph.secret_key.load(ctx, blob2)
    assert np.array_equal(sk3.decrypt(ctx, c), want)
    assert sk3.save() == blob2
    with pytest.raises(ValueError):
        ph.secret_key.load(ctx, blob[:-8])                      # truncated
    bad = bytearray(blob)
    bad[8:16] = struct.pack("<Q", 2 * N)
    with pytest.raises(ValueError):
        ph.secret_key.load(ctx, bytes(bad))                     # other ring degree
