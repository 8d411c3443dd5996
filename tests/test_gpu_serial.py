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
