"""Edge cases of the evaluator entry points on the GPU:

- rotate_inplace without a key for the step falls back to the non-adjacent form
  of the step (rotate_internal, src/evaluate.cu:1633-1668): rotate by 3 with keys
  only for +-1, 2, 4 runs rotate(4) then rotate(-1), each an apply_galois
  (src/evaluate.cu:1567-1630); checked against the oracle's apply_galois chain;
- apply_galois, rotate and hoisting refuse a ciphertext of size 1 before any
  kernel runs (the reference's check is `encrypted.size() != 2`,
  src/evaluate.cu:1571-1573);
- decryption of a size-1 ciphertext is c0 alone (needed_sk_power 0 in
  ckks_decrypt, src/secretkey.cu:533-569);
- BFV hps_overq_leveled: a key switch that drops levels is defined at the first
  data level only (src/eval_key_switch.cu:110-150 reads size_Q limbs of c2), so a
  ciphertext switched below it is refused by relinearize, apply_galois, rotate and
  hoisting, with nothing launched."""
import struct

import numpy as np
import pytest

import oracle_bind as ob
import pyphantom as ph
from gpu_helpers import make_ctx, rand_ct, rand_key, say, upload_ct

pytestmark = pytest.mark.gpu

N = 1 << 12


@pytest.fixture(scope="module")
def ckks():
    ob.set_threads(8)
    return make_ctx(N, [60, 50, 50, 50, 50, 60, 60], 2)


def test_rotate_naf_fallback_bit_exact(ckks):
    ctx, octx, mods = ckks
    say("case: rotate by 3 through the NAF fallback (keys for +-1, 2, 4)")
    Ql = ctx.size_Q
    rng = np.random.default_rng(31)
    key_steps = [1, -1, 2, 4]
    elts = [ph.galois_elt_from_step(s, N) for s in key_steps]
    keys = np.stack([rand_key(rng, mods, N, ctx.dnum) for _ in elts])
    gk = ph.galois_key(ctx, elts, host=keys)
    a = rand_ct(rng, mods[:Ql], N)
    c = upload_ct(ctx, a, 1)
    assert ph.galois_elt_from_step(3, N) not in elts
    ph.rotate_inplace(ctx, c, 3, gk)
    # naf(3) = [-1, 4] in the reference's order of increasing magnitude
    want = a
    for st in (-1, 4):
        i = key_steps.index(st)
        want = octx.apply_galois(Ql, want, elts[i], keys[i])
    assert np.array_equal(c.download(N), want)
    # a step whose NAF is a single term (no key for it) is refused
    c8 = upload_ct(ctx, a, 1)
    with pytest.raises(ValueError):
        ph.rotate_inplace(ctx, c8, 8, gk)
    assert np.array_equal(c8.download(N), a)


def test_size1_refused_by_galois_paths(ckks):
    ctx, octx, mods = ckks
    say("case: size-1 ciphertexts refused by apply_galois / rotate / hoisting")
    Ql = ctx.size_Q
    rng = np.random.default_rng(32)
    elt = ph.galois_elt_from_step(1, N)
    gk = ph.galois_key(ctx, [elt], host=np.stack([rand_key(rng, mods, N, ctx.dnum)]))
    a = rand_ct(rng, mods[:Ql], N, size=1)
    c = upload_ct(ctx, a, 1)
    with pytest.raises(ValueError):
        ph.apply_galois_inplace(ctx, c, elt, gk)
    with pytest.raises(ValueError):
        ph.rotate_inplace(ctx, c, 1, gk)
    with pytest.raises(ValueError):
        ph.hoisting_inplace(ctx, c, gk, [1])
    assert np.array_equal(c.download(N), a)


def test_size1_decrypt_is_c0(ckks):
    ctx, octx, mods = ckks
    say("case: decrypt of a size-1 ciphertext")
    Ql = ctx.size_Q
    rng = np.random.default_rng(33)
    sk = ph.secret_key(ctx, seed=5)
    a = rand_ct(rng, mods[:Ql], N, size=1)
    c = upload_ct(ctx, a, 1)
    got = sk.decrypt(ctx, c)
    assert np.array_equal(got, octx.ntt_inverse(a[0], 0))


T = 65537
NB = 1 << 13
KB = 4


@pytest.fixture(scope="module")
def leveled():
    mods = ph.coeff_modulus_create(NB, [54] * KB + [60])
    parms = ph.EncryptionParameters(ph.SCHEME_BFV)
    parms.set_poly_modulus_degree(NB)
    parms.set_coeff_modulus(mods)
    parms.set_special_modulus_size(1)
    parms.set_plain_modulus(T)
    parms.set_mul_tech(ph.MUL_TECH_HPS_OVERQ_LEVELED)
    parms.set_galois_elts([ph.galois_elt_from_step(1, NB)])
    ctx = ph.context(parms)
    octx = ob.Ctx(13, mods, 1)
    return ctx, octx, mods


def with_degree(ctx, arr, deg):
    c = ph.ciphertext(ctx, 1, arr.shape[0])
    c.upload(arr)
    blob = bytearray(c.save())
    blob[48:56] = struct.pack("<Q", deg)
    return ph.ciphertext.load(ctx, bytes(blob))


def test_leveled_below_first_level_refused(leveled):
    ctx, octx, mods = leveled
    say("case: hps_overq_leveled key switch below chain 1 refused")
    deg = 7
    assert ob.hps_levels_to_drop(octx, T, deg - 1) > 0
    assert ob.hps_levels_to_drop(octx, T, deg - 1, key_switch=True) > 0
    rng = np.random.default_rng(34)
    elt = ph.galois_elt_from_step(1, NB)
    gk = ph.galois_key(ctx, [elt], host=np.stack([rand_key(rng, mods, NB, ctx.dnum)]))
    rk = ph.relin_key(ctx, host=rand_key(rng, mods, NB, ctx.dnum))
    c3 = with_degree(ctx, rand_ct(rng, mods[:KB], NB, size=3), deg)
    ph.mod_switch_to_next_inplace(ctx, c3)
    assert c3.chain_index() == 2
    before = c3.download(NB)
    with pytest.raises(ValueError):
        ph.relinearize_inplace(ctx, c3, rk)
    assert np.array_equal(c3.download(NB), before)
    c2 = with_degree(ctx, rand_ct(rng, mods[:KB], NB), deg)
    ph.mod_switch_to_next_inplace(ctx, c2)
    before = c2.download(NB)
    with pytest.raises(ValueError):
        ph.apply_galois_inplace(ctx, c2, elt, gk)
    with pytest.raises(ValueError):
        ph.rotate_inplace(ctx, c2, 1, gk)
    with pytest.raises(ValueError):
        ph.hoisting_inplace(ctx, c2, gk, [1])
    assert np.array_equal(c2.download(NB), before)
    # without dropped levels (a fresh ciphertext) the same calls run below chain 1
    f = with_degree(ctx, rand_ct(rng, mods[:KB], NB), 1)
    ph.mod_switch_to_next_inplace(ctx, f)
    ph.apply_galois_inplace(ctx, f, elt, gk)
    assert f.chain_index() == 2


@pytest.mark.parametrize("n", [2, 4096, 4096 + 6, 3 * 4096 + 2, 1 << 20])
def test_copy_kernel_sizes(n):
    """the streaming copy the bench calibrates its counters with: whole 32 KiB
    chunks and a partial last chunk copy every word and nothing beyond"""
    say(f"case: copy_u64 n={n}")
    rng = np.random.default_rng(n)
    src = rng.integers(0, 1 << 63, n, dtype=np.uint64)
    guard = np.full(8, 0x5A5A5A5A5A5A5A5A, dtype=np.uint64)
    a = ph.DeviceBuffer.from_array(src)
    b = ph.DeviceBuffer.from_array(np.concatenate([np.zeros(n, dtype=np.uint64), guard]))
    ph.copy_u64(b, a, n)
    ph.device_synchronize()
    out = b.to_array((n + 8,))
    assert np.array_equal(out[:n], src)
    assert np.array_equal(out[n:], guard)
