"""BFV decryption on the GPU (PhantomSecretKey::decrypt of a BFV ciphertext,
src/secretkey.cu:571-635): bit-exact against the oracle's scale-and-round of the
same phase for BEHZ and the four HPS forms, on uniform ciphertexts (sizes 2 and
3), and semantically: an encryption of m decrypts to m, and
multiply_and_relin of encryptions of m1, m2 decrypts to m1 m2 mod (X^N + 1, t)."""
import numpy as np
import pytest

import oracle_bind as ob
import pyphantom as ph
from bigint_ref import primes_below
from gpu_helpers import say

pytestmark = pytest.mark.gpu

N = 4096
T = 65537

# (name, data prime bits, data primes, t bits or None for 65537, mul_tech)
CASES = [
    ("behz", 54, 3, None, ph.MUL_TECH_BEHZ),
    ("hps_small_lazy", 30, 3, None, ph.MUL_TECH_HPS),
    ("hps_small", 40, 3, None, ph.MUL_TECH_HPS),
    ("hps_large_lazy", 54, 4, None, ph.MUL_TECH_HPS_OVERQ),
    ("hps_large", 54, 4, 30, ph.MUL_TECH_HPS),
]


def make(bits, k, tbits, tech):
    data = primes_below(bits, N, k)
    special = [p for p in primes_below(60, N, 2) if p not in data][:1]
    mods = data + special
    t = T if tbits is None else primes_below(tbits, 1, 1)[0]
    parms = ph.EncryptionParameters(ph.SCHEME_BFV)
    parms.set_poly_modulus_degree(N)
    parms.set_coeff_modulus(mods)
    parms.set_special_modulus_size(1)
    parms.set_plain_modulus(t)
    parms.set_mul_tech(tech)
    return ph.context(parms), ob.Ctx(N.bit_length() - 1, mods, 1), data, t


def mulmod(a, b, moduli):
    qs = np.array(moduli, dtype=object)[:, None]
    return ((a.astype(object) * b.astype(object)) % qs).astype(np.uint64)


def phase(octx, data, c, s_ntt):
    """c0 + c1 s (+ c2 s^2), coefficient form, through the oracle's NTT"""
    k = len(data)
    acc = c[0].astype(object)
    s_pow = s_ntt[:k]
    for j in range(1, c.shape[0]):
        acc = acc + octx.ntt_inverse(mulmod(octx.ntt_forward(c[j], 0), s_pow, data), 0).astype(object)
        s_pow = mulmod(s_pow, s_ntt[:k], data)
    return (acc % np.array(data, dtype=object)[:, None]).astype(np.uint64)


@pytest.mark.parametrize("name,bits,k,tbits,tech", CASES)
def test_bfv_decrypt_bit_exact(name, bits, k, tbits, tech):
    say(f"case: bfv decrypt {name}")
    ob.set_threads(8)
    ctx, octx, data, t = make(bits, k, tbits, tech)
    sk = ph.secret_key(ctx, seed=41)
    s_ntt = sk.ntt_form()
    rng = np.random.default_rng(bits + 3 * k)
    for size in (2, 3):
        c = np.stack([np.stack([rng.integers(0, q, N, dtype=np.uint64) for q in data]) for _ in range(size)])
        ct = ph.ciphertext(ctx, 1, size)
        assert not ct.is_ntt_form()
        ct.upload(c)
        got = sk.decrypt_bfv(ctx, ct)
        want = ob.bfv_decrypt(octx, t, phase(octx, data, c, s_ntt), tech == ph.MUL_TECH_BEHZ)
        assert np.array_equal(got, want), (name, size)


def negacyclic_mod_t(a, b):
    octx_t = ob.Ctx(N.bit_length() - 1, [T, primes_below(60, N, 1)[0]], 1)
    fa = octx_t.ntt_forward(np.array(a, np.uint64)[None], 0)
    fb = octx_t.ntt_forward(np.array(b, np.uint64)[None], 0)
    return octx_t.ntt_inverse((fa.astype(object) * fb.astype(object) % T).astype(np.uint64), 0)[0]


@pytest.mark.parametrize("tech", [ph.MUL_TECH_BEHZ, ph.MUL_TECH_HPS, ph.MUL_TECH_HPS_OVERQ])
def test_bfv_encrypt_multiply_decrypt(tech):
    say(f"case: bfv encrypt/multiply/decrypt tech={tech}")
    ctx, octx, data, t = make(54, 4, None, tech)
    sk = ph.secret_key(ctx, seed=42)
    pk = sk.gen_publickey(ctx)
    rk = sk.gen_relinkey(ctx)
    rng = np.random.default_rng(tech)
    m1 = rng.integers(0, T, N).astype(np.uint64)
    m2 = rng.integers(0, T, N).astype(np.uint64)
    c1 = pk.encrypt_asymmetric(ctx, m1, 1, seed=3)
    c2 = pk.encrypt_asymmetric(ctx, m2, 1, seed=4)
    assert np.array_equal(sk.decrypt_bfv(ctx, c1) % T, m1)
    ph.multiply_and_relin_inplace(ctx, c1, c2, rk)
    assert np.array_equal(sk.decrypt_bfv(ctx, c1) % T, negacyclic_mod_t(m1, m2))


def test_bfv_decrypt_rejects():
    ctx, octx, data, t = make(54, 3, None, ph.MUL_TECH_BEHZ)
    sk = ph.secret_key(ctx, seed=43)
    ct = ph.ciphertext(ctx, 1, 2)
    ct.set_ntt_form(True)
    with pytest.raises(ValueError):
        sk.decrypt_bfv(ctx, ct)
    # a CKKS context has no plaintext modulus to decrypt to
    parms = ph.EncryptionParameters(ph.SCHEME_CKKS)
    parms.set_poly_modulus_degree(N)
    parms.set_coeff_modulus(ph.coeff_modulus_create(N, [50, 50, 60]))
    parms.set_special_modulus_size(1)
    cctx = ph.context(parms)
    csk = ph.secret_key(cctx, seed=44)
    cct = ph.ciphertext(cctx, 1, 2)
    with pytest.raises(ValueError):
        csk.decrypt_bfv(cctx, cct)


def test_bfv_mod_switch_floor_divides():
    """BFV mod_switch_to_next: each coefficient's CRT integer X becomes
    floor(X / q_last) in the remaining primes (the reference's divide_and_round_q_last
    kernel floors); mod_switch_to_inplace two levels down is floor(X / (q_k q_{k-1}))"""
    from bigint_ref import crt
    say("case: bfv mod switch")
    ctx, octx, data, t = make(54, 4, None, ph.MUL_TECH_BEHZ)
    rng = np.random.default_rng(11)
    c = np.stack([np.stack([rng.integers(0, q, N, dtype=np.uint64) for q in data]) for _ in range(2)])
    ct = ph.ciphertext(ctx, 1, 2)
    ct.upload(c)
    d = ph.mod_switch_to_next(ctx, ct)
    assert d.chain_index() == 2 and not d.is_ntt_form()
    got = d.download(N)
    ph.mod_switch_to_inplace(ctx, ct, 3)
    got2 = ct.download(N)
    for p in range(2):
        for n in range(0, N, 7):
            X = crt([int(c[p, i, n]) for i in range(4)], data)[0]
            assert [int(v) for v in got[p, :, n]] == [(X // data[3]) % q for q in data[:3]], (p, n)
            assert [int(v) for v in got2[p, :, n]] == [(X // (data[3] * data[2])) % q for q in data[:2]], (p, n)


@pytest.mark.parametrize("tech", [ph.MUL_TECH_BEHZ, ph.MUL_TECH_HPS])
def test_bfv_decrypt_below_the_top_level(tech):
    """decryption at chain 2 and 3 (the level's own constants; HPS's qMSB still over
    all of Q): bit-exact against the oracle on uniform ciphertexts, and an
    encryption switched down by mod_switch_to_next still decrypts to m"""
    say(f"case: bfv decrypt below the top level tech={tech}")
    ctx, octx, data, t = make(54, 4, None, tech)
    sk = ph.secret_key(ctx, seed=45)
    s_ntt = sk.ntt_form()
    rng = np.random.default_rng(tech + 20)
    for chain in (2, 3):
        kl = 5 - chain
        c = np.stack([np.stack([rng.integers(0, q, N, dtype=np.uint64) for q in data[:kl]]) for _ in range(2)])
        ct = ph.ciphertext(ctx, chain, 2)
        ct.upload(c)
        want = ob.bfv_decrypt(octx, t, phase(octx, data[:kl], c, s_ntt), tech == ph.MUL_TECH_BEHZ)
        assert np.array_equal(sk.decrypt_bfv(ctx, ct), want), chain
    pk = sk.gen_publickey(ctx)
    m = rng.integers(0, T, N).astype(np.uint64)
    e = pk.encrypt_asymmetric(ctx, m, 1, seed=8)
    ph.mod_switch_to_inplace(ctx, e, 3)
    assert np.array_equal(sk.decrypt_bfv(ctx, e) % T, m)
