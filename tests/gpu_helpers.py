"""Shared helpers for the GPU parity tests."""
import sys

import numpy as np

import oracle_bind as ob
import pyphantom as ph


def say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


def make_ctx(N, bits, size_P, galois_elts=None, scheme=ph.SCHEME_CKKS):
    mods = ph.coeff_modulus_create(N, bits)
    parms = ph.EncryptionParameters(scheme)
    parms.set_poly_modulus_degree(N)
    parms.set_coeff_modulus(mods)
    parms.set_special_modulus_size(size_P)
    if galois_elts:
        parms.set_galois_elts(galois_elts)
    ctx = ph.context(parms)
    logN = N.bit_length() - 1
    octx = ob.Ctx(logN, mods, size_P)
    return ctx, octx, mods


def rand_limbs(rng, moduli, N):
    return np.stack([rng.integers(0, q, N, dtype=np.uint64) for q in moduli]).astype(np.uint64)


def rand_ct(rng, moduli, N, size=2):
    return np.stack([rand_limbs(rng, moduli, N) for _ in range(size)])


def rand_key(rng, moduli, N, dnum):
    return np.stack([np.stack([rand_limbs(rng, moduli, N) for _ in range(2)]) for _ in range(dnum)])


def upload_ct(ctx, arr, chain_index):
    ct = ph.ciphertext(ctx, chain_index, arr.shape[0])
    ct.upload(arr)
    return ct
