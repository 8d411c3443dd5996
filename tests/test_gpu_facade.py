"""The C++ facade (include/phantom.hpp) run as a PhantomFHE program would run it:
tests/facade/facade_test (built by build() with tests/facade/Makefile) calls
multiply_and_relin, multiply + relinearize, rescale_to_next, rotate_inplace,
apply_galois, hoisting, keyswitch_inplace and the 2-D NTT through the reference's
class names and signatures, on seeded residues at N = 2^12 (5 data limbs, 2
special, dnum 3); every output is compared here bit for bit with the oracle."""
import os
import subprocess

import numpy as np
import pytest

import oracle_bind as ob
from gpu_helpers import say

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
BIN = os.path.join(ROOT, "tests", "facade", "facade_test")


def load(d, name, shape):
    return np.fromfile(os.path.join(d, name), dtype=np.uint64).reshape(shape)


def add_mod(a, b, mods):
    m = np.array(mods[:a.shape[-2]], dtype=object)[:, None]
    return ((a.astype(object) + b.astype(object)) % m).astype(np.uint64)


def test_facade_program_bit_exact(tmp_path):
    assert os.path.exists(BIN), "tests/facade/facade_test is not built (run build())"
    say("case: C++ facade program, N=4096, 5+2 limbs")
    r = subprocess.run([BIN, str(tmp_path)], capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stderr + r.stdout
    assert "facade: ok" in r.stdout
    N, Q, P = 4096, 5, 2
    mods = [int(v) for v in np.fromfile(tmp_path / "moduli.bin", dtype=np.uint64)]
    QP, dnum = Q + P, 3
    elts = [int(v) for v in np.fromfile(tmp_path / "elts.bin", dtype=np.uint64)]
    d = str(tmp_path)
    ct1, ct2 = load(d, "ct1.bin", (2, Q, N)), load(d, "ct2.bin", (2, Q, N))
    rk = load(d, "rk.bin", (dnum, 2, QP, N))
    gk = load(d, "gk.bin", (len(elts), dnum, 2, QP, N))
    c2 = load(d, "c2.bin", (Q, N))
    octx = ob.Ctx(12, mods, P)
    mr = octx.multiply_and_relin(Q, ct1, ct2, rk)
    assert np.array_equal(load(d, "mr.bin", (2, Q, N)), mr)
    rs = octx.rescale(Q, mr)
    assert np.array_equal(load(d, "rs.bin", (2, Q - 1, N)), rs)
    assert np.array_equal(load(d, "rot.bin", (2, Q - 1, N)), octx.apply_galois(Q - 1, rs, elts[0], gk[0]))
    assert np.array_equal(load(d, "gal.bin", (2, Q, N)), octx.apply_galois(Q, ct1, elts[2], gk[2]))
    assert np.array_equal(load(d, "hoist.bin", (2, Q, N)), octx.hoisting(Q, ct1, elts[:2], gk[:2]))
    assert np.array_equal(load(d, "ks.bin", (2, Q, N)), add_mod(ct1, octx.keyswitch(Q, c2, rk), mods))
    assert np.array_equal(load(d, "ntt.bin", (Q, N)), octx.ntt_forward(np.ascontiguousarray(ct1[0]), 0))
