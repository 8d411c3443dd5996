"""FastBConv fused with the forward NTT's column pass (bconv_ntt.hip): modup,
moddown, multiply_and_relin, rotation and batch plans bit-exact against the
oracle with the fused launch on and off, at every ring degree it serves
(2^12 .. 2^16), over prime sets whose 64-bit-integer and FP64 outputs are not
multiples of four (padded groups)."""
import numpy as np
import pytest

import pyphantom as ph
from gpu_helpers import make_ctx, rand_ct, rand_key, rand_limbs, say, upload_ct

pytestmark = pytest.mark.gpu

# (N, bit sizes, special primes): 60-bit primes take the integer butterflies,
# 50-bit ones the FP64 butterflies; class sizes per conversion vary by digit
SHAPES = [
    (1 << 12, [60, 50, 50, 50, 50, 50, 60, 60], 2),
    (1 << 13, [60] + [50] * 6 + [60] * 3, 3),
    (1 << 14, [60, 60] + [50] * 8 + [60] * 5, 5),
    (1 << 15, [60] + [50] * 10 + [60] * 4, 4),
    (1 << 16, [60] + [50] * 12 + [60] * 6, 6),
]


def _with_fused(value, fn):
    old = ph.tune_get("bconv_fused")
    try:
        ph.tune_set("bconv_fused", value)
        return fn()
    finally:
        ph.tune_set("bconv_fused", old)


@pytest.mark.parametrize("fused", [1, 0])
@pytest.mark.parametrize("N,bits,P", SHAPES)
def test_pieces_and_ops(N, bits, P, fused):
    say(f"case: bconv fused={fused} N={N} L={len(bits) - P} P={P}")
    ctx, octx, mods = make_ctx(N, bits, P)
    rng = np.random.default_rng(N + P + fused)
    key = rand_key(rng, mods, N, ctx.dnum)
    rk = ph.relin_key(ctx, host=key)

    def run():
        for chain in (1, 3):
            Ql = ctx.limbs_at(chain)
            QlP = Ql + P
            beta = octx.beta(Ql)
            c2 = rand_limbs(rng, mods[:Ql], N)
            c2b = ph.DeviceBuffer.from_array(c2)
            tb = ph.DeviceBuffer(beta * QlP * N)
            ph.modup(ctx, chain, c2b, tb)
            ph.device_synchronize()
            assert np.array_equal(tb.to_array((beta, QlP, N)), octx.modup(Ql, c2)), ("modup", chain)
            cx = np.stack([rand_limbs(rng, [mods[i] if i < Ql else mods[len(bits) - P + i - Ql]
                                            for i in range(QlP)], N) for _ in range(2)])
            cxb = ph.DeviceBuffer.from_array(cx)
            ob_ = ph.DeviceBuffer(2 * Ql * N)
            ph.moddown_from_ntt(ctx, chain, cxb, ob_, 2)
            ph.device_synchronize()
            out = ob_.to_array((2, Ql, N))
            for h in range(2):
                assert np.array_equal(out[h], octx.moddown(Ql, cx[h])), ("moddown", chain, h)
            a, b = rand_ct(rng, mods[:Ql], N), rand_ct(rng, mods[:Ql], N)
            c1, cb = upload_ct(ctx, a, chain), upload_ct(ctx, b, chain)
            ph.multiply_and_relin_inplace(ctx, c1, cb, rk)
            assert np.array_equal(c1.download(N), octx.multiply_and_relin(Ql, a, b, key)), ("mulrelin", chain)
        # a batch plan over items of three levels
        xs, ys, cxs, cys, Qls = [], [], [], [], []
        for chain in (1, 2, 4):
            Ql = ctx.limbs_at(chain)
            x, y = rand_ct(rng, mods[:Ql], N), rand_ct(rng, mods[:Ql], N)
            xs.append(x)
            ys.append(y)
            Qls.append(Ql)
            cxs.append(upload_ct(ctx, x, chain))
            cys.append(upload_ct(ctx, y, chain))
        bt = ph.batch.multiply_and_relin(ctx, cxs, cys, rk)
        bt.run()
        ph.device_synchronize()
        for i in range(3):
            assert np.array_equal(cxs[i].download(N), octx.multiply_and_relin(Qls[i], xs[i], ys[i], key)), i
        # a rotation (the key switch of pi(c1))
        elt = ph.galois_elt_from_step(5, N)
        gk = ph.galois_key(ctx, [elt], host=key[None])
        Ql = ctx.limbs_at(2)
        a = rand_ct(rng, mods[:Ql], N)
        c = upload_ct(ctx, a, 2)
        ph.rotate_inplace(ctx, c, 5, gk)
        assert np.array_equal(c.download(N), octx.apply_galois(Ql, a, elt, key))

    _with_fused(fused, run)


def test_extreme_inputs_fused():
    """All inputs q - 1 (every byte and digit at its largest) through the fused
    modup and moddown at N = 2^16 with 16-limb digits."""
    N = 1 << 16
    say(f"case: bconv fused extreme N={N}")
    ctx, octx, mods = make_ctx(N, [60] + [50] * 15 + [60] * 16, 16)
    Ql, P = 16, 16
    beta = octx.beta(Ql)
    QlP = Ql + P

    def run():
        qs = np.array(mods[:Ql], np.uint64)[:, None]
        c2 = np.broadcast_to(qs - 1, (Ql, N)).copy()
        tb = ph.DeviceBuffer(beta * QlP * N)
        ph.modup(ctx, 1, ph.DeviceBuffer.from_array(c2), tb)
        ph.device_synchronize()
        assert np.array_equal(tb.to_array((beta, QlP, N)), octx.modup(Ql, c2))
        qlp = np.array(mods, np.uint64)[:, None]
        cx = np.broadcast_to(qlp - 1, (QlP, N)).copy()
        ob_ = ph.DeviceBuffer(Ql * N)
        ph.moddown_from_ntt(ctx, 1, ph.DeviceBuffer.from_array(cx), ob_, 1)
        ph.device_synchronize()
        assert np.array_equal(ob_.to_array((Ql, N)), octx.moddown(Ql, cx))

    _with_fused(1, run)


@pytest.mark.parametrize("N", [1 << 12, 1 << 14])
@pytest.mark.parametrize("P", [1, 2, 5, 8, 13, 16])
def test_special_modulus_sizes(N, P):
    """Digits of every width the fused launch takes (1 .. 16 limbs), with data and
    special primes of both arithmetics: multiply_and_relin at the top level and two
    levels down, bit-exact against the oracle."""
    say(f"case: bconv fused N={N} P={P}")
    L = 2 * P + 3
    bits = [60] + [50] * (L - 1) + [60] * P
    ctx, octx, mods = make_ctx(N, bits, P)
    rng = np.random.default_rng(N + 31 * P)
    key = rand_key(rng, mods, N, ctx.dnum)
    rk = ph.relin_key(ctx, host=key)
    for chain in (1, 3):
        Ql = ctx.limbs_at(chain)
        a, b = rand_ct(rng, mods[:Ql], N), rand_ct(rng, mods[:Ql], N)
        c1, c2 = upload_ct(ctx, a, chain), upload_ct(ctx, b, chain)
        ph.multiply_and_relin_inplace(ctx, c1, c2, rk)
        assert np.array_equal(c1.download(N), octx.multiply_and_relin(Ql, a, b, key)), chain
