"""The allocation-free multiply_and_relin_inplace (phx_multiply_and_relin_inplace_noalloc):
bit-exact against phx_multiply_and_relin_inplace on two calls in a row, the second on a
smaller input (a lower level), at N = 2^16 with 35 data limbs + 12 special limbs, dnum 3;
a shape whose workspace exceeds the static scratch is refused with nothing launched."""
import numpy as np
import pytest

import pyphantom as ph
from gpu_helpers import make_ctx, rand_ct, rand_key, say, upload_ct

pytestmark = pytest.mark.gpu

BITS = [60] + [50] * 34 + [60] * 12


@pytest.fixture(scope="module")
def big():
    ctx, _, mods = make_ctx(1 << 16, BITS, 12)
    rng = np.random.default_rng(700)
    rk = ph.relin_key(ctx, host=rand_key(rng, mods, 1 << 16, ctx.dnum))
    return ctx, mods, rk


def test_noalloc_equals_allocating_twice(big):
    ctx, mods, rk = big
    N = ctx.N
    rng = np.random.default_rng(701)
    for chain in (1, 15):       # 35 limbs, then 21 limbs
        say(f"case: no-alloc multiply_and_relin N=2^16 chain {chain}")
        L = ctx.limbs_at(chain)
        a, b = rand_ct(rng, mods[:L], N), rand_ct(rng, mods[:L], N)
        x1, y1 = upload_ct(ctx, a, chain), upload_ct(ctx, b, chain)
        x2, y2 = upload_ct(ctx, a, chain), upload_ct(ctx, b, chain)
        x1.set_scale(2.0 ** 40)
        y1.set_scale(2.0 ** 40)
        x2.set_scale(2.0 ** 40)
        y2.set_scale(2.0 ** 40)
        ph.multiply_and_relin_inplace(ctx, x1, y1, rk)
        ph.multiply_and_relin_inplace_noalloc(ctx, x2, y2, rk)
        ph.device_synchronize()
        assert np.array_equal(x2.download(N), x1.download(N)), chain
        assert x2.scale() == x1.scale() and x2.size() == 2 and x2.chain_index() == chain
    # squaring (one ciphertext as both operands)
    L = ctx.limbs_at(20)
    a = rand_ct(rng, mods[:L], N)
    s1, s2 = upload_ct(ctx, a, 20), upload_ct(ctx, a, 20)
    ph.multiply_and_relin_inplace(ctx, s1, s1, rk)
    ph.multiply_and_relin_inplace_noalloc(ctx, s2, s2, rk)
    assert np.array_equal(s2.download(N), s1.download(N))


def test_noalloc_refuses_beyond_scratch():
    say("case: no-alloc multiply_and_relin refuses a shape above its scratch")
    N = 1 << 17
    ctx, _, mods = make_ctx(N, [60] + [50] * 34 + [60] * 12, 12)
    rng = np.random.default_rng(702)
    rk = ph.relin_key(ctx, host=rand_key(rng, mods, N, ctx.dnum))
    a = rand_ct(rng, mods[:35], N)
    x = upload_ct(ctx, a, 1)
    with pytest.raises(ValueError):
        ph.multiply_and_relin_inplace_noalloc(ctx, x, x, rk)
    assert np.array_equal(x.download(N), a)
