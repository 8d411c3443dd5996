"""Batched hot path (phx_batch_*): every item bit-exact against the oracle,
for mixed levels, squaring, mixed rotation steps, and hipGraph replays."""
import numpy as np
import pytest

import oracle_bind as ob
import pyphantom as ph
from gpu_helpers import make_ctx, rand_ct, rand_key, say, upload_ct

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def small():
    ob.set_threads(8)
    ctx, octx, mods = make_ctx(1 << 13, [60, 50, 50, 50, 50, 50, 50, 60, 60, 60], 3)
    rng = np.random.default_rng(7)
    key = rand_key(rng, mods, ctx.N, ctx.dnum)
    return ctx, octx, mods, key


def test_batch_multiply_and_relin_mixed_levels(small):
    ctx, octx, mods, key = small
    N = ctx.N
    say("case: batch mulrelin mixed levels")
    rng = np.random.default_rng(1)
    chains = [1, 1, 3, 7, 2, 1]
    data = []
    for c in chains:
        Ql = ctx.limbs_at(c)
        data.append((c, Ql, rand_ct(rng, mods[:Ql], N), rand_ct(rng, mods[:Ql], N)))
    rk = ph.relin_key(ctx, host=key)
    c1 = [upload_ct(ctx, a, c) for c, _, a, _ in data]
    c2 = [upload_ct(ctx, b, c) for c, _, _, b in data]
    c2[4] = c1[4]   # squaring item: operand aliases the output
    bt = ph.batch.multiply_and_relin(ctx, c1, c2, rk)
    bt.run()
    ph.device_synchronize()
    for i, (c, Ql, a, b) in enumerate(data):
        other = a if i == 4 else b
        assert np.array_equal(c1[i].download(N), octx.multiply_and_relin(Ql, a, other, key)), i


def test_batch_rotate_mixed_steps(small):
    ctx, octx, mods, key = small
    N = ctx.N
    say("case: batch rotate")
    steps_set = [1, 2, -1, 5]
    elts = [ph.galois_elt_from_step(s, N) for s in steps_set]
    rng = np.random.default_rng(2)
    keys = [rand_key(rng, mods, N, ctx.dnum) for _ in elts]
    gk = ph.galois_key(ctx, elts, host=np.stack(keys))
    chains = [1, 2, 1, 5, 3, 1, 1]
    steps = [1, 2, -1, 5, 1, 5, 2]
    data = [(c, ctx.limbs_at(c), rand_ct(rng, mods[:ctx.limbs_at(c)], N)) for c in chains]
    cts = [upload_ct(ctx, a, c) for c, _, a in data]
    bt = ph.batch.rotate(ctx, cts, steps, gk)
    bt.run()
    ph.device_synchronize()
    for i, ((c, Ql, a), st) in enumerate(zip(data, steps)):
        k = keys[steps_set.index(st)]
        want = octx.apply_galois(Ql, a, ph.galois_elt_from_step(st, N), k)
        assert np.array_equal(cts[i].download(N), want), i


def test_batch_graph_replays(small):
    ctx, octx, mods, key = small
    N = ctx.N
    say("case: batch graph replays")
    rng = np.random.default_rng(3)
    B, chain = 4, 1
    Ql = ctx.limbs_at(chain)
    rk = ph.relin_key(ctx, host=key)
    a = [rand_ct(rng, mods[:Ql], N) for _ in range(B)]
    b = [rand_ct(rng, mods[:Ql], N) for _ in range(B)]
    c1 = [upload_ct(ctx, x, chain) for x in a]
    c2 = [upload_ct(ctx, x, chain) for x in b]
    want = [octx.multiply_and_relin(Ql, a[i], b[i], key) for i in range(B)]
    stream = ph.stream_create()
    bt = ph.batch.multiply_and_relin(ctx, c1, c2, rk)
    bt.capture(stream)
    for rep in range(3):
        for i in range(B):     # eager work between replays: reset the inputs
            c1[i].upload(a[i])
        ph.device_synchronize()
        bt.replay(stream)
        ph.synchronize(stream)
        for i in range(B):
            assert np.array_equal(c1[i].download(N), want[i]), (rep, i)


def test_batch_rejects_missing_key_and_aliasing(small):
    ctx, octx, mods, key = small
    N = ctx.N
    rng = np.random.default_rng(4)
    gk = ph.galois_key(ctx, [ph.galois_elt_from_step(1, N)], host=rand_key(rng, mods, N, ctx.dnum)[None])
    c = upload_ct(ctx, rand_ct(rng, mods[:7], N), 1)
    with pytest.raises(ValueError):
        ph.batch.rotate(ctx, [c], [3], gk)
    rk = ph.relin_key(ctx, host=key)
    x = upload_ct(ctx, rand_ct(rng, mods[:7], N), 1)
    y = upload_ct(ctx, rand_ct(rng, mods[:7], N), 1)
    with pytest.raises(ValueError):
        ph.batch.multiply_and_relin(ctx, [x, y], [y, x], rk)


def test_batch_n16_q35_bit_exact():
    """N=2^16, 35 + 12 limbs, dnum 3: four pairs in one batch, graph replay."""
    ob.set_threads(16)
    N = 1 << 16
    say("case: batch N=2^16, 35 + 12 limbs")
    ctx, octx, mods = make_ctx(N, [60] + [50] * 34 + [60] * 12, 12)
    rng = np.random.default_rng(11)
    key = rand_key(rng, mods, N, 3)
    rk = ph.relin_key(ctx, host=key)
    B = 4
    a = [rand_ct(rng, mods[:35], N) for _ in range(B)]
    b = [rand_ct(rng, mods[:35], N) for _ in range(B)]
    c1 = [upload_ct(ctx, x, 1) for x in a]
    c2 = [upload_ct(ctx, x, 1) for x in b]
    stream = ph.stream_create()
    bt = ph.batch.multiply_and_relin(ctx, c1, c2, rk)
    bt.capture(stream)
    bt.replay(stream)
    ph.synchronize(stream)
    for i in (0, 3):
        assert np.array_equal(c1[i].download(N), octx.multiply_and_relin(35, a[i], b[i], key)), i


def test_batch_one_launch_per_stage_across_levels(small):
    """a plan over items of four levels (with a squaring) issues one launch per
    pipeline stage, the same count as a one-level plan, and stays bit-exact; so
    does a rotation plan over two levels (one launch more: the permutation)"""
    ctx, octx, mods, key = small
    N = ctx.N
    say("case: batch over four levels, one launch per stage")
    rng = np.random.default_rng(21)
    chains = [1, 3, 1, 7]
    data = [(c, ctx.limbs_at(c), rand_ct(rng, mods[:ctx.limbs_at(c)], N),
             rand_ct(rng, mods[:ctx.limbs_at(c)], N)) for c in chains]
    rk = ph.relin_key(ctx, host=key)
    c1 = [upload_ct(ctx, a, c) for c, _, a, _ in data]
    c2 = [upload_ct(ctx, b, c) for c, _, _, b in data]
    c2[2] = c1[2]
    bt = ph.batch.multiply_and_relin(ctx, c1, c2, rk)
    one = ph.batch.multiply_and_relin(ctx, [upload_ct(ctx, data[0][2], 1)], [upload_ct(ctx, data[0][3], 1)], rk)
    # 11 launches per branch: the tensor product is the first load of the modup INTT;
    # four items run as two concurrent branches of two
    assert one.branches() == 1 and one.launches() == 11
    assert bt.branches() == 2 and bt.launches() == 11 * 2
    bt.run()
    ph.device_synchronize()
    for i, (c, Ql, a, b) in enumerate(data):
        other = a if i == 2 else b
        assert np.array_equal(c1[i].download(N), octx.multiply_and_relin(Ql, a, other, key)), i
    elt = ph.galois_elt_from_step(2, N)
    gk = ph.galois_key(ctx, [elt], host=key[None])
    x, y = rand_ct(rng, mods[:5], N), rand_ct(rng, mods[:2], N)
    cx = [upload_ct(ctx, x, 3), upload_ct(ctx, y, 6)]
    bt = ph.batch.rotate(ctx, cx, [2, 2], gk)
    assert bt.branches() == 1 and bt.launches() == 12
    bt.run()
    ph.device_synchronize()
    assert np.array_equal(cx[0].download(N), octx.apply_galois(5, x, elt, key))
    assert np.array_equal(cx[1].download(N), octx.apply_galois(2, y, elt, key))


ROTATION_STEPS = [1, 2, 3, 4, 5, 7, 8, 16, 32, 64, 128, 256, 512, 1024, -1, -5]


def test_128_ciphertexts_batched_rotations_bit_exact():
    """128 ciphertexts at N=2^16 with 5..35 limbs, each rotated by a step drawn
    from 16 resident rotation keys, in one captured batch plan; sampled items
    (smallest and largest level among them) bit-exact against the oracle."""
    ob.set_threads(16)
    N = 1 << 16
    say("case: batched rotations (128 cts, 16 keys)")
    ctx, octx, mods = make_ctx(N, [60] + [50] * 34 + [60] * 12, 12)
    elts = [ph.galois_elt_from_step(s, N) for s in ROTATION_STEPS]
    rng = np.random.default_rng(33)
    key_seeds = [1000 + i for i in range(len(elts))]

    def key_of(i):
        return rand_key(np.random.default_rng(key_seeds[i]), mods, N, 3)

    gk = ph.galois_key(ctx, elts, host=np.stack([key_of(i) for i in range(len(elts))]))
    limbs = rng.integers(5, 36, 128)
    steps_idx = rng.integers(0, len(elts), 128)
    chains = [36 - int(l) for l in limbs]           # chain c holds 35 - (c - 1) limbs
    data_seeds = [5000 + i for i in range(128)]
    cts = []
    for i in range(128):
        a = rand_ct(np.random.default_rng(data_seeds[i]), mods[:int(limbs[i])], N)
        cts.append(upload_ct(ctx, a, chains[i]))
    stream = ph.stream_create()
    bt = ph.batch.rotate(ctx, cts, [ROTATION_STEPS[k] for k in steps_idx], gk)
    # one launch per stage and branch over all 31 levels
    assert bt.branches() == 2 and bt.launches() == 12 * 2
    bt.capture(stream)
    bt.replay(stream)
    ph.synchronize(stream)
    order = np.argsort(limbs)
    for i in [int(order[0]), int(order[-1]), 0, 17, 64, 127]:
        Ql = int(limbs[i])
        a = rand_ct(np.random.default_rng(data_seeds[i]), mods[:Ql], N)
        k = int(steps_idx[i])
        want = octx.apply_galois(Ql, a, elts[k], key_of(k))
        assert np.array_equal(cts[i].download(N), want), (i, Ql, ROTATION_STEPS[k])


def test_joined_batch_two_ring_degrees():
    """CKKS plans of N=2^13 and N=2^14 joined into one plan and one graph:
    every item bit-exact against the oracle, on eager runs and on graph replays
    with eager work between them."""
    ob.set_threads(8)
    say("case: joined ckks batch N=2^13 + 2^14")
    parts, checks = [], []
    for N, bits, P in [(1 << 13, [60, 50, 50, 50, 60, 60], 2), (1 << 14, [60] + [50] * 6 + [60, 60, 60], 3)]:
        ctx, octx, mods = make_ctx(N, bits, P)
        rng = np.random.default_rng(N)
        key = rand_key(rng, mods, N, ctx.dnum)
        rk = ph.relin_key(ctx, host=key)
        chains = [1, 2, 1]
        data = [(c, ctx.limbs_at(c), rand_ct(rng, mods[:ctx.limbs_at(c)], N),
                 rand_ct(rng, mods[:ctx.limbs_at(c)], N)) for c in chains]
        c1 = [upload_ct(ctx, a, c) for c, _, a, _ in data]
        c2 = [upload_ct(ctx, b, c) for c, _, _, b in data]
        parts.append(ph.batch.multiply_and_relin(ctx, c1, c2, rk))
        checks.append((ctx, octx, key, data, c1))
    # three items per ring degree: not split; the two parts run as two branches,
    # each in its own arena
    assert [p.branches() for p in parts] == [1, 1]
    both = sum(p.scratch_bytes() for p in parts)
    jb = ph.batch.join(parts)
    assert all(p._h is None for p in parts)
    assert jb.branches() == 2 and jb.scratch_bytes() == both
    stream = ph.stream_create()

    def verify():
        for ctx, octx, key, data, c1 in checks:
            for i, (c, Ql, a, b) in enumerate(data):
                assert np.array_equal(c1[i].download(ctx.N), octx.multiply_and_relin(Ql, a, b, key)), (ctx.N, i)

    def reset():
        for ctx, octx, key, data, c1 in checks:
            for i, (c, Ql, a, b) in enumerate(data):
                c1[i].upload(a)

    jb.run(stream)
    ph.synchronize(stream)
    verify()
    jb.capture(stream)
    for _ in range(2):
        reset()
        ph.device_synchronize()
        jb.replay(stream)
        ph.synchronize(stream)
        verify()


def test_join_rejects_repeated_part(small):
    ctx, octx, mods, key = small
    rk = ph.relin_key(ctx, host=key)
    rng = np.random.default_rng(9)
    a = upload_ct(ctx, rand_ct(rng, mods[:ctx.size_Q], ctx.N), 1)
    b = upload_ct(ctx, rand_ct(rng, mods[:ctx.size_Q], ctx.N), 1)
    bt = ph.batch.multiply_and_relin(ctx, [a], [b], rk)
    with pytest.raises(ValueError):
        ph.batch.join([bt, bt])
    bt.run()   # untouched by the failed join
    ph.device_synchronize()


@pytest.fixture
def branches_switch():
    old = ph.tune_get("batch_branches")
    yield lambda v: ph.tune_set("batch_branches", v)
    ph.tune_set("batch_branches", old)


def test_batch_branches_split_and_join_bit_exact(small, branches_switch):
    """the same items as one branch, split into 2 and 3 concurrent branches of
    about equal work (mixed levels, a squaring), and two split batches joined
    (flattened: four parts over two branches sharing arenas): every item
    bit-exact against the oracle on eager runs and on graph replays with eager
    work between them"""
    ctx, octx, mods, key = small
    N = ctx.N
    rng = np.random.default_rng(31)
    rk = ph.relin_key(ctx, host=key)
    chains = [1, 7, 1, 3, 2, 1, 5, 1]
    data = [(c, ctx.limbs_at(c), rand_ct(rng, mods[:ctx.limbs_at(c)], N),
             rand_ct(rng, mods[:ctx.limbs_at(c)], N)) for c in chains]
    want = [octx.multiply_and_relin(Ql, a, a if i == 3 else b, key) for i, (c, Ql, a, b) in enumerate(data)]
    stream = ph.stream_create()

    def operands(lo, hi):
        c1 = [upload_ct(ctx, data[i][2], data[i][0]) for i in range(lo, hi)]
        c2 = [c1[i - lo] if i == 3 else upload_ct(ctx, data[i][3], data[i][0]) for i in range(lo, hi)]
        return c1, c2

    def check(c1, lo, tag):
        for j, c in enumerate(c1):
            assert np.array_equal(c.download(N), want[lo + j]), (tag, lo + j)

    def reset(c1, lo):
        for j, c in enumerate(c1):
            c.upload(data[lo + j][2])

    for nb in (1, 2, 3):
        branches_switch(nb)
        say(f"case: batch branches={nb}")
        c1, c2 = operands(0, len(data))
        bt = ph.batch.multiply_and_relin(ctx, c1, c2, rk)
        assert bt.branches() == nb and bt.launches() == 11 * nb
        bt.run(stream)
        ph.synchronize(stream)
        check(c1, 0, ("eager", nb))
        bt.capture(stream)
        for rep in range(2):
            reset(c1, 0)
            ph.device_synchronize()
            bt.replay(stream)
            ph.synchronize(stream)
            check(c1, 0, ("replay", nb, rep))
    branches_switch(2)
    say("case: join of two split batches")
    a1, a2 = operands(0, 4)
    b1, b2 = operands(4, 8)
    p = ph.batch.multiply_and_relin(ctx, a1, a2, rk)
    q = ph.batch.multiply_and_relin(ctx, b1, b2, rk)
    assert p.branches() == q.branches() == 2
    jb = ph.batch.join([p, q])
    assert p._h is None and q._h is None
    assert jb.branches() == 2 and jb.launches() == 4 * 11
    jb.capture(stream)
    for rep in range(2):
        reset(a1, 0)
        reset(b1, 4)
        ph.device_synchronize()
        jb.replay(stream)
        ph.synchronize(stream)
        check(a1, 0, ("join", rep))
        check(b1, 4, ("join", rep))
