"""Serving loop (phx_server_*): continuous batching over two ring degrees,
Poisson arrivals, admission by memory. Every request's result is bit-exact
with the same chain run alone through the plain evaluator calls; the
reservation never exceeds the budget."""
import time

import numpy as np
import pytest

import pyphantom as ph
from gpu_helpers import make_ctx, rand_ct, rand_key, say, upload_ct

pytestmark = pytest.mark.gpu

STEPS = [1, 2, 3, -1, 5]


@pytest.fixture(scope="module")
def lanes():
    out = []
    for lane, N in enumerate((1 << 12, 1 << 13)):
        ctx, _, mods = make_ctx(N, [60] + [40] * 9 + [60] * 3, 3)
        rng = np.random.default_rng(100 + lane)
        rk = ph.relin_key(ctx, host=rand_key(rng, mods, N, ctx.dnum))
        elts = [ph.galois_elt_from_step(s, N) for s in STEPS]
        gk = ph.galois_key(ctx, elts, host=np.stack([rand_key(rng, mods, N, ctx.dnum) for _ in elts]))
        out.append((ctx, rk, gk, mods))
    return out


def alone(ctx, rk, gk, ct, steps):
    for st in steps:
        ph.multiply_and_relin_inplace(ctx, ct, ct, rk)
        ct = ph.rescale_to_next(ctx, ct)
        ph.rotate_inplace(ctx, ct, st, gk)
    ph.device_synchronize()
    return ct


def make_requests(lanes, n, seed):
    rng = np.random.default_rng(seed)
    reqs = []
    for i in range(n):
        lane = int(rng.integers(0, 2))
        ctx, _, _, mods = lanes[lane]
        k = int(rng.integers(1, 5))
        chain = int(rng.integers(1, ctx.size_Q - k + 1))
        Ql = ctx.limbs_at(chain)
        data = rand_ct(rng, mods[:Ql], ctx.N)
        steps = [STEPS[j] for j in rng.integers(0, len(STEPS), k)]
        reqs.append((lane, chain, data, steps))
    return reqs


def test_serving_mixed_equals_alone(lanes):
    say("case: serving mixed vs alone (2 ring degrees, Poisson arrivals)")
    reqs = make_requests(lanes, 24, 7)
    srv = ph.server([(c, r, g) for c, r, g, _ in lanes], mem_budget=1 << 30, max_active=64, depth=2)
    rng = np.random.default_rng(8)
    arrivals = np.cumsum(rng.exponential(1 / 400.0, len(reqs)))
    t0 = srv.now()
    ids, i = [], 0
    while i < len(reqs):
        now = srv.now() - t0
        while i < len(reqs) and arrivals[i] <= now:
            lane, chain, data, steps = reqs[i]
            ct = upload_ct(lanes[lane][0], data, chain)
            ct.set_scale(2.0 ** 40)
            ids.append(srv.submit(lane, ct, steps, t0 + arrivals[i]))
            i += 1
        if srv.step() == 0:
            time.sleep(0.0005)
    srv.drain()
    st = srv.stats()
    say(f"serving stats: {st}")
    assert st["completed"] == len(reqs) and st["peak_active"] > 1
    for rid, (lane, chain, data, steps) in zip(ids, reqs):
        assert srv.status(rid) == 2
        got, lat = srv.take(rid)
        assert lat > 0
        ctx, rk, gk, _ = lanes[lane]
        ref = upload_ct(ctx, data, chain)
        ref.set_scale(2.0 ** 40)
        ref = alone(ctx, rk, gk, ref, steps)
        assert got.chain_index() == chain + len(steps)
        assert np.array_equal(got.download(ctx.N), ref.download(ctx.N)), (rid, lane, chain, steps)
        assert got.scale() == pytest.approx(ref.scale())
        assert srv.status(rid) == 3


def test_serving_admission_by_memory(lanes):
    say("case: serving admission by memory")
    ctx, rk, gk, mods = lanes[1]
    N = ctx.N
    # one request at chain 1: state 2 x 10 limbs + rotation scratch of one item
    per = 2 * 10 * N * 8 + N * 8 * (10 * 3 + 4 * 13 + 2 * 13 + 2 * 10)
    srv = ph.server([(c, r, g) for c, r, g, _ in lanes], mem_budget=int(2.5 * per), max_active=64, depth=2)
    rng = np.random.default_rng(9)
    ids = []
    for _ in range(6):
        ct = upload_ct(ctx, rand_ct(rng, mods[:10], N), 1)
        ids.append(srv.submit(1, ct, [1, 2]))
    srv.step()
    assert [srv.status(r) for r in ids].count(1) == 2       # only two fit
    # results must be taken to free their state; take as they complete
    done, spins = 0, 0
    while done < len(ids):
        spins += 1
        assert spins < 100000, "requests did not complete"
        if srv.step() == 0:
            time.sleep(0.0002)
        for r in ids:
            if srv.status(r) == 2:
                srv.take(r)
                done += 1
    st = srv.stats()
    assert st["peak_reserved_bytes"] <= st["budget_bytes"] and st["peak_active"] == 2


def test_serving_rejects_bad_requests(lanes):
    ctx, rk, gk, mods = lanes[0]
    srv = ph.server([(c, r, g) for c, r, g, _ in lanes], mem_budget=1 << 30)
    ct = upload_ct(ctx, rand_ct(np.random.default_rng(1), mods[:10], ctx.N), 1)
    with pytest.raises(ValueError):
        srv.submit(0, ct, [4])                  # no key for step 4
    with pytest.raises(ValueError):
        srv.submit(0, ct, [1] * 10)             # past the end of the chain
    with pytest.raises(ValueError):
        srv.submit(1, ct, [1])                  # wrong ring degree for the lane
    tiny = ph.server([(c, r, g) for c, r, g, _ in lanes], mem_budget=1 << 10)
    with pytest.raises(ValueError):
        tiny.submit(0, ct, [1])                 # larger than the whole budget


# ---- the serving workload at its full shape: N in {2^15, 2^16}, 35 data limbs
# (60 + 34 x 50 bit) + 12 special (60 bit), dnum 3, start level 10..35 limbs,
# chains of k in 1..8 links, Poisson arrivals ----
SHAPE_BITS = [60] + [50] * 34 + [60] * 12
SHAPE_STEPS = [1, 2, -1, 5, 16, -7]


@pytest.fixture(scope="module")
def shape_lanes():
    import oracle_bind as ob

    ob.set_threads(16)
    out = []
    for lane, N in enumerate((1 << 15, 1 << 16)):
        ctx, octx, mods = make_ctx(N, SHAPE_BITS, 12)
        rng = np.random.default_rng(500 + lane)
        rkey = rand_key(rng, mods, N, ctx.dnum)
        rk = ph.relin_key(ctx, host=rkey)
        elts = [ph.galois_elt_from_step(s, N) for s in SHAPE_STEPS]
        gkeys = [rand_key(rng, mods, N, ctx.dnum) for _ in elts]
        gk = ph.galois_key(ctx, elts, host=np.stack(gkeys))
        out.append((ctx, rk, gk, mods, octx, rkey, dict(zip(SHAPE_STEPS, zip(elts, gkeys)))))
    return out


def oracle_chain(octx, rkey, gkeys, data, steps):
    """the chain on the oracle: (multiply_and_relin of ct with itself, rescale,
    apply_galois with the step's key) per link (src/evaluate.cu:1061-1104,
    1545-1565, 1567-1630)"""
    ct = data
    for st in steps:
        Ql = ct.shape[1]
        ct = octx.multiply_and_relin(Ql, ct, ct, rkey)
        ct = octx.rescale(Ql, ct)
        elt, key = gkeys[st]
        ct = octx.apply_galois(Ql - 1, ct, elt, key)
    return ct


def test_serving_full_shape_mixed_equals_alone_and_oracle(shape_lanes):
    say("case: serving at N in {2^15, 2^16}, 10..35 limbs, k in 1..8, Poisson 400/s")
    rng = np.random.default_rng(21)
    n_req = 24
    reqs = []
    for _ in range(n_req):
        lane = int(rng.integers(0, 2))
        ctx, _, _, mods = shape_lanes[lane][:4]
        k = int(rng.integers(1, 9))
        limbs = int(rng.integers(max(10, k + 1), 36))
        chain = ctx.size_Q - limbs + 1
        data = rand_ct(rng, mods[:limbs], ctx.N)
        steps = [SHAPE_STEPS[j] for j in rng.integers(0, len(SHAPE_STEPS), k)]
        reqs.append((lane, chain, data, steps))
    assert {r[0] for r in reqs} == {0, 1} and max(len(r[3]) for r in reqs) >= 7
    srv = ph.server([(c, r, g) for c, r, g in (l[:3] for l in shape_lanes)], mem_budget=24 << 30,
                    max_active=256, depth=2)
    arrivals = np.cumsum(rng.exponential(1 / 400.0, n_req))
    t0 = srv.now()
    ids, i = [], 0
    while i < n_req:
        now = srv.now() - t0
        while i < n_req and arrivals[i] <= now:
            lane, chain, data, steps = reqs[i]
            ct = upload_ct(shape_lanes[lane][0], data, chain)
            ct.set_scale(2.0 ** 50)
            ids.append(srv.submit(lane, ct, steps, t0 + arrivals[i]))
            i += 1
        if srv.step() == 0:
            time.sleep(0.0005)
    srv.drain()
    st = srv.stats()
    say(f"serving stats: {st}")
    assert st["completed"] == n_req and st["peak_active"] > 1
    results = []
    for rid, (lane, chain, data, steps) in zip(ids, reqs):
        got, lat = srv.take(rid)
        ctx, rk, gk = shape_lanes[lane][:3]
        ref = upload_ct(ctx, data, chain)
        ref.set_scale(2.0 ** 50)
        ref = alone(ctx, rk, gk, ref, steps)
        g = got.download(ctx.N)
        assert got.chain_index() == chain + len(steps)
        assert np.array_equal(g, ref.download(ctx.N)), (rid, lane, chain, steps)
        results.append(g)
    # two requests per lane against the oracle's chain (the cheapest ones: the
    # oracle's key switch at N = 2^16 takes about a second per link)
    for lane in (0, 1):
        idx = [j for j in range(n_req) if reqs[j][0] == lane]
        idx.sort(key=lambda j: len(reqs[j][3]) * reqs[j][2].shape[1])
        for j in idx[:2]:
            _, chain, data, steps = reqs[j]
            octx, rkey, gkeys = shape_lanes[lane][4:]
            say(f"case: oracle chain lane {lane} limbs {data.shape[1]} steps {steps}")
            want = oracle_chain(octx, rkey, gkeys, data, steps)
            assert np.array_equal(results[j], want), (j, lane, steps)
