"""Benchmark: CKKS HMult+relin per second at N=2^16, L=35 (12 special primes,
dnum=3), bit-exact path, on one MI355X; plus the NTT's HBM throughput.

A step is one pass of the hot path over one batch: B ciphertext pairs, each
multiplied and relinearised with the same resident relinearisation key
(plain path: one ciphertext pair per call). Inputs are synthetic, seeded,
uniform residues already resident in HBM when timing starts.

Prints one JSON line (rank 0). `--gpus N` with N > 1 exits: the library runs
on a single GPU.
"""
import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
for p in (os.path.join(ROOT, "phantom-fhe_amd"), os.path.join(ROOT, "tests"), ROOT):
    if p not in sys.path:
        sys.path.insert(0, p)

import numpy as np  # noqa: E402

HBM_PEAK_GBS = 8000.0
BITS_N16_Q35_P12 = [60] + [50] * 34 + [60] * 12


def log(msg):
    print(msg, file=sys.stderr, flush=True)


def make_context(ph, N, bits, size_P):
    mods = ph.coeff_modulus_create(N, bits)
    parms = ph.EncryptionParameters(ph.SCHEME_CKKS)
    parms.set_poly_modulus_degree(N)
    parms.set_coeff_modulus(mods)
    parms.set_special_modulus_size(size_P)
    return ph.context(parms), mods


def rand_ct(rng, mods, N):
    return np.stack([np.stack([rng.integers(0, q, N, dtype=np.uint64) for q in mods]) for _ in range(2)])


def bench_hmult(ph, ctx, mods, N, batch, steps, warmup, stream):
    rng = np.random.default_rng(1234)
    Ql = ctx.size_Q
    key = np.stack([np.stack([np.stack([rng.integers(0, q, N, dtype=np.uint64) for q in mods])
                              for _ in range(2)]) for _ in range(ctx.dnum)])
    rlk = ph.relin_key(ctx, host=key)
    del key
    a = [ph.ciphertext(ctx, 1, 2) for _ in range(batch)]
    b = [ph.ciphertext(ctx, 1, 2) for _ in range(batch)]
    for i in range(batch):
        a[i].upload(rand_ct(rng, mods[:Ql], N))
        b[i].upload(rand_ct(rng, mods[:Ql], N))

    def step():
        for i in range(batch):
            ph.multiply_and_relin_inplace(ctx, a[i], b[i], rlk, stream)

    for _ in range(warmup):
        step()
    ph.synchronize(stream)
    ev0, ev1 = ph.Event(), ph.Event()
    ph.device_synchronize()
    t0 = time.perf_counter()
    ev0.record(stream)
    for _ in range(steps):
        step()
    ev1.record(stream)
    ph.synchronize(stream)
    ph.device_synchronize()
    wall = time.perf_counter() - t0
    ms = ev0.elapsed_ms(ev1)
    return ms, wall


def bench_hmult_batched(ph, ctx, mods, N, batch, steps, warmup, stream):
    """The same work as bench_hmult, as one batch plan replayed from a hipGraph."""
    rng = np.random.default_rng(4321)
    Ql = ctx.size_Q
    key = np.stack([np.stack([np.stack([rng.integers(0, q, N, dtype=np.uint64) for q in mods])
                              for _ in range(2)]) for _ in range(ctx.dnum)])
    rlk = ph.relin_key(ctx, host=key)
    del key
    a = [ph.ciphertext(ctx, 1, 2) for _ in range(batch)]
    b = [ph.ciphertext(ctx, 1, 2) for _ in range(batch)]
    for i in range(batch):
        a[i].upload(rand_ct(rng, mods[:Ql], N))
        b[i].upload(rand_ct(rng, mods[:Ql], N))
    bt = ph.batch.multiply_and_relin(ctx, a, b, rlk)
    bt.capture(stream)
    for _ in range(warmup):
        bt.replay(stream)
    ph.synchronize(stream)
    ev0, ev1 = ph.Event(), ph.Event()
    ev0.record(stream)
    for _ in range(steps):
        bt.replay(stream)
    ev1.record(stream)
    ph.synchronize(stream)
    return ev0.elapsed_ms(ev1), bt.scratch_bytes()


def bench_bfv_mixed_rings(ph, stream, n_ct=64, reps=3):
    """BFV multiply + relin (BEHZ) over 64 ciphertexts with N uniform in
    {2^13, 2^14, 2^15} (4/8/16 limbs of 54-bit primes, t = 65537, P = 1): one
    multiply_and_relin_inplace call per ciphertext against one plan per ring
    degree joined into one captured graph."""
    limbs = {1 << 13: 4, 1 << 14: 8, 1 << 15: 16}
    rng = np.random.default_rng(404)
    Ns = [int(x) for x in rng.choice(list(limbs), n_ct)]
    ctxs, rks = {}, {}
    for N, k in limbs.items():
        mods = ph.coeff_modulus_create(N, [54] * k + [60])
        parms = ph.EncryptionParameters(ph.SCHEME_BFV)
        parms.set_poly_modulus_degree(N)
        parms.set_coeff_modulus(mods)
        parms.set_special_modulus_size(1)
        parms.set_plain_modulus(65537)
        parms.set_mul_tech(ph.MUL_TECH_BEHZ)
        ctx = ph.context(parms)
        key = np.stack([np.stack([np.stack([rng.integers(0, q, N, dtype=np.uint64) for q in mods])
                                  for _ in range(2)]) for _ in range(ctx.dnum)])
        ctxs[N] = (ctx, mods)
        rks[N] = ph.relin_key(ctx, host=key)
    a, b = [], []
    for N in Ns:
        ctx, mods = ctxs[N]
        for lst in (a, b):
            c = ph.ciphertext(ctx, 1, 2)
            c.upload(np.stack([np.stack([rng.integers(0, q, N, dtype=np.uint64) for q in mods[:limbs[N]]])
                               for _ in range(2)]))
            lst.append(c)
    ev0, ev1 = ph.Event(), ph.Event()
    for i, N in enumerate(Ns):      # warm-up
        ph.multiply_and_relin_inplace(ctxs[N][0], a[i], b[i], rks[N], stream)
    ph.synchronize(stream)
    ev0.record(stream)
    for _ in range(reps):
        for i, N in enumerate(Ns):
            ph.multiply_and_relin_inplace(ctxs[N][0], a[i], b[i], rks[N], stream)
    ev1.record(stream)
    ph.synchronize(stream)
    plain = reps * n_ct / (ev0.elapsed_ms(ev1) / 1e3)
    parts = []
    for N in limbs:
        idx = [i for i, n in enumerate(Ns) if n == N]
        if idx:
            parts.append(ph.batch.multiply_and_relin(ctxs[N][0], [a[i] for i in idx], [b[i] for i in idx], rks[N]))
    jb = ph.batch.join(parts)
    jb.capture(stream)
    jb.replay(stream)
    ph.synchronize(stream)
    ev0.record(stream)
    for _ in range(reps):
        jb.replay(stream)
    ev1.record(stream)
    ph.synchronize(stream)
    batched = reps * n_ct / (ev0.elapsed_ms(ev1) / 1e3)
    return {"workload": f"{n_ct} BFV ciphertexts (BEHZ), N uniform in 2^13..2^15 with 4/8/16 limbs, t=65537",
            "per_N": {str(N): Ns.count(N) for N in limbs},
            "plain_op_per_s": round(plain, 1), "batched_op_per_s": round(batched, 1),
            "batched_over_plain": round(batched / plain, 3), "scratch_GiB": round(jb.scratch_bytes() / 2**30, 2)}


ROT_STEPS = [1, 2, 3, 4, 5, 7, 8, 16, 32, 64, 128, 256, 512, 1024, -1, -5]


def bench_rotations(ph, ctx, mods, N, stream, reps=3, n_ct=128):
    """Batched CKKS rotations: n_ct ciphertexts with 5..35 limbs, each rotated
    by a step from 16 resident rotation keys; one captured batch plan replayed,
    against one rotate_inplace call per ciphertext."""
    rng = np.random.default_rng(77)
    sk = ph.secret_key(ctx, seed=77)
    gk = sk.create_galois_keys(ctx, [ph.galois_elt_from_step(st, N) for st in ROT_STEPS], seed=78)
    limbs = rng.integers(5, 36, n_ct)
    steps = [ROT_STEPS[k] for k in rng.integers(0, len(ROT_STEPS), n_ct)]
    cts = []
    for l in limbs:
        c = ph.ciphertext(ctx, 36 - int(l), 2)
        c.upload(rand_ct(rng, mods[:int(l)], N))
        cts.append(c)
    ev0, ev1 = ph.Event(), ph.Event()
    for c, st in zip(cts, steps):     # warm-up
        ph.rotate_inplace(ctx, c, st, gk, stream)
    ph.synchronize(stream)
    ev0.record(stream)
    for _ in range(reps):
        for c, st in zip(cts, steps):
            ph.rotate_inplace(ctx, c, st, gk, stream)
    ev1.record(stream)
    ph.synchronize(stream)
    plain = reps * n_ct / (ev0.elapsed_ms(ev1) / 1e3)
    bt = ph.batch.rotate(ctx, cts, steps, gk)
    bt.capture(stream)
    bt.replay(stream)
    ph.synchronize(stream)
    ev0.record(stream)
    for _ in range(reps):
        bt.replay(stream)
    ev1.record(stream)
    ph.synchronize(stream)
    batched = reps * n_ct / (ev0.elapsed_ms(ev1) / 1e3)
    return {"workload": f"{n_ct} ciphertexts, N=2^16, 5..35 limbs, 16 resident rotation keys",
            "plain_rot_per_s": round(plain, 1), "batched_rot_per_s": round(batched, 1),
            "batched_over_plain": round(batched / plain, 3), "mean_limbs": float(np.mean(limbs)),
            "scratch_GiB": round(bt.scratch_bytes() / 2**30, 2)}


def bench_serving(ph, lanes, qps, n_req, seed=55):
    """Serving loop: Poisson arrivals at `qps` of chains of k in 1..8 links
    (HMult+relin, rescale, rotate by a random step), N in {2^15, 2^16}, start
    level 10..35 limbs; keys resident; p50/p99 latency beside throughput."""
    rng = np.random.default_rng(seed)
    pool = []        # per lane: ciphertexts at every start level, uploaded once
    for ctx, mods, _, _ in lanes:
        full = rand_ct(rng, mods[:35], ctx.N)
        per = {}
        for limbs in range(10, 36):
            c = ph.ciphertext(ctx, 36 - limbs, 2)
            c.upload(np.ascontiguousarray(full[:, :limbs]))
            c.set_scale(2.0 ** 50)
            per[limbs] = c
        pool.append(per)
    srv = ph.server([(ctx, rk, gk) for ctx, _, rk, gk in lanes], mem_budget=64 << 30, max_active=512, depth=2)
    reqs = []
    for _ in range(n_req):
        k = int(rng.integers(1, 9))
        limbs = int(rng.integers(max(10, k + 1), 36))
        reqs.append((int(rng.integers(0, len(lanes))), limbs, [ROT_STEPS[j] for j in rng.integers(0, 16, k)]))
    arrivals = np.cumsum(rng.exponential(1.0 / qps, n_req))
    t0 = srv.now()
    ids, i, done = [], 0, 0
    while done < n_req:
        now = srv.now() - t0
        while i < n_req and arrivals[i] <= now:
            lane, limbs, steps = reqs[i]
            ids.append(srv.submit(lane, pool[lane][limbs], steps, t0 + arrivals[i]))
            i += 1
        issued = srv.step()
        for rid in ids[done:]:
            if srv.status(rid) != 2:
                break
            srv.take(rid)
            done += 1
        if issued == 0 and i < n_req:
            time.sleep(max(0.0, min(0.001, arrivals[i] - (srv.now() - t0))))
    st = srv.stats()
    return {"workload": f"{n_req} requests, Poisson {qps}/s, k in 1..8 links of (HMult+relin, rescale, rotate), "
                        f"N in {{2^15, 2^16}}, start 10..35 limbs, 16 resident rotation keys per N",
            "p50_ms": round(st["p50_ms"], 3), "p99_ms": round(st["p99_ms"], 3), "mean_ms": round(st["mean_ms"], 3),
            "req_per_s": round(st["req_per_s"], 1), "ops_per_s": round(st["ops_per_s"], 1),
            "peak_active": st["peak_active"], "ticks": st["ticks"],
            "peak_reserved_GiB": round(st["peak_reserved_bytes"] / 2**30, 2)}


def ckks_error_report(ph, ctx, scale=2.0 ** 50, step=5, seed=20):
    """Real CKKS encryptions of seeded messages (|x|, |y| <= sqrt(2)): max slot
    error against FP64 after each stage of x*y -> relin -> rescale -> rotate."""
    N = ctx.N
    enc = ph.ckks_encoder(ctx)
    rng = np.random.default_rng(seed)
    x = rng.uniform(-1, 1, N // 2) + 1j * rng.uniform(-1, 1, N // 2)
    y = rng.uniform(-1, 1, N // 2) + 1j * rng.uniform(-1, 1, N // 2)
    sk = ph.secret_key(ctx, seed=seed)
    rk = sk.gen_relinkey(ctx, seed=seed + 1)
    gk = sk.create_galois_keys(ctx, [ph.galois_elt_from_step(step, N)], seed=seed + 2)
    cx = sk.encrypt_symmetric(ctx, enc.encode(x, scale, 1), 1, seed=seed + 3)
    cy = sk.encrypt_symmetric(ctx, enc.encode(y, scale, 1), 1, seed=seed + 4)
    cx.set_scale(scale)
    cy.set_scale(scale)

    def err(ct, want):
        got = enc.decode(sk.decrypt(ctx, ct), ct.chain_index(), ct.scale())
        return float(np.max(np.abs(got - want)))

    out = {"scale_log2": 50, "fresh": err(cx, x)}
    ph.multiply_and_relin_inplace(ctx, cx, cy, rk)
    out["mul_relin"] = err(cx, x * y)
    r = ph.rescale_to_next(ctx, cx)
    out["rescale"] = err(r, x * y)
    ph.rotate_inplace(ctx, r, step, gk)
    out["rotate"] = err(r, np.roll(x * y, -step))
    return {k: (float(f"{v:.3e}") if isinstance(v, float) else v) for k, v in out.items()}


def bench_ntt(ph, ctx, N, L, npolys, iters, stream, warmup=80):
    """Forward NTT over npolys x L limbs in one launch pair (pass 1 + pass 2):
    mean over iters pairs back to back after warmup pairs, and the fastest single pair."""
    rng = np.random.default_rng(7)
    mods = ctx.coeff_modulus[:L]
    host = np.concatenate([np.stack([rng.integers(0, q, N, dtype=np.uint64) for q in mods])
                           for _ in range(npolys)])
    buf = ph.DeviceBuffer.from_array(host)
    del host
    # warm-up of about 30 ms of transforms: in the first milliseconds of sustained
    # load the launch pairs run up to 20% slower for a while (a transient of the
    # GPU's clock and power management, seen in kernel traces), then settle
    for _ in range(warmup):
        ph.ntt_forward_batch(ctx, buf, L, 0, npolys, L * N, stream)
    ph.synchronize(stream)
    # mean: iters launch pairs back to back between two events (an event recorded
    # between launches leaves a gap of a few microseconds before the next one)
    e0, e1 = ph.Event(), ph.Event()
    e0.record(stream)
    for _ in range(iters):
        ph.ntt_forward_batch(ctx, buf, L, 0, npolys, L * N, stream)
    e1.record(stream)
    ph.synchronize(stream)
    mean = e0.elapsed_ms(e1) / iters
    # fastest single pair, each bracketed by its own events
    evs = [ph.Event() for _ in range(iters + 1)]
    evs[0].record(stream)
    for i in range(iters):
        ph.ntt_forward_batch(ctx, buf, L, 0, npolys, L * N, stream)
        evs[i + 1].record(stream)
    ph.synchronize(stream)
    per = [evs[i].elapsed_ms(evs[i + 1]) for i in range(iters)]
    return float(mean), float(np.min(per))


def bench_copy(ph, n_u64, iters, stream):
    """Plain streaming copy kernel of known size: the bandwidth a kernel that
    reads and writes each byte once actually reaches on this GPU."""
    a = ph.DeviceBuffer(n_u64)
    b = ph.DeviceBuffer(n_u64)
    ph.copy_u64(b, a, n_u64, stream)
    ph.synchronize(stream)
    e0, e1 = ph.Event(), ph.Event()
    e0.record(stream)
    for _ in range(iters):
        ph.copy_u64(b, a, n_u64, stream)
    e1.record(stream)
    ph.synchronize(stream)
    ms = e0.elapsed_ms(e1) / iters
    return 2 * n_u64 * 8 / (ms / 1e3) / 1e9


def cpu_baseline(N, bits, size_P, budget_s):
    import oracle_bind as ob

    # the host's CPU share for one GPU is 16 threads; all of them when fewer exist
    host_cpus = os.cpu_count() or 1
    threads = max(1, min(16, host_cpus))
    ob.set_threads(threads)
    mods = ob.coeff_modulus_create(N, bits)
    octx = ob.Ctx(N.bit_length() - 1, mods, size_P)
    rng = np.random.default_rng(99)
    Ql = len(bits) - size_P
    dnum = (Ql + size_P - 1) // size_P
    a, b = rand_ct(rng, mods[:Ql], N), rand_ct(rng, mods[:Ql], N)
    key = np.stack([np.stack([np.stack([rng.integers(0, q, N, dtype=np.uint64) for q in mods])
                              for _ in range(2)]) for _ in range(dnum)])
    n = 0
    t0 = time.perf_counter()
    while True:
        octx.multiply_and_relin(Ql, a, b, key)
        n += 1
        el = time.perf_counter() - t0
        if el >= budget_s:
            break
    # the same op on one thread (one op: ~10-20 s on one core)
    ob.set_threads(1)
    t1 = time.perf_counter()
    octx.multiply_and_relin(Ql, a, b, key)
    el1 = time.perf_counter() - t1
    ob.set_threads(threads)
    return {"value": n / el, "unit": "HMult+relin/s", "cores": threads, "kind": "port",
            "host_cpus": host_cpus, "value_1thread": 1.0 / el1,
            "sample": f"{n} HMult+relin at N=2^16, L=35+12, dnum=3 on the oracle (CPU restatement of the "
                      f"reference algorithm, {threads} threads), {el:.1f} s; plus one op on 1 thread, {el1:.1f} s"}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--batch", type=int, default=64)
    ap.add_argument("--ntt-polys", type=int, default=16)
    ap.add_argument("--cpu-seconds", type=float, default=12.0)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--ntt-only", action="store_true",
                    help="only the forward NTT over 47 x 16 limbs and the copy kernel (for counter runs)")
    ap.add_argument("--rotations-only", action="store_true",
                    help="only the batched-rotation comparison (128 ciphertexts, 16 keys); one JSON line")
    ap.add_argument("--hmult-only", action="store_true",
                    help="only the plain and batched HMult+relin (for kernel-trace runs)")
    ap.add_argument("--serve-qps", type=float, default=400.0)
    ap.add_argument("--serve-requests", type=int, default=400)
    ap.add_argument("--serve-sweep", type=float, nargs="*", default=[200.0, 400.0, 800.0, 1600.0],
                    help="offered loads (requests/s) of the serving sweep")
    ap.add_argument("--serve-sweep-requests", type=int, default=300)
    ap.add_argument("--no-autotune", action="store_true",
                    help="skip the autotuner (kernel variants stay at the process-wide switches)")
    ap.add_argument("--tune", action="append", default=[], metavar="KEY=VALUE",
                    help="kernel-variant switch (bconv_mfma, ntt_fp64, ntt_fused, ntt_chunk_limbs, ...)")
    args = ap.parse_args()
    if args.gpus != 1:
        print(json.dumps({"error": "the library runs on one GPU only; --gpus must be 1"}))
        sys.exit(2)

    import pyphantom as ph

    tune = {}
    for kv in args.tune:
        k, v = kv.split("=", 1)
        ph.tune_set(k, int(v))
        tune[k] = int(v)
    N = 1 << 16
    ctx, mods = make_context(ph, N, BITS_N16_Q35_P12, 12)
    stream = ph.stream_create()
    autotune = None
    if not args.no_autotune and not args.tune and not args.ntt_only and not args.rotations_only:
        autotune = ctx.autotune()
        log(f"bench: autotune {autotune}")
    if args.rotations_only:
        print(json.dumps(bench_rotations(ph, ctx, mods, N, stream)), flush=True)
        return
    if args.ntt_only:
        L = ctx.size_Q + ctx.size_P
        ntt_ms, ntt_min = bench_ntt(ph, ctx, N, L, args.ntt_polys, 30, stream)
        copy_gbs = bench_copy(ph, 1 << 27, 10, stream)
        print(json.dumps({"ntt_limbs": L * args.ntt_polys, "ntt_ms": ntt_ms, "ntt_bytes_algorithmic":
                          2 * N * 8 * L * args.ntt_polys, "copy_bytes_per_launch": 2 * (1 << 27) * 8,
                          "copy_GBps": copy_gbs}), flush=True)
        return
    log(f"bench: HMult+relin N=2^16 L=35 P=12 dnum={ctx.dnum} batch={args.batch} steps={args.steps}")
    ms_plain, wall = bench_hmult(ph, ctx, mods, N, args.batch, args.steps, args.warmup, stream)
    ops = args.batch * args.steps
    plain = ops / (ms_plain / 1e3)
    log(f"bench: plain   {ops} ops in {ms_plain:.2f} ms (events), {wall * 1e3:.2f} ms wall -> {plain:.1f} op/s")
    ms, scratch = bench_hmult_batched(ph, ctx, mods, N, args.batch, args.steps, args.warmup, stream)
    value = ops / (ms / 1e3)
    log(f"bench: batched {ops} ops in {ms:.2f} ms (graph replays) -> {value:.1f} op/s "
        f"({value / plain:.2f}x plain; scratch {scratch / 2**30:.2f} GiB)")
    if args.hmult_only:
        print(json.dumps({"plain_op_per_s": round(plain, 1), "batched_op_per_s": round(value, 1),
                          "ops": ops, "batch": args.batch, "steps": args.steps, "warmup": args.warmup}), flush=True)
        return

    # dominant kernel family: the forward NTT (two passes) over 47 x P limbs, at the
    # default launch shape (the one the committed counter runs measured)
    L = ctx.size_Q + ctx.size_P
    tuned = ctx.tuning()
    ctx.set_tuning(**{k: -1 for k in tuned})
    ntt_ms, ntt_min = bench_ntt(ph, ctx, N, L, args.ntt_polys, 30, stream)
    ctx.set_tuning(**tuned)
    ntt_bytes = 2 * N * 8 * L * args.ntt_polys
    achieved = ntt_bytes / (ntt_ms / 1e3) / 1e9
    log(f"bench: NTT {L}x{args.ntt_polys} limbs: {ntt_ms:.3f} ms -> {achieved:.1f} GB/s algorithmic")
    copy_gbs = bench_copy(ph, 1 << 27, 10, stream)
    log(f"bench: copy kernel (1 GiB read + 1 GiB write): {copy_gbs:.1f} GB/s")
    traffic = None
    tpath = os.path.join(ROOT, "profiles", "ntt_fwd_traffic.json")
    if os.path.exists(tpath):
        traffic = json.load(open(tpath)).get("bytes_per_launch")

    rot = bench_rotations(ph, ctx, mods, N, stream)
    log(f"bench: rotations {rot}")
    bfv4 = bench_bfv_mixed_rings(ph, stream)
    log(f"bench: bfv mixed rings {bfv4}")
    serving = None
    if args.serve_requests > 0:
        lanes = []
        for n_ in (1 << 15, 1 << 16):
            c_, m_ = (ctx, mods) if n_ == N else make_context(ph, n_, BITS_N16_Q35_P12, 12)
            sk_ = ph.secret_key(c_, seed=90)
            lanes.append((c_, m_, sk_.gen_relinkey(c_, seed=91),
                          sk_.create_galois_keys(c_, [ph.galois_elt_from_step(st, n_) for st in ROT_STEPS], seed=92)))
        serving = bench_serving(ph, lanes, args.serve_qps, args.serve_requests)
        log(f"bench: serving {serving}")
        # offered-load sweep: p50/p99 and throughput at each load, the saturation rate
        sweep = []
        for qps in args.serve_sweep:
            r = bench_serving(ph, lanes, qps, args.serve_sweep_requests, seed=56)
            sweep.append({"qps": qps, "p50_ms": r["p50_ms"], "p99_ms": r["p99_ms"], "req_per_s": r["req_per_s"],
                          "ops_per_s": r["ops_per_s"], "peak_active": r["peak_active"]})
            log(f"bench: serving sweep {sweep[-1]}")
        if sweep:
            serving["sweep"] = sweep
            serving["saturation_req_per_s"] = max(x["req_per_s"] for x in sweep)
    ckks_err = ckks_error_report(ph, ctx)
    log(f"bench: CKKS max slot error vs FP64 (scale 2^50): {ckks_err}")

    cpu = None
    if not args.no_cpu:
        log("bench: CPU baseline (oracle) ...")
        cpu = cpu_baseline(N, BITS_N16_Q35_P12, 12, args.cpu_seconds)

    # algorithmic bytes of one HMult+relin with the key streamed per op:
    # (4*35 in + 2*35 out + 3*2*47 key) limbs x N x 8 B
    op_bytes = (4 * 35 + 2 * 35 + 3 * 2 * 47) * N * 8
    out = {
        "metric": "HMult+relin/sec at N=2^16 L=35 (dnum=3), bit-exact",
        "value": round(value, 2),
        "unit": "HMult+relin/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(ms / args.steps, 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "u64",
        "data": "synthetic (seeded uniform residues, resident in HBM)",
        "config": {"workload": "CKKS HMult+relin, N=65536, 35 data limbs (60+34x50 bit) + 12 special (60 bit), "
                               "dnum=3; batched path: one batch plan over all pairs of a step, replayed as a "
                               "captured hipGraph",
                   "batch_per_step": args.batch, "batch_branches": ph.tune_get("batch_branches"),
                   "mode": "batched", "tune": tune},
        "autotune": autotune,
        "plain_value": round(plain, 2),
        "plain_ms_per_step": round(ms_plain / args.steps, 4),
        "batched_over_plain": round(value / plain, 3),
        "op_algorithmic_GBps": round(op_bytes * value / 1e9, 1),
        "ntt": {"limbs": L * args.ntt_polys, "N": N, "ms": round(ntt_ms, 4), "ms_min": round(ntt_min, 4),
                "GBps_algorithmic": round(achieved, 1)},
        "copy_kernel_GBps": round(copy_gbs, 1),
        "roofline": {"bound": "hbm", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": round(achieved / HBM_PEAK_GBS, 4), "traffic": traffic,
                     "kernel": "forward NTT (k_fwd_cols + k_fwd_rows), algorithmic bytes 2*N*8 per limb"},
        "cpu_baseline": cpu,
        "ckks_error_vs_fp64": ckks_err,
        "rotations": rot,
        "bfv_mixed_rings": bfv4,
        "serving": serving,
    }
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
