"""The autotuner on the GPU: it times the kernel variants on HMult+relin,
applies a valid choice to the context, stores it in the cache, a context
created with PHX_TUNE_CACHE naming that cache applies it, and results under
every variant it may pick stay bit-exact against the oracle."""
import os

import numpy as np
import pytest

import oracle_bind as ob
import pyphantom as ph
from gpu_helpers import make_ctx, rand_ct, rand_key, say, upload_ct

pytestmark = pytest.mark.gpu

BITS = [60] + [50] * 6 + [60, 60]


def check_mul_relin(ctx, octx, mods, seed):
    N = ctx.N
    Ql = ctx.size_Q
    rng = np.random.default_rng(seed)
    a, b = rand_ct(rng, mods[:Ql], N), rand_ct(rng, mods[:Ql], N)
    key = rand_key(rng, mods, N, ctx.dnum)
    rk = ph.relin_key(ctx, host=key)
    c1 = ph.ciphertext(ctx, 1, 2)
    c1.upload(a)
    c2 = ph.ciphertext(ctx, 1, 2)
    c2.upload(b)
    ph.multiply_and_relin_inplace(ctx, c1, c2, rk)
    assert np.array_equal(c1.download(N), octx.multiply_and_relin(Ql, a, b, key))


def test_autotune_applies_and_caches(tmp_path):
    ob.set_threads(8)
    N = 1 << 14
    say("case: autotune N=2^14 L=9")
    ctx, octx, mods = make_ctx(N, BITS, 2)
    assert ctx.tuning() == {k: -1 for k in ph.TUNE_KEYS}
    path = str(tmp_path / "tune.json")
    rep = ctx.autotune(path)
    say(f"autotune: {rep}")
    ch = rep["choice"]
    assert ch["ntt_fused"] in (0, 1, 2, 4) and ch["bconv_mfma"] in (0, 1)
    assert ch["ntt_chunk_limbs"] >= 0 and ch["ntt_fused_min_limbs"] >= 0
    assert ch["ntt_cols_threads"] in (0, 256, 512) and ch["ntt_rows_threads"] in (128, 256, 512)
    assert ch["bconv_wgs_per_cu"] in (-1, 0, 1, 2)
    assert ch["bconv_fused"] in (-1, 0, 1)
    assert rep["trials"] >= 12 and rep["best_us"] > 0 and rep["default_us"] > 0
    # the report lists every candidate with its timing, each axis among them
    assert len(rep["timings"]) == rep["trials"]
    for axis in ("ntt_rows_threads=128", "ntt_rows_threads=512", "ntt_cols_threads=512", "bconv_mfma=0",
                 "bconv_wgs_per_cu=1", "ntt_fused=2", "ntt_chunk_limbs=128", "bconv_fused=0"):
        assert any(axis in t["choice"] for t in rep["timings"]), axis
    # the winner is no slower than the default beyond timing noise
    assert rep["best_us"] <= 1.1 * rep["default_us"]
    assert ctx.tuning() == ch
    assert ph.tune_cache_lookup(path, ph.SCHEME_CKKS, N, ctx.size_Q, ctx.size_P) == ch
    check_mul_relin(ctx, octx, mods, 1)
    # a new context of the same shape picks the entry up at creation
    os.environ["PHX_TUNE_CACHE"] = path
    try:
        ctx2, _, _ = make_ctx(N, BITS, 2)
    finally:
        del os.environ["PHX_TUNE_CACHE"]
    assert ctx2.tuning() == ch
    ctx3, _, _ = make_ctx(N, BITS, 2)
    assert ctx3.tuning()["ntt_fused"] == -1
    assert ctx3.tune_load(path) and ctx3.tuning() == ch


@pytest.mark.parametrize("choice", [
    dict(ntt_fused=0, ntt_fused_min_limbs=0, ntt_chunk_limbs=128, bconv_mfma=1),
    dict(ntt_fused=1, ntt_fused_min_limbs=0, ntt_chunk_limbs=0, bconv_mfma=0),
    dict(ntt_fused=4, ntt_fused_min_limbs=0, ntt_chunk_limbs=0, bconv_mfma=1),
    dict(ntt_fused=0, ntt_cols_threads=512, ntt_rows_threads=128, bconv_mfma=1, bconv_wgs_per_cu=1),
    dict(ntt_fused=0, ntt_cols_threads=256, ntt_rows_threads=512, bconv_mfma=1, bconv_wgs_per_cu=2),
    dict(ntt_fused=0, ntt_rows_threads=128, bconv_mfma=1, bconv_fused=1),
    dict(ntt_fused=0, ntt_rows_threads=512, bconv_mfma=1, bconv_fused=1),
    dict(ntt_fused=0, bconv_mfma=1, bconv_fused=0),
])
def test_per_context_choice_bit_exact(choice):
    """a context's own choice drives its launches (process-wide switches untouched),
    plain and in a batch plan over two levels"""
    N = 1 << 14
    say(f"case: per-context tuning {choice}")
    ctx, octx, mods = make_ctx(N, BITS, 2)
    ctx.set_tuning(**choice)
    before = {k: ph.tune_get(k) for k in ph.TUNE_KEYS}
    check_mul_relin(ctx, octx, mods, 2)
    x = np.stack([np.random.default_rng(3).integers(0, q, N, dtype=np.uint64) for q in mods])
    buf = ph.DeviceBuffer.from_array(x)
    ph.ntt_forward(ctx, buf, len(mods), 0)
    ph.device_synchronize()
    assert np.array_equal(buf.to_array(x.shape), octx.ntt_forward(x, 0))
    ph.ntt_inverse(ctx, buf, len(mods), 0)
    ph.device_synchronize()
    assert np.array_equal(buf.to_array(x.shape), x)
    rng = np.random.default_rng(4)
    key = rand_key(rng, mods, N, ctx.dnum)
    rk = ph.relin_key(ctx, host=key)
    items = [(1, ctx.limbs_at(1)), (3, ctx.limbs_at(3))]
    data = [(c, L, rand_ct(rng, mods[:L], N), rand_ct(rng, mods[:L], N)) for c, L in items]
    c1 = [upload_ct(ctx, a, c) for c, _, a, _ in data]
    c2 = [upload_ct(ctx, b, c) for c, _, _, b in data]
    bt = ph.batch.multiply_and_relin(ctx, c1, c2, rk)
    bt.run()
    ph.device_synchronize()
    for i, (c, L, a, b) in enumerate(data):
        assert np.array_equal(c1[i].download(N), octx.multiply_and_relin(L, a, b, key)), i
    assert {k: ph.tune_get(k) for k in before} == before


@pytest.mark.parametrize("logN", [11, 12, 13, 15, 16, 17])
@pytest.mark.parametrize("cols,rows", [(256, 128), (512, 256), (512, 512), (0, 256)])
def test_tile_shapes_every_degree(logN, cols, rows):
    """every tile shape at every ring degree (shapes whose tile exceeds a limb fall
    back to a smaller one): forward and inverse NTT bit-exact against the oracle"""
    N = 1 << logN
    say(f"case: NTT tile shape cols={cols} rows={rows} N=2^{logN}")
    ctx, octx, mods = make_ctx(N, [60, 50, 60], 1)
    ctx.set_tuning(ntt_cols_threads=cols, ntt_rows_threads=rows)
    x = np.stack([np.random.default_rng(logN).integers(0, q, N, dtype=np.uint64) for q in mods])
    buf = ph.DeviceBuffer.from_array(x)
    ph.ntt_forward(ctx, buf, len(mods), 0)
    ph.device_synchronize()
    assert np.array_equal(buf.to_array(x.shape), octx.ntt_forward(x, 0))
    ph.ntt_inverse(ctx, buf, len(mods), 0)
    ph.device_synchronize()
    assert np.array_equal(buf.to_array(x.shape), x)


def test_set_tuning_rejects():
    ctx, _, _ = make_ctx(1 << 12, [60, 50, 60], 1)
    with pytest.raises(ValueError):
        ctx.set_tuning(ntt_fused=3)
    with pytest.raises(ValueError):
        ctx.set_tuning(bconv_mfma=2)
    with pytest.raises(ValueError):
        ctx.set_tuning(colour=1)
    with pytest.raises(ValueError):
        ctx.set_tuning(ntt_cols_threads=128)
    with pytest.raises(ValueError):
        ctx.set_tuning(ntt_rows_threads=1024)
    with pytest.raises(ValueError):
        ctx.set_tuning(bconv_wgs_per_cu=3)
    with pytest.raises(ValueError):
        ctx.set_tuning(bconv_fused=2)


def test_autotune_bfv_leveled_context():
    """a BFV hps_overq_leveled context has no batch plans: the autotuner times the
    plain multiply_and_relin only and still applies a valid choice"""
    N = 1 << 13
    say("case: autotune bfv hps_overq_leveled")
    mods = ph.coeff_modulus_create(N, [54] * 4 + [60])
    parms = ph.EncryptionParameters(ph.SCHEME_BFV)
    parms.set_poly_modulus_degree(N)
    parms.set_coeff_modulus(mods)
    parms.set_special_modulus_size(1)
    parms.set_plain_modulus(65537)
    parms.set_mul_tech(ph.MUL_TECH_HPS_OVERQ_LEVELED)
    ctx = ph.context(parms)
    rep = ctx.autotune()
    assert rep["trials"] >= 2 and rep["best_us"] > 0
    assert set(ctx.tuning()) == set(ph.TUNE_KEYS)
