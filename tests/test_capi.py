"""CPU-side checks of the C ABI: the library loads, exports every symbol that
include/phantom_amd.h declares, and its host-only entry points agree with the
oracle. No GPU call is made here."""
import ctypes
import os
import re

import numpy as np
import pytest

import oracle_bind as ob
import pyphantom as ph

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_symbols():
    txt = open(os.path.join(ROOT, "include", "phantom_amd.h")).read()
    txt = re.sub(r"/\*.*?\*/", "", txt, flags=re.S)
    return sorted(set(re.findall(r"\b(phx_[a-z0-9_]+)\s*\(", txt)))


def test_header_declares_api():
    syms = declared_symbols()
    for must in ["phx_multiply_and_relin_inplace", "phx_rotate_inplace", "phx_keyswitch_inplace",
                 "phx_ntt_forward", "phx_ntt_inverse", "phx_modup", "phx_moddown_from_ntt",
                 "phx_context_create", "phx_rescale_to_next", "phx_apply_galois_inplace"]:
        assert must in syms


def test_library_exports_every_declared_symbol():
    L = ph.lib()
    for s in declared_symbols():
        assert hasattr(L, s), s
    assert set(declared_symbols()) == set(ph.exported_symbols())


def test_coeff_modulus_create_matches_oracle():
    for N, bits in [(4096, [60, 60, 60]), (65536, [60] + [50] * 34 + [60] * 12), (8192, [54] * 5)]:
        assert ph.coeff_modulus_create(N, bits) == ob.coeff_modulus_create(N, bits)


def test_galois_elt_matches_oracle():
    for N in (4096, 65536):
        for step in [0, 1, 2, 7, -1, -5, 100]:
            assert ph.galois_elt_from_step(step, N) == ob.galois_elt_from_step(step, N)


def test_invalid_arguments_raise_value_error():
    with pytest.raises(ValueError):
        ph.coeff_modulus_create(1000, [60])
    with pytest.raises(ValueError):
        ph.galois_elt_from_step(1 << 20, 4096)
    assert "step count too large" in ph.lib().phx_last_error().decode()


def test_tune_switches_roundtrip():
    assert ph.tune_get("bconv_mfma") in (0, 1)
    old = ph.tune_get("ntt_fp64")
    ph.tune_set("ntt_fp64", 0)
    assert ph.tune_get("ntt_fp64") == 0
    ph.tune_set("ntt_fp64", old)
    with pytest.raises(ValueError):
        ph.tune_set("no_such_knob", 1)


def test_tune_cache_roundtrip(tmp_path):
    """the autotuner's JSON cache: store, look up, replace one entry, keep others"""
    path = str(tmp_path / "tune.json")
    assert ph.tune_cache_lookup(path, ph.SCHEME_CKKS, 65536, 35, 12) is None
    a = dict(ntt_fused=0, ntt_fused_min_limbs=0, ntt_chunk_limbs=256, bconv_mfma=1, ntt_cols_threads=512,
             ntt_rows_threads=128, bconv_wgs_per_cu=2, bconv_fused=0)
    b = dict(ntt_fused=2, ntt_fused_min_limbs=64, ntt_chunk_limbs=0, bconv_mfma=0, ntt_cols_threads=-1,
             ntt_rows_threads=256, bconv_wgs_per_cu=0, bconv_fused=1)
    ph.tune_cache_store(path, ph.SCHEME_CKKS, 65536, 35, 12, a, 340.5)
    ph.tune_cache_store(path, ph.SCHEME_BFV, 8192, 4, 1, b, 12.0)
    assert ph.tune_cache_lookup(path, ph.SCHEME_CKKS, 65536, 35, 12) == a
    assert ph.tune_cache_lookup(path, ph.SCHEME_BFV, 8192, 4, 1) == b
    assert ph.tune_cache_lookup(path, ph.SCHEME_CKKS, 8192, 4, 1) is None
    assert ph.tune_cache_lookup(path, ph.SCHEME_CKKS, 65536, 35, 11) is None
    ph.tune_cache_store(path, ph.SCHEME_CKKS, 65536, 35, 12, b, 300.0)
    assert ph.tune_cache_lookup(path, ph.SCHEME_CKKS, 65536, 35, 12) == b
    assert ph.tune_cache_lookup(path, ph.SCHEME_BFV, 8192, 4, 1) == b
    import json
    doc = json.load(open(path))
    assert len(doc["entries"]) == 2
    assert {e["key"] for e in doc["entries"]} == {"scheme=ckks,N=65536,Q=35,P=12", "scheme=bfv,N=8192,Q=4,P=1"}
    # an entry written before the tile-shape axes existed reads them as -1
    with open(path, "w") as f:
        f.write('{"entries": [\n  {"key": "scheme=ckks,N=4096,Q=2,P=1", "ntt_fused": 1, "ntt_fused_min_limbs": 0, '
                '"ntt_chunk_limbs": 0, "bconv_mfma": 1, "hmult_relin_us": 10}\n]}\n')
    old = ph.tune_cache_lookup(path, ph.SCHEME_CKKS, 4096, 2, 1)
    assert old["ntt_fused"] == 1 and old["ntt_cols_threads"] == old["ntt_rows_threads"] == -1
    assert old["bconv_fused"] == -1
    with open(path, "w") as f:
        f.write('{"entries": [\n  {"key": "scheme=ckks,N=4096,Q=2,P=1", "ntt_fused": 1}\n]}\n')
    with pytest.raises(ValueError):
        ph.tune_cache_lookup(path, ph.SCHEME_CKKS, 4096, 2, 1)


def test_cpp_facade_header_compiles(tmp_path):
    """include/phantom.hpp (the reference's class names over the C ABI) compiles as a
    translation unit of its own with the host compiler"""
    import subprocess
    src = tmp_path / "tu.cpp"
    src.write_text('#include "phantom.hpp"\nint main() { phantom::EncryptionParameters p(phantom::scheme_type::ckks);'
                   ' (void)p; return 0; }\n')
    r = subprocess.run(["g++", "-std=c++17", "-fsyntax-only", "-Wall", "-Wextra", "-Werror",
                        "-I" + os.path.join(ROOT, "include"), "-I/opt/rocm/include", "-D__HIP_PLATFORM_AMD__",
                        str(src)], capture_output=True, text=True)
    assert r.returncode == 0, r.stderr
