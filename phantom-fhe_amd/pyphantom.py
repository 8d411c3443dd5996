"""pyphantom -- Python binding of libphantom_amd.so over its C ABI.

Mirrors the names of the reference's pybind11 module (python/src/binding.cu:
context, ciphertext, multiply_and_relin, relinearize, rescale_to_next,
apply_galois, rotate) and the reference's exception types: a failed call
raises ValueError (std::invalid_argument), RuntimeError otherwise.

The module loads the in-tree HIP library and fails loudly when it is absent;
there is no CPU fallback.
"""
import ctypes
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libphantom_amd.so")

SCHEME_BGV, SCHEME_BFV, SCHEME_CKKS = 1, 2, 3
MUL_TECH_NONE, MUL_TECH_BEHZ, MUL_TECH_HPS, MUL_TECH_HPS_OVERQ, MUL_TECH_HPS_OVERQ_LEVELED = 0, 1, 2, 3, 4

_u64p = ctypes.POINTER(ctypes.c_uint64)
_u32p = ctypes.POINTER(ctypes.c_uint32)
_vp = ctypes.c_void_p
_sz = ctypes.c_size_t
_dp = ctypes.POINTER(ctypes.c_double)
_lib = None


class PhxError(RuntimeError):
    pass


class _Params(ctypes.Structure):
    _fields_ = [
        ("scheme", ctypes.c_int),
        ("mul_tech", ctypes.c_int),
        ("poly_modulus_degree", ctypes.c_uint64),
        ("coeff_modulus", _u64p),
        ("coeff_modulus_size", _sz),
        ("special_modulus_size", _sz),
        ("plain_modulus", ctypes.c_uint64),
        ("galois_elts", _u32p),
        ("galois_elts_size", _sz),
    ]


_SIGS = {
    "phx_last_error": (ctypes.c_char_p, []),
    "phx_coeff_modulus_create": (ctypes.c_int, [ctypes.c_uint64, ctypes.POINTER(ctypes.c_int), _sz, _u64p]),
    "phx_galois_elt_from_step": (ctypes.c_int, [ctypes.c_int, ctypes.c_uint64, _u32p]),
    "phx_context_create": (ctypes.c_int, [ctypes.POINTER(_Params), ctypes.POINTER(_vp)]),
    "phx_context_destroy": (ctypes.c_int, [_vp]),
    "phx_context_info": (ctypes.c_int, [_vp, _u64p, ctypes.POINTER(_sz), ctypes.POINTER(_sz), ctypes.POINTER(_sz)]),
    "phx_context_limbs_at": (ctypes.c_int, [_vp, _sz, ctypes.POINTER(_sz)]),
    "phx_ciphertext_create": (ctypes.c_int, [ctypes.POINTER(_vp)]),
    "phx_ciphertext_destroy": (ctypes.c_int, [_vp]),
    "phx_ciphertext_resize": (ctypes.c_int, [_vp, _vp, _sz, _sz, _vp]),
    "phx_ciphertext_upload": (ctypes.c_int, [_vp, _u64p, _sz, _vp]),
    "phx_ciphertext_download": (ctypes.c_int, [_vp, _u64p, _sz, _vp]),
    "phx_ciphertext_copy": (ctypes.c_int, [_vp, _vp, _vp, _vp]),
    "phx_ciphertext_info": (ctypes.c_int, [_vp, ctypes.POINTER(_sz), ctypes.POINTER(_sz), ctypes.POINTER(_sz),
                                           ctypes.POINTER(ctypes.c_double), ctypes.POINTER(ctypes.c_int)]),
    "phx_ciphertext_set_scale": (ctypes.c_int, [_vp, ctypes.c_double]),
    "phx_ciphertext_data": (ctypes.c_int, [_vp, ctypes.POINTER(_u64p)]),
    "phx_secret_key_create": (ctypes.c_int, [_vp, ctypes.c_uint64, ctypes.POINTER(_vp)]),
    "phx_secret_key_create_from_coeffs": (ctypes.c_int, [_vp, ctypes.POINTER(ctypes.c_int8), ctypes.POINTER(_vp)]),
    "phx_secret_key_download": (ctypes.c_int, [_vp, _vp, _u64p]),
    "phx_secret_key_destroy": (ctypes.c_int, [_vp]),
    "phx_relin_key_create": (ctypes.c_int, [_vp, _vp, ctypes.c_uint64, ctypes.POINTER(_vp)]),
    "phx_relin_key_create_from_host": (ctypes.c_int, [_vp, _u64p, _sz, ctypes.POINTER(_vp)]),
    "phx_relin_key_download": (ctypes.c_int, [_vp, _vp, _u64p]),
    "phx_relin_key_destroy": (ctypes.c_int, [_vp]),
    "phx_galois_key_create": (ctypes.c_int, [_vp, _vp, _u32p, _sz, ctypes.c_uint64, ctypes.POINTER(_vp)]),
    "phx_galois_key_create_from_host": (ctypes.c_int, [_vp, _u32p, _sz, _u64p, _sz, ctypes.POINTER(_vp)]),
    "phx_galois_key_destroy": (ctypes.c_int, [_vp]),
    "phx_encrypt_symmetric": (ctypes.c_int, [_vp, _vp, _u64p, _sz, ctypes.c_uint64, _vp, _vp]),
    "phx_decrypt": (ctypes.c_int, [_vp, _vp, _vp, _u64p, _vp]),
    "phx_ciphertext_set_ntt_form": (ctypes.c_int, [_vp, ctypes.c_int]),
    "phx_hoisting_inplace": (ctypes.c_int, [_vp, _vp, _vp, ctypes.POINTER(ctypes.c_int), _sz, _vp]),
    "phx_ciphertext_save_size": (ctypes.c_int, [_vp, ctypes.POINTER(_sz)]),
    "phx_ciphertext_save": (ctypes.c_int, [_vp, _vp, _sz]),
    "phx_ciphertext_load": (ctypes.c_int, [_vp, _vp, _vp, _sz]),
    "phx_ciphertext_save_symmetric_size": (ctypes.c_int, [_vp, ctypes.POINTER(_sz)]),
    "phx_ciphertext_save_symmetric": (ctypes.c_int, [_vp, _vp, _vp, _sz]),
    "phx_ciphertext_load_symmetric": (ctypes.c_int, [_vp, _vp, _vp, _sz]),
    "phx_sample_uniform_poly": (ctypes.c_int, [_vp, ctypes.c_char_p, _sz, _vp, _vp]),
    "phx_encrypt_symmetric_seeded": (ctypes.c_int, [_vp, _vp, _u64p, _sz, ctypes.c_char_p, ctypes.c_uint64, _vp,
                                                    _vp]),
    "phx_secret_key_save_size": (ctypes.c_int, [_vp, _vp, ctypes.POINTER(_sz)]),
    "phx_secret_key_save": (ctypes.c_int, [_vp, _vp, _vp, _sz]),
    "phx_secret_key_load": (ctypes.c_int, [_vp, _vp, _sz, ctypes.POINTER(_vp)]),
    "phx_decrypt_bfv": (ctypes.c_int, [_vp, _vp, _vp, _u64p, _vp]),
    "phx_public_key_create": (ctypes.c_int, [_vp, _vp, ctypes.c_char_p, ctypes.c_uint64, ctypes.POINTER(_vp)]),
    "phx_public_key_destroy": (ctypes.c_int, [_vp]),
    "phx_public_key_download": (ctypes.c_int, [_vp, _vp, _u64p]),
    "phx_public_key_save_size": (ctypes.c_int, [_vp, ctypes.POINTER(_sz)]),
    "phx_public_key_save": (ctypes.c_int, [_vp, _vp, _sz]),
    "phx_public_key_load": (ctypes.c_int, [_vp, _vp, _sz, ctypes.POINTER(_vp)]),
    "phx_encrypt_asymmetric": (ctypes.c_int, [_vp, _vp, _u64p, _sz, ctypes.c_uint64, _vp, _vp]),
    "phx_encrypt_asymmetric_with_samples": (ctypes.c_int, [_vp, _vp, _u64p, _sz, ctypes.POINTER(ctypes.c_int8),
                                                           ctypes.POINTER(ctypes.c_int8),
                                                           ctypes.POINTER(ctypes.c_int8), ctypes.c_uint64, _vp, _vp]),
    "phx_relin_key_save_size": (ctypes.c_int, [_vp, _vp, ctypes.POINTER(_sz)]),
    "phx_relin_key_save": (ctypes.c_int, [_vp, _vp, _vp, _sz]),
    "phx_relin_key_load": (ctypes.c_int, [_vp, _vp, _sz, ctypes.POINTER(_vp)]),
    "phx_galois_key_save_size": (ctypes.c_int, [_vp, _vp, ctypes.POINTER(_sz)]),
    "phx_galois_key_save": (ctypes.c_int, [_vp, _vp, _vp, _sz]),
    "phx_galois_key_load": (ctypes.c_int, [_vp, _u32p, _sz, _vp, _sz, ctypes.POINTER(_vp)]),
    "phx_server_create": (ctypes.c_int, [ctypes.POINTER(_vp), ctypes.POINTER(_vp), ctypes.POINTER(_vp), _sz, _sz,
                                         _sz, _sz, _vp, ctypes.POINTER(_vp)]),
    "phx_server_now": (ctypes.c_int, [_vp, _dp]),
    "phx_server_submit": (ctypes.c_int, [_vp, _sz, _vp, ctypes.POINTER(ctypes.c_int), _sz, ctypes.c_double,
                                         _u64p]),
    "phx_server_step": (ctypes.c_int, [_vp, ctypes.POINTER(_sz)]),
    "phx_server_drain": (ctypes.c_int, [_vp]),
    "phx_server_status": (ctypes.c_int, [_vp, ctypes.c_uint64, ctypes.POINTER(ctypes.c_int)]),
    "phx_server_take": (ctypes.c_int, [_vp, ctypes.c_uint64, _vp, _dp]),
    "phx_server_stats_get": (ctypes.c_int, [_vp, _vp]),
    "phx_server_destroy": (ctypes.c_int, [_vp]),
    "phx_ckks_encode": (ctypes.c_int, [_vp, _dp, _dp, _sz, ctypes.c_double, _sz, _u64p]),
    "phx_ckks_decode": (ctypes.c_int, [_vp, _u64p, _sz, ctypes.c_double, _dp, _dp]),
    "phx_multiply_and_relin_inplace": (ctypes.c_int, [_vp, _vp, _vp, _vp, _vp]),
    "phx_multiply_and_relin_inplace_noalloc": (ctypes.c_int, [_vp, _vp, _vp, _vp, _vp]),
    "phx_batch_launches": (ctypes.c_int, [_vp, ctypes.POINTER(_sz)]),
    "phx_batch_branches": (ctypes.c_int, [_vp, ctypes.POINTER(_sz)]),
    "phx_ciphertext_clone": (ctypes.c_int, [_vp, ctypes.POINTER(_vp), _vp]),
    "phx_multiply_inplace": (ctypes.c_int, [_vp, _vp, _vp, _vp]),
    "phx_relinearize_inplace": (ctypes.c_int, [_vp, _vp, _vp, _vp]),
    "phx_rescale_to_next": (ctypes.c_int, [_vp, _vp, _vp, _vp]),
    "phx_mod_switch_to_next": (ctypes.c_int, [_vp, _vp, _vp, _vp]),
    "phx_mod_switch_to_inplace": (ctypes.c_int, [_vp, _vp, _sz, _vp]),
    "phx_apply_galois_inplace": (ctypes.c_int, [_vp, _vp, ctypes.c_uint32, _vp, _vp]),
    "phx_rotate_inplace": (ctypes.c_int, [_vp, _vp, ctypes.c_int, _vp, _vp]),
    "phx_keyswitch_inplace": (ctypes.c_int, [_vp, _vp, _u64p, _vp, ctypes.c_int, _vp]),
    "phx_ntt_forward": (ctypes.c_int, [_vp, _u64p, _sz, _sz, _vp]),
    "phx_ntt_inverse": (ctypes.c_int, [_vp, _u64p, _sz, _sz, _vp]),
    "phx_ntt_forward_batch": (ctypes.c_int, [_vp, _u64p, _sz, _sz, _sz, _sz, _vp]),
    "phx_ntt_inverse_batch": (ctypes.c_int, [_vp, _u64p, _sz, _sz, _sz, _sz, _vp]),
    "phx_modup": (ctypes.c_int, [_vp, _sz, _u64p, _u64p, _vp]),
    "phx_key_inner_product": (ctypes.c_int, [_vp, _sz, _u64p, _vp, _u64p, _vp]),
    "phx_moddown_from_ntt": (ctypes.c_int, [_vp, _sz, _u64p, _u64p, _sz, _vp]),
    "phx_malloc": (ctypes.c_int, [_sz, ctypes.POINTER(_vp)]),
    "phx_free": (ctypes.c_int, [_vp]),
    "phx_memcpy_htod": (ctypes.c_int, [_vp, _vp, _sz, _vp]),
    "phx_memcpy_dtoh": (ctypes.c_int, [_vp, _vp, _sz, _vp]),
    "phx_stream_synchronize": (ctypes.c_int, [_vp]),
    "phx_device_synchronize": (ctypes.c_int, []),
    "phx_stream_create": (ctypes.c_int, [ctypes.POINTER(_vp)]),
    "phx_stream_destroy": (ctypes.c_int, [_vp]),
    "phx_event_create": (ctypes.c_int, [ctypes.POINTER(_vp)]),
    "phx_event_record": (ctypes.c_int, [_vp, _vp]),
    "phx_event_elapsed_ms": (ctypes.c_int, [_vp, _vp, ctypes.POINTER(ctypes.c_float)]),
    "phx_event_destroy": (ctypes.c_int, [_vp]),
    "phx_device_count": (ctypes.c_int, [ctypes.POINTER(ctypes.c_int)]),
    "phx_copy_u64": (ctypes.c_int, [_u64p, _u64p, _sz, _vp]),
    "phx_batch_multiply_and_relin_create": (ctypes.c_int, [_vp, ctypes.POINTER(_vp), ctypes.POINTER(_vp), _sz, _vp,
                                                           ctypes.POINTER(_vp)]),
    "phx_batch_rotate_create": (ctypes.c_int, [_vp, ctypes.POINTER(_vp), ctypes.POINTER(ctypes.c_int), _sz, _vp,
                                               ctypes.POINTER(_vp)]),
    "phx_batch_join": (ctypes.c_int, [ctypes.POINTER(_vp), _sz, ctypes.POINTER(_vp)]),
    "phx_batch_run": (ctypes.c_int, [_vp, _vp]),
    "phx_batch_capture": (ctypes.c_int, [_vp, _vp]),
    "phx_batch_replay": (ctypes.c_int, [_vp, _vp]),
    "phx_batch_scratch_bytes": (ctypes.c_int, [_vp, ctypes.POINTER(_sz)]),
    "phx_batch_destroy": (ctypes.c_int, [_vp]),
    "phx_context_get_tuning": (ctypes.c_int, [_vp, ctypes.POINTER(ctypes.c_int)]),
    "phx_context_set_tuning": (ctypes.c_int, [_vp, ctypes.POINTER(ctypes.c_int)]),
    "phx_autotune": (ctypes.c_int, [_vp, ctypes.c_char_p, _vp, _vp]),
    "phx_tune_load": (ctypes.c_int, [_vp, ctypes.c_char_p, ctypes.POINTER(ctypes.c_int)]),
    "phx_tune_cache_store": (ctypes.c_int, [ctypes.c_char_p, ctypes.c_int, ctypes.c_uint64, _sz, _sz,
                                            ctypes.POINTER(ctypes.c_int), ctypes.c_double]),
    "phx_tune_cache_lookup": (ctypes.c_int, [ctypes.c_char_p, ctypes.c_int, ctypes.c_uint64, _sz, _sz,
                                             ctypes.POINTER(ctypes.c_int), ctypes.POINTER(ctypes.c_int)]),
    "phx_tune_set": (ctypes.c_int, [ctypes.c_char_p, ctypes.c_int]),
    "phx_tune_get": (ctypes.c_int, [ctypes.c_char_p, ctypes.POINTER(ctypes.c_int)]),
}


def exported_symbols():
    return list(_SIGS.keys())


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise ImportError(f"{LIB_PATH} is missing: build it with __graft_entry__.build() "
                              "(make -C phantom-fhe_amd); there is no CPU fallback")
        L = ctypes.CDLL(LIB_PATH)
        for name, (res, args) in _SIGS.items():
            f = getattr(L, name)
            f.restype = res
            f.argtypes = args
        _lib = L
    return _lib


def _check(rc):
    if rc != 0:
        msg = lib().phx_last_error().decode()
        if rc == -1:
            raise ValueError(msg)
        raise PhxError(msg)


def _p(a):
    assert a.dtype == np.uint64 and a.flags["C_CONTIGUOUS"]
    return a.ctypes.data_as(_u64p)


def _stream(s):
    return None if s is None else s


# ------------------------------------------------------------------ params
def coeff_modulus_create(N, bit_sizes):
    """CoeffModulus::Create (src/host/modulus.cu:80-111)."""
    b = (ctypes.c_int * len(bit_sizes))(*bit_sizes)
    out = np.zeros(len(bit_sizes), np.uint64)
    _check(lib().phx_coeff_modulus_create(N, b, len(bit_sizes), _p(out)))
    return [int(x) for x in out]


def galois_elt_from_step(step, N):
    out = ctypes.c_uint32()
    _check(lib().phx_galois_elt_from_step(step, N, ctypes.byref(out)))
    return out.value


class EncryptionParameters:
    def __init__(self, scheme=SCHEME_CKKS):
        self.scheme = scheme
        self.mul_tech = MUL_TECH_HPS if scheme == SCHEME_BFV else MUL_TECH_NONE
        self.poly_modulus_degree = 0
        self.coeff_modulus = []
        self.special_modulus_size = 1
        self.plain_modulus = 0
        self.galois_elts = []

    def set_poly_modulus_degree(self, n):
        self.poly_modulus_degree = n

    def set_coeff_modulus(self, m):
        self.coeff_modulus = [int(x) for x in m]

    def set_special_modulus_size(self, s):
        self.special_modulus_size = s

    def set_galois_elts(self, e):
        self.galois_elts = list(e)

    def set_mul_tech(self, t):
        self.mul_tech = t

    def set_plain_modulus(self, t):
        self.plain_modulus = t


class context:
    """PhantomContext (include/context.cuh:133-273)."""

    def __init__(self, parms):
        self._mods = np.array(parms.coeff_modulus, np.uint64)
        self._elts = np.array(parms.galois_elts or [0], np.uint32)
        p = _Params(parms.scheme, parms.mul_tech, parms.poly_modulus_degree, _p(self._mods), len(self._mods),
                    parms.special_modulus_size, parms.plain_modulus,
                    self._elts.ctypes.data_as(_u32p) if parms.galois_elts else None, len(parms.galois_elts))
        h = _vp()
        _check(lib().phx_context_create(ctypes.byref(p), ctypes.byref(h)))
        self._h = h
        self.coeff_modulus = [int(x) for x in parms.coeff_modulus]
        N = ctypes.c_uint64()
        sq, sp, dn = _sz(), _sz(), _sz()
        _check(lib().phx_context_info(h, ctypes.byref(N), ctypes.byref(sq), ctypes.byref(sp), ctypes.byref(dn)))
        self.N, self.size_Q, self.size_P, self.dnum = N.value, sq.value, sp.value, dn.value
        self.size_QP = self.size_Q + self.size_P

    def limbs_at(self, chain_index):
        out = _sz()
        _check(lib().phx_context_limbs_at(self._h, chain_index, ctypes.byref(out)))
        return out.value

    def tuning(self):
        """kernel variants of this context (-1 = the process-wide switch)"""
        c = (ctypes.c_int * len(TUNE_KEYS))()
        _check(lib().phx_context_get_tuning(self._h, c))
        return dict(zip(TUNE_KEYS, list(c)))

    def set_tuning(self, **choice):
        cur = self.tuning()
        for k, v in choice.items():
            if k not in cur:
                raise ValueError(f"unknown tuning key: {k}")
            cur[k] = int(v)
        c = (ctypes.c_int * len(TUNE_KEYS))(*[cur[k] for k in TUNE_KEYS])
        _check(lib().phx_context_set_tuning(self._h, c))

    def autotune(self, cache_path=None, stream=None):
        """times the kernel variants on HMult+relin at the top level, applies the
        fastest and (with cache_path) stores it; returns the report"""
        r = _TuneReport()
        _check(lib().phx_autotune(self._h, cache_path.encode() if cache_path else None, ctypes.byref(r),
                                  _stream(stream)))
        import json
        return dict(choice=dict(zip(TUNE_KEYS, list(r.choice))), default_us=r.default_us, best_us=r.best_us,
                    trials=r.trials, timings=json.loads(r.trials_json.decode() or "[]"))

    def tune_load(self, cache_path):
        found = ctypes.c_int()
        _check(lib().phx_tune_load(self._h, cache_path.encode(), ctypes.byref(found)))
        return bool(found.value)

    def __del__(self):
        if getattr(self, "_h", None) and _lib is not None:
            _lib.phx_context_destroy(self._h)
            self._h = None


class ciphertext:
    """PhantomCiphertext (include/ciphertext.h:7-308)."""

    def __init__(self, ctx=None, chain_index=None, size=2, stream=None):
        h = _vp()
        _check(lib().phx_ciphertext_create(ctypes.byref(h)))
        self._h = h
        if ctx is not None and chain_index is not None:
            self.resize(ctx, chain_index, size, stream)

    def resize(self, ctx, chain_index, size, stream=None):
        _check(lib().phx_ciphertext_resize(ctx._h, self._h, chain_index, size, _stream(stream)))

    def clone(self, stream=None):
        """deep copy (the reference's copy constructor)"""
        c = ciphertext.__new__(ciphertext)
        h = _vp()
        _check(lib().phx_ciphertext_clone(self._h, ctypes.byref(h), _stream(stream)))
        c._h = h
        return c

    def info(self):
        c, s, L = _sz(), _sz(), _sz()
        sc = ctypes.c_double()
        ntt = ctypes.c_int()
        _check(lib().phx_ciphertext_info(self._h, ctypes.byref(c), ctypes.byref(s), ctypes.byref(L),
                                         ctypes.byref(sc), ctypes.byref(ntt)))
        return dict(chain_index=c.value, size=s.value, coeff_modulus_size=L.value, scale=sc.value,
                    is_ntt_form=bool(ntt.value))

    def chain_index(self):
        return self.info()["chain_index"]

    def size(self):
        return self.info()["size"]

    def coeff_modulus_size(self):
        return self.info()["coeff_modulus_size"]

    def scale(self):
        return self.info()["scale"]

    def set_scale(self, v):
        _check(lib().phx_ciphertext_set_scale(self._h, v))

    def save(self):
        """PhantomCiphertext::save byte stream"""
        n = _sz()
        _check(lib().phx_ciphertext_save_size(self._h, ctypes.byref(n)))
        buf = ctypes.create_string_buffer(n.value)
        _check(lib().phx_ciphertext_save(self._h, buf, n.value))
        return buf.raw

    @staticmethod
    def load(ctx, data):
        ct = ciphertext()
        _check(lib().phx_ciphertext_load(ctx._h, ct._h, data, len(data)))
        return ct

    def save_symmetric(self, ctx):
        """PhantomCiphertext::save_symmetric byte stream (c0 + seed)"""
        n = _sz()
        _check(lib().phx_ciphertext_save_symmetric_size(self._h, ctypes.byref(n)))
        buf = ctypes.create_string_buffer(n.value)
        _check(lib().phx_ciphertext_save_symmetric(ctx._h, self._h, buf, n.value))
        return buf.raw

    @staticmethod
    def load_symmetric(ctx, data):
        ct = ciphertext()
        _check(lib().phx_ciphertext_load_symmetric(ctx._h, ct._h, data, len(data)))
        return ct

    def set_ntt_form(self, v):
        _check(lib().phx_ciphertext_set_ntt_form(self._h, int(bool(v))))

    def is_ntt_form(self):
        return bool(self.info()["is_ntt_form"])

    def upload(self, arr, stream=None):
        a = np.ascontiguousarray(arr, np.uint64)
        _check(lib().phx_ciphertext_upload(self._h, _p(a), a.size, _stream(stream)))

    def download(self, N, stream=None):
        i = self.info()
        out = np.zeros((i["size"], i["coeff_modulus_size"], N), np.uint64)
        _check(lib().phx_ciphertext_download(self._h, _p(out), out.size, _stream(stream)))
        return out

    def data_ptr(self):
        p = _u64p()
        _check(lib().phx_ciphertext_data(self._h, ctypes.byref(p)))
        return p

    def copy_from(self, ctx, other, stream=None):
        _check(lib().phx_ciphertext_copy(ctx._h, self._h, other._h, _stream(stream)))

    def __del__(self):
        if getattr(self, "_h", None) and _lib is not None:
            _lib.phx_ciphertext_destroy(self._h)
            self._h = None


class _ServerStats(ctypes.Structure):
    _fields_ = [(n, ctypes.c_uint64) for n in ("submitted", "admitted", "completed", "ticks", "ops")] + \
               [(n, ctypes.c_double) for n in ("p50_ms", "p99_ms", "mean_ms", "max_ms", "span_s", "req_per_s",
                                               "ops_per_s")] + \
               [(n, ctypes.c_uint64) for n in ("peak_reserved_bytes", "budget_bytes", "peak_active")]


class server:
    """Continuous-batching request loop (phx_server_*): request = ciphertext and
    rotation steps; link i is rotate(rescale(ct * ct relinearised), steps[i])."""

    def __init__(self, lanes, mem_budget, max_active=256, depth=2, stream=None):
        n = len(lanes)
        ctxs = (_vp * n)(*[c._h for c, _, _ in lanes])
        rks = (_vp * n)(*[r._h for _, r, _ in lanes])
        gks = (_vp * n)(*[g._h for _, _, g in lanes])
        self._keep = lanes
        self.stream = stream if stream is not None else stream_create()
        h = _vp()
        _check(lib().phx_server_create(ctxs, rks, gks, n, int(mem_budget), int(max_active), int(depth),
                                       self.stream, ctypes.byref(h)))
        self._h = h

    def now(self):
        t = ctypes.c_double()
        _check(lib().phx_server_now(self._h, ctypes.byref(t)))
        return t.value

    def submit(self, lane, ct, steps, arrival=-1.0):
        st = (ctypes.c_int * len(steps))(*steps)
        rid = ctypes.c_uint64()
        _check(lib().phx_server_submit(self._h, lane, ct._h, st, len(steps), float(arrival), ctypes.byref(rid)))
        return rid.value

    def step(self):
        n = _sz()
        _check(lib().phx_server_step(self._h, ctypes.byref(n)))
        return n.value

    def drain(self):
        _check(lib().phx_server_drain(self._h))

    def status(self, rid):
        v = ctypes.c_int()
        _check(lib().phx_server_status(self._h, rid, ctypes.byref(v)))
        return v.value

    def take(self, rid):
        out = ciphertext()
        lat = ctypes.c_double()
        _check(lib().phx_server_take(self._h, rid, out._h, ctypes.byref(lat)))
        return out, lat.value

    def stats(self):
        st = _ServerStats()
        _check(lib().phx_server_stats_get(self._h, ctypes.byref(st)))
        return {n: getattr(st, n) for n, _ in st._fields_}

    def __del__(self):
        if getattr(self, "_h", None) and _lib is not None:
            _lib.phx_server_destroy(self._h)
            self._h = None


class ckks_encoder:
    """PhantomCKKSEncoder (include/ckks.h): slot j <-> m(zeta^(5^j)), host side."""

    def __init__(self, ctx):
        self._ctx = ctx

    def slot_count(self):
        return self._ctx.N // 2

    def encode(self, values, scale, chain_index=1):
        """values (complex or real, at most N/2) -> coefficient residues [Ql][N]."""
        v = np.asarray(values)
        re = np.ascontiguousarray(v.real, np.float64)
        im = np.ascontiguousarray(v.imag if np.iscomplexobj(v) else np.zeros(len(v)), np.float64)
        L = self._ctx.limbs_at(chain_index)
        out = np.zeros((L, self._ctx.N), np.uint64)
        _check(lib().phx_ckks_encode(self._ctx._h, re.ctypes.data_as(_dp), im.ctypes.data_as(_dp), len(v),
                                     float(scale), chain_index, _p(out)))
        return out

    def decode(self, coeffs, chain_index, scale):
        """coefficient residues [Ql][N] -> N/2 complex slot values."""
        a = np.ascontiguousarray(coeffs, np.uint64)
        n = self._ctx.N // 2
        re, im = np.zeros(n), np.zeros(n)
        _check(lib().phx_ckks_decode(self._ctx._h, _p(a), chain_index, float(scale), re.ctypes.data_as(_dp),
                                     im.ctypes.data_as(_dp)))
        return re + 1j * im


class secret_key:
    """PhantomSecretKey (include/secretkey.h): ternary secret, NTT form."""

    def __init__(self, ctx, seed=1, coeffs=None):
        h = _vp()
        if coeffs is not None:
            c = np.ascontiguousarray(coeffs, np.int8)
            _check(lib().phx_secret_key_create_from_coeffs(ctx._h, c.ctypes.data_as(ctypes.POINTER(ctypes.c_int8)),
                                                           ctypes.byref(h)))
        else:
            _check(lib().phx_secret_key_create(ctx._h, seed, ctypes.byref(h)))
        self._h = h
        self._ctx = ctx

    def ntt_form(self):
        out = np.zeros((self._ctx.size_QP, self._ctx.N), np.uint64)
        _check(lib().phx_secret_key_download(self._ctx._h, self._h, _p(out)))
        return out

    def gen_relinkey(self, ctx, seed=2):
        return relin_key(ctx, sk=self, seed=seed)

    def create_galois_keys(self, ctx, elts, seed=3):
        return galois_key(ctx, elts, sk=self, seed=seed)

    def encrypt_symmetric(self, ctx, plain_limbs, chain_index, seed=4, stream=None):
        a = np.ascontiguousarray(plain_limbs, np.uint64)
        ct = ciphertext()
        _check(lib().phx_encrypt_symmetric(ctx._h, self._h, _p(a), chain_index, seed, ct._h, _stream(stream)))
        return ct

    def encrypt_symmetric_seeded(self, ctx, plain_limbs, chain_index, seed_a, seed_e=5, stream=None):
        """c1 from the seeded uniform sampler (64-byte seed_a), kept for save_symmetric"""
        if len(seed_a) != 64:
            raise ValueError("the seed is 64 bytes")
        a = np.ascontiguousarray(plain_limbs, np.uint64)
        ct = ciphertext()
        _check(lib().phx_encrypt_symmetric_seeded(ctx._h, self._h, _p(a), chain_index, bytes(seed_a), seed_e, ct._h,
                                                  _stream(stream)))
        return ct

    def gen_publickey(self, ctx, seed_a=None, seed_e=6):
        """PhantomSecretKey::gen_publickey: pk = (-(a s + e), a) at the key level"""
        return public_key(ctx, self, seed_a=seed_a, seed_e=seed_e)

    def save(self):
        """PhantomSecretKey::save byte stream"""
        n = _sz()
        _check(lib().phx_secret_key_save_size(self._ctx._h, self._h, ctypes.byref(n)))
        buf = ctypes.create_string_buffer(n.value)
        _check(lib().phx_secret_key_save(self._ctx._h, self._h, buf, n.value))
        return buf.raw

    @classmethod
    def load(cls, ctx, data):
        k = cls.__new__(cls)
        h = _vp()
        _check(lib().phx_secret_key_load(ctx._h, data, len(data), ctypes.byref(h)))
        k._h, k._ctx = h, ctx
        return k

    def decrypt(self, ctx, ct, stream=None):
        L = ct.coeff_modulus_size()
        out = np.zeros((L, ctx.N), np.uint64)
        _check(lib().phx_decrypt(ctx._h, self._h, ct._h, _p(out), _stream(stream)))
        return out

    def decrypt_bfv(self, ctx, ct, stream=None):
        """PhantomSecretKey::decrypt of a BFV ciphertext: the plaintext's N coefficients"""
        out = np.zeros(ctx.N, np.uint64)
        _check(lib().phx_decrypt_bfv(ctx._h, self._h, ct._h, _p(out), _stream(stream)))
        return out

    def __del__(self):
        if getattr(self, "_h", None) and _lib is not None:
            _lib.phx_secret_key_destroy(self._h)
            self._h = None


class public_key:
    """PhantomPublicKey (include/secretkey.h:17-100): (-(a s + e), a) over QP, NTT form."""

    def __init__(self, ctx, sk, seed_a=None, seed_e=6):
        if seed_a is not None and len(seed_a) != 64:
            raise ValueError("the seed is 64 bytes")
        h = _vp()
        _check(lib().phx_public_key_create(ctx._h, sk._h, None if seed_a is None else bytes(seed_a), seed_e,
                                           ctypes.byref(h)))
        self._h, self._ctx = h, ctx

    def data(self):
        out = np.zeros((2, self._ctx.size_QP, self._ctx.N), np.uint64)
        _check(lib().phx_public_key_download(self._ctx._h, self._h, _p(out)))
        return out

    def encrypt_asymmetric(self, ctx, plain, chain_index=1, seed=7, u=None, e=None, scale=None, stream=None):
        """PhantomPublicKey::encrypt_asymmetric. plain: CKKS coefficient residues [size_Q][N]
        (ckks_encoder.encode), BFV N coefficients mod t. u, e = (e0, e1): optional int8 samples.
        scale: the CKKS plaintext's scale, carried by the ciphertext."""
        a = np.ascontiguousarray(plain, np.uint64)
        ct = ciphertext()
        if u is None:
            _check(lib().phx_encrypt_asymmetric(ctx._h, self._h, _p(a), chain_index, seed, ct._h, _stream(stream)))
        else:
            i8 = ctypes.POINTER(ctypes.c_int8)
            uu = np.ascontiguousarray(u, np.int8)
            e0 = np.ascontiguousarray(e[0], np.int8)
            e1 = np.ascontiguousarray(e[1], np.int8)
            for v in (uu, e0, e1):
                if v.size != ctx.N:
                    raise ValueError("samples are N int8 coefficients each")
            _check(lib().phx_encrypt_asymmetric_with_samples(ctx._h, self._h, _p(a), chain_index,
                                                             uu.ctypes.data_as(i8), e0.ctypes.data_as(i8),
                                                             e1.ctypes.data_as(i8), seed, ct._h, _stream(stream)))
        if scale is not None:
            ct.set_scale(scale)
        return ct

    def save(self):
        """PhantomPublicKey::save byte stream"""
        n = _sz()
        _check(lib().phx_public_key_save_size(self._h, ctypes.byref(n)))
        buf = ctypes.create_string_buffer(n.value)
        _check(lib().phx_public_key_save(self._h, buf, n.value))
        return buf.raw

    @classmethod
    def load(cls, ctx, data):
        k = cls.__new__(cls)
        h = _vp()
        _check(lib().phx_public_key_load(ctx._h, data, len(data), ctypes.byref(h)))
        k._h, k._ctx = h, ctx
        return k

    def __del__(self):
        if getattr(self, "_h", None) and _lib is not None:
            _lib.phx_public_key_destroy(self._h)
            self._h = None


class relin_key:
    """PhantomRelinKey: [dnum][2][QP][N]."""

    def __init__(self, ctx, sk=None, seed=2, host=None):
        h = _vp()
        if host is not None:
            a = np.ascontiguousarray(host, np.uint64)
            _check(lib().phx_relin_key_create_from_host(ctx._h, _p(a), a.shape[0], ctypes.byref(h)))
            self.dnum = a.shape[0]
        else:
            _check(lib().phx_relin_key_create(ctx._h, sk._h, seed, ctypes.byref(h)))
            self.dnum = ctx.dnum
        self._h = h
        self._ctx = ctx

    def download(self):
        out = np.zeros((self.dnum, 2, self._ctx.size_QP, self._ctx.N), np.uint64)
        _check(lib().phx_relin_key_download(self._ctx._h, self._h, _p(out)))
        return out

    def save(self):
        """PhantomRelinKey::save byte stream"""
        n = _sz()
        _check(lib().phx_relin_key_save_size(self._ctx._h, self._h, ctypes.byref(n)))
        buf = ctypes.create_string_buffer(n.value)
        _check(lib().phx_relin_key_save(self._ctx._h, self._h, buf, n.value))
        return buf.raw

    @classmethod
    def load(cls, ctx, data):
        k = cls.__new__(cls)
        h = _vp()
        _check(lib().phx_relin_key_load(ctx._h, data, len(data), ctypes.byref(h)))
        k._h, k._ctx = h, ctx
        k.dnum = int.from_bytes(data[:8], "little")
        return k

    def __del__(self):
        if getattr(self, "_h", None) and _lib is not None:
            _lib.phx_relin_key_destroy(self._h)
            self._h = None


class galois_key:
    """PhantomGaloisKey: one key-switching key per Galois element."""

    def __init__(self, ctx, elts, sk=None, seed=3, host=None):
        e = np.array(elts, np.uint32)
        h = _vp()
        if host is not None:
            a = np.ascontiguousarray(host, np.uint64)
            _check(lib().phx_galois_key_create_from_host(ctx._h, e.ctypes.data_as(_u32p), len(e), _p(a),
                                                         a.shape[1], ctypes.byref(h)))
        else:
            _check(lib().phx_galois_key_create(ctx._h, sk._h, e.ctypes.data_as(_u32p), len(e), seed,
                                               ctypes.byref(h)))
        self._h = h
        self._ctx = ctx
        self.elts = [int(x) for x in elts]

    def save(self):
        """PhantomGaloisKey::save byte stream"""
        n = _sz()
        _check(lib().phx_galois_key_save_size(self._ctx._h, self._h, ctypes.byref(n)))
        buf = ctypes.create_string_buffer(n.value)
        _check(lib().phx_galois_key_save(self._ctx._h, self._h, buf, n.value))
        return buf.raw

    @classmethod
    def load(cls, ctx, elts, data):
        g = cls.__new__(cls)
        e = np.array(elts, np.uint32)
        h = _vp()
        _check(lib().phx_galois_key_load(ctx._h, e.ctypes.data_as(_u32p), len(e), data, len(data), ctypes.byref(h)))
        g._h, g._ctx, g.elts = h, ctx, [int(x) for x in elts]
        return g

    def __del__(self):
        if getattr(self, "_h", None) and _lib is not None:
            _lib.phx_galois_key_destroy(self._h)
            self._h = None


# ------------------------------------------------------------------ evaluator
def multiply_and_relin_inplace(ctx, ct1, ct2, rlk, stream=None):
    _check(lib().phx_multiply_and_relin_inplace(ctx._h, ct1._h, ct2._h, rlk._h, _stream(stream)))


def multiply_and_relin_inplace_noalloc(ctx, ct1, ct2, rlk, stream=None):
    """CKKS multiply_and_relin_inplace in the library's static device scratch (no allocation)"""
    _check(lib().phx_multiply_and_relin_inplace_noalloc(ctx._h, ct1._h, ct2._h, rlk._h, _stream(stream)))


def multiply_and_relin(ctx, ct1, ct2, rlk, stream=None):
    out = ciphertext()
    out.copy_from(ctx, ct1, stream)
    multiply_and_relin_inplace(ctx, out, ct2, rlk, stream)
    return out


def multiply_inplace(ctx, ct1, ct2, stream=None):
    _check(lib().phx_multiply_inplace(ctx._h, ct1._h, ct2._h, _stream(stream)))


def relinearize_inplace(ctx, ct, rlk, stream=None):
    _check(lib().phx_relinearize_inplace(ctx._h, ct._h, rlk._h, _stream(stream)))


def rescale_to_next(ctx, ct, stream=None):
    out = ciphertext()
    _check(lib().phx_rescale_to_next(ctx._h, ct._h, out._h, _stream(stream)))
    return out


def mod_switch_to_next(ctx, ct, stream=None):
    """the ciphertext at the next level: CKKS drops the last limb (NTT form, scale
    kept); BFV divides by the last prime (coefficient form, floor division as the
    reference's divide_and_round_q_last)"""
    out = ciphertext()
    _check(lib().phx_mod_switch_to_next(ctx._h, ct._h, out._h, _stream(stream)))
    return out


def mod_switch_to_next_inplace(ctx, ct, stream=None):
    _check(lib().phx_mod_switch_to_next(ctx._h, ct._h, ct._h, _stream(stream)))


def mod_switch_to_inplace(ctx, ct, chain_index, stream=None):
    _check(lib().phx_mod_switch_to_inplace(ctx._h, ct._h, chain_index, _stream(stream)))


def rescale_to_next_inplace(ctx, ct, stream=None):
    _check(lib().phx_rescale_to_next(ctx._h, ct._h, ct._h, _stream(stream)))


def apply_galois_inplace(ctx, ct, elt, gk, stream=None):
    _check(lib().phx_apply_galois_inplace(ctx._h, ct._h, elt, gk._h, _stream(stream)))


def rotate_inplace(ctx, ct, step, gk, stream=None):
    _check(lib().phx_rotate_inplace(ctx._h, ct._h, step, gk._h, _stream(stream)))


def hoisting_inplace(ctx, ct, gk, steps, stream=None):
    """hoisting_inplace: ct <- sum of the rotations of ct by steps (one shared modup)"""
    st = (ctypes.c_int * len(steps))(*steps)
    _check(lib().phx_hoisting_inplace(ctx._h, ct._h, gk._h, st, len(steps), _stream(stream)))


def keyswitch_inplace(ctx, ct, c2_dev, key, is_relin=True, stream=None):
    _check(lib().phx_keyswitch_inplace(ctx._h, ct._h, c2_dev, key._h, int(is_relin), _stream(stream)))


# ------------------------------------------------------------------ batches
class batch:
    """Batched hot path over a fixed list of ciphertexts (phx_batch_*)."""

    def __init__(self, h, keep):
        self._h = h
        self._keep = keep   # keep the ciphertexts/keys alive while the batch lives

    @classmethod
    def multiply_and_relin(cls, ctx, ct1s, ct2s, rlk):
        a = (_vp * len(ct1s))(*[c._h for c in ct1s])
        b = (_vp * len(ct2s))(*[c._h for c in ct2s])
        h = _vp()
        _check(lib().phx_batch_multiply_and_relin_create(ctx._h, a, b, len(ct1s), rlk._h, ctypes.byref(h)))
        return cls(h, (ctx, list(ct1s), list(ct2s), rlk))

    @classmethod
    def rotate(cls, ctx, cts, steps, gk):
        a = (_vp * len(cts))(*[c._h for c in cts])
        st = (ctypes.c_int * len(steps))(*steps)
        h = _vp()
        _check(lib().phx_batch_rotate_create(ctx._h, a, st, len(cts), gk._h, ctypes.byref(h)))
        return cls(h, (ctx, list(cts), gk))

    @classmethod
    def join(cls, parts):
        """one batch running `parts` (any contexts; joined batches are flattened) on
        concurrent branches; the parts are consumed"""
        arr = (_vp * len(parts))(*[p._h for p in parts])
        h = _vp()
        _check(lib().phx_batch_join(arr, len(parts), ctypes.byref(h)))
        keep = []
        for p in parts:
            keep.append(p._keep)
            p._h = None
        return cls(h, tuple(keep))

    def run(self, stream=None):
        _check(lib().phx_batch_run(self._h, _stream(stream)))

    def capture(self, stream):
        _check(lib().phx_batch_capture(self._h, stream))

    def replay(self, stream=None):
        _check(lib().phx_batch_replay(self._h, _stream(stream)))

    def launches(self):
        """kernel launches of one run (kernel nodes of the captured graph)"""
        n = _sz()
        _check(lib().phx_batch_launches(self._h, ctypes.byref(n)))
        return n.value

    def branches(self):
        """concurrent branches the batch runs on (1: the caller's stream alone)"""
        n = _sz()
        _check(lib().phx_batch_branches(self._h, ctypes.byref(n)))
        return n.value

    def scratch_bytes(self):
        v = _sz()
        _check(lib().phx_batch_scratch_bytes(self._h, ctypes.byref(v)))
        return v.value

    def __del__(self):
        if getattr(self, "_h", None) and _lib is not None:
            _lib.phx_batch_destroy(self._h)
            self._h = None


# ------------------------------------------------------------------ device buffers
class DeviceBuffer:
    """Plain device allocation of u64 elements (hipMalloc)."""

    def __init__(self, n_elems):
        h = _vp()
        _check(lib().phx_malloc(max(1, n_elems) * 8, ctypes.byref(h)))
        self._h = h
        self.n = n_elems

    @classmethod
    def from_array(cls, arr, stream=None):
        a = np.ascontiguousarray(arr, np.uint64)
        b = cls(a.size)
        _check(lib().phx_memcpy_htod(b._h, a.ctypes.data_as(_vp), a.size * 8, _stream(stream)))
        return b

    def ptr(self):
        return ctypes.cast(self._h, _u64p)

    def to_array(self, shape=None, stream=None):
        out = np.zeros(self.n, np.uint64)
        _check(lib().phx_memcpy_dtoh(out.ctypes.data_as(_vp), self._h, self.n * 8, _stream(stream)))
        return out.reshape(shape) if shape is not None else out

    def __del__(self):
        if getattr(self, "_h", None) and _lib is not None:
            _lib.phx_device_synchronize()
            _lib.phx_free(self._h)
            self._h = None


def ntt_forward(ctx, buf, L, start=0, stream=None):
    _check(lib().phx_ntt_forward(ctx._h, buf.ptr(), L, start, _stream(stream)))


def ntt_inverse(ctx, buf, L, start=0, stream=None):
    _check(lib().phx_ntt_inverse(ctx._h, buf.ptr(), L, start, _stream(stream)))


def ntt_forward_batch(ctx, buf, L, start, npolys, poly_stride, stream=None):
    _check(lib().phx_ntt_forward_batch(ctx._h, buf.ptr(), L, start, npolys, poly_stride, _stream(stream)))


def ntt_inverse_batch(ctx, buf, L, start, npolys, poly_stride, stream=None):
    _check(lib().phx_ntt_inverse_batch(ctx._h, buf.ptr(), L, start, npolys, poly_stride, _stream(stream)))


def stream_create():
    h = _vp()
    _check(lib().phx_stream_create(ctypes.byref(h)))
    return h


def modup(ctx, chain_index, c2_buf, t_buf, stream=None):
    _check(lib().phx_modup(ctx._h, chain_index, c2_buf.ptr(), t_buf.ptr(), _stream(stream)))


def key_inner_product(ctx, chain_index, t_buf, key, cx_buf, stream=None):
    _check(lib().phx_key_inner_product(ctx._h, chain_index, t_buf.ptr(), key._h, cx_buf.ptr(), _stream(stream)))


def moddown_from_ntt(ctx, chain_index, cx_buf, out_buf, npolys=1, stream=None):
    _check(lib().phx_moddown_from_ntt(ctx._h, chain_index, cx_buf.ptr(), out_buf.ptr(), npolys, _stream(stream)))


def synchronize(stream=None):
    _check(lib().phx_stream_synchronize(_stream(stream)))


def device_synchronize():
    _check(lib().phx_device_synchronize())


class Event:
    def __init__(self):
        h = _vp()
        _check(lib().phx_event_create(ctypes.byref(h)))
        self._h = h

    def record(self, stream=None):
        _check(lib().phx_event_record(self._h, _stream(stream)))

    def elapsed_ms(self, end):
        ms = ctypes.c_float()
        _check(lib().phx_event_elapsed_ms(self._h, end._h, ctypes.byref(ms)))
        return ms.value

    def __del__(self):
        if getattr(self, "_h", None) and _lib is not None:
            _lib.phx_event_destroy(self._h)
            self._h = None


TUNE_KEYS = ("ntt_fused", "ntt_fused_min_limbs", "ntt_chunk_limbs", "bconv_mfma", "ntt_cols_threads",
             "ntt_rows_threads", "bconv_wgs_per_cu", "bconv_fused")


class _TuneReport(ctypes.Structure):
    _fields_ = [("choice", ctypes.c_int * len(TUNE_KEYS)), ("default_us", ctypes.c_double),
                ("best_us", ctypes.c_double), ("trials", ctypes.c_uint32), ("trials_json", ctypes.c_char * 8192)]


def tune_cache_store(path, scheme, N, size_Q, size_P, choice, us=0.0):
    """writes one autotuner cache entry (host only)"""
    c = (ctypes.c_int * len(TUNE_KEYS))(*[int(choice.get(k, -1)) for k in TUNE_KEYS])
    _check(lib().phx_tune_cache_store(path.encode(), scheme, N, size_Q, size_P, c, float(us)))


def tune_cache_lookup(path, scheme, N, size_Q, size_P):
    """the cached choice for a shape, or None (host only)"""
    c = (ctypes.c_int * len(TUNE_KEYS))()
    found = ctypes.c_int()
    _check(lib().phx_tune_cache_lookup(path.encode(), scheme, N, size_Q, size_P, c, ctypes.byref(found)))
    return dict(zip(TUNE_KEYS, list(c))) if found.value else None


def tune_set(key, value):
    _check(lib().phx_tune_set(key.encode(), int(value)))


def tune_get(key):
    v = ctypes.c_int()
    _check(lib().phx_tune_get(key.encode(), ctypes.byref(v)))
    return v.value


def sample_uniform_poly(ctx, seed, coeff_mod_size, out, stream=None):
    """out (DeviceBuffer, [coeff_mod_size][N]) <- the seeded uniform sampler for a 64-byte seed"""
    if len(seed) != 64:
        raise ValueError("the seed is 64 bytes")
    _check(lib().phx_sample_uniform_poly(ctx._h, bytes(seed), coeff_mod_size, out.ptr(), _stream(stream)))


def copy_u64(dst, src, n, stream=None):
    _check(lib().phx_copy_u64(dst.ptr(), src.ptr(), n, _stream(stream)))


def device_count():
    n = ctypes.c_int()
    rc = lib().phx_device_count(ctypes.byref(n))
    return n.value if rc == 0 else 0
