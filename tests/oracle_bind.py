"""ctypes binding of the CPU oracle (oracle/liboracle.so).

Test infrastructure: only tests/, the smoke entry and the benchmark's CPU
baseline import this module. It loads the oracle as the checker.
"""
import ctypes
import os
import subprocess

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
ORACLE_DIR = os.path.join(ROOT, "oracle")
LIB_PATH = os.path.join(ORACLE_DIR, "liboracle.so")

_u64p = ctypes.POINTER(ctypes.c_uint64)
_u32p = ctypes.POINTER(ctypes.c_uint32)
_sz = ctypes.c_size_t
_lib = None


def build():
    subprocess.run(["make", "-s", "-C", ORACLE_DIR], check=True)


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            build()
        L = ctypes.CDLL(LIB_PATH)
        L.orc_ctx_create.restype = ctypes.c_void_p
        L.orc_ctx_create.argtypes = [ctypes.c_int, _u64p, _sz, _sz]
        L.orc_ctx_destroy.argtypes = [ctypes.c_void_p]
        L.orc_ntt_forward.argtypes = [ctypes.c_void_p, _u64p, _sz, _sz]
        L.orc_ntt_inverse.argtypes = [ctypes.c_void_p, _u64p, _sz, _sz]
        L.orc_get_primes.argtypes = [ctypes.c_uint64, ctypes.c_int, _sz, _u64p]
        L.orc_coeff_modulus_create.argtypes = [ctypes.c_uint64, ctypes.POINTER(ctypes.c_int), _sz, _u64p]
        L.orc_minimal_root.restype = ctypes.c_uint64
        L.orc_minimal_root.argtypes = [ctypes.c_uint64, ctypes.c_uint64]
        L.orc_is_prime.argtypes = [ctypes.c_uint64]
        L.orc_shoup.restype = ctypes.c_uint64
        L.orc_shoup.argtypes = [ctypes.c_uint64, ctypes.c_uint64]
        L.orc_ntt_tables.argtypes = [ctypes.c_uint64, ctypes.c_int, _u64p, _u64p, _u64p, _u64p, _u64p]
        L.orc_ctx_root_powers.restype = _u64p
        L.orc_ctx_root_powers.argtypes = [ctypes.c_void_p, _sz]
        L.orc_fastbconv.argtypes = [_sz, _u64p, _sz, _u64p, _sz, _u64p, _u64p]
        L.orc_beta.restype = _sz
        L.orc_beta.argtypes = [ctypes.c_void_p, _sz]
        L.orc_modup.argtypes = [ctypes.c_void_p, _sz, _u64p, _u64p]
        L.orc_inner_product.argtypes = [ctypes.c_void_p, _sz, _u64p, _u64p, _sz, _u64p]
        L.orc_moddown.argtypes = [ctypes.c_void_p, _sz, _u64p, _u64p]
        L.orc_keyswitch.argtypes = [ctypes.c_void_p, _sz, _u64p, _u64p, _sz, _u64p]
        L.orc_tensor.argtypes = [ctypes.c_void_p, _sz, _u64p, _u64p, _u64p]
        L.orc_multiply_and_relin.argtypes = [ctypes.c_void_p, _sz, _u64p, _u64p, _u64p, _sz, _u64p]
        L.orc_rescale.argtypes = [ctypes.c_void_p, _sz, _u64p, _sz, _u64p]
        L.orc_galois_elt_from_step.restype = ctypes.c_uint32
        L.orc_galois_elt_from_step.argtypes = [ctypes.c_int, _sz]
        L.orc_galois_perm_table.argtypes = [ctypes.c_int, ctypes.c_uint32, _u32p]
        L.orc_apply_galois.argtypes = [ctypes.c_void_p, _sz, _u64p, ctypes.c_uint32, _u64p, _sz, _u64p]
        L.orc_make_kswitch_key.argtypes = [ctypes.c_void_p, _u64p, _u64p, _u64p, _u64p, _sz, _u64p]
        L.orc_set_threads.argtypes = [ctypes.c_int]
        L.orc_behz_create.restype = ctypes.c_void_p
        L.orc_behz_create.argtypes = [ctypes.c_void_p, ctypes.c_uint64]
        L.orc_behz_destroy.argtypes = [ctypes.c_void_p]
        L.orc_behz_bsk_size.restype = _sz
        L.orc_behz_bsk_size.argtypes = [ctypes.c_void_p]
        L.orc_behz_bsk.argtypes = [ctypes.c_void_p, _u64p]
        L.orc_bfv_multiply.argtypes = [ctypes.c_void_p, ctypes.c_void_p, _u64p, _u64p, _u64p]
        L.orc_bfv_relinearize.argtypes = [ctypes.c_void_p, _u64p, _u64p, _sz, _u64p]
        L.orc_hoisting.argtypes = [ctypes.c_void_p, _sz, _u64p, _u32p, _sz, _u64p, _sz, _u64p]
        L.orc_hoisting_bfv.argtypes = [ctypes.c_void_p, _sz, _u64p, _u32p, _sz, _u64p, _sz, _u64p]
        L.orc_hps_create.restype = ctypes.c_void_p
        L.orc_hps_create.argtypes = [ctypes.c_void_p, ctypes.c_uint64]
        L.orc_hps_create2.restype = ctypes.c_void_p
        L.orc_hps_create2.argtypes = [ctypes.c_void_p, ctypes.c_uint64, ctypes.c_int]
        L.orc_hps_modswitch.argtypes = [ctypes.c_void_p, ctypes.c_void_p, _u64p, _u64p]
        L.orc_bfv_multiply_hps2.argtypes = [ctypes.c_void_p, ctypes.c_void_p, _u64p, _u64p, ctypes.c_int, _u64p]
        L.orc_hps_create_level.restype = ctypes.c_void_p
        L.orc_hps_create_level.argtypes = [ctypes.c_void_p, ctypes.c_uint64, ctypes.c_size_t]
        L.orc_hps_levels_to_drop.restype = ctypes.c_size_t
        L.orc_hps_levels_to_drop.argtypes = [ctypes.c_void_p, ctypes.c_uint64, ctypes.c_size_t, ctypes.c_int,
                                             ctypes.c_int]
        L.orc_hps_scale_q_ql.argtypes = [ctypes.c_void_p, ctypes.c_size_t, _u64p, _u64p]
        L.orc_hps_expand.argtypes = [ctypes.c_void_p, ctypes.c_size_t, _u64p, _u64p]
        L.orc_bfv_multiply_hps_level.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_size_t, _u64p, _u64p,
                                                 ctypes.c_int, ctypes.c_size_t, _u64p]
        L.orc_hps_destroy.argtypes = [ctypes.c_void_p]
        L.orc_bfv_decrypt.argtypes = [ctypes.c_void_p, ctypes.c_size_t, ctypes.c_uint64, ctypes.c_int, _u64p, _u64p]
        L.orc_salsa_quarterround.argtypes = [ctypes.POINTER(ctypes.c_uint32)]
        L.orc_salsa_hash.argtypes = [ctypes.c_char_p, ctypes.c_char_p]
        L.orc_sample_uniform_seeded.argtypes = [ctypes.c_void_p, ctypes.c_char_p, _sz, _u64p]
        L.orc_hps_r_size.restype = _sz
        L.orc_hps_r_size.argtypes = [ctypes.c_void_p]
        L.orc_hps_r.argtypes = [ctypes.c_void_p, _u64p]
        L.orc_bfv_multiply_hps.argtypes = [ctypes.c_void_p, ctypes.c_void_p, _u64p, _u64p, _u64p]
        _dp = ctypes.POINTER(ctypes.c_double)
        L.orc_fma_chain.argtypes = [_u64p, _sz, _sz, _dp, ctypes.c_double, _dp]
        L.orc_bconv_hps.argtypes = [_sz, _u64p, _sz, _u64p, _sz, _u64p, _u64p, _dp]
        L.orc_hps_scale_round.argtypes = [ctypes.c_void_p, ctypes.c_void_p, _u64p, _u64p, _dp]
        L.orc_hps_frac.argtypes = [ctypes.c_void_p, _dp, _dp]
        L.orc_hps_scale_q_ql_nu.argtypes = [ctypes.c_void_p, ctypes.c_size_t, _u64p, _u64p, _dp]
        _lib = L
    return _lib


def _p(a):
    assert a.dtype == np.uint64 and a.flags["C_CONTIGUOUS"], (a.dtype, a.flags)
    return a.ctypes.data_as(_u64p)


def _d(a):
    assert a.dtype == np.float64 and a.flags["C_CONTIGUOUS"], (a.dtype, a.flags)
    return a.ctypes.data_as(ctypes.POINTER(ctypes.c_double))


def fma_chain(x, w, init):
    """the oracle's FP64 sum of the HPS path: fma((double) x_i, w_i, acc) in order,
    x [nterms][N] -> [N]"""
    x = np.ascontiguousarray(x, np.uint64)
    w = np.ascontiguousarray(w, np.float64)
    out = np.zeros(x.shape[1], np.float64)
    lib().orc_fma_chain(_p(x), x.shape[0], x.shape[1], _d(w), float(init), _d(out))
    return out


def bconv_hps(ib, ob, x):
    """bConv_HPS [len(ib)][N] -> ([len(ob)][N], the FP64 sum per coefficient)"""
    x = np.ascontiguousarray(x, np.uint64)
    N = x.shape[1]
    ib_a, ob_a = np.array(ib, np.uint64), np.array(ob, np.uint64)
    out = np.zeros((len(ob), N), np.uint64)
    s = np.zeros(N, np.float64)
    lib().orc_bconv_hps(N, _p(ib_a), len(ib), _p(ob_a), len(ob), _p(x), _p(out), _d(s))
    return out, s


def set_threads(n):
    lib().orc_set_threads(int(n))


def get_primes(N, bits, count):
    out = np.zeros(count, np.uint64)
    assert lib().orc_get_primes(N, bits, count, _p(out)) == 0
    return [int(x) for x in out]


def coeff_modulus_create(N, bits):
    b = (ctypes.c_int * len(bits))(*bits)
    out = np.zeros(len(bits), np.uint64)
    assert lib().orc_coeff_modulus_create(N, b, len(bits), _p(out)) == 0
    return [int(x) for x in out]


def minimal_root(q, degree):
    return int(lib().orc_minimal_root(q, degree))


def is_prime(v):
    return bool(lib().orc_is_prime(v))


def ntt_tables(q, logN):
    N = 1 << logN
    rp, rps, irp, irps = (np.zeros(N, np.uint64) for _ in range(4))
    ninv = np.zeros(1, np.uint64)
    lib().orc_ntt_tables(q, logN, _p(rp), _p(rps), _p(irp), _p(irps), _p(ninv))
    return rp, rps, irp, irps, int(ninv[0])


def galois_elt_from_step(step, N):
    return int(lib().orc_galois_elt_from_step(step, N))


def galois_perm_table(logN, elt):
    out = np.zeros(1 << logN, np.uint32)
    lib().orc_galois_perm_table(logN, elt, out.ctypes.data_as(_u32p))
    return out


def fastbconv(N, ibase, obase, x):
    ib = np.array(ibase, np.uint64)
    ob = np.array(obase, np.uint64)
    x = np.ascontiguousarray(x, np.uint64)
    out = np.zeros((len(obase), N), np.uint64)
    lib().orc_fastbconv(N, _p(ib), len(ibase), _p(ob), len(obase), _p(x), _p(out))
    return out


class Ctx:
    """Oracle context over the QP chain (size_Q data primes then size_P special)."""

    def __init__(self, logN, moduli, size_P):
        self.logN = logN
        self.N = 1 << logN
        self.moduli = [int(m) for m in moduli]
        self.size_QP = len(moduli)
        self.size_P = size_P
        self.size_Q = self.size_QP - size_P
        m = np.array(self.moduli, np.uint64)
        self._h = lib().orc_ctx_create(logN, _p(m), self.size_QP, size_P)

    def __del__(self):
        if getattr(self, "_h", None):
            lib().orc_ctx_destroy(self._h)
            self._h = None

    def beta(self, size_Ql):
        return int(lib().orc_beta(self._h, size_Ql))

    def ntt_forward(self, data, start=0):
        a = np.ascontiguousarray(data, np.uint64).copy()
        L = a.size // self.N
        lib().orc_ntt_forward(self._h, _p(a), L, start)
        return a

    def ntt_inverse(self, data, start=0):
        a = np.ascontiguousarray(data, np.uint64).copy()
        L = a.size // self.N
        lib().orc_ntt_inverse(self._h, _p(a), L, start)
        return a

    def root_powers(self, i):
        ptr = lib().orc_ctx_root_powers(self._h, i)
        return np.ctypeslib.as_array(ptr, shape=(self.N,)).copy()

    def modup(self, size_Ql, c2):
        beta = self.beta(size_Ql)
        out = np.zeros((beta, size_Ql + self.size_P, self.N), np.uint64)
        c2 = np.ascontiguousarray(c2, np.uint64)
        assert lib().orc_modup(self._h, size_Ql, _p(c2), _p(out)) == 0
        return out

    def inner_product(self, size_Ql, t, key):
        dnum = key.shape[0]
        out = np.zeros((2, size_Ql + self.size_P, self.N), np.uint64)
        assert lib().orc_inner_product(self._h, size_Ql, _p(np.ascontiguousarray(t)), _p(key), dnum, _p(out)) == 0
        return out

    def moddown(self, size_Ql, cx_i):
        out = np.zeros((size_Ql, self.N), np.uint64)
        assert lib().orc_moddown(self._h, size_Ql, _p(np.ascontiguousarray(cx_i)), _p(out)) == 0
        return out

    def keyswitch(self, size_Ql, c2, key):
        out = np.zeros((2, size_Ql, self.N), np.uint64)
        rc = lib().orc_keyswitch(self._h, size_Ql, _p(np.ascontiguousarray(c2)), _p(key), key.shape[0], _p(out))
        assert rc == 0
        return out

    def tensor(self, size_Ql, ct1, ct2):
        out = np.zeros((3, size_Ql, self.N), np.uint64)
        lib().orc_tensor(self._h, size_Ql, _p(np.ascontiguousarray(ct1)), _p(np.ascontiguousarray(ct2)), _p(out))
        return out

    def multiply_and_relin(self, size_Ql, ct1, ct2, key):
        out = np.zeros((2, size_Ql, self.N), np.uint64)
        rc = lib().orc_multiply_and_relin(self._h, size_Ql, _p(np.ascontiguousarray(ct1)),
                                          _p(np.ascontiguousarray(ct2)), _p(key), key.shape[0], _p(out))
        assert rc == 0
        return out

    def rescale(self, size_Ql, ct):
        ct = np.ascontiguousarray(ct, np.uint64)
        size = ct.shape[0]
        out = np.zeros((size, size_Ql - 1, self.N), np.uint64)
        assert lib().orc_rescale(self._h, size_Ql, _p(ct), size, _p(out)) == 0
        return out

    def apply_galois(self, size_Ql, ct, elt, key):
        out = np.zeros((2, size_Ql, self.N), np.uint64)
        rc = lib().orc_apply_galois(self._h, size_Ql, _p(np.ascontiguousarray(ct)), elt, _p(key),
                                    key.shape[0], _p(out))
        assert rc == 0
        return out

    def hoisting(self, size_Ql, ct, elts, keys, bfv=False):
        """sum of the rotations by elts; keys [n][dnum][2][QP][N]; bfv: coefficient form"""
        out = np.zeros((2, size_Ql, self.N), np.uint64)
        e = np.array(elts, np.uint32)
        keys = np.ascontiguousarray(keys, np.uint64)
        fn = lib().orc_hoisting_bfv if bfv else lib().orc_hoisting
        rc = fn(self._h, size_Ql, _p(np.ascontiguousarray(ct, np.uint64)), e.ctypes.data_as(_u32p),
                                len(e), _p(keys), keys.shape[1], _p(out))
        assert rc == 0
        return out

    def make_kswitch_key(self, s_ntt, s2_ntt, a, e):
        dnum = a.shape[0]
        key = np.zeros((dnum, 2, self.size_QP, self.N), np.uint64)
        lib().orc_make_kswitch_key(self._h, _p(np.ascontiguousarray(s_ntt)), _p(np.ascontiguousarray(s2_ntt)),
                                   _p(np.ascontiguousarray(a)), _p(np.ascontiguousarray(e)), dnum, _p(key))
        return key


class Behz:
    """BEHZ constants of an oracle context's top data level, BFV multiply."""

    def __init__(self, octx, t):
        self.octx = octx
        self.t = t
        self._h = lib().orc_behz_create(octx._h, t)

    def __del__(self):
        if getattr(self, "_h", None) and _lib is not None:
            _lib.orc_behz_destroy(self._h)
            self._h = None

    def bsk(self):
        n = lib().orc_behz_bsk_size(self._h)
        out = np.zeros(n, np.uint64)
        lib().orc_behz_bsk(self._h, _p(out))
        return [int(x) for x in out]

    def multiply(self, ct1, ct2, square=False):
        """BEHZ squaring is the product of the operand with itself (square is ignored)"""
        k = self.octx.size_QP - self.octx.size_P
        out = np.zeros((3, k, self.octx.N), np.uint64)
        rc = lib().orc_bfv_multiply(self.octx._h, self._h, _p(np.ascontiguousarray(ct1, np.uint64)),
                                    _p(np.ascontiguousarray(ct2, np.uint64)), _p(out))
        assert rc == 0
        return out


def bfv_relinearize(octx, ct3, key):
    k = octx.size_QP - octx.size_P
    out = np.zeros((2, k, octx.N), np.uint64)
    rc = lib().orc_bfv_relinearize(octx._h, _p(np.ascontiguousarray(ct3, np.uint64)), _p(key), key.shape[0],
                                   _p(out))
    assert rc == 0
    return out


class Hps:
    """HPS constants of an oracle context's top data level, BFV multiply."""

    def __init__(self, octx, t, overq=False):
        self.octx = octx
        self.t = t
        self.overq = overq
        self._h = lib().orc_hps_create2(octx._h, t, 1 if overq else 0)

    def __del__(self):
        if getattr(self, "_h", None) and _lib is not None:
            _lib.orc_hps_destroy(self._h)
            self._h = None

    def r(self):
        n = lib().orc_hps_r_size(self._h)
        out = np.zeros(n, np.uint64)
        lib().orc_hps_r(self._h, _p(out))
        return [int(x) for x in out]

    def multiply(self, ct1, ct2, square=False):
        k = self.octx.size_QP - self.octx.size_P
        out = np.zeros((3, k, self.octx.N), np.uint64)
        rc = lib().orc_bfv_multiply_hps2(self.octx._h, self._h, _p(np.ascontiguousarray(ct1, np.uint64)),
                                         _p(np.ascontiguousarray(ct2, np.uint64)), 1 if square else 0, _p(out))
        assert rc == 0
        return out

    def frac(self):
        """(tRSHatInvModsDivsFrac [k], tQlSlHatInvModsDivsFrac [nr] or None)"""
        k = self.octx.size_QP - self.octx.size_P if not isinstance(self, HpsLevel) else self.octx.size_Q - self.d
        fq = np.zeros(k, np.float64)
        fr = np.zeros(len(self.r()), np.float64)
        lib().orc_hps_frac(self._h, _d(fq), _d(fr))
        return fq, (fr if self.overq else None)

    def scale_round(self, d):
        """the scale-and-round step of one polynomial d [k + nr][N]: (out, nu); out is
        [nr][N] (QR -> R) for hps, [k][N] (QlRl -> Ql) for hps_overq and levels"""
        d = np.ascontiguousarray(d, np.uint64)
        nr = len(self.r())
        k = d.shape[0] - nr
        out = np.zeros((k if self.overq else nr, self.octx.N), np.uint64)
        nu = np.zeros(self.octx.N, np.float64)
        lib().orc_hps_scale_round(self.octx._h, self._h, _p(d), _p(out), _d(nu))
        return out, nu

    def modswitch(self, x):
        """bConv_BEHZ_var1 Q -> R of x [k][N] (over-Q tool)"""
        out = np.zeros((len(self.r()), self.octx.N), np.uint64)
        lib().orc_hps_modswitch(self.octx._h, self._h, _p(np.ascontiguousarray(x, np.uint64)), _p(out))
        return out


class HpsLevel(Hps):
    """hps_overq_leveled tool of the level with d dropped data primes."""

    def __init__(self, octx, t, d):
        self.octx = octx
        self.t = t
        self.overq = True
        self.d = d
        self._h = lib().orc_hps_create_level(octx._h, t, d)
        assert self._h

    def multiply(self, ct1, ct2, square=False, expand_polys=3):
        """[3][size_Q][N]; limbs >= Ql of an unexpanded polynomial are 0"""
        k = self.octx.size_Q
        out = np.zeros((3, k, self.octx.N), np.uint64)
        rc = lib().orc_bfv_multiply_hps_level(self.octx._h, self._h, self.d, _p(np.ascontiguousarray(ct1, np.uint64)),
                                              _p(np.ascontiguousarray(ct2, np.uint64)), 1 if square else 0,
                                              expand_polys, _p(out))
        assert rc == 0
        return out

    def modswitch(self, x):
        """bConv_BEHZ_var1 from all of Q to Rl of x [size_Q][N]"""
        out = np.zeros((len(self.r()), self.octx.N), np.uint64)
        lib().orc_hps_modswitch(self.octx._h, self._h, _p(np.ascontiguousarray(x, np.uint64)), _p(out))
        return out


def bfv_decrypt(octx, t, phase, behz):
    """BFV scale-and-round of a coefficient-form phase [size_Ql][N] (the level of
    its first size_Ql primes) -> N plaintext coefficients"""
    out = np.zeros(octx.N, np.uint64)
    phase = np.ascontiguousarray(phase, np.uint64)
    assert lib().orc_bfv_decrypt(octx._h, phase.shape[0], t, int(behz), _p(phase), _p(out)) == 0
    return out


def hps_levels_to_drop(octx, t, depth, key_switch=False, asymmetric=False):
    """FindLevelsToDrop"""
    return int(lib().orc_hps_levels_to_drop(octx._h, t, depth, int(key_switch), int(asymmetric)))


def hps_scale_q_ql(octx, d, x):
    """scaleAndRound_HPS_Q_Ql: [size_Q][N] -> [size_Q - d][N]"""
    out = np.zeros((octx.size_Q - d, octx.N), np.uint64)
    lib().orc_hps_scale_q_ql(octx._h, d, _p(np.ascontiguousarray(x, np.uint64)), _p(out))
    return out


def hps_scale_q_ql_nu(octx, d, x):
    """scaleAndRound_HPS_Q_Ql with its FP64 sum: ([size_Q - d][N], nu [N])"""
    out = np.zeros((octx.size_Q - d, octx.N), np.uint64)
    nu = np.zeros(octx.N, np.float64)
    lib().orc_hps_scale_q_ql_nu(octx._h, d, _p(np.ascontiguousarray(x, np.uint64)), _p(out), _d(nu))
    return out, nu


def hps_expand(octx, d, x):
    """ExpandCRTBasis_Ql_Q: [size_Q - d][N] -> [size_Q][N]"""
    out = np.zeros((octx.size_Q, octx.N), np.uint64)
    xx = np.zeros((octx.size_Q, octx.N), np.uint64)
    xx[:octx.size_Q - d] = x
    lib().orc_hps_expand(octx._h, d, _p(xx), _p(out))
    return out


def salsa_quarterround(y):
    a = (ctypes.c_uint32 * 4)(*y)
    lib().orc_salsa_quarterround(a)
    return list(a)


def salsa_hash(data):
    assert len(data) == 64
    out = ctypes.create_string_buffer(64)
    lib().orc_salsa_hash(bytes(data), out)
    return list(out.raw)


def sample_uniform_seeded(octx, seed, L):
    """the reference's seeded uniform sampler, [L][N]"""
    out = np.zeros((L, octx.N), np.uint64)
    lib().orc_sample_uniform_seeded(octx._h, bytes(seed), L, _p(out))
    return out
