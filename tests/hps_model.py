"""Exact-rational model of the FP64 arithmetic in the HPS paths.

The reference evaluates its HPS sums as `acc += static_cast<double>(x) * w`
(src/rns_bconv.cu:170-195; src/rns.cu:1708-1714, 1757-1763). Compiled by nvcc
(which contracts a multiply feeding an add into one fused multiply-add by
default) each step is acc <- fma(fl(x), w, acc), with fl the IEEE conversion of
a 64-bit integer to double (round to nearest, ties to even). The model below
evaluates exactly that with Python's exact rationals: every product and sum
is formed exactly, then rounded once by Python's correctly rounded int/int
division (round half to even) -- the IEEE definition of fma, independent of
any FPU or libm. The constants are the reference's `(double) m / (double) q`
(src/rns.cu:743-748, 848-854; src/host/rns.cu:319-323), again as exact
integers rounded once per operation.

With the sums fixed this way, the HPS steps are integer functions of their
inputs, and the model composes them from textbook CRT arithmetic:
  bConv_HPS:        out = sum_i y_i Qhat_i - v Q,  v = llround(fma chain of y_i fl(1/q_i))
  scale-and-round:  out = floor(t X / D + 1/2) + (trunc(nu) - floor(nu_exact))
where D is the divisor (Q for QR -> R, R for QlRl -> Ql, Q/Ql for Q -> Ql), nu
the fma chain and nu_exact the same sum over the exact fractions.
"""
from fractions import Fraction as F
import math


def fl(x):
    """IEEE round-to-nearest-even conversion of an integer to double"""
    return float(int(x))


def fma(a, b, c):
    """IEEE fused multiply-add of doubles: the exact a*b + c, rounded once"""
    return float(F(a) * F(b) + F(c))


def chain(xs, ws, init):
    acc = float(init)
    for x, w in zip(xs, ws):
        acc = fma(fl(x), w, acc)
    return acc


def prod(ms):
    p = 1
    for m in ms:
        p *= m
    return p


def crt(res, ms):
    M = prod(ms)
    x = 0
    for r, m in zip(res, ms):
        Mi = M // m
        x += int(r) * Mi * pow(Mi, -1, m)
    return x % M


def frac_consts(num, ss, which):
    """(fl(m_i) / fl(s_i), m_i / s_i) for i in which, m_i = (num [(S/s_i)^-1]_{s_i}) mod s_i,
    S = prod(ss): the reference's tRSHatInvModsDivsFrac / tQlSlHatInvModsDivsFrac"""
    S = prod(ss)
    out_fp, out_ex = [], []
    for i in which:
        s = ss[i]
        m = num * pow((S // s) % s, -1, s) % s
        out_fp.append(fl(m) / fl(s))
        out_ex.append(F(m, s))
    return out_fp, out_ex


def fma_error_bound(xs, ws_fp, ws_ex, init):
    """a rigorous bound on |chain(xs, ws_fp, init) - (init + sum x_i w_ex_i)|: the
    constants' and conversions' errors plus half an ulp of every partial sum"""
    b = F(0)
    acc = F(init)
    for x, wf, we in zip(xs, ws_fp, ws_ex):
        b += abs(F(fl(x)) * F(wf) - int(x) * we)
        acc += F(fl(x)) * F(wf)
        e = math.frexp(float(abs(acc)) * (1 + 2.0 ** -40) + 2.0 ** -1074)[1]
        b += F(2) ** (e - 54)
    return b


def bconv_hps(ib, ob, xcol):
    """bConv_HPS of one coefficient: (out residues, fp sum, v)"""
    Q = prod(ib)
    y = [int(x) * pow((Q // q) % q, -1, q) % q for x, q in zip(xcol, ib)]
    s = chain(y, [1.0 / fl(q) for q in ib], 0.0)
    v = math.floor(F(s) + F(1, 2))           # llround: half away from zero, s >= 0
    val = sum(yi * (Q // q) for yi, q in zip(y, ib)) - v * Q
    return [val % p for p in ob], s, v, val


def scale_round(xq, xr, qs, rs, num, which_frac, t_over):
    """one coefficient of a scale-and-round: X = CRT over qs + rs of (xq, xr); the
    fractions come from the `which_frac` side ('q' or 'r') with numerator num;
    result = floor(t_over(X) + 1/2) + delta, where t_over(X) is the exact rational
    t X / divisor. Returns (value, nu, nu_exact)"""
    ss = list(qs) + list(rs)
    k = len(qs)
    idx = range(k) if which_frac == "q" else range(k, len(ss))
    fp, ex = frac_consts(num, ss, idx)
    xs = list(xq) if which_frac == "q" else list(xr)
    nu = chain(xs, fp, 0.5)
    nu_ex = F(1, 2) + sum(int(x) * e for x, e in zip(xs, ex))
    X = crt(list(xq) + list(xr), ss)
    val = math.floor(t_over(X) + F(1, 2)) + (int(nu) - math.floor(nu_ex))
    return val, nu, nu_ex


def llround(x):
    """C llround of a double: nearest integer, halves away from zero (exact)"""
    v = F(x)
    n = math.floor(abs(v) + F(1, 2))
    return n if v >= 0 else -n


def decrypt_hps(phase, qs, q_max, t, fused=True, reverse=False):
    """The reference's HPS decryption scale-and-round (src/rns.cu:1519-1632, forms
    chosen as src/rns.cu:1634-1698 chooses them, constants as src/rns.cu:583-678)
    of phase = [k][N] residues modulo the level's primes qs; q_max is the largest
    prime of the whole chain. Every double operation is rounded once, exactly as
    IEEE defines it: the fma chain as above, the other sums, products, the
    division of the constants and 1/t by Python's correctly rounded floats; the
    64-bit integer sum wraps as a uint64_t does. fused=False (a rounded product
    added with a second rounding) and reverse=True (the primes summed in reverse
    order) are the variants the tests must tell apart."""
    k, N = len(qs), len(phase[0])
    qmsb, sizeqmsb, tmsb = int(q_max).bit_length(), k.bit_length(), int(t).bit_length()
    large = qmsb + sizeqmsb >= 52
    h = qmsb >> 1
    lazy = (h + tmsb + sizeqmsb < 52) if large else (qmsb + tmsb + sizeqmsb < 52)
    Q = prod(qs)
    om, fr, omb, frb = [], [], [], []
    for q in qs:
        v = pow((Q // q) % q, -1, q)
        big = t * v
        om.append((big // q) % t)
        fr.append(fl(big % q) / fl(q))
        vb = (v << h) % q
        bigb = t * vb
        omb.append((bigb // q) % t)
        frb.append(fl(bigb % q) / fl(q))
    tinv = 1.0 / fl(t)
    mask = (1 << 64) - 1
    out = []
    def mac(x, w, acc):
        return fma(fl(x), w, acc) if fused else acc + fl(x) * w

    order = list(range(k))[::-1] if reverse else list(range(k))
    for n in range(N):
        fs, isum = 0.0, 0
        for i in order:
            x = int(phase[i][n])
            if not large:
                fs = mac(x, fr[i], fs)
                isum = (isum + (x * om[i] if lazy else x * om[i] % t)) & mask
            else:
                hi, lo = x >> h, x & ((1 << h) - 1)
                fs = mac(lo, fr[i], fs)
                fs = mac(hi, frb[i], fs)
                isum = (isum + (lo * om[i] if lazy else lo * om[i] % t)) & mask
                isum = (isum + (hi * omb[i] if lazy else hi * omb[i] % t)) & mask
        fs = fs + fl(isum)
        quot = int(fs * tinv)
        fs = fs - fl((t * quot) & mask)
        out.append(llround(fs) & mask)
    return out
