"""Writes tests/golden/golden_small.json: small known-answer vectors computed
from closed forms with Python integers (tests/bigint_ref.py), independent of
the C oracle and of the GPU library.

The reference ships no golden vectors or known-answer tests for this path,
so these come from the mathematics that fixes each output:
  * primes: largest p < 2^b with p = 1 mod 2N (scan + Miller-Rabin)
  * psi: smallest primitive 2N-th root (brute force over x < q is too slow
    for 50/60-bit q, so psi is taken as the minimum over the odd powers of a
    root found by search, which is the same set)
  * NTT: evaluation at psi^(2 brv(k) + 1)
  * FastBConv: sum_i [x_i qhat_i^-1]_{q_i} * qhat_i mod p
  * rescale: floor(X / q_last) mod q_j of the CRT integer X

Run: python tests/golden/make_golden.py
"""
import json
import os
import random
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.dirname(HERE))
import bigint_ref as br  # noqa: E402


def min_root(q, degree):
    quot = (q - 1) // degree
    g = None
    for x in range(2, 1000):
        c = pow(x, quot, q)
        if pow(c, degree // 2, q) == q - 1:
            g = c
            break
    return min(pow(g, 2 * k + 1, q) for k in range(degree // 2))


def main():
    rng = random.Random(20261203)
    logN, N = 6, 64
    q60 = br.primes_below(60, N, 3)
    q50 = br.primes_below(50, N, 2)
    out = {"logN": logN, "N": N, "ntt": [], "fastbconv": None, "rescale": None}
    for q in [q60[0], q60[2], q50[0]]:
        psi = min_root(q, 2 * N)
        a = [rng.randrange(q) for _ in range(N)]
        out["ntt"].append({"q": q, "psi": psi, "input": a, "output": br.ntt_closed(a, q, psi, logN)})
    # FastBConv from {q50[0], q50[1], q60[1]} to {q60[0], q60[2]}
    ib = [q50[0], q50[1], q60[1]]
    ob = [q60[0], q60[2]]
    x = [[rng.randrange(q) for _ in range(N)] for q in ib]
    Q = 1
    for q in ib:
        Q *= q
    res = [[0] * N for _ in ob]
    for k in range(N):
        acc = 0
        for i, qi in enumerate(ib):
            qh = Q // qi
            acc += (x[i][k] * pow(qh, -1, qi) % qi) * qh
        for o, p in enumerate(ob):
            res[o][k] = acc % p
    out["fastbconv"] = {"ibase": ib, "obase": ob, "input": x, "output": res}
    # rescale (coefficient domain): floor(X / q_last) mod q_j
    mods = [q60[0], q50[0], q50[1]]
    X = [rng.randrange(mods[0] * mods[1] * mods[2]) for _ in range(N)]
    limbs = [[v % m for v in X] for m in mods]
    out_r = [[(v // mods[-1]) % m for v in X] for m in mods[:-1]]
    out["rescale"] = {"moduli": mods, "input": limbs, "output": out_r}
    with open(os.path.join(HERE, "golden_small.json"), "w") as f:
        json.dump(out, f)


if __name__ == "__main__":
    main()
