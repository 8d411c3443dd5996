"""Closed forms in pure-Python integers, used to pin the oracle.

Each function is a textbook definition, not a restatement of the reference's
loop structure: the NTT as polynomial evaluation at the odd powers of psi in
bit-reversed order, base conversion through the CRT integer, rescale as a
floor division of the CRT integer, and so on.
"""


def brv(x, bits):
    r = 0
    for _ in range(bits):
        r = (r << 1) | (x & 1)
        x >>= 1
    return r


def is_prime(n):
    if n < 2:
        return False
    for p in (2, 3, 5, 7, 11, 13, 17, 19, 23, 29, 31, 37):
        if n % p == 0:
            return n == p
    d, s = n - 1, 0
    while d % 2 == 0:
        d //= 2
        s += 1
    for a in (2, 3, 5, 7, 11, 13, 17, 19, 23, 29, 31, 37):
        x = pow(a, d, n)
        if x in (1, n - 1):
            continue
        for _ in range(s - 1):
            x = x * x % n
            if x == n - 1:
                break
        else:
            return False
    return True


def primes_below(bits, N, count):
    """largest primes p < 2^bits with p = 1 mod 2N, descending (scan downward)"""
    out, v = [], (1 << bits) - 2 * N + 1
    while len(out) < count:
        if is_prime(v):
            out.append(v)
        v -= 2 * N
    return out


def minimal_root_bruteforce(q, degree):
    """smallest x with x^(degree/2) = -1 mod q (primitive degree-th root)"""
    for x in range(2, q):
        if pow(x, degree // 2, q) == q - 1:
            return x
    return None


def ntt_closed(a, q, psi, logN):
    """X[k] = sum_j a_j psi^((2 brv(k) + 1) j): evaluation at the odd powers of
    psi, written in bit-reversed order"""
    N = 1 << logN
    out = []
    for k in range(N):
        r = pow(psi, 2 * brv(k, logN) + 1, q)
        acc, pw = 0, 1
        for j in range(N):
            acc += int(a[j]) * pw
            pw = pw * r % q
        out.append(acc % q)
    return out


def ntt_closed_at(a, q, psi, logN, k):
    r = pow(psi, 2 * brv(k, logN) + 1, q)
    acc, pw = 0, 1
    for j in range(1 << logN):
        acc += int(a[j]) * pw
        pw = pw * r % q
    return acc % q


def negacyclic_mul(a, b, q):
    N = len(a)
    c = [0] * N
    for i in range(N):
        ai = int(a[i])
        if not ai:
            continue
        for j in range(N):
            k = i + j
            if k < N:
                c[k] += ai * int(b[j])
            else:
                c[k - N] -= ai * int(b[j])
    return [x % q for x in c]


def crt(res, moduli):
    """res[i] residues mod moduli[i] (one coefficient) -> integer in [0, prod)"""
    Q = 1
    for m in moduli:
        Q *= m
    x = 0
    for r, m in zip(res, moduli):
        Mi = Q // m
        x += int(r) * Mi * pow(Mi, -1, m)
    return x % Q, Q


def galois_coeff(a, g, q):
    """sigma_g: X -> X^g on a coefficient vector mod X^N + 1"""
    N = len(a)
    out = [0] * N
    for i in range(N):
        e = (i * g) % (2 * N)
        if e < N:
            out[e] = (out[e] + int(a[i])) % q
        else:
            out[e - N] = (out[e - N] - int(a[i])) % q
    return out
