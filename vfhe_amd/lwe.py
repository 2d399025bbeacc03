"""Host mirror of vfhe.mlwe.lwe (LWE_Key, LWE) and of mlwe_extract_LWE.

Follows modules/mlwe/python/src/vfhe/mlwe/lwe.py:17-140 over the semantics of
modules/mlwe/c/src/lwe.c:6-160 and mlwe.c:40-66.  An LWE sample is l rows
(one per prime) of n integers plus l values b: b_i = <a_i, s> + m_i + e mod q_i
with a_i drawn independently per limb and one error e for all limbs
(lwe.c:95-121), phase_i = b_i - <a_i, s> (lwe.c:141-156).  These are a few
hundred words of scalar work per sample with no place on the device path, so
they stay on the host as integer arrays; randomness comes from the seeded host
sampler of vfhe_amd.arith (the reference draws from its C PRNG).  A key built
from an array has error width 0 (lwe_alloc_key leaves it unset).
"""
from __future__ import annotations

import numpy as np

from . import arith
from .arith import crt, repr


def _rows_mod(values, primes):
    return [[int(v) % q for v in values] for q in primes]


class LWE_Key:
    def __init__(self, ring, sec_sigma=None, err_sigma=None, sparse_h=None, key=None, n=None):
        self.ring = ring
        self.n = n if n is not None else ring.N
        self.l = ring.ell
        self.sigma = 0.0
        if key is not None:
            self.set_s(key)
        elif sparse_h is not None and err_sigma is not None:
            # lwe.c:72-93: h non-zero entries, each +1 or -1
            s = [0] * self.n
            placed = 0
            while placed < sparse_h:
                k = int(arith._rng.integers(0, self.n))
                if s[k]:
                    continue
                s[k] = 1 if int(arith._rng.integers(0, 2)) else -1
                placed += 1
            self.set_s(s)
            self.sigma = err_sigma
        elif sec_sigma is not None and err_sigma is not None:
            # lwe.c:56-70: rounded Gaussian entries
            self.set_s([int(np.rint(x)) for x in arith._rng.normal(0.0, sec_sigma, self.n)])
            self.sigma = err_sigma
        else:
            self.set_s([0] * self.n)

    def set_s(self, key):
        assert len(key) == self.n
        self.key = [int(v) for v in key]
        self.s = _rows_mod(self.key, self.ring.primes[: self.l])

    def get_s(self) -> list[int]:
        """lwe.py:56-64: limb 0 with the sign recovered."""
        q0 = self.ring.primes[0]
        return [v - q0 if v > q0 // 2 else v for v in self.s[0]]


class LWE:
    def __init__(self, ring, m=None, key: LWE_Key | None = None, is_trivial=False, n=None, a=None, b=None):
        self.ring = ring
        self.n = n if n is not None else ring.N
        self.l = ring.ell
        primes = ring.primes[: self.l]
        if a is not None:
            self.a, self.b = [list(map(int, row)) for row in a], [int(x) for x in b]
            self.n = len(self.a[0])
        elif key is not None and m is not None:
            self.n = key.n
            e = int(np.rint(arith._rng.normal(0.0, key.sigma))) if key.sigma > 0 else 0
            self.a, self.b = [], []
            for i, q in enumerate(primes):
                row = [int(x) for x in arith._rng.integers(0, q, self.n, dtype=np.uint64)]
                dot = sum(x * y for x, y in zip(row, key.s[i])) % q
                self.a.append(row)
                self.b.append((dot + int(m[i]) + e) % q)
        else:
            self.a = [[0] * self.n for _ in primes]
            self.b = [int(m[i]) % q for i, q in enumerate(primes)] if (is_trivial and m is not None) else [0] * self.l

    def phase(self, key: LWE_Key, recompose: bool = False):
        primes = self.ring.primes[: self.l]
        out = [(self.b[i] - sum(x * y for x, y in zip(self.a[i], key.s[i]))) % q for i, q in enumerate(primes)]
        return crt(out, primes) if recompose else out

    def subto(self, other: "LWE"):
        primes = self.ring.primes[: self.l]
        for i, q in enumerate(primes):
            self.a[i] = [(x - y) % q for x, y in zip(self.a[i], other.a[i])]
            self.b[i] = (self.b[i] - other.b[i]) % q

    def get_a(self) -> list[list[int]]:
        return [list(row) for row in self.a]

    def get_b(self) -> list[int]:
        return list(self.b)


def extract_lwe(c, idx: int = 0) -> LWE:
    """mlwe_extract_LWE (mlwe.c:40-66): coefficient idx of the phase of an MLWE sample as an LWE
    sample of dimension r*N: a'[k*N + i] = a_k[idx - i] for i <= idx, -a_k[N + idx - i] otherwise."""
    c.to_coeff()
    N, r = c.ring.N, c.r
    rows_b = c.get_b_poly().get_coeff_matrix(repr.coeff)
    rows_a = [c.get_a_poly(k).get_coeff_matrix(repr.coeff) for k in range(r)]
    primes = c.ring.primes
    a, b = [], []
    for j, q in enumerate(primes):
        row = []
        for k in range(r):
            ak = [int(x) for x in rows_a[k][j]]
            row += [ak[idx - i] if i <= idx else (q - ak[N + idx - i]) % q for i in range(N)]
        a.append(row)
        b.append(int(rows_b[j][idx]))
    return LWE(c.ring, a=a, b=b)
