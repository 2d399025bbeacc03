"""RNS parameter generation: the reference's prime rule and primality test.

Restates Ring.gen_prime / Ring.gen_primes
(modules/arith/python/src/vfhe/arith/polynomial.py:151-170) and the
deterministic Miller-Rabin of modules/arith/python/src/vfhe/arith/number_theory.py:3-46,
so that a pool built here holds exactly the primes, in exactly the order, that
the reference's Ring registers (limb index = registration order,
modules/arith/python/src/vfhe/arith/ntt.py:20-43).
"""
from __future__ import annotations

_BASES = (2, 3, 5, 7, 11, 13, 17, 19, 23, 29, 31, 37)


def is_prime(n: int) -> bool:
    """Miller-Rabin with the first twelve prime bases: exact for every n < 2^64."""
    if n < 2:
        return False
    for p in _BASES:
        if n == p:
            return True
        if n % p == 0:
            return False
    d, s = n - 1, 0
    while d % 2 == 0:
        d //= 2
        s += 1
    for a in _BASES:
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


def gen_prime(rou_order: int, prime_size: int, exclude=()) -> int:
    """Largest prime a*rou_order + 1 below 2^prime_size with a odd, skipping the
    first candidate and anything in `exclude` (polynomial.py:151-160)."""
    a = ((2 ** prime_size - 1) // rou_order) | 1
    a -= 2
    while True:
        c = a * rou_order + 1
        if c not in exclude and is_prime(c):
            return c
        a -= 2


def gen_primes(N: int, prime_sizes, split_degree: int = 1) -> list[int]:
    """Ring.gen_primes (polynomial.py:162-170): one prime per size, distinct, q = 1 mod 2N/split."""
    primes: list[int] = []
    for b in prime_sizes:
        primes.append(gen_prime(2 * N // split_degree, b, primes))
    return primes
