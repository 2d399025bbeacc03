"""Thin numpy-facing helpers over the engine's C ABI.

`Pool` owns one incNTT (the reference's prime pool for a fixed N and split
degree, modules/arith/c/src/rns_polynomial.c:7-61) and moves limb-major numpy
arrays in and out of engine polynomials through the host mirror rows
(`coeffs[i]`) plus the two residency calls of the engine.
"""
from __future__ import annotations

import ctypes

import numpy as np

from . import _native
from ._native import RNSMLWE, RNSPolynomial, U64P


def lib():
    return _native.load()


def _u64(a) -> np.ndarray:
    return np.ascontiguousarray(a, dtype=np.uint64)


class Pool:
    def __init__(self, N: int, primes, split_degree: int = 1):
        self.N = int(N)
        self.split = int(split_degree)
        self.primes = [int(p) for p in primes]
        self.l = len(self.primes)
        arr = _u64(self.primes)
        self.h = lib().new_incomplete_ntt_list(arr.ctypes.data_as(U64P), self.split, self.N, self.l)

    def extend(self, primes):
        arr = _u64([int(p) for p in primes])
        lib().incNTT_extend_with_primes(self.h, arr.ctypes.data_as(U64P), len(arr))
        self.primes += [int(p) for p in primes]
        self.l = len(self.primes)

    def full_mask(self, l: int | None = None) -> int:
        return (1 << (self.l if l is None else l)) - 1

    # ---- polynomials ------------------------------------------------------------
    def poly(self, data=None, mask: int | None = None):
        m = self.full_mask() if mask is None else mask
        p = lib().polynomial_new_RNS_polynomial(self.N, m, self.h)
        if data is not None:
            self.write(p, data)
        return p

    def free(self, p):
        lib().free_RNS_polynomial(p)

    @staticmethod
    def struct(p) -> RNSPolynomial:
        return ctypes.cast(p, ctypes.POINTER(RNSPolynomial)).contents

    def mask(self, p) -> int:
        return int(self.struct(p).rns_mask)

    def set_mask(self, p, m: int):
        self.struct(p).rns_mask = m

    def write(self, p, data):
        """Copy a [l][N] (or [allocated_l][N]) array into the host rows and mark them dirty."""
        s = self.struct(p)
        data = _u64(data)
        rows = min(int(s.allocated_l), data.shape[0])
        for i in range(rows):
            ctypes.memmove(s.coeffs[i], data[i].ctypes.data, self.N * 8)
        lib().polynomial_mark_host_dirty(p)

    def read(self, p) -> np.ndarray:
        """All allocated rows, after waiting for the device."""
        lib().polynomial_sync_to_host(p)
        s = self.struct(p)
        out = np.empty((int(s.allocated_l), self.N), dtype=np.uint64)
        for i in range(int(s.allocated_l)):
            ctypes.memmove(out[i].ctypes.data, s.coeffs[i], self.N * 8)
        return out

    # ---- MLWE samples -----------------------------------------------------------
    def sample(self, r: int, mask: int | None = None, data=None):
        m = self.full_mask() if mask is None else mask
        c = lib().mlwe_alloc_RNS_sample(self.N, r, m, self.h)
        if data is not None:
            self.write_sample(c, data)
        return c

    @staticmethod
    def comps(c) -> list:
        s = ctypes.cast(c, ctypes.POINTER(RNSMLWE)).contents
        return [s.a[i] for i in range(int(s.r))] + [s.b]

    def write_sample(self, c, data):
        for p, d in zip(self.comps(c), data):
            self.write(p, d)

    def read_sample(self, c) -> np.ndarray:
        return np.stack([self.read(p) for p in self.comps(c)])

    def sample_masks(self, c) -> list[int]:
        return [self.mask(p) for p in self.comps(c)]

    def set_sample_mask(self, c, m: int):
        for p in self.comps(c):
            self.set_mask(p, m)


def random_rows(rng: np.random.Generator, primes, N: int, mask: int | None = None) -> np.ndarray:
    """Uniform canonical residues, one row per prime (zero rows outside mask)."""
    out = np.zeros((len(primes), N), dtype=np.uint64)
    for i, q in enumerate(primes):
        if mask is None or (mask >> i) & 1:
            out[i] = rng.integers(0, int(q), N, dtype=np.uint64)
    return out
