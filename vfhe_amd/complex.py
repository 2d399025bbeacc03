"""Host mirror of vfhe.arith.complex (ComplexRing, ComplexPolynomial) on the hip engine.

Follows modules/arith/python/src/vfhe/arith/complex.py:21-162: the CKKS
"special" roots of unity are generated in 100-bit precision (mpmath) exactly as
the reference does, handed to the engine's load_rous_CT / load_rous_GS, and every
transform (bit reversal, CT_NR, GS_RN, scaling, rounding into RNS) runs on the
device through the engine's C ABI.  A ComplexPolynomial is a float64 buffer of
2N values, real parts then imaginary parts, as in the reference.
"""
from __future__ import annotations

import ctypes
from functools import lru_cache
from math import log2

import numpy as np

from ._native import load as _load


def _brev(x: int, size: int) -> int:
    return int(bin(x)[2:].rjust(size, "0")[::-1], 2) if size > 0 else 0


@lru_cache(maxsize=None)
def gen_special_rous_hp(N: int) -> tuple:
    """complex.py:91-106: result[2^(k-1) + i] = rou^((5^brev(i, k-1) * N / 2^(k+1)) mod 2N),
    rou = exp(2 pi i / 2N), evaluated with 100-bit mpmath and rounded to complex doubles."""
    import mpmath as mp
    N = int(N)
    mp.mp.prec = 100
    rou = mp.exp(2 * mp.pi * 1j / (2 * mp.mpf(N)))
    result = [1] * N
    for k in range(int(log2(N)) - 1, 0, -1):
        for i in range(2 ** (k - 1)):
            expo = ((5 ** _brev(i, k - 1)) * N // (2 ** (k + 1))) % (2 * N)
            result[2 ** (k - 1) + i] = mp.power(rou, expo)
    return tuple(complex(v) for v in result)


def _dp(a: np.ndarray):
    return a.ctypes.data_as(ctypes.POINTER(ctypes.c_double))


class ComplexRing:
    def __init__(self, N: int, special_rous: bool = True):
        self.lib = _load()
        self.N = N
        self.logN = int(log2(N))
        self.N2 = 2 * N
        if special_rous:
            rous = gen_special_rous_hp(self.N2)
            re = np.array([v.real for v in rous])
            im = np.array([v.imag for v in rous])
            self.CT_rous = self.lib.load_rous_CT(_dp(re), _dp(im), self.N2)
            inv = [v ** -1 for v in rous]
            ire = np.array([v.real for v in inv])
            iim = np.array([v.imag for v in inv])
            self.GS_rous = self.lib.load_rous_GS(_dp(ire), _dp(iim), self.N2)

    def alloc_polynomial(self) -> np.ndarray:
        return np.zeros(self.N2, dtype=np.float64)

    def exp_complex_polys_ifft_scale_round_to_RNS_batch(self, e_polys, ring, temp_delta: float):
        """complex.py:46-73: IFFT, scale by temp_delta, round into NTT-form RNS polynomials."""
        from .arith import Polynomial, repr
        n = len(e_polys)
        if n == 0:
            return []
        res = [Polynomial(ring) for _ in range(n)]
        rows = (ctypes.c_void_p * n)(*[p.obj.ctypes.data for p in e_polys])
        outs = (ctypes.c_void_p * n)(*[p.obj for p in res])
        self.lib.complex_polys_ifft_scale_round_to_RNS_batch(rows, outs, n, self.N, self.logN, self.GS_rous,
                                                             ctypes.c_double(temp_delta))
        for p in res:
            p.repr = repr.ntt
        return res


class ComplexPolynomial:
    def __init__(self, ring: ComplexRing):
        self.ring = ring
        self.obj = ring.alloc_polynomial()

    def __iter__(self):
        N = self.ring.N
        return iter(self.obj[:N] + 1j * self.obj[N:])

    def IFFT(self):
        lib, N = self.ring.lib, self.ring.N
        lib.bit_reverse_array(_dp(self.obj), N, self.ring.logN)
        lib.GS_RN(_dp(self.obj), self.ring.GS_rous, N)
        lib.complex_poly_scale_double(_dp(self.obj), ctypes.c_double(1.0 / N), N)

    def FFT(self):
        lib, N = self.ring.lib, self.ring.N
        lib.CT_NR(_dp(self.obj), self.ring.CT_rous, N)
        lib.bit_reverse_array(_dp(self.obj), N, self.ring.logN)

    def __imul__(self, other):
        if type(other) is int or type(other) is float:
            self.ring.lib.complex_poly_scale_double(_dp(self.obj), ctypes.c_double(float(other)), self.ring.N)
            return self
        raise NotImplementedError(f"cannot scale by {type(other).__name__}")

    def __setitem__(self, idx, val):
        if type(val) in (float, int):
            self.obj[idx] = val
            self.obj[idx + self.ring.N] = 0.0
        elif isinstance(val, complex):
            self.obj[idx] = float(val.real)
            self.obj[idx + self.ring.N] = float(val.imag)
        else:
            raise NotImplementedError(f"cannot assign a {type(val).__name__}")

    def from_array(self, v) -> "ComplexPolynomial":
        v = np.asarray(v, dtype=np.complex128)
        self.obj[:len(v)] = v.real
        self.obj[self.ring.N:self.ring.N + len(v)] = v.imag
        return self

    def round_to_RNS_cpp(self, ring):
        """complex.py:158-162: round the 2N values into an NTT-form RNS polynomial."""
        from .arith import Polynomial, repr
        res = Polynomial(ring)
        self.ring.lib.complex_poly_round_to_RNS(res.obj, _dp(self.obj), ring.N)
        res.repr = repr.ntt
        return res
