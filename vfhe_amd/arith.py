"""Host-side mirror of vfhe.arith (Ring, Polynomial, NTT_Processor) on the hip engine.

Same class names, constructor forms, method names and argument meanings as
modules/arith/python/src/vfhe/arith/{polynomial.py,ntt.py}; every coefficient
operation is one C-ABI call into libvfhe_hip.so, the call the reference makes
through cffi.  Differences, all on the host side:
  * coefficient reads/writes go through polynomial_sync_to_host /
    polynomial_mark_host_dirty (the engine keeps limbs in HBM);
  * sampling draws from a seeded numpy generator (the reference's C PRNG is
    engine-specific and not a parity target) and uploads the residues.
"""
from __future__ import annotations

import ctypes
import math
from enum import Enum

import numpy as np

from . import _native
from .params import gen_prime as _gen_prime
from .params import is_prime  # noqa: F401  (re-exported like vfhe.arith.number_theory)

U64P = _native.U64P
repr = Enum("Polynomial Representation", ["empty", "ntt", "coeff"])  # noqa: A001 (reference name)

_rng = np.random.default_rng(0x5EED)


def set_seed(seed: int) -> None:
    """Seed the host sampler (the test-only analogue of vfhe_prng_set_deterministic_seed)."""
    global _rng
    _rng = np.random.default_rng(seed)


def lib():
    return _native.load()


def next_power_of_2(x):
    return 1 << math.ceil(math.log2(x))


def crt(values, moduli):
    """modules/arith/python/src/vfhe/arith/number_theory.py:49-60."""
    if len(moduli) == 1:
        return values[0] % moduli[0]
    M = math.prod(moduli)
    return sum(v * (M // m) * pow(M // m, -1, m) for v, m in zip(values, moduli)) % M


class NTT_Processor:
    """Process-global prime pool per (N, split_degree) (ntt.py:10-62): limb index = registration order."""

    def __init__(self):
        self.incNTTs = {}
        self.primes = {}
        self.prime_to_index = {}
        self.conversion_params_cache = {}

    def register_ring_primes(self, primes, N, split_degree):
        key = (N, split_degree)
        if key not in self.incNTTs:
            self.primes[key] = list(primes)
            self.prime_to_index[key] = {p: i for i, p in enumerate(primes)}
            arr = np.array(list(primes), dtype=np.uint64)
            self.incNTTs[key] = lib().new_incomplete_ntt_list(arr.ctypes.data_as(U64P), split_degree, N, len(primes))
        else:
            pmap = self.prime_to_index[key]
            new = [p for p in primes if p not in pmap]
            if new:
                start = len(self.primes[key])
                arr = np.array(new, dtype=np.uint64)
                lib().incNTT_extend_with_primes(self.incNTTs[key], arr.ctypes.data_as(U64P), len(new))
                for i, p in enumerate(new):
                    self.primes[key].append(p)
                    pmap[p] = start + i
        return [self.prime_to_index[key][p] for p in primes]

    def get_conversion_params(self, N, split_degree, in_mask, out_mask):
        key = (N, split_degree, in_mask, out_mask, tuple(self.primes[(N, split_degree)]))
        if key not in self.conversion_params_cache:
            self.conversion_params_cache[key] = lib().init_base_conversion_params(
                self.incNTTs[(N, split_degree)], in_mask, out_mask)
        return self.conversion_params_cache[key]

    def reset(self):
        """Forget every pool (the reference's tests reset the global pool between tests, conftest.py:63-85)."""
        self.__init__()


NTT_processor_instance = NTT_Processor()


class Ring:
    """polynomial.py:20-197."""

    def __init__(self, N, mod_size=None, split_degree=None, primes=None, mask=None,
                 prime_size: int | list[int] = 49, exceptional_set_size=128) -> None:
        assert (mod_size is not None) or (type(prime_size) is list) or (primes is not None) or (mask is not None)
        if mask is not None:
            assert primes is not None, "must provide primes when mask is given"
            tsd = split_degree or next_power_of_2(exceptional_set_size / min(math.ceil(math.log2(p)) for p in primes))
            pmap = NTT_processor_instance.prime_to_index[(N, tsd)]
            primes = [p for p in primes if (mask >> pmap[p]) & 1]
            prime_size = [math.ceil(math.log2(p)) for p in primes]
        if primes is not None and prime_size is None:
            prime_size = [math.ceil(math.log2(i)) for i in primes]
        if isinstance(prime_size, list):
            self.ell = len(prime_size)
            self.prime_size = prime_size
            self.smallest_prime = min(prime_size)
        else:
            self.ell = math.ceil(mod_size / prime_size)
            self.prime_size = [prime_size] * self.ell
            self.smallest_prime = prime_size
        if not split_degree:
            split_degree = next_power_of_2(exceptional_set_size / self.smallest_prime)
        self.split_degree = split_degree
        self.N = N
        self.byte_size = N * self.ell * 8
        if primes:
            assert len(primes) >= self.ell
            self.primes = list(primes[: self.ell])
        else:
            self.primes = self.gen_primes()
        self.q_l = math.prod(self.primes)
        self.bit_size = math.ceil(math.log2(self.q_l))
        self.prime_indices = NTT_processor_instance.register_ring_primes(self.primes, N, split_degree)
        self.NTT = NTT_processor_instance.incNTTs[(N, split_degree)]
        self.mask = sum(1 << i for i in self.prime_indices)

    def _ntt_l(self):
        return ctypes.cast(self.NTT, ctypes.POINTER(_native.IncNTT)).contents.l

    def get_rou_matrix(self):
        w = lib().incNTT_get_rou_matrix(self.NTT)
        return [[w[i][k] for k in range(self.N // self.split_degree)] for i in self.prime_indices]

    def quotient_ring(self, mod_size=None, ell=None, mask=None):
        assert (mod_size is not None) ^ (ell is not None) ^ (mask is not None)
        if mod_size is not None:
            return Ring(self.N, mod_size, self.split_degree, primes=self.primes)
        if ell is not None:
            return Ring(self.N, prime_size=self.prime_size[:ell], split_degree=self.split_degree, primes=self.primes)
        return Ring(self.N, mask=mask, split_degree=self.split_degree, primes=self.primes)

    def is_quotient_ring(self, parent: Ring):
        return parent.mask & self.mask == self.mask

    def modulus_ratio(self, other_ring: Ring, return_pointer: bool = False):
        assert other_ring.is_quotient_ring(self)
        sm = self.mask & ~other_ring.mask
        delta = math.prod(p for p, idx in zip(self.primes, self.prime_indices) if (sm >> idx) & 1)
        if return_pointer:
            arr = np.zeros(self._ntt_l(), dtype=np.uint64)
            for p, idx in zip(self.primes, self.prime_indices):
                arr[idx] = delta % p
            return arr
        return delta

    @staticmethod
    def gen_prime(rou_order, prime_size, exclude_list=None):
        return _gen_prime(rou_order, prime_size, exclude_list or [])

    def gen_primes(self):
        primes = []
        for b in self.prime_size:
            primes.append(Ring.gen_prime(2 * self.N // self.split_degree, b, primes))
        return primes

    def alloc_polynomial(self):
        return lib().polynomial_new_RNS_polynomial(self.N, self.mask, self.NTT)

    def scalar_array(self, value):
        arr = np.zeros(self._ntt_l(), dtype=np.uint64)
        vals = [value % p for p in self.primes] if isinstance(value, int) else list(value)
        for v, idx in zip(vals, self.prime_indices):
            arr[idx] = v
        return arr

    def random_element(self, ntt=True):
        return Polynomial(self).sample_uniform(ntt)

    def random_gaussian_element(self, sigma, ntt=True):
        return Polynomial(self).sample_gaussian(sigma, ntt)

    def __eq__(self, other):
        return isinstance(other, Ring) and self.N == other.N and self.mask == other.mask \
            and self.split_degree == other.split_degree

    def __hash__(self):
        return hash((self.N, self.mask, self.split_degree))


class Polynomial:
    """polynomial.py:203-520: one engine RNS_Polynomial plus its representation tag."""

    def __init__(self, ring: Ring, repr=repr.empty) -> None:  # noqa: A002
        self.ring = ring
        self.obj = ring.alloc_polynomial()
        self.repr = repr

    def __del__(self):
        try:
            lib().free_RNS_polynomial(self.obj)
        except Exception:  # interpreter teardown
            pass

    @property
    def struct(self):
        return ctypes.cast(self.obj, ctypes.POINTER(_native.RNSPolynomial)).contents

    @property
    def rns_mask(self):
        return int(self.struct.rns_mask)

    # ---- construction ----------------------------------------------------------
    def from_array(self, array):
        array = list(array) + [0] * (self.ring.N - len(array))
        arr = np.array([x & 0xFFFFFFFFFFFFFFFF for x in array], dtype=np.uint64)
        lib().int_array_to_RNS(self.obj, arr.ctypes.data_as(U64P))
        self.repr = repr.ntt
        return self

    def from_bigint_array(self, array):
        n_l = self.ring._ntt_l()
        rows = (U64P * n_l)()
        keep = []
        for p, idx in zip(self.ring.primes, self.ring.prime_indices):
            row = np.array([v % p for v in array], dtype=np.uint64)
            keep.append(row)
            rows[idx] = row.ctypes.data_as(U64P)
        lib().array_to_RNS(self.obj, rows)
        self.repr = repr.ntt
        return self

    def get_coeff_matrix(self, repr=repr.coeff):  # noqa: A002
        if self.repr != repr:
            self.to_repr(repr)
        lib().polynomial_sync_to_host(self.obj)
        s = self.struct
        S, n = self.ring.split_degree, self.ring.N // self.ring.split_degree
        j = np.arange(self.ring.N)
        src = (j & (S - 1)) * n + j // S
        out = []
        for idx in self.ring.prime_indices:
            row = np.ctypeslib.as_array(s.coeffs[idx], shape=(self.ring.N,)).copy()
            out.append([int(v) for v in row[src]])
        return out

    def from_coeff_matrix(self, matrix, repr=repr.coeff):  # noqa: A002
        lib().polynomial_sync_to_host(self.obj)
        s = self.struct
        S, n = self.ring.split_degree, self.ring.N // self.ring.split_degree
        j = np.arange(self.ring.N)
        dst = (j & (S - 1)) * n + j // S
        for k, idx in enumerate(self.ring.prime_indices):
            row = np.ctypeslib.as_array(s.coeffs[idx], shape=(self.ring.N,))
            row[dst] = np.array([int(v) for v in matrix[k]], dtype=np.uint64)
        lib().polynomial_mark_host_dirty(self.obj)
        self.repr = repr
        return self

    # ---- domain -------------------------------------------------------------------
    def to_NTT(self):
        if self.repr == repr.ntt:
            return
        lib().polynomial_RNSc_to_RNS(self.obj, self.obj)
        self.repr = repr.ntt

    def to_coeff(self):
        if self.repr == repr.coeff:
            return
        lib().polynomial_RNS_to_RNSc(self.obj, self.obj)
        self.repr = repr.coeff

    def to_repr(self, r):
        if r == repr.ntt:
            self.to_NTT()
        if r == repr.coeff:
            self.to_coeff()

    # ---- sampling (host RNG, seeded) ------------------------------------------------
    def sample_uniform(self, ntt=True):
        rows = [[int(v) for v in _rng.integers(0, p, self.ring.N, dtype=np.uint64)] for p in self.ring.primes]
        self.from_coeff_matrix(rows, repr.coeff)
        self.repr = repr.ntt if ntt else repr.coeff
        return self

    def sample_gaussian(self, sigma, ntt=True):
        e = np.rint(_rng.normal(0.0, sigma, self.ring.N)).astype(np.int64)
        self.from_coeff_matrix([[int(v) % p for v in e] for p in self.ring.primes], repr.coeff)
        if ntt:
            self.to_NTT()
        return self

    # ---- arithmetic -----------------------------------------------------------------
    def _binary(self, other, fn_ntt, fn_coeff):
        if self.repr != other.repr:
            self.to_NTT()
            other.to_NTT()
        res = Polynomial(self.ring)
        getattr(lib(), fn_ntt if self.repr == repr.ntt else fn_coeff)(res.obj, self.obj, other.obj)
        res.repr = self.repr
        return res

    def __add__(self, other):
        if isinstance(other, int):
            res = self.copy()
            fn = "polynomial_RNS_add_integer" if self.repr == repr.ntt else "polynomial_RNSc_add_integer"
            getattr(lib(), fn)(res.obj, res.obj, other & 0xFFFFFFFFFFFFFFFF)
            return res
        return self._binary(other, "polynomial_add_RNS_polynomial", "polynomial_add_RNSc_polynomial")

    def __sub__(self, other):
        return self._binary(other, "polynomial_sub_RNS_polynomial", "polynomial_sub_RNSc_polynomial")

    def __neg__(self):
        res = Polynomial(self.ring)
        lib().polynomial_RNSc_negate(res.obj, self.obj)
        res.repr = self.repr
        return res

    def __mul__(self, other):
        if isinstance(other, int):
            res = Polynomial(self.ring)
            lib().polynomial_scale_RNSc_polynomial(res.obj, self.obj, other % (1 << 64))
            res.repr = self.repr
            return res
        if isinstance(other, list):
            res = Polynomial(self.ring)
            arr = self.ring.scalar_array(other)
            lib().polynomial_scale_RNS_polynomial_RNS(res.obj, self.obj, arr.ctypes.data_as(U64P))
            res.repr = self.repr
            return res
        self.to_NTT()
        other.to_NTT()
        res = Polynomial(self.ring)
        lib().polynomial_mul_RNS_polynomial(res.obj, self.obj, other.obj)
        res.repr = repr.ntt
        return res

    def automorphism(self, gen):
        assert gen < self.ring.N * 2
        res = Polynomial(self.ring)
        self.to_coeff()
        lib().polynomial_RNSc_permute(res.obj, self.obj, gen)
        res.repr = repr.coeff
        return res

    def copy(self):
        res = Polynomial(self.ring)
        lib().polynomial_copy_RNS_polynomial(res.obj, self.obj)
        res.repr = self.repr
        return res

    __copy__ = copy

    # ---- tower ------------------------------------------------------------------------
    def lift_to(self, ring=None, out=None, params=None):
        self.to_coeff()
        out_ = out if out is not None else Polynomial(ring)
        if params is None:
            params = NTT_processor_instance.get_conversion_params(self.ring.N, self.ring.split_degree,
                                                                  self.rns_mask, out_.rns_mask)
        lib().polynomial_base_conversion_RNSc(out_.obj, self.obj, params)
        out_.repr = repr.coeff
        return out_

    def base_extend(self, ring=None, out=None):
        out_ = out if out is not None else Polynomial(ring)
        assert self.ring.is_quotient_ring(out_.ring)
        return self.lift_to(out=out_)

    def mod_reduce(self, ring=None, out=None):
        self.to_coeff()
        out_ = out if out is not None else Polynomial(ring)
        assert out_.ring.is_quotient_ring(self.ring)
        lib().polynomial_RNSc_mod_reduce(out_.obj, self.obj)
        out_.repr = repr.coeff
        return out_

    def floor_division(self, ring: Ring):
        assert ring.ell < self.ring.ell and ring.is_quotient_ring(self.ring)
        self.to_coeff()
        lib().polynomial_floor_division_RNSc_wo_free(self.obj, self.rns_mask & ~ring.mask)
        self.ring = ring
        return self

    def round_division(self, ring: Ring):
        assert ring.ell < self.ring.ell and ring.is_quotient_ring(self.ring)
        self.to_coeff()
        lib().polynomial_round_division_RNSc_wo_free(self.obj, self.rns_mask & ~ring.mask)
        self.ring = ring
        return self

    def scaled_lift(self, ring: Ring, delta=None):
        self.to_coeff()
        out = Polynomial(ring)
        lib().polynomial_RNSc_scaled_lift(out.obj, self.obj,
                                          delta.ctypes.data_as(U64P) if delta is not None else None)
        out.repr = repr.coeff
        return out

    def __mod__(self, value: int):
        self.to_coeff()
        res = Polynomial(self.ring)
        lib().polynomial_RNSc_mod_reduce_lifted(res.obj, self.obj, self.ring.primes.index(value))
        res.repr = repr.coeff
        return res

    # ---- views ------------------------------------------------------------------------
    def __iter__(self):
        self.to_coeff()
        return iter(self.get_coeff_matrix(repr.coeff))

    def get_polynomial(self, signed: bool = False) -> list:
        self.to_coeff()
        rns = self.get_coeff_matrix(repr.coeff)
        if self.ring.ell == 1:
            res = list(rns[0])
        else:
            res = [crt([rns[i][j] for i in range(self.ring.ell)], self.ring.primes) for j in range(self.ring.N)]
        if not signed:
            return res
        q = self.ring.q_l
        return [v - q if v > q // 2 else v for v in res]

    def __eq__(self, value) -> bool:
        if isinstance(value, int):
            rows = self.get_coeff_matrix(repr.coeff)
            return all(r[0] == value % p and not any(r[1:]) for r, p in zip(rows, self.ring.primes))
        self.to_repr(value.repr)
        return bool(lib().polynomial_eq(self.obj, value.obj))

    def __hash__(self):
        raise TypeError("not a Python hashable object")
