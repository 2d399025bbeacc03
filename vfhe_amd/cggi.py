"""CGGI16 functional bootstrapping, RNS variant, on the hip engine, batch first.

The reference bootstraps one LWE ciphertext at a time
(modules/fhe/python/src/vfhe/fhe/cggi16.py:46-140): an initial rotation of the
test vector, then one CMUX per LWE coefficient, then sample extraction.  Here
a whole batch of LWE ciphertexts moves through the bootstrapping key together:
the rotation amounts of every input are computed on the host
(`Bootstrapper.amounts`, exact 128-bit arithmetic in the engine), the accumulators live in HBM, and
`vfhe_hip_blind_rotate_tv_batch` starts every accumulator from the test vector and
steps all of them through the key indices: at N = 1024 in one launch that keeps each
accumulator in LDS, otherwise with one monomial-product launch and one batched CMUX
(mgsw.c:51-71) per index.  The
result of each accumulator is the reference's, bit for bit, since every
accumulator sees the same sequence of engine operations the reference makes.

LWE samples and keys are the engine's host-resident `LWE` / `LWE_Key` structs
(lwe.c semantics, include/vfhe_hip.h), wrapped by `LweKey` / `LweBatch`.

Restated from the reference because they fix bits of the result:
  * the MGSW plaintext layout (`gsw_plaintexts`, mgsw.py:17-53);
  * the rotation amounts from an LWE sample (`rotation_amounts`, cggi16.py:54-86),
    with Python's exact-rational true division and round-half-even;
  * the test-vector packing of a lookup table (`pack_lut`, cggi16.py:92-116).
"""
from __future__ import annotations

import ctypes
import math

import numpy as np

from . import _native
from .arith import Polynomial, crt
from .fhe import Batch, Chain, SecretKey, _handles

U64 = _native.U64
U64P = _native.U64P
_MARK = np.uint64(0xFFFFFFFFFFFFFFFF)


def lib():
    return _native.load()


class _LweStruct(ctypes.Structure):
    """struct _LWE (modules/mlwe/python/cdef/mlwe.cdef:18-24)."""
    _fields_ = [("a", ctypes.POINTER(U64P)), ("b", U64P), ("n", U64), ("l", U64), ("ntt", ctypes.c_void_p)]


class _LweKeyStruct(ctypes.Structure):
    """struct _LWE_Key (modules/mlwe/python/cdef/mlwe.cdef:11-17)."""
    _fields_ = [("s", ctypes.POINTER(U64P)), ("n", U64), ("l", U64), ("ntt", ctypes.c_void_p),
                ("sigma", ctypes.c_double)]


# ---------------------------------------------------------------------------------------------
# LWE
# ---------------------------------------------------------------------------------------------
class LweKey:
    """An LWE secret of n signed coefficients over the first l limbs of a ring's pool."""

    def __init__(self, ring, coeffs, sigma: float = 0.0, limbs: int | None = None):
        self.ring = ring
        self.coeffs = [int(v) for v in coeffs]
        self.n = len(self.coeffs)
        self.l = ring.ell if limbs is None else limbs
        self.h = lib().lwe_alloc_key(self.n, self.l, ring.NTT)
        k = ctypes.cast(self.h, ctypes.POINTER(_LweKeyStruct)).contents
        k.sigma = float(sigma)
        for j, q in enumerate(ring.primes[: self.l]):
            row = np.array([v % q for v in self.coeffs], dtype=np.uint64)
            ctypes.memmove(k.s[j], row.ctypes.data, 8 * self.n)

    def __del__(self):
        try:
            lib().free_lwe_key(self.h)
        except Exception:
            pass


class LweBatch:
    """Engine LWE samples (host arrays a[limb][n], b[limb])."""

    def __init__(self, ring, handles, n: int, l: int):
        self.ring, self.handles, self.n, self.l = ring, list(handles), n, l

    @classmethod
    def encrypt(cls, key: LweKey, messages: list[int]) -> "LweBatch":
        """b = <a, s> + m + e per limb (lwe.c:97-123); m an integer modulo the limbs' product."""
        hs = []
        for m in messages:
            ms = np.array([int(m) % q for q in key.ring.primes[: key.l]], dtype=np.uint64)
            hs.append(lib().lwe_new_sample(ms.ctypes.data_as(U64P), key.h))
        return cls(key.ring, hs, key.n, key.l)

    def __del__(self):
        try:
            for h in self.handles:
                lib().free_lwe_sample(h)
        except Exception:
            pass

    def __len__(self):
        return len(self.handles)

    def arrays(self, k: int):
        """(a as l x n uint64, b as l uint64) of member k."""
        s = ctypes.cast(self.handles[k], ctypes.POINTER(_LweStruct)).contents
        l, n = int(s.l), int(s.n)
        a = np.stack([np.ctypeslib.as_array(s.a[j], shape=(n,)).copy() for j in range(l)])
        b = np.ctypeslib.as_array(s.b, shape=(l,)).copy()
        return a, b

    def phases(self, key: LweKey) -> list[int]:
        """b - <a, s> per member (lwe.c:144-158), recomposed over the limbs by CRT."""
        primes = self.ring.primes[: self.l]
        out = []
        buf = np.zeros(self.l, dtype=np.uint64)
        for h in self.handles:
            lib().lwe_phase(buf.ctypes.data_as(U64P), h, key.h)
            out.append(crt([int(v) for v in buf], primes))
        return out


def extract(batch: Batch, idx: int = 0) -> LweBatch:
    """Coefficient idx of every member's phase as an LWE sample of dimension r*N (mlwe.c:40-66),
    the whole batch read back at once (vfhe_hip_extract_lwe_batch)."""
    batch.to_coeff()
    outs = (ctypes.c_void_p * max(1, len(batch)))()
    lib().vfhe_hip_extract_lwe_batch(_handles(batch.handles), len(batch), idx, outs)
    return LweBatch(batch.ring, [outs[k] for k in range(len(batch))], batch.rank * batch.ring.N, batch.ring.ell)


# ---------------------------------------------------------------------------------------------
# MGSW encryptions
# ---------------------------------------------------------------------------------------------
def gsw_plaintexts(key: SecretKey, msg: Polynomial, ring, ell: int, special_modulus: int) -> list[Polynomial]:
    """The (r+1)*ell plaintexts of an MGSW encryption of msg on `ring`
    (modules/mlwe/python/src/vfhe/mlwe/mgsw.py:17-53): for every key component j and digit i,
    -s_j * msg * f_i, then for every digit i, msg * f_i, where f_i holds P mod q_i at the ring's
    i-th limb and 0 elsewhere (P the key-switching primes' product).  NTT form."""
    msg.to_NTT()
    factors = [[special_modulus % p if k == i else 0 for k, p in enumerate(ring.primes)] for i in range(ell)]
    out = []
    for j in range(key.chain.rank):
        sm = -(key.poly(j, ring) * msg)
        out += [sm * f for f in factors]
    out += [msg * f for f in factors]
    return out


class BootstrapKey:
    """One MGSW encryption per LWE key coefficient, each (r+1)*ell NTT-form samples on the
    extended ring of level 0 (cggi16.py:25-44)."""

    def __init__(self, chain: Chain, lwe_key: LweKey, out_key: SecretKey, ell: int | None = None):
        self.chain = chain
        self.ring = chain.ext[0]
        self.ell = ell if ell else chain.levels[0].ell
        self.n = lwe_key.n
        N = chain.N
        self.rows = []
        for s in lwe_key.coeffs:
            m = Polynomial(self.ring).from_array([s] + [0] * (N - 1))
            plains = gsw_plaintexts(out_key, m, self.ring, self.ell, chain.special_modulus)
            for p in plains:
                p.to_coeff()
            b = Batch.encrypt(chain, out_key, plains, ring=self.ring)
            b.to_ntt()
            self.rows.append(b)
        self.arrays = [_handles(b.handles) for b in self.rows]
        self.table = _handles([ctypes.cast(a, ctypes.c_void_p).value for a in self.arrays])


# ---------------------------------------------------------------------------------------------
# Bootstrapping
# ---------------------------------------------------------------------------------------------
def rotation_amounts(a: np.ndarray, b: np.ndarray, primes: list[int], N: int, torus_base: int):
    """Rotation of the test vector and per-coefficient exponents of one LWE sample
    (cggi16.py:54-86): with q the limbs' product and b, a_i recomposed by CRT,
    rot_b = round((b + q // (4 torus_base)) * 2N / q) mod 2N, the initial exponent
    2N - rot_b, and a_i -> round(a_i * 2N / q) mod 2N; the divisions are Python's correctly
    rounded true division and the rounding is half to even, as there."""
    q = math.prod(primes)
    bv = crt([int(v) for v in b], primes)
    off = q // (4 * torus_base)
    rot_b = round(((bv + off) * (2 * N)) / q) % (2 * N)
    steps = []
    for i in range(a.shape[1]):
        ai = crt([int(a[j, i]) for j in range(a.shape[0])], primes)
        steps.append(round((ai * (2 * N)) / q) % (2 * N))
    return (2 * N - rot_b) % (2 * N), steps


def pack_lut(chain: Chain, lut: list[int], size: int, prec: int) -> Batch:
    """Trivial encryption (a = 0) of the packed lookup table (cggi16.py:92-116): entry k scaled to
    lut[k] * q // 2^prec (q the level-0 modulus) fills N // size consecutive coefficients."""
    ring = chain.levels[0]
    N, q = ring.N, ring.q_l
    scaled = [(v * q) // (1 << prec) for v in lut]
    coeffs = [0] * N
    for i in range(N):
        k = i // (N // size)
        if k < len(scaled):
            coeffs[i] = scaled[k]
    poly = Polynomial(ring).from_bigint_array(coeffs)
    tv = Batch.alloc(chain, 0, 1, form="ntt")
    lib().mlwe_RNS_trivial_sample_of_zero(tv.handles[0])
    lib().mlwe_RNS_add_polynomial(tv.handles[0], tv.handles[0], poly.obj)
    return tv


class Bootstrapper:
    """Functional bootstrapping of LWE batches into the level-0 ring of `chain`."""

    def __init__(self, chain: Chain, bk: BootstrapKey):
        self.chain, self.bk = chain, bk

    def blind_rotate(self, tv: Batch, lwes: LweBatch, torus_base: int, amounts=None, acc: Batch | None = None) -> Batch:
        """The accumulators X^(phase-dependent) * tv of every input (cggi16.py:46-90), coefficient
        form on the level-0 ring.  amounts: (rot_tv, steps) from `amounts` when already computed;
        acc: an output batch of len(lwes) samples (every row is written).  The work is queued on
        the device; the call does not wait for it (at N = 1024)."""
        chain, N = self.chain, self.chain.N
        assert lwes.n == self.bk.n
        rots, steps = amounts if amounts is not None else self.amounts(lwes, torus_base)
        tv.to_coeff()
        if acc is None:
            acc = self.alloc_accumulators(len(lwes))
        assert len(acc) == len(lwes)
        lib().vfhe_hip_blind_rotate_tv_batch(_handles(acc.handles), len(acc), tv.handles[0],
                                             rots.ctypes.data_as(U64P), self.bk.table, self.bk.n,
                                             steps.ctypes.data_as(U64P), self.bk.ell, chain.special)
        return acc

    def alloc_accumulators(self, count: int) -> Batch:
        """`count` uninitialised accumulators on the level-0 ring, coefficient form."""
        return Batch.alloc(self.chain, 0, count, form="coeff", uninit=True)

    def amounts(self, lwes: LweBatch, torus_base: int):
        """(rot_tv per input, steps as count x n) of cggi16.py:54-86, computed by the engine's
        exact 128-bit helper; the rare entries it marks as too close to a rounding half for
        its bound, or a modulus too wide for it, go through `rotation_amounts`."""
        n, N = lwes.n, self.chain.N
        rots = np.zeros(len(lwes), dtype=np.uint64)
        steps = np.zeros(len(lwes) * n, dtype=np.uint64)
        marked = lib().vfhe_hip_cggi_rotations(_handles(lwes.handles), len(lwes), N, torus_base,
                                               rots.ctypes.data_as(U64P), steps.ctypes.data_as(U64P))
        if marked:
            primes = lwes.ring.primes[: lwes.l]
            st2 = steps.reshape(len(lwes), n)
            todo = range(len(lwes)) if marked < 0 else [
                k for k in range(len(lwes)) if rots[k] == _MARK or (st2[k] == _MARK).any()]
            for k in todo:
                rot, st = rotation_amounts(*lwes.arrays(k), primes, N, torus_base)
                rots[k] = rot
                st2[k] = st
        return rots, steps

    def bootstrap(self, tv: Batch, lwes: LweBatch, torus_base: int) -> LweBatch:
        """Blind rotation, then extraction of coefficient 0 (cggi16.py:118-140)."""
        return extract(self.blind_rotate(tv, lwes, torus_base), 0)


def lwe_key_of(key: SecretKey, ring) -> LweKey:
    """The LWE secret an extracted sample decrypts under (the key's coefficients, concatenated)."""
    return LweKey(ring, key.lwe_coeffs())
