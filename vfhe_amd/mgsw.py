"""Host mirror of vfhe.mlwe.mgsw (MGSW_Scheme, MGSW, CMUX, NCMUX) on the hip engine.

Follows modules/mlwe/python/src/vfhe/mlwe/mgsw.py:10-140.  An MGSW ciphertext is
the reference's list of (r+1)*ell MLWE samples on the scheme's special ring:
for every key component s_j and digit i an encryption of -s_j*m*(P mod q_i at
limb i), then for every digit an encryption of m*(P mod q_i at limb i)
(mgsw.py:17-53).  The external product, CMUX and NCMUX are the engine's
mgsw_* entry points; `cmux_batch` runs many independent CMUXes in one launch
sequence (vfhe_hip_mgsw_cmux_batch).
"""
from __future__ import annotations

from . import _native
from .arith import Polynomial, lib, repr
from .mlwe import MLWE, MLWE_Key, MLWE_Set


class MGSW_Scheme:
    def __init__(self, mlwe_scheme, ell: int | None = None):
        self.mlwe_scheme = mlwe_scheme
        self.ell = ell if ell else mlwe_scheme.rings[0].ell
        self.ring = mlwe_scheme.special_rings[0]

    def encrypt(self, msg: Polynomial, key: MLWE_Key, lvl: int = 0) -> "MGSW":
        """mgsw.py:17-53.  Only limb i of each digit's plaintext is non-zero, so computing
        s_j * m on the special ring gives the reference's plaintexts limb for limb."""
        scheme = self.mlwe_scheme
        special_q = self.ring.modulus_ratio(scheme.rings[lvl])
        if msg.ring != self.ring:
            msg = msg.base_extend(self.ring)
        key_special = MLWE_Key(key.key, key.sigma_err, scheme, ring=self.ring)

        def digit_factor(i):
            return [0] * i + [special_q % self.ring.primes[i]] + [0] * (self.ring.ell - 1 - i)

        result = []
        for j in range(scheme.r):
            sm = -(key_special.poly[j] * msg)
            for i in range(self.ell - lvl):
                out = MLWE(scheme, ring=self.ring)
                scheme.sample(sm * digit_factor(i), key_special, out=out)
                result.append(out)
        for i in range(self.ell - lvl):
            out = MLWE(scheme, ring=self.ring)
            scheme.sample(msg * digit_factor(i), key_special, out=out)
            result.append(out)
        return MGSW(self, obj=result)


class MGSW:
    def __init__(self, scheme: MGSW_Scheme, obj: list | None = None):
        self.scheme = scheme
        self.obj = obj if obj else [MLWE(scheme.mlwe_scheme, ring=scheme.ring)
                                    for _ in range(scheme.mlwe_scheme.r * scheme.ell + scheme.ell)]
        self._arr = None

    def to_NTT(self):
        for c in self.obj:
            c.to_NTT()

    def to_coeff(self):
        for c in self.obj:
            c.to_coeff()
        self._arr = None

    def handles(self):
        """The `RNS_MLWE *` array the engine takes (entries in NTT form)."""
        self.to_NTT()
        if self._arr is None:
            self._arr = _native.voidp_array([c.obj for c in self.obj])
        return self._arr

    def external_product(self, other: MLWE) -> MLWE:
        """mgsw.py:77-98: result in NTT form on the scheme's level-0 ring."""
        res = MLWE(self.scheme.mlwe_scheme)
        arr = self.handles()
        other.to_coeff()
        lib().mgsw_external_product(res.obj, arr, other.obj, self.scheme.ell,
                                    self.scheme.mlwe_scheme.special_primes)
        res.repr = repr.ntt
        return res

    def __mul__(self, other):
        if isinstance(other, MLWE):
            return self.external_product(other)
        return NotImplemented


def CMUX(in1: MLWE, in2: MLWE, selector: MGSW) -> MLWE:
    """mgsw.py:107-128: selector * (in2 - in1) + in1, in NTT form."""
    res = MLWE(in1.scheme)
    in1.to_coeff()
    in2.to_coeff()
    arr = selector.handles()
    lib().mgsw_CMUX(res.obj, in1.obj, in2.obj, arr, selector.scheme.ell, in1.scheme.special_primes)
    res.repr = repr.ntt
    return res


def NCMUX(in1: MLWE, in2: MLWE, selector: MGSW, aut_minus1) -> MLWE:
    """mgsw.py:131-158: CMUX(in1, automorphism_{2N-1}(in2), selector)."""
    aut_minus1 = aut_minus1 if isinstance(aut_minus1, MLWE_Set) else aut_minus1[in1.lvl]
    res = MLWE(in1.scheme, lvl=in1.lvl)
    in1.to_coeff()
    in2.to_coeff()
    arr = selector.handles()
    lib().mgsw_NCMUX(res.obj, in1.obj, in2.obj, arr, aut_minus1.obj, selector.scheme.ell,
                     in1.scheme.special_primes)
    res.repr = repr.ntt
    return res


def cmux_batch(ins1: list, ins2: list, selectors: list) -> list:
    """CMUX over independent triples in one launch sequence; results in NTT form."""
    n = len(ins1)
    assert len(ins2) == n and len(selectors) == n
    if n == 0:
        return []
    outs = [MLWE(c.scheme) for c in ins1]
    for c in ins1 + ins2:
        c.to_coeff()
    sels = [s.handles() for s in selectors]
    scheme = selectors[0].scheme
    lib().vfhe_hip_mgsw_cmux_batch(_native.voidp_array([o.obj for o in outs]),
                                   _native.voidp_array([c.obj for c in ins1]),
                                   _native.voidp_array([c.obj for c in ins2]),
                                   _native.voidp_array(sels), n, scheme.ell, ins1[0].scheme.special_primes)
    for o in outs:
        o.repr = repr.ntt
    return outs
