"""Host-side mirror of vfhe.mlwe (MLWE_Scheme, MLWE_Key, MLWE_Set, MLWE) on the hip engine.

Follows modules/mlwe/python/src/vfhe/mlwe/mlwe.py: the level chain and special
rings (:38-93), sparse keys (:130-143), gadget key-switch keys (:145-176), the
relinearization key layout with NULL pass-through slots (:178-236), and the
ciphertext operations, each a call into the engine's C ABI.  Encryption draws
its randomness from the seeded host sampler of vfhe_amd.arith; the secret key's
support comes from that sampler too (the reference draws it from `secrets`).
"""
from __future__ import annotations

import ctypes
import functools
import operator

import numpy as np

from . import _native, arith
from .arith import Polynomial, Ring, lib, repr


class MLWE_Scheme:
    ciphertext_type: type

    def __init__(self, rings, special_primes=0, special_rings=None, max_lvl=None, module_rank=1):
        if isinstance(rings, Ring):
            max_lvl = max_lvl if max_lvl is not None else len(rings.primes) - 1
            if special_primes == 0:
                self.rings = [rings.quotient_ring(ell=rings.ell - i) for i in range(max_lvl)]
                self.special_rings = self.rings
            else:
                self.rings = [rings.quotient_ring(ell=rings.ell - i)
                              for i in range(special_primes, special_primes + max_lvl)]
                special_mask = ((1 << special_primes) - 1) << self.rings[0].ell
                self.special_rings = [rings.quotient_ring(mask=r.mask | special_mask) for r in self.rings]
        else:
            max_lvl = max_lvl if max_lvl is not None else len(rings)
            special_rings = special_rings if special_rings is not None else rings
            self.rings = [rings[i] for i in range(max_lvl)]
            self.special_rings = [special_rings[i] for i in range(max_lvl)]
        self.r = module_rank
        self.N = self.rings[0].N
        self.ell = len(self.rings)
        self.max_lvl = max_lvl
        self.rlk = None

    @property
    def ring(self) -> Ring:
        return self.rings[0]

    @property
    def extended_rank(self) -> int:
        return self.r * (self.r + 3) // 2

    @property
    def special_primes(self) -> int:
        return self.special_rings[0].ell - self.rings[0].ell

    def level_of_ring(self, ring: Ring, strict: bool = True) -> int:
        for lvl, r in enumerate(self.rings):
            if r.mask == ring.mask:
                return lvl
        if strict:
            raise ValueError("ring is not one of the scheme's levels")
        return -1

    def key_gen_sparse(self, h, sigma_err, ternary=True):
        """mlwe.py:130-143 with the seeded host sampler instead of `secrets`."""
        assert self.N * self.r >= h
        key = [[0] * self.N for _ in range(self.r)]
        placed, sign = 0, 1
        while placed < h:
            k = int(arith._rng.integers(0, self.N * self.r))
            if key[k // self.N][k % self.N]:
                continue
            key[k // self.N][k % self.N] = sign
            if ternary:
                sign = -sign
            placed += 1
        return MLWE_Key(key, sigma_err, self)

    def _gen_ksk_components(self, key_out, key_poly, lvl):
        special = self.special_rings[lvl]
        special_q = special.modulus_ratio(self.rings[lvl])
        key_out_special = MLWE_Key(key_out.key, key_out.sigma_err, self, ring=special)
        result = []
        for poly_j in key_poly:
            if poly_j.ring != special:
                poly_j = Polynomial(special).from_bigint_array(poly_j.get_polynomial(signed=True))
            comps = []
            for i in range(special.ell):
                sf = [0] * i + [special_q % special.primes[i]] + [0] * (special.ell - 1 - i)
                out = MLWE(self, lvl=lvl, ring=special)
                self.sample(poly_j * sf, key_out_special, out=out)
                comps.append(out)
            result.append(comps)
        return result

    def gen_ksk_for_level(self, key_out, key_in, lvl):
        key_poly = key_in if isinstance(key_in, list) else key_in.poly
        return MLWE_Set(self._gen_ksk_components(key_out, key_poly, lvl))

    def gen_rlk_for_level(self, key_out, quad_polys, lvl):
        return MLWE_Set(self._gen_ksk_components(key_out, quad_polys, lvl) + [None] * self.r)

    def quadratic_key_polys(self, key):
        return [-(key.poly[i] * key.poly[j]) for i in range(self.r) for j in range(i, self.r)]

    def gen_rlk(self, key_out, quad_polys, lvl=None):
        quad_polys = quad_polys if isinstance(quad_polys, list) else self.quadratic_key_polys(quad_polys)
        assert len(quad_polys) == self.r * (self.r + 1) // 2
        if lvl is not None:
            return self.gen_rlk_for_level(key_out, quad_polys, lvl)
        return [self.gen_rlk_for_level(key_out, quad_polys, lv) for lv in range(len(self.rings))]

    def gen_ksk(self, key_out, key_in, lvl=None):
        key_poly = key_in if isinstance(key_in, list) else key_in.poly
        if lvl is not None:
            return self.gen_ksk_for_level(key_out, key_poly, lvl)
        return [self.gen_ksk_for_level(key_out, key_poly, lv) for lv in range(len(self.rings))]

    def gen_ksk_automorphism(self, key_out, key_in, g, lvl=None):
        return self.gen_ksk(key_out, [p.automorphism(g) for p in key_in.poly], lvl)

    def keyswitch(self, c, ksk):
        ksk = ksk if isinstance(ksk, MLWE_Set) else ksk[c.lvl]
        out = c.new_like(lvl=c.lvl, ring=self.special_rings[c.lvl])
        c.to_coeff()
        lib().mlwe_RNSc_GHS_hybrid_keyswitch(out.obj, c.obj, ksk.obj, c.lvl)
        out.repr = repr.coeff
        out.ring = self.rings[c.lvl]
        return out

    def automorphism(self, c, gen, ksk):
        ksk = ksk if isinstance(ksk, MLWE_Set) else ksk[c.lvl]
        out = c.new_like(lvl=c.lvl, ring=self.special_rings[c.lvl])
        c.to_coeff()
        lib().mlwe_automorphism_RNSc_GHS(out.obj, c.obj, gen, ksk.obj, c.lvl)
        out.repr = repr.coeff
        out.ring = self.rings[c.lvl]
        return out

    def sample(self, msg: Polynomial, key, out=None, lvl=None):
        """Encryption (mlwe.c:190-241): b = e + sum_i s_i a_i (NTT form), back to coefficients, b += m."""
        if out is None:
            out = self.ciphertext_type(self, lvl=lvl)
        ring = out.ring
        msg.to_coeff()
        e = Polynomial(ring).sample_gaussian(key.sigma_err, ntt=True)
        b = e
        kr = key.at_ring(ring)
        for i in range(self.r):
            a = Polynomial(ring).sample_uniform(ntt=True)
            t = Polynomial(ring)
            lib().polynomial_mul_RNS_polynomial(t.obj, kr.poly[i].obj, a.obj)
            t.repr = repr.ntt
            b = b + t
            a.to_coeff()
            lib().polynomial_copy_RNS_polynomial(out.obj_a_i(i), a.obj)
        b.to_coeff()
        m = msg
        if msg.ring != ring:
            m = Polynomial(ring).from_bigint_array(msg.get_polynomial(signed=True))
            m.to_coeff()
        b = b + m
        lib().polynomial_copy_RNS_polynomial(out.obj_b(), b.obj)
        out.repr = repr.coeff
        return out

    def phase(self, rlwe, key, out=None):
        out = out if out is not None else Polynomial(rlwe.ring)
        kr = key.at_ring(rlwe.ring)
        rlwe.to_NTT()
        lib().mlwe_RNS_phase(out.obj, rlwe.obj, kr.obj)
        out.repr = repr.ntt
        return out

    def tensor_product(self, in1, in2):
        in1.to_NTT()
        in2.to_NTT()
        outs = [Polynomial(in1.ring, repr.ntt) for _ in range(self.extended_rank + 1)]
        lib().mlwe_tensor_product(_native.voidp_array([p.obj for p in outs]), in1.obj, in2.obj)
        return outs

    def multiply(self, in1, in2, ksk=None):
        assert in1.ring == in2.ring and in1.lvl == in2.lvl
        in1.to_NTT()
        in2.to_NTT()
        if ksk is None:
            out = in1.new_like(lvl=in1.lvl, rank=self.extended_rank)
            out.is_extended = True
            lib().mlwe_multiply(out.obj, in1.obj, in2.obj, None)
            out.repr = repr.ntt
            return out
        ksk = ksk if isinstance(ksk, MLWE_Set) else ksk[in1.lvl]
        out = in1.new_like(lvl=in1.lvl, ring=self.special_rings[in1.lvl])
        lib().mlwe_multiply(out.obj, in1.obj, in2.obj, ksk.obj)
        out.repr = repr.ntt
        out.ring = self.rings[in1.lvl]
        return out

    def relinearize(self, c_ext, ksk):
        ksk = ksk if isinstance(ksk, MLWE_Set) else ksk[c_ext.lvl]
        out = c_ext.new_like(lvl=c_ext.lvl, ring=self.special_rings[c_ext.lvl], rank=self.r)
        c_ext.to_coeff()
        lib().mlwe_RNSc_GHS_hybrid_keyswitch(out.obj, c_ext.obj, ksk.obj, c_ext.lvl)
        out.repr = repr.coeff
        out.ring = self.rings[c_ext.lvl]
        return out


class MLWE_Key:
    def __init__(self, key, sigma_err, scheme, ring=None):
        assert len(key) == scheme.r
        self.key, self.scheme, self.sigma_err = key, scheme, sigma_err
        self.ring = ring if ring is not None else scheme.rings[0]
        flat = np.array([x & 0xFFFFFFFFFFFFFFFF for x in functools.reduce(operator.iadd, [list(k) for k in key], [])],
                        dtype=np.uint64)
        self.obj = lib().mlwe_new_RNS_key_from_array(flat.ctypes.data_as(_native.U64P), self.ring.N, scheme.r,
                                                     self.ring.ell, self.ring.NTT, sigma_err)
        self.poly = [Polynomial(self.ring).from_array(s) for s in key]
        for p in self.poly:
            p.to_NTT()
        s_rns = ctypes.cast(self.obj, ctypes.POINTER(_KeyStruct)).contents.s_RNS
        for i in range(scheme.r):
            lib().polynomial_copy_RNS_polynomial(s_rns[i], self.poly[i].obj)
        self._at = {}

    def at_ring(self, ring):
        if ring == self.ring:
            return self
        k = (ring.N, ring.mask)
        if k not in self._at:
            self._at[k] = MLWE_Key(self.key, self.sigma_err, self.scheme, ring=ring)
        return self._at[k]

    def __del__(self):
        try:
            lib().free_mlwe_RNS_key(self.obj)
        except Exception:
            pass


class _KeyStruct(ctypes.Structure):
    """struct _RNS_MLWE_Key (modules/mlwe/python/cdef/mlwe.cdef:33-40)."""
    _fields_ = [("s", _native.VPP), ("s_RNS", _native.VPP), ("N", _native.U64), ("l", _native.U64),
                ("r", _native.U64), ("sigma", ctypes.c_double)]


class MLWE_Set:
    """A key-switch key set: void*[] of deep-copied MLWE arrays, NULL for pass-through (mlwe.py:521-555)."""

    def __init__(self, mlwe=None):
        if mlwe is None:
            return
        self.mlwe = mlwe
        self.dim = 2
        handles = []
        self._arrays = []
        for comp in mlwe:
            if comp is None:
                handles.append(None)
                continue
            for x in comp:
                x.to_NTT()
            arr = lib().mlwe_create_copy_array(_native.voidp_array([x.obj for x in comp]), len(comp))
            self._arrays.append((arr, len(comp)))
            handles.append(ctypes.cast(arr, ctypes.c_void_p).value)
        self.obj = _native.voidp_array(handles)

    def __del__(self):
        for arr, n in getattr(self, "_arrays", []):
            try:
                lib().free_RNS_mlwe_array(n, arr)
            except Exception:
                pass


class MLWE:
    """A ciphertext: one engine RNS_MLWE (mlwe.py:558-740)."""

    def __init__(self, scheme, lvl=None, ring=None, rank=None):
        if lvl is None:
            lvl = scheme.level_of_ring(ring, strict=False) if ring is not None else 0
        if ring is None:
            ring = scheme.rings[lvl]
        self.r = rank if rank is not None else scheme.r
        self.obj = lib().mlwe_alloc_RNS_sample(ring.N, self.r, ring.mask, ring.NTT)
        self.ring, self.scheme, self.repr, self.lvl = ring, scheme, repr.empty, lvl
        self.is_extended = False

    def __del__(self):
        try:
            lib().free_mlwe_RNS_sample(self.obj)
        except Exception:
            pass

    @property
    def ell(self):
        return self.ring.ell

    def new_like(self, lvl=None, ring=None, rank=None):
        out = type(self)(self.scheme, lvl=self.lvl if lvl is None else lvl, ring=ring, rank=rank)
        out._inherit(self)
        return out

    def _inherit(self, other):
        pass

    def _s(self):
        return ctypes.cast(self.obj, ctypes.POINTER(_native.RNSMLWE)).contents

    def obj_a_i(self, j):
        return self._s().a[j]

    def obj_b(self):
        return self._s().b

    def to_NTT(self):
        if self.repr != repr.ntt:
            lib().mlwe_RNSc_to_RNS(self.obj, self.obj)
            self.repr = repr.ntt

    def to_coeff(self):
        if self.repr != repr.coeff:
            lib().mlwe_RNS_to_RNSc(self.obj, self.obj)
            self.repr = repr.coeff

    def copy(self):
        res = self.new_like()
        lib().mlwe_copy_RNS_sample(res.obj, self.obj)
        res.repr = self.repr
        return res

    def copy_from(self, other):
        lib().mlwe_copy_RNS_sample(self.obj, other.obj)
        self.repr = other.repr

    def get_a_poly(self, j):
        res = Polynomial(self.ring)
        lib().polynomial_copy_RNS_polynomial(res.obj, self.obj_a_i(j))
        res.repr = self.repr
        return res

    def get_b_poly(self):
        res = Polynomial(self.ring)
        lib().polynomial_copy_RNS_polynomial(res.obj, self.obj_b())
        res.repr = self.repr
        return res

    def round_division(self, ring=None, lvl=None):
        assert (ring is None) != (lvl is None)
        if lvl is None:
            lvl = self.scheme.level_of_ring(ring)
        else:
            ring = self.scheme.rings[lvl]
        assert ring.is_quotient_ring(self.ring)
        self.to_coeff()
        lib().mlwe_round_division_RNSc(self.obj, self.ring.mask ^ ring.mask)
        self.lvl, self.ring = lvl, ring
        return self

    def __add__(self, other):
        if isinstance(other, Polynomial):
            return self._poly_op(other, "mlwe_RNS_add_polynomial", "mlwe_add_RNSc_polynomial")
        return self._poly_op(other, "mlwe_add_RNSc_sample", "mlwe_add_RNSc_sample")

    def __sub__(self, other):
        if isinstance(other, Polynomial):
            return self._poly_op(other, "mlwe_RNS_sub_polynomial", "mlwe_sub_RNSc_polynomial")
        return self._poly_op(other, "mlwe_sub_RNSc_sample", "mlwe_sub_RNSc_sample")

    def _poly_op(self, other, fn_ntt, fn_coeff):
        if self.repr != other.repr:
            self.to_NTT()
            other.to_NTT()
        res = self.new_like()
        getattr(lib(), fn_ntt if self.repr == repr.ntt else fn_coeff)(res.obj, self.obj, other.obj)
        res.repr = self.repr
        return res

    def __mul__(self, other):
        """mlwe.py:809-830: by a ciphertext (relinearized product), a polynomial of this ring or a
        quotient of it (base-extended first), or a per-limb scalar array."""
        if isinstance(other, MLWE):
            assert self.scheme.rlk is not None, "Relinearization key (rlk) must be set in the scheme"
            return self.scheme.multiply(self, other, self.scheme.rlk)
        if isinstance(other, Polynomial):
            res = self.new_like()
            if other.ring != self.ring:
                assert other.ring.is_quotient_ring(self.ring)
                other = other.base_extend(self.ring)
            self.to_NTT()
            other.to_NTT()
            lib().mlwe_RNS_mul_by_poly(res.obj, self.obj, other.obj)
            res.repr = repr.ntt
            return res
        res = self.new_like()
        if res.obj != self.obj:
            lib().mlwe_copy_RNS_sample(res.obj, self.obj)
        arr = np.ascontiguousarray(other, dtype=np.uint64)
        lib().mlwe_scale_RNS_mlwe_RNS(res.obj, arr.ctypes.data_as(_native.U64P))
        res.repr = self.repr
        return res

    __rmul__ = __mul__


MLWE_Scheme.ciphertext_type = MLWE
