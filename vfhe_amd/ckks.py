"""Host-side mirror of vfhe.fhe.ckks (CKKS_Scheme, CKKS_Ciphertext) on the hip engine.

The homomorphic path (multiply, relinearize, rescale, rotate, key-switch)
follows modules/fhe/python/src/vfhe/fhe/ckks.py and runs on the engine, and so
does the encoder: the special-root complex FFT of complex.c on the device
(vfhe_amd.complex), slot j at zeta^(5^j), zeta = exp(i pi / N), with the
reference's rotation convention: rotate(ct, k) moves slot i + k to slot i
(test_ckks.py:60-74).
"""
from __future__ import annotations

import math

from .arith import Polynomial, Ring, crt, repr
from .complex import ComplexPolynomial, ComplexRing
from .mlwe import MLWE, MLWE_Scheme


class CKKS_Scheme(MLWE_Scheme):
    def __init__(self, rings, scaling_factor: float = 2**30, module_rank: int = 1, special_primes: int = 0,
                 special_rings=None):
        super().__init__(rings, special_primes=special_primes, special_rings=special_rings, module_rank=module_rank)
        self.scaling_factor = scaling_factor
        self.complex_ring = ComplexRing(self.ring.N // 2, True)

    def encode(self, values) -> Polynomial:
        """ckks.py:37-56: slots -> IFFT (bit reversal, GS_RN, 1/N) -> scale -> round into RNS (NTT form)."""
        assert len(values) == self.ring.N // 2
        c_poly = ComplexPolynomial(self.complex_ring)
        c_poly.from_array(values)
        c_poly.IFFT()
        c_poly *= self.scaling_factor
        poly = c_poly.round_to_RNS_cpp(self.ring)
        poly.struct.rns_mask = self.ring.mask
        return poly

    def decode(self, poly: Polynomial, scaling_factor: float | None = None):
        """ckks.py:58-97: CRT-reconstruct the signed coefficients, pack (c_i, c_{i+N/2}) as
        complex values, FFT (CT_NR, bit reversal), divide by the scaling factor."""
        scaling_factor = self.scaling_factor if scaling_factor is None else scaling_factor
        mask = poly.rns_mask
        rows = poly.get_coeff_matrix(repr.coeff)
        act = [(rows[i], p) for i, (p, idx) in enumerate(zip(poly.ring.primes, poly.ring.prime_indices))
               if (mask >> idx) & 1]
        mods = [p for _, p in act]
        Q = math.prod(mods)
        coeffs = []
        for j in range(poly.ring.N):
            v = crt([r[j] for r, _ in act], mods)
            coeffs.append(float(v - Q if v >= Q // 2 else v))
        c_poly = ComplexPolynomial(self.complex_ring)
        half = poly.ring.N // 2
        for i in range(half):
            c_poly[i] = complex(coeffs[i], coeffs[i + half])
        c_poly.FFT()
        return [val / scaling_factor for val in c_poly]

    def encrypt(self, message: Polynomial, key):
        out = CKKS_Ciphertext(self)
        self.sample(message, key, out=out)
        return out

    def decrypt(self, ciphertext, key) -> Polynomial:
        return self.phase(ciphertext, key)

    def rotate(self, ciphertext, k: int, ksk):
        N = self.ring.N
        return self.automorphism(ciphertext, pow(5, k % (N // 2), 2 * N), ksk)

    def gen_rotation_key(self, key, k: int):
        N = self.ring.N
        return self.gen_ksk_automorphism(key, key, pow(5, k % (N // 2), 2 * N))

    def rescale(self, ciphertext):
        """ckks.py:128-148: round-divide by the primes the next level drops."""
        lvl, nxt = ciphertext.lvl, ciphertext.lvl + 1
        assert nxt < len(self.rings)
        cur, nring = ciphertext.ring, self.rings[nxt]
        assert nring.is_quotient_ring(cur), "non-nested levels need rational rescaling"
        dropped = math.prod(p for p in cur.primes if p not in nring.primes)
        ciphertext.round_division(lvl=nxt)
        ciphertext.delta = ciphertext.delta / dropped
        return ciphertext


class CKKS_Ciphertext(MLWE):
    def __init__(self, scheme, lvl=None, ring=None, rank=None):
        super().__init__(scheme, lvl=lvl, ring=ring, rank=rank)
        self.delta = scheme.scaling_factor

    def _inherit(self, other):
        if isinstance(other, CKKS_Ciphertext):
            self.delta = other.delta

    def rescale(self):
        return self.scheme.rescale(self)

    def __mul__(self, other):
        if isinstance(other, CKKS_Ciphertext):
            base = MLWE.__mul__(self, other)
            prod = CKKS_Ciphertext(self.scheme, lvl=base.lvl)
            prod.copy_from(base)
            prod.delta = self.delta * other.delta
            return self.scheme.rescale(prod)
        if isinstance(other, Polynomial):
            base = MLWE.__mul__(self, other)
            prod = CKKS_Ciphertext(self.scheme, lvl=base.lvl)
            prod.copy_from(base)
            prod.delta = self.delta * self.scheme.scaling_factor
            return self.scheme.rescale(prod)
        raise NotImplementedError


CKKS_Scheme.ciphertext_type = CKKS_Ciphertext
__all__ = ["CKKS_Scheme", "CKKS_Ciphertext", "Ring"]
