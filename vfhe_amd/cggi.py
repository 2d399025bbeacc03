"""Host mirror of vfhe.fhe.cggi16 (CGGI16 functional bootstrap) on the hip engine.

Follows modules/fhe/python/src/vfhe/fhe/cggi16.py:14-140: bootstrap key = one
MGSW encryption of each LWE key coefficient (:25-44), blind rotation = an
initial X^-round(2N(b + offset)/Q) rotation of the test vector, then for every
input coefficient a_i a CMUX between the accumulator and its X^round(2N a_i/Q)
rotation selected by bk[i] (:46-90), LUT packing (:92-116), sample extraction
(:118-140).  The arithmetic steps are the engine's mlwe_RNSc_mul_by_xai and
mgsw_CMUX; `blind_rotate_batch` advances many independent bootstraps one key
index at a time, so each step is one batched inverse NTT and one batched CMUX
(vfhe_hip_ntt_batch, vfhe_hip_mgsw_cmux_batch) over all of them.
"""
from __future__ import annotations

import math

from . import _native
from .arith import Polynomial, crt, lib, repr
from .lwe import LWE, LWE_Key, extract_lwe
from .mgsw import CMUX, MGSW_Scheme, cmux_batch
from .mlwe import MLWE, MLWE_Key


class CGGI16_Key:
    def __init__(self):
        self.bk = []
        self.b_prec = 0


def _extract_lwe_key(key: MLWE_Key) -> LWE_Key:
    """mlwe.py:513-518: the key polynomials' coefficients, concatenated."""
    coeffs = [c for comp in key.key for c in comp]
    return LWE_Key(ring=key.scheme.rings[0], key=coeffs, n=len(coeffs))


class CGGI16:
    def __init__(self, scheme, gsw_ell: int | None = None):
        self.scheme = scheme
        self.mgsw_scheme = MGSW_Scheme(scheme, ell=gsw_ell)
        self.ring = scheme.ring

    def generate_bootstrap_key(self, input_key, output_key: MLWE_Key) -> CGGI16_Key:
        """cggi16.py:25-44."""
        bk = CGGI16_Key()
        lwe_key = _extract_lwe_key(input_key) if isinstance(input_key, MLWE_Key) else input_key
        N = self.ring.N
        for s_i in lwe_key.get_s():
            poly = Polynomial(self.ring).from_array([s_i] + [0] * (N - 1))
            mgsw = self.mgsw_scheme.encrypt(poly, output_key)
            mgsw.to_NTT()
            bk.bk.append(mgsw)
        return bk

    # -- rotation amounts, exactly as cggi16.py:54-86 computes them ------------------
    def _rotations(self, rlwe_in: LWE, torus_base: int):
        N = self.ring.N
        primes = rlwe_in.ring.primes[: rlwe_in.l]
        q_lwe = math.prod(primes)
        b_val = crt(rlwe_in.get_b(), primes)
        a_limbs = rlwe_in.get_a()
        a_vals = [crt([a_limbs[j][i] for j in range(rlwe_in.l)], primes) for i in range(rlwe_in.n)]
        prec_offset = q_lwe // (4 * torus_base)
        rot_b = round(((b_val + prec_offset) * (2 * N)) / q_lwe) % (2 * N)
        rot_tv = (2 * N - rot_b) % (2 * N)
        steps = [round((a * (2 * N)) / q_lwe) % (2 * N) for a in a_vals]
        return rot_tv, steps

    def functional_bootstrap_wo_extract(self, out: MLWE, tv: MLWE, rlwe_in: LWE, bk: CGGI16_Key, torus_base: int):
        """cggi16.py:46-90: blind rotation of tv by the phase of rlwe_in, result in out."""
        rot_tv, steps = self._rotations(rlwe_in, torus_base)
        out.to_coeff()
        tv.to_coeff()
        lib().mlwe_RNSc_mul_by_xai(out.obj, tv.obj, rot_tv)
        out.repr = repr.coeff
        for i, a_i in enumerate(steps):
            if a_i == 0:
                continue
            out.to_coeff()
            rotated = MLWE(self.scheme)
            lib().mlwe_RNSc_mul_by_xai(rotated.obj, out.obj, a_i)
            rotated.repr = repr.coeff
            res = CMUX(out, rotated, bk.bk[i])
            out.copy_from(res)

    def blind_rotate_batch(self, tvs: list, lwes: list, bk: CGGI16_Key, torus_base: int) -> list:
        """functional_bootstrap_wo_extract for many inputs at once: the accumulators advance
        together through the key indices; step i is one batched inverse NTT of the accumulators
        that rotate at i, their rotations, and one batched CMUX selected by bk.bk[i]."""
        accs = []
        plans = []
        for tv, lw in zip(tvs, lwes):
            rot_tv, steps = self._rotations(lw, torus_base)
            acc = MLWE(self.scheme)
            tv.to_coeff()
            lib().mlwe_RNSc_mul_by_xai(acc.obj, tv.obj, rot_tv)
            acc.repr = repr.coeff
            accs.append(acc)
            plans.append(steps)
        n = max((len(p) for p in plans), default=0)
        for i in range(n):
            idx = [k for k, p in enumerate(plans) if i < len(p) and p[i] != 0]
            if not idx:
                continue
            ntt_rows = [accs[k] for k in idx if accs[k].repr != repr.coeff]
            if ntt_rows:
                comps = []
                for c in ntt_rows:
                    comps += [c.obj_a_i(j) for j in range(c.r)] + [c.obj_b()]
                lib().vfhe_hip_ntt_batch(_native.voidp_array(comps), len(comps), 1)
                for c in ntt_rows:
                    c.repr = repr.coeff
            rotated = []
            for k in idx:
                r = MLWE(self.scheme)
                lib().mlwe_RNSc_mul_by_xai(r.obj, accs[k].obj, plans[k][i])
                r.repr = repr.coeff
                rotated.append(r)
            res = cmux_batch([accs[k] for k in idx], rotated, [bk.bk[i]] * len(idx))
            for k, r in zip(idx, res):
                accs[k] = r
        return accs

    def LUT_packing(self, lut: list, size: int, LUT_prec: int) -> MLWE:
        """cggi16.py:92-116: a trivial encryption of the scaled table, each entry repeated N/size times."""
        rlwe_tv = MLWE(self.scheme)
        N = self.ring.N
        q = self.ring.quotient_ring(ell=self.ring.ell).q_l
        normal_ring = self.scheme.rings[0]
        scaled_lut = [(val * q) // (1 << LUT_prec) for val in lut]
        coeffs = [0] * N
        for i in range(N):
            k = i // (N // size)
            if k < len(scaled_lut):
                coeffs[i] = scaled_lut[k]
        poly = Polynomial(normal_ring).from_bigint_array(coeffs)
        lib().mlwe_RNS_trivial_sample_of_zero(rlwe_tv.obj)
        rlwe_tv.repr = repr.ntt
        return rlwe_tv + poly

    def functional_bootstrap(self, out: LWE | None, tv: MLWE, rlwe_in: LWE, bk: CGGI16_Key, torus_base: int) -> LWE:
        """cggi16.py:118-140: blind rotation, then extraction of coefficient 0 (dimension r*N)."""
        rotated = MLWE(self.scheme)
        self.functional_bootstrap_wo_extract(rotated, tv, rlwe_in, bk, torus_base)
        res = extract_lwe(rotated, 0)
        if out is not None:
            out.a, out.b, out.n, out.l = res.a, res.b, res.n, res.l
            return out
        return res
