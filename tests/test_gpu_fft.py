"""GPU parity of the CKKS encoder FFT (complex.c) with the reference engines' own builds.

"fft.fma" = 1 reproduces the avx512ifma engine's fused complex product and
= 0 the portable engine's, so each is compared bit for bit with the reference
build that matches it, and within 2^-40 (relative to the largest output) with
the other.  The batched IFFT -> scale -> round -> RNS/NTT entry point is checked
against the same composition made of the reference FFT, round-half-even and
the oracle's NTT.
"""
import ctypes

import numpy as np
import pytest

from oracle import oracle as O
from vfhe_amd import _native, runtime as rt
from vfhe_amd.complex import gen_special_rous_hp

pytestmark = pytest.mark.gpu


def dp(a):
    return a.ctypes.data_as(ctypes.POINTER(ctypes.c_double))


def roots(n):
    r = gen_special_rous_hp(2 * n)
    re, im = np.array([v.real for v in r]), np.array([v.imag for v in r])
    ire, iim = np.array([(1 / v).real for v in r]), np.array([(1 / v).imag for v in r])
    return re, im, ire, iim


@pytest.fixture(params=O.ref_engines())
def engine(request, lib):
    old = lib.vfhe_hip_get_option(b"fft.fma")
    lib.vfhe_hip_set_option(b"fft.fma", 1 if request.param == "avx512ifma" else 0)
    yield request.param
    lib.vfhe_hip_set_option(b"fft.fma", old)


@pytest.mark.parametrize("n", [16, 2048, 4096, 32768])
def test_fft_bit_exact_with_matching_engine(lib, engine, n):
    print(f"[fft] engine={engine} n={n}", flush=True)
    L = O.ref_complex(engine)
    re, im, ire, iim = roots(n)
    ref_ct, ref_gs = L.load_rous_CT(dp(re), dp(im), 2 * n), L.load_rous_GS(dp(ire), dp(iim), 2 * n)
    hip_ct, hip_gs = lib.load_rous_CT(dp(re), dp(im), 2 * n), lib.load_rous_GS(dp(ire), dp(iim), 2 * n)
    rng = np.random.default_rng(n)
    x0 = rng.standard_normal(2 * n) * 1e6
    for fn, rt_, ht_ in (("CT_NR", ref_ct, hip_ct), ("GS_RN", ref_gs, hip_gs)):
        a = O.aligned_f64(2 * n); a[:] = x0
        b = x0.copy()
        getattr(L, fn)(dp(a), rt_, n)
        getattr(lib, fn)(dp(b), ht_, n)
        assert np.array_equal(a.view(np.uint64), b.view(np.uint64)), fn
    a = O.aligned_f64(2 * n); a[:] = x0
    b = x0.copy()
    L.bit_reverse_array(dp(a), n, int(np.log2(n)))
    lib.bit_reverse_array(dp(b), n, int(np.log2(n)))
    assert np.array_equal(a, b)
    L.complex_poly_scale_double(dp(a), 1.0 / 3.0, n)
    lib.complex_poly_scale_double(dp(b), 1.0 / 3.0, n)
    assert np.array_equal(a.view(np.uint64), b.view(np.uint64))


def test_fft_other_engine_within_2_40(lib):
    engines = O.ref_engines()
    n = 4096
    re, im, _, _ = roots(n)
    x0 = np.random.default_rng(9).standard_normal(2 * n)
    hip_ct = lib.load_rous_CT(dp(re), dp(im), 2 * n)
    for e in engines:
        L = O.ref_complex(e)
        a = O.aligned_f64(2 * n); a[:] = x0
        L.CT_NR(dp(a), L.load_rous_CT(dp(re), dp(im), 2 * n), n)
        for fma in (0, 1):
            lib.vfhe_hip_set_option(b"fft.fma", fma)
            b = x0.copy()
            lib.CT_NR(dp(b), hip_ct, n)
            assert np.max(np.abs(a - b)) <= 2 ** -40 * np.max(np.abs(a)), (e, fma)
    lib.vfhe_hip_set_option(b"fft.fma", 1)


@pytest.mark.parametrize("N", [8192, 65536])
def test_ifft_scale_round_to_rns_batch(lib, engine, N):
    """complex.c:455-500 (per row: bit reversal, GS_RN, scale by delta/n, llrint, int_array_to_RNS)."""
    n = N // 2
    L = O.ref_complex(engine)
    re, im, ire, iim = roots(n)
    ref_gs = L.load_rous_GS(dp(ire), dp(iim), 2 * n)
    hip_gs = lib.load_rous_GS(dp(ire), dp(iim), 2 * n)
    primes = O.gen_primes(N, [50] * 3)
    pool = rt.Pool(N, primes)
    rng = np.random.default_rng(N + 1)
    count, delta = 3, float(2 ** 40)
    slots = [rng.standard_normal(2 * n) for _ in range(count)]
    rows = [s.copy() for s in slots]
    outs = [pool.poly() for _ in range(count)]
    lib.complex_polys_ifft_scale_round_to_RNS_batch(
        _native.voidp_array([r.ctypes.data for r in rows]), _native.voidp_array(outs), count, n, int(np.log2(n)),
        hip_gs, ctypes.c_double(delta))
    for k in range(count):
        a = O.aligned_f64(2 * n); a[:] = slots[k]
        L.bit_reverse_array(dp(a), n, int(np.log2(n)))
        L.GS_RN(dp(a), ref_gs, n)
        L.complex_poly_scale_double(dp(a), delta / n, n)
        assert np.array_equal(rows[k].view(np.uint64), a.view(np.uint64)), k   # rows hold the transform
        ints = np.rint(a[:N]).astype(np.int64)
        got = pool.read(outs[k])
        for i, q in enumerate(primes):
            exp = np.array([int(v) % q for v in ints], dtype=np.uint64)
            O.lib().o_ntt_forward(O.ptr(exp), N, q)
            assert np.array_equal(got[i], exp), (k, i)
    # the single-row rounding entry point (complex.c:39-50)
    p = pool.poly()
    v = rng.standard_normal(N) * 1e9
    v[:4] = [0.5, 1.5, -0.5, -2.5]          # ties round to even
    lib.complex_poly_round_to_RNS(p, dp(v), N)
    got = pool.read(p)
    ints = np.rint(v).astype(np.int64)
    assert list(ints[:4]) == [0, 2, 0, -2]
    for i, q in enumerate(primes):
        exp = np.array([int(x) % q for x in ints], dtype=np.uint64)
        O.lib().o_ntt_forward(O.ptr(exp), N, q)
        assert np.array_equal(got[i], exp)
    for o in outs + [p]:
        pool.free(o)
