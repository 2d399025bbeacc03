"""GPU parity: monomial products, MGSW external product, CMUX and NCMUX against the oracle.

Keys are uniform random NTT-form samples (bit-exactness does not depend on
validity; the oracle's CMUX is pinned by decryption in tests/test_oracle.py).
Layout follows mgsw.c:6-88: (r+1) * ell entries, ell = gadget rows per input
component, special primes on top of the level primes.
"""
import numpy as np
import pytest

from oracle import oracle as O
from vfhe_amd import _native, runtime as rt

pytestmark = pytest.mark.gpu


def setup(N, L, K, seed):
    primes = O.gen_primes(N, [50] * (L + K))
    return primes, rt.Pool(N, primes), O.Ctx(N, primes), np.random.default_rng(seed)


def rows(rng, primes, N, mask):
    return rt.random_rows(rng, primes, N, mask)


@pytest.mark.parametrize("N", [1024, 16384])
def test_mul_by_xai(lib, N):
    primes, pool, c, rng = setup(N, 3, 0, N)
    m = 0b111
    x = rows(rng, primes, N, m)
    p = pool.poly(x)
    out = pool.poly()
    for a in (0, 1, 8, 13, N - 1, N, N + 5, 2 * N - 1, 2 * N + 3):
        for minus1, fn in ((False, lib.polynomial_RNSc_mul_by_xai), (True, lib.polynomial_RNSc_mul_by_xai_minus1)):
            fn(out, p, a)
            ref = np.zeros_like(x)
            c.mul_by_xai(ref, x, m, a, minus1)
            assert pool.mask(out) == m
            assert np.array_equal(pool.read(out), ref), (a, minus1)
    # the MLWE form runs it on every component
    ct = pool.sample(1, m, np.stack([x, rows(rng, primes, N, m)]))
    o = pool.sample(1, m)
    lib.mlwe_RNSc_mul_by_xai(o, ct, 7)
    d = pool.read_sample(ct)
    for k in range(2):
        ref = np.zeros_like(x)
        c.mul_by_xai(ref, d[k], m, 7, False)
        assert np.array_equal(pool.read_sample(o)[k], ref)


def mgsw_key(pool, primes, rng, N, ext, r, ell):
    data = np.stack([np.stack([rows(rng, primes, N, ext) for _ in range(r + 1)]) for _ in range((r + 1) * ell)])
    entries = [pool.sample(r, ext, data[k]) for k in range((r + 1) * ell)]
    return np.ascontiguousarray(data), entries, _native.voidp_array(entries)


@pytest.mark.parametrize("N,L,K,sp", [(4096, 3, 1, 1), (8192, 4, 2, 2), (16384, 3, 1, 0)])
def test_external_product_and_cmux(lib, N, L, K, sp):
    primes, pool, c, rng = setup(N, L, K, 7 * N + L)
    lvl, ext, r = (1 << L) - 1, (1 << (L + K)) - 1, 1
    data, entries, arr = mgsw_key(pool, primes, rng, N, ext, r, L)
    d1 = np.stack([rows(rng, primes, N, lvl) for _ in range(r + 1)])
    d2 = np.stack([rows(rng, primes, N, lvl) for _ in range(r + 1)])
    c1, c2 = pool.sample(r, lvl, d1), pool.sample(r, lvl, d2)
    print(f"[mgsw] N={N} L={L} K={K} sp={sp}", flush=True)
    out = pool.sample(r, ext)
    lib.mgsw_external_product(out, arr, c1, L, sp)
    ref = np.zeros((r + 1, L + K, N), dtype=np.uint64)
    m = c.external_product(ref, r, d1, lvl, data, L, sp, ext)
    assert pool.sample_masks(out) == [m] * (r + 1)
    got = pool.read_sample(out)
    for i in range(L + K):
        if (m >> i) & 1:
            assert np.array_equal(got[:, i], ref[:, i]), i
    # CMUX, single and batched
    outs = [pool.sample(r, ext) for _ in range(3)]
    lib.mgsw_CMUX(outs[0], c1, c2, arr, L, sp)
    lib.vfhe_hip_mgsw_cmux_batch(_native.voidp_array(outs[1:]), _native.voidp_array([c1, c2]),
                                 _native.voidp_array([c2, c1]), _native.voidp_array([arr, arr]), 2, L, sp)
    for o, (a, b) in zip(outs, ((d1, d2), (d1, d2), (d2, d1))):
        ref = np.zeros((r + 1, L + K, N), dtype=np.uint64)
        m = c.cmux(ref, r, a, b, lvl, data, L, sp, ext)
        assert pool.sample_masks(o) == [m] * (r + 1)
        got = pool.read_sample(o)
        for i in range(L + K):
            if (m >> i) & 1:
                assert np.array_equal(got[:, i], ref[:, i]), i


def test_ncmux(lib):
    """mgsw.c:73-88: CMUX(in1, automorphism_{2N-1}(in2)) with a key-switching key for the automorphism."""
    N, L, K = 4096, 3, 1
    primes, pool, c, rng = setup(N, L, K, 99)
    lvl, ext, r = (1 << L) - 1, (1 << (L + K)) - 1, 1
    data, entries, arr = mgsw_key(pool, primes, rng, N, ext, r, L)
    kdata = np.stack([np.stack([rows(rng, primes, N, ext) for _ in range(r + 1)]) for _ in range(L)])
    kentries = [pool.sample(r, ext, kdata[d]) for d in range(L)]
    karr = _native.voidp_array(kentries)
    ksk = _native.voidp_array([karr])
    d1 = np.stack([rows(rng, primes, N, lvl) for _ in range(r + 1)])
    d2 = np.stack([rows(rng, primes, N, lvl) for _ in range(r + 1)])
    c1, c2 = pool.sample(r, lvl, d1), pool.sample(r, lvl, d2)
    out = pool.sample(r, ext)
    lib.mgsw_NCMUX(out, c1, c2, arr, ksk, L, K)
    tmp = np.zeros((r + 1, L + K, N), dtype=np.uint64)
    mt = c.automorphism(tmp, d2, r, lvl, 2 * N - 1, [np.ascontiguousarray(kdata)], ext)
    ref = np.zeros((r + 1, L + K, N), dtype=np.uint64)
    m = c.cmux(ref, r, d1, tmp, mt & lvl, data, L, K, ext)
    assert pool.sample_masks(out) == [m] * (r + 1)
    got = pool.read_sample(out)
    for i in range(L + K):
        if (m >> i) & 1:
            assert np.array_equal(got[:, i], ref[:, i]), i
