"""The allocation-free NTT batch (vfhe_hip_ntt_batch_noalloc) against the batch it mirrors
(vfhe_hip_ntt_batch), bit for bit, on two calls in a row (the second smaller, with a ragged
mask), in both directions; the engine's block allocator sees no request during those calls,
and a batch larger than the scratch array is refused with nothing changed."""
import numpy as np
import pytest

from oracle import oracle as O
from vfhe_amd import _native, runtime as rt

pytestmark = pytest.mark.gpu


def _allocs(lib):
    return lib.vfhe_hip_get_option(b"stat.cache_hits") + lib.vfhe_hip_get_option(b"stat.cache_misses")


def test_noalloc_matches_ntt_batch(lib):
    print("[noalloc] xcd placement available:", lib.vfhe_hip_get_option(b"ntt.xcd_available"), flush=True)
    for call, (N, L, P, mask) in enumerate([(65536, 24, 6, None), (8192, 5, 3, 0b10110)]):
        print(f"[noalloc] call {call}: N={N} L={L} polys={P}", flush=True)
        primes = O.gen_primes(N, [50] * L)
        pool = rt.Pool(N, primes)
        rng = np.random.default_rng(N + call)
        xs = [rt.random_rows(rng, primes, N, mask) for _ in range(P)]
        a = [pool.poly(x, mask) for x in xs]
        b = [pool.poly(x, mask) for x in xs]
        aa, ab = _native.voidp_array(a), _native.voidp_array(b)
        lib.vfhe_hip_synchronize()
        for direction in (0, 1):
            n0 = _allocs(lib)
            assert lib.vfhe_hip_ntt_batch_noalloc(aa, P, direction) == 0
            assert _allocs(lib) == n0, "the allocation-free call reached the block allocator"
            lib.vfhe_hip_ntt_batch(ab, P, direction)
            lib.vfhe_hip_synchronize()
            for k in range(P):
                assert pool.mask(a[k]) == pool.mask(b[k])
                assert np.array_equal(pool.read(a[k]), pool.read(b[k])), (call, direction, k)
        if call == 1:
            ref = xs[0].copy()
            O.Ctx(N, primes).to_ntt(ref, ref, mask)
            lib.vfhe_hip_ntt_batch_noalloc(_native.voidp_array(a[:1]), 1, 0)
            got = pool.read(a[0])
            for i in range(L):
                if (mask >> i) & 1:
                    assert np.array_equal(got[i], ref[i]), i
        for p in a + b:
            pool.free(p)


def test_noalloc_refuses_oversized_batch(lib):
    N, L, P = 4096, 28, 700      # 19,600 limb rows: more than the scratch array holds
    primes = O.gen_primes(N, [50] * L)
    pool = rt.Pool(N, primes)
    rng = np.random.default_rng(7)
    x = rt.random_rows(rng, primes, N)
    ps = [pool.poly(x) for _ in range(P)]
    lib.vfhe_hip_synchronize()
    assert lib.vfhe_hip_ntt_batch_noalloc(_native.voidp_array(ps), P, 0) == -1
    assert np.array_equal(pool.read(ps[0]), x) and np.array_equal(pool.read(ps[-1]), x)
    for p in ps:
        pool.free(p)
