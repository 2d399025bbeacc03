"""Slot operations on NTT-form polynomials (rns_polynomial.c:1228-1300, 1383-1429) on the device
against the oracle's restatements: broadcast, rotation and copy of slots with split degree 1 and 4
and a ragged limb mask, and the slot-wise inverse (status -2 on a zero slot, -1 for split > 1,
whose slots are extension-field elements outside this engine)."""
import numpy as np
import pytest

from oracle import oracle as O
from vfhe_amd import runtime as rt

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("N,split", [(256, 1), (4096, 1), (256, 4), (4096, 4)])
def test_slot_moves_match_oracle(lib, N, split):
    L = 4
    primes = O.gen_primes(N, [50] * L, split)
    pool = rt.Pool(N, primes, split)
    rng = np.random.default_rng(N + split)
    mask = 0b1011
    x = rt.random_rows(rng, primes, N, mask)
    ps = N // split
    p = pool.poly(x, mask)
    for slot in (0, 1, ps - 1):
        o = pool.poly(None, mask)
        lib.polynomial_RNS_broadcast_slot(o, p, slot)
        assert pool.mask(o) == mask
        assert np.array_equal(pool.read(o)[:L], O.slot_broadcast(x, mask, split, slot)[:L]), ("broadcast", slot)
        pool.free(o)
    for rot in (1, 3, ps - 1):
        o = pool.poly(None, mask)
        lib.polynomial_RNS_rotate_slot(o, p, rot)
        assert np.array_equal(pool.read(o)[:L], O.slot_rotate(x, mask, split, rot)[:L]), ("rotate", rot)
        pool.free(o)
    y = rt.random_rows(rng, primes, N, mask)
    o = pool.poly(y, mask)
    lib.polynomial_RNS_copy_slot(o, 2, p, ps - 2)
    assert np.array_equal(pool.read(o)[:L], O.slot_copy(y, 2, x, mask, split, ps - 2)[:L])
    pool.free(o)
    pool.free(p)


@pytest.mark.parametrize("N", [256, 8192])
def test_slot_inverse_matches_oracle(lib, N):
    L = 3
    primes = O.gen_primes(N, [50] * L)
    pool = rt.Pool(N, primes)
    rng = np.random.default_rng(N)
    mask = 0b101
    x = rt.random_rows(rng, primes, N, mask)
    x[x == 0] = 1
    p, o = pool.poly(x, mask), pool.poly(None, mask)
    assert lib.polynomial_RNS_inverse(o, p) == 0
    st, ref = O.slot_inverse(x, mask, primes)
    assert st == 0 and np.array_equal(pool.read(o)[:L], ref[:L])
    x[2, 17] = 0
    pool.write(p, x)
    assert lib.polynomial_RNS_inverse(o, p) == -2
    pool.free(p)
    pool.free(o)
    primes4 = O.gen_primes(256, [50] * 2, 4)
    pool4 = rt.Pool(256, primes4, 4)
    a, b = pool4.poly(rt.random_rows(rng, primes4, 256)), pool4.poly()
    assert lib.polynomial_RNS_inverse(b, a) == -1
    pool4.free(a)
    pool4.free(b)
