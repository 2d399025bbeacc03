"""GPU parity: RNS-polynomial operations of the hip engine against the oracle.

Covers the rns_polynomial.c entry points on the path: pointwise products (with
the split-degree wrap), additive/scalar ops, ModUp digit lift, sequential
rounded/floor division (rescale / ModDown), fast base extension, scaled lift
and the automorphism.  Inputs are seeded uniform residues; every comparison is
bit-exact on the active limbs, and the output masks follow the reference.
"""
import numpy as np
import pytest

from oracle import oracle as O
from vfhe_amd import runtime as rt

pytestmark = pytest.mark.gpu

CASES = [(4096, 5, 1), (1024, 4, 4), (16, 3, 1)]


def setup(N, L, split, seed=1):
    primes = O.gen_primes(N, [50] * (L - 1) + [49], split)
    return primes, rt.Pool(N, primes, split), O.Ctx(N, primes, split), np.random.default_rng(seed)


def active_eq(got, ref, mask):
    for i in range(ref.shape[0]):
        if (mask >> i) & 1:
            assert np.array_equal(got[i], ref[i]), f"limb {i}"


@pytest.mark.parametrize("N,L,split", CASES)
def test_products(lib, N, L, split):
    primes, pool, octx, rng = setup(N, L, split)
    m1, m2 = pool.full_mask(), pool.full_mask() & ~0b100
    a, b, c = (rt.random_rows(rng, primes, N) for _ in range(3))
    pa, pb, pc = pool.poly(a, m1), pool.poly(b, m2), pool.poly(c, m1)
    out = pool.poly()
    for mode, fn in ((0, "polynomial_mul_RNS_polynomial"), (1, "polynomial_mul_addto_RNS_polynomial"),
                     (2, "polynomial_mul_subto_RNS_polynomial")):
        pool.write(out, c)
        ref = c.copy()
        mref = octx.mul(ref, a, m1, b, m2, mode)
        getattr(lib, fn)(out, pa, pb)
        assert pool.mask(out) == mref
        active_eq(pool.read(out), ref, mref)
    # multo: out *= in
    ref = c.copy()
    mref = octx.mul(ref, c, m1, a, m1, 0)
    lib.polynomial_multo_RNS_polynomial(pc, pa)
    assert pool.mask(pc) == mref
    active_eq(pool.read(pc), ref, mref)


@pytest.mark.parametrize("N,L,split", CASES)
def test_additive_and_scalar_ops(lib, N, L, split):
    primes, pool, octx, rng = setup(N, L, split, 2)
    m = pool.full_mask() & ~0b1
    a, b = rt.random_rows(rng, primes, N), rt.random_rows(rng, primes, N)
    pa, pb, out = pool.poly(a, m), pool.poly(b, pool.full_mask()), pool.poly()
    for sub, fn in ((False, "polynomial_add_RNSc_polynomial"), (True, "polynomial_sub_RNS_polynomial")):
        ref = np.zeros_like(a)
        mref = octx.addsub(ref, a, m, b, pool.full_mask(), sub)
        getattr(lib, fn)(out, pa, pb)
        assert pool.mask(out) == mref
        active_eq(pool.read(out), ref, mref)
    ref = np.zeros_like(a)
    octx.negate(ref, a, m)
    lib.polynomial_RNSc_negate(out, pa)
    active_eq(pool.read(out), ref, m)
    s = 0x9876543210ABCDEF
    ref = np.zeros_like(a)
    octx.scale(ref, a, m, s)
    lib.polynomial_scale_RNSc_polynomial(out, pa, s)
    active_eq(pool.read(out), ref, m)
    ref = b.copy()
    pool.write(out, b)
    octx.scale(ref, a, m, s, addto=True)
    lib.polynomial_scale_addto_RNSc_polynomial(out, pa, s)
    active_eq(pool.read(out), ref, m)
    svec = np.array([(s * (i + 3)) & 0xFFFFFFFFFFFFFFFF for i in range(L)], dtype=np.uint64)
    ref = np.zeros_like(a)
    octx.scale_rns(ref, a, m, svec)
    lib.polynomial_scale_RNS_polynomial_RNS(out, pa, svec.ctypes.data_as(rt.U64P))
    active_eq(pool.read(out), ref, m)
    for v in (12345, (-777) & 0xFFFFFFFFFFFFFFFF, 0, 2**63 + 5):
        for ntt_form, fn in ((False, "polynomial_RNSc_add_integer"), (True, "polynomial_RNS_add_integer")):
            ref = np.zeros_like(a)
            octx.add_integer(ref, a, m, v, ntt_form)
            getattr(lib, fn)(out, pa, v)
            active_eq(pool.read(out), ref, m)


@pytest.mark.parametrize("N,L,split", CASES)
def test_mod_reduce_lifted(lib, N, L, split):
    """ModUp digit (rns_polynomial.c:824-843): every target limb gets in[idx] mod q_i."""
    primes, pool, octx, rng = setup(N, L, split, 3)
    a = rt.random_rows(rng, primes, N)
    pa = pool.poly(a)
    for idx in range(L):
        out_mask = pool.full_mask() & ~(1 << ((idx + 1) % L))
        out = pool.poly(None, out_mask)
        ref = np.zeros_like(a)
        octx.mod_reduce_lifted(ref, out_mask, a, idx)
        lib.polynomial_RNSc_mod_reduce_lifted(out, pa, idx)
        active_eq(pool.read(out), ref, out_mask)
        pool.free(out)


@pytest.mark.parametrize("N,L,split", CASES + [(65536, 28, 1)])
@pytest.mark.parametrize("rnd", [True, False])
def test_division(lib, N, L, split, rnd):
    """Sequential round/floor division (rns_polynomial.c:905-976) by one, several and all-but-one primes."""
    primes, pool, octx, rng = setup(N, L, split, 4)
    full = pool.full_mask()
    divides = [1 << (L - 1), (1 << (L - 1)) | (1 << 1), full & ~1]
    if L >= 28:
        divides = [1 << 23, 0xF << 24]
    for dm in divides:
        a = rt.random_rows(rng, primes, N)
        p = pool.poly(a, full)
        ref = a.copy()
        mref = octx.division(ref, full, dm, rnd)
        fn = "polynomial_round_division_RNSc_wo_free" if rnd else "polynomial_floor_division_RNSc_wo_free"
        getattr(lib, fn)(p, dm)
        assert pool.mask(p) == mref
        got = pool.read(p)
        active_eq(got, ref, mref)
        for i in range(L):
            if (dm >> i) & 1:
                assert not got[i].any(), "divided limb is zeroed"
        pool.free(p)


@pytest.mark.parametrize("N,L,split", CASES)
def test_base_conversion_and_scaled_lift(lib, N, L, split):
    primes, pool, octx, rng = setup(N, L, split, 5)
    in_mask = 0b11
    out_mask = pool.full_mask()
    a = rt.random_rows(rng, primes, N, in_mask)
    pa = pool.poly(a, in_mask)
    out = pool.poly(None, out_mask)
    ref = np.zeros_like(a)
    octx.base_conversion(ref, out_mask, a, in_mask)
    params = lib.init_base_conversion_params(pool.h, in_mask, out_mask)
    lib.polynomial_base_conversion_RNSc(out, pa, params)
    assert pool.mask(out) == out_mask
    active_eq(pool.read(out), ref, out_mask)
    out2 = pool.poly(None, out_mask)
    lib.polynomial_base_conversion_RNSc(out2, pa, None)
    active_eq(pool.read(out2), ref, out_mask)
    lib.free_base_conversion_params(params)
    # scaled lift with computed and with explicit factors
    d = octx.scaling_factors(in_mask, out_mask)
    dl = np.zeros(L, dtype=np.uint64)
    lib.rns_compute_scaling_factors(dl.ctypes.data_as(rt.U64P), pool.h, in_mask, out_mask)
    assert np.array_equal(d, dl)
    ref = np.zeros_like(a)
    octx.scaled_lift(ref, out_mask, a, in_mask)
    lib.polynomial_RNSc_scaled_lift(out, pa, None)
    active_eq(pool.read(out), ref, out_mask)
    lib.polynomial_RNSc_scaled_lift(out2, pa, dl.ctypes.data_as(rt.U64P))
    active_eq(pool.read(out2), ref, out_mask)


@pytest.mark.parametrize("N,L,split", CASES + [(65536, 4, 1)])
def test_permute(lib, N, L, split):
    primes, pool, octx, rng = setup(N, L, split, 6)
    m = pool.full_mask() & ~0b10
    a = rt.random_rows(rng, primes, N, m)
    pa = pool.poly(a, m)
    out = pool.poly()
    for gen in (5, 3, 2 * N - 1, pow(5, 7, 2 * N)):
        ref = np.zeros_like(a)
        octx.permute(ref, a, m, gen)
        lib.polynomial_RNSc_permute(out, pa, gen)
        assert pool.mask(out) == m
        got = pool.read(out)
        assert np.array_equal(got[:L], ref), gen
