"""Trace and LWE packing (mlwe.c:328-445) on the device.

mlwe_partial_trace / mlwe_trace are sums of automorphisms X -> X^g, each key-switched back:
checked bit for bit against the same sequence made of the reference-named entry points
(mlwe_automorphism_RNSc_GHS, mlwe_add_RNSc_sample), and by decryption with valid Galois keys:
the trace over g = 2^k + 1 (k = 1..log N) maps sum_j m_j X^j to N m_0.  The packing key-switch
folds LWE samples into one MLWE sample whose phase holds the samples' phases as coefficients
0..size-1; with a valid packing key (digit keys of the LWE secret's coefficients as constant
polynomials) every coefficient decrypts to its message."""
import ctypes

import numpy as np
import pytest

from vfhe_amd import _native, arith
from vfhe_amd.arith import Polynomial, Ring
from vfhe_amd.cggi import LweBatch, LweKey
from vfhe_amd.fhe import Batch, Chain, SecretKey, SwitchKey, galois_key

pytestmark = pytest.mark.gpu
N = 256


@pytest.fixture
def ghs():
    arith.NTT_processor_instance.reset()
    _native.load().vfhe_prng_set_deterministic_seed(0x7ACE)
    Rq = Ring(N, prime_size=[45, 45, 45, 50], split_degree=1)
    chain = Chain.nested(Rq, special=1)
    key = SecretKey.sparse(chain, N // 8, 3.2, np.random.default_rng(11))
    return Rq, Rq.quotient_ring(ell=1), chain, key


def _enc(chain, Rp, m, key):
    delta = chain.levels[0].modulus_ratio(Rp, return_pointer=True)
    return Batch.encrypt(chain, key, [m.scaled_lift(chain.levels[0], delta=delta)])


def test_trace_decrypts_and_equals_composition(ghs):
    lib = _native.load()
    _, Rp, chain, key = ghs
    logN = N.bit_length() - 1
    gens = [(1 << (logN - i + 1)) + 1 for i in range(1, logN + 1)]
    gks = [galois_key(chain, key, g) for g in gens]
    ksks = _native.voidp_array([ctypes.cast(k.at(0), ctypes.c_void_p).value for k in gks])
    r = np.random.default_rng(3)
    m = Polynomial(Rp).from_array([int(x) for x in r.integers(-5, 6, N)])
    ct = _enc(chain, Rp, m, key)
    ct.to_coeff()
    out = Batch.alloc(chain, 0, 1, ring=chain.ext[0], form="coeff")
    lib.mlwe_trace(out.handles[0], ct.handles[0], ksks, 0)
    out.ring = chain.levels[0]
    got = out.decrypt(key)[0].round_division(Rp).get_polynomial(signed=True)
    m0 = m.get_polynomial(signed=True)[0]
    assert got[0] == N * m0 and not any(got[1:])
    # the same sum made of the reference-named entry points, one call at a time
    tmp = ct.copy()
    step = Batch.alloc(chain, 0, 1, ring=chain.ext[0], form="coeff")
    for g, k in zip(gens, gks):
        lib.mlwe_automorphism_RNSc_GHS(step.handles[0], tmp.handles[0], g, k.at(0), 0)
        lib.mlwe_add_RNSc_sample(tmp.handles[0], tmp.handles[0], step.handles[0])
    out.to_coeff()
    for p, q in zip(out.polys(0), tmp.polys(0)):
        assert p == q
    # a partial trace over the first three elements, against the same composition
    part = Batch.alloc(chain, 0, 1, ring=chain.ext[0], form="coeff")
    g3 = np.array(gens[:3], dtype=np.uint64)
    lib.mlwe_partial_trace(part.handles[0], ct.handles[0], g3.ctypes.data_as(_native.U64P),
                           _native.voidp_array([ctypes.cast(k.at(0), ctypes.c_void_p).value for k in gks[:3]]), 3, 0)
    tmp = ct.copy()
    for g, k in zip(gens[:3], gks[:3]):
        lib.mlwe_automorphism_RNSc_GHS(step.handles[0], tmp.handles[0], g, k.at(0), 0)
        lib.mlwe_add_RNSc_sample(tmp.handles[0], tmp.handles[0], step.handles[0])
    for p, q in zip(part.polys(0), tmp.polys(0)):
        assert p == q


@pytest.mark.parametrize("size", [1, 5, N])
def test_packing_keyswitch_decrypts(ghs, size):
    lib = _native.load()
    _, Rp, chain, key = ghs
    level = chain.levels[0]
    n = 64
    rng = np.random.default_rng(size)
    s_lwe = [int(v) for v in rng.integers(-1, 2, n)]
    lkey = LweKey(level, s_lwe)
    # packing key: for each LWE coefficient, the digit keys of the constant polynomial s_i
    pk = SwitchKey(chain, key, [Polynomial(chain.span).from_array([s] + [0] * (N - 1)) for s in s_lwe])
    p = Rp.primes[0]
    delta = level.q_l // p
    msgs = [int(v) for v in rng.integers(0, p, size)]
    lwes = LweBatch.encrypt(lkey, [m * delta for m in msgs])
    out = Batch.alloc(chain, 0, 1, form="ntt")   # the target mask: the level's limbs
    lib.mlwe_full_packing_keyswitch(out.handles[0], _native.voidp_array(lwes.handles), size, pk.at(0), 0)
    got = out.decrypt(key)[0].round_division(Rp).get_polynomial()
    assert got[:size] == [m % p for m in msgs]
    assert not any(got[size:])
