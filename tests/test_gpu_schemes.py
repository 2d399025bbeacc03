"""The reference's Python-level tests, run on the hip engine through the host mirror.

Ports modules/arith/python/test/test_arith.py, modules/mlwe/python/test/test_mlwe.py
and modules/fhe/python/test/test_ckks.py: same parameters, same assertions
(exact decryption after round division, error bounds, CKKS tolerances), with
keys and encryption randomness from the seeded host sampler.
"""
import math
import random

import numpy as np
import pytest

from vfhe_amd import arith
from vfhe_amd.arith import Polynomial, Ring, repr
from vfhe_amd.ckks import CKKS_Scheme
from vfhe_amd.mlwe import MLWE_Scheme

pytestmark = pytest.mark.gpu


@pytest.fixture(autouse=True)
def fresh_pool():
    """Reset the global prime pool and the sampler between tests (reference conftest.py:45-85)."""
    arith.NTT_processor_instance.reset()
    arith.set_seed(0xC0FFEE)
    yield


# --- test_arith.py ------------------------------------------------------------------------------
NA = 16
rng = random.Random(0xC0FFEE)


def negacyclic_mul(a, b, q, n):
    out = [0] * n
    for i in range(n):
        for j in range(n):
            k = (i + j) % n
            s = (a[i] * b[j]) % q
            out[k] = (out[k] + (q - s if i + j >= n else s)) % q
    return out


@pytest.fixture
def ring():
    return Ring(NA, prime_size=[30, 30], split_degree=1)


def test_ring_and_roundtrip(ring):
    assert ring.ell == 2
    v = [i * 7 + 1 for i in range(NA)]
    assert Polynomial(ring).from_array(v).get_polynomial() == [x % ring.q_l for x in v]


def test_multiply_matches_schoolbook(ring):
    a_c = [i + 1 for i in range(NA)]
    b_c = [3 * i + 2 for i in range(NA)]
    a, b = Polynomial(ring).from_array(a_c), Polynomial(ring).from_array(b_c)
    assert (a * b).get_polynomial() == negacyclic_mul(a_c, b_c, ring.q_l, NA)


def test_add_sub_negate_scale(ring):
    a_c = [i * 7 + 1 for i in range(NA)]
    b_c = [i * i + 3 for i in range(NA)]
    a, b = Polynomial(ring).from_array(a_c), Polynomial(ring).from_array(b_c)
    q = ring.q_l
    assert (a + b).get_polynomial() == [(x + y) % q for x, y in zip(a_c, b_c)]
    assert (a - b).get_polynomial() == [(x - y) % q for x, y in zip(a_c, b_c)]
    assert (-a).get_polynomial() == [(-x) % q for x in a_c]
    assert (a * 5).get_polynomial() == [(5 * x) % q for x in a_c]
    assert (a + 9).get_polynomial() == [(a_c[0] + 9) % q] + [x % q for x in a_c[1:]]


def test_ntt_roundtrip(ring):
    v = [rng.randrange(ring.q_l) for _ in range(NA)]
    p = Polynomial(ring).from_array(v)
    p.to_coeff()
    p.to_NTT()
    p.to_coeff()
    assert p.get_polynomial() == [x % ring.q_l for x in v]


def test_automorphism_composition(ring):
    v = [rng.randrange(ring.q_l) for _ in range(NA)]
    p = Polynomial(ring).from_array(v)
    assert p.automorphism(3).automorphism(11).get_polynomial() == p.get_polynomial()


def test_copy_and_eq(ring):
    a = ring.random_element()
    assert a == a.copy()


def test_split_degree_ring_product():
    """Default split degree (incomplete NTT): the ring product still equals the schoolbook one."""
    r = Ring(64, prime_size=[49, 49])
    assert r.split_degree == 4
    a_c = [rng.randrange(1000) for _ in range(64)]
    b_c = [rng.randrange(1000) for _ in range(64)]
    a, b = Polynomial(r).from_array(a_c), Polynomial(r).from_array(b_c)
    assert (a * b).get_polynomial() == negacyclic_mul(a_c, b_c, r.q_l, 64)


# --- test_mlwe.py -------------------------------------------------------------------------------
N = 256


@pytest.fixture
def bv():
    Rq = Ring(N, prime_size=[45, 45, 45], split_degree=1)
    return Rq, Rq.quotient_ring(ell=1), MLWE_Scheme(Rq, special_primes=0, module_rank=1)


@pytest.fixture
def ghs():
    Rq = Ring(N, prime_size=[45, 45, 45, 50], split_degree=1)
    return Rq, Rq.quotient_ring(ell=1), MLWE_Scheme(Rq, special_primes=1, module_rank=1)


def enc(scheme, Rp, m, key):
    delta = scheme.rings[0].modulus_ratio(Rp, return_pointer=True)
    return scheme.sample(m.scaled_lift(scheme.rings[0], delta=delta), key)


def _mul_error(Rq, Rp, scheme, m_out, m1, m2):
    ell_ns = Rq.ell - scheme.special_primes
    P = math.prod(Rq.primes[Rp.ell: ell_ns])
    P_poly = Polynomial(Rp).from_bigint_array([P] + [0] * (Rp.N - 1))
    diff = (m_out - m1 * m2 * P_poly).get_polynomial(signed=True)
    return max(abs(c) for c in diff)


def test_encrypt_decrypt_add_sub_mul(bv):
    _, Rp, scheme = bv
    key = scheme.key_gen_sparse(N // 8, 3.2)
    m0, m1 = Rp.random_element(), Rp.random_element()
    c0, c1 = enc(scheme, Rp, m0, key), enc(scheme, Rp, m1, key)
    assert scheme.phase(c0, key).round_division(Rp) == m0
    assert scheme.phase(c1, key).round_division(Rp) == m1
    assert scheme.phase(c0 + c1, key).round_division(Rp) == m0 + m1
    assert scheme.phase(c0 - c1, key).round_division(Rp) == m0 - m1
    z = Rp.random_element()
    assert scheme.phase(c0 * z, key).round_division(Rp) == m0 * z


@pytest.mark.parametrize("fixture", ["bv", "ghs"])
def test_keyswitch(fixture, request):
    _, Rp, scheme = request.getfixturevalue(fixture)
    key = scheme.key_gen_sparse(N // 8, 3.2)
    m0 = Rp.random_element()
    c0 = enc(scheme, Rp, m0, key)
    key2 = scheme.key_gen_sparse(N // 8, 3.2)
    c_out = scheme.keyswitch(c0, scheme.gen_ksk(key2, key))
    assert scheme.phase(c_out, key2).round_division(Rp) == m0


def test_ghs_automorphism(ghs):
    _, Rp, scheme = ghs
    key = scheme.key_gen_sparse(N // 8, 3.2)
    m0 = Rp.random_element()
    c0 = enc(scheme, Rp, m0, key)
    c_out = scheme.automorphism(c0, 5, scheme.gen_ksk_automorphism(key, key, 5))
    assert scheme.phase(c_out, key).round_division(Rp) == m0.automorphism(5)


@pytest.mark.parametrize("fixture", ["bv", "ghs"])
def test_mlwe_multiplication(fixture, request):
    Rq, Rp, scheme = request.getfixturevalue(fixture)
    key = scheme.key_gen_sparse(N // 8, 3.2)
    s0 = key.poly[0]
    scheme.rlk = scheme.gen_rlk(key, [-(s0 * s0)])
    r = np.random.default_rng(1)
    m1 = Polynomial(Rp).from_array([int(x) for x in r.integers(-1, 2, N)])
    m2 = Polynomial(Rp).from_array([int(x) for x in r.integers(-1, 2, N)])
    c1, c2 = enc(scheme, Rp, m1, key), enc(scheme, Rp, m2, key)
    m_out = scheme.phase(c1 * c2, key).round_division(Rp)
    assert _mul_error(Rq, Rp, scheme, m_out, m1, m2) < 1000


@pytest.mark.parametrize("fixture", ["bv", "ghs"])
def test_mlwe_multiplication_deferred_relinearization(fixture, request):
    Rq, Rp, scheme = request.getfixturevalue(fixture)
    key = scheme.key_gen_sparse(N // 8, 3.2)
    s0 = key.poly[0]
    rlk = scheme.gen_rlk(key, [-(s0 * s0)])
    r = np.random.default_rng(2)
    m1 = Polynomial(Rp).from_array([int(x) for x in r.integers(-1, 2, N)])
    m2 = Polynomial(Rp).from_array([int(x) for x in r.integers(-1, 2, N)])
    c1, c2 = enc(scheme, Rp, m1, key), enc(scheme, Rp, m2, key)
    c_ext = scheme.multiply(c1, c2, None)
    assert c_ext.is_extended and c_ext.r == scheme.extended_rank
    c_rel = scheme.relinearize(c_ext, rlk)
    assert c_rel.r == scheme.r
    m_out = scheme.phase(c_rel, key).round_division(Rp)
    assert _mul_error(Rq, Rp, scheme, m_out, m1, m2) < 1000


@pytest.mark.parametrize("r, N_r", [(1, 256), (2, 128), (4, 64)])
def test_tensor_product_slot_layout(r, N_r):
    Rq = Ring(N_r, prime_size=[45, 45, 45], split_degree=1)
    Rp = Rq.quotient_ring(ell=1)
    scheme = MLWE_Scheme(Rq, special_primes=0, module_rank=r)
    key = scheme.key_gen_sparse(N_r * r // 8, 3.2)
    c1, c2 = enc(scheme, Rp, Rp.random_element(), key), enc(scheme, Rp, Rp.random_element(), key)
    slots = scheme.tensor_product(c1, c2)
    assert len(slots) == scheme.extended_rank + 1
    ext_key = scheme.quadratic_key_polys(key) + key.poly
    ext_phase = slots[-1]
    for slot, t in zip(slots[:-1], ext_key):
        ext_phase = ext_phase - slot * t
    expected = scheme.phase(c1, key) * scheme.phase(c2, key)
    ext_phase.to_coeff()
    expected.to_coeff()
    assert ext_phase == expected


# --- test_ckks.py -------------------------------------------------------------------------------

def rand_values(n, seed=5):
    r = np.random.default_rng(seed)
    return [complex(x, y) for x, y in zip(r.uniform(-1, 1, n), r.uniform(-1, 1, n))]


def test_encode_decode():
    scheme = CKKS_Scheme(Ring(N, 300, split_degree=1), scaling_factor=2**30)
    values = rand_values(N // 2)
    dec = scheme.decode(scheme.encode(values))
    assert all(abs(v - d) < 1e-3 for v, d in zip(values, dec))


def test_encrypt_decrypt():
    scheme = CKKS_Scheme(Ring(N, 300, split_degree=1), scaling_factor=2**30)
    key = scheme.key_gen_sparse(N // 8, 3.2)
    values = rand_values(N // 2)
    dec = scheme.decode(scheme.decrypt(scheme.encrypt(scheme.encode(values), key), key))
    assert all(abs(v - d) < 0.05 for v, d in zip(values, dec))


@pytest.mark.parametrize("k", [1, 3, 7])
def test_rotation(k):
    scheme = CKKS_Scheme(Ring(N, 300, split_degree=1), scaling_factor=2**25, special_primes=1)
    key = scheme.key_gen_sparse(N // 8, 3.2)
    values = rand_values(N // 2)
    ct = scheme.encrypt(scheme.encode(values), key)
    dec = scheme.decode(scheme.decrypt(scheme.rotate(ct, k, scheme.gen_rotation_key(key, k)), key))
    M = N // 2
    assert all(abs(dec[i] - values[(i + k) % M]) < 0.05 for i in range(M))


def test_keyswitch_ghs():
    scheme = CKKS_Scheme(Ring(N, 300, split_degree=1), scaling_factor=2**25, special_primes=1)
    key_in, key_out = scheme.key_gen_sparse(N // 8, 3.2), scheme.key_gen_sparse(N // 8, 3.2)
    values = rand_values(N // 2)
    ct = scheme.encrypt(scheme.encode(values), key_in)
    dec = scheme.decode(scheme.decrypt(scheme.keyswitch(ct, scheme.gen_ksk(key_out, key_in)), key_out))
    assert all(abs(v - d) < 0.05 for v, d in zip(values, dec))


@pytest.mark.parametrize("special_primes", [0, 1])
def test_ciphertext_multiplication(special_primes):
    scheme = CKKS_Scheme(Ring(N, 300, split_degree=1), scaling_factor=2**49, special_primes=special_primes)
    key = scheme.key_gen_sparse(N // 8, 3.2)
    s0 = key.poly[0]
    scheme.rlk = scheme.gen_rlk(key, [-(s0 * s0)])
    v1 = [complex(0.5, 0.5) if i == 0 else 0 for i in range(N // 2)]
    v2 = [complex(0.4, -0.4) if i == 0 else 0 for i in range(N // 2)]
    c1, c2 = scheme.encrypt(scheme.encode(v1), key), scheme.encrypt(scheme.encode(v2), key)
    c_mul = c1 * c2
    dec = scheme.decode(scheme.decrypt(c_mul, key), scaling_factor=c_mul.delta)
    assert all(abs(a * b - d) < 0.05 for a, b, d in zip(v1, v2, dec))


def test_ciphertext_multiplication_dense_slots():
    """All slots populated, uniform in the unit disk (the benchmarked CKKS message distribution)."""
    scheme = CKKS_Scheme(Ring(N, 300, split_degree=1), scaling_factor=2**40, special_primes=1)
    key = scheme.key_gen_sparse(N // 8, 3.2)
    s0 = key.poly[0]
    scheme.rlk = scheme.gen_rlk(key, [-(s0 * s0)])
    r = np.random.default_rng(3)
    rad, ang = np.sqrt(r.uniform(0, 1, N // 2)), r.uniform(0, 2 * np.pi, N // 2)
    v1 = list(rad * np.exp(1j * ang))
    v2 = list(np.roll(v1, 7))
    c1, c2 = scheme.encrypt(scheme.encode(v1), key), scheme.encrypt(scheme.encode(v2), key)
    c_mul = c1 * c2
    dec = scheme.decode(scheme.decrypt(c_mul, key), scaling_factor=c_mul.delta)
    assert max(abs(a * b - d) for a, b, d in zip(v1, v2, dec)) < 1e-3
    assert c_mul.ring.ell == scheme.rings[0].ell - 1
