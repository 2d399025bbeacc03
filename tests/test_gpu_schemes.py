"""The reference's Python-level tests, run on the hip engine through the host layer.

Ports modules/arith/python/test/test_arith.py, modules/mlwe/python/test/test_mlwe.py
and modules/fhe/python/test/test_ckks.py (including the rational-rescale cases,
test_ckks.py:137-268): same ring parameters, same assertions (exact decryption
after round division, error bounds, CKKS tolerances).  The level chains, keys
and ciphertexts are vfhe_amd.fhe's (Chain, SecretKey, SwitchKey, Batch, Ckks);
secrets come from a seeded generator and encryption noise from the engine's
ChaCha20 sampler under a fixed seed.
"""
import math
import random

import numpy as np
import pytest

from vfhe_amd import _native, arith
from vfhe_amd.arith import Polynomial, Ring, repr
from vfhe_amd.fhe import Batch, Chain, Ckks, SecretKey, galois_key, relinearization_key, switching_key

pytestmark = pytest.mark.gpu


@pytest.fixture(autouse=True)
def fresh_pool():
    """Reset the global prime pool and the sampler between tests (reference conftest.py:45-85)."""
    arith.NTT_processor_instance.reset()
    arith.set_seed(0xC0FFEE)
    _native.load().vfhe_prng_set_deterministic_seed(0xC0FFEE)
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


def keygen(chain, seed=7, h=N // 8, ternary=True):
    return SecretKey.sparse(chain, h, 3.2, np.random.default_rng(seed), ternary=ternary)


@pytest.fixture
def bv():
    Rq = Ring(N, prime_size=[45, 45, 45], split_degree=1)
    return Rq, Rq.quotient_ring(ell=1), Chain.nested(Rq, special=0)


@pytest.fixture
def ghs():
    Rq = Ring(N, prime_size=[45, 45, 45, 50], split_degree=1)
    return Rq, Rq.quotient_ring(ell=1), Chain.nested(Rq, special=1)


def enc(chain, Rp, msgs, key):
    """Encryptions of messages of Rp scaled by Q/p into level 0 (test_mlwe.py's `encrypt`)."""
    delta = chain.levels[0].modulus_ratio(Rp, return_pointer=True)
    return Batch.encrypt(chain, key, [m.scaled_lift(chain.levels[0], delta=delta) for m in msgs])


def phase(b, key, k=0):
    return b.decrypt(key)[k]


def _mul_error(chain, Rp, m_out, m1, m2):
    P = chain.levels[0].modulus_ratio(Rp)
    P_poly = Polynomial(Rp).from_bigint_array([P] + [0] * (Rp.N - 1))
    diff = (m_out - m1 * m2 * P_poly).get_polynomial(signed=True)
    return max(abs(c) for c in diff)


def test_encrypt_decrypt_add_sub_mul(bv):
    _, Rp, chain = bv
    key = keygen(chain)
    m0, m1 = Rp.random_element(), Rp.random_element()
    c = enc(chain, Rp, [m0, m1], key)
    c0, c1 = c.split()
    assert phase(c0, key).round_division(Rp) == m0
    assert phase(c1, key).round_division(Rp) == m1
    assert phase(c0 + c1, key).round_division(Rp) == m0 + m1
    assert phase(c0 - c1, key).round_division(Rp) == m0 - m1
    z = Rp.random_element()
    assert phase(c0.mul_plain([z]), key).round_division(Rp) == m0 * z


@pytest.mark.parametrize("fixture", ["bv", "ghs"])
def test_keyswitch(fixture, request):
    _, Rp, chain = request.getfixturevalue(fixture)
    key = keygen(chain, 1)
    m0 = Rp.random_element()
    c0 = enc(chain, Rp, [m0], key)
    key2 = keygen(chain, 2)
    c_out = c0.keyswitch(switching_key(chain, key2, key))
    assert phase(c_out, key2).round_division(Rp) == m0


def test_ghs_automorphism(ghs):
    _, Rp, chain = ghs
    key = keygen(chain)
    m0 = Rp.random_element()
    c0 = enc(chain, Rp, [m0], key)
    c_out = c0.automorphism(5, galois_key(chain, key, 5))
    assert phase(c_out, key).round_division(Rp) == m0.automorphism(5)


@pytest.mark.parametrize("fixture", ["bv", "ghs"])
def test_mlwe_multiplication(fixture, request):
    _, Rp, chain = request.getfixturevalue(fixture)
    key = keygen(chain)
    rlk = relinearization_key(chain, key)
    r = np.random.default_rng(1)
    m1 = Polynomial(Rp).from_array([int(x) for x in r.integers(-1, 2, N)])
    m2 = Polynomial(Rp).from_array([int(x) for x in r.integers(-1, 2, N)])
    c1, c2 = enc(chain, Rp, [m1], key), enc(chain, Rp, [m2], key)
    m_out = phase(c1.multiply(c2, rlk), key).round_division(Rp)
    assert _mul_error(chain, Rp, m_out, m1, m2) < 1000


@pytest.mark.parametrize("fixture", ["bv", "ghs"])
def test_mlwe_multiplication_deferred_relinearization(fixture, request):
    _, Rp, chain = request.getfixturevalue(fixture)
    key = keygen(chain)
    rlk = relinearization_key(chain, key)
    r = np.random.default_rng(2)
    m1 = Polynomial(Rp).from_array([int(x) for x in r.integers(-1, 2, N)])
    m2 = Polynomial(Rp).from_array([int(x) for x in r.integers(-1, 2, N)])
    c1, c2 = enc(chain, Rp, [m1], key), enc(chain, Rp, [m2], key)
    c_ext = c1.tensor(c2)
    assert c_ext.rank == chain.extended_rank
    c_rel = c_ext.keyswitch(rlk)
    assert c_rel.rank == chain.rank
    m_out = phase(c_rel, key).round_division(Rp)
    assert _mul_error(chain, Rp, m_out, m1, m2) < 1000


@pytest.mark.parametrize("r, N_r", [(1, 256), (2, 128), (4, 64)])
def test_tensor_product_slot_layout(r, N_r):
    Rq = Ring(N_r, prime_size=[45, 45, 45], split_degree=1)
    Rp = Rq.quotient_ring(ell=1)
    chain = Chain.nested(Rq, special=0, rank=r)
    key = SecretKey.sparse(chain, N_r * r // 8, 3.2, np.random.default_rng(r))
    delta = chain.levels[0].modulus_ratio(Rp, return_pointer=True)
    c = Batch.encrypt(chain, key, [Rp.random_element().scaled_lift(chain.levels[0], delta=delta) for _ in range(2)])
    c1, c2 = c.split()
    slots = c1.tensor(c2)
    comps = slots.polys(0)
    assert len(comps) == chain.extended_rank + 1
    ring = chain.levels[0]
    ext_key = key.quadratic(ring) + [key.poly(i, ring) for i in range(r)]
    ext_phase = comps[-1]
    for slot, t in zip(comps[:-1], ext_key):
        ext_phase = ext_phase - slot * t
    expected = phase(c1, key) * phase(c2, key)
    ext_phase.to_coeff()
    expected.to_coeff()
    assert ext_phase == expected


def test_batched_operations_equal_member_by_member(ghs):
    """A batch's multiply / key-switch / automorphism equal the same operation on each member alone."""
    _, Rp, chain = ghs
    key = keygen(chain)
    rlk = relinearization_key(chain, key)
    gk = galois_key(chain, key, 5)
    c = enc(chain, Rp, [Rp.random_element() for _ in range(3)], key)
    d = enc(chain, Rp, [Rp.random_element() for _ in range(3)], key)
    both = [(c.multiply(d, rlk), lambda x, y: x.multiply(y, rlk)), (c.automorphism(5, gk), lambda x, y: x.automorphism(5, gk))]
    for whole, one in both:
        for k, (x, y) in enumerate(zip(c.split(), d.split())):
            alone = one(x, y)
            for p, q in zip(whole.polys(k), alone.polys(0)):
                assert p == q, k


# --- test_ckks.py -------------------------------------------------------------------------------

def rand_values(n, seed=5):
    r = np.random.default_rng(seed)
    return [complex(int(x), int(y)) for x, y in zip(r.integers(-2, 3, n), r.integers(-2, 3, n))]


def ckks(special=1, scale=2**25, ring=None):
    chain = Chain.nested(ring if ring is not None else Ring(N, 300, split_degree=1), special=special)
    return chain, Ckks(chain, scale)


def test_encode_decode():
    chain, ck = ckks()
    values = rand_values(N // 2)
    dec = ck.decode(ck.encode(values))
    assert all(abs(v - d) < 1e-3 for v, d in zip(values, dec))


def test_encrypt_decrypt():
    chain, ck = ckks()
    key = keygen(chain)
    values = rand_values(N // 2)
    dec = ck.decrypt(key, ck.encrypt(key, [values]))[0]
    assert all(abs(v - d) < 0.05 for v, d in zip(values, dec))


def test_encode_many_equals_encode():
    chain, ck = ckks()
    vs = [rand_values(N // 2, s) for s in (1, 2, 3)]
    for p, v in zip(ck.encode_many(vs), vs):
        assert p == ck.encode(v)


@pytest.mark.parametrize("k", [1, 3, 7])
def test_rotation(k):
    chain, ck = ckks()
    key = keygen(chain)
    values = rand_values(N // 2)
    ct = ck.encrypt(key, [values])
    dec = ck.decrypt(key, ck.rotate(ct, k, ck.rotation_key(key, k)))[0]
    M = N // 2
    assert all(abs(dec[i] - values[(i + k) % M]) < 0.05 for i in range(M))


def test_keyswitch_ghs():
    chain, ck = ckks()
    key_in, key_out = keygen(chain, 1), keygen(chain, 2)
    values = rand_values(N // 2)
    ct = ck.encrypt(key_in, [values])
    dec = ck.decrypt(key_out, ct.keyswitch(switching_key(chain, key_out, key_in)))[0]
    assert all(abs(v - d) < 0.05 for v, d in zip(values, dec))


@pytest.mark.parametrize("special", [0, 1])
def test_ciphertext_multiplication(special):
    chain, ck = ckks(special, 2**49)
    key = keygen(chain)
    rlk = relinearization_key(chain, key)
    v1 = [complex(0.5, 0.5) if i == 0 else 0 for i in range(N // 2)]
    v2 = [complex(0.4, -0.4) if i == 0 else 0 for i in range(N // 2)]
    c1, c2 = ck.encrypt(key, [v1]), ck.encrypt(key, [v2])
    c_mul = ck.multiply(c1, c2, rlk)
    assert c_mul.lvl == 1
    dec = ck.decrypt(key, c_mul)[0]
    assert all(abs(a * b - d) < 0.05 for a, b, d in zip(v1, v2, dec))


def test_ciphertext_plaintext_multiplication():
    chain, ck = ckks(0, 2**49)
    key = keygen(chain)
    v1 = [complex(0.5, 0.5) if i == 0 else 0 for i in range(N // 2)]
    v2 = [complex(0.4, -0.4) if i == 0 else 0 for i in range(N // 2)]
    c1 = ck.encrypt(key, [v1])
    c_mul = ck.mul_plain(c1, [ck.encode(v2)])
    dec = ck.decrypt(key, c_mul)[0]
    assert all(abs(a * b - d) < 0.05 for a, b, d in zip(v1, v2, dec))


def test_ciphertext_multiplication_dense_slots():
    """All slots populated, uniform in the unit disk (the benchmarked CKKS message distribution),
    a batch of three products in one engine call."""
    chain, ck = ckks(1, 2**40)
    key = keygen(chain)
    rlk = relinearization_key(chain, key)
    r = np.random.default_rng(3)
    v1s, v2s = [], []
    for _ in range(3):
        rad, ang = np.sqrt(r.uniform(0, 1, N // 2)), r.uniform(0, 2 * np.pi, N // 2)
        v = list(rad * np.exp(1j * ang))
        v1s.append(v)
        v2s.append(list(np.roll(v, 7)))
    c_mul = ck.multiply(ck.encrypt(key, v1s), ck.encrypt(key, v2s), rlk)
    for v1, v2, dec in zip(v1s, v2s, ck.decrypt(key, c_mul)):
        assert max(abs(a * b - d) for a, b, d in zip(v1, v2, dec)) < 1e-3
    assert c_mul.ring.ell == chain.levels[0].ell - 1


def _subring(Rq, *indices):
    mask = 0
    for i in indices:
        mask |= 1 << Rq.prime_indices[i]
    return Rq.quotient_ring(mask=mask)


def test_rational_rescale_shared_primes():
    """test_ckks.py:137-163: level 1 = {0,1,3} is not a sub-ring of level 0 = {0,1,2}."""
    Rq = Ring(N, split_degree=1, prime_size=[50, 50, 50, 30, 50])
    chain = Chain([_subring(Rq, 0, 1, 2), _subring(Rq, 0, 1, 3)],
                  [_subring(Rq, 0, 1, 2, 4), _subring(Rq, 0, 1, 3, 4)])
    ck = Ckks(chain, 2**50)
    key = keygen(chain)
    values = rand_values(N // 2)
    ct = ck.encrypt(key, [values])
    assert ct.lvl == 0
    res = ck.rational_rescale(ct)
    assert res.lvl == 1 and res.ring == chain.levels[1]
    dec = ck.decrypt(key, res)[0]
    assert all(abs(v - d) < 0.05 for v, d in zip(values, dec))


def test_rational_rescale_disjoint_primes():
    """test_ckks.py:166-190: levels {2,3} -> {0,1} share no prime."""
    Rq = Ring(N, split_degree=1, prime_size=[40, 40, 51, 58, 50])
    chain = Chain([_subring(Rq, 2, 3), _subring(Rq, 0, 1), _subring(Rq, 2)],
                  [_subring(Rq, 2, 3, 4), _subring(Rq, 0, 1, 4), _subring(Rq, 2, 4)])
    ck = Ckks(chain, 2**50)
    key = keygen(chain)
    values = rand_values(N // 2)
    res = ck.rational_rescale(ck.encrypt(key, [values]))
    assert res.lvl == 1
    dec = ck.decrypt(key, res)[0]
    assert all(abs(v - d) < 0.05 for v, d in zip(values, dec))


def test_multiplication_with_rational_rescale():
    """test_ckks.py:221-268 on a hand-built non-nested chain (the reference derives its chain with
    residue_selection, a module outside the path): levels {52,51} -> {37,37} -> {45} bits, each
    rescale dividing by about the 2^29 scale, plus a 50-bit special prime; (v0 v1)(v2 v3) has
    multiplicative depth 2 and every rescale is a rational one."""
    Rq = Ring(N, split_degree=1, prime_size=[52, 51, 37, 37, 45, 50])
    levels = [_subring(Rq, 0, 1), _subring(Rq, 2, 3), _subring(Rq, 4)]
    chain = Chain(levels, [_subring(Rq, 0, 1, 5), _subring(Rq, 2, 3, 5), _subring(Rq, 4, 5)])
    assert not levels[1].is_quotient_ring(levels[0]) and not levels[2].is_quotient_ring(levels[1])
    ck = Ckks(chain, 2**29)
    key = keygen(chain)
    rlk = relinearization_key(chain, key)
    v = [rand_values(N // 2, s) for s in range(4)]
    c = ck.encrypt(key, v)
    c0, c1, c2, c3 = c.split()
    p01, p23 = ck.multiply(c0, c1, rlk), ck.multiply(c2, c3, rlk)
    assert p01.lvl == 1 and p23.lvl == 1
    res = ck.multiply(p01, p23, rlk)
    assert res.lvl == 2
    dec = ck.decrypt(key, res)[0]
    assert all(abs(a * b * cc * d - x) < 0.05 for a, b, cc, d, x in zip(*v, dec))


# --- ciphertext and key files (the engine's block format; the reference has none) --------------
def test_ciphertext_and_key_files_round_trip(tmp_path):
    """A ciphertext batch and a relinearization key written to files and read back: the rows are
    equal bit for bit, the domain is kept, and a product with the loaded key and ciphertexts
    equals the product with the originals; a corrupted, truncated or other-ring block is refused."""
    chain, ck = ckks(1, 2**40)
    key = keygen(chain)
    rlk = relinearization_key(chain, key)
    vals = [rand_values(N // 2) for _ in range(3)]
    cts = ck.encrypt(key, vals)
    cts.to_ntt()
    cts.save(str(tmp_path / "cts.bin"))
    rlk.save(str(tmp_path / "rlk.bin"))
    back = Batch.load(str(tmp_path / "cts.bin"), chain, cts.lvl, delta=cts.delta)
    assert back.form == "ntt" and len(back) == 3
    for k in range(3):
        assert back.polys(k) == cts.polys(k)
    rlk2 = type(rlk).load(str(tmp_path / "rlk.bin"), chain)
    a = ck.multiply(cts, cts, rlk)
    b = ck.multiply(back, back, rlk2)
    for k in range(3):
        assert a.polys(k) == b.polys(k)
    dec = ck.decrypt(key, b)
    for v, d in zip(vals, dec):
        assert all(abs(x * x - y) < 0.05 for x, y in zip(v, d))
    data = bytearray(cts.to_bytes())
    data[-20] ^= 1
    with pytest.raises(ValueError, match="check word"):
        Batch.from_bytes(bytes(data), chain, cts.lvl)
    with pytest.raises(ValueError, match="truncated"):
        Batch.from_bytes(cts.to_bytes()[:-8], chain, cts.lvl)
    forged = np.frombuffer(cts.to_bytes(), dtype=np.uint64).copy()
    forged[2] = 1 << 31                        # a header whose sample count the data cannot hold
    with pytest.raises(ValueError, match="truncated"):
        Batch.from_bytes(forged.tobytes(), chain, cts.lvl)
    other = Ring(2 * N, prime_size=[40, 40], split_degree=1)   # another prime pool (another N)
    with pytest.raises(ValueError, match="another ring"):
        Batch.from_bytes(cts.to_bytes(), chain, cts.lvl, ring=other)
