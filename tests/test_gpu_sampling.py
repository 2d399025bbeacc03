"""On-device sampling: ChaCha20 keystream, uniform and Gaussian polynomials, MLWE samples.

The reference draws randomness from AES-CTR or BLAKE3 (prng.c:91-108), so no
engine reproduces another's stream; what is pinned here is the keystream
(RFC 8439 section 2.3.2 known answer), the exact map from keystream to
residues under a deterministic seed, the distributions, and the mask and
decryption behaviour of the MLWE sampling entry points (mlwe.c:70-83, 190-300).
"""
import ctypes
import math

import numpy as np
import pytest

from oracle import oracle as O
from vfhe_amd import _native, runtime as rt

pytestmark = pytest.mark.gpu

M32 = 0xFFFFFFFF


def _rotl(x, r):
    return ((x << r) | (x >> (32 - r))) & M32


def chacha20_block(key, ctr, nonce):
    """RFC 8439 block function with a 64-bit counter (words 12-13) and 64-bit nonce (words 14-15)."""
    st = [0x61707865, 0x3320646E, 0x79622D32, 0x6B206574] + list(key) + \
         [ctr & M32, ctr >> 32, nonce & M32, nonce >> 32]
    x = list(st)

    def qr(a, b, c, d):
        x[a] = (x[a] + x[b]) & M32; x[d] = _rotl(x[d] ^ x[a], 16)
        x[c] = (x[c] + x[d]) & M32; x[b] = _rotl(x[b] ^ x[c], 12)
        x[a] = (x[a] + x[b]) & M32; x[d] = _rotl(x[d] ^ x[a], 8)
        x[c] = (x[c] + x[d]) & M32; x[b] = _rotl(x[b] ^ x[c], 7)

    for _ in range(10):
        qr(0, 4, 8, 12); qr(1, 5, 9, 13); qr(2, 6, 10, 14); qr(3, 7, 11, 15)
        qr(0, 5, 10, 15); qr(1, 6, 11, 12); qr(2, 7, 8, 13); qr(3, 4, 9, 14)
    return [(a + b) & M32 for a, b in zip(x, st)]


def seed_key(seed):
    """splitmix64 expansion of a deterministic seed into the 8 key words."""
    words, s = [], seed
    for _ in range(4):
        s = (s + 0x9E3779B97F4A7C15) & 0xFFFFFFFFFFFFFFFF
        z = s
        z = ((z ^ (z >> 30)) * 0xBF58476D1CE4E5B9) & 0xFFFFFFFFFFFFFFFF
        z = ((z ^ (z >> 27)) * 0x94D049BB133111EB) & 0xFFFFFFFFFFFFFFFF
        z ^= z >> 31
        words += [z & M32, z >> 32]
    return words


def test_chacha20_rfc8439_known_answer(lib):
    key = (ctypes.c_uint32 * 8)(*[int.from_bytes(bytes(range(4 * i, 4 * i + 4)), "little") for i in range(8)])
    out = (ctypes.c_uint32 * 16)()
    # RFC 8439 2.3.2: counter 1, nonce 00:00:00:09:00:00:00:4a:00:00:00:00 -> words 12..15
    lib.vfhe_hip_chacha20_blocks(key, 0x4A000000, 0x0900000000000001, 1, out)
    exp = [0xE4E7F110, 0x15593BD1, 0x1FDD0F50, 0xC47120A3, 0xC7F4D1C7, 0x0368C033, 0x9AAA2204, 0x4E6CD4C3,
           0x466482D2, 0x09AA9F07, 0x05D7C214, 0xA2028BD9, 0xD19C12B5, 0xB94E16DE, 0xE883D0CB, 0x4E3C50A2]
    assert list(out) == exp
    assert chacha20_block(list(key), 0x0900000000000001, 0x4A000000) == exp
    many = (ctypes.c_uint32 * (16 * 300))()
    lib.vfhe_hip_chacha20_blocks(key, 7, 5, 300, many)
    for b in (0, 1, 299):
        assert list(many[16 * b:16 * b + 16]) == chacha20_block(list(key), 5 + b, 7)


def test_uniform_residues_follow_keystream(lib):
    """Row k of call c: coefficient 8b+j = (u64 word pair j of block k*N/8+b) * q >> 64."""
    N = 256
    primes = O.gen_primes(N, [50, 40, 60])
    pool = rt.Pool(N, primes)
    p = pool.poly(mask=0b101)
    lib.vfhe_prng_set_deterministic_seed(1234)
    lib.polynomial_gen_random_RNSc_polynomial(p)
    rows = pool.read(p)
    key = seed_key(1234)
    nb = N // 8
    for k, limb in enumerate([0, 2]):
        q = primes[limb]
        for b in (0, 3, nb - 1):
            blk = chacha20_block(key, k * nb + b, 0)
            for j in range(8):
                r = blk[2 * j] | (blk[2 * j + 1] << 32)
                assert int(rows[limb][8 * b + j]) == (r * q) >> 64
    assert not rows[1].any()
    # a second call takes the next nonce; reseeding repeats the first call
    lib.polynomial_gen_random_RNSc_polynomial(p)
    again = pool.read(p)
    assert not np.array_equal(again, rows)
    lib.vfhe_prng_set_deterministic_seed(1234)
    lib.polynomial_gen_random_RNSc_polynomial(p)
    assert np.array_equal(pool.read(p), rows)
    lib.vfhe_prng_clear_deterministic_seed()
    lib.polynomial_gen_random_RNSc_polynomial(p)
    assert not np.array_equal(pool.read(p), rows)
    pool.free(p)


@pytest.mark.parametrize("N", [4096, 65536])
def test_uniform_distribution(lib, N):
    primes = O.gen_primes(N, [50] * 4)
    pool = rt.Pool(N, primes)
    p = pool.poly()
    lib.vfhe_prng_set_deterministic_seed(N)
    lib.polynomial_gen_random_RNSc_polynomial(p)
    lib.vfhe_prng_clear_deterministic_seed()
    rows = pool.read(p)
    for q, row in zip(primes, rows):
        x = row.astype(np.float64) / q
        assert (row < np.uint64(q)).all()
        assert abs(x.mean() - 0.5) < 6 * math.sqrt(1 / 12 / N)
        hist = np.bincount((x * 16).astype(int), minlength=16)
        chi2 = float(((hist - N / 16) ** 2 / (N / 16)).sum())
        assert chi2 < 60.0, chi2        # 15 degrees of freedom: P(chi2 > 60) < 1e-6
    pool.free(p)


def test_gaussian_same_noise_on_every_limb(lib):
    N, sigma = 16384, 3.2
    primes = O.gen_primes(N, [50, 45, 60])
    pool = rt.Pool(N, primes)
    p = pool.poly()
    lib.polynomial_gen_gaussian_RNSc_polynomial(p, ctypes.c_double(sigma))
    rows = pool.read(p)
    vals = []
    for q, row in zip(primes, rows):
        v = row.astype(object)
        vals.append(np.array([int(a) if a < q // 2 else int(a) - q for a in v], dtype=np.int64))
    assert np.array_equal(vals[0], vals[1]) and np.array_equal(vals[0], vals[2])
    e = vals[0]
    assert abs(e.mean()) < 6 * sigma / math.sqrt(N)
    assert abs(e.std() - sigma) < 0.1 * sigma
    assert np.abs(e).max() < 12 * sigma
    pool.free(p)


def test_mlwe_sampling_decrypts(lib):
    """mlwe_RNSc_sample(out, key, m) then phase == m + e with |e| ~ sigma (mlwe.c:236-252)."""
    N, l, r = 1024, 3, 2
    primes = O.gen_primes(N, [50] * l)
    pool = rt.Pool(N, primes)
    key = lib.mlwe_new_RNS_gaussian_key(N, r, l, ctypes.c_double(1.0), pool.h, ctypes.c_double(3.2))
    assert key
    kst = ctypes.cast(key, ctypes.POINTER(_native.RNSMLWEKey)).contents
    assert (kst.N, kst.l, kst.r) == (N, l, r)
    s0 = np.ctypeslib.as_array(kst.s[0].contents.coeffs, shape=(N,)).view(np.int64)
    assert np.abs(s0).max() <= 8 and np.abs(s0).max() > 0
    full = (1 << l) - 1
    msg = rt.random_rows(np.random.default_rng(5), primes, N)
    m = pool.poly(msg)
    ct = pool.sample(r, full)
    lib.mlwe_RNSc_sample(ct, key, m)
    lib.mlwe_RNSc_to_RNS(ct, ct)
    ph = pool.poly()
    lib.mlwe_RNS_phase(ph, ct, key)
    lib.polynomial_RNS_to_RNSc(ph, ph)
    diff = pool.poly()
    lib.polynomial_sub_RNSc_polynomial(diff, ph, m)
    d = pool.read(diff)
    for q, row in zip(primes, d):
        c = np.array([int(a) if a < q // 2 else int(a) - q for a in row])
        assert np.abs(c).max() < 40
        assert c.std() > 1.0
    # mlwe.c:281-293: new samples of zero carry mask = key->l
    z = lib.mlwe_new_RNS_sample_of_zero(key)
    assert pool.sample_masks(z) == [l] * (r + 1)
    t = lib.mlwe_new_RNS_trivial_sample_of_zero(N, r, full, pool.h)
    assert not pool.read_sample(t).any()
    for c in (ct, z, t):
        lib.free_mlwe_RNS_sample(c)
    for p in (m, ph, diff):
        pool.free(p)
    lib.free_mlwe_RNS_key(key)
