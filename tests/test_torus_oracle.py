"""Pins the torus oracle (oracle/torus.py) before it checks the GPU's CGGI over the 32-bit torus.

There is no reference implementation of this variant (the reference's CGGI16 runs over RNS
primes), so every check is against something the oracle does not itself compute: direct
evaluation of polynomials at the roots of X^N + 1, schoolbook products in Python integers,
exact rationals for the rounding, recomposition of the digits, and decryption of every
entry of a random lookup table after a full bootstrap.  One test plants a mistake (a cyclic
instead of negacyclic product) in a copy of an oracle routine to show the check catches it.
"""
from fractions import Fraction
import random

import numpy as np
import pytest

from oracle import torus as T
from vfhe_amd.torus import TorusKeys, TorusParams, pack_lut

R = random.Random(0x7043)


@pytest.mark.parametrize("N", [16, 64, 1024])
def test_fft_is_evaluation_at_the_negacyclic_roots(N):
    rng = np.random.default_rng(N)
    p = rng.integers(-(1 << 31), 1 << 31, size=N).astype(np.float64)
    X = T.negacyclic_fft(p)
    M = N // 2
    for k in range(0, M, max(1, M // 16)):
        x = np.exp(1j * np.pi / N) * np.exp(-2j * np.pi * k / M)
        assert abs(x ** N + 1) < 1e-9
        direct = np.polyval(p[::-1], x)
        assert abs(X[k] - direct) <= 1e-12 * np.abs(p).sum() * N ** 0.5, (N, k)


@pytest.mark.parametrize("N", [16, 64])
def test_product_equals_schoolbook(N):
    for _ in range(5):
        a = [R.randrange(-64, 64) for _ in range(N)]
        b = [R.randrange(-(1 << 31), 1 << 31) for _ in range(N)]
        assert list(T.negacyclic_mul(np.array(a), np.array(b))) == T.schoolbook_mul(a, b)


def test_product_n1024_equals_integer_convolution():
    rng = np.random.default_rng(5)
    N = 1024
    d = rng.integers(-64, 64, size=(6, N))
    b = rng.integers(-(1 << 31), 1 << 31, size=(6, N))
    want = np.zeros(N, dtype=np.int64)
    for r in range(6):
        full = np.convolve(d[r], b[r])                  # exact in int64: < 2^48 per term sum
        want += full[:N]
        want[: N - 1] -= full[N:]
    got = T.round_exact(T.negacyclic_ifft((T.negacyclic_fft(d) * T.negacyclic_fft(b)).sum(axis=0)))
    assert np.array_equal(got, want)


def test_planted_cyclic_product_fails():
    N = 16
    a = [R.randrange(-64, 64) for _ in range(N)]
    b = [R.randrange(-1000, 1000) for _ in range(N)]
    cyclic = np.rint(np.fft.ifft(np.fft.fft(a) * np.fft.fft(b)).real).astype(np.int64)
    assert list(cyclic) != T.schoolbook_mul(a, b)


@pytest.mark.parametrize("ell,B", [(3, 7), (2, 10), (4, 8), (1, 16)])
def test_decomposition_recomposes_and_is_bounded(ell, B):
    xs = [0, 1, (1 << 32) - 1, 1 << 31, (1 << 31) - 1] + [R.randrange(1 << 32) for _ in range(2000)]
    sh = 32 - B * ell
    for x in xs[:5]:
        xs.append((x + (1 << max(sh - 1, 0))) & ((1 << 32) - 1))   # around the rounding points
    x = np.array(xs, dtype=np.int64)
    d = T.decompose(x, ell, B)
    assert d.shape == (ell, len(xs))
    assert d.min() >= -(1 << (B - 1)) and d.max() < (1 << (B - 1))
    rec = sum(d[l].astype(object) * (1 << (32 - B * (l + 1))) for l in range(ell))
    for xi, ri in zip(xs, rec):
        rounded = ((xi + (1 << (sh - 1))) >> sh << sh) if sh > 0 else xi
        assert (ri - rounded) % (1 << 32) == 0, (xi, ri)


def test_mod_switch_matches_exact_rationals_with_ties_to_even():
    two_n = 2048
    xs = [0, 1, (1 << 32) - 1, 1 << 20, 3 << 20, 5 << 20, (1 << 32) + (1 << 20)]
    xs += [R.randrange(1 << 33) for _ in range(2000)]
    xs += [(R.randrange(1 << 12) << 21) | (1 << 20) for _ in range(200)]   # exact halves
    got = T.mod_switch(np.array(xs, dtype=np.uint64), two_n)
    for x, g in zip(xs, got):
        assert g == round(Fraction(x * two_n, 1 << 32)) % two_n, x


def test_mul_by_xai_equals_monomial_product():
    N = 32
    p = [R.randrange(1 << 32) for _ in range(N)]
    for a in [0, 1, 5, N - 1, N, N + 3, 2 * N - 1]:
        mono = [0] * N
        mono[a % N] = 1 if a < N else -1
        want = [v % (1 << 32) for v in T.schoolbook_mul(mono, p)]
        assert list(T.mul_by_xai(np.array(p), a)) == want, a


def test_extraction_keeps_the_phase_of_coefficient_zero():
    rng = np.random.default_rng(9)
    N = 64
    acc = rng.integers(0, 1 << 32, size=(2, N))
    s = rng.integers(0, 2, size=N)
    ph = (acc[1].astype(object) - np.array(T.schoolbook_mul(list(acc[0]), list(s)), dtype=object)) % (1 << 32)
    assert T.phase(T.extract(acc), s) == int(ph[0])


def test_external_product_multiplies_the_phase_by_the_key_bit():
    p = TorusParams(n=4)
    keys = TorusKeys(p, seed=3)
    kf = T.key_fft(keys.bootstrap_key())
    rng = np.random.default_rng(4)
    a = rng.integers(0, 1 << 32, size=p.N)
    m = rng.integers(0, 1 << 32, size=p.N)
    acc = np.stack([a, (T.negacyclic_mul(T.center(a), keys.glwe) + m) & (T.Q - 1)])
    for i in range(p.n):
        out = T.external_product(kf[i], acc, p.ell, p.base_log)
        ph = (out[1] - T.negacyclic_mul(T.center(out[0]), keys.glwe)) & (T.Q - 1)
        err = T.center(ph - keys.lwe[i] * m)
        assert np.abs(err).max() < 1 << 22, i   # decomposition and key noise only


def test_oracle_bootstrap_decrypts_every_lut_entry():
    """test_cggi16.py:20-62's check on the torus: each message m bootstraps to lut[m]."""
    p = TorusParams(n=48)
    keys = TorusKeys(p, seed=11)
    kf = T.key_fft(keys.bootstrap_key())
    prec, size = 5, 16
    lut = [R.randrange(16) for _ in range(size)]
    tv = T.pack_lut(lut, size, prec, p.N)
    assert np.array_equal(tv, pack_lut(lut, size, prec, p.N).astype(np.int64))
    msgs = list(range(16))
    lw = keys.encrypt(msgs, prec)
    for k, m in enumerate(msgs):
        ex = T.extract(T.blind_rotate(lw[k].astype(np.int64), tv, kf, p.ell, p.base_log, 1 << (prec - 1)))
        ph = T.phase(ex, keys.glwe)
        assert ((ph * (1 << prec) + (1 << 31)) >> 32) % (1 << prec) == lut[m], m


def test_c_port_equals_numpy_oracle():
    """oracle/torus_port.c (the CPU baseline) computes the same bootstrap bit for bit, threaded."""
    p = TorusParams(n=24)
    keys = TorusKeys(p, seed=51)
    bsk = keys.bootstrap_key()
    kf = T.key_fft(bsk)
    tv = T.pack_lut([(5 * m + 2) % 16 for m in range(16)], 16, 5, p.N)
    lw = keys.encrypt(list(range(16)), 5)
    lw[3, :-1] = 1 << 20                          # exact halves in every amount
    got = T.port_bootstrap(lw, tv, bsk, p.ell, p.base_log, 16, threads=4)
    for k in range(len(lw)):
        want = T.extract(T.blind_rotate(lw[k].astype(np.int64), tv, kf, p.ell, p.base_log, 16))
        assert np.array_equal(got[k].astype(np.int64), want), k
