"""GPU parity: CGGI over the 32-bit torus (vfhe_amd.torus, vfhe_hip_torus_*) against the oracle.

No reference implementation exists for this variant (the reference's CGGI16 runs over RNS
primes); the oracle (oracle/torus.py, pinned in tests/test_torus_oracle.py) is the checker.
Tolerances: the transform within 2^-40 of the float64 CPU transform, relative to the largest
output; every external product's float sum lies far closer than 1/2 to its exact integer, so
accumulators and extracted samples are compared bit for bit; and every output of the
n = 630, N = 1024, 1024-sample batch decrypts exactly (test_cggi16.py:20-62's assertion).
"""
import numpy as np
import pytest

from oracle import torus as T
from vfhe_amd import torus as VT
from vfhe_amd.torus import TorusBatch, TorusBootstrapper, TorusKeys, TorusParams

pytestmark = pytest.mark.gpu

PREC, SIZE = 5, 16          # 4-bit messages plus a padding bit, as test_cggi16.py
TORUS_BASE = 1 << (PREC - 1)


def test_transform_within_2_pow_minus_40_of_cpu():
    rng = np.random.default_rng(1)
    polys = np.concatenate([
        rng.integers(0, 1 << 32, size=(24, 1024)),                # torus rows (read as signed)
        rng.integers(-64, 64, size=(8, 1024)) & 0xFFFFFFFF,       # digit rows
        np.zeros((1, 1024), dtype=np.int64),
        np.full((1, 1024), 0x80000000, dtype=np.int64),           # -2^31 everywhere
    ]).astype(np.uint32)
    print("[torus] transform of", len(polys), "polynomials", flush=True)
    got = VT.torus_fft(polys)
    want = T.negacyclic_fft(T.center(polys))
    for k in range(len(polys)):
        scale = max(np.abs(want[k]).max(), 1.0)
        err = np.abs(got[k] - want[k]).max() / scale
        assert err <= 2.0 ** -40, (k, err)


def _setup(n, seed):
    p = TorusParams(n=n)
    keys = TorusKeys(p, seed=seed)
    bsk = keys.bootstrap_key()
    return p, keys, bsk, T.key_fft(bsk), TorusBootstrapper(bsk, p)


def test_blind_rotation_bit_exact_against_oracle():
    p, keys, bsk, kf, boot = _setup(24, 5)
    rng = np.random.default_rng(6)
    lut = rng.integers(0, 16, size=SIZE)
    tv = VT.pack_lut(lut, SIZE, PREC, p.N)
    msgs = np.arange(16)
    lw = keys.encrypt(msgs, PREC)
    lw[3, :-1] = 0                            # every amount zero: the rotated test vector alone
    lw[4, :-1] = 1 << 20                      # exact halves, rounded to even (0)
    lw[5, :-1] = 3 << 20                      # exact halves, rounded to even (2)
    print("[torus] blind rotation, n=24, 16 samples", flush=True)
    acc = TorusBatch.zeros(len(msgs), 2 * p.N)
    out = boot.bootstrap(TorusBatch.upload(lw), TorusBatch.upload(tv), TORUS_BASE, acc=acc)
    got_acc, got_out = acc.read(), out.read()
    for k in range(len(msgs)):
        want = T.blind_rotate(lw[k].astype(np.int64), tv.astype(np.int64), kf, p.ell, p.base_log, TORUS_BASE)
        assert np.array_equal(got_acc[k].reshape(2, p.N), want), k
        assert np.array_equal(got_out[k], T.extract(want)), k
    dec = keys.decrypt(got_out, PREC)
    for k in (0, 1, 2, 6, 7, 8, 9, 10, 11, 12, 13, 14, 15):
        assert dec[k] == lut[msgs[k]], k


def test_outputs_only_or_accumulators_only():
    p, keys, bsk, kf, boot = _setup(8, 7)
    tv = TorusBatch.upload(VT.pack_lut(list(range(16)), SIZE, PREC, p.N))
    lw = TorusBatch.upload(keys.encrypt([3, 9], PREC))
    acc = TorusBatch.zeros(2, 2 * p.N)
    print("[torus] accumulators only", flush=True)
    rc = VT.lib().vfhe_hip_torus_bootstrap(None, acc.handle, boot.handle, lw.handle, tv.handle, TORUS_BASE)
    assert rc == 0
    out = boot.bootstrap(lw, tv, TORUS_BASE)
    a, o = acc.read(), out.read()
    for k in range(2):
        assert np.array_equal(o[k], T.extract(a[k].reshape(2, p.N).astype(np.int64)))


def test_bad_arguments_are_rejected():
    p, keys, bsk, kf, boot = _setup(8, 8)
    L = VT.lib()
    tv = TorusBatch.upload(np.zeros(p.N, dtype=np.uint32))
    lw = TorusBatch.upload(keys.encrypt([1], PREC))
    wrong = TorusBatch.zeros(1, p.n + 2)
    out = TorusBatch.zeros(1, p.N + 1)
    small = TorusBatch.zeros(1, p.N)
    assert L.vfhe_hip_torus_bootstrap(out.handle, None, boot.handle, wrong.handle, tv.handle, TORUS_BASE) == -2
    assert L.vfhe_hip_torus_bootstrap(small.handle, None, boot.handle, lw.handle, tv.handle, TORUS_BASE) == -3
    assert L.vfhe_hip_torus_bootstrap(out.handle, None, boot.handle, lw.handle, tv.handle, 0) == -1
    assert L.vfhe_hip_torus_bootstrap(None, None, boot.handle, lw.handle, tv.handle, TORUS_BASE) == -1
    bad = np.zeros((2, 6, 2, 512), dtype=np.uint32)
    assert not L.vfhe_hip_torus_key_new(bad.ctypes.data, 2, 512, 3, 7)
    assert not L.vfhe_hip_torus_key_new(bad.ctypes.data, 2, 1024, 3, 11)
    assert not L.vfhe_hip_torus_batch_new(None, 0, 5)


def test_bootstrap_batch_n630_N1024_1024_samples():
    """1024 uniform 4-bit messages through the identity table at n = 630, N = 1024, base 2^7,
    3 levels: every output decrypts exactly; two are equal bit for bit to the oracle's."""
    p, keys, bsk, kf, boot = _setup(630, 12)
    rng = np.random.default_rng(13)
    msgs = rng.integers(0, 16, size=1024)
    tv = VT.pack_lut(list(range(16)), SIZE, PREC, p.N)
    lw = keys.encrypt(msgs, PREC)
    print("[torus] bootstrapping 1024 samples, n=630, N=1024", flush=True)
    out = boot.bootstrap(TorusBatch.upload(lw), TorusBatch.upload(tv), TORUS_BASE).read()
    dec = keys.decrypt(out, PREC)
    assert np.array_equal(dec, msgs), np.nonzero(dec != msgs)
    for k in (0, 777):
        want = T.extract(T.blind_rotate(lw[k].astype(np.int64), tv.astype(np.int64), kf, p.ell, p.base_log,
                                        TORUS_BASE))
        assert np.array_equal(out[k], want), k


@pytest.mark.parametrize("ell,base_log", [(1, 10), (2, 9), (4, 8)])
def test_other_gadgets_bit_exact(ell, base_log):
    """Every compiled level count (the kernel is instantiated for ell = 1..4) against the oracle,
    with digit widths whose products keep the float sums far inside the exact-integer range.
    One 10-bit level keeps too few bits for these messages to decrypt (its rounding noise is
    about 2^-6 of the torus), so that case checks the bits only."""
    p = TorusParams(n=10, ell=ell, base_log=base_log)
    keys = TorusKeys(p, seed=20 + ell)
    bsk = keys.bootstrap_key()
    kf = T.key_fft(bsk)
    boot = TorusBootstrapper(bsk, p)
    tv = VT.pack_lut([(3 * m + 1) % 16 for m in range(16)], SIZE, PREC, p.N)
    msgs = np.array([0, 5, 11, 15])
    lw = keys.encrypt(msgs, PREC)
    print(f"[torus] ell={ell} base_log={base_log}: 4 samples, n=10", flush=True)
    acc = TorusBatch.zeros(len(msgs), 2 * p.N)
    out = boot.bootstrap(TorusBatch.upload(lw), TorusBatch.upload(tv), TORUS_BASE, acc=acc).read()
    got = acc.read()
    for k in range(len(msgs)):
        want = T.blind_rotate(lw[k].astype(np.int64), tv.astype(np.int64), kf, ell, base_log, TORUS_BASE)
        assert np.array_equal(got[k].reshape(2, p.N), want), k
        assert np.array_equal(out[k], T.extract(want)), k
    if ell > 1:
        assert np.array_equal(keys.decrypt(out, PREC), (3 * msgs + 1) % 16)


def test_single_sample_and_other_torus_base():
    """One sample, 3-bit messages (torus_base 8, prec 4): the rounding offset follows torus_base."""
    p, keys, bsk, kf, boot = _setup(16, 31)
    prec, size = 4, 8
    tv = VT.pack_lut([7 - m for m in range(8)], size, prec, p.N)
    lw = keys.encrypt([6], prec)
    print("[torus] one sample, torus_base 8", flush=True)
    acc = TorusBatch.zeros(1, 2 * p.N)
    out = boot.bootstrap(TorusBatch.upload(lw), TorusBatch.upload(tv), 8, acc=acc).read()
    want = T.blind_rotate(lw[0].astype(np.int64), tv.astype(np.int64), kf, p.ell, p.base_log, 8)
    assert np.array_equal(acc.read()[0].reshape(2, p.N), want)
    assert keys.decrypt(out, prec)[0] == 1


def test_batch_larger_than_one_residency_round():
    """2,500 samples (more workgroups than the device holds at once) into one reused output batch:
    every output decrypts, spot samples equal the oracle, and a second call overwrites the first."""
    p, keys, bsk, kf, boot = _setup(6, 41)
    rng = np.random.default_rng(42)
    tv = TorusBatch.upload(VT.pack_lut(list(range(16)), SIZE, PREC, p.N))
    out = TorusBatch.zeros(2500, p.N + 1)
    for rep in range(2):
        msgs = rng.integers(0, 16, size=2500)
        lw = keys.encrypt(msgs, PREC)
        print(f"[torus] 2500 samples, call {rep}", flush=True)
        got = boot.bootstrap(TorusBatch.upload(lw), tv, TORUS_BASE, out=out).read()
        assert np.array_equal(keys.decrypt(got, PREC), msgs)
        tvh = VT.pack_lut(list(range(16)), SIZE, PREC, p.N).astype(np.int64)
        for k in (0, 1023, 1024, 2499):
            want = T.extract(T.blind_rotate(lw[k].astype(np.int64), tvh, kf, p.ell, p.base_log, TORUS_BASE))
            assert np.array_equal(got[k], want), (rep, k)
