"""Serving runtime: results are the same alone and in the mix (bit-exact).

A seeded Poisson mix of CKKS multiply+relinearize+rescale and rotation requests
of two ring sizes and five levels runs through the engine's serving runtime
(continuous batching on several streams); every output is then recomputed alone
through the reference-named entry points (mlwe_multiply + round division,
ckks.py:128-148; mlwe_automorphism_RNSc_GHS, mlwe.c:314-326) and compared
bit for bit, and one request of each kind is checked against the oracle.
"""
import numpy as np
import pytest

from oracle import oracle as O
from vfhe_amd import _native, serving

pytestmark = pytest.mark.gpu

NS, L, K, LEVELS, STEPS = (8192, 16384), 8, 2, (4, 8), 2


@pytest.fixture(scope="module")
def rings(lib):
    rs = {N: serving.Ring(lib, N, L=L, K=K, rot_steps=STEPS, levels=LEVELS) for N in NS}
    yield rs
    lib.vfhe_hip_synchronize()
    for r in rs.values():
        r.free()


def alone(lib, ring, req):
    op, in1, in2, ksk, gen, dm = ring.args(req)
    out = ring.pool.sample(1, ring.pool.mask(_native_b(in1)))
    if op == serving.REQ_MULTIPLY:
        lib.mlwe_multiply(out, in1, in2, ksk)
        lib.mlwe_RNS_to_RNSc(out, out)
        lib.mlwe_round_division_RNSc(out, dm)
    else:
        lib.mlwe_automorphism_RNSc_GHS(out, in1, gen, ksk, 0)
    return out


def _native_b(c):
    import ctypes
    return ctypes.cast(c, ctypes.POINTER(_native.RNSMLWE)).contents.b


def test_results_same_alone_and_in_mix(lib, rings):
    # a burst (about 100 requests within 2 ms) so that requests queue behind running batches
    wl = serving.make_workload(50000.0, 0.002, NS, levels=LEVELS, rot_steps=STEPS, seed=3)
    assert len(wl) > 60
    print(f"[serving] {len(wl)} requests", flush=True)
    res = serving.run(lib, rings, wl, memory_budget=1 << 30, lanes=4, max_batch=8, keep_outputs=True)
    print({k: v for k, v in res.items() if k != "outputs"}, flush=True)
    assert res["requests"] == len(wl) and res["pending"] == 0
    assert res["batches"] < len(wl)          # requests were batched
    assert res["max_lanes_busy"] >= 2        # batches were in flight on several streams at once
    for k, (req, out) in enumerate(zip(wl, res["outputs"])):
        ring = rings[req.N]
        ref = alone(lib, ring, req)
        masks = ring.pool.sample_masks(out)
        assert masks == ring.pool.sample_masks(ref), k
        rows = [i for i in range(len(ring.primes)) if (masks[0] >> i) & 1]
        assert np.array_equal(ring.pool.read_sample(out)[:, rows], ring.pool.read_sample(ref)[:, rows]), (k, req)
        lib.free_mlwe_RNS_sample(ref)
        lib.free_mlwe_RNS_sample(out)


def test_budget_smaller_than_one_request_still_progresses(lib, rings):
    wl = serving.make_workload(2000.0, 0.01, NS[:1], levels=LEVELS, rot_steps=STEPS, seed=4)
    res = serving.run(lib, rings, wl, memory_budget=1, lanes=2, max_batch=4)
    assert res["requests"] == len(wl) and res["pending"] == 0
    assert res["mean_batch"] == 1.0          # nothing fits the budget, so each runs alone


def test_one_of_each_against_oracle(lib, rings):
    N = NS[0]
    ring = rings[N]
    c = O.Ctx(N, ring.primes)
    lvl = 5
    lm, ext = (1 << lvl) - 1, ((1 << lvl) - 1) | ring.special
    c1, c2 = ring.cts[lvl]
    x1, x2 = ring.pool.read_sample(c1), ring.pool.read_sample(c2)

    def key_np(arr):
        return np.ascontiguousarray(np.stack([ring.pool.read_sample(arr[d]) for d in range(lvl)]))

    # multiply + relinearize + rescale by the top level prime
    req = serving.Request(0.0, serving.REQ_MULTIPLY, N, lvl, 0)
    got = alone(lib, ring, req)
    out = np.zeros((2, len(ring.primes), N), dtype=np.uint64)
    m = c.multiply(out, x1, x2, 1, lm, [key_np(ring.rlk_top), None], ext)
    for comp in range(2):
        c.to_coeff(out[comp], out[comp], m)
        mm = c.division(out[comp], m, 1 << (lvl - 1))
    assert ring.pool.sample_masks(got) == [mm, mm]
    g = ring.pool.read_sample(got)
    for i in range(len(ring.primes)):
        if (mm >> i) & 1:
            assert np.array_equal(g[:, i], out[:, i]), i
    lib.free_mlwe_RNS_sample(got)
    # rotation by 5^(2^1)
    req = serving.Request(0.0, serving.REQ_ROTATE, N, lvl, 1)
    got = alone(lib, ring, req)
    out = np.zeros((2, len(ring.primes), N), dtype=np.uint64)
    m = c.automorphism(out, x1, 1, lm, ring.rot_gens[1], [key_np(ring.rot_top[1])], ext)
    assert ring.pool.sample_masks(got) == [m, m]
    g = ring.pool.read_sample(got)
    for i in range(len(ring.primes)):
        if (m >> i) & 1:
            assert np.array_equal(g[:, i], out[:, i]), i
    lib.free_mlwe_RNS_sample(got)


def test_submit_output_still_in_flight_on_lane0(lib, rings):
    """A ciphertext produced asynchronously on the caller's lane (lane 0, no synchronisation) is
    submitted at once: the batch's lane waits for the submit-time event, so the served result
    equals the one computed after a synchronisation."""
    N, lvl = NS[1], LEVELS[1]
    ring = rings[N]
    op, c1, c2, ksk, gen, dm = ring.args(serving.Request(0.0, serving.REQ_MULTIPLY, N, lvl, 0))
    srv = serving.Server(lib, 1 << 30, lanes=3, max_batch=4)
    try:
        x = ring.pool.sample(1, ring.pool.mask(_native_b(c1)))
        lib.mlwe_multiply(x, c1, c2, ksk)               # enqueued on lane 0, not waited for
        rid = srv.submit(op, x, c2, ksk, gen, dm)
        done = []
        while not done:
            done = srv.poll(wait=True)
        assert [d[0] for d in done] == [rid]
        got = done[0][1]
        lib.vfhe_hip_synchronize()
        ref = ring.pool.sample(1, ring.pool.mask(_native_b(c1)))
        lib.mlwe_multiply(ref, x, c2, ksk)
        lib.mlwe_RNS_to_RNSc(ref, ref)
        lib.mlwe_round_division_RNSc(ref, dm)
        masks = ring.pool.sample_masks(got)
        assert masks == ring.pool.sample_masks(ref)
        rows = [i for i in range(len(ring.primes)) if (masks[0] >> i) & 1]
        assert np.array_equal(ring.pool.read_sample(got)[:, rows], ring.pool.read_sample(ref)[:, rows])
        for c in (got, ref, x):
            lib.free_mlwe_RNS_sample(c)
    finally:
        lib.vfhe_hip_synchronize()
        srv.close()


BIG_NS, BIG_L, BIG_K, BIG_LEVELS = (32768, 65536), 24, 4, (22, 24)


@pytest.fixture(scope="module")
def big_rings(lib):
    rs = {N: serving.Ring(lib, N, L=BIG_L, K=BIG_K, rot_steps=STEPS, levels=BIG_LEVELS) for N in BIG_NS}
    yield rs
    lib.vfhe_hip_synchronize()
    for r in rs.values():
        r.free()


@pytest.mark.slow
def test_large_rings_same_alone_and_in_mix(lib, big_rings):
    """The benchmarked upper end of the mix: N = 2^15 and 2^16 at levels 22..24 of a 24 + 4 prime ring.
    Every output of the served mix equals the same request run alone, bit for bit."""
    wl = serving.make_workload(20000.0, 0.003, BIG_NS, levels=BIG_LEVELS, rot_steps=STEPS, seed=5)
    assert len(wl) > 30 and any(r.N == 65536 and r.level == 24 for r in wl)
    print(f"[serving large] {len(wl)} requests", flush=True)
    res = serving.run(lib, big_rings, wl, memory_budget=8 << 30, lanes=4, max_batch=8, keep_outputs=True)
    print({k: v for k, v in res.items() if k != "outputs"}, flush=True)
    assert res["requests"] == len(wl) and res["pending"] == 0
    assert res["batches"] < len(wl)
    for k, (req, out) in enumerate(zip(wl, res["outputs"])):
        ring = big_rings[req.N]
        ref = alone(lib, ring, req)
        masks = ring.pool.sample_masks(out)
        assert masks == ring.pool.sample_masks(ref), k
        rows = [i for i in range(len(ring.primes)) if (masks[0] >> i) & 1]
        assert np.array_equal(ring.pool.read_sample(out)[:, rows], ring.pool.read_sample(ref)[:, rows]), (k, req)
        lib.free_mlwe_RNS_sample(ref)
        lib.free_mlwe_RNS_sample(out)


@pytest.mark.slow
def test_level24_n65536_against_oracle(lib, big_rings):
    """A served multiply + relinearize + rescale at N = 2^16, level 24, against the oracle bit for bit."""
    N, lvl = 65536, 24
    ring = big_rings[N]
    c = O.Ctx(N, ring.primes)
    lm, ext = (1 << lvl) - 1, ((1 << lvl) - 1) | ring.special
    c1, c2 = ring.cts[lvl]
    x1, x2 = ring.pool.read_sample(c1), ring.pool.read_sample(c2)
    key = np.ascontiguousarray(np.stack([ring.pool.read_sample(ring.rlk_top[d]) for d in range(lvl)]))
    wl = [serving.Request(0.0, serving.REQ_MULTIPLY, N, lvl, 0)]
    print("[serving large] one multiply at N=65536 level 24 against the oracle", flush=True)
    res = serving.run(lib, big_rings, wl, memory_budget=8 << 30, lanes=2, max_batch=4, keep_outputs=True)
    got = res["outputs"][0]
    out = np.zeros((2, len(ring.primes), N), dtype=np.uint64)
    m = c.multiply(out, x1, x2, 1, lm, [key, None], ext)
    for comp in range(2):
        c.to_coeff(out[comp], out[comp], m)
        mm = c.division(out[comp], m, 1 << (lvl - 1))
    assert ring.pool.sample_masks(got) == [mm, mm]
    g = ring.pool.read_sample(got)
    for i in range(len(ring.primes)):
        if (mm >> i) & 1:
            assert np.array_equal(g[:, i], out[:, i]), i
    lib.free_mlwe_RNS_sample(got)
