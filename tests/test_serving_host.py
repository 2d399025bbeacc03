"""Serving loop host logic (CPU): workload generation and latency statistics."""
import numpy as np

from vfhe_amd import serving


def test_workload_is_seeded_poisson():
    a = serving.make_workload(1000.0, 2.0, [8192, 16384, 32768, 65536], seed=7)
    b = serving.make_workload(1000.0, 2.0, [8192, 16384, 32768, 65536], seed=7)
    assert a == b
    assert 1800 < len(a) < 2200
    ts = np.array([r.t for r in a])
    assert (np.diff(ts) > 0).all() and ts[-1] < 2.0
    gaps = np.diff(ts)
    assert abs(gaps.mean() - 1e-3) < 1e-4 and abs(gaps.std() / gaps.mean() - 1.0) < 0.1
    assert {r.N for r in a} == {8192, 16384, 32768, 65536}
    assert min(r.level for r in a) == 4 and max(r.level for r in a) == 24
    frac = np.mean([r.op == serving.REQ_MULTIPLY for r in a])
    assert 0.45 < frac < 0.55
    assert {r.step for r in a if r.op == serving.REQ_ROTATE} == {0, 1, 2, 3}
    assert serving.make_workload(1000.0, 2.0, [8192], seed=8) != a


def test_percentiles():
    lat = [i / 1000 for i in range(1, 101)]   # 1..100 ms
    p = serving.percentiles(lat)
    assert abs(p["p50_ms"] - 50.5) < 1e-6 and abs(p["p99_ms"] - 99.01) < 1e-6
    assert serving.percentiles([]) == {}
