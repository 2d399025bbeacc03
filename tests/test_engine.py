"""Engine picker and capability probe (CPU): _vfhe_native.py:29-56, cpu_probe.c:14-33."""
import pytest

from vfhe_amd import _native, engine


def test_capability_probe_semantics(lib):
    assert lib.vfhe_cpu_supports(b"") == 1          # no requirement
    assert lib.vfhe_cpu_supports(b"not-a-feature") == 0
    assert lib.vfhe_cpu_supports(b"hip") in (0, 1)  # 0 here: no device visible


def test_pin_unknown_engine_raises(monkeypatch):
    monkeypatch.setenv("VFHE_ENGINE", "cuda")
    with pytest.raises(RuntimeError, match="unknown VFHE_ENGINE"):
        engine.select()


def test_pin_uninstalled_engine_raises():
    with pytest.raises(RuntimeError, match="not installed"):
        engine.select("portable")


def test_hip_requires_a_device(lib):
    if lib.vfhe_cpu_supports(b"hip"):
        assert engine.select("hip") == "hip"
    else:
        with pytest.raises(RuntimeError, match="lacks hip"):
            engine.select("hip")
