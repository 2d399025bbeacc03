import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs the hip engine's kernels)")
    config.addinivalue_line("markers", "slow: full-size configuration (seconds of oracle time)")


@pytest.fixture(scope="session")
def lib():
    from vfhe_amd import _native
    return _native.load()


@pytest.fixture(scope="session")
def oracle():
    from oracle import oracle as O
    return O


def golden_ntt_mod_cases():
    """Cases of tests/golden/ref_ntt_mod.npz: (n, q, x, y, expected dict) from the reference kernels."""
    import importlib.util

    import numpy as np
    here = os.path.join(ROOT, "tests", "golden")
    spec = importlib.util.spec_from_file_location("make_golden", os.path.join(here, "make_golden.py"))
    mg = importlib.util.module_from_spec(spec)
    spec.loader.exec_module(mg)
    data = np.load(os.path.join(here, "ref_ntt_mod.npz"))
    cases = []
    for k in range(len(mg.CASES)):
        n, q = int(data[f"n_{k}"][0]), int(data[f"q_{k}"][0])
        x, y = mg.inputs(k, n, q)
        exp = {name: data[f"{name}_{k}"] for name in ("fwd", "rev", "mul", "add", "sub")}
        cases.append((n, q, x, y, exp))
    return cases
