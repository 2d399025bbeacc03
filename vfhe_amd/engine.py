"""Engine picker with the `hip` row: the VFHE_ENGINE mechanism of the reference.

Restates modules/misc/python/src/_vfhe_native.py:29-56 over an engine table
that lists `hip` first (requires a visible gfx950 device, answered by the
engine's own vfhe_cpu_supports("hip")), then the reference's CPU engines.
`VFHE_ENGINE=hip` pins the hip engine; an unknown or non-runnable pin raises as
the reference does.  Only the hip engine is loadable from this repository: the
CPU engines are the reference's own extensions.
"""
from __future__ import annotations

import os

from . import _native

# (name, capability it requires), best first -- tools/_engines.py:91-126 plus the hip row
ENGINES = [("hip", "hip"), ("avx512ifma", "avx512ifma"), ("portable", None)]


def _installed(name: str) -> bool:
    return name == "hip" and os.path.exists(_native.LIB_PATH)


def _runnable(requires: str | None) -> bool:
    if not os.path.exists(_native.LIB_PATH):
        return False
    return bool(_native.load().vfhe_cpu_supports((requires or "").encode()))


def select(pin: str | None = None) -> str:
    pin = pin if pin is not None else (os.environ.get("VFHE_ENGINE") or "")
    if pin:
        for name, requires in ENGINES:
            if name != pin:
                continue
            if not _installed(name):
                raise RuntimeError(f"VFHE_ENGINE={pin}, but it is not installed")
            if not _runnable(requires):
                raise RuntimeError(f"VFHE_ENGINE={pin}, but this host lacks {requires}")
            return name
        known = ", ".join(n for n, _ in ENGINES)
        raise RuntimeError(f"unknown VFHE_ENGINE '{pin}'; known engines: {known}")
    for name, requires in ENGINES:
        if _installed(name) and _runnable(requires):
            return name
    raise RuntimeError("no installed engine runs on this host")


def load(pin: str | None = None):
    """The selected engine's library handle (only `hip` is provided here)."""
    name = select(pin)
    assert name == "hip"
    return _native.load()
