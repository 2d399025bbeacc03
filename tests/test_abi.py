"""The C ABI boundary (CPU only: loads the library, makes no GPU call).

Every function include/vfhe_hip.h declares is exported by libvfhe_hip.so and
bound by vfhe_amd._native; the public struct prefixes keep the reference's
field layout (modules/arith/python/cdef/arith.cdef:10-23, mlwe.cdef:27-31).
"""
import ctypes
import os
import re
import subprocess

from vfhe_amd import _native

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = os.path.join(ROOT, "include", "vfhe_hip.h")


def declared_functions():
    text = open(HEADER).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    names = re.findall(r"^\s*[A-Za-z_][\w\s\*]*?\b([A-Za-z_]\w*)\s*\([^;{]*\)\s*;", text, flags=re.M)
    return sorted(set(n for n in names if n not in ("if", "while")))


def test_header_parses_to_many_functions():
    names = declared_functions()
    assert len(names) > 100
    for must in ("ntt_forward", "polynomial_RNSc_to_RNS", "polynomial_round_division_RNSc_wo_free",
                 "mlwe_RNSc_GHS_hybrid_keyswitch", "mlwe_multiply", "polynomial_RNSc_mod_reduce_lifted"):
        assert must in names


def test_every_declared_symbol_is_exported():
    out = subprocess.run(["nm", "-D", "--defined-only", _native.LIB_PATH], capture_output=True, text=True, check=True)
    exported = {line.split()[-1] for line in out.stdout.splitlines() if " T " in line}
    missing = [n for n in declared_functions() if n not in exported]
    assert not missing, missing


def test_binding_covers_header():
    assert set(declared_functions()) == set(_native.symbols())


def test_library_loads_and_reports_engine():
    lib = _native.load()
    assert lib.vfhe_engine_active() == b"hip"
    # number theory runs on the host: no GPU needed
    assert lib.is_prime(562949954142209)
    assert lib.inverse_mod(3, 7) == 5
    assert lib.next_special_prime(1 << 40, 1024, True) % 2048 == 1
    assert lib.mlwe_extended_rank(1) == 2 and lib.mlwe_extended_rank(2) == 5


def test_struct_layout_matches_reference_cdef():
    P = _native.RNSPolynomial
    assert P.coeffs.offset == 0 and P.ntt.offset == 8 and P.rns_mask.offset == 16 and P.allocated_l.offset == 24
    T = _native.IncNTT
    assert [T.ntt.offset, T.split_degree.offset, T.w.offset, T.N.offset, T.l.offset] == [0, 8, 16, 24, 32]
    M = _native.RNSMLWE
    assert [M.a.offset, M.b.offset, M.r.offset] == [0, 8, 16]
    assert ctypes.sizeof(M) == 24
