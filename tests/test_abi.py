"""The C ABI boundary (CPU only: loads the library, makes no GPU call).

Every function include/vfhe_hip.h declares is exported by libvfhe_hip.so and
bound by vfhe_amd._native; the public struct prefixes keep the reference's
field layout (modules/arith/python/cdef/arith.cdef:10-23, mlwe.cdef:27-31).
"""
import ctypes
import os
import re
import subprocess

import pytest

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


# Functions the reference cdefs declare that this engine deliberately does not export: they are
# off the RNS polynomial path (hashing needs BLAKE3, an empty submodule in the reference; mp
# polynomials, prime fields and GP25 belong to the proof-system and GP25 stacks), or libc's own.
ABSENT = {
    "polynomial_RNS_get_hash": "BLAKE3 hashing (rns_polynomial.c:884-903), not on the path",
    "polynomial_RNS_get_hash_p": "BLAKE3 hashing, not on the path",
    "get_mp_vector_size": "mp polynomials (proof-system stack)",
    "mp_scale": "mp polynomials", "mp_sub": "mp polynomials", "mp_load": "mp polynomials",
    "load_m512": "mp polynomials (AVX-512 register loader)",
    "new_mp_polynomial": "mp polynomials", "mp_polynomial_from_RNS": "mp polynomials",
    "mp_polynomial_to_RNSc": "mp polynomials",
    "field_new_proc": "prime fields (proof-system stack)", "field_ext_add": "prime fields",
    "field_ext_sub": "prime fields", "field_ext_neg": "prime fields", "field_ext_mul": "prime fields",
    "field_ext_pow": "prime fields", "field_ext_inv": "prime fields",
    "field_sample_random_element": "prime fields", "field_hash_element": "prime fields (BLAKE3)",
    "field_ext_is_equal": "prime fields", "field_base_conversion": "prime fields",
    "gp25_RGSW_monomial_mul": "GP25 scheme", "gp25_RGSW_monomial_mul_mt": "GP25 scheme",
    "gp25_sub_a_mt": "GP25 scheme",
    "free": "libc",
}
REF = "/root/reference/modules"
CDEFS = ["arith/python/cdef/arith.cdef", "mlwe/python/cdef/mlwe.cdef", "fhe/python/cdef/fhe.cdef",
         "misc/python/cdef/misc.cdef", "misc/python/cdef/vfhe_cpu.cdef"]


def _arity(params: str) -> int:
    params = params.strip()
    if params in ("", "void"):
        return 0
    depth, n = 0, 1
    for ch in params:
        depth += ch in "(["
        depth -= ch in ")]"
        n += ch == "," and depth == 0
    return n


def _signatures(text: str) -> dict:
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    text = re.sub(r"//[^\n]*", "", text)
    return {name: _arity(params) for name, params in
            re.findall(r"\b([A-Za-z_]\w*)\s*\(([^;{()]*(?:\([^()]*\)[^;{()]*)*)\)\s*;", text)
            if name not in ("if", "while", "sizeof")}


@pytest.mark.skipif(not os.path.isdir(REF), reason="reference sources absent")
def test_reference_cdefs_are_covered():
    """Every function of the reference's cdefs (arith, mlwe, fhe, misc, vfhe_cpu) is declared in
    include/vfhe_hip.h with the same number of arguments, or is listed above as deliberately
    absent; a cffi module over the covered cdefs therefore links against this library."""
    ours = _signatures(open(HEADER).read())
    missing, arity = [], []
    for rel in CDEFS:
        for name, n in _signatures(open(os.path.join(REF, rel)).read()).items():
            if name in ABSENT:
                assert name not in ours, f"{name} is exported but listed as absent"
                continue
            if name not in ours:
                missing.append(f"{rel}:{name}")
            elif ours[name] != n:
                arity.append(f"{name}: reference {n}, here {ours[name]}")
    assert not missing, missing
    assert not arity, arity


def test_scalar_modq_helpers_against_bigint_and_reference():
    """add_modq / sub_modq / negate_modq (mod.c:54-62) are host arithmetic: the engine's exports
    against Python integers and, where oracle/_ref is built, the reference's own functions, over
    edge values (0, 1, q - 1) and random residues of 30..62-bit moduli.  mul_modq needs an NTT_proc
    and is checked with the device tests."""
    import random
    lib = ctypes.CDLL(_native.LIB_PATH)
    U = ctypes.c_uint64
    for name, nargs in (("add_modq", 3), ("sub_modq", 3), ("negate_modq", 2)):
        getattr(lib, name).restype = U
        getattr(lib, name).argtypes = [U] * nargs
    refs = []
    try:
        from oracle import oracle as O
        for eng in O.ref_engines():
            R = O.ref_lib(eng)
            for name, nargs in (("add_modq", 3), ("sub_modq", 3), ("negate_modq", 2)):
                getattr(R, name).restype = U
                getattr(R, name).argtypes = [U] * nargs
            refs.append(R)
    except Exception:
        refs = []
    rnd = random.Random(54)
    for bits in (30, 50, 60, 62):
        q = (1 << bits) - rnd.randrange(1, 1 << 20) * 2 - 1
        vals = [0, 1, q - 1, q // 2] + [rnd.randrange(q) for _ in range(200)]
        for a in vals:
            b = rnd.choice(vals)
            want = ((a + b) % q, (a - b) % q, (-a) % q)
            got = (lib.add_modq(a, b, q), lib.sub_modq(a, b, q), lib.negate_modq(a, q))
            assert got == want, (bits, a, b)
            for R in refs:
                assert (R.add_modq(a, b, q), R.sub_modq(a, b, q), R.negate_modq(a, q)) == want
