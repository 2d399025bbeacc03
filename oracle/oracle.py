"""ctypes bindings to the CPU oracle.

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and the
cpu_baseline leg of bench.py, always as the checker (or the timed CPU
reference), never as part of the engine.

Two libraries are bound here:
  * liboracle.so -- vfhe_oracle.c, a scalar restatement of the reference path;
  * _ref/libvfheref_<engine>.so -- the reference's own NTT and modular kernels
    (modules/arith/c/src/{ntt*,mod*,number_theory}.c) compiled unmodified by
    oracle/Makefile, for the engines "portable" and "avx512ifma".
"""
from __future__ import annotations

import ctypes
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
U64 = ctypes.c_uint64
U64P = ctypes.POINTER(ctypes.c_uint64)
VP = ctypes.c_void_p

O_MUL, O_MUL_ADDTO, O_MUL_SUBTO, O_SCALE, O_FMA, O_ADD, O_SUB, O_NEGATE = range(8)
O_REDUCE, O_REDUCE_SIGNED, O_ADD_SCALAR, O_SUB_SCALAR = range(8, 12)

_lib = None


def build() -> None:
    """Compile liboracle.so and (when the reference sources are present) oracle/_ref."""
    subprocess.run(["make", "-s", "-C", HERE, "all"], check=True)


def ptr(a: np.ndarray):
    assert a.dtype == np.uint64 and a.flags["C_CONTIGUOUS"]
    return a.ctypes.data_as(U64P)


def lib():
    global _lib
    if _lib is None:
        path = os.path.join(HERE, "liboracle.so")
        if not os.path.exists(path):
            raise RuntimeError("oracle/liboracle.so is not built (run `make -C oracle`)")
        L = ctypes.CDLL(path)
        sig = {
            "o_power_mod": (U64, [U64, U64, U64]),
            "o_inverse_mod": (U64, [U64, U64]),
            "o_is_prime": (ctypes.c_int, [U64]),
            "o_next_special_prime": (U64, [U64, U64, ctypes.c_int]),
            "o_gen_prime": (U64, [U64, U64, U64P, U64]),
            "o_find_psi": (U64, [U64, U64]),
            "o_ntt_forward": (None, [U64P, U64, U64]),
            "o_ntt_inverse": (None, [U64P, U64, U64]),
            "o_eltwise": (None, [ctypes.c_int, U64P, U64P, U64P, U64, U64, U64]),
            "o_reduce_mp": (None, [U64P, U64P, U64P, U64, U64]),
            "o_ctx_new": (VP, [U64, U64, U64P, U64]),
            "o_ctx_free": (None, [VP]),
            "o_ctx_psi": (U64, [VP, U64]),
            "o_ctx_w": (None, [VP, U64, U64P]),
            "o_to_ntt": (U64, [VP, U64P, U64P, U64]),
            "o_to_coeff": (U64, [VP, U64P, U64P, U64]),
            "o_mul": (U64, [VP, U64P, U64P, U64, U64P, U64, ctypes.c_int]),
            "o_addsub": (U64, [VP, U64P, U64P, U64, U64P, U64, ctypes.c_int]),
            "o_negate": (U64, [VP, U64P, U64P, U64]),
            "o_scale": (U64, [VP, U64P, U64P, U64, U64, ctypes.c_int]),
            "o_scale_rns": (U64, [VP, U64P, U64P, U64, U64P]),
            "o_add_integer": (U64, [VP, U64P, U64P, U64, U64, ctypes.c_int]),
            "o_mod_reduce_lifted": (None, [VP, U64P, U64, U64P, U64]),
            "o_division": (U64, [VP, U64P, U64, U64, ctypes.c_int]),
            "o_base_conversion": (U64, [VP, U64P, U64, U64P, U64]),
            "o_scaling_factors": (None, [VP, U64P, U64, U64]),
            "o_scaled_lift": (None, [VP, U64P, U64, U64P, U64, U64P]),
            "o_permute": (U64, [VP, U64P, U64P, U64, U64]),
            "o_extended_rank": (U64, [U64]),
            "o_tensor_product": (U64, [VP, U64P, U64P, U64P, U64, U64]),
            "o_ghs_keyswitch": (U64, [VP, U64P, U64, U64P, U64, U64, ctypes.POINTER(U64P), U64]),
            "o_multiply": (U64, [VP, U64P, U64P, U64P, U64, U64, ctypes.POINTER(U64P), U64]),
            "o_automorphism": (U64, [VP, U64P, U64P, U64, U64, U64, ctypes.POINTER(U64P), U64]),
            "o_mlwe_round_division": (U64, [VP, U64P, U64, U64, U64]),
            "o_mul_by_xai": (U64, [VP, U64P, U64P, U64, U64, ctypes.c_int]),
            "o_external_product": (U64, [VP, U64P, U64, U64P, U64, U64P, U64, U64, U64]),
            "o_cmux": (U64, [VP, U64P, U64, U64P, U64P, U64, U64P, U64, U64, U64]),
        }
        for name, (res, args) in sig.items():
            f = getattr(L, name)
            f.restype = res
            f.argtypes = args
        _lib = L
    return _lib


def gen_prime(rou_order: int, prime_size: int, exclude=()) -> int:
    ex = np.array(list(exclude) or [0], dtype=np.uint64)
    return int(lib().o_gen_prime(rou_order, prime_size, ptr(ex), len(exclude)))


def gen_primes(N: int, sizes, split_degree: int = 1) -> list[int]:
    """Ring.gen_primes (polynomial.py:162-170): one prime per size, excluding earlier ones."""
    primes: list[int] = []
    for b in sizes:
        primes.append(gen_prime(2 * N // split_degree, b, primes))
    return primes


def ksk_array(entries):
    """A ksk[] argument: list of (None | uint64 ndarray of shape [digits, r+1, l, N])."""
    arr = (U64P * len(entries))()
    for i, e in enumerate(entries):
        arr[i] = ptr(e) if e is not None else None
    return arr


class Ctx:
    """One RNS pool (the reference's incNTT for a fixed N, split degree and prime list)."""

    def __init__(self, N: int, primes, split_degree: int = 1):
        self.N, self.split, self.l = N, split_degree, len(primes)
        self.primes = [int(p) for p in primes]
        pa = np.array(self.primes, dtype=np.uint64)
        self.h = lib().o_ctx_new(N, split_degree, ptr(pa), self.l)

    def __del__(self):
        if getattr(self, "h", None):
            lib().o_ctx_free(self.h)
            self.h = None

    def zeros(self, npoly: int = 1) -> np.ndarray:
        shape = (self.l, self.N) if npoly == 1 else (npoly, self.l, self.N)
        return np.zeros(shape, dtype=np.uint64)

    def psi(self, i: int) -> int:
        return int(lib().o_ctx_psi(self.h, i))

    def w(self, i: int) -> np.ndarray:
        out = np.zeros(self.N // self.split, dtype=np.uint64)
        lib().o_ctx_w(self.h, i, ptr(out))
        return out

    def mask_all(self) -> int:
        return (1 << self.l) - 1

    def to_ntt(self, out, inp, mask):
        return int(lib().o_to_ntt(self.h, ptr(out), ptr(inp), mask))

    def to_coeff(self, out, inp, mask):
        return int(lib().o_to_coeff(self.h, ptr(out), ptr(inp), mask))

    def mul(self, out, a, ma, b, mb, mode=0):
        return int(lib().o_mul(self.h, ptr(out), ptr(a), ma, ptr(b), mb, mode))

    def addsub(self, out, a, ma, b, mb, sub=False):
        return int(lib().o_addsub(self.h, ptr(out), ptr(a), ma, ptr(b), mb, int(sub)))

    def negate(self, out, a, m):
        return int(lib().o_negate(self.h, ptr(out), ptr(a), m))

    def scale(self, out, a, m, s, addto=False):
        return int(lib().o_scale(self.h, ptr(out), ptr(a), m, s, int(addto)))

    def scale_rns(self, out, a, m, s):
        return int(lib().o_scale_rns(self.h, ptr(out), ptr(a), m, ptr(s)))

    def add_integer(self, out, a, m, v, ntt_form):
        return int(lib().o_add_integer(self.h, ptr(out), ptr(a), m, v & 0xFFFFFFFFFFFFFFFF, int(ntt_form)))

    def mod_reduce_lifted(self, out, out_mask, inp, idx):
        lib().o_mod_reduce_lifted(self.h, ptr(out), out_mask, ptr(inp), idx)

    def division(self, p, mask, divide_mask, round_=True):
        return int(lib().o_division(self.h, ptr(p), mask, divide_mask, int(round_)))

    def base_conversion(self, out, out_mask, inp, in_mask):
        return int(lib().o_base_conversion(self.h, ptr(out), out_mask, ptr(inp), in_mask))

    def scaling_factors(self, in_mask, out_mask):
        d = np.zeros(self.l, dtype=np.uint64)
        lib().o_scaling_factors(self.h, ptr(d), in_mask, out_mask)
        return d

    def scaled_lift(self, out, out_mask, inp, in_mask, delta=None):
        lib().o_scaled_lift(self.h, ptr(out), out_mask, ptr(inp), in_mask, ptr(delta) if delta is not None else None)

    def permute(self, out, inp, m, gen):
        return int(lib().o_permute(self.h, ptr(out), ptr(inp), m, gen))

    def tensor_product(self, out, in1, in2, r, mask):
        return int(lib().o_tensor_product(self.h, ptr(out), ptr(in1), ptr(in2), r, mask))

    def ghs_keyswitch(self, out, r, inp, Rin, in_mask, ksk, ext_mask):
        return int(lib().o_ghs_keyswitch(self.h, ptr(out), r, ptr(inp), Rin, in_mask, ksk_array(ksk), ext_mask))

    def multiply(self, out, in1, in2, r, mask, rlk, ext_mask):
        return int(lib().o_multiply(self.h, ptr(out), ptr(in1), ptr(in2), r, mask, ksk_array(rlk), ext_mask))

    def automorphism(self, out, inp, r, mask, gen, ksk, ext_mask):
        return int(lib().o_automorphism(self.h, ptr(out), ptr(inp), r, mask, gen, ksk_array(ksk), ext_mask))

    def mul_by_xai(self, out, inp, m, a, minus1=False):
        return int(lib().o_mul_by_xai(self.h, ptr(out), ptr(inp), m, a, int(minus1)))

    def external_product(self, out, r, inp, in_mask, mgsw, ell, special_primes, ext_mask):
        return int(lib().o_external_product(self.h, ptr(out), r, ptr(inp), in_mask, ptr(mgsw), ell, special_primes,
                                            ext_mask))

    def cmux(self, out, r, in1, in2, mask, mgsw, ell, special_primes, ext_mask):
        return int(lib().o_cmux(self.h, ptr(out), r, ptr(in1), ptr(in2), mask, ptr(mgsw), ell, special_primes,
                                ext_mask))

    def mlwe_round_division(self, ct, r, mask, divide_mask):
        return int(lib().o_mlwe_round_division(self.h, ptr(ct), r, mask, divide_mask))


# ---------------------------------------------------------------------------
# The reference's own kernels (oracle/_ref)
# ---------------------------------------------------------------------------

def cpu_has_ifma() -> bool:
    try:
        with open("/proc/cpuinfo") as f:
            return "avx512ifma" in f.read()
    except OSError:
        return False


_ref = {}


def ref_engines() -> list[str]:
    engines = []
    for e in ("avx512ifma", "portable"):
        if e == "avx512ifma" and not cpu_has_ifma():
            continue
        if os.path.exists(os.path.join(HERE, "_ref", f"libvfheref_{e}.so")):
            engines.append(e)
    return engines


def ref_lib(engine: str):
    if engine not in _ref:
        L = ctypes.CDLL(os.path.join(HERE, "_ref", f"libvfheref_{engine}.so"))
        L.ntt_new_proc.restype = VP
        L.ntt_new_proc.argtypes = [U64, U64]
        L.ntt_free_proc.argtypes = [VP]
        for f in (L.ntt_forward, L.ntt_reverse):
            f.argtypes = [U64P, U64P, VP]
            f.restype = None
        for name in ("mod_eltwise_mul", "mod_eltwise_mul_addto", "mod_eltwise_mul_subto",
                     "mod_eltwise_add", "mod_eltwise_sub"):
            getattr(L, name).argtypes = [U64P, U64P, U64P, U64, VP]
        for name in ("mod_eltwise_scale", "mod_eltwise_fma", "mod_eltwise_add_scalar",
                     "mod_eltwise_sub_scalar"):
            getattr(L, name).argtypes = [U64P, U64P, U64, U64, VP]
        for name in ("mod_eltwise_negate", "mod_eltwise_reduce", "mod_eltwise_reduce_signed"):
            getattr(L, name).argtypes = [U64P, U64P, U64, VP]
        L.mod_reduce_array_mp.argtypes = [U64P, U64P, U64P, U64, VP]
        L.next_special_prime.restype = U64
        L.next_special_prime.argtypes = [U64, U64, ctypes.c_bool]
        L.is_prime.restype = ctypes.c_bool
        L.is_prime.argtypes = [U64]
        L.inverse_mod.restype = U64
        L.inverse_mod.argtypes = [U64, U64]
        L.ref_ckks_multiply.restype = ctypes.c_int
        L.ref_ckks_multiply.argtypes = [U64, U64, U64, U64P, U64P, U64P, U64P, U64P, U64, U64, ctypes.c_int]
        _ref[engine] = L
    return _ref[engine]


def ref_ckks_multiply(engine: str, N: int, L: int, K: int, primes, ct1: np.ndarray, ct2: np.ndarray,
                      rlk: np.ndarray, threads: int = 1, count: int | None = None) -> np.ndarray:
    """CKKS multiply + relinearize + rescale composed from the reference's own kernels
    (oracle/ref_ckks.c): ct1, ct2 (n, 2, L, N) in NTT form, rlk (L, 2, L+K, N) in NTT form;
    returns the n products (n, 2, L-1, N) in coefficient form.  count > n repeats the inputs
    (pair k = input k mod n) to make a timing sample of `count` products on `threads` threads."""
    n = ct1.shape[0]
    count = n if count is None else count
    a1 = aligned(ct1.size)
    a1[:] = ct1.reshape(-1)
    a2 = aligned(ct2.size)
    a2[:] = ct2.reshape(-1)
    k = aligned(rlk.size)
    k[:] = rlk.reshape(-1)
    out = aligned(n * 2 * (L - 1) * N)
    pr = np.array([int(p) for p in primes[: L + K]], dtype=np.uint64)
    rc = ref_lib(engine).ref_ckks_multiply(N, L, K, ptr(pr), ptr(a1), ptr(a2), ptr(k), ptr(out), n, count, threads)
    assert rc == 0
    return out.reshape(n, 2, L - 1, N).copy()


def aligned(n: int) -> np.ndarray:
    """A zeroed uint64 array whose data is 64-byte aligned (the AVX-512 kernels need it)."""
    raw = np.zeros(n + 8, dtype=np.uint64)
    off = (-raw.ctypes.data % 64) // 8
    return raw[off:off + n]


_cx = {}


def ref_complex(engine: str):
    """The reference's complex FFT (complex.c) built into oracle/_ref for `engine`."""
    if engine not in _cx:
        path = os.path.join(HERE, "_ref", f"libvfheref_complex_{engine}.so")
        if not os.path.exists(path):
            raise RuntimeError(f"{path} is not built (run `make -C oracle` where the reference is present)")
        L = ctypes.CDLL(path)
        DP = ctypes.POINTER(ctypes.c_double)
        for name, res, args in (("load_rous_CT", VP, [DP, DP, U64]), ("load_rous_GS", VP, [DP, DP, U64]),
                                ("CT_NR", None, [DP, VP, U64]), ("GS_RN", None, [DP, VP, U64]),
                                ("bit_reverse_array", None, [DP, U64, ctypes.c_uint32]),
                                ("complex_poly_scale_double", None, [DP, ctypes.c_double, U64])):
            f = getattr(L, name)
            f.restype, f.argtypes = res, args
        _cx[engine] = L
    return _cx[engine]


def aligned_f64(n: int) -> np.ndarray:
    """A 64-byte aligned float64 buffer (the AVX-512 FFT uses aligned loads)."""
    raw = np.zeros(n + 8, dtype=np.float64)
    off = (-raw.ctypes.data % 64) // 8
    return raw[off:off + n]


def dptr(a: np.ndarray):
    assert a.dtype == np.float64 and a.flags["C_CONTIGUOUS"]
    return a.ctypes.data_as(ctypes.POINTER(ctypes.c_double))


class RefProc:
    """One reference NTT_proc (ntt_new_proc(n, q)) for calling the reference kernels."""

    def __init__(self, engine: str, n: int, q: int):
        self.L = ref_lib(engine)
        self.n, self.q = n, q
        self.h = self.L.ntt_new_proc(n, q)

    def __del__(self):
        if getattr(self, "h", None):
            self.L.ntt_free_proc(self.h)
            self.h = None

    def forward(self, x: np.ndarray) -> np.ndarray:
        src, out = aligned(self.n), aligned(self.n)
        src[:] = x
        self.L.ntt_forward(ptr(out), ptr(src), self.h)
        return out.copy()

    def reverse(self, x: np.ndarray) -> np.ndarray:
        src, out = aligned(self.n), aligned(self.n)
        src[:] = x
        self.L.ntt_reverse(ptr(out), ptr(src), self.h)
        return out.copy()


# ---------------------------------------------------------------------------
# Slot operations on NTT-form rows (numpy restatements; rows are [limb][N], a row of
# N = split * ps values holds `split` blocks of ps slots)
# ---------------------------------------------------------------------------

def slot_broadcast(rows: np.ndarray, mask: int, split: int, slot: int) -> np.ndarray:
    """polynomial_RNS_broadcast_slot (modules/arith/c/src/rns_polynomial.c:1228-1245): every slot of
    each block takes the value of slot `slot` of that block, on the active limbs."""
    out = np.zeros_like(rows)
    ps = rows.shape[1] // split
    for i in range(rows.shape[0]):
        if (mask >> i) & 1:
            for j in range(split):
                out[i, j * ps:(j + 1) * ps] = rows[i, j * ps + slot]
    return out


def slot_rotate(rows: np.ndarray, mask: int, split: int, rot: int) -> np.ndarray:
    """polynomial_RNS_rotate_slot (rns_polynomial.c:1262-1284): out[k] = in[(k + rot) mod ps] in each block."""
    out = np.zeros_like(rows)
    ps = rows.shape[1] // split
    for i in range(rows.shape[0]):
        if (mask >> i) & 1:
            for j in range(split):
                out[i, j * ps:(j + 1) * ps] = np.roll(rows[i, j * ps:(j + 1) * ps], -rot)
    return out


def slot_copy(out_rows: np.ndarray, dst: int, rows: np.ndarray, mask: int, split: int, src: int) -> np.ndarray:
    """polynomial_RNS_copy_slot (rns_polynomial.c:1286-1300): slot dst of each block of out takes slot src
    of the same block of in, on in's active limbs; the rest of out is kept."""
    out = out_rows.copy()
    ps = rows.shape[1] // split
    for i in range(rows.shape[0]):
        if (mask >> i) & 1:
            for j in range(split):
                out[i, j * ps + dst] = rows[i, j * ps + src]
    return out


def slot_inverse(rows: np.ndarray, mask: int, primes) -> tuple[int, np.ndarray]:
    """polynomial_RNS_inverse at split degree 1 (rns_polynomial.c:1383-1429): the slot-wise inverse mod
    q_i on every active limb, or status -2 when a slot is zero (returns (status, rows))."""
    out = np.zeros_like(rows)
    for i, q in enumerate(primes):
        if (mask >> i) & 1:
            row = [int(v) for v in rows[i]]
            if any(v == 0 for v in row):
                return -2, out
            out[i] = np.array([pow(v, -1, int(q)) for v in row], dtype=np.uint64)
    return 0, out
