"""CPU oracle of CGGI functional bootstrapping over the 32-bit torus (numpy).

TEST INFRASTRUCTURE ONLY: imported by tests/ and the cpu_baseline leg of bench.py,
always as the checker (or the timed CPU baseline), never as part of the engine.

There is no reference counterpart: the reference's CGGI16 bootstraps over RNS primes
with the NTT (modules/fhe/python/src/vfhe/fhe/cggi16.py:46-140).  This restates the
same bootstrap structure on the discretised torus Z/2^32 --
  * rotation amounts: rot_b = round((b + q // (4 torus_base)) 2N / q) mod 2N,
    a_i -> round(a_i 2N / q) mod 2N, Python's round (ties to even) on a double quotient
    (cggi16.py:60-62, 77), here with q = 2^32 where the quotient is exact;
  * the accumulator starts as X^(2N - rot_b) tv (cggi16.py:64-71) and for each key index
    with a non-zero amount becomes CMUX(acc, X^a acc, bk[i]) = acc + bk[i] [x] (X^a acc - acc)
    (cggi16.py:74-90, mgsw.c:89-119's CMUX);
  * the test vector packs lut[i // (N / size)] * q // 2^prec (cggi16.py:92-116);
  * extraction of coefficient 0 (cggi16.py:118-140, mlwe.c:40-66's layout)
-- with the external product and signed gadget decomposition of TFHE (Chillotti, Gama,
Georgieva, Izabachene, J. Cryptology 33 (2020): Lemma 3.7, Theorem 3.13, Algorithm 3).

The external product is computed exactly in integers: digit polynomials (|d| <= 2^(B-1))
times key rows (32-bit, centred) accumulate below 2^51, so a float64 FFT product rounds
to the exact integer when its error stays below 1/2; `negacyclic_mul` asserts a margin
of 1/4 on every coefficient, and tests pin it against schoolbook convolution.

Pinned by tests/test_torus_oracle.py: the FFT against direct evaluation at the roots;
the product against schoolbook integer convolution; decomposition by recomposition and
digit bounds; the rounding of the rotation amounts against exact rationals at ties;
the bootstrap by decryption of every LUT entry.
"""
from __future__ import annotations

import ctypes
import os

import numpy as np

Q = 1 << 32
MASK = np.uint64(Q - 1)


def center(x: np.ndarray) -> np.ndarray:
    """Torus values (uint32 / uint64 < 2^32) as signed int64 in [-2^31, 2^31)."""
    x = np.asarray(x, dtype=np.int64) & (Q - 1)
    return np.where(x >= Q // 2, x - Q, x)


def negacyclic_fft(p: np.ndarray) -> np.ndarray:
    """p(zeta w^k), k < N/2, for real polynomials p (..., N): zeta = e^(i pi/N),
    w = e^(-2 pi i/(N/2)) -- the folded, twisted N/2-point DFT."""
    p = np.asarray(p, dtype=np.float64)
    N = p.shape[-1]
    M = N // 2
    zeta = np.exp(1j * np.pi * np.arange(M) / N)
    return np.fft.fft((p[..., :M] + 1j * p[..., M:]) * zeta, axis=-1)


def negacyclic_ifft(X: np.ndarray) -> np.ndarray:
    """Inverse of negacyclic_fft (real coefficients, float64)."""
    M = X.shape[-1]
    N = 2 * M
    z = np.fft.ifft(X, axis=-1) * np.exp(-1j * np.pi * np.arange(M) / N)
    return np.concatenate([z.real, z.imag], axis=-1)


def round_exact(y: np.ndarray) -> np.ndarray:
    """Nearest integers of a float result that must be within 1/4 of them."""
    r = np.rint(y)
    assert np.abs(y - r).max(initial=0.0) < 0.25, "float product too far from an integer"
    return r.astype(np.int64)


def negacyclic_mul(a: np.ndarray, b: np.ndarray) -> np.ndarray:
    """Exact a * b mod X^N + 1 for integer polynomials (..., N) whose product stays below 2^51."""
    return round_exact(negacyclic_ifft(negacyclic_fft(a) * negacyclic_fft(b)))


def schoolbook_mul(a, b) -> list[int]:
    """a * b mod X^N + 1 in Python integers (small N)."""
    N = len(a)
    out = [0] * N
    for i in range(N):
        for j in range(N):
            if i + j < N:
                out[i + j] += int(a[i]) * int(b[j])
            else:
                out[i + j - N] -= int(a[i]) * int(b[j])
    return out


def mod_switch(x, two_n: int) -> np.ndarray:
    """round(x * 2N / 2^32) mod 2N, ties to even (cggi16.py:62,77 with q = 2^32).  x < 2^33."""
    x = np.asarray(x, dtype=np.uint64)
    shift = np.uint64(32 - (int(two_n).bit_length() - 1))
    q = x >> shift
    rem = x & ((np.uint64(1) << shift) - np.uint64(1))
    half = np.uint64(1) << (shift - np.uint64(1))
    up = (rem > half) | ((rem == half) & ((q & np.uint64(1)) == np.uint64(1)))
    return ((q + up.astype(np.uint64)) % np.uint64(two_n)).astype(np.int64)


def decompose(x: np.ndarray, ell: int, base_log: int) -> np.ndarray:
    """Signed digits (ell, ...) of torus values x: x ~ sum_l d_l 2^(32 - base_log (l+1)),
    d_l in [-2^(B-1), 2^(B-1)), x rounded to ell * B bits first (Lemma 3.7)."""
    B = base_log
    off = sum(1 << (32 - B * l + B - 1) for l in range(1, ell + 1))
    if B * ell < 32:
        off += 1 << (32 - B * ell - 1)
    y = (np.asarray(x, dtype=np.int64) + off) & (Q - 1)
    return np.stack([((y >> (32 - B * l)) & ((1 << B) - 1)) - (1 << (B - 1)) for l in range(1, ell + 1)])


def mul_by_xai(p: np.ndarray, a: int) -> np.ndarray:
    """X^a * p mod X^N + 1 (a in [0, 2N)), torus coefficients as int64 mod 2^32."""
    N = p.shape[-1]
    a %= 2 * N
    sign = 1
    if a >= N:
        a -= N
        sign = -1
    r = np.roll(np.asarray(p, dtype=np.int64), a, axis=-1)
    r[..., :a] *= -1
    return (sign * r) & (Q - 1)


def key_fft(bsk: np.ndarray) -> np.ndarray:
    """The key rows (n, 2 ell, 2, N) centred and transformed."""
    return negacyclic_fft(center(bsk))


def external_product(kf: np.ndarray, acc: np.ndarray, ell: int, base_log: int) -> np.ndarray:
    """bk [x] acc for one GGSW key (2 ell, 2, N/2 transformed rows) and acc (2, N):
    sum over components c and levels l of dec_l(acc_c) * row (c ell + l), mod 2^32."""
    digits = decompose(acc, ell, base_log)                  # (ell, 2, N)
    df = negacyclic_fft(digits.transpose(1, 0, 2))          # (2, ell, N/2): row c*ell + l
    df = df.reshape(2 * ell, -1)
    prod = (df[:, None, :] * kf).sum(axis=0)                # (2, N/2)
    return round_exact(negacyclic_ifft(prod)) & (Q - 1)


def rotation_amounts(lwe: np.ndarray, N: int, torus_base: int):
    """(rot_tv, steps) of one sample (n + 1 values, b last)."""
    b = int(lwe[-1])
    rot_b = int(mod_switch(b + Q // (4 * torus_base), 2 * N))
    return (2 * N - rot_b) % (2 * N), mod_switch(lwe[:-1], 2 * N)


def blind_rotate(lwe: np.ndarray, tv: np.ndarray, kf: np.ndarray, ell: int, base_log: int,
                 torus_base: int) -> np.ndarray:
    """The accumulator (2, N) of one sample after the blind rotation."""
    N = tv.shape[-1]
    rot_tv, steps = rotation_amounts(lwe, N, torus_base)
    acc = np.zeros((2, N), dtype=np.int64)
    acc[1] = mul_by_xai(tv, rot_tv)
    for i, a in enumerate(steps):
        if a == 0:
            continue
        diff = (mul_by_xai(acc, int(a)) - acc) & (Q - 1)
        acc = (acc + external_product(kf[i], diff, ell, base_log)) & (Q - 1)
    return acc


def extract(acc: np.ndarray) -> np.ndarray:
    """The LWE sample (N + 1) of coefficient 0: a'_0 = a_0, a'_j = -a_(N-j), b' = b_0."""
    a = acc[0]
    out = np.empty(a.shape[-1] + 1, dtype=np.int64)
    out[0] = a[0]
    out[1:-1] = (-a[:0:-1]) & (Q - 1)
    out[-1] = acc[1][0]
    return out


def pack_lut(lut, size: int, prec: int, N: int) -> np.ndarray:
    """Test vector coefficients lut[i // (N / size)] * 2^32 // 2^prec (cggi16.py:92-116)."""
    scaled = [(int(v) * Q) >> prec for v in lut]
    out = np.zeros(N, dtype=np.int64)
    for i in range(N):
        idx = i // (N // size)
        if idx < len(scaled):
            out[i] = scaled[idx]
    return out & (Q - 1)


def phase(sample: np.ndarray, key: np.ndarray) -> int:
    """b - <a, s> mod 2^32 of one LWE sample."""
    a = np.asarray(sample[:-1], dtype=np.int64)
    return int((int(sample[-1]) - int((a * np.asarray(key, dtype=np.int64)).sum())) % Q)


_port = None


def _port_lib():
    global _port
    if _port is None:
        path = os.path.join(os.path.dirname(os.path.abspath(__file__)), "liboracle_torus.so")
        if not os.path.exists(path):
            raise RuntimeError("oracle/liboracle_torus.so is not built (run `make -C oracle`)")
        _port = ctypes.CDLL(path)
        VP, U64, I = ctypes.c_void_p, ctypes.c_uint64, ctypes.c_int
        _port.torus_port_new.restype = VP
        _port.torus_port_new.argtypes = [VP, I, I, I, I]
        _port.torus_port_free.restype = None
        _port.torus_port_free.argtypes = [VP]
        _port.torus_port_run.restype = I
        _port.torus_port_run.argtypes = [VP, VP, U64, VP, U64, VP, I]
    return _port


class PortKey:
    """The key rows transformed once by torus_port.c (the C form of this bootstrap)."""

    def __init__(self, bsk: np.ndarray, ell: int, base_log: int):
        key = np.ascontiguousarray(np.asarray(bsk, dtype=np.int64) & (Q - 1), dtype=np.uint32)
        self.n, self.N = key.shape[0], key.shape[-1]
        assert key.shape == (self.n, 2 * ell, 2, self.N)
        self.h = _port_lib().torus_port_new(key.ctypes.data, self.n, self.N, ell, base_log)
        assert self.h

    def bootstrap(self, lwe: np.ndarray, tv: np.ndarray, torus_base: int, threads: int = 1) -> np.ndarray:
        """Extracted samples (count, N + 1) of every sample (count, n + 1), on `threads` threads."""
        lw = np.ascontiguousarray(np.asarray(lwe, dtype=np.int64) & (Q - 1), dtype=np.uint32)
        if lw.ndim == 1:
            lw = lw[None, :]
        assert lw.shape[1] == self.n + 1
        tvu = np.ascontiguousarray(np.asarray(tv, dtype=np.int64) & (Q - 1), dtype=np.uint32)
        out = np.empty((lw.shape[0], self.N + 1), dtype=np.uint32)
        rc = _port_lib().torus_port_run(self.h, lw.ctypes.data, lw.shape[0], tvu.ctypes.data, torus_base,
                                        out.ctypes.data, threads)
        assert rc == 0, rc
        return out

    def __del__(self):
        try:
            _port_lib().torus_port_free(self.h)
        except Exception:
            pass


def port_bootstrap(lwe: np.ndarray, tv: np.ndarray, bsk: np.ndarray, ell: int, base_log: int, torus_base: int,
                   threads: int = 1) -> np.ndarray:
    """PortKey(bsk).bootstrap(...) in one call."""
    return PortKey(bsk, ell, base_log).bootstrap(lwe, tv, torus_base, threads)
