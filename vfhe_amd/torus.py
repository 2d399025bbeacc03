"""CGGI functional bootstrapping over the 32-bit torus on the hip engine.

No counterpart in the reference, whose CGGI16 (modules/fhe/python/src/vfhe/fhe/
cggi16.py:46-140) bootstraps over RNS primes with the NTT.  The same functional
bootstrap runs here on the discretised torus Z/2^32 with the external product
through an FP64 negacyclic FFT (vfhe_hip_torus_* in include/vfhe_hip.h): LWE
dimension n, one GLWE ring of N = 1024, signed gadget decomposition in base
2^base_log with ell levels.  The defaults are n = 630, N = 1024, base 2^7, 3 levels
(TFHE's gate-bootstrapping parameters).

Host side: key generation and encryption (numpy; the body a * s of a GLWE sample is a
binary-key negacyclic product, taken through a float64 FFT whose result is checked to
lie within 1/4 of integers), LUT packing (cggi16.py:92-116 with q = 2^32), device
batches, the bootstrap call, and decryption helpers.
"""
from __future__ import annotations

from dataclasses import dataclass

import numpy as np

from . import _native

Q = 1 << 32


def lib():
    return _native.load()


@dataclass(frozen=True)
class TorusParams:
    n: int = 630
    N: int = 1024
    ell: int = 3
    base_log: int = 7
    lwe_sigma: float = 2.0 ** -15      # standard deviations as fractions of the torus
    glwe_sigma: float = 2.0 ** -25


def _u32(a) -> np.ndarray:
    return np.ascontiguousarray(np.asarray(a, dtype=np.int64) & (Q - 1), dtype=np.uint32)


def _gauss(rng: np.random.Generator, sigma: float, shape) -> np.ndarray:
    return np.rint(rng.normal(0.0, sigma * Q, size=shape)).astype(np.int64)


def _binary_negacyclic(a: np.ndarray, s: np.ndarray) -> np.ndarray:
    """a * s mod X^N + 1 for torus polynomials a (..., N) and a binary key s (N), exact:
    |a| < 2^31 centred, so every coefficient is below 2^41 and the float64 FFT product
    lands within 1/4 of it (checked)."""
    a = np.asarray(a, dtype=np.int64) & (Q - 1)
    a = np.where(a >= Q // 2, a - Q, a).astype(np.float64)
    N = a.shape[-1]
    M = N // 2
    zeta = np.exp(1j * np.pi * np.arange(M) / N)
    fa = np.fft.fft((a[..., :M] + 1j * a[..., M:]) * zeta, axis=-1)
    sf = np.asarray(s, dtype=np.float64)
    fs = np.fft.fft((sf[:M] + 1j * sf[M:]) * zeta)
    z = np.fft.ifft(fa * fs, axis=-1) * np.conj(zeta)
    y = np.concatenate([z.real, z.imag], axis=-1)
    r = np.rint(y)
    if np.abs(y - r).max(initial=0.0) >= 0.25:
        raise ArithmeticError("negacyclic product not exact in float64")
    return r.astype(np.int64) & (Q - 1)


class TorusKeys:
    """An LWE key (n bits) and a GLWE key (N bits), uniform binary, from a seeded generator."""

    def __init__(self, params: TorusParams = TorusParams(), seed: int = 0):
        self.params = params
        self.rng = np.random.default_rng(seed)
        self.lwe = self.rng.integers(0, 2, size=params.n, dtype=np.int64)
        self.glwe = self.rng.integers(0, 2, size=params.N, dtype=np.int64)

    def encrypt(self, msgs, prec: int) -> np.ndarray:
        """LWE samples (count, n + 1) of messages m < 2^prec encoded as round(m 2^32 / 2^prec)
        (test_cggi16.py's mod_switch), noise lwe_sigma."""
        p = self.params
        msgs = np.asarray(msgs, dtype=np.int64)
        a = self.rng.integers(0, Q, size=(len(msgs), p.n), dtype=np.int64)
        e = _gauss(self.rng, p.lwe_sigma, len(msgs))
        m = (msgs * Q + (1 << (prec - 1))) >> prec
        b = ((a * self.lwe).sum(axis=1) + e + m) & (Q - 1)
        return _u32(np.concatenate([a, b[:, None]], axis=1))

    def bootstrap_key(self) -> np.ndarray:
        """n GGSW encryptions of the LWE key bits under the GLWE key: (n, 2 ell, 2, N) uint32,
        row c * ell + l = a GLWE sample of 0 plus s_i 2^(32 - base_log (l+1)) on component c."""
        p = self.params
        rows = 2 * p.ell
        a = self.rng.integers(0, Q, size=(p.n, rows, p.N), dtype=np.int64)
        e = _gauss(self.rng, p.glwe_sigma, (p.n, rows, p.N))
        body = (_binary_negacyclic(a.reshape(-1, p.N), self.glwe).reshape(a.shape) + e) & (Q - 1)
        out = np.stack([a, body], axis=2)                        # (n, rows, 2, N)
        for c in range(2):
            for l in range(p.ell):
                g = 1 << (32 - p.base_log * (l + 1))
                out[:, c * p.ell + l, c, 0] += self.lwe * g
        return _u32(out)

    def phases(self, samples: np.ndarray) -> np.ndarray:
        """b - <a, s> mod 2^32 of extracted samples (count, N + 1) under the GLWE key."""
        a = np.asarray(samples[:, :-1], dtype=np.int64)
        return (np.asarray(samples[:, -1], dtype=np.int64) - (a * self.glwe).sum(axis=1)) & (Q - 1)

    def decrypt(self, samples: np.ndarray, prec: int) -> np.ndarray:
        """round(phase 2^prec / 2^32) mod 2^prec (test_cggi16.py:62's mod_switch)."""
        ph = self.phases(samples)
        return ((ph * (1 << prec) + (Q >> 1)) >> 32) & ((1 << prec) - 1)


def pack_lut(lut, size: int, prec: int, N: int = 1024) -> np.ndarray:
    """The test vector of a LUT: coefficient i = lut[i // (N / size)] 2^32 // 2^prec
    (cggi16.py:92-116 with q = 2^32)."""
    scaled = [(int(v) << 32) >> prec for v in lut]
    out = np.zeros(N, dtype=np.int64)
    for i in range(N):
        idx = i // (N // size)
        if idx < len(scaled):
            out[i] = scaled[idx]
    return _u32(out)


class TorusBatch:
    """count x dim torus values resident on the device."""

    def __init__(self, handle, count: int, dim: int):
        if not handle:
            raise ValueError(f"torus batch of {count} x {dim} rejected")
        self.handle, self.count, self.dim = handle, count, dim

    @classmethod
    def upload(cls, values: np.ndarray) -> "TorusBatch":
        v = _u32(values)
        if v.ndim == 1:
            v = v[None, :]
        return cls(lib().vfhe_hip_torus_batch_new(v.ctypes.data, v.shape[0], v.shape[1]), v.shape[0], v.shape[1])

    @classmethod
    def zeros(cls, count: int, dim: int) -> "TorusBatch":
        return cls(lib().vfhe_hip_torus_batch_new(None, count, dim), count, dim)

    def read(self) -> np.ndarray:
        out = np.empty((self.count, self.dim), dtype=np.uint32)
        if lib().vfhe_hip_torus_batch_read(self.handle, out.ctypes.data) != 0:
            raise RuntimeError("torus batch read failed")
        return out

    def free(self) -> None:
        if self.handle:
            lib().vfhe_hip_torus_batch_free(self.handle)
            self.handle = None

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass


class TorusBootstrapper:
    """The bootstrapping key on the device and the batched functional bootstrap."""

    def __init__(self, bsk: np.ndarray, params: TorusParams = TorusParams()):
        p = params
        bsk = _u32(bsk)
        if bsk.shape != (p.n, 2 * p.ell, 2, p.N):
            raise ValueError(f"bootstrapping key shape {bsk.shape}")
        self.params = p
        self.handle = lib().vfhe_hip_torus_key_new(bsk.ctypes.data, p.n, p.N, p.ell, p.base_log)
        if not self.handle:
            raise ValueError("bootstrapping key rejected (N must be 1024, ell * base_log <= 32)")

    def bootstrap(self, lwes: TorusBatch, tv: TorusBatch, torus_base: int, out: TorusBatch | None = None,
                  acc: TorusBatch | None = None) -> TorusBatch:
        """Extracted samples (count x (N + 1)) of the functional bootstrap of every input;
        `acc` (count x 2N), when given, receives the accumulators.  Queued on the device."""
        p = self.params
        if out is None:
            out = TorusBatch.zeros(lwes.count, p.N + 1)
        rc = lib().vfhe_hip_torus_bootstrap(out.handle, acc.handle if acc is not None else None, self.handle,
                                            lwes.handle, tv.handle, torus_base)
        if rc != 0:
            raise ValueError(f"vfhe_hip_torus_bootstrap rejected its arguments ({rc})")
        return out

    def free(self) -> None:
        if self.handle:
            lib().vfhe_hip_torus_key_free(self.handle)
            self.handle = None

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass


def torus_fft(polys: np.ndarray) -> np.ndarray:
    """The device's negacyclic transform of polynomials (count, 1024) (signed 32-bit
    coefficients): (count, 512) complex, natural order."""
    v = _u32(polys)
    if v.ndim == 1:
        v = v[None, :]
    out = np.empty((v.shape[0], v.shape[1] // 2, 2), dtype=np.float64)
    if lib().vfhe_hip_torus_fft(v.ctypes.data, v.shape[0], v.shape[1], out.ctypes.data) != 0:
        raise ValueError("vfhe_hip_torus_fft rejected its arguments")
    return out[..., 0] + 1j * out[..., 1]

