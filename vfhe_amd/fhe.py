"""Leveled RLWE and CKKS evaluation on the hip engine, organised around ciphertext batches.

The drop-in boundary is the engine's C ABI (include/vfhe_hip.h); this module is
the project's host layer over it, used by the tests, the bench and the serving
workload.  Its unit of work is a *batch*: B ciphertexts of one rank at one level,
each an engine `RNS_MLWE` whose limbs stay in HBM, all in one domain.  Every
operation on a batch is one batched engine call (domain conversion of all
components, the relinearized multiply with its rescale, key-switch,
automorphism) or a loop of device calls over its members; nothing is copied to
the host except by `decrypt`/`decode`.

  Chain      levels of an RNS modulus chain (and their key-switching extensions)
  SecretKey  a rank-r secret: the engine key handle plus per-ring NTT polynomials
  SwitchKey  a gadget key-switching key, encrypted once over the limbs of every
             extended ring of the chain and handed to the engine per level as
             views of those entries
  Batch      B ciphertexts at one level
  Ckks       the CKKS encoder and scale bookkeeping on top of Batch

Two pieces restate the reference because they fix bits of every key-switched
result: the nested level chain (`nested_levels`, mlwe.py:57-73) and the gadget
plaintext factor of a key-switching digit (`gadget_factor`, mlwe.py:156-176).
Randomness (secret support, encryption noise, uniform masks) comes from the
engine's ChaCha20 sampler and a seeded numpy generator; it is not a parity
target (the reference's own engines draw from different generators).
"""
from __future__ import annotations

import ctypes
import math

import numpy as np

from . import _native
from .arith import NTT_processor_instance, Polynomial, Ring, crt, repr
from .complex import ComplexPolynomial, ComplexRing

U64P = _native.U64P


def lib():
    return _native.load()


def _handles(xs) -> ctypes.Array:
    return _native.voidp_array(list(xs))


# ---------------------------------------------------------------------------------------------
# Level chains
# ---------------------------------------------------------------------------------------------
def nested_levels(top: Ring, special: int = 0, levels: int | None = None):
    """The level chain a single ring defines (modules/mlwe/python/src/vfhe/mlwe/mlwe.py:57-73).

    Level i keeps the first ell - special - i primes of `top` (the top `special` primes are the
    key-switching primes); its extended ring is the level's primes plus the special primes.  The
    default number of levels is len(top.primes) - 1.  Returns (level rings, extended rings)."""
    n = levels if levels is not None else len(top.primes) - 1
    lv = [top.quotient_ring(ell=top.ell - special - i) for i in range(n)]
    if special == 0:
        return lv, lv
    smask = ((1 << special) - 1) << lv[0].ell
    return lv, [top.quotient_ring(mask=r.mask | smask) for r in lv]


def _union(rings) -> Ring:
    """The ring over the union of the rings' primes, in pool (limb index) order."""
    r0 = rings[0]
    mask = 0
    for r in rings:
        mask |= r.mask
    if mask == r0.mask:
        return r0
    index = {}
    for r in rings:
        index.update(zip(r.primes, r.prime_indices))
    primes = sorted(index, key=index.get)
    return Ring(r0.N, mask=mask, primes=primes, split_degree=r0.split_degree)


class Chain:
    """Levels 0..n-1 of a modulus chain on one engine prime pool, with the extended ring of each
    level (level primes plus the key-switching primes; the level ring itself without them)."""

    def __init__(self, levels: list[Ring], ext: list[Ring] | None = None, rank: int = 1):
        self.levels = list(levels)
        self.ext = list(ext) if ext is not None else list(levels)
        assert len(self.ext) == len(self.levels)
        self.rank = rank
        self.N = self.levels[0].N
        specials = {e.mask & ~lv.mask for lv, e in zip(self.levels, self.ext)}
        assert len(specials) == 1, "every level must share the key-switching primes"
        self.special_mask = specials.pop()
        for lv, e in zip(self.levels, self.ext):
            assert lv.is_quotient_ring(e) and not (lv.mask & self.special_mask)
        self.special = bin(self.special_mask).count("1")
        # every limb a key-switching key may be asked for, at any level
        self.span = _union(self.ext)

    @classmethod
    def nested(cls, top: Ring, special: int = 0, levels: int | None = None, rank: int = 1) -> "Chain":
        lv, ext = nested_levels(top, special, levels)
        return cls(lv, ext, rank)

    def __len__(self):
        return len(self.levels)

    def level_of(self, ring: Ring) -> int:
        for i, r in enumerate(self.levels):
            if r.mask == ring.mask:
                return i
        raise ValueError("ring is not a level of this chain")

    @property
    def special_modulus(self) -> int:
        return math.prod(p for p, i in zip(self.span.primes, self.span.prime_indices)
                         if (self.special_mask >> i) & 1)

    @property
    def extended_rank(self) -> int:
        """Components of a product before relinearization: r(r+1)/2 quadratic + r linear."""
        return self.rank * (self.rank + 3) // 2


def gadget_factor(ring: Ring, limb: int, special_modulus: int) -> list[int]:
    """Plaintext factor of the key-switching digit for pool limb `limb`
    (modules/mlwe/python/src/vfhe/mlwe/mlwe.py:156-176): P mod q at that limb and 0 at every other
    limb of `ring`, P the product of the special primes; the digit encrypts source * factor, i.e.
    P * source times the RNS indicator of the limb, so any sub-ring holding the limb sees the same
    plaintext restricted to its limbs."""
    return [special_modulus % p if idx == limb else 0 for p, idx in zip(ring.primes, ring.prime_indices)]


# ---------------------------------------------------------------------------------------------
# Keys
# ---------------------------------------------------------------------------------------------
class SecretKey:
    """A rank-r secret with small signed coefficients, held as one engine key over every limb of
    the pool (so it serves every ring of the chain) plus NTT-form polynomials per ring."""

    def __init__(self, chain: Chain, coeffs: list[list[int]], sigma: float):
        assert len(coeffs) == chain.rank and all(len(c) == chain.N for c in coeffs)
        self.chain, self.sigma = chain, float(sigma)
        self.coeffs = [[int(v) for v in c] for c in coeffs]
        pool = chain.levels[0].NTT
        n_limbs = int(ctypes.cast(pool, ctypes.POINTER(_native.IncNTT)).contents.l)
        flat = np.array([v & 0xFFFFFFFFFFFFFFFF for c in self.coeffs for v in c], dtype=np.uint64)
        self.h = lib().mlwe_new_RNS_key_from_array(flat.ctypes.data_as(U64P), chain.N, chain.rank, n_limbs,
                                                   pool, self.sigma)
        self._polys = {}

    @classmethod
    def sparse(cls, chain: Chain, h: int, sigma: float, rng: np.random.Generator, ternary: bool = True):
        """h non-zero coefficients over the r*N positions, alternately +1 and -1 (all +1 when not
        ternary), at positions drawn without repetition."""
        n = chain.N * chain.rank
        assert 0 < h <= n
        pos = rng.choice(n, size=h, replace=False)
        flat = [0] * n
        for k, p in enumerate(pos):
            flat[int(p)] = -1 if (ternary and k % 2) else 1
        return cls(chain, [flat[i * chain.N:(i + 1) * chain.N] for i in range(chain.rank)], sigma)

    def __del__(self):
        try:
            lib().free_mlwe_RNS_key(self.h)
        except Exception:
            pass

    def poly(self, i: int, ring: Ring) -> Polynomial:
        """s_i on `ring`, NTT form (cached)."""
        k = (i, ring.mask)
        if k not in self._polys:
            p = Polynomial(ring).from_array(self.coeffs[i])
            p.to_NTT()
            self._polys[k] = p
        return self._polys[k]

    def quadratic(self, ring: Ring) -> list[Polynomial]:
        """The tensor product's quadratic key terms -(s_i s_j), i <= j lexicographic
        (the slot order of mlwe_tensor_product, mlwe.c:684-708)."""
        r = self.chain.rank
        return [-(self.poly(i, ring) * self.poly(j, ring)) for i in range(r) for j in range(i, r)]

    def lwe_coeffs(self) -> list[int]:
        """The LWE secret of an extracted sample (mlwe.c:40-66): the components' coefficients in order."""
        return [v for c in self.coeffs for v in c]


# ---------------------------------------------------------------------------------------------
# Ciphertext and key files (the engine's block format, vfhe_hip_mlwe_serialize)
# ---------------------------------------------------------------------------------------------
_FORM_CODES = {"ntt": 1, "coeff": 2}
_KEY_MAGIC = 0x3159454B45484656   # "VFHEKEY1"


def _samples_to_bytes(handles, form: str | None) -> bytes:
    arr = _handles(handles)
    size = int(lib().vfhe_hip_mlwe_serialized_size(arr, len(handles)))
    buf = np.empty(size // 8, dtype=np.uint64)
    n = lib().vfhe_hip_mlwe_serialize(arr, len(handles), _FORM_CODES.get(form, 0),
                                      buf.ctypes.data_as(ctypes.c_void_p), size)
    if n != size:
        raise ValueError("samples of different rings, ranks or limb masks cannot share one block")
    return buf.tobytes()


def _samples_from_bytes(data: bytes, ring: Ring, max_count: int = 1 << 20):
    """(handles, form) of a block read onto `ring`; ValueError names what is wrong."""
    if len(data) % 8 or len(data) < 72:
        raise ValueError("truncated block")
    words = np.frombuffer(data, dtype=np.uint64).copy()   # 8-byte aligned for the engine
    count = int(words[2])
    out = (ctypes.c_void_p * max(1, min(count, max_count)))()
    form = ctypes.c_uint32(0)
    n = lib().vfhe_hip_mlwe_deserialize(words.ctypes.data_as(ctypes.c_void_p), words.nbytes, ring.NTT, out,
                                        max(1, min(count, max_count)), ctypes.byref(form))
    if n < 0:
        raise ValueError({-1: "not a vfhe-hip sample block, or its check word does not match",
                          -2: "the block belongs to another ring (N, split degree or primes differ)",
                          -3: "more samples than allowed", -4: "truncated block"}.get(int(n), "bad block"))
    return [out[k] for k in range(n)], {1: "ntt", 2: "coeff"}.get(form.value)


class SwitchKey:
    """Gadget key-switching key from the sources to `target`.

    `sources[k]` is the key component that input component k is multiplied by (a relinearization
    key has the quadratic terms and then None for each linear component, which the key-switch
    passes through).  Every source gets one digit per limb of the chain's span, each an encryption
    under `target` of source * gadget_factor, generated once; level l uses the digits of its active
    limbs in ascending order, as views (shared device rows) restricted to its extended ring."""

    def __init__(self, chain: Chain, target: SecretKey, sources: list):
        self.chain = chain
        span, P = chain.span, chain.special_modulus
        self.digits = []        # per source: list of engine samples (NTT form, on span) or None
        for src in sources:
            if src is None:
                self.digits.append(None)
                continue
            if src.ring != span:
                src = Polynomial(span).from_bigint_array(src.get_polynomial(signed=True))
            src.to_NTT()
            plains = []
            for limb in span.prime_indices:
                m = src * gadget_factor(span, limb, P)
                m.to_coeff()
                plains.append(m)
            b = Batch.encrypt(chain, target, plains, ring=span)
            b.to_ntt()
            self.digits.append(b)
        self._views = {}

    def at(self, lvl: int) -> ctypes.Array:
        """The `RNS_MLWE **ksk` array the engine takes at level `lvl`."""
        if lvl not in self._views:
            chain = self.chain
            level, ext = chain.levels[lvl], chain.ext[lvl]
            # the engine consumes one digit per active limb of the input, in ascending limb order
            pos = [chain.span.prime_indices.index(i) for i in sorted(level.prime_indices)]
            arrays, entries = [], []
            for d in self.digits:
                if d is None:
                    entries.append(None)
                    continue
                sel = _handles([d.handles[k] for k in pos])
                v = lib().vfhe_hip_mlwe_view_array(sel, len(pos), ext.mask)
                arrays.append((v, len(pos)))
                entries.append(ctypes.cast(v, ctypes.c_void_p).value)
            self._views[lvl] = (_handles(entries), arrays)
        return self._views[lvl][0]

    def to_bytes(self) -> bytes:
        """"VFHEKEY1", the source count, then per source a length (0: a pass-through source) and
        the engine block of its digits (NTT form, on the chain's span)."""
        parts = [np.array([_KEY_MAGIC, len(self.digits)], dtype=np.uint64).tobytes()]
        for d in self.digits:
            blob = b"" if d is None else _samples_to_bytes(d.handles, "ntt")
            parts.append(np.array([len(blob)], dtype=np.uint64).tobytes())
            parts.append(blob)
        return b"".join(parts)

    def save(self, path: str):
        with open(path, "wb") as f:
            f.write(self.to_bytes())

    @classmethod
    def from_bytes(cls, data: bytes, chain: Chain) -> "SwitchKey":
        """The key written by to_bytes, on `chain` (whose span must hold the same primes)."""
        head = np.frombuffer(data[:16], dtype=np.uint64)
        if head.size != 2 or int(head[0]) != _KEY_MAGIC:
            raise ValueError("not a vfhe-hip switching key")
        self = cls.__new__(cls)
        self.chain, self.digits, self._views = chain, [], {}
        off = 16
        for _ in range(int(head[1])):
            if off + 8 > len(data):
                raise ValueError("truncated switching key")
            n = int(np.frombuffer(data[off:off + 8], dtype=np.uint64)[0])
            off += 8
            if n == 0:
                self.digits.append(None)
                continue
            hs, _ = _samples_from_bytes(data[off:off + n], chain.span)
            off += n
            self.digits.append(Batch(chain, 0, hs, "ntt", chain.rank, chain.span))
        return self

    @classmethod
    def load(cls, path: str, chain: Chain) -> "SwitchKey":
        with open(path, "rb") as f:
            return cls.from_bytes(f.read(), chain)

    def __del__(self):
        try:
            lib().vfhe_hip_synchronize()
            for _, arrays in getattr(self, "_views", {}).values():
                for v, n in arrays:
                    lib().free_RNS_mlwe_array(n, v)
        except Exception:
            pass


def relinearization_key(chain: Chain, key: SecretKey) -> SwitchKey:
    """Switches the quadratic components of a product back to `key`; linear ones pass through."""
    return SwitchKey(chain, key, key.quadratic(chain.span) + [None] * chain.rank)


def switching_key(chain: Chain, target: SecretKey, source: SecretKey) -> SwitchKey:
    return SwitchKey(chain, target, [source.poly(i, chain.span) for i in range(chain.rank)])


def galois_key(chain: Chain, key: SecretKey, gen: int) -> SwitchKey:
    """Switches a ciphertext permuted by X -> X^gen (mlwe.c:314-326) back to `key`."""
    return SwitchKey(chain, key, [key.poly(i, chain.span).copy().automorphism(gen) for i in range(chain.rank)])


# ---------------------------------------------------------------------------------------------
# Ciphertext batches
# ---------------------------------------------------------------------------------------------
class Batch:
    """B ciphertexts of one rank at level `lvl` (on `ring`, the level ring unless stated), each an
    engine sample, all in one domain ('ntt' or 'coeff')."""

    def __init__(self, chain: Chain, lvl: int, handles: list, form: str, rank: int, ring: Ring | None = None,
                 delta: float | None = None):
        self.chain, self.lvl, self.handles, self.form, self.rank = chain, lvl, list(handles), form, rank
        self.ring = ring if ring is not None else chain.levels[lvl]
        self.delta = delta

    @classmethod
    def alloc(cls, chain: Chain, lvl: int, count: int, rank: int | None = None, ring: Ring | None = None,
              form: str = "ntt", delta: float | None = None, uninit: bool = False) -> "Batch":
        """`count` samples on the level's ring, zeroed; uninit: not cleared (outputs written
        before they are read), allocated in one engine call."""
        ring = ring if ring is not None else chain.levels[lvl]
        rank = rank if rank is not None else chain.rank
        if uninit:
            arr = (ctypes.c_void_p * max(1, count))()
            lib().vfhe_hip_mlwe_alloc_RNS_samples_uninit(arr, count, ring.N, rank, ring.mask, ring.NTT)
            hs = [arr[k] for k in range(count)]
        else:
            hs = [lib().mlwe_alloc_RNS_sample(ring.N, rank, ring.mask, ring.NTT) for _ in range(count)]
        return cls(chain, lvl, hs, form, rank, ring, delta)

    @classmethod
    def encrypt(cls, chain: Chain, key: SecretKey, plains: list, lvl: int = 0, ring: Ring | None = None,
                delta: float | None = None) -> "Batch":
        """Fresh encryptions b = sum_i a_i s_i + e + m (mlwe.c:190-241) of coefficient-form
        plaintexts on the batch's ring; a plaintext of another ring is re-expressed there from its
        signed representative.  Result in coefficient form."""
        out = cls.alloc(chain, lvl, len(plains), ring=ring, form="coeff", delta=delta)
        for h, m in zip(out.handles, plains):
            if m.ring != out.ring:
                m = Polynomial(out.ring).from_bigint_array(m.get_polynomial(signed=True))
            m.to_coeff()
            lib().mlwe_RNSc_sample(h, key.h, m.obj)
        return out

    def __del__(self):
        try:
            if self.handles:
                lib().vfhe_hip_free_mlwe_samples(_handles(self.handles), len(self.handles))
        except Exception:
            pass

    def __len__(self):
        return len(self.handles)

    def _like(self, handles, form=None, lvl=None, ring=None, rank=None, delta="same") -> "Batch":
        return Batch(self.chain, self.lvl if lvl is None else lvl, handles, self.form if form is None else form,
                     self.rank if rank is None else rank, self.ring if ring is None else ring,
                     self.delta if delta == "same" else delta)

    def components(self) -> list:
        out = []
        for h in self.handles:
            s = ctypes.cast(h, ctypes.POINTER(_native.RNSMLWE)).contents
            out += [s.a[i] for i in range(int(s.r))] + [s.b]
        return out

    def _to(self, form: str):
        if self.form != form and self.handles:
            comps = self.components()
            lib().vfhe_hip_ntt_batch(_handles(comps), len(comps), 0 if form == "ntt" else 1)
        self.form = form
        return self

    def to_ntt(self):
        return self._to("ntt")

    def to_coeff(self):
        return self._to("coeff")

    def split(self) -> list["Batch"]:
        """One-ciphertext batches sharing this batch's samples (the parent keeps ownership)."""
        return [_Borrowed(self, h) for h in self.handles]

    def polys(self, k: int) -> list[Polynomial]:
        """Copies of member k's components a_0..a_{r-1}, b as polynomials of the batch's ring."""
        s = ctypes.cast(self.handles[k], ctypes.POINTER(_native.RNSMLWE)).contents
        out = []
        for h in [s.a[i] for i in range(int(s.r))] + [s.b]:
            p = Polynomial(self.ring)
            lib().polynomial_copy_RNS_polynomial(p.obj, h)
            p.repr = repr.ntt if self.form == "ntt" else repr.coeff
            out.append(p)
        return out

    def copy(self) -> "Batch":
        out = Batch.alloc(self.chain, self.lvl, len(self), self.rank, self.ring, self.form, self.delta)
        for d, s in zip(out.handles, self.handles):
            lib().mlwe_copy_RNS_sample(d, s)
        return out

    # ---- decryption ---------------------------------------------------------------------
    def to_bytes(self) -> bytes:
        """The batch as one engine block (vfhe_hip_mlwe_serialize), its domain recorded."""
        return _samples_to_bytes(self.handles, self.form)

    def save(self, path: str):
        with open(path, "wb") as f:
            f.write(self.to_bytes())

    @classmethod
    def from_bytes(cls, data: bytes, chain: Chain, lvl: int = 0, ring: Ring | None = None,
                   delta: float | None = None) -> "Batch":
        """A batch written by to_bytes, read onto `ring` (the level ring unless stated)."""
        ring = ring if ring is not None else chain.levels[lvl]
        hs, form = _samples_from_bytes(data, ring)
        rank = int(np.frombuffer(data[32:40], dtype=np.uint64)[0])
        return cls(chain, lvl, hs, form or "ntt", rank, ring, delta)

    @classmethod
    def load(cls, path: str, chain: Chain, lvl: int = 0, ring: Ring | None = None,
             delta: float | None = None) -> "Batch":
        with open(path, "rb") as f:
            return cls.from_bytes(f.read(), chain, lvl, ring, delta)

    def decrypt(self, key: SecretKey) -> list[Polynomial]:
        """The phases b - sum_i a_i s_i (mlwe.c:243-252), NTT form, one polynomial per member."""
        self.to_ntt()
        out = []
        for h in self.handles:
            p = Polynomial(self.ring)
            lib().mlwe_RNS_phase(p.obj, h, key.h)
            p.repr = repr.ntt
            out.append(p)
        return out

    # ---- linear operations ---------------------------------------------------------------
    def _pairwise(self, other: "Batch", fn: str) -> "Batch":
        assert len(other) == len(self) and other.ring == self.ring and other.rank == self.rank
        other._to(self.form)
        out = Batch.alloc(self.chain, self.lvl, len(self), self.rank, self.ring, self.form, self.delta)
        for o, a, b in zip(out.handles, self.handles, other.handles):
            getattr(lib(), fn)(o, a, b)
        return out

    def __add__(self, other: "Batch") -> "Batch":
        return self._pairwise(other, "mlwe_add_RNSc_sample")

    def __sub__(self, other: "Batch") -> "Batch":
        return self._pairwise(other, "mlwe_sub_RNSc_sample")

    def mul_plain(self, polys: list[Polynomial]) -> "Batch":
        """Member k times polys[k] (a polynomial of the batch's ring or of a sub-ring, which is
        base-extended first), NTT form (mlwe.c:254-261)."""
        self.to_ntt()
        out = Batch.alloc(self.chain, self.lvl, len(self), self.rank, self.ring, "ntt", self.delta)
        for o, h, p in zip(out.handles, self.handles, polys):
            if p.ring != self.ring:
                assert p.ring.is_quotient_ring(self.ring)
                p = p.base_extend(self.ring)
            p.to_NTT()
            lib().mlwe_RNS_mul_by_poly(o, h, p.obj)
        return out

    # ---- products and key-switching -----------------------------------------------------------
    def tensor(self, other: "Batch") -> "Batch":
        """The unrelinearized products (mlwe_multiply without a key, mlwe.c:711-760): rank
        r(r+3)/2, NTT form."""
        assert len(other) == len(self) and other.lvl == self.lvl and other.rank == self.rank
        self.to_ntt()
        other.to_ntt()
        out = Batch.alloc(self.chain, self.lvl, len(self), self.chain.extended_rank, self.ring, "ntt", self.delta)
        lib().vfhe_hip_mlwe_multiply_batch(_handles(out.handles), _handles(self.handles), _handles(other.handles),
                                           len(self), None, 0)
        return out

    def multiply(self, other: "Batch", rlk: SwitchKey, divide_mask: int = 0, graph: bool = False) -> "Batch":
        """Relinearized products of the members pairwise, one batched engine call
        (mlwe_multiply, mlwe.c:711-760, for all of them).  divide_mask = 0: NTT form on the level
        ring; otherwise the result is also round-divided by those primes (the CKKS rescale) and left
        in coefficient form on the remaining limbs.  graph: replay the captured launch sequence of
        this batch shape (vfhe_hip_mlwe_multiply_batch_graph)."""
        assert len(other) == len(self) and other.lvl == self.lvl and other.rank == self.rank == self.chain.rank
        self.to_ntt()
        other.to_ntt()
        ext = self.chain.ext[self.lvl]
        out = Batch.alloc(self.chain, self.lvl, len(self), self.rank, ext, "ntt", self.delta)
        fn = lib().vfhe_hip_mlwe_multiply_batch_graph if graph else lib().vfhe_hip_mlwe_multiply_batch
        fn(_handles(out.handles), _handles(self.handles), _handles(other.handles), len(self), rlk.at(self.lvl),
           divide_mask)
        out.ring = self.ring
        if divide_mask:
            out.form = "coeff"
            out.ring = Ring(self.ring.N, mask=self.ring.mask & ~divide_mask, primes=self.ring.primes,
                            split_degree=self.ring.split_degree)
        return out

    def keyswitch(self, ksk: SwitchKey) -> "Batch":
        """GHS hybrid key-switch of every member (mlwe.c:557-598), coefficient form on the level ring;
        the input may be a product of extended rank (relinearization)."""
        self.to_coeff()
        ext = self.chain.ext[self.lvl]
        out = Batch.alloc(self.chain, self.lvl, len(self), self.chain.rank, ext, "coeff", self.delta)
        lib().vfhe_hip_mlwe_keyswitch_batch(_handles(out.handles), _handles(self.handles), len(self), ksk.at(self.lvl))
        out.ring = self.ring
        return out

    def automorphism(self, gen: int, ksk: SwitchKey) -> "Batch":
        """X -> X^gen on every member, then key-switched back (mlwe.c:314-326), coefficient form."""
        self.to_coeff()
        ext = self.chain.ext[self.lvl]
        out = Batch.alloc(self.chain, self.lvl, len(self), self.rank, ext, "coeff", self.delta)
        lib().vfhe_hip_mlwe_automorphism_batch(_handles(out.handles), _handles(self.handles), len(self), gen,
                                               ksk.at(self.lvl))
        out.ring = self.ring
        return out

    def round_divide(self, to: Ring, lvl: int | None = None) -> "Batch":
        """In place: round-divide every member by the primes of its ring missing from `to`
        (mlwe_round_division_RNSc, mlwe.c:762-772); `to` must be a sub-ring."""
        assert to.is_quotient_ring(self.ring) and to.mask != self.ring.mask
        self.to_coeff()
        drop = self.ring.mask & ~to.mask
        for h in self.handles:
            lib().mlwe_round_division_RNSc(h, drop)
        self.ring = to
        if lvl is not None:
            self.lvl = lvl
        return self


class _Borrowed(Batch):
    """A one-member view of another batch's sample (never freed here).  Its domain is the parent's:
    converting the view converts the whole parent, so every member keeps one domain."""

    def __init__(self, parent: Batch, h):
        self._parent = parent
        super().__init__(parent.chain, parent.lvl, [h], parent.form, parent.rank, parent.ring, parent.delta)

    @property
    def form(self):
        return self._parent.form

    @form.setter
    def form(self, value):
        if value != self._parent.form:
            self._parent._to(value)

    def _to(self, form: str):
        self._parent._to(form)
        return self

    def __del__(self):
        pass


# ---------------------------------------------------------------------------------------------
# CKKS
# ---------------------------------------------------------------------------------------------
class Ckks:
    """CKKS on a chain: the special-root encoder (slot j at zeta^(5^j), zeta = exp(i pi / N)), the
    scale carried by each batch (`Batch.delta`), and the level moves of a rescale.

    Rotation follows the reference's convention: rotating by k moves slot i + k to slot i, i.e. the
    automorphism X -> X^(5^k mod 2N) (modules/fhe/python/src/vfhe/fhe/ckks.py:110-126)."""

    def __init__(self, chain: Chain, scale: float):
        self.chain, self.scale = chain, float(scale)
        self.slots = chain.N // 2
        self.cring = ComplexRing(self.slots, True)

    # ---- encoding ---------------------------------------------------------------------------
    def encode(self, values, ring: Ring | None = None, scale: float | None = None) -> Polynomial:
        """Slots -> bit reversal, GS_RN inverse FFT, 1/n -> times the scale -> rounded into RNS
        (NTT form), all on the device (complex.c)."""
        assert len(values) == self.slots
        c = ComplexPolynomial(self.cring).from_array(values)
        c.IFFT()
        c *= float(self.scale if scale is None else scale)
        ring = ring if ring is not None else self.chain.levels[0]
        poly = c.round_to_RNS_cpp(ring)
        poly.struct.rns_mask = ring.mask   # the rounding fills every pool limb (ckks.py:54)
        return poly

    def encode_many(self, vectors: list, ring: Ring | None = None) -> list[Polynomial]:
        """encode() of many vectors in one batched device call."""
        cps = [ComplexPolynomial(self.cring).from_array(v) for v in vectors]
        ring = ring if ring is not None else self.chain.levels[0]
        polys = self.cring.exp_complex_polys_ifft_scale_round_to_RNS_batch(cps, ring, self.scale)
        for p in polys:
            p.struct.rns_mask = ring.mask
        return polys

    def decode(self, poly: Polynomial, scale: float | None = None) -> list[complex]:
        """Centred CRT reconstruction of the coefficients over the active limbs, the pairs
        (c_i, c_{i+N/2}) as complex values, CT_NR forward FFT and bit reversal on the device,
        divided by the scale."""
        scale = self.scale if scale is None else scale
        mask = poly.rns_mask
        rows = poly.get_coeff_matrix(repr.coeff)
        act = [(rows[k], p) for k, (p, idx) in enumerate(zip(poly.ring.primes, poly.ring.prime_indices))
               if (mask >> idx) & 1]
        mods = [p for _, p in act]
        Q = math.prod(mods)
        n = poly.ring.N
        vals = []
        for j in range(n):
            v = crt([r[j] for r, _ in act], mods)
            vals.append(float(v - Q if v >= Q // 2 else v))
        c = ComplexPolynomial(self.cring)
        c.obj[: self.slots] = vals[: self.slots]
        c.obj[self.slots:] = vals[self.slots:]
        c.FFT()
        return [v / scale for v in c]

    # ---- encryption ---------------------------------------------------------------------------
    def encrypt(self, key: SecretKey, vectors: list, lvl: int = 0) -> Batch:
        ring = self.chain.levels[lvl]
        return Batch.encrypt(self.chain, key, [self.encode(v, ring) for v in vectors], lvl=lvl, delta=self.scale)

    def decrypt(self, key: SecretKey, batch: Batch) -> list[list[complex]]:
        return [self.decode(p, batch.delta) for p in batch.decrypt(key)]

    # ---- level moves ----------------------------------------------------------------------------
    def rescale(self, b: Batch) -> Batch:
        """To the next level (modules/fhe/python/src/vfhe/fhe/ckks.py:128-148): round-divide by the
        primes it drops when the next level is a sub-ring, a rational rescale otherwise."""
        nxt = b.lvl + 1
        assert nxt < len(self.chain), "no lower level to rescale into"
        to = self.chain.levels[nxt]
        if not to.is_quotient_ring(b.ring):
            return self.rational_rescale(b)
        dropped = math.prod(p for p in b.ring.primes if p not in to.primes)
        b.round_divide(to, nxt)
        b.delta = b.delta / dropped
        return b

    def rational_rescale(self, b: Batch) -> Batch:
        """To the next level when it is not a sub-ring (ckks.py:150-202): every component is lifted
        to the union of both levels' primes scaled by the primes it gains (F16,
        polynomial_RNSc_scaled_lift, which also clears the new limbs), then round-divided by the
        primes the next level lacks; the scale becomes delta * (gained primes) / (dropped primes)."""
        nxt = b.lvl + 1
        assert nxt < len(self.chain), "no lower level to rescale into"
        start, end = b.ring, self.chain.levels[nxt]
        union = _union([start, end])
        b.to_coeff()
        out = Batch.alloc(self.chain, nxt, len(b), b.rank, end, "coeff")
        lifted = Polynomial(union)
        for src, dst in zip(b.handles, out.handles):
            s = ctypes.cast(src, ctypes.POINTER(_native.RNSMLWE)).contents
            d = ctypes.cast(dst, ctypes.POINTER(_native.RNSMLWE)).contents
            for sp, dp in [(s.a[i], d.a[i]) for i in range(int(s.r))] + [(s.b, d.b)]:
                lifted.struct.rns_mask = union.mask   # the division below narrowed it
                lib().polynomial_RNSc_scaled_lift(lifted.obj, sp, None)
                lib().polynomial_round_division_RNSc_wo_free(lifted.obj, union.mask & ~end.mask)
                lib().polynomial_copy_RNS_polynomial(dp, lifted.obj)
        gained = math.prod(p for p in end.primes if p not in start.primes)
        dropped = math.prod(p for p in start.primes if p not in end.primes)
        out.delta = b.delta * gained / dropped
        return out

    # ---- evaluation -----------------------------------------------------------------------------
    def multiply(self, a: Batch, b: Batch, rlk: SwitchKey, graph: bool = False) -> Batch:
        """Relinearized product, then the rescale to the next level.  When the next level is the
        current one minus one prime the engine fuses the rescale into the multiply's ModDown
        (k_moddown_rescale); otherwise the product is rescaled afterwards."""
        delta = a.delta * b.delta
        nxt = a.lvl + 1
        to = self.chain.levels[nxt] if nxt < len(self.chain) else None
        if to is not None and to.is_quotient_ring(a.ring):
            drop = a.ring.mask & ~to.mask
            out = a.multiply(b, rlk, drop, graph=graph)
            out.ring, out.lvl = to, nxt
            out.delta = delta / math.prod(p for p in a.ring.primes if p not in to.primes)
            return out
        out = a.multiply(b, rlk)
        out.delta = delta
        return self.rescale(out)

    def mul_plain(self, a: Batch, polys: list[Polynomial]) -> Batch:
        """Product with encoded plaintexts (at the scheme's scale), then the rescale."""
        out = a.mul_plain(polys)
        out.delta = a.delta * self.scale
        return self.rescale(out)

    def galois_element(self, k: int) -> int:
        return pow(5, k % self.slots, 2 * self.chain.N)

    def rotation_key(self, key: SecretKey, k: int) -> SwitchKey:
        return galois_key(self.chain, key, self.galois_element(k))

    def rotate(self, b: Batch, k: int, rk: SwitchKey) -> Batch:
        return b.automorphism(self.galois_element(k), rk)


__all__ = ["Chain", "SecretKey", "SwitchKey", "Batch", "Ckks", "nested_levels", "gadget_factor",
           "relinearization_key", "switching_key", "galois_key", "NTT_processor_instance"]
