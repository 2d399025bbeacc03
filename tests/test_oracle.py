"""Pins the CPU oracle (oracle/vfhe_oracle.c) before it is trusted as the checker.

Each block checks the oracle against something it does not itself compute:
the reference's own test vectors and known answers, the reference's compiled
kernels (oracle/_ref), closed forms, and big-integer CRT arithmetic.  A few
tests plant a mistake in a copy of an oracle result to show the check would
catch it.
"""
import math
import random

import numpy as np
import pytest

from oracle import oracle as O
from vfhe_amd import params

R = random.Random(0x5EED)


# --- number theory: modules/arith/c/test/test_number_theory.c ---------------------------------

def test_number_theory_known_answers():
    L = O.lib()
    assert L.o_is_prime(562949954142209)            # test_number_theory.c:14-18
    assert not L.o_is_prime(100)
    n, start = 1024, 1 << 40                         # :20-29
    sp = int(L.o_next_special_prime(start, n, 1))
    assert sp > start and L.o_is_prime(sp) and sp % (2 * n) == 1 and sp % (4 * n) != 1
    sp2 = int(L.o_next_special_prime(start, n, 0))   # :31-37
    assert sp2 > start and L.o_is_prime(sp2) and sp2 % (2 * n) == 1
    p = 1073739937                                   # :49-58
    for a in range(1, 1000):
        assert int(L.o_inverse_mod(a, p)) * a % p == 1
    # root of unity for q = 769, n = 16 (:39-47): the order-2n root rule of ntt.c:271-281
    psi = int(L.o_find_psi(8, 769))
    assert pow(psi, 8, 769) == 768


def test_primality_matches_python_bigint_rule():
    """Oracle Miller-Rabin == the package's restatement == trial division on a sample."""
    L = O.lib()
    for _ in range(300):
        n = R.randrange(2, 10**6)
        td = n > 1 and all(n % d for d in range(2, int(math.isqrt(n)) + 1))
        assert bool(L.o_is_prime(n)) == td == params.is_prime(n)
    for _ in range(200):
        n = R.randrange(2**40, 2**62) | 1
        assert bool(L.o_is_prime(n)) == params.is_prime(n)


def test_gen_prime_rule():
    """Ring.gen_prime (polynomial.py:151-160): descending odd multiples, first candidate skipped."""
    for N in (4096, 65536):
        ps = O.gen_primes(N, [50] * 6)
        assert ps == params.gen_primes(N, [50] * 6)
        order = 2 * N
        a0 = ((2**50 - 1) // order) | 1
        prev = (a0 - 2) * order + 1 + 2 * order
        for p in ps:
            assert p < 2**50 and p % order == 1 and ((p - 1) // order) % 2 == 1 and params.is_prime(p)
            assert p < prev
            # no prime of the rule's form lies strictly between consecutive outputs
            for c in range(prev - 2 * order, p, -2 * order):
                assert not params.is_prime(c)
            prev = p
    assert len(set(ps)) == len(ps)


def test_psi_known_answers():
    """psi values printed by the reference's own ntt_new_proc (built from its sources) at n = 4096."""
    assert O.lib().o_find_psi(4096, 562949953216513) == 169658170966992
    assert O.lib().o_find_psi(4096, 1125899906424833) == 371347710277771


# --- NTT ----------------------------------------------------------------------------------------

def brv(x, bits):
    return int(format(x, f"0{bits}b")[::-1], 2)


@pytest.mark.parametrize("n", [16, 64, 256])
def test_ntt_closed_form(n):
    """NTT(x)[i] = sum_j x_j psi^(j (2 brv(i) + 1)) mod q, brute force with Python ints."""
    for q in (O.gen_prime(2 * n, 50), int(O.lib().o_next_special_prime(1 << 60, n, 1)), 7681 if n <= 256 else 12289):
        if (q - 1) % (2 * n):
            continue
        psi = int(O.lib().o_find_psi(n, q))
        x = [R.randrange(q) for _ in range(n)]
        a = np.array(x, dtype=np.uint64)
        O.lib().o_ntt_forward(O.ptr(a), n, q)
        logn = n.bit_length() - 1
        for i in range(n):
            e = 2 * brv(i, logn) + 1
            assert int(a[i]) == sum(x[j] * pow(psi, j * e, q) for j in range(n)) % q
        # a planted mistake -- natural instead of bit-reversed output order -- fails the same check
        assert any(int(a[i]) != sum(x[j] * pow(psi, j * (2 * i + 1), q) for j in range(n)) % q
                   for i in range(1, n))


@pytest.mark.parametrize("engine", O.ref_engines())
def test_ntt_matches_reference_kernels(engine):
    """Bit-exact against the reference's own ntt_forward / ntt_reverse (oracle/_ref)."""
    rng = np.random.default_rng(3)
    for n in (64, 1024, 4096, 16384, 65536):
        for q in (O.gen_prime(2 * n, 50), O.gen_prime(2 * n, 40),
                  int(O.lib().o_next_special_prime(1 << 30, n, 1)), int(O.lib().o_next_special_prime(1 << 60, n, 1))):
            ref = O.RefProc(engine, n, q)
            x = rng.integers(0, q, n, dtype=np.uint64)
            a = x.copy()
            O.lib().o_ntt_forward(O.ptr(a), n, q)
            assert np.array_equal(a, ref.forward(x)), (engine, n, q)
            b = a.copy()
            O.lib().o_ntt_inverse(O.ptr(b), n, q)
            assert np.array_equal(b, x)
            y = rng.integers(0, q, n, dtype=np.uint64)
            c = y.copy()
            O.lib().o_ntt_inverse(O.ptr(c), n, q)
            assert np.array_equal(c, ref.reverse(y)), (engine, n, q)


def test_ntt_and_eltwise_match_golden_vectors():
    """tests/golden/ref_ntt_mod.npz holds the reference kernels' outputs (no reference needed here)."""
    from conftest import golden_ntt_mod_cases
    ops = {"mul": O.O_MUL, "add": O.O_ADD, "sub": O.O_SUB}
    for n, q, x, y, exp in golden_ntt_mod_cases():
        a = x.copy()
        O.lib().o_ntt_forward(O.ptr(a), n, q)
        assert np.array_equal(a, exp["fwd"]), (n, q)
        b = y.copy()
        O.lib().o_ntt_inverse(O.ptr(b), n, q)
        assert np.array_equal(b, exp["rev"]), (n, q)
        for name, op in ops.items():
            o = np.zeros(n, dtype=np.uint64)
            O.lib().o_eltwise(op, O.ptr(o), O.ptr(x), O.ptr(y), 0, n, q)
            assert np.array_equal(o, exp[name]), (name, n, q)


def test_negacyclic_convolution():
    """test_ntt.c:51-104 and test_arith.py:31-53: pointwise product of NTTs == schoolbook mod X^n + 1."""
    for n, bits in ((64, 20), (64, 62), (256, 50)):
        q = int(O.lib().o_next_special_prime(1 << bits, n, 1))
        a = [(i + 123) % q for i in range(n)]
        b = [(2 * i + 456) % q for i in range(n)]
        ref = [0] * n
        for i in range(n):
            for j in range(n):
                k = i + j
                if k < n:
                    ref[k] = (ref[k] + a[i] * b[j]) % q
                else:
                    ref[k - n] = (ref[k - n] - a[i] * b[j]) % q
        na, nb = np.array(a, dtype=np.uint64), np.array(b, dtype=np.uint64)
        O.lib().o_ntt_forward(O.ptr(na), n, q)
        O.lib().o_ntt_forward(O.ptr(nb), n, q)
        nr = np.zeros(n, dtype=np.uint64)
        O.lib().o_eltwise(O.O_MUL, O.ptr(nr), O.ptr(na), O.ptr(nb), 0, n, q)
        O.lib().o_ntt_inverse(O.ptr(nr), n, q)
        assert [int(v) for v in nr] == ref


# --- elementwise: test_mod.c:21-85 oracles with Python big ints, and the reference kernels -------

@pytest.mark.parametrize("bits", [10, 20, 30, 40, 50, 60, 62])
def test_eltwise_against_bigint_and_reference(bits):
    n = 1024
    q = int(O.lib().o_next_special_prime(1 << bits, n, 1))
    i = np.arange(n, dtype=np.uint64)
    in1 = ((np.uint64(0x123456789ABCDEF) ^ (i * np.uint64(0x1337BEEF))) % np.uint64(q)).astype(np.uint64)
    in2 = ((np.uint64(0xFEDCBA987654321) ^ (i * np.uint64(0xDEADBEEF))) % np.uint64(q)).astype(np.uint64)
    s = 0x123456789ABCDEF
    big = {
        O.O_MUL: lambda o, a, b: a * b % q,
        O.O_MUL_ADDTO: lambda o, a, b: (o + a * b) % q,
        O.O_MUL_SUBTO: lambda o, a, b: (o - a * b) % q,
        O.O_SCALE: lambda o, a, b: a * (s % q) % q,
        O.O_FMA: lambda o, a, b: (o + a * (s % q)) % q,
        O.O_ADD: lambda o, a, b: (a + b) % q,
        O.O_SUB: lambda o, a, b: (a - b) % q,
        O.O_NEGATE: lambda o, a, b: (-a) % q,
        O.O_REDUCE: lambda o, a, b: a % q,
        O.O_ADD_SCALAR: lambda o, a, b: (a + s % q) % q,
        O.O_SUB_SCALAR: lambda o, a, b: (a - s % q) % q,
    }
    start = in2.copy()
    for op, f in big.items():
        out = start.copy()
        O.lib().o_eltwise(op, O.ptr(out), O.ptr(in1), O.ptr(in2), s, n, q)
        exp = [f(int(o), int(a), int(b)) for o, a, b in zip(start, in1, in2)]
        assert [int(v) for v in out] == exp, op
    sig = np.array([0, -(2**63), -1, 1] + [(1 if k % 2 else -1) * k * 0x12345678 for k in range(4, n)], dtype=np.int64)
    out = np.zeros(n, dtype=np.uint64)
    O.lib().o_eltwise(O.O_REDUCE_SIGNED, O.ptr(out), O.ptr(sig.view(np.uint64)), O.ptr(in2), 0, n, q)
    assert [int(v) for v in out] == [int(v) % q for v in sig]
    for engine in O.ref_engines():
        if bits < 20:
            continue
        L = O.ref_lib(engine)
        proc = L.ntt_new_proc(n, q)
        for op, name in ((O.O_MUL, "mod_eltwise_mul"), (O.O_ADD, "mod_eltwise_add"), (O.O_SUB, "mod_eltwise_sub")):
            a, b, o1, o2 = O.aligned(n), O.aligned(n), O.aligned(n), O.aligned(n)
            a[:], b[:] = in1, in2
            getattr(L, name)(O.ptr(o1), O.ptr(a), O.ptr(b), n, proc)
            O.lib().o_eltwise(op, O.ptr(o2), O.ptr(a), O.ptr(b), 0, n, q)
            assert np.array_equal(o1, o2), (engine, name, q)
        L.ntt_free_proc(proc)


# --- RNS layer: big-integer CRT closed forms ----------------------------------------------------

def crt(res, mods):
    M = math.prod(mods)
    return sum(r * (M // m) * pow(M // m, -1, m) for r, m in zip(res, mods)) % M


def small_ctx(N=16, L=5, seed=1):
    primes = O.gen_primes(N, [50] * L)
    return primes, O.Ctx(N, primes), np.random.default_rng(seed)


@pytest.mark.parametrize("rnd", [True, False])
def test_division_closed_form(rnd):
    """Dividing by the primes of D in ascending order equals, per coefficient X (CRT value in [0, Q)),
    applying X <- floor((X + h_p) / p) with h_p = floor(p/2) (round) or 0 (floor), one prime at a time."""
    primes, c, rng = small_ctx()
    L, N = len(primes), 16
    for dm in (0b10000, 0b11000, 0b01010, 0b11110):
        x = np.stack([rng.integers(0, q, N, dtype=np.uint64) for q in primes])
        y = x.copy()
        m = c.division(y, (1 << L) - 1, dm, rnd)
        assert m == ((1 << L) - 1) & ~dm
        for k in range(N):
            mods = list(primes)
            X = crt([int(x[i][k]) for i in range(L)], mods)
            for idx in range(L):
                if not (dm >> idx) & 1:
                    continue
                p = primes[idx]
                X = (X + (p // 2 if rnd else 0)) // p
                mods.remove(p)
                X %= math.prod(mods)
            for i in range(L):
                if (m >> i) & 1:
                    assert int(y[i][k]) == X % primes[i]
                else:
                    assert int(y[i][k]) == 0
    # a planted mistake -- the rounding offset h_p dropped (or added for floor) -- fails the check
    x = np.stack([rng.integers(0, q, N, dtype=np.uint64) for q in primes])
    y = x.copy()
    c.division(y, (1 << L) - 1, 0b11000, rnd)
    mism = 0
    for k in range(N):
        X = crt([int(x[i][k]) for i in range(L)], primes)
        for p in (primes[3], primes[4]):
            X = (X + (0 if rnd else p // 2)) // p
        X %= math.prod(primes[:3])
        mism += any(int(y[i][k]) != X % primes[i] for i in range(3))
    assert mism > 0


@pytest.mark.parametrize("bits", [20, 31, 40, 49, 50, 58, 62])
def test_reduce_mp_against_bigint_and_reference(bits):
    """o_reduce_mp (mod_reduce_array_mp, mod.c:261-276): (hi*2^64 + lo) mod q for every 128-bit input.
    Checked against Python big ints; against the reference's portable mod_reduce_array_mp
    (mod_portable.c:153-161, exact for every input) on all inputs; and against both builds of oracle/_ref
    on the reference test's own inputs (test_mod.c:167-176).  The avx512ifma build's kernel for q < 2^32
    (mod32.c:366-411) forms (hi mod q)*(2^64 mod q) + lo in 64 bits, which wraps once lo + (q-1)^2
    reaches 2^64; the reference's test inputs stay below that, and the oracle follows the exact
    portable definition."""
    n = 512
    q = int(O.lib().o_next_special_prime(1 << bits, n, 1))
    rng = np.random.default_rng(bits)
    hi = rng.integers(0, 2**63, n, dtype=np.uint64) * np.uint64(2) + rng.integers(0, 2, n, dtype=np.uint64)
    lo = rng.integers(0, 2**63, n, dtype=np.uint64) * np.uint64(2) + rng.integers(0, 2, n, dtype=np.uint64)
    # edge values: 0, 2^64 - 1 halves, multiples of q, q - 1, the largest 128-bit value
    edges = [(0, 0), (0, 1), (0, q - 1), (0, q), (1, 0), (2**64 - 1, 2**64 - 1), (q - 1, 2**64 - 1),
             (0, 2**64 - 1), (2**63, 0), ((q << 1) >> 64, (q << 1) & (2**64 - 1))]
    for k, (h, l_) in enumerate(edges):
        hi[k], lo[k] = h, l_

    def check(hi, lo, engines):
        out = np.zeros(n, dtype=np.uint64)
        O.lib().o_reduce_mp(O.ptr(out), O.ptr(hi), O.ptr(lo), n, q)
        exp = [((int(h) << 64) | int(l_)) % q for h, l_ in zip(hi, lo)]
        assert [int(v) for v in out] == exp
        for engine in engines:
            L = O.ref_lib(engine)
            proc = L.ntt_new_proc(n, q)
            o, h, l_ = O.aligned(n), O.aligned(n), O.aligned(n)
            h[:], l_[:] = hi, lo
            L.mod_reduce_array_mp(O.ptr(o), O.ptr(h), O.ptr(l_), n, proc)
            L.ntt_free_proc(proc)
            assert np.array_equal(o, out), (engine, q)
        return exp

    exp = check(hi, lo, [e for e in O.ref_engines() if e == "portable" or q >= 2**32])
    i = np.arange(n, dtype=np.uint64)
    check(np.uint64(0xAAAAAAAAAAAAAAAA) ^ (i * np.uint64(0x11111111)),
          np.uint64(0x5555555555555555) ^ (i * np.uint64(0x22222222)), O.ref_engines())
    # a planted mistake -- the high word dropped -- fails the big-int check
    bad = (lo % np.uint64(q)).astype(np.uint64)
    assert [int(v) for v in bad] != exp


@pytest.mark.parametrize("split", [1, 4])
def test_add_integer_closed_forms(split):
    """o_add_integer (rns_polynomial.c:474-514).  Coefficient form: the CRT value of coefficient 0
    moves by v (signed, reduced mod Q) and nothing else changes.  NTT form: adding v to every slot of
    the first block is what the NTT makes of adding v to coefficient 0 (X^0 evaluates to 1 at every
    root), for positive and negative v."""
    N = 32
    primes = O.gen_primes(N, [50] * 4, split)
    c = O.Ctx(N, primes, split)
    L = len(primes)
    Q = math.prod(primes)
    rng = np.random.default_rng(split)
    for m in (0b1111, 0b1011):
        act = [i for i in range(L) if (m >> i) & 1]
        Qm = math.prod(primes[i] for i in act)
        for v in (0, 1, 7, 2**63 - 1, -1, -5, -(2**62), -primes[0], -(2**63), primes[1] * 3):
            x = np.stack([rng.integers(0, q, N, dtype=np.uint64) for q in primes])
            y = np.zeros_like(x)
            assert c.add_integer(y, x, m, v, False) == m
            X0 = crt([int(x[i][0]) for i in act], [primes[i] for i in act])
            Y0 = crt([int(y[i][0]) for i in act], [primes[i] for i in act])
            assert Y0 == (X0 + v) % Qm, v
            for i in act:
                assert np.array_equal(y[i][1:], x[i][1:])
            # NTT form: NTT(x + v) == NTT(x) (+) v on the first block
            xn, yn = x.copy(), y.copy()
            c.to_ntt(xn, xn, m)
            c.to_ntt(yn, yn, m)
            zn = np.zeros_like(xn)
            c.add_integer(zn, xn, m, v, True)
            n_blk = N // split
            for i in act:
                assert np.array_equal(zn[i][:n_blk], yn[i][:n_blk]), (v, i)
                assert np.array_equal(zn[i][n_blk:], xn[i][n_blk:]), (v, i)
    # a planted mistake -- a negative v taken as its unsigned 64-bit value -- fails the CRT check
    x = np.stack([rng.integers(0, q, N, dtype=np.uint64) for q in primes])
    y = np.zeros_like(x)
    c.add_integer(y, x, 0b1111, 2**64 - 5, False)   # what a sign-blind oracle would compute for v = -5
    Y0 = crt([int(y[i][0]) for i in range(L)], primes)
    X0 = crt([int(x[i][0]) for i in range(L)], primes)
    assert Y0 != (X0 + 2**64 - 5) % Q


def test_mod_reduce_lifted_closed_form():
    primes, c, rng = small_ctx()
    x = np.stack([rng.integers(0, q, 16, dtype=np.uint64) for q in primes])
    out = np.zeros_like(x)
    c.mod_reduce_lifted(out, 0b11011, x, 2)
    for i, q in enumerate(primes):
        exp = [int(v) % q for v in x[2]] if (0b11011 >> i) & 1 else [0] * 16
        assert [int(v) for v in out[i]] == exp


def test_base_conversion_closed_form():
    """out_p = sum_j [x_j Dhat_j]_{q_j} (D/q_j) mod p, i.e. X + u*D for some 0 <= u < w (no overflow correction)."""
    primes, c, rng = small_ctx()
    in_mask, out_mask = 0b00111, 0b11111
    x = np.zeros((5, 16), dtype=np.uint64)
    for i in range(3):
        x[i] = rng.integers(0, primes[i], 16, dtype=np.uint64)
    out = np.zeros_like(x)
    c.base_conversion(out, out_mask, x, in_mask)
    D = primes[:3]
    Dp = math.prod(D)
    for k in range(16):
        X = crt([int(x[i][k]) for i in range(3)], D)
        y = sum(int(x[j][k]) * pow(Dp // D[j], -1, D[j]) % D[j] * (Dp // D[j]) for j in range(3))
        assert (y - X) % Dp == 0 and 0 <= (y - X) // Dp < 3
        for t in (3, 4):
            assert int(out[t][k]) == y % primes[t]
        for i in range(3):
            assert out[i][k] == x[i][k]


def test_scaled_lift_closed_form():
    primes, c, rng = small_ctx()
    in_mask, out_mask = 0b00111, 0b11111
    x = np.stack([rng.integers(0, q, 16, dtype=np.uint64) for q in primes])
    out = np.zeros_like(x)
    c.scaled_lift(out, out_mask, x, in_mask)
    delta = primes[3] * primes[4]
    for i in range(5):
        for k in range(16):
            exp = int(x[i][k]) * delta % primes[i] if i < 3 else 0
            assert int(out[i][k]) == exp


@pytest.mark.parametrize("split", [1, 4])
def test_permute_closed_form(split):
    """X -> X^gen on the integer polynomial (test_arith.py:83-88 composition, plus the direct map)."""
    N = 16
    primes = O.gen_primes(N, [50] * 2, split)
    c = O.Ctx(N, primes, split)
    n = N // split
    rng = np.random.default_rng(9)
    coeffs = [[R.randrange(q) for _ in range(N)] for q in primes]
    x = np.zeros((2, N), dtype=np.uint64)
    for i in range(2):
        for j in range(N):
            x[i][(j % split) * n + j // split] = coeffs[i][j]
    out = np.zeros_like(x)
    back = np.zeros_like(x)
    gen = 3
    c.permute(out, x, 0b11, gen)
    c.permute(back, out, 0b11, 11)     # 3 * 11 = 33 = 1 mod 2N
    assert np.array_equal(back, x)
    for i, q in enumerate(primes):
        exp = [0] * N
        for j in range(N):
            e = j * gen % (2 * N)
            exp[e % N] = (coeffs[i][j] if e < N else -coeffs[i][j]) % q
        got = [int(out[i][(j % split) * n + j // split]) for j in range(N)]
        assert got == exp
    del rng


def test_split_product_is_ring_product():
    """With split degree S, NTT-domain block products with the w wrap equal the product in Z_q[X]/(X^N+1)."""
    N, S = 16, 4
    primes = O.gen_primes(N, [50] * 2, S)
    c = O.Ctx(N, primes, S)
    n = N // S
    a = [[R.randrange(q) for _ in range(N)] for q in primes]
    b = [[R.randrange(q) for _ in range(N)] for q in primes]

    def place(v):
        out = np.zeros((2, N), dtype=np.uint64)
        for i in range(2):
            for j in range(N):
                out[i][(j % S) * n + j // S] = v[i][j]
        return out

    pa, pb = place(a), place(b)
    c.to_ntt(pa, pa, 0b11)
    c.to_ntt(pb, pb, 0b11)
    pr = np.zeros_like(pa)
    c.mul(pr, pa, 0b11, pb, 0b11, 0)
    c.to_coeff(pr, pr, 0b11)
    for i, q in enumerate(primes):
        ref = [0] * N
        for x in range(N):
            for y in range(N):
                k = x + y
                if k < N:
                    ref[k] = (ref[k] + a[i][x] * b[i][y]) % q
                else:
                    ref[k - N] = (ref[k - N] - a[i][x] * b[i][y]) % q
        got = [int(pr[i][(j % S) * n + j // S]) for j in range(N)]
        assert got == ref


def test_incntt_w_table():
    """incNTT.w[k] = psi^(2 brv(k) + 1) (rns_polynomial.c:21-27), i.e. the slot value of X."""
    N = 64
    primes = O.gen_primes(N, [50] * 2)
    c = O.Ctx(N, primes)
    for i, q in enumerate(primes):
        psi = c.psi(i)
        w = c.w(i)
        x = np.zeros(N, dtype=np.uint64)
        x[1] = 1  # the polynomial X
        O.lib().o_ntt_forward(O.ptr(x), N, q)
        assert np.array_equal(x, w)
        assert all(int(w[k]) == pow(psi, 2 * brv(k, 6) + 1, q) for k in range(N))


# --- MLWE: tensor slot identity and decryption with valid keys -----------------------------------

def test_tensor_product_slot_identity():
    """test_mlwe.py:183-207: b1b2 - sum_slots slot*key == phase(c1)*phase(c2) (r = 2, NTT domain)."""
    N, r = 16, 2
    primes = O.gen_primes(N, [50] * 2)
    c = O.Ctx(N, primes)
    rng = np.random.default_rng(4)
    m = 0b11

    def rnd():
        return np.stack([rng.integers(0, q, N, dtype=np.uint64) for q in primes])

    s = [rnd() for _ in range(r)]
    c1 = np.stack([rnd() for _ in range(r + 1)])
    c2 = np.stack([rnd() for _ in range(r + 1)])
    R = int(O.lib().o_extended_rank(r))
    out = np.zeros((R + 1, 2, N), dtype=np.uint64)
    c.tensor_product(out, c1, c2, r, m)
    tmp = np.zeros((2, N), dtype=np.uint64)

    def phase(ct):
        acc = ct[r].copy()
        for i in range(r):
            c.mul(tmp, ct[i], m, s[i], m, 0)
            c.addsub(acc, acc, m, tmp, m, True)
        return acc

    exp = np.zeros((2, N), dtype=np.uint64)
    c.mul(exp, phase(c1), m, phase(c2), m, 0)
    keys = []
    for i in range(r):
        for j in range(i, r):
            kk = np.zeros((2, N), dtype=np.uint64)
            c.mul(kk, s[i], m, s[j], m, 0)
            c.negate(kk, kk, m)
            keys.append(kk)
    keys += s
    acc = out[R].copy()
    for k in range(R):
        c.mul(tmp, out[k], m, keys[k], m, 0)
        c.addsub(acc, acc, m, tmp, m, True)
    assert np.array_equal(acc, exp)


# --- key-switch, relinearization and rotation with valid keys ---------------------------------------
# Pinned by decryption: for a key-switching key of s' under s (mlwe.py:156-176 layout, digit j =
# an encryption of P * s' placed on limb j only), phase_s(KS(c)) == b - c_from * s' up to the
# key-switch noise; a wrong sign, a dropped digit or a wrong digit index leaves noise of size Q.

KS_N, KS_LVL, KS_SPECIAL = 32, 3, 1


def _ks_setup(seed):
    primes = O.gen_primes(KS_N, [50] * (KS_LVL + KS_SPECIAL))
    c = O.Ctx(KS_N, primes)
    lvl, ext = (1 << KS_LVL) - 1, (1 << (KS_LVL + KS_SPECIAL)) - 1
    rng = np.random.default_rng(seed)
    return c, primes, lvl, ext, rng


def _small(c, primes, rng, bound):
    v = rng.integers(-bound, bound + 1, KS_N)
    return np.stack([np.array([int(x) % q for x in v], dtype=np.uint64) for q in primes]), v


def _ntt(c, a):
    out = a.copy()
    c.to_ntt(out, a, c.mask_all())
    return out


def _gen_ksk(c, primes, rng, s_ntt, s_from_ntt, lvl):
    """digit j (one per level limb): a uniform, b = a*s + e + (P mod q_j) * s' on limb j only."""
    P = math.prod(primes[KS_LVL:])
    m = c.mask_all()
    digits = []
    for j in range(KS_LVL):
        a = np.stack([rng.integers(0, q, KS_N, dtype=np.uint64) for q in primes])
        e, _ = _small(c, primes, rng, 3)
        b = np.zeros_like(a)
        c.mul(b, a, m, s_ntt, m, 0)
        c.addsub(b, b, m, _ntt(c, e), m)
        t = np.zeros_like(a)
        t[j] = s_from_ntt[j]
        c.scale(t, t, 1 << j, P % primes[j])
        c.addsub(b, b, m, t, m)
        digits.append(np.stack([a, b]))
    return np.ascontiguousarray(np.stack(digits))


def _phase_coeff(c, ct_ntt, s_ntt, mask):
    """b - a*s in NTT form, returned in coefficient form."""
    tmp = np.zeros_like(ct_ntt[0])
    c.mul(tmp, ct_ntt[0], mask, s_ntt, mask, 0)
    ph = ct_ntt[1].copy()
    c.addsub(ph, ph, mask, tmp, mask, True)
    out = np.zeros_like(ph)
    c.to_coeff(out, ph, mask)
    return out


def _centered_max(c, primes, rows, mask):
    mods = [q for i, q in enumerate(primes) if (mask >> i) & 1]
    Q = math.prod(mods)
    worst = 0
    for k in range(KS_N):
        v = crt([int(rows[i][k]) for i in range(len(primes)) if (mask >> i) & 1], mods)
        worst = max(worst, min(v, Q - v))
    return worst


def _keyswitch_error(c, primes, lvl, ext, rng, s, s_from, ksk):
    """|phase_s(KS(c_from, b)) - (b - c_from * s')| for a random coefficient-form input."""
    cin = np.stack([np.stack([rng.integers(0, q, KS_N, dtype=np.uint64) if i < KS_LVL else np.zeros(KS_N, np.uint64)
                              for i, q in enumerate(primes)]) for _ in range(2)])
    out = np.zeros((2, len(primes), KS_N), dtype=np.uint64)
    mo = c.ghs_keyswitch(out, 1, cin, 1, lvl, [ksk], ext)
    assert mo == lvl
    s_ntt = _ntt(c, s)
    ph = _phase_coeff(c, np.stack([_ntt(c, out[0]), _ntt(c, out[1])]), s_ntt, lvl)
    exp_ntt = np.zeros_like(s)
    c.mul(exp_ntt, _ntt(c, cin[0]), lvl, _ntt(c, s_from), lvl, 0)
    bb = _ntt(c, cin[1])
    c.addsub(exp_ntt, bb, lvl, exp_ntt, lvl, True)
    exp = np.zeros_like(exp_ntt)
    c.to_coeff(exp, exp_ntt, lvl)
    diff = np.zeros_like(exp)
    c.addsub(diff, ph, lvl, exp, lvl, True)
    return _centered_max(c, primes, diff, lvl)


def test_keyswitch_decrypts_with_valid_key():
    """mlwe.c:557-598 / gadget.c:22-37 restated: phase preserved up to |e| < 2^16 (Q ~ 2^150)."""
    c, primes, lvl, ext, rng = _ks_setup(11)
    s, _ = _small(c, primes, rng, 1)
    s_from, _ = _small(c, primes, rng, 1)
    ksk = _gen_ksk(c, primes, rng, _ntt(c, s), _ntt(c, s_from), lvl)
    assert _keyswitch_error(c, primes, lvl, ext, rng, s, s_from, ksk) < 2**16
    # planted mistakes: digits in the wrong order, and a key for the wrong source secret
    assert _keyswitch_error(c, primes, lvl, ext, rng, s, s_from, np.ascontiguousarray(ksk[::-1])) > 2**100
    other, _ = _small(c, primes, rng, 1)
    assert _keyswitch_error(c, primes, lvl, ext, rng, s, other, ksk) > 2**100


def test_multiply_relinearizes_with_valid_key():
    """mlwe.c:711-760: phase_s(multiply(c1, c2)) == m1 * m2 (negacyclic) up to key-switch noise."""
    c, primes, lvl, ext, rng = _ks_setup(12)
    s, sv = _small(c, primes, rng, 1)
    s_ntt = _ntt(c, s)
    neg_s2 = np.zeros_like(s)
    c.mul(neg_s2, s_ntt, ext, s_ntt, ext, 0)
    c.negate(neg_s2, neg_s2, ext)          # quadratic slot key -s*s (tensor slot convention)
    ksk = _gen_ksk(c, primes, rng, s_ntt, neg_s2, lvl)
    msgs, cts = [], []
    for _ in range(2):
        m, mv = _small(c, primes, rng, 1000)
        a = np.stack([rng.integers(0, q, KS_N, dtype=np.uint64) if i < KS_LVL else np.zeros(KS_N, np.uint64)
                      for i, q in enumerate(primes)])
        b = np.zeros_like(a)
        c.mul(b, a, lvl, s_ntt, lvl, 0)
        c.addsub(b, b, lvl, _ntt(c, m), lvl)
        msgs.append(mv)
        cts.append(np.stack([a, b]))
    out = np.zeros((2, len(primes), KS_N), dtype=np.uint64)
    mo = c.multiply(out, cts[0], cts[1], 1, lvl, [ksk, None], ext)
    assert mo == lvl
    ph = _phase_coeff(c, out, s_ntt, lvl)
    prod = [0] * KS_N
    for i in range(KS_N):
        for j in range(KS_N):
            k, v = i + j, int(msgs[0][i]) * int(msgs[1][j])
            if k < KS_N:
                prod[k] += v
            else:
                prod[k - KS_N] -= v
    exp = np.stack([np.array([v % q for v in prod], dtype=np.uint64) for q in primes])
    diff = np.zeros_like(exp)
    c.addsub(diff, ph, lvl, exp, lvl, True)
    assert _centered_max(c, primes, diff, lvl) < 2**16


def test_automorphism_decrypts_with_galois_key():
    """mlwe.c:314-326: phase_s(automorphism(c, g)) == sigma_g(phase_s(c)) with a key for sigma_g(s)."""
    c, primes, lvl, ext, rng = _ks_setup(13)
    g = 5
    s, _ = _small(c, primes, rng, 1)
    s_ntt = _ntt(c, s)
    sg = np.zeros_like(s)
    c.permute(sg, s, ext, g)
    ksk = _gen_ksk(c, primes, rng, s_ntt, _ntt(c, sg), lvl)
    m, _ = _small(c, primes, rng, 1000)
    a = np.stack([rng.integers(0, q, KS_N, dtype=np.uint64) if i < KS_LVL else np.zeros(KS_N, np.uint64)
                  for i, q in enumerate(primes)])
    b = np.zeros_like(a)
    c.mul(b, a, lvl, s_ntt, lvl, 0)
    c.addsub(b, b, lvl, _ntt(c, m), lvl)
    ct = np.stack([np.zeros_like(a), np.zeros_like(a)])
    c.to_coeff(ct[0], a, lvl)
    c.to_coeff(ct[1], b, lvl)
    out = np.zeros((2, len(primes), KS_N), dtype=np.uint64)
    assert c.automorphism(out, ct, 1, lvl, g, [ksk], ext) == lvl
    ph = _phase_coeff(c, np.stack([_ntt(c, out[0]), _ntt(c, out[1])]), s_ntt, lvl)
    exp = np.zeros_like(m)
    c.permute(exp, m, lvl, g)
    diff = np.zeros_like(exp)
    c.addsub(diff, ph, lvl, exp, lvl, True)
    assert _centered_max(c, primes, diff, lvl) < 2**16


# --- monomial products and MGSW ----------------------------------------------------------------

def test_mul_by_xai_closed_form():
    """rns_polynomial.c:1048-1168: X^a * x and (X^a - 1) * x in Z_q[X]/(X^N + 1), every a mod 2N."""
    N = 16
    primes = O.gen_primes(N, [50, 40])
    c = O.Ctx(N, primes)
    rng = np.random.default_rng(21)
    x = np.stack([rng.integers(0, q, N, dtype=np.uint64) for q in primes])
    for a in (0, 1, 5, 8, N - 1, N, N + 3, 2 * N - 1, 2 * N + 2):
        for minus1 in (False, True):
            out = np.zeros_like(x)
            c.mul_by_xai(out, x, 0b11, a, minus1)
            for i, q in enumerate(primes):
                exp = [0] * N
                for k in range(N):
                    e = (k + a) % (2 * N)
                    if e < N:
                        exp[e] += int(x[i][k])
                    else:
                        exp[e - N] -= int(x[i][k])
                if minus1:
                    exp = [v - int(x[i][k]) for k, v in enumerate(exp)]
                assert [int(v) for v in out[i]] == [v % q for v in exp], (a, minus1)


def _gen_mgsw(c, primes, rng, s_ntt, bit, lvl):
    """MGSW of `bit` under s (rank 1, ell = level limbs, mgsw.c layout): entry (j, k) is a sample
    whose phase is bit * P * [limb k] * (-s for j = 0 (the a component), 1 for j = 1 (b))."""
    P = math.prod(primes[KS_LVL:])
    m = c.mask_all()
    neg_s = np.zeros_like(s_ntt)
    c.negate(neg_s, s_ntt, m)
    one = np.stack([np.ones(KS_N, dtype=np.uint64) for _ in primes])   # NTT of the constant 1
    ents = []
    for factor in (neg_s, one):
        for k in range(KS_LVL):
            a = np.stack([rng.integers(0, q, KS_N, dtype=np.uint64) for q in primes])
            e, _ = _small(c, primes, rng, 3)
            b = np.zeros_like(a)
            c.mul(b, a, m, s_ntt, m, 0)
            c.addsub(b, b, m, _ntt(c, e), m)
            if bit:
                t = np.zeros_like(a)
                t[k] = factor[k]
                c.scale(t, t, 1 << k, P % primes[k])
                c.addsub(b, b, m, t, m)
            ents.append(np.stack([a, b]))
    return np.ascontiguousarray(np.stack(ents))


def test_cmux_selects_with_valid_mgsw():
    """mgsw.c:51-71 restated: phase_s(CMUX(c0, c1, MGSW(bit))) == phase_s(c_bit) up to noise."""
    c, primes, lvl, ext, rng = _ks_setup(14)
    s, _ = _small(c, primes, rng, 1)
    s_ntt = _ntt(c, s)
    cts, msgs = [], []
    for _ in range(2):
        m, _ = _small(c, primes, rng, 1000)
        a = np.stack([rng.integers(0, q, KS_N, dtype=np.uint64) if i < KS_LVL else np.zeros(KS_N, np.uint64)
                      for i, q in enumerate(primes)])
        b = np.zeros_like(a)
        c.mul(b, a, lvl, s_ntt, lvl, 0)
        c.addsub(b, b, lvl, _ntt(c, m), lvl)
        ct = np.stack([np.zeros_like(a), np.zeros_like(a)])
        c.to_coeff(ct[0], a, lvl)
        c.to_coeff(ct[1], b, lvl)
        cts.append(ct)
        msgs.append(m)
    for bit in (0, 1):
        mg = _gen_mgsw(c, primes, rng, s_ntt, bit, lvl)
        out = np.zeros((2, len(primes), KS_N), dtype=np.uint64)
        mo = c.cmux(out, 1, cts[0], cts[1], lvl, mg, KS_LVL, KS_SPECIAL, ext)
        assert mo == lvl
        ph = _phase_coeff(c, out, s_ntt, lvl)
        diff = np.zeros_like(ph)
        c.addsub(diff, ph, lvl, msgs[bit], lvl, True)
        assert _centered_max(c, primes, diff, lvl) < 2**8, bit
        # planted mistake: the other message is far away
        c.addsub(diff, ph, lvl, msgs[1 - bit], lvl, True)
        assert _centered_max(c, primes, diff, lvl) > 2**8


@pytest.mark.skipif(not O.ref_engines(), reason="oracle/_ref not built")
@pytest.mark.parametrize("N,L,K", [(4096, 4, 1), (8192, 5, 2)])
def test_multiply_rescale_matches_reference_kernels(N, L, K):
    """The oracle's mlwe_multiply + CKKS rescale (its restatement of mlwe.c:557-760, gadget.c:22-37
    and rns_polynomial.c:939-976) equals the same sequence composed from the reference engine's own
    kernels (oracle/ref_ckks.c over oracle/_ref), bit for bit, for every engine built here; the
    oracle fed its key digits in reverse order (a planted indexing slip) does not."""
    primes = O.gen_primes(N, [50] * (L + K))
    c = O.Ctx(N, primes)
    rng = np.random.default_rng(N + L)
    lvl, ext = (1 << L) - 1, (1 << (L + K)) - 1

    def rows(mask):
        out = np.zeros((L + K, N), dtype=np.uint64)
        for i, q in enumerate(primes):
            if (mask >> i) & 1:
                out[i] = rng.integers(0, q, N, dtype=np.uint64)
        return out

    kdata = np.stack([np.stack([rows(ext), rows(ext)]) for _ in range(L)])
    pairs = [(np.stack([rows(lvl), rows(lvl)]), np.stack([rows(lvl), rows(lvl)])) for _ in range(2)]
    top = 1 << (L - 1)
    expect = []
    for d1, d2 in pairs:
        ref = np.zeros((2, L + K, N), dtype=np.uint64)
        m = c.multiply(ref, d1, d2, 1, lvl, [kdata, None], ext)
        for comp in range(2):
            c.to_coeff(ref[comp], ref[comp], m)
        m2 = c.mlwe_round_division(ref, 1, m, top)
        assert m2 == lvl & ~top
        expect.append(ref[:, : L - 1])
    ct1 = np.stack([d1[:, :L] for d1, _ in pairs])
    ct2 = np.stack([d2[:, :L] for _, d2 in pairs])
    for eng in O.ref_engines():
        got = O.ref_ckks_multiply(eng, N, L, K, primes, ct1, ct2, kdata, threads=2)
        for k in range(len(pairs)):
            assert np.array_equal(got[k], expect[k]), (eng, k)
    d1, d2 = pairs[0]
    ref = np.zeros((2, L + K, N), dtype=np.uint64)
    m = c.multiply(ref, d1, d2, 1, lvl, [kdata[::-1].copy(), None], ext)
    for comp in range(2):
        c.to_coeff(ref[comp], ref[comp], m)
    c.mlwe_round_division(ref, 1, m, top)
    assert not np.array_equal(got[0], ref[:, : L - 1])


def test_slot_restatements_pinned():
    """Slot operations of the oracle against what fixes them: rotation by 1 is the index shift
    k -> k+1 within each block and rotating by r then ps - r is the identity; a broadcast equals
    copying the slot into every position; the slot-wise inverse times the input is 1 mod q, and a
    zero slot reports -2.  Inactive limbs stay zero."""
    q = [97, 193]
    split, ps = 2, 8
    rows = np.zeros((3, split * ps), dtype=np.uint64)
    rows[0] = np.arange(split * ps) % 97 + 1
    rows[1] = (np.arange(split * ps) * 5 + 3) % 193 + 1
    mask = 0b011
    rot1 = O.slot_rotate(rows, mask, split, 1)
    for j in range(split):
        for k in range(ps):
            assert rot1[0, j * ps + k] == rows[0, j * ps + (k + 1) % ps]
    assert np.array_equal(O.slot_rotate(O.slot_rotate(rows, mask, split, 3), mask, split, ps - 3), rows)
    b = O.slot_broadcast(rows, mask, split, 5)
    c = np.zeros_like(rows)
    for dst in range(ps):
        c = O.slot_copy(c, dst, rows, mask, split, 5)
    assert np.array_equal(b, c)
    assert not b[2].any() and not rot1[2].any()
    st, inv = O.slot_inverse(rows, mask, q + [1])
    assert st == 0
    for i in range(2):
        assert all((int(a) * int(x)) % q[i] == 1 for a, x in zip(rows[i], inv[i]))
    rows[1, 3] = 0
    assert O.slot_inverse(rows, mask, q + [1])[0] == -2
