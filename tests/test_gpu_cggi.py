"""CGGI16 functional bootstrap on the hip engine (vfhe_amd.cggi).

Ports modules/fhe/python/test/test_cggi16.py:20-62 (same ring, key weights, LUT
and message precision): bootstrapping key (MGSW encryptions of the LWE key),
LUT packing, blind rotation and sample extraction, checked by decrypting the
extracted LWE sample and reading the lookup table's entry for the message.
The batched blind rotation (vfhe_hip_blind_rotate_batch) is checked bit for bit
against the reference's one-at-a-time sequence made of the reference-named
entry points (mlwe_RNSc_mul_by_xai, mgsw_CMUX, mlwe_RNS_to_RNSc), whose pieces
are pinned against the oracle in test_gpu_mgsw.py.  The last test runs the
benchmark's bootstrap shape: LWE dimension 630, N = 1024, 1024 ciphertexts of
4-bit messages, identity table, every output decrypted exactly.
"""
import ctypes
import random

import numpy as np
import pytest

from vfhe_amd import _native, arith
from vfhe_amd.arith import Ring
from vfhe_amd.cggi import BootstrapKey, Bootstrapper, LweBatch, LweKey, extract, lwe_key_of, pack_lut, rotation_amounts
from vfhe_amd.cggi import _LweStruct as LweStruct
from vfhe_amd.fhe import Batch, Chain, SecretKey, _handles

pytestmark = pytest.mark.gpu

MSG_PREC = 5


def mod_switch(v, q, p):
    return round((v * p) / q) % p


def binary_key(rng, n, h):
    s = [0] * n
    for k in rng.choice(n, size=h, replace=False):
        s[int(k)] = 1
    return s


@pytest.fixture(scope="module")
def setup():
    arith.NTT_processor_instance.reset()
    lib = _native.load()
    lib.vfhe_prng_set_deterministic_seed(0xB007C0DE)
    rng = np.random.default_rng(0xB007C0DE)
    rnd = random.Random(0xB007C0DE)
    Rq = Ring(256, prime_size=[50, 50, 50], split_degree=1)
    in_ring = Rq.quotient_ring(ell=2)
    lwe_key = LweKey(in_ring, binary_key(rng, 256, 17))           # key_gen_sparse(17, ternary=False)
    chain = Chain.nested(Rq, special=1, levels=1)
    out_key = SecretKey(chain, [binary_key(rng, 256, 64)], 3.2)    # key_gen_sparse(64, ternary=False)
    print("[cggi] bootstrap key: 256 MGSW encryptions, N=256", flush=True)
    bk = BootstrapKey(chain, lwe_key, out_key)
    lut_size = 1 << (MSG_PREC - 1)
    lut = [rnd.randint(0, (1 << (MSG_PREC - 1)) - 1) for _ in range(lut_size)]
    tv = pack_lut(chain, lut, lut_size, MSG_PREC)
    return dict(in_ring=in_ring, lwe_key=lwe_key, chain=chain, out_key=out_key, bk=bk, lut=lut, tv=tv,
                boot=Bootstrapper(chain, bk), out_lwe=lwe_key_of(out_key, in_ring))


def encrypt_msgs(s, msgs):
    q = s["in_ring"].q_l
    return LweBatch.encrypt(s["lwe_key"], [mod_switch(v, 1 << MSG_PREC, q) for v in msgs])


def test_lwe_phase_and_extraction(setup):
    """lwe.c:97-158: phase(sample(m)) == m under an error-free key; mlwe.c:40-66: the extracted
    sample's phase is coefficient idx of the MLWE phase."""
    s = setup
    q = s["in_ring"].q_l
    msgs = [mod_switch(v, 1 << MSG_PREC, q) for v in (0, 5, 15)]
    assert LweBatch.encrypt(s["lwe_key"], msgs).phases(s["lwe_key"]) == [m % q for m in msgs]
    chain, key = s["chain"], s["out_key"]
    msg = arith.Polynomial(chain.levels[0]).from_array([(7 * i) % 97 for i in range(256)])
    ct = Batch.encrypt(chain, key, [msg])
    coeffs = ct.decrypt(key)[0].get_polynomial()
    for idx in (0, 1, 100, 255):
        assert extract(ct, idx).phases(s["out_lwe"])[0] == coeffs[idx]
    # the batched extraction (one read-back) equals mlwe_extract_LWE sample by sample
    cts = Batch.encrypt(chain, key, [msg, msg, msg])
    cts.to_coeff()
    got = extract(cts, 7)
    lib = _native.load()
    for k, h in enumerate(cts.handles):
        one = LweBatch(got.ring, [lib.mlwe_extract_LWE(h, 7)], got.n, got.l)
        a1, b1 = got.arrays(k)
        a2, b2 = one.arrays(0)
        assert np.array_equal(a1, a2) and np.array_equal(b1, b2), k


def test_functional_bootstrap(setup):
    """test_cggi16.py:20-62: each bootstrapped sample decrypts to lut[msg]."""
    s = setup
    msgs = [0, 3, 7, 12, 15]
    print(f"[cggi] bootstrap of {len(msgs)} messages", flush=True)
    out = s["boot"].bootstrap(s["tv"], encrypt_msgs(s, msgs), torus_base=1 << (MSG_PREC - 1))
    q = s["in_ring"].q_l
    for v, ph in zip(msgs, out.phases(s["out_lwe"])):
        assert mod_switch(ph, q, 1 << MSG_PREC) == s["lut"][v], v


def _reference_sequence(s, lwes, k, torus_base):
    """The reference's functional_bootstrap_wo_extract (cggi16.py:46-90) for input k, one
    reference-named engine call at a time."""
    lib = _native.load()
    chain, bk, tv = s["chain"], s["bk"], s["tv"]
    rot, steps = rotation_amounts(*lwes.arrays(k), lwes.ring.primes[: lwes.l], chain.N, torus_base)
    tv.to_coeff()
    out = Batch.alloc(chain, 0, 1, form="coeff")
    lib.mlwe_RNSc_mul_by_xai(out.handles[0], tv.handles[0], rot)
    for i, a in enumerate(steps):
        if a == 0:
            continue
        out.to_coeff()
        rotated = Batch.alloc(chain, 0, 1, form="coeff")
        lib.mlwe_RNSc_mul_by_xai(rotated.handles[0], out.handles[0], a)
        res = Batch.alloc(chain, 0, 1, form="ntt")
        lib.mgsw_CMUX(res.handles[0], out.handles[0], rotated.handles[0], bk.arrays[i], bk.ell, chain.special)
        lib.mlwe_copy_RNS_sample(out.handles[0], res.handles[0])
        out.form = "ntt"
    return out.to_coeff()


def test_blind_rotate_batch_matches_reference_sequence(setup):
    s = setup
    tb = 1 << (MSG_PREC - 1)
    lwes = encrypt_msgs(s, [1, 9, 14])
    print("[cggi] batched blind rotation of 3 inputs against the one-call-at-a-time sequence", flush=True)
    got = s["boot"].blind_rotate(s["tv"], lwes, tb)
    for k in range(len(lwes)):
        ref = _reference_sequence(s, lwes, k, tb)
        for p, r in zip(got.polys(k), ref.polys(0)):
            assert p == r, k


def test_rotation_amounts_engine_helper(setup):
    """vfhe_hip_cggi_rotations (cggi16.py:54-86 in 128-bit integers) equals the exact-rational
    restatement on random samples, and marks coefficients whose quotient sits next to a
    rounding half (set here by hand: a_0 = floor((2k+1) q / 4N)), which Bootstrapper.amounts then
    takes through the restatement."""
    s = setup
    tb = 1 << (MSG_PREC - 1)
    lwes = encrypt_msgs(s, [0, 1, 2, 7, 8, 15, 3, 11])
    primes = lwes.ring.primes[: lwes.l]
    N, q = s["chain"].N, s["in_ring"].q_l
    st = ctypes.cast(lwes.handles[0], ctypes.POINTER(LweStruct)).contents
    for k, i in ((3, 0), (200, 5)):
        v = ((2 * k + 1) * q) // (4 * N)
        for j, p in enumerate(primes):
            st.a[j][i] = v % p
    lib = _native.load()
    n = lwes.n
    rots = np.zeros(len(lwes), dtype=np.uint64)
    steps = np.zeros(len(lwes) * n, dtype=np.uint64)
    marked = lib.vfhe_hip_cggi_rotations(_handles(lwes.handles), len(lwes), N, tb,
                                         rots.ctypes.data_as(_native.U64P), steps.ctypes.data_as(_native.U64P))
    assert marked == 2
    assert steps[0] == steps[5] == np.uint64(2**64 - 1)
    rots, steps = s["boot"].amounts(lwes, tb)
    for k in range(len(lwes)):
        rot, ref = rotation_amounts(*lwes.arrays(k), primes, N, tb)
        assert int(rots[k]) == rot, k
        assert [int(v) for v in steps[k * n:(k + 1) * n]] == ref, k


@pytest.fixture(scope="module")
def setup1024():
    """N = 1024 (the fused one-launch blind rotation's shape), a short LWE key (n = 40)."""
    arith.NTT_processor_instance.reset()
    lib = _native.load()
    lib.vfhe_prng_set_deterministic_seed(0x1024)
    rng = np.random.default_rng(0x1024)
    Rq = Ring(1024, prime_size=[50, 50, 50], split_degree=1)
    in_ring = Rq.quotient_ring(ell=2)
    lwe_key = LweKey(in_ring, [int(b) for b in rng.integers(0, 2, 40)])
    chain = Chain.nested(Rq, special=1, levels=1)
    out_key = SecretKey(chain, [binary_key(rng, 1024, 64)], 3.2)
    bk = BootstrapKey(chain, lwe_key, out_key)
    lut = list(range(1 << (MSG_PREC - 1)))
    tv = pack_lut(chain, lut, len(lut), MSG_PREC)
    return dict(in_ring=in_ring, lwe_key=lwe_key, chain=chain, out_key=out_key, bk=bk, lut=lut, tv=tv,
                boot=Bootstrapper(chain, bk), out_lwe=lwe_key_of(out_key, in_ring))


def test_fused_blind_rotation_matches_steps_and_reference_sequence(setup1024):
    """br.fused = 1 (every key index in one launch, accumulators resident in LDS) against
    br.fused = 0 (a monomial-product launch and a batched CMUX per key index) and against the
    reference's one-call-at-a-time sequence, bit for bit; the outputs decrypt to lut[msg]."""
    s = setup1024
    lib = _native.load()
    tb = 1 << (MSG_PREC - 1)
    msgs = [0, 5, 9, 14, 3]
    lwes = encrypt_msgs(s, msgs)
    saved = lib.vfhe_hip_get_option(b"br.fused")
    got = {}
    try:
        for fused in (1, 0):   # one launch, a launch sequence per index
            assert lib.vfhe_hip_set_option(b"br.fused", fused) == 0
            n0 = lib.vfhe_hip_get_option(b"stat.br_fused_launches")
            print(f"[cggi] N=1024 blind rotation of {len(msgs)} inputs, br.fused={fused}", flush=True)
            got[fused] = s["boot"].blind_rotate(s["tv"], lwes, tb)
            assert (lib.vfhe_hip_get_option(b"stat.br_fused_launches") - n0) == (1 if fused else 0)
    finally:
        lib.vfhe_hip_set_option(b"br.fused", saved)
    for k in range(len(msgs)):
        for other in (0,):
            for p, r in zip(got[1].polys(k), got[other].polys(k)):
                assert p == r, (k, other)
    # vfhe_hip_blind_rotate_batch from accumulators already rotated (the fused kernel's other start)
    rots, steps = s["boot"].amounts(lwes, tb)
    acc = Batch.alloc(s["chain"], 0, len(msgs), form="coeff")
    for h, rot in zip(acc.handles, rots):
        lib.mlwe_RNSc_mul_by_xai(h, s["tv"].handles[0], int(rot))
    lib.vfhe_hip_blind_rotate_batch(_handles(acc.handles), len(acc), s["bk"].table, s["bk"].n,
                                    steps.ctypes.data_as(_native.U64P), s["bk"].ell, s["chain"].special)
    for k in range(len(msgs)):
        for p, r in zip(acc.polys(k), got[1].polys(k)):
            assert p == r, k
    for k in (0, 3):
        ref = _reference_sequence(s, lwes, k, tb)
        for p, r in zip(got[1].polys(k), ref.polys(0)):
            assert p == r, k
    q = s["in_ring"].q_l
    phases = extract(got[1], 0).phases(s["out_lwe"])
    assert [mod_switch(ph, q, 1 << MSG_PREC) for ph in phases] == [s["lut"][v] for v in msgs]


def test_blind_rotations_queued_back_to_back(setup1024):
    """Two fused blind rotations queued without a wait between them (precomputed amounts, given
    accumulators, the second batch's host work done while the first runs) equal the same
    rotations run one at a time, bit for bit."""
    s = setup1024
    lib = _native.load()
    tb = 1 << (MSG_PREC - 1)
    boot = s["boot"]
    a_msgs, b_msgs = [1, 6, 11], [2, 15, 7, 4]
    la, lb = encrypt_msgs(s, a_msgs), encrypt_msgs(s, b_msgs)
    print("[cggi] two blind rotations queued back to back", flush=True)
    acc_a = boot.alloc_accumulators(len(a_msgs))
    boot.blind_rotate(s["tv"], la, tb, amounts=boot.amounts(la, tb), acc=acc_a)
    acc_b = boot.blind_rotate(s["tv"], lb, tb, amounts=boot.amounts(lb, tb))
    lib.vfhe_hip_synchronize()
    one_a = boot.blind_rotate(s["tv"], la, tb)
    one_b = boot.blind_rotate(s["tv"], lb, tb)
    for got, want, n in ((acc_a, one_a, len(a_msgs)), (acc_b, one_b, len(b_msgs))):
        for k in range(n):
            for p, r in zip(got.polys(k), want.polys(k)):
                assert p == r, k
    q = s["in_ring"].q_l
    phases = extract(acc_b, 0).phases(s["out_lwe"])
    assert [mod_switch(ph, q, 1 << MSG_PREC) for ph in phases] == [s["lut"][v] for v in b_msgs]


def test_bootstrap_batch_shape_n630_N1024():
    """The bootstrap benchmark's shape: LWE n = 630 (binary key), GLWE N = 1024 (rank 1, two
    50-bit level primes + one special), 1024 LWE ciphertexts of uniform 4-bit messages, identity
    table; every output decrypts to its message."""
    arith.NTT_processor_instance.reset()
    lib = _native.load()
    lib.vfhe_prng_set_deterministic_seed(0x630)
    rng = np.random.default_rng(0x630)
    Rq = Ring(1024, prime_size=[50, 50, 50], split_degree=1)
    in_ring = Rq.quotient_ring(ell=2)
    lwe_key = LweKey(in_ring, [int(b) for b in rng.integers(0, 2, 630)])
    chain = Chain.nested(Rq, special=1, levels=1)
    out_key = SecretKey(chain, [binary_key(rng, 1024, 64)], 3.2)
    print("[cggi] bootstrap key: 630 MGSW encryptions, N=1024", flush=True)
    bk = BootstrapKey(chain, lwe_key, out_key)
    lut = list(range(1 << (MSG_PREC - 1)))
    tv = pack_lut(chain, lut, len(lut), MSG_PREC)
    msgs = [int(v) for v in rng.integers(0, 16, 1024)]
    q = in_ring.q_l
    lwes = LweBatch.encrypt(lwe_key, [mod_switch(v, 1 << MSG_PREC, q) for v in msgs])
    print("[cggi] bootstrapping 1024 ciphertexts", flush=True)
    out = Bootstrapper(chain, bk).bootstrap(tv, lwes, torus_base=1 << (MSG_PREC - 1))
    phases = out.phases(lwe_key_of(out_key, in_ring))
    got = [mod_switch(ph, q, 1 << MSG_PREC) for ph in phases]
    assert got == msgs
