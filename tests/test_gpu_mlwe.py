"""GPU parity: tensor product, GHS key-switch, relinearized multiply, CKKS rescale
and automorphism of the hip engine against the oracle, bit-exact.

Key-switch keys here are uniform random NTT-form samples: bit-exactness of the
arithmetic does not depend on the key being a valid encryption (decryption
correctness with real keys is covered by tests/test_gpu_ckks.py).  The pool
layout follows modules/mlwe/python/src/vfhe/mlwe/mlwe.py:57-73: level-0 primes
first, the special primes on top.
"""
import numpy as np
import pytest

from oracle import oracle as O
from vfhe_amd import _native, runtime as rt

pytestmark = pytest.mark.gpu


class Setup:
    def __init__(self, N, L, K, seed, split=1, bits=50):
        self.N, self.L, self.K = N, L, K
        self.primes = O.gen_primes(N, [bits] * (L + K), split)
        self.pool = rt.Pool(N, self.primes, split)
        self.octx = O.Ctx(N, self.primes, split)
        self.rng = np.random.default_rng(seed)
        self.lvl_mask = (1 << L) - 1
        self.ext_mask = (1 << (L + K)) - 1
        self.keep = []

    def rows(self, mask):
        return rt.random_rows(self.rng, self.primes, self.N, mask)

    def ct(self, r, mask):
        data = np.stack([self.rows(mask) for _ in range(r + 1)])
        return data, self.pool.sample(r, mask, data)

    def key(self, ndig, r):
        """One key-switch key: ndig entries of (r+1) NTT-form polys on the extended mask."""
        data = np.stack([np.stack([self.rows(self.ext_mask) for _ in range(r + 1)]) for _ in range(ndig)])
        entries = [self.pool.sample(r, self.ext_mask, data[d]) for d in range(ndig)]
        arr = _native.voidp_array(entries)
        self.keep += [entries, arr]
        return data, arr

    def outer(self, keys):
        arr = _native.voidp_array([k[1] if k is not None else None for k in keys])
        self.keep.append(arr)
        return arr

    def check(self, got, ref, mask):
        for c in range(ref.shape[0]):
            for i in range(ref.shape[1]):
                if (mask >> i) & 1:
                    assert np.array_equal(got[c][i], ref[c][i]), f"component {c} limb {i}"


@pytest.mark.parametrize("r", [1, 2])
def test_tensor_product(lib, r):
    s = Setup(4096, 4, 0, 11)
    m = s.lvl_mask
    d1, c1 = s.ct(r, m)
    d2, c2 = s.ct(r, m)
    R = int(lib.mlwe_extended_rank(r))
    outs = [s.pool.poly() for _ in range(R + 1)]
    lib.mlwe_tensor_product(_native.voidp_array(outs), c1, c2)
    ref = np.zeros((R + 1, s.L, s.N), dtype=np.uint64)
    mref = s.octx.tensor_product(ref, d1, d2, r, m)
    got = np.stack([s.pool.read(p) for p in outs])
    s.check(got, ref, mref)


@pytest.mark.parametrize("N,L,K,r,relin", [(4096, 4, 1, 1, True), (4096, 3, 2, 1, False),
                                            (2048, 3, 2, 2, True), (1024, 2, 1, 1, False)])
def test_ghs_keyswitch(lib, N, L, K, r, relin):
    s = Setup(N, L, K, 12 + N + r)
    Rin = int(lib.mlwe_extended_rank(r)) if relin else r
    nkeys = Rin - r if relin else Rin
    keys = [s.key(L, r) for _ in range(nkeys)] + ([None] * r if relin else [])
    din, cin = s.ct(Rin, s.lvl_mask)
    out = s.pool.sample(r, s.ext_mask)
    lib.mlwe_RNSc_GHS_hybrid_keyswitch(out, cin, s.outer(keys), 0)
    ref = np.zeros((r + 1, s.L + s.K, s.N), dtype=np.uint64)
    mref = s.octx.ghs_keyswitch(ref, r, din, Rin, s.lvl_mask, [k[0] if k else None for k in keys], s.ext_mask)
    assert s.pool.sample_masks(out) == [mref] * (r + 1)
    s.check(s.pool.read_sample(out), ref, mref)


@pytest.mark.parametrize("N,L,K,r,relin,fused,bits", [
    (8192, 3, 1, 1, True, 1, 50),      # column pass + fused row-pass/inner-product kernel
    (8192, 3, 1, 1, True, 0, 50),      # two-pass NTT + FP64 inner product
    (16384, 2, 2, 2, True, 1, 50),     # three output components: FP64 inner product
    (8192, 3, 1, 1, True, 1, 60),      # primes above 2^50: integer path throughout
    (32768, 3, 2, 1, True, 1, 50),     # fused at 2^15 (7-stage column pass)
    (65536, 2, 1, 1, False, 1, 50),    # largest fused size
    (131072, 2, 1, 1, False, 1, 50),   # largest ring (2^17): generic passes + FP64 inner product
])
def test_keyswitch_schedules(lib, N, L, K, r, relin, fused, bits):
    """Every key-switch schedule (mlwe.c:557-598) against the oracle, bit-exact."""
    old = lib.vfhe_hip_get_option(b"ks.fused")
    lib.vfhe_hip_set_option(b"ks.fused", fused)
    try:
        s = Setup(N, L, K, 100 + N + r + bits, bits=bits)
        Rin = int(lib.mlwe_extended_rank(r)) if relin else r
        nkeys = Rin - r if relin else Rin
        keys = [s.key(L, r) for _ in range(nkeys)] + ([None] * r if relin else [])
        din, cin = s.ct(Rin, s.lvl_mask)
        out = s.pool.sample(r, s.ext_mask)
        print(f"[keyswitch] N={N} L={L} K={K} r={r} fused={fused} bits={bits}", flush=True)
        lib.mlwe_RNSc_GHS_hybrid_keyswitch(out, cin, s.outer(keys), 0)
        ref = np.zeros((r + 1, s.L + s.K, s.N), dtype=np.uint64)
        mref = s.octx.ghs_keyswitch(ref, r, din, Rin, s.lvl_mask, [k[0] if k else None for k in keys], s.ext_mask)
        assert s.pool.sample_masks(out) == [mref] * (r + 1)
        s.check(s.pool.read_sample(out), ref, mref)
    finally:
        lib.vfhe_hip_set_option(b"ks.fused", old)


def test_multiply_and_batched_rescale(lib):
    """mlwe_multiply (mlwe.c:711-760) and the fused multiply+relinearize+rescale batch."""
    s = Setup(4096, 4, 1, 21)
    r, m = 1, s.lvl_mask
    rlk = [s.key(s.L, r), None]
    rlk_arr = s.outer(rlk)
    rlk_np = [rlk[0][0], None]
    cts = [(s.ct(r, m), s.ct(r, m)) for _ in range(3)]
    top = 1 << (s.L - 1)
    outs_single, outs_batch = [], []
    for (d1, c1), (d2, c2) in cts:
        o = s.pool.sample(r, s.ext_mask)
        lib.mlwe_multiply(o, c1, c2, rlk_arr)
        outs_single.append(o)
        outs_batch.append(s.pool.sample(r, s.ext_mask))
    lib.vfhe_hip_mlwe_multiply_batch(_native.voidp_array(outs_batch), _native.voidp_array([c[0][1] for c in cts]),
                                     _native.voidp_array([c[1][1] for c in cts]), len(cts), rlk_arr, top)
    for k, ((d1, c1), (d2, c2)) in enumerate(cts):
        ref = np.zeros((r + 1, s.L + s.K, s.N), dtype=np.uint64)
        mref = s.octx.multiply(ref, d1, d2, r, m, rlk_np, s.ext_mask)
        assert mref == m
        s.check(s.pool.read_sample(outs_single[k]), ref, mref)
        # CKKS rescale (ckks.py:128-148): to coefficient form, round-divide by the top prime
        for c in range(r + 1):
            s.octx.to_coeff(ref[c], ref[c], mref)
        m2 = s.octx.mlwe_round_division(ref, r, mref, top)
        assert s.pool.sample_masks(outs_batch[k]) == [m2] * (r + 1)
        s.check(s.pool.read_sample(outs_batch[k]), ref, m2)


@pytest.mark.parametrize("K", [1, 2, 3, 4])
def test_moddown_rescale_fp64_against_integer_and_oracle(lib, K):
    """The FP64 ModDown + pass-through + rescale kernel (ks.mdr_fp64 = 1: the sequential round
    divisions folded into per-limb affine constants) against the integer kernel (ks.mdr_fp64 = 0)
    and, for the first pair, the oracle: 1 to 4 special primes, a batch of 3."""
    s = Setup(4096, 5, K, 40 + K)
    r, m = 1, s.lvl_mask
    rlk = [s.key(s.L, r), None]
    rlk_arr = s.outer(rlk)
    cts = [(s.ct(r, m), s.ct(r, m)) for _ in range(3)]
    top = 1 << (s.L - 1)
    got = {}
    saved = lib.vfhe_hip_get_option(b"ks.mdr_fp64")
    try:
        for fp in (1, 0):
            assert lib.vfhe_hip_set_option(b"ks.mdr_fp64", fp) == 0
            outs = [s.pool.sample(r, s.ext_mask) for _ in cts]
            n0 = lib.vfhe_hip_get_option(b"stat.mdr_fp64_launches")
            lib.vfhe_hip_mlwe_multiply_batch(_native.voidp_array(outs), _native.voidp_array([c[0][1] for c in cts]),
                                             _native.voidp_array([c[1][1] for c in cts]), len(cts), rlk_arr, top)
            assert (lib.vfhe_hip_get_option(b"stat.mdr_fp64_launches") > n0) == bool(fp)
            got[fp] = [s.pool.read_sample(o) for o in outs]
            masks = [s.pool.sample_masks(o) for o in outs]
    finally:
        lib.vfhe_hip_set_option(b"ks.mdr_fp64", saved)
    m2 = m & ~top
    assert masks == [[m2] * (r + 1)] * len(cts)
    for a, b in zip(got[1], got[0]):
        s.check(a, b, m2)
    (d1, _), (d2, _) = cts[0]
    ref = np.zeros((r + 1, s.L + s.K, s.N), dtype=np.uint64)
    mref = s.octx.multiply(ref, d1, d2, r, m, [rlk[0][0], None], s.ext_mask)
    for c in range(r + 1):
        s.octx.to_coeff(ref[c], ref[c], mref)
    assert s.octx.mlwe_round_division(ref, r, mref, top) == m2
    s.check(got[1][0], ref, m2)


@pytest.mark.parametrize("rescale", [False, True])
def test_multiply_pass_through_folded_into_key_mac(lib, rescale):
    """ks.fold = 1: the tensor's linear and constant slots stay in NTT form and enter the fused
    key inner product as P * d (k_ntt16_row_mac), instead of being added after the ModDown.
    Bit-equal to ks.fold = 0 and to the oracle (mlwe.c:711-760, then ckks.py:128-148)."""
    s = Setup(8192, 4, 2, 77 + rescale)
    r, m = 1, s.lvl_mask
    rlk = [s.key(s.L, r), None]
    rlk_arr = s.outer(rlk)
    cts = [(s.ct(r, m), s.ct(r, m)) for _ in range(3)]
    top = (1 << (s.L - 1)) if rescale else 0
    got = {}
    saved = lib.vfhe_hip_get_option(b"ks.fold")
    try:
        for fold in (1, 0):
            assert lib.vfhe_hip_set_option(b"ks.fold", fold) == 0
            outs = [s.pool.sample(r, s.ext_mask) for _ in cts]
            n0 = lib.vfhe_hip_get_option(b"stat.ks_fold_jobs")
            lib.vfhe_hip_mlwe_multiply_batch(_native.voidp_array(outs), _native.voidp_array([c[0][1] for c in cts]),
                                             _native.voidp_array([c[1][1] for c in cts]), len(cts), rlk_arr, top)
            assert (lib.vfhe_hip_get_option(b"stat.ks_fold_jobs") - n0) == (len(cts) if fold else 0)
            got[fold] = [s.pool.read_sample(o) for o in outs]
    finally:
        lib.vfhe_hip_set_option(b"ks.fold", saved)
    m2 = m & ~top
    for a, b in zip(got[1], got[0]):
        s.check(a, b, m2)
    (d1, _), (d2, _) = cts[1]
    ref = np.zeros((r + 1, s.L + s.K, s.N), dtype=np.uint64)
    mref = s.octx.multiply(ref, d1, d2, r, m, [rlk[0][0], None], s.ext_mask)
    if rescale:
        for c in range(r + 1):
            s.octx.to_coeff(ref[c], ref[c], mref)
        mref = s.octx.mlwe_round_division(ref, r, mref, top)
    assert mref == m2
    s.check(got[1][1], ref, m2)


def test_automorphism(lib):
    s = Setup(2048, 3, 1, 31)
    r = 1
    keys = [s.key(s.L, r)]
    din, cin = s.ct(r, s.lvl_mask)
    out = s.pool.sample(r, s.ext_mask)
    gen = pow(5, 3, 2 * s.N)
    lib.mlwe_automorphism_RNSc_GHS(out, cin, gen, s.outer(keys), 0)
    ref = np.zeros((r + 1, s.L + s.K, s.N), dtype=np.uint64)
    mref = s.octx.automorphism(ref, din, r, s.lvl_mask, gen, [keys[0][0]], s.ext_mask)
    s.check(s.pool.read_sample(out), ref, mref)


@pytest.mark.slow
def test_ckks_config3_shape_multiply_rescale(lib):
    """The benchmarked CKKS shape: N = 2^16, L = 24 level limbs + K = 4 special limbs, rank 1."""
    s = Setup(65536, 24, 4, 41)
    r, m = 1, s.lvl_mask
    rlk = [s.key(s.L, r), None]
    rlk_arr = s.outer(rlk)
    (d1, c1), (d2, c2) = s.ct(r, m), s.ct(r, m)
    top = 1 << (s.L - 1)
    o = s.pool.sample(r, s.ext_mask)
    lib.vfhe_hip_mlwe_multiply_batch(_native.voidp_array([o]), _native.voidp_array([c1]), _native.voidp_array([c2]),
                                     1, rlk_arr, top)
    ref = np.zeros((r + 1, s.L + s.K, s.N), dtype=np.uint64)
    mref = s.octx.multiply(ref, d1, d2, r, m, [rlk[0][0], None], s.ext_mask)
    for c in range(r + 1):
        s.octx.to_coeff(ref[c], ref[c], mref)
    m2 = s.octx.mlwe_round_division(ref, r, mref, top)
    s.check(s.pool.read_sample(o), ref, m2)


def _oracle_multiply_rescale(s, d1, d2, rlk_np, top):
    """Oracle: mlwe_multiply (mlwe.c:711-760), then INTT and round division by the top prime (ckks.py:128-148)."""
    r, m = 1, s.lvl_mask
    ref = np.zeros((r + 1, s.L + s.K, s.N), dtype=np.uint64)
    mref = s.octx.multiply(ref, d1, d2, r, m, rlk_np, s.ext_mask)
    for c in range(r + 1):
        s.octx.to_coeff(ref[c], ref[c], mref)
    m2 = s.octx.mlwe_round_division(ref, r, mref, top)
    return ref, m2


@pytest.mark.slow
def test_ckks_config3_batch32_multiply_rescale(lib):
    """The benchmarked batch: 32 ciphertext pairs at N = 2^16, L = 24, K = 4 in one batched call (the
    key-switch row pass remaps workgroups across XCDs only when several ciphertexts share a key block).
    Every output equals the same pair multiplied alone (batch of one), and two outputs -- the first and
    the last -- equal the oracle bit for bit."""
    import concurrent.futures as cf
    B = 32
    s = Setup(65536, 24, 4, 43)
    r, m = 1, s.lvl_mask
    rlk = [s.key(s.L, r), None]
    rlk_arr = s.outer(rlk)
    pairs = [(s.ct(r, m), s.ct(r, m)) for _ in range(B)]
    top = 1 << (s.L - 1)
    outs = [s.pool.sample(r, s.ext_mask) for _ in range(B)]
    print(f"[ckks batch] B={B} N=65536 L=24 K=4", flush=True)
    lib.vfhe_hip_mlwe_multiply_batch(_native.voidp_array(outs), _native.voidp_array([p[0][1] for p in pairs]),
                                     _native.voidp_array([p[1][1] for p in pairs]), B, rlk_arr, top)
    with cf.ThreadPoolExecutor(2) as ex:   # the oracle runs beside the GPU checks (ctypes drops the GIL)
        futs = {k: ex.submit(_oracle_multiply_rescale, s, pairs[k][0][0], pairs[k][1][0], [rlk[0][0], None], top)
                for k in (0, B - 1)}
        one = s.pool.sample(r, s.ext_mask)
        for k in range(B):
            lib.vfhe_hip_mlwe_multiply_batch(_native.voidp_array([one]), _native.voidp_array([pairs[k][0][1]]),
                                             _native.voidp_array([pairs[k][1][1]]), 1, rlk_arr, top)
            got, alone = s.pool.read_sample(outs[k]), s.pool.read_sample(one)
            assert s.pool.sample_masks(outs[k]) == s.pool.sample_masks(one)
            s.check(got, alone, s.pool.sample_masks(one)[0])
        for k, f in futs.items():
            ref, m2 = f.result()
            assert s.pool.sample_masks(outs[k]) == [m2] * (r + 1)
            s.check(s.pool.read_sample(outs[k]), ref, m2)
    lib.free_mlwe_RNS_sample(one)
    for o in outs:
        lib.free_mlwe_RNS_sample(o)
    for (_, c1), (_, c2) in pairs:
        lib.free_mlwe_RNS_sample(c1)
        lib.free_mlwe_RNS_sample(c2)


@pytest.mark.parametrize("N,L,split", [(4096, 4, 1), (65536, 3, 1), (4096, 3, 4)])
@pytest.mark.parametrize("mode,fn", [(0, "mlwe_RNS_mul_by_poly"), (1, "mlwe_RNS_mul_addto_by_poly"),
                                     (2, "mlwe_RNS_mul_subto_by_poly")])
def test_mlwe_mul_by_poly(lib, N, L, split, mode, fn):
    """mlwe_RNS_mul{,_addto,_subto}_by_poly (mlwe.c:254-279): each of the r+1 components times one
    NTT-form polynomial (polynomial_mul{,_addto,_subto}_RNS_polynomial, rns_polynomial.c:311-435; with
    split degree 4 the products wrap through w), against the oracle, bit-exact.  The output's mask is
    the input's AND the polynomial's."""
    r = 2
    s = Setup(N, L, 0, 71 + N + mode + split, split=split)
    mc, mp = s.lvl_mask, s.lvl_mask & ~0b10 if L > 2 else s.lvl_mask
    din, cin = s.ct(r, mc)
    pdat = s.rows(mp)
    poly = s.pool.poly(pdat, mp)
    dout, cout = s.ct(r, s.lvl_mask)          # starting values for the accumulating forms
    print(f"[{fn}] N={N} L={L} split={split}", flush=True)
    getattr(lib, fn)(cout, cin, poly)
    ref = dout.copy()
    for c in range(r + 1):
        mo = s.octx.mul(ref[c], din[c], mc, pdat, mp, mode)
    got = s.pool.read_sample(cout)
    assert s.pool.sample_masks(cout) == [mo] * (r + 1)
    s.check(got, ref, mo)
    lib.free_mlwe_RNS_sample(cin)
    lib.free_mlwe_RNS_sample(cout)
    s.pool.free(poly)


def test_multiply_autotuner_adopts_a_timed_schedule(lib):
    """vfhe_hip_autotune_multiply times its 12 schedules on a batch, adopts the fastest (only the
    options it tunes change), and a multiply afterwards still equals the oracle."""
    import ctypes
    s = Setup(8192, 4, 2, 91)
    r, m = 1, s.lvl_mask
    rlk = [s.key(s.L, r), None]
    rlk_arr = s.outer(rlk)
    cts = [(s.ct(r, m), s.ct(r, m)) for _ in range(2)]
    top = 1 << (s.L - 1)
    keys = (b"ntt.col_wide", b"ntt.tile_grid", b"ks.fold", b"ntt.mode", b"ks.fused", b"ks.mdr_fp64")
    before = {k: lib.vfhe_hip_get_option(k) for k in keys}
    outs = [s.pool.sample(r, s.ext_mask) for _ in cts]
    us = (ctypes.c_double * 12)()
    try:
        best = lib.vfhe_hip_autotune_multiply(_native.voidp_array(outs), _native.voidp_array([c[0][1] for c in cts]),
                                              _native.voidp_array([c[1][1] for c in cts]), len(cts), rlk_arr, top, 1, us)
        print("[autotune multiply] best", best, [round(u, 1) for u in us], flush=True)
        assert 0 <= best < 12 and all(u > 0 for u in us)
        after = {k: lib.vfhe_hip_get_option(k) for k in keys}
        for k in (b"ntt.mode", b"ks.fused", b"ks.mdr_fp64"):
            assert after[k] == before[k]
        assert after[b"ntt.col_wide"] == best % 3 and after[b"ks.fold"] == (0 if best // 6 else 1)
        o = s.pool.sample(r, s.ext_mask)
        lib.vfhe_hip_mlwe_multiply_batch(_native.voidp_array([o]), _native.voidp_array([cts[0][0][1]]),
                                         _native.voidp_array([cts[0][1][1]]), 1, rlk_arr, top)
        (d1, _), (d2, _) = cts[0]
        ref = np.zeros((r + 1, s.L + s.K, s.N), dtype=np.uint64)
        mref = s.octx.multiply(ref, d1, d2, r, m, [rlk[0][0], None], s.ext_mask)
        for c in range(r + 1):
            s.octx.to_coeff(ref[c], ref[c], mref)
        m2 = s.octx.mlwe_round_division(ref, r, mref, top)
        s.check(s.pool.read_sample(o), ref, m2)
    finally:
        for k, v in before.items():
            lib.vfhe_hip_set_option(k, v)
