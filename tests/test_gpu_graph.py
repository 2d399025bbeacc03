"""Graph-captured CKKS multiply batch (vfhe_hip_mlwe_multiply_batch_graph) against the
eager batched entry point it mirrors (vfhe_hip_mlwe_multiply_batch), bit for bit.

The first call of a shape runs eagerly and captures the launch sequence; every later call
replays the graph with work lists rebuilt for its own operands.  Each replay is checked
against the eager entry point on the same operands, with other eager work and device
synchronisations between the replays; one replay is also checked against the oracle."""
import numpy as np
import pytest

from oracle import oracle as O
from vfhe_amd import _native, runtime as rt

pytestmark = pytest.mark.gpu


def _stat(lib, name):
    return lib.vfhe_hip_get_option(f"stat.{name}".encode())


class Ring:
    def __init__(self, lib, N, L, K, seed):
        self.lib, self.N, self.L, self.K = lib, N, L, K
        self.primes = O.gen_primes(N, [50] * (L + K))
        self.pool = rt.Pool(N, self.primes)
        self.rng = np.random.default_rng(seed)
        self.lvl, self.ext = (1 << L) - 1, (1 << (L + K)) - 1
        kdata = np.stack([np.stack([self.rows(self.ext), self.rows(self.ext)]) for _ in range(L)])
        self.kdata = kdata
        self.entries = [self.pool.sample(1, self.ext, kdata[d]) for d in range(L)]
        self.key = _native.voidp_array(self.entries)
        self.rlk = _native.voidp_array([self.key, None])

    def rows(self, mask):
        return rt.random_rows(self.rng, self.primes, self.N, mask)

    def ct(self):
        d = np.stack([self.rows(self.lvl), self.rows(self.lvl)])
        return d, self.pool.sample(1, self.lvl, d)


@pytest.mark.parametrize("N,L,K,B,calls", [(8192, 6, 2, 4, 4), (65536, 24, 4, 2, 3)])
def test_graph_replays_match_eager(lib, N, L, K, B, calls):
    g = Ring(lib, N, L, K, N + B)
    top = 1 << (L - 1)
    cap0, rep0 = _stat(lib, "graph_captures"), _stat(lib, "graph_replays")
    for call in range(calls):
        pairs = [(g.ct(), g.ct()) for _ in range(B)]
        a1 = _native.voidp_array([p[0][1] for p in pairs])
        a2 = _native.voidp_array([p[1][1] for p in pairs])
        og = [g.pool.sample(1, g.ext) for _ in range(B)]
        oe = [g.pool.sample(1, g.ext) for _ in range(B)]
        print(f"[graph] N={N} L={L} K={K} B={B} call={call}", flush=True)
        lib.vfhe_hip_mlwe_multiply_batch_graph(_native.voidp_array(og), a1, a2, B, g.rlk, top)
        lib.vfhe_hip_mlwe_multiply_batch(_native.voidp_array(oe), a1, a2, B, g.rlk, top)
        lib.vfhe_hip_synchronize()
        for k in range(B):
            assert g.pool.sample_masks(og[k]) == g.pool.sample_masks(oe[k])
            m = g.pool.sample_masks(oe[k])[0]
            got, exp = g.pool.read_sample(og[k]), g.pool.read_sample(oe[k])
            rows = [i for i in range(L + K) if (m >> i) & 1]
            assert np.array_equal(got[:, rows], exp[:, rows]), (call, k)
        if call == 1 and N <= 8192:
            # one replayed output against the oracle
            c = O.Ctx(N, g.primes)
            (d1, _), (d2, _) = pairs[0]
            ref = np.zeros((2, L + K, N), dtype=np.uint64)
            mref = c.multiply(ref, d1, d2, 1, g.lvl, [g.kdata, None], g.ext)
            for comp in range(2):
                c.to_coeff(ref[comp], ref[comp], mref)
            m2 = c.mlwe_round_division(ref, 1, mref, top)
            got = g.pool.read_sample(og[0])
            for i in range(L + K):
                if (m2 >> i) & 1:
                    assert np.array_equal(got[:, i], ref[:, i]), i
        for p in og + oe:
            lib.free_mlwe_RNS_sample(p)
        for (_, c1), (_, c2) in pairs:
            lib.free_mlwe_RNS_sample(c1)
            lib.free_mlwe_RNS_sample(c2)
    print("[graph] captures", _stat(lib, "graph_captures") - cap0, "replays", _stat(lib, "graph_replays") - rep0,
          "fallbacks", _stat(lib, "graph_fallbacks"), "why", _stat(lib, "graph_why"), _stat(lib, "graph_why_want"),
          _stat(lib, "graph_why_got"), flush=True)
    assert _stat(lib, "graph_captures") - cap0 == 1
    assert _stat(lib, "graph_replays") - rep0 == calls - 1
    lib.vfhe_hip_graph_clear()
    assert _stat(lib, "graph_plans") == 0


def test_graph_new_shape_captures_again(lib):
    """A batch of another level is another plan; the first plan keeps replaying correctly."""
    N, L, K = 8192, 5, 1
    g = Ring(lib, N, L, K, 77)
    cap0 = _stat(lib, "graph_captures")
    for lvl in (5, 4, 5, 4):
        m = (1 << lvl) - 1
        d = [np.stack([g.rows(m), g.rows(m)]) for _ in range(4)]
        cs = [g.pool.sample(1, m, x) for x in d]
        og, oe = g.pool.sample(1, g.ext), g.pool.sample(1, g.ext)
        special = g.ext ^ g.lvl
        view = lib.vfhe_hip_mlwe_view_array(g.key, lvl, m | special)   # the key at this level
        rlk = _native.voidp_array([view, None])
        lib.vfhe_hip_mlwe_multiply_batch_graph(_native.voidp_array([og]), _native.voidp_array([cs[0]]),
                                               _native.voidp_array([cs[1]]), 1, rlk, 1 << (lvl - 1))
        lib.vfhe_hip_mlwe_multiply_batch(_native.voidp_array([oe]), _native.voidp_array([cs[0]]),
                                         _native.voidp_array([cs[1]]), 1, rlk, 1 << (lvl - 1))
        mm = g.pool.sample_masks(oe)[0]
        assert g.pool.sample_masks(og)[0] == mm
        rows = [i for i in range(L + K) if (mm >> i) & 1]
        assert np.array_equal(g.pool.read_sample(og)[:, rows], g.pool.read_sample(oe)[:, rows]), lvl
        for p in cs + [og, oe]:
            lib.free_mlwe_RNS_sample(p)
        lib.vfhe_hip_synchronize()
        lib.free_RNS_mlwe_array(lvl, view)
    assert _stat(lib, "graph_captures") - cap0 == 2
    lib.vfhe_hip_graph_clear()


@pytest.mark.parametrize("N,L,P", [(16384, 6, 5), (65536, 24, 4)])
def test_fused_product_graph_replays_match_eager(lib, N, L, P):
    """vfhe_hip_ntt_mul_intt_batch_graph: a capture, then replays on new operands, each equal bit for
    bit to the eager vfhe_hip_ntt_mul_intt_batch and (first replay, one pair) to the oracle, with eager
    work and synchronisations between the replays."""
    primes = O.gen_primes(N, [50] * L)
    pool = rt.Pool(N, primes)
    octx = O.Ctx(N, primes)
    rng = np.random.default_rng(N + P)
    cap0, rep0 = _stat(lib, "graph_captures"), _stat(lib, "graph_replays")
    for call in range(3):
        xs = [rt.random_rows(rng, primes, N) for _ in range(P)]
        ys = [rt.random_rows(rng, primes, N) for _ in range(P)]
        px, py = [pool.poly(x) for x in xs], [pool.poly(y) for y in ys]
        zg, ze = [pool.poly() for _ in range(P)], [pool.poly() for _ in range(P)]
        print(f"[fused graph] N={N} L={L} pairs={P} call={call}", flush=True)
        lib.vfhe_hip_ntt_mul_intt_batch_graph(_native.voidp_array(zg), _native.voidp_array(px),
                                              _native.voidp_array(py), P)
        lib.vfhe_hip_ntt_mul_intt_batch(_native.voidp_array(ze), _native.voidp_array(px),
                                        _native.voidp_array(py), P)
        lib.vfhe_hip_synchronize()
        for a, b in zip(zg, ze):
            assert np.array_equal(pool.read(a), pool.read(b)), call
        if call == 1:
            x, y = xs[0].copy(), ys[0].copy()
            m = pool.full_mask()
            octx.to_ntt(x, x, m)
            octx.to_ntt(y, y, m)
            z = np.zeros_like(x)
            octx.mul(z, x, m, y, m, 0)
            octx.to_coeff(z, z, m)
            assert np.array_equal(pool.read(zg[0]), z)
        for p in px + py + zg + ze:
            pool.free(p)
    assert _stat(lib, "graph_captures") - cap0 == 1
    assert _stat(lib, "graph_replays") - rep0 == 2


@pytest.mark.parametrize("N,L,P", [(4096, 4, 6), (65536, 24, 3)])
def test_ntt_batch_graph_replays_match_eager(lib, N, L, P):
    """vfhe_hip_ntt_batch_graph: forward and inverse batches captured once and replayed on new
    rows, each equal to the oracle's transform and the round trip the identity."""
    primes = O.gen_primes(N, [50] * L)
    pool = rt.Pool(N, primes)
    octx = O.Ctx(N, primes)
    rng = np.random.default_rng(N + L)
    rep0 = _stat(lib, "graph_replays")
    for call in range(3):
        xs = [rt.random_rows(rng, primes, N) for _ in range(P)]
        ps = [pool.poly(x) for x in xs]
        arr = _native.voidp_array(ps)
        print(f"[ntt graph] N={N} L={L} polys={P} call={call}", flush=True)
        lib.vfhe_hip_ntt_batch_graph(arr, P, 0)
        lib.vfhe_hip_synchronize()
        for p, x in zip(ps, xs):
            r = x.copy()
            octx.to_ntt(r, r, pool.full_mask())
            assert np.array_equal(pool.read(p), r), call
        lib.vfhe_hip_ntt_batch_graph(arr, P, 1)
        lib.vfhe_hip_synchronize()
        for p, x in zip(ps, xs):
            assert np.array_equal(pool.read(p), x), call
        for p in ps:
            pool.free(p)
    assert _stat(lib, "graph_replays") - rep0 == 4
