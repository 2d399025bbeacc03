"""Serving loop for CKKS requests of mixed ring size and level on the hip engine.

A request is a homomorphic multiply + relinearize + rescale, or a rotation
(automorphism X -> X^(5^k) + key-switch), on a ciphertext of ring size
N in {2^13 .. 2^16} at a level of 4..24 limbs.  Ciphertexts and evaluation keys
live in HBM for the whole run; keys are stored once per ring at the top level
and every lower level uses a view of them (vfhe_hip_mlwe_view_array).  The
request loop itself is the engine's serving runtime (vfhe_hip_server_*):
admission by memory footprint, continuous batching per request class, batches
on several HIP streams.  This module builds the resident state, generates
Poisson arrivals from a seed, drives the runtime in real time and reports
throughput and latency percentiles.

Orchestration follows the reference's CKKS level bookkeeping
(modules/fhe/python/src/vfhe/fhe/ckks.py:128-148: rescale drops the top prime of
the level; modules/mlwe/python/src/vfhe/mlwe/mlwe.py:65-73: special primes
follow the level primes in the ring).
"""
from __future__ import annotations

import ctypes
import sys
import threading
import time
from dataclasses import dataclass

import numpy as np

REQ_MULTIPLY, REQ_ROTATE = 0, 1


@dataclass(frozen=True)
class Request:
    t: float          # arrival time, seconds from the start of the run
    op: int           # REQ_MULTIPLY or REQ_ROTATE
    N: int
    level: int        # active level limbs
    step: int         # rotation step (index into the ring's rotation key set)


def make_workload(rate: float, duration: float, Ns, levels=(4, 24), p_multiply: float = 0.5,
                  rot_steps: int = 4, seed: int = 0) -> list[Request]:
    """Poisson arrivals at `rate` per second for `duration` seconds; N, level, op and step uniform."""
    rng = np.random.default_rng(seed)
    out, t = [], 0.0
    while True:
        t += float(rng.exponential(1.0 / rate))
        if t >= duration:
            return out
        op = REQ_MULTIPLY if rng.random() < p_multiply else REQ_ROTATE
        out.append(Request(t, op, int(rng.choice(list(Ns))), int(rng.integers(levels[0], levels[1] + 1)),
                           int(rng.integers(0, rot_steps))))


def percentiles(lat_s, qs=(50, 99)) -> dict:
    a = np.asarray(lat_s, dtype=np.float64) * 1e3
    return {f"p{q}_ms": round(float(np.percentile(a, q)), 3) for q in qs} if len(a) else {}


class Ring:
    """Resident state of one ring size: pool, keys (top level) and level views, input ciphertexts."""

    def __init__(self, lib, N: int, L: int = 24, K: int = 4, rot_steps: int = 4, levels=(4, 24)):
        from . import _native, params, runtime as rt
        self.lib, self.N, self.L, self.K = lib, N, L, K
        self.primes = params.gen_primes(N, [50] * (L + K))
        self.pool = rt.Pool(N, self.primes)
        self.special = ((1 << (L + K)) - 1) ^ ((1 << L) - 1)
        full = (1 << (L + K)) - 1
        self._keep = []

        def random_key():
            # digits = L entries of (r=1) samples on the full ring, uniform residues (on the device)
            entries = []
            for _ in range(L):
                c = self.pool.sample(1, full)
                for p in self.pool.comps(c):
                    lib.polynomial_gen_random_RNSc_polynomial(p)
                entries.append(c)
            self._keep += entries
            return _native.voidp_array(entries)

        self.rlk_top = random_key()
        self.rot_gens = [pow(5, 1 << k, 2 * N) for k in range(rot_steps)]
        self.rot_top = [random_key() for _ in range(rot_steps)]
        self.views, self.rlk, self.rot, self.cts = {}, {}, {}, {}
        for lvl in range(levels[0], levels[1] + 1):
            m = ((1 << lvl) - 1) | self.special
            v = lib.vfhe_hip_mlwe_view_array(self.rlk_top, lvl, m)
            self.views[lvl] = [v] + [lib.vfhe_hip_mlwe_view_array(k, lvl, m) for k in self.rot_top]
            self.rlk[lvl] = _native.voidp_array([v, None])
            self.rot[lvl] = [_native.voidp_array([k]) for k in self.views[lvl][1:]]
            lm = (1 << lvl) - 1
            cs = []
            for _ in range(2):
                c = self.pool.sample(1, lm)
                for p in self.pool.comps(c):
                    lib.polynomial_gen_random_RNSc_polynomial(p)
                cs.append(c)
            self.cts[lvl] = cs
        lib.vfhe_hip_synchronize()

    def args(self, req: Request):
        """(op, in1, in2, ksk, gen, divide_mask) of a request on this ring."""
        c1, c2 = self.cts[req.level]
        if req.op == REQ_MULTIPLY:
            return REQ_MULTIPLY, c1, c2, self.rlk[req.level], 0, 1 << (req.level - 1)
        return REQ_ROTATE, c1, None, self.rot[req.level][req.step], self.rot_gens[req.step], 0

    def free(self):
        """Views first (they borrow the keys' rows), then keys and ciphertexts."""
        for lvl, vs in self.views.items():
            for v in vs:
                self.lib.free_RNS_mlwe_array(lvl, v)
        self.views.clear()
        for c in self._keep + [c for cs in self.cts.values() for c in cs]:
            self.lib.free_mlwe_RNS_sample(c)
        self._keep, self.cts = [], {}


class Server:
    """ctypes wrapper of the engine's serving runtime."""

    def __init__(self, lib, memory_budget: int, lanes: int = 4, max_batch: int = 32):
        self.lib = lib
        self.h = lib.vfhe_hip_server_new(memory_budget, lanes, max_batch)
        if not self.h:
            raise ValueError("invalid server parameters")

    def submit(self, op, in1, in2, ksk, gen, divide_mask) -> int:
        rid = self.lib.vfhe_hip_server_submit(self.h, op, in1, in2, ksk, gen, divide_mask)
        if not rid:
            raise ValueError("request rejected")
        return rid

    def poll(self, max_n: int = 256, wait: bool = True):
        """Completed requests: (id, output, latency s, (queued s, enqueued s, device + completion s))."""
        ids = (ctypes.c_uint64 * max_n)()
        outs = (ctypes.c_void_p * max_n)()
        lat = (ctypes.c_double * max_n)()
        parts = (ctypes.c_double * (3 * max_n))()
        n = self.lib.vfhe_hip_server_poll_timed(self.h, ids, outs, lat, parts, max_n, int(wait))
        return [(int(ids[k]), outs[k], float(lat[k]), (parts[3 * k], parts[3 * k + 1], parts[3 * k + 2]))
                for k in range(n)]

    def free_outputs(self, samples):
        """Frees a list of output samples with one engine call."""
        arr = (ctypes.c_void_p * len(samples))(*samples)
        self.lib.vfhe_hip_free_mlwe_samples(arr, len(samples))

    def stats(self) -> dict:
        s = (ctypes.c_double * 10)()
        self.lib.vfhe_hip_server_stats(self.h, s)
        return {"batches": int(s[0]), "requests": int(s[1]), "mean_batch": round(s[2], 2),
                "peak_inflight_bytes": int(s[3]), "max_lanes_busy": int(s[4]), "pending": int(s[5]),
                "host_enqueue_s": round(s[6], 4), "device_batch_s": round(s[7], 4), "server_s": round(s[8], 4),
                "max_batch_enqueue_ms": round(s[9] * 1e3, 3)}

    def close(self):
        if self.h:
            self.lib.vfhe_hip_server_free(self.h)
            self.h = None


def cache_counters(lib) -> dict:
    """The engine's device block cache counters (hits, misses, out-of-memory flushes, parked bytes)."""
    return {k: int(lib.vfhe_hip_get_option(f"stat.cache_{k}".encode())) for k in ("hits", "misses", "flushes",
                                                                                   "parked_bytes")}


def run(lib, rings: dict, workload: list[Request], memory_budget: int, lanes: int = 4, max_batch: int = 32,
        keep_outputs: bool = False) -> dict:
    """Replay `workload` open-loop in real time through a fresh server; return throughput, latencies, stats.

    A request's latency runs from its scheduled arrival to its completion: the
    runtime's submit-to-completion time plus how late the replay submitted it,
    so a client that falls behind shows up in the latencies (`service_p99_ms`
    is the runtime's part alone).  With keep_outputs the output samples are
    returned by request index (the caller frees them); otherwise each poll's
    outputs are freed together as they complete."""
    srv = Server(lib, memory_budget, lanes, max_batch)
    n = len(workload)
    lat = [None] * n
    svc = [None] * n
    parts = [None] * n
    cache0 = cache_counters(lib)
    late = [0.0] * n
    outs = [None] * n
    index = {}
    done_t = [0.0]
    switch = sys.getswitchinterval()
    sys.setswitchinterval(2e-4)
    t0 = time.perf_counter()

    def collector():
        got = 0
        while got < n:
            batch = srv.poll(wait=True)
            if not batch:                 # nothing pending yet: the submitter is between arrivals
                time.sleep(0.0002)
                continue
            frees = []
            for rid, out, l, pt in batch:
                k = index.pop(rid, None)
                while k is None:          # submit() returns after the runtime may already have finished it
                    time.sleep(0)
                    k = index.pop(rid, None)
                svc[k] = l
                parts[k] = pt
                lat[k] = l + late[k]
                if keep_outputs:
                    outs[k] = out
                else:
                    frees.append(out)
            if frees:
                srv.free_outputs(frees)
            got += len(batch)
        done_t[0] = time.perf_counter()

    th = threading.Thread(target=collector, daemon=True)
    th.start()
    try:
        k = 0
        while k < n:
            now = time.perf_counter() - t0
            wait = workload[k].t - now
            if wait > 0:
                time.sleep(wait - 2e-4 if wait > 4e-4 else 0)
                continue
            while k < n and workload[k].t <= now:     # every request that is due
                req = workload[k]
                ts = time.perf_counter() - t0
                rid = srv.submit(*rings[req.N].args(req))
                late[k] = max(ts - req.t, 0.0)
                index[rid] = k
                k += 1
        th.join()
    finally:
        sys.setswitchinterval(switch)
    stats = srv.stats()
    srv.close()
    span = max(done_t[0] - t0, 1e-9)
    res = {"requests": n, "seconds": round(span, 4), "requests_per_s": round(n / span, 2),
           "offered_per_s": round(n / max(workload[-1].t, 1e-9), 2) if n else 0.0}
    res.update(percentiles([x for x in lat if x is not None]))
    sv = [x for x in svc if x is not None]
    if sv:
        res["service_p99_ms"] = round(float(np.percentile(np.asarray(sv) * 1e3, 99)), 3)
        res["max_lateness_ms"] = round(max(late) * 1e3, 3)
        res["client_late_p99_ms"] = round(float(np.percentile(np.asarray(late) * 1e3, 99)), 3)
    pt = np.array([p for p in parts if p is not None]) * 1e3
    if len(pt):
        # where the tail comes from: each phase's p99, and the phases of the slowest request
        worst = int(np.argmax([(s or 0.0) + lt for s, lt in zip(svc, late)]))
        res["phase_p99_ms"] = {name: round(float(np.percentile(pt[:, i], 99)), 3)
                               for i, name in enumerate(("queued", "enqueue", "device"))}
        if parts[worst] is not None:
            res["slowest_request_ms"] = {"client_late": round(late[worst] * 1e3, 3),
                                         **{name: round(parts[worst][i] * 1e3, 3)
                                            for i, name in enumerate(("queued", "enqueue", "device"))}}
    cache1 = cache_counters(lib)
    res["block_cache"] = {k: cache1[k] - cache0[k] for k in ("hits", "misses", "flushes")}
    res["block_cache"]["parked_bytes"] = cache1["parked_bytes"]
    res.update(stats)
    if keep_outputs:
        res["outputs"] = outs
    return res
