#!/usr/bin/env python3
"""Headline benchmark of the hip engine on one MI355X.

Metric: NTT round-trips/s at N = 2^16, L = 24 (one round trip = forward plus
inverse transform of every limb of one polynomial), on the batched workload of
64 polynomials whose 24 limbs are ~50-bit primes q = 1 mod 2^17 from
Ring.gen_prime.  A step is one pass of the hot path over the batch: one
forward launch sequence and one inverse launch sequence over all 64 x 24
limb rows, inputs already resident in HBM.

Also reported on the same line:
  * roofline  -- the NTT pass kernels' algorithmic HBM rate, timed live with HIP
                 events on the engine stream (vfhe_hip_profile_*), against 8 TB/s;
  * cpu_baseline -- the reference's own ntt_forward/ntt_reverse (compiled from
                 its sources into oracle/_ref) on this host's cores, bounded sample;
  * fused     -- z = INTT(NTT(x) * NTT(y)) over 64 pairs (products/s);
  * plain     -- the same round trips as one C-ABI call per polynomial, with and
                 without host residency;
  * ckks      -- CKKS multiply + relinearize + rescale at N = 2^16, L = 24 + 4
                 special limbs, batch of 32 ciphertext pairs (mults/s, p50/p99 over
                 50 batches), eager and graph-replayed, with two products checked
                 bit for bit against the reference kernels' composition and that
                 composition timed on the host's CPUs;
  * cggi      -- CGGI16 functional bootstrap of 1024 LWE ciphertexts (n = 630)
                 into N = 1024 (bootstraps/s, every output decrypted);
  * serving   -- Poisson CKKS requests of mixed N and level (requests/s, p50/p99).
"""
from __future__ import annotations

import argparse
import ctypes
import json
import os
import sys
import threading
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)

N, L, POLYS = 1 << 16, 24, 64
HBM_PEAK_GBS = 8000.0
FP64_PEAK_TFLOPS = 78.6    # MI355X vector FP64
PROF_TAGS = 6


def log(*a):
    print("[bench]", *a, file=sys.stderr, flush=True)


def parse():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=5)
    ap.add_argument("--cpu-seconds", type=float, default=1.5, help="wall time of the CPU baseline sample")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-ckks", action="store_true")
    ap.add_argument("--ckks-batch", type=int, default=32)
    ap.add_argument("--ckks-steps", type=int, default=50)
    ap.add_argument("--no-cggi", action="store_true")
    ap.add_argument("--cggi-steps", type=int, default=3)
    ap.add_argument("--only-cggi", action="store_true", help="run the bootstrap leg alone (profiling)")
    ap.add_argument("--no-torus", action="store_true")
    ap.add_argument("--torus-steps", type=int, default=20)
    ap.add_argument("--only-torus", action="store_true", help="run the 32-bit-torus bootstrap leg alone (profiling)")
    ap.add_argument("--no-serving", action="store_true")
    ap.add_argument("--serving-seconds", type=float, default=1.5, help="wall time per offered load")
    ap.add_argument("--serving-lanes", type=int, default=7,
                    help="HIP streams the serving runtime uses (1..7; beyond 4 only while the queue is short)")
    ap.add_argument("--only-ntt", action="store_true",
                    help="run the batched NTT round trip alone (the command profiled with rocprofv3)")
    ap.add_argument("--only-ckks", action="store_true", help="run the CKKS leg alone (profiling)")
    ap.add_argument("--only-serving", action="store_true", help="run the serving leg alone (profiling)")
    ap.add_argument("--only-fused", action="store_true", help="run the fused product z = INTT(NTT(x) NTT(y)) alone (profiling)")
    ap.add_argument("--no-autotune", action="store_true", help="keep the default NTT schedule")
    ap.add_argument("--calibrate-copy", action="store_true",
                    help="also copy the 64-polynomial batch once per step (known-size copy for counter scaling)")
    return ap.parse_args()


class Prof:
    def __init__(self, lib):
        self.lib = lib

    def __enter__(self):
        self.lib.vfhe_hip_profile_start()
        return self

    def __exit__(self, *exc):
        ms = (ctypes.c_double * PROF_TAGS)()
        n = (ctypes.c_uint64 * PROF_TAGS)()
        self.lib.vfhe_hip_profile_stop(ms, n)
        self.ms = list(ms)
        self.launches = [int(x) for x in n]


def timed(lib, fn, steps, warmup):
    for _ in range(warmup):
        fn()
    lib.vfhe_hip_synchronize()
    t0 = time.perf_counter()
    for _ in range(steps):
        fn()
    lib.vfhe_hip_synchronize()
    return (time.perf_counter() - t0) / steps


def cpu_info() -> dict:
    """The host CPU as /proc/cpuinfo describes it: model, logical CPUs, physical cores, SMT."""
    model, cores, logical = "?", set(), 0
    try:
        phys = core = None
        with open("/proc/cpuinfo") as f:
            for line in f:
                k, _, v = line.partition(":")
                k, v = k.strip(), v.strip()
                if k == "processor":
                    logical += 1
                elif k == "model name":
                    model = v
                elif k == "physical id":
                    phys = v
                elif k == "core id":
                    core = v
                    cores.add((phys, core))
    except OSError:
        pass
    usable = len(os.sched_getaffinity(0)) if hasattr(os, "sched_getaffinity") else (os.cpu_count() or 1)
    ncores = len(cores) or logical
    return {"model": model, "logical_cpus": logical, "physical_cores": ncores,
            "smt": round(logical / ncores, 2) if ncores else None, "usable_cpus": usable}


def cpu_threads() -> int:
    return max(1, len(os.sched_getaffinity(0)) if hasattr(os, "sched_getaffinity") else (os.cpu_count() or 1))


def cpu_baseline(primes, seconds):
    """Reference ntt_forward + ntt_reverse on every CPU this process may run on (oracle/_ref)."""
    from oracle import oracle as O

    engines = O.ref_engines()
    if not engines:
        return None
    eng = engines[0]
    procs = [O.RefProc(eng, N, q) for q in primes]
    threads = cpu_threads()
    counts = [0] * threads
    rng = np.random.default_rng(5)
    xs = [rng.integers(0, q, N, dtype=np.uint64) for q in primes]
    stop = time.perf_counter() + seconds
    lib = procs[0].L

    def work(t):
        a, b = O.aligned(N), O.aligned(N)
        k = t
        while time.perf_counter() < stop:
            limb = k % L
            a[:] = xs[limb]
            lib.ntt_forward(O.ptr(b), O.ptr(a), procs[limb].h)
            lib.ntt_reverse(O.ptr(a), O.ptr(b), procs[limb].h)
            counts[t] += 1
            k += threads

    t0 = time.perf_counter()
    ths = [threading.Thread(target=work, args=(t,)) for t in range(threads)]
    for th in ths:
        th.start()
    for th in ths:
        th.join()
    wall = time.perf_counter() - t0
    limb_rt = sum(counts)
    return {
        "value": limb_rt / L / wall,
        "unit": "round-trips/s",
        "cores": threads,
        "kind": "reference",
        "cpu": cpu_info(),
        "sample": f"{limb_rt} limb round trips (ntt_forward+ntt_reverse, reference {eng} engine, N=2^16, "
                  f"the 24 ~50-bit primes) over {wall:.2f} s on {threads} threads; value = limb round trips / 24 / s",
    }


def main():
    args = parse()
    if args.gpus != 1:
        print("bench.py measures one GPU only: the engine runs on a single device and has no "
              "multi-GPU path; run with --gpus 1", file=sys.stderr)
        sys.exit(2)

    from vfhe_amd import _native, params, runtime as rt

    lib = _native.load()
    if args.only_ckks:
        print(json.dumps(bench_ckks(lib, _native, params, rt, args)), flush=True)
        return
    if args.only_serving:
        print(json.dumps(bench_serving(lib, args)), flush=True)
        return
    if args.only_cggi:
        print(json.dumps(bench_cggi(lib, args)), flush=True)
        return
    if args.only_torus:
        print(json.dumps(bench_cggi_torus(lib, args)), flush=True)
        return
    primes = params.gen_primes(N, [50] * L)
    log("device", lib.vfhe_hip_device_name().decode(), "primes", primes[0], "...", primes[-1])
    sched = {k: int(lib.vfhe_hip_get_option(k.encode())) for k in
             ("ntt.mode", "ntt.xcd_available", "ntt.xcd_lag", "ntt.xcd_occupancy", "ntt.chunk_mb", "ntt.tile_grid",
              "ntt.col_wide")}
    log("schedule", sched, "probe", [int(lib.vfhe_hip_get_option(f"probe.xcd{k}".encode())) for k in range(8)])
    pool = rt.Pool(N, primes)
    tuned = None
    if not (args.no_autotune or args.only_ntt or args.only_fused or args.only_ckks or args.only_serving or args.only_cggi):
        us = (ctypes.c_double * 9)()
        best = lib.vfhe_hip_autotune_ntt(pool.h, POLYS * L, 2, us)
        tuned = {"chosen": int(best), "us_per_round_trip": [round(x, 1) for x in us]}
        sched = {k: int(lib.vfhe_hip_get_option(k.encode())) for k in sched}
        log("autotune:", tuned, "->", sched)
    rng = np.random.default_rng(20261202)
    polys = []
    for _ in range(POLYS):
        polys.append(pool.poly(rt.random_rows(rng, primes, N)))
    arr = _native.voidp_array(polys)
    ref0 = pool.read(polys[0])

    if args.only_fused:
        ys = [pool.poly(rt.random_rows(rng, primes, N)) for _ in range(POLYS)]
        zs = [pool.poly() for _ in range(POLYS)]
        ax, ay, az = _native.voidp_array(polys), _native.voidp_array(ys), _native.voidp_array(zs)
        with Prof(lib) as fprof:
            fused_dt = timed(lib, lambda: lib.vfhe_hip_ntt_mul_intt_batch(az, ax, ay, POLYS), args.steps, args.warmup)
        graph_dt = timed(lib, lambda: lib.vfhe_hip_ntt_mul_intt_batch_graph(az, ax, ay, POLYS), args.steps, args.warmup)
        print(json.dumps({"fused_ms_per_batch": fused_dt * 1e3, "products_per_s": POLYS / fused_dt,
                          "graph_ms_per_batch": graph_dt * 1e3, "graph_products_per_s": POLYS / graph_dt,
                          "ntt_ms": fprof.ms[0], "launches": fprof.launches[0]}), flush=True)
        return

    def round_trip():
        lib.vfhe_hip_ntt_batch(arr, POLYS, 0)
        lib.vfhe_hip_ntt_batch(arr, POLYS, 1)

    log("round trip: warmup", args.warmup, "steps", args.steps)
    for _ in range(args.warmup):
        round_trip()
    lib.vfhe_hip_synchronize()
    with Prof(lib) as prof:
        t0 = time.perf_counter()
        for _ in range(args.steps):
            round_trip()
        lib.vfhe_hip_synchronize()
        dt = (time.perf_counter() - t0) / args.steps
    assert np.array_equal(pool.read(polys[0]), ref0), "round trip changed the data"

    def round_trip_graph():
        lib.vfhe_hip_ntt_batch_graph(arr, POLYS, 0)
        lib.vfhe_hip_ntt_batch_graph(arr, POLYS, 1)

    graph_dt = None
    if not args.only_ntt:   # (the profiled --only-ntt command runs the eager passes alone)
        graph_dt = timed(lib, round_trip_graph, args.steps, args.warmup)
        assert np.array_equal(pool.read(polys[0]), ref0), "graph round trip changed the data"
        log(f"round trip (graph replay): {graph_dt * 1e3:.3f} ms/step, {POLYS / graph_dt:.1f} round-trips/s")
    value = POLYS / dt
    ntt_ms, ntt_launches = prof.ms[0], prof.launches[0]
    avg_launch_s = ntt_ms / 1e3 / max(1, ntt_launches)
    # Algorithmic bytes of a step: every limb row is read and written once per direction
    # (16*N bytes per limb per transform); the step's pass launches share them evenly.
    step_bytes = POLYS * L * 2 * 16 * N
    bytes_per_launch = step_bytes * args.steps / max(1, ntt_launches)
    achieved = bytes_per_launch / avg_launch_s / 1e9
    # HBM bytes per launch from the committed FETCH_SIZE / WRITE_SIZE measurement: its ratio to
    # the algorithmic bytes of the measured launch, applied to this run's launch size (the
    # schedule may split the batch into sub-batches).
    traffic = None
    tpath = os.path.join(ROOT, "profiles", "ntt_pass_traffic.json")
    if os.path.exists(tpath):
        with open(tpath) as f:
            t = json.load(f)
        if t.get("hbm_bytes_per_launch") and t.get("algorithmic_bytes_per_launch"):
            traffic = int(round(t["hbm_bytes_per_launch"] / t["algorithmic_bytes_per_launch"] * bytes_per_launch))
    log(f"round trip: {dt * 1e3:.3f} ms/step, {value:.1f} round-trips/s, NTT launches {ntt_launches}, "
        f"avg {avg_launch_s * 1e3:.3f} ms, {achieved:.0f} GB/s algorithmic")

    if args.calibrate_copy:
        dst = [pool.poly() for _ in range(POLYS)]
        for _ in range(args.steps):
            for d, p in zip(dst, polys):
                lib.polynomial_copy_RNS_polynomial(d, p)
        lib.vfhe_hip_synchronize()
        log(f"copy calibration: {args.steps} x {POLYS} copies of {L * N * 8} bytes")
    if args.only_ntt:
        print(json.dumps({"schedule": sched, "round_trip_ms_per_step": dt * 1e3, "value": value, "ntt_launches": ntt_launches,
                          "avg_launch_ms": avg_launch_s * 1e3, "achieved_gbs": achieved}), flush=True)
        return

    # fused product z = INTT(NTT(x) * NTT(y))
    xs = polys[:POLYS]
    ys = [pool.poly(rt.random_rows(rng, primes, N)) for _ in range(POLYS)]
    zs = [pool.poly() for _ in range(POLYS)]
    ax, ay, az = _native.voidp_array(xs), _native.voidp_array(ys), _native.voidp_array(zs)
    fused_tuned = None
    if not args.no_autotune:
        fus = (ctypes.c_double * 6)()
        fbest = lib.vfhe_hip_autotune_product(az, ax, ay, POLYS, 3, fus)
        fused_tuned = {"candidate": int(fbest), "us": [round(u, 1) for u in fus]}
        log("fused product autotune:", fused_tuned)
    fused_dt = timed(lib, lambda: lib.vfhe_hip_ntt_mul_intt_batch(az, ax, ay, POLYS), args.steps, 2)
    log(f"fused product: {fused_dt * 1e3:.3f} ms per 64 products")
    # the same through a graph replay (one host launch per batch), checked against the eager output
    zg = [pool.poly() for _ in range(POLYS)]
    azg = _native.voidp_array(zg)
    fused_graph_dt = timed(lib, lambda: lib.vfhe_hip_ntt_mul_intt_batch_graph(azg, ax, ay, POLYS), args.steps, 2)
    fused_graph_equal = all(np.array_equal(pool.read(a), pool.read(b)) for a, b in zip(zg[:4], zs[:4]))
    log(f"fused product (graph): {fused_graph_dt * 1e3:.3f} ms per 64 products, equal: {fused_graph_equal}")

    # plain one-call-per-polynomial forms (device resident, and host resident)
    def plain_dev():
        for p in polys:
            lib.polynomial_RNSc_to_RNS(p, p)
            lib.vfhe_hip_synchronize()
            lib.polynomial_RNS_to_RNSc(p, p)
            lib.vfhe_hip_synchronize()

    plain_dev_dt = timed(lib, plain_dev, max(2, args.steps // 4), 1)
    host_polys = polys[:4]

    def plain_host():
        for p in host_polys:
            lib.polynomial_mark_host_dirty(p)
            lib.polynomial_RNSc_to_RNS(p, p)
            lib.polynomial_sync_to_host(p)
            lib.polynomial_mark_host_dirty(p)
            lib.polynomial_RNS_to_RNSc(p, p)
            lib.polynomial_sync_to_host(p)

    plain_host_dt = timed(lib, plain_host, 2, 1) / len(host_polys) * POLYS
    log(f"plain: device-resident {POLYS / plain_dev_dt:.1f} rt/s, host-resident {POLYS / plain_host_dt:.1f} rt/s")
    for p in ys + zs:
        pool.free(p)

    ckks = None
    if not args.no_ckks:
        ckks = bench_ckks(lib, _native, params, rt, args)

    cggi = None
    if not args.no_cggi:
        cggi = bench_cggi(lib, args)

    torus = None
    if not args.no_torus:
        torus = bench_cggi_torus(lib, args)

    serve = None
    if not args.no_serving:
        # the earlier legs' device blocks (ModUp buffers, bootstrap temporaries) go back to the
        # allocator, so the serving mix starts as it would in its own process
        lib.vfhe_hip_synchronize()
        lib.vfhe_hip_release_cached()
        serve = bench_serving(lib, args)

    # per-operation rates behind the choice of FP64 modular arithmetic (q < 2^50) over a
    # 64-bit integer Shoup product: each kind is 8 dependent-free chains per thread on
    # 8 workgroups of 256 threads per CU
    rates = (ctypes.c_double * 6)()
    lib.vfhe_hip_probe_alu_rates(rates)
    alu = {"unit": "Gop/s", "fp64_fma": round(rates[0] / 1e9, 1), "u32_mul_lo_add": round(rates[1] / 1e9, 1),
           "u32_mul_hi": round(rates[2] / 1e9, 1), "u32x32_to_64_mad": round(rates[3] / 1e9, 1),
           "modmul_fp64": round(rates[4] / 1e9, 1), "modmul_shoup_u64": round(rates[5] / 1e9, 1),
           "modmul_fp64_over_shoup": round(rates[4] / rates[5], 2)}
    log("alu rates (Gop/s):", alu)

    cpu = None
    if not args.no_cpu:
        log("cpu baseline:", args.cpu_seconds, "s")
        cpu = cpu_baseline(primes, args.cpu_seconds)

    line = {
        "metric": "NTT round-trips/s at N=2^16,L=24 (% HBM roofline); CKKS mults/s + p99 latency",
        "value": round(value, 2),
        "unit": "round-trips/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(dt * 1e3, 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "u64",
        "data": "synthetic",
        "config": {"workload": "batched NTT round trip, 64 polynomials x 24 limbs, N=2^16, ~50-bit primes "
                               "q=1 mod 2^17 (Ring.gen_prime), uniform residues, seed 20261202",
                   "N": N, "L": L, "polys": POLYS, "schedule": sched, "autotune": tuned},
        "roofline": {"bound": "hbm", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": round(achieved / HBM_PEAK_GBS, 4), "traffic": traffic,
                     "kernel": {2: "k_ntt16p (pipelined FP64 radix-16 column and row passes)",
                                0: "k_ntt16 (FP64 radix-16 column and row passes)",
                                1: "k_ntt_xcd (persistent per-XCD column and row passes)"}.get(sched["ntt.mode"], "?"),
                     "launches": ntt_launches,
                     "avg_launch_ms": round(avg_launch_s * 1e3, 4),
                     "algorithmic_bytes_per_launch": int(bytes_per_launch)},
        "cpu_baseline": cpu,
        "graph": {"round_trips_per_s": round(POLYS / graph_dt, 2), "ms_per_step": round(graph_dt * 1e3, 4),
                  "note": "the same batches through vfhe_hip_ntt_batch_graph (one graph replay per direction)"},
        "fused": {"unit": "products/s", "value": round(POLYS / fused_dt, 2), "ms_per_batch": round(fused_dt * 1e3, 4),
                  "graph": {"value": round(POLYS / fused_graph_dt, 2), "ms_per_batch": round(fused_graph_dt * 1e3, 4),
                            "equals_eager": fused_graph_equal},
                  "autotune": fused_tuned,
                  "workload": "z=INTT(NTT(x)*NTT(y)), 64 pairs x 24 limbs, N=2^16"},
        "plain": {"device_resident_round_trips_per_s": round(POLYS / plain_dev_dt, 2),
                  "host_resident_round_trips_per_s": round(POLYS / plain_host_dt, 2),
                  "batched_over_device_plain": round(plain_dev_dt / dt, 2),
                  "batched_over_host_plain": round(plain_host_dt / dt, 2)},
        "ckks": ckks,
        "cggi": cggi,
        "cggi_torus": torus,
        "serving": serve,
        "alu": alu,
    }
    print(json.dumps(line), flush=True)


def bench_ckks(lib, _native, params, rt, args):
    """CKKS multiply + relinearize + rescale, N=2^16, L=24 level limbs + 4 special limbs, r=1, batch of
    32 pairs: the eager batched entry point and its graph-replayed form; throughput with --ckks-steps
    batches issued back to back, and p50/p99 latency with each batch waited for; parity of two
    products against the reference kernels' composition, and the same composition timed on the
    host's CPUs."""
    K, B = 4, args.ckks_batch
    primes = params.gen_primes(N, [50] * (L + K))
    pool = rt.Pool(N, primes)
    lvl, ext = (1 << L) - 1, (1 << (L + K)) - 1
    rng = np.random.default_rng(99)
    log("ckks: building relinearization key (24 digits x 2 x 28 limbs)")
    kdata = np.stack([np.stack([rt.random_rows(rng, primes, N, ext) for _ in range(2)]) for _ in range(L)])
    entries = [pool.sample(1, ext, kdata[d]) for d in range(L)]
    key = _native.voidp_array(entries)
    rlk = _native.voidp_array([key, None])
    keep = []   # plaintext rows of the first two pairs, for the parity check
    in1, in2 = [], []
    for k in range(B):
        d1 = np.stack([rt.random_rows(rng, primes, N, lvl) for _ in range(2)])
        d2 = np.stack([rt.random_rows(rng, primes, N, lvl) for _ in range(2)])
        in1.append(pool.sample(1, lvl, d1))
        in2.append(pool.sample(1, lvl, d2))
        if k < 2:
            keep.append((d1[:, :L].copy(), d2[:, :L].copy()))
    outs = [pool.sample(1, ext) for _ in range(B)]
    a1, a2, ao = _native.voidp_array(in1), _native.voidp_array(in2), _native.voidp_array(outs)
    top = 1 << (L - 1)

    def run(fn, steps):
        lat, enq = [], []
        for _ in range(steps):
            for c in outs:
                pool.set_sample_mask(c, ext)
            t0 = time.perf_counter()
            fn(ao, a1, a2, B, rlk, top)
            t1 = time.perf_counter()
            lib.vfhe_hip_synchronize()
            lat.append(time.perf_counter() - t0)
            enq.append(t1 - t0)
        return np.array(lat) * 1e3, np.array(enq) * 1e3

    def pipelined(fn, steps):
        """mults/s with the batches issued back to back on the lane (one wait at the end): the
        host's enqueue of batch i+1 overlaps batch i on the device, as in a serving loop."""
        lib.vfhe_hip_synchronize()
        t0 = time.perf_counter()
        for _ in range(steps):
            for c in outs:
                pool.set_sample_mask(c, ext)
            fn(ao, a1, a2, B, rlk, top)
        lib.vfhe_hip_synchronize()
        return B * steps / (time.perf_counter() - t0)

    # the multiply's schedule (column-pass width, inverse-pass grid, pass-through fold), timed on
    # these operands by the engine's tuner (every candidate is bit-exact)
    tuned = None
    if not args.no_autotune:
        us = (ctypes.c_double * 12)()
        best = lib.vfhe_hip_autotune_multiply(ao, a1, a2, B, rlk, top, 2, us)
        tuned = {"candidate": int(best), "us": [round(u, 1) for u in us],
                 "ntt.col_wide": lib.vfhe_hip_get_option(b"ntt.col_wide"),
                 "ntt.tile_grid": lib.vfhe_hip_get_option(b"ntt.tile_grid"),
                 "ks.fold": lib.vfhe_hip_get_option(b"ks.fold")}
        log("ckks autotune:", tuned)
    run(lib.vfhe_hip_mlwe_multiply_batch, 2)
    with Prof(lib) as prof:
        lat, enq = run(lib.vfhe_hip_mlwe_multiply_batch, args.ckks_steps)
    piped = pipelined(lib.vfhe_hip_mlwe_multiply_batch, args.ckks_steps)
    # parity of the eager result: two products against the reference kernels (oracle/_ref)
    parity = None
    cpu = None
    try:
        from oracle import oracle as O
        engines = O.ref_engines()
    except Exception:
        engines = []
    got = [pool.read_sample(outs[k]) for k in range(len(keep))]
    if engines:
        ct1 = np.stack([d1 for d1, _ in keep])
        ct2 = np.stack([d2 for _, d2 in keep])
        ref = O.ref_ckks_multiply(engines[0], N, L, K, primes, ct1, ct2, kdata, threads=2)
        same = all(np.array_equal(got[k][:, : L - 1], ref[k]) for k in range(len(keep)))
        parity = f"{'bit-exact' if same else 'MISMATCH'}: {len(keep)} products vs the reference {engines[0]} " \
                 f"kernels composed in the reference's call order (oracle/ref_ckks.c)"
        log("ckks parity:", parity)
        if not args.no_cpu:
            T = cpu_threads()
            log(f"ckks cpu baseline: {T} products on {T} threads")
            t0 = time.perf_counter()
            O.ref_ckks_multiply(engines[0], N, L, K, primes, ct1, ct2, kdata, threads=T, count=T)
            wall = time.perf_counter() - t0
            cpu = {"value": round(T / wall, 3), "unit": "mults/s", "cores": T, "kind": "reference",
                   "cpu": cpu_info(),
                   "sample": f"{T} multiply+relinearize+rescale products at N=2^16, L=24+4 (one per thread, "
                             f"reference {engines[0]} kernels in the reference's call order, oracle/ref_ckks.c) "
                             f"in {wall:.2f} s"}
            log("ckks cpu baseline:", cpu)
    # graph-replayed form: the first call captures the batch's launch sequence
    g0 = {k: int(lib.vfhe_hip_get_option(f"stat.graph_{k}".encode())) for k in ("captures", "replays", "fallbacks")}
    run(lib.vfhe_hip_mlwe_multiply_batch_graph, 2)
    glat, genq = run(lib.vfhe_hip_mlwe_multiply_batch_graph, args.ckks_steps)
    gpiped = pipelined(lib.vfhe_hip_mlwe_multiply_batch_graph, args.ckks_steps)
    g1 = {k: int(lib.vfhe_hip_get_option(f"stat.graph_{k}".encode())) - g0[k] for k in g0}
    ggot = [pool.read_sample(outs[k]) for k in range(len(keep))]
    graph_same = all(np.array_equal(ggot[k], got[k]) for k in range(len(keep)))
    lib.vfhe_hip_graph_clear()
    one = []
    o1 = _native.voidp_array(outs[:1])
    for _ in range(3):
        pool.set_sample_mask(outs[0], ext)
        t0 = time.perf_counter()
        lib.vfhe_hip_mlwe_multiply_batch(o1, _native.voidp_array(in1[:1]), _native.voidp_array(in2[:1]), 1, rlk, top)
        lib.vfhe_hip_synchronize()
        one.append(time.perf_counter() - t0)
    names = ["ntt_fp64", "ntt_int", "eltwise", "division", "keyswitch_mac", "other"]
    res = {
        "unit": "mults/s",
        "value": round(gpiped, 2),
        "value_form": "graph-replayed (the batch's launch sequence captured once, replayed per batch)",
        "eager": {"value": round(piped, 2), "value_one_batch_at_a_time": round(B / float(np.median(lat) / 1e3), 2)},
        "batch": B,
        "batches_timed": int(args.ckks_steps),
        "batch_ms_p50": round(float(np.percentile(lat, 50)), 3),
        "batch_ms_p99": round(float(np.percentile(lat, 99)), 3),
        "request_latency_p99_ms": round(float(np.percentile(lat, 99)), 3),
        "host_enqueue_ms_p50": round(float(np.percentile(enq, 50)), 3),
        "graph": {"value": round(gpiped, 2), "value_one_batch_at_a_time": round(B / float(np.median(glat) / 1e3), 2),
                  "batch_ms_p50": round(float(np.percentile(glat, 50)), 3),
                  "batch_ms_p99": round(float(np.percentile(glat, 99)), 3),
                  "host_enqueue_ms_p50": round(float(np.percentile(genq, 50)), 3), "counters": g1,
                  "equals_eager": bool(graph_same)},
        "unbatched_mults_per_s": round(1.0 / float(np.median(one)), 2),
        "kernel_ms_per_batch": {n: round(m / args.ckks_steps, 3) for n, m in zip(names, prof.ms)},
        "parity": parity,
        "autotune": tuned,
        "cpu_baseline": cpu,
        "workload": "multiply+relinearize+rescale, N=2^16, L=24+4 special, rank 1, batch of 32 pairs, "
                    "random rlk (throughput does not depend on key validity); value: the graph-replayed form "
                    "with batches issued back to back; batch_ms_*: the eager form, each batch waited for",
    }
    log("ckks:", res)
    for c in in1 + in2 + outs + entries:
        lib.free_mlwe_RNS_sample(c)
    return res


def bench_cggi(lib, args):
    """CGGI16 functional bootstrap of 1024 LWE ciphertexts (n = 630, binary key) into N = 1024 (rank 1,
    two 50-bit level primes + one 50-bit special prime, the reference's RNS variant), 4-bit messages,
    identity table: bootstraps/s over a stream of batches (two accumulator batches used in turn, the
    host computing batch i+1's rotation amounts while batch i runs, each batch extracted) and one
    batch at a time, every output of the first batch decrypted."""
    from vfhe_amd.arith import Ring
    from vfhe_amd.cggi import BootstrapKey, Bootstrapper, LweBatch, LweKey, extract, lwe_key_of, pack_lut
    from vfhe_amd.fhe import Chain, SecretKey

    prec, n, Nc, count = 5, 630, 1024, 1024
    lib.vfhe_prng_set_deterministic_seed(630)
    rng = np.random.default_rng(630)
    Rq = Ring(Nc, prime_size=[50, 50, 50], split_degree=1)
    in_ring = Rq.quotient_ring(ell=2)
    lwe_key = LweKey(in_ring, [int(b) for b in rng.integers(0, 2, n)])
    chain = Chain.nested(Rq, special=1, levels=1)
    sk = [0] * Nc
    for k in rng.choice(Nc, size=64, replace=False):
        sk[int(k)] = 1
    out_key = SecretKey(chain, [sk], 3.2)
    t0 = time.perf_counter()
    bk = BootstrapKey(chain, lwe_key, out_key)
    keygen_s = time.perf_counter() - t0
    tv = pack_lut(chain, list(range(1 << (prec - 1))), 1 << (prec - 1), prec)
    q = in_ring.q_l
    msgs = [int(v) for v in rng.integers(0, 16, count)]
    lwes = LweBatch.encrypt(lwe_key, [round((v * q) / (1 << prec)) % q for v in msgs])
    boot = Bootstrapper(chain, bk)
    out_key_lwe = lwe_key_of(out_key, in_ring)
    times, rot_times, errors = [], [], 0
    tb = 1 << (prec - 1)
    for rep in range(args.cggi_steps + 1):
        lib.vfhe_hip_synchronize()
        t0 = time.perf_counter()
        acc = boot.blind_rotate(tv, lwes, tb)
        lib.vfhe_hip_synchronize()
        t1 = time.perf_counter()
        out = extract(acc, 0)
        t2 = time.perf_counter()
        if rep == 0:
            got = [round((ph * (1 << prec)) / q) % (1 << prec) for ph in out.phases(out_key_lwe)]
            errors = sum(1 for a, b in zip(got, msgs) if a != b)
            continue
        rot_times.append(t1 - t0)
        times.append(t2 - t0)
        del acc, out
    t = float(np.median(times))
    # a stream of batches into two accumulator batches used in turn (as a serving loop keeps its
    # buffers): while batch i runs on the device the host computes batch i+1's rotation amounts,
    # then reads batch i back
    steps_p = max(2, args.cggi_steps)
    accs = [boot.alloc_accumulators(count) for _ in range(2)]
    lib.vfhe_hip_synchronize()
    t0 = time.perf_counter()
    nxt = boot.amounts(lwes, tb)
    for i in range(steps_p):
        boot.blind_rotate(tv, lwes, tb, amounts=nxt, acc=accs[i & 1])
        nxt = boot.amounts(lwes, tb)
        out = extract(accs[i & 1], 0)
        del out
    lib.vfhe_hip_synchronize()
    t_stream = (time.perf_counter() - t0) / steps_p
    del accs
    # where a batch's time goes: the host's rotation amounts, the device's kernels, the extraction
    t0 = time.perf_counter()
    boot.amounts(lwes, 1 << (prec - 1))
    amounts_ms = (time.perf_counter() - t0) * 1e3
    with Prof(lib) as prof:
        acc = boot.blind_rotate(tv, lwes, 1 << (prec - 1))
        lib.vfhe_hip_synchronize()
    t0 = time.perf_counter()
    extract(acc, 0)
    extract_ms = (time.perf_counter() - t0) * 1e3
    del acc
    res = {"unit": "bootstraps/s", "value": round(count / t_stream, 2), "ms_per_batch": round(t_stream * 1e3, 2),
           "one_batch_at_a_time": {"value": round(count / t, 2), "ms_per_batch": round(t * 1e3, 2)},
           "blind_rotation_ms_per_batch": round(float(np.median(rot_times)) * 1e3, 2), "batch": count,
           "split_ms": {"host_rotation_amounts": round(amounts_ms, 2), "device_kernels": round(sum(prof.ms), 2),
                        "extraction": round(extract_ms, 2)},
           "decrypt_errors": errors, "bootstrap_key_s": round(keygen_s, 2),
           "workload": "CGGI16 functional bootstrap (RNS variant of the reference, cggi16.py:46-140): LWE n=630 "
                       "(binary key), GLWE N=1024 rank 1 over two 50-bit primes + one special, gadget = one digit "
                       "per limb (ell=2), 1024 LWE ciphertexts of uniform 4-bit messages, identity table; the "
                       "32-bit-torus FFT variant has no counterpart in the reference"}
    log("cggi:", res)
    return res


def torus_flops_per_index(p) -> int:
    """FP64 operations of one sample at one key index in k_torus_blind_rotate, counted from
    its code (an FMA counts two; the pass-1 twiddles are formed once per kernel, outside the
    loop).  Per transform, per 8 points (one lane): three DFT-8 (56 each), the seven slot
    factors, pass 1's eight and pass 2's seven twiddle products (complex products of 6): 300.
    Per key row and output component one complex multiply-add per point (8); per output
    component the partial swap's complex add and the rounding adds (2 + 2 per point)."""
    M = p.N // 2
    fft = M // 8 * (3 * 56 + (7 + 8 + 7) * 6)
    rows = 2 * p.ell
    return (rows + 2) * fft + rows * 2 * 8 * M + 2 * (2 + 2) * M


def bench_cggi_torus(lib, args):
    """CGGI over the 32-bit torus with an FP64 FFT: 1024 LWE samples (n = 630) of uniform 4-bit
    messages through the identity table into N = 1024, base 2^7, 3 levels; inputs, test vector
    and outputs resident on the device; bootstraps/s over --torus-steps batches."""
    from vfhe_amd import torus as VT

    prec, count = 5, 1024
    p = VT.TorusParams()
    t0 = time.perf_counter()
    keys = VT.TorusKeys(p, seed=630)
    bsk = keys.bootstrap_key()
    boot = VT.TorusBootstrapper(bsk, p)
    keygen_s = time.perf_counter() - t0
    msgs = np.random.default_rng(631).integers(0, 16, size=count)
    samples = keys.encrypt(msgs, prec)
    tv_h = VT.pack_lut(list(range(16)), 16, prec, p.N)
    lw, tv = VT.TorusBatch.upload(samples), VT.TorusBatch.upload(tv_h)
    out = VT.TorusBatch.zeros(count, p.N + 1)
    boot.bootstrap(lw, tv, 1 << (prec - 1), out=out)
    errors = int((keys.decrypt(out.read(), prec) != msgs).sum())

    def step():
        boot.bootstrap(lw, tv, 1 << (prec - 1), out=out)

    t = timed(lib, step, args.torus_steps, 2)
    with Prof(lib) as prof:
        for _ in range(5):
            step()
        lib.vfhe_hip_synchronize()
    kernel_ms = prof.ms[5] / max(prof.launches[5], 1)
    flops = torus_flops_per_index(p) * p.n * count
    tflops = flops / (kernel_ms * 1e-3) / 1e12
    res = {"unit": "bootstraps/s", "value": round(count / t, 2), "ms_per_batch": round(t * 1e3, 3),
           "kernel_ms": round(kernel_ms, 3), "batch": count, "decrypt_errors": errors,
           "fp64": {"achieved": round(tflops, 2), "peak": FP64_PEAK_TFLOPS, "unit": "TFLOP/s",
                    "frac": round(tflops / FP64_PEAK_TFLOPS, 4), "flops_per_batch": int(flops),
                    "kernel": "k_torus_blind_rotate"},
           "bootstrap_key_s": round(keygen_s, 2),
           "workload": "CGGI functional bootstrap over the 32-bit torus, FP64 negacyclic FFT of N/2 points: "
                       "LWE n=630 (binary key, sigma 2^-15), GLWE N=1024 rank 1 (binary key, sigma 2^-25), "
                       "gadget base 2^7 x 3 levels, 1024 samples of uniform 4-bit messages, identity table; "
                       "no counterpart in the reference (its CGGI16 is the RNS variant)"}
    if not args.no_cpu:
        # the same bootstrap in C (oracle/torus_port.c, checked bit for bit against oracle/torus.py)
        # on every host thread, over a growing prefix of the batch until it runs a third of
        # --cpu-seconds
        from oracle import torus as OT
        threads = cpu_threads()
        port = OT.PortKey(bsk, p.ell, p.base_log)   # the key's transform, outside the timing
        k, dt = threads, 0.0
        while True:
            t0 = time.perf_counter()
            port.bootstrap(samples[:k], tv_h, 1 << (prec - 1), threads=threads)
            dt = time.perf_counter() - t0
            if dt >= args.cpu_seconds / 3 or k >= count:
                break
            k = min(count, 2 * k)
        res["cpu_baseline"] = {"value": round(k / dt, 2), "unit": "bootstraps/s", "cores": threads, "kind": "port",
                               "cpu": cpu_info(),
                               "sample": f"{k} samples of the batch through oracle/torus_port.c (C, double FFT, "
                                         f"{threads} threads)"}
    log("cggi torus:", res)
    lw.free(), tv.free(), out.free(), boot.free()
    return res


def bench_serving(lib, args):
    """Serving loop: Poisson CKKS requests (multiply+relinearize+rescale or rotation), N in 2^13..2^16,
    level uniform in 4..24, keys and ciphertexts resident in HBM, continuous batching on 4 streams."""
    from vfhe_amd import serving
    Ns = [1 << 13, 1 << 14, 1 << 15, 1 << 16]
    log("serving: building resident keys and ciphertexts for N =", Ns)
    t0 = time.perf_counter()
    rings = {n: serving.Ring(lib, n, L=24, K=4, rot_steps=4, levels=(4, 24)) for n in Ns}
    log(f"serving: setup {time.perf_counter() - t0:.1f} s")
    budget, lanes, max_batch = 32 << 30, args.serving_lanes, 32
    burst = serving.make_workload(40000.0, 0.2, Ns, seed=1)
    sat = serving.run(lib, rings, burst, budget, lanes, max_batch)
    cap = sat["requests_per_s"]
    log(f"serving: saturated throughput {cap:.0f} requests/s ({len(burst)} requests)", sat)
    loads = []
    for k, frac in enumerate((0.25, 0.5, 0.75, 0.95)):
        wl = serving.make_workload(frac * cap, args.serving_seconds, Ns, seed=10 + k)
        r = serving.run(lib, rings, wl, budget, lanes, max_batch)
        r["offered_fraction"] = frac
        log("serving:", r)
        loads.append(r)
    lib.vfhe_hip_synchronize()
    for ring in rings.values():
        ring.free()
    return {"unit": "requests/s", "saturated_requests_per_s": cap, "saturated": sat, "loads": loads, "lanes": lanes,
            "max_batch": max_batch, "memory_budget_bytes": budget,
            "workload": "Poisson arrivals; 50% multiply+relinearize+rescale, 50% rotation by 5^(2^k), k<4; "
                        "N uniform in {2^13..2^16}, level uniform in [4,24] limbs + 4 special limbs; "
                        "random evaluation keys and ciphertexts resident in HBM"}


if __name__ == "__main__":
    main()
