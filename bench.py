#!/usr/bin/env python3
"""Ajtai commitment throughput on one MI355X (d = 64, 32-bit q, kappa = 16, n = 2^15).

One step = one plain drop-in commitment t = A*s through gh_commit with device
buffers already resident in HBM: twisted NTT of the 2^15 witness polynomials,
the streaming MAC against the 134 MB NTT-domain key, and the inverse transform
of the 16 output polynomials. Three resident keys are used in rotation so that
A (134 MB each, 402 MB together) is read from HBM and not from the 256 MB
Infinity Cache; the single-key (cache-warm) rate is reported beside it.

Prints one JSON line. Extra fields: p50/p99 latency of individually
synchronised commits, the roofline of the dominant kernel (the MAC, timed with
HIP events on the library stream), the CPU baseline (the reference's AVX2
prepared-commit path restated in C, run on this host), and the batched
negacyclic-product rate (2^18 pairs).
"""
import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

Q = 4294966657
D = 64
KAPPA = 16
N = 1 << 15
HBM_PEAK_GBS = 8000.0
# algorithmic bytes per commitment: A_ntt (kappa*n*d*4) + s (n*d*4) + t (kappa*d*4)
ALGO_BYTES = KAPPA * N * D * 4 + N * D * 4 + KAPPA * D * 4


def witness(rng, n):
    x = rng.integers(-128, 128, size=(n, D), dtype=np.int64)
    return np.where(x < 0, Q + x, x).astype(np.uint32)


def batched_ns(ctx, keys, s_dev, log, batch=32, reps=6):
    """The batched form at the same shape: `batch` commitments per gh_commit_batch
    call against one key (the key is loaded once per group of witnesses), keys
    in rotation across calls. Item 0 is checked against the plain form."""
    outs = [ctx.alloc(KAPPA * D * 4) for _ in range(batch)]
    wit = [s_dev[i % len(s_dev)] for i in range(batch)]
    lens = [N] * batch
    keys[0].commit_batch_device(wit, lens, outs)
    probe = ctx.alloc(KAPPA * D * 4)
    keys[0].commit_device(s_dev[0], N, probe)
    ctx.synchronize()
    same = bool(np.array_equal(outs[0].download(np.uint32, (KAPPA, D)),
                               probe.download(np.uint32, (KAPPA, D))))
    a = time.perf_counter()
    for r in range(reps):
        keys[r % len(keys)].commit_batch_device(wit, lens, outs)
    ctx.synchronize()
    dt = time.perf_counter() - a
    rate = batch * reps / dt
    log(f"batched ({batch} per call, same shape): {rate:,.0f} commits/s; item 0 == plain: {same}")
    for b in outs + [probe]:
        b.free()
    return {"commits_per_s": round(rate, 1), "batch": batch, "us_per_batch": round(dt / reps * 1e6, 1),
            "item0_equals_plain": same}


def c4_goldilocks(ctx, la, log, count=2000, kappa=16, max_cols=1 << 14):
    """2000 Goldilocks commitments (kappa = 16, d = 64) with log-uniform witness
    lengths in [2^8, 2^14] against one shared key: per-call form (one
    gh_commit_batch call per commitment) versus one batched call."""
    p = la.GOLDILOCKS
    rng = np.random.default_rng(44)
    ring = la.Ring(ctx, p, D)
    hi = rng.integers(0, 1 << 32, size=(kappa, max_cols, D), dtype=np.uint64)
    lo = rng.integers(0, 1 << 32, size=(kappa, max_cols, D), dtype=np.uint64)
    a = ((hi << np.uint64(32)) | lo) % np.uint64(p)
    del hi, lo
    key = la.Key(ring, a, domain=la.DOMAIN_NTT)
    pool_polys = 1 << 15
    t = rng.integers(-1, 2, size=(pool_polys, D), dtype=np.int64)
    pool = np.where(t < 0, np.uint64(p - 1), t.astype(np.uint64))
    pool_dev = ctx.to_device(pool)
    lens = np.floor(2.0 ** rng.uniform(8, 14, size=count)).astype(np.int64)
    offs = rng.integers(0, pool_polys - lens + 1)
    wptr = [pool_dev.ptr.value + int(o) * D * 8 for o in offs]
    out = ctx.alloc(count * kappa * D * 8)
    optr = [out.ptr.value + i * kappa * D * 8 for i in range(count)]
    # warm-up and parity spot check (two short items against the oracle)
    key.commit_batch_device(wptr, lens, optr)
    ctx.synchronize()
    res = out.download(np.uint64, (count, kappa, D))
    from oracle import gridref
    ok = True
    for i in np.argsort(lens)[:2]:
        n, o = int(lens[i]), int(offs[i])
        z = np.zeros((kappa, 1, D), dtype=np.uint64)
        exp = gridref.commit(p, D, a[:, :n], z, pool[o:o + n], np.zeros((1, D), dtype=np.uint64),
                             key_domain=1)
        ok = ok and bool(np.array_equal(res[i], exp))
    reps = 3
    t0 = time.perf_counter()
    for _ in range(reps):
        key.commit_batch_device(wptr, lens, optr)
    ctx.synchronize()
    batched = count * reps / (time.perf_counter() - t0)
    t0 = time.perf_counter()
    for i in range(count):
        key.commit_batch_device([wptr[i]], [lens[i]], [optr[i]])
    ctx.synchronize()
    plain = count / (time.perf_counter() - t0)
    log(f"C4 Goldilocks x{count}: batched {batched:,.0f} commits/s, per-call {plain:,.0f} "
        f"commits/s; oracle spot check {ok}")
    key.close()
    pool_dev.free()
    out.free()
    ring.close()
    return {"batched_commits_per_s": round(batched, 1), "per_call_commits_per_s": round(plain, 1),
            "mean_len": float(lens.mean()), "oracle_spot_check": ok}


def main():
    ap = argparse.ArgumentParser(description=__doc__)
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=400)
    ap.add_argument("--warmup", type=int, default=40)
    ap.add_argument("--keys", type=int, default=3, help="resident keys used in rotation")
    ap.add_argument("--cpu-seconds", type=float, default=10.0)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-extras", action="store_true")
    ap.add_argument("--no-check", action="store_true")
    args = ap.parse_args()
    world = int(os.environ.get("WORLD_SIZE", "1"))
    if args.gpus != 1 or world != 1:
        print(f"bench.py: this build measures one GPU only (asked for --gpus {args.gpus}, "
              f"WORLD_SIZE={world}); the multi-GPU sharding is described, not run.",
              file=sys.stderr, flush=True)
        sys.exit(2)

    import lattica_amd as la

    log = lambda *a: print("[bench]", *a, file=sys.stderr, flush=True)  # noqa: E731
    rng = np.random.default_rng(20260517)
    ctx = la.Context(0)
    ring = la.Ring(ctx, Q, D)
    log(f"device 0: {ctx.num_cus} CUs; building {args.keys} resident keys "
        f"({KAPPA}x{N} polys, {KAPPA * N * D * 4 / 2**20:.0f} MiB each)")
    keys, a_host0 = [], None
    for k in range(args.keys):
        a = rng.integers(0, Q, size=(KAPPA, N, D), dtype=np.uint32)  # uniform NTT-domain key
        keys.append(la.Key(ring, a, domain=la.DOMAIN_NTT))
        if k == 0:
            a_host0 = a
        else:
            del a
    nw = 8
    s_host = [witness(rng, N) for _ in range(nw)]
    s_dev = [ctx.to_device(s) for s in s_host]
    t_dev = [ctx.alloc(KAPPA * D * 4) for _ in range(nw)]

    def step(i, key_idx=None):
        key = keys[i % len(keys) if key_idx is None else key_idx]
        key.commit_device(s_dev[i % nw], N, t_dev[i % nw])

    for i in range(args.warmup):
        step(i)
    ctx.synchronize()

    if not args.no_check:
        from oracle import gridref
        step(0, 0)
        ctx.synchronize()
        got = t_dev[0].download(np.uint32, (KAPPA, D)).astype(np.uint64)
        z = np.zeros((KAPPA, 1, D), dtype=np.uint64)
        exp = gridref.commit(Q, D, a_host0.astype(np.uint64), z, s_host[0].astype(np.uint64),
                             np.zeros((1, D), dtype=np.uint64), key_domain=1)
        if not np.array_equal(got, exp):
            log("PARITY FAILURE: GPU commitment differs from the oracle")
            sys.exit(1)
        log("parity: commitment 0 bit-exact vs oracle")

    # ---- timed region: plain per-call commitments, keys in rotation ----
    ctx.synchronize()
    t0 = time.perf_counter()
    for i in range(args.steps):
        step(i)
    ctx.synchronize()
    elapsed = time.perf_counter() - t0
    value = args.steps / elapsed
    log(f"plain: {value:,.0f} commits/s ({1e6 / value:.1f} us/commit)")

    tuning = ctx.get_tuning()
    if tuning["fused"]:
        mac_kernel = "k_commit_fused_q32"  # witness transform + MAC in one kernel
    else:
        mac_kernel = "k_mac4_q32" if tuning["mac_variant"] == 1 else "k_mac_q32"
    # HBM bytes per dispatch of that kernel from the committed rocprofv3 --pmc summary
    traffic = None
    tpath = os.path.join(ROOT, "profiles", "traffic_commit_q32_k16_n32768.json")
    if os.path.exists(tpath):
        with open(tpath) as fh:
            tdoc = json.load(fh)
        if tdoc.get("kernel", "").startswith(mac_kernel):
            traffic = tdoc["traffic_bytes_per_dispatch"]
    # ---- roofline: MAC kernel duration from HIP events on the library stream ----
    ctx.profile(True)
    for i in range(args.steps):
        step(i)
    prof = ctx.profile_read()
    ctx.profile(False)
    mac_ms = prof[1][0] / max(prof[1][1], 1)
    ntt_ms = prof[0][0] / max(prof[0][1], 1)
    fin_ms = prof[2][0] / max(prof[2][1], 1)
    achieved = ALGO_BYTES / (mac_ms * 1e-3) / 1e9
    log(f"phases (avg): witness NTT {ntt_ms * 1e3:.1f} us, MAC {mac_ms * 1e3:.1f} us, "
        f"finalize {fin_ms * 1e3:.1f} us; MAC {achieved:,.0f} GB/s algorithmic")

    # ---- latency: individually synchronised commits ----
    lat = []
    for i in range(max(50, args.steps // 4)):
        a = time.perf_counter()
        step(i)
        ctx.synchronize()
        lat.append((time.perf_counter() - a) * 1e6)
    lat = np.array(lat)

    extra = {}
    if not args.no_extras:
        # cache-warm rate: one key only (A stays in the Infinity Cache)
        ctx.synchronize()
        a = time.perf_counter()
        for i in range(args.steps):
            step(i, 0)
        ctx.synchronize()
        extra["single_key_commits_per_s"] = round(args.steps / (time.perf_counter() - a), 1)
        # measured HBM ceiling: streaming read / copy over 1 GiB (beyond the Infinity Cache)
        rd, cp = ctx.measure_bandwidth(1 << 30, 5)
        extra["hbm_ceiling_GBps"] = {"stream_read": round(rd, 1), "stream_copy": round(cp, 1)}
        log(f"measured HBM ceiling: read {rd:,.0f} GB/s, copy {cp:,.0f} GB/s")
        # batched negacyclic product, 2^18 pairs (device buffers)
        pairs = 1 << 18
        pa = ctx.to_device(rng.integers(0, Q, size=(pairs, D), dtype=np.uint32))
        pb = ctx.to_device(rng.integers(0, Q, size=(pairs, D), dtype=np.uint32))
        po = ctx.alloc(pairs * D * 4)
        L = la.lib()
        for _ in range(3):
            L.gh_negacyclic_mul(ring.h, pa.ptr, pb.ptr, po.ptr, pairs, la.MEM_DEVICE)
        ctx.synchronize()
        reps = 20
        a = time.perf_counter()
        for _ in range(reps):
            L.gh_negacyclic_mul(ring.h, pa.ptr, pb.ptr, po.ptr, pairs, la.MEM_DEVICE)
        ctx.synchronize()
        dt = (time.perf_counter() - a) / reps
        extra["negacyclic_2^18_pairs"] = {
            "pairs_per_s": round(pairs / dt, 1), "us_per_batch": round(dt * 1e6, 1),
            "algorithmic_GBps": round(3 * pairs * D * 4 / dt / 1e9, 1)}
        for b in (pa, pb, po):
            b.free()
        extra["batched_same_shape"] = batched_ns(ctx, keys, s_dev, log)
        extra["c4_goldilocks_2000"] = c4_goldilocks(ctx, la, log)

    cpu = None
    if not args.no_cpu:
        from oracle import gridref
        threads_all = min(16, os.cpu_count() or 1)
        c = gridref.Avx2Commit(Q, 64, a_host0.astype(np.uint64), D)
        c.load(s_host[0].astype(np.uint64))
        c.run(1)  # warm-up
        n_done, a = 0, time.perf_counter()
        while True:
            c.run(1)
            n_done += 1
            if time.perf_counter() - a >= args.cpu_seconds:
                break
        cpu_rate = n_done / (time.perf_counter() - a)
        c.run(threads_all)
        n_all, a = 0, time.perf_counter()
        while time.perf_counter() - a < args.cpu_seconds / 2:
            c.run(threads_all)
            n_all += 1
        cpu_all = n_all / (time.perf_counter() - a)
        c.close()
        cpu = {"value": round(cpu_rate, 3), "unit": "commits/s", "cores": 1, "kind": "port",
               "sample": f"{n_done} prepared commits (kappa=16, n=2^15, d=64, q32) of the reference's "
                         f"AVX2 L=u64 path restated in C (oracle/gridref_avx2.c), 1 thread",
               "all_cores": {"value": round(cpu_all, 3), "threads": threads_all,
                             "note": "this build's row/column split over threads"}}
        log(f"cpu baseline: {cpu_rate:.2f} commits/s on 1 core, {cpu_all:.2f} on {threads_all}")

    out = {
        "metric": "Ajtai commits/sec (d=64)",
        "value": round(value, 1),
        "unit": "commits/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(1e3 * elapsed / args.steps, 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "u32",
        "data": "synthetic: seeded uniform NTT-domain key mod q, witness uniform in [-2^7, 2^7)",
        "config": {"workload": "ajtai_commit_q32_d64_kappa16_n32768", "q": Q, "d": D,
                   "kappa": KAPPA, "n": N, "form": "plain per-call gh_commit, resident buffers",
                   "keys_in_rotation": args.keys},
        "latency_us": {"p50": round(float(np.percentile(lat, 50)), 1),
                       "p99": round(float(np.percentile(lat, 99)), 1)},
        "roofline": {"bound": "hbm", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS,
                     "unit": "GB/s", "frac": round(achieved / HBM_PEAK_GBS, 4), "traffic": traffic,
                     "kernel": mac_kernel, "kernel_us": round(mac_ms * 1e3, 2),
                     "algorithmic_bytes": ALGO_BYTES,
                     "frac_of_measured_read_ceiling":
                         round(achieved / extra["hbm_ceiling_GBps"]["stream_read"], 4)
                         if "hbm_ceiling_GBps" in extra else None},
        "tuning": tuning,
        "phases_us": {"witness_ntt": round(ntt_ms * 1e3, 2), "mac": round(mac_ms * 1e3, 2),
                      "finalize": round(fin_ms * 1e3, 2)},
        "cpu_baseline": cpu,
        "extra": extra,
    }
    print(json.dumps(out), flush=True)
    for k in keys:
        k.close()
    for b in s_dev + t_dev:
        b.free()
    ring.close()
    ctx.close()


if __name__ == "__main__":
    main()
