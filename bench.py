#!/usr/bin/env python3
"""Ajtai commitment throughput on one MI355X (d = 64, 32-bit q, kappa = 16, n = 2^15).

One step = one plain drop-in commitment t = A*s through gh_commit with device
buffers already resident in HBM: twisted NTT of the 2^15 witness polynomials,
the streaming MAC against the 134 MB NTT-domain key, and the inverse transform
of the 16 output polynomials. Three resident keys are used in rotation so that
A (134 MB each, 402 MB together) is read from HBM and not from the 256 MB
Infinity Cache; the single-key (cache-warm) rate is reported beside it.

Prints one JSON line. Extra fields: p50/p99 latency of individually
synchronised commits, the roofline of the dominant kernel (the fused
transform+MAC kernel, timed with HIP events on the library stream), the CPU
baseline (the reference's AVX2 prepared-commit path restated in C, run on this
host), and the side workloads, each checked against the oracle or the plain
form: the batched form at the same shape, kappa 8 x n 256 commitments
(`small`), 2^18 twisted forward and inverse transforms in both fields (`ntt`),
2^18 negacyclic products (`negacyclic`), LaBRADOR-scale inner
commitments, Gram matrix, JL projection and the steps around them
(`labrador`), 2000 Goldilocks commitments batched, graph-replayed and per call
(`goldilocks`), the serving loop against the plain per-call form on a Poisson
stream of mixed requests (`serving`), and key persistence and RingVec wire
encodings at the headline shape (`wire`). `--only NAME` runs one side workload
alone and prints its JSON (for profiling).
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


def batched_same_shape(ctx, keys, s_dev, log, batch=32, reps=6):
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


def goldilocks_batch(ctx, la, log, count=2000, kappa=16, max_cols=1 << 14):
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
    # the same batch recorded once into a HIP graph and replayed
    plan = key.batch_plan(wptr, lens, optr)
    plan.launch()
    ctx.synchronize()
    graph_same = bool(np.array_equal(out.download(np.uint64, (count, kappa, D)), res))
    t0 = time.perf_counter()
    for _ in range(reps):
        plan.launch()
    ctx.synchronize()
    graph = count * reps / (time.perf_counter() - t0)
    plan.close()
    # the one-launch form (tuning batch_fused = 1), graph-replayed the same way
    ctx.set_tuning(batch_fused=1, batch_fused_cols=128)
    plan = key.batch_plan(wptr, lens, optr)
    plan.launch()
    ctx.synchronize()
    fused_same = bool(np.array_equal(out.download(np.uint64, (count, kappa, D)), res))
    t0 = time.perf_counter()
    for _ in range(reps):
        plan.launch()
    ctx.synchronize()
    fused = count * reps / (time.perf_counter() - t0)
    plan.close()
    ctx.set_tuning(batch_fused=0)
    t0 = time.perf_counter()
    for i in range(count):
        key.commit_batch_device([wptr[i]], [lens[i]], [optr[i]])
    ctx.synchronize()
    plain = count / (time.perf_counter() - t0)
    res = {"batched_commits_per_s": round(batched, 1),
           "graph_replay_commits_per_s": round(graph, 1), "graph_equals_batched": graph_same,
           "one_launch_graph_commits_per_s": round(fused, 1),
           "one_launch_equals_batched": fused_same,
           "per_call_commits_per_s": round(plain, 1),
           "mean_len": float(lens.mean()), "oracle_spot_check": ok}
    log(f"Goldilocks batch x{count}: batched {batched:,.0f} commits/s, graph replay {graph:,.0f} "
        f"commits/s (== batched: {graph_same}), one-launch graph {fused:,.0f} commits/s "
        f"(== batched: {fused_same}), per-call {plain:,.0f} commits/s; oracle spot check {ok}")
    if CPU_ON:
        # the reference's AVX2 prepared commit on the wide Goldilocks path
        # (L = u64, u64_arith.rs:193-357), restated: a random sample of the
        # items on one core, a larger one over all cores (whole commitments per
        # thread), scaled to the batch by witness length (the work is linear in it)
        res_all = out.download(np.uint64, (count, kappa, D))
        c = gridref.Avx2Commit(p, 64, a, D)
        pick = rng.choice(count, 40, replace=False)
        c.load_items(pool, offs[pick], lens[pick])
        t1 = once_seconds(lambda: c.run_items(1))
        cpu_ok = bool(np.array_equal(c.items_result(), res_all[pick]))
        thr = cpu_threads()
        pick_all = rng.choice(count, min(count, 40 * thr), replace=False)
        c.load_items(pool, offs[pick_all], lens[pick_all])
        t_all = once_seconds(lambda: c.run_items(thr))
        c.close()
        cpu_1 = count / (t1 * float(lens.sum()) / float(lens[pick].sum()))
        cpu_all = count / (t_all * float(lens.sum()) / float(lens[pick_all].sum()))
        log(f"  CPU AVX2 {cpu_1:,.1f} commits/s on 1 core, {cpu_all:,.1f} on {thr} "
            f"(sample == GPU: {cpu_ok})")
        res["cpu_1core"] = {"commits_per_s": round(cpu_1, 2), "kind": "port",
                            "path": "prepared Ajtai commit, AVX2 L = u64 wide Montgomery "
                                    "(oracle/gridref_avx2.c)",
                            "sample": f"40 random items ({int(lens[pick].sum())} witness "
                                      f"polynomials, {t1:.2f} s), scaled to the batch's "
                                      f"{int(lens.sum())} by witness length",
                            "equals_gpu": cpu_ok}
        res["cpu_all_cores"] = {"commits_per_s": round(cpu_all, 2), "threads": thr,
                                "sample": f"{len(pick_all)} random items, scaled the same way",
                                "note": "one commitment per thread at a time"}
    key.close()
    pool_dev.free()
    out.free()
    ring.close()
    return res


def ct_pair_sum(a, b, q):
    """sum over rows g of the constant term of a_g * b_g in Z_q[X]/(X^64 + 1), q < 2^32."""
    a = np.asarray(a).reshape(-1, 64).astype(np.uint64)
    b = np.asarray(b).reshape(-1, 64).astype(np.uint64)
    qq = np.uint64(q)
    first = (a[:, 0] * b[:, 0]) % qq
    rest = np.zeros(a.shape[0], dtype=np.uint64)
    for k in range(1, 64):
        rest += (a[:, k] * b[:, 64 - k]) % qq
    return (int(first.sum()) - int(rest.sum())) % q


def host_cpu():
    """The host's CPU model, logical CPUs and the CPUs this process may run on."""
    model = "unknown"
    try:
        with open("/proc/cpuinfo") as fh:
            for line in fh:
                if line.startswith("model name"):
                    model = line.split(":", 1)[1].strip()
                    break
    except OSError:
        pass
    return {"model": model, "logical_cpus": os.cpu_count(),
            "usable_cpus": len(os.sched_getaffinity(0))}


# CPU baselines inside the side workloads (off with --no-cpu, e.g. for profiling)
CPU_ON = True


def cpu_threads():
    """Host threads for the all-core CPU figures (at most 16: the GPU box's CPU share)."""
    return max(1, min(16, len(os.sched_getaffinity(0))))


def once_seconds(fn):
    """Seconds of one call of fn (after the caller's own warm-up)."""
    a = time.perf_counter()
    fn()
    return time.perf_counter() - a


def cpu_seconds(fn, min_s=0.5):
    """Seconds per call of fn on this host (the oracle, one core), at least min_s of runs."""
    fn()
    n, a = 0, time.perf_counter()
    while True:
        fn()
        n += 1
        if time.perf_counter() - a >= min_s:
            return (time.perf_counter() - a) / n


def ntt_batch(ctx, la, log, count=1 << 18, reps=20):
    """Batched twisted transforms (TwistedNttPlan forward and inverse,
    poly/ntt.rs:75-104) of 2^18 polynomials in place on resident device
    buffers, q32 and Goldilocks: one gh_twisted_forward and one
    gh_twisted_inverse call each; a sample checked against the oracle and the
    round trip checked whole. CPU: the reference's AVX2 transforms, restated,
    on a sample."""
    from oracle import gridref
    L = la.lib()
    res = {}
    for name, q in (("q32", Q), ("goldilocks", la.GOLDILOCKS)):
        r = la.Ring(ctx, q, D)
        dt = np.uint32 if q < (1 << 32) else np.uint64
        rng = np.random.default_rng(33)
        h = (rng.integers(0, 1 << 62, size=(count, D), dtype=np.uint64) % np.uint64(q)).astype(dt)
        buf = ctx.to_device(h)
        L.gh_twisted_forward(r.h, buf.ptr, count, la.MEM_DEVICE)
        ctx.synchronize()
        fwd = buf.download(dt, (count, D))
        idx = rng.choice(count, 64, replace=False)
        ok = bool(np.array_equal(fwd[idx].astype(np.uint64),
                                 gridref.twisted_forward(q, D, h[idx].astype(np.uint64))))
        L.gh_twisted_inverse(r.h, buf.ptr, count, la.MEM_DEVICE)
        ctx.synchronize()
        ok = ok and bool(np.array_equal(buf.download(dt, (count, D)), h))
        us = {}
        for kind, fn in (("forward", L.gh_twisted_forward), ("inverse", L.gh_twisted_inverse)):
            for _ in range(3):
                fn(r.h, buf.ptr, count, la.MEM_DEVICE)
            ctx.synchronize()
            a = time.perf_counter()
            for _ in range(reps):
                fn(r.h, buf.ptr, count, la.MEM_DEVICE)
            ctx.synchronize()
            us[kind] = (time.perf_counter() - a) / reps * 1e6
        buf.free()
        r.close()
        nbytes = 2 * count * D * np.dtype(dt).itemsize  # read + written
        leg = {"polys": count, "forward_us": round(us["forward"], 1), "inverse_us": round(us["inverse"], 1),
               "forward_GBps": round(nbytes / us["forward"] / 1e3, 1),
               "inverse_GBps": round(nbytes / us["inverse"] / 1e3, 1), "oracle_sample_and_round_trip": ok}
        log(f"twisted NTT {name} x{count}: forward {us['forward']:,.1f} us, inverse {us['inverse']:,.1f} us "
            f"({nbytes / us['forward'] / 1e3:,.0f} / {nbytes / us['inverse'] / 1e3:,.0f} GB/s); oracle sample "
            f"and round trip {ok}")
        if CPU_ON:
            cpu_n = 1 << 14
            m = gridref.Avx2Twisted(q, 64)
            m.load(h[:cpu_n].astype(np.uint64))
            m.run(False)
            f1 = cpu_n / once_seconds(lambda: m.run(False))
            thr = cpu_threads()
            fa = cpu_n / once_seconds(lambda: m.run(False, thr))
            m.close()
            log(f"  CPU AVX2 forward {f1 / 1e6:.3f} M polys/s on 1 core, {fa / 1e6:.3f} M on {thr}")
            leg["cpu_1core"] = {"polys_per_s": round(f1, 1), "kind": "port",
                                "path": "TwistedNttPlan forward through the AVX2 NTT, L = u64 (oracle/gridref_avx2.c)",
                                "sample": f"{cpu_n} polynomials"}
            leg["cpu_all_cores"] = {"polys_per_s": round(fa, 1), "threads": thr}
        res[name] = leg
    return res


def negacyclic_batch(ctx, la, ring, log, pairs=1 << 18, reps=20):
    """Batched negacyclic products a*b mod (X^64 + 1), 2^18 pairs of uniform
    residues on resident device buffers (one gh_negacyclic_mul call); a sample
    is checked against the oracle."""
    rng = np.random.default_rng(22)
    ha = rng.integers(0, Q, size=(pairs, D), dtype=np.uint32)
    hb = rng.integers(0, Q, size=(pairs, D), dtype=np.uint32)
    pa, pb = ctx.to_device(ha), ctx.to_device(hb)
    po = ctx.alloc(pairs * D * 4)
    L = la.lib()
    for _ in range(3):
        L.gh_negacyclic_mul(ring.h, pa.ptr, pb.ptr, po.ptr, pairs, la.MEM_DEVICE)
    ctx.synchronize()
    a = time.perf_counter()
    for _ in range(reps):
        L.gh_negacyclic_mul(ring.h, pa.ptr, pb.ptr, po.ptr, pairs, la.MEM_DEVICE)
    ctx.synchronize()
    dt = (time.perf_counter() - a) / reps
    from oracle import gridref
    idx = rng.choice(pairs, 64, replace=False)
    got = po.download(np.uint32, (pairs, D))[idx].astype(np.uint64)
    ok = bool(np.array_equal(got, gridref.negacyclic_mul(Q, D, ha[idx].astype(np.uint64),
                                                         hb[idx].astype(np.uint64))))
    for b in (pa, pb, po):
        b.free()
    res = {"pairs_per_s": round(pairs / dt, 1), "us_per_batch": round(dt * 1e6, 1),
           "algorithmic_GBps": round(3 * pairs * D * 4 / dt / 1e9, 1), "oracle_sample": ok}
    log(f"negacyclic x{pairs}: {dt * 1e6:,.1f} us ({pairs / dt / 1e9:.2f} G pairs/s, "
        f"{3 * pairs * D * 4 / dt / 1e9:,.0f} GB/s algorithmic); oracle sample {ok}")
    if CPU_ON:
        # the reference's CPU path: negacyclic_mul_coeffs through its AVX2
        # twisted NTT (poly/ring.rs:83-101, simd/montgomery.rs:688-890),
        # restated, on a sample of the pairs; the scalar oracle beside it
        cpu_n = 1 << 16
        m = gridref.Avx2Negacyclic(Q, 64, D)
        m.load(ha[:cpu_n].astype(np.uint64), hb[:cpu_n].astype(np.uint64))
        m.run(1)
        cpu_1 = cpu_n / once_seconds(lambda: m.run(1))
        thr = cpu_threads()
        cpu_all = cpu_n / once_seconds(lambda: m.run(thr))
        sel = idx < cpu_n
        cpu_ok = bool(np.array_equal(m.result()[idx[sel]], got[sel])) if np.any(sel) else None
        m.close()
        sa, sb = ha[:4096].astype(np.uint64), hb[:4096].astype(np.uint64)
        oracle_pairs_s = 4096 / cpu_seconds(lambda: gridref.negacyclic_mul(Q, D, sa, sb))
        log(f"  CPU AVX2 path {cpu_1 / 1e6:.3f} M pairs/s on 1 core, {cpu_all / 1e6:.3f} M on "
            f"{thr}; scalar oracle {oracle_pairs_s / 1e6:.3f} M on 1 core")
        res["cpu_1core"] = {"pairs_per_s": round(cpu_1, 1), "kind": "port",
                            "path": "negacyclic_mul_coeffs via the AVX2 twisted NTT, L = u64 "
                                    "(oracle/gridref_avx2.c)",
                            "sample": f"{cpu_n} of the pairs", "equals_gpu": cpu_ok}
        res["cpu_all_cores"] = {"pairs_per_s": round(cpu_all, 1), "threads": thr,
                                "note": "pairs split over threads (this build's parallelisation)"}
        res["cpu_oracle_scalar_1core"] = {"pairs_per_s": round(oracle_pairs_s, 1),
                                          "sample": "4096 pairs through the scalar oracle"}
    return res


def commit_small(ctx, la, ring, log, batch=1024, reps=200):
    """Plain Ajtai commitment at kappa = 8, n = 256 (uniform key, ternary witness):
    per-call rate (latency-bound: A is 512 KB and stays in cache) and the batched
    rate (`batch` commitments per gh_commit_batch call)."""
    rng = np.random.default_rng(11)
    kappa, n = 8, 256
    a = rng.integers(0, Q, size=(kappa, n, D), dtype=np.uint32)
    key = la.Key(ring, a, domain=la.DOMAIN_NTT)
    t = rng.integers(-1, 2, size=(batch, n, D), dtype=np.int64)
    s_dev = ctx.to_device(np.where(t < 0, Q - 1, t).astype(np.uint32))
    o_dev = ctx.alloc(batch * kappa * D * 4)
    for _ in range(10):
        key.commit_device(s_dev, n, o_dev)
    ctx.synchronize()
    t0 = time.perf_counter()
    for _ in range(reps):
        key.commit_device(s_dev, n, o_dev)
    ctx.synchronize()
    per_call = reps / (time.perf_counter() - t0)
    wptr = [s_dev.ptr.value + i * n * D * 4 for i in range(batch)]
    optr = [o_dev.ptr.value + i * kappa * D * 4 for i in range(batch)]
    key.commit_batch_device(wptr, [n] * batch, optr)
    ctx.synchronize()
    t0 = time.perf_counter()
    for _ in range(10):
        key.commit_batch_device(wptr, [n] * batch, optr)
    ctx.synchronize()
    batched = batch * 10 / (time.perf_counter() - t0)
    res = {"per_call_commits_per_s": round(per_call, 1), "batched_commits_per_s": round(batched, 1),
           "batch": batch}
    log(f"commit kappa 8 x n 256: per-call {per_call:,.0f} commits/s, batched ({batch}/call) "
        f"{batched:,.0f} commits/s")
    if CPU_ON:
        # the reference's AVX2 prepared commit (L = u64), restated: 256 of the
        # commitments, one core and all cores (independent commitments per thread)
        from oracle import gridref
        gpu = o_dev.download(np.uint32, (batch, kappa, D))
        pool = np.where(t < 0, Q - 1, t).astype(np.uint64).reshape(batch * n, D)
        cpu_n = 256
        c = gridref.Avx2Commit(Q, 64, a.astype(np.uint64), D)
        c.load_items(pool[:cpu_n * n], [i * n for i in range(cpu_n)], [n] * cpu_n)
        c.run_items(1)
        cpu_1 = cpu_n / once_seconds(lambda: c.run_items(1))
        thr = cpu_threads()
        cpu_all = cpu_n / once_seconds(lambda: c.run_items(thr))
        cpu_ok = bool(np.array_equal(c.items_result(), gpu[:cpu_n].astype(np.uint64)))
        c.close()
        log(f"  CPU AVX2 {cpu_1:,.0f} commits/s on 1 core, {cpu_all:,.0f} on {thr} "
            f"(== GPU: {cpu_ok})")
        res["cpu_1core"] = {"commits_per_s": round(cpu_1, 1), "kind": "port",
                            "path": "prepared Ajtai commit, AVX2 L = u64 (oracle/gridref_avx2.c)",
                            "sample": f"{cpu_n} of the commitments", "equals_gpu": cpu_ok}
        res["cpu_all_cores"] = {"commits_per_s": round(cpu_all, 1), "threads": thr,
                                "note": "one commitment per thread at a time"}
    key.close()
    s_dev.free()
    o_dev.free()
    return res


def labrador_calls(ctx, la, ring, log, reps=10):
    """LaBRADOR scale: r = 32 witness vectors x 1024 polynomials (n = 2^15 in all),
    coefficients in [-2^7, 2^7), kappa = 16. Times the 32 inner commitments
    t_i = A*s_i (gh_inner_commit), the full Gram matrix (528 polynomials,
    gh_gram), the 256 x 2^21 ternary JL projection (gh_jl_project), the step
    after it (the 256 JL evaluations gh_jl_evals, the 2^23 conjugated JL
    polynomials gh_jl_conjugate, the flat int8 entries gh_jl_extract_flat) and
    the step around the commitment (h garbage, amortization z = sum c_i s_i,
    balanced decomposition of the 512 t polynomials), each on resident device
    buffers (the JL evaluations also on uniform residues, which need every
    digit plane where the witness needs two). The oracle checks one Gram entry; the JL p mod q must equal the
    exact sum mod q and the constant terms of the JL evaluations."""
    rng = np.random.default_rng(33)
    r_, n_, rows = 32, 1024, 256
    a = rng.integers(0, Q, size=(KAPPA, n_, D), dtype=np.uint32)
    key = la.Key(ring, a, domain=la.DOMAIN_NTT)
    s = witness(rng, r_ * n_).reshape(r_, n_, D)
    s_dev = ctx.to_device(s)
    t_dev = ctx.alloc(r_ * KAPPA * D * 4)
    npairs = r_ * (r_ + 1) // 2
    g_dev = ctx.alloc(npairs * D * 4)
    pm_dev, pe_dev = ctx.alloc(rows * 4), ctx.alloc(rows * 8)
    ev_dev = ctx.alloc(rows * D * 4)
    conj_dev = ctx.alloc(rows * r_ * n_ * D * 4)          # 2 GiB of conjugated polynomials
    flat_dev = ctx.alloc(rows * r_ * n_ * D)              # 512 MiB of int8 entries
    h_dev = ctx.alloc(npairs * D * 4)
    c = witness(rng, r_)
    c_dev = ctx.to_device(c)
    z_dev = ctx.alloc(n_ * D * 4)
    dec_dev = ctx.alloc(r_ * KAPPA * 4 * D * 4)
    scale = rng.integers(0, Q, size=(rows, D), dtype=np.uint32)  # aggregation scale polys
    sc_dev = ctx.to_device(scale)
    phi_dev = ctx.to_device(np.zeros((r_, n_, D), dtype=np.uint32))
    su_dev = ctx.to_device(rng.integers(0, Q, size=(r_, n_, D), dtype=np.uint32))
    evu_dev = ctx.alloc(rows * D * 4)
    seed = np.frombuffer(bytes(range(32)), dtype=np.uint8).copy()
    L = la.lib()
    import ctypes
    calls = {
        "inner_commit_32x1024": lambda: L.gh_inner_commit(key.h, s_dev.ptr, r_, n_, la.DOMAIN_COEFF,
                                                          t_dev.ptr, la.MEM_DEVICE),
        "gram_32x1024": lambda: L.gh_gram(ring.h, s_dev.ptr, r_, n_, la.DOMAIN_COEFF, g_dev.ptr,
                                          la.MEM_DEVICE),
        "jl_256x2^21": lambda: L.gh_jl_project(ring.h, ctypes.c_void_p(seed.ctypes.data), rows,
                                               s_dev.ptr, r_, n_ * D, pm_dev.ptr,
                                               ctypes.c_void_p(pe_dev.ptr.value), None,
                                               la.MEM_DEVICE),
        "jl_evals_256x32x1024": lambda: L.gh_jl_evals(ring.h, ctypes.c_void_p(seed.ctypes.data),
                                                      rows, s_dev.ptr, r_, n_ * D, ev_dev.ptr,
                                                      la.MEM_DEVICE),
        # the same on residues drawn uniformly mod q: every digit plane is live
        "jl_evals_256x32x1024_uniform": lambda: L.gh_jl_evals(
            ring.h, ctypes.c_void_p(seed.ctypes.data), rows, su_dev.ptr, r_, n_ * D, evu_dev.ptr,
            la.MEM_DEVICE),
        "jl_aggregate_256x32x1024": lambda: L.gh_jl_aggregate(
            ring.h, ctypes.c_void_p(seed.ctypes.data), rows, sc_dev.ptr, r_, n_ * D, phi_dev.ptr,
            la.MEM_DEVICE),
        "jl_conjugate_256x32x1024": lambda: L.gh_jl_conjugate(
            ring.h, ctypes.c_void_p(seed.ctypes.data), None, rows, r_, n_ * D, conj_dev.ptr,
            la.MEM_DEVICE),
        "jl_extract_flat_256x2^21": lambda: L.gh_jl_extract_flat(
            ring.h, ctypes.c_void_p(seed.ctypes.data), rows, r_, n_ * D, flat_dev.ptr,
            la.MEM_DEVICE),
        "garbage_h_32x1024": lambda: L.gh_garbage_h(ring.h, s_dev.ptr, s_dev.ptr, r_, n_,
                                                    la.DOMAIN_COEFF, h_dev.ptr, la.MEM_DEVICE),
        "amortize_32x1024": lambda: L.gh_amortize(ring.h, c_dev.ptr, s_dev.ptr, r_, n_,
                                                  la.DOMAIN_COEFF, z_dev.ptr, la.MEM_DEVICE),
        "decompose_t_512_base2^8_4limbs": lambda: L.gh_decompose_balanced(
            ring.h, t_dev.ptr, r_ * KAPPA, 256, 4, dec_dev.ptr, la.MEM_DEVICE),
    }
    res = {}
    for name, fn in calls.items():
        for _ in range(2):
            st = fn()
            if st:
                raise la.GridHipError(st, name)
        ctx.synchronize()
        t0 = time.perf_counter()
        for _ in range(reps):
            fn()
        ctx.synchronize()
        res[name] = {"us_per_call": round((time.perf_counter() - t0) / reps * 1e6, 1)}
    res["inner_commit_32x1024"]["commits_per_s"] = round(
        r_ * 1e6 / res["inner_commit_32x1024"]["us_per_call"], 1)
    from oracle import gridref
    g = g_dev.download(np.uint32, (npairs, D))
    exp = gridref.gram(Q, D, np.stack([s[3], s[17]]).astype(np.uint64))
    idx = r_ * 3 - 3 * 2 // 2 + (17 - 3)
    pm = pm_dev.download(np.uint32, (rows,)).astype(np.int64)
    pe = pe_dev.download(np.int64, (rows,))
    res["oracle_gram_entry"] = bool(np.array_equal(g[idx].astype(np.uint64), exp[1]))
    res["jl_modq_matches_exact"] = bool(np.array_equal(pm, pe % Q))
    ev = ev_dev.download(np.uint32, (rows, D)).astype(np.int64)
    res["jl_evals_constant_terms_equal_p"] = bool(np.array_equal(ev[:, 0], pm))
    # aggregation from phi = 0: sum_ij ct(phi_ij s_ij) = sum_m ct(scale_m jl_eval[m])
    phi_dev.upload(np.zeros((r_, n_, D), dtype=np.uint32))
    st = calls["jl_aggregate_256x32x1024"]()
    if st:
        raise la.GridHipError(st, "gh_jl_aggregate")
    phi = phi_dev.download(np.uint32, (r_ * n_, D))
    res["jl_aggregate_pairs_with_evals"] = bool(
        ct_pair_sum(phi, s.reshape(-1, D), Q) == ct_pair_sum(scale, ev, Q))
    if CPU_ON:
        # The reference's CPU path on this host. Inner commitments and the Gram
        # matrix in full: the reference's own algorithm multiplies in the
        # coefficient domain, one negacyclic_mul_coeffs per ring product
        # (crs.rs:211-216 -> RingMat::mul_slice; garbage.rs:76-93 ->
        # Ring::dot_product), through its AVX2 twisted NTT (restated in
        # oracle/gridref_avx2.c), and, labelled apart, the prepared (NTT-domain
        # key) commitments and the NTT-once Gram of the scalar oracle. The JL
        # projection in full through the oracle (project_and_extract_seq is scalar
        # code, jl.rs:388-508). Evaluations and aggregation on bounded samples.
        s64 = s.astype(np.uint64)
        seed_b = bytes(range(32))
        thr = cpu_threads()
        t_gpu = t_dev.download(np.uint32, (r_, KAPPA, D)).astype(np.uint64)
        dots = gridref.Avx2Dots(Q, 64, D)
        # inner commitments, coefficient-domain key (its words taken as coefficients)
        a_coef = a.astype(np.uint64)
        dots.load(a_coef, s64, [row for p_ in range(r_) for row in range(KAPPA)],
                  [p_ for p_ in range(r_) for row in range(KAPPA)])
        icc_s = once_seconds(lambda: dots.run(1))
        icc_all = once_seconds(lambda: dots.run(thr))
        zk = np.zeros((KAPPA, 1, D), dtype=np.uint64)
        zr = np.zeros((1, D), dtype=np.uint64)
        icc_ok = bool(np.array_equal(dots.result()[:KAPPA],
                                     gridref.commit(Q, D, a_coef, zk, s64[0], zr)))
        pairs = [(i, j) for i in range(r_) for j in range(i, r_)]
        dots.load(s64, s64, [i for i, _ in pairs], [j for _, j in pairs])
        gc_s = once_seconds(lambda: dots.run(1))
        gc_all = once_seconds(lambda: dots.run(thr))
        gc_ok = bool(np.array_equal(dots.result(), g.astype(np.uint64)))
        dots.close()
        # prepared commitments (NTT-domain key), AVX2 L = u64, all 32 parts
        c = gridref.Avx2Commit(Q, 64, a.astype(np.uint64), D)
        c.load_items(s64.reshape(r_ * n_, D), [i * n_ for i in range(r_)], [n_] * r_)
        ic_s = once_seconds(lambda: c.run_items(1))
        ic_all = once_seconds(lambda: c.run_items(thr))
        ic_ok = bool(np.array_equal(c.items_result(), t_gpu))
        c.close()
        g_s = once_seconds(lambda: gridref.gram(Q, D, s64))
        jl_s = once_seconds(lambda: gridref.jl_project(Q, seed_b, rows, s64.reshape(r_, -1), D))
        ev_s = cpu_seconds(lambda: gridref.jl_evals(Q, D, seed_b, rows, s64[:1, :4])) / 4 * r_ * n_
        # the aggregation's JL term is rows x r x n ring products (aggregation.rs:237-245);
        # timed as 16384 of the oracle's NTT products, conjugation and sums not counted
        pa = rng.integers(0, Q, size=(16384, D), dtype=np.uint64)
        pb = rng.integers(0, Q, size=(16384, D), dtype=np.uint64)
        ag_s = cpu_seconds(lambda: gridref.negacyclic_mul(Q, D, pa, pb)) / 16384 * rows * r_ * n_
        res["cpu_1core_us"] = {
            "inner_commit_32x1024": round(icc_s * 1e6, 1),
            "gram_32x1024": round(gc_s * 1e6, 1),
            "jl_256x2^21": round(jl_s * 1e6, 1),
            "path": "the reference's algorithm: inner commitments and Gram with every ring "
                    "product a negacyclic_mul_coeffs through the AVX2 twisted NTT (L = u64), "
                    "in full (524288 and 540672 products); JL projection in full through "
                    "the scalar oracle",
            "checks": {"inner_commit_part0_equals_oracle": icc_ok, "gram_equals_gpu": gc_ok},
            "prepared_key_ntt_once": {"inner_commit_32x1024": round(ic_s * 1e6, 1),
                                      "gram_32x1024": round(g_s * 1e6, 1),
                                      "inner_commit_equals_gpu": ic_ok,
                                      "note": "NTT-domain key: AVX2 prepared commit of all 32 "
                                              "parts; Gram with each witness transformed once "
                                              "(scalar oracle)"},
            "jl_evals_256x32x1024": round(ev_s * 1e6, 1),
            "jl_aggregate_256x32x1024": round(ag_s * 1e6, 1),
            "kind": "port",
            "sample": "inner commitments, Gram, prepared commitments and JL projection: the "
                      "whole call; evaluations over 4 polynomials of 1 part and aggregation as "
                      "16384 NTT products, scaled to the call"}
        res["cpu_all_cores_us"] = {"threads": thr,
                                   "inner_commit_32x1024": round(icc_all * 1e6, 1),
                                   "gram_32x1024": round(gc_all * 1e6, 1),
                                   "prepared_inner_commit_32x1024": round(ic_all * 1e6, 1),
                                   "note": "products / parts split over threads (this build's "
                                           "parallelisation)"}
    log(f"LaBRADOR r 32 x 1024: inner commits {res['inner_commit_32x1024']['us_per_call']:,.0f} us "
        f"({res['inner_commit_32x1024']['commits_per_s']:,.0f} commits/s), Gram "
        f"{res['gram_32x1024']['us_per_call']:,.0f} us, JL {res['jl_256x2^21']['us_per_call']:,.0f} us; "
        f"oracle Gram entry {res['oracle_gram_entry']}, JL mod q == exact mod q "
        f"{res['jl_modq_matches_exact']}")
    cpu1 = res.get("cpu_1core_us")
    if cpu1:
        log(f"  CPU 1 core (reference algorithm, AVX2): inner commits "
            f"{cpu1['inner_commit_32x1024'] / 1e3:,.1f} ms, Gram {cpu1['gram_32x1024'] / 1e3:,.1f} ms, "
            f"JL {cpu1['jl_256x2^21'] / 1e3:,.1f} ms; prepared key / NTT once: inner commits "
            f"{cpu1['prepared_key_ntt_once']['inner_commit_32x1024'] / 1e3:,.1f} ms, Gram "
            f"{cpu1['prepared_key_ntt_once']['gram_32x1024'] / 1e3:,.1f} ms; checks {cpu1['checks']}, "
            f"prepared == GPU {cpu1['prepared_key_ntt_once']['inner_commit_equals_gpu']}")
    log(f"  JL evals {res['jl_evals_256x32x1024']['us_per_call']:,.0f} us (ct == p "
        f"{res['jl_evals_constant_terms_equal_p']}; uniform residues "
        f"{res['jl_evals_256x32x1024_uniform']['us_per_call']:,.0f} us), JL aggregation "
        f"{res['jl_aggregate_256x32x1024']['us_per_call']:,.0f} us (pairs with evals "
        f"{res['jl_aggregate_pairs_with_evals']}), conjugated polys "
        f"{res['jl_conjugate_256x32x1024']['us_per_call']:,.0f} us, flat "
        f"{res['jl_extract_flat_256x2^21']['us_per_call']:,.0f} us, h garbage "
        f"{res['garbage_h_32x1024']['us_per_call']:,.0f} us, amortize "
        f"{res['amortize_32x1024']['us_per_call']:,.0f} us, decompose "
        f"{res['decompose_t_512_base2^8_4limbs']['us_per_call']:,.0f} us")
    key.close()
    for b in (s_dev, t_dev, g_dev, pm_dev, pe_dev, ev_dev, conj_dev, flat_dev, h_dev, c_dev,
              z_dev, dec_dev, sc_dev, phi_dev, su_dev, evu_dev):
        b.free()
    return res


def wire_formats(ctx, la, ring, log, reps=5):
    """Canonical wire formats and key persistence at the headline shape: the
    kappa 16 x 2^15 q32 key saved as LinearCommitmentKey<TwistedNttPoly> bytes
    (134 MB) and loaded back into a resident key, and a 2^15-polynomial witness
    RingVec (8 MB) encoded from and decoded into device memory. Host bytes are
    pageable numpy buffers, so the rates include the PCIe copies; the loaded
    key's commitment must equal the original's."""
    import ctypes
    rng = np.random.default_rng(91)
    a = rng.integers(0, Q, size=(KAPPA, N, D), dtype=np.uint32)
    key = la.Key(ring, a, domain=la.DOMAIN_NTT)
    res = {}
    t0 = time.perf_counter()
    for _ in range(reps):
        blob = key.save(la.DOMAIN_NTT)
    dt = (time.perf_counter() - t0) / reps
    res["key_save_ntt"] = {"ms": round(dt * 1e3, 2), "bytes": len(blob),
                           "GBps": round(len(blob) / dt / 1e9, 2)}
    t0 = time.perf_counter()
    for _ in range(reps):
        k2 = ring.load_key(blob, la.DOMAIN_NTT)
        k2.close()
    dt = (time.perf_counter() - t0) / reps
    res["key_load_ntt"] = {"ms": round(dt * 1e3, 2), "GBps": round(len(blob) / dt / 1e9, 2)}
    k2 = ring.load_key(blob, la.DOMAIN_NTT)
    s = witness(rng, N)
    res["loaded_key_commits_equal"] = bool(np.array_equal(k2.commit(s), key.commit(s)))
    k2.close()
    # RingVec of the witness, device words <-> host bytes
    s_dev = ctx.to_device(s)
    out_dev = ctx.alloc(s.nbytes)
    L = la.lib()
    size = ring.wire_size(la.WIRE_RINGVEC, N)
    enc = np.zeros(size, dtype=np.uint8)
    wr = ctypes.c_uint64()
    t0 = time.perf_counter()
    for _ in range(reps * 4):
        st = L.gh_wire_encode(ring.h, la.WIRE_RINGVEC, N, 1, s_dev.ptr, la.MEM_DEVICE,
                              enc.ctypes.data, size, ctypes.byref(wr))
        if st:
            raise la.GridHipError(st, "gh_wire_encode")
    dt = (time.perf_counter() - t0) / (reps * 4)
    res["ringvec_encode_2^15_polys"] = {"us": round(dt * 1e6, 1), "GBps": round(size / dt / 1e9, 2)}
    t0 = time.perf_counter()
    for _ in range(reps * 4):
        st = L.gh_wire_decode(ring.h, la.WIRE_RINGVEC, la.DOMAIN_COEFF, enc.ctypes.data, size, N,
                              1, out_dev.ptr, la.MEM_DEVICE, None)
        if st:
            raise la.GridHipError(st, "gh_wire_decode")
    ctx.synchronize()
    dt = (time.perf_counter() - t0) / (reps * 4)
    res["ringvec_decode_2^15_polys"] = {"us": round(dt * 1e6, 1), "GBps": round(size / dt / 1e9, 2)}
    res["ringvec_round_trip_equal"] = bool(np.array_equal(out_dev.download(np.uint32, s.shape), s))
    log(f"wire: key save {res['key_save_ntt']['ms']:.1f} ms ({res['key_save_ntt']['GBps']:.1f} GB/s), "
        f"load {res['key_load_ntt']['ms']:.1f} ms ({res['key_load_ntt']['GBps']:.1f} GB/s), loaded key "
        f"commits equal {res['loaded_key_commits_equal']}; RingVec 2^15 polys encode "
        f"{res['ringvec_encode_2^15_polys']['us']:,.0f} us, decode "
        f"{res['ringvec_decode_2^15_polys']['us']:,.0f} us, round trip "
        f"{res['ringvec_round_trip_equal']}")
    key.close()
    s_dev.free()
    out_dev.free()
    return res


def serving_cpu(reqs, host_keys, host_pools, plain_bytes, served_bytes, la, log, sample=64,
                seed=56):
    """The same trace through the reference's CPU path, one request at a time
    on one core: commit requests as AVX2 prepared commitments of their parts
    (L = u64; the wide Montgomery path for Goldilocks), Gram requests as the
    reference's coefficient-domain dot products (every ring product a
    negacyclic_mul_coeffs through the AVX2 twisted NTT). Timed on a random
    sample of the requests, scaled to the trace by request count; each CPU
    result is compared with the plain GPU result's bytes."""
    from oracle import gridref
    rng = np.random.default_rng(seed)
    pick = sorted(rng.choice(len(reqs), min(sample, len(reqs)), replace=False))
    qs = {"q32": Q, "gold": la.GOLDILOCKS}
    commit_h = {n: gridref.Avx2Commit(qs[n], 64, host_keys[n].astype(np.uint64), D)
                for n in ("q32", "gold")}
    dots = {n: gridref.Avx2Dots(qs[n], 64, D) for n in ("q32", "gold")}
    times, ok, commits = [], True, 0
    for i in pick:
        q = reqs[i]
        w = host_pools[q["name"]][q["off"]:q["off"] + q["r"] * q["n"]]
        if q["op"] == la.OP_COMMIT:
            h = commit_h[q["name"]]
            h.load_items(w, [k * q["n"] for k in range(q["r"])], [q["n"]] * q["r"])
            times.append(once_seconds(lambda: h.run_items(1)))
            out = h.items_result()
            commits += q["r"]
        else:
            parts = w.reshape(q["r"], q["n"], D)
            prs = [(a_, b_) for a_ in range(q["r"]) for b_ in range(a_, q["r"])]
            h = dots[q["name"]]
            h.load(parts, parts, [a_ for a_, _ in prs], [b_ for _, b_ in prs])
            times.append(once_seconds(lambda: h.run(1)))
            out = h.result()
        dt = np.uint32 if q["wb"] == 4 else np.uint64
        gpu = plain_bytes[q["o"]:q["o"] + q["ob"]].view(dt).astype(np.uint64)
        ok = ok and bool(np.array_equal(out.reshape(-1), gpu))
    for h in list(commit_h.values()) + list(dots.values()):
        h.close()
    t = np.array(times)
    scale = len(reqs) / len(pick)
    total_commits = sum(q["r"] for q in reqs if q["op"] == la.OP_COMMIT)
    est = float(t.sum()) * scale
    log(f"serving CPU 1 core (sample of {len(pick)}): {len(reqs) / est:,.1f} req/s, "
        f"{total_commits / est:,.1f} commits/s; per-request service p50 "
        f"{np.percentile(t, 50) * 1e3:,.2f} ms p99 {np.percentile(t, 99) * 1e3:,.2f} ms; "
        f"== GPU {ok}")
    return {"requests_per_s": round(len(reqs) / est, 2),
            "commits_per_s": round(total_commits / est, 2), "kind": "port",
            "service_p50_us": round(float(np.percentile(t, 50)) * 1e6, 1),
            "service_p99_us": round(float(np.percentile(t, 99)) * 1e6, 1),
            "path": "AVX2 prepared commitments (L = u64) and the reference's "
                    "coefficient-domain Gram (oracle/gridref_avx2.c), one request at a time",
            "sample": f"{len(pick)} random requests of the trace ({commits} commitments), "
                      f"scaled to its {len(reqs)} requests",
            "equals_gpu": ok}


def serving(ctx, la, log, count=1000, seed=55):
    """Serving: a Poisson arrival stream of mixed commit and Gram requests, half
    q32 and half Goldilocks (d = 64, kappa = 16 keys of 2^15 columns, both
    resident), witness sizes log-uniform in [2^8, 2^15] polynomials split into
    r in {1..32} parts, coefficients uniform in [-2^7, 2^7), witnesses resident
    in HBM. The same trace runs through the plain per-call form (one
    gh_commit per commitment, one gh_gram per Gram request, each synchronised,
    served in arrival order) and through gh_server, at two offered rates (half
    and three times the plain form's saturated request rate) and saturated (all
    requests at once). Latency = completion - arrival. Every served result is
    compared with the plain result of the same request."""
    from lattica_amd import CommitItem, Request
    import ctypes
    rng = np.random.default_rng(seed)
    kappa, cols, pool_polys = 16, 1 << 15, 1 << 15
    rings = {"q32": la.Ring(ctx, Q, D), "gold": la.Ring(ctx, la.GOLDILOCKS, D)}
    keys, pools, host_keys, host_pools = {}, {}, {}, {}
    for name, ring in rings.items():
        if name == "q32":
            a = rng.integers(0, Q, size=(kappa, cols, D), dtype=np.uint32)
        else:
            hi = rng.integers(0, 1 << 32, size=(kappa, cols, D), dtype=np.uint64)
            a = ((hi << np.uint64(32)) | rng.integers(0, 1 << 32, size=(kappa, cols, D),
                                                       dtype=np.uint64)) % np.uint64(la.GOLDILOCKS)
            del hi
        keys[name] = la.Key(ring, a, domain=la.DOMAIN_NTT)
        host_keys[name] = a
        x = rng.integers(-128, 128, size=(pool_polys, D), dtype=np.int64)
        lift = np.where(x < 0, np.uint64(ring.q) - (-x).astype(np.uint64), x.astype(np.uint64))
        pools[name] = ctx.to_device(lift.astype(ring.dtype))
        host_pools[name] = lift
    reqs, out_bytes = [], 0
    for i in range(count):
        name = "q32" if rng.random() < 0.5 else "gold"
        ring = rings[name]
        wb = np.dtype(ring.dtype).itemsize
        op = la.OP_COMMIT if rng.random() < 0.5 else la.OP_GRAM
        r = int(rng.integers(1, 33))
        total = int(np.floor(2.0 ** rng.uniform(8, 15)))
        n = max(1, total // r)
        off = int(rng.integers(0, pool_polys - r * n + 1))
        ob = r * kappa * D * wb if op == la.OP_COMMIT else r * (r + 1) // 2 * D * wb
        reqs.append({"name": name, "op": op, "r": r, "n": n, "w": pools[name].ptr.value + off * D * wb,
                     "off": off, "o": out_bytes, "ob": ob, "wb": wb})
        out_bytes += (ob + 255) // 256 * 256
    outs = {"plain": ctx.alloc(out_bytes), "served": ctx.alloc(out_bytes)}
    n_commits = sum(q["r"] for q in reqs if q["op"] == la.OP_COMMIT)
    L = la.lib()

    # plain per-call form: prebuilt one-item descriptors, no per-call Python building
    def plain_calls(dst):
        base = outs[dst].ptr.value
        calls = []
        for q in reqs:
            ring, key = rings[q["name"]], keys[q["name"]]
            if q["op"] == la.OP_COMMIT:
                items = []
                for i in range(q["r"]):
                    it = (CommitItem * 1)()
                    it[0].witness = q["w"] + i * q["n"] * D * q["wb"]
                    it[0].len = q["n"]
                    it[0].t_out = base + q["o"] + i * kappa * D * q["wb"]
                    items.append(it)
                calls.append(("c", key.h, items))
            else:
                calls.append(("g", ring.h, (ctypes.c_void_p(q["w"]), q["r"], q["n"],
                                            ctypes.c_void_p(base + q["o"]))))
        return calls

    pcalls = plain_calls("plain")

    def run_plain(arrivals):
        lat = np.zeros(count)
        t0 = time.perf_counter()
        for i, c in enumerate(pcalls):
            ta = t0 + arrivals[i]
            while time.perf_counter() < ta:
                pass
            if c[0] == "c":
                for it in c[2]:
                    L.gh_commit_batch(c[1], it, 1, la.MEM_DEVICE)
            else:
                w, r, n, g = c[2]
                L.gh_gram(c[1], w, r, n, la.DOMAIN_COEFF, g, la.MEM_DEVICE)
            ctx.synchronize()
            lat[i] = time.perf_counter() - ta
        return lat, time.perf_counter() - t0

    sbase = outs["served"].ptr.value
    sreqs = []
    for q in reqs:
        rq = Request()
        rq.ops, rq.r, rq.n, rq.mem = q["op"], q["r"], q["n"], la.MEM_DEVICE
        rq.ring = rings[q["name"]].h.value
        rq.key = keys[q["name"]].h.value if q["op"] == la.OP_COMMIT else None
        rq.witness = q["w"]
        if q["op"] == la.OP_COMMIT:
            rq.t_out = sbase + q["o"]
        else:
            rq.g_out = sbase + q["o"]
        sreqs.append(rq)

    def run_served(srv, arrivals):
        tks = []
        tk = ctypes.c_uint64()
        t0 = time.perf_counter()
        for i, rq in enumerate(sreqs):
            ta = t0 + arrivals[i]
            while time.perf_counter() < ta:
                pass
            st = L.gh_server_submit(srv.h, ctypes.byref(rq), ctypes.byref(tk))
            if st:
                raise la.GridHipError(st, "gh_server_submit")
            tks.append(tk.value)
        lat = np.zeros(count)
        for i, t in enumerate(tks):
            status, lat_us, _, _ = srv.wait(t)
            if status:
                raise la.GridHipError(status, "served request")
            lat[i] = lat_us * 1e-6
        return lat, time.perf_counter() - t0

    def summary(lat, elapsed):
        return {"p50_us": round(float(np.percentile(lat, 50)) * 1e6, 1),
                "p99_us": round(float(np.percentile(lat, 99)) * 1e6, 1),
                "commits_per_s": round(n_commits / elapsed, 1),
                "requests_per_s": round(count / elapsed, 1)}

    zeros = np.zeros(count)
    run_plain(zeros)  # warm-up (plans, workspaces)
    lat, el = run_plain(zeros)
    sat_plain = summary(lat, el)
    cap = count / el
    srv = la.Server(ctx, mem_bytes=8 << 30)
    run_served(srv, zeros)  # warm-up
    lat, el = run_served(srv, zeros)
    sat_served = summary(lat, el)
    res = {"requests": count, "commits": n_commits,
           "grams": sum(1 for q in reqs if q["op"] == la.OP_GRAM),
           "saturated": {"plain": sat_plain, "served": sat_served}}
    for label, mult in (("offered_0.5x_plain_capacity", 0.5), ("offered_3x_plain_capacity", 3.0)):
        rate = mult * cap
        arrivals = np.cumsum(rng.exponential(1.0 / rate, size=count))
        lp, ep = run_plain(arrivals)
        ls, es = run_served(srv, arrivals)
        res[label] = {"offered_requests_per_s": round(rate, 1), "plain": summary(lp, ep),
                      "served": summary(ls, es)}
        log(f"serving {label} ({rate:,.0f} req/s): plain p50 {res[label]['plain']['p50_us']:,.0f} us "
            f"p99 {res[label]['plain']['p99_us']:,.0f} us {res[label]['plain']['commits_per_s']:,.0f} "
            f"commits/s | served p50 {res[label]['served']['p50_us']:,.0f} us p99 "
            f"{res[label]['served']['p99_us']:,.0f} us {res[label]['served']['commits_per_s']:,.0f} "
            f"commits/s")
    res["server"] = {k: v for k, v in srv.stats().items()
                     if k in ("waves", "max_wave_seen", "mem_peak", "device_ms", "failed")}
    srv.close()
    ctx.synchronize()
    a = outs["plain"].download(np.uint8, (out_bytes,))
    b = outs["served"].download(np.uint8, (out_bytes,))
    res["served_equals_plain"] = bool(np.array_equal(a, b))
    if CPU_ON:
        res["cpu_1core"] = serving_cpu(reqs, host_keys, host_pools, a, b, la, log)
    log(f"serving saturated: plain {sat_plain['commits_per_s']:,.0f} commits/s "
        f"({sat_plain['requests_per_s']:,.0f} req/s), served {sat_served['commits_per_s']:,.0f} "
        f"commits/s ({sat_served['requests_per_s']:,.0f} req/s, p99 {sat_served['p99_us']:,.0f} us); "
        f"served == plain for all {count}: {res['served_equals_plain']}")
    for b in list(outs.values()) + list(pools.values()):
        b.free()
    for k in keys.values():
        k.close()
    for r in rings.values():
        r.close()
    return res


def main():
    ap = argparse.ArgumentParser(description=__doc__)
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=2000)
    ap.add_argument("--warmup", type=int, default=1000,
                    help="untimed steps first; about 30 ms, so the clocks have ramped")
    ap.add_argument("--keys", type=int, default=3, help="resident keys used in rotation")
    ap.add_argument("--cpu-seconds", type=float, default=10.0)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-extras", action="store_true")
    ap.add_argument("--no-check", action="store_true")
    ap.add_argument("--only", choices=["small", "ntt", "negacyclic", "labrador", "goldilocks",
                                       "serving", "wire", "hbm", "issue", "valu"],
                    help="run one side workload alone (for profiling) and print its JSON")
    args = ap.parse_args()
    global CPU_ON
    CPU_ON = not args.no_cpu
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
    if args.only:
        legs = {"small": lambda: commit_small(ctx, la, ring, log),
                "ntt": lambda: ntt_batch(ctx, la, log),
                "negacyclic": lambda: negacyclic_batch(ctx, la, ring, log),
                "labrador": lambda: labrador_calls(ctx, la, ring, log),
                "goldilocks": lambda: goldilocks_batch(ctx, la, log),
                "serving": lambda: serving(ctx, la, log),
                "wire": lambda: wire_formats(ctx, la, ring, log),
                # streaming read and copy probes of known size (1 GiB): the
                # calibration of rocprofv3's FETCH_SIZE / WRITE_SIZE
                "hbm": lambda: dict(zip(("read_GBps", "copy_GBps"),
                                        ctx.measure_bandwidth(1 << 30, 3))),
                # int8 MFMA and VALU issue rates, alone and interleaved: what
                # bounds the JL projection
                "issue": lambda: {str(w): ctx.measure_issue(w) for w in (1, 2, 3, 4)},
                # single VALU instruction kinds: what the multiply-heavy
                # transforms and multiply-accumulates are bounded by
                "valu": lambda: {str(w): ctx.measure_valu(w) for w in (1, 2, 4)}}
        print(json.dumps({"workload": args.only, "result": legs[args.only]()}), flush=True)
        ring.close()
        ctx.close()
        return
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

    # warm-up steps directly before the timed ones
    for i in range(args.warmup):
        step(i)
    ctx.synchronize()

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
    if tuning["fused"] == 2 and KAPPA % 16 == 0:
        mac_kernel = "k_commit_rows_q32"  # witness transform + MAC, waves split the rows
    elif tuning["fused"]:
        mac_kernel = "k_commit_fused_q32"  # witness transform + MAC, waves split the columns
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
        # the same with 5 keys in rotation (670 MB of keys against the 256 MB
        # Infinity Cache): equal to the 3-key rate when 3 already defeat the cache
        more = [la.Key(ring, rng.integers(0, Q, size=(KAPPA, N, D), dtype=np.uint32),
                       domain=la.DOMAIN_NTT) for _ in range(max(0, 5 - len(keys)))]
        five = (keys + more)[:5]
        nrot = max(args.steps, 200)
        for i in range(50):
            five[i % 5].commit_device(s_dev[i % nw], N, t_dev[i % nw])
        ctx.synchronize()
        a = time.perf_counter()
        for i in range(nrot):
            five[i % 5].commit_device(s_dev[i % nw], N, t_dev[i % nw])
        ctx.synchronize()
        r5 = nrot / (time.perf_counter() - a)
        a = time.perf_counter()
        for i in range(nrot):
            step(i)
        ctx.synchronize()
        r3 = nrot / (time.perf_counter() - a)
        extra["keys_in_rotation_commits_per_s"] = {"1": extra["single_key_commits_per_s"],
                                                   str(len(keys)): round(r3, 1),
                                                   "5": round(r5, 1), "commits_timed": nrot}
        for k in more:
            k.close()
        log(f"keys in rotation: 1 -> {extra['single_key_commits_per_s']:,.0f}, {len(keys)} -> "
            f"{r3:,.0f}, 5 -> {r5:,.0f} commits/s")
        # measured HBM ceiling: streaming read / copy over 1 GiB, and a streaming
        # read of one commitment's bytes (three buffers in turn, so none is
        # served from the Infinity Cache), launch ramp and tail included
        rd, cp = ctx.measure_bandwidth(1 << 30, 5)
        rd_at, _ = ctx.measure_bandwidth(ALGO_BYTES, 20)
        extra["hbm_ceiling_GBps"] = {"stream_read_1GiB": round(rd, 1),
                                     "stream_copy_1GiB": round(cp, 1),
                                     "stream_read_commit_size": round(rd_at, 1)}
        log(f"measured HBM ceiling: read {rd:,.0f} GB/s, copy {cp:,.0f} GB/s (1 GiB); "
            f"read {rd_at:,.0f} GB/s at {ALGO_BYTES / 1e6:.1f} MB")
        # VALU counters of the C2 and C4 kernels from the committed rocprofv3
        # summary (not measured in this run)
        vpath = os.path.join(ROOT, "profiles", "valu_c2_c4.json")
        valu = {}
        if os.path.exists(vpath):
            with open(vpath) as fh:
                valu = json.load(fh)
        extra["twisted_ntt_2^18_polys"] = ntt_batch(ctx, la, log)
        extra["negacyclic_q32_2^18_pairs"] = negacyclic_batch(ctx, la, ring, log)
        if "negacyclic_q32_2^18_pairs" in valu:
            vp = valu["negacyclic_q32_2^18_pairs"]
            leg = extra["negacyclic_q32_2^18_pairs"]
            leg["valu_profile"] = vp
            # the profiled VALU floor against this run's time per call
            if vp.get("valu_roofline_us"):
                leg["valu_roofline_frac_this_run"] = round(vp["valu_roofline_us"] / leg["us_per_batch"], 3)
        extra["batched_same_shape"] = batched_same_shape(ctx, keys, s_dev, log)
        extra["commit_q32_k8_n256"] = commit_small(ctx, la, ring, log)
        extra["labrador_q32_r32_n1024"] = labrador_calls(ctx, la, ring, log)
        extra["goldilocks_batch_2000"] = goldilocks_batch(ctx, la, log)
        if "goldilocks_batch_2000" in valu:
            extra["goldilocks_batch_2000"]["valu_profile"] = valu["goldilocks_batch_2000"]
        extra["serving_mixed_1000"] = serving(ctx, la, log)
        extra["wire_formats"] = wire_formats(ctx, la, ring, log)

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
        # the reference's L = u32 limb (8 lanes, the wide Montgomery path for q >= 2^31)
        c32 = gridref.Avx2Commit(Q, 32, a_host0.astype(np.uint64), D)
        c32.load(s_host[0].astype(np.uint64))
        c32.run(1)
        n32, a = 0, time.perf_counter()
        while time.perf_counter() - a < args.cpu_seconds / 2:
            c32.run(1)
            n32 += 1
        cpu_l32 = n32 / (time.perf_counter() - a)
        c32.close()
        cpu = {"value": round(cpu_rate, 3), "unit": "commits/s", "cores": 1, "kind": "port",
               "sample": f"{n_done} prepared commits (kappa=16, n=2^15, d=64, q32) of the reference's "
                         f"AVX2 L=u64 path restated in C (oracle/gridref_avx2.c), 1 thread",
               "all_cores": {"value": round(cpu_all, 3), "threads": threads_all,
                             "note": "this build's row/column split over threads"},
               "limb_u32_1core": {"value": round(cpu_l32, 3), "commits": n32,
                                  "note": "the reference's L = u32 path, 1 thread: 8-lane "
                                          "transforms, scalar wide Montgomery products in the "
                                          "dot product for q >= 2^31"},
               "host": host_cpu()}
        log(f"cpu baseline: {cpu_rate:.2f} commits/s on 1 core (L = u64), {cpu_l32:.2f} "
            f"(L = u32), {cpu_all:.2f} on {threads_all} threads; {cpu['host']['model']}")

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
                     "frac_of_measured_read_ceiling_1GiB":
                         round(achieved / extra["hbm_ceiling_GBps"]["stream_read_1GiB"], 4)
                         if "hbm_ceiling_GBps" in extra else None,
                     "frac_of_measured_read_ceiling_commit_size":
                         round(achieved / extra["hbm_ceiling_GBps"]["stream_read_commit_size"], 4)
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
