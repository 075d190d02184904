"""Autotuner for the launch shapes of libgridhip.so.

Sweeps the tunable launch parameters of a context (gh_ctx_set_tuning) on a
representative workload, times each setting with the library's own event
timing plus wall clock, checks that every setting gives bit-identical results,
and writes the fastest setting to tuning.json next to the library. Context()
applies that file when it exists.

    python -m lattica_amd.autotune [--quick]

Workload: plain Ajtai commitments at kappa = 16, n = 2^15, d = 64, q32, with
three resident keys in rotation (so the key is read from HBM, not from the
Infinity Cache), the shape the library is deployed for.
"""
from __future__ import annotations

import argparse
import itertools
import json
import os
import sys
import time

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
TUNING_FILE = os.path.join(HERE, "tuning.json")


def _witness(rng, q, n, d=64):
    x = rng.integers(-128, 128, size=(n, d), dtype=np.int64)
    return np.where(x < 0, q + x, x).astype(np.uint32)


def tune_commit(ctx, kappa=16, n=1 << 15, reps=200, quick=False, log=print):
    import lattica_amd as la
    q = la.Q32
    ring = la.Ring(ctx, q, 64)
    rng = np.random.default_rng(7)
    keys = [la.Key(ring, rng.integers(0, q, size=(kappa, n, 64), dtype=np.uint32),
                   domain=la.DOMAIN_NTT) for _ in range(3)]
    ws = [ctx.to_device(_witness(rng, q, n)) for _ in range(4)]
    ts = [ctx.alloc(kappa * 64 * 4) for _ in range(4)]

    bpcs = [1, 2] if quick else [1, 2, 3, 4]
    configs = []
    # fused transform+MAC, row split (16-row blocks of four waves): key
    # prefetch depth and blocks per CU
    for bpc, depth in itertools.product(bpcs, [1, 2, 3, 4]):
        configs.append({"fused": 2, "fused_depth": depth, "mac_variant": 1, "mac_waves": 4,
                        "mac_blocks_per_cu": bpc, "quad_max": 1 << 17})
    # fused transform+MAC, column split (16-byte loads): block shape and depth
    for waves, bpc, depth in itertools.product([16, 8, 4], bpcs, [1, 2, 3, 4]):
        configs.append({"fused": 1, "fused_depth": depth, "mac_variant": 1, "mac_waves": waves,
                        "mac_blocks_per_cu": bpc, "quad_max": 1 << 17})
    # separate kernels: MAC load width, block shape, witness transform kernel
    for variant, waves, bpc, quad in itertools.product([1, 0], [16, 8], bpcs, [1 << 17, 0]):
        if variant == 0 and waves != 16:
            continue  # the dword kernel has a fixed 16-wave block
        configs.append({"fused": 0, "mac_variant": variant, "mac_waves": waves,
                        "mac_blocks_per_cu": bpc, "quad_max": quad})
    # one second of commitments first, so the first configurations are not
    # measured while the clocks ramp
    a = time.perf_counter()
    i = 0
    while time.perf_counter() - a < 1.0:
        keys[i % 3].commit_device(ws[i % 4], n, ts[i % 4])
        i += 1
        if i % 256 == 0:
            ctx.synchronize()
    ctx.synchronize()
    ref = None
    results = []
    for cfg in configs:
        ctx.set_tuning(**cfg)
        for i in range(6):
            keys[i % 3].commit_device(ws[i % 4], n, ts[i % 4])
        ctx.synchronize()
        keys[0].commit_device(ws[0], n, ts[0])
        ctx.synchronize()
        out = ts[0].download(np.uint32, (kappa, 64))
        if ref is None:
            ref = out
        elif not np.array_equal(ref, out):
            raise RuntimeError(f"tuning {cfg} changed the result")
        a = time.perf_counter()
        for i in range(reps):
            keys[i % 3].commit_device(ws[i % 4], n, ts[i % 4])
        ctx.synchronize()
        us = (time.perf_counter() - a) / reps * 1e6
        ctx.profile(True)
        for i in range(reps // 2):
            keys[i % 3].commit_device(ws[i % 4], n, ts[i % 4])
        prof = ctx.profile_read()
        ctx.profile(False)
        phase_us = {ph: 1e3 * ms / max(c, 1) for ph, (ms, c) in prof.items()}
        results.append({"config": cfg, "us_per_commit": round(us, 2),
                        "ntt_us": round(phase_us[0], 2), "mac_us": round(phase_us[1], 2),
                        "finalize_us": round(phase_us[2], 2)})
        log(f"{cfg} -> {us:7.2f} us/commit (ntt {phase_us[0]:.1f}, mac {phase_us[1]:.1f}, "
            f"fin {phase_us[2]:.1f})")
    for k in keys:
        k.close()
    for b in ws + ts:
        b.free()
    ring.close()
    results.sort(key=lambda r: r["us_per_commit"])
    return results


def tune_batch(ctx, quick=False, log=print):
    """Sweep chunk width x batch_group on two batched workloads: (a) 16 q32
    commitments at kappa = 16, n = 2^15 per call; (b) 2000 Goldilocks
    commitments, kappa = 16, log-uniform lengths in [2^8, 2^14] (the C4
    workload). The group is shared; each field gets the chunk width fastest
    for its workload at that group, and the group with the least summed time
    wins. Results must not change with the tuning. Returns (sweep, best)."""
    import lattica_amd as la
    rng = np.random.default_rng(9)
    work = []
    # (a) q32 same shape
    rq = la.Ring(ctx, la.Q32, 64)
    kq = la.Key(rq, rng.integers(0, la.Q32, size=(16, 1 << 15, 64), dtype=np.uint32),
                domain=la.DOMAIN_NTT)
    wq = [ctx.to_device(_witness(rng, la.Q32, 1 << 15)) for _ in range(4)]
    oq = [ctx.alloc(16 * 64 * 4) for _ in range(16)]
    work.append(("q32_16x32768", kq, [wq[i % 4] for i in range(16)], [1 << 15] * 16, oq, 4 * 16 * 64))
    # (b) Goldilocks, C4 length distribution
    p = la.GOLDILOCKS
    rg = la.Ring(ctx, p, 64)
    hi = rng.integers(0, 1 << 32, size=(16, 1 << 14, 64), dtype=np.uint64)
    lo = rng.integers(0, 1 << 32, size=(16, 1 << 14, 64), dtype=np.uint64)
    kg = la.Key(rg, ((hi << np.uint64(32)) | lo) % np.uint64(p), domain=la.DOMAIN_NTT)
    del hi, lo
    t = rng.integers(-1, 2, size=(1 << 15, 64), dtype=np.int64)
    pool = ctx.to_device(np.where(t < 0, np.uint64(p - 1), t.astype(np.uint64)))
    n_items = 2000
    lens = np.floor(2.0 ** rng.uniform(8, 14, size=n_items)).astype(np.int64)
    offs = rng.integers(0, (1 << 15) - lens + 1)
    og = ctx.alloc(n_items * 16 * 64 * 8)
    work.append(("gold_c4_2000", kg, [pool.ptr.value + int(o) * 512 for o in offs], list(lens),
                 [og.ptr.value + i * 16 * 64 * 8 for i in range(n_items)], None))

    chunks = [64, 128, 256] if quick else [32, 64, 128, 256, 512]
    groups = [4, 8, 16]
    results = []
    refs = {}
    for cc, g in itertools.product(chunks, groups):
        ctx.set_tuning(batch_chunk_cols=cc, batch_chunk_cols_gold=cc, batch_group=g)
        total_us, per, raw = 0.0, {}, {}
        for name, key, ws, ls, os_, probe in work:
            key.commit_batch_device(ws, ls, os_)
            ctx.synchronize()
            out0 = (os_[0].download(np.uint32, (16, 64)) if probe else
                    og.download(np.uint64, (4, 16, 64)))
            if name in refs and not np.array_equal(refs[name], out0):
                raise RuntimeError(f"batch tuning {cc, g} changed the result of {name}")
            refs.setdefault(name, out0)
            reps = 3
            a = time.perf_counter()
            for _ in range(reps):
                key.commit_batch_device(ws, ls, os_)
            ctx.synchronize()
            us = (time.perf_counter() - a) / reps * 1e6
            per[name] = round(us / len(ls), 3)
            raw[name] = us
            total_us += us
        results.append({"config": {"batch_chunk_cols": cc, "batch_group": g},
                        "us_per_commit": per, "total_us": round(total_us, 1), "us": raw})
        log(f"batch chunk {cc:4d} group {g:2d} -> {per}")
    for b in wq + oq + [pool, og]:
        b.free()
    kq.close()
    kg.close()
    rq.close()
    rg.close()
    best = None

    def narrowest_near_best(rows, name, tol=1.02):
        # among chunk widths within 2 % of the fastest, the narrowest: it gives
        # small batches and the server's waves more blocks at the same speed
        t = min(r["us"][name] for r in rows)
        near = [r for r in rows if r["us"][name] <= t * tol]
        return min(near, key=lambda r: r["config"]["batch_chunk_cols"])

    for g in groups:
        rows = [r for r in results if r["config"]["batch_group"] == g]
        rq = narrowest_near_best(rows, "q32_16x32768")
        rg_ = narrowest_near_best(rows, "gold_c4_2000")
        total = rq["us"]["q32_16x32768"] + rg_["us"]["gold_c4_2000"]
        if best is None or total < best["total_us"]:
            best = {"config": {"batch_chunk_cols": rq["config"]["batch_chunk_cols"],
                               "batch_chunk_cols_gold": rg_["config"]["batch_chunk_cols"],
                               "batch_group": g},
                    "us_per_commit": {"q32_16x32768": rq["us_per_commit"]["q32_16x32768"],
                                      "gold_c4_2000": rg_["us_per_commit"]["gold_c4_2000"]},
                    "total_us": round(total, 1)}
    for r in results:
        del r["us"]
    results.sort(key=lambda r: r["total_us"])
    return results, best


def main(argv=None):
    ap = argparse.ArgumentParser(description=__doc__)
    ap.add_argument("--quick", action="store_true")
    ap.add_argument("--out", default=TUNING_FILE)
    args = ap.parse_args(argv)
    import lattica_amd as la
    ctx = la.Context(0, apply_tuning=False)
    res = tune_commit(ctx, quick=args.quick, log=lambda m: print(m, flush=True))
    best = res[0]
    ctx.set_tuning(**best["config"])
    bres, bbest = tune_batch(ctx, quick=args.quick, log=lambda m: print(m, flush=True))
    cfg = dict(best["config"])
    cfg.update(bbest["config"])
    doc = {"device_cus": ctx.num_cus, "workload": "commit_q32_d64_kappa16_n32768",
           "best": cfg, "best_us_per_commit": best["us_per_commit"],
           "best_batch": bbest, "sweep": res, "batch_sweep": bres}
    with open(args.out, "w") as fh:
        json.dump(doc, fh, indent=1)
    print(json.dumps({"best": best}), flush=True)
    ctx.close()


if __name__ == "__main__":
    sys.exit(main())
