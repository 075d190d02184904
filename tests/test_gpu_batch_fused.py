"""The one-launch Goldilocks batch (k_batch_fused_gold, tuning batch_fused = 1)
against the three-kernel batch (batch_fused = 0) and the oracle.

Every row count from 1 to 20 that the reference's benches use or that leaves a
ragged row block, ragged lengths (1, 15, 16, 17, ranges split mid sub-chunk),
batch sizes that leave a short last group, several column-range widths, and
the completion counters: repeated eager batches and graph replays reuse them,
so each launch must leave them at zero.
"""
import numpy as np
import pytest

from conftest import GOLDILOCKS, rand_residues, small_signed

pytestmark = pytest.mark.gpu


def _batch(ctx, key, ws, lens, fused, cols=None):
    """Device batch of witnesses ws under the given batch_fused setting."""
    ring = key.ring
    saved = ctx.get_tuning()
    ctx.set_tuning(batch_fused=fused, **({"batch_fused_cols": cols} if cols else {}))
    bufs = [ctx.to_device(w) for w in ws]
    outs = [ctx.alloc(key.rows * 64 * 8) for _ in ws]
    try:
        key.commit_batch_device(bufs, lens, outs)
        ctx.synchronize()
        return [o.download(ring.dtype, (key.rows, 64)) for o in outs]
    finally:
        ctx.set_tuning(batch_fused=saved["batch_fused"],
                       batch_fused_cols=saved["batch_fused_cols"])
        for b in bufs + outs:
            b.free()


@pytest.mark.parametrize("rows", [1, 3, 4, 8, 16, 20])
def test_fused_batch_equals_three_kernel_batch(oracle, ctx, ring_gold, rows):
    import lattica_amd as la
    p = GOLDILOCKS
    rng = np.random.default_rng(900 + rows)
    cols = 600
    a = rand_residues(rng, p, (rows, cols, 64))
    key = la.Key(ring_gold, a, domain=la.DOMAIN_NTT)
    lens = [600, 1, 15, 16, 17, 255, 256, 257, 511, 2, 33]
    ws = [small_signed(rng, p, (n, 64), -1, 2).astype(np.uint64) for n in lens]
    try:
        three = _batch(ctx, key, ws, lens, 0)
        for rc in (32, 96, 256):
            fused = _batch(ctx, key, ws, lens, 1, rc)
            for i in range(len(lens)):
                assert np.array_equal(fused[i], three[i]), (rows, rc, i, lens[i])
        z = np.zeros((rows, 1, 64), dtype=np.uint64)
        zr = np.zeros((1, 64), dtype=np.uint64)
        for i in (0, 1, 4, 9):
            exp = oracle.commit(p, 64, a[:, :lens[i]], z, ws[i], zr, key_domain=1)
            assert np.array_equal(three[i], exp), (rows, i)
    finally:
        key.close()


def test_fused_batch_counters_survive_repeats_and_replays(ctx, ring_gold):
    """Eager batches of different sizes back to back with no synchronize,
    then a plan replayed four times with new witness contents: every result
    equals the three-kernel batch on the same inputs."""
    import lattica_amd as la
    p = GOLDILOCKS
    rng = np.random.default_rng(77)
    rows, cols = 16, 1024
    key = la.Key(ring_gold, rand_residues(rng, p, (rows, cols, 64)), domain=la.DOMAIN_NTT)
    sizes = [9, 2, 37, 5]
    batches = []
    for n in sizes:
        lens = [int(x) for x in rng.integers(1, cols + 1, size=n)]
        ws = [small_signed(rng, p, (m, 64), -1, 2).astype(np.uint64) for m in lens]
        batches.append((lens, ws))
    try:
        want = [_batch(ctx, key, ws, lens, 0) for lens, ws in batches]
        ctx.set_tuning(batch_fused=1)
        held = []
        for lens, ws in batches:  # back to back, one synchronize at the end
            bufs = [ctx.to_device(w) for w in ws]
            outs = [ctx.alloc(rows * 64 * 8) for _ in ws]
            key.commit_batch_device(bufs, lens, outs)
            held.append((bufs, outs))
        ctx.synchronize()
        for (bufs, outs), exp in zip(held, want):
            for o, e in zip(outs, exp):
                assert np.array_equal(o.download(np.uint64, (rows, 64)), e)
            for b in bufs + outs:
                b.free()
        lens = [1024, 700, 700, 3, 64, 1000, 17]
        wbufs = [ctx.alloc(n * 64 * 8) for n in lens]
        outs = [ctx.alloc(rows * 64 * 8) for _ in lens]
        plan = key.batch_plan(wbufs, lens, outs)
        try:
            for rep in range(4):
                ws = [small_signed(rng, p, (n, 64), -1, 2).astype(np.uint64) for n in lens]
                for b, w in zip(wbufs, ws):
                    b.upload(w)
                plan.launch()
                ctx.synchronize()
                exp = _batch(ctx, key, ws, lens, 0)
                ctx.set_tuning(batch_fused=1)
                for i in range(len(lens)):
                    assert np.array_equal(outs[i].download(np.uint64, (rows, 64)), exp[i]), (rep, i)
        finally:
            plan.close()
            for b in wbufs + outs:
                b.free()
    finally:
        ctx.set_tuning(batch_fused=0)
        key.close()


def test_fused_batch_tuning_values_checked(ctx):
    import lattica_amd as la
    for bad in ({"batch_fused": 2}, {"batch_fused": -1}, {"batch_fused_cols": 16},
                {"batch_fused_cols": 100}, {"batch_fused_cols": 8192}):
        with pytest.raises(la.GridHipError):
            ctx.set_tuning(**bad)
