"""The grouped MACs with their LDS ring (k_mac_batch_q32_ring and
k_mac_batch_gold_ring, tuning batch_mac_ring = 2..4) against the
register-pipelined kernels (batch_mac_ring = 0) and the oracle, for q32 and
Goldilocks keys.

Full-range residues in key and witness (the exact sums' largest terms), ragged
lengths (1, 2, 3, 5 and lengths around the chunk width), a batch whose last
group is short, chunk widths below, at and above the ring depth, row counts
that take the ring (multiples of 16) or fall back (20), and batches with
enough groups that each block runs through several chunks.
"""
import numpy as np
import pytest

from conftest import rand_residues

pytestmark = pytest.mark.gpu


def ring_for(name, ring_q32, ring_gold):
    return ring_q32 if name == "q32" else ring_gold


def _batch(ctx, key, ws, lens, ring, cols, group=8):
    ring_field = key.ring
    saved = ctx.get_tuning()
    ctx.set_tuning(batch_mac_ring=ring, batch_chunk_cols=cols, batch_chunk_cols_gold=cols,
                   q32_ring_group=group)
    bufs = [ctx.to_device(ring_field.words(w)) for w in ws]
    wb = np.dtype(ring_field.dtype).itemsize
    outs = [ctx.alloc(key.rows * 64 * wb) for _ in ws]
    try:
        key.commit_batch_device(bufs, lens, outs)
        ctx.synchronize()
        return [o.download(ring_field.dtype, (key.rows, 64)) for o in outs]
    finally:
        ctx.set_tuning(batch_mac_ring=saved["batch_mac_ring"],
                       q32_ring_group=saved["q32_ring_group"],
                       batch_chunk_cols=saved["batch_chunk_cols"],
                       batch_chunk_cols_gold=saved["batch_chunk_cols_gold"])
        for b in bufs + outs:
            b.free()


@pytest.mark.parametrize("name", ["q32", "goldilocks"])
@pytest.mark.parametrize("rows", [16, 32, 20])
def test_ring_mac_equals_register_mac(oracle, ctx, ring_q32, ring_gold, name, rows):
    import lattica_amd as la
    r = ring_for(name, ring_q32, ring_gold)
    p = r.q
    rng = np.random.default_rng(1300 + rows)
    cols = 700
    a = rand_residues(rng, p, (rows, cols, 64))
    key = la.Key(r, r.words(a), domain=la.DOMAIN_NTT)
    lens = [700, 1, 2, 3, 5, 63, 64, 65, 129, 300, 17]
    ws = [rand_residues(rng, p, (n, 64)) for n in lens]
    try:
        groups = (4, 8) if name == "q32" else (4,)
        for cc in (16, 64, 256):
            base = _batch(ctx, key, ws, lens, 0, cc)
            for depth in (2, 3, 4):
                for g in groups:
                    got = _batch(ctx, key, ws, lens, depth, cc, g)
                    for i in range(len(lens)):
                        assert np.array_equal(got[i], base[i]), (name, rows, cc, depth, g, i, lens[i])
        z = np.zeros((rows, 1, 64), dtype=np.uint64)
        zr = np.zeros((1, 64), dtype=np.uint64)
        got = _batch(ctx, key, ws, lens, 4, 64)
        for i in (0, 1, 3, 6, 10):
            exp = oracle.commit(p, 64, a[:, :lens[i]], z, ws[i], zr, key_domain=1)
            assert np.array_equal(got[i].astype(np.uint64), exp), (name, rows, i)
    finally:
        key.close()


def test_ring_tuning_values_checked(ctx):
    import lattica_amd as la
    for bad in (1, 5, -1):
        with pytest.raises(la.GridHipError):
            ctx.set_tuning(batch_mac_ring=bad)
    assert ctx.get_tuning()["batch_mac_ring"] in (0, 2, 3, 4)
    for bad in (2, 6, 16):
        with pytest.raises(la.GridHipError):
            ctx.set_tuning(q32_ring_group=bad)


@pytest.mark.parametrize("name", ["q32", "goldilocks"])
def test_ring_mac_chunk_runs(oracle, ctx, ring_q32, ring_gold, name):
    # enough groups that each block runs through several chunks of its group
    # (runs per group = about 8 * 2 * num_cus / groups, below the chunk count)
    import lattica_amd as la
    r = ring_for(name, ring_q32, ring_gold)
    p = r.q
    rng = np.random.default_rng(1400)
    rows, cols = 16, 600
    a = rand_residues(rng, p, (rows, cols, 64))
    key = la.Key(r, r.words(a), domain=la.DOMAIN_NTT)
    lens = [int(x) for x in rng.integers(1, cols + 1, size=1200)]
    lens[0] = cols
    ws = [rand_residues(rng, p, (n, 64)) for n in lens]
    try:
        for cc in (16, 32):
            groups = (len(lens) + 3) // 4
            assert -(-8 * 2 * ctx.num_cus // groups) < -(-cols // cc)  # runs of several chunks
            base = _batch(ctx, key, ws, lens, 0, cc)
            for depth in (2, 4):
                for g in ((4, 8) if name == "q32" else (4,)):
                    got = _batch(ctx, key, ws, lens, depth, cc, g)
                    for i in range(len(lens)):
                        assert np.array_equal(got[i], base[i]), (name, cc, depth, g, i, lens[i])
        z = np.zeros((rows, 1, 64), dtype=np.uint64)
        zr = np.zeros((1, 64), dtype=np.uint64)
        for i in (0, 1, 777):
            exp = oracle.commit(p, 64, a[:, :lens[i]], z, ws[i], zr, key_domain=1)
            assert np.array_equal(got[i].astype(np.uint64), exp), (name, i)
    finally:
        key.close()
