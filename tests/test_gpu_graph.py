"""Batch plans (gh_batch_plan_*): a batch recorded once into a HIP graph and
replayed. Every replay must equal the eager batched call (gh_commit_batch) on
the witness contents current at that replay, with eager library work and
synchronisations between replays; small items are checked against the oracle.
"""
import numpy as np
import pytest

from conftest import rand_residues, small_signed

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("name", ["q32", "goldilocks"])
def test_graph_replays_equal_eager_batches(oracle, ctx, ring_q32, ring_gold, name):
    import lattica_amd as la
    ring = ring_q32 if name == "q32" else ring_gold
    rng = np.random.default_rng(31 if name == "q32" else 32)
    rows, cols = 16, 700
    a = rand_residues(rng, ring.q, (rows, cols, 64))
    key = la.Key(ring, a, domain=la.DOMAIN_NTT)
    lens = [700, 1, 64, 65, 333, 129, 700, 2, 400]
    wb = np.dtype(ring.dtype).itemsize
    wbufs = [ctx.alloc(n * 64 * wb) for n in lens]
    plan_out = [ctx.alloc(rows * 64 * wb) for _ in lens]
    eager_out = [ctx.alloc(rows * 64 * wb) for _ in lens]
    plan = key.batch_plan(wbufs, lens, plan_out)
    other = key.commit_batch([small_signed(rng, ring.q, (50, 64), -128, 128)])  # eager work
    try:
        for rep in range(3):
            ws = [small_signed(rng, ring.q, (n, 64), -128, 128).astype(ring.dtype) for n in lens]
            for b, w in zip(wbufs, ws):
                b.upload(w)
            plan.launch()
            ctx.synchronize()
            key.commit_batch_device(wbufs, lens, eager_out)  # eager batch on the same data
            again = key.commit_batch([small_signed(rng, ring.q, (50, 64), -128, 128)])
            assert again[0].shape == other[0].shape
            ctx.synchronize()
            for i, n in enumerate(lens):
                got = plan_out[i].download(ring.dtype, (rows, 64))
                assert np.array_equal(got, eager_out[i].download(ring.dtype, (rows, 64))), (rep, i)
                if n <= 2:
                    z = np.zeros((rows, 1, 64), dtype=np.uint64)
                    exp = oracle.commit(ring.q, 64, a[:, :n], z, ws[i].astype(np.uint64),
                                        np.zeros((1, 64), dtype=np.uint64), key_domain=1)
                    assert np.array_equal(got.astype(np.uint64), exp), (rep, i)
    finally:
        plan.close()
        for b in wbufs + plan_out + eager_out:
            b.free()
        key.close()


def test_graph_plan_rejects_bad_items(ctx, ring_q32):
    import lattica_amd as la
    rng = np.random.default_rng(3)
    key = la.Key(ring_q32, rand_residues(rng, ring_q32.q, (16, 8, 64)), domain=la.DOMAIN_NTT)
    w, t = ctx.alloc(16 * 64 * 4), ctx.alloc(16 * 64 * 4)
    try:
        with pytest.raises(la.GridHipError) as e:
            key.batch_plan([w], [9], [t])  # more columns than the key holds
        assert e.value.name == "DIMENSION_MISMATCH"
    finally:
        w.free()
        t.free()
        key.close()


@pytest.mark.parametrize("name", ["q32", "goldilocks"])
@pytest.mark.parametrize("kappa", [1, 3, 6])
def test_plans_and_batches_take_any_row_count(oracle, ctx, ring_q32, ring_gold, name, kappa):
    """Keys with kappa not a multiple of 4 (the reference benches use kappa = 1):
    the grouped batch, its graph replay and the per-call form agree with the
    oracle."""
    import lattica_amd as la
    ring = ring_q32 if name == "q32" else ring_gold
    rng = np.random.default_rng(33 + kappa)
    a = rand_residues(rng, ring.q, (kappa, 300, 64))
    key = la.Key(ring, a, domain=la.DOMAIN_NTT)
    lens = [300, 1, 65, 200, 64, 7, 299, 128, 3]
    wb = np.dtype(ring.dtype).itemsize
    ws = [small_signed(rng, ring.q, (n, 64), -128, 128).astype(ring.dtype) for n in lens]
    wbufs = [ctx.to_device(w) for w in ws]
    plan_out = [ctx.alloc(kappa * 64 * wb) for _ in lens]
    plan = key.batch_plan(wbufs, lens, plan_out)
    try:
        eager = key.commit_batch(ws)
        for rep in range(3):
            plan.launch()
            ctx.synchronize()
            for i, n in enumerate(lens):
                got = plan_out[i].download(ring.dtype, (kappa, 64))
                assert np.array_equal(got, eager[i]), (rep, i)
        for i, n in enumerate(lens):
            z = np.zeros((kappa, 1, 64), dtype=np.uint64)
            exp = oracle.commit(ring.q, 64, a[:, :n], z, ws[i].astype(np.uint64),
                                np.zeros((1, 64), dtype=np.uint64), key_domain=1)
            assert np.array_equal(eager[i].astype(np.uint64), exp), i
            assert np.array_equal(key.commit(ws[i]) if n == 300 else key.commit_batch([ws[i]])[0],
                                  eager[i]), i
    finally:
        plan.close()
        for b in wbufs + plan_out:
            b.free()
        key.close()
