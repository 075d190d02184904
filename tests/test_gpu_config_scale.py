"""Parity at the configured sizes, against the oracle.

* LaBRADOR-scale witness, 32-bit q, kappa 16, r = 32 parts x 1024 polynomials
  (n = 2^15) with coefficients in [-2^7, 2^7): all 32 inner commitments
  t_i = A * s_i (labrador/src/crs.rs:211-216), the full Gram upper triangle
  (main_protocol/garbage.rs:76-93) and the 256-row JL projection over all
  2^21 columns (labrador/src/jl.rs:388-508), p mod q and the exact centred p.
* 2000 Goldilocks commitments, kappa 16, witness lengths log-uniform in
  [2^8, 2^14], ternary coefficients, one shared NTT-domain key of 2^14
  columns: the eager batch, three replays of its HIP graph and the per-call
  form agree on every item, and 20 items (the longest among them) equal the
  oracle.
"""
import numpy as np
import pytest

from conftest import GOLDILOCKS, Q32, rand_residues, small_signed

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def c3_witness():
    rng = np.random.default_rng(3303)
    q = Q32
    a = rand_residues(rng, q, (16, 1024, 64))
    s = small_signed(rng, q, (32, 1024, 64), -128, 128)
    return a, s


def test_c3_all_inner_commitments(oracle, ring_q32, c3_witness):
    import lattica_amd as la
    a, s = c3_witness
    print("case: inner commitments kappa 16, r 32 x 1024", flush=True)
    key = la.Key(ring_q32, a)
    try:
        got = key.inner_commit(s)
    finally:
        key.close()
    z = np.zeros((16, 1, 64), dtype=np.uint64)
    zr = np.zeros((1, 64), dtype=np.uint64)
    for i in range(32):
        exp = oracle.commit(Q32, 64, a, z, s[i], zr)
        assert np.array_equal(got[i].astype(np.uint64), exp), i


def test_c3_full_gram(oracle, ring_q32, c3_witness):
    _, s = c3_witness
    print("case: Gram r 32 x 1024", flush=True)
    got = ring_q32.gram(s)
    exp = oracle.gram(Q32, 64, s)
    assert got.shape == (528, 64)
    assert np.array_equal(got.astype(np.uint64), exp)


def test_c3_jl_projection_all_columns(oracle, ring_q32, c3_witness):
    """256 rows over 32 parts x 65536 columns: the matrix-core path."""
    _, s = c3_witness
    print("case: JL projection 256 x 2^21", flush=True)
    seed = bytes((7 * i + 3) & 0xFF for i in range(32))
    parts = s.reshape(32, 1024 * 64)
    pm, pe, _ = ring_q32.jl_project(seed, 256, parts)
    om, oe, _ = oracle.jl_project(Q32, seed, 256, parts, 64)
    assert np.array_equal(pe, oe)
    assert np.array_equal(pm.astype(np.uint64), om)


def test_c4_goldilocks_2000_commitments(oracle, ctx, ring_gold):
    import lattica_amd as la
    p = GOLDILOCKS
    rng = np.random.default_rng(4404)
    kappa, cols, count = 16, 1 << 14, 2000
    print("case: 2000 Goldilocks commitments, kappa 16, lengths in [2^8, 2^14]", flush=True)
    a = rand_residues(rng, p, (kappa, cols, 64))
    key = la.Key(ring_gold, a, domain=la.DOMAIN_NTT)
    pool_polys = 1 << 15
    t = rng.integers(-1, 2, size=(pool_polys, 64), dtype=np.int64)
    pool = np.where(t < 0, np.uint64(p - 1), t.astype(np.uint64))
    lens = np.floor(2.0 ** rng.uniform(8, 14, size=count)).astype(np.int64)
    lens[int(rng.integers(0, count))] = cols  # the longest the key allows
    offs = rng.integers(0, pool_polys - lens + 1)
    pool_dev = ctx.to_device(pool)
    wptr = [pool_dev.ptr.value + int(o) * 64 * 8 for o in offs]
    outs = [ctx.alloc(count * kappa * 64 * 8) for _ in range(3)]
    optr = [[b.ptr.value + i * kappa * 64 * 8 for i in range(count)] for b in outs]
    plan = None
    try:
        key.commit_batch_device(wptr, lens, optr[0])
        ctx.synchronize()
        eager = outs[0].download(np.uint64, (count, kappa, 64))
        plan = key.batch_plan(wptr, lens, optr[1])
        for rep in range(3):
            plan.launch()
            ctx.synchronize()
            graph = outs[1].download(np.uint64, (count, kappa, 64))
            assert np.array_equal(graph, eager), rep
        for i in range(count):  # per call: one-item device batches, finalizes held back
            key.commit_batch_device([wptr[i]], [lens[i]], [optr[2][i]])
        ctx.synchronize()
        per_call = outs[2].download(np.uint64, (count, kappa, 64))
        assert np.array_equal(per_call, eager)
        sample = {int(np.argmax(lens))} | {int(x) for x in rng.choice(count, 19, replace=False)}
        z = np.zeros((kappa, 1, 64), dtype=np.uint64)
        zr = np.zeros((1, 64), dtype=np.uint64)
        for i in sorted(sample):
            n, o = int(lens[i]), int(offs[i])
            exp = oracle.commit(p, 64, a[:, :n], z, pool[o:o + n], zr, key_domain=1)
            assert np.array_equal(eager[i], exp), (i, n)
    finally:
        if plan is not None:
            plan.close()
        for b in outs:
            b.free()
        pool_dev.free()
        key.close()
