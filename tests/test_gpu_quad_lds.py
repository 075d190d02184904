"""The four-lane transforms with their 4-point stage over the lanes done
through LDS (tuning quad_lds = 1, the default: quad_fwd64_lds and
quad_inv64_lds) against the lane-exchange form (quad_lds = 0) and the oracle:
twisted transforms, the quad negacyclic product, plain commitments (the q32
row-split kernel with 64 columns per block and with a ragged last block, the
Goldilocks per-call kernel) and batches (the segmented witness transform), for
q32 and Goldilocks.

Counts that leave partial quads and blocks (1, 65, 1000); full-range residues
with the largest in the first polynomial; round trips.
"""
import numpy as np
import pytest

import lattica_amd as la
from conftest import rand_residues

pytestmark = pytest.mark.gpu


def ring_for(name, ring_q32, ring_gold):
    return ring_q32 if name == "q32" else ring_gold


def _with(ctx, fn, **kv):
    saved = ctx.get_tuning()
    ctx.set_tuning(**kv)
    try:
        return fn()
    finally:
        ctx.set_tuning(**{k: saved[k] for k in kv})


def test_quad_lds_is_default_and_checked(ctx):
    assert ctx.get_tuning()["quad_lds"] == 1
    for bad in (-1, 2):
        with pytest.raises(la.GridHipError):
            ctx.set_tuning(quad_lds=bad)


@pytest.mark.parametrize("name", ["q32", "goldilocks"])
@pytest.mark.parametrize("count", [1, 65, 1000])
def test_transforms(oracle, ctx, ring_q32, ring_gold, name, count):
    r = ring_for(name, ring_q32, ring_gold)
    rng = np.random.default_rng(4100 + count)
    x = rand_residues(rng, r.q, (count, 64))
    x[0] = r.q - 1
    xw = r.words(x)
    fwd = _with(ctx, lambda: r.twisted_forward(xw), quad_lds=1, xform_quad=1)
    assert np.array_equal(fwd.astype(np.uint64), oracle.twisted_forward(r.q, 64, x))
    assert np.array_equal(fwd, _with(ctx, lambda: r.twisted_forward(xw), quad_lds=0, xform_quad=1))
    assert np.array_equal(_with(ctx, lambda: r.twisted_inverse(fwd), quad_lds=1, xform_quad=1), xw)
    y = r.words(rand_residues(rng, r.q, (count, 64)))
    assert np.array_equal(_with(ctx, lambda: r.twisted_inverse(y), quad_lds=1, xform_quad=1),
                          _with(ctx, lambda: r.twisted_inverse(y), quad_lds=0, xform_quad=1))


@pytest.mark.parametrize("name", ["q32", "goldilocks"])
@pytest.mark.parametrize("count", [1, 65, 1000])
def test_negacyclic(oracle, ctx, ring_q32, ring_gold, name, count):
    r = ring_for(name, ring_q32, ring_gold)
    rng = np.random.default_rng(4200 + count)
    a, b = rand_residues(rng, r.q, (count, 64)), rand_residues(rng, r.q, (count, 64))
    a[0], b[0] = r.q - 1, r.q - 1
    got = _with(ctx, lambda: r.negacyclic_mul(a, b), quad_lds=1, negacyclic_quad=1)
    assert np.array_equal(got, _with(ctx, lambda: r.negacyclic_mul(a, b), quad_lds=0, negacyclic_quad=1))
    assert np.array_equal(got.astype(np.uint64), oracle.negacyclic_mul(r.q, 64, a, b))
    # the default (-1) takes the quad form for both fields
    assert np.array_equal(got, _with(ctx, lambda: r.negacyclic_mul(a, b), quad_lds=1, negacyclic_quad=-1))


@pytest.mark.parametrize("name", ["q32", "goldilocks"])
@pytest.mark.parametrize("msg", [1000, 1 << 14])
def test_commit(oracle, ctx, ring_q32, ring_gold, name, msg):
    r = ring_for(name, ring_q32, ring_gold)
    rng = np.random.default_rng(4300 + msg)
    rows = 16
    am = rand_residues(rng, r.q, (rows, msg, 64))
    m = rand_residues(rng, r.q, (msg, 64))
    m[0] = r.q - 1
    key = la.Key(r, am)
    try:
        got = _with(ctx, lambda: key.commit(m), quad_lds=1)
        assert np.array_equal(got, _with(ctx, lambda: key.commit(m), quad_lds=0))
        exp = oracle.commit(r.q, 64, am, np.zeros((rows, 1, 64), dtype=np.uint64), m,
                            np.zeros((1, 64), dtype=np.uint64))
        assert np.array_equal(got.astype(np.uint64), exp)
    finally:
        key.close()


@pytest.mark.parametrize("name", ["q32", "goldilocks"])
def test_batch(oracle, ctx, ring_q32, ring_gold, name):
    r = ring_for(name, ring_q32, ring_gold)
    rng = np.random.default_rng(4400)
    rows, n = 16, 1200
    am = rand_residues(rng, r.q, (rows, n, 64))
    lens = [1, 17, 300, 1200, 65, 1000, 7, 640, 1199]
    ws = [rand_residues(rng, r.q, (k, 64)) for k in lens]
    key = la.Key(r, am)
    try:
        got = _with(ctx, lambda: key.commit_batch(ws), quad_lds=1)
        ref = _with(ctx, lambda: key.commit_batch(ws), quad_lds=0)
        for g, e in zip(got, ref):
            assert np.array_equal(g, e)
        for i in (0, 3):
            k = lens[i]
            exp = oracle.commit(r.q, 64, am[:, :k], np.zeros((rows, 1, 64), dtype=np.uint64), ws[i],
                                np.zeros((1, 64), dtype=np.uint64))
            assert np.array_equal(got[i].astype(np.uint64), exp)
    finally:
        key.close()
