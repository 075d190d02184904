"""Twisted transforms with four lanes per polynomial at every size
(k_twisted_fwd_quad, k_twisted_inv_quad; tuning xform_quad = 1) against the
oracle and the other kernels (xform_quad = 0), for q32 and Goldilocks.

Counts that leave a partial block (1, 3, 65, 1000) and one above quad_max
(2^17 + 5), where xform_quad = 0 takes the polynomial-per-thread kernels in
both directions; the largest residues in the first polynomial; round trips.
"""
import numpy as np
import pytest

from conftest import rand_residues

pytestmark = pytest.mark.gpu


def ring_for(name, ring_q32, ring_gold):
    return ring_q32 if name == "q32" else ring_gold


def _xf(ctx, r, x, quad, inverse):
    saved = ctx.get_tuning()["xform_quad"]
    ctx.set_tuning(xform_quad=quad)
    try:
        return r.twisted_inverse(x) if inverse else r.twisted_forward(x)
    finally:
        ctx.set_tuning(xform_quad=saved)


@pytest.mark.parametrize("name", ["q32", "goldilocks"])
@pytest.mark.parametrize("count", [1, 3, 65, 1000])
def test_quad_transforms_match_oracle(oracle, ctx, ring_q32, ring_gold, name, count):
    r = ring_for(name, ring_q32, ring_gold)
    rng = np.random.default_rng(3000 + count)
    x = rand_residues(rng, r.q, (count, 64))
    x[0] = r.q - 1
    xw = r.words(x)
    fwd = _xf(ctx, r, xw, 1, False)
    assert np.array_equal(fwd.astype(np.uint64), oracle.twisted_forward(r.q, 64, x))
    assert np.array_equal(fwd, _xf(ctx, r, xw, 0, False))
    inv = _xf(ctx, r, fwd, 1, True)
    assert np.array_equal(inv, xw)
    y = r.words(rand_residues(rng, r.q, (count, 64)))
    assert np.array_equal(_xf(ctx, r, y, 1, True), _xf(ctx, r, y, 0, True))


@pytest.mark.parametrize("name", ["q32", "goldilocks"])
def test_quad_transforms_above_quad_max(oracle, ctx, ring_q32, ring_gold, name):
    r = ring_for(name, ring_q32, ring_gold)
    count = ctx.get_tuning()["quad_max"] + 5
    rng = np.random.default_rng(31)
    x = r.words(rand_residues(rng, r.q, (count, 64)))
    fwd = _xf(ctx, r, x, 1, False)
    assert np.array_equal(fwd, _xf(ctx, r, x, 0, False))
    inv = _xf(ctx, r, fwd, 1, True)
    assert np.array_equal(inv, x)
    assert np.array_equal(inv, _xf(ctx, r, fwd, 0, True))
    idx = rng.choice(count, 50, replace=False)
    assert np.array_equal(fwd[idx].astype(np.uint64), oracle.twisted_forward(r.q, 64, x[idx].astype(np.uint64)))


def test_xform_quad_values_checked(ctx):
    import lattica_amd as la
    assert ctx.get_tuning()["xform_quad"] in (-1, 0, 1)
    for bad in (2, -2):
        with pytest.raises(la.GridHipError):
            ctx.set_tuning(xform_quad=bad)
