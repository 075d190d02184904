"""The negacyclic product with one pair per quad of lanes (k_negacyclic_quad,
tuning negacyclic_quad = 1) against the golden vectors, the oracle and the
one-pair-per-thread kernel (negacyclic_quad = 0), for q32 and Goldilocks.

Counts that leave a partial block (1, 3, 65, 1000) and one that makes every
quad loop over several pairs (2^18 for q32); the largest residues in the
first pair; X * X^63 = -1 and the zero product.
"""
import numpy as np
import pytest

from conftest import rand_residues

pytestmark = pytest.mark.gpu


def ring_for(name, ring_q32, ring_gold):
    return ring_q32 if name == "q32" else ring_gold


def _mul(ctx, r, a, b, quad):
    saved = ctx.get_tuning()["negacyclic_quad"]
    ctx.set_tuning(negacyclic_quad=quad)
    try:
        return r.negacyclic_mul(a, b)
    finally:
        ctx.set_tuning(negacyclic_quad=saved)


@pytest.mark.parametrize("name", ["q32", "goldilocks"])
def test_quad_golden_and_edges(golden, ctx, ring_q32, ring_gold, name):
    g = golden[name]
    r = ring_for(name, ring_q32, ring_gold)
    out = _mul(ctx, r, np.asarray(g["mul_a"], dtype=np.uint64), np.asarray(g["mul_b"], dtype=np.uint64), 1)
    assert np.array_equal(out, r.words(np.asarray(g["mul_out"], dtype=np.uint64)))
    a = np.zeros((1, 64), dtype=np.uint64); a[0, 1] = 1
    b = np.zeros((1, 64), dtype=np.uint64); b[0, 63] = 1
    out = _mul(ctx, r, a, b, 1).astype(np.uint64)
    assert int(out[0, 0]) == r.q - 1 and not out[0, 1:].any()
    z = np.zeros((2, 64), dtype=np.uint64)
    assert not _mul(ctx, r, z, rand_residues(np.random.default_rng(1), r.q, (2, 64)), 1).any()


@pytest.mark.parametrize("name", ["q32", "goldilocks"])
@pytest.mark.parametrize("count", [1, 3, 65, 1000])
def test_quad_matches_oracle_and_thread_kernel(oracle, ctx, ring_q32, ring_gold, name, count):
    r = ring_for(name, ring_q32, ring_gold)
    rng = np.random.default_rng(2000 + count)
    a, b = rand_residues(rng, r.q, (count, 64)), rand_residues(rng, r.q, (count, 64))
    a[0], b[0] = r.q - 1, r.q - 1
    got = _mul(ctx, r, a, b, 1)
    assert np.array_equal(got, _mul(ctx, r, a, b, 0))
    assert np.array_equal(got.astype(np.uint64), oracle.negacyclic_mul(r.q, 64, a, b))


def test_quad_c2_shape(oracle, ctx, ring_q32):
    # 2^18 pairs: every quad runs several pairs of the grid-stride loop
    r = ring_q32
    count = 1 << 18
    rng = np.random.default_rng(7)
    a = r.words(rand_residues(rng, r.q, (count, 64)))
    b = r.words(rand_residues(rng, r.q, (count, 64)))
    got = _mul(ctx, r, a, b, 1)
    assert np.array_equal(got, _mul(ctx, r, a, b, 0))
    idx = rng.choice(count, 100, replace=False)
    assert np.array_equal(got[idx].astype(np.uint64),
                          oracle.negacyclic_mul(r.q, 64, a[idx].astype(np.uint64), b[idx].astype(np.uint64)))


def test_quad_tuning_values_checked(ctx):
    import lattica_amd as la
    assert ctx.get_tuning()["negacyclic_quad"] in (-1, 0, 1)
    for bad in (2, -2):
        with pytest.raises(la.GridHipError):
            ctx.set_tuning(negacyclic_quad=bad)
