"""JL rows for aggregation through the C ABI: flat extraction, conjugated
polynomials and the prover's JL evaluations against the oracle, bit for bit.

Sizes cross the kernels' tiles: 256-polynomial extraction blocks, 4- and
64-polynomial conjugation tiles, 8-polynomial matrix-core rounds and the
per-block split of the evaluation GEMM. At the LaBRADOR scale (32 parts of
1024 polynomials) the evaluations are checked through their constant terms,
which must equal the projection p.
"""
import numpy as np
import pytest

from conftest import GOLDILOCKS, Q32, rand_residues

pytestmark = pytest.mark.gpu

SEED = bytes(range(7, 39))


@pytest.mark.parametrize("rows,num_parts,n", [(256, 2, 5), (256, 3, 300), (256, 2, 512),
                                              (254, 1, 256), (253, 1, 7), (100, 2, 9),
                                              (7, 3, 257), (1, 1, 1)])
def test_extract_flat_matches_oracle(oracle, ring_q32, rows, num_parts, n):
    got = ring_q32.jl_extract_flat(SEED, rows, num_parts, n * 64)
    want = oracle.jl_extract_flat(SEED, rows, num_parts, n * 64, 64)
    assert np.array_equal(got, want)


@pytest.mark.parametrize("rows,n", [(256, 70), (256, 256), (255, 9), (200, 33)])
def test_project_flat_output_matches_oracle(oracle, ring_q32, rows, n):
    """256-row projections run on the matrix cores, others on the row-per-thread kernel."""
    rng = np.random.default_rng(4 + n)
    parts = rand_residues(rng, Q32, (3, n * 64))
    pm, pe, fl = ring_q32.jl_project(SEED, rows, parts, want_flat=True)
    opm, ope, ofl = oracle.jl_project(Q32, SEED, rows, parts, 64, want_flat=True)
    assert np.array_equal(pm.astype(np.uint64), opm) and np.array_equal(pe, ope)
    assert np.array_equal(fl, ofl)


@pytest.mark.parametrize("n", [1, 9, 300])
def test_project_goldilocks_matches_oracle(oracle, ring_gold, n):
    rng = np.random.default_rng(40 + n)
    parts = rand_residues(rng, GOLDILOCKS, (2, n * 64))
    pm, _, _ = ring_gold.jl_project(SEED, 256, parts, want_exact=False)
    opm, _, _ = oracle.jl_project(GOLDILOCKS, SEED, 256, parts, 64)
    assert np.array_equal(pm.astype(np.uint64), opm)


def test_project_extreme_words(oracle, ring_q32):
    """Centred lifts at the edges: 0, 1, q - 1, (q - 1) / 2, (q + 1) / 2 in 5 digits."""
    q = Q32
    edges = np.array([0, 1, q - 1, (q - 1) // 2, (q + 1) // 2, 2139062143, 2139062144],
                     dtype=np.uint64)
    rng = np.random.default_rng(5)
    parts = edges[rng.integers(0, len(edges), size=(2, 50 * 64))]
    pm, pe, _ = ring_q32.jl_project(SEED, 256, parts)
    opm, ope, _ = oracle.jl_project(q, SEED, 256, parts, 64)
    assert np.array_equal(pm.astype(np.uint64), opm) and np.array_equal(pe, ope)


@pytest.mark.parametrize("name", ["q32", "gold"])
@pytest.mark.parametrize("rows,num_parts,n", [(256, 2, 5), (256, 1, 70), (9, 3, 66), (256, 2, 129)])
def test_conjugate_matches_oracle(oracle, ring_q32, ring_gold, name, rows, num_parts, n):
    ring = ring_q32 if name == "q32" else ring_gold
    flat = oracle.jl_extract_flat(SEED, rows, num_parts, n * 64, 64)
    want = oracle.jl_conjugate_flat(ring.q, flat, rows, num_parts, n, 64)
    from_flat = ring.jl_conjugate(rows, num_parts, n * 64, flat=flat)
    from_seed = ring.jl_conjugate(rows, num_parts, n * 64, seed=SEED)
    assert np.array_equal(from_flat.astype(np.uint64), want)
    assert np.array_equal(from_seed.astype(np.uint64), want)


@pytest.mark.parametrize("name", ["q32", "gold"])
@pytest.mark.parametrize("num_parts,n", [(1, 1), (2, 3), (3, 17), (4, 300)])
def test_jl_evals_match_oracle(oracle, ring_q32, ring_gold, name, num_parts, n):
    ring = ring_q32 if name == "q32" else ring_gold
    rng = np.random.default_rng(num_parts * 1000 + n)
    s = rand_residues(rng, ring.q, (num_parts, n, 64))
    got = ring.jl_evals(SEED, s)
    want = oracle.jl_evals(ring.q, 64, SEED, 256, s)
    assert np.array_equal(got.astype(np.uint64), want)


@pytest.mark.parametrize("name", ["q32", "gold"])
def test_jl_evals_extreme_words(oracle, ring_q32, ring_gold, name):
    """Witness words at the edges of the digit windows: 0, 1, q - 1, q / 2 and
    the lifts' switch points."""
    ring = ring_q32 if name == "q32" else ring_gold
    q = ring.q
    edges = [0, 1, 2, q - 1, q - 2, q // 2, q // 2 + 1]
    edges += ([2139062143, 2139062144] if name == "q32"
              else [0x7F7F7F7F7F7F7F7F, 0x7F7F7F7F7F7F7F80])
    rng = np.random.default_rng(77)
    s = np.array(edges, dtype=np.uint64)[rng.integers(0, len(edges), size=(2, 40, 64))]
    got = ring.jl_evals(SEED, s)
    want = oracle.jl_evals(q, 64, SEED, 256, s)
    assert np.array_equal(got.astype(np.uint64), want)


@pytest.mark.parametrize("name", ["q32", "gold"])
def test_jl_evals_constant_terms_at_labrador_scale(ring_q32, ring_gold, name):
    ring = ring_q32 if name == "q32" else ring_gold
    rng = np.random.default_rng(123)
    num_parts, n = 32, 1024
    s = rand_residues(rng, ring.q, (num_parts, n * 64)).astype(ring.dtype)
    ev = ring.jl_evals(SEED, s)
    pm, _, _ = ring.jl_project(SEED, 256, s, want_exact=(name == "q32"))
    assert np.array_equal(ev[:, 0], pm)


def test_jl_evals_device_memory_and_errors(ctx, oracle, ring_q32):
    import ctypes

    import lattica_amd as la
    rng = np.random.default_rng(8)
    s = rand_residues(rng, Q32, (2, 9 * 64)).astype(np.uint32)
    inb = la.DeviceBuffer(ctx, s.nbytes)
    outb = la.DeviceBuffer(ctx, 256 * 64 * 4)
    try:
        inb.upload(s)
        sd = np.frombuffer(SEED, dtype=np.uint8).copy()
        L = la.lib()
        assert L.gh_jl_evals(ring_q32.h, sd.ctypes.data, 256, inb.ptr, 2, 9 * 64, outb.ptr,
                             la.MEM_DEVICE) == 0
        got = outb.download(np.uint32, (256, 64))
        want = oracle.jl_evals(Q32, 64, SEED, 256, s.astype(np.uint64).reshape(2, 9, 64))
        assert np.array_equal(got.astype(np.uint64), want)
        out = np.zeros((256, 64), dtype=np.uint32)
        assert L.gh_jl_evals(ring_q32.h, sd.ctypes.data, 128, s.ctypes.data, 2, 9 * 64,
                             out.ctypes.data, la.MEM_HOST) == -20   # JLProfile: 256 rows
        assert L.gh_jl_evals(ring_q32.h, sd.ctypes.data, 256, s.ctypes.data, 2, 9 * 64 - 1,
                             out.ctypes.data, la.MEM_HOST) == -20   # cols % d
        bad = s.copy()
        bad[0, 0] = Q32
        assert L.gh_jl_evals(ring_q32.h, sd.ctypes.data, 256, bad.ctypes.data, 2, 9 * 64,
                             out.ctypes.data, la.MEM_HOST) == -21   # not canonical
    finally:
        inb.free()
        outb.free()
