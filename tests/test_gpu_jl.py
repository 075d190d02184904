"""JL rows for aggregation through the C ABI: flat extraction, conjugated
polynomials, the prover's JL evaluations and the aggregation's JL term against
the oracle, bit for bit.

Sizes cross the kernels' tiles: 256-polynomial extraction blocks, 4- and
64-polynomial conjugation tiles, 8-polynomial matrix-core rounds and the
per-block split of the evaluation GEMM, and the aggregation's 32-polynomial
tiles and 128-polynomial blocks. At the LaBRADOR scale (32 parts of 1024
polynomials) the evaluations are checked through their constant terms, which
must equal the projection p, and the aggregation through its pairing with the
witness, which must equal the evaluations' pairing with the scale.
"""
import numpy as np
import pytest

from conftest import GOLDILOCKS, Q32, rand_residues

pytestmark = pytest.mark.gpu

SEED = bytes(range(7, 39))


@pytest.mark.parametrize("rows,num_parts,n", [(256, 2, 5), (256, 3, 300), (256, 2, 512),
                                              (254, 1, 256), (253, 3, 1024), (256, 2, 272),
                                              (253, 1, 7), (100, 2, 9), (7, 3, 257), (1, 1, 1)])
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


def _ct_sum(a, b, q):
    """sum over polynomials of the constant term of a_g * b_g in Z_q[X]/(X^64 + 1), q < 2^32."""
    a = a.reshape(-1, 64).astype(np.uint64)
    b = b.reshape(-1, 64).astype(np.uint64)
    qq = np.uint64(q)
    first = (a[:, 0] * b[:, 0]) % qq
    rest = np.zeros(a.shape[0], dtype=np.uint64)
    for k in range(1, 64):
        rest += (a[:, k] * b[:, 64 - k]) % qq
    return (int(first.sum()) - int(rest.sum())) % q


@pytest.mark.parametrize("name", ["q32", "gold"])
@pytest.mark.parametrize("num_parts,n", [(1, 1), (2, 3), (1, 40), (3, 50), (2, 129)])
def test_jl_aggregate_matches_oracle(oracle, ring_q32, ring_gold, name, num_parts, n):
    """Sizes cross the 32-polynomial tiles and 128-polynomial blocks, with part
    boundaries inside a tile; phi starts nonzero (the caller's agg_phi)."""
    ring = ring_q32 if name == "q32" else ring_gold
    rng = np.random.default_rng(num_parts * 7 + n)
    scale = rand_residues(rng, ring.q, (256, 64))
    phi0 = rand_residues(rng, ring.q, (num_parts, n, 64))
    got = ring.jl_aggregate(SEED, scale, phi0)
    want = oracle.jl_aggregate(ring.q, 64, SEED, 256, scale, num_parts, n, phi=phi0)
    assert np.array_equal(got.astype(np.uint64), want)


@pytest.mark.parametrize("name", ["q32", "gold"])
def test_jl_aggregate_extreme_scale_words(oracle, ring_q32, ring_gold, name):
    ring = ring_q32 if name == "q32" else ring_gold
    q = ring.q
    edges = [0, 1, 2, q - 1, q - 2, q // 2, q // 2 + 1]
    edges += ([2139062143, 2139062144] if name == "q32"
              else [0x7F7F7F7F7F7F7F7F, 0x7F7F7F7F7F7F7F80])
    rng = np.random.default_rng(78)
    scale = np.array(edges, dtype=np.uint64)[rng.integers(0, len(edges), size=(256, 64))]
    phi0 = np.array(edges, dtype=np.uint64)[rng.integers(0, len(edges), size=(2, 20, 64))]
    got = ring.jl_aggregate(SEED, scale, phi0)
    want = oracle.jl_aggregate(q, 64, SEED, 256, scale, 2, 20, phi=phi0)
    assert np.array_equal(got.astype(np.uint64), want)


@pytest.mark.parametrize("name", ["q32", "gold"])
def test_jl_aggregate_is_linear_in_the_scale(ring_q32, ring_gold, name):
    """Two calls with scales c and c' on the same phi equal one call with c + c'."""
    ring = ring_q32 if name == "q32" else ring_gold
    q = ring.q
    rng = np.random.default_rng(79)
    c1 = rand_residues(rng, q, (256, 64))
    c2 = rand_residues(rng, q, (256, 64))
    csum = np.array([(int(a) + int(b)) % q for a, b in zip(c1.ravel(), c2.ravel())],
                    dtype=np.uint64).reshape(256, 64)
    phi0 = rand_residues(rng, q, (5, 77, 64))
    two = ring.jl_aggregate(SEED, c2, ring.jl_aggregate(SEED, c1, phi0))
    one = ring.jl_aggregate(SEED, csum, phi0)
    assert np.array_equal(two, one)


def test_jl_aggregate_pairs_with_the_evaluations_at_labrador_scale(ring_q32):
    """r 32 x 1024: sum_ij ct(phi_ij s_ij) = sum_m ct(scale_m jl_eval[m]), with
    phi from zero, both sides on the device."""
    ring = ring_q32
    q = ring.q
    rng = np.random.default_rng(124)
    num_parts, n = 32, 1024
    s = rand_residues(rng, q, (num_parts, n, 64)).astype(np.uint32)
    scale = rand_residues(rng, q, (256, 64)).astype(np.uint32)
    phi = ring.jl_aggregate(SEED, scale, np.zeros((num_parts, n, 64), dtype=np.uint32))
    ev = ring.jl_evals(SEED, s)
    assert _ct_sum(phi, s, q) == _ct_sum(scale, ev, q)


def test_jl_aggregate_device_memory_and_errors(ctx, oracle, ring_q32):
    import lattica_amd as la
    rng = np.random.default_rng(80)
    scale = rand_residues(rng, Q32, (256, 64)).astype(np.uint32)
    phi0 = rand_residues(rng, Q32, (2, 9, 64)).astype(np.uint32)
    sb = la.DeviceBuffer(ctx, scale.nbytes)
    pb = la.DeviceBuffer(ctx, phi0.nbytes)
    try:
        sb.upload(scale)
        pb.upload(phi0)
        sd = np.frombuffer(SEED, dtype=np.uint8).copy()
        L = la.lib()
        assert L.gh_jl_aggregate(ring_q32.h, sd.ctypes.data, 256, sb.ptr, 2, 9 * 64, pb.ptr,
                                 la.MEM_DEVICE) == 0
        got = pb.download(np.uint32, (2, 9, 64))
        want = oracle.jl_aggregate(Q32, 64, SEED, 256, scale, 2, 9, phi=phi0)
        assert np.array_equal(got.astype(np.uint64), want)
        ph = phi0.copy()
        assert L.gh_jl_aggregate(ring_q32.h, sd.ctypes.data, 128, scale.ctypes.data, 2, 9 * 64,
                                 ph.ctypes.data, la.MEM_HOST) == -20   # JLProfile: 256 rows
        assert L.gh_jl_aggregate(ring_q32.h, sd.ctypes.data, 256, scale.ctypes.data, 2, 9 * 64 - 1,
                                 ph.ctypes.data, la.MEM_HOST) == -20   # cols % d
        bad = scale.copy()
        bad[3, 5] = Q32
        assert L.gh_jl_aggregate(ring_q32.h, sd.ctypes.data, 256, bad.ctypes.data, 2, 9 * 64,
                                 ph.ctypes.data, la.MEM_HOST) == -21   # not canonical
        bad = phi0.copy()
        bad[1, 8, 63] = 0xFFFFFFFF
        assert L.gh_jl_aggregate(ring_q32.h, sd.ctypes.data, 256, scale.ctypes.data, 2, 9 * 64,
                                 bad.ctypes.data, la.MEM_HOST) == -21
        assert np.array_equal(ph, phi0)   # rejected calls leave phi alone
        assert L.gh_jl_aggregate(ring_q32.h, sd.ctypes.data, 256, scale.ctypes.data, 0, 9 * 64,
                                 None, la.MEM_HOST) == 0          # nothing to aggregate
    finally:
        sb.free()
        pb.free()


def _lifted(q, vals):
    """Signed integers as residues mod q."""
    v = np.asarray(vals, dtype=object)
    return np.array([int(x) % q for x in v.ravel()], dtype=np.uint64).reshape(v.shape)


@pytest.mark.parametrize("name", ["q32", "gold"])
@pytest.mark.parametrize("case", ["byte", "minus128", "short", "one_large", "wide", "mixed"])
def test_jl_evals_by_witness_digit_span(oracle, ring_q32, ring_gold, name, case):
    """Each evaluation block skips the digit planes past its witness words'
    span: small words (one digit), -128 (its negation needs two), 16-bit words
    (two), one full-width word among small ones (all, in one block only),
    40-bit words (Goldilocks: six of eight, both passes), and polynomials of
    every width at random over 32 blocks (the reduction then sums a different
    set of blocks for each digit plane)."""
    ring = ring_q32 if name == "q32" else ring_gold
    q = ring.q
    rng = np.random.default_rng({"byte": 1, "minus128": 2, "short": 3, "one_large": 4,
                                 "wide": 5, "mixed": 6}[case])
    shape = (4, 64, 64) if case == "mixed" else (3, 37, 64)
    if case == "mixed":
        bits = rng.choice([7, 15, 23, 31, 40, 62], size=shape[:2])
        if name == "q32":
            bits = np.minimum(bits, 31)
        lim = (1 << bits.astype(object))[..., None]
        s = _lifted(q, np.vectorize(lambda b: int(rng.integers(-b, b)), otypes=[object])(
            np.broadcast_to(lim, shape)))
    elif case == "byte":
        s = _lifted(q, rng.integers(-128, 128, size=shape))
    elif case == "minus128":
        s = _lifted(q, np.full(shape, -128))
    elif case == "short":
        s = _lifted(q, rng.integers(-(1 << 15), 1 << 15, size=shape))
    elif case == "one_large":
        s = _lifted(q, rng.integers(-3, 4, size=shape))
        s[2, 36, 63] = q // 2
    else:
        s = _lifted(q, rng.integers(-(1 << 40), 1 << 40, size=shape)) if name == "gold" else \
            rand_residues(rng, q, shape)
    got = ring.jl_evals(SEED, s)
    want = oracle.jl_evals(q, 64, SEED, 256, s)
    assert np.array_equal(got.astype(np.uint64), want)


def test_jl_evals_small_witness_at_labrador_scale(ring_q32):
    """r 32 x 1024 with coefficients in [-2^7, 2^7) (one digit): the constant
    terms are still the projection."""
    rng = np.random.default_rng(125)
    s = _lifted(Q32, rng.integers(-128, 128, size=(32, 1024 * 64))).astype(np.uint32)
    ev = ring_q32.jl_evals(SEED, s)
    pm, _, _ = ring_q32.jl_project(SEED, 256, s)
    assert np.array_equal(ev[:, 0], pm)
