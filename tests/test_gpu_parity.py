"""HIP path (through the C ABI) versus the oracle and the golden vectors.

Bit-exact equality everywhere: every output is a canonical residue of an exact
ring computation. Sizes cross the kernels' tile and block boundaries; the
full-size north-star commitment (kappa = 16, n = 2^15) is compared exactly.
"""
import ctypes

import numpy as np
import pytest

from conftest import GOLDILOCKS, Q32, rand_residues, small_signed

pytestmark = pytest.mark.gpu


def arr(x):
    return np.asarray(x, dtype=np.uint64)


def ring_for(name, ring_q32, ring_gold):
    return ring_q32 if name == "q32" else ring_gold


# ---- ring plan -------------------------------------------------------------

@pytest.mark.parametrize("name", ["q32", "goldilocks"])
def test_ring_info_matches_reference_roots(golden, ring_q32, ring_gold, name):
    g = golden[name]
    r = ring_for(name, ring_q32, ring_gold)
    assert (r.q, r.d) == (g["q"], 64)
    assert r.info.generator == g["generator"]
    assert r.info.psi == g["psi"]
    assert r.info.omega == g["omega"]
    assert r.info.d_inv == g["d_inv"]


def test_ring_errors_map_to_ntt_error(ctx):
    import lattica_amd as la
    with pytest.raises(la.GridHipError) as e:
        la.Ring(ctx, 17, 12)
    assert e.value.name == "LENGTH_NOT_POW2"
    with pytest.raises(la.GridHipError) as e:
        la.Ring(ctx, 17, 32)  # F17 has no 32nd root (poly/ntt.rs:675-682)
    assert e.value.name == "UNSUPPORTED_SIZE"
    with pytest.raises(la.GridHipError) as e:
        la.Ring(ctx, 17, 16)  # no 32nd root for the twist (poly/ntt.rs:684-693)
    assert e.value.name == "UNSUPPORTED_TWIST"
    with pytest.raises(la.GridHipError) as e:
        la.Ring(ctx, 184683593729, 64)  # valid NTT prime > 2^32, not Goldilocks
    assert e.value.name == "UNSUPPORTED_MODULUS"
    with pytest.raises(la.GridHipError) as e:
        la.Ring(ctx, 17, 8)  # a valid ring the GPU kernels do not cover (d != 64)
    assert e.value.name == "UNSUPPORTED_SIZE"


# ---- transforms ------------------------------------------------------------

@pytest.mark.parametrize("name", ["q32", "goldilocks"])
def test_ntt_golden(golden, ring_q32, ring_gold, name):
    g = golden[name]
    r = ring_for(name, ring_q32, ring_gold)
    x = r.words(arr(g["ntt_in"]))
    assert np.array_equal(r.twisted_forward(x), r.words(arr(g["twisted_out"])))
    assert np.array_equal(r.ntt_forward(x), r.words(arr(g["cyclic_out"])))
    assert np.array_equal(r.twisted_inverse(r.words(arr(g["twisted_out"]))), x)
    assert np.array_equal(r.ntt_inverse(r.words(arr(g["cyclic_out"]))), x)


@pytest.mark.parametrize("name", ["q32", "goldilocks"])
@pytest.mark.parametrize("count", [1, 63, 64, 65, 257, 3001])
def test_ntt_batches_match_oracle(oracle, ring_q32, ring_gold, name, count):
    r = ring_for(name, ring_q32, ring_gold)
    rng = np.random.default_rng(count)
    x = rand_residues(rng, r.q, (count, 64))
    assert np.array_equal(r.twisted_forward(x).astype(np.uint64), oracle.twisted_forward(r.q, 64, x))
    assert np.array_equal(r.twisted_inverse(x).astype(np.uint64), oracle.twisted_inverse(r.q, 64, x))
    assert np.array_equal(r.ntt_forward(x).astype(np.uint64), oracle.ntt_forward(r.q, 64, x))
    assert np.array_equal(r.ntt_inverse(x).astype(np.uint64), oracle.ntt_inverse(r.q, 64, x))


@pytest.mark.parametrize("name", ["q32", "goldilocks"])
def test_twisted_forward_extreme_words(oracle, ring_q32, ring_gold, name):
    """Words at the edges of the reductions (0, 1, q - 1, q - 2, 2^31, 2^32 - 1,
    2^32, 2^63 where below q) in every slot pattern: the Goldilocks transform
    multiplies by its roots as shifts, whose carries these words exercise."""
    r = ring_for(name, ring_q32, ring_gold)
    q = r.q
    edges = [v for v in (0, 1, q - 1, q - 2, 1 << 31, (1 << 32) - 1, 1 << 32, 1 << 63,
                         q - (1 << 32), (q + 1) // 2) if 0 <= v < q]
    rng = np.random.default_rng(5)
    x = np.array(edges, dtype=np.uint64)[rng.integers(0, len(edges), size=(97, 64))]
    x[0] = q - 1
    x[1] = 0
    x[1, 7] = q - 1
    got = r.twisted_forward(r.words(x)).astype(np.uint64)
    assert np.array_equal(got, oracle.twisted_forward(q, 64, x))
    assert np.array_equal(r.twisted_inverse(r.words(got)).astype(np.uint64), x)


@pytest.mark.parametrize("name", ["q32", "goldilocks"])
def test_large_batch_round_trip_and_spot_check(oracle, ctx, ring_q32, ring_gold, name):
    # enough polynomials for the wide-tile kernels; round trip over all of them,
    # exact comparison on a sample
    r = ring_for(name, ring_q32, ring_gold)
    count = 256 * ctx.num_cus * 2 + 17
    rng = np.random.default_rng(99)
    x = r.words(rand_residues(rng, r.q, (count, 64)))
    y = r.twisted_forward(x)
    assert np.array_equal(r.twisted_inverse(y), x)
    idx = rng.choice(count, 200, replace=False)
    assert np.array_equal(y[idx].astype(np.uint64), oracle.twisted_forward(r.q, 64, x[idx]))


@pytest.mark.parametrize("name", ["q32", "goldilocks"])
def test_negacyclic_golden_and_edges(golden, ring_q32, ring_gold, name):
    g = golden[name]
    r = ring_for(name, ring_q32, ring_gold)
    out = r.negacyclic_mul(arr(g["mul_a"]), arr(g["mul_b"]))
    assert np.array_equal(out, r.words(arr(g["mul_out"])))
    # X * X^63 = -1
    a = np.zeros((1, 64), dtype=np.uint64); a[0, 1] = 1
    b = np.zeros((1, 64), dtype=np.uint64); b[0, 63] = 1
    out = r.negacyclic_mul(a, b).astype(np.uint64)
    assert int(out[0, 0]) == r.q - 1 and not out[0, 1:].any()
    z = np.zeros((2, 64), dtype=np.uint64)
    assert not r.negacyclic_mul(z, rand_residues(np.random.default_rng(1), r.q, (2, 64))).any()


@pytest.mark.parametrize("name", ["q32", "goldilocks"])
@pytest.mark.parametrize("count", [1, 65, 1000])
def test_negacyclic_matches_oracle(oracle, ring_q32, ring_gold, name, count):
    r = ring_for(name, ring_q32, ring_gold)
    rng = np.random.default_rng(1000 + count)
    a, b = rand_residues(rng, r.q, (count, 64)), rand_residues(rng, r.q, (count, 64))
    # the largest residues in the first pair (every modular correction taken)
    a[0], b[0] = r.q - 1, r.q - 1
    assert np.array_equal(r.negacyclic_mul(a, b).astype(np.uint64), oracle.negacyclic_mul(r.q, 64, a, b))


@pytest.mark.parametrize("name", ["q32", "goldilocks"])
def test_negacyclic_c2_shape_sample(oracle, ctx, ring_q32, ring_gold, name):
    # the batched product at the C2 scale (2^18 pairs: several tiles per
    # block); exact on a sample, and commutativity a*b == b*a on all pairs
    r = ring_for(name, ring_q32, ring_gold)
    count = 1 << 18
    rng = np.random.default_rng(7)
    a = r.words(rand_residues(rng, r.q, (count, 64)))
    b = r.words(rand_residues(rng, r.q, (count, 64)))
    ab = r.negacyclic_mul(a, b)
    assert np.array_equal(ab, r.negacyclic_mul(b, a))
    idx = rng.choice(count, 300, replace=False)
    assert np.array_equal(ab[idx].astype(np.uint64),
                          oracle.negacyclic_mul(r.q, 64, a[idx].astype(np.uint64), b[idx].astype(np.uint64)))


@pytest.mark.parametrize("name", ["q32", "goldilocks"])
def test_pointwise_mul(oracle, ring_q32, ring_gold, name):
    r = ring_for(name, ring_q32, ring_gold)
    rng = np.random.default_rng(4)
    a, b = rand_residues(rng, r.q, (5, 64)), rand_residues(rng, r.q, (5, 64))
    got = r.pointwise_mul(a, b).astype(np.uint64)
    exp = np.array([[int(x) * int(y) % r.q for x, y in zip(ra, rb)] for ra, rb in zip(a, b)],
                   dtype=np.uint64)
    assert np.array_equal(got, exp)


def test_non_canonical_input_rejected(ring_q32):
    import lattica_amd as la
    x = np.zeros((1, 64), dtype=np.uint32)
    x[0, 5] = Q32  # == q
    with pytest.raises(la.GridHipError) as e:
        ring_q32.twisted_forward(x)
    assert e.value.name == "NOT_CANONICAL"


# ---- commitment ------------------------------------------------------------

@pytest.mark.parametrize("name", ["q32", "goldilocks"])
def test_commit_golden(golden, ring_q32, ring_gold, name):
    import lattica_amd as la
    g = golden[name]
    r = ring_for(name, ring_q32, ring_gold)
    c = g["commit"]
    key = la.Key(r, arr(c["a_msg"]), arr(c["a_open"]))
    t = key.commit(arr(c["m"]), arr(c["r"]))
    assert np.array_equal(t, r.words(arr(c["t"])))
    key.close()


@pytest.mark.parametrize("name", ["q32", "goldilocks"])
@pytest.mark.parametrize("rows,msg,opn", [(1, 1, 0), (3, 17, 2), (5, 100, 0), (16, 333, 1), (8, 256, 0)])
def test_commit_matches_oracle(oracle, ring_q32, ring_gold, name, rows, msg, opn):
    import lattica_amd as la
    r = ring_for(name, ring_q32, ring_gold)
    rng = np.random.default_rng(rows * 1000 + msg)
    am = rand_residues(rng, r.q, (rows, msg, 64))
    ao = rand_residues(rng, r.q, (rows, max(opn, 1), 64))
    m = small_signed(rng, r.q, (msg, 64), -1, 2)  # ternary witness (C1 / C4)
    rr = small_signed(rng, r.q, (max(opn, 1), 64), -1, 2)
    if opn == 0:
        ao0 = np.zeros((rows, 1, 64), dtype=np.uint64)
        exp = oracle.commit(r.q, 64, am, ao0, m, np.zeros((1, 64), dtype=np.uint64))
        key = la.Key(r, am)
        got = key.commit(m)
    else:
        exp = oracle.commit(r.q, 64, am, ao, m, rr)
        key = la.Key(r, am, ao)
        got = key.commit(m, rr)
    assert np.array_equal(got.astype(np.uint64), exp)
    key.close()


def test_commit_ntt_domain_key_and_witness(oracle, ring_q32):
    import lattica_amd as la
    r = ring_q32
    rng = np.random.default_rng(77)
    am_ntt = rand_residues(rng, r.q, (4, 50, 64))
    s = small_signed(rng, r.q, (50, 64), -128, 128)
    z = np.zeros((4, 1, 64), dtype=np.uint64)
    exp = oracle.commit(r.q, 64, am_ntt, z, s, np.zeros((1, 64), dtype=np.uint64), key_domain=1)
    key = la.Key(r, am_ntt, domain=la.DOMAIN_NTT)
    assert np.array_equal(key.commit(s).astype(np.uint64), exp)
    # witness already in the NTT domain (commit_with_opening_ntt path)
    s_ntt = r.twisted_forward(s)
    assert np.array_equal(key.commit(s_ntt, domain=la.DOMAIN_NTT).astype(np.uint64), exp)
    key.close()


def test_commit_dimension_mismatch(ring_q32):
    import lattica_amd as la
    key = la.Key(ring_q32, np.zeros((2, 3, 64), dtype=np.uint32))
    with pytest.raises(la.GridHipError) as e:
        key.commit(np.zeros((4, 64), dtype=np.uint32))
    assert e.value.name == "DIMENSION_MISMATCH"
    key.close()


def test_north_star_commit_full_size(oracle, ctx, ring_q32):
    """kappa = 16, n = 2^15, d = 64, q32, s in [-2^7, 2^7): exact against the oracle,
    through resident device buffers (the benchmarked path)."""
    import lattica_amd as la
    r = ring_q32
    rows, n = 16, 1 << 15
    rng = np.random.default_rng(2024)
    a_ntt = rand_residues(rng, r.q, (rows, n, 64))
    s = small_signed(rng, r.q, (n, 64), -128, 128)
    z = np.zeros((rows, 1, 64), dtype=np.uint64)
    exp = oracle.commit(r.q, 64, a_ntt, z, s, np.zeros((1, 64), dtype=np.uint64), key_domain=1)
    key = la.Key(r, a_ntt.astype(np.uint32), domain=la.DOMAIN_NTT)
    s_dev = ctx.to_device(s.astype(np.uint32))
    t_dev = ctx.alloc(rows * 64 * 4)
    key.commit_device(s_dev, n, t_dev)
    ctx.synchronize()
    got = t_dev.download(np.uint32, (rows, 64))
    assert np.array_equal(got.astype(np.uint64), exp)
    # same result through the host-memory entry point
    assert np.array_equal(key.commit(s.astype(np.uint32)).astype(np.uint64), exp)
    s_dev.free(); t_dev.free(); key.close()


@pytest.fixture(scope="module")
def full_key(oracle, ring_q32):
    """The kappa 16 x 2^15 key of the headline shape and one commitment of it."""
    import lattica_amd as la
    rng = np.random.default_rng(62)
    a = rand_residues(rng, Q32, (16, 1 << 15, 64))
    s = small_signed(rng, Q32, (1 << 15, 64), -128, 128)
    z = np.zeros((16, 1, 64), dtype=np.uint64)
    zr = np.zeros((1, 64), dtype=np.uint64)
    exp = oracle.commit(Q32, 64, a, z, s, zr, key_domain=1)
    exp_tail = oracle.commit(Q32, 64, a[:, :32763], z, s[:32763], zr, key_domain=1)
    key = la.Key(ring_q32, a, domain=la.DOMAIN_NTT)
    yield key, s, exp, exp_tail
    key.close()


@pytest.mark.parametrize("fused,waves,bpc", [(1, 4, 1), (1, 4, 2), (1, 4, 4), (1, 8, 1),
                                             (1, 8, 2), (2, 4, 1), (2, 4, 2), (2, 4, 3)])
def test_unrolled_sweeps_at_the_headline_shape(ctx, full_key, fused, waves, bpc):
    """The fused kernels' unrolled sweeps (column split: 8, 16 and 32 columns per
    wave; row split: 64 or 128 columns per block, unrolled for 64) at every
    depth, on the full 2^15 columns and on 32763 (a last block 5 columns short)."""
    key, s, exp, exp_tail = full_key
    saved = ctx.get_tuning()
    try:
        for depth in (1, 2, 3, 4):
            ctx.set_tuning(fused=fused, mac_waves=waves, mac_blocks_per_cu=bpc, fused_depth=depth)
            assert np.array_equal(key.commit(s).astype(np.uint64), exp), depth
            # a one-item batch takes the plain path over the key's first 32763 columns
            got = key.commit_batch([s[:32763]])[0]
            assert np.array_equal(got.astype(np.uint64), exp_tail), depth
    finally:
        ctx.set_tuning(**saved)


def test_every_tuning_setting_gives_the_same_commitment(oracle, ctx, ring_q32):
    import lattica_amd as la
    r = ring_q32
    rng = np.random.default_rng(61)
    rows, n = 16, 3001
    a = rand_residues(rng, r.q, (rows, n, 64))
    s = small_signed(rng, r.q, (n, 64), -128, 128)
    z = np.zeros((rows, 1, 64), dtype=np.uint64)
    exp = oracle.commit(r.q, 64, a, z, s, np.zeros((1, 64), dtype=np.uint64), key_domain=1)
    key = la.Key(r, a, domain=la.DOMAIN_NTT)
    saved = ctx.get_tuning()
    try:
        for fused, variant, waves, bpc, quad, depth in [
                (0, 0, 16, 1, 0, 1), (0, 1, 4, 1, 0, 1), (0, 1, 8, 2, 1 << 17, 1),
                (0, 1, 16, 4, 1 << 17, 1), (0, 1, 16, 8, 0, 1), (1, 1, 4, 1, 0, 1),
                (1, 1, 8, 2, 0, 1), (1, 1, 16, 1, 0, 1), (1, 1, 16, 8, 0, 1), (1, 1, 4, 3, 0, 1),
                (1, 1, 8, 3, 0, 1), (1, 1, 4, 1, 0, 2), (1, 1, 4, 2, 0, 2), (1, 1, 4, 3, 0, 2),
                (1, 1, 8, 1, 0, 2), (1, 1, 16, 2, 0, 2), (1, 1, 4, 2, 0, 3), (1, 1, 4, 2, 0, 4),
                (1, 1, 16, 1, 0, 4), (2, 1, 4, 1, 0, 1), (2, 1, 4, 2, 0, 2), (2, 1, 4, 3, 0, 3),
                (2, 1, 4, 2, 0, 4), (2, 1, 16, 8, 0, 2), (2, 1, 8, 4, 0, 1)]:
            ctx.set_tuning(fused=fused, mac_variant=variant, mac_waves=waves,
                           mac_blocks_per_cu=bpc, quad_max=quad, fused_depth=depth)
            got = key.commit(s).astype(np.uint64)
            assert np.array_equal(got, exp), (fused, variant, waves, bpc, quad, depth)
        with pytest.raises(la.GridHipError):
            ctx.set_tuning(mac_waves=5)
    finally:
        ctx.set_tuning(**saved)
        key.close()


def test_back_to_back_device_commits_pipelined(oracle, ctx, ring_q32):
    """Many device-memory commits enqueued without waiting (the finalize of one
    overlaps the MAC of the next), two keys alternating, outputs reused, and a
    non-commit call in between: every result equals its oracle value."""
    import lattica_amd as la
    r = ring_q32
    rng = np.random.default_rng(88)
    rows, n = 16, 700
    a = [rand_residues(rng, r.q, (rows, n, 64)) for _ in range(2)]
    keys = [la.Key(r, x, domain=la.DOMAIN_NTT) for x in a]
    ws = [small_signed(rng, r.q, (n, 64), -128, 128) for _ in range(5)]
    wd = [ctx.to_device(w.astype(np.uint32)) for w in ws]
    outs = [ctx.alloc(rows * 64 * 4) for _ in range(3)]
    z = np.zeros((rows, 1, 64), dtype=np.uint64)
    exp = {(k, i): oracle.commit(r.q, 64, a[k], z, ws[i], np.zeros((1, 64), dtype=np.uint64),
                                 key_domain=1) for k in range(2) for i in range(5)}
    seq = [(i % 2, i % 5) for i in range(12)]
    got = {}
    for j, (k, i) in enumerate(seq):
        keys[k].commit_device(wd[i], n, outs[j % 3])
        if j % 3 == 2:
            # a non-commit call on another buffer, then read the three outputs
            la.lib().gh_twisted_forward(r.h, wd[4].ptr, n, la.MEM_DEVICE)
            ws[4] = r.twisted_forward(ws[4]).astype(np.uint64)
            for jj in range(j - 2, j + 1):
                got[jj] = outs[jj % 3].download(np.uint32, (rows, 64)).astype(np.uint64)
    for j, (k, i) in enumerate(seq):
        # witness 4 was transformed in place after each group of three commits
        if i == 4:
            continue
        assert np.array_equal(got[j], exp[(k, i)]), j
    for b in wd + outs:
        b.free()
    for k in keys:
        k.close()


def test_commit_batch_uneven_lengths(ring_q32, ring_gold):
    import lattica_amd as la
    for r in (ring_q32, ring_gold):
        rng = np.random.default_rng(31)
        cols = 300
        a = rand_residues(rng, r.q, (16, cols, 64))
        key = la.Key(r, a)
        lens = [1, 7, 64, 65, 299, 300, 128]
        ws = [small_signed(rng, r.q, (n, 64), -1, 2) for n in lens]
        got = key.commit_batch(ws)
        for n, w, t in zip(lens, ws, got):
            k1 = la.Key(r, a[:, :n])
            assert np.array_equal(t, k1.commit(w))
            k1.close()
        key.close()


@pytest.mark.parametrize("name", ["q32", "goldilocks"])
def test_commit_batch_device_many_items_all_tunings(oracle, ctx, ring_q32, ring_gold, name):
    """Log-uniform lengths (the C4 shape, scaled down), device buffers, every
    batch tuning: each item equals the oracle commitment with the key prefix."""
    import lattica_amd as la
    r = ring_for(name, ring_q32, ring_gold)
    rng = np.random.default_rng(404)
    rows, cols, count = 16, 600, 37
    a = rand_residues(rng, r.q, (rows, cols, 64))
    key = la.Key(r, a, domain=la.DOMAIN_NTT)
    lens = np.floor(2.0 ** rng.uniform(0, np.log2(cols), size=count)).astype(int)
    lens[0], lens[1] = cols, 1
    ws = [small_signed(rng, r.q, (n, 64), -1, 2) for n in lens]
    wd = [ctx.to_device(r.words(w)) for w in ws]
    outs = [ctx.alloc(rows * 64 * r.info.word_bytes) for _ in lens]
    z = np.zeros((rows, 1, 64), dtype=np.uint64)
    exp = [oracle.commit(r.q, 64, a[:, :n], z, w, np.zeros((1, 64), dtype=np.uint64), key_domain=1)
           for n, w in zip(lens, ws)]
    saved = ctx.get_tuning()
    try:
        for group, chunk in [(8, 256), (4, 16), (16, 100), (8, 4096)]:
            # each field reads its own width; the other field's is set elsewhere
            ctx.set_tuning(batch_group=group, batch_chunk_cols=chunk if name == "q32" else 64,
                           batch_chunk_cols_gold=chunk if name != "q32" else 64)
            key.commit_batch_device(wd, lens, outs)
            ctx.synchronize()
            for i in range(count):
                got = outs[i].download(r.dtype, (rows, 64)).astype(np.uint64)
                assert np.array_equal(got, exp[i]), (group, chunk, i, lens[i])
    finally:
        ctx.set_tuning(**saved)
        for b in wd + outs:
            b.free()
        key.close()


@pytest.mark.parametrize("name", ["q32", "goldilocks"])
@pytest.mark.parametrize("rows,msg,opn", [(4, 3, 2), (16, 100, 37), (8, 1, 1)])
def test_bdlop_matches_oracle(oracle, ring_q32, ring_gold, name, rows, msg, opn):
    """u = A_mask r, v = A_msg m + A_open r (commit/src/bdlop/mod.rs:34-54), both
    domains, against the oracle's linear commitment applied to each part."""
    import lattica_amd as la
    r = ring_for(name, ring_q32, ring_gold)
    rng = np.random.default_rng(rows * 100 + msg + opn)
    a_mask = rand_residues(rng, r.q, (rows, opn, 64))
    a_msg = rand_residues(rng, r.q, (rows, msg, 64))
    a_open = rand_residues(rng, r.q, (rows, opn, 64))
    m = small_signed(rng, r.q, (msg, 64), -128, 128)
    rr = small_signed(rng, r.q, (opn, 64), -1, 2)
    mask = la.Key(r, a_mask)
    key = la.Key(r, a_msg, a_open)
    u, v = key.bdlop(mask, m, rr)
    z = np.zeros((rows, 1, 64), dtype=np.uint64)
    z1 = np.zeros((1, 64), dtype=np.uint64)
    assert np.array_equal(u.astype(np.uint64), oracle.commit(r.q, 64, a_mask, z, rr, z1))
    assert np.array_equal(v.astype(np.uint64), oracle.commit(r.q, 64, a_msg, a_open, m, rr))
    # NTT-domain inputs give the same commitment (commit_with_opening_ntt)
    u2, v2 = key.bdlop(mask, r.twisted_forward(m), r.twisted_forward(rr), domain=la.DOMAIN_NTT)
    assert np.array_equal(u2, u) and np.array_equal(v2, v)
    mask.close()
    key.close()


def test_bdlop_additive_homomorphism_and_errors(ring_q32):
    """commit(m1, r1) + commit(m2, r2) = commit(m1 + m2, r1 + r2)
    (bdlop/tests.rs:498-528), and the dimension rules."""
    import lattica_amd as la
    r = ring_q32
    q = r.q
    rng = np.random.default_rng(5)
    mask = la.Key(r, rand_residues(rng, q, (8, 6, 64)))
    key = la.Key(r, rand_residues(rng, q, (8, 10, 64)), rand_residues(rng, q, (8, 6, 64)))
    m1, m2 = rand_residues(rng, q, (10, 64)), rand_residues(rng, q, (10, 64))
    r1, r2 = rand_residues(rng, q, (6, 64)), rand_residues(rng, q, (6, 64))
    u1, v1 = key.bdlop(mask, m1, r1)
    u2, v2 = key.bdlop(mask, m2, r2)
    u3, v3 = key.bdlop(mask, (m1 + m2) % np.uint64(q), (r1 + r2) % np.uint64(q))
    add = lambda x, y: ((x.astype(np.uint64) + y.astype(np.uint64)) % np.uint64(q))  # noqa: E731
    assert np.array_equal(add(u1, u2), u3.astype(np.uint64))
    assert np.array_equal(add(v1, v2), v3.astype(np.uint64))
    with pytest.raises(la.GridHipError) as e:
        key.bdlop(mask, m1[:9], r1)
    assert e.value.name == "DIMENSION_MISMATCH"
    with pytest.raises(la.GridHipError) as e:
        key.bdlop(mask, m1, r1[:5])
    assert e.value.name == "DIMENSION_MISMATCH"
    mask.close()
    key.close()


def test_inner_commit_matches_oracle(oracle, ring_q32):
    import lattica_amd as la
    r = ring_q32
    rng = np.random.default_rng(12)
    rows, parts, n = 16, 4, 40
    a = rand_residues(rng, r.q, (rows, n, 64))
    s = small_signed(rng, r.q, (parts, n, 64), -128, 128)
    key = la.Key(r, a)
    got = key.inner_commit(s)
    z = np.zeros((rows, 1, 64), dtype=np.uint64)
    for i in range(parts):
        exp = oracle.commit(r.q, 64, a, z, s[i], np.zeros((1, 64), dtype=np.uint64))
        assert np.array_equal(got[i].astype(np.uint64), exp)
    key.close()


# ---- Gram ------------------------------------------------------------------

@pytest.mark.parametrize("name", ["q32", "goldilocks"])
def test_gram_golden(golden, ring_q32, ring_gold, name):
    g = golden[name]
    r = ring_for(name, ring_q32, ring_gold)
    assert np.array_equal(r.gram(arr(g["gram"]["s"])), r.words(arr(g["gram"]["g"])))


@pytest.mark.parametrize("name", ["q32", "goldilocks"])
@pytest.mark.parametrize("rr,n", [(1, 1), (5, 33), (9, 64), (32, 16)])
def test_gram_matches_oracle(oracle, ring_q32, ring_gold, name, rr, n):
    r = ring_for(name, ring_q32, ring_gold)
    rng = np.random.default_rng(rr * 100 + n)
    s = small_signed(rng, r.q, (rr, n, 64), -128, 128)
    assert np.array_equal(r.gram(s).astype(np.uint64), oracle.gram(r.q, 64, s))


def test_gram_zero_and_empty(ring_q32):
    assert not ring_q32.gram(np.zeros((3, 5, 64), dtype=np.uint32)).any()
    g = ring_q32.gram(np.zeros((4, 0, 64), dtype=np.uint32))
    assert g.shape == (10, 64) and not g.any()


def test_gram_symmetry_full_size_c3(oracle, ring_q32):
    # C3 scale: r = 32 parts x 1024 polys. Check the diagonal entries and a
    # few off-diagonal ones exactly against the oracle on those parts.
    r = ring_q32
    rng = np.random.default_rng(55)
    s = small_signed(rng, r.q, (32, 1024, 64), -128, 128).astype(np.uint32)
    g = r.gram(s)
    assert g.shape == (528, 64)

    def idx(i, j):
        return 32 * i - i * (i - 1) // 2 + (j - i)

    for (i, j) in [(0, 0), (0, 31), (7, 19), (31, 31)]:
        sub = np.stack([s[i], s[j]]).astype(np.uint64)
        exp = oracle.gram(r.q, 64, sub)  # [(i,i), (i,j), (j,j)]
        assert np.array_equal(g[idx(i, j)].astype(np.uint64), exp[1] if i != j else exp[0])


# ---- JL --------------------------------------------------------------------

def test_jl_golden(golden, ring_q32):
    j = golden["jl"]
    pm, pe, _ = ring_q32.jl_project(bytes.fromhex(j["seed"]), j["rows"], arr(j["parts"]))
    assert pm.astype(np.uint64).tolist() == j["p_modq"]
    assert pe.tolist() == j["p_exact"]
    s = golden["jl_small"]
    pm, pe, fl = ring_q32.jl_project(bytes.fromhex(s["seed"]), s["rows"], arr(s["parts"]),
                                     want_flat=True)
    assert pm.astype(np.uint64).tolist() == s["p_modq"]
    assert pe.tolist() == s["p_exact"]
    assert fl.tolist() == s["flat"]


@pytest.mark.parametrize("rows,parts,cols", [(256, 3, 64 * 7), (256, 1, 64 * 300), (17, 2, 128)])
def test_jl_matches_oracle(oracle, ring_q32, rows, parts, cols):
    rng = np.random.default_rng(rows + parts + cols)
    seed = rng.integers(0, 256, 32, dtype=np.uint8).tobytes()
    x = small_signed(rng, Q32, (parts, cols), -128, 128)
    pm, pe, _ = ring_q32.jl_project(seed, rows, x)
    em, ee, _ = oracle.jl_project(Q32, seed, rows, x, 64)
    assert np.array_equal(pm.astype(np.uint64), em)
    assert np.array_equal(pe, ee)


def test_jl_goldilocks_matches_oracle(oracle, ring_gold):
    rng = np.random.default_rng(3)
    seed = bytes(range(32))
    x = rand_residues(rng, GOLDILOCKS, (2, 256))
    pm, _, _ = ring_gold.jl_project(seed, 256, x, want_exact=False)
    em, _, _ = oracle.jl_project(GOLDILOCKS, seed, 256, x, 64)
    assert np.array_equal(pm, em)
