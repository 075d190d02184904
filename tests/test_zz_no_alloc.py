"""gh_commit_noalloc against gh_commit (device memory, coefficient domain).

Two calls in a row per ring, the second with a smaller key and witness, each
compared bit for bit with gh_commit on the same inputs; the first q32 call is
also checked against the oracle. Then: a commitment held back by gh_commit
and overwritten by gh_commit_noalloc into the same output ends as the second;
and the rejections (a key whose partial sums exceed the 4 MiB array, an
opening, a length mismatch, rows % 4 != 0) leave the context working.
"""
import numpy as np
import pytest

from conftest import GOLDILOCKS, Q32, rand_residues, small_signed

pytestmark = pytest.mark.gpu


def _pair(ctx, key, s, words):
    """(gh_commit, gh_commit_noalloc) of witness s against key, device memory."""
    n = s.shape[0]
    s_dev = ctx.to_device(key.ring.words(s))
    t_ref = ctx.alloc(key.rows * 64 * words)
    t_na = ctx.alloc(key.rows * 64 * words)
    try:
        key.commit_device(s_dev, n, t_ref)
        ctx.synchronize()
        key.commit_noalloc(s_dev, n, t_na)
        ctx.synchronize()
        dt = np.uint32 if words == 4 else np.uint64
        return (t_ref.download(dt, (key.rows, 64)), t_na.download(dt, (key.rows, 64)))
    finally:
        s_dev.free()
        t_ref.free()
        t_na.free()


@pytest.mark.parametrize("ring_name", ["q32", "gold"])
def test_noalloc_two_calls_match_commit(oracle, ctx, ring_q32, ring_gold, ring_name):
    import lattica_amd as la
    ring, q, words = (ring_q32, Q32, 4) if ring_name == "q32" else (ring_gold, GOLDILOCKS, 8)
    rng = np.random.default_rng(7707 if ring_name == "q32" else 7708)
    shapes = [(16, 1 << 15), (8, 1000)] if ring_name == "q32" else [(16, 1 << 14), (4, 300)]
    for call, (rows, cols) in enumerate(shapes):
        print(f"case: noalloc {ring_name} call {call} rows {rows} cols {cols}", flush=True)
        a = rand_residues(rng, q, (rows, cols, 64))
        s = small_signed(rng, q, (cols, 64), -128, 128)
        key = la.Key(ring, a)
        try:
            ref, got = _pair(ctx, key, s, words)
        finally:
            key.close()
        assert np.array_equal(got, ref), (ring_name, call)
        if call == 0 and ring_name == "q32":
            z = np.zeros((rows, 1, 64), dtype=np.uint64)
            zr = np.zeros((1, 64), dtype=np.uint64)
            exp = oracle.commit(q, 64, a, z, s, zr)
            assert np.array_equal(got.astype(np.uint64), exp)


def test_noalloc_after_held_back_commit(ctx, ring_q32):
    """gh_commit into t (its finalize held back), then gh_commit_noalloc into
    the same t: the held-back finalize is issued first, so t ends as the
    second commitment."""
    import lattica_amd as la
    print("case: noalloc after a held-back commit into the same output", flush=True)
    rng = np.random.default_rng(7709)
    rows, cols = 16, 4096
    a = rand_residues(rng, Q32, (rows, cols, 64))
    s1 = small_signed(rng, Q32, (cols, 64), -128, 128)
    s2 = small_signed(rng, Q32, (cols, 64), -128, 128)
    key = la.Key(ring_q32, a)
    d1, d2 = ctx.to_device(ring_q32.words(s1)), ctx.to_device(ring_q32.words(s2))
    t = ctx.alloc(rows * 64 * 4)
    t_ref = ctx.alloc(rows * 64 * 4)
    try:
        key.commit_device(d2, cols, t_ref)
        ctx.synchronize()
        exp = t_ref.download(np.uint32, (rows, 64))
        key.commit_device(d1, cols, t)
        key.commit_noalloc(d2, cols, t)
        ctx.synchronize()
        assert np.array_equal(t.download(np.uint32, (rows, 64)), exp)
    finally:
        for b in (d1, d2, t, t_ref):
            b.free()
        key.close()


def test_noalloc_rejections(ctx, ring_q32, ring_gold):
    import lattica_amd as la
    print("case: noalloc rejections", flush=True)
    rng = np.random.default_rng(7710)
    # Goldilocks, 4 rows, 2^17 + 1 columns: at least 2049 blocks x 4 rows x
    # 64 words x 8 bytes = 4196352 bytes, beyond the 4 MiB array
    rows, cols = 4, (1 << 17) + 1
    a = np.zeros((rows, cols, 64), dtype=np.uint64)
    key = la.Key(ring_gold, a)
    s = ctx.to_device(np.zeros((cols, 64), dtype=np.uint64))
    t = ctx.alloc(rows * 64 * 8)
    try:
        with pytest.raises(la.GridHipError) as e:
            key.commit_noalloc(s, cols, t)
        assert e.value.name == "INVALID_ARGUMENT"
        with pytest.raises(la.GridHipError) as e:
            key.commit_noalloc(s, cols - 1, t)
        assert e.value.name == "DIMENSION_MISMATCH"
    finally:
        key.close()
    # one column fewer than the limit's witness still fits: 2048 blocks
    key = la.Key(ring_gold, a[:, : cols - 1])
    try:
        key.commit_noalloc(s, cols - 1, t)
        ctx.synchronize()
    finally:
        key.close()
    del a
    # rows % 4 != 0, and a key with an opening part
    a3 = rand_residues(rng, Q32, (3, 64, 64))
    key = la.Key(ring_q32, a3)
    try:
        with pytest.raises(la.GridHipError) as e:
            key.commit_noalloc(s, 64, t)
        assert e.value.name == "INVALID_ARGUMENT"
    finally:
        key.close()
    a4 = rand_residues(rng, Q32, (4, 64, 64))
    key = la.Key(ring_q32, a4[:, :48], a4[:, 48:])
    try:
        with pytest.raises(la.GridHipError) as e:
            key.commit_noalloc(s, 48, t)
        assert e.value.name == "INVALID_ARGUMENT"
    finally:
        key.close()
    # the context still commits after the rejections
    key = la.Key(ring_q32, a4)
    sm = small_signed(rng, Q32, (64, 64), -128, 128)
    try:
        ref, got = _pair(ctx, key, sm, 4)
        assert np.array_equal(got, ref)
    finally:
        key.close()
        s.free()
        t.free()
