"""Serving loop (gh_server_*) through the C ABI.

Every request's result must equal the plain per-call result on the same inputs
bit for bit (gh_commit on the key's first n columns via a one-item
gh_commit_batch, gh_gram), whether the request runs alone or in a mix of other
requests of both rings, both kinds and many sizes; small cases are also
checked against the oracle.
"""
import numpy as np
import pytest

from conftest import rand_residues, small_signed

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def keys(ring_q32, ring_gold):
    import lattica_amd as la
    rng = np.random.default_rng(404)
    kq = la.Key(ring_q32, rand_residues(rng, ring_q32.q, (16, 2048, 64)), domain=la.DOMAIN_NTT)
    kg = la.Key(ring_gold, rand_residues(rng, ring_gold.q, (8, 1024, 64)), domain=la.DOMAIN_NTT)
    yield {"q32": kq, "gold": kg}
    kq.close()
    kg.close()


def make_requests(rng, rings, keys, count, max_r=6, max_log_n=10):
    import lattica_amd as la
    reqs = []
    for i in range(count):
        name = "q32" if rng.random() < 0.5 else "gold"
        ring, key = rings[name], keys[name]
        ops = int(rng.choice([la.OP_COMMIT, la.OP_GRAM, la.OP_COMMIT | la.OP_GRAM]))
        r = int(rng.integers(1, max_r + 1))
        n = int(np.floor(2.0 ** rng.uniform(0, max_log_n)))
        n = min(n, key.msg_len)
        w = small_signed(rng, ring.q, (r, n, 64), -128, 128).astype(ring.dtype)
        reqs.append((name, ops, r, n, w))
    return reqs


def plain_result(rings, keys, req):
    import lattica_amd as la
    name, ops, r, n, w = req
    ring, key = rings[name], keys[name]
    t = np.stack([key.commit_batch([w[i]])[0] for i in range(r)]) if ops & la.OP_COMMIT else None
    g = ring.gram(w) if ops & la.OP_GRAM else None
    return t, g


def submit(srv, rings, keys, req):
    name, ops, r, n, w = req
    return srv.submit(ops, rings[name], w, r, n, key=keys[name])


def test_server_alone_equals_mix_equals_plain(ctx, ring_q32, ring_gold, keys):
    import lattica_amd as la
    rings = {"q32": ring_q32, "gold": ring_gold}
    rng = np.random.default_rng(2024)
    reqs = make_requests(rng, rings, keys, 48)
    plain = [plain_result(rings, keys, q) for q in reqs]
    srv = la.Server(ctx, mem_bytes=1 << 30)
    try:
        alone = []
        for q in reqs:
            st, lat, t, g = srv.wait(submit(srv, rings, keys, q))
            assert st == 0 and lat > 0
            alone.append((t, g))
        tickets = [submit(srv, rings, keys, q) for q in reqs]  # all in flight together
        mixed = [srv.wait(tk) for tk in tickets]
        stats = srv.stats()
    finally:
        srv.close()
    for q, p, a, m in zip(reqs, plain, alone, mixed):
        assert m[0] == 0
        for k in (0, 1):
            if p[k] is None:
                assert a[k] is None and m[k + 2] is None
            else:
                assert np.array_equal(a[k], p[k]), (q[:4], k)
                assert np.array_equal(m[k + 2], p[k]), (q[:4], k)
    assert stats["completed"] == 2 * len(reqs) and stats["failed"] == 0
    assert stats["max_wave_seen"] > 1  # the mix really was batched
    assert stats["mem_in_use"] == 0


def test_server_matches_oracle(oracle, ctx, ring_q32, ring_gold):
    import lattica_amd as la
    rng = np.random.default_rng(77)
    srv = la.Server(ctx, mem_bytes=64 << 20)
    try:
        for ring, rows in ((ring_q32, 8), (ring_gold, 4)):
            a = rand_residues(rng, ring.q, (rows, 40, 64))
            key = la.Key(ring, a)
            w = small_signed(rng, ring.q, (3, 29, 64), -128, 128)
            st, _, t, g = srv.wait(srv.submit(la.OP_COMMIT | la.OP_GRAM, ring, w, 3, 29, key=key))
            assert st == 0
            z = np.zeros((rows, 1, 64), dtype=np.uint64)
            for i in range(3):
                exp = oracle.commit(ring.q, 64, a[:, :29], z, w[i], np.zeros((1, 64), np.uint64))
                assert np.array_equal(t[i].astype(np.uint64), exp)
            assert np.array_equal(g.astype(np.uint64), oracle.gram(ring.q, 64, w))
            key.close()
    finally:
        srv.close()


def test_server_device_memory_requests(ctx, ring_q32, ring_gold, keys):
    import lattica_amd as la
    rings = {"q32": ring_q32, "gold": ring_gold}
    rng = np.random.default_rng(5)
    reqs = make_requests(rng, rings, keys, 24)
    plain = [plain_result(rings, keys, q) for q in reqs]
    bufs = []
    srv = la.Server(ctx, mem_bytes=256 << 20, slots=3)
    try:
        tickets = []
        for (name, ops, r, n, w), _ in zip(reqs, plain):
            ring, key = rings[name], keys[name]
            wd = ctx.to_device(w)
            td = ctx.alloc(r * key.rows * 64 * w.itemsize) if ops & la.OP_COMMIT else None
            gd = ctx.alloc(r * (r + 1) // 2 * 64 * w.itemsize) if ops & la.OP_GRAM else None
            bufs.append((wd, td, gd))
            tickets.append(srv.submit(ops, ring, wd, r, n, key=key, t_out=td, g_out=gd,
                                      mem=la.MEM_DEVICE))
        for tk in tickets:
            assert srv.wait(tk)[0] == 0
    finally:
        srv.close()
    for (name, ops, r, n, w), (pt, pg), (wd, td, gd) in zip(reqs, plain, bufs):
        key = keys[name]
        if td is not None:
            assert np.array_equal(td.download(w.dtype, (r, key.rows, 64)), pt)
        if gd is not None:
            assert np.array_equal(gd.download(w.dtype, (r * (r + 1) // 2, 64)), pg)
        for b in (wd, td, gd):
            if b is not None:
                b.free()


def test_server_admits_by_footprint(ctx, ring_q32, keys):
    """An arena that holds only a few requests at a time: everything still
    completes, in arrival order of admission, never above the arena size."""
    import lattica_amd as la
    key = keys["q32"]
    rng = np.random.default_rng(9)
    n, r = 1024, 2
    # footprint of one commit request: NTT witness + partials + staged in/out, about 1.3 MB
    arena = 4 << 20
    srv = la.Server(ctx, mem_bytes=arena)
    ws = [small_signed(rng, ring_q32.q, (r, n, 64), -128, 128).astype(np.uint32) for _ in range(12)]
    try:
        tickets = [srv.submit(la.OP_COMMIT, ring_q32, w, r, n, key=key) for w in ws]
        outs = [srv.wait(tk) for tk in tickets]
        stats = srv.stats()
        with pytest.raises(la.GridHipError) as e:  # can never fit
            srv.submit(la.OP_COMMIT, ring_q32,
                       np.zeros((8, 2048, 64), dtype=np.uint32), 8, 2048, key=key)
        assert e.value.name == "OOM"
    finally:
        srv.close()
    assert 0 < stats["mem_peak"] <= arena
    assert stats["max_wave_seen"] < 12
    for w, (st, _, t, _) in zip(ws, outs):
        assert st == 0
        assert np.array_equal(t[1], key.commit_batch([w[1]])[0])


def test_server_rejects_bad_requests(ctx, ring_q32, ring_gold, keys):
    import lattica_amd as la
    srv = la.Server(ctx, mem_bytes=16 << 20)
    kq = keys["q32"]
    w = np.zeros((1, 4, 64), dtype=np.uint32)
    try:
        cases = [
            (dict(ops=la.OP_COMMIT, ring=ring_q32, witness=np.zeros((1, 4096, 64), np.uint32),
                  r=1, n=4096, key=kq), "DIMENSION_MISMATCH"),  # more columns than the key holds
            (dict(ops=la.OP_COMMIT, ring=ring_gold, witness=w.astype(np.uint64), r=1, n=4,
                  key=kq), "INVALID_ARGUMENT"),  # key of another ring
            (dict(ops=4, ring=ring_q32, witness=w, r=1, n=4, key=kq), "INVALID_ARGUMENT"),
            (dict(ops=la.OP_GRAM, ring=ring_q32, witness=w, r=0, n=4), "DIMENSION_MISMATCH"),
            (dict(ops=la.OP_GRAM, ring=ring_q32, witness=np.full((1, 4, 64), ring_q32.q,
                                                                   dtype=np.uint32), r=1, n=4),
             "NOT_CANONICAL"),
        ]
        for kw, name in cases:
            with pytest.raises(la.GridHipError) as e:
                srv.submit(**kw)
            assert e.value.name == name, kw
        assert srv.stats()["submitted"] == 0
    finally:
        srv.close()


@pytest.mark.parametrize("kappa", [1, 2, 3, 5])
def test_server_keys_of_any_row_count(ctx, ring_q32, ring_gold, kappa):
    """Keys whose row count is not a multiple of 4 (the reference's own Ajtai
    benches use kappa = 1) are served; each result equals the plain call and
    the oracle."""
    import lattica_amd as la
    from oracle import gridref
    rng = np.random.default_rng(410 + kappa)
    srv = la.Server(ctx, mem_bytes=64 << 20)
    keys, tickets = [], []
    try:
        for ring in (ring_q32, ring_gold):
            a = rand_residues(rng, ring.q, (kappa, 300, 64))
            key = la.Key(ring, a, domain=la.DOMAIN_NTT)
            keys.append(key)
            for n, r in ((300, 2), (17, 3), (1, 1), (129, 1)):
                w = small_signed(rng, ring.q, (r, n, 64), -128, 128).astype(ring.dtype)
                tickets.append((ring, key, a, w, srv.submit(la.OP_COMMIT, ring, w, r, n, key=key)))
        for ring, key, a, w, tk in tickets:
            st, _, t, _ = srv.wait(tk)
            assert st == 0
            for i in range(w.shape[0]):
                n = w.shape[1]
                z = np.zeros((kappa, 1, 64), dtype=np.uint64)
                exp = gridref.commit(ring.q, 64, a[:, :n], z, w[i].astype(np.uint64),
                                     np.zeros((1, 64), dtype=np.uint64), key_domain=1)
                assert np.array_equal(t[i].astype(np.uint64), exp), (ring.q, n, i)
                assert np.array_equal(t[i], key.commit_batch([w[i]])[0])
    finally:
        srv.close()
        for k in keys:
            k.close()
