"""Back-to-back device-memory commitments. A device-memory gh_commit holds its
finalize back; the next such commitment runs it as the first blocks of its
own launch, and any other call on the context issues it first. Whatever
follows a commitment, its result must equal the host-memory commitment of the
same inputs (which finalizes in place)."""
import numpy as np
import pytest

from conftest import Q32, rand_residues, small_signed

pytestmark = pytest.mark.gpu


def _key(la, ring, rng, rows, cols):
    a = rand_residues(rng, ring.q, (rows, cols, 64)).astype(ring.dtype)
    return la.Key(ring, a, domain=la.DOMAIN_NTT)


def _case(ctx, ring, key, rng):
    w = small_signed(rng, ring.q, (key.msg_len, 64), -128, 128).astype(ring.dtype)
    want = key.commit(w)
    return ctx.to_device(w), ctx.alloc(key.rows * 64 * ring.info.word_bytes), want


@pytest.mark.parametrize("name", ["q32", "goldilocks"])
def test_back_to_back_commitments_of_mixed_shapes(ctx, ring_q32, ring_gold, name):
    """Keys of different row counts and lengths, so consecutive commitments
    need partial-sum areas of different sizes (the largest does not fit beside
    a pending one and takes the flush path); one synchronize at the end."""
    import lattica_amd as la
    ring = ring_q32 if name == "q32" else ring_gold
    rng = np.random.default_rng(61)
    keys = [_key(la, ring, rng, 16, 300), _key(la, ring, rng, 8, 300),
            _key(la, ring, rng, 16, 4096), _key(la, ring, rng, 4, 40),
            _key(la, ring, rng, 32, 1000), _key(la, ring, rng, 12, 1)]
    cases = []
    try:
        for i in range(17):
            key = keys[i % len(keys)]
            s, t, want = _case(ctx, ring, key, rng)
            cases.append((key, s, t, want))
        for key, s, t, _ in cases:
            key.commit_device(s, key.msg_len, t)
        ctx.synchronize()
        for i, (key, s, t, want) in enumerate(cases):
            got = t.download(ring.dtype, (key.rows, 64))
            assert np.array_equal(got, want), i
    finally:
        for _, s, t, _ in cases:
            s.free()
            t.free()
        for k in keys:
            k.close()


def test_rings_of_different_moduli_interleaved(ctx):
    """The held-back finalize carries its own ring's modulus and tables, and a
    Goldilocks commitment after a q32 one (or before) issues the other field's
    pending finalize first."""
    import lattica_amd as la
    rng = np.random.default_rng(62)
    rings = [la.Ring(ctx, q, 64) for q in (Q32, 12289, la.GOLDILOCKS, 8380417)]
    keys = [_key(la, r, rng, 16, 257) for r in rings]
    cases = []
    try:
        for i in range(12):
            key = keys[i % 4]
            cases.append((key,) + _case(ctx, key.ring, key, rng))
        for key, s, t, _ in cases:
            key.commit_device(s, key.msg_len, t)
        ctx.synchronize()
        for i, (key, s, t, want) in enumerate(cases):
            assert np.array_equal(t.download(key.ring.dtype, (key.rows, 64)), want), i
    finally:
        for _, s, t, _ in cases:
            s.free()
            t.free()
        for k in keys:
            k.close()
        for r in rings:
            r.close()


def test_every_follow_up_call_sees_the_finished_commitment(ctx, ring_q32):
    """A device commitment followed directly by: a download of its output, a
    transform of its output, a host-memory commitment, a key destroy, a free of
    another buffer, and a batch; no synchronize in between."""
    import lattica_amd as la
    rng = np.random.default_rng(63)
    key = _key(la, ring_q32, rng, 16, 512)
    other = _key(la, ring_q32, rng, 8, 64)
    bufs = []
    try:
        # download straight after the commitment
        s, t, want = _case(ctx, ring_q32, key, rng)
        bufs += [s, t]
        key.commit_device(s, key.msg_len, t)
        assert np.array_equal(t.download(ring_q32.dtype, (16, 64)), want)

        # forward transform of the output, in place on the device
        s2, t2, want2 = _case(ctx, ring_q32, key, rng)
        bufs += [s2, t2]
        key.commit_device(s2, key.msg_len, t2)
        L = la.lib()
        assert L.gh_twisted_forward(ring_q32.h, t2.ptr, 16, la.MEM_DEVICE) == 0
        got = t2.download(ring_q32.dtype, (16, 64))
        assert np.array_equal(got, ring_q32.twisted_forward(want2))

        # a host-memory commitment after a device one
        s3, t3, want3 = _case(ctx, ring_q32, key, rng)
        bufs += [s3, t3]
        key.commit_device(s3, key.msg_len, t3)
        w4 = small_signed(rng, ring_q32.q, (other.msg_len, 64), -1, 2).astype(np.uint32)
        want4 = other.commit(w4)
        assert np.array_equal(other.commit(w4), want4)
        assert np.array_equal(t3.download(ring_q32.dtype, (16, 64)), want3)

        # destroying another key, then freeing another buffer
        s5, t5, want5 = _case(ctx, ring_q32, key, rng)
        bufs += [s5, t5]
        key.commit_device(s5, key.msg_len, t5)
        other.close()
        other = None
        spare = ctx.alloc(4096)
        key.commit_device(s5, key.msg_len, t5)
        spare.free()
        assert np.array_equal(t5.download(ring_q32.dtype, (16, 64)), want5)

        # a batch after a device commitment
        s6, t6, want6 = _case(ctx, ring_q32, key, rng)
        bufs += [s6, t6]
        key.commit_device(s6, key.msg_len, t6)
        wb = [small_signed(rng, ring_q32.q, (n, 64), -1, 2).astype(np.uint32) for n in (512, 100)]
        outs = key.commit_batch(wb)
        assert np.array_equal(outs[0], key.commit(wb[0]))
        assert np.array_equal(t6.download(ring_q32.dtype, (16, 64)), want6)
    finally:
        for b in bufs:
            b.free()
        key.close()
        if other is not None:
            other.close()


def test_profiled_commitments(ctx, ring_q32):
    """With phase timing on, the MAC phase is recorded per commitment and the
    held-back finalizes that run on their own (the last one) as phase 2."""
    import lattica_amd as la
    rng = np.random.default_rng(64)
    key = _key(la, ring_q32, rng, 16, 1024)
    s, t, want = _case(ctx, ring_q32, key, rng)
    try:
        ctx.profile(True)
        for _ in range(5):
            key.commit_device(s, key.msg_len, t)
        prof = ctx.profile_read()
        ctx.profile(False)
        assert prof[1][1] == 5 and prof[2][1] >= 1
        assert np.array_equal(t.download(np.uint32, (16, 64)), want)
    finally:
        s.free()
        t.free()
        key.close()


@pytest.mark.parametrize("name", ["q32", "goldilocks"])
def test_single_item_device_batches_back_to_back(ctx, ring_q32, ring_gold, name):
    """A one-item device-memory gh_commit_batch takes the per-call path with its
    finalize held back: back-to-back single items of different lengths, then a
    grouped batch (which issues the pending finalize first)."""
    import lattica_amd as la
    ring = ring_q32 if name == "q32" else ring_gold
    rng = np.random.default_rng(65)
    key = _key(la, ring, rng, 16, 700)
    lens = [700, 1, 333, 64, 699]
    ws = [small_signed(rng, ring.q, (n, 64), -128, 128).astype(ring.dtype) for n in lens]
    bufs = []
    try:
        outs = []
        for w in ws:
            s = ctx.to_device(w)
            t = ctx.alloc(16 * 64 * ring.info.word_bytes)
            bufs += [s, t]
            outs.append(t)
            key.commit_batch_device([s], [w.shape[0]], [t])
        grouped = key.commit_batch(ws)  # the grouped kernels: an independent path
        ctx.synchronize()
        for i in range(len(ws)):
            assert np.array_equal(outs[i].download(ring.dtype, (16, 64)), grouped[i]), i
    finally:
        for b in bufs:
            b.free()
        key.close()


def _commit_raw(la, key, s_ptr, n, t_ptr):
    assert la.lib().gh_commit(key.h, s_ptr, n, None, 0, la.DOMAIN_COEFF, t_ptr,
                              la.MEM_DEVICE) == 0


@pytest.mark.parametrize("name", ["q32", "goldilocks"])
def test_commitment_of_a_commitment(ctx, ring_q32, ring_gold, name):
    """The witness of a device commitment is the previous commitment's output
    (whole, and as the tail of a longer witness), with no synchronize in
    between: the held-back finalize that writes it must complete before the
    next launch reads it."""
    import lattica_amd as la
    ring = ring_q32 if name == "q32" else ring_gold
    wb = ring.info.word_bytes
    rng = np.random.default_rng(66)
    k1 = _key(la, ring, rng, 16, 600)
    k2 = _key(la, ring, rng, 16, 16)       # commits t1 itself
    k3 = _key(la, ring, rng, 16, 40 + 16)  # commits [w3 || t2]
    w1 = small_signed(rng, ring.q, (600, 64), -128, 128).astype(ring.dtype)
    w3 = rand_residues(rng, ring.q, (40, 64)).astype(ring.dtype)
    s1 = ctx.to_device(w1)
    t1 = ctx.alloc(16 * 64 * wb)
    s3 = ctx.alloc(56 * 64 * wb)
    s3.upload(w3)
    t2_in_s3 = s3.offset(40 * 64 * wb)
    t3 = ctx.alloc(16 * 64 * wb)
    t4 = ctx.alloc(16 * 64 * wb)
    try:
        for _ in range(2):
            _commit_raw(la, k1, s1.ptr, 600, t1.ptr)
            _commit_raw(la, k2, t1.ptr, 16, t2_in_s3)   # witness = t1 (pending)
            _commit_raw(la, k3, s3.ptr, 56, t3.ptr)     # witness tail = t2 (pending)
            _commit_raw(la, k2, t3.ptr, 16, t4.ptr)     # witness = t3 (pending)
            ctx.synchronize()
            e1 = k1.commit(w1)
            e2 = k2.commit(e1)
            e3 = k3.commit(np.concatenate([w3, e2]))
            e4 = k2.commit(e3)
            assert np.array_equal(t1.download(ring.dtype, (16, 64)), e1)
            assert np.array_equal(s3.download(ring.dtype, (56, 64))[40:], e2)
            assert np.array_equal(t3.download(ring.dtype, (16, 64)), e3)
            assert np.array_equal(t4.download(ring.dtype, (16, 64)), e4)
    finally:
        for b in (s1, t1, s3, t3, t4):
            b.free()
        for k in (k1, k2, k3):
            k.close()


def _hip():
    import ctypes
    import glob
    for cand in ["libamdhip64.so", *sorted(glob.glob("/opt/rocm*/lib/libamdhip64.so*"))]:
        try:
            return ctypes.CDLL(cand)
        except OSError:
            continue
    pytest.skip("libamdhip64 not loadable")


@pytest.mark.parametrize("name", ["q32", "goldilocks"])
def test_defer_off_completes_in_stream_order(ctx, ring_q32, ring_gold, name):
    """With defer_finalize = 0 a device commitment is complete once the device
    is synchronized through HIP itself (no library call in between)."""
    import ctypes
    import lattica_amd as la
    ring = ring_q32 if name == "q32" else ring_gold
    rng = np.random.default_rng(67)
    key = _key(la, ring, rng, 16, 1000)
    hip = _hip()
    bufs = []
    ctx.set_tuning(defer_finalize=0)
    try:
        assert ctx.get_tuning()["defer_finalize"] == 0
        cases = [_case(ctx, ring, key, rng) for _ in range(3)]
        for s, t, _ in cases:
            bufs += [s, t]
        for s, t, _ in cases:
            key.commit_device(s, key.msg_len, t)
        assert hip.hipDeviceSynchronize() == 0
        for s, t, want in cases:
            out = np.empty((16, 64), dtype=ring.dtype)
            assert hip.hipMemcpy(ctypes.c_void_p(out.ctypes.data), t.ptr,
                                 ctypes.c_size_t(out.nbytes), 2) == 0  # DeviceToHost
            assert np.array_equal(out, want)
    finally:
        ctx.set_tuning(defer_finalize=1)
        for b in bufs:
            b.free()
        key.close()


def test_turning_defer_off_completes_a_held_back_step(ctx, ring_q32):
    import ctypes
    import lattica_amd as la
    rng = np.random.default_rng(68)
    key = _key(la, ring_q32, rng, 16, 300)
    hip = _hip()
    s, t, want = _case(ctx, ring_q32, key, rng)
    try:
        key.commit_device(s, key.msg_len, t)  # held back
        ctx.set_tuning(defer_finalize=0)      # enqueues it
        assert hip.hipDeviceSynchronize() == 0
        out = np.empty((16, 64), dtype=np.uint32)
        assert hip.hipMemcpy(ctypes.c_void_p(out.ctypes.data), t.ptr,
                             ctypes.c_size_t(out.nbytes), 2) == 0
        assert np.array_equal(out, want)
    finally:
        ctx.set_tuning(defer_finalize=1)
        s.free()
        t.free()
        key.close()


@pytest.mark.parametrize("stage1", [-1, 0, 1, 3, 8])
@pytest.mark.parametrize("name", ["q32", "goldilocks"])
def test_long_pipeline_of_varied_commitments(oracle, ctx, ring_q32, ring_gold, name, stage1):
    """Forty device commitments back to back on one ring (one-item device
    batches, which take the per-call path with its finalize held back), no
    synchronize in between: lengths from 1 to 4096 columns and row counts from
    4 to 64, so the held-back stages of one commitment sometimes fit the next
    launches and sometimes are issued on their own. One output buffer is
    reused by three commitments in a row (the last one must win). Every
    placement of the first held-back stage (stage1_tasks_per_block)."""
    ctx.set_tuning(stage1_tasks_per_block=stage1)
    try:
        _long_pipeline(oracle, ctx, ring_q32 if name == "q32" else ring_gold)
    finally:
        ctx.set_tuning(stage1_tasks_per_block=-1)


def _long_pipeline(oracle, ctx, ring):
    import lattica_amd as la
    wb = ring.info.word_bytes
    rng = np.random.default_rng(69)
    shapes = [(16, 4096), (16, 512), (4, 300), (64, 1000), (16, 1), (8, 2048), (32, 4096),
              (12, 777), (16, 4096)]
    keys, mats = [], []
    for r, c in shapes:
        a = rand_residues(rng, ring.q, (r, c, 64)).astype(ring.dtype)
        keys.append(la.Key(ring, a, domain=la.DOMAIN_NTT))
        mats.append(a)
    shared = ctx.alloc(64 * 64 * wb)
    cases, bufs = [], [shared]
    try:
        for i in range(40):
            ki = int(rng.integers(0, len(keys)))
            key = keys[ki]
            n = int(rng.integers(1, key.msg_len + 1)) if i % 3 else key.msg_len
            w = small_signed(rng, ring.q, (n, 64), -128, 128).astype(ring.dtype)
            s = ctx.to_device(w)
            bufs.append(s)
            if i in (20, 21, 22):
                t = shared
            else:
                t = ctx.alloc(key.rows * 64 * wb)
                bufs.append(t)
            cases.append((ki, n, w, s, t, i))
        for ki, n, w, s, t, _ in cases:
            keys[ki].commit_batch_device([s], [n], [t])
        ctx.synchronize()
        for ki, n, w, s, t, i in cases:
            if t is shared and i != 22:
                continue
            key = keys[ki]
            got = t.download(ring.dtype, (key.rows, 64))
            if n <= 64:
                z = np.zeros((key.rows, 1, 64), dtype=np.uint64)
                want = oracle.commit(ring.q, 64, mats[ki][:, :n].astype(np.uint64), z,
                                     w.astype(np.uint64), np.zeros((1, 64), dtype=np.uint64),
                                     key_domain=1)
                assert np.array_equal(got.astype(np.uint64), want), (i, key.rows, n)
            else:
                assert np.array_equal(got, key.commit_batch([w])[0]), (i, key.rows, n)
    finally:
        for b in bufs:
            b.free()
        for k in keys:
            k.close()


@pytest.fixture(scope="module")
def rotation(ctx, ring_q32, ring_gold):
    """Per ring: three kappa-16 NTT-domain keys over the full key width (2^15
    q32 columns, 2^14 Goldilocks) and four resident witnesses."""
    import lattica_amd as la
    made = {}
    for name, ring, cols in (("q32", ring_q32, 1 << 15), ("goldilocks", ring_gold, 1 << 14)):
        rng = np.random.default_rng(4040 + cols)
        keys = [la.Key(ring, rand_residues(rng, ring.q, (16, cols, 64)).astype(ring.dtype),
                       domain=la.DOMAIN_NTT) for _ in range(3)]
        ws = [ctx.to_device(small_signed(rng, ring.q, (cols, 64), -128, 128).astype(ring.dtype))
              for _ in range(4)]
        made[name] = (ring, cols, keys, ws)
    yield made
    for ring, cols, keys, ws in made.values():
        for b in ws:
            b.free()
        for k in keys:
            k.close()


@pytest.mark.parametrize("stage1", [-1, 0, 1, 2, 4, 8])
@pytest.mark.parametrize("name", ["q32", "goldilocks"])
def test_headline_shape_back_to_back_every_stage1_placement(ctx, rotation, name, stage1):
    """Twelve device commitments at kappa 16 over the full key width, three
    keys in rotation, no synchronize in between, for each
    stage1_tasks_per_block: every output equals the one of the same
    commitment made alone with deferral off."""
    ring, cols, keys, ws = rotation[name]
    wb = ring.info.word_bytes
    outs = [ctx.alloc(16 * 64 * wb) for _ in range(12)]
    ref = ctx.alloc(16 * 64 * wb)
    try:
        ctx.set_tuning(stage1_tasks_per_block=stage1)
        for i in range(12):
            keys[i % 3].commit_device(ws[i % 4], cols, outs[i])
        ctx.synchronize()
        got = [o.download(ring.dtype, (16, 64)) for o in outs]
        ctx.set_tuning(defer_finalize=0)
        for i in range(12):
            keys[i % 3].commit_device(ws[i % 4], cols, ref)
            ctx.synchronize()
            assert np.array_equal(got[i], ref.download(ring.dtype, (16, 64))), (i, stage1)
    finally:
        ctx.set_tuning(defer_finalize=1, stage1_tasks_per_block=-1)
        for b in outs + [ref]:
            b.free()
