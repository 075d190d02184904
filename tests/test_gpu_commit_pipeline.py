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
