"""Ring slice hooks on the device (gh_add_assign, gh_sub_assign, gh_scalar_mul,
gh_add_assign_scaled, gh_dot_product) against the oracle's restatement of
PrimeField's hooks (algebra/src/arith/prime.rs:433-491), bit-exact.

Lengths cross the 16-byte vector width and the grid stride; device slices are
also taken at a one-word offset (the scalar path) and with src == dst.
"""
import ctypes

import numpy as np
import pytest

from conftest import GOLDILOCKS, Q32, rand_residues

pytestmark = pytest.mark.gpu

OPS = ("add", "sub", "scalar_mul", "add_scaled")

# inputs of the reference's slice-hook tests (prime.rs:945-1224) on the rings
# the device runs (q < 2^32 with a 128th root of unity, and Goldilocks)
KATS = [
    (12289, [1, 7, 12_000, 3, 11], [2, 9, 1_111, 14, 6], 19),
    (3329, [1, 7, 3_000, 3, 11], [2, 9, 1_111, 14, 6], 19),
    (GOLDILOCKS, [1, (1 << 32) - 1, GOLDILOCKS - 2, 123_456_789, 9_876_543_210],
     [2, (1 << 32) + 7, GOLDILOCKS - 3, 987_654_321, 77_777_777], (1 << 32) + 13),
    (8380417, [1, 7, 8_000_000, 3, 11], [2, 9, 111_111, 14, 6], 19),
]


def _host(ring, op, d, s, c):
    if op == "add":
        return ring.add_assign(d, s)
    if op == "sub":
        return ring.sub_assign(d, s)
    if op == "scalar_mul":
        return ring.scalar_mul(d, c)
    return ring.add_assign_scaled(d, s, c)


@pytest.fixture(scope="module")
def rings(ctx):
    import lattica_amd as la
    made = {}
    yield lambda q: made.setdefault(q, la.Ring(ctx, q, 64))
    for r in made.values():
        r.close()


@pytest.mark.parametrize("q,lhs,rhs,scalar", KATS)
def test_reference_inputs(rings, oracle, q, lhs, rhs, scalar):
    ring = rings(q)
    lhs, rhs = np.array(lhs, dtype=np.uint64), np.array(rhs, dtype=np.uint64)
    for op in OPS:
        got = _host(ring, op, lhs, rhs, scalar)
        exp = oracle.slice_op(q, op, lhs, rhs, scalar)
        assert np.array_equal(got.astype(np.uint64), exp), op
    assert ring.dot_product(lhs, rhs) == oracle.dot_product(q, lhs, rhs)


@pytest.mark.parametrize("name", ["q32", "goldilocks"])
@pytest.mark.parametrize("n", [0, 1, 3, 5, 64, 1001, 4099, (1 << 20) + 3])
def test_host_memory_matches_oracle(rings, oracle, name, n):
    q = Q32 if name == "q32" else GOLDILOCKS
    ring = rings(q)
    rng = np.random.default_rng(n * 7 + len(name))
    d = rand_residues(rng, q, (n,))
    s = rand_residues(rng, q, (n,))
    for c in (0, 1, q - 1, int(rand_residues(rng, q, (1,))[0])):
        for op in OPS:
            got = _host(ring, op, d, s, c).astype(np.uint64)
            assert np.array_equal(got, oracle.slice_op(q, op, d, s, c)), (op, c)
    assert ring.dot_product(d, s) == oracle.dot_product(q, d, s)


@pytest.mark.parametrize("name", ["q32", "goldilocks"])
def test_dot_product_all_carries(rings, name):
    # every product (q-1)^2 = 1 mod q is the largest a word pair gives, so the
    # per-thread 64-bit sums carry as often as they can; the result is n mod q
    q = Q32 if name == "q32" else GOLDILOCKS
    ring = rings(q)
    n = (1 << 22) + 5
    a = np.full(n, q - 1, dtype=ring.dtype)
    assert ring.dot_product(a, a) == n % q


@pytest.mark.parametrize("name", ["q32", "goldilocks"])
@pytest.mark.parametrize("offset", [0, 1])
def test_device_memory_offsets_and_aliasing(ctx, rings, oracle, name, offset):
    import lattica_amd as la
    q = Q32 if name == "q32" else GOLDILOCKS
    ring = rings(q)
    wb = np.dtype(ring.dtype).itemsize
    n = 70001
    rng = np.random.default_rng(11 + offset)
    d = rand_residues(rng, q, (n,))
    s = rand_residues(rng, q, (n,))
    c = int(rand_residues(rng, q, (1,))[0])
    buf_d = la.DeviceBuffer(ctx, (n + 2) * wb)
    buf_s = la.DeviceBuffer(ctx, (n + 2) * wb)
    buf_o = la.DeviceBuffer(ctx, 256)
    try:
        pd, ps = buf_d.offset(offset * wb), buf_s.offset(offset * wb)
        L = la.lib()
        for op in OPS:
            host = np.zeros(n + 2, dtype=ring.dtype)
            host[offset:offset + n] = d
            buf_d.upload(host)
            host[offset:offset + n] = s
            buf_s.upload(host)
            if op == "add":
                st = L.gh_add_assign(ring.h, pd, ps, n, la.MEM_DEVICE)
            elif op == "sub":
                st = L.gh_sub_assign(ring.h, pd, ps, n, la.MEM_DEVICE)
            elif op == "scalar_mul":
                st = L.gh_scalar_mul(ring.h, pd, c, n, la.MEM_DEVICE)
            else:
                st = L.gh_add_assign_scaled(ring.h, pd, ps, c, n, la.MEM_DEVICE)
            assert st == 0, op
            ctx.synchronize()
            got = buf_d.download(ring.dtype, (n + 2,))
            assert np.array_equal(got[offset:offset + n].astype(np.uint64),
                                  oracle.slice_op(q, op, d, s, c)), op
            # the words around the slice are untouched
            assert got[:offset].tolist() == [0] * offset and got[offset + n:].tolist() == [0] * (2 - offset)
        # src == dst: x + x, then x + c x
        host = np.zeros(n + 2, dtype=ring.dtype)
        host[offset:offset + n] = d
        buf_d.upload(host)
        assert L.gh_add_assign(ring.h, pd, pd, n, la.MEM_DEVICE) == 0
        assert L.gh_add_assign_scaled(ring.h, pd, pd, c, n, la.MEM_DEVICE) == 0
        ctx.synchronize()
        got = buf_d.download(ring.dtype, (n + 2,))[offset:offset + n].astype(np.uint64)
        x = oracle.slice_op(q, "add", d, d)
        x = oracle.slice_op(q, "add_scaled", x, x, c)
        assert np.array_equal(got, x)
        # dot product into a device word
        buf_d.upload(np.concatenate([np.zeros(offset, ring.dtype), d.astype(ring.dtype)]))
        buf_s.upload(np.concatenate([np.zeros(offset, ring.dtype), s.astype(ring.dtype)]))
        assert L.gh_dot_product(ring.h, pd, ps, n, buf_o.ptr, la.MEM_DEVICE) == 0
        ctx.synchronize()
        assert int(buf_o.download(ring.dtype, (1,))[0]) == oracle.dot_product(q, d, s)
    finally:
        buf_d.free()
        buf_s.free()
        buf_o.free()


def test_rejections(rings):
    import lattica_amd as la
    ring = rings(Q32)
    good = np.arange(10, dtype=np.uint32)
    bad = good.copy()
    bad[7] = Q32
    L = la.lib()
    with pytest.raises(la.GridHipError) as e:
        ring.scalar_mul(good, Q32)
    assert e.value.name == "NOT_CANONICAL"
    with pytest.raises(la.GridHipError) as e:
        ring.add_assign_scaled(good, good, Q32 + 5)
    assert e.value.name == "NOT_CANONICAL"
    with pytest.raises(la.GridHipError) as e:
        ring.add_assign(bad, good)
    assert e.value.name == "NOT_CANONICAL"
    with pytest.raises(la.GridHipError) as e:
        ring.sub_assign(good, bad)
    assert e.value.name == "NOT_CANONICAL"
    with pytest.raises(la.GridHipError) as e:
        ring.dot_product(good, bad)
    assert e.value.name == "NOT_CANONICAL"
    # null pointers with a nonzero length, and a bad memory flag
    assert L.gh_add_assign(ring.h, None, ctypes.c_void_p(good.ctypes.data), 10, la.MEM_HOST) == -20
    assert L.gh_dot_product(ring.h, ctypes.c_void_p(good.ctypes.data),
                            ctypes.c_void_p(good.ctypes.data), 10, None, la.MEM_HOST) == -20
    assert L.gh_scalar_mul(ring.h, ctypes.c_void_p(good.ctypes.data), 3, 10, 7) == -20
    # empty slices are fine, and the empty dot product is zero
    assert ring.add_assign(good[:0], good[:0]).size == 0
    assert ring.dot_product(good[:0], good[:0]) == 0
