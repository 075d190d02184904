"""The int8 matrix-core JL kernels sum in int32, and their launch shapes keep
every such sum exact: each product is a JL entry in {-1, 0, 1} times a signed
base-256 digit in [-128, 127], so a sum of K products stays within K * 128,
and the shape must give K * 128 <= 2^31 - 1 for every accepted size (up to
2^36 witness polynomials, 65535 parts x 2^20). Host arithmetic only: no GPU.
"""
import numpy as np
import pytest

INT32_MAX = (1 << 31) - 1
MAX_POLYS = 1 << 36


def totals():
    """Every total up to 2^16, every power of two up to 2^36 with its
    neighbours, the chunk and cap boundaries, and a seeded log-uniform sample
    up to 2^36."""
    out = set(range(0, 1 << 16))
    for e in range(17, 37):
        for d in (-3, -1, 0, 1, 3):
            v = (1 << e) + d
            if 0 <= v <= MAX_POLYS:
                out.add(v)
    for cus in (1, 80, 256, 304, 1024):
        for per in (65535 * 4 * cus, 262136 * min(cus, 1024)):
            for k in (1, 2, 3, 1000):
                for d in (-1, 0, 1):
                    v = per * k + d
                    if 0 <= v <= MAX_POLYS:
                        out.add(v)
    rng = np.random.default_rng(5)
    out.update(int(x) for x in np.exp2(rng.uniform(16, 36, size=20000)).astype(np.uint64))
    out.add(MAX_POLYS)
    return sorted(out)


@pytest.mark.parametrize("num_cus", [1, 80, 256, 304])
def test_projection_shape_keeps_int32_exact(num_cus):
    import lattica_amd as la
    for total in totals():
        s = la.jl_launch_shape(la.JL_KERNEL_PROJECT, total, num_cus)
        ppw = s["per_unit"]
        assert s["terms"] == 4 * 64 * ppw and s["bound"] == s["terms"] * 128
        assert s["bound"] <= INT32_MAX, total
        # four waves per block, whole polynomials per wave: every polynomial covered
        assert s["blocks"] * 4 * ppw >= total, total
        assert s["blocks"] >= 1 and s["chunks"] == 1
        assert s["blocks"] <= 0x7FFFFFFF


@pytest.mark.parametrize("num_cus", [1, 80, 256, 304, 2048])
def test_evaluation_shape_keeps_int32_exact(num_cus):
    import lattica_amd as la
    for total in totals():
        s = la.jl_launch_shape(la.JL_KERNEL_EVALS, total, num_cus)
        ppb = s["per_unit"]
        assert ppb % 8 == 0 and ppb >= 8
        assert s["terms"] == 64 * ppb and s["bound"] == s["terms"] * 128
        assert s["bound"] <= INT32_MAX, total
        assert s["blocks"] <= min(num_cus, 1024)
        # the chunks cover the witness, each chunk's blocks cover the chunk
        assert s["chunks"] * s["chunk_polys"] >= total
        assert s["blocks"] * ppb >= min(total, s["chunk_polys"])
        assert s["chunk_polys"] <= s["blocks"] * ppb or total < s["chunk_polys"]


def test_bound_is_tight_at_the_cap():
    """At the largest sizes the shapes sit exactly at their caps, whose bounds
    are the stated constants."""
    import lattica_amd as la
    p = la.jl_launch_shape(la.JL_KERNEL_PROJECT, MAX_POLYS, 256)
    assert p["per_unit"] == 65535 and p["bound"] == 2147450880
    e = la.jl_launch_shape(la.JL_KERNEL_EVALS, MAX_POLYS, 256)
    assert e["per_unit"] == 262136 and e["bound"] == 2147418112
    assert e["chunks"] == -(-MAX_POLYS // (256 * 262136))


def test_shape_query_rejects_bad_arguments():
    import ctypes
    import lattica_amd as la
    s = la.JlShape()
    L = la.lib()
    assert L.gh_jl_launch_shape(0, MAX_POLYS + 1, 256, ctypes.byref(s)) == -20
    assert L.gh_jl_launch_shape(2, 10, 256, ctypes.byref(s)) == -20
    assert L.gh_jl_launch_shape(0, 10, 0, ctypes.byref(s)) == -20
    assert L.gh_jl_launch_shape(0, 10, 256, None) == -20
