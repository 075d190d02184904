import json
import os
import sys

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)

Q32 = 4294966657
GOLDILOCKS = 0xFFFFFFFF00000001


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs the HIP library on cuda:0)")


@pytest.fixture(scope="session")
def golden():
    with open(os.path.join(ROOT, "tests", "golden", "golden.json")) as fh:
        return json.load(fh)


@pytest.fixture(scope="session")
def oracle():
    from oracle import gridref
    gridref.lib()
    return gridref


@pytest.fixture(scope="session")
def ctx():
    import lattica_amd as la
    c = la.Context(0)
    yield c
    c.close()


@pytest.fixture(scope="session")
def ring_q32(ctx):
    import lattica_amd as la
    r = la.Ring(ctx, Q32, 64)
    yield r
    r.close()


@pytest.fixture(scope="session")
def ring_gold(ctx):
    import lattica_amd as la
    r = la.Ring(ctx, GOLDILOCKS, 64)
    yield r
    r.close()


def rand_residues(rng, q, shape):
    """Uniform residues in [0, q) as uint64 (q up to 2^64)."""
    hi = rng.integers(0, 1 << 32, size=shape, dtype=np.uint64)
    lo = rng.integers(0, 1 << 32, size=shape, dtype=np.uint64)
    v = (hi << np.uint64(32)) | lo
    if q < (1 << 64):
        v = v % np.uint64(q)
    return v


def small_signed(rng, q, shape, lo, hi):
    """Integers uniform in [lo, hi) as residues mod q."""
    x = rng.integers(lo, hi, size=shape, dtype=np.int64)
    return np.where(x < 0, np.uint64(q) - (-x).astype(np.uint64), x.astype(np.uint64)).astype(np.uint64)
