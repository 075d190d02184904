"""LaBRADOR step around the commitment on the GPU versus the oracle: balanced
decomposition (main_protocol/mod.rs:258-320), garbage h (garbage.rs:101-133),
amortization (amortization.rs:61-84), and the outer commitments as gh_commit
(crs.rs:273-304). Bit-exact on both rings."""
import numpy as np
import pytest

from conftest import GOLDILOCKS, Q32, rand_residues, small_signed

pytestmark = pytest.mark.gpu


def ring_for(name, ring_q32, ring_gold):
    return ring_q32 if name == "q32" else ring_gold


@pytest.mark.parametrize("name", ["q32", "goldilocks"])
@pytest.mark.parametrize("base,limbs", [(2, 1), (16, 3), (1 << 12, 3), (1 << 32, 2), (7, 40)])
def test_decompose_balanced_matches_oracle(oracle, ring_q32, ring_gold, name, base, limbs):
    r = ring_for(name, ring_q32, ring_gold)
    rng = np.random.default_rng(base + limbs)
    polys = rand_residues(rng, r.q, (37, 64))
    polys[0, :8] = [0, 1, r.q - 1, r.q // 2, r.q // 2 + 1, base // 2, r.q - base // 2, base % r.q]
    got = r.decompose_balanced(polys, base, limbs)
    assert np.array_equal(got.astype(np.uint64), oracle.decompose_balanced(r.q, 64, polys, base, limbs))


def test_decompose_rejects_bad_base(ring_q32):
    import lattica_amd as la
    for base, limbs in [(1, 2), ((1 << 32) + 1, 2), (16, 0)]:
        with pytest.raises(la.GridHipError) as e:
            ring_q32.decompose_balanced(np.zeros((1, 64), dtype=np.uint32), base, limbs)
        assert e.value.name == "INVALID_PARAMETERS"


@pytest.mark.parametrize("name", ["q32", "goldilocks"])
@pytest.mark.parametrize("rr,n", [(1, 1), (3, 17), (6, 64), (9, 5)])
def test_garbage_h_matches_oracle(oracle, ring_q32, ring_gold, name, rr, n):
    import lattica_amd as la
    r = ring_for(name, ring_q32, ring_gold)
    rng = np.random.default_rng(rr * 31 + n)
    phi = rand_residues(rng, r.q, (rr, n, 64))
    s = small_signed(rng, r.q, (rr, n, 64), -128, 128)
    got = r.garbage_h(phi, s)
    assert np.array_equal(got.astype(np.uint64), oracle.garbage_h(r.q, 64, phi, s))
    got_ntt = r.garbage_h(r.twisted_forward(phi.reshape(-1, 64)).reshape(phi.shape),
                          r.twisted_forward(s.reshape(-1, 64)).reshape(s.shape), domain=la.DOMAIN_NTT)
    assert np.array_equal(got_ntt, got)
    # phi = s gives the Gram matrix
    assert np.array_equal(r.garbage_h(s, s), r.gram(s))


@pytest.mark.parametrize("name", ["q32", "goldilocks"])
@pytest.mark.parametrize("rr,n", [(1, 1), (4, 16), (7, 33), (32, 40)])
def test_amortize_matches_oracle(oracle, ring_q32, ring_gold, name, rr, n):
    import lattica_amd as la
    r = ring_for(name, ring_q32, ring_gold)
    rng = np.random.default_rng(rr * 7 + n)
    c = small_signed(rng, r.q, (rr, 64), -2, 3)
    s = small_signed(rng, r.q, (rr, n, 64), -128, 128)
    got = r.amortize(c, s)
    assert np.array_equal(got.astype(np.uint64), oracle.amortize(r.q, 64, c, s))
    got_ntt = r.amortize(r.twisted_forward(c), r.twisted_forward(s.reshape(-1, 64)).reshape(s.shape),
                         domain=la.DOMAIN_NTT)
    assert np.array_equal(got_ntt, got)


def test_outer_commitments_are_linear_commitments(oracle, ring_q32):
    """u1 = B t + C g (crs.rs:293-301) is gh_commit with key [B | C]; u2 = D h."""
    import lattica_amd as la
    r = ring_q32
    rng = np.random.default_rng(17)
    kappa1, nt, ng = 8, 24, 10
    b = rand_residues(rng, r.q, (kappa1, nt, 64))
    c = rand_residues(rng, r.q, (kappa1, ng, 64))
    t = small_signed(rng, r.q, (nt, 64), -8, 9)
    g = small_signed(rng, r.q, (ng, 64), -8, 9)
    key = la.Key(r, b, c)
    u1 = key.commit(t, g)
    assert np.array_equal(u1.astype(np.uint64), oracle.commit(r.q, 64, b, c, t, g))
    key.close()
