"""Pins the oracle's JL aggregation rows (conjugated polynomials, JL
evaluations) before the library is compared against them.

Checks come from what the mathematics fixes: sigma_{-1} is the automorphism
X -> X^{-1}, so a conjugated polynomial evaluated at a root w of X^d + 1 equals
the original at w^{-1}; jl_eval is a sum of ring products, so it evaluates to
the matching sum of products of evaluations; and its constant terms are the
projection p (labrador/src/jl.rs:855-955), which the JL golden vectors pin.
"""
import numpy as np
import pytest

from conftest import GOLDILOCKS, Q32, rand_residues

SEED = bytes(range(7, 39))


def poly_eval(coeffs, x, q):
    acc = 0
    for c in reversed([int(v) for v in coeffs]):
        acc = (acc * x + c) % q
    return acc


def entries(flat, rows, num_parts, d, n):
    return np.asarray(flat, dtype=np.int64).reshape(rows, num_parts, d, n)


@pytest.mark.parametrize("q", [Q32, GOLDILOCKS])
def test_conjugate_is_the_inverse_automorphism(oracle, q):
    rows, num_parts, n, d = 6, 2, 3, 64
    flat = oracle.jl_extract_flat(SEED, rows, num_parts, n * d, d)
    conj = oracle.jl_conjugate_flat(q, flat, rows, num_parts, n, d)
    e = entries(flat, rows, num_parts, d, n)
    psi = oracle.root_of_unity(q, 2 * d)       # a root of X^d + 1
    psi_inv = pow(psi, 2 * d - 1, q)
    for m in range(rows):
        for i in range(num_parts):
            for j in range(n):
                pi = [int(v) % q for v in e[m, i, :, j]]
                assert poly_eval(conj[m, i, j], psi, q) == poly_eval(pi, psi_inv, q)
                assert poly_eval(conj[m, i, j], psi_inv, q) == poly_eval(pi, psi, q)


def test_extract_flat_entries_and_density(oracle):
    flat = oracle.jl_extract_flat(SEED, 256, 2, 1024, 64)
    assert set(np.unique(flat).tolist()) <= {-1, 0, 1}
    frac = np.bincount(flat.astype(np.int64) + 1, minlength=3) / flat.size
    assert abs(frac[1] - 0.5) < 0.01 and abs(frac[0] - 0.25) < 0.01   # jl.rs:573-597
    # the entries do not depend on the witness: same as jl_project's flat
    q = Q32
    parts = rand_residues(np.random.default_rng(1), q, (2, 1024))
    _, _, f2 = oracle.jl_project(q, SEED, 256, parts, 64, want_flat=True)
    assert np.array_equal(flat, f2)


@pytest.mark.parametrize("q", [Q32, GOLDILOCKS])
def test_jl_evals_constant_terms_are_the_projection(oracle, q):
    rng = np.random.default_rng(5)
    num_parts, n, d = 3, 5, 64
    s = rand_residues(rng, q, (num_parts, n, d))
    ev = oracle.jl_evals(q, d, SEED, 256, s)
    pm, _, _ = oracle.jl_project(q, SEED, 256, s.reshape(num_parts, n * d), d)
    assert np.array_equal(ev[:, 0], pm)


def test_jl_evals_reference_round_trip_parameters(oracle):
    """jl.rs:855-955: F_4294967291, d = 4, 2 parts of 2 polynomials, witness
    ((pi * 997 + idx * 311 + 42) mod 100): ct(sum sigma(pi) * s) = p."""
    q, d, num_parts, n = 4294967291, 4, 2, 2
    s = np.array([[[(pi * 997 + (j * d + k) * 311 + 42) % 100 for k in range(d)]
                   for j in range(n)] for pi in range(num_parts)], dtype=np.uint64)
    ev = oracle.jl_evals(q, d, SEED, 256, s, naive=True)
    pm, _, _ = oracle.jl_project(q, SEED, 256, s.reshape(num_parts, n * d), d)
    assert np.array_equal(ev[:, 0], pm)


@pytest.mark.parametrize("q", [Q32, GOLDILOCKS])
def test_jl_evals_evaluation_homomorphism(oracle, q):
    """jl_eval[m](w) = sum_ij pi_ij(w^-1) * s_ij(w) at a root w of X^d + 1."""
    rng = np.random.default_rng(9)
    rows, num_parts, n, d = 256, 2, 3, 64
    s = rand_residues(rng, q, (num_parts, n, d))
    ev = oracle.jl_evals(q, d, SEED, rows, s)
    e = entries(oracle.jl_extract_flat(SEED, rows, num_parts, n * d, d), rows, num_parts, d, n)
    psi = oracle.root_of_unity(q, 2 * d)
    psi_inv = pow(psi, 2 * d - 1, q)
    sv = [[poly_eval(s[i, j], psi, q) for j in range(n)] for i in range(num_parts)]
    for m in range(0, rows, 37):
        want = 0
        for i in range(num_parts):
            for j in range(n):
                pi = [int(v) % q for v in e[m, i, :, j]]
                want = (want + poly_eval(pi, psi_inv, q) * sv[i][j]) % q
        assert poly_eval(ev[m], psi, q) == want, m


def ct_product(a, b, q):
    """Constant term of a * b in Z_q[X]/(X^d + 1)."""
    a = [int(v) for v in a]
    b = [int(v) for v in b]
    d = len(a)
    return (a[0] * b[0] - sum(a[k] * b[d - k] for k in range(1, d))) % q


@pytest.mark.parametrize("q", [Q32, GOLDILOCKS])
def test_jl_aggregate_evaluation_homomorphism(oracle, q):
    """aggregation.rs:237-245 at a root w of X^d + 1: phi_ij(w) = phi0_ij(w) +
    sum_m scale_m(w) * pi_ij^(m)(w^-1), with pi read from the flat entries."""
    rng = np.random.default_rng(21)
    rows, num_parts, n, d = 256, 2, 3, 64
    scale = rand_residues(rng, q, (rows, d))
    phi0 = rand_residues(rng, q, (num_parts, n, d))
    phi = oracle.jl_aggregate(q, d, SEED, rows, scale, num_parts, n, phi=phi0)
    e = entries(oracle.jl_extract_flat(SEED, rows, num_parts, n * d, d), rows, num_parts, d, n)
    psi = oracle.root_of_unity(q, 2 * d)
    psi_inv = pow(psi, 2 * d - 1, q)
    sc = [poly_eval(scale[m], psi, q) for m in range(rows)]
    for i in range(num_parts):
        for j in range(n):
            want = poly_eval(phi0[i, j], psi, q)
            for m in range(rows):
                pi = [int(v) % q for v in e[m, i, :, j]]
                want = (want + sc[m] * poly_eval(pi, psi_inv, q)) % q
            assert poly_eval(phi[i, j], psi, q) == want, (i, j)


def test_jl_aggregate_constant_scale_gives_the_projection(oracle):
    """With scale_m = omega_m (constants) the aggregated phi pairs with the
    witness to sum_m omega_m p[m], the check the verifier makes on b''
    (aggregation.rs:158-178)."""
    q = Q32
    rng = np.random.default_rng(22)
    rows, num_parts, n, d = 256, 3, 2, 64
    omega = rand_residues(rng, q, (rows,))
    scale = np.zeros((rows, d), dtype=np.uint64)
    scale[:, 0] = omega
    s = rand_residues(rng, q, (num_parts, n, d))
    phi = oracle.jl_aggregate(q, d, SEED, rows, scale, num_parts, n)
    pm, _, _ = oracle.jl_project(q, SEED, rows, s.reshape(num_parts, n * d), d)
    lhs = sum(ct_product(phi[i, j], s[i, j], q) for i in range(num_parts) for j in range(n)) % q
    rhs = sum(int(omega[m]) * int(pm[m]) for m in range(rows)) % q
    assert lhs == rhs


@pytest.mark.parametrize("q", [Q32, GOLDILOCKS])
def test_jl_aggregate_pairs_with_the_jl_evaluations(oracle, q):
    """sum_ij ct(phi_ij * s_ij) = sum_m ct(scale_m * jl_eval[m]): the two sums
    regroup the same products sigma(pi_ij^(m)) * scale_m * s_ij."""
    rng = np.random.default_rng(23)
    rows, num_parts, n, d = 256, 2, 3, 64
    scale = rand_residues(rng, q, (rows, d))
    s = rand_residues(rng, q, (num_parts, n, d))
    phi = oracle.jl_aggregate(q, d, SEED, rows, scale, num_parts, n)
    ev = oracle.jl_evals(q, d, SEED, rows, s)
    lhs = sum(ct_product(phi[i, j], s[i, j], q) for i in range(num_parts) for j in range(n)) % q
    rhs = sum(ct_product(scale[m], ev[m], q) for m in range(rows)) % q
    assert lhs == rhs


def test_jl_aggregate_naive_products_agree(oracle):
    q = Q32
    rng = np.random.default_rng(24)
    scale = rand_residues(rng, q, (256, 64))
    a = oracle.jl_aggregate(q, 64, SEED, 256, scale, 1, 2)
    b = oracle.jl_aggregate(q, 64, SEED, 256, scale, 1, 2, naive=True)
    assert np.array_equal(a, b)
