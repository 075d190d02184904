"""Pins the oracle (oracle/gridref.c) before anything is compared against it.

Each check is chosen so a plausible restatement mistake fails it: golden vectors
from closed forms and schoolbook sums (tests/golden/make_golden.py), the
reference's own known answers, RFC 8439 vectors, round trips, and invariants.
"""
import numpy as np
import pytest

from conftest import GOLDILOCKS, Q32, rand_residues, small_signed

F17 = 17
F3329 = 3329
F12289 = 12289


def arr(x):
    return np.asarray(x, dtype=np.uint64)


# ---- field arithmetic KATs (algebra/src/arith/prime.rs) --------------------

def test_f17_add_sub_mul_inv(oracle):
    # prime.rs:923-930 (add/sub), :1226-1231 (mul), :1233-1238 (inv)
    assert oracle.add(F17, 13, 10) == 6
    assert oracle.sub(F17, 13, 10) == 3
    assert oracle.sub(F17, 10, 13) == 14
    assert oracle.mul(F17, 5, 4) == 3
    assert oracle.inv(F17, 5) == 7


def test_goldilocks_near_modulus(oracle):
    # prime.rs:932-955: add/sub/mul against 128-bit integer arithmetic
    p = GOLDILOCKS
    pairs = [(p - 1, p - 2), ((1 << 32) - 1, (1 << 32) + 7), (123456789, 987654321),
             (p - (1 << 32), 1 << 32)]
    for a, b in pairs:
        assert oracle.add(p, a, b) == (a + b) % p
        assert oracle.sub(p, a, b) == (a + p - b) % p
        assert oracle.mul(p, a, b) == (a * b) % p


def test_near_2pow32_u32_boundary(oracle):
    # prime.rs:1481-1501 (4294967291, values near 2^32)
    q = 4294967291
    assert oracle.mul(q, q - 1, 2) == (2 * (q - 1)) % q
    assert oracle.add(q, q - 1, 2) == 1


def test_montgomery_words_and_qinv(oracle):
    # R = 2^64 (the default limb, prime.rs:26): to_montgomery(a) = a * 2^64 mod q
    for q in (F17, Q32, GOLDILOCKS, 184683593729):
        qi = oracle.q_inv_neg(q)
        assert (q * qi) % (1 << 64) == (1 << 64) - 1  # q * (-q^-1) = -1 mod 2^64
        for a in (0, 1, 2, q - 1, q // 3):
            assert oracle.to_montgomery(q, a) == (a << 64) % q
            assert oracle.from_montgomery(q, oracle.to_montgomery(q, a)) == a


def test_random_mul_matches_bigint(oracle):
    rng = np.random.default_rng(3)
    for q in (Q32, GOLDILOCKS, F12289, 184683593729):
        for a, b in rand_residues(rng, q, (50, 2)):
            assert oracle.mul(q, int(a), int(b)) == int(a) * int(b) % q


# ---- roots (prime.rs:566-607, :645-687) ------------------------------------

def test_roots_match_golden(oracle, golden):
    for name in ("q32", "goldilocks"):
        g = golden[name]
        assert oracle.find_generator(g["q"]) == g["generator"]
        assert oracle.root_of_unity(g["q"], 128) == g["psi"]
        assert oracle.root_of_unity(g["q"], 64) == g["omega"]
    assert golden["q32"]["q"] == Q32
    # Goldilocks: omega = 2^39 (a power of two), psi^2 = omega
    assert golden["goldilocks"]["omega"] == 1 << 39


def test_max_ntt_size(oracle):
    # prime.rs:1300-1306 (F12289: 4096) and q32: q-1 = 2^7 * 3 * 641 * 17449
    assert oracle.max_ntt_size(F12289) == 4096
    assert oracle.max_ntt_size(Q32) == 128
    assert oracle.max_ntt_size(GOLDILOCKS) == 1 << 32


def test_ntt_error_codes(oracle):
    # poly/ntt.rs:659-694: length 3 -> LengthNotPowerOfTwo; F17 length 32 ->
    # UnsupportedSize; F17 twist of length 16 -> UnsupportedTwist
    assert oracle.xform_status("ntt_forward", F17, 3, [1, 1, 1]) == -2
    assert oracle.xform_status("ntt_forward", F17, 32, [1] * 32) == -3
    assert oracle.xform_status("twisted_forward", F17, 16, [1] * 16) == -4
    assert oracle.xform_status("ntt_forward", F17, 16, [1] * 16) == 0


# ---- transforms against closed forms ---------------------------------------

@pytest.mark.parametrize("name", ["q32", "goldilocks"])
def test_twisted_and_cyclic_ntt_golden(oracle, golden, name):
    g = golden[name]
    x = arr(g["ntt_in"])
    assert np.array_equal(oracle.twisted_forward(g["q"], 64, x), arr(g["twisted_out"]))
    assert np.array_equal(oracle.ntt_forward(g["q"], 64, x), arr(g["cyclic_out"]))
    assert np.array_equal(oracle.twisted_inverse(g["q"], 64, arr(g["twisted_out"])), x)
    assert np.array_equal(oracle.ntt_inverse(g["q"], 64, arr(g["cyclic_out"])), x)


def test_round_trips_other_primes(oracle):
    # prime.rs:1310-1352 and poly/ntt.rs:361-417 round trips
    x = arr(list(range(1, 9)))
    assert np.array_equal(oracle.ntt_inverse(F3329, 8, oracle.ntt_forward(F3329, 8, x)), x)
    assert np.array_equal(oracle.twisted_inverse(F17, 8, oracle.twisted_forward(F17, 8, x)), x)
    rng = np.random.default_rng(5)
    for q, n in ((F12289, 256), (184683593729, 256), (GOLDILOCKS, 512)):
        v = rand_residues(rng, q, n)
        assert np.array_equal(oracle.ntt_inverse(q, n, oracle.ntt_forward(q, n, v)), v)


@pytest.mark.parametrize("name", ["q32", "goldilocks"])
def test_negacyclic_golden(oracle, golden, name):
    g = golden[name]
    out = oracle.negacyclic_mul(g["q"], 64, arr(g["mul_a"]), arr(g["mul_b"]))
    assert np.array_equal(out, arr(g["mul_out"]))
    naive = oracle.negacyclic_mul(g["q"], 64, arr(g["mul_a"]), arr(g["mul_b"]), naive=True)
    assert np.array_equal(naive, arr(g["mul_out"]))


def test_negacyclic_ntt_vs_naive_random(oracle):
    # poly/ntt.rs:483-594: NTT product equals the naive product (F12289 n=256, F17 n=8)
    rng = np.random.default_rng(11)
    for q, n in ((F12289, 256), (F17, 8), (Q32, 64), (GOLDILOCKS, 64)):
        a, b = rand_residues(rng, q, (3, n)), rand_residues(rng, q, (3, n))
        assert np.array_equal(oracle.negacyclic_mul(q, n, a, b),
                              oracle.negacyclic_mul(q, n, a, b, naive=True))


def test_x_times_x_pow_dminus1_is_minus_one(oracle):
    # X * X^(d-1) = X^d = -1 in Z_q[X]/(X^d + 1): the sign a dropped twist would lose
    for q in (Q32, GOLDILOCKS):
        a = np.zeros(64, dtype=np.uint64); a[1] = 1
        b = np.zeros(64, dtype=np.uint64); b[63] = 1
        out = oracle.negacyclic_mul(q, 64, a, b)
        assert int(out[0]) == q - 1 and not out[1:].any()


# ---- commitment ------------------------------------------------------------

def test_f17_linear_commitment_kat(oracle, golden):
    # commit/src/linear.rs:1048-1083: t = [4, 10] (scalar ring: d = 1)
    k = golden["f17_linear_kat"]
    am = arr(k["a_msg"]).reshape(2, 2, 1)
    ao = arr(k["a_open"]).reshape(2, 3, 1)
    t = oracle.commit(F17, 1, am, ao, arr(k["m"]).reshape(2, 1), arr(k["r"]).reshape(3, 1))
    assert t.ravel().tolist() == [4, 10]


def test_f17_prepared_matches_coefficient_path(oracle, golden):
    # commit/src/linear.rs:1128-1178 with the schoolbook expectation
    k = golden["f17_prepared_vs_coeff"]
    args = (F17, 8, arr(k["a_msg"]), arr(k["a_open"]), arr(k["m"]), arr(k["r"]))
    assert np.array_equal(oracle.commit(*args), arr(k["t"]))
    assert np.array_equal(oracle.commit(*args, path="coeff"), arr(k["t"]))


@pytest.mark.parametrize("name", ["q32", "goldilocks"])
def test_commit_golden(oracle, golden, name):
    g = golden[name]
    c = g["commit"]
    args = (g["q"], 64, arr(c["a_msg"]), arr(c["a_open"]), arr(c["m"]), arr(c["r"]))
    assert np.array_equal(oracle.commit(*args), arr(c["t"]))
    assert np.array_equal(oracle.commit(*args, path="coeff"), arr(c["t"]))
    # NTT-domain key gives the same commitment
    amn = oracle.twisted_forward(g["q"], 64, arr(c["a_msg"]).reshape(-1, 64)).reshape(2, 3, 64)
    aon = oracle.twisted_forward(g["q"], 64, arr(c["a_open"]).reshape(-1, 64)).reshape(2, 2, 64)
    t = oracle.commit(g["q"], 64, amn, aon, arr(c["m"]), arr(c["r"]), key_domain=1)
    assert np.array_equal(t, arr(c["t"]))


def test_commit_linearity(oracle):
    rng = np.random.default_rng(21)
    q = Q32
    am = rand_residues(rng, q, (3, 5, 64))
    ao = np.zeros((3, 1, 64), dtype=np.uint64)
    m1, m2 = rand_residues(rng, q, (5, 64)), rand_residues(rng, q, (5, 64))
    z = np.zeros((1, 64), dtype=np.uint64)
    t1 = oracle.commit(q, 64, am, ao, m1, z)
    t2 = oracle.commit(q, 64, am, ao, m2, z)
    t12 = oracle.commit(q, 64, am, ao, (m1 + m2) % np.uint64(q), z)
    assert np.array_equal(t12, (t1 + t2) % np.uint64(q))


# ---- Gram ------------------------------------------------------------------

@pytest.mark.parametrize("name", ["q32", "goldilocks"])
def test_gram_golden(oracle, golden, name):
    g = golden[name]
    s = arr(g["gram"]["s"])
    assert np.array_equal(oracle.gram(g["q"], 64, s, ntt_once=False), arr(g["gram"]["g"]))
    assert np.array_equal(oracle.gram(g["q"], 64, s, ntt_once=True), arr(g["gram"]["g"]))


def test_gram_zero_witness(oracle):
    # garbage.rs:204-215
    g = oracle.gram(Q32, 64, np.zeros((3, 4, 64), dtype=np.uint64))
    assert g.shape == (6, 64) and not g.any()


# ---- ChaCha20 + JL ---------------------------------------------------------

def test_chacha20_rfc8439(oracle, golden):
    c = golden["chacha20"]
    key = bytes.fromhex(c["key"])
    blk = oracle.chacha20_block(key, c["counter_232"], bytes.fromhex(c["nonce_232"]))
    assert blk.hex() == c["block_232"]
    # RFC 8439 2.3.2 printed serialized block, first and last 16 bytes
    assert blk[:16].hex() == "10f1e7e4d13b5915500fdd1fa32071c4"
    assert blk[-16:].hex() == "b5129cd1de164eb9cbd083e8a2503c4e"
    ks = oracle.chacha20_block(key, c["counter_242"], bytes.fromhex(c["nonce_242"]))
    # RFC 8439 2.4.2: first keystream block, first 16 bytes
    assert ks[:16].hex() == "224f51f3401bd9e12fde276fb8631ded"


def test_jl_golden(oracle, golden):
    j = golden["jl"]
    pm, pe, _ = oracle.jl_project(j["q"], bytes.fromhex(j["seed"]), j["rows"], arr(j["parts"]), 64)
    assert pm.tolist() == j["p_modq"]
    assert pe.tolist() == j["p_exact"]
    s = golden["jl_small"]
    pm, pe, flat = oracle.jl_project(s["q"], bytes.fromhex(s["seed"]), s["rows"], arr(s["parts"]),
                                     64, want_flat=True)
    assert pm.tolist() == s["p_modq"] and pe.tolist() == s["p_exact"]
    assert flat.tolist() == s["flat"]


def test_jl_invariants(oracle):
    rng = np.random.default_rng(8)
    q, seed = Q32, bytes(range(32))
    a = small_signed(rng, q, (2, 256), -128, 128)
    b = small_signed(rng, q, (2, 256), -128, 128)
    pa, ea, fa = oracle.jl_project(q, seed, 256, a, 64, want_flat=True)
    pb, eb, _ = oracle.jl_project(q, seed, 256, b, 64)
    pab, eab, _ = oracle.jl_project(q, seed, 256, (a + b) % np.uint64(q), 64)
    # linearity (jl.rs:753-772) when no coefficient wraps
    assert np.array_equal(eab, ea + eb)
    assert np.array_equal(pab, (pa + pb) % np.uint64(q))
    # zero witness -> 0 (jl.rs:600-646)
    pz, ez, _ = oracle.jl_project(q, seed, 256, np.zeros((2, 256), dtype=np.uint64), 64)
    assert not pz.any() and not ez.any()
    # density: P[+1] = P[-1] = 1/4, P[0] = 1/2 (jl.rs:573-597)
    f = fa.astype(np.int64)
    n = f.size
    assert abs((f == 1).sum() / n - 0.25) < 0.01 and abs((f == -1).sum() / n - 0.25) < 0.01
    # part position matters (jl.rs:775-792): swapping parts changes p
    ps, _, _ = oracle.jl_project(q, seed, 256, a[::-1].copy(), 64)
    assert not np.array_equal(ps, pa)


# ---- AVX2 CPU-baseline restatement agrees with the scalar oracle -----------

@pytest.mark.parametrize("limb", [64, 32])
def test_avx2_commit_matches_oracle(oracle, limb):
    if not oracle.lib().gr_avx2_available():
        pytest.skip("no AVX2 on this host")
    rng = np.random.default_rng(limb)
    q = Q32
    a = rand_residues(rng, q, (4, 9, 64))
    s = rand_residues(rng, q, (9, 64))
    c = oracle.Avx2Commit(q, limb, a)
    c.load(s)
    c.run(threads=2)
    z = np.zeros((4, 1, 64), dtype=np.uint64)
    ref = oracle.commit(q, 64, a, z, s, np.zeros((1, 64), dtype=np.uint64), key_domain=1)
    assert np.array_equal(c.result(), ref)
    c.close()


@pytest.mark.parametrize("q", [Q32, GOLDILOCKS])
def test_avx2_prefix_commits_match_oracle(oracle, q):
    """Commits of witnesses shorter than the key (its first n columns), on
    the q < 2^32 path and the wide Goldilocks path (u64_arith.rs:193-357)."""
    if not oracle.lib().gr_avx2_available():
        pytest.skip("no AVX2 on this host")
    rng = np.random.default_rng(7)
    a = rand_residues(rng, q, (5, 23, 64))
    c = oracle.Avx2Commit(q, 64, a)
    z = np.zeros((5, 1, 64), dtype=np.uint64)
    zr = np.zeros((1, 64), dtype=np.uint64)
    for n in (23, 1, 7, 16):
        s = rand_residues(rng, q, (n, 64))
        c.load(s)
        c.run(threads=1 + n % 3)
        ref = oracle.commit(q, 64, a[:, :n], z, s, zr, key_domain=1)
        assert np.array_equal(c.result(), ref), n
    c.close()


def test_avx2_wide_words_at_the_extremes(oracle):
    """Goldilocks words next to 0, 2^32 and p, where the 64 x 64 products'
    carries (mul_epu64_wide_avx2) and the wide reduction's overflow branch
    (montgomery_reduce_wide) are taken."""
    if not oracle.lib().gr_avx2_available():
        pytest.skip("no AVX2 on this host")
    q = GOLDILOCKS
    edge = np.array([0, 1, 2, (1 << 32) - 1, 1 << 32, (1 << 32) + 1, (1 << 63) - 1, 1 << 63,
                     q - 2, q - 1, q // 2, q // 2 + 1, 0xFFFFFFFF00000000, 0xFFFFFFFE00000001,
                     (1 << 64) - (1 << 33), 12345678901234567890 % q], dtype=np.uint64)
    rng = np.random.default_rng(8)
    a = edge[rng.integers(0, len(edge), size=(3, 4, 64))]
    s = edge[rng.integers(0, len(edge), size=(4, 64))]
    c = oracle.Avx2Commit(q, 64, a)
    c.load(s)
    c.run()
    z = np.zeros((3, 1, 64), dtype=np.uint64)
    ref = oracle.commit(q, 64, a, z, s, np.zeros((1, 64), dtype=np.uint64), key_domain=1)
    assert np.array_equal(c.result(), ref)
    c.close()


@pytest.mark.parametrize("q,limb", [(Q32, 64), (Q32, 32), (GOLDILOCKS, 64), (12289, 64)])
def test_avx2_negacyclic_matches_schoolbook(oracle, q, limb):
    """negacyclic_mul_coeffs through the AVX2 twisted NTT against the
    definition (and the scalar oracle on more pairs)."""
    if not oracle.lib().gr_avx2_available():
        pytest.skip("no AVX2 on this host")
    rng = np.random.default_rng(9)
    a = rand_residues(rng, q, (40, 64))
    b = rand_residues(rng, q, (40, 64))
    m = oracle.Avx2Negacyclic(q, limb)
    m.load(a, b)
    m.run(threads=3)
    got = m.result()
    m.close()
    for i in range(3):
        want = _schoolbook(q, [int(x) for x in a[i]], [int(x) for x in b[i]])
        assert [int(x) for x in got[i]] == want, i
    assert np.array_equal(got, oracle.negacyclic_mul(q, 64, a, b))


@pytest.mark.parametrize("q", [Q32, GOLDILOCKS])
def test_avx2_independent_items_match_oracle(oracle, q):
    """Independent prefix commitments drawn from a pool, several threads."""
    if not oracle.lib().gr_avx2_available():
        pytest.skip("no AVX2 on this host")
    rng = np.random.default_rng(11)
    a = rand_residues(rng, q, (4, 30, 64))
    pool = small_signed(rng, q, (50, 64), -1, 2)
    lens = [30, 1, 17, 8, 29, 2]
    offs = [0, 49, 10, 40, 21, 3]
    c = oracle.Avx2Commit(q, 64, a)
    c.load_items(pool, offs, lens)
    c.run_items(threads=3)
    got = c.items_result()
    c.close()
    z = np.zeros((4, 1, 64), dtype=np.uint64)
    zr = np.zeros((1, 64), dtype=np.uint64)
    for i, (o, n) in enumerate(zip(offs, lens)):
        exp = oracle.commit(q, 64, a[:, :n], z, pool[o:o + n], zr, key_domain=1)
        assert np.array_equal(got[i], exp), i


@pytest.mark.parametrize("q", [Q32, GOLDILOCKS])
def test_avx2_coefficient_domain_gram_and_inner_commit(oracle, q):
    """The reference-faithful coefficient-domain Gram matrix and inner
    commitments (one NTT product per term) through the AVX2 restatement equal
    the scalar oracle's Gram and commitment."""
    if not oracle.lib().gr_avx2_available():
        pytest.skip("no AVX2 on this host")
    rng = np.random.default_rng(10)
    r, n, rows = 4, 6, 3
    s = small_signed(rng, q, (r, n, 64), -128, 128)
    pairs = [(i, j) for i in range(r) for j in range(i, r)]
    m = oracle.Avx2Dots(q, 64)
    m.load(s, s, [i for i, _ in pairs], [j for _, j in pairs])
    m.run(threads=2)
    assert np.array_equal(m.result(), oracle.gram(q, 64, s))
    a = rand_residues(rng, q, (rows, n, 64))
    m.load(a, s, [row for part in range(r) for row in range(rows)],
           [part for part in range(r) for row in range(rows)])
    m.run()
    got = m.result().reshape(r, rows, 64)
    m.close()
    z = np.zeros((rows, 1, 64), dtype=np.uint64)
    for part in range(r):
        exp = oracle.commit(q, 64, a, z, s[part], np.zeros((1, 64), dtype=np.uint64))
        assert np.array_equal(got[part], exp), part


# ---- LaBRADOR step around the commitment ---------------------------------------

def _schoolbook(q, a, b):
    """negacyclic product mod (X^len + 1), by definition."""
    n = len(a)
    out = [0] * n
    for i in range(n):
        for j in range(n):
            k = i + j
            if k < n:
                out[k] = (out[k] + a[i] * b[j]) % q
            else:
                out[k - n] = (out[k - n] - a[i] * b[j]) % q
    return out


def test_decompose_balanced_reference_cases(oracle):
    # decompose_coeff_balanced round trips (main_protocol/mod.rs:1156-1188):
    # q = 12289, base 8, 8 limbs, all v < 256 and the near-q values
    q, base, limbs = 12289, 8, 8
    vals = list(range(256)) + [q - 10, q - 1, q // 2, q // 2 + 1, 0, 1]
    vals += [0] * (-len(vals) % 4)
    polys = np.array(vals, dtype=np.uint64).reshape(-1, 4)
    out = oracle.decompose_balanced(q, 4, polys, base, limbs)
    for p in range(polys.shape[0]):
        for k in range(4):
            centred = [int(x) if int(x) <= q // 2 else int(x) - q for x in out[p, :, k]]
            assert all(abs(c) <= base // 2 for c in centred)
            assert sum(c * base ** l for l, c in enumerate(centred)) % q == int(polys[p, k])
    # the poly case of mod.rs:1190-1236: c_i = (193 i + 7) mod 12289, 8 limbs
    poly = np.array([[(193 * i + 7) % q for i in range(64)]], dtype=np.uint64)
    out = oracle.decompose_balanced(q, 64, poly, base, limbs)
    for i in range(64):
        centred = [int(x) if int(x) <= q // 2 else int(x) - q for x in out[0, :, i]]
        assert sum(c * base ** l for l, c in enumerate(centred)) % q == int(poly[0, i])


def test_decompose_balanced_boundaries(oracle):
    # centered_mod (mod.rs:277-289) keeps +b/2 for a non-negative carry and -b/2
    # for a negative one; truncating division carries the sign
    q = 4294966657
    cases = {100: [4, 6, 0], 15: [-1, 1, 0], q - 1: [-1, 0, 0], 8: [8, 0, 0], q - 8: [-8, 0, 0],
             24: [8, 1, 0], q - 24: [-8, -1, 0], 0: [0, 0, 0]}
    vals = list(cases)
    out = oracle.decompose_balanced(q, 8, np.array([vals], dtype=np.uint64), 16, 3)
    for k, v in enumerate(vals):
        got = [int(x) if int(x) <= q // 2 else int(x) - q for x in out[0, :, k]]
        assert got == cases[v], (v, got)


def test_garbage_h_pins(oracle):
    # zero phi -> zero (garbage.rs:218-231); phi = s -> h_ij = <s_i, s_j> (the
    # Gram matrix from the NTT-once path); schoolbook brute force over F17, len 4
    rng = np.random.default_rng(3)
    q = 4294966657
    s = rng.integers(0, q, size=(3, 5, 64), dtype=np.uint64)
    assert not oracle.garbage_h(q, 64, np.zeros_like(s), s).any()
    assert np.array_equal(oracle.garbage_h(q, 64, s, s), oracle.gram(q, 64, s, ntt_once=True))
    q, d, r, n = 17, 4, 2, 3
    phi = rng.integers(0, q, size=(r, n, d), dtype=np.uint64)
    s = rng.integers(0, q, size=(r, n, d), dtype=np.uint64)
    inv2 = (q + 1) // 2
    exp = []
    for i in range(r):
        for j in range(i, r):
            acc = [0] * d
            for k in range(n):
                t1 = _schoolbook(q, [int(x) for x in phi[i, k]], [int(x) for x in s[j, k]])
                t2 = _schoolbook(q, [int(x) for x in phi[j, k]], [int(x) for x in s[i, k]])
                acc = [(a + b + c) % q for a, b, c in zip(acc, t1, t2)]
            exp.append([a * inv2 % q for a in acc])
    assert oracle.garbage_h(q, d, phi, s).tolist() == exp


def test_amortize_pins(oracle):
    # zero witness -> zero (amortization.rs:172-196); c_i = 1 -> z = sum_i s_i;
    # schoolbook brute force over F17, len 4
    rng = np.random.default_rng(4)
    q = 4294966657
    s = rng.integers(0, q, size=(4, 6, 64), dtype=np.uint64)
    c = rng.integers(0, q, size=(4, 64), dtype=np.uint64)
    assert not oracle.amortize(q, 64, c, np.zeros_like(s)).any()
    one = np.zeros((4, 64), dtype=np.uint64)
    one[:, 0] = 1
    exp = [[sum(int(s[i, k, e]) for i in range(4)) % q for e in range(64)] for k in range(6)]
    assert oracle.amortize(q, 64, one, s).tolist() == exp
    q, d, r, n = 17, 4, 3, 2
    s = rng.integers(0, q, size=(r, n, d), dtype=np.uint64)
    c = rng.integers(0, q, size=(r, d), dtype=np.uint64)
    exp = []
    for k in range(n):
        acc = [0] * d
        for i in range(r):
            t = _schoolbook(q, [int(x) for x in c[i]], [int(x) for x in s[i, k]])
            acc = [(a + b) % q for a, b in zip(acc, t)]
        exp.append(acc)
    assert oracle.amortize(q, d, c, s).tolist() == exp


# ---- Ring slice hooks (prime.rs:433-491) ------------------------------------

# The inputs of the reference's slice-hook tests; each asserts hook == the
# scalar operation, whose value is the closed form below.
SLICE_KATS = [
    # prime.rs:945-972 (add/sub) and :1163-1190 (mul hooks, scalar 19)
    (F12289, [1, 7, 12_000, 3, 11], [2, 9, 1_111, 14, 6], 19),
    # prime.rs:1055-1095: every hook, scalar 19
    (F3329, [1, 7, 3_000, 3, 11], [2, 9, 1_111, 14, 6], 19),
    # prime.rs:1097-1124 (add/sub above 2^32)
    (4294967311, [1, 7, 4_000_000_000, 3, 11], [2, 9, 111_111_111, 14, 6], 5),
    # prime.rs:1126-1160 (Goldilocks, scalar 2^32 + 13)
    (GOLDILOCKS, [1, (1 << 32) - 1, GOLDILOCKS - 2, 123_456_789, 9_876_543_210],
     [2, (1 << 32) + 7, GOLDILOCKS - 3, 987_654_321, 77_777_777], (1 << 32) + 13),
    # prime.rs:1192-1224 (q = 8380417, scalar 19)
    (8380417, [1, 7, 8_000_000, 3, 11], [2, 9, 111_111, 14, 6], 19),
]


def _closed(q, op, d, s, c):
    if op == "add":
        return [(x + y) % q for x, y in zip(d, s)]
    if op == "sub":
        return [(x - y) % q for x, y in zip(d, s)]
    if op == "scalar_mul":
        return [(x * c) % q for x in d]
    return [(x + y * c) % q for x, y in zip(d, s)]


@pytest.mark.parametrize("q,lhs,rhs,scalar", SLICE_KATS)
def test_slice_hooks_reference_inputs(oracle, q, lhs, rhs, scalar):
    for op in ("add", "sub", "scalar_mul", "add_scaled"):
        got = oracle.slice_op(q, op, lhs, rhs, scalar)
        assert [int(v) for v in got] == _closed(q, op, lhs, rhs, scalar), op
    assert oracle.dot_product(q, lhs, rhs) == sum(x * y for x, y in zip(lhs, rhs)) % q


@pytest.mark.parametrize("q", [F17, F12289, Q32, GOLDILOCKS])
@pytest.mark.parametrize("n", [0, 1, 63, 64, 65, 130])
def test_slice_hooks_random_against_integers(oracle, q, n):
    # lengths around the 64-word chunk of add_assign_scaled_slice / dot_product
    rng = np.random.default_rng(n + q % 1000)
    d = rand_residues(rng, q, (n,))
    s = rand_residues(rng, q, (n,))
    dl, sl = [int(x) for x in d], [int(x) for x in s]
    for c in (0, 1, q - 1, int(rand_residues(rng, q, (1,))[0])):
        for op in ("add", "sub", "scalar_mul", "add_scaled"):
            got = oracle.slice_op(q, op, d, s, c)
            assert [int(v) for v in got] == _closed(q, op, dl, sl, c), (op, c)
    assert oracle.dot_product(q, d, s) == sum(x * y for x, y in zip(dl, sl)) % q


def test_avx2_twisted_matches_oracle():
    # the CPU baseline of the transform leg: the AVX2 restatement of
    # TwistedNttPlan equals the scalar oracle, both limb widths and fields
    import numpy as np
    from oracle import gridref
    rng = np.random.default_rng(77)
    for q, limbs in ((4294966657, (64, 32)), (0xffffffff00000001, (64,))):
        x = rng.integers(0, 1 << 63, size=(40, 64), dtype=np.uint64) % np.uint64(q)
        for limb in limbs:
            m = gridref.Avx2Twisted(q, limb)
            m.load(x)
            m.run(False)
            assert np.array_equal(m.result(), gridref.twisted_forward(q, 64, x))
            m.run(True, 2)
            assert np.array_equal(m.result(), x)
            m.close()
