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
