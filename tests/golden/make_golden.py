"""Writes tests/golden/golden.json from first principles, in pure Python.

Nothing here imports the oracle or the GPU library: every expected value comes
from a definition that a plausible implementation mistake cannot satisfy —
  * the modulus, generator and roots from a primality test, a factorisation of
    q-1 and the reference's smallest-generator rule
    (algebra/src/arith/prime.rs:645-687; Goldilocks pinned to 7);
  * twisted NTT outputs from the closed form X_k = sum_i a_i psi^(i(2k+1)),
    cyclic ones from X_k = sum_i a_i omega^(ik) (slot order of
    algebra/src/poly/ntt.rs:75-87 and arith/ntt.rs:292-312);
  * products from the schoolbook negacyclic rule (poly/ring.rs:60-77);
  * commitments and Gram entries from schoolbook sums over R_q;
  * ChaCha20 from RFC 8439 section 2.3, JL entries from the 2-bit decode rule
    of labrador/src/jl.rs:7-10, 388-508.
The reference's own known answers are included verbatim where it has them:
the F17 linear commitment (commit/src/linear.rs:1048-1083) and the closed-form
inputs of the prepared-vs-coefficient test (linear.rs:1128-1178).

Inputs come from SplitMix64 with fixed seeds. Run: python tests/golden/make_golden.py
"""
from __future__ import annotations

import json
import os

HERE = os.path.dirname(os.path.abspath(__file__))
GOLDILOCKS = 0xFFFFFFFF00000001


class SplitMix64:
    def __init__(self, seed: int):
        self.s = seed & 0xFFFFFFFFFFFFFFFF

    def next(self) -> int:
        self.s = (self.s + 0x9E3779B97F4A7C15) & 0xFFFFFFFFFFFFFFFF
        z = self.s
        z = ((z ^ (z >> 30)) * 0xBF58476D1CE4E5B9) & 0xFFFFFFFFFFFFFFFF
        z = ((z ^ (z >> 27)) * 0x94D049BB133111EB) & 0xFFFFFFFFFFFFFFFF
        return z ^ (z >> 31)

    def below(self, q: int) -> int:
        return self.next() % q  # bias is irrelevant for test inputs


def is_prime(n: int) -> bool:
    if n < 2:
        return False
    for p in (2, 3, 5, 7, 11, 13, 17, 19, 23, 29, 31, 37):
        if n % p == 0:
            return n == p
    d, s = n - 1, 0
    while d % 2 == 0:
        d //= 2
        s += 1
    for a in (2, 3, 5, 7, 11, 13, 17, 19, 23, 29, 31, 37):
        x = pow(a, d, n)
        if x in (1, n - 1):
            continue
        for _ in range(s - 1):
            x = x * x % n
            if x == n - 1:
                break
        else:
            return False
    return True


def prime_factors(n: int):
    out, d = [], 2
    while d * d <= n:
        if n % d == 0:
            out.append(d)
            while n % d == 0:
                n //= d
        d += 1
    if n > 1:
        out.append(n)
    return out


def generator(q: int) -> int:
    if q == GOLDILOCKS:
        return 7
    fs = prime_factors(q - 1)
    g = 2
    while any(pow(g, (q - 1) // f, q) == 1 for f in fs):
        g += 1
    return g


def root(q: int, n: int) -> int:
    assert (q - 1) % n == 0
    return pow(generator(q), (q - 1) // n, q)


def twisted_ntt(q, d, a):
    psi = root(q, 2 * d)
    return [sum(a[i] * pow(psi, i * (2 * k + 1), q) for i in range(d)) % q for k in range(d)]


def cyclic_ntt(q, d, a):
    w = root(q, d)
    return [sum(a[i] * pow(w, i * k, q) for i in range(d)) % q for k in range(d)]


def negacyclic(q, d, a, b):
    out = [0] * d
    for i in range(d):
        for j in range(d):
            if i + j < d:
                out[i + j] += a[i] * b[j]
            else:
                out[i + j - d] -= a[i] * b[j]
    return [x % q for x in out]


def poly_add(q, a, b):
    return [(x + y) % q for x, y in zip(a, b)]


def commit(q, d, a_msg, a_open, m, r):
    rows = len(a_msg)
    t = []
    for row in range(rows):
        acc = [0] * d
        for j, p in enumerate(a_msg[row]):
            acc = poly_add(q, acc, negacyclic(q, d, p, m[j]))
        for j, p in enumerate(a_open[row]):
            acc = poly_add(q, acc, negacyclic(q, d, p, r[j]))
        t.append(acc)
    return t


def gram(q, d, s):
    out = []
    r = len(s)
    for i in range(r):
        for j in range(i, r):
            acc = [0] * d
            for x, y in zip(s[i], s[j]):
                acc = poly_add(q, acc, negacyclic(q, d, x, y))
            out.append(acc)
    return out


# ---- ChaCha20, RFC 8439 section 2.3 -----------------------------------------

def _rotl(x, n):
    return ((x << n) | (x >> (32 - n))) & 0xFFFFFFFF


def chacha20_block(key: bytes, counter: int, nonce: bytes) -> bytes:
    c = [0x61707865, 0x3320646E, 0x79622D32, 0x6B206574]
    k = [int.from_bytes(key[4 * i:4 * i + 4], "little") for i in range(8)]
    n = [int.from_bytes(nonce[4 * i:4 * i + 4], "little") for i in range(3)]
    s = c + k + [counter & 0xFFFFFFFF] + n
    x = list(s)

    def qr(a, b, cc, dd):
        x[a] = (x[a] + x[b]) & 0xFFFFFFFF; x[dd] = _rotl(x[dd] ^ x[a], 16)
        x[cc] = (x[cc] + x[dd]) & 0xFFFFFFFF; x[b] = _rotl(x[b] ^ x[cc], 12)
        x[a] = (x[a] + x[b]) & 0xFFFFFFFF; x[dd] = _rotl(x[dd] ^ x[a], 8)
        x[cc] = (x[cc] + x[dd]) & 0xFFFFFFFF; x[b] = _rotl(x[b] ^ x[cc], 7)

    for _ in range(10):
        qr(0, 4, 8, 12); qr(1, 5, 9, 13); qr(2, 6, 10, 14); qr(3, 7, 11, 15)
        qr(0, 5, 10, 15); qr(1, 6, 11, 12); qr(2, 7, 8, 13); qr(3, 4, 9, 14)
    return b"".join(((x[i] + s[i]) & 0xFFFFFFFF).to_bytes(4, "little") for i in range(16))


def keystream(key: bytes, nonce: bytes, nbytes: int) -> bytes:
    out = bytearray()
    ctr = 0
    while len(out) < nbytes:
        out += chacha20_block(key, ctr, nonce)
        ctr += 1
    return bytes(out[:nbytes])


def jl(q, d, seed: bytes, rows: int, parts):
    """p mod q, exact centred sum and the flat int8 layout [m][part][coeff][poly]."""
    num_parts, cols = len(parts), len(parts[0])
    n = cols // d
    pk = (rows + 3) // 4
    p_exact = [0] * rows
    flat = [0] * (rows * num_parts * cols)
    for part_idx, part in enumerate(parts):
        nonce = part_idx.to_bytes(4, "little") + bytes(8)
        ks = keystream(seed, nonce, cols * pk)
        for col in range(cols):
            v = part[col] % q
            centred = v - q if v > q // 2 else v
            for m in range(rows):
                code = (ks[col * pk + m // 4] >> (2 * (m % 4))) & 3
                e = 1 if code == 0 else (-1 if code == 1 else 0)
                p_exact[m] += e * centred
                flat[m * num_parts * cols + (part_idx * d + col % d) * n + col // d] = e
    return [x % q for x in p_exact], p_exact, flat


def rand_polys(rng, q, count, d):
    return [[rng.below(q) for _ in range(d)] for _ in range(count)]


def main():
    # the 32-bit modulus: largest prime below 2^32 with q = 1 (mod 128)
    q32 = (1 << 32) - 1
    while not (q32 % 128 == 1 and is_prime(q32)):
        q32 -= 1
    gold = {}
    rings = {"q32": q32, "goldilocks": GOLDILOCKS}
    for name, q in rings.items():
        d = 64
        rng = SplitMix64(0x5EED0000 + d + (q & 0xFFFF))
        g = generator(q)
        psi, omega = root(q, 2 * d), root(q, d)
        polys = rand_polys(rng, q, 3, d)
        edge = [[q - 1] * d, [0] * d]
        edge[1][0] = 1
        inputs = polys + edge
        a = rand_polys(rng, q, 2, d)
        b = rand_polys(rng, q, 2, d)
        a.append([q - 1] * d)
        b.append([q - 1] * d)
        gold[name] = {
            "q": q, "d": d, "generator": g, "psi": psi, "omega": omega,
            "d_inv": pow(d, q - 2, q),
            "ntt_in": inputs,
            "twisted_out": [twisted_ntt(q, d, p) for p in inputs],
            "cyclic_out": [cyclic_ntt(q, d, p) for p in inputs],
            "mul_a": a, "mul_b": b,
            "mul_out": [negacyclic(q, d, x, y) for x, y in zip(a, b)],
        }
        # small commitment: 2 rows, 3 message + 2 opening columns
        am = [rand_polys(rng, q, 3, d) for _ in range(2)]
        ao = [rand_polys(rng, q, 2, d) for _ in range(2)]
        m = rand_polys(rng, q, 3, d)
        r = rand_polys(rng, q, 2, d)
        gold[name]["commit"] = {"a_msg": am, "a_open": ao, "m": m, "r": r,
                                "t": commit(q, d, am, ao, m, r)}
        s = [rand_polys(rng, q, 2, d) for _ in range(3)]
        gold[name]["gram"] = {"s": s, "g": gram(q, d, s)}

    # JL at q32: 256 rows, 2 parts x 128 columns (n = 2 polys of d = 64),
    # coefficients uniform in [-2^7, 2^7) as residues
    rng = SplitMix64(0x31)
    seed = bytes((7 * i + 3) & 0xFF for i in range(32))
    parts = [[(rng.below(256) - 128) % q32 for _ in range(128)] for _ in range(2)]
    p_modq, p_exact, flat = jl(q32, 64, seed, 256, parts)
    gold["jl"] = {"q": q32, "d": 64, "rows": 256, "seed": seed.hex(), "parts": parts,
                  "p_modq": p_modq, "p_exact": p_exact,
                  "flat_sha_src": None}
    # a smaller flat fixture (rows = 10: straddles keystream blocks, pk = 3)
    parts_s = [[rng.below(q32) for _ in range(64)] for _ in range(3)]
    pm_s, pe_s, flat_s = jl(q32, 64, seed, 10, parts_s)
    gold["jl_small"] = {"q": q32, "d": 64, "rows": 10, "seed": seed.hex(), "parts": parts_s,
                        "p_modq": pm_s, "p_exact": pe_s, "flat": flat_s}
    del gold["jl"]["flat_sha_src"]

    # RFC 8439 test vectors: 2.3.2 (block function) and 2.4.2 (keystream)
    key = bytes(range(32))
    gold["chacha20"] = {
        "key": key.hex(),
        "nonce_232": bytes([0, 0, 0, 9, 0, 0, 0, 0x4A, 0, 0, 0, 0]).hex(),
        "counter_232": 1,
        "block_232": chacha20_block(key, 1, bytes([0, 0, 0, 9, 0, 0, 0, 0x4A, 0, 0, 0, 0])).hex(),
        "nonce_242": bytes([0, 0, 0, 0, 0, 0, 0, 0x4A, 0, 0, 0, 0]).hex(),
        "counter_242": 1,
        "keystream_242_first_block": chacha20_block(
            key, 1, bytes([0, 0, 0, 0, 0, 0, 0, 0x4A, 0, 0, 0, 0])).hex(),
    }

    # The reference's F17 known answers.
    gold["f17_linear_kat"] = {  # commit/src/linear.rs:1048-1083, scalar ring (d = 1)
        "q": 17, "a_msg": [[1, 2], [3, 4]], "a_open": [[1, 0, 1], [0, 1, 1]],
        "m": [5, 6], "r": [1, 2, 3], "t": [4, 10],
    }
    q, d = 17, 8
    one = [1] + [0] * 7
    zero = [0] * 8

    def f(k):
        return [(k * i + 1) % q for i in range(d)]

    am = [[one, zero], [f(1), f(2)]]
    ao = [[zero, one], [f(3), f(4)]]
    m = [one, f(5)]
    r = [zero, f(6)]
    gold["f17_prepared_vs_coeff"] = {  # commit/src/linear.rs:1128-1178 inputs
        "q": q, "d": d, "a_msg": am, "a_open": ao, "m": m, "r": r,
        "t": commit(q, d, am, ao, m, r),
    }
    with open(os.path.join(HERE, "golden.json"), "w") as fh:
        json.dump(gold, fh, separators=(",", ":"))
    print("wrote", os.path.join(HERE, "golden.json"))


if __name__ == "__main__":
    main()
