"""Pins the wire-format restatement (oracle/wire.py) before the library is
compared against it.

Expected bytes are assembled here with struct from the layout rules of the
reference's serialize crate, independently of the oracle's numpy packing; the
reference's own serialization tests give the error cases.
"""
import struct

import numpy as np
import pytest

from conftest import GOLDILOCKS, Q32, rand_residues

from oracle import wire

F17 = 17


def kind(fn, *args, **kw):
    with pytest.raises(wire.WireError) as e:
        fn(*args, **kw)
    return e.value.kind


def test_element_widths():
    # prime.rs:54-58: ceil(bits(q - 1) / 8); prime.rs:1380 (F17 -> 1 byte)
    for q, eb in ((F17, 1), (251, 1), (257, 2), (12289, 2), (65521, 2), (8380417, 3),
                  (Q32, 4), (4294967291, 4), (GOLDILOCKS, 8)):
        assert wire.elem_bytes(q) == eb, q


def test_f17_round_trip_and_out_of_range():
    # prime.rs:1377-1394: every value round-trips in one byte; 17u64 LE is rejected
    for v in range(17):
        b = wire.encode_elems(F17, [v])
        assert b == bytes([v])
        vals, used = wire.decode_poly(F17, 1, b)
        assert (int(vals[0]), used) == (v, 1)
    assert kind(wire.decode_poly, F17, 1, (17).to_bytes(8, "little")) == "InvalidData"


def test_ringvec_and_ringmat_layouts():
    # lattice/types.rs:578-598 (round trips of RingVec [1,2,3] and RingMat 2x2)
    v = [[1], [2], [3]]
    b = wire.encode_ringvec(F17, v)
    assert b == struct.pack("<Q", 3) + bytes([1, 2, 3])
    polys, used = wire.decode_ringvec(F17, 1, b)
    assert used == len(b) and polys.reshape(-1).tolist() == [1, 2, 3]
    m = [[1], [2], [3], [4]]
    b = wire.encode_ringmat(F17, 2, 2, m)
    assert b == struct.pack("<QQ", 2, 2) + bytes([1, 2, 3, 4])
    rows, cols, polys, used = wire.decode_ringmat(F17, 1, b)
    assert (rows, cols, used) == (2, 2, len(b)) and polys.reshape(-1).tolist() == [1, 2, 3, 4]
    # Vec<poly>: u32 prefix (serialize/src/lib.rs:210-225)
    assert wire.encode_vec(F17, v) == struct.pack("<I", 3) + bytes([1, 2, 3])


def test_ringmat_overflow_shape_is_invalid_data():
    # lattice/types.rs:600-607
    b = struct.pack("<QQ", 2 ** 64 - 1, 2)
    assert kind(wire.decode_ringmat, F17, 1, b) == "InvalidData"


def test_truncated_coefficient_poly_is_unexpected_end():
    # poly/ring.rs:661-676: Poly8 over F17, one byte short
    b = wire.encode_polys(F17, list(range(8)))
    assert b == bytes(range(8))
    assert kind(wire.decode_poly, F17, 8, b[:-1]) == "UnexpectedEnd"


def test_corrupt_and_truncated_twisted_poly():
    # poly/twisted_ntt.rs:915-946: F8380417 N=256 with byte 2 forced to 0xFF,
    # and an F17 N=8 evaluation vector one element short
    q = 8380417
    rng = np.random.default_rng(5)
    vals = rand_residues(rng, q, (256,))
    b = bytearray(wire.encode_polys(q, vals))
    assert len(b) == 3 * 256
    b[2] = 0xFF
    assert kind(wire.decode_poly, q, 256, bytes(b), ntt=True) == "InvalidData"
    b = wire.encode_polys(F17, [3, 1, 4, 1, 5, 9, 2, 6])
    assert kind(wire.decode_poly, F17, 8, b[:-1], ntt=True) == "UnexpectedEnd"


def test_linear_key_row_mismatch_is_invalid():
    # commit/src/linear.rs:1100-1124: a_msg 1 x 2, a_open 2 x 3 over F17
    a_msg = np.array([[[1], [2]]], dtype=np.uint64)
    a_open = np.array([[[3], [4], [5]], [[6], [7], [8]]], dtype=np.uint64)
    b = wire.encode_linear_key(F17, a_msg, a_open)
    assert b == (struct.pack("<QQ", 1, 2) + bytes([1, 2]) + struct.pack("<QQ", 2, 3) +
                 bytes([3, 4, 5, 6, 7, 8]))
    assert kind(wire.decode_linear_key, F17, 1, b) == "InvalidData"
    m, o, used = wire.decode_linear_key(F17, 1, b, validate=False)
    assert used == len(b) and m.shape == (1, 2, 1) and o.shape == (2, 3, 1)


def test_wide_elements_are_little_endian():
    for q in (Q32, GOLDILOCKS, 12289, 8380417):
        eb = wire.elem_bytes(q)
        vals = [0, 1, q - 1, (q - 1) // 3]
        assert wire.encode_elems(q, vals) == b"".join(v.to_bytes(eb, "little") for v in vals)
    assert kind(wire.decode_poly, Q32, 1, Q32.to_bytes(4, "little")) == "InvalidData"
    assert kind(wire.decode_poly, GOLDILOCKS, 1, GOLDILOCKS.to_bytes(8, "little")) == "InvalidData"


def test_exact_consumption():
    b = wire.encode_ringvec(F17, [[1], [2]])
    assert wire.exact(wire.decode_ringvec(F17, 1, b), b).reshape(-1).tolist() == [1, 2]
    assert kind(wire.exact, wire.decode_ringvec(F17, 1, b + b"\0"), b + b"\0") == "InvalidData"


@pytest.mark.parametrize("ntt", [False, True])
def test_fast_decoder_matches_element_loop(ntt):
    """_fast_polys (one pass over a payload) against _polys (the reference's
    loop per polynomial) on every truncation and one corruption at a time."""
    q, d, count = 12289, 4, 3
    rng = np.random.default_rng(11)
    base = wire.encode_polys(q, rand_residues(rng, q, (count, d)))

    def outcome(fn, data):
        try:
            p, used = fn(q, d, data, 0, count, ntt)
            return ("ok", p.tolist(), used)
        except wire.WireError as e:
            return (e.kind,)

    cases = []
    for cut in range(len(base) + 1):
        cases.append(base[:cut])
        for pos in range(0, cut, 2):  # high byte of an element: 0xFF makes it >= q
            b = bytearray(base[:cut])
            if pos + 1 < cut:
                b[pos + 1] = 0xFF
                cases.append(bytes(b))
    for data in cases:
        assert outcome(wire._fast_polys, data) == outcome(wire._polys, data)


def test_truncation_precedence_by_domain():
    """A bad element inside the polynomial that is cut short: the coefficient
    domain reports it (element by element), the NTT domain reports the end."""
    q, d = 12289, 4
    good = wire.encode_polys(q, [[1, 2, 3, 4], [5, 6, 7, 8]])
    bad = bytearray(good)
    bad[2 * 4 + 1] = 0xFF          # second polynomial, first element
    cut = bytes(bad[:-2])           # second polynomial incomplete
    assert kind(wire.decode_ringvec, q, d, struct.pack("<Q", 2) + cut, 0, False) == "InvalidData"
    assert kind(wire.decode_ringvec, q, d, struct.pack("<Q", 2) + cut, 0, True) == "UnexpectedEnd"
    bad0 = bytearray(good)
    bad0[1] = 0xFF                  # first (complete) polynomial
    cut0 = bytes(bad0[:-2])
    for ntt in (False, True):
        assert kind(wire.decode_ringvec, q, d, struct.pack("<Q", 2) + cut0, 0, ntt) == "InvalidData"
