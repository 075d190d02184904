"""Restatement of the reference's canonical encodings for the ring objects.

TEST INFRASTRUCTURE ONLY. The oracle is the CPU restatement of the reference
path that the HIP library is checked against; only tests/, the smoke check and
bench.py's cpu_baseline leg load it.

Encodings (serialize/src/lib.rs:1-7): fixed-width little-endian integers,
usize as u64, Vec<T> with a u32 length prefix. Decoders follow the
reference's loops so that their errors come in the same order:
  * PrimeField (algebra/src/arith/prime.rs:54-58, :690-721): ceil(bits(q-1)/8)
    bytes; fewer bytes -> UnexpectedEnd; a value >= q -> InvalidData;
  * CyclotomicPolyRing (poly/ring.rs:469-505): N elements decoded one by one;
  * TwistedNttPoly (poly/twisted_ntt.rs:647-685): the length of all N
    elements is checked before any element;
  * RingVec (lattice/types.rs:183-219): u64 count, then the entries;
  * RingMat (lattice/types.rs:440-490): u64 rows, u64 cols, rows*cols checked
    for overflow (InvalidData), then the entries row-major;
  * Vec<T> (serialize/src/lib.rs:210-245): u32 count, then the entries;
  * LinearCommitmentKey (commit/src/linear.rs:169-192): RingMat a_msg then
    RingMat a_open; valid when both have the same row count;
  * deserialize_exact (serialize/src/lib.rs:79-83): trailing bytes -> InvalidData.
A polynomial is a list of d elements; d = 1 gives the reference's
RingVec<F>/RingMat<F> of bare field elements.

Storage reservation (try_reserve_exact) is not modelled beyond the shape
overflow the reference checks explicitly; counts whose byte size needs more
than 63 bits are reported as InvalidData, as the library does.
"""
from __future__ import annotations

import struct

import numpy as np


class WireError(Exception):
    """SerializationError (serialize/src/lib.rs:28-34): kind is 'UnexpectedEnd'
    or 'InvalidData'."""

    def __init__(self, kind: str, msg: str = ""):
        super().__init__(f"{kind}: {msg}" if msg else kind)
        self.kind = kind


def elem_bytes(q: int) -> int:
    """PrimeField::SERIALIZED_BYTES (prime.rs:54-58)."""
    return ((q - 1).bit_length() + 7) // 8


# ---- encoders -----------------------------------------------------------------

def encode_elems(q: int, values) -> bytes:
    """Field elements back to back (prime.rs:696-702): the low bytes of the LE u64."""
    v = np.asarray(values, dtype=np.uint64).reshape(-1)
    if v.size and int(v.max()) >= q:
        raise ValueError("non-canonical element")
    eb = elem_bytes(q)
    raw = v.astype("<u8").view(np.uint8).reshape(-1, 8)[:, :eb]
    return raw.tobytes()


def encode_polys(q: int, polys) -> bytes:
    """Polynomials back to back ([T; N] / CyclotomicPolyRing / TwistedNttPoly)."""
    return encode_elems(q, polys)


def encode_ringvec(q: int, polys) -> bytes:
    p = np.asarray(polys, dtype=np.uint64)
    count = p.shape[0] if p.ndim else 0
    return struct.pack("<Q", count) + encode_elems(q, p)


def encode_ringmat(q: int, rows: int, cols: int, polys) -> bytes:
    p = np.asarray(polys, dtype=np.uint64)
    assert p.size == 0 or p.shape[0] == rows * cols
    return struct.pack("<QQ", rows, cols) + encode_elems(q, p)


def encode_vec(q: int, polys) -> bytes:
    p = np.asarray(polys, dtype=np.uint64)
    count = p.shape[0] if p.ndim else 0
    if count > 0xFFFFFFFF:
        raise WireError("InvalidData", "Vec length exceeds u32")
    return struct.pack("<I", count) + encode_elems(q, p)


def encode_linear_key(q: int, a_msg, a_open) -> bytes:
    """a_msg: (rows, msg_len, d), a_open: (rows, open_len, d)."""
    a_msg = np.asarray(a_msg, dtype=np.uint64)
    a_open = np.asarray(a_open, dtype=np.uint64)
    out = encode_ringmat(q, a_msg.shape[0], a_msg.shape[1], a_msg.reshape(-1, a_msg.shape[-1]))
    out += encode_ringmat(q, a_open.shape[0], a_open.shape[1],
                          a_open.reshape(-1, a_open.shape[-1]) if a_open.size else
                          np.zeros((0, a_msg.shape[-1]), np.uint64))
    return out


# ---- decoders -----------------------------------------------------------------

def _elements(q: int, data: bytes, off: int, count: int):
    """`count` elements from data[off:] in order: the first element that is
    short raises UnexpectedEnd, the first >= q InvalidData (prime.rs:705-721).
    Returns (values, bytes used)."""
    eb = elem_bytes(q)
    avail = max(0, (len(data) - off)) // eb
    n = min(count, avail)
    raw = np.frombuffer(data, dtype=np.uint8, count=n * eb, offset=off).reshape(n, eb)
    vals = np.zeros(n, dtype=np.uint64)
    for b in range(eb):
        vals |= raw[:, b].astype(np.uint64) << np.uint64(8 * b)
    bad = np.nonzero(vals >= np.uint64(q))[0]
    if bad.size:
        raise WireError("InvalidData", f"value {int(vals[bad[0]])} >= modulus {q}")
    if n < count:
        raise WireError("UnexpectedEnd")
    return vals, count * eb


def decode_poly(q: int, d: int, data: bytes, off: int = 0, ntt: bool = False):
    """One polynomial: CyclotomicPolyRing element by element, or TwistedNttPoly
    with its length checked first."""
    if ntt and len(data) - off < elem_bytes(q) * d:
        raise WireError("UnexpectedEnd")
    return _elements(q, data, off, d)


def _polys(q, d, data, off, count, ntt):
    out = np.zeros((count, d), dtype=np.uint64)
    used = 0
    for i in range(count):
        out[i], u = decode_poly(q, d, data, off + used, ntt)
        used += u
    return out, used


def _fast_polys(q, d, data, off, count, ntt):
    """Same result and errors as _polys without a Python loop per polynomial:
    a whole payload that is present decodes at once; a short one falls back to
    the element order of the reference."""
    eb = elem_bytes(q)
    if len(data) - off >= count * d * eb:
        vals, used = _elements(q, data, off, count * d)
        return vals.reshape(count, d), used
    if not ntt:
        _elements(q, data, off, count * d)  # raises in the reference's order
    whole = (len(data) - off) // (d * eb)
    _elements(q, data, off, whole * d)      # whole polynomials first
    raise WireError("UnexpectedEnd")


def _shape_ok(q: int, d: int, items: int, hdr: int):
    if items * d * elem_bytes(q) + hdr >= 1 << 63:
        raise WireError("InvalidData", "shape too large")


def decode_ringvec(q: int, d: int, data: bytes, off: int = 0, ntt: bool = False):
    if len(data) - off < 8:
        raise WireError("UnexpectedEnd")
    (count,) = struct.unpack_from("<Q", data, off)
    _shape_ok(q, d, count, 8)
    polys, used = _fast_polys(q, d, data, off + 8, count, ntt)
    return polys, 8 + used


def decode_ringmat(q: int, d: int, data: bytes, off: int = 0, ntt: bool = False):
    """-> (rows, cols, polys (rows*cols, d), used); overflow of rows*cols is
    InvalidData (lattice/types.rs:464-470)."""
    if len(data) - off < 16:
        raise WireError("UnexpectedEnd")
    rows, cols = struct.unpack_from("<QQ", data, off)
    if rows * cols >= 1 << 64:
        raise WireError("InvalidData", "matrix shape too large")
    _shape_ok(q, d, rows * cols, 16)
    polys, used = _fast_polys(q, d, data, off + 16, rows * cols, ntt)
    return rows, cols, polys, 16 + used


def decode_vec(q: int, d: int, data: bytes, off: int = 0, ntt: bool = False):
    if len(data) - off < 4:
        raise WireError("UnexpectedEnd")
    (count,) = struct.unpack_from("<I", data, off)
    _shape_ok(q, d, count, 4)
    polys, used = _fast_polys(q, d, data, off + 4, count, ntt)
    return polys, 4 + used


def decode_linear_key(q: int, d: int, data: bytes, ntt: bool = False, validate: bool = True):
    """-> (a_msg (rows, msg_len, d), a_open (rows, open_len, d), used).
    validate: deserialize_and_validate (serialize/src/lib.rs:106-110) with
    LinearCommitmentKey's Valid (commit/src/linear.rs:188-192)."""
    r1, c1, m, u1 = decode_ringmat(q, d, data, 0, ntt)
    r2, c2, o, u2 = decode_ringmat(q, d, data, u1, ntt)
    if validate and r1 != r2:
        raise WireError("InvalidData", "deserialized value is invalid")
    return m.reshape(r1, c1, d), o.reshape(r2, c2, d), u1 + u2


def exact(result_and_used, data: bytes):
    """deserialize_exact: the value when every byte was used, else InvalidData."""
    *value, used = result_and_used
    if used != len(data):
        raise WireError("InvalidData", "trailing bytes after deserialized value")
    return value[0] if len(value) == 1 else tuple(value)
