"""Canonical wire formats, key persistence and wire requests through the C ABI.

Every encoding must equal the oracle's (oracle/wire.py, pinned against the
reference's serialization tests) byte for byte; every decode must return the
same words or the same error kind, for the path's two rings and for rings
whose elements are 2 and 3 bytes wide. Keys saved and loaded again must commit
exactly as the key they came from.
"""
import struct
import zlib

import numpy as np
import pytest

from conftest import GOLDILOCKS, Q32, rand_residues, small_signed

pytestmark = pytest.mark.gpu

ERR = {"UnexpectedEnd": "UNEXPECTED_END", "InvalidData": "INVALID_DATA"}


@pytest.fixture(scope="module")
def rings(ctx, ring_q32, ring_gold):
    import lattica_amd as la
    r12 = la.Ring(ctx, 12289, 64)     # 2-byte elements
    r23 = la.Ring(ctx, 8380417, 64)   # 3-byte elements
    yield {"q32": ring_q32, "gold": ring_gold, "12289": r12, "8380417": r23}
    r12.close()
    r23.close()


def oracle_encode(w, kind, q, words, rows, cols=1):
    import lattica_amd as la
    if kind == la.WIRE_POLYS:
        return w.encode_polys(q, words)
    if kind == la.WIRE_RINGVEC:
        return w.encode_ringvec(q, words)
    if kind == la.WIRE_RINGMAT:
        return w.encode_ringmat(q, rows, cols, words)
    if kind == la.WIRE_VEC:
        return w.encode_vec(q, words)
    return struct.pack("<I", rows) + w.encode_elems(q, words)


def oracle_decode(w, kind, q, d, data, ntt, rows=None):
    """-> ('ok', words, used) or (error kind,)"""
    import lattica_amd as la
    try:
        if kind == la.WIRE_POLYS:
            p, used = w._fast_polys(q, d, data, 0, rows, ntt)
            return ("ok", p, used)
        if kind == la.WIRE_RINGVEC:
            p, used = w.decode_ringvec(q, d, data, 0, ntt)
            return ("ok", p, used)
        if kind == la.WIRE_RINGMAT:
            _, _, p, used = w.decode_ringmat(q, d, data, 0, ntt)
            return ("ok", p, used)
        if kind == la.WIRE_VEC:
            p, used = w.decode_vec(q, d, data, 0, ntt)
            return ("ok", p, used)
        if len(data) < 4:
            raise w.WireError("UnexpectedEnd")
        (n,) = struct.unpack_from("<I", data, 0)
        v, used = w._elements(q, data, 4, n)
        return ("ok", v, 4 + used)
    except w.WireError as e:
        return (e.kind,)


def lib_decode(ring, kind, data, ntt, rows, cols):
    import lattica_amd as la
    try:
        words, used = ring.wire_decode(kind, data, la.DOMAIN_NTT if ntt else la.DOMAIN_COEFF,
                                       rows=rows, cols=cols, exact=False)
        return ("ok", words, used)
    except la.GridHipError as e:
        return (e.name,)


KINDS = ["POLYS", "RINGVEC", "RINGMAT", "VEC", "ELEMS"]


def shape_for(kind, n):
    return (n // 4, 4) if kind == "RINGMAT" else (n, 1)


@pytest.mark.parametrize("name", ["q32", "gold", "12289", "8380417"])
@pytest.mark.parametrize("kname", KINDS)
def test_encode_matches_oracle_and_round_trips(rings, name, kname):
    import lattica_amd as la
    from oracle import wire
    ring = rings[name]
    kind = getattr(la, "WIRE_" + kname)
    rng = np.random.default_rng(zlib.crc32(f"{name}/{kname}".encode()))
    for n in (0, 1, 4, 67, 1000):
        rows, cols = shape_for(kname, n)
        if kname == "ELEMS":
            words = rand_residues(rng, ring.q, (rows,))
        else:
            words = rand_residues(rng, ring.q, (rows * cols, 64))
        enc = ring.wire_encode(kind, words, rows=rows, cols=cols)
        assert enc == oracle_encode(wire, kind, ring.q, words, rows, cols), (name, kname, n)
        assert ring.wire_size(kind, rows, cols) == len(enc)
        if kind != la.WIRE_POLYS:
            assert ring.wire_peek(kind, enc)[:2] == (rows, cols)
        for domain in (la.DOMAIN_COEFF, la.DOMAIN_NTT):
            back, used = ring.wire_decode(kind, enc, domain, rows=rows, cols=cols)
            assert used == len(enc)
            assert np.array_equal(back.reshape(-1), words.reshape(-1).astype(ring.dtype))


@pytest.mark.parametrize("name", ["q32", "gold", "8380417"])
def test_device_memory_encode_and_decode(ctx, rings, name):
    import lattica_amd as la
    from oracle import wire
    ring = rings[name]
    rng = np.random.default_rng(7)
    rows = 33
    words = rand_residues(rng, ring.q, (rows, 64)).astype(ring.dtype)
    buf = la.DeviceBuffer(ctx, words.nbytes)
    out = la.DeviceBuffer(ctx, words.nbytes)
    try:
        buf.upload(words)
        size = ring.wire_size(la.WIRE_RINGVEC, rows)
        enc = np.zeros(size, dtype=np.uint8)
        wr = __import__("ctypes").c_uint64()
        L = la.lib()
        assert L.gh_wire_encode(ring.h, la.WIRE_RINGVEC, rows, 1, buf.ptr, la.MEM_DEVICE,
                                enc.ctypes.data, size, __import__("ctypes").byref(wr)) == 0
        assert enc.tobytes() == wire.encode_ringvec(ring.q, words)
        assert L.gh_wire_decode(ring.h, la.WIRE_RINGVEC, la.DOMAIN_COEFF, enc.ctypes.data, size,
                                rows, 1, out.ptr, la.MEM_DEVICE, None) == 0
        assert np.array_equal(out.download(ring.dtype, (rows, 64)), words)
    finally:
        buf.free()
        out.free()


@pytest.mark.parametrize("name", ["q32", "gold", "12289", "8380417"])
@pytest.mark.parametrize("kname", ["POLYS", "RINGVEC", "RINGMAT", "VEC", "ELEMS"])
def test_decode_errors_follow_the_reference_order(rings, name, kname):
    """Every truncation point, and a corrupted element before, inside or after
    the cut, in both domains: the library's error (or words) equals the oracle's."""
    import lattica_amd as la
    from oracle import wire
    ring = rings[name]
    kind = getattr(la, "WIRE_" + kname)
    eb = wire.elem_bytes(ring.q)
    rng = np.random.default_rng(31)
    rows, cols = shape_for(kname, 4)
    per = 1 if kname == "ELEMS" else 64
    words = rand_residues(rng, ring.q, (rows * cols, per) if per > 1 else (rows,))
    enc = oracle_encode(wire, kind, ring.q, words, rows, cols)
    hdr = len(enc) - rows * cols * per * eb
    cuts = sorted(set([0, 1, hdr, hdr + 1, hdr + eb, hdr + 64 * eb - 1, hdr + 64 * eb,
                       hdr + 64 * eb + 3 * eb + 1, len(enc) - 1, len(enc), len(enc) + 1]))
    bad_elems = [0, 63, 64, 64 + 5, rows * cols * per - 1] if per > 1 else [0, 1, rows - 1]
    seen = set()
    for cut in cuts:
        for bad in [None] + bad_elems:
            data = bytearray(enc + b"\x00")[:cut] if cut > len(enc) else bytearray(enc[:cut])
            if bad is not None:  # every byte 0xFF: a value >= q for each ring here
                pos = hdr + bad * eb
                if pos >= len(data):
                    continue
                for b in range(pos, min(pos + eb, len(data))):
                    data[b] = 0xFF
            data = bytes(data)
            for ntt in (False, True):
                exp = oracle_decode(wire, kind, ring.q, 64, data, ntt, rows)
                got = lib_decode(ring, kind, data, ntt, rows, cols)
                if exp[0] == "ok":
                    assert got[0] == "ok", (cut, bad, ntt, got)
                    assert got[2] == exp[2]
                    assert np.array_equal(got[1].reshape(-1), exp[1].reshape(-1).astype(ring.dtype))
                else:
                    assert got[0] == ERR[exp[0]], (cut, bad, ntt, exp, got)
                seen.add(got[0])
    assert {"ok", "UNEXPECTED_END", "INVALID_DATA"} <= seen


def test_exact_consumption_and_shape_errors(ring_q32):
    import lattica_amd as la
    rng = np.random.default_rng(3)
    words = rand_residues(rng, ring_q32.q, (5, 64))
    enc = ring_q32.wire_encode(la.WIRE_RINGVEC, words)
    with pytest.raises(la.GridHipError) as e:  # deserialize_exact: trailing bytes
        ring_q32.wire_decode(la.WIRE_RINGVEC, enc + b"\x07", rows=5, cols=1)
    assert e.value.name == "INVALID_DATA"
    _, used = ring_q32.wire_decode(la.WIRE_RINGVEC, enc + b"\x07", rows=5, cols=1, exact=False)
    assert used == len(enc)
    with pytest.raises(la.GridHipError) as e:  # header holds 5, the caller expects 4
        ring_q32.wire_decode(la.WIRE_RINGVEC, enc, rows=4, cols=1)
    assert e.value.name == "DIMENSION_MISMATCH"
    with pytest.raises(la.GridHipError) as e:  # lattice/types.rs:600-607
        ring_q32.wire_peek(la.WIRE_RINGMAT, struct.pack("<QQ", 2 ** 64 - 1, 2))
    assert e.value.name == "INVALID_DATA"
    with pytest.raises(la.GridHipError) as e:
        ring_q32.wire_encode(la.WIRE_RINGVEC, np.full((1, 64), Q32, dtype=np.uint64))
    assert e.value.name == "NOT_CANONICAL"
    with pytest.raises(la.GridHipError) as e:
        ring_q32.wire_size(la.WIRE_VEC, 2 ** 32, 1)  # Vec lengths are u32
    assert e.value.name == "INVALID_DATA"


def test_large_key_shaped_matrix_round_trip(ring_q32):
    """A 16 x 4096 RingMat (16 MiB of words): encode equals the oracle's
    packing, decode returns the words."""
    import lattica_amd as la
    from oracle import wire
    rng = np.random.default_rng(8)
    words = rand_residues(rng, ring_q32.q, (16 * 4096, 64))
    enc = ring_q32.wire_encode(la.WIRE_RINGMAT, words, rows=16, cols=4096)
    assert enc == wire.encode_ringmat(ring_q32.q, 16, 4096, words)
    back, _ = ring_q32.wire_decode(la.WIRE_RINGMAT, enc, la.DOMAIN_NTT)
    assert np.array_equal(back, words.astype(np.uint32))


# ---- prepared-key persistence ------------------------------------------------

@pytest.mark.parametrize("name", ["q32", "gold"])
@pytest.mark.parametrize("open_len", [0, 3])
def test_key_save_matches_oracle_encoding(oracle, rings, name, open_len):
    import lattica_amd as la
    from oracle import wire
    ring = rings[name]
    rng = np.random.default_rng(50 + open_len)
    rows, msg_len = 8, 37
    a_msg = rand_residues(rng, ring.q, (rows, msg_len, 64))
    a_open = rand_residues(rng, ring.q, (rows, open_len, 64))
    key = la.Key(ring, a_msg, a_open if open_len else None, domain=la.DOMAIN_COEFF)
    try:
        coeff = key.save(la.DOMAIN_COEFF)
        assert coeff == wire.encode_linear_key(ring.q, a_msg, a_open)
        ntt_msg = oracle.twisted_forward(ring.q, 64, a_msg.reshape(-1, 64)).reshape(a_msg.shape)
        ntt_open = (oracle.twisted_forward(ring.q, 64, a_open.reshape(-1, 64)).reshape(a_open.shape)
                    if open_len else a_open)
        ntt = key.save(la.DOMAIN_NTT)
        assert ntt == wire.encode_linear_key(ring.q, ntt_msg, ntt_open)
    finally:
        key.close()


@pytest.mark.parametrize("name", ["q32", "gold"])
def test_key_load_commits_like_the_original(rings, name):
    import lattica_amd as la
    ring = rings[name]
    rng = np.random.default_rng(60)
    rows, msg_len, open_len = 16, 300, 5
    a_msg = rand_residues(rng, ring.q, (rows, msg_len, 64))
    a_open = rand_residues(rng, ring.q, (rows, open_len, 64))
    m = small_signed(rng, ring.q, (msg_len, 64), -8, 8)
    r = small_signed(rng, ring.q, (open_len, 64), -2, 3)
    key = la.Key(ring, a_msg, a_open, domain=la.DOMAIN_COEFF)
    try:
        t0 = key.commit(m, r)
        for dom in (la.DOMAIN_NTT, la.DOMAIN_COEFF):
            blob = key.save(dom)
            k2 = ring.load_key(blob, dom)
            try:
                assert (k2.rows, k2.msg_len, k2.open_len) == (rows, msg_len, open_len)
                assert np.array_equal(k2.commit(m, r), t0)
                assert k2.save(dom) == blob
            finally:
                k2.close()
    finally:
        key.close()


def test_key_load_from_oracle_bytes_and_errors(oracle, ring_q32):
    import lattica_amd as la
    from oracle import wire
    q = ring_q32.q
    rng = np.random.default_rng(70)
    a_msg = rand_residues(rng, q, (4, 9, 64))
    a_open = rand_residues(rng, q, (4, 2, 64))
    m = small_signed(rng, q, (9, 64), -8, 8)
    r = small_signed(rng, q, (2, 64), -2, 3)
    blob = wire.encode_linear_key(q, a_msg, a_open)
    k = ring_q32.load_key(blob, la.DOMAIN_COEFF)
    try:
        expect = oracle.commit(q, 64, a_msg, a_open, m, r)
        assert np.array_equal(k.commit(m, r), np.asarray(expect).astype(np.uint32))
    finally:
        k.close()

    def err(data, domain=la.DOMAIN_COEFF, exact=True):
        with pytest.raises(la.GridHipError) as e:
            ring_q32.load_key(data, domain, exact).close()
        return e.value.name

    # commit/src/linear.rs:1100-1124: row counts differ
    bad_rows = wire.encode_linear_key(q, a_msg, a_open[:3])
    assert err(bad_rows) == "INVALID_DATA"
    assert err(blob + b"\x00") == "INVALID_DATA"          # trailing byte
    k = ring_q32.load_key(blob + b"\x00", la.DOMAIN_COEFF, exact=False)
    assert k.consumed == len(blob)
    k.close()
    for cut in (0, 15, 16, 100, 16 + 4 * 9 * 256, 16 + 4 * 9 * 256 + 15, len(blob) - 1):
        exp = None
        try:
            wire.decode_linear_key(q, 64, blob[:cut])
        except wire.WireError as e:
            exp = ERR[e.kind]
        assert err(blob[:cut]) == exp, cut
    corrupt = bytearray(blob)
    corrupt[16:20] = b"\xff" * 4                            # first element of a_msg >= q
    assert err(bytes(corrupt)) == "INVALID_DATA"
    assert err(bytes(corrupt[:200])) == "INVALID_DATA"      # before the end is reported
    empty_msg = wire.encode_linear_key(q, np.zeros((4, 0, 64), np.uint64), a_open)
    assert err(empty_msg) == "INVALID_PARAMETERS"           # CommitmentDimensions::validate


# ---- wire requests on the server ---------------------------------------------

def wire_witness(w, q, parts):
    """LabradorWitness::parts as Vec<Vec<poly>>."""
    out = struct.pack("<I", len(parts))
    for p in parts:
        out += w.encode_vec(q, p)
    return out


def test_server_wire_requests_equal_plain_encodings(ctx, ring_q32, ring_gold):
    import lattica_amd as la
    from oracle import wire
    rng = np.random.default_rng(90)
    rings = {"q32": ring_q32, "gold": ring_gold}
    keys = {name: la.Key(rg, rand_residues(rng, rg.q, (8, 512, 64)), domain=la.DOMAIN_NTT)
            for name, rg in rings.items()}
    srv = la.Server(ctx, mem_bytes=1 << 28)
    try:
        reqs = []
        for i in range(24):
            name = "q32" if i % 2 == 0 else "gold"
            ring = rings[name]
            r, n = int(rng.integers(1, 6)), int(rng.integers(1, 300))
            s = small_signed(rng, ring.q, (r, n, 64), -100, 100)
            ops = [la.OP_COMMIT, la.OP_GRAM, la.OP_COMMIT | la.OP_GRAM][i % 3]
            reqs.append((name, ops, s))
        bad_index = 5
        tickets = []
        for i, (name, ops, s) in enumerate(reqs):
            blob = bytearray(wire_witness(wire, rings[name].q, list(s)))
            if i == bad_index:  # the last element of the last part >= q
                eb = rings[name].info.wire_bytes
                blob[-eb:] = b"\xff" * eb
            tickets.append(srv.submit_wire(ops, rings[name], bytes(blob), key=keys[name]))
        results = [srv.wait(tk) for tk in tickets]
        stats = srv.stats()
    finally:
        srv.close()
    for i, ((name, ops, s), (st, lat, t, g)) in enumerate(zip(reqs, results)):
        ring, key = rings[name], keys[name]
        if i == bad_index:
            assert st == -31  # GH_E_INVALID_DATA
            continue
        assert st == 0
        if ops & la.OP_COMMIT:
            plain = [key.commit_batch([s[k]])[0] for k in range(s.shape[0])]
            exp = struct.pack("<I", s.shape[0]) + b"".join(
                wire.encode_ringvec(ring.q, p.astype(np.uint64)) for p in plain)
            assert t == exp, i
        else:
            assert t is None
        if ops & la.OP_GRAM:
            assert g == wire.encode_vec(ring.q, ring.gram(s).astype(np.uint64)), i
        else:
            assert g is None
    assert stats["failed"] == 1 and stats["completed"] == len(reqs) - 1
    for k in keys.values():
        k.close()


def test_server_wire_structure_errors(ctx, ring_q32):
    import lattica_amd as la
    from oracle import wire
    rng = np.random.default_rng(91)
    key = la.Key(ring_q32, rand_residues(rng, ring_q32.q, (4, 64, 64)), domain=la.DOMAIN_NTT)
    srv = la.Server(ctx, mem_bytes=1 << 26)
    try:
        s = small_signed(rng, ring_q32.q, (3, 10, 64), -5, 5)
        good = wire_witness(wire, ring_q32.q, list(s))

        def err(blob, **kw):
            with pytest.raises(la.GridHipError) as e:
                srv.submit_wire(la.OP_COMMIT, ring_q32, blob, key=key, **kw)
            return e.value.name

        assert err(good[:3]) == "UNEXPECTED_END"
        assert err(good[:-1]) == "UNEXPECTED_END"
        assert err(good + b"\x00") == "INVALID_DATA"
        uneven = wire_witness(wire, ring_q32.q, [s[0], s[1][:9]])
        assert err(uneven) == "DIMENSION_MISMATCH"
        assert err(good, r=2) == "DIMENSION_MISMATCH"
        assert err(struct.pack("<I", 0)) == "DIMENSION_MISMATCH"
        st, _, t, _ = srv.wait(srv.submit_wire(la.OP_COMMIT, ring_q32, good, key=key, r=3, n=10))
        assert st == 0 and len(t) == 4 + 3 * (8 + 4 * 64 * 4)
    finally:
        srv.close()
        key.close()
