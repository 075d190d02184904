"""ctypes binding of the oracle (oracle/liboracle.so).

TEST INFRASTRUCTURE ONLY. The oracle is the CPU restatement of the reference
path that the HIP library is checked against; only tests/, the smoke check and
bench.py's cpu_baseline leg load it. Arrays are numpy uint64 canonical residues.
"""
from __future__ import annotations

import ctypes
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB_PATH = os.path.join(_HERE, "liboracle.so")
_lib = None

U64P = ctypes.POINTER(ctypes.c_uint64)
I64P = ctypes.POINTER(ctypes.c_int64)
I8P = ctypes.POINTER(ctypes.c_int8)
U8P = ctypes.POINTER(ctypes.c_uint8)


def build() -> str:
    """Compile liboracle.so with the committed Makefile."""
    subprocess.run(["make", "-s", "-C", _HERE], check=True)
    return _LIB_PATH


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(_LIB_PATH):
            build()
        L = ctypes.CDLL(_LIB_PATH)
        u64, i32, sz = ctypes.c_uint64, ctypes.c_int, ctypes.c_size_t
        for name in ("gr_add", "gr_sub", "gr_mul", "gr_pow"):
            getattr(L, name).restype = u64
            getattr(L, name).argtypes = [u64, u64, u64]
        for name in ("gr_inv", "gr_to_montgomery", "gr_from_montgomery"):
            getattr(L, name).restype = u64
            getattr(L, name).argtypes = [u64, u64]
        L.gr_q_inv_neg.restype = u64
        L.gr_q_inv_neg.argtypes = [u64]
        L.gr_slice_op.restype = i32
        L.gr_slice_op.argtypes = [u64, i32, U64P, U64P, u64, sz]
        L.gr_dot_product.restype = u64
        L.gr_dot_product.argtypes = [u64, U64P, U64P, sz]
        L.gr_max_ntt_size.restype = u64
        L.gr_max_ntt_size.argtypes = [u64]
        L.gr_find_generator.restype = i32
        L.gr_find_generator.argtypes = [u64, U64P]
        L.gr_root_of_unity.restype = i32
        L.gr_root_of_unity.argtypes = [u64, u64, U64P]
        for name in ("gr_ntt_forward", "gr_ntt_inverse", "gr_twisted_forward", "gr_twisted_inverse"):
            getattr(L, name).restype = i32
            getattr(L, name).argtypes = [u64, u64, U64P, sz]
        for name in ("gr_negacyclic_mul", "gr_negacyclic_mul_naive"):
            getattr(L, name).restype = i32
            getattr(L, name).argtypes = [u64, u64, U64P, U64P, U64P, sz]
        L.gr_commit_prepared.restype = i32
        L.gr_commit_prepared.argtypes = [u64, u64, ctypes.c_uint32, U64P, ctypes.c_uint32, U64P,
                                         ctypes.c_uint32, i32, U64P, U64P, U64P]
        L.gr_commit_coeff.restype = i32
        L.gr_commit_coeff.argtypes = [u64, u64, ctypes.c_uint32, U64P, ctypes.c_uint32, U64P,
                                      ctypes.c_uint32, U64P, U64P, U64P]
        L.gr_gram.restype = i32
        L.gr_gram.argtypes = [u64, u64, ctypes.c_uint32, ctypes.c_uint32, U64P, U64P, i32]
        L.gr_decompose_balanced.restype = i32
        L.gr_decompose_balanced.argtypes = [u64, u64, sz, U64P, u64, ctypes.c_uint32, U64P]
        L.gr_garbage_h.restype = i32
        L.gr_garbage_h.argtypes = [u64, u64, ctypes.c_uint32, ctypes.c_uint32, U64P, U64P, U64P]
        L.gr_amortize.restype = i32
        L.gr_amortize.argtypes = [u64, u64, ctypes.c_uint32, ctypes.c_uint32, U64P, U64P, U64P]
        L.gr_jl_project.restype = i32
        L.gr_jl_project.argtypes = [u64, U8P, ctypes.c_uint32, ctypes.c_uint32, u64, u64, U64P,
                                    U64P, I64P, I8P]
        L.gr_chacha20_block.restype = None
        L.gr_chacha20_block.argtypes = [U8P, ctypes.c_uint32, U8P, U8P]
        L.gr_avx2_available.restype = i32
        L.gr_avx2_commit.restype = i32
        L.gr_avx2_commit.argtypes = [u64, i32, ctypes.c_uint32, ctypes.c_uint32, U64P, U64P, U64P, i32]
        L.gr_avx2_prepare.restype = ctypes.c_void_p
        L.gr_avx2_prepare.argtypes = [u64, i32, ctypes.c_uint32, ctypes.c_uint32, ctypes.c_uint32, U64P]
        L.gr_avx2_load_witness.restype = None
        L.gr_avx2_load_witness.argtypes = [ctypes.c_void_p, U64P]
        L.gr_avx2_commit_run.restype = None
        L.gr_avx2_commit_run.argtypes = [ctypes.c_void_p, i32]
        L.gr_avx2_result.restype = None
        L.gr_avx2_result.argtypes = [ctypes.c_void_p, U64P]
        L.gr_avx2_free.restype = None
        L.gr_avx2_free.argtypes = [ctypes.c_void_p]
        L.gr_avx2_set_len.restype = i32
        L.gr_avx2_set_len.argtypes = [ctypes.c_void_p, ctypes.c_uint32]
        L.gr_avx2_load_witness_len.restype = None
        L.gr_avx2_load_witness_len.argtypes = [ctypes.c_void_p, U64P, ctypes.c_uint32]
        L.gr_avx2_negacyclic_load.restype = i32
        L.gr_avx2_negacyclic_load.argtypes = [ctypes.c_void_p, U64P, U64P, ctypes.c_size_t]
        L.gr_avx2_negacyclic_run.restype = None
        L.gr_avx2_negacyclic_run.argtypes = [ctypes.c_void_p, i32]
        L.gr_avx2_negacyclic_result.restype = None
        L.gr_avx2_negacyclic_result.argtypes = [ctypes.c_void_p, U64P]
        L.gr_avx2_xform_run.restype = None
        L.gr_avx2_xform_run.argtypes = [ctypes.c_void_p, i32, i32]
        L.gr_avx2_xform_result.restype = None
        L.gr_avx2_xform_result.argtypes = [ctypes.c_void_p, U64P]
        U32P = ctypes.POINTER(ctypes.c_uint32)
        L.gr_avx2_dots_load.restype = i32
        L.gr_avx2_dots_load.argtypes = [ctypes.c_void_p, U64P, ctypes.c_uint32, U64P,
                                        ctypes.c_uint32, ctypes.c_uint32, U32P, U32P,
                                        ctypes.c_uint32]
        L.gr_avx2_dots_run.restype = None
        L.gr_avx2_dots_run.argtypes = [ctypes.c_void_p, i32]
        L.gr_avx2_dots_result.restype = None
        L.gr_avx2_dots_result.argtypes = [ctypes.c_void_p, U64P]
        L.gr_avx2_items_load.restype = i32
        L.gr_avx2_items_load.argtypes = [ctypes.c_void_p, U64P, ctypes.c_size_t, U64P, U32P,
                                         ctypes.c_uint32]
        L.gr_avx2_items_run.restype = None
        L.gr_avx2_items_run.argtypes = [ctypes.c_void_p, i32]
        L.gr_avx2_items_result.restype = None
        L.gr_avx2_items_result.argtypes = [ctypes.c_void_p, U64P]
        _lib = L
    return _lib


class OracleError(RuntimeError):
    def __init__(self, code: int, what: str):
        super().__init__(f"{what}: status {code}")
        self.code = code


def _p(a: np.ndarray, t=U64P):
    return a.ctypes.data_as(t)


def _u64(a) -> np.ndarray:
    return np.ascontiguousarray(np.asarray(a, dtype=np.uint64))


def _check(code: int, what: str):
    if code != 0:
        raise OracleError(code, what)


def add(q, a, b):
    return int(lib().gr_add(q, a, b))


def sub(q, a, b):
    return int(lib().gr_sub(q, a, b))


def mul(q, a, b):
    return int(lib().gr_mul(q, a, b))


SLICE_OPS = {"add": 0, "sub": 1, "scalar_mul": 2, "add_scaled": 3}


def slice_op(q, op: str, dst, src=None, scalar: int = 0):
    """Ring slice hooks of PrimeField (prime.rs:433-491) on flat word arrays:
    add / sub (dst op src), scalar_mul (dst * scalar), add_scaled (dst + src * scalar)."""
    d = _u64(dst).copy().reshape(-1)
    s = _u64(src).reshape(-1) if src is not None else np.zeros(1, dtype=np.uint64)
    _check(lib().gr_slice_op(q, SLICE_OPS[op], _p(d), _p(s), int(scalar), d.size), "slice_op")
    return d


def dot_product(q, a, b) -> int:
    """Ring::dot_product of PrimeField (prime.rs:467-491)."""
    a, b = _u64(a).reshape(-1), _u64(b).reshape(-1)
    assert a.size == b.size
    return int(lib().gr_dot_product(q, _p(a), _p(b), a.size))


def pow_(q, a, e):
    return int(lib().gr_pow(q, a, e))


def inv(q, a):
    return int(lib().gr_inv(q, a))


def to_montgomery(q, a):
    return int(lib().gr_to_montgomery(q, a))


def from_montgomery(q, a):
    return int(lib().gr_from_montgomery(q, a))


def q_inv_neg(q):
    return int(lib().gr_q_inv_neg(q))


def max_ntt_size(q):
    return int(lib().gr_max_ntt_size(q))


def find_generator(q) -> int:
    out = ctypes.c_uint64()
    _check(lib().gr_find_generator(q, ctypes.byref(out)), "find_generator")
    return out.value


def root_of_unity(q, n) -> int:
    out = ctypes.c_uint64()
    _check(lib().gr_root_of_unity(q, n, ctypes.byref(out)), "root_of_unity")
    return out.value


def _xform(name, q, d, values):
    v = _u64(values).copy()
    count = v.size // d if d else 0
    _check(getattr(lib(), name)(q, d, _p(v), count), name)
    return v


def ntt_forward(q, d, values):
    return _xform("gr_ntt_forward", q, d, values)


def ntt_inverse(q, d, values):
    return _xform("gr_ntt_inverse", q, d, values)


def twisted_forward(q, d, values):
    return _xform("gr_twisted_forward", q, d, values)


def twisted_inverse(q, d, values):
    return _xform("gr_twisted_inverse", q, d, values)


def xform_status(name, q, d, values) -> int:
    v = _u64(values).copy()
    return int(getattr(lib(), "gr_" + name)(q, d, _p(v), max(1, v.size // max(d, 1))))


def negacyclic_mul(q, d, a, b, naive=False):
    a, b = _u64(a), _u64(b)
    out = np.zeros_like(a)
    fn = lib().gr_negacyclic_mul_naive if naive else lib().gr_negacyclic_mul
    _check(fn(q, d, _p(a), _p(b), _p(out), a.size // d), "negacyclic_mul")
    return out


def commit(q, d, a_msg, a_open, message, randomness, key_domain=0, path="prepared"):
    """t = A_msg*m + A_open*r. a_msg: (rows, msg_len, d), a_open: (rows, open_len, d)."""
    a_msg, a_open = _u64(a_msg), _u64(a_open)
    m, r = _u64(message), _u64(randomness)
    rows, msg_len = a_msg.shape[0], a_msg.shape[1]
    open_len = a_open.shape[1]
    out = np.zeros((rows, d), dtype=np.uint64)
    if path == "prepared":
        st = lib().gr_commit_prepared(q, d, rows, _p(a_msg), msg_len, _p(a_open), open_len,
                                      key_domain, _p(m), _p(r), _p(out))
    else:
        st = lib().gr_commit_coeff(q, d, rows, _p(a_msg), msg_len, _p(a_open), open_len,
                                   _p(m), _p(r), _p(out))
    _check(st, "commit")
    return out


def gram(q, d, s, ntt_once=True):
    """s: (r, n, d) -> (r(r+1)/2, d) upper triangle, row-major."""
    s = _u64(s)
    r, n = s.shape[0], s.shape[1]
    out = np.zeros((r * (r + 1) // 2, d), dtype=np.uint64)
    _check(lib().gr_gram(q, d, r, n, _p(s), _p(out), 1 if ntt_once else 0), "gram")
    return out


def decompose_balanced(q, d, polys, base, limbs):
    """polys: (m, d) -> (m, limbs, d) balanced base-`base` limbs stored mod q."""
    polys = _u64(polys)
    m = polys.shape[0]
    out = np.zeros((m, limbs, d), dtype=np.uint64)
    _check(lib().gr_decompose_balanced(q, d, m, _p(polys), base, limbs, _p(out)),
           "decompose_balanced")
    return out


def garbage_h(q, d, phi, s):
    """phi, s: (r, n, d) -> (r(r+1)/2, d): (<phi_i, s_j> + <phi_j, s_i>) / 2."""
    phi, s = _u64(phi), _u64(s)
    r, n = s.shape[0], s.shape[1]
    out = np.zeros((r * (r + 1) // 2, d), dtype=np.uint64)
    _check(lib().gr_garbage_h(q, d, r, n, _p(phi), _p(s), _p(out)), "garbage_h")
    return out


def amortize(q, d, c, s):
    """c: (r, d), s: (r, n, d) -> z: (n, d) = sum_i c_i * s_i."""
    c, s = _u64(c), _u64(s)
    r, n = s.shape[0], s.shape[1]
    out = np.zeros((n, d), dtype=np.uint64)
    _check(lib().gr_amortize(q, d, r, n, _p(c), _p(s), _p(out)), "amortize")
    return out


def jl_project(q, seed: bytes, rows, parts, d, want_flat=False):
    """parts: (num_parts, cols) residues. Returns (p_modq, p_exact, flat|None)."""
    parts = _u64(parts)
    num_parts, cols = parts.shape
    sd = np.frombuffer(bytes(seed), dtype=np.uint8).copy()
    pm = np.zeros(rows, dtype=np.uint64)
    pe = np.zeros(rows, dtype=np.int64)
    flat = np.zeros(rows * num_parts * cols, dtype=np.int8) if want_flat else None
    _check(lib().gr_jl_project(q, _p(sd, U8P), rows, num_parts, cols, d, _p(parts), _p(pm),
                               _p(pe, I64P), _p(flat, I8P) if want_flat else None), "jl_project")
    return pm, pe, flat


def jl_extract_flat(seed: bytes, rows, num_parts, cols, d):
    """JLMatrix::extract_jl_rows_flat (labrador/src/jl.rs:311-383): the entries
    alone, [m][part][coeff][poly] int8. The entries do not depend on the
    witness, so this is jl_project's flat output over a zero witness."""
    _, _, flat = jl_project(4294966657, seed, rows, np.zeros((num_parts, cols), np.uint64), d,
                            want_flat=True)
    return flat


def _addmod(a, b, q):
    """(a + b) mod q elementwise for uint64 arrays of residues, q up to 2^64."""
    s = a + b  # wraps mod 2^64
    wrap = (s < a) | (s >= np.uint64(q))
    return np.where(wrap, s - np.uint64(q), s)


def jl_conjugate_flat(q, flat, rows, num_parts, n, d):
    """jl_rows_flat_to_conjugated_polys (labrador/src/main_protocol/mod.rs:885-923):
    flat [m][part][coeff][poly] -> (rows, num_parts, n, d) polynomials with
    coefficient k = entry(src) for k = 0 and q - entry(src) otherwise (0 stays
    0), src = 0 for k = 0 else d - k, entry lifted to [0, q)."""
    f = np.asarray(flat, dtype=np.int64).reshape(rows, num_parts, d, n)
    src = np.array([0] + [d - k for k in range(1, d)])
    e = f[:, :, src, :]                       # (rows, parts, d(k), n)
    u = np.where(e < 0, np.uint64(q) - np.uint64(1), e.astype(np.uint64))  # v in {-1, 0, 1}
    neg = np.where(u == 0, np.uint64(0), np.uint64(q) - u)
    out = np.where((np.arange(d) == 0)[None, None, :, None], u, neg)
    return np.ascontiguousarray(out.transpose(0, 1, 3, 2)).astype(np.uint64)


def jl_evals(q, d, seed: bytes, rows, parts_polys, naive=False):
    """jl_eval[m] = sum_i sum_j conj(pi_i^(m))_j * s_ij over Z_q[X]/(X^d + 1)
    (labrador/src/main_protocol/step_prover.rs:196-205): the conjugated JL
    polynomials times the witness polynomials through the oracle's negacyclic
    product, summed in the reference's order. parts_polys: (num_parts, n, d)."""
    s = _u64(parts_polys)
    num_parts, n = s.shape[0], s.shape[1]
    flat = jl_extract_flat(seed, rows, num_parts, n * d, d)
    conj = jl_conjugate_flat(q, flat, rows, num_parts, n, d)
    b = np.broadcast_to(s[None], conj.shape).reshape(-1, d)
    prod = negacyclic_mul(q, d, conj.reshape(-1, d), np.ascontiguousarray(b), naive=naive)
    prod = prod.reshape(rows, num_parts * n, d)
    acc = np.zeros((rows, d), dtype=np.uint64)
    for c in range(num_parts * n):
        acc = _addmod(acc, prod[:, c, :], q)
    return acc


def jl_aggregate(q, d, seed: bytes, rows, scale, num_parts, n, phi=None, naive=False):
    """The JL term of aggregate_step2_flat (labrador/src/main_protocol/aggregation.rs:237-245):
    phi[i][j] += scale_m * conj(pi_i^(m))_j summed over the rows m, with the
    conjugated JL polynomials of jl_rows_flat_to_conjugated_polys
    (main_protocol/mod.rs:885-923) and the oracle's negacyclic product, in the
    reference's order (m outer, then i, j). scale: (rows, d) polynomials, the
    caller's jl_scale_m = sum_k beta_k * omega_k[m]; phi: (num_parts, n, d) or
    None for zeros. Returns the new phi."""
    sc = _u64(scale).reshape(rows, d)
    acc = (np.zeros((num_parts * n, d), dtype=np.uint64) if phi is None
           else _u64(phi).reshape(num_parts * n, d).copy())
    flat = jl_extract_flat(seed, rows, num_parts, n * d, d)
    conj = jl_conjugate_flat(q, flat, rows, num_parts, n, d).reshape(rows * num_parts * n, d)
    a = np.repeat(sc, num_parts * n, axis=0)     # scale_m beside each (m, i, j)
    prod = negacyclic_mul(q, d, a, conj, naive=naive).reshape(rows, num_parts * n, d)
    for m in range(rows):
        acc = _addmod(acc, prod[m], q)
    return acc.reshape(num_parts, n, d)


def chacha20_block(key: bytes, counter: int, nonce: bytes) -> bytes:
    k = np.frombuffer(bytes(key), dtype=np.uint8).copy()
    n = np.frombuffer(bytes(nonce), dtype=np.uint8).copy()
    out = np.zeros(64, dtype=np.uint8)
    lib().gr_chacha20_block(_p(k, U8P), counter, _p(n, U8P), _p(out, U8P))
    return out.tobytes()


class Avx2Commit:
    """The reference's AVX2 prepared-commit path, restated (CPU baseline)."""

    def __init__(self, q, limb_bits, a_ntt, d=64):
        a_ntt = _u64(a_ntt)
        self.rows, self.n = a_ntt.shape[0], a_ntt.shape[1]
        self.d = d
        self._keep = a_ntt
        self.h = lib().gr_avx2_prepare(q, limb_bits, self.rows, self.n, d, _p(a_ntt))
        if not self.h:
            raise OracleError(-3, "avx2 prepare")

    def load(self, s):
        """Witness of n polynomials (the key's width), or fewer: a prefix commit."""
        s = _u64(s)
        n_use = s.shape[0]
        _check(lib().gr_avx2_set_len(self.h, n_use), "avx2 set_len")
        lib().gr_avx2_load_witness_len(self.h, _p(s), n_use)

    def run(self, threads=1):
        lib().gr_avx2_commit_run(self.h, threads)

    def result(self):
        out = np.zeros((self.rows, self.d), dtype=np.uint64)
        lib().gr_avx2_result(self.h, _p(out))
        return out

    def load_items(self, pool, offs, lens):
        """Independent commitments: item i = pool[offs[i] : offs[i] + lens[i]]."""
        pool = _u64(pool)
        offs = np.ascontiguousarray(offs, dtype=np.uint64)
        lens = np.ascontiguousarray(lens, dtype=np.uint32)
        self._items = (offs, lens)
        u32p = ctypes.POINTER(ctypes.c_uint32)
        _check(lib().gr_avx2_items_load(self.h, _p(pool), pool.shape[0], _p(offs),
                                        lens.ctypes.data_as(u32p), len(lens)), "avx2 items load")

    def run_items(self, threads=1):
        lib().gr_avx2_items_run(self.h, threads)

    def items_result(self):
        out = np.zeros((len(self._items[1]), self.rows, self.d), dtype=np.uint64)
        lib().gr_avx2_items_result(self.h, _p(out))
        return out

    def close(self):
        if self.h:
            lib().gr_avx2_free(self.h)
            self.h = None

    def __del__(self):
        self.close()


class Avx2Dots:
    """Sums of coefficient-domain negacyclic products through the AVX2 twisted
    NTT: the reference's Ring::dot_product / RingMat::mul_slice over
    CyclotomicPolyRing, each product a full negacyclic_mul_coeffs (CPU
    baseline of the Gram matrix, garbage.rs:76-93, and of the inner
    commitments, crs.rs:211-216)."""

    def __init__(self, q, limb_bits, d=64):
        self.d = d
        empty = np.zeros(1, dtype=np.uint64)
        self.h = lib().gr_avx2_prepare(q, limb_bits, 0, 0, d, _p(empty))
        if not self.h:
            raise OracleError(-3, "avx2 prepare")
        self.ntasks = 0

    def load(self, x, y, tx, ty):
        """x: (nx, len, d), y: (ny, len, d); output t = sum_k x[tx[t]][k] * y[ty[t]][k]."""
        x, y = _u64(x), _u64(y)
        tx = np.ascontiguousarray(tx, dtype=np.uint32)
        ty = np.ascontiguousarray(ty, dtype=np.uint32)
        self.ntasks = len(tx)
        u32p = ctypes.POINTER(ctypes.c_uint32)
        self._keep = (tx, ty)
        _check(lib().gr_avx2_dots_load(self.h, _p(x), x.shape[0], _p(y), y.shape[0], x.shape[1],
                                       tx.ctypes.data_as(u32p), ty.ctypes.data_as(u32p),
                                       self.ntasks), "avx2 dots load")

    def run(self, threads=1):
        lib().gr_avx2_dots_run(self.h, threads)

    def result(self):
        out = np.zeros((self.ntasks, self.d), dtype=np.uint64)
        lib().gr_avx2_dots_result(self.h, _p(out))
        return out

    def close(self):
        if self.h:
            lib().gr_avx2_free(self.h)
            self.h = None

    def __del__(self):
        self.close()


class Avx2Twisted:
    """The reference's twisted transforms (TwistedNttPlan, poly/ntt.rs:75-104)
    through its AVX2 NTT kernels, restated, in place over a batch of
    polynomials (CPU baseline)."""

    def __init__(self, q, limb_bits, d=64):
        self.inner = Avx2Negacyclic(q, limb_bits, d)

    def load(self, x):
        self.inner.load(x, x)

    def run(self, inverse=False, threads=1):
        lib().gr_avx2_xform_run(self.inner.h, 1 if inverse else 0, threads)

    def result(self):
        out = np.zeros((self.inner.count, self.inner.d), dtype=np.uint64)
        lib().gr_avx2_xform_result(self.inner.h, _p(out))
        return out

    def close(self):
        self.inner.close()


class Avx2Negacyclic:
    """The reference's negacyclic product through its AVX2 twisted NTT
    (poly/ring.rs:83-101), restated (CPU baseline)."""

    def __init__(self, q, limb_bits, d=64):
        self.d = d
        empty = np.zeros(1, dtype=np.uint64)
        self.h = lib().gr_avx2_prepare(q, limb_bits, 0, 0, d, _p(empty))
        if not self.h:
            raise OracleError(-3, "avx2 prepare")
        self.count = 0

    def load(self, a, b):
        a, b = _u64(a), _u64(b)
        self.count = a.shape[0]
        _check(lib().gr_avx2_negacyclic_load(self.h, _p(a), _p(b), self.count), "avx2 load")

    def run(self, threads=1):
        lib().gr_avx2_negacyclic_run(self.h, threads)

    def result(self):
        out = np.zeros((self.count, self.d), dtype=np.uint64)
        lib().gr_avx2_negacyclic_result(self.h, _p(out))
        return out

    def close(self):
        if self.h:
            lib().gr_avx2_free(self.h)
            self.h = None

    def __del__(self):
        self.close()
