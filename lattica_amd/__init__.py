"""lattica_amd — MI355X ring-arithmetic path of Gridland behind a C ABI.

The product is libgridhip.so (HIP kernels for gfx950 + the extern "C" API in
include/gridhip.h). This module is a thin ctypes binding over that ABI with the
reference's vocabulary (rings, twisted NTT, prepared commitment keys, Gram,
JL), used by the tests and the benchmark. It never computes anything itself:
if the shared library is missing, importing `lib()` raises.
"""
from __future__ import annotations

import ctypes
import os

import numpy as np

__all__ = [
    "GridHipError", "lib", "library_path", "Context", "Ring", "Key", "DeviceBuffer", "Server",
    "BatchPlan",
    "OP_COMMIT", "OP_GRAM",
    "WIRE_POLYS", "WIRE_RINGVEC", "WIRE_RINGMAT", "WIRE_VEC", "WIRE_ELEMS",
    "Q32", "GOLDILOCKS", "MEM_HOST", "MEM_DEVICE", "DOMAIN_COEFF", "DOMAIN_NTT", "STATUS",
]

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB_PATH = os.path.join(_HERE, "libgridhip.so")

# the 32-bit modulus of the benchmark configurations: largest prime < 2^32 with q = 1 mod 128
Q32 = 4294966657
GOLDILOCKS = 0xFFFFFFFF00000001

MEM_HOST, MEM_DEVICE = 0, 1
OP_COMMIT, OP_GRAM = 1, 2
DOMAIN_COEFF, DOMAIN_NTT = 0, 1
WIRE_POLYS, WIRE_RINGVEC, WIRE_RINGMAT, WIRE_VEC, WIRE_ELEMS = 0, 1, 2, 3, 4

STATUS = {
    0: "OK",
    -1: "LENGTH_MISMATCH", -2: "LENGTH_NOT_POW2", -3: "UNSUPPORTED_SIZE", -4: "UNSUPPORTED_TWIST",
    -10: "INVALID_PARAMETERS", -11: "DIMENSION_MISMATCH", -12: "OPENING_NORM_EXCEEDED",
    -13: "INVALID_MESSAGE_ENCODING", -14: "INVALID_OPENING",
    -20: "INVALID_ARGUMENT", -21: "NOT_CANONICAL", -22: "UNSUPPORTED_MODULUS", -23: "DEVICE",
    -24: "OOM", -25: "BUSY", -26: "NO_DEVICE", -30: "UNEXPECTED_END", -31: "INVALID_DATA",
}

# every symbol include/gridhip.h declares
EXPORTS = (
    "gh_status_string", "gh_version", "gh_device_count", "gh_ctx_create", "gh_ctx_destroy",
    "gh_ctx_synchronize", "gh_ctx_num_cus", "gh_ctx_profile_enable", "gh_ctx_profile_read",
    "gh_ctx_set_tuning", "gh_ctx_get_tuning", "gh_measure_bandwidth", "gh_alloc", "gh_free",
    "gh_upload", "gh_download",
    "gh_check_canonical", "gh_ring_create", "gh_ring_destroy", "gh_ring_query", "gh_ntt_forward",
    "gh_ntt_inverse", "gh_twisted_forward", "gh_twisted_inverse", "gh_negacyclic_mul",
    "gh_pointwise_mul", "gh_key_create", "gh_key_destroy", "gh_key_dims", "gh_commit",
    "gh_commit_batch", "gh_gram", "gh_inner_commit", "gh_jl_project",
    "gh_server_create", "gh_server_destroy", "gh_server_submit", "gh_server_wait",
    "gh_server_poll", "gh_server_query", "gh_batch_plan_create", "gh_batch_plan_launch",
    "gh_batch_plan_destroy", "gh_bdlop_commit", "gh_decompose_balanced", "gh_garbage_h",
    "gh_amortize", "gh_wire_size", "gh_wire_encode", "gh_wire_peek", "gh_wire_decode",
    "gh_key_save", "gh_key_load", "gh_jl_extract_flat", "gh_jl_conjugate", "gh_jl_evals",
    "gh_jl_aggregate", "gh_jl_launch_shape", "gh_commit_noalloc", "gh_measure_issue",
    "gh_add_assign", "gh_sub_assign", "gh_scalar_mul", "gh_add_assign_scaled", "gh_dot_product",
    "gh_measure_valu",
)


class GridHipError(RuntimeError):
    def __init__(self, code: int, what: str = ""):
        name = STATUS.get(code, str(code))
        super().__init__(f"{what}: GH_E_{name} ({code})" if what else f"GH_E_{name} ({code})")
        self.code = code
        self.name = name


class RingInfo(ctypes.Structure):
    _fields_ = [
        ("q", ctypes.c_uint64), ("d", ctypes.c_uint32), ("word_bytes", ctypes.c_uint32),
        ("kind", ctypes.c_uint32), ("wire_bytes", ctypes.c_uint32), ("generator", ctypes.c_uint64),
        ("psi", ctypes.c_uint64), ("omega", ctypes.c_uint64), ("d_inv", ctypes.c_uint64),
    ]


class CommitItem(ctypes.Structure):
    _fields_ = [("witness", ctypes.c_void_p), ("len", ctypes.c_uint32),
                ("reserved", ctypes.c_uint32), ("t_out", ctypes.c_void_p)]


class ServerConfig(ctypes.Structure):
    _fields_ = [("mem_bytes", ctypes.c_uint64), ("max_inflight", ctypes.c_uint32),
                ("max_wave", ctypes.c_uint32), ("slots", ctypes.c_uint32),
                ("max_pending", ctypes.c_uint32)]


class Request(ctypes.Structure):
    _fields_ = [("ops", ctypes.c_uint32), ("r", ctypes.c_uint32), ("n", ctypes.c_uint32),
                ("mem", ctypes.c_int32), ("ring", ctypes.c_void_p), ("key", ctypes.c_void_p),
                ("witness", ctypes.c_void_p), ("t_out", ctypes.c_void_p),
                ("g_out", ctypes.c_void_p), ("wire", ctypes.c_uint32),
                ("reserved", ctypes.c_uint32), ("witness_len", ctypes.c_uint64),
                ("t_cap", ctypes.c_uint64), ("g_cap", ctypes.c_uint64)]


class JlShape(ctypes.Structure):
    _fields_ = [("chunk_polys", ctypes.c_uint64), ("per_unit", ctypes.c_uint64),
                ("blocks", ctypes.c_uint64), ("chunks", ctypes.c_uint64),
                ("terms", ctypes.c_uint64), ("bound", ctypes.c_uint64)]


JL_KERNEL_PROJECT, JL_KERNEL_EVALS = 0, 1


def jl_launch_shape(kernel: int, total_polys: int, num_cus: int) -> dict:
    """Launch shape and int32 bound of a matrix-core JL kernel (host arithmetic)."""
    s = JlShape()
    _check(lib().gh_jl_launch_shape(kernel, total_polys, num_cus, ctypes.byref(s)),
           "gh_jl_launch_shape")
    return {k: getattr(s, k) for k, _ in JlShape._fields_}


class ServerStats(ctypes.Structure):
    _fields_ = [("submitted", ctypes.c_uint64), ("completed", ctypes.c_uint64),
                ("failed", ctypes.c_uint64), ("waves", ctypes.c_uint64),
                ("commits", ctypes.c_uint64), ("grams", ctypes.c_uint64),
                ("mem_in_use", ctypes.c_uint64), ("mem_peak", ctypes.c_uint64),
                ("pending", ctypes.c_uint32), ("admitted", ctypes.c_uint32),
                ("max_wave_seen", ctypes.c_uint32), ("reserved", ctypes.c_uint32),
                ("device_ms", ctypes.c_double)]


_lib = None


def library_path() -> str:
    return _LIB_PATH


def lib():
    """Load libgridhip.so. Raises if it has not been built: there is no fallback."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(_LIB_PATH):
        raise RuntimeError(
            f"{_LIB_PATH} is missing: build it with `python -m lattica_amd.build` "
            "(no CPU fallback exists by design)")
    L = ctypes.CDLL(_LIB_PATH)
    vp, i32, u32, u64, sz = ctypes.c_void_p, ctypes.c_int, ctypes.c_uint32, ctypes.c_uint64, ctypes.c_size_t
    sig = {
        "gh_status_string": (ctypes.c_char_p, [i32]),
        "gh_version": (i32, []),
        "gh_device_count": (i32, [ctypes.POINTER(i32)]),
        "gh_ctx_create": (i32, [i32, ctypes.POINTER(vp)]),
        "gh_ctx_destroy": (i32, [vp]),
        "gh_ctx_synchronize": (i32, [vp]),
        "gh_ctx_num_cus": (i32, [vp, ctypes.POINTER(i32)]),
        "gh_ctx_profile_enable": (i32, [vp, i32]),
        "gh_ctx_set_tuning": (i32, [vp, ctypes.c_char_p, ctypes.c_int64]),
        "gh_measure_issue": (i32, [vp, i32, ctypes.POINTER(ctypes.c_double)]),
        "gh_measure_valu": (i32, [vp, i32, ctypes.POINTER(ctypes.c_double)]),
        "gh_measure_bandwidth": (i32, [vp, sz, i32, ctypes.POINTER(ctypes.c_double),
                                       ctypes.POINTER(ctypes.c_double)]),
        "gh_ctx_get_tuning": (i32, [vp, ctypes.c_char_p, ctypes.POINTER(ctypes.c_int64)]),
        "gh_ctx_profile_read": (i32, [vp, i32, ctypes.POINTER(ctypes.c_double), ctypes.POINTER(u64)]),
        "gh_alloc": (i32, [vp, sz, ctypes.POINTER(vp)]),
        "gh_free": (i32, [vp, vp]),
        "gh_upload": (i32, [vp, vp, vp, sz]),
        "gh_download": (i32, [vp, vp, vp, sz]),
        "gh_check_canonical": (i32, [vp, vp, sz, i32]),
        "gh_ring_create": (i32, [vp, u64, u32, ctypes.POINTER(vp)]),
        "gh_ring_destroy": (i32, [vp]),
        "gh_ring_query": (i32, [vp, ctypes.POINTER(RingInfo)]),
        "gh_ntt_forward": (i32, [vp, vp, sz, i32]),
        "gh_ntt_inverse": (i32, [vp, vp, sz, i32]),
        "gh_twisted_forward": (i32, [vp, vp, sz, i32]),
        "gh_twisted_inverse": (i32, [vp, vp, sz, i32]),
        "gh_negacyclic_mul": (i32, [vp, vp, vp, vp, sz, i32]),
        "gh_pointwise_mul": (i32, [vp, vp, vp, vp, sz, i32]),
        "gh_add_assign": (i32, [vp, vp, vp, sz, i32]),
        "gh_sub_assign": (i32, [vp, vp, vp, sz, i32]),
        "gh_scalar_mul": (i32, [vp, vp, u64, sz, i32]),
        "gh_add_assign_scaled": (i32, [vp, vp, vp, u64, sz, i32]),
        "gh_dot_product": (i32, [vp, vp, vp, sz, vp, i32]),
        "gh_key_create": (i32, [vp, u32, vp, u32, vp, u32, i32, i32, ctypes.POINTER(vp)]),
        "gh_key_destroy": (i32, [vp]),
        "gh_key_dims": (i32, [vp, ctypes.POINTER(u32), ctypes.POINTER(u32), ctypes.POINTER(u32)]),
        "gh_commit": (i32, [vp, vp, u32, vp, u32, i32, vp, i32]),
        "gh_commit_noalloc": (i32, [vp, vp, u32, vp]),
        "gh_commit_batch": (i32, [vp, ctypes.POINTER(CommitItem), sz, i32]),
        "gh_gram": (i32, [vp, vp, u32, u32, i32, vp, i32]),
        "gh_inner_commit": (i32, [vp, vp, u32, u32, i32, vp, i32]),
        "gh_jl_project": (i32, [vp, vp, u32, vp, u32, u64, vp, vp, vp, i32]),
        "gh_server_create": (i32, [vp, ctypes.POINTER(ServerConfig), ctypes.POINTER(vp)]),
        "gh_server_destroy": (i32, [vp]),
        "gh_server_submit": (i32, [vp, ctypes.POINTER(Request), ctypes.POINTER(u64)]),
        "gh_server_wait": (i32, [vp, u64, ctypes.POINTER(i32), ctypes.POINTER(ctypes.c_double)]),
        "gh_server_poll": (i32, [vp, u64, ctypes.POINTER(i32), ctypes.POINTER(i32),
                                 ctypes.POINTER(ctypes.c_double)]),
        "gh_server_query": (i32, [vp, ctypes.POINTER(ServerStats)]),
        "gh_batch_plan_create": (i32, [vp, ctypes.POINTER(CommitItem), sz, ctypes.POINTER(vp)]),
        "gh_batch_plan_launch": (i32, [vp]),
        "gh_batch_plan_destroy": (i32, [vp]),
        "gh_bdlop_commit": (i32, [vp, vp, vp, u32, vp, u32, i32, vp, vp, i32]),
        "gh_decompose_balanced": (i32, [vp, vp, sz, u64, u32, vp, i32]),
        "gh_garbage_h": (i32, [vp, vp, vp, u32, u32, i32, vp, i32]),
        "gh_amortize": (i32, [vp, vp, vp, u32, u32, i32, vp, i32]),
        "gh_wire_size": (i32, [vp, i32, u64, u64, ctypes.POINTER(u64)]),
        "gh_wire_encode": (i32, [vp, i32, u64, u64, vp, i32, vp, u64, ctypes.POINTER(u64)]),
        "gh_wire_peek": (i32, [vp, i32, vp, u64, ctypes.POINTER(u64), ctypes.POINTER(u64),
                               ctypes.POINTER(u64)]),
        "gh_wire_decode": (i32, [vp, i32, i32, vp, u64, u64, u64, vp, i32, ctypes.POINTER(u64)]),
        "gh_key_save": (i32, [vp, i32, vp, u64, ctypes.POINTER(u64)]),
        "gh_key_load": (i32, [vp, vp, u64, i32, ctypes.POINTER(u64), ctypes.POINTER(vp)]),
        "gh_jl_extract_flat": (i32, [vp, vp, u32, u32, u64, vp, i32]),
        "gh_jl_conjugate": (i32, [vp, vp, vp, u32, u32, u64, vp, i32]),
        "gh_jl_evals": (i32, [vp, vp, u32, vp, u32, u64, vp, i32]),
        "gh_jl_aggregate": (i32, [vp, vp, u32, vp, u32, u64, vp, i32]),
        "gh_jl_launch_shape": (i32, [i32, u64, i32, ctypes.POINTER(JlShape)]),
    }
    for name, (res, args) in sig.items():
        fn = getattr(L, name)
        fn.restype = res
        fn.argtypes = args
    _lib = L
    return L


def _check(code: int, what: str):
    if code != 0:
        raise GridHipError(code, what)


def _ptr(a: np.ndarray):
    return ctypes.c_void_p(a.ctypes.data)


def device_count() -> int:
    n = ctypes.c_int(0)
    _check(lib().gh_device_count(ctypes.byref(n)), "gh_device_count")
    return n.value


class Context:
    """gh_ctx: one device, one stream, grow-only workspace."""

    def __init__(self, device: int = 0, apply_tuning: bool = True):
        h = ctypes.c_void_p()
        _check(lib().gh_ctx_create(device, ctypes.byref(h)), "gh_ctx_create")
        self.h = h
        self.device = device
        self.tuning_source = None
        if apply_tuning:
            self._apply_tuning_file()

    def _apply_tuning_file(self):
        """Apply lattica_amd/tuning.json (written by lattica_amd.autotune) if present."""
        import json
        path = os.path.join(_HERE, "tuning.json")
        if not os.path.exists(path):
            return
        with open(path) as fh:
            doc = json.load(fh)
        if doc.get("device_cus") not in (None, self.num_cus):
            return  # tuned on a different device
        self.set_tuning(**doc.get("best", {}))
        self.tuning_source = path

    @property
    def num_cus(self) -> int:
        n = ctypes.c_int()
        _check(lib().gh_ctx_num_cus(self.h, ctypes.byref(n)), "gh_ctx_num_cus")
        return n.value

    def synchronize(self):
        _check(lib().gh_ctx_synchronize(self.h), "gh_ctx_synchronize")

    TUNING_KEYS = ("mac_variant", "mac_waves", "mac_blocks_per_cu", "quad_max", "fused", "fused_depth",
                   "batch_chunk_cols", "batch_chunk_cols_gold", "batch_group", "defer_finalize",
                   "stage1_tasks_per_block", "batch_fused", "batch_fused_cols", "batch_mac_ring",
                   "negacyclic_quad", "q32_ring_group", "quad_lds", "xform_quad")

    def set_tuning(self, **kv):
        for k, v in kv.items():
            _check(lib().gh_ctx_set_tuning(self.h, k.encode(), int(v)), f"gh_ctx_set_tuning({k})")

    def get_tuning(self) -> dict:
        out = {}
        for k in self.TUNING_KEYS:
            v = ctypes.c_int64()
            _check(lib().gh_ctx_get_tuning(self.h, k.encode(), ctypes.byref(v)), "gh_ctx_get_tuning")
            out[k] = v.value
        return out

    def measure_bandwidth(self, nbytes: int = 1 << 30, reps: int = 5):
        """(read GB/s, copy GB/s) of streaming kernels over nbytes of HBM."""
        r, c = ctypes.c_double(), ctypes.c_double()
        _check(lib().gh_measure_bandwidth(self.h, nbytes, reps, ctypes.byref(r), ctypes.byref(c)),
               "gh_measure_bandwidth")
        return r.value, c.value

    VALU_KINDS = ("v_add_u32", "v_mul_hi_u32", "v_mul_lo_u32", "v_mad_u64_u32",
                  "v_lshl_add_u64", "v_fma_f64", "sub_fix_3op", "v_cndmask_b32",
                  "v_xor_b32", "v_alignbit_b32")

    def measure_valu(self, waves_per_simd: int = 4) -> dict:
        """gh_measure_valu: G wave instructions per second of single VALU
        instruction kinds, and each against v_add_u32."""
        g = (ctypes.c_double * len(self.VALU_KINDS))()
        _check(lib().gh_measure_valu(self.h, waves_per_simd, g), "gh_measure_valu")
        rates = {k: round(g[i], 1) for i, k in enumerate(self.VALU_KINDS)}
        return {"waves_per_simd": waves_per_simd, "G_wave_inst_per_s": rates,
                "slots_vs_add": {k: round(g[0] / g[i], 2) for i, k in enumerate(self.VALU_KINDS)
                                 if g[i] > 0}}

    def measure_issue(self, waves_per_simd: int = 2) -> dict:
        """gh_measure_issue: microseconds per probe launch, the int8 MFMA
        rates (TOP/s, two ops per multiply-add), the VALU instruction rate
        (wave instructions per second) and the mixed loop's time against the
        sum and the larger of its two parts."""
        us = (ctypes.c_double * 4)()
        _check(lib().gh_measure_issue(self.h, waves_per_simd, us), "gh_measure_issue")
        waves = self.num_cus * waves_per_simd * 4
        iters = 4096
        ops_mfma = waves * iters * 4 * 2 * 32 * 32 * 32  # int8 ops (2 per multiply-add)
        valu = waves * iters * 48                           # wave-level VALU instructions
        return {"waves_per_simd": waves_per_simd,
                "us": {"mfma_32x32x32": us[0], "mfma_16x16x64": us[1], "valu_chacha": us[2],
                       "mixed": us[3]},
                "mfma_32x32x32_tops": ops_mfma / (us[0] * 1e-6) / 1e12,
                "mfma_16x16x64_tops": ops_mfma / (us[1] * 1e-6) / 1e12,
                "valu_ginst_per_s": valu / (us[2] * 1e-6) / 1e9,
                "mixed_over_sum": us[3] / (us[0] + us[2]),
                "mixed_over_max": us[3] / max(us[0], us[2])}

    def profile(self, on: bool):
        _check(lib().gh_ctx_profile_enable(self.h, 1 if on else 0), "gh_ctx_profile_enable")

    def profile_read(self):
        """{phase: (total_ms, count)} for phases 0 (witness NTT), 1 (MAC), 2 (finalize)."""
        out = {}
        for ph in (0, 1, 2):
            ms, n = ctypes.c_double(), ctypes.c_uint64()
            _check(lib().gh_ctx_profile_read(self.h, ph, ctypes.byref(ms), ctypes.byref(n)),
                   "gh_ctx_profile_read")
            out[ph] = (ms.value, n.value)
        return out

    def alloc(self, nbytes: int) -> "DeviceBuffer":
        return DeviceBuffer(self, nbytes)

    def to_device(self, arr: np.ndarray) -> "DeviceBuffer":
        arr = np.ascontiguousarray(arr)
        buf = DeviceBuffer(self, arr.nbytes)
        buf.upload(arr)
        return buf

    def close(self):
        if self.h:
            lib().gh_ctx_destroy(self.h)
            self.h = None

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()


class DeviceBuffer:
    def __init__(self, ctx: Context, nbytes: int):
        self.ctx = ctx
        self.nbytes = int(nbytes)
        p = ctypes.c_void_p()
        _check(lib().gh_alloc(ctx.h, self.nbytes, ctypes.byref(p)), "gh_alloc")
        self.ptr = p

    def upload(self, arr: np.ndarray):
        arr = np.ascontiguousarray(arr)
        assert arr.nbytes <= self.nbytes
        _check(lib().gh_upload(self.ctx.h, self.ptr, _ptr(arr), arr.nbytes), "gh_upload")

    def download(self, dtype, shape) -> np.ndarray:
        out = np.empty(shape, dtype=dtype)
        assert out.nbytes <= self.nbytes
        _check(lib().gh_download(self.ctx.h, _ptr(out), self.ptr, out.nbytes), "gh_download")
        return out

    def offset(self, nbytes: int) -> ctypes.c_void_p:
        return ctypes.c_void_p(self.ptr.value + int(nbytes))

    def free(self):
        if self.ptr:
            lib().gh_free(self.ctx.h, self.ptr)
            self.ptr = None


class Ring:
    """gh_ring: Z_q[X]/(X^d + 1) with the reference's root selection."""

    def __init__(self, ctx: Context, q: int, d: int = 64):
        h = ctypes.c_void_p()
        _check(lib().gh_ring_create(ctx.h, q, d, ctypes.byref(h)), "gh_ring_create")
        self.h = h
        self.ctx = ctx
        info = RingInfo()
        _check(lib().gh_ring_query(h, ctypes.byref(info)), "gh_ring_query")
        self.info = info
        self.q, self.d = int(info.q), int(info.d)
        self.dtype = np.uint32 if info.word_bytes == 4 else np.uint64

    def words(self, a) -> np.ndarray:
        return np.ascontiguousarray(np.asarray(a).astype(self.dtype, copy=False))

    def _inplace(self, fn, values):
        v = self.words(values).copy()
        _check(getattr(lib(), fn)(self.h, _ptr(v), v.size // self.d, MEM_HOST), fn)
        return v

    def ntt_forward(self, values):
        return self._inplace("gh_ntt_forward", values)

    def ntt_inverse(self, values):
        return self._inplace("gh_ntt_inverse", values)

    def twisted_forward(self, values):
        return self._inplace("gh_twisted_forward", values)

    def twisted_inverse(self, values):
        return self._inplace("gh_twisted_inverse", values)

    def negacyclic_mul(self, a, b):
        a, b = self.words(a), self.words(b)
        out = np.empty_like(a)
        _check(lib().gh_negacyclic_mul(self.h, _ptr(a), _ptr(b), _ptr(out), a.size // self.d,
                                       MEM_HOST), "gh_negacyclic_mul")
        return out

    def pointwise_mul(self, a, b):
        a, b = self.words(a), self.words(b)
        out = np.empty_like(a)
        _check(lib().gh_pointwise_mul(self.h, _ptr(a), _ptr(b), _ptr(out), a.size // self.d,
                                      MEM_HOST), "gh_pointwise_mul")
        return out

    # Ring slice hooks on flat word arrays (any length); each returns a new array
    def add_assign(self, dst, src):
        d, s = self.words(dst).copy(), self.words(src)
        _check(lib().gh_add_assign(self.h, _ptr(d), _ptr(s), d.size, MEM_HOST), "gh_add_assign")
        return d

    def sub_assign(self, dst, src):
        d, s = self.words(dst).copy(), self.words(src)
        _check(lib().gh_sub_assign(self.h, _ptr(d), _ptr(s), d.size, MEM_HOST), "gh_sub_assign")
        return d

    def scalar_mul(self, dst, scalar: int):
        d = self.words(dst).copy()
        _check(lib().gh_scalar_mul(self.h, _ptr(d), int(scalar), d.size, MEM_HOST), "gh_scalar_mul")
        return d

    def add_assign_scaled(self, dst, src, scalar: int):
        d, s = self.words(dst).copy(), self.words(src)
        _check(lib().gh_add_assign_scaled(self.h, _ptr(d), _ptr(s), int(scalar), d.size, MEM_HOST),
               "gh_add_assign_scaled")
        return d

    def dot_product(self, a, b) -> int:
        a, b = self.words(a), self.words(b)
        out = np.zeros(1, dtype=self.dtype)
        _check(lib().gh_dot_product(self.h, _ptr(a), _ptr(b), a.size, _ptr(out), MEM_HOST),
               "gh_dot_product")
        return int(out[0])

    def gram(self, s, domain=DOMAIN_COEFF):
        s = self.words(s)
        r, n = s.shape[0], s.shape[1]
        out = np.zeros((r * (r + 1) // 2, self.d), dtype=self.dtype)
        _check(lib().gh_gram(self.h, _ptr(s), r, n, domain, _ptr(out), MEM_HOST), "gh_gram")
        return out

    def decompose_balanced(self, polys, base: int, limbs: int):
        """(m, d) -> (m, limbs, d): gh_decompose_balanced."""
        v = self.words(polys)
        out = np.zeros((v.shape[0], limbs, self.d), dtype=self.dtype)
        _check(lib().gh_decompose_balanced(self.h, _ptr(v), v.shape[0], base, limbs, _ptr(out),
                                           MEM_HOST), "gh_decompose_balanced")
        return out

    def garbage_h(self, phi, s, domain=DOMAIN_COEFF):
        """(r, n, d) x (r, n, d) -> (r(r+1)/2, d): gh_garbage_h."""
        phi, s = self.words(phi), self.words(s)
        r, n = s.shape[0], s.shape[1]
        out = np.zeros((r * (r + 1) // 2, self.d), dtype=self.dtype)
        _check(lib().gh_garbage_h(self.h, _ptr(phi), _ptr(s), r, n, domain, _ptr(out), MEM_HOST),
               "gh_garbage_h")
        return out

    def amortize(self, c, s, domain=DOMAIN_COEFF):
        """c: (r, d), s: (r, n, d) -> z: (n, d) = sum_i c_i * s_i: gh_amortize."""
        c, s = self.words(c), self.words(s)
        r, n = s.shape[0], s.shape[1]
        out = np.zeros((n, self.d), dtype=self.dtype)
        _check(lib().gh_amortize(self.h, _ptr(c), _ptr(s), r, n, domain, _ptr(out), MEM_HOST),
               "gh_amortize")
        return out

    def jl_project(self, seed: bytes, rows: int, parts, want_exact=True, want_flat=False):
        parts = self.words(parts)
        num_parts, cols = parts.shape
        sd = np.frombuffer(bytes(seed), dtype=np.uint8).copy()
        pm = np.zeros(rows, dtype=self.dtype)
        pe = np.zeros(rows, dtype=np.int64) if want_exact else None
        fl = np.zeros(rows * num_parts * cols, dtype=np.int8) if want_flat else None
        _check(lib().gh_jl_project(self.h, _ptr(sd), rows, _ptr(parts), num_parts, cols, _ptr(pm),
                                   _ptr(pe) if pe is not None else None,
                                   _ptr(fl) if fl is not None else None, MEM_HOST),
               "gh_jl_project")
        return pm, pe, fl

    # ---- JL rows for aggregation ----
    def jl_extract_flat(self, seed: bytes, rows: int, num_parts: int, cols: int) -> np.ndarray:
        """gh_jl_extract_flat: int8 entries, [m][part][coeff][poly]."""
        sd = np.frombuffer(bytes(seed), dtype=np.uint8).copy()
        out = np.zeros(rows * num_parts * cols, dtype=np.int8)
        _check(lib().gh_jl_extract_flat(self.h, _ptr(sd), rows, num_parts, cols,
                                        _ptr(out) if out.size else None, MEM_HOST),
               "gh_jl_extract_flat")
        return out

    def jl_conjugate(self, rows: int, num_parts: int, cols: int, flat=None, seed: bytes = None):
        """gh_jl_conjugate: (rows, num_parts, cols / d, d) conjugated JL polynomials,
        from a flat int8 array or straight from the seed."""
        n = cols // self.d
        out = np.zeros((rows, num_parts, n, self.d), dtype=self.dtype)
        sd = np.frombuffer(bytes(seed), dtype=np.uint8).copy() if seed is not None else None
        fl = np.ascontiguousarray(flat, dtype=np.int8) if flat is not None else None
        _check(lib().gh_jl_conjugate(self.h, _ptr(sd) if sd is not None else None,
                                     _ptr(fl) if fl is not None else None, rows, num_parts, cols,
                                     _ptr(out) if out.size else None, MEM_HOST),
               "gh_jl_conjugate")
        return out

    def jl_evals(self, seed: bytes, parts, rows: int = 256) -> np.ndarray:
        """gh_jl_evals: parts (num_parts, cols) or (num_parts, n, d) -> (rows, d)."""
        p = self.words(parts)
        num_parts = p.shape[0]
        cols = p.size // max(1, num_parts)
        sd = np.frombuffer(bytes(seed), dtype=np.uint8).copy()
        out = np.zeros((rows, self.d), dtype=self.dtype)
        _check(lib().gh_jl_evals(self.h, _ptr(sd), rows, _ptr(p) if p.size else None, num_parts,
                                 cols, _ptr(out), MEM_HOST), "gh_jl_evals")
        return out

    def jl_aggregate(self, seed: bytes, scale, phi, rows: int = 256) -> np.ndarray:
        """gh_jl_aggregate: phi (num_parts, n, d) + sum_m scale_m * sigma(pi^(m));
        scale (rows, d). Returns the new phi; the argument is not modified."""
        sc = self.words(scale)
        ph = self.words(phi).copy()
        num_parts = ph.shape[0] if ph.ndim == 3 else 1
        cols = ph.size // max(1, num_parts)
        sd = np.frombuffer(bytes(seed), dtype=np.uint8).copy()
        _check(lib().gh_jl_aggregate(self.h, _ptr(sd), rows, _ptr(sc), num_parts, cols,
                                     _ptr(ph) if ph.size else None, MEM_HOST), "gh_jl_aggregate")
        return ph

    # ---- canonical wire formats (gh_wire_*) ----
    def wire_size(self, kind: int, rows: int, cols: int = 1) -> int:
        out = ctypes.c_uint64()
        _check(lib().gh_wire_size(self.h, kind, rows, cols, ctypes.byref(out)), "gh_wire_size")
        return out.value

    def wire_encode(self, kind: int, items, rows: int = None, cols: int = 1) -> bytes:
        """Ring words (..., d) or, for WIRE_ELEMS, (rows,) -> encoded bytes."""
        v = self.words(items)
        if rows is None:
            rows = v.size // (1 if kind == WIRE_ELEMS else self.d) // cols
        size = self.wire_size(kind, rows, cols)
        out = np.zeros(max(1, size), dtype=np.uint8)
        wr = ctypes.c_uint64()
        _check(lib().gh_wire_encode(self.h, kind, rows, cols, _ptr(v) if v.size else None,
                                    MEM_HOST, _ptr(out), size, ctypes.byref(wr)),
               "gh_wire_encode")
        return out[:wr.value].tobytes()

    def wire_peek(self, kind: int, data: bytes, rows: int = 0):
        """-> (rows, cols, total_bytes) from the header."""
        buf = np.frombuffer(data, dtype=np.uint8) if len(data) else np.zeros(1, np.uint8)
        r, c, t = ctypes.c_uint64(rows), ctypes.c_uint64(1), ctypes.c_uint64()
        _check(lib().gh_wire_peek(self.h, kind, _ptr(buf), len(data), ctypes.byref(r),
                                  ctypes.byref(c), ctypes.byref(t)), "gh_wire_peek")
        return r.value, c.value, t.value

    def wire_decode(self, kind: int, data: bytes, domain=DOMAIN_COEFF, rows: int = None,
                    cols: int = None, exact: bool = True):
        """Encoded bytes -> ring words, shape (rows*cols, d) or (rows,) for WIRE_ELEMS.
        Returns (words, consumed); exact=True asks for deserialize_exact."""
        if rows is None or cols is None:
            pr, pc, _ = self.wire_peek(kind, data, rows or 0)
            rows = pr if rows is None else rows
            cols = pc if cols is None else cols
        per = 1 if kind == WIRE_ELEMS else self.d
        out = np.zeros((rows * cols * per,), dtype=self.dtype)
        buf = np.frombuffer(data, dtype=np.uint8) if len(data) else np.zeros(1, np.uint8)
        used = ctypes.c_uint64()
        _check(lib().gh_wire_decode(self.h, kind, domain, _ptr(buf), len(data), rows, cols,
                                    _ptr(out) if out.size else None, MEM_HOST,
                                    None if exact else ctypes.byref(used)), "gh_wire_decode")
        shape = (rows,) if kind == WIRE_ELEMS else (rows * cols, self.d)
        return out.reshape(shape), (len(data) if exact else used.value)

    def load_key(self, data: bytes, domain=DOMAIN_NTT, exact: bool = True) -> "Key":
        """gh_key_load: a resident key from LinearCommitmentKey bytes."""
        return Key.load(self, data, domain, exact)

    def close(self):
        if self.h:
            lib().gh_ring_destroy(self.h)
            self.h = None


class Key:
    """gh_key: resident NTT-domain commitment key A = [A_msg | A_open]."""

    def __init__(self, ring: Ring, a_msg, a_open=None, domain=DOMAIN_COEFF):
        a_msg = ring.words(a_msg)
        rows, msg_len = a_msg.shape[0], a_msg.shape[1]
        if a_open is None:
            a_open_arr, open_len = None, 0
        else:
            a_open_arr = ring.words(a_open)
            open_len = a_open_arr.shape[1]
        h = ctypes.c_void_p()
        _check(lib().gh_key_create(ring.h, rows, _ptr(a_msg), msg_len,
                                   _ptr(a_open_arr) if a_open_arr is not None else None, open_len,
                                   domain, MEM_HOST, ctypes.byref(h)), "gh_key_create")
        self.h = h
        self.ring = ring
        self.rows, self.msg_len, self.open_len = rows, msg_len, open_len

    @classmethod
    def from_device(cls, ring: Ring, rows: int, a_msg_dev: DeviceBuffer, msg_len: int,
                    domain=DOMAIN_NTT):
        self = cls.__new__(cls)
        h = ctypes.c_void_p()
        _check(lib().gh_key_create(ring.h, rows, a_msg_dev.ptr, msg_len, None, 0, domain,
                                   MEM_DEVICE, ctypes.byref(h)), "gh_key_create")
        self.h, self.ring = h, ring
        self.rows, self.msg_len, self.open_len = rows, msg_len, 0
        return self

    def commit(self, message, randomness=None, domain=DOMAIN_COEFF):
        ring = self.ring
        m = ring.words(message)
        r = ring.words(randomness) if randomness is not None else None
        out = np.zeros((self.rows, ring.d), dtype=ring.dtype)
        _check(lib().gh_commit(self.h, _ptr(m), m.shape[0], _ptr(r) if r is not None else None,
                               r.shape[0] if r is not None else 0, domain, _ptr(out), MEM_HOST),
               "gh_commit")
        return out

    def commit_device(self, s_dev: DeviceBuffer, n: int, t_dev: DeviceBuffer, domain=DOMAIN_COEFF):
        _check(lib().gh_commit(self.h, s_dev.ptr, n, None, 0, domain, t_dev.ptr, MEM_DEVICE),
               "gh_commit")

    def commit_noalloc(self, s_dev, n: int, t_dev):
        """gh_commit_noalloc: the device commitment with its partial sums in the
        library's static per-device array; s_dev and t_dev are DeviceBuffers
        or raw device addresses."""
        sp = s_dev.ptr if isinstance(s_dev, DeviceBuffer) else ctypes.c_void_p(s_dev)
        tp = t_dev.ptr if isinstance(t_dev, DeviceBuffer) else ctypes.c_void_p(t_dev)
        _check(lib().gh_commit_noalloc(self.h, sp, n, tp), "gh_commit_noalloc")

    def commit_batch(self, witnesses):
        ring = self.ring
        ws = [ring.words(w) for w in witnesses]
        outs = [np.zeros((self.rows, ring.d), dtype=ring.dtype) for _ in ws]
        items = (CommitItem * max(1, len(ws)))()
        for i, (w, o) in enumerate(zip(ws, outs)):
            items[i].witness = w.ctypes.data
            items[i].len = w.shape[0]
            items[i].t_out = o.ctypes.data
        _check(lib().gh_commit_batch(self.h, items, len(ws), MEM_HOST), "gh_commit_batch")
        return outs

    def commit_batch_device(self, witnesses, lens, outs):
        """Device-memory batch: witnesses/outs are DeviceBuffers (or raw addresses)."""
        items = (CommitItem * max(1, len(lens)))()
        for i, (w, n, o) in enumerate(zip(witnesses, lens, outs)):
            items[i].witness = w.ptr.value if isinstance(w, DeviceBuffer) else int(w)
            items[i].len = int(n)
            items[i].t_out = o.ptr.value if isinstance(o, DeviceBuffer) else int(o)
        _check(lib().gh_commit_batch(self.h, items, len(lens), MEM_DEVICE), "gh_commit_batch")

    def bdlop(self, mask: "Key", message, randomness, domain=DOMAIN_COEFF):
        """gh_bdlop_commit with this key as [A_msg | A_open]: returns (u, v)."""
        ring = self.ring
        m, r = ring.words(message), ring.words(randomness)
        u = np.zeros((mask.rows, ring.d), dtype=ring.dtype)
        v = np.zeros((self.rows, ring.d), dtype=ring.dtype)
        _check(lib().gh_bdlop_commit(mask.h, self.h, _ptr(m), m.shape[0], _ptr(r), r.shape[0],
                                     domain, _ptr(u), _ptr(v), MEM_HOST), "gh_bdlop_commit")
        return u, v

    def batch_plan(self, witnesses, lens, outs) -> "BatchPlan":
        """gh_batch_plan over device-memory items (DeviceBuffers or addresses)."""
        return BatchPlan(self, witnesses, lens, outs)

    def save(self, domain=DOMAIN_NTT) -> bytes:
        """gh_key_save: LinearCommitmentKey bytes of this key in `domain`."""
        size = ctypes.c_uint64()
        _check(lib().gh_key_save(self.h, domain, None, 0, ctypes.byref(size)), "gh_key_save")
        out = np.zeros(max(1, size.value), dtype=np.uint8)
        _check(lib().gh_key_save(self.h, domain, _ptr(out), size.value, ctypes.byref(size)),
               "gh_key_save")
        return out[:size.value].tobytes()

    @classmethod
    def load(cls, ring: Ring, data: bytes, domain=DOMAIN_NTT, exact: bool = True) -> "Key":
        buf = np.frombuffer(data, dtype=np.uint8) if len(data) else np.zeros(1, np.uint8)
        h = ctypes.c_void_p()
        used = ctypes.c_uint64()
        _check(lib().gh_key_load(ring.h, _ptr(buf), len(data), domain,
                                 None if exact else ctypes.byref(used), ctypes.byref(h)),
               "gh_key_load")
        self = cls.__new__(cls)
        self.h, self.ring = h, ring
        rows, ml, ol = ctypes.c_uint32(), ctypes.c_uint32(), ctypes.c_uint32()
        _check(lib().gh_key_dims(h, ctypes.byref(rows), ctypes.byref(ml), ctypes.byref(ol)),
               "gh_key_dims")
        self.rows, self.msg_len, self.open_len = rows.value, ml.value, ol.value
        self.consumed = len(data) if exact else used.value
        return self

    def inner_commit(self, s, domain=DOMAIN_COEFF):
        ring = self.ring
        s = ring.words(s)
        r, n = s.shape[0], s.shape[1]
        out = np.zeros((r, self.rows, ring.d), dtype=ring.dtype)
        _check(lib().gh_inner_commit(self.h, _ptr(s), r, n, domain, _ptr(out), MEM_HOST),
               "gh_inner_commit")
        return out

    def close(self):
        if self.h:
            lib().gh_key_destroy(self.h)
            self.h = None


class BatchPlan:
    """gh_batch_plan: a fixed device-memory batch replayed as one graph launch."""

    def __init__(self, key: "Key", witnesses, lens, outs):
        items = (CommitItem * max(1, len(lens)))()
        for i, (w, n, o) in enumerate(zip(witnesses, lens, outs)):
            items[i].witness = _addr(w)
            items[i].len = int(n)
            items[i].t_out = _addr(o)
        h = ctypes.c_void_p()
        _check(lib().gh_batch_plan_create(key.h, items, len(lens), ctypes.byref(h)),
               "gh_batch_plan_create")
        self.h, self.key = h, key

    def launch(self):
        _check(lib().gh_batch_plan_launch(self.h), "gh_batch_plan_launch")

    def close(self):
        if self.h:
            lib().gh_batch_plan_destroy(self.h)
            self.h = None


def _addr(x) -> int:
    if isinstance(x, DeviceBuffer):
        return x.ptr.value
    if isinstance(x, ctypes.c_void_p):
        return x.value
    return int(x)


class Server:
    """gh_server: continuous batching of commit and Gram requests on a context.

    submit() takes host arrays (mem=MEM_HOST; outputs are allocated here and
    returned by wait()) or device addresses (mem=MEM_DEVICE; the caller passes
    t_out / g_out addresses)."""

    def __init__(self, ctx: Context, mem_bytes=0, max_inflight=0, max_wave=0, slots=0,
                 max_pending=0):
        cfg = ServerConfig(mem_bytes, max_inflight, max_wave, slots, max_pending)
        h = ctypes.c_void_p()
        _check(lib().gh_server_create(ctx.h, ctypes.byref(cfg), ctypes.byref(h)),
               "gh_server_create")
        self.h, self.ctx = h, ctx
        self._held = {}

    def submit(self, ops, ring: Ring, witness, r: int, n: int, key: "Key" = None,
               t_out=None, g_out=None, mem=MEM_HOST) -> int:
        req = Request()
        req.ops, req.r, req.n, req.mem = ops, r, n, mem
        req.ring = ring.h.value
        req.key = key.h.value if key is not None else None
        held = None
        if mem == MEM_HOST:
            w = ring.words(witness)
            if w.size < r * n * ring.d:
                raise ValueError("witness holds fewer than r * n polynomials")
            t = (np.zeros((r, key.rows, ring.d), dtype=ring.dtype)
                 if ops & OP_COMMIT else None)
            g = (np.zeros((r * (r + 1) // 2, ring.d), dtype=ring.dtype)
                 if ops & OP_GRAM else None)
            req.witness = w.ctypes.data
            req.t_out = t.ctypes.data if t is not None else None
            req.g_out = g.ctypes.data if g is not None else None
            held = (w, t, g)
        else:
            req.witness = _addr(witness)
            req.t_out = _addr(t_out) if t_out is not None else None
            req.g_out = _addr(g_out) if g_out is not None else None
        tk = ctypes.c_uint64()
        _check(lib().gh_server_submit(self.h, ctypes.byref(req), ctypes.byref(tk)),
               "gh_server_submit")
        if held is not None:
            self._held[tk.value] = held
        return tk.value

    def submit_wire(self, ops, ring: Ring, witness: bytes, key: "Key" = None, r: int = 0,
                    n: int = 0) -> int:
        """A wire request: witness is LabradorWitness::parts encoded as Vec<Vec<poly>>;
        wait() returns the encoded t_vecs / Gram polynomials as bytes."""
        req = Request()
        req.ops, req.r, req.n, req.mem = ops, r, n, MEM_HOST
        req.ring = ring.h.value
        req.key = key.h.value if key is not None else None
        w = np.frombuffer(witness, dtype=np.uint8) if len(witness) else np.zeros(1, np.uint8)
        if len(witness) >= 4:  # output sizes from the part count; the library checks the rest
            parts = int.from_bytes(witness[:4], "little")
        else:
            parts = 0
        eb = ring.info.wire_bytes
        t_cap = 4 + parts * (8 + key.rows * ring.d * eb) if (ops & OP_COMMIT and key) else 0
        g_cap = 4 + parts * (parts + 1) // 2 * ring.d * eb if ops & OP_GRAM else 0
        t = np.zeros(max(1, t_cap), dtype=np.uint8) if ops & OP_COMMIT else None
        g = np.zeros(max(1, g_cap), dtype=np.uint8) if ops & OP_GRAM else None
        req.witness = w.ctypes.data
        req.witness_len = len(witness)
        req.wire = 1
        req.t_out = t.ctypes.data if t is not None else None
        req.g_out = g.ctypes.data if g is not None else None
        req.t_cap, req.g_cap = t_cap, g_cap
        tk = ctypes.c_uint64()
        _check(lib().gh_server_submit(self.h, ctypes.byref(req), ctypes.byref(tk)),
               "gh_server_submit")
        self._held[tk.value] = (w, t[:t_cap].view() if t is not None else None,
                                g[:g_cap].view() if g is not None else None, True)
        return tk.value

    def _outputs(self, ticket):
        held = self._held.pop(ticket, (None, None, None))
        if len(held) == 4:  # wire request: encoded bytes
            return (held[1].tobytes() if held[1] is not None else None,
                    held[2].tobytes() if held[2] is not None else None)
        return held[1], held[2]

    def wait(self, ticket: int):
        """-> (status, latency_us, t_out, g_out); outputs are None for device requests."""
        st, lat = ctypes.c_int(0), ctypes.c_double(0)
        _check(lib().gh_server_wait(self.h, ticket, ctypes.byref(st), ctypes.byref(lat)),
               "gh_server_wait")
        t, g = self._outputs(ticket)
        return st.value, lat.value, t, g

    def poll(self, ticket: int):
        """-> None while pending, else as wait()."""
        done, st, lat = ctypes.c_int(0), ctypes.c_int(0), ctypes.c_double(0)
        _check(lib().gh_server_poll(self.h, ticket, ctypes.byref(done), ctypes.byref(st),
                                    ctypes.byref(lat)), "gh_server_poll")
        if not done.value:
            return None
        t, g = self._outputs(ticket)
        return st.value, lat.value, t, g

    def stats(self) -> dict:
        s = ServerStats()
        _check(lib().gh_server_query(self.h, ctypes.byref(s)), "gh_server_stats")
        return {name: getattr(s, name) for name, _ in ServerStats._fields_ if name != "reserved"}

    def close(self):
        if self.h:
            lib().gh_server_destroy(self.h)
            self.h = None
            self._held.clear()
