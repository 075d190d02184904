"""Build recipe for libgridhip.so (HIP kernels + C ABI) for gfx950.

Run as `python -m lattica_amd.build` or through __graft_entry__.build(). The
library is built in-tree next to this file so it travels with the checkout.
"""
from __future__ import annotations

import os
import shutil
import subprocess
import sys
import tempfile

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
LIB = os.path.join(HERE, "libgridhip.so")
ARCH = os.environ.get("GRIDHIP_ARCH", "gfx950")

SOURCES = ["gh_plan.cpp", "gh_capi.cpp", "gh_kernels.hip", "gh_labrador.hip", "gh_batch.hip",
           "gh_server.cpp", "gh_wire.hip", "gh_jl.hip", "gh_slice.hip"]

COMMON = [
    "-O3", "-std=c++17", "-fPIC", "-Wall", "-Wno-unused-function",
    f"--offload-arch={ARCH}",
    # never emit packed FP32 ops (none are wanted in this integer code)
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-fvisibility=hidden",
]


def _hipcc() -> str:
    for cand in (os.environ.get("HIPCC"), "/opt/rocm/bin/hipcc", shutil.which("hipcc")):
        if cand and os.path.exists(cand):
            return cand
    raise RuntimeError("hipcc not found")


def build(verbose: bool = False) -> str:
    from concurrent.futures import ThreadPoolExecutor

    hipcc = _hipcc()
    with tempfile.TemporaryDirectory(prefix="gridhip_") as tmp:
        jobs = []
        for src in SOURCES:
            path = os.path.join(CSRC, src)
            obj = os.path.join(tmp, os.path.splitext(src)[0] + ".o")
            lang = ["-x", "hip"] if src.endswith(".hip") else []
            jobs.append((src, obj, [hipcc, *COMMON, *lang, "-c", path, "-o", obj]))

        def compile_one(job):
            src, obj, cmd = job
            if verbose:
                print(" ".join(cmd), flush=True)
            res = subprocess.run(cmd, capture_output=True, text=True)
            if res.returncode != 0:
                sys.stderr.write(res.stdout + res.stderr)
                raise RuntimeError(f"hipcc failed on {src}")
            return obj

        with ThreadPoolExecutor(max_workers=min(len(jobs), os.cpu_count() or 1)) as ex:
            objs = list(ex.map(compile_one, jobs))
        tmp_lib = LIB + ".tmp"
        cmd = [hipcc, f"--offload-arch={ARCH}", "-shared", "-fPIC", "-o", tmp_lib, *objs]
        res = subprocess.run(cmd, capture_output=True, text=True)
        if res.returncode != 0:
            sys.stderr.write(res.stdout + res.stderr)
            raise RuntimeError("link failed")
        os.replace(tmp_lib, LIB)
    return LIB


if __name__ == "__main__":
    print(build(verbose="-v" in sys.argv))
