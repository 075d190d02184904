"""The C-ABI library builds, loads and exports every symbol include/gridhip.h declares.

No compute calls here: this suite runs without a GPU.
"""
import os
import re
import subprocess

import pytest

from conftest import ROOT

HEADER = os.path.join(ROOT, "include", "gridhip.h")


def declared_symbols():
    text = open(HEADER).read()
    return sorted(set(re.findall(r"GH_API\s+[\w\s\*]+?\b(gh_\w+)\s*\(", text)))


def test_header_declares_expected_surface():
    import lattica_amd as la
    assert declared_symbols() == sorted(la.EXPORTS)


def test_library_exports_every_declared_symbol():
    import lattica_amd as la
    path = la.library_path()
    if not os.path.exists(path):
        pytest.fail("libgridhip.so is not built (run python -m lattica_amd.build)")
    out = subprocess.run(["nm", "-D", "--defined-only", path], capture_output=True, text=True,
                         check=True).stdout
    exported = set(re.findall(r"\sT\s(gh_\w+)$", out, flags=re.M))
    missing = [s for s in declared_symbols() if s not in exported]
    assert not missing, missing
    # nothing beyond the declared ABI leaks out
    extra = sorted(s for s in exported if s not in set(declared_symbols()))
    assert not extra, extra


def test_library_loads_and_reports_status_strings():
    import lattica_amd as la
    L = la.lib()
    assert L.gh_version() >= 1
    assert L.gh_status_string(0) == b"ok"
    assert b"twist" in L.gh_status_string(-4)
    assert L.gh_status_string(-11) == b"dimension mismatch"


def test_null_arguments_rejected_without_device():
    import ctypes

    import lattica_amd as la
    L = la.lib()
    assert L.gh_ctx_create(0, None) == -20
    assert L.gh_ring_create(None, 17, 8, ctypes.byref(ctypes.c_void_p())) == -20
    assert L.gh_commit(None, None, 0, None, 0, 0, None, 0) == -20
    assert L.gh_device_count(None) == -20


def test_no_device_is_an_error_not_a_fallback():
    import ctypes

    import lattica_amd as la
    n = ctypes.c_int(-1)
    assert la.lib().gh_device_count(ctypes.byref(n)) == 0
    if n.value == 0:
        h = ctypes.c_void_p()
        assert la.lib().gh_ctx_create(0, ctypes.byref(h)) == -26  # GH_E_NO_DEVICE
