"""The Rust binding written out in INTEGRATION.md declares exactly the C ABI
of include/gridhip.h: the same function names with the same parameter
counts. Text comparison only (no Rust toolchain here)."""
import os
import re

from conftest import ROOT


def _c_decls():
    text = open(os.path.join(ROOT, "include", "gridhip.h")).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    out = {}
    for m in re.finditer(r"GH_API\s+[^;]*?\b(gh_\w+)\s*\(([^)]*)\)\s*;", text, flags=re.S):
        params = m.group(2).strip()
        out[m.group(1)] = 0 if params in ("", "void") else len(params.split(","))
    return out


def _rust_decls():
    text = open(os.path.join(ROOT, "INTEGRATION.md")).read()
    blocks = re.findall(r"```rust\n(.*?)```", text, flags=re.S)
    assert blocks, "INTEGRATION.md holds no rust block"
    ffi = next(b for b in blocks if 'extern "C"' in b)
    out = {}
    for m in re.finditer(r"pub fn (gh_\w+)\s*\(([^)]*)\)", ffi, flags=re.S):
        params = [p for p in m.group(2).split(",") if p.strip()]
        out[m.group(1)] = len(params)
    return out


def test_rust_binding_mirrors_the_header():
    c, rust = _c_decls(), _rust_decls()
    assert len(c) > 50
    missing = sorted(set(c) - set(rust))
    extra = sorted(set(rust) - set(c))
    assert not missing, f"declared in gridhip.h, absent from the Rust binding: {missing}"
    assert not extra, f"in the Rust binding, not in gridhip.h: {extra}"
    arity = {k: (c[k], rust[k]) for k in c if c[k] != rust[k]}
    assert not arity, f"parameter counts differ (header, rust): {arity}"
