"""The C ABI library loads and exports every entry point include/opsec_hip.h
declares (no compute calls; CPU only)."""
import ctypes
import os
import re
import subprocess

import pytest

from opsec_amd import lib

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_functions(header="opsec_hip.h"):
    text = open(os.path.join(ROOT, "include", header)).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(opsec_[a-z0-9_]+)\s*\(", text)))


def test_header_declarations_are_bound():
    decl = declared_functions()
    assert len(decl) > 30
    assert set(decl) == set(lib.exported_symbols())
    assert set(declared_functions("opsec_hip_diag.h")) == set(lib.diag_symbols())


def test_library_exports_every_symbol():
    L = ctypes.CDLL(lib.LIB_PATH)
    missing = [f for f in declared_functions() + declared_functions("opsec_hip_diag.h") if not hasattr(L, f)]
    assert not missing, missing


def test_library_exports_nothing_undeclared():
    # the dynamic symbol table holds exactly the two headers' functions
    out = subprocess.run(["nm", "-D", "--defined-only", lib.LIB_PATH], capture_output=True, text=True, check=True)
    names = {ln.split()[-1] for ln in out.stdout.splitlines() if ln.split()[-2:-1] in (["T"], ["t"])}
    declared = set(declared_functions()) | set(declared_functions("opsec_hip_diag.h"))
    assert names == declared, sorted(names ^ declared)


def test_abi_version_and_defaults():
    L = lib.load()
    assert L.opsec_abi_version() == 1
    q = lib.Quad()
    assert L.opsec_quad_defaults(lib.OPSEC_QUAD_GM, ctypes.byref(q)) == 0
    assert (q.epsrel, q.epsabs, q.maxeval) == (1e-5, 1e-10, 50000000)
    assert L.opsec_quad_defaults(7, ctypes.byref(q)) == -1


def test_no_device_reports_enodev_or_works():
    # Without a GPU the context cannot be created and the error is explicit.
    L = lib.load()
    h = ctypes.c_void_p()
    rc = L.opsec_ctx_create(0, ctypes.byref(h))
    if rc == 0:
        L.opsec_ctx_destroy(h)
        pytest.skip("a GPU is present")
    assert rc == -6
