"""The C-ABI library loads and exports every entry point include/lmx.h
declares (no device calls: runs without a GPU)."""
import ctypes
import os
import re

from lmx import _lib

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_symbols():
    text = open(os.path.join(ROOT, "include", "lmx.h")).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    names = set(re.findall(r"\b(lmx_[a-z0-9_]+)\s*\(", text))
    return sorted(names)


def test_library_loads_and_exports_every_declared_symbol():
    lib = ctypes.CDLL(_lib.LIB_PATH)
    names = declared_symbols()
    assert len(names) >= 30
    missing = [n for n in names if not hasattr(lib, n)]
    assert not missing, missing


def test_binding_covers_header():
    assert set(declared_symbols()) == set(_lib.SIGNATURES)


def test_abi_version_and_errors_without_device_calls():
    L = _lib.lib()
    assert L.lmx_abi_version() == 1
    # argument validation happens before any HIP call
    h = ctypes.c_void_p()
    dims = (ctypes.c_int32 * 4)(4, 4, 0, 4)
    rc = L.lmx_context_create(ctypes.byref(h), dims, None, None)
    assert rc == _lib.LMX_ERR_ARG
    assert b"positive" in L.lmx_last_error()
    p = _lib.default_params()
    assert p.rtol == 1e-10 and p.atol == 0.0 and p.maxiter == 5000
