"""The C-ABI library loads and exports every entry point include/lmx.h
declares (no device calls: runs without a GPU)."""
import ctypes
import os
import re
import shutil
import subprocess
import tempfile

import pytest

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
    # operator entry points reject a NULL operator before touching the device
    for name in ("lmx_wilson_apply", "lmx_wilson_apply_f32"):
        rc = getattr(L, name)(None, None, None, 0, 1)
        assert rc != _lib.LMX_OK, name


C_CONSUMER = r"""
#include "lmx.h"
#include <stdio.h>
/* a C consumer of the ABI: every public type and a few entry points by address */
int main(void) {
    lmx_cg_params p;
    lmx_cg_result r;
    lmx_rng_key k = {1235u, 0u, 0u, 0u, 0u, 0u};
    lmx_rng_draws d = {0, 0, 0, 0, 0, 0, 17u, 0};
    lmx_tune_result t;
    int (*create)(lmx_context**, const int32_t*, const double*, void*) = lmx_context_create;
    int (*fill)(lmx_context*, void*, int, int, const lmx_rng_key*, double, int, const int32_t*, const int32_t*) =
        lmx_spinor_randomize;
    int (*solve)(lmx_wilson*, void*, const void*, int, void*, size_t, const lmx_cg_params*, lmx_cg_result*) =
        lmx_cg_batched;
    lmx_cg_params_default(&p);
    (void)r; (void)k; (void)d; (void)t; (void)create; (void)fill; (void)solve;
    printf("%d %d %g\n", lmx_abi_version(), LMX_H16, p.rtol);
    return p.maxiter == 5000 ? 0 : 1;
}
"""


def test_header_is_a_c_interface():
    # plain C99 (what cgo, ctypes or Julia's ccall bind against), warnings as
    # errors, linked against the library; run only for the host-side defaults
    gcc = shutil.which("gcc")
    if gcc is None:
        pytest.skip("gcc not available")
    with tempfile.TemporaryDirectory() as d:
        src = os.path.join(d, "consumer.c")
        exe = os.path.join(d, "consumer")
        open(src, "w").write(C_CONSUMER)
        lib_dir = os.path.dirname(_lib.LIB_PATH)
        subprocess.run([gcc, "-std=c99", "-Wall", "-Wextra", "-Werror", "-I", os.path.join(ROOT, "include"), src,
                        "-o", exe, "-L", lib_dir, "-llmx", "-Wl,-rpath," + lib_dir], check=True)
        out = subprocess.run([exe], capture_output=True, text=True, check=True).stdout.split()
        assert out == ["1", "3", "1e-10"]
