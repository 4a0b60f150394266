"""ctypes binding of liblmx.so (the C ABI declared in include/lmx.h).

The library is loaded from this package directory.  There is no CPU
fallback: if liblmx.so is missing, or if a device entry point is called
without a GPU, an exception is raised.
"""
from __future__ import annotations

import ctypes
import os

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "liblmx.so")

LMX_OK, LMX_ERR_ARG, LMX_ERR_ALIAS, LMX_ERR_HIP, LMX_ERR_NOMEM, LMX_ERR_UNSUPPORTED, LMX_ERR_WORKSPACE = range(7)
LMX_F64, LMX_F32, LMX_F16, LMX_H16 = 0, 1, 2, 3
LMX_DAGGER, LMX_DONLY = 1, 2
LMX_CG_CONVERGED, LMX_CG_MAXITER, LMX_CG_NONPOSITIVE_CURVATURE, LMX_CG_BREAKDOWN = 0, 1, 2, 3
REASONS = {0: "converged", 1: "maximum_iterations", 2: "nonpositive_curvature", 3: "numerical_breakdown"}
LMX_RNG_PCG32, LMX_RNG_XOSHIRO256PP, LMX_RNG_PHILOX4X32 = 0, 1, 2
LMX_ELEM_F32, LMX_ELEM_F64, LMX_ELEM_C64, LMX_ELEM_C128 = 0, 1, 2, 3
LMX_FILL_UNIFORM, LMX_FILL_GAUSSIAN = 0, 1


class LmxError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__(f"lmx error {code}: {msg}")
        self.code = code


class ArgumentError(LmxError, ValueError):
    """Mirror of Julia's ArgumentError (bad geometry, aliasing, tolerances)."""


class CGParams(ctypes.Structure):
    _fields_ = [("rtol", ctypes.c_double), ("atol", ctypes.c_double), ("maxiter", ctypes.c_int32),
                ("precision", ctypes.c_int32), ("reliable_delta", ctypes.c_double),
                ("reliable_maxiter", ctypes.c_int32), ("check_every", ctypes.c_int32),
                ("use_graph", ctypes.c_int32), ("rhs_block", ctypes.c_int32), ("reserved", ctypes.c_int32 * 4)]


class CGResultC(ctypes.Structure):
    _fields_ = [("converged", ctypes.c_int32), ("iterations", ctypes.c_int32), ("reason", ctypes.c_int32),
                ("reliable_updates", ctypes.c_int32), ("residual_norm", ctypes.c_double),
                ("relative_residual", ctypes.c_double), ("true_relative_residual", ctypes.c_double)]


class TuneResult(ctypes.Structure):
    _fields_ = [("rhs_block", ctypes.c_int32), ("block_size", ctypes.c_int32), ("ms_per_apply", ctypes.c_double)]


class RNGKey(ctypes.Structure):
    _fields_ = [("seed", ctypes.c_uint64), ("sweep", ctypes.c_uint64), ("direction", ctypes.c_uint32),
                ("color", ctypes.c_uint32), ("subgroup", ctypes.c_uint32), ("reserved", ctypes.c_uint32)]


class RNGDraws(ctypes.Structure):
    _fields_ = [("words", ctypes.c_int32), ("uniform32", ctypes.c_int32), ("uniform64", ctypes.c_int32),
                ("normal32", ctypes.c_int32), ("normal64", ctypes.c_int32), ("bounded", ctypes.c_int32),
                ("upper", ctypes.c_uint32), ("reserved", ctypes.c_int32)]


APPLY_FN = ctypes.CFUNCTYPE(ctypes.c_int, ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p)

# name -> (restype, argtypes); the exported surface of include/lmx.h
P = ctypes.c_void_p
I = ctypes.c_int
D = ctypes.c_double
SIGNATURES = {
    "lmx_last_error": (ctypes.c_char_p, []),
    "lmx_abi_version": (I, []),
    "lmx_context_create": (I, [ctypes.POINTER(P), ctypes.POINTER(ctypes.c_int32), P, P]),
    "lmx_context_create_nc": (I, [ctypes.POINTER(P), ctypes.POINTER(ctypes.c_int32), P, P, I]),
    "lmx_context_destroy": (I, [P]),
    "lmx_context_set_stream": (I, [P, P]),
    "lmx_context_synchronize": (I, [P]),
    "lmx_context_volume": (ctypes.c_int64, [P]),
    "lmx_spinor_bytes": (ctypes.c_size_t, [P, I, I]),
    "lmx_links_bytes": (ctypes.c_size_t, [P, I]),
    "lmx_clover_bytes": (ctypes.c_size_t, [P, I]),
    "lmx_spinor_import": (I, [P, P, P, I, I]),
    "lmx_spinor_export": (I, [P, P, I, P, I]),
    "lmx_links_import": (I, [P, P, ctypes.POINTER(P), I]),
    "lmx_links_export": (I, [P, ctypes.POINTER(P), I, P]),
    "lmx_axpby": (I, [P, D, P, D, P, I]),
    "lmx_dot": (I, [P, P, P, ctypes.POINTER(D)]),
    "lmx_scale": (I, [P, P, P, D, D, I]),
    "lmx_wilson_create": (I, [P, ctypes.POINTER(P), P, D, D]),
    "lmx_wilson_destroy": (I, [P]),
    "lmx_wilson_set_kappa": (I, [P, D]),
    "lmx_wilson_update_links": (I, [P, P]),
    "lmx_clover_export": (I, [P, P]),
    "lmx_wilson_apply": (I, [P, P, P, I, I]),
    "lmx_wilson_apply_f32": (I, [P, P, P, I, I]),
    "lmx_wilson_apply_padded": (I, [P, P, P, I, I, I]),
    "lmx_wilson_apply_padded_f32": (I, [P, P, P, I, I, I]),
    "lmx_ddagd_apply": (I, [P, P, P, P, I]),
    "lmx_cg_params_default": (None, [ctypes.POINTER(CGParams)]),
    "lmx_cg_ddagd": (I, [P, P, P, P, P, P, P, ctypes.POINTER(CGParams), ctypes.POINTER(CGResultC)]),
    "lmx_cg_ddagd_noalloc": (I, [P, P, P, P, P, P, P, ctypes.POINTER(CGParams), ctypes.POINTER(CGResultC)]),
    "lmx_cg_ddagd_f32": (I, [P, P, P, P, P, P, P, ctypes.POINTER(CGParams), ctypes.POINTER(CGResultC)]),
    "lmx_cg_ddagd_padded": (I, [P, P, P, I, P, ctypes.c_size_t, ctypes.POINTER(CGParams), ctypes.POINTER(CGResultC)]),
    "lmx_cg_ddagd_padded_f32": (I, [P, P, P, I, P, ctypes.c_size_t, ctypes.POINTER(CGParams),
                                    ctypes.POINTER(CGResultC)]),
    "lmx_cg_generic": (I, [P, APPLY_FN, P, P, P, P, P, P, ctypes.POINTER(CGParams), ctypes.POINTER(CGResultC)]),
    "lmx_cg_batched_workspace": (I, [P, I, ctypes.POINTER(CGParams), ctypes.POINTER(ctypes.c_size_t)]),
    "lmx_cg_batched": (I, [P, P, P, I, P, ctypes.c_size_t, ctypes.POINTER(CGParams), ctypes.POINTER(CGResultC)]),
    "lmx_wilson_pullback": (I, [P, P, P, P, D]),
    "lmx_ta_project": (I, [P, P, P, P]),
    "lmx_wilson_force": (I, [P, P, P, P, P, P, P, ctypes.POINTER(CGParams), ctypes.POINTER(CGResultC)]),
    "lmx_autotune": (I, [P, I, I, ctypes.c_char_p, ctypes.POINTER(TuneResult)]),
    "lmx_inner_kernel_times": (I, [P, I, I, I, I, ctypes.POINTER(D)]),
    "lmx_field_pad": (I, [P, P, P, I, I, I, I]),
    "lmx_field_gather": (I, [P, P, P, I, I, I, I]),
    "lmx_site_rng_draw": (I, [P, ctypes.POINTER(RNGKey), I, ctypes.c_uint64, ctypes.c_int32, ctypes.POINTER(RNGDraws),
                              P, P, P, P, P, P]),
    "lmx_randomize_padded": (I, [P, P, I, I, I, I, I, ctypes.POINTER(RNGKey), D, I, ctypes.POINTER(ctypes.c_int32),
                                 ctypes.POINTER(ctypes.c_int32)]),
    "lmx_spinor_randomize": (I, [P, P, I, I, ctypes.POINTER(RNGKey), D, I, ctypes.POINTER(ctypes.c_int32),
                                 ctypes.POINTER(ctypes.c_int32)]),
}

_lib = None


def lib():
    """Load liblmx.so (raises if it has not been built)."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError(f"liblmx.so not built at {LIB_PATH}: run __graft_entry__.build()")
        L = ctypes.CDLL(LIB_PATH)
        for name, (res, args) in SIGNATURES.items():
            f = getattr(L, name)
            f.restype = res
            f.argtypes = args
        _lib = L
    return _lib


def check(rc: int) -> None:
    if rc != LMX_OK:
        msg = lib().lmx_last_error().decode()
        if rc in (LMX_ERR_ARG, LMX_ERR_ALIAS):
            raise ArgumentError(rc, msg)
        raise LmxError(rc, msg)


def default_params(**kw) -> CGParams:
    p = CGParams()
    lib().lmx_cg_params_default(ctypes.byref(p))
    for k, v in kw.items():
        setattr(p, k, v)
    return p
