"""Generic lattice fields in the reference's own storage, on the device.

Mirrors of LatticeMatrix(A, dim, PEs) and gather_matrix on one rank
(src/LatticeMatrices_core.jl:405-461, 936-993) for any NC1 x NC2 matrix per
site and element type Float32 / Float64 / ComplexF32 / ComplexF64.  Arrays are
device tensors in C order: the core array A[NC1, NC2, L1, ..., L4] is
[L4, L3, L2, L1, NC2, NC1], the padded storage [L4+2nw, ..., L1+2nw, NC2, NC1].
"""
from __future__ import annotations

import ctypes

from . import _lib
from ._lib import ArgumentError, check, lib

try:
    import torch
except ImportError:  # pragma: no cover
    torch = None

ELEM = {}
if torch is not None:
    ELEM = {torch.float32: _lib.LMX_ELEM_F32, torch.float64: _lib.LMX_ELEM_F64,
            torch.complex64: _lib.LMX_ELEM_C64, torch.complex128: _lib.LMX_ELEM_C128}


def _check(t, lattice, lead):
    if not (isinstance(t, torch.Tensor) and t.is_contiguous() and t.device == lattice.device and t.dtype in ELEM):
        raise ArgumentError(_lib.LMX_ERR_ARG, "fields are contiguous Float32, Float64, ComplexF32 or ComplexF64 "
                                              "device tensors")
    if t.dim() != 6 or tuple(t.shape[:4]) != lead:
        raise ArgumentError(_lib.LMX_ERR_ARG, f"array shape {tuple(t.shape)} is not [*{lead}, NC2, NC1]")
    return int(t.shape[5]), int(t.shape[4])


def lattice_matrix(lattice, A, nw: int = 1):
    """LatticeMatrix(A, dim, PEs; nw, phases) on one rank: the padded storage
    of the core array A, its halo filled as set_halo! does with the lattice's
    boundary phases.  Returns a new device tensor."""
    nc1, nc2 = _check(A, lattice, tuple(reversed(lattice.dims)))
    out = torch.empty(tuple(d + 2 * int(nw) for d in reversed(lattice.dims)) + (nc2, nc1), dtype=A.dtype,
                      device=A.device)
    check(lib().lmx_field_pad(lattice.ctx, ctypes.c_void_p(out.data_ptr()), ctypes.c_void_p(A.data_ptr()),
                              ELEM[A.dtype], nc1, nc2, int(nw)))
    return out


def gather_matrix(lattice, C, nw: int = 1):
    """gather_matrix(ls) on one rank: the interior of the padded storage C as
    a new core-array device tensor."""
    nc1, nc2 = _check(C, lattice, tuple(d + 2 * int(nw) for d in reversed(lattice.dims)))
    out = torch.empty(tuple(reversed(lattice.dims)) + (nc2, nc1), dtype=C.dtype, device=C.device)
    check(lib().lmx_field_gather(lattice.ctx, ctypes.c_void_p(out.data_ptr()), ctypes.c_void_p(C.data_ptr()),
                                 ELEM[C.dtype], nc1, nc2, int(nw)))
    return out
