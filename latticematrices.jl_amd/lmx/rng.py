"""Site random-number streams and global-site lattice fills on the GPU.

Mirror of the reference's site RNG API (src/utilities/randomgenerator.jl)
and of randomize_matrix! / randomize_gaussian_matrix!
(src/LinearAlgebras/linearalgebra_D.jl:2550-2660).  Every zero-based global
site id (x fastest) owns one stream keyed by an RNGStreamKey, so a field
does not depend on how the lattice is split.  Integer and uniform draws are
bit-identical to the reference; normal draws agree to the last bits of the
device log / sqrt / sincos.
"""
from __future__ import annotations

import ctypes
from dataclasses import dataclass

import numpy as np

from . import _lib
from ._lib import ArgumentError, check, lib
from .fields import ELEM
from .lattice import SpinorField

try:
    import torch
except ImportError:  # pragma: no cover
    torch = None

# _DEFAULT_RANDOMIZE_SEED, linearalgebra_D.jl:2539
DEFAULT_RANDOMIZE_SEED = 0x12345678ABCDEF01
ALGORITHMS = {"pcg32": _lib.LMX_RNG_PCG32, "xoshiro256++": _lib.LMX_RNG_XOSHIRO256PP,
              "philox4x32": _lib.LMX_RNG_PHILOX4X32}


@dataclass(frozen=True)
class RNGStreamKey:
    """RNGStreamKey(seed, sweep=0, direction=0, color=0, subgroup=0)
    (randomgenerator.jl:98-118): seed and sweep are UInt64, the other tags
    UInt32; values outside those ranges are rejected (Julia's InexactError)."""
    seed: int
    sweep: int = 0
    direction: int = 0
    color: int = 0
    subgroup: int = 0

    def __post_init__(self):
        for name, bits in (("seed", 64), ("sweep", 64), ("direction", 32), ("color", 32), ("subgroup", 32)):
            v = getattr(self, name)
            if not isinstance(v, (int, np.integer)) or not 0 <= int(v) < (1 << bits):
                raise ArgumentError(_lib.LMX_ERR_ARG, f"RNGStreamKey.{name} = {v!r} is not a UInt{bits}")

    def _c(self) -> _lib.RNGKey:
        return _lib.RNGKey(int(self.seed), int(self.sweep), int(self.direction), int(self.color),
                           int(self.subgroup), 0)


def _algorithm(name: str) -> int:
    if name not in ALGORITHMS:
        raise ArgumentError(_lib.LMX_ERR_ARG, f"unknown site RNG algorithm {name!r} (one of {sorted(ALGORITHMS)})")
    return ALGORITHMS[name]


def site_rng_draw(lattice, key: RNGStreamKey, nsites: int, *, algorithm: str = "pcg32", first_site: int = 0,
                  words: int = 0, uniform32: int = 0, uniform64: int = 0, normal32: int = 0, normal64: int = 0,
                  bounded: int = 0, upper: int = 17) -> dict:
    """For sites first_site .. first_site + nsites - 1: site_rng(key, site,
    algorithm) (randomgenerator.jl:152-200; the reference's default algorithm
    is PCG32) followed by the given numbers of rand_u32, rand_uniform(Float32),
    rand_uniform(Float64), rand_normal_pair(Float32), rand_normal_pair(Float64)
    values and rand_bounded(upper) draws, in that order, on the device.
    Returns host arrays [nsites, count] (the reference's [count, nsites])."""
    counts = dict(words=(words, torch.int32), uniform32=(uniform32, torch.float32),
                  uniform64=(uniform64, torch.float64), normal32=(normal32, torch.float32),
                  normal64=(normal64, torch.float64), bounded=(bounded, torch.int32))
    outs = {k: torch.zeros((int(nsites), max(int(n), 0)), dtype=dt, device=lattice.device) for k, (n, dt) in counts.items()}
    n = _lib.RNGDraws(int(words), int(uniform32), int(uniform64), int(normal32), int(normal64), int(bounded),
                      int(upper) & 0xFFFFFFFF, 0)

    def p(k):
        return ctypes.c_void_p(outs[k].data_ptr() if outs[k].numel() else 0)

    check(lib().lmx_site_rng_draw(lattice.ctx, ctypes.byref(key._c()), _algorithm(algorithm), int(first_site),
                                  int(nsites), ctypes.byref(n), p("words"), p("uniform32"), p("uniform64"),
                                  p("normal32"), p("normal64"), p("bounded")))
    lattice.synchronize()
    res = {k: v.cpu().numpy() for k, v in outs.items()}
    res["words"] = res["words"].view(np.uint32)
    res["bounded"] = res["bounded"].view(np.uint32)
    return res


_ELEM = ELEM


def _placement(global_offset, global_dims):
    if (global_offset is None) != (global_dims is None):
        raise ArgumentError(_lib.LMX_ERR_ARG, "global_offset and global_dims are given together or not at all")
    if global_offset is None:
        return None, None
    if len(global_offset) != 4 or len(global_dims) != 4:
        raise ArgumentError(_lib.LMX_ERR_ARG, "global_offset and global_dims need 4 entries")
    return (ctypes.c_int32 * 4)(*[int(v) for v in global_offset]), (ctypes.c_int32 * 4)(*[int(v) for v in global_dims])


def _fill(C, dist, keys, sigma, rng_algorithm, lattice, nw, global_offset, global_dims):
    alg = _algorithm(rng_algorithm)
    off, gd = _placement(global_offset, global_dims)
    if isinstance(C, SpinorField):
        if len(keys) != C.nrhs:
            raise ArgumentError(_lib.LMX_ERR_ARG, f"{C.nrhs} right-hand sides need {C.nrhs} keys, got {len(keys)}")
        arr = (_lib.RNGKey * max(len(keys), 1))(*[k._c() for k in keys])
        check(lib().lmx_spinor_randomize(C.lattice.ctx, C.ptr, C.nrhs, dist, arr, float(sigma), alg, off, gd))
        return C
    if lattice is None:
        raise ArgumentError(_lib.LMX_ERR_ARG, "a reference-storage array needs its lattice")
    if len(keys) != 1:
        raise ArgumentError(_lib.LMX_ERR_ARG, "a LatticeMatrix array takes one key")
    if not (isinstance(C, torch.Tensor) and C.is_contiguous() and C.device == lattice.device and C.dtype in _ELEM):
        raise ArgumentError(_lib.LMX_ERR_ARG, "random lattice fills support contiguous Float32, Float64, "
                                              "ComplexF32 and ComplexF64 device arrays")
    padded = tuple(d + 2 * int(nw) for d in reversed(lattice.dims))
    if C.dim() != 6 or tuple(C.shape[:4]) != padded:
        raise ArgumentError(_lib.LMX_ERR_ARG, f"array shape {tuple(C.shape)} is not [*{padded}, NC2, NC1] "
                                              f"(C order of A[NC1, NC2, L1+2nw, ..., L4+2nw])")
    nc2, nc1 = int(C.shape[4]), int(C.shape[5])
    check(lib().lmx_randomize_padded(lattice.ctx, ctypes.c_void_p(C.data_ptr()), _ELEM[C.dtype], nc1, nc2, int(nw),
                                     dist, ctypes.byref(keys[0]._c()), float(sigma), alg, off, gd))
    return C


def _keys(key):
    if key is None:
        return None
    return list(key) if isinstance(key, (list, tuple)) else [key]


def randomize_matrix_(C, key=None, *, seed: int = DEFAULT_RANDOMIZE_SEED, sweep: int = 0, direction: int = 0,
                      color: int = 0, subgroup: int = 0, rng_algorithm: str = "philox4x32", lattice=None, nw: int = 0,
                      global_offset=None, global_dims=None):
    """randomize_matrix!(C, key; rng_algorithm) (linearalgebra_D.jl:2560-2597):
    uniform values in [-0.5, 0.5), complex entries (re, im) from consecutive
    draws.  C is a SpinorField (one key per right-hand side) or a device
    tensor in the reference's padded storage (C order [L4+2nw, L3+2nw, L2+2nw,
    L1+2nw, NC2, NC1]) with its lattice and nw; a whole-lattice fill refills
    the halo as set_halo! does.  global_offset / global_dims place the lattice
    as a box of a larger one."""
    keys = _keys(key) or [RNGStreamKey(seed, sweep, direction, color, subgroup)]
    return _fill(C, _lib.LMX_FILL_UNIFORM, keys, 1.0, rng_algorithm, lattice, nw, global_offset, global_dims)


def randomize_gaussian_matrix_(C, key=None, *, sigma: float = 1.0, seed: int = DEFAULT_RANDOMIZE_SEED,
                               sweep: int = 0, direction: int = 0, color: int = 0, subgroup: int = 0,
                               rng_algorithm: str = "philox4x32", lattice=None, nw: int = 0, global_offset=None,
                               global_dims=None):
    """randomize_gaussian_matrix!(C, key; sigma, rng_algorithm)
    (linearalgebra_D.jl:2610-2653): zero-mean normals of standard deviation
    sigma (>= 0), complex entries (re, im) from one Box-Muller pair, real
    entries using both values of a pair.  Arguments as randomize_matrix_."""
    if not float(sigma) >= 0.0:
        raise ArgumentError(_lib.LMX_ERR_ARG, "sigma must be non-negative")
    keys = _keys(key) or [RNGStreamKey(seed, sweep, direction, color, subgroup)]
    return _fill(C, _lib.LMX_FILL_GAUSSIAN, keys, sigma, rng_algorithm, lattice, nw, global_offset, global_dims)
