"""ctypes front end of the CPU oracle (liblmx_oracle.so).

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and the
cpu_baseline leg of bench.py, always as the checker (or as the timed CPU
restatement for cpu_baseline).  The product package never imports this module.

Array conventions (numpy, C order, complex128):
  spinor  psi[V, 4, 3]        (site, spin, colour)  == the reference's
                              (3,4,X,Y,Z,T) column-major array without halo
  links   U[4, V, 3, 3]       (mu, site, row, col)
  clover  F[6, V, 3, 3]       planes (12,13,14,23,24,34)
Sites are x-fastest: site = x + Lx*(y + Ly*(z + Lz*t)).
"""
from __future__ import annotations

import ctypes
import os
from dataclasses import dataclass

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB_PATH = os.path.join(_HERE, "liblmx_oracle.so")

CONVERGED, MAXITER, NONPOSITIVE_CURVATURE, BREAKDOWN = 0, 1, 2, 3
REASONS = {CONVERGED: "converged", MAXITER: "maximum_iterations",
           NONPOSITIVE_CURVATURE: "nonpositive_curvature", BREAKDOWN: "numerical_breakdown"}

OP_DDAGD_WILSON, OP_DDAGD_CLOVER, OP_SCALE, OP_DDAGD_SCALE = 0, 1, 2, 3


class _OrcOp(ctypes.Structure):
    _fields_ = [("kind", ctypes.c_int), ("U", ctypes.c_void_p), ("F", ctypes.c_void_p),
                ("L", ctypes.c_int * 4), ("phases", ctypes.c_double * 8),
                ("kappa", ctypes.c_double), ("csw", ctypes.c_double),
                ("scale_re", ctypes.c_double), ("scale_im", ctypes.c_double),
                ("tmp", ctypes.c_void_p), ("nc", ctypes.c_int)]


class _OrcResult(ctypes.Structure):
    _fields_ = [("converged", ctypes.c_int), ("iterations", ctypes.c_int),
                ("residual_norm", ctypes.c_double), ("relative_residual", ctypes.c_double),
                ("reason", ctypes.c_int)]


@dataclass
class CGResult:
    converged: bool
    iterations: int
    residual_norm: float
    relative_residual: float
    reason: str


_lib = None


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(_LIB_PATH):
            raise RuntimeError(f"oracle library not built: {_LIB_PATH} (run __graft_entry__.build())")
        L = ctypes.CDLL(_LIB_PATH)
        P = ctypes.c_void_p
        L.orc_wilson.argtypes = [P, P, P, P, P, ctypes.c_double, ctypes.c_int, ctypes.c_int]
        L.orc_dot.argtypes = [P, P, ctypes.c_long, ctypes.c_int, P]
        L.orc_axpby.argtypes = [ctypes.c_double, P, ctypes.c_double, P, ctypes.c_long]
        L.orc_clover_field_strength.argtypes = [P, P, P]
        L.orc_wilson_clover.argtypes = [P, P, P, P, P, P, ctypes.c_double, ctypes.c_double, ctypes.c_int]
        L.orc_cg.argtypes = [P, P, P, P, P, ctypes.POINTER(_OrcOp), ctypes.c_long, ctypes.c_double,
                             ctypes.c_double, ctypes.c_int, ctypes.c_int, ctypes.POINTER(_OrcResult)]
        L.orc_op_apply.argtypes = [ctypes.POINTER(_OrcOp), P, P, ctypes.c_long]
        L.orc_wilson_pullback.argtypes = [P, P, P, P, P, ctypes.c_double]
        L.orc_ta_force.argtypes = [P, P, P, ctypes.c_long]
        L.orc_dsfdu.argtypes = [P, P, P, P, P, P, P, ctypes.c_double, ctypes.c_double, ctypes.c_int,
                                ctypes.c_int, P, ctypes.POINTER(_OrcResult)]
        L.orc_clover_pullback.argtypes = [P, P, P, P, P, P, ctypes.c_double, ctypes.c_double]
        L.orc_dsfdu_clover.argtypes = [P, P, P, P, P, P, P, P, ctypes.c_double, ctypes.c_double, ctypes.c_double,
                                       ctypes.c_int, ctypes.c_int, P, ctypes.POINTER(_OrcResult)]
        U64, U32 = ctypes.c_uint64, ctypes.c_uint32
        L.orc_splitmix64.argtypes, L.orc_splitmix64.restype = [U64], U64
        L.orc_rng_stream_seed.argtypes, L.orc_rng_stream_seed.restype = [U64, U64, U32, U32, U32], U64
        L.orc_pcg32_step.argtypes, L.orc_pcg32_step.restype = [ctypes.POINTER(U64), U64], U32
        L.orc_philox4x32_10.argtypes = [P, P, P]
        L.orc_site_rng_draw.argtypes = [U64, ctypes.c_int, U64, ctypes.c_int, P, U32, P, P, P, P, P, P]
        L.orc_randomize_fill.argtypes = [P, ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_int, P, P, P,
                                         ctypes.c_int, ctypes.c_double, U64, ctypes.c_int]
        L.orc_dot_n.argtypes = [P, P, ctypes.c_long, ctypes.c_int, ctypes.c_int, P]
        L.orc_wilson_nc.argtypes = [P, P, P, P, P, ctypes.c_double, ctypes.c_int, ctypes.c_int, ctypes.c_int]
        L.orc_clover_field_strength_nc.argtypes = [P, P, P, ctypes.c_int]
        L.orc_wilson_clover_nc.argtypes = [P, P, P, P, P, P, ctypes.c_double, ctypes.c_double, ctypes.c_int,
                                           ctypes.c_int]
        L.orc_wilson_pullback_nc.argtypes = [P, P, P, P, P, ctypes.c_double, ctypes.c_int]
        L.orc_refcpu_cg_ddagd.argtypes = [P, P, P, P, P, P, P, P, P, ctypes.c_double, ctypes.c_double, ctypes.c_int,
                                          ctypes.POINTER(_OrcResult)]
        L.orc_refcpu_pad_spinor.argtypes = [P, P, P, P]
        L.orc_refcpu_pad_links.argtypes = [P, P, P]
        L.orc_refcpu_unpad_spinor.argtypes = [P, P, P]
        L.orc_num_threads.restype = ctypes.c_int
        L.orc_set_num_threads.argtypes = [ctypes.c_int]
        _lib = L
    return _lib


def _c(a, shape=None):
    a = np.ascontiguousarray(a, dtype=np.complex128)
    if shape is not None and a.shape != shape:
        a = a.reshape(shape)
    return a


def _ptr(a):
    return a.ctypes.data_as(ctypes.c_void_p)


def _dims(L):
    return (ctypes.c_int * 4)(*[int(v) for v in L])


def _phases(phases):
    ph = np.zeros(8, dtype=np.float64)
    for d in range(4):
        z = complex(phases[d]) if phases is not None else 1.0 + 0j
        ph[2 * d], ph[2 * d + 1] = z.real, z.imag
    return ph


def volume(L):
    return int(np.prod([int(v) for v in L]))


def set_num_threads(n: int) -> None:
    lib().orc_set_num_threads(int(n))


def num_threads() -> int:
    return int(lib().orc_num_threads())


def wilson(U, psi, L, kappa, phases=(1, 1, 1, -1), dagger=False, donly=False):
    """D_W psi (dagger: D_W^dagger psi; donly: hopping-only, coefficient 1/2)."""
    V = volume(L)
    U = _c(U, (4, V, 3, 3))
    psi = _c(psi, (V, 4, 3))
    out = np.empty_like(psi)
    ph = _phases(phases)
    lib().orc_wilson(_ptr(out), _ptr(psi), _ptr(U), _dims(L), _ptr(ph), float(kappa), int(dagger), int(donly))
    return out


def dot(a, b, order=1):
    a = _c(a)
    b = _c(b)
    V = a.size // 12
    out = np.zeros(2, dtype=np.float64)
    lib().orc_dot(_ptr(a), _ptr(b), V, int(order), _ptr(out))
    return complex(out[0], out[1])


def clover_field_strength(U, L):
    V = volume(L)
    U = _c(U, (4, V, 3, 3))
    F = np.empty((6, V, 3, 3), dtype=np.complex128)
    lib().orc_clover_field_strength(_ptr(F), _ptr(U), _dims(L))
    return F


def wilson_clover(U, F, psi, L, kappa, csw, phases=(1, 1, 1, -1), dagger=False):
    V = volume(L)
    U = _c(U, (4, V, 3, 3))
    F = _c(F, (6, V, 3, 3))
    psi = _c(psi, (V, 4, 3))
    out = np.empty_like(psi)
    ph = _phases(phases)
    lib().orc_wilson_clover(_ptr(out), _ptr(psi), _ptr(U), _ptr(F), _dims(L), _ptr(ph), float(kappa),
                            float(csw), int(dagger))
    return out


class NormalOp:
    """Operator description for the oracle CG: D^dagger D of Wilson/clover, a
    complex scale (mul!(y, s, x)) or D^dagger D of a scale."""

    def __init__(self, kind, L, U=None, F=None, kappa=0.0, csw=0.0, phases=(1, 1, 1, -1), scale=0j, nc=3):
        self.V = volume(L)
        self.nc = int(nc)
        self._keep = []
        self.s = _OrcOp()
        self.s.kind = kind
        self.s.nc = self.nc
        if U is not None:
            U = _c(U, (4, self.V, nc, nc))
            self._keep.append(U)
            self.s.U = U.ctypes.data
        if F is not None:
            F = _c(F, (6, self.V, nc, nc))
            self._keep.append(F)
            self.s.F = F.ctypes.data
        for d in range(4):
            self.s.L[d] = int(L[d])
        ph = _phases(phases)
        for k in range(8):
            self.s.phases[k] = ph[k]
        self.s.kappa, self.s.csw = float(kappa), float(csw)
        self.s.scale_re, self.s.scale_im = complex(scale).real, complex(scale).imag
        tmp = np.zeros((self.V, 4, nc), dtype=np.complex128)
        self._keep.append(tmp)
        self.s.tmp = tmp.ctypes.data

    def apply(self, x):
        x = _c(x, (self.V, 4, self.nc))
        y = np.empty_like(x)
        lib().orc_op_apply(ctypes.byref(self.s), _ptr(y), _ptr(x), self.V)
        return y


def cg(op: NormalOp, b, x0=None, rtol=1e-10, atol=0.0, maxiter=5000, order=1):
    """cg!(x, A, b, r, p, Ap; rtol, atol, maxiter); returns (x, CGResult)."""
    V = op.V
    b = _c(b, (V, 4, op.nc))
    x = np.zeros_like(b) if x0 is None else _c(x0, (V, 4, op.nc)).copy()
    r = np.zeros_like(b)
    p = np.zeros_like(b)
    Ap = np.zeros_like(b)
    res = _OrcResult()
    lib().orc_cg(_ptr(x), _ptr(b), _ptr(r), _ptr(p), _ptr(Ap), ctypes.byref(op.s), V, float(rtol), float(atol),
                 int(maxiter), int(order), ctypes.byref(res))
    return x, CGResult(bool(res.converged), int(res.iterations), float(res.residual_norm),
                       float(res.relative_residual), REASONS[int(res.reason)])


# ---------------------------------------------------------------------------
# Generic NC (lmx_nc_oracle.c): the reference's kernels for NC != 3.  Arrays
# as above with nc in place of 3: psi[V, 4, nc], U[4, V, nc, nc],
# F[6, V, nc, nc].
# ---------------------------------------------------------------------------
def wilson_nc(U, psi, L, kappa, phases=(1, 1, 1, -1), dagger=False, donly=False):
    """D psi for any NC by the reference's generic kernels (NC = 3 runs them too,
    although the reference itself uses its half-spin kernels there)."""
    V = volume(L)
    psi = _c(psi)
    nc = psi.shape[-1]
    psi = psi.reshape(V, 4, nc)
    U = _c(U, (4, V, nc, nc))
    out = np.empty_like(psi)
    lib().orc_wilson_nc(_ptr(out), _ptr(psi), _ptr(U), _dims(L), _ptr(_phases(phases)), float(kappa), nc,
                        int(dagger), int(donly))
    return out


def clover_field_strength_nc(U, L):
    V = volume(L)
    U = _c(U)
    nc = U.shape[-1]
    U = U.reshape(4, V, nc, nc)
    F = np.empty((6, V, nc, nc), dtype=np.complex128)
    lib().orc_clover_field_strength_nc(_ptr(F), _ptr(U), _dims(L), nc)
    return F


def wilson_clover_nc(U, F, psi, L, kappa, csw, phases=(1, 1, 1, -1), dagger=False):
    V = volume(L)
    psi = _c(psi)
    nc = psi.shape[-1]
    psi = psi.reshape(V, 4, nc)
    U = _c(U, (4, V, nc, nc))
    F = _c(F, (6, V, nc, nc))
    out = np.empty_like(psi)
    lib().orc_wilson_clover_nc(_ptr(out), _ptr(psi), _ptr(U), _ptr(F), _dims(L), _ptr(_phases(phases)),
                               float(kappa), float(csw), nc, int(dagger))
    return out


def wilson_pullback_nc(dres, psi, L, coefficient, phases=(1, 1, 1, -1), dU=None):
    V = volume(L)
    psi = _c(psi)
    nc = psi.shape[-1]
    psi = psi.reshape(V, 4, nc)
    dres = _c(dres, (V, 4, nc))
    dU = np.zeros((4, V, nc, nc), dtype=np.complex128) if dU is None else _c(dU, (4, V, nc, nc)).copy()
    lib().orc_wilson_pullback_nc(_ptr(dU), _ptr(dres), _ptr(psi), _dims(L), _ptr(_phases(phases)),
                                 float(coefficient), nc)
    return dU


def dot_n(a, b, nc, order=1):
    a = _c(a)
    b = _c(b)
    V = a.size // (4 * nc)
    out = np.zeros(2, dtype=np.float64)
    lib().orc_dot_n(_ptr(a), _ptr(b), V, 4 * nc, int(order), _ptr(out))
    return complex(out[0], out[1])


def wilson_pullback(dres, psi, L, coefficient, phases=(1, 1, 1, -1), dU=None):
    V = volume(L)
    dres = _c(dres, (V, 4, 3))
    psi = _c(psi, (V, 4, 3))
    dU = np.zeros((4, V, 3, 3), dtype=np.complex128) if dU is None else _c(dU, (4, V, 3, 3)).copy()
    ph = _phases(phases)
    lib().orc_wilson_pullback(_ptr(dU), _ptr(dres), _ptr(psi), _dims(L), _ptr(ph), float(coefficient))
    return dU


def clover_pullback(U, dres, psi, L, coefficient, csw, phases=(1, 1, 1, -1), dU=None):
    """wilson_clover_link_pullback! with this library's coefficient convention:
    dU += coefficient * hopping pullback + clover part with coefficient * csw."""
    V = volume(L)
    U = _c(U, (4, V, 3, 3))
    dres = _c(dres, (V, 4, 3))
    psi = _c(psi, (V, 4, 3))
    dU = np.zeros((4, V, 3, 3), dtype=np.complex128) if dU is None else _c(dU, (4, V, 3, 3)).copy()
    ph = _phases(phases)
    lib().orc_clover_pullback(_ptr(dU), _ptr(U), _ptr(dres), _ptr(psi), _dims(L), _ptr(ph), float(coefficient),
                              float(csw))
    return dU


def dsfdu_clover(U, phi, L, kappa, csw, phases=(1, 1, 1, -1), rtol=1e-12, maxiter=20000, order=1):
    """dSFdU for the Wilson-clover operator: (G, eta, chi, CGResult)."""
    V = volume(L)
    U = _c(U, (4, V, 3, 3))
    F = clover_field_strength(U, L)
    phi = _c(phi, (V, 4, 3))
    G = np.empty((4, V, 3, 3), dtype=np.complex128)
    eta = np.empty((V, 4, 3), dtype=np.complex128)
    chi = np.empty_like(eta)
    work = np.empty((4, V, 4, 3), dtype=np.complex128)
    ph = _phases(phases)
    res = _OrcResult()
    lib().orc_dsfdu_clover(_ptr(G), _ptr(eta), _ptr(chi), _ptr(U), _ptr(F), _ptr(phi), _dims(L), _ptr(ph),
                           float(kappa), float(csw), float(rtol), int(maxiter), int(order), _ptr(work),
                           ctypes.byref(res))
    return G, eta, chi, CGResult(bool(res.converged), int(res.iterations), float(res.residual_norm),
                                 float(res.relative_residual), REASONS[int(res.reason)])


def ta_force(U, G):
    U = _c(U)
    G = _c(G, U.shape)
    F = np.empty_like(U)
    lib().orc_ta_force(_ptr(F), _ptr(U), _ptr(G), U.size // 9)
    return F


def dsfdu(U, phi, L, kappa, phases=(1, 1, 1, -1), rtol=1e-12, maxiter=20000, order=1):
    """Returns (G, eta, chi, CGResult) with G = dS/dU in the Enzyme convention."""
    V = volume(L)
    U = _c(U, (4, V, 3, 3))
    phi = _c(phi, (V, 4, 3))
    G = np.empty((4, V, 3, 3), dtype=np.complex128)
    eta = np.empty((V, 4, 3), dtype=np.complex128)
    chi = np.empty_like(eta)
    work = np.empty((4, V, 4, 3), dtype=np.complex128)
    ph = _phases(phases)
    res = _OrcResult()
    lib().orc_dsfdu(_ptr(G), _ptr(eta), _ptr(chi), _ptr(U), _ptr(phi), _dims(L), _ptr(ph), float(kappa),
                    float(rtol), int(maxiter), int(order), _ptr(work), ctypes.byref(res))
    return G, eta, chi, CGResult(bool(res.converged), int(res.iterations), float(res.residual_norm),
                                 float(res.relative_residual), REASONS[int(res.reason)])


# ---------------------------------------------------------------------------
# Site random-number streams and lattice fills (lmx_rng_oracle.c)
# ---------------------------------------------------------------------------
RNG_ALGORITHMS = {"pcg32": 0, "xoshiro256++": 1, "philox4x32": 2}
_ELEM = {np.dtype(np.float32): 0, np.dtype(np.float64): 1, np.dtype(np.complex64): 2, np.dtype(np.complex128): 3}


def splitmix64(x: int) -> int:
    return int(lib().orc_splitmix64(x))


def rng_stream_seed(seed, sweep=0, direction=0, color=0, subgroup=0) -> int:
    """_rng_stream_seed(RNGStreamKey(seed, sweep, direction, color, subgroup))."""
    return int(lib().orc_rng_stream_seed(seed, sweep, direction, color, subgroup))


def pcg32_step(state: int, inc: int):
    st = ctypes.c_uint64(state)
    out = lib().orc_pcg32_step(ctypes.byref(st), inc)
    return int(st.value), int(out)


def philox4x32_10(ctr, key):
    c = np.asarray(ctr, dtype=np.uint32)
    k = np.asarray(key, dtype=np.uint32)
    o = np.zeros(4, dtype=np.uint32)
    lib().orc_philox4x32_10(_ptr(c), _ptr(k), _ptr(o))
    return tuple(int(v) for v in o)


def site_rng_draw(key, algorithm, first_site, nsites, words=0, uniform32=0, uniform64=0, normal32=0, normal64=0,
                  bounded=0, upper=17):
    """The per-site draw sequence of test/site_rng_gpu.jl:7-42; each output is
    [nsites, count] (the reference's [count, nsites] column-major)."""
    counts = np.array([words, uniform32, uniform64, normal32, normal64, bounded], dtype=np.int32)
    out = {"words": np.zeros((nsites, words), np.uint32), "uniform32": np.zeros((nsites, uniform32), np.float32),
           "uniform64": np.zeros((nsites, uniform64), np.float64), "normal32": np.zeros((nsites, normal32), np.float32),
           "normal64": np.zeros((nsites, normal64), np.float64), "bounded": np.zeros((nsites, bounded), np.uint32)}
    lib().orc_site_rng_draw(rng_stream_seed(*key), RNG_ALGORITHMS[algorithm], first_site, nsites, _ptr(counts),
                            upper, *[_ptr(out[k]) for k in ("words", "uniform32", "uniform64", "normal32", "normal64",
                                                            "bounded")])
    return out


def randomize_fill(dtype, nc1, nc2, L, key, gaussian=False, sigma=1.0, algorithm="philox4x32", offset=None,
                   gdims=None):
    """randomize_matrix! / randomize_gaussian_matrix! on the core of a field of
    local extents L (any dimension) placed at offset in a lattice of extents
    gdims (default: L is the whole lattice).  Returns [V, nc2, nc1] (site
    x-fastest), i.e. the reference's [nc1, nc2, L...] column-major core."""
    dt = np.dtype(dtype)
    D = len(L)
    Lc = (ctypes.c_int * D)(*[int(v) for v in L])
    off = (ctypes.c_int * D)(*([0] * D if offset is None else [int(v) for v in offset]))
    gd = (ctypes.c_int * D)(*[int(v) for v in (L if gdims is None else gdims)])
    out = np.zeros((volume_any(L), nc2, nc1), dtype=dt)
    lib().orc_randomize_fill(_ptr(out), _ELEM[dt], nc1, nc2, D, Lc, off, gd, int(bool(gaussian)), float(sigma),
                             rng_stream_seed(*key), RNG_ALGORITHMS[algorithm])
    return out


def volume_any(L):
    return int(np.prod([int(v) for v in L]))


# ---------------------------------------------------------------------------
# The reference's CPU algorithm (padded storage, halo refresh per operator
# application, unfused BLAS-1): the CPU baseline of bench.py
# ---------------------------------------------------------------------------
def refcpu_cg_ddagd(U, b, L, kappa, phases=(1, 1, 1, -1), rtol=1e-10, maxiter=5000):
    """cg! of D^dagger D as the reference runs it on the CPU, from a zero start.
    Returns (x [V, 4, 3], CGResult, seconds of the solve itself)."""
    import time
    V = volume(L)
    Vp = int(np.prod([int(d) + 2 for d in L]))
    U = _c(U, (4, V, 3, 3))
    b = _c(b, (V, 4, 3))
    ph = _phases(phases)
    Up = [np.zeros((Vp, 3, 3), dtype=np.complex128) for _ in range(4)]
    uptr = (ctypes.c_void_p * 4)(*[u.ctypes.data for u in Up])
    lib().orc_refcpu_pad_links(uptr, _ptr(U), _dims(L))
    bp = np.zeros((Vp, 4, 3), dtype=np.complex128)
    lib().orc_refcpu_pad_spinor(_ptr(bp), _ptr(b), _dims(L), _ptr(ph))
    x, r, p, Ap, tmp = (np.zeros_like(bp) for _ in range(5))
    res = _OrcResult()
    t0 = time.perf_counter()
    lib().orc_refcpu_cg_ddagd(_ptr(x), _ptr(bp), _ptr(r), _ptr(p), _ptr(Ap), _ptr(tmp), uptr, _dims(L), _ptr(ph),
                              float(kappa), float(rtol), int(maxiter), ctypes.byref(res))
    dt = time.perf_counter() - t0
    xc = np.zeros((V, 4, 3), dtype=np.complex128)
    lib().orc_refcpu_unpad_spinor(_ptr(xc), _ptr(x), _dims(L))
    return xc, CGResult(bool(res.converged), int(res.iterations), float(res.residual_norm),
                        float(res.relative_residual), REASONS[int(res.reason)]), dt
