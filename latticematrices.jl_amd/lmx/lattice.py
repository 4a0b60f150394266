"""Host-side mirror of the reference's LatticeMatrix operator API for the
Wilson(-clover) D^dagger D path, backed by liblmx.so.

Reference interface (LatticeMatrices.jl) -> this module:
  LatticeMatrix(NC1, NC2, dim, gsize, PEs; phases)  -> Lattice + SpinorField / GaugeField
  WilsonDiracOperator4D(U, kappa)                   -> WilsonDiracOperator4D(U, kappa)
  WilsonDiracOperator4D_Donly(U)                    -> WilsonDiracOperator4D_Donly(U)
  WilsonDiracCloverOperator4D(U, kappa, cSW)        -> WilsonDiracCloverOperator4D(U, kappa, cSW)
  adjoint(D) / D'                                   -> adjoint(D)
  mul!(C, A, psi)                                   -> mul_(C, A, psi)
  DdagDOp(D, temp)                                  -> DdagDOp(D, temp)
  cg!(x, A, b, r, p, Ap; rtol, atol, maxiter)       -> cg_(x, A, b, r, p, Ap, rtol=, atol=, maxiter=)
  solve!(x, A::DdagDOp, b, r, p, Ap; kw...)         -> solve_(...)
  pseudofermion_action(D, phi, eta, Deta, r, p, Ap) -> pseudofermion_action(...)
  dSFdU(dfdU, D, phi)                               -> dSFdU(dfdU, D, phi)
  axpby!, dot, clear_matrix!, substitute!           -> axpby_, dot, clear_matrix_, substitute_
Errors: ArgumentError for bad geometry, aliasing and tolerances (as the
reference); CG non-convergence is a CGResult status, not an exception.

Fields live on the GPU (torch tensors used as device memory) in the
library's structure-of-arrays layout; numpy arrays in the reference's
column-major layout ([V,4,3] spinors, [4,V,3,3] links, optionally with a halo
of width nw) are converted on the device by the C ABI.
"""
from __future__ import annotations

import ctypes
from dataclasses import dataclass
from numbers import Number

import numpy as np

from . import _lib
from ._lib import ArgumentError, check, lib

try:
    import torch
except ImportError:  # pragma: no cover - torch is part of the image
    torch = None


def _require_gpu():
    if torch is None or not torch.cuda.is_available():
        raise RuntimeError("liblmx requires a ROCm GPU (torch.cuda.is_available() is False)")


def _ptr(t) -> int:
    return t.data_ptr()


class Lattice:
    """Geometry, spinor boundary phases and the HIP stream of one lattice
    (the single-rank LatticeMatrix construction, src/LatticeMatrices_core.jl:297-343).
    nc: colours of its fields (NC1 of LatticeMatrix; 3, or 2 for the
    reference's generic-NC kernels)."""

    def __init__(self, dims, phases=(1, 1, 1, 1), device=None, nc: int = 3):
        _require_gpu()
        dims = tuple(int(d) for d in dims)
        if len(dims) != 4:
            raise ArgumentError(_lib.LMX_ERR_ARG, f"global size must have 4 entries, got {len(dims)}")
        if len(phases) != 4:
            raise ArgumentError(_lib.LMX_ERR_ARG, f"phases must have 4 entries, got {len(phases)}")
        self.dims = dims
        self.phases = tuple(complex(p) for p in phases)
        self.device = torch.device("cuda", torch.cuda.current_device() if device is None else device)
        self.V = int(np.prod(dims))
        self.nc = int(nc)
        self.ncomp = 4 * self.nc  # complex spinor components per site
        with torch.cuda.device(self.device):
            self.stream = torch.cuda.current_stream(self.device)
        ph = (ctypes.c_double * 8)(*[v for p in self.phases for v in (p.real, p.imag)])
        d = (ctypes.c_int32 * 4)(*dims)
        h = ctypes.c_void_p()
        check(lib().lmx_context_create_nc(ctypes.byref(h), d, ph, ctypes.c_void_p(self.stream.cuda_stream),
                                          self.nc))
        self._ctx = h

    @property
    def ctx(self):
        return self._ctx

    def synchronize(self):
        check(lib().lmx_context_synchronize(self._ctx))

    def __del__(self):
        try:
            if getattr(self, "_ctx", None):
                lib().lmx_context_destroy(self._ctx)
                self._ctx = None
        except Exception:
            pass

    # -- field factories -------------------------------------------------
    def spinor(self, nrhs: int = 1) -> "SpinorField":
        return SpinorField(self, nrhs)

    def spinor_from_numpy(self, arr, nw: int = 0) -> "SpinorField":
        return SpinorField.from_numpy(self, arr, nw)

    def gauge_from_numpy(self, U, nw: int = 0) -> "GaugeField":
        return GaugeField.from_numpy(self, U, nw)


class SpinorField:
    """nrhs Dirac spinor fields (nc colours x 4 spins per site), FP64, in the
    device layout double2[nrhs][4 nc][V]."""

    def __init__(self, lattice: Lattice, nrhs: int = 1, data=None):
        self.lattice = lattice
        self.nrhs = int(nrhs)
        if data is None:
            data = torch.zeros((self.nrhs, lattice.ncomp, lattice.V, 2), dtype=torch.float64, device=lattice.device)
        self.data = data

    @property
    def ptr(self):
        return ctypes.c_void_p(_ptr(self.data))

    def similar(self) -> "SpinorField":
        return SpinorField(self.lattice, self.nrhs)

    @classmethod
    def from_numpy(cls, lattice: Lattice, arr, nw: int = 0) -> "SpinorField":
        """arr: complex [V,4,nc] / [nrhs,V,4,nc] (nw = 0) or the padded reference
        array(s) [nrhs, Lt+2nw, Lz+2nw, Ly+2nw, Lx+2nw, 4, nc] in C order."""
        a = np.ascontiguousarray(arr, dtype=np.complex128)
        per = int(np.prod([d + 2 * nw for d in lattice.dims])) * lattice.ncomp
        if a.size % per != 0:
            raise ArgumentError(_lib.LMX_ERR_ARG, f"array of {a.size} elements is not a multiple of {per}")
        nrhs = a.size // per
        f = cls(lattice, nrhs)
        src = torch.from_numpy(a.reshape(-1)).to(lattice.device)
        check(lib().lmx_spinor_import(lattice.ctx, f.ptr, ctypes.c_void_p(_ptr(src)), int(nw), nrhs))
        lattice.synchronize()
        return f

    def to_padded(self, nw: int = 0):
        L = self.lattice
        n = self.nrhs * int(np.prod([d + 2 * nw for d in L.dims])) * L.ncomp
        dst = torch.empty(n, dtype=torch.complex128, device=L.device)
        check(lib().lmx_spinor_export(L.ctx, ctypes.c_void_p(_ptr(dst)), int(nw), self.ptr, self.nrhs))
        L.synchronize()
        return dst.cpu().numpy()

    def to_numpy(self):
        """[V,4,nc] (nrhs == 1) or [nrhs,V,4,nc] complex128."""
        a = self.to_padded(0).reshape(self.nrhs, self.lattice.V, 4, self.lattice.nc)
        return a[0] if self.nrhs == 1 else a

    def rhs(self, i: int) -> "SpinorField":
        """View of right-hand side i as a single field."""
        return SpinorField(self.lattice, 1, self.data[i:i + 1])


class GaugeField:
    """The four link fields U_1..U_4 (nc x nc each), FP64, device layout double2[4][nc nc][V]."""

    def __init__(self, lattice: Lattice, data=None):
        self.lattice = lattice
        if data is None:
            data = torch.zeros((4, lattice.nc * lattice.nc, lattice.V, 2), dtype=torch.float64, device=lattice.device)
        self.data = data
        self._core_epoch = 0

    @property
    def ptr(self):
        return ctypes.c_void_p(_ptr(self.data))

    def mark_halo_dirty_(self):
        """mark_halo_dirty!(U) (src/LatticeMatrices_core.jl:196-203): record a
        write to the links that torch's in-place operations do not count (one
        through a raw device pointer, or through `tensor.data`), so that
        mul_cached_clover_ refreshes the operator's derived copies."""
        self._core_epoch += 1

    @property
    def core_epoch(self):
        """The links' version: (explicit epoch, torch's in-place version
        counter of .data, device address of .data).  Any in-place torch write,
        a replaced .data tensor or mark_halo_dirty_ changes it (the reference's
        halo_epoch.core, src/LatticeMatrices_core.jl:196-214)."""
        return (self._core_epoch, self.data._version, self.data.data_ptr())

    @classmethod
    def from_numpy(cls, lattice: Lattice, U, nw: int = 0) -> "GaugeField":
        """U: complex [4, V, nc, nc] (mu, site, row, col) with nw = 0, or four padded
        reference link arrays (any container of 4 arrays of V_padded*nc*nc values
        in the reference's column-major element order when nw > 0)."""
        g = cls(lattice)
        nc = lattice.nc
        if nw == 0 and not isinstance(U, (list, tuple)):
            a = np.ascontiguousarray(U, dtype=np.complex128).reshape(4, lattice.V, nc, nc)
            # reference element order within a site is column-major (row fastest)
            arrs = [np.ascontiguousarray(np.swapaxes(a[mu], -1, -2)).reshape(-1) for mu in range(4)]
        else:
            arrs = [np.ascontiguousarray(u, dtype=np.complex128).reshape(-1) for u in U]
        dev = [torch.from_numpy(x).to(lattice.device) for x in arrs]
        ptrs = (ctypes.c_void_p * 4)(*[_ptr(t) for t in dev])
        check(lib().lmx_links_import(lattice.ctx, g.ptr, ptrs, int(nw)))
        lattice.synchronize()
        return g

    def to_numpy(self):
        """[4, V, nc, nc] complex128 (mu, site, row, col)."""
        L = self.lattice
        nc = L.nc
        dev = [torch.empty(L.V * nc * nc, dtype=torch.complex128, device=L.device) for _ in range(4)]
        ptrs = (ctypes.c_void_p * 4)(*[_ptr(t) for t in dev])
        check(lib().lmx_links_export(L.ctx, ptrs, 0, self.ptr))
        L.synchronize()
        out = np.stack([t.cpu().numpy().reshape(L.V, nc, nc) for t in dev])
        return np.ascontiguousarray(np.swapaxes(out, -1, -2))


# ---------------------------------------------------------------------------
# Operators
# ---------------------------------------------------------------------------
class WilsonDiracOperator4D:
    """D = 1 - kappa sum_mu [U_mu(x)(1-g_mu) psi(x+mu) + U_mu(x-mu)^dagger (1+g_mu) psi(x-mu)]
    (src/Operators/WilsonDiracOperator.jl:2-35,57-76).  csw != 0 gives the
    Wilson-clover operator (WilsonDiracCloverOperator4D)."""

    def __init__(self, U: GaugeField, kappa: float, csw: float = 0.0):
        self.lattice = U.lattice
        self.U = U
        self.csw = float(csw)
        h = ctypes.c_void_p()
        check(lib().lmx_wilson_create(self.lattice.ctx, ctypes.byref(h), U.ptr, float(kappa), float(csw)))
        self._op = h
        self._kappa = float(kappa)
        # refreshes of the derived copies (clover cache, low-precision links),
        # and the links and their version those copies were built from
        # (CloverCacheState, src/Operators/WilsonDiracCloverOperator.jl:417-432)
        self.cache_epoch = 0
        self._source = (U, U.core_epoch)

    @property
    def kappa(self):
        return self._kappa

    @kappa.setter
    def kappa(self, value):
        check(lib().lmx_wilson_set_kappa(self._op, float(value)))
        self._kappa = float(value)

    def update_links(self, U: GaugeField | None = None):
        """update_clover!(operator[, U]) (src/Operators/WilsonDiracCloverOperator.jl:468-479)."""
        if U is not None:
            if U.lattice.dims != self.lattice.dims:
                raise ArgumentError(_lib.LMX_ERR_ARG, "links of another lattice geometry")
            self.U = U
        check(lib().lmx_wilson_update_links(self._op, self.U.ptr))
        self.cache_epoch += 1
        self._source = (self.U, self.U.core_epoch)

    def clover_planes(self):
        """The six clover field-strength planes [6, V, nc, nc] (CloverFieldStrength4D)."""
        return self.clover_planes_device().cpu().numpy()

    def clover_planes_device(self):
        """The same planes as a complex128 device tensor [6, V, nc, nc] (plane, site, row, col)."""
        L = self.lattice
        out = torch.empty((6, L.V, L.nc, L.nc), dtype=torch.complex128, device=L.device)
        check(lib().lmx_clover_export(self._op, ctypes.c_void_p(_ptr(out))))
        L.synchronize()
        return out

    def __del__(self):
        try:
            if getattr(self, "_op", None):
                lib().lmx_wilson_destroy(self._op)
                self._op = None
        except Exception:
            pass


def WilsonDiracCloverOperator4D(U: GaugeField, kappa: float, cSW: float) -> WilsonDiracOperator4D:
    """D = D_W - kappa cSW sum_{mu<nu} gamma_mu gamma_nu F_mu,nu
    (src/Operators/WilsonDiracCloverOperator.jl:403-456)."""
    return WilsonDiracOperator4D(U, kappa, cSW)


class CloverFieldStrength4D:
    """CloverFieldStrength4D(U) (src/Operators/WilsonDiracCloverOperator.jl:1-33,
    386-399): the six planes (12, 13, 14, 23, 24, 34) of F = (Q - Q^dagger)/8,
    Q the four-leaf plaquette sum, built on the device by the operator's
    clover kernel.  len(F) == 6; F[i] is a complex128 device tensor
    [V, 3, 3] (site, row, col)."""

    def __init__(self, U: GaugeField):
        self.lattice = U.lattice
        self.components = ()
        update_clover_(self, U)

    def __len__(self):
        return 6

    def __getitem__(self, i):
        return self.components[i]


def update_clover_(target, U: GaugeField | None = None):
    """update_clover!(F::CloverFieldStrength4D, U) recomputes the six planes
    from U (src/Operators/WilsonDiracCloverOperator.jl:392-399);
    update_clover!(operator[, U]) refreshes an operator's cache, optionally
    from new links (:462-479)."""
    if isinstance(target, CloverFieldStrength4D):
        if not isinstance(U, GaugeField):
            raise TypeError("update_clover_(F, U) needs the links U as a GaugeField")
        if U.lattice.dims != target.lattice.dims:
            raise ArgumentError(_lib.LMX_ERR_ARG, "links of another lattice geometry")
        op = WilsonDiracOperator4D(U, 0.125, 1.0)  # any nonzero cSW builds the field strength
        planes = op.clover_planes_device()
        del op
        target.components = tuple(planes[i] for i in range(6))
        return target
    if isinstance(target, WilsonDiracOperator4D):
        target.update_links(U)
        return target
    raise TypeError(f"update_clover_ needs a CloverFieldStrength4D or a Wilson-clover operator, not "
                    f"{type(target).__name__}")


def _ensure_clover_cache_current(cache: WilsonDiracOperator4D, U: GaugeField) -> WilsonDiracOperator4D:
    """_ensure_clover_cache_current! (src/Operators/WilsonDiracCloverOperator.jl:603-623):
    refresh (update_clover!(cache, U)) unless the cache was built from this
    very GaugeField at its current version."""
    if not isinstance(cache, WilsonDiracOperator4D):
        raise TypeError(f"not a Wilson(-clover) operator: {type(cache).__name__}")
    if not isinstance(U, GaugeField):
        raise TypeError(f"links must be a GaugeField, not {type(U).__name__}")
    src, epoch = cache._source
    if src is U and epoch == U.core_epoch:
        return cache
    cache.update_links(U)
    return cache


def mul_cached_clover_(result: SpinorField, cache: WilsonDiracOperator4D, U: GaugeField,
                       psi: SpinorField) -> SpinorField:
    """mul_cached_clover!(result, cache, U1, U2, U3, U4, psi)
    (src/Operators/WilsonDiracCloverOperator.jl:625-651): apply the operator
    with the links taken explicitly from U (U1..U4 are its four directions),
    refreshing the clover cache first when U is not the GaugeField the cache
    was built from or has been written since."""
    _ensure_clover_cache_current(cache, U)
    return mul_(result, cache, psi)


def mul_cached_clover_adjoint_(result: SpinorField, cache: WilsonDiracOperator4D, U: GaugeField,
                               psi: SpinorField) -> SpinorField:
    """mul_cached_clover_adjoint!(result, cache, U1, U2, U3, U4, psi)
    (src/Operators/WilsonDiracCloverOperator.jl:653-668): the adjoint, with the
    same refreshed cache."""
    _ensure_clover_cache_current(cache, U)
    return mul_(result, adjoint(cache), psi)


class WilsonDiracOperator4D_Donly:
    """(1/2) sum_mu [U(1-g)psi(x+mu) + U^dagger(1+g)psi(x-mu)] (:522-551)."""

    def __init__(self, U: GaugeField):
        self._w = WilsonDiracOperator4D(U, 0.0)
        self.U = U
        self.lattice = U.lattice


@dataclass
class Adjoint:
    parent: object


def adjoint(A):
    """adjoint(D) is involutive; adjoint(DdagDOp) is itself (DiracOperators.jl:102)."""
    if isinstance(A, Adjoint):
        return A.parent
    if isinstance(A, DdagDOp):
        return A
    return Adjoint(A)


class FieldPool:
    """PreallocatedArray with get_block / unused! (the reference's work-field
    pool): fields made by `make` on first use, handed out and returned by
    index; the pool grows when every field is in use.  A pool over a fixed
    list (FieldPool.of) cannot grow."""

    def __init__(self, make, num: int = 0):
        self._make = make
        self._fields = [None] * int(num)
        self._used = [False] * int(num)

    @classmethod
    def of(cls, fields):
        pool = cls(None)
        pool._fields = list(fields)
        pool._used = [False] * len(pool._fields)
        return pool

    def __len__(self):
        return len(self._fields)

    def get_block(self, n: int | None = None):
        """One field and its index (n None), or lists of n fields and indices."""
        k = 1 if n is None else int(n)
        idx = [i for i, u in enumerate(self._used) if not u][:k]
        while len(idx) < k:
            if self._make is None:
                raise ArgumentError(_lib.LMX_ERR_ARG, f"work-field pool of {len(self._fields)} fields has fewer "
                                                      f"than {k} free")
            self._fields.append(None)
            self._used.append(False)
            idx.append(len(self._fields) - 1)
        for i in idx:
            if self._fields[i] is None:
                self._fields[i] = self._make()
            self._used[i] = True
        blk = [self._fields[i] for i in idx]
        return (blk[0], idx[0]) if n is None else (blk, idx)

    def unused_(self, idx):
        for i in ([idx] if isinstance(idx, int) else idx):
            self._used[i] = False


class DiracOp:
    """DiracOp(U, apply, apply_dag, p, phi; numtemp=4, numphitemp=4)
    (src/Operators/DiracOperators.jl:21-46,48-77): an operator given by two
    callbacks, apply(y, U, x, p, phitemps, temps) for D and apply_dag(...) for
    D^dagger, each called by mul_ with four borrowed spinor work fields
    (similar to phi) and four borrowed link work fields (GaugeField; U and the
    link temporaries hold all four directions U1..U4 of the reference's
    call).  adjoint(DiracOp) applies apply_dag; DdagDOp(D) without a temp
    borrows one from D.phitemps."""

    def __init__(self, U: GaugeField, apply, apply_dag, p, phi: SpinorField, numtemp: int = 4,
                 numphitemp: int = 4):
        self.U, self.apply, self.apply_dag, self.p = U, apply, apply_dag, p
        self.temps = FieldPool(lambda: GaugeField(U.lattice), numtemp)
        self.phitemps = FieldPool(phi.similar, numphitemp)


@dataclass
class DdagDOp:
    """Allocation-free normal operator D^dagger D with a caller-owned temp
    (src/Operators/DiracOperators.jl:93-121).  DdagDOp(D) with a DiracOp and
    no temp borrows the temporary from D.phitemps on each application (the
    compatibility constructor, :98)."""
    D: object
    temp: SpinorField | None = None

    def __post_init__(self):
        if self.temp is None and not isinstance(self.D, DiracOp):
            raise TypeError("DdagDOp(D) without a temporary field needs a DiracOp")


def _wilson_object(A):
    """The WilsonDiracOperator4D behind a Wilson-family operator or its adjoint."""
    if isinstance(A, Adjoint):
        return _wilson_object(A.parent)
    if isinstance(A, WilsonDiracOperator4D_Donly):
        return A._w
    return A if isinstance(A, WilsonDiracOperator4D) else None


def _wilson_handle(A):
    if isinstance(A, WilsonDiracOperator4D):
        return A._op, 0
    if isinstance(A, WilsonDiracOperator4D_Donly):
        return A._w._op, _lib.LMX_DONLY
    if isinstance(A, Adjoint):
        h, f = _wilson_handle(A.parent)
        return h, f | _lib.LMX_DAGGER
    raise TypeError(f"not a Wilson operator: {type(A).__name__}")


def mul_(C: SpinorField, A, psi: SpinorField) -> SpinorField:
    """mul!(C, A, psi) for the Wilson family, DdagDOp, or a complex scalar
    (mul!(C, a::Number, x), src/LinearAlgebras/mul.jl:2-9)."""
    if C.nrhs != psi.nrhs:
        raise ArgumentError(_lib.LMX_ERR_ARG, "C and psi must hold the same number of fields")
    if isinstance(A, Number):
        a = complex(A)
        check(lib().lmx_scale(C.lattice.ctx, C.ptr, psi.ptr, a.real, a.imag, C.nrhs))
        return C
    if isinstance(A, DiracOp) or (isinstance(A, Adjoint) and isinstance(A.parent, DiracOp)):
        D = A if isinstance(A, DiracOp) else A.parent
        fn = D.apply if D is A else D.apply_dag
        temps, it = D.temps.get_block(4)
        phis, ip = D.phitemps.get_block(4)
        try:
            fn(C, D.U, psi, D.p, phis, temps)
        finally:
            D.temps.unused_(it)
            D.phitemps.unused_(ip)
        return C
    if isinstance(A, DdagDOp) and A.temp is None:
        temp, i = A.D.phitemps.get_block()
        try:
            mul_(temp, A.D, psi)
            mul_(C, adjoint(A.D), temp)
        finally:
            A.D.phitemps.unused_(i)
        return C
    if isinstance(A, DdagDOp):
        if A.temp is psi or A.temp is C:
            raise ArgumentError(_lib.LMX_ERR_ALIAS, "DdagDOp temporary field must not alias its input or output")
        if isinstance(A.D, WilsonDiracOperator4D):
            check(lib().lmx_ddagd_apply(A.D._op, C.ptr, psi.ptr, A.temp.ptr, C.nrhs))
            return C
        mul_(A.temp, A.D, psi)
        mul_(C, adjoint(A.D), A.temp)
        return C
    if isinstance(A, ScaledOperator):
        return mul_(C, A.scale, psi)
    if isinstance(A, Adjoint) and isinstance(A.parent, ScaledOperator):
        return mul_(C, complex(A.parent.scale).conjugate(), psi)
    h, flags = _wilson_handle(A)
    check(lib().lmx_wilson_apply(h, C.ptr, psi.ptr, flags, C.nrhs))
    return C


def spinor_f32(lattice: Lattice, nrhs: int = 1):
    """A zero single-precision spinor field: float32 device tensor [nrhs, 12, V, 2]
    (component spin*3+colour, (re, im)) -- LatticeMatrix{ComplexF32} in the
    library's layout."""
    return torch.zeros((nrhs, 12, lattice.V, 2), dtype=torch.float32, device=lattice.device)


def spinor_f32_from_numpy(lattice: Lattice, arr):
    """[V, 4, 3] or [nrhs, V, 4, 3] complex (site, spin, colour) -> float32 [nrhs, 12, V, 2]."""
    a = np.asarray(arr, dtype=np.complex64)
    if a.ndim == 3:
        a = a[None]
    if a.shape[1:] != (lattice.V, 4, 3):
        raise ArgumentError(_lib.LMX_ERR_ARG, f"spinor array of shape {a.shape} does not match V = {lattice.V}")
    a = np.ascontiguousarray(a.reshape(a.shape[0], lattice.V, 12).transpose(0, 2, 1))
    return torch.view_as_real(torch.from_numpy(a)).contiguous().to(lattice.device)


def spinor_f32_to_numpy(t):
    """float32 [nrhs, 12, V, 2] -> complex64 [nrhs, V, 4, 3]."""
    c = torch.view_as_complex(t.detach().contiguous()).cpu().numpy()
    return np.ascontiguousarray(c.transpose(0, 2, 1)).reshape(c.shape[0], c.shape[2], 4, 3)


def _refresh_low_precision(op):
    """The single-precision and mixed paths read the operator's rounded link
    (and clover) copies, built from the links when the operator was created
    or last updated; the FP64 path reads the links themselves.  Rebuild the
    copies when the links object or its version (GaugeField.core_epoch: an
    in-place torch write, a replaced .data, mark_halo_dirty_) changed since,
    so that every precision applies the same operator."""
    src, epoch = op._source
    if src is not op.U or epoch != op.U.core_epoch:
        op.update_links()


def mul_f32_(C, A, psi):
    """mul!(C, A, psi) on single-precision fields (ComplexF32,
    test/wilson_halfspin.jl:105,124-130): C and psi are float32 device tensors
    [nrhs, 12, V, 2] (spinor_f32); A is a Wilson(-clover) operator, its
    adjoint, or the hopping-only operator.  FP32 links (the operator's,
    rounded) and FP32 arithmetic."""
    h, flags = _wilson_handle(A)
    for t in (C, psi):
        if not (isinstance(t, torch.Tensor) and t.dtype == torch.float32 and t.is_contiguous() and t.dim() == 4
                and t.shape[1:] == (12, t.shape[2], 2)):
            raise ArgumentError(_lib.LMX_ERR_ARG, "single-precision fields must be contiguous float32 [nrhs, 12, V, 2]")
    if C.shape != psi.shape:
        raise ArgumentError(_lib.LMX_ERR_ARG, "C and psi must have the same shape")
    op = A.parent if isinstance(A, Adjoint) else A
    lat = op.lattice
    if C.shape[2] != lat.V or C.device != lat.device:
        raise ArgumentError(_lib.LMX_ERR_ARG, "fields do not match the operator's lattice")
    _refresh_low_precision(_wilson_object(A))
    check(lib().lmx_wilson_apply_f32(h, ctypes.c_void_p(_ptr(C)), ctypes.c_void_p(_ptr(psi)), flags, C.shape[0]))
    return C


def mul_padded_(C, A, psi, nw: int):
    """mul!(C, A, psi) on arrays in the reference's own storage (no
    conversion): C and psi are device complex128 (ComplexF64) or complex64
    (ComplexF32) tensors laid out as LatticeMatrix.A, i.e. C-order [nrhs?,
    Lt+2nw, Lz+2nw, Ly+2nw, Lx+2nw, 4, nc] (column-major [nc,4,Lx+2nw,...]).
    psi's halo is ignored (in-kernel wrap with the lattice's phases); only C's
    core sites are written.  A is a Wilson(-clover) operator, its adjoint, or
    the hopping-only operator; complex64 fields use FP32 links and arithmetic."""
    h, flags = _wilson_handle(A)
    lat = (A.parent if isinstance(A, Adjoint) else A).lattice
    per = int(np.prod([d + 2 * nw for d in lat.dims])) * lat.ncomp
    dt = psi.dtype if isinstance(psi, torch.Tensor) else None
    for t in (C, psi):
        if not (isinstance(t, torch.Tensor) and t.dtype in (torch.complex128, torch.complex64) and t.dtype == dt
                and t.is_contiguous() and t.device == lat.device):
            raise ArgumentError(_lib.LMX_ERR_ARG, "padded fields must be contiguous device tensors, both complex128 "
                                                  "or both complex64")
        if t.numel() % per != 0:
            raise ArgumentError(_lib.LMX_ERR_ARG, f"array of {t.numel()} elements is not a multiple of {per}")
    if C.numel() != psi.numel():
        raise ArgumentError(_lib.LMX_ERR_ARG, "C and psi must hold the same number of fields")
    if dt == torch.complex64:
        _refresh_low_precision(_wilson_object(A))
    fn = lib().lmx_wilson_apply_padded if dt == torch.complex128 else lib().lmx_wilson_apply_padded_f32
    check(fn(h, ctypes.c_void_p(_ptr(C)), ctypes.c_void_p(_ptr(psi)), int(nw), flags, psi.numel() // per))
    return C


@dataclass
class ScaledOperator:
    """A = scale * I with a device apply (the _CGScaledOperator of test/cg.jl:1-10)."""
    scale: complex


# ---------------------------------------------------------------------------
# BLAS-1
# ---------------------------------------------------------------------------
def axpby_(a: float, X: SpinorField, b: float, Y: SpinorField) -> SpinorField:
    """Y = a X + b Y (src/LinearAlgebras/linearalgebra_D.jl:2-23)."""
    check(lib().lmx_axpby(Y.lattice.ctx, float(a), X.ptr, float(b), Y.ptr, Y.nrhs))
    return Y


def dot(A: SpinorField, B: SpinorField) -> complex:
    """sum conj(A) B (src/LinearAlgebras/linearalgebra_D.jl:3325-3341)."""
    out = (ctypes.c_double * 2)()
    check(lib().lmx_dot(A.lattice.ctx, A.ptr, B.ptr, out))
    return complex(out[0], out[1])


def add_matrix_(C: SpinorField, A: SpinorField, alpha: complex = 1.0) -> SpinorField:
    """C += alpha * A (add_matrix!, src/LinearAlgebras/linearalgebra_D.jl:2882-2911:
    u += alpha * v per entry, unfused).  Real alpha is axpby(alpha, A, 1, C)
    (1 * C is exact and the sum commutes); complex alpha forms the product with
    the library's complex scale first, then adds."""
    z = complex(alpha)
    if z.imag == 0.0:
        return axpby_(z.real, A, 1.0, C)
    t = A.similar()
    check(lib().lmx_scale(A.lattice.ctx, t.ptr, A.ptr, z.real, z.imag, A.nrhs))
    return axpby_(1.0, t, 1.0, C)


def clear_matrix_(A: SpinorField) -> SpinorField:
    A.data.zero_()
    return A


def substitute_(A: SpinorField, B: SpinorField) -> SpinorField:
    A.data.copy_(B.data)
    return A


# ---------------------------------------------------------------------------
# Solvers
# ---------------------------------------------------------------------------
@dataclass
class CGResult:
    """CGResult (src/LinearAlgebras/cg.jl:7-13) + true residual / reliable updates."""
    converged: bool
    iterations: int
    residual_norm: float
    relative_residual: float
    reason: str
    true_relative_residual: float = -1.0
    reliable_updates: int = 0

    @classmethod
    def from_c(cls, r):
        return cls(bool(r.converged), int(r.iterations), float(r.residual_norm), float(r.relative_residual),
                   _lib.REASONS[int(r.reason)], float(r.true_relative_residual), int(r.reliable_updates))


def _aliases(fields):
    names = ("x", "b", "r", "p", "Ap")
    for i in range(len(fields)):
        for j in range(i + 1, len(fields)):
            if fields[i] is fields[j] or fields[i].data.data_ptr() == fields[j].data.data_ptr():
                raise ArgumentError(_lib.LMX_ERR_ALIAS, f"CG fields {names[i]} and {names[j]} must not alias")


def _cg_f32(x, A, b, r, p, Ap, params) -> CGResult:
    """cg! on single-precision fields (float32 [1, 12, V, 2], spinor_f32):
    A must be DdagDOp of a Wilson(-clover) operator; its temp is used when it
    is such a field, else one is allocated."""
    names = ("x", "b", "r", "p", "Ap")
    fields = (x, b, r, p, Ap)
    if not isinstance(A, DdagDOp) or not isinstance(A.D, WilsonDiracOperator4D):
        raise ArgumentError(_lib.LMX_ERR_UNSUPPORTED, "cg_ on single-precision fields is built for DdagDOp of a "
                                                      "Wilson(-clover) operator")
    lat = A.D.lattice
    for n, t in zip(names, fields):
        if not (isinstance(t, torch.Tensor) and t.dtype == torch.float32 and t.is_contiguous()
                and tuple(t.shape) == (1, 12, lat.V, 2) and t.device == lat.device):
            raise ArgumentError(_lib.LMX_ERR_ARG, f"CG field {n} must be a contiguous float32 [1, 12, V, 2] field "
                                                  "on the operator's device")
    for i in range(5):
        for j in range(i + 1, 5):
            if fields[i].data_ptr() == fields[j].data_ptr():
                raise ArgumentError(_lib.LMX_ERR_ALIAS, f"CG fields {names[i]} and {names[j]} must not alias")
    tmp = A.temp if (isinstance(A.temp, torch.Tensor) and A.temp.dtype == torch.float32
                     and tuple(A.temp.shape) == (1, 12, lat.V, 2)) else spinor_f32(lat, 1)
    _refresh_low_precision(A.D)
    res = _lib.CGResultC()
    check(lib().lmx_cg_ddagd_f32(A.D._op, ctypes.c_void_p(_ptr(x)), ctypes.c_void_p(_ptr(b)),
                                 ctypes.c_void_p(_ptr(r)), ctypes.c_void_p(_ptr(p)), ctypes.c_void_p(_ptr(Ap)),
                                 ctypes.c_void_p(_ptr(tmp)), ctypes.byref(params), ctypes.byref(res)))
    return CGResult.from_c(res)


def cg_(x, A, b, r, p, Ap, rtol=1e-10, atol=0.0, maxiter=5000, check_every=0) -> CGResult:
    """cg!(x, A, b, r, p, Ap; rtol, atol, maxiter) (src/LinearAlgebras/cg.jl:73-137).
    The element type follows the fields: SpinorField (ComplexF64), or
    single-precision fields (spinor_f32 tensors, ComplexF32) for D^dagger D of
    a Wilson(-clover) operator."""
    if int(maxiter) < 0:
        raise ArgumentError(_lib.LMX_ERR_ARG, "maxiter must be nonnegative")
    params = _lib.default_params(rtol=float(rtol), atol=float(atol), maxiter=int(maxiter),
                                 check_every=int(check_every))
    if isinstance(x, torch.Tensor):
        return _cg_f32(x, A, b, r, p, Ap, params)
    _aliases((x, b, r, p, Ap))
    res = _lib.CGResultC()
    if isinstance(A, DdagDOp) and isinstance(A.D, WilsonDiracOperator4D) and A.temp is not None:
        check(lib().lmx_cg_ddagd(A.D._op, x.ptr, b.ptr, r.ptr, p.ptr, Ap.ptr, A.temp.ptr, ctypes.byref(params),
                                 ctypes.byref(res)))
        return CGResult.from_c(res)
    lattice = x.lattice
    shapes = {f.data.data_ptr(): f.data for f in (x, b, r, p, Ap)}

    def apply(user, out, inp, stream):
        try:
            o = SpinorField(lattice, 1, _wrap_ptr(out, lattice, shapes))
            i = SpinorField(lattice, 1, _wrap_ptr(inp, lattice, shapes))
            mul_(o, A, i)
            return 0
        except Exception:  # pragma: no cover - reported as a status
            return 1

    cb = _lib.APPLY_FN(apply)
    check(lib().lmx_cg_generic(lattice.ctx, cb, None, x.ptr, b.ptr, r.ptr, p.ptr, Ap.ptr, ctypes.byref(params),
                               ctypes.byref(res)))
    return CGResult.from_c(res)


def _wrap_ptr(ptr, lattice, cache):
    """Recover the SpinorField tensor for a device pointer handed back by the
    generic CG callback (one of the caller's five fields)."""
    return cache[ptr]


def cg(x, A, b, temps, eps=1e-10, maxsteps=5000, verboselevel=2):
    """cg(x, A, b, temps; eps, maxsteps, verboselevel)
    (src/LinearAlgebras/cg.jl:140-198), the compatibility interface: r, p and
    Ap are borrowed from temps (a FieldPool such as DiracOp.phitemps, or a
    list of spinor fields), x holds the initial guess; returns None on
    convergence and raises RuntimeError otherwise."""
    pool = temps if isinstance(temps, FieldPool) else FieldPool.of(temps)
    (r, p, Ap), idx = pool.get_block(3)
    try:
        res = cg_(x, A, b, r, p, Ap, rtol=eps, atol=0.0, maxiter=maxsteps)
    finally:
        pool.unused_(idx)
    if res.converged:
        if verboselevel >= 3:
            print(f"Converged at {res.iterations}-th step. eps: {res.relative_residual}")
        return None
    raise RuntimeError(f"The CG is not converged! with maxsteps = {maxsteps}\nresidual is {res.relative_residual}\n"
                       f"reason is {res.reason}\nmaxsteps should be larger.")


def solve_(x, A: DdagDOp, b, r=None, p=None, Ap=None, verboselevel=2, **kw):
    """solve!(x, A::DdagDOp, b, r, p, Ap; kwargs...) (DiracOperators.jl:124-132),
    returning the CGResult; solve!(x, A::DdagDOp{<:DiracOp}, b) (:134-143), the
    compatibility form without work fields, borrows them from A.D.phitemps
    and returns None on convergence (raises otherwise)."""
    if r is None and p is None and Ap is None:
        if not (isinstance(A, DdagDOp) and isinstance(A.D, DiracOp)):
            raise TypeError("solve_ without work fields needs a DdagDOp of a DiracOp")
        return cg(x, A, b, A.D.phitemps, verboselevel=verboselevel)
    return cg_(x, A, b, r, p, Ap, **kw)


def pseudofermion_action(D, phi, eta=None, Deta=None, r=None, p=None, Ap=None, **kw) -> float:
    """real(dot(phi, (D^dagger D)^-1 phi)); eta is the initial guess and the
    solution (src/Operators/DiracOperators.jl:146-162).  pseudofermion_action(D,
    phi) with a DiracOp borrows eta and the work fields from D.phitemps
    (:164-180)."""
    if eta is None and Deta is None and r is None and p is None and Ap is None:
        if not isinstance(D, DiracOp):
            raise TypeError("pseudofermion_action(D, phi) without work fields needs a DiracOp")
        eta, i = D.phitemps.get_block()
        try:
            solve_(eta, DdagDOp(D), phi)
            return dot(phi, eta).real
        finally:
            D.phitemps.unused_(i)
    res = solve_(eta, DdagDOp(D, Deta), phi, r, p, Ap, **kw)
    if not res.converged:
        raise RuntimeError(f"CG failed with reason {res.reason} after {res.iterations} iterations; "
                           f"relative residual = {res.relative_residual}")
    return dot(phi, eta).real


def dSFdU(dfdU: GaugeField, D: WilsonDiracOperator4D, phi: SpinorField, force: GaugeField | None = None,
          rtol=1e-10, maxiter=5000):
    """dSFdU(dfdU, D, phi) (ext/LatticeMatricesEnzymeExt.jl:395-460): dfdU <-
    dS/dU of S = phi^dagger (D^dagger D)^-1 phi in the Enzyme convention; with
    `force`, also the su(3) force TA(U dfdU^dagger).  Returns (CGResult, eta, chi)."""
    L = D.lattice
    eta, chi = SpinorField(L), SpinorField(L)
    work = torch.empty((4, L.ncomp, L.V, 2), dtype=torch.float64, device=L.device)
    params = _lib.default_params(rtol=float(rtol), maxiter=int(maxiter))
    res = _lib.CGResultC()
    check(lib().lmx_wilson_force(D._op, dfdU.ptr, force.ptr if force is not None else None, phi.ptr, eta.ptr,
                                 chi.ptr, ctypes.c_void_p(_ptr(work)), ctypes.byref(params), ctypes.byref(res)))
    return CGResult.from_c(res), eta, chi


def wilson_pullback_(dU: GaugeField, D: WilsonDiracOperator4D, dresult: SpinorField, psi: SpinorField,
                     coefficient: float) -> GaugeField:
    """The Enzyme reverse-rule link cotangent of mul!(result, D, psi)
    (ext/AD/WilsonDiracOperator.jl:62-91,137-143): dU += coefficient * ..."""
    check(lib().lmx_wilson_pullback(D._op, dU.ptr, dresult.ptr, psi.ptr, float(coefficient)))
    return dU


def ta_project_(F: GaugeField, U: GaugeField, G: GaugeField) -> GaugeField:
    check(lib().lmx_ta_project(U.lattice.ctx, F.ptr, U.ptr, G.ptr))
    return F


_PREC = {"f64": _lib.LMX_F64, "f32": _lib.LMX_F32, "f16": _lib.LMX_F16, "h16": _lib.LMX_H16}


def cg_batched_(x: SpinorField, D: WilsonDiracOperator4D, b: SpinorField, precision: str = "f32", rtol=1e-10,
                atol=0.0, maxiter=20000, reliable_delta=0.1, reliable_maxiter=100, check_every=0,
                use_graph=False, rhs_block=0, workspace=None) -> list:
    """Solve D^dagger D x_i = b_i for the x.nrhs right-hand sides at once.

    precision: "f64" (plain cg! per RHS), "f32" (FP32 inner CG), "f16"
    (FP16-stored links, FP32 spinors and math) or "h16" (FP16 links, inner
    vectors as 16-bit fixed point with a per-site scale, FP32 math; x_inner
    stays FP32); the mixed modes use FP64
    reliable updates, FP64 reductions and FP64 solution accumulation, and
    stop on the FP64 true residual.  x holds the initial guess.  rhs_block
    encodes (z-ranges) << 20 | (t-slices per workgroup) << 12 | (RHS per
    launch) << 4 | (RHS per work item, 1, 2 or 4), zero fields automatic;
    bit 29 selects the LDS-staged stencils (one RHS per work item), bit 28
    the RHS-loop stencils (FP16 links, RHS per launch only); 0 uses the
    autotuned or default shape.  Returns one CGResult per RHS."""
    if x.nrhs != b.nrhs:
        raise ArgumentError(_lib.LMX_ERR_ARG, "x and b must hold the same number of right-hand sides")
    if x.data.data_ptr() == b.data.data_ptr():
        raise ArgumentError(_lib.LMX_ERR_ALIAS, "CG fields x and b must not alias")
    params = _lib.default_params(rtol=float(rtol), atol=float(atol), maxiter=int(maxiter),
                                 precision=_PREC[precision], reliable_delta=float(reliable_delta),
                                 reliable_maxiter=int(reliable_maxiter), check_every=int(check_every),
                                 use_graph=int(bool(use_graph)), rhs_block=int(rhs_block))
    if precision != "f64":
        _refresh_low_precision(D)
    nbytes = ctypes.c_size_t()
    check(lib().lmx_cg_batched_workspace(D._op, x.nrhs, ctypes.byref(params), ctypes.byref(nbytes)))
    if workspace is None or workspace.numel() < nbytes.value:
        key = (x.nrhs, precision)
        cache = D.__dict__.setdefault("_ws", {})
        workspace = cache.get(key)
        if workspace is None or workspace.numel() < nbytes.value:
            workspace = torch.empty(max(nbytes.value, 1), dtype=torch.uint8, device=x.lattice.device)
            cache[key] = workspace
    res = (_lib.CGResultC * x.nrhs)()
    check(lib().lmx_cg_batched(D._op, x.ptr, b.ptr, x.nrhs, ctypes.c_void_p(workspace.data_ptr()),
                               nbytes.value, ctypes.byref(params), res))
    return [CGResult.from_c(r) for r in res]


def inner_kernel_times(D: WilsonDiracOperator4D, precision: str, nrhs: int, rhs_block: int = 0, groups: int = 4):
    """Microseconds per launch of the mixed solver's two inner stencils at a
    launch shape (0: the autotuned one): {"D": D with |Dp|^2, "Ddag": D^dagger
    with the r / x update}, HIP events on the library stream."""
    _refresh_low_precision(D)
    us = (ctypes.c_double * 2)()
    check(lib().lmx_inner_kernel_times(D._op, _PREC[precision], int(nrhs), int(rhs_block), int(groups), us))
    return {"D": us[0], "Ddag": us[1]}


def autotune(D: WilsonDiracOperator4D, precision: str, nrhs: int, cache_path: str | None = None):
    """Pick the RHS blocking of the mixed solver's dslash (lmx_autotune)."""
    _refresh_low_precision(D)
    out = _lib.TuneResult()
    check(lib().lmx_autotune(D._op, _PREC[precision], int(nrhs),
                             cache_path.encode() if cache_path else None, ctypes.byref(out)))
    return {"rhs_block": int(out.rhs_block), "block_size": int(out.block_size), "ms_per_apply": out.ms_per_apply}
