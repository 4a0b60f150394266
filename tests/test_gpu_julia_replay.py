"""Replays, through ctypes, the exact C-ABI call sequence of the Julia binding
julia/LatticeMatricesLmx.jl on the reference's own storage (LatticeMatrix.A
arrays, halo-padded, column-major), and checks each method's result against
the oracle:

  context / handle     lmx_context_create_nc, lmx_links_import (the four padded
                       link arrays), lmx_wilson_create, lmx_spinor_bytes(6)
  mul!(C, D, psi)      lmx_wilson_apply_padded (D, D', Donly, Donly', clover)
  cg!(x, DdagDOp(D, temp), b, r, p, Ap)
                       lmx_cg_ddagd_padded (x, b padded; the six work fields in
                       the operator's workspace); x's halo is then set_halo!'s;
                       on LatticeMatrix{ComplexF32}: lmx_cg_ddagd_padded_f32
  link write           lmx_links_import + lmx_wilson_update_links
  reverse rule of mul! (dU += pullback)
                       lmx_links_import (dU), lmx_spinor_import (dresult, psi),
                       lmx_wilson_pullback, lmx_links_export
  dSFdU(dfdU, D, phi)  lmx_spinor_import (phi), lmx_wilson_force (F = NULL),
                       lmx_links_export
NC = 3 (the half-spin kernels, Haar links, nw = 1) and, for mul! and cg!,
NC = 2 (the generic kernels, the reference's SU(2) fixture).
"""
import ctypes

import numpy as np
import pytest
import torch

import dense_reference as dr
import oracle as orc
import su2_fixtures as fx
from lmx import _lib
from lmx import generators as gen

pytestmark = pytest.mark.gpu

P = ctypes.c_void_p
PHASES = (1, 1, 1, -1)
REASONS = ("converged", "maximum_iterations", "nonpositive_curvature", "numerical_breakdown")


def lib():
    return _lib.lib()


def padded_spinor(core, L, nw, dev):
    """LatticeMatrix.A of a spinor: C order [Lt+2nw, ..., Lx+2nw, 4, nc], halo as set_halo!."""
    return torch.from_numpy(np.ascontiguousarray(dr.halo_fill(core, L, nw, PHASES))).to(dev)


def padded_links(U, L, nw, dev):
    """The four LatticeMatrix.A link arrays: C order [..., col, row] (column-major [nc, nc, ...]), periodic halo."""
    nc = U.shape[-1]
    out = []
    for mu in range(4):
        cm = np.swapaxes(U[mu], -1, -2)  # [V, col, row]
        out.append(torch.from_numpy(np.ascontiguousarray(dr.halo_fill(cm, L, nw, (1, 1, 1, 1)))).to(dev))
    return out


def core(t, L, nw):
    a = t.cpu().numpy()
    if nw:
        a = a[nw:-nw, nw:-nw, nw:-nw, nw:-nw]
    return np.ascontiguousarray(a)


class Binding:
    """The state the Julia module keeps: a context per lattice and an operator
    handle (links in the library layout, lmx_wilson, six-spinor workspace)."""

    def __init__(self, L, nc, Upad, nw, kappa, csw, dev):
        self.L, self.nc, self.nw, self.dev = L, nc, nw, dev
        V = int(np.prod(L))
        ph = (ctypes.c_double * 8)(*[v for p in PHASES for v in (float(p), 0.0)])
        h = P()
        stream = torch.cuda.current_stream(dev).cuda_stream
        _lib.check(lib().lmx_context_create_nc(ctypes.byref(h), (ctypes.c_int32 * 4)(*L), ph, P(stream), nc))
        self.ctx = h
        self.links = torch.zeros(2 * 4 * nc * nc * V, dtype=torch.float64, device=dev)
        self.import_links(self.links, Upad)
        op = P()
        _lib.check(lib().lmx_wilson_create(self.ctx, ctypes.byref(op), P(self.links.data_ptr()), kappa, csw))
        self.op = op
        self.n1 = lib().lmx_spinor_bytes(self.ctx, _lib.LMX_F64, 1)
        self.work = torch.zeros(lib().lmx_spinor_bytes(self.ctx, _lib.LMX_F64, 6), dtype=torch.uint8, device=dev)

    def import_links(self, dst, Upad):
        ptrs = (P * 4)(*[t.data_ptr() for t in Upad])
        _lib.check(lib().lmx_links_import(self.ctx, P(dst.data_ptr()), ptrs, self.nw))

    def mul(self, C, psi, flags):
        _lib.check(lib().lmx_wilson_apply_padded(self.op, P(C.data_ptr()), P(psi.data_ptr()), self.nw, flags, 1))

    def cg(self, x, b, rtol=1e-10, atol=0.0, maxiter=5000):
        params = _lib.default_params(rtol=rtol, atol=atol, maxiter=maxiter)
        res = _lib.CGResultC()
        _lib.check(lib().lmx_cg_ddagd_padded(self.op, P(x.data_ptr()), P(b.data_ptr()), self.nw,
                                             P(self.work.data_ptr()), self.work.numel(), ctypes.byref(params),
                                             ctypes.byref(res)))
        return res

    def cg_f32(self, x, b, rtol=1e-5, atol=0.0, maxiter=5000):
        """cg! on LatticeMatrix{ComplexF32} (complex64 arrays), the same workspace."""
        params = _lib.default_params(rtol=rtol, atol=atol, maxiter=maxiter)
        res = _lib.CGResultC()
        _lib.check(lib().lmx_cg_ddagd_padded_f32(self.op, P(x.data_ptr()), P(b.data_ptr()), self.nw,
                                                 P(self.work.data_ptr()), self.work.numel(), ctypes.byref(params),
                                                 ctypes.byref(res)))
        return res

    def slot(self, i):
        return P(self.work.data_ptr() + i * self.n1)

    def import_spinor(self, i, psi):
        _lib.check(lib().lmx_spinor_import(self.ctx, self.slot(i), P(psi.data_ptr()), self.nw, 1))
        return self.slot(i)

    def export_links(self, dUpad, G):
        ptrs = (P * 4)(*[t.data_ptr() for t in dUpad])
        _lib.check(lib().lmx_links_export(self.ctx, ptrs, self.nw, P(G.data_ptr())))

    def add_link_pullback(self, dUpad, dres, psi, coefficient):
        V = int(np.prod(self.L))
        G = torch.zeros(2 * 4 * self.nc * self.nc * V, dtype=torch.float64, device=self.dev)
        self.import_links(G, dUpad)
        pd = self.import_spinor(0, dres)
        pp = self.import_spinor(1, psi)
        _lib.check(lib().lmx_wilson_pullback(self.op, P(G.data_ptr()), pd, pp, coefficient))
        self.export_links(dUpad, G)

    def dsfdu(self, dfdU, phi, rtol=1e-10, maxiter=5000):
        V = int(np.prod(self.L))
        G = torch.zeros(2 * 4 * self.nc * self.nc * V, dtype=torch.float64, device=self.dev)
        work = torch.zeros(4 * self.n1, dtype=torch.uint8, device=self.dev)
        pphi = self.import_spinor(0, phi)
        params = _lib.default_params(rtol=rtol, maxiter=maxiter)
        res = _lib.CGResultC()
        _lib.check(lib().lmx_wilson_force(self.op, P(G.data_ptr()), None, pphi, self.slot(1), self.slot(2),
                                          P(work.data_ptr()), ctypes.byref(params), ctypes.byref(res)))
        self.export_links(dfdU, G)
        return res

    def close(self):
        lib().lmx_context_synchronize(self.ctx)
        lib().lmx_wilson_destroy(self.op)
        lib().lmx_context_destroy(self.ctx)


def links_from_padded(Upad, L, nw):
    return np.stack([np.swapaxes(core(t, L, nw).reshape(-1, t.shape[-1], t.shape[-1]), -1, -2) for t in Upad])


@pytest.mark.parametrize("nc", [3, 2])
def test_replay_mul_and_cg(gpu, nc):
    nw = 1
    if nc == 3:
        L, kappa, csw = (4, 4, 4, 8), 0.125, 1.0
        U = gen.links(L, "haar", 515)
        rng = np.random.default_rng(516)
    else:
        L, kappa, csw = (2, 2, 2, 4), 0.12, 1.17
        U = fx.clover_test_links(L)
        rng = np.random.default_rng(517)
    V = int(np.prod(L))
    print(f"binding replay: NC={nc} on {L}, nw={nw}", flush=True)
    psi = rng.standard_normal((V, 4, nc)) + 1j * rng.standard_normal((V, 4, nc))
    Upad = padded_links(U, L, nw, gpu)
    for c in (0.0, csw):
        B = Binding(L, nc, Upad, nw, kappa, c, gpu)
        try:
            psi_d = padded_spinor(psi, L, nw, gpu)
            C = torch.zeros_like(psi_d)
            Fo = (orc.clover_field_strength(U, L) if nc == 3 else orc.clover_field_strength_nc(U, L)) if c else None
            for flags in ((0, 1, 2, 3) if c == 0.0 else (0, 1)):
                B.mul(C, psi_d, flags)
                got = core(C, L, nw).reshape(V, 4, nc)
                dag, donly = bool(flags & 1), bool(flags & 2)
                if c:
                    ref = (orc.wilson_clover(U, Fo, psi, L, kappa, c, PHASES, dag) if nc == 3 else
                           orc.wilson_clover_nc(U, Fo, psi, L, kappa, c, PHASES, dag))
                else:
                    ref = (orc.wilson(U, psi, L, kappa, PHASES, dag, donly) if nc == 3 else
                           orc.wilson_nc(U, psi, L, kappa, PHASES, dag, donly))
                assert np.array_equal(got, ref), f"csw={c} flags={flags}"
            # cg!(x, DdagDOp(D, temp), b, r, p, Ap) on the padded storage, zero start
            x = torch.zeros_like(psi_d)
            res = B.cg(x, psi_d)
            kind = orc.OP_DDAGD_CLOVER if c else orc.OP_DDAGD_WILSON
            op = orc.NormalOp(kind, L, U=U, F=Fo, kappa=kappa, csw=c, phases=PHASES, nc=nc)
            xo, ro = orc.cg(op, psi, rtol=1e-10, order=1)
            print(f"  csw={c}: cg! iterations {res.iterations} (oracle {ro.iterations}), "
                  f"{REASONS[res.reason]}", flush=True)
            assert res.converged and res.iterations == ro.iterations and REASONS[res.reason] == ro.reason
            assert np.array_equal(core(x, L, nw).reshape(V, 4, nc), xo)
            assert np.array_equal(x.cpu().numpy(), dr.halo_fill(xo, L, nw, PHASES))  # halo as set_halo!
        finally:
            B.close()


def test_replay_link_write_pullback_and_dsfdu(gpu):
    nw = 1
    L, kappa = (4, 4, 4, 8), 0.125
    V = int(np.prod(L))
    U = gen.links(L, "haar", 615)
    U2 = gen.links(L, "haar", 616)
    rng = np.random.default_rng(617)
    psi = rng.standard_normal((V, 4, 3)) + 1j * rng.standard_normal((V, 4, 3))
    dres = rng.standard_normal((V, 4, 3)) + 1j * rng.standard_normal((V, 4, 3))
    Upad = padded_links(U, L, nw, gpu)
    B = Binding(L, 3, Upad, nw, kappa, 0.0, gpu)
    try:
        # a write to the links (new core epoch): re-import and update_links, then mul!
        for mu, t in enumerate(padded_links(U2, L, nw, gpu)):
            Upad[mu].copy_(t)
        B.import_links(B.links, Upad)
        _lib.check(lib().lmx_wilson_update_links(B.op, P(B.links.data_ptr())))
        psi_d = padded_spinor(psi, L, nw, gpu)
        C = torch.zeros_like(psi_d)
        B.mul(C, psi_d, 0)
        assert np.array_equal(core(C, L, nw).reshape(V, 4, 3), orc.wilson(U2, psi, L, kappa, PHASES))
        # the reverse rule's link part: dU += pullback(dresult, psi; -kappa), dU nonzero beforehand
        dU0 = 0.01 * (rng.standard_normal((4, V, 3, 3)) + 1j * rng.standard_normal((4, V, 3, 3)))
        dUpad = padded_links(dU0, L, nw, gpu)
        B.add_link_pullback(dUpad, padded_spinor(dres, L, nw, gpu), psi_d, -kappa)
        got = links_from_padded(dUpad, L, nw)
        assert np.array_equal(got, orc.wilson_pullback(dres, psi, L, -kappa, PHASES, dU=dU0))
        assert np.array_equal(dUpad[2].cpu().numpy(),
                              dr.halo_fill(np.swapaxes(got[2], -1, -2), L, nw, (1, 1, 1, 1)))  # periodic halo
        # dSFdU(dfdU, D, phi) with phi = D^dagger eta
        eta = rng.standard_normal((V, 4, 3)) + 1j * rng.standard_normal((V, 4, 3))
        phi = orc.wilson(U2, eta, L, kappa, PHASES, dagger=True)
        dfdU = padded_links(np.zeros((4, V, 3, 3), np.complex128), L, nw, gpu)
        res = B.dsfdu(dfdU, padded_spinor(phi, L, nw, gpu), rtol=1e-12, maxiter=20000)
        Go, _, _, ro = orc.dsfdu(U2, phi, L, kappa, PHASES, rtol=1e-12, maxiter=20000)
        print(f"dSFdU: CG iterations {res.iterations} (oracle {ro.iterations})", flush=True)
        assert res.converged and res.iterations == ro.iterations
        assert np.array_equal(links_from_padded(dfdU, L, nw), Go)
    finally:
        B.close()


@pytest.mark.parametrize("csw", [0.0, 1.0])
def test_replay_cg_f32(gpu, csw):
    # cg! on LatticeMatrix{ComplexF32} (complex64 padded arrays): FP32 iterates,
    # so the bar is the FP64 true residual of the solution, the iteration
    # count near the FP64 oracle's at the same rtol, and x's halo as set_halo!
    nw, L, kappa, rtol = 1, (4, 4, 4, 8), 0.12, 1e-5
    V = int(np.prod(L))
    U = gen.links(L, "haar", 715)
    rng = np.random.default_rng(716)
    b = (rng.standard_normal((V, 4, 3)) + 1j * rng.standard_normal((V, 4, 3))).astype(np.complex64)
    B = Binding(L, 3, padded_links(U, L, nw, gpu), nw, kappa, csw, gpu)
    try:
        b_d = padded_spinor(b, L, nw, gpu)
        assert b_d.dtype == torch.complex64
        x = torch.zeros_like(b_d)
        res = B.cg_f32(x, b_d, rtol=rtol)
        Fo = orc.clover_field_strength(U, L) if csw else None
        op = orc.NormalOp(orc.OP_DDAGD_CLOVER if csw else orc.OP_DDAGD_WILSON, L, U=U, F=Fo, kappa=kappa, csw=csw,
                          phases=PHASES)
        b64 = b.astype(np.complex128)
        xo, ro = orc.cg(op, b64, rtol=rtol)
        xs = core(x, L, nw).reshape(V, 4, 3)
        true_rel = np.linalg.norm(op.apply(xs.astype(np.complex128)) - b64) / np.linalg.norm(b64)
        print(f"  csw={csw}: ComplexF32 cg! iterations {res.iterations} (FP64 oracle {ro.iterations}), FP64 true rel "
              f"{true_rel:.3e}", flush=True)
        assert res.converged and abs(res.iterations - ro.iterations) <= 3 and true_rel <= 3 * rtol
        assert np.array_equal(x.cpu().numpy(), dr.halo_fill(xs, L, nw, PHASES).astype(np.complex64))
    finally:
        B.close()
