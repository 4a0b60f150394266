"""The larger workload configurations on the GPU.

- Multi-RHS batch: 24^3x48, 12 spin-colour diluted Gaussian sources plus an
  uneven second batch of 4 Gaussian sources; FP32 spinors, FP16-stored links,
  FP32 arithmetic, FP64 reductions and reliable updates (precision "f16"),
  kappa 0.1245, rtol 1e-10, the RHS shape autotuned over {1,2,4,6,12} x
  {1,2,4}, hipGraph-captured iterations.  The configuration leaves the gauge
  ensemble open: weak-field links (eps 0.2, the 32^4 ensemble) are used.
- Force: 16^4 Haar links, kappa 0.125, phi = D^dagger eta,
  S = phi^dagger (D^dagger D)^-1 phi and dS/dU by the hand-written adjoint,
  against central finite differences (eps 1e-5 along 8 random su(3)
  directions at each of 16 random links) to 1e-8, with the local difference
  S(U') - S(U) = -X'^dagger (A' - A) X (X = A^-1 phi, X' = A'^-1 phi) so that
  the difference is not lost to cancellation.  Known answer: S = |eta|^2.
- Largest volume: 48^3x96 Wilson-clover, cSW 1.0, weak-field eps 0.15,
  kappa 0.1235, 4 Gaussian sources, mixed 16-bit/FP32/FP64 multi-RHS CG to
  1e-10.

At these volumes the FP64 reference is the library's FP64 path, which is
bit-identical to the oracle on every case the oracle runs in seconds
(test_gpu_parity.py); true residuals are recomputed independently with the
FP64 operator, and the solution error against the FP64 CG solution is
printed per right-hand side."""
import time

import numpy as np
import pytest
import torch

import lmx
from lmx import generators as gen

pytestmark = pytest.mark.gpu


def true_residuals(D, x, b_np):
    """|b - D^dagger D x| / |b| per RHS with the FP64 operator."""
    lat = x.lattice
    Ax, tmp = lat.spinor(x.nrhs), lat.spinor(x.nrhs)
    lmx.mul_(Ax, lmx.DdagDOp(D, tmp), x)
    out = []
    A = Ax.to_numpy().reshape(x.nrhs, -1)
    B = b_np.reshape(x.nrhs, -1)
    for i in range(x.nrhs):
        out.append(float(np.linalg.norm(B[i] - A[i]) / np.linalg.norm(B[i])))
    return out


def fp64_solution(D, b_field, i):
    lat = b_field.lattice
    x, r, p, Ap, tmp = (lat.spinor() for _ in range(5))
    res = lmx.cg_(x, lmx.DdagDOp(D, tmp), b_field.rhs(i), r, p, Ap, rtol=1e-12, maxiter=20000)
    assert res.converged
    return x.to_numpy()


def test_batch_24x48_spin_colour_plus_uneven(gpu):
    # the 12 spin-colour sources and the uneven second batch of 4 Gaussian
    # sources are one 16-RHS solve: the autotuned launch splits the RHS into
    # groups (12 + 4, 6 + 6 + 4, ... -- a ragged last group), retires each RHS
    # on its own true residual, and replays captured iterations
    L, kappa, rtol = (24, 24, 24, 48), 0.1245, 1e-10
    print(f"batch L={L} kappa={kappa}", flush=True)
    U = gen.links(L, "weak", 500, eps=0.2)
    lat = lmx.Lattice(L, (1, 1, 1, -1))
    D = lmx.WilsonDiracOperator4D(lat.gauge_from_numpy(U), kappa)
    b = np.concatenate([gen.spin_colour_sources(L, 501), gen.gaussian_spinor(L, 502, nrhs=4)])
    n = b.shape[0]
    assert n == 16
    tune = lmx.autotune(D, "f16", n)
    bg = lat.spinor_from_numpy(b)
    x = lat.spinor(n)
    t0 = time.perf_counter()
    res = lmx.cg_batched_(x, D, bg, precision="f16", rtol=rtol, use_graph=True)
    torch.cuda.synchronize()
    dt = time.perf_counter() - t0
    tr = true_residuals(D, x, b)
    errs = []
    for i in (0, 12):  # one spin-colour and one Gaussian source against FP64 CG
        x64 = fp64_solution(D, bg, i)
        errs.append(float(np.linalg.norm(x.to_numpy()[i] - x64) / np.linalg.norm(x64)))
    blk = tune["rhs_block"]
    print(f"  16 RHS: shape TL={(blk >> 12) & 0xff} ZS={(blk >> 20) & 0xf} LDS={(blk >> 29) & 1} per={(blk >> 4) & 0xff} nr={blk & 15}, {dt:.2f} s "
          f"({dt / n:.3f} s/RHS), iterations {[r.iterations for r in res]}, "
          f"updates {[r.reliable_updates for r in res]}, max true rel {max(tr):.3e}, "
          f"error vs FP64 (RHS 0, 12) {errs[0]:.3e} {errs[1]:.3e}", flush=True)
    for r, t in zip(res, tr):
        assert r.converged and r.reason == "converged"
        assert t <= rtol * 1.05
        assert abs(t - r.true_relative_residual) <= 1e-3 * rtol + 1e-6 * t
    assert max(errs) <= 1e-7
    # the same 16 RHS in launches of 12 + 4 (a ragged last group), eager: the
    # grouping and the launch shape change which work item computes what, not
    # the arithmetic or the reduction order, so the iterates are identical
    x2 = lat.spinor(n)
    res2 = lmx.cg_batched_(x2, D, bg, precision="f16", rtol=rtol, use_graph=False, rhs_block=(12 << 4) | 1)
    assert [r.iterations for r in res2] == [r.iterations for r in res]
    assert np.array_equal(x2.to_numpy(), x.to_numpy())


def _set_link(G, mu, s, M):
    v = torch.from_numpy(np.stack([M.real.reshape(9), M.imag.reshape(9)], axis=-1)).to(G.data.device)
    G.data[mu, :, s, :] = v


def test_force_16x4_finite_differences(gpu):
    L, kappa, eps = (16, 16, 16, 16), 0.125, 1e-5
    print(f"force L={L}", flush=True)
    U = gen.links(L, "haar", 600)
    eta = gen.gaussian_spinor(L, 601)
    lat = lmx.Lattice(L, (1, 1, 1, -1))
    Ug = lat.gauge_from_numpy(U)
    D = lmx.WilsonDiracOperator4D(Ug, kappa)
    phi = lat.spinor()
    lmx.mul_(phi, lmx.adjoint(D), lat.spinor_from_numpy(eta))
    G, F = lmx.GaugeField(lat), lmx.GaugeField(lat)
    res, X, chi = lmx.dSFdU(G, D, phi, force=F, rtol=1e-13, maxiter=20000)
    assert res.converged
    S = lmx.dot(phi, X).real
    assert abs(S - np.vdot(eta, eta).real) <= 1e-11 * S  # S = |eta|^2 for phi = D^dagger eta
    Gn, Fn = G.to_numpy(), F.to_numpy()
    AX, tmpA = lat.spinor(), lat.spinor()
    lmx.mul_(AX, lmx.DdagDOp(D, tmpA), X)
    AXn = AX.to_numpy()
    rng = np.random.default_rng(602)
    picks = [(int(rng.integers(0, lat.V)), int(rng.integers(0, 4))) for _ in range(16)]
    dirs = gen.random_su3_algebra(8, 603)
    Up = lmx.GaugeField(lat, Ug.data.clone())
    Dp = lmx.WilsonDiracOperator4D(Up, kappa)
    xp, r, p, Ap, tmp, out = (lat.spinor() for _ in range(6))
    worst = 0.0
    for s, mu in picks:
        for k in range(8):
            Xd = dirs[k]
            dS = []
            for sign in (1.0, -1.0):
                _set_link(Up, mu, s, gen.expm_antihermitian(sign * eps * Xd) @ U[mu, s])
                lmx.substitute_(xp, X)  # warm start from the unperturbed solution
                rp = lmx.cg_(xp, lmx.DdagDOp(Dp, tmp), phi, r, p, Ap, rtol=1e-13, maxiter=20000)
                assert rp.converged
                lmx.mul_(out, lmx.DdagDOp(Dp, tmp), X)
                d = out.to_numpy() - AXn  # exactly zero away from the perturbed link
                dS.append(-np.vdot(xp.to_numpy(), d).real)
            _set_link(Up, mu, s, U[mu, s])
            fd = (dS[0] - dS[1]) / (2 * eps)
            an_G = np.sum(np.conj(Gn[mu, s]) * (Xd @ U[mu, s])).real
            an_F = np.trace(Fn[mu, s] @ Xd).real
            assert abs(an_G - an_F) <= 1e-12 * abs(an_G)
            rel = abs(fd - an_G) / abs(an_G)
            worst = max(worst, rel)
            assert rel <= 1e-8, (s, mu, k, fd, an_G)
    print(f"  128 link directions: worst |FD - adjoint| / |adjoint| = {worst:.2e}", flush=True)


def test_clover_48x96_mixed_4rhs(gpu):
    L, kappa, csw, rtol = (48, 48, 48, 96), 0.1235, 1.0, 1e-10
    print(f"clover L={L}", flush=True)
    t0 = time.perf_counter()
    U = gen.links(L, "weak", 700, eps=0.15)
    b = gen.gaussian_spinor(L, 701, nrhs=4)
    print(f"  inputs {time.perf_counter() - t0:.1f} s", flush=True)
    lat = lmx.Lattice(L, (1, 1, 1, -1))
    Ug = lat.gauge_from_numpy(U)
    del U
    D = lmx.WilsonDiracCloverOperator4D(Ug, kappa, csw)
    bg = lat.spinor_from_numpy(b)
    x = lat.spinor(4)
    tune = lmx.autotune(D, "h16", 4)
    blk = tune["rhs_block"]
    print(f"  autotuned shape TL={(blk >> 12) & 0xff} ZS={(blk >> 20) & 0xf} LDS={(blk >> 29) & 1} per={(blk >> 4) & 0xff} nr={blk & 15}: fused D / D^dagger pair "
          f"{tune['ms_per_apply'] * 1e3:.0f} us per application per RHS", flush=True)
    # the FP64 clover D_W against the HBM roofline: links 576 + psi 192 +
    # clover cache 432 + out 192 = 1392 algorithmic bytes per site
    y = lat.spinor()
    b0 = bg.rhs(0)
    lmx.mul_(y, D, b0)
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    e0.record()
    for _ in range(10):
        lmx.mul_(y, D, b0)
    e1.record()
    torch.cuda.synchronize()
    us = e0.elapsed_time(e1) / 10 * 1e3
    gbs = 1392 * lat.V / (us * 1e-6) / 1e9
    print(f"  FP64 clover D_W: {us:.0f} us per application, {gbs:.0f} GB/s algorithmic = {gbs / 8000:.3f} of the "
          f"8 TB/s HBM peak", flush=True)
    del y
    lmx.cg_batched_(x, D, bg, precision="h16", rtol=1e-3)  # warm up
    lmx.clear_matrix_(x)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    res = lmx.cg_batched_(x, D, bg, precision="h16", rtol=rtol, use_graph=True)
    torch.cuda.synchronize()
    tm = (time.perf_counter() - t0) / 4
    tr = true_residuals(D, x, b)
    x64, r_, p_, Ap_, tmp_ = (lat.spinor() for _ in range(5))
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    r64 = lmx.cg_(x64, lmx.DdagDOp(D, tmp_), bg.rhs(0), r_, p_, Ap_, rtol=rtol, maxiter=20000)
    torch.cuda.synchronize()
    t64 = time.perf_counter() - t0
    err0 = float(np.linalg.norm(x.to_numpy()[0] - x64.to_numpy()) / np.linalg.norm(x64.to_numpy()))
    print(f"  mixed h16 x4: {tm:.3f} s/RHS, iterations {[q.iterations for q in res]}, "
          f"updates {[q.reliable_updates for q in res]}, true rel {[f'{t:.2e}' for t in tr]}; "
          f"FP64 CG {t64:.3f} s/RHS ({r64.iterations} iterations): speedup {t64 / tm:.2f}x; "
          f"RHS0 error vs FP64 {err0:.2e}", flush=True)
    for q, t in zip(res, tr):
        assert q.converged and t <= rtol * 1.05
    assert r64.converged
    assert err0 <= 1e-7
