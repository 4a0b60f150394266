"""Plain FP64 vs mixed-precision D^dagger D CG at 32^4 on the GPU.

Workload: 32^4 Wilson D^dagger D, weak-field SU(3) links U = exp(i 0.2 H)
(H traceless Hermitian Gaussian, seeded -- the benchmark's ensemble),
kappa 0.124, phases (1, 1, 1, -1), the reference benchmark's Gaussian sources
(randomize_gaussian_matrix! with sigma sqrt(1/2), Philox, seed 1235 + r,
benchmark/ddagd_cg/ddagd_cg.jl:773-774, generated on the device), rtol 1e-10,
zero start.

- The plain FP64 CG (the reference algorithm cg!) on RHS 0 gives the oracle's
  iteration count and solution bit for bit (the C/OpenMP oracle runs the same
  32^4 solve on the host in ~25 s), and its true residual
  |b - D^dagger D x| / |b| recomputed with the FP64 operator
  (true_relative_residual!, benchmark/ddagd_cg/ddagd_cg.jl:646-650) is within
  the tolerance.
- The mixed solvers (16-bit inner spinors + FP16 links, and FP32 spinors +
  FP32 links) on 4 RHS with graph-captured iterations stop on the FP64 true
  residual: recomputed independently it is <= rtol (the reference harness
  accepts 10 rtol, :828-829), and each solution is within 1e-8 of the FP64
  CG's solution of the same RHS.
"""
import numpy as np
import pytest

import oracle as orc
import lmx
from lmx import generators as gen

pytestmark = pytest.mark.gpu

L = (32, 32, 32, 32)
KAPPA, RTOL, NRHS = 0.124, 1e-10, 4


@pytest.fixture(scope="module")
def problem(gpu):
    print(f"32^4 setup: weak-field links eps 0.2, {NRHS} Philox sources", flush=True)
    U = gen.links(L, "weak", 20240601, eps=0.2)
    lat = lmx.Lattice(L, (1, 1, 1, -1))
    D = lmx.WilsonDiracOperator4D(lat.gauge_from_numpy(U), KAPPA)
    bg = lat.spinor(NRHS)
    lmx.randomize_gaussian_matrix_(bg, [lmx.RNGStreamKey(1235 + r) for r in range(NRHS)], sigma=0.5 ** 0.5)
    b = bg.to_numpy()
    x64 = []
    res64 = []
    for i in range(NRHS):
        x, r, p, Ap, tmp = (lat.spinor() for _ in range(5))
        res64.append(lmx.cg_(x, lmx.DdagDOp(D, tmp), bg.rhs(i), r, p, Ap, rtol=RTOL, maxiter=20000))
        x64.append(x.to_numpy())  # [V, 4, 3] for one RHS
    return {"U": U, "lat": lat, "D": D, "bg": bg, "b": b, "x64": np.stack(x64), "res64": res64}


def true_residuals(D, x, b):
    lat = x.lattice
    Ax, tmp = lat.spinor(x.nrhs), lat.spinor(x.nrhs)
    lmx.mul_(Ax, lmx.DdagDOp(D, tmp), x)
    A = Ax.to_numpy().reshape(x.nrhs, -1)
    B = b.reshape(x.nrhs, -1)
    return [float(np.linalg.norm(B[i] - A[i]) / np.linalg.norm(B[i])) for i in range(x.nrhs)]


def test_fp64_cg_32x4_matches_oracle(problem):
    P = problem
    lat = P["lat"]
    res = P["res64"]
    tr = []
    for i in range(NRHS):
        xi = lat.spinor_from_numpy(P["x64"][i])
        tr.append(true_residuals(P["D"], xi, P["b"][i:i + 1])[0])
    print(f"FP64 CG: iterations {[r.iterations for r in res]}, true rel {['%.3e' % t for t in tr]}", flush=True)
    for r, t in zip(res, tr):
        assert r.converged and r.reason == "converged"
        assert t <= RTOL * 1.05
    # the C/OpenMP oracle: same iterations, same solution bit for bit (RHS 0)
    print("oracle FP64 CG at 32^4 (host) ...", flush=True)
    op = orc.NormalOp(orc.OP_DDAGD_WILSON, L, U=P["U"], kappa=KAPPA)
    xo, ro = orc.cg(op, P["b"][0], rtol=RTOL, order=1)
    print(f"oracle: iterations {ro.iterations}, reason {ro.reason}", flush=True)
    assert ro.converged and res[0].iterations == ro.iterations
    assert np.array_equal(P["x64"][0], xo)


@pytest.mark.parametrize("precision", ["h16", "f32"])
def test_mixed_32x4_against_fp64(problem, precision):
    P = problem
    lat, D, bg = P["lat"], P["D"], P["bg"]
    print(f"mixed {precision} x{NRHS} at 32^4, graph", flush=True)
    x = lat.spinor(NRHS)
    res = lmx.cg_batched_(x, D, bg, precision=precision, rtol=RTOL, use_graph=True)
    tr = true_residuals(D, x, P["b"])
    xs = x.to_numpy()
    errs = [float(np.linalg.norm(xs[i] - P["x64"][i]) / np.linalg.norm(P["x64"][i])) for i in range(NRHS)]
    print(f"  iterations {[r.iterations for r in res]} (FP64 {[r.iterations for r in P['res64']]}), updates "
          f"{[r.reliable_updates for r in res]}, true rel {['%.3e' % t for t in tr]}, error vs FP64 "
          f"{['%.2e' % e for e in errs]}", flush=True)
    for r, t, e in zip(res, tr, errs):
        assert r.converged and r.reason == "converged"
        assert t <= RTOL * 1.05
        assert abs(t - r.true_relative_residual) <= 1e-3 * RTOL + 1e-6 * t
        assert e <= 1e-8
