"""The reference's pre-v1 compatibility interfaces on the device: the
pool-based cg (test/cg.jl:96-108), DiracOp callback operators with the
one-argument DdagDOp, the work-field-free solve! and pseudofermion_action
(test/cg.jl:110-137), the adjoint identities (test/regressions.jl:150-160),
and a DiracOp whose callbacks apply the Wilson operator."""
import numpy as np
import pytest
import torch

import lmx
from lmx import generators as gen

pytestmark = pytest.mark.gpu

L = (4, 4, 4, 4)


def _close(a, b, tol):
    return torch.allclose(a, b, rtol=tol, atol=tol)


def test_pool_based_cg(gpu):
    lat = lmx.Lattice(L, (1, 1, 1, -1))
    b = lat.spinor_from_numpy(gen.gaussian_spinor(L, 930))
    x = lat.spinor()
    temps = lmx.FieldPool(b.similar, 3)
    assert lmx.cg(x, lmx.ScaledOperator(4.0), b, temps, eps=1e-13, maxsteps=10) is None
    assert _close(x.data, b.data / 4, 2e-13)
    assert not any(temps._used)
    x2 = lat.spinor()
    assert lmx.cg(x2, lmx.ScaledOperator(4.0), b, [b.similar() for _ in range(3)], eps=1e-13) is None
    assert torch.equal(x2.data, x.data)


def test_pre_v1_dirac_solver(gpu):
    lat = lmx.Lattice(L, (1, 1, 1, -1))
    b = lat.spinor_from_numpy(gen.gaussian_spinor(L, 931))
    scale = 2 + 1j

    def apply_scale(y, U, x, coefficient, phitemps, temps):
        assert len(phitemps) == 4 and len(temps) == 4
        lmx.mul_(y, coefficient, x)

    def apply_scale_adjoint(y, U, x, coefficient, phitemps, temps):
        lmx.mul_(y, coefficient.conjugate(), x)

    D = lmx.DiracOp(lmx.GaugeField(lat), apply_scale, apply_scale_adjoint, scale, b, numtemp=1, numphitemp=4)
    N = lmx.DdagDOp(D)
    out = b.similar()
    lmx.mul_(out, N, b)
    assert _close(out.data, abs(scale) ** 2 * b.data, 1e-14)
    x = b.similar()
    assert lmx.solve_(x, N, b, verboselevel=2) is None
    assert _close(x.data, b.data / abs(scale) ** 2, 2e-13)
    expected = lmx.dot(b, b).real / abs(scale) ** 2
    assert abs(lmx.pseudofermion_action(D, b) - expected) <= 2e-13 * expected
    assert not any(D.phitemps._used) and not any(D.temps._used)
    assert lmx.adjoint(lmx.adjoint(D)) is D
    N2 = lmx.DdagDOp(D, b.similar())
    assert lmx.adjoint(N2) is N2 and lmx.adjoint(lmx.adjoint(N2)) is N2


def test_dirac_op_over_the_wilson_operator(gpu):
    kappa = 0.125
    lat = lmx.Lattice(L, (1, 1, 1, -1))
    U = lat.gauge_from_numpy(gen.links(L, "haar", 932))
    W = lmx.WilsonDiracOperator4D(U, kappa)
    b = lat.spinor_from_numpy(gen.gaussian_spinor(L, 933))
    D = lmx.DiracOp(U, lambda y, U_, x, p, ph, t: lmx.mul_(y, p, x),
                    lambda y, U_, x, p, ph, t: lmx.mul_(y, lmx.adjoint(p), x), W, b)
    for op_callback, op_direct in ((D, W), (lmx.adjoint(D), lmx.adjoint(W))):
        y1, y2 = b.similar(), b.similar()
        lmx.mul_(y1, op_callback, b)
        lmx.mul_(y2, op_direct, b)
        assert torch.equal(y1.data, y2.data)
    x = b.similar()
    assert lmx.solve_(x, lmx.DdagDOp(D), b) is None
    Ax, tmp = b.similar(), b.similar()
    lmx.mul_(Ax, lmx.DdagDOp(W, tmp), x)
    rel = float(torch.linalg.vector_norm(Ax.data - b.data) / torch.linalg.vector_norm(b.data))
    print(f"DiracOp(Wilson) legacy solve: true relative residual {rel:.2e}", flush=True)
    assert rel < 1e-9
    x3 = b.similar()
    with pytest.raises(RuntimeError, match="not converged"):
        lmx.cg(x3, lmx.DdagDOp(D), b, D.phitemps, eps=1e-10, maxsteps=3)
    assert not any(D.phitemps._used)
