"""Pin the oracle's cg! restatement with the reference's known answers
(test/cg.jl) and with identities of the D^dagger D solve."""
import numpy as np
import pytest

import oracle as orc
from lmx import generators as gen

L1 = (1, 1, 1, 4)  # the 4-site ramp of test/cg.jl:15-18, as spinor fields


def ramp(L):
    V = int(np.prod(L))
    return (np.arange(1, V * 12 + 1, dtype=np.float64) + 0j).reshape(V, 4, 3)


def test_scaled_operator_one_iteration():
    # test/cg.jl:29-39: A = 4I converges in exactly one iteration to b/4
    b = ramp(L1)
    op = orc.NormalOp(orc.OP_SCALE, L1, scale=4.0)
    x, res = orc.cg(op, b, rtol=1e-13, maxiter=10)
    assert res.converged and res.iterations == 1 and res.reason == "converged"
    assert np.max(np.abs(x - b / 4)) <= 2e-13 * np.max(np.abs(b / 4))


def test_zero_rhs_zero_iterations():
    # test/cg.jl:41-49
    op = orc.NormalOp(orc.OP_SCALE, L1, scale=4.0)
    x, res = orc.cg(op, np.zeros((4, 4, 3), complex), rtol=1e-13, maxiter=10)
    assert res.converged and res.iterations == 0
    assert res.residual_norm == 0.0 and res.relative_residual == 0.0


def test_maxiter_zero_and_indefinite():
    b = ramp(L1)
    _, res = orc.cg(orc.NormalOp(orc.OP_SCALE, L1, scale=4.0), b, maxiter=0)
    assert not res.converged and res.reason == "maximum_iterations"
    _, res = orc.cg(orc.NormalOp(orc.OP_SCALE, L1, scale=-1.0), b, maxiter=10)
    assert not res.converged and res.reason == "nonpositive_curvature"


def test_ddagd_scale_known_answer():
    # test/cg.jl:76-91: DdagDOp((2+i) I) = 5 I, solve gives b/5; action |b|^2/5
    b = ramp(L1)
    op = orc.NormalOp(orc.OP_DDAGD_SCALE, L1, scale=2 + 1j)
    assert np.allclose(op.apply(b), 5 * b, rtol=0, atol=1e-12)
    x, res = orc.cg(op, b, rtol=1e-13, maxiter=10)
    assert res.converged
    assert np.max(np.abs(x - b / 5)) <= 2e-13 * np.max(np.abs(b / 5))
    action = np.vdot(b, x).real
    assert abs(action - np.vdot(b, b).real / 5) <= 2e-13 * action


@pytest.mark.parametrize("order", [0, 1])
def test_wilson_ddagd_solve_config1(order):
    # 4^4 Haar, kappa 0.125, phases (1,1,1,-1), rtol 1e-10 from zero
    L = (4, 4, 4, 4)
    U = gen.links(L, "haar", 1234)
    eta = gen.gaussian_spinor(L, 1235)
    phi = orc.wilson(U, eta, L, 0.125, dagger=True)   # phi = D^dagger eta
    op = orc.NormalOp(orc.OP_DDAGD_WILSON, L, U=U, kappa=0.125)
    x, res = orc.cg(op, phi, rtol=1e-12, maxiter=5000, order=order)
    assert res.converged
    true_res = np.linalg.norm(op.apply(x) - phi) / np.linalg.norm(phi)
    assert true_res <= 1e-11
    # the solution of D^dagger D x = D^dagger eta is x = D^-1 eta, so D x = eta
    Dx = orc.wilson(U, x, L, 0.125)
    assert np.linalg.norm(Dx - eta) <= 1e-9 * np.linalg.norm(eta)
    # S = phi^dagger (D^dagger D)^-1 phi = eta^dagger eta
    S = np.vdot(phi, x).real
    assert abs(S - np.vdot(eta, eta).real) <= 1e-10 * S


def test_orders_agree_on_solution():
    L = (4, 4, 4, 4)
    U = gen.links(L, "haar", 99)
    b = gen.gaussian_spinor(L, 100)
    op = orc.NormalOp(orc.OP_DDAGD_WILSON, L, U=U, kappa=0.125)
    x0, r0 = orc.cg(op, b, rtol=1e-10, order=0)
    x1, r1 = orc.cg(op, b, rtol=1e-10, order=1)
    assert r0.iterations == r1.iterations
    assert np.max(np.abs(x0 - x1)) <= 1e-12 * np.max(np.abs(x0))


def test_reference_cpu_path_matches_oracle_cg():
    # bench.py's CPU baseline (padded storage, set_halo! per operator
    # application, unfused BLAS-1, per-thread reductions) solves the same
    # system as the oracle's CG: same iterations, the same solution to rounding
    L = (4, 4, 4, 8)
    U = gen.links(L, "haar", 91)
    b = gen.gaussian_spinor(L, 92)
    x, res, _ = orc.refcpu_cg_ddagd(U, b, L, 0.125)
    op = orc.NormalOp(orc.OP_DDAGD_WILSON, L, U=U, kappa=0.125)
    xo, ro = orc.cg(op, b, rtol=1e-10, order=0)
    assert res.converged and res.iterations == ro.iterations
    assert np.max(np.abs(x - xo)) <= 1e-11 * np.max(np.abs(xo))
    assert np.linalg.norm(op.apply(x) - b) <= 1.01e-10 * np.linalg.norm(b)
