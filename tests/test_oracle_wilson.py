"""Pin the CPU oracle's D_W / D_W^dagger / Donly against what fixes them
independently of the oracle's own code:
  - the dense gamma-matrix restatement of test/wilson_halfspin.jl:8-42,
  - the free-field plane-wave closed form (U = 1),
  - the adjoint identity, gamma5-hermiticity, gauge covariance,
and show, by planting plausible mistakes in the oracle's inputs/conventions,
that these checks fail when a term is dropped, a sign flipped or an operand
transposed."""
import numpy as np
import pytest

import dense_reference as dr
import oracle as orc
from lmx import generators as gen


def rand_c(rng, shape):
    return rng.random(shape) + 1j * rng.random(shape)


def close(a, b, tol):
    return np.max(np.abs(a - b)) <= tol * max(1.0, np.max(np.abs(b)))


@pytest.mark.parametrize("L", [(2, 2, 2, 2), (4, 2, 3, 2), (4, 4, 4, 4)])
@pytest.mark.parametrize("dagger,donly", [(False, False), (True, False), (False, True), (True, True)])
def test_oracle_matches_dense_reference(L, dagger, donly):
    # test/wilson_halfspin.jl:44-105: rand links and spinors, phases (1,1,1,-1), kappa 0.117, tol 2e-12
    rng = np.random.default_rng(11)
    V = int(np.prod(L))
    U = rand_c(rng, (4, V, 3, 3))
    psi = rand_c(rng, (V, 4, 3))
    phases = (1, 1, 1, -1)
    got = orc.wilson(U, psi, L, 0.117, phases, dagger, donly)
    ref = dr.wilson(U, psi, L, 0.117, phases, dagger, donly)
    assert close(got, ref, 2e-12)


def test_general_complex_phases():
    rng = np.random.default_rng(3)
    L = (3, 2, 2, 4)
    V = int(np.prod(L))
    U = rand_c(rng, (4, V, 3, 3))
    psi = rand_c(rng, (V, 4, 3))
    phases = (np.exp(0.3j), 1, -1j, -1)
    for dag in (False, True):
        assert close(orc.wilson(U, psi, L, 0.13, phases, dag), dr.wilson(U, psi, L, 0.13, phases, dag), 1e-13)


@pytest.mark.parametrize("n", [(0, 0, 0, 0), (1, 0, 2, 1), (3, 1, 0, 2)])
def test_free_field_closed_form(n):
    L = (4, 4, 4, 6)
    V = int(np.prod(L))
    kappa = 0.124
    U = gen.links(L, "unit", 0)
    ph, Dt = dr.free_field_eigen(L, n, kappa)
    rng = np.random.default_rng(5)
    u = rand_c(rng, (4, 3))
    psi = ph[:, None, None] * u[None]
    got = orc.wilson(U, psi, L, kappa)
    expect = ph[:, None, None] * (Dt @ u)[None]
    assert close(got, expect, 1e-13)
    # D^dagger D eigenvalue (1 - 2k sum cos p)^2 + 4 k^2 sum sin^2 p
    gotdd = orc.wilson(U, got, L, kappa, dagger=True)
    p = [(2 * np.pi * n[m] + (np.pi if m == 3 else 0)) / L[m] for m in range(4)]
    lam = (1 - 2 * kappa * sum(np.cos(p))) ** 2 + 4 * kappa ** 2 * sum(np.sin(p) ** 2)
    assert close(gotdd, lam * psi, 1e-13)


def test_adjoint_identity_and_gamma5_hermiticity():
    L = (4, 4, 4, 4)
    V = int(np.prod(L))
    U = gen.links(L, "haar", 7)
    psi = gen.gaussian_spinor(L, 8)
    chi = gen.gaussian_spinor(L, 9)
    D = orc.wilson(U, psi, L, 0.125)
    Dd = orc.wilson(U, chi, L, 0.125, dagger=True)
    lhs = np.vdot(chi, D)
    rhs = np.vdot(Dd, psi)
    assert abs(lhs - rhs) <= 1e-12 * abs(lhs)
    # D^dagger = gamma5 D gamma5 (test/wilson_clover.jl:268-276)
    g5 = dr.gamma5(orc.wilson(U, dr.gamma5(psi), L, 0.125))
    assert close(orc.wilson(U, psi, L, 0.125, dagger=True), g5, 1e-14)


def test_gauge_covariance():
    L = (4, 2, 2, 4)
    V = int(np.prod(L))
    U = gen.links(L, "haar", 21)
    psi = gen.gaussian_spinor(L, 22)
    Om = gen.haar_su3(V, 23)
    Ut, psit = dr.gauge_transform(U, psi, L, Om)
    for dag in (False, True):
        a = orc.wilson(Ut, psit, L, 0.12, dagger=dag)
        b = np.einsum("vij,vsj->vsi", Om, orc.wilson(U, psi, L, 0.12, dagger=dag))
        assert close(a, b, 1e-13)


def test_planted_mistakes_are_detected():
    """Each plausible oracle mistake is emulated on the CPU and must fail a pinning check."""
    L = (4, 2, 2, 4)
    V = int(np.prod(L))
    rng = np.random.default_rng(1)
    U = gen.links(L, "haar", 31)
    psi = gen.gaussian_spinor(L, 32)
    ref = dr.wilson(U, psi, L, 0.12)
    good = orc.wilson(U, psi, L, 0.12)
    assert close(good, ref, 1e-13)
    # wrong sign of the projector (forward uses 1+gamma): the D^dagger kernel in place of D
    assert not close(orc.wilson(U, psi, L, 0.12, dagger=True), ref, 1e-6)
    # transposed link operand
    assert not close(orc.wilson(np.swapaxes(U, -1, -2), psi, L, 0.12), ref, 1e-6)
    # dropped boundary phase (periodic t)
    assert not close(orc.wilson(U, psi, L, 0.12, phases=(1, 1, 1, 1)), ref, 1e-6)
    # dropped direction (no mu=2 hop)
    U2 = U.copy()
    U2[2] = 0
    assert not close(orc.wilson(U2, psi, L, 0.12), ref, 1e-6)
    # wrong sign of kappa
    assert not close(orc.wilson(U, psi, L, -0.12), ref, 1e-6)


def test_dot_orders_agree():
    L = (8, 4, 4, 4)
    a = gen.gaussian_spinor(L, 1)
    b = gen.gaussian_spinor(L, 2)
    exact = np.vdot(a, b)
    for order in (0, 1):
        d = orc.dot(a, b, order)
        assert abs(d - exact) <= 1e-13 * abs(exact)
