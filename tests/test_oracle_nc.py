"""Pin the oracle's generic-NC kernels (lmx_nc_oracle.c: the reference's
Wilson, clover and link-pullback kernels for NC != 3) with what fixes them
independently of the oracle's code:
  - the dense restatement of test/wilson_clover.jl:47-112 (any NC) on the
    reference's own closed-form SU(2) fields (:12-45), at the reference's
    tolerances (F 2e-12, D 3e-12);
  - F anti-Hermitian, zero for unit links; cSW = 0 equal to Wilson;
  - the adjoint identity and gamma5-Hermiticity (:253-275), gauge covariance
    under the test's SU(2) transformation (:277-301);
  - the nw = 0 fixture of test/nw0.jl:213-254 (D, D^dagger, Donly, Donly^dagger);
  - the generic order at NC = 3 against the half-spin oracle (same operator,
    different rounding);
  - exact central differences of Re <w, D(U) psi> in single link entries
    for the pullback, and the reference's directional difference (:187-217);
  - planted mistakes (projector sign, transposed links, dropped phase, cSW
    sign) are detected.
"""
import numpy as np
import pytest

import dense_reference as dr
import oracle as orc
import su2_fixtures as fx

L2 = (2, 2, 2, 2)
PHASES = (1, 1, 1, -1)
KAPPA, CSW = 0.12, 1.17


def close(a, b, tol):
    return np.max(np.abs(a - b)) <= tol * max(1.0, np.max(np.abs(b)))


def test_clover_field_strength_su2_matches_dense():
    U = fx.clover_test_links(L2)
    F = orc.clover_field_strength_nc(U, L2)
    ref = dr.field_strength(U, L2)
    assert close(F, ref, 2e-12)
    assert np.array_equal(np.conj(np.swapaxes(F, -1, -2)), -F)  # anti-Hermitian as stored
    unit = np.broadcast_to(np.eye(2, dtype=np.complex128), (4, 16, 2, 2)).copy()
    assert not np.any(orc.clover_field_strength_nc(unit, L2))


@pytest.mark.parametrize("dagger", [False, True])
def test_wilson_clover_su2_matches_dense(dagger):
    U = fx.clover_test_links(L2)
    psi = fx.clover_test_spinor(L2)
    F = orc.clover_field_strength_nc(U, L2)
    got = orc.wilson_clover_nc(U, F, psi, L2, KAPPA, CSW, PHASES, dagger)
    ref = dr.wilson_clover(U, psi, L2, KAPPA, CSW, PHASES, dagger)
    assert close(got, ref, 3e-12)
    # cSW = 0 is the Wilson operator
    w = orc.wilson_nc(U, psi, L2, KAPPA, PHASES, dagger)
    assert close(orc.wilson_clover_nc(U, F, psi, L2, KAPPA, 0.0, PHASES, dagger), w, 2e-12)
    assert close(w, dr.wilson(U, psi, L2, KAPPA, PHASES, dagger), 2e-12)


def test_adjoint_and_gamma5_hermiticity_su2():
    U = fx.clover_test_links(L2)
    psi = fx.clover_test_spinor(L2)
    chi = fx.clover_test_spinor(L2, offset=0.19)
    F = orc.clover_field_strength_nc(U, L2)
    D = lambda v, dag: orc.wilson_clover_nc(U, F, v, L2, KAPPA, CSW, PHASES, dag)
    lhs = np.vdot(chi, D(psi, False))
    rhs = np.vdot(D(chi, True), psi)
    assert abs(lhs - rhs) <= 5e-11 + 5e-12 * abs(lhs)
    assert close(D(psi, True), dr.gamma5(D(dr.gamma5(psi), False)), 5e-12)


def test_gauge_covariance_su2():
    U = fx.clover_test_links(L2)
    psi = fx.clover_test_spinor(L2)
    Om = fx.clover_test_transformations(L2)
    Ut, psit = dr.gauge_transform(U, psi, L2, Om)
    F = orc.clover_field_strength_nc(U, L2)
    Ft = orc.clover_field_strength_nc(Ut, L2)
    got = orc.wilson_clover_nc(Ut, Ft, psit, L2, KAPPA, CSW, PHASES)
    expect = np.einsum("vij,vsj->vsi", Om, orc.wilson_clover_nc(U, F, psi, L2, KAPPA, CSW, PHASES))
    assert close(got, expect, 8e-12)


@pytest.mark.parametrize("dagger,donly", [(False, False), (True, False), (False, True), (True, True)])
def test_nw0_fixture_matches_dense(dagger, donly):
    # test/nw0.jl:213-254: global size (4, 2, 2, 2), NC = 2, kappa 0.12
    L = (4, 2, 2, 2)
    U = fx.nw0_links(L)
    psi = fx.nw0_spinor(L)
    got = orc.wilson_nc(U, psi, L, 0.12, PHASES, dagger, donly)
    ref = dr.wilson(U, psi, L, 0.12, PHASES, dagger, donly)
    assert close(got, ref, 1e-14)


@pytest.mark.parametrize("dagger,donly", [(False, False), (True, True)])
def test_generic_order_at_nc3_agrees_with_half_spin_oracle(dagger, donly):
    rng = np.random.default_rng(7)
    L = (4, 2, 2, 4)
    V = int(np.prod(L))
    U = rng.standard_normal((4, V, 3, 3)) + 1j * rng.standard_normal((4, V, 3, 3))
    psi = rng.standard_normal((V, 4, 3)) + 1j * rng.standard_normal((V, 4, 3))
    a = orc.wilson_nc(U, psi, L, 0.117, PHASES, dagger, donly)
    b = orc.wilson(U, psi, L, 0.117, PHASES, dagger, donly)
    assert close(a, b, 1e-14) and not np.array_equal(a, b)  # same operator, another rounding


def test_complex_phases_nc2():
    rng = np.random.default_rng(9)
    L = (3, 2, 2, 4)
    V = int(np.prod(L))
    U = rng.standard_normal((4, V, 2, 2)) + 1j * rng.standard_normal((4, V, 2, 2))
    psi = rng.standard_normal((V, 4, 2)) + 1j * rng.standard_normal((V, 4, 2))
    phases = (np.exp(0.3j), 1, -1j, -1)
    for dag in (False, True):
        assert close(orc.wilson_nc(U, psi, L, 0.13, phases, dag), dr.wilson(U, psi, L, 0.13, phases, dag), 1e-13)


def test_link_pullback_exact_central_differences_nc2():
    # Re <w, D(U) psi> is affine in every link entry: central differences are exact
    rng = np.random.default_rng(5)
    L = (2, 2, 2, 4)
    V = int(np.prod(L))
    U = rng.standard_normal((4, V, 2, 2)) + 1j * rng.standard_normal((4, V, 2, 2))
    psi = rng.standard_normal((V, 4, 2)) + 1j * rng.standard_normal((V, 4, 2))
    w = rng.standard_normal((V, 4, 2)) + 1j * rng.standard_normal((V, 4, 2))
    G = orc.wilson_pullback_nc(w, psi, L, -KAPPA, PHASES)

    def f(links):
        return float(np.real(np.vdot(w, orc.wilson_nc(links, psi, L, KAPPA, PHASES))))

    for (mu, s, r, c) in [(0, 0, 0, 1), (3, V - 1, 1, 0), (2, 5, 1, 1), (1, 9, 0, 0), (3, 0, 0, 0)]:
        for unit in (1.0, 1j):
            Up, Um = U.copy(), U.copy()
            Up[mu, s, r, c] += unit
            Um[mu, s, r, c] -= unit
            fd = (f(Up) - f(Um)) / 2
            # G = dS/dRe U + i dS/dIm U
            an = G[mu, s, r, c].real if unit == 1.0 else G[mu, s, r, c].imag
            assert abs(fd - an) <= 1e-11 * max(1.0, abs(fd))


def test_link_pullback_reference_directional_difference_nc2():
    # test/wilson_clover.jl:187-217 with the Wilson operator: directions are
    # _clover_test_links, eps 1e-6, rtol 3e-7
    U = fx.clover_test_links(L2)
    psi = fx.clover_test_spinor(L2)
    chi = fx.clover_test_spinor(L2, offset=0.19)
    dirs = fx.clover_test_links(L2)
    G = orc.wilson_pullback_nc(chi, psi, L2, -KAPPA, PHASES)
    eps = 1e-6
    f = lambda links: float(np.real(np.vdot(chi, orc.wilson_nc(links, psi, L2, KAPPA, PHASES))))
    fd = (f(U + eps * dirs) - f(U - eps * dirs)) / (2 * eps)
    an = float(np.real(np.vdot(G, dirs)))
    assert abs(an - fd) <= 3e-6 + 3e-7 * abs(fd)


def test_cg_nc2_solves_ddagger_d():
    # phi = D^dagger eta: the solution x of D^dagger D x = phi has D x = eta and S = |eta|^2
    L = (2, 2, 2, 4)
    U = fx.clover_test_links(L)
    eta = fx.clover_test_spinor(L, offset=0.37)
    F = orc.clover_field_strength_nc(U, L)
    phi = orc.wilson_clover_nc(U, F, eta, L, KAPPA, CSW, PHASES, dagger=True)
    op = orc.NormalOp(orc.OP_DDAGD_CLOVER, L, U=U, F=F, kappa=KAPPA, csw=CSW, phases=PHASES, nc=2)
    x, res = orc.cg(op, phi, rtol=1e-12)
    assert res.converged
    assert close(orc.wilson_clover_nc(U, F, x, L, KAPPA, CSW, PHASES), eta, 1e-10)
    assert abs(np.real(np.vdot(phi, x)) - np.real(np.vdot(eta, eta))) <= 1e-10 * np.real(np.vdot(eta, eta))
    assert abs(orc.dot_n(phi, x, 2).real - np.real(np.vdot(phi, x))) <= 1e-12 * abs(np.vdot(phi, x))


def test_planted_mistakes_detected_nc2():
    U = fx.clover_test_links(L2)
    psi = fx.clover_test_spinor(L2)
    ref = dr.wilson(U, psi, L2, KAPPA, PHASES)
    good = orc.wilson_nc(U, psi, L2, KAPPA, PHASES)
    assert close(good, ref, 2e-12)
    # wrong projector sign (D^dagger for D), transposed links, dropped phase, kappa sign
    assert not close(orc.wilson_nc(U, psi, L2, KAPPA, PHASES, dagger=True), ref, 1e-6)
    assert not close(orc.wilson_nc(np.swapaxes(U, -1, -2), psi, L2, KAPPA, PHASES), ref, 1e-6)
    assert not close(orc.wilson_nc(U, psi, L2, KAPPA, (1, 1, 1, 1)), ref, 1e-6)
    assert not close(orc.wilson_nc(U, psi, L2, -KAPPA, PHASES), ref, 1e-6)
    F = orc.clover_field_strength_nc(U, L2)
    refc = dr.wilson_clover(U, psi, L2, KAPPA, CSW, PHASES)
    assert not close(orc.wilson_clover_nc(U, F, psi, L2, KAPPA, -CSW, PHASES), refc, 1e-6)
    assert not close(orc.wilson_clover_nc(U, np.swapaxes(F, -1, -2), psi, L2, KAPPA, CSW, PHASES), refc, 1e-6)
