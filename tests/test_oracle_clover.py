"""Pin the oracle's clover field strength and Wilson-clover apply against the
dense restatement of test/wilson_clover.jl:47-112 and its invariants
(:180 anti-Hermitian F, :225-250 cSW=0 and unit gauge, :268-303
gamma5-hermiticity and gauge covariance)."""
import numpy as np

import dense_reference as dr
import oracle as orc
from lmx import generators as gen


def close(a, b, tol):
    return np.max(np.abs(a - b)) <= tol * max(1.0, np.max(np.abs(b)))


def test_field_strength_matches_dense():
    L = (4, 2, 3, 4)
    rng = np.random.default_rng(4)
    V = int(np.prod(L))
    U = rng.random((4, V, 3, 3)) + 1j * rng.random((4, V, 3, 3))  # non-unitary, as rand() in the reference test
    assert close(orc.clover_field_strength(U, L), dr.field_strength(U, L), 3e-12)


def test_field_strength_antihermitian_and_unit_gauge():
    L = (4, 4, 4, 4)
    U = gen.links(L, "haar", 5)
    F = orc.clover_field_strength(U, L)
    assert close(F, -np.conj(np.swapaxes(F, -1, -2)), 1e-15)
    Fu = orc.clover_field_strength(gen.links(L, "unit", 0), L)
    assert np.max(np.abs(Fu)) == 0.0


def test_wilson_clover_matches_dense():
    L = (4, 2, 2, 4)
    U = gen.links(L, "haar", 6)
    psi = gen.gaussian_spinor(L, 7)
    F = orc.clover_field_strength(U, L)
    for dag in (False, True):
        got = orc.wilson_clover(U, F, psi, L, 0.117, 1.17, dagger=dag)
        ref = dr.wilson_clover(U, psi, L, 0.117, 1.17, dagger=dag)
        assert close(got, ref, 3e-12)


def test_csw_zero_is_wilson_and_gamma5_hermiticity():
    L = (4, 4, 2, 4)
    U = gen.links(L, "weak", 8, eps=0.15)
    psi = gen.gaussian_spinor(L, 9)
    F = orc.clover_field_strength(U, L)
    assert close(orc.wilson_clover(U, F, psi, L, 0.12, 0.0), orc.wilson(U, psi, L, 0.12), 0.0)
    Dd = orc.wilson_clover(U, F, psi, L, 0.12, 1.0, dagger=True)
    g5Dg5 = dr.gamma5(orc.wilson_clover(U, F, dr.gamma5(psi), L, 0.12, 1.0))
    assert close(Dd, g5Dg5, 1e-14)
    chi = gen.gaussian_spinor(L, 10)
    lhs = np.vdot(chi, orc.wilson_clover(U, F, psi, L, 0.12, 1.0))
    rhs = np.vdot(orc.wilson_clover(U, F, chi, L, 0.12, 1.0, dagger=True), psi)
    assert abs(lhs - rhs) <= 1e-12 * abs(lhs)


def test_gauge_covariance_clover():
    L = (4, 2, 2, 4)
    V = int(np.prod(L))
    U = gen.links(L, "haar", 11)
    psi = gen.gaussian_spinor(L, 12)
    Om = gen.haar_su3(V, 13)
    Ut, psit = dr.gauge_transform(U, psi, L, Om)
    F, Ft = orc.clover_field_strength(U, L), orc.clover_field_strength(Ut, L)
    assert close(Ft, Om[None] @ F @ np.conj(np.swapaxes(Om, -1, -2))[None], 1e-13)
    a = orc.wilson_clover(Ut, Ft, psit, L, 0.12, 1.3)
    b = np.einsum("vij,vsj->vsi", Om, orc.wilson_clover(U, F, psi, L, 0.12, 1.3))
    assert close(a, b, 1e-13)


def test_planted_clover_mistakes_detected():
    L = (4, 2, 2, 4)
    U = gen.links(L, "haar", 14)
    psi = gen.gaussian_spinor(L, 15)
    F = orc.clover_field_strength(U, L)
    ref = dr.wilson_clover(U, psi, L, 0.12, 1.0)
    assert close(orc.wilson_clover(U, F, psi, L, 0.12, 1.0), ref, 1e-12)
    # planes in the wrong order (13 <-> 14)
    Fb = F[[0, 2, 1, 3, 4, 5]]
    assert not close(orc.wilson_clover(U, Fb, psi, L, 0.12, 1.0), ref, 1e-6)
    # F transposed (a wrong index order in the leaf products)
    assert not close(orc.wilson_clover(U, np.swapaxes(F, -1, -2), psi, L, 0.12, 1.0), ref, 1e-6)
    # wrong sign of cSW
    assert not close(orc.wilson_clover(U, F, psi, L, 0.12, -1.0), ref, 1e-6)
