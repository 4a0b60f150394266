"""Pin the oracle's link pullback and dSFdU force by finite differences.

1. The pullback with seed w is the Enzyme-convention gradient of the real
   linear functional f(U) = Re <w, D(U) psi>: checked entry by entry by
   central differences in Re and Im of single link entries (f is affine in U,
   so the difference is exact up to rounding) -- the linear-functional form of
   test/wilson_dirac_ad.jl:85-104 and src/ND.jl:280-361.
2. dSFdU: dS/de of S = phi^dagger (D^dagger D)^-1 phi along U -> exp(eX) U,
   X in su(3), equals Re tr(F X) with F = TA(U G^dagger); the finite
   difference uses the local form S(U') - S(U) = -X'^dagger (A' - A) X."""
import numpy as np
import pytest

import oracle as orc
from lmx import generators as gen

L = (4, 4, 4, 4)
KAPPA = 0.125


def f_linear(U, w, psi):
    return np.vdot(w, orc.wilson(U, psi, L, KAPPA)).real


def test_pullback_is_gradient_of_linear_functional():
    U = gen.links(L, "haar", 41)
    psi = gen.gaussian_spinor(L, 42)
    w = gen.gaussian_spinor(L, 43)
    G = orc.wilson_pullback(w, psi, L, -KAPPA)
    rng = np.random.default_rng(44)
    V = int(np.prod(L))
    eps = 1e-3
    # include boundary sites in t (antiperiodic) and x
    sites = [0, V - 1, 63, 64 * 3 + 5] + list(rng.integers(0, V, 4))
    for s in sites:
        for mu in range(4):
            r, c = rng.integers(0, 3, 2)
            for part, unit in ((0, 1.0), (1, 1j)):
                Up, Um = U.copy(), U.copy()
                Up[mu, s, r, c] += eps * unit
                Um[mu, s, r, c] -= eps * unit
                fd = (f_linear(Up, w, psi) - f_linear(Um, w, psi)) / (2 * eps)
                an = G[mu, s, r, c].real if part == 0 else G[mu, s, r, c].imag
                assert abs(fd - an) <= 1e-9 * max(1.0, abs(an)), (s, mu, r, c, part, fd, an)


def local_delta_S(U, Up, phi, X, rtol):
    """S(Up) - S(U) = -X'^dagger (A' - A) X with X = A^-1 phi, X' = A'^-1 phi."""
    opp = orc.NormalOp(orc.OP_DDAGD_WILSON, L, U=Up, kappa=KAPPA)
    op = orc.NormalOp(orc.OP_DDAGD_WILSON, L, U=U, kappa=KAPPA)
    Xp, res = orc.cg(opp, phi, x0=X, rtol=rtol, maxiter=20000)
    assert res.converged
    dAX = opp.apply(X) - op.apply(X)
    return -np.vdot(Xp, dAX).real


def test_dsfdu_matches_finite_differences():
    U = gen.links(L, "haar", 51)
    eta = gen.gaussian_spinor(L, 52)
    phi = orc.wilson(U, eta, L, KAPPA, dagger=True)
    G, X, chi, res = orc.dsfdu(U, phi, L, KAPPA, rtol=1e-13)
    assert res.converged
    F = orc.ta_force(U, G)
    # F is traceless anti-Hermitian
    assert np.max(np.abs(F + np.conj(np.swapaxes(F, -1, -2)))) <= 1e-12 * np.max(np.abs(F))
    assert np.max(np.abs(np.trace(F, axis1=-2, axis2=-1))) <= 1e-12 * np.max(np.abs(F))
    V = int(np.prod(L))
    rng = np.random.default_rng(53)
    eps = 1e-5
    dirs = gen.random_su3_algebra(4, 54)
    for k in range(4):
        s = int(rng.integers(0, V))
        mu = int(rng.integers(0, 4))
        Xd = dirs[k]
        Up, Um = U.copy(), U.copy()
        Up[mu, s] = gen.expm_antihermitian(eps * Xd) @ U[mu, s]
        Um[mu, s] = gen.expm_antihermitian(-eps * Xd) @ U[mu, s]
        fd = (local_delta_S(U, Up, phi, X, 1e-13) - local_delta_S(U, Um, phi, X, 1e-13)) / (2 * eps)
        an_G = np.sum(np.conj(G[mu, s]) * (Xd @ U[mu, s])).real
        an_F = np.trace(F[mu, s] @ Xd).real
        assert abs(an_G - an_F) <= 1e-12 * abs(an_G)
        assert abs(fd - an_G) <= 1e-8 * abs(an_G), (k, fd, an_G)


def test_planted_force_mistakes_detected():
    """Wrong coefficient sign or a dropped backward-hop term in the pullback fails the FD check."""
    U = gen.links(L, "haar", 61)
    psi = gen.gaussian_spinor(L, 62)
    w = gen.gaussian_spinor(L, 63)
    G = orc.wilson_pullback(w, psi, L, -KAPPA)
    wrong_sign = orc.wilson_pullback(w, psi, L, +KAPPA)
    forward_only = orc.wilson_pullback(w, np.zeros_like(psi), L, -KAPPA)  # drops psi(x) terms
    s, mu, r, c, eps = 5, 1, 0, 2, 1e-3
    Up, Um = U.copy(), U.copy()
    Up[mu, s, r, c] += eps
    Um[mu, s, r, c] -= eps
    fd = (f_linear(Up, w, psi) - f_linear(Um, w, psi)) / (2 * eps)
    assert abs(fd - G[mu, s, r, c].real) <= 1e-9 * max(1, abs(fd))
    assert abs(fd - wrong_sign[mu, s, r, c].real) > 1e-6
    assert abs(fd - forward_only[mu, s, r, c].real) > 1e-6 or abs(fd) < 1e-12


# ---------------------------------------------------------------------------
# Wilson-clover: wilson_clover_link_pullback! (src/Operators/WilsonCloverPullback.jl)
# ---------------------------------------------------------------------------
CSW = 1.3


def f_clover(U, w, psi, kappa=KAPPA, csw=CSW):
    F = orc.clover_field_strength(U, L)
    return np.vdot(w, orc.wilson_clover(U, F, psi, L, kappa, csw)).real


def test_clover_pullback_is_gradient_of_linear_functional():
    # Re <w, D_clover(U) psi> is real-affine in any single link entry (each
    # leaf of the clover holds a link at most once), so central differences
    # are exact up to rounding; boundary sites in t (antiperiodic) and x included
    U = gen.links(L, "haar", 141)
    psi = gen.gaussian_spinor(L, 142)
    w = gen.gaussian_spinor(L, 143)
    G = orc.clover_pullback(U, w, psi, L, -KAPPA, CSW)
    rng = np.random.default_rng(144)
    V = int(np.prod(L))
    eps = 1e-3
    sites = [0, V - 1, 63, 64 * 3 + 5] + list(rng.integers(0, V, 3))
    worst = 0.0
    for s in sites:
        for mu in range(4):
            r, c = rng.integers(0, 3, 2)
            for part, unit in ((0, 1.0), (1, 1j)):
                Up, Um = U.copy(), U.copy()
                Up[mu, s, r, c] += eps * unit
                Um[mu, s, r, c] -= eps * unit
                fd = (f_clover(Up, w, psi) - f_clover(Um, w, psi)) / (2 * eps)
                an = G[mu, s, r, c].real if part == 0 else G[mu, s, r, c].imag
                worst = max(worst, abs(fd - an) / max(1.0, abs(an)))
                assert abs(fd - an) <= 1e-9 * max(1.0, abs(an)), (s, mu, r, c, part, fd, an)
    # the clover part is not negligible at this cSW: dropping it fails
    Gw = orc.wilson_pullback(w, psi, L, -KAPPA)
    assert np.max(np.abs(G - Gw)) > 1e-3 * np.max(np.abs(G))


def test_planted_clover_pullback_mistakes_detected():
    U = gen.links(L, "haar", 151)
    psi = gen.gaussian_spinor(L, 152)
    w = gen.gaussian_spinor(L, 153)
    good = orc.clover_pullback(U, w, psi, L, -KAPPA, CSW)
    wrong_sign = orc.clover_pullback(U, w, psi, L, -KAPPA, -CSW)
    no_clover = orc.clover_pullback(U, w, psi, L, -KAPPA, 0.0)
    s, mu, r, c, eps = 77, 2, 1, 0, 1e-3
    Up, Um = U.copy(), U.copy()
    Up[mu, s, r, c] += eps
    Um[mu, s, r, c] -= eps
    fd = (f_clover(Up, w, psi) - f_clover(Um, w, psi)) / (2 * eps)
    assert abs(fd - good[mu, s, r, c].real) <= 1e-9 * max(1, abs(fd))
    assert abs(fd - wrong_sign[mu, s, r, c].real) > 1e-6
    assert abs(fd - no_clover[mu, s, r, c].real) > 1e-6


def local_delta_S_clover(U, Up, phi, X, rtol):
    Fp, F = orc.clover_field_strength(Up, L), orc.clover_field_strength(U, L)
    opp = orc.NormalOp(orc.OP_DDAGD_CLOVER, L, U=Up, F=Fp, kappa=KAPPA, csw=CSW)
    op = orc.NormalOp(orc.OP_DDAGD_CLOVER, L, U=U, F=F, kappa=KAPPA, csw=CSW)
    Xp, res = orc.cg(opp, phi, x0=X, rtol=rtol, maxiter=20000)
    assert res.converged
    dAX = opp.apply(X) - op.apply(X)
    return -np.vdot(Xp, dAX).real


def test_dsfdu_clover_matches_finite_differences():
    U = gen.links(L, "weak", 161, eps=0.3)
    eta = gen.gaussian_spinor(L, 162)
    F = orc.clover_field_strength(U, L)
    phi = orc.wilson_clover(U, F, eta, L, KAPPA, CSW, dagger=True)
    G, X, chi, res = orc.dsfdu_clover(U, phi, L, KAPPA, CSW, rtol=1e-13)
    assert res.converged
    # S = phi^dagger (D^dagger D)^-1 phi = |eta|^2 for phi = D^dagger eta
    assert abs(np.vdot(phi, X).real - np.vdot(eta, eta).real) <= 1e-11 * np.vdot(eta, eta).real
    Fo = orc.ta_force(U, G)
    V = int(np.prod(L))
    rng = np.random.default_rng(163)
    eps = 1e-5
    dirs = gen.random_su3_algebra(4, 164)
    for k in range(4):
        s = int(rng.integers(0, V))
        mu = int(rng.integers(0, 4))
        Xd = dirs[k]
        Up, Um = U.copy(), U.copy()
        Up[mu, s] = gen.expm_antihermitian(eps * Xd) @ U[mu, s]
        Um[mu, s] = gen.expm_antihermitian(-eps * Xd) @ U[mu, s]
        fd = (local_delta_S_clover(U, Up, phi, X, 1e-13) - local_delta_S_clover(U, Um, phi, X, 1e-13)) / (2 * eps)
        an_G = np.sum(np.conj(G[mu, s]) * (Xd @ U[mu, s])).real
        an_F = np.trace(Fo[mu, s] @ Xd).real
        assert abs(an_G - an_F) <= 1e-12 * abs(an_G)
        assert abs(fd - an_G) <= 1e-8 * abs(an_G), (k, fd, an_G)
