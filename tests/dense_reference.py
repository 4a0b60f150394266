"""Dense numpy restatements of the reference's in-test references.

These follow the reference tests' own formulation (explicit gamma matrices,
explicit 3x3 products, phase^wraps), not the half-spin kernels, so they pin
the oracle independently:
  _wilson_halfspin_reference     test/wilson_halfspin.jl:1-42
  _clover_test_field_strength    test/wilson_clover.jl:47-78
  _clover_test_reference         test/wilson_clover.jl:80-112
Arrays: spinor [V,4,3] (site, spin, colour); links [4,V,3,3] (mu, site, row, col);
sites x-fastest.
"""
from __future__ import annotations

import numpy as np

# src/Operators/Operators.jl:44-67
GAMMA = np.array([
    [[0, 0, 0, -1j], [0, 0, -1j, 0], [0, 1j, 0, 0], [1j, 0, 0, 0]],
    [[0, 0, 0, -1], [0, 0, 1, 0], [0, 1, 0, 0], [-1, 0, 0, 0]],
    [[0, 0, -1j, 0], [0, 0, 0, 1j], [1j, 0, 0, 0], [0, -1j, 0, 0]],
    [[0, 0, -1, 0], [0, 0, 0, -1], [-1, 0, 0, 0], [0, -1, 0, 0]],
], dtype=np.complex128)
GAMMA5 = GAMMA[0] @ GAMMA[1] @ GAMMA[2] @ GAMMA[3]
PLANES = ((0, 1), (0, 2), (0, 3), (1, 2), (1, 3), (2, 3))


def grid(L, a):
    """[V, ...] -> [Lt, Lz, Ly, Lx, ...] (x fastest)."""
    return a.reshape((L[3], L[2], L[1], L[0]) + a.shape[1:])


def flat(L, a):
    V = int(np.prod(L))
    return a.reshape((V,) + a.shape[4:])


_AX = {0: 3, 1: 2, 2: 1, 3: 0}  # lattice direction -> grid axis


def shift(L, a, mu, amount, phases=None):
    """a(x + amount*mu) with phase^wraps (amount = +-1), on the grid layout."""
    ax = _AX[mu]
    out = np.roll(a, -amount, axis=ax)
    if phases is not None and complex(phases[mu]) != 1:
        ph = complex(phases[mu]) ** (1 if amount > 0 else -1)
        idx = [slice(None)] * out.ndim
        idx[ax] = L[mu] - 1 if amount > 0 else 0
        out = out.copy()
        out[tuple(idx)] *= ph
    return out


def wilson(U, psi, L, kappa, phases=(1, 1, 1, -1), dagger=False, donly=False):
    """test/wilson_halfspin.jl:8-42 in numpy."""
    Ug = grid(L, np.asarray(U).transpose(1, 0, 2, 3))  # [..., mu, 3, 3]
    P = grid(L, np.asarray(psi))                        # [..., 4, 3]
    out = np.zeros_like(P) if donly else P.copy()
    I4 = np.eye(4, dtype=np.complex128)
    plus_sign = 1 if dagger else -1
    for mu in range(4):
        Pp = shift(L, P, mu, +1, phases)
        Pm = shift(L, P, mu, -1, phases)
        Um = shift(L, Ug[..., mu, :, :], mu, -1)
        proj_p = I4 + plus_sign * GAMMA[mu]
        proj_m = I4 - plus_sign * GAMMA[mu]
        # reference: U psi_ref P^T with psi_ref[c, s]; here psi[s, c] -> P psi U^T
        hop = np.einsum("st,...tc,...dc->...sd", proj_p, Pp, Ug[..., mu, :, :])
        hop = hop + np.einsum("st,...tc,...cd->...sd", proj_m, Pm, np.conj(Um))  # U^dagger[d,c] = conj(U[c,d])
        if donly:
            out += 0.5 * hop
        else:
            out -= kappa * hop
    return flat(L, out)


def field_strength(U, L):
    """test/wilson_clover.jl:47-78: F[plane] = (Q - Q^dagger)/8, four-leaf Q."""
    Ug = grid(L, np.asarray(U).transpose(1, 0, 2, 3))
    dag = lambda m: np.conj(np.swapaxes(m, -1, -2))
    F = []
    for mu, nu in PLANES:
        Umu, Unu = Ug[..., mu, :, :], Ug[..., nu, :, :]
        s = lambda a, d, k: shift(L, a, d, k)
        Q = (Umu @ s(Unu, mu, 1) @ dag(s(Umu, nu, 1)) @ dag(Unu)
             + Unu @ dag(s(s(Umu, nu, 1), mu, -1)) @ dag(s(Unu, mu, -1)) @ s(Umu, mu, -1)
             + dag(s(Umu, mu, -1)) @ dag(s(s(Unu, mu, -1), nu, -1)) @ s(s(Umu, mu, -1), nu, -1) @ s(Unu, nu, -1)
             + dag(s(Unu, nu, -1)) @ s(Umu, nu, -1) @ s(s(Unu, nu, -1), mu, 1) @ dag(Umu))
        F.append(flat(L, 0.125 * (Q - dag(Q))))
    return np.stack(F)


def wilson_clover(U, psi, L, kappa, csw, phases=(1, 1, 1, -1), dagger=False):
    """test/wilson_clover.jl:80-112 (the adjoint adds the same Hermitian term)."""
    out = wilson(U, psi, L, kappa, phases, dagger=dagger)
    F = field_strength(U, L)
    P = np.asarray(psi)
    for pl, (mu, nu) in enumerate(PLANES):
        gg = GAMMA[mu] @ GAMMA[nu]
        # reference: F psi_ref (g_mu g_nu)^T -> here (g g) psi F^T
        out = out - kappa * csw * np.einsum("st,vtc,vdc->vsd", gg, P, F[pl])
    return out


def gamma5(psi):
    return np.einsum("st,...tc->...sc", GAMMA5, psi)


def gauge_transform(U, psi, L, Omega):
    """U_mu(x) -> Om(x) U_mu(x) Om(x+mu)^dagger, psi(x) -> Om(x) psi(x)."""
    Ug = grid(L, np.asarray(U).transpose(1, 0, 2, 3))
    Og = grid(L, Omega)
    out = []
    for mu in range(4):
        Op = shift(L, Og, mu, 1)
        out.append(flat(L, Og @ Ug[..., mu, :, :] @ np.conj(np.swapaxes(Op, -1, -2))))
    Ut = np.stack(out)
    psit = np.einsum("vij,vsj->vsi", Omega, psi)
    return Ut, psit


def free_field_eigen(L, n, kappa, phases=(1, 1, 1, -1)):
    """Plane wave e^{i p.x} u with p_mu = (2 pi n_mu + theta_mu)/L_mu,
    theta = arg(phase); returns (phase field [V], Dtilde(p) [4,4]) with
    D psi = e^{ipx} Dtilde u for U = 1:  Dtilde = 1 - kappa sum (2cos p - 2i g sin p)."""
    p = []
    for mu in range(4):
        theta = np.angle(complex(phases[mu]))
        p.append((2 * np.pi * n[mu] + theta) / L[mu])
    x = np.indices((L[3], L[2], L[1], L[0]))  # t, z, y, x
    coords = [x[3], x[2], x[1], x[0]]
    ph = np.exp(1j * sum(p[mu] * coords[mu] for mu in range(4))).reshape(-1)
    Dt = np.eye(4, dtype=np.complex128)
    for mu in range(4):
        Dt = Dt - kappa * (2 * np.cos(p[mu]) * np.eye(4) - 2j * GAMMA[mu] * np.sin(p[mu]))
    return ph, Dt


def halo_fill(core, dims, nw, phases):
    """set_halo! on one rank (src/LatticeMatrices_core.jl:707-723) of a core
    array [L4, L3, L2, L1, NC2, NC1] (C order): periodic copy into a halo of
    width nw, plus ghosts times the phase, minus ghosts times its inverse,
    direction by direction (a corner takes the phases of every direction it
    wraps across)."""
    a = np.asarray(core).reshape(dims[3], dims[2], dims[1], dims[0], *np.asarray(core).shape[-2:])
    if nw == 0:
        return a.copy()
    a = np.pad(a, [(nw, nw)] * 4 + [(0, 0), (0, 0)], mode="wrap")
    for d in range(4):
        ph = phases[d]
        if ph == 1:
            continue
        ax = 3 - d
        lo = [slice(None)] * 6
        hi = [slice(None)] * 6
        lo[ax] = slice(0, nw)
        hi[ax] = slice(dims[d] + nw, dims[d] + 2 * nw)
        a[tuple(lo)] = a[tuple(lo)] * a.dtype.type(1 / ph)
        a[tuple(hi)] = a[tuple(hi)] * a.dtype.type(ph)
    return a
