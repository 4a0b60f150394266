"""The reference tests' own closed-form NC = 2 inputs, as numpy arrays.

  clover_test_links / clover_test_spinor / clover_test_transformations
        test/wilson_clover.jl:12-45 and :125-150 (SU(2) links from angles and
        phases of the 1-based site coordinate, ramp spinors)
  nw0_links / nw0_spinor
        test/nw0.jl:213-229 (diagonal links 1 + 0.1 d + 0.01 i x1, the spinor
        1..N in column-major order)
Layouts as in oracle.py with NC = 2: spinor [V, 4, nc] (site x-fastest,
spin, colour), links [4, V, nc, nc] (mu, site, row, col).
"""
from __future__ import annotations

import itertools

import numpy as np


def sites(L):
    """1-based coordinates (x1, x2, x3, x4) of every site, x fastest."""
    for t, z, y, x in itertools.product(range(1, L[3] + 1), range(1, L[2] + 1), range(1, L[1] + 1),
                                        range(1, L[0] + 1)):
        yield (x, y, z, t)


def su2(angle, phase):
    """_clover_test_su2 (test/wilson_clover.jl:12-17)."""
    c, s, z = np.cos(angle), np.sin(angle), np.exp(1j * phase)
    return np.array([[c, s * z], [-s * np.conj(z), c]], dtype=np.complex128)


def clover_test_links(L):
    """_clover_test_links (test/wilson_clover.jl:19-31)."""
    V = int(np.prod(L))
    U = np.zeros((4, V, 2, 2), dtype=np.complex128)
    for s, x in enumerate(sites(L)):
        coordinate = x[0] + 2 * x[1] + 3 * x[2] + 5 * x[3]
        for mu in range(1, 5):
            U[mu - 1, s] = su2(0.071 * (coordinate + 2 * mu), 0.113 * (2 * coordinate - mu))
    return U


def clover_test_spinor(L, offset=0.0):
    """_clover_test_spinor (test/wilson_clover.jl:33-45)."""
    V = int(np.prod(L))
    psi = np.zeros((V, 4, 2), dtype=np.complex128)
    for s, x in enumerate(sites(L)):
        coordinate = x[0] + 2 * x[1] + 3 * x[2] + 5 * x[3]
        for spin in range(1, 5):
            for color in range(1, 3):
                psi[s, spin - 1, color - 1] = (offset + 0.031 * (color + 2 * spin + coordinate)
                                               + 1j * 0.047 * (2 * color - spin + coordinate))
    return psi


def clover_test_transformations(L):
    """The SU(2) gauge transformation of _clover_test_gauge_transform
    (test/wilson_clover.jl:125-135), [V, 2, 2]."""
    V = int(np.prod(L))
    Om = np.zeros((V, 2, 2), dtype=np.complex128)
    for s, x in enumerate(sites(L)):
        coordinate = x[0] + 3 * x[1] + 4 * x[2] + 7 * x[3]
        Om[s] = su2(0.053 * coordinate, 0.097 * (coordinate + 1))
    return Om


def nw0_links(L, nc=2):
    """gauge_arrays of the "nw=0 Wilson operators" test (test/nw0.jl:219-225):
    U_d[ic, ic, x] = 1 + 0.1 d + 0.01 i x1, off-diagonals zero."""
    V = int(np.prod(L))
    U = np.zeros((4, V, nc, nc), dtype=np.complex128)
    for s, x in enumerate(sites(L)):
        for d in range(1, 5):
            for ic in range(nc):
                U[d - 1, s, ic, ic] = 1 + 0.1 * d + 0.01j * x[0]
    return U


def nw0_spinor(L, nc=2):
    """psiarray = reshape(ComplexF64.(1:(NC*4*V)), NC, 4, L...) (test/nw0.jl:226-227):
    column-major, so element (colour, spin, site) holds 1 + colour + nc*spin + 4*nc*site."""
    V = int(np.prod(L))
    return (np.arange(1, nc * 4 * V + 1, dtype=np.float64).reshape(V, 4, nc) + 0j).astype(np.complex128)


def random_su2(n, seed):
    """n Haar-random SU(2) matrices [n, 2, 2]: a0 + i a.sigma with a uniform on S^3."""
    rng = np.random.default_rng(seed)
    a = rng.standard_normal((n, 4))
    a /= np.linalg.norm(a, axis=1, keepdims=True)
    m = np.empty((n, 2, 2), dtype=np.complex128)
    m[:, 0, 0] = a[:, 0] + 1j * a[:, 3]
    m[:, 0, 1] = a[:, 2] + 1j * a[:, 1]
    m[:, 1, 0] = -a[:, 2] + 1j * a[:, 1]
    m[:, 1, 1] = a[:, 0] - 1j * a[:, 3]
    return m
