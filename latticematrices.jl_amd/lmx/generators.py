"""Seeded synthetic inputs for the Wilson(-clover) D^dagger D path.

All generators are deterministic functions of (shape, seed) built on numpy's
PCG64 Generator, so the device path and the CPU oracle see identical inputs.

Array conventions (complex128, C order):
  links   U[4, V, 3, 3]   (mu, site, row, col)
  spinor  psi[V, 4, 3]    (site, spin, colour); sites x-fastest.
"""
from __future__ import annotations

import numpy as np

# Gell-Mann matrices lambda_1..lambda_8
_GM = np.zeros((8, 3, 3), dtype=np.complex128)
_GM[0][0, 1] = _GM[0][1, 0] = 1
_GM[1][0, 1], _GM[1][1, 0] = -1j, 1j
_GM[2][0, 0], _GM[2][1, 1] = 1, -1
_GM[3][0, 2] = _GM[3][2, 0] = 1
_GM[4][0, 2], _GM[4][2, 0] = -1j, 1j
_GM[5][1, 2] = _GM[5][2, 1] = 1
_GM[6][1, 2], _GM[6][2, 1] = -1j, 1j
_GM[7] = np.diag([1, 1, -2]) / np.sqrt(3)
GELL_MANN = _GM


def volume(L) -> int:
    return int(np.prod([int(v) for v in L]))


def _rng(seed):
    return np.random.Generator(np.random.PCG64(int(seed)))


def haar_su3(n: int, seed: int, chunk: int = 1 << 20) -> np.ndarray:
    """n Haar-random SU(3) matrices: Z = (G1 + i G2)/sqrt(2) with G ~ N(0,1) per
    entry, Q,R = qr(Z), U = Q diag(R_ii/|R_ii|), U <- U / det(U)^(1/3)
    (principal branch).  Returns complex128 [n, 3, 3]."""
    rng = _rng(seed)
    out = np.empty((n, 3, 3), dtype=np.complex128)
    for s in range(0, n, chunk):
        m = min(chunk, n - s)
        g = rng.standard_normal((m, 3, 3, 2))
        z = (g[..., 0] + 1j * g[..., 1]) / np.sqrt(2.0)
        q, r = np.linalg.qr(z)
        d = np.diagonal(r, axis1=1, axis2=2)
        q = q * (d / np.abs(d))[:, None, :]
        det = np.linalg.det(q)
        q = q / (det ** (1.0 / 3.0))[:, None, None]
        out[s:s + m] = q
    return out


def su3_exp_i_hermitian(H: np.ndarray) -> np.ndarray:
    """exp(i H) for Hermitian H[..., 3, 3] via the eigen-decomposition."""
    w, v = np.linalg.eigh(H)
    return (v * np.exp(1j * w)[..., None, :]) @ np.conj(np.swapaxes(v, -1, -2))


def weak_field_su3(n: int, eps: float, seed: int, chunk: int = 1 << 20) -> np.ndarray:
    """n weak-field SU(3) links U = exp(i eps H), H = sum_a h_a lambda_a / 2
    with h_a ~ N(0,1) (traceless Hermitian Gaussian).  complex128 [n, 3, 3]."""
    rng = _rng(seed)
    out = np.empty((n, 3, 3), dtype=np.complex128)
    for s in range(0, n, chunk):
        m = min(chunk, n - s)
        h = rng.standard_normal((m, 8))
        H = np.einsum("na,aij->nij", h, _GM) * 0.5
        out[s:s + m] = su3_exp_i_hermitian(eps * H)
    return out


def links(L, kind: str, seed: int, eps: float = 0.2) -> np.ndarray:
    """Gauge field U[4, V, 3, 3]; kind in {'haar', 'weak', 'unit'}."""
    V = volume(L)
    if kind == "haar":
        return haar_su3(4 * V, seed).reshape(4, V, 3, 3)
    if kind == "weak":
        return weak_field_su3(4 * V, eps, seed).reshape(4, V, 3, 3)
    if kind == "unit":
        return np.broadcast_to(np.eye(3, dtype=np.complex128), (4, V, 3, 3)).copy()
    raise ValueError(f"unknown link kind {kind!r}")


def gaussian_spinor(L, seed: int, sigma: float = np.sqrt(0.5), nrhs: int | None = None) -> np.ndarray:
    """Complex Gaussian spinor(s), re and im ~ N(0, sigma^2) (sigma = sqrt(1/2)
    as benchmark/ddagd_cg/ddagd_cg.jl:773-774).  [V,4,3] or [nrhs,V,4,3]."""
    V = volume(L)
    rng = _rng(seed)
    shape = (V, 4, 3) if nrhs is None else (nrhs, V, 4, 3)
    g = rng.standard_normal(shape + (2,)) * sigma
    return g[..., 0] + 1j * g[..., 1]


def spin_colour_sources(L, seed: int, sigma: float = np.sqrt(0.5)) -> np.ndarray:
    """12 spin-colour diluted Gaussian sources [12, V, 4, 3]: source k = spin*3
    + colour carries Gaussian noise on that component only."""
    V = volume(L)
    rng = _rng(seed)
    out = np.zeros((12, V, 4, 3), dtype=np.complex128)
    for k in range(12):
        g = rng.standard_normal((V, 2)) * sigma
        out[k, :, k // 3, k % 3] = g[:, 0] + 1j * g[:, 1]
    return out


def random_su3_algebra(n: int, seed: int) -> np.ndarray:
    """n random su(3) directions X = i sum_a x_a lambda_a / 2, normalised so
    that tr(X^dagger X) = 1.  complex128 [n, 3, 3] (anti-Hermitian, traceless)."""
    rng = _rng(seed)
    x = rng.standard_normal((n, 8))
    X = 1j * np.einsum("na,aij->nij", x, _GM) * 0.5
    nrm = np.sqrt(np.einsum("nij,nij->n", np.conj(X), X).real)
    return X / nrm[:, None, None]


def expm_antihermitian(X: np.ndarray) -> np.ndarray:
    """exp(X) for anti-Hermitian X (X = i H with H = -i X Hermitian)."""
    return su3_exp_i_hermitian(-1j * X)
