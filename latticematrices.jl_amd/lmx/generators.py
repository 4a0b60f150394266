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


def su3_exp_i_traceless(Q: np.ndarray) -> np.ndarray:
    """exp(i Q) for traceless Hermitian Q[..., 3, 3] in closed form:
    exp(iQ) = f0 + f1 Q + f2 Q^2 with the Cayley-Hamilton coefficients of
    Morningstar & Peardon, Phys. Rev. D 69, 054501 (2004), eqs. (23)-(34)."""
    Q2 = Q @ Q
    c0 = (np.einsum("...ij,...ji->...", Q2, Q).real) / 3.0          # det Q
    c1 = 0.5 * np.einsum("...ii->...", Q2).real                       # tr Q^2 / 2
    neg = c0 < 0
    c0a = np.abs(c0)
    c0max = 2.0 * (c1 / 3.0) ** 1.5
    ratio = np.where(c0max > 0, c0a / np.where(c0max > 0, c0max, 1.0), 0.0)
    theta = np.arccos(np.clip(ratio, -1.0, 1.0))
    u = np.sqrt(c1 / 3.0) * np.cos(theta / 3.0)
    w = np.sqrt(c1) * np.sin(theta / 3.0)
    w2 = w * w
    small = np.abs(w) < 0.05
    xi0 = np.where(small, 1 - w2 / 6 * (1 - w2 / 20 * (1 - w2 / 42)), np.sin(w) / np.where(small, 1.0, w))
    e2 = np.exp(2j * u)
    em = np.exp(-1j * u)
    cw = np.cos(w)
    h0 = (u * u - w2) * e2 + em * (8 * u * u * cw + 2j * u * (3 * u * u + w2) * xi0)
    h1 = 2 * u * e2 - em * (2 * u * cw - 1j * (3 * u * u - w2) * xi0)
    h2 = e2 - em * (cw + 3j * u * xi0)
    den = 9 * u * u - w2
    tiny = den < 1e-12
    den = np.where(tiny, 1.0, den)
    f0, f1, f2 = h0 / den, h1 / den, h2 / den
    # c0 < 0: f_j(-c0) = (-1)^j conj(f_j(c0))
    f0 = np.where(neg, np.conj(f0), f0)
    f1 = np.where(neg, -np.conj(f1), f1)
    f2 = np.where(neg, np.conj(f2), f2)
    # Q -> 0: second-order Taylor (|Q| below 1e-6 only)
    f0 = np.where(tiny, 1.0 - c1, f0)
    f1 = np.where(tiny, 1.0 + 0j, f1)
    f2 = np.where(tiny, -0.5 + 0j, f2)
    I = np.eye(3, dtype=np.complex128)
    return f0[..., None, None] * I + f1[..., None, None] * Q + f2[..., None, None] * Q2


def weak_field_su3(n: int, eps: float, seed: int, chunk: int = 1 << 20) -> np.ndarray:
    """n weak-field SU(3) links U = exp(i eps H), H = sum_a h_a lambda_a / 2
    with h_a ~ N(0,1) (traceless Hermitian Gaussian).  complex128 [n, 3, 3]."""
    rng = _rng(seed)
    out = np.empty((n, 3, 3), dtype=np.complex128)
    for s in range(0, n, chunk):
        m = min(chunk, n - s)
        h = rng.standard_normal((m, 8))
        H = np.einsum("na,aij->nij", h, _GM) * 0.5
        out[s:s + m] = su3_exp_i_traceless(eps * H)
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
