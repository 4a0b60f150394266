"""The reference's own deterministic inputs through the HIP path.

test/wilson_dirac_ad.jl:13-18 (_wilson_dirac_ad_values) builds links and
spinors from closed-form ramps on a 2x2x2x2 lattice with nw = 1, kappa 0.117
and phases (1, 1, 1, -1) (:25-51), applies D, and checks the mul! reverse
rule: dpsi = D^dagger left (:80-84) and the link cotangent against central
finite differences along fixed link directions (:86-104, eps 1e-6,
rtol 2e-7).  Here the same inputs go through the C ABI on the reference's
own padded storage and the library layout: D, D^dagger and the hopping-only
pair bit for bit against the oracle, the reverse rule's dpsi and link
cotangent bit for bit against the oracle, and the link cotangent's
directional derivative against the reference's finite difference.
"""
import numpy as np
import pytest
import torch

import oracle as orc
import lmx

pytestmark = pytest.mark.gpu

GS = (2, 2, 2, 2)
KAPPA = 0.117
PHASES = (1.0, 1.0, 1.0, -1.0)


def ad_values(n1, n2, global_size, offset):
    """_wilson_dirac_ad_values(Val(N1), Val(N2), global_size, offset),
    test/wilson_dirac_ad.jl:13-18, as a column-major array [n1, n2, x, y, z, t]."""
    count = n1 * n2 * int(np.prod(global_size))
    values = np.arange(1, count + 1, dtype=np.float64).reshape((n1, n2) + tuple(global_size), order="F")
    scale = count + offset
    return (values + offset) / scale + 1j * (values[:, ::-1] / (2 * scale))


def as_spinor(a):
    """[nc, 4, x, y, z, t] column-major -> [V, 4, nc] (site x-fastest, spin, colour)."""
    nc = a.shape[0]
    return np.ascontiguousarray(np.transpose(a, (5, 4, 3, 2, 1, 0))).reshape(-1, 4, nc)


def as_link(a):
    """[nc, nc, x, y, z, t] -> [V, nc, nc] (site, row, col)."""
    nc = a.shape[0]
    return np.ascontiguousarray(np.transpose(a, (5, 4, 3, 2, 0, 1))).reshape(-1, nc, nc)


def padded(core, nw, phases):
    """The core spinor [V, 4, 3] as LatticeMatrix.A, C order [t, z, y, x, 4, 3] with
    the halo filled as set_halo! does (the dense restatement's halo fill)."""
    import dense_reference as dr
    return dr.halo_fill(core, GS, nw, phases)


@pytest.fixture(scope="module")
def inputs(gpu):
    U = np.stack([as_link(ad_values(3, 3, GS, 7 * mu)) for mu in range(1, 5)])
    dirs = np.stack([as_link(ad_values(3, 3, GS, 11 * mu + 3)) for mu in range(1, 5)])
    psi = as_spinor(ad_values(3, 4, GS, 5))
    left = as_spinor(ad_values(3, 4, GS, 13))
    return U, dirs, psi, left


def test_ramp_inputs_apply_bit_exact(inputs):
    U, dirs, psi, left = inputs
    lat = lmx.Lattice(GS, PHASES)
    Ug = lat.gauge_from_numpy(U)
    D = lmx.WilsonDiracOperator4D(Ug, KAPPA)
    x = lat.spinor_from_numpy(psi)
    y = lat.spinor()
    for flags in (0, 1, 2, 3):
        A = lmx.WilsonDiracOperator4D_Donly(Ug) if flags & 2 else D
        if flags & 1:
            A = lmx.adjoint(A)
        lmx.mul_(y, A, x)
        ref = orc.wilson(U, psi, GS, KAPPA, PHASES, dagger=bool(flags & 1), donly=bool(flags & 2))
        assert np.array_equal(y.to_numpy(), ref), f"flags={flags}"
        # the same on the reference's padded storage (nw = 1), halo filled as set_halo!
        pad = padded(psi, 1, PHASES)
        psi_d = torch.from_numpy(np.ascontiguousarray(pad)).to(lat.device)
        C_d = torch.zeros_like(psi_d)
        lmx.mul_padded_(C_d, A, psi_d, 1)
        got = C_d.cpu().numpy()[1:3, 1:3, 1:3, 1:3].reshape(-1, 4, 3)
        assert np.array_equal(got, ref), f"padded flags={flags}"


def test_ramp_inputs_reverse_rule(inputs):
    # loss = real(dot(left, D psi)) (:1-4): dresult = left, dpsi = D^dagger left,
    # dlinks = the pullback of dresult and psi with coefficient -kappa
    U, dirs, psi, left = inputs
    lat = lmx.Lattice(GS, PHASES)
    Ug = lat.gauge_from_numpy(U)
    D = lmx.WilsonDiracOperator4D(Ug, KAPPA)
    lg = lat.spinor_from_numpy(left)
    dpsi = lat.spinor()
    lmx.mul_(dpsi, lmx.adjoint(D), lg)
    assert np.array_equal(dpsi.to_numpy(), orc.wilson(U, left, GS, KAPPA, PHASES, dagger=True))
    dU = lmx.GaugeField(lat)
    lmx.wilson_pullback_(dU, D, lg, lat.spinor_from_numpy(psi), -KAPPA)
    G = dU.to_numpy()
    assert np.array_equal(G, orc.wilson_pullback(left, psi, GS, -KAPPA, PHASES))
    # directional derivative against the reference's central difference (eps 1e-6)
    eps = 1e-6

    def loss(links):
        return float(np.real(np.vdot(left, orc.wilson(links, psi, GS, KAPPA, PHASES))))

    fd = (loss(U + eps * dirs) - loss(U - eps * dirs)) / (2 * eps)
    directional = float(np.real(np.vdot(G, dirs)))
    print(f"ramp inputs: directional {directional:.15e} finite difference {fd:.15e}", flush=True)
    assert abs(directional - fd) <= 2e-6 + 2e-7 * abs(fd)
