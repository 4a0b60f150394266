"""The pseudofermion action as a torch autograd function whose backward is
the hand-written adjoint (lmx.PseudofermionAction).  The finite-difference
accuracy of the adjoint itself is pinned in test_gpu_configs.py (16^4, 128
link directions, 1e-8) and test_gpu_parity.py (against the oracle); here the
wiring: the autograd gradient is the adjoint's G bit for bit, the gradient
convention (re, im storage) and sign agree with central differences of S in
single real entries of U and phi, and S = |eta|^2 for phi = D^dagger eta."""
import numpy as np
import pytest
import torch

import lmx
from lmx import generators as gen

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("csw", [0.0, 1.0])
def test_autograd_is_the_adjoint(gpu, csw):
    L, kappa = (4, 4, 4, 4), 0.125
    print(f"autograd csw={csw}", flush=True)
    U = gen.links(L, "haar", 900)
    eta = gen.gaussian_spinor(L, 901)
    lat = lmx.Lattice(L, (1, 1, 1, -1))
    Ug = lat.gauge_from_numpy(U)
    D = lmx.WilsonDiracOperator4D(Ug, kappa, csw)
    phi = lat.spinor()
    lmx.mul_(phi, lmx.adjoint(D), lat.spinor_from_numpy(eta))
    links = Ug.data.clone().requires_grad_(True)
    ph = phi.data.clone().requires_grad_(True)
    S = lmx.PseudofermionAction.apply(links, ph, lat, kappa, csw, 1e-13, 20000)
    S.backward()
    assert abs(S.item() - np.vdot(eta, eta).real) <= 1e-11 * S.item()  # S = |eta|^2
    G = lmx.GaugeField(lat)
    res, X, _ = lmx.dSFdU(G, D, phi, rtol=1e-13, maxiter=20000)
    assert torch.equal(links.grad, G.data)
    assert torch.equal(ph.grad, 2.0 * X.data)

    def S_of(lk, pp):
        return lmx.PseudofermionAction.apply(lk, pp, lat, kappa, csw, 1e-13, 20000).item()

    rng = np.random.default_rng(902)
    eps = 1e-4
    worst = 0.0
    for _ in range(6):  # single real entries of U: (mu, element, site, re/im)
        idx = (int(rng.integers(4)), int(rng.integers(9)), int(rng.integers(lat.V)), int(rng.integers(2)))
        lp, lm = Ug.data.clone(), Ug.data.clone()
        lp[idx] += eps
        lm[idx] -= eps
        fd = (S_of(lp, phi.data) - S_of(lm, phi.data)) / (2 * eps)
        g = links.grad[idx].item()
        worst = max(worst, abs(fd - g) / max(abs(g), 1.0))
    for _ in range(4):  # single real entries of phi
        idx = (0, int(rng.integers(12)), int(rng.integers(lat.V)), int(rng.integers(2)))
        pp, pm = phi.data.clone(), phi.data.clone()
        pp[idx] += eps
        pm[idx] -= eps
        fd = (S_of(Ug.data, pp) - S_of(Ug.data, pm)) / (2 * eps)
        g = ph.grad[idx].item()
        worst = max(worst, abs(fd - g) / max(abs(g), 1.0))
    print(f"  worst |FD - autograd| / max(|autograd|, 1) = {worst:.2e}", flush=True)
    assert worst <= 1e-7  # measured 2e-9 (Wilson), 4e-9 (clover)


def test_autograd_composes(gpu):
    # the action inside a larger torch expression: the chain rule scales the adjoint
    L, kappa = (4, 4, 2, 4), 0.12
    U = gen.links(L, "weak", 910, eps=0.2)
    lat = lmx.Lattice(L, (1, 1, 1, -1))
    Ug = lat.gauge_from_numpy(U)
    phi = lat.spinor_from_numpy(gen.gaussian_spinor(L, 911))
    links = Ug.data.clone().requires_grad_(True)
    S = lmx.pseudofermion_action_ad(lmx.GaugeField(lat, links), phi, kappa)
    (3.0 * S + 1.0).backward()
    G = lmx.GaugeField(lat)
    lmx.dSFdU(G, lmx.WilsonDiracOperator4D(Ug, kappa), phi, rtol=1e-12, maxiter=20000)
    assert torch.equal(links.grad, 3.0 * G.data)
