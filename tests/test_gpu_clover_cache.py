"""mul_cached_clover! / mul_cached_clover_adjoint! with explicit links, and
their reverse rule, after test/wilson_clover_ad.jl:94-199 ("cached
Wilson-clover explicit-link pullback"): the cached apply equals mul! of an
operator built on the same links (here bit for bit: the same kernels and the
same clover build), the cache is refreshed exactly when the links object or
its contents change, and the autograd gradients are the hand-written
pullbacks, with the link gradient checked against central differences."""
import numpy as np
import pytest
import torch

import lmx
from lmx import generators as gen

pytestmark = pytest.mark.gpu

L, KAPPA, CSW = (4, 4, 4, 8), 0.125, 1.3


def _fresh(lat, U, kappa=KAPPA, csw=CSW):
    return lmx.WilsonDiracCloverOperator4D(lmx.GaugeField(lat, U.data.clone()), kappa, csw)


def _apply(D, psi, dagger=False):
    out = psi.similar()
    lmx.mul_(out, lmx.adjoint(D) if dagger else D, psi)
    return out.data


def test_cached_apply_refreshes_on_link_changes(gpu):
    lat = lmx.Lattice(L, (1, 1, 1, -1))
    U1 = lat.gauge_from_numpy(gen.links(L, "haar", 910))
    psi = lat.spinor_from_numpy(gen.gaussian_spinor(L, 911))
    op = lmx.WilsonDiracCloverOperator4D(U1, KAPPA, CSW)
    res = lat.spinor()
    e0 = op.cache_epoch
    lmx.mul_cached_clover_(res, op, U1, psi)
    assert op.cache_epoch == e0  # the links the cache was built from: no refresh
    assert torch.equal(res.data, _apply(op, psi))

    U2 = lat.gauge_from_numpy(gen.links(L, "haar", 912))  # another links object
    lmx.mul_cached_clover_(res, op, U2, psi)
    e1 = op.cache_epoch
    assert e1 == e0 + 1
    assert torch.equal(res.data, _apply(_fresh(lat, U2), psi))
    assert not torch.equal(res.data, _apply(_fresh(lat, U1), psi))
    lmx.mul_cached_clover_(res, op, U2, psi)
    assert op.cache_epoch == e1

    # an in-place torch write to the links (add_matrix!-like) is seen
    d = torch.from_numpy(gen.random_su3_algebra(4 * lat.V, 913).reshape(4, lat.V, 9)).to(gpu)
    d = torch.view_as_real(d.to(torch.complex128)).permute(0, 2, 1, 3).contiguous()
    U2.data.add_(1e-3 * d)
    lmx.mul_cached_clover_(res, op, U2, psi)
    e2 = op.cache_epoch
    assert e2 == e1 + 1
    assert torch.equal(res.data, _apply(_fresh(lat, U2), psi))

    # a write torch does not count needs mark_halo_dirty! (as a write through
    # LatticeMatrix.A does in the reference)
    U2.data.data.mul_(-1.0)
    lmx.mul_cached_clover_(res, op, U2, psi)
    assert op.cache_epoch == e2
    U2.mark_halo_dirty_()
    lmx.mul_cached_clover_(res, op, U2, psi)
    assert op.cache_epoch == e2 + 1
    assert torch.equal(res.data, _apply(_fresh(lat, U2), psi))

    # the adjoint entry point with a third links object
    U3 = lat.gauge_from_numpy(gen.links(L, "haar", 914))
    op2 = lmx.WilsonDiracCloverOperator4D(U1, KAPPA, CSW)
    lmx.mul_cached_clover_adjoint_(res, op2, U3, psi)
    assert op2.cache_epoch == 1
    assert torch.equal(res.data, _apply(_fresh(lat, U3), psi, dagger=True))

    with pytest.raises(TypeError):
        lmx.mul_cached_clover_(res, op, U3.data, psi)
    with pytest.raises(lmx.ArgumentError):
        other = lmx.Lattice((8, 4, 4, 4), (1, 1, 1, -1))
        lmx.mul_cached_clover_(res, op, lmx.GaugeField(other), psi)


def test_cached_apply_autograd_is_the_pullback(gpu):
    lat = lmx.Lattice(L, (1, 1, 1, -1))
    U = lat.gauge_from_numpy(gen.links(L, "haar", 920))
    psi = lat.spinor_from_numpy(gen.gaussian_spinor(L, 921, nrhs=2))
    left = lat.spinor_from_numpy(gen.gaussian_spinor(L, 922, nrhs=2))
    op = lmx.WilsonDiracCloverOperator4D(lat.gauge_from_numpy(gen.links(L, "haar", 923)), KAPPA, CSW)
    Ug = lmx.GaugeField(lat, U.data.clone().requires_grad_(True))
    pg = lmx.SpinorField(lat, 2, psi.data.clone().requires_grad_(True))
    out = lmx.mul_cached_clover_ad(op, Ug, pg)
    loss = (left.data * out).sum()  # Re <left, D psi> summed over the two fields
    loss.backward()

    ref = _fresh(lat, U)
    assert torch.equal(out.detach(), _apply(ref, psi))
    dU = lmx.GaugeField(lat)
    for i in range(2):
        lmx.wilson_pullback_(dU, ref, left.rhs(i), psi.rhs(i), -KAPPA)
    assert torch.equal(Ug.data.grad, dU.data)
    assert torch.equal(pg.data.grad, _apply(ref, left, dagger=True))

    def loss_of(links):
        G = lmx.GaugeField(lat, links)
        res = psi.similar()
        lmx.mul_cached_clover_(res, op, G, psi)  # a new links object each time: refreshed
        return float((left.data * res.data).sum())

    rng = np.random.default_rng(924)
    eps, worst = 1e-4, 0.0
    for _ in range(8):  # single real entries (mu, element, site, re/im), clover and hopping terms alike
        idx = (int(rng.integers(4)), int(rng.integers(9)), int(rng.integers(lat.V)), int(rng.integers(2)))
        lp, lm = U.data.clone(), U.data.clone()
        lp[idx] += eps
        lm[idx] -= eps
        fd = (loss_of(lp) - loss_of(lm)) / (2 * eps)
        g = Ug.data.grad[idx].item()
        worst = max(worst, abs(fd - g) / max(abs(g), 1.0))
    print(f"cached clover autograd: worst |FD - grad| / max(|grad|, 1) = {worst:.2e}", flush=True)
    assert worst < 1e-6


@pytest.mark.parametrize("csw", [0.0, CSW])
def test_operator_apply_autograd(gpu, csw):
    # the mul! reverse rules of the Wilson and Wilson-clover operators
    lat = lmx.Lattice(L, (1, 1, 1, -1))
    U = lat.gauge_from_numpy(gen.links(L, "haar", 930))
    psi = lat.spinor_from_numpy(gen.gaussian_spinor(L, 931))
    left = lat.spinor_from_numpy(gen.gaussian_spinor(L, 932))
    links = U.data.clone().requires_grad_(True)
    D = lmx.WilsonDiracOperator4D(lmx.GaugeField(lat, links), KAPPA, csw)
    pg = lmx.SpinorField(lat, 1, psi.data.clone().requires_grad_(True))
    out = lmx.wilson_apply_ad(D, pg)
    (left.data * out).sum().backward()
    ref = _fresh(lat, U, csw=csw)
    assert torch.equal(out.detach(), _apply(ref, psi))
    dU = lmx.GaugeField(lat)
    lmx.wilson_pullback_(dU, ref, left, psi, -KAPPA)
    assert torch.equal(links.grad, dU.data)
    assert torch.equal(pg.data.grad, _apply(ref, left, dagger=True))


def test_clover_field_strength(gpu):
    # CloverFieldStrength4D(U) and update_clover!(F, U): the oracle's planes
    # bit for bit (as the operator's cache, test_gpu_parity), and refreshed
    import oracle as orc
    Lf = (4, 4, 2, 4)
    lat = lmx.Lattice(Lf, (1, 1, 1, -1))
    U1, U2 = gen.links(Lf, "weak", 940, eps=0.15), gen.links(Lf, "haar", 941)
    F = lmx.CloverFieldStrength4D(lat.gauge_from_numpy(U1))
    assert len(F) == 6
    for U in (U1, U2):
        if U is U2:
            assert lmx.update_clover_(F, lat.gauge_from_numpy(U2)) is F
        want = orc.clover_field_strength(U, Lf)
        got = np.stack([F[i].cpu().numpy() for i in range(6)])
        assert np.array_equal(got, want)
        assert torch.allclose(F[3], -F[3].conj().transpose(-1, -2), rtol=0, atol=1e-15)  # anti-Hermitian
    with pytest.raises(TypeError):
        lmx.update_clover_(F, U2)
    with pytest.raises(TypeError):
        lmx.update_clover_(object(), None)
