"""NC = 2 fields through the HIP path (the reference's generic-NC kernels,
src/Operators/WilsonDiracOperator.jl:80-147,444-511,554-621,665-733;
src/Operators/WilsonDiracCloverOperator.jl:57-118,481-525;
ext/AD/WilsonDiracOperator.jl:27-53), against the oracle bit for bit and
against the dense restatement at the reference tests' tolerances, on the
reference's own inputs:
  - the closed-form SU(2) fields of test/wilson_clover.jl:12-45 (lattice
    2^4, kappa 0.12, cSW 1.17, phases (1, 1, 1, -1)): clover planes, D and
    D^dagger with and without the clover term, the adjoint identity,
    gamma5-Hermiticity, gauge covariance, the link pullback and its
    directional difference (:187-217);
  - the nw = 0 fixture of test/nw0.jl:213-254 (4 x 2 x 2 x 2, D, D^dagger,
    Donly, Donly^dagger on the reference's padded storage with nw = 0 and
    nw = 1, equal to each other and to the oracle);
  - the FP64 D^dagger D CG on NC = 2 fields (also allocation-free) against
    the oracle's CG: same iterations, identical solution.
"""
import ctypes

import numpy as np
import pytest
import torch

import dense_reference as dr
import oracle as orc
import lmx
import su2_fixtures as fx
from lmx import _lib

pytestmark = pytest.mark.gpu

L2 = (2, 2, 2, 2)
PHASES = (1, 1, 1, -1)
KAPPA, CSW = 0.12, 1.17


def close(a, b, tol):
    return np.max(np.abs(a - b)) <= tol * max(1.0, np.max(np.abs(b)))


@pytest.fixture(scope="module")
def su2(gpu):
    lat = lmx.Lattice(L2, PHASES, nc=2)
    U = fx.clover_test_links(L2)
    Ug = lat.gauge_from_numpy(U)
    assert np.array_equal(Ug.to_numpy(), U)
    return lat, U, Ug


def test_clover_planes_and_apply_su2(su2):
    lat, U, Ug = su2
    psi = fx.clover_test_spinor(L2)
    D = lmx.WilsonDiracOperator4D(Ug, KAPPA, CSW)
    F = D.clover_planes()
    Fo = orc.clover_field_strength_nc(U, L2)
    assert np.array_equal(F, Fo)
    assert close(F, dr.field_strength(U, L2), 2e-12)
    x = lat.spinor_from_numpy(psi)
    assert np.array_equal(x.to_numpy(), psi)
    y = lat.spinor()
    for dag in (False, True):
        lmx.mul_(y, lmx.adjoint(D) if dag else D, x)
        got = y.to_numpy()
        assert np.array_equal(got, orc.wilson_clover_nc(U, Fo, psi, L2, KAPPA, CSW, PHASES, dag)), f"dagger={dag}"
        assert close(got, dr.wilson_clover(U, psi, L2, KAPPA, CSW, PHASES, dag), 3e-12)


@pytest.mark.parametrize("flags", [0, 1, 2, 3])
def test_wilson_and_donly_su2(su2, flags):
    lat, U, Ug = su2
    psi = fx.clover_test_spinor(L2, offset=0.19)
    A = lmx.WilsonDiracOperator4D_Donly(Ug) if flags & 2 else lmx.WilsonDiracOperator4D(Ug, KAPPA)
    if flags & 1:
        A = lmx.adjoint(A)
    y = lat.spinor()
    lmx.mul_(y, A, lat.spinor_from_numpy(psi))
    ref = orc.wilson_nc(U, psi, L2, KAPPA, PHASES, dagger=bool(flags & 1), donly=bool(flags & 2))
    assert np.array_equal(y.to_numpy(), ref)
    assert close(y.to_numpy(), dr.wilson(U, psi, L2, KAPPA, PHASES, bool(flags & 1), bool(flags & 2)), 2e-12)


def test_adjoint_gamma5_and_gauge_covariance_su2(su2):
    lat, U, Ug = su2
    psi = fx.clover_test_spinor(L2)
    chi = fx.clover_test_spinor(L2, offset=0.19)
    D = lmx.WilsonDiracOperator4D(Ug, KAPPA, CSW)

    def apply(op, v):
        out = lat.spinor()
        lmx.mul_(out, op, lat.spinor_from_numpy(v))
        return out.to_numpy()

    lhs = np.vdot(chi, apply(D, psi))
    rhs = np.vdot(apply(lmx.adjoint(D), chi), psi)
    assert abs(lhs - rhs) <= 5e-11 + 5e-12 * abs(lhs)
    assert close(apply(lmx.adjoint(D), psi), dr.gamma5(apply(D, dr.gamma5(psi))), 5e-12)
    Om = fx.clover_test_transformations(L2)
    Ut, psit = dr.gauge_transform(U, psi, L2, Om)
    Dt = lmx.WilsonDiracOperator4D(lat.gauge_from_numpy(Ut), KAPPA, CSW)
    expect = np.einsum("vij,vsj->vsi", Om, apply(D, psi))
    assert close(apply(Dt, psit), expect, 8e-12)


def test_link_pullback_su2(su2):
    lat, U, Ug = su2
    psi = fx.clover_test_spinor(L2)
    chi = fx.clover_test_spinor(L2, offset=0.19)
    D = lmx.WilsonDiracOperator4D(Ug, KAPPA)
    dU = lmx.GaugeField(lat)
    lmx.wilson_pullback_(dU, D, lat.spinor_from_numpy(chi), lat.spinor_from_numpy(psi), -KAPPA)
    G = dU.to_numpy()
    assert np.array_equal(G, orc.wilson_pullback_nc(chi, psi, L2, -KAPPA, PHASES))
    # the reference's directional central difference (test/wilson_clover.jl:187-217, eps 1e-6)
    dirs = fx.clover_test_links(L2)
    eps = 1e-6
    f = lambda links: float(np.real(np.vdot(chi, orc.wilson_nc(links, psi, L2, KAPPA, PHASES))))
    fd = (f(U + eps * dirs) - f(U - eps * dirs)) / (2 * eps)
    an = float(np.real(np.vdot(G, dirs)))
    print(f"SU(2) pullback: directional {an:.15e} finite difference {fd:.15e}", flush=True)
    assert abs(an - fd) <= 3e-6 + 3e-7 * abs(fd)
    # the clover link pullback is built for NC = 3 only: a clear error
    Dc = lmx.WilsonDiracOperator4D(Ug, KAPPA, CSW)
    with pytest.raises(lmx.LmxError) as e:
        lmx.wilson_pullback_(lmx.GaugeField(lat), Dc, lat.spinor_from_numpy(chi), lat.spinor_from_numpy(psi), -KAPPA)
    assert e.value.code == _lib.LMX_ERR_UNSUPPORTED


@pytest.mark.parametrize("flags", [0, 1, 2, 3])
def test_nw0_fixture_padded_storage(gpu, flags):
    # test/nw0.jl:213-254: (4, 2, 2, 2), NC = 2, kappa 0.12; nw = 0 and nw = 1 storage agree
    L = (4, 2, 2, 2)
    lat = lmx.Lattice(L, PHASES, nc=2)
    U = fx.nw0_links(L)
    psi = fx.nw0_spinor(L)
    Ug = lat.gauge_from_numpy(U)
    A = lmx.WilsonDiracOperator4D_Donly(Ug) if flags & 2 else lmx.WilsonDiracOperator4D(Ug, 0.12)
    if flags & 1:
        A = lmx.adjoint(A)
    ref = orc.wilson_nc(U, psi, L, 0.12, PHASES, dagger=bool(flags & 1), donly=bool(flags & 2))
    assert close(ref, dr.wilson(U, psi, L, 0.12, PHASES, bool(flags & 1), bool(flags & 2)), 1e-14)
    results = []
    for nw in (0, 1):
        pad = dr.halo_fill(psi, L, nw, PHASES)
        psi_d = torch.from_numpy(np.ascontiguousarray(pad)).to(lat.device)
        C_d = torch.zeros_like(psi_d)
        lmx.mul_padded_(C_d, A, psi_d, nw)
        core = C_d.cpu().numpy()
        if nw:
            core = core[1:-1, 1:-1, 1:-1, 1:-1]
        results.append(np.ascontiguousarray(core).reshape(-1, 4, 2))
    assert np.array_equal(results[0], results[1])
    assert np.array_equal(results[0], ref)


@pytest.mark.parametrize("csw", [0.0, 1.0])
def test_cg_nc2_matches_oracle(gpu, csw):
    L = (4, 4, 4, 8)
    V = int(np.prod(L))
    print(f"NC=2 D^dagger D CG on {L}, Haar SU(2) links, csw {csw}", flush=True)
    U = fx.random_su2(4 * V, 77).reshape(4, V, 2, 2)
    rng = np.random.default_rng(78)
    b = (rng.standard_normal((V, 4, 2)) + 1j * rng.standard_normal((V, 4, 2))) * np.sqrt(0.5)
    lat = lmx.Lattice(L, PHASES, nc=2)
    D = lmx.WilsonDiracOperator4D(lat.gauge_from_numpy(U), 0.124, csw)
    bg = lat.spinor_from_numpy(b)
    x, r, p, Ap, tmp = (lat.spinor() for _ in range(5))
    res = lmx.cg_(x, lmx.DdagDOp(D, tmp), bg, r, p, Ap, rtol=1e-10)
    kind = orc.OP_DDAGD_CLOVER if csw else orc.OP_DDAGD_WILSON
    F = orc.clover_field_strength_nc(U, L) if csw else None
    op = orc.NormalOp(kind, L, U=U, F=F, kappa=0.124, csw=csw, phases=PHASES, nc=2)
    xo, ro = orc.cg(op, b, rtol=1e-10, order=1)
    print(f"  iterations {res.iterations} (oracle {ro.iterations}), relative residual {res.relative_residual:.3e}",
          flush=True)
    assert res.converged and ro.converged and res.iterations == ro.iterations
    assert np.array_equal(x.to_numpy(), xo)
    # the allocation-free entry point on the same NC = 2 problem
    x2, r2, p2, Ap2, t2 = (lat.spinor() for _ in range(5))
    params = _lib.default_params(rtol=1e-10)
    rc = _lib.CGResultC()
    _lib.check(_lib.lib().lmx_cg_ddagd_noalloc(D._op, x2.ptr, bg.ptr, r2.ptr, p2.ptr, Ap2.ptr, t2.ptr,
                                                ctypes.byref(params), ctypes.byref(rc)))
    assert rc.iterations == res.iterations and np.array_equal(x2.to_numpy(), xo)


def test_nc2_unsupported_paths_raise(su2):
    lat, U, Ug = su2
    D = lmx.WilsonDiracOperator4D(Ug, KAPPA)
    x = lat.spinor(2)
    with pytest.raises(lmx.LmxError) as e:
        lmx.cg_batched_(lat.spinor(2), D, x, precision="h16")
    assert e.value.code == _lib.LMX_ERR_UNSUPPORTED
    with pytest.raises(lmx.LmxError) as e:
        lmx.autotune(D, "h16", 2)
    assert e.value.code == _lib.LMX_ERR_UNSUPPORTED
