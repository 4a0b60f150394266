"""GPU parity of the HIP library (through the C ABI) against the CPU oracle.

FP64 bar: relative error <= 1e-13 (the FP64 tolerance of the path); the FP64
stencil, clover, BLAS-1 and CG kernels evaluate the reference's operations in
the reference's order, so they are also checked for bit equality."""
import numpy as np
import pytest
import torch

import dense_reference as dr
import oracle as orc
import lmx
from lmx import generators as gen

pytestmark = pytest.mark.gpu

TOL = 1e-13


def rel(a, b):
    return float(np.max(np.abs(a - b)) / max(np.max(np.abs(b)), 1e-300))


def rand_c(rng, shape):
    return rng.random(shape) + 1j * rng.random(shape)


CASES = [
    ((2, 2, 2, 2), "rand", (1, 1, 1, -1)),
    ((4, 2, 3, 2), "rand", (1, 1, 1, -1)),
    ((4, 4, 4, 4), "haar", (1, 1, 1, -1)),
    ((3, 2, 2, 4), "rand", (np.exp(0.3j), 1, -1j, -1)),
    ((8, 4, 6, 2), "haar", (-1, 1, 1, -1)),
    # 3-D volume a multiple of 256: the t-streaming kernel (auto shape
    # TL = 2 t-slices per workgroup, 4 z-ranges at 16^3 x 32)
    ((8, 8, 4, 8), "rand", (-1, 1, 1, -1)),
    ((16, 16, 16, 32), "haar", (1, 1, 1, -1)),
]


def make_links(L, kind, seed):
    V = int(np.prod(L))
    if kind == "rand":
        return rand_c(np.random.default_rng(seed), (4, V, 3, 3))
    return gen.links(L, kind, seed)


@pytest.mark.parametrize("L,kind,phases", CASES)
@pytest.mark.parametrize("flags", [0, 1, 2, 3])
def test_wilson_apply_matches_oracle(gpu, L, kind, phases, flags):
    print(f"wilson parity L={L} kind={kind} flags={flags}", flush=True)
    U = make_links(L, kind, 1)
    psi = rand_c(np.random.default_rng(2), (int(np.prod(L)), 4, 3))
    lat = lmx.Lattice(L, phases)
    Ug = lat.gauge_from_numpy(U)
    D = lmx.WilsonDiracOperator4D(Ug, 0.117)
    A = lmx.WilsonDiracOperator4D_Donly(Ug) if flags & 2 else D
    if flags & 1:
        A = lmx.adjoint(A)
    x = lat.spinor_from_numpy(psi)
    y = lat.spinor()
    lmx.mul_(y, A, x)
    got = y.to_numpy()
    ref = orc.wilson(U, psi, L, 0.117, phases, dagger=bool(flags & 1), donly=bool(flags & 2))
    assert rel(got, ref) <= TOL
    assert np.array_equal(got, ref), f"not bit-exact: max rel {rel(got, ref):.3e}"


def test_layout_roundtrip_with_halo(gpu):
    L = (4, 3, 2, 4)
    phases = (1, 1j, 1, -1)
    V = int(np.prod(L))
    psi = gen.gaussian_spinor(L, 3)
    lat = lmx.Lattice(L, phases)
    x = lat.spinor_from_numpy(psi)
    pad = x.to_padded(1).reshape(L[3] + 2, L[2] + 2, L[1] + 2, L[0] + 2, 4, 3)
    core = dr.grid(L, psi)
    # expected padded array: periodic extension with phase^wraps per direction
    idx = [np.arange(-1, L[d] + 1) for d in range(4)]
    exp = np.empty_like(pad)
    for it, t in enumerate(idx[3]):
        for iz, z in enumerate(idx[2]):
            for iy, y in enumerate(idx[1]):
                for ix, xx in enumerate(idx[0]):
                    c = (xx % L[0], y % L[1], z % L[2], t % L[3])
                    w = [(v - (v % L[d])) // L[d] for d, v in enumerate((xx, y, z, t))]
                    f = np.prod([complex(phases[d]) ** w[d] for d in range(4)])
                    exp[it, iz, iy, ix] = core[c[3], c[2], c[1], c[0]] * f
    assert np.array_equal(pad, exp)
    # import from the padded array ignores the halo
    y = lat.spinor_from_numpy(pad, nw=1)
    assert np.array_equal(y.to_numpy(), psi)
    # links round trip
    U = gen.links(L, "haar", 4)
    assert np.array_equal(lat.gauge_from_numpy(U).to_numpy(), U)


def test_blas_matches_oracle(gpu):
    L = (8, 4, 4, 6)
    a = gen.gaussian_spinor(L, 5)
    b = gen.gaussian_spinor(L, 6)
    lat = lmx.Lattice(L, (1, 1, 1, -1))
    A, B = lat.spinor_from_numpy(a), lat.spinor_from_numpy(b)
    d = lmx.dot(A, B)
    ref = orc.dot(a, b, order=1)
    assert d == ref
    lmx.axpby_(0.37, A, -1.25, B)
    assert np.array_equal(B.to_numpy(), 0.37 * a + (-1.25) * b)
    # add_matrix!(C, A, alpha): u += alpha * v per entry, every product and sum
    # rounded separately (numpy's elementwise float64 operations)
    B = lat.spinor_from_numpy(b)
    lmx.add_matrix_(B, A, 0.75)
    assert np.array_equal(B.to_numpy(), b + 0.75 * a)
    B = lat.spinor_from_numpy(b)
    lmx.add_matrix_(B, A, 0.3 - 1.1j)
    ar, ai = 0.3, -1.1
    pr = ar * a.real - ai * a.imag
    pi = ar * a.imag + ai * a.real
    assert np.array_equal(B.to_numpy(), (b.real + pr) + 1j * (b.imag + pi))


def test_clover_matches_oracle(gpu):
    L = (4, 4, 2, 4)
    U = gen.links(L, "weak", 7, eps=0.15)
    psi = gen.gaussian_spinor(L, 8)
    lat = lmx.Lattice(L, (1, 1, 1, -1))
    Ug = lat.gauge_from_numpy(U)
    D = lmx.WilsonDiracCloverOperator4D(Ug, 0.1235, 1.0)
    F = orc.clover_field_strength(U, L)
    Fg = D.clover_planes()
    assert rel(Fg, F) <= TOL
    assert np.array_equal(Fg, F)
    x = lat.spinor_from_numpy(psi)
    y = lat.spinor()
    for dag in (False, True):
        lmx.mul_(y, lmx.adjoint(D) if dag else D, x)
        ref = orc.wilson_clover(U, F, psi, L, 0.1235, 1.0, dagger=dag)
        got = y.to_numpy()
        assert rel(got, ref) <= TOL
        assert np.array_equal(got, ref)


def test_cg_config1_matches_oracle(gpu):
    # 4^4 Haar, kappa 0.125, (1,1,1,-1), rtol 1e-10 from zero
    L = (4, 4, 4, 4)
    U = gen.links(L, "haar", 1234)
    b = gen.gaussian_spinor(L, 1235)
    op = orc.NormalOp(orc.OP_DDAGD_WILSON, L, U=U, kappa=0.125)
    xo, ro = orc.cg(op, b, rtol=1e-10, order=1)
    lat = lmx.Lattice(L, (1, 1, 1, -1))
    D = lmx.WilsonDiracOperator4D(lat.gauge_from_numpy(U), 0.125)
    bg = lat.spinor_from_numpy(b)
    x, r, p, Ap, tmp = (lat.spinor() for _ in range(5))
    res = lmx.cg_(x, lmx.DdagDOp(D, tmp), bg, r, p, Ap, rtol=1e-10)
    assert res.converged and ro.converged
    assert res.iterations == ro.iterations
    xg = x.to_numpy()
    assert rel(xg, xo) <= TOL
    assert np.array_equal(xg, xo)
    assert res.residual_norm == ro.residual_norm
    true_res = np.linalg.norm(op.apply(xg) - b) / np.linalg.norm(b)
    assert true_res <= 1e-10 * 1.01


def test_cg_known_answers_generic(gpu):
    # test/cg.jl:29-64,76-95 through the generic-operator CG
    L = (1, 1, 1, 4)
    b = (np.arange(1, 49, dtype=np.float64) + 0j).reshape(4, 4, 3)
    lat = lmx.Lattice(L)
    bg = lat.spinor_from_numpy(b)
    x, r, p, Ap = (lat.spinor() for _ in range(4))
    A = lmx.ScaledOperator(4.0)
    res = lmx.cg_(x, A, bg, r, p, Ap, rtol=1e-13, maxiter=10)
    assert res.converged and res.iterations == 1 and res.reason == "converged"
    assert rel(x.to_numpy(), b / 4) <= 2e-13
    lmx.clear_matrix_(x)
    zb = lat.spinor()
    res = lmx.cg_(x, A, zb, r, p, Ap, rtol=1e-13, maxiter=10)
    assert res.converged and res.iterations == 0 and res.residual_norm == 0 and res.relative_residual == 0
    lmx.clear_matrix_(x)
    res = lmx.cg_(x, A, bg, r, p, Ap, maxiter=0)
    assert not res.converged and res.reason == "maximum_iterations"
    lmx.clear_matrix_(x)
    res = lmx.cg_(x, lmx.ScaledOperator(-1.0), bg, r, p, Ap, maxiter=10)
    assert not res.converged and res.reason == "nonpositive_curvature"
    with pytest.raises(lmx.ArgumentError):
        lmx.cg_(x, A, bg, x, p, Ap)
    with pytest.raises(lmx.ArgumentError):
        lmx.cg_(x, A, bg, r, p, Ap, rtol=-1)
    with pytest.raises(lmx.ArgumentError):
        lmx.cg_(x, A, bg, r, p, Ap, maxiter=-1)
    # DdagDOp((2+i) I) = 5 I
    tmp, out = lat.spinor(), lat.spinor()
    N = lmx.DdagDOp(lmx.ScaledOperator(2 + 1j), tmp)
    lmx.mul_(out, N, bg)
    assert rel(out.to_numpy(), 5 * b) <= 1e-15
    lmx.clear_matrix_(x)
    res = lmx.solve_(x, N, bg, r, p, Ap, rtol=1e-13, maxiter=10)
    assert res.converged and rel(x.to_numpy(), b / 5) <= 2e-13
    with pytest.raises(lmx.ArgumentError):
        lmx.mul_(out, lmx.DdagDOp(lmx.ScaledOperator(2 + 1j), bg), bg)


def test_force_matches_oracle(gpu):
    L = (4, 4, 4, 4)
    U = gen.links(L, "haar", 51)
    eta = gen.gaussian_spinor(L, 52)
    phi = orc.wilson(U, eta, L, 0.125, dagger=True)
    Go, Xo, chio, ro = orc.dsfdu(U, phi, L, 0.125, rtol=1e-10, order=1)
    Fo = orc.ta_force(U, Go)
    lat = lmx.Lattice(L, (1, 1, 1, -1))
    Ug = lat.gauge_from_numpy(U)
    D = lmx.WilsonDiracOperator4D(Ug, 0.125)
    G, F = lmx.GaugeField(lat), lmx.GaugeField(lat)
    res, eta_g, chi_g = lmx.dSFdU(G, D, lat.spinor_from_numpy(phi), force=F, rtol=1e-10)
    assert res.converged and res.iterations == ro.iterations
    assert rel(eta_g.to_numpy(), Xo) <= TOL
    assert rel(G.to_numpy(), Go) <= TOL
    assert rel(F.to_numpy(), Fo) <= TOL
    assert np.array_equal(G.to_numpy(), Go)


def test_pullback_matches_oracle(gpu):
    L = (4, 2, 4, 4)
    U = gen.links(L, "haar", 71)
    psi = gen.gaussian_spinor(L, 72)
    w = gen.gaussian_spinor(L, 73)
    lat = lmx.Lattice(L, (1, 1, 1, -1))
    Ug = lat.gauge_from_numpy(U)
    D = lmx.WilsonDiracOperator4D(Ug, 0.125)
    dU = lmx.GaugeField(lat)
    lmx.wilson_pullback_(dU, D, lat.spinor_from_numpy(w), lat.spinor_from_numpy(psi), -0.125)
    ref = orc.wilson_pullback(w, psi, L, -0.125)
    assert np.array_equal(dU.to_numpy(), ref)


# Forced streaming shapes (t-slices per workgroup TL, 3-D sub-ranges ZS) on
# small lattices: every shape must give the same bits as the oracle.
STREAM_SHAPES = [((8, 8, 4, 16), 4, 1), ((16, 16, 8, 8), 8, 2), ((16, 8, 8, 12), 3, 4), ((8, 8, 8, 4), 4, 2)]


@pytest.mark.parametrize("L,TL,ZS", STREAM_SHAPES)
def test_stream_shapes_bit_exact(gpu, monkeypatch, L, TL, ZS):
    print(f"stream shape L={L} TL={TL} ZS={ZS}", flush=True)
    monkeypatch.setenv("LMX_STREAM_TL", str(TL))
    monkeypatch.setenv("LMX_STREAM_ZS", str(ZS))
    phases = (1, -1, 1, -1)
    U = gen.links(L, "haar", 21)
    psi = gen.gaussian_spinor(L, 22)
    lat = lmx.Lattice(L, phases)
    Ug = lat.gauge_from_numpy(U)
    D = lmx.WilsonDiracOperator4D(Ug, 0.121)
    x = lat.spinor_from_numpy(psi)
    y = lat.spinor()
    for flags in (0, 1, 2, 3):
        A = lmx.WilsonDiracOperator4D_Donly(Ug) if flags & 2 else D
        if flags & 1:
            A = lmx.adjoint(A)
        lmx.mul_(y, A, x)
        ref = orc.wilson(U, psi, L, 0.121, phases, dagger=bool(flags & 1), donly=bool(flags & 2))
        assert np.array_equal(y.to_numpy(), ref), f"flags={flags} max rel {rel(y.to_numpy(), ref):.3e}"
    # clover on the same shape
    Dc = lmx.WilsonDiracCloverOperator4D(Ug, 0.121, 1.3)
    F = orc.clover_field_strength(U, L)
    for dag in (False, True):
        lmx.mul_(y, lmx.adjoint(Dc) if dag else Dc, x)
        ref = orc.wilson_clover(U, F, psi, L, 0.121, 1.3, dagger=dag, phases=phases)
        assert np.array_equal(y.to_numpy(), ref), f"clover dag={dag}"


def test_cg_stream_path_bit_exact(gpu, monkeypatch):
    # CG with the fused p.Ap epilogue in the streaming kernel (block partials
    # indexed by the 256 lexicographic sites each step covers)
    monkeypatch.setenv("LMX_STREAM_TL", "4")
    monkeypatch.setenv("LMX_STREAM_ZS", "2")
    L = (16, 8, 4, 8)
    U = gen.links(L, "haar", 61)
    b = gen.gaussian_spinor(L, 62)
    for csw in (0.0, 1.0):
        F = orc.clover_field_strength(U, L) if csw else None
        op = orc.NormalOp(orc.OP_DDAGD_CLOVER if csw else orc.OP_DDAGD_WILSON, L, U=U, F=F, kappa=0.12, csw=csw)
        xo, ro = orc.cg(op, b, rtol=1e-10, order=1)
        lat = lmx.Lattice(L, (1, 1, 1, -1))
        D = lmx.WilsonDiracOperator4D(lat.gauge_from_numpy(U), 0.12, csw)
        bg = lat.spinor_from_numpy(b)
        x, r, p, Ap, tmp = (lat.spinor() for _ in range(5))
        res = lmx.cg_(x, lmx.DdagDOp(D, tmp), bg, r, p, Ap, rtol=1e-10)
        print(f"  csw={csw} iterations gpu={res.iterations} oracle={ro.iterations}", flush=True)
        assert res.converged and res.iterations == ro.iterations
        assert np.array_equal(x.to_numpy(), xo)
        assert res.residual_norm == ro.residual_norm


def test_full_volume_bit_exact(gpu):
    # the benchmark volume 32^4 (automatic shape: 16 t-slices per workgroup,
    # 4 z-ranges), weak-field links eps = 0.2, kappa = 0.124
    L = (32, 32, 32, 32)
    print("full volume 32^4 D and D^dagger", flush=True)
    U = gen.links(L, "weak", 11, eps=0.2)
    psi = gen.gaussian_spinor(L, 12)
    lat = lmx.Lattice(L, (1, 1, 1, -1))
    D = lmx.WilsonDiracOperator4D(lat.gauge_from_numpy(U), 0.124)
    x = lat.spinor_from_numpy(psi)
    y = lat.spinor()
    for dag in (False, True):
        lmx.mul_(y, lmx.adjoint(D) if dag else D, x)
        ref = orc.wilson(U, psi, L, 0.124, dagger=dag)
        assert np.array_equal(y.to_numpy(), ref), f"dagger={dag} max rel {rel(y.to_numpy(), ref):.3e}"


@pytest.mark.parametrize("nw,phases,csw", [(1, (1, 1, 1, -1), 0.0), (2, (np.exp(0.4j), -1, 1j, -1), 0.0),
                                           (1, (1, 1, 1, -1), 1.2)])
def test_apply_on_reference_storage(gpu, nw, phases, csw):
    # mul! directly on LatticeMatrix.A-layout arrays: garbage in psi's halo is
    # never read, C's halo is left untouched, the core is bit-exact
    import torch
    L = (4, 6, 2, 8)
    print(f"padded apply nw={nw} phases={phases} csw={csw}", flush=True)
    U = gen.links(L, "haar", 81)
    rng = np.random.default_rng(82)
    P = [d + 2 * nw for d in L]
    nrhs = 2
    shape = (nrhs, P[3], P[2], P[1], P[0], 4, 3)
    pad = rng.standard_normal(shape) + 1j * rng.standard_normal(shape)  # halo: garbage
    core = (slice(None),) + tuple(slice(nw, nw + d) for d in reversed(L)) + (slice(None), slice(None))
    psi_core = np.ascontiguousarray(pad[core]).reshape(nrhs, -1, 4, 3)
    lat = lmx.Lattice(L, phases)
    Ug = lat.gauge_from_numpy(U)
    D = lmx.WilsonDiracOperator4D(Ug, 0.118, csw)
    F = orc.clover_field_strength(U, L) if csw else None
    psi_d = torch.from_numpy(pad.copy()).to(lat.device)
    sentinel = np.full(shape, 7.0 - 3.0j)
    for flags in ((0, 1, 2, 3) if csw == 0 else (0, 1)):
        A = lmx.WilsonDiracOperator4D_Donly(Ug) if flags & 2 else D
        if flags & 1:
            A = lmx.adjoint(A)
        C_d = torch.from_numpy(sentinel.copy()).to(lat.device)
        lmx.mul_padded_(C_d, A, psi_d, nw)
        C = C_d.cpu().numpy()
        for i in range(nrhs):
            if csw:
                ref = orc.wilson_clover(U, F, psi_core[i], L, 0.118, csw, phases=phases, dagger=bool(flags & 1))
            else:
                ref = orc.wilson(U, psi_core[i], L, 0.118, phases, dagger=bool(flags & 1), donly=bool(flags & 2))
            got = np.ascontiguousarray(C[i][core[1:]]).reshape(-1, 4, 3)
            assert np.array_equal(got, ref), f"flags={flags} rhs={i} max rel {rel(got, ref):.3e}"
        halo = np.ones(shape, dtype=bool)
        halo[core] = False
        assert np.all(C[halo] == sentinel[halo])
    with pytest.raises(lmx.ArgumentError):
        lmx.mul_padded_(psi_d, D, psi_d, nw)


def test_clover_pullback_matches_oracle(gpu):
    # wilson_clover_link_pullback! (hopping + clover field-strength chain rule)
    L = (4, 2, 4, 6)
    U = gen.links(L, "haar", 171)
    psi = gen.gaussian_spinor(L, 172)
    w = gen.gaussian_spinor(L, 173)
    lat = lmx.Lattice(L, (1, 1, 1, -1))
    D = lmx.WilsonDiracCloverOperator4D(lat.gauge_from_numpy(U), 0.12, 1.3)
    dU = lmx.GaugeField(lat)
    lmx.wilson_pullback_(dU, D, lat.spinor_from_numpy(w), lat.spinor_from_numpy(psi), -0.12)
    ref = orc.clover_pullback(U, w, psi, L, -0.12, 1.3)
    got = dU.to_numpy()
    assert rel(got, ref) <= TOL
    assert np.array_equal(got, ref), f"max rel {rel(got, ref):.3e}"


def test_clover_force_matches_oracle(gpu):
    L = (4, 4, 4, 4)
    U = gen.links(L, "weak", 181, eps=0.3)
    eta = gen.gaussian_spinor(L, 182)
    F = orc.clover_field_strength(U, L)
    phi = orc.wilson_clover(U, F, eta, L, 0.125, 1.0, dagger=True)
    Go, Xo, chio, ro = orc.dsfdu_clover(U, phi, L, 0.125, 1.0, rtol=1e-10, order=1)
    Fo = orc.ta_force(U, Go)
    lat = lmx.Lattice(L, (1, 1, 1, -1))
    Ug = lat.gauge_from_numpy(U)
    D = lmx.WilsonDiracCloverOperator4D(Ug, 0.125, 1.0)
    G, Fg = lmx.GaugeField(lat), lmx.GaugeField(lat)
    res, eta_g, chi_g = lmx.dSFdU(G, D, lat.spinor_from_numpy(phi), force=Fg, rtol=1e-10)
    assert res.converged and res.iterations == ro.iterations
    assert np.array_equal(eta_g.to_numpy(), Xo)
    assert np.array_equal(chi_g.to_numpy(), chio)
    assert np.array_equal(G.to_numpy(), Go)
    assert rel(Fg.to_numpy(), Fo) <= TOL


@pytest.mark.parametrize("L,phases", [((8, 8, 4, 8), (1, 1, 1, -1)), ((4, 2, 3, 2), (1, 1, 1, -1)),
                                      ((3, 2, 2, 4), (np.exp(0.3j), 1, -1j, -1))])
@pytest.mark.parametrize("nrhs", [2, 3, 4])
def test_multi_rhs_fp64_apply_bit_exact(gpu, L, phases, nrhs):
    # several RHS in one FP64 launch (two per work item for Wilson, a ragged
    # last group for odd counts; streaming, per-site and complex-phase
    # kernels) equal the one-RHS applies and the oracle bit for bit
    print(f"multi-RHS FP64 L={L} nrhs={nrhs}", flush=True)
    U = make_links(L, "rand", 11)
    V = int(np.prod(L))
    psi = rand_c(np.random.default_rng(12), (nrhs, V, 4, 3))
    lat = lmx.Lattice(L, phases)
    Ug = lat.gauge_from_numpy(U)
    D = lmx.WilsonDiracOperator4D(Ug, 0.117)
    Dc = lmx.WilsonDiracCloverOperator4D(Ug, 0.117, 1.0)
    x = lat.spinor_from_numpy(psi)
    for flags in range(4):
        A = lmx.WilsonDiracOperator4D_Donly(Ug) if flags & 2 else D
        if flags & 1:
            A = lmx.adjoint(A)
        y = lat.spinor(nrhs)
        lmx.mul_(y, A, x)
        got = y.to_numpy()
        for r in range(nrhs):
            y1 = lat.spinor()
            lmx.mul_(y1, A, x.rhs(r))
            assert np.array_equal(got[r], y1.to_numpy()), (flags, r)
        ref = orc.wilson(U, psi[nrhs - 1], L, 0.117, phases, dagger=bool(flags & 1), donly=bool(flags & 2))
        assert np.array_equal(got[nrhs - 1], ref), flags
    for dag in (False, True):  # clover keeps one RHS per work item
        y = lat.spinor(nrhs)
        lmx.mul_(y, lmx.adjoint(Dc) if dag else Dc, x)
        got = y.to_numpy()
        for r in range(nrhs):
            y1 = lat.spinor()
            lmx.mul_(y1, lmx.adjoint(Dc) if dag else Dc, x.rhs(r))
            assert np.array_equal(got[r], y1.to_numpy()), ("clover", dag, r)


F32_CASES = [
    ((4, 2, 3, 2), "rand", (1, 1, 1, -1)),
    ((3, 2, 2, 4), "rand", (np.exp(0.3j), 1, -1j, -1)),  # complex phases: per-site kernel
    ((8, 8, 4, 8), "haar", (-1, 1, 1, -1)),              # streaming kernel
]


@pytest.mark.parametrize("L,kind,phases", F32_CASES)
@pytest.mark.parametrize("csw", [0.0, 1.3])
def test_single_precision_apply(gpu, L, kind, phases, csw):
    # test/wilson_halfspin.jl:105-121 at ComplexF32: D, D^dagger, the
    # hopping-only operator and its adjoint, and the clover operator within
    # 2e-5 of the FP64 oracle, and <chi, D psi> = <D^dagger chi, psi>
    kappa, tol = 0.117, 2e-5
    print(f"FP32 apply L={L} kind={kind} csw={csw}", flush=True)
    U = make_links(L, kind, 11)
    V = int(np.prod(L))
    psi = rand_c(np.random.default_rng(12), (V, 4, 3))
    chi = rand_c(np.random.default_rng(13), (V, 4, 3))
    lat = lmx.Lattice(L, phases)
    Ug = lat.gauge_from_numpy(U)
    D = lmx.WilsonDiracOperator4D(Ug, kappa, csw)
    F = orc.clover_field_strength(U, L) if csw else None
    x = lmx.spinor_f32_from_numpy(lat, psi)
    y = lmx.spinor_f32(lat)
    worst = 0.0
    for flags in (0, 1, 2, 3):
        A = lmx.WilsonDiracOperator4D_Donly(Ug) if flags & 2 else D
        if flags & 1:
            A = lmx.adjoint(A)
        lmx.mul_f32_(y, A, x)
        got = lmx.spinor_f32_to_numpy(y)[0]
        if csw and not flags & 2:
            ref = orc.wilson_clover(U, F, psi, L, kappa, csw, phases, dagger=bool(flags & 1))
        else:
            ref = orc.wilson(U, psi, L, kappa, phases, dagger=bool(flags & 1), donly=bool(flags & 2))
        e = float(np.linalg.norm(got - ref) / np.linalg.norm(ref))
        worst = max(worst, e)
        assert e <= tol, (flags, e)
    # adjoint identity in FP32
    yc = lmx.spinor_f32(lat)
    lmx.mul_f32_(y, D, x)
    lmx.mul_f32_(yc, lmx.adjoint(D), lmx.spinor_f32_from_numpy(lat, chi))
    lhs = np.vdot(chi.astype(np.complex64), lmx.spinor_f32_to_numpy(y)[0])
    rhs = np.vdot(lmx.spinor_f32_to_numpy(yc)[0], psi.astype(np.complex64))
    assert abs(lhs - rhs) <= tol * abs(lhs)
    print(f"  worst relative error vs FP64 oracle {worst:.2e}", flush=True)
    with pytest.raises(lmx.ArgumentError):
        lmx.mul_f32_(y, D, y)  # aliasing
    with pytest.raises(lmx.ArgumentError):
        lmx.mul_f32_(y, D, lmx.spinor_f32(lat).double().float()[:, :6].contiguous())


def test_single_precision_apply_many_rhs(gpu):
    # several fields in one call equal the fields applied one by one
    L = (8, 8, 4, 8)
    U = make_links(L, "haar", 21)
    V = int(np.prod(L))
    psi = rand_c(np.random.default_rng(22), (3, V, 4, 3))
    lat = lmx.Lattice(L, (1, 1, 1, -1))
    D = lmx.WilsonDiracCloverOperator4D(lat.gauge_from_numpy(U), 0.12, 1.1)
    x = lmx.spinor_f32_from_numpy(lat, psi)
    y = lmx.spinor_f32(lat, 3)
    lmx.mul_f32_(y, lmx.adjoint(D), x)
    for r in range(3):
        y1 = lmx.spinor_f32(lat)
        lmx.mul_f32_(y1, lmx.adjoint(D), x[r:r + 1].contiguous())
        assert torch.equal(y[r:r + 1], y1)


@pytest.mark.parametrize("nw,phases,csw", [(1, (1, 1, 1, -1), 0.0), (2, (np.exp(0.4j), -1, 1j, -1), 0.0),
                                           (1, (1, 1, 1, -1), 1.2)])
def test_apply_on_single_precision_reference_storage(gpu, nw, phases, csw):
    # the same on LatticeMatrix{ComplexF32}.A arrays: FP32 arithmetic within the
    # reference's 2e-5 (test/wilson_halfspin.jl:105), halo never read or written
    L = (4, 6, 2, 8)
    print(f"padded FP32 apply nw={nw} phases={phases} csw={csw}", flush=True)
    U = gen.links(L, "haar", 83)
    rng = np.random.default_rng(84)
    P = [d + 2 * nw for d in L]
    nrhs = 2
    shape = (nrhs, P[3], P[2], P[1], P[0], 4, 3)
    pad = (rng.standard_normal(shape) + 1j * rng.standard_normal(shape)).astype(np.complex64)
    core = (slice(None),) + tuple(slice(nw, nw + d) for d in reversed(L)) + (slice(None), slice(None))
    psi_core = np.ascontiguousarray(pad[core]).reshape(nrhs, -1, 4, 3).astype(np.complex128)
    lat = lmx.Lattice(L, phases)
    Ug = lat.gauge_from_numpy(U)
    D = lmx.WilsonDiracOperator4D(Ug, 0.118, csw)
    F = orc.clover_field_strength(U, L) if csw else None
    psi_d = torch.from_numpy(pad.copy()).to(lat.device)
    sentinel = np.full(shape, 7.0 - 3.0j, dtype=np.complex64)
    worst = 0.0
    for flags in ((0, 1, 2, 3) if csw == 0 else (0, 1)):
        A = lmx.WilsonDiracOperator4D_Donly(Ug) if flags & 2 else D
        if flags & 1:
            A = lmx.adjoint(A)
        C_d = torch.from_numpy(sentinel.copy()).to(lat.device)
        lmx.mul_padded_(C_d, A, psi_d, nw)
        C = C_d.cpu().numpy()
        for i in range(nrhs):
            if csw:
                ref = orc.wilson_clover(U, F, psi_core[i], L, 0.118, csw, phases=phases, dagger=bool(flags & 1))
            else:
                ref = orc.wilson(U, psi_core[i], L, 0.118, phases, dagger=bool(flags & 1), donly=bool(flags & 2))
            got = np.ascontiguousarray(C[i][core[1:]]).reshape(-1, 4, 3)
            e = float(np.linalg.norm(got - ref) / np.linalg.norm(ref))
            worst = max(worst, e)
            assert e <= 2e-5, (flags, i, e)
        halo = np.ones(shape, dtype=bool)
        halo[core] = False
        assert np.all(C[halo] == sentinel[halo])
    print(f"  worst relative error vs FP64 oracle {worst:.2e}", flush=True)
    with pytest.raises(lmx.ArgumentError):  # mixed element types
        lmx.mul_padded_(torch.zeros(shape, dtype=torch.complex128, device=lat.device), D, psi_d, nw)
