"""Mixed-precision and multi-RHS D^dagger D CG on the GPU against the FP64
oracle.  The mixed solvers stop on the FP64 true residual, so the bar is
|b - A x| / |b| <= rtol (checked independently on the host with the oracle's
operator) and the solution error against the oracle's FP64 solution, which
is bounded by cond(A) * rtol; both are printed per RHS."""
import numpy as np
import pytest

import oracle as orc
import lmx
from lmx import generators as gen

pytestmark = pytest.mark.gpu


def setup(L, kind, seed, kappa, csw=0.0, nrhs=3):
    U = gen.links(L, kind, seed, eps=0.2)
    b = gen.gaussian_spinor(L, seed + 1, nrhs=nrhs)
    lat = lmx.Lattice(L, (1, 1, 1, -1))
    Ug = lat.gauge_from_numpy(U)
    D = lmx.WilsonDiracOperator4D(Ug, kappa, csw)
    return U, b, lat, Ug, D


def oracle_check(U, L, kappa, b, x, rtol, csw=0.0):
    F = orc.clover_field_strength(U, L) if csw else None
    op = orc.NormalOp(orc.OP_DDAGD_CLOVER if csw else orc.OP_DDAGD_WILSON, L, U=U, F=F, kappa=kappa, csw=csw)
    out = []
    for i in range(b.shape[0]):
        res = np.linalg.norm(op.apply(x[i]) - b[i]) / np.linalg.norm(b[i])
        xo, ro = orc.cg(op, b[i], rtol=1e-12, maxiter=20000)
        err = np.linalg.norm(x[i] - xo) / np.linalg.norm(xo)
        out.append((res, err))
    return out


def test_h16_projection_at_full_scale(gpu):
    # Every component of every site equal (RHS 0: 1 + i; RHS 1: 1 - i): the
    # 16-bit fixed-point values all sit at the format's maximum, 32752, and
    # the projected half-spinors, sums of two components, reach 2 x 32752 =
    # 65504, the largest finite FP16 value, in the FP16 pair products of the
    # link multiplication.  With a maximum of 32767 (65534) they would round
    # to infinity and the solve would break down at its first iteration.
    L, kappa, rtol = (8, 8, 4, 8), 0.124, 1e-10
    U, b, lat, Ug, D = setup(L, "weak", 341, kappa, nrhs=2)
    b = np.empty_like(b)
    b[0] = 1 + 1j
    b[1] = 1 - 1j
    print(f"h16 at full scale L={L}", flush=True)
    x = lat.spinor(2)
    res = lmx.cg_batched_(x, D, lat.spinor_from_numpy(b), precision="h16", rtol=rtol, use_graph=False)
    checks = oracle_check(U, L, kappa, b, x.to_numpy(), rtol)
    for r, (tr, err) in zip(res, checks):
        print(f"  iters={r.iterations} reason={r.reason} true_rel={tr:.3e} err_vs_fp64={err:.3e}", flush=True)
        assert r.converged and r.reason == "converged"
        assert tr <= rtol * 1.05 and err <= 1e-7


@pytest.mark.parametrize("precision", ["f32", "f16", "h16"])
@pytest.mark.parametrize("use_graph", [False, True])
def test_mixed_batched_converges_to_fp64_true_residual(gpu, precision, use_graph):
    L, kappa, rtol = (8, 8, 8, 8), 0.124, 1e-10
    U, b, lat, Ug, D = setup(L, "weak", 300, kappa)
    print(f"mixed {precision} graph={use_graph} L={L}", flush=True)
    x = lat.spinor(3)
    bg = lat.spinor_from_numpy(b)
    res = lmx.cg_batched_(x, D, bg, precision=precision, rtol=rtol, use_graph=use_graph)
    xs = x.to_numpy()
    checks = oracle_check(U, L, kappa, b, xs, rtol)
    for r, (tr, err) in zip(res, checks):
        print(f"  iters={r.iterations} updates={r.reliable_updates} true_rel={tr:.3e} "
              f"reported={r.true_relative_residual:.3e} err_vs_fp64={err:.3e}", flush=True)
        assert r.converged and r.reason == "converged"
        assert tr <= rtol * 1.05
        assert abs(tr - r.true_relative_residual) <= 1e-3 * rtol + 1e-6 * tr
        assert err <= 1e-7


@pytest.mark.parametrize("precision,L,nrhs", [("f32", (8, 4, 4, 8), 2), ("h16", (8, 8, 4, 8), 5)])
def test_graph_and_eager_agree(gpu, precision, L, nrhs):
    # every solve replays its graph many times with host polls in between;
    # the replays must give the eager launches' result bit for bit (8x8x4 is
    # a streaming shape, 8x4x4 the per-site kernel)
    kappa = 0.12
    print(f"graph vs eager {precision} L={L} nrhs={nrhs}", flush=True)
    U, b, lat, Ug, D = setup(L, "haar", 310, kappa, nrhs=nrhs)
    bg = lat.spinor_from_numpy(b)
    x1, x2 = lat.spinor(nrhs), lat.spinor(nrhs)
    r1 = lmx.cg_batched_(x1, D, bg, precision=precision, use_graph=False, check_every=5)
    r2 = lmx.cg_batched_(x2, D, bg, precision=precision, use_graph=True, check_every=5)
    assert all(r.converged for r in r1 + r2)
    assert [r.iterations for r in r1] == [r.iterations for r in r2]
    assert [r.reliable_updates for r in r1] == [r.reliable_updates for r in r2]
    assert np.array_equal(x1.to_numpy(), x2.to_numpy())


@pytest.mark.parametrize("rhs_block", [(1 << 4) | 1, (4 << 4) | 2, (5 << 4) | 1, (3 << 4) | 2, (4 << 4) | 4])
def test_rhs_blocking_uneven_batch(gpu, rhs_block):
    # 5 RHS split unevenly by the launch shape; every shape gives the same iterates
    L, kappa = (4, 4, 4, 8), 0.12
    U, b, lat, Ug, D = setup(L, "haar", 320, kappa, nrhs=5)
    bg = lat.spinor_from_numpy(b)
    x = lat.spinor(5)
    res = lmx.cg_batched_(x, D, bg, precision="f16", rhs_block=rhs_block)
    ref = lat.spinor(5)
    rref = lmx.cg_batched_(ref, D, bg, precision="f16", rhs_block=(1 << 4) | 1)
    assert [r.iterations for r in res] == [r.iterations for r in rref]
    assert np.array_equal(x.to_numpy(), ref.to_numpy())
    assert all(r.converged for r in res)


def test_plain_fp64_batched_matches_single(gpu):
    L, kappa = (4, 4, 4, 4), 0.125
    U, b, lat, Ug, D = setup(L, "haar", 330, kappa, nrhs=2)
    bg = lat.spinor_from_numpy(b)
    x = lat.spinor(2)
    res = lmx.cg_batched_(x, D, bg, precision="f64", rtol=1e-10)
    op = orc.NormalOp(orc.OP_DDAGD_WILSON, L, U=U, kappa=kappa)
    for i in range(2):
        xo, ro = orc.cg(op, b[i], rtol=1e-10, order=1)
        assert res[i].iterations == ro.iterations
        assert np.array_equal(x.to_numpy()[i], xo)


def test_mixed_edge_cases(gpu):
    L, kappa = (4, 4, 4, 4), 0.12
    U, b, lat, Ug, D = setup(L, "haar", 340, kappa, nrhs=3)
    b[1] = 0.0  # a zero right-hand side converges at iteration 0
    bg = lat.spinor_from_numpy(b)
    x = lat.spinor(3)
    res = lmx.cg_batched_(x, D, bg, precision="f32")
    assert res[1].converged and res[1].iterations == 0 and res[1].residual_norm == 0.0
    assert res[0].converged and res[2].converged
    assert np.all(x.to_numpy()[1] == 0)
    # maxiter reached: status, not an exception
    x2 = lat.spinor(3)
    res = lmx.cg_batched_(x2, D, bg, precision="f32", maxiter=3, check_every=1)
    assert res[0].reason == "maximum_iterations" and not res[0].converged
    with pytest.raises(lmx.ArgumentError):
        lmx.cg_batched_(x2, D, x2, precision="f32")


def test_mixed_clover(gpu):
    L, kappa, csw, rtol = (8, 4, 4, 8), 0.1235, 1.0, 1e-10
    U = gen.links(L, "weak", 350, eps=0.15)
    b = gen.gaussian_spinor(L, 351, nrhs=2)
    lat = lmx.Lattice(L, (1, 1, 1, -1))
    D = lmx.WilsonDiracCloverOperator4D(lat.gauge_from_numpy(U), kappa, csw)
    bg = lat.spinor_from_numpy(b)
    x = lat.spinor(2)
    res = lmx.cg_batched_(x, D, bg, precision="f16", rtol=rtol)
    checks = oracle_check(U, L, kappa, b, x.to_numpy(), rtol, csw=csw)
    for r, (tr, err) in zip(res, checks):
        print(f"  clover iters={r.iterations} true_rel={tr:.3e} err={err:.3e}", flush=True)
        assert r.converged and tr <= rtol * 1.05 and err <= 1e-7


@pytest.mark.parametrize("precision", ["f16", "h16"])
def test_mixed_stream_shape(gpu, monkeypatch, precision):
    # forced multi-slice streaming shape with the fused p.Ap on D^dagger,
    # 4 RHS at 2 per work item, graph replay
    monkeypatch.setenv("LMX_STREAM_TL", "4")
    monkeypatch.setenv("LMX_STREAM_ZS", "2")
    L, kappa, rtol = (8, 8, 8, 8), 0.124, 1e-10
    U, b, lat, Ug, D = setup(L, "weak", 360, kappa, nrhs=4)
    x = lat.spinor(4)
    res = lmx.cg_batched_(x, D, lat.spinor_from_numpy(b), precision=precision, rtol=rtol, use_graph=True,
                          rhs_block=(4 << 4) | 2)
    checks = oracle_check(U, L, kappa, b, x.to_numpy(), rtol)
    for r, (tr, err) in zip(res, checks):
        print(f"  {precision} iters={r.iterations} updates={r.reliable_updates} true_rel={tr:.3e} err={err:.3e}",
              flush=True)
        assert r.converged and tr <= rtol * 1.05 and err <= 1e-7


def test_many_rhs_split_over_launches_h16(gpu):
    # 14 RHS in launches of 6, 6, 2 at 2 per work item, against one launch of
    # 14 at 1 per work item: identical iterates (each RHS is computed and
    # reduced independently of the grouping), all converged to the true residual
    L, kappa, rtol = (8, 8, 4, 8), 0.124, 1e-10
    U, b, lat, Ug, D = setup(L, "weak", 370, kappa, nrhs=14)
    bg = lat.spinor_from_numpy(b)
    x1, x2 = lat.spinor(14), lat.spinor(14)
    r1 = lmx.cg_batched_(x1, D, bg, precision="h16", rtol=rtol, use_graph=True, rhs_block=(6 << 4) | 2)
    r2 = lmx.cg_batched_(x2, D, bg, precision="h16", rtol=rtol, use_graph=False, rhs_block=(14 << 4) | 1)
    assert [r.iterations for r in r1] == [r.iterations for r in r2]
    assert np.array_equal(x1.to_numpy(), x2.to_numpy())
    checks = oracle_check(U, L, kappa, b[:3], x1.to_numpy()[:3], rtol)
    for r, (tr, err) in zip(r1, checks):
        assert r.converged and tr <= rtol * 1.05 and err <= 1e-7
    assert all(r.converged for r in r1)


@pytest.mark.parametrize("check_every", [1, 3, 6, 7])
def test_batched_x_update_any_poll_interval(gpu, check_every):
    # the x update groups iterations by four (the first three defer alpha p
    # into their ring slots, the fourth applies them); poll intervals that
    # split groups, graph (rounded up to a whole group) or eager, must not
    # change the iterates, and reliable updates land inside groups
    L, kappa = (8, 4, 4, 8), 0.12
    U, b, lat, Ug, D = setup(L, "weak", 380, kappa, nrhs=2)
    bg = lat.spinor_from_numpy(b)
    xs = []
    for graph in (False, True):
        x = lat.spinor(2)
        res = lmx.cg_batched_(x, D, bg, precision="h16", use_graph=graph, check_every=check_every)
        assert all(r.converged for r in res)
        xs.append((x.to_numpy(), [r.iterations for r in res]))
    assert xs[0][1] == xs[1][1]
    assert np.array_equal(xs[0][0], xs[1][0])
    # stopped by maxiter with one, two or three deferred terms: all are folded in
    A = orc.NormalOp(orc.OP_DDAGD_WILSON, L, U=U, kappa=kappa)
    for maxiter in (5, 6, 7):
        x = lat.spinor(2)
        res = lmx.cg_batched_(x, D, bg, precision="h16", maxiter=maxiter, check_every=1)
        assert res[0].reason == "maximum_iterations"
        tr = np.linalg.norm(A.apply(x.to_numpy()[0]) - b[0]) / np.linalg.norm(b[0])
        assert abs(tr - res[0].true_relative_residual) <= 1e-6 * tr


def test_rhs_count_limits(gpu):
    L, kappa = (4, 4, 4, 8), 0.12
    U, b, lat, Ug, D = setup(L, "weak", 390, kappa, nrhs=64)
    bg = lat.spinor_from_numpy(b)
    x = lat.spinor(64)
    res = lmx.cg_batched_(x, D, bg, precision="h16", use_graph=True)
    assert len(res) == 64 and all(r.converged for r in res)
    b65 = lat.spinor(65)
    with pytest.raises(lmx.ArgumentError):
        lmx.cg_batched_(lat.spinor(65), D, b65, precision="h16")


def test_z_range_shapes_are_bit_identical(gpu):
    # the z-range count (rhs_block bits 20-23) only changes which workgroup
    # computes which site; arithmetic and reduction order are the same
    L, kappa = (16, 8, 8, 8), 0.12
    print(f"z-range shapes L={L}", flush=True)
    U, b, lat, Ug, D = setup(L, "weak", 320, kappa, nrhs=3)
    bg = lat.spinor_from_numpy(b)
    ref = None
    for zs in (1, 2, 4):
        x = lat.spinor(3)
        res = lmx.cg_batched_(x, D, bg, precision="h16", use_graph=True, rhs_block=(zs << 20) | (4 << 12) | (3 << 4) | 1)
        assert all(r.converged for r in res)
        got = ([r.iterations for r in res], x.to_numpy())
        if ref is None:
            ref = got
        else:
            assert got[0] == ref[0] and np.array_equal(got[1], ref[1]), zs


@pytest.mark.parametrize("check_every", [1, 10])
@pytest.mark.parametrize("maxiter", [5, 6, 7, 13])
def test_graph_stops_at_maxiter(gpu, maxiter, check_every):
    # cg! stops after exactly maxiter iterations with reason maximum_iterations
    # (src/LinearAlgebras/cg.jl:105,135-136); graph replays run whole chunks,
    # so the device retires each RHS in the iteration that reaches maxiter and
    # the replayed remainder of the chunk leaves it untouched: same iterate as
    # the eager launches, its inner solution folded in, its true residual
    # reported
    L, kappa = (8, 4, 4, 8), 0.12
    print(f"graph maxiter={maxiter} check_every={check_every}", flush=True)
    U, b, lat, Ug, D = setup(L, "weak", 385, kappa, nrhs=2)
    bg = lat.spinor_from_numpy(b)
    A = orc.NormalOp(orc.OP_DDAGD_WILSON, L, U=U, kappa=kappa)
    x, xe = lat.spinor(2), lat.spinor(2)
    res = lmx.cg_batched_(x, D, bg, precision="h16", maxiter=maxiter, check_every=check_every, use_graph=True)
    rese = lmx.cg_batched_(xe, D, bg, precision="h16", maxiter=maxiter, check_every=check_every, use_graph=False)
    xs = x.to_numpy()
    for i, (r, re_) in enumerate(zip(res, rese)):
        assert r.iterations == maxiter and re_.iterations == maxiter
        assert r.reason == "maximum_iterations" and not r.converged
        assert r.reliable_updates == re_.reliable_updates
        tr = np.linalg.norm(A.apply(xs[i]) - b[i]) / np.linalg.norm(b[i])
        assert abs(tr - r.true_relative_residual) <= 1e-6 * tr
    assert np.array_equal(xs, xe.to_numpy())


def test_low_precision_copies_follow_in_place_link_writes(gpu):
    # the FP64 path reads the links themselves, the single-precision and mixed
    # paths the operator's rounded copies: after an in-place write to the
    # links every path must apply the new operator
    L, kappa = (8, 4, 4, 8), 0.12
    U, b, lat, Ug, D = setup(L, "weak", 395, kappa, nrhs=2)
    bg = lat.spinor_from_numpy(b)
    lmx.cg_batched_(lat.spinor(2), D, bg, precision="h16")  # builds the copies from the first links
    psi32 = lmx.spinor_f32_from_numpy(lat, b)
    out32 = lmx.spinor_f32(lat, 2)
    lmx.mul_f32_(out32, D, psi32)
    G2 = lat.gauge_from_numpy(gen.links(L, "weak", 396, eps=0.2))
    Ug.data.copy_(G2.data)  # in place: same tensor, new version
    lmx.mul_f32_(out32, D, psi32)
    out64 = lat.spinor(2)
    lmx.mul_(out64, D, bg)
    got, ref = lmx.spinor_f32_to_numpy(out32), out64.to_numpy()
    assert np.max(np.abs(got - ref)) <= 1e-5 * np.max(np.abs(ref))
    D2 = lmx.WilsonDiracOperator4D(G2, kappa)
    x1, x2 = lat.spinor(2), lat.spinor(2)
    r1 = lmx.cg_batched_(x1, D, bg, precision="h16")
    r2 = lmx.cg_batched_(x2, D2, bg, precision="h16")
    assert [r.iterations for r in r1] == [r.iterations for r in r2]
    assert np.array_equal(x1.to_numpy(), x2.to_numpy())
    assert all(r.converged for r in r1)


RHS_LOOP = 1 << 28


@pytest.mark.parametrize("precision,per,csw", [("h16", 5, 0.0), ("h16", 2, 0.0), ("f16", 5, 0.0), ("h16", 5, 1.0)])
def test_rhs_loop_shape(gpu, precision, per, csw):
    # the RHS-loop shape (each work item applies its site's FP16 links to every
    # RHS of the launch) against the FP64 oracle, on an uneven batch of 5 RHS
    # split into launches of `per`, with graphs and with clover
    L, kappa, rtol = (8, 8, 4, 8), 0.124, 1e-10
    U, b, lat, Ug, D = setup(L, "weak", 330, kappa, csw=csw, nrhs=5)
    print(f"RHS-loop {precision} per={per} csw={csw} L={L}", flush=True)
    bg = lat.spinor_from_numpy(b)
    x = lat.spinor(5)
    res = lmx.cg_batched_(x, D, bg, precision=precision, rtol=rtol, use_graph=True,
                          rhs_block=RHS_LOOP | (per << 4) | 1)
    ref = lat.spinor(5)
    rref = lmx.cg_batched_(ref, D, bg, precision=precision, rtol=rtol, rhs_block=(5 << 4) | 1)
    checks = oracle_check(U, L, kappa, b, x.to_numpy(), rtol, csw=csw)
    for r, q, (tr, err) in zip(res, rref, checks):
        print(f"  iters={r.iterations} (per-RHS shape {q.iterations}) true_rel={tr:.3e} err_vs_fp64={err:.3e}",
              flush=True)
        assert r.converged and tr <= rtol * 1.05 and err <= 1e-7
        assert abs(r.iterations - q.iterations) <= 3
    # the inner kernels alone at this shape: timed without error
    t = lmx.inner_kernel_times(D, precision, 5, RHS_LOOP | (per << 4) | 1, groups=1)
    assert t["D"] > 0 and t["Ddag"] > 0


LDS_STAGE = 1 << 29


@pytest.mark.parametrize("precision,csw,L", [("h16", 0.0, (32, 16, 2, 8)), ("h16", 1.0, (32, 16, 2, 8)),
                                             ("f16", 0.0, (32, 16, 2, 8)), ("f32", 1.0, (32, 16, 2, 8)),
                                             ("h16", 0.0, (24, 8, 4, 8)), ("f32", 1.0, (8, 8, 8, 4))])
def test_lds_stage_shape(gpu, precision, csw, L):
    # the LDS-staged stencils (neighbours inside the workgroup's 256-site
    # block, the t+1 and own spinors from LDS) repeat the streaming kernel's
    # arithmetic, hop order and block partials, so the whole solve is
    # bit-identical to the streaming shape at one RHS per work item.
    # 32x16x2x8: two blocks per z-plane (y neighbours inside the block and
    # across its edge); 24x8x4x8: blocks cut x-rows (x neighbours in the next
    # block); 8x8x8x4: four z-planes per block (z neighbours in the block).
    # 3 RHS in launches of 2 leave a ragged last launch.
    kappa, rtol = 0.124, 1e-10
    U, b, lat, Ug, D = setup(L, "weak", 331, kappa, csw=csw, nrhs=3)
    print(f"LDS-staged {precision} csw={csw} L={L}", flush=True)
    bg = lat.spinor_from_numpy(b)
    shape = (min(4, L[3]) << 12) | (2 << 4) | 1
    x, xs = lat.spinor(3), lat.spinor(3)
    res = lmx.cg_batched_(x, D, bg, precision=precision, rtol=rtol, use_graph=True, rhs_block=LDS_STAGE | shape)
    ref = lmx.cg_batched_(xs, D, bg, precision=precision, rtol=rtol, use_graph=True, rhs_block=shape)
    assert [r.iterations for r in res] == [q.iterations for q in ref]
    assert np.array_equal(x.to_numpy(), xs.to_numpy())
    checks = oracle_check(U, L, kappa, b, x.to_numpy(), rtol, csw=csw)
    for r, (tr, err) in zip(res, checks):
        print(f"  iters={r.iterations} true_rel={tr:.3e} err_vs_fp64={err:.3e}", flush=True)
        assert r.converged and tr <= rtol * 1.05 and err <= 1e-7
    t = lmx.inner_kernel_times(D, precision, 3, LDS_STAGE | shape, groups=1)
    assert t["D"] > 0 and t["Ddag"] > 0
