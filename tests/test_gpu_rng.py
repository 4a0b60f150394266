"""Site random-number streams and lattice fills on the GPU against the oracle.

Integer words, uniforms and bounded integers must be bit-identical to the
oracle; normals (device log / sqrt / sincos against the host C library) must
agree within the reference's own GPU tolerances, rtol = atol = 2e-12 for
Float64 and 2e-6 (draws) / 3e-6 (fills) for Float32 (test/site_rng_gpu.jl:100-106,
test/random_fill_gpu.jl:66,112); the largest differences seen are printed."""
import math

import numpy as np
import pytest
import torch

import oracle as orc
import lmx
from dense_reference import halo_fill

pytestmark = pytest.mark.gpu

ALGS = ("pcg32", "xoshiro256++", "philox4x32")
TOL32, TOL64 = 3e-6, 2e-12


def _close(a, b, tol):
    a = np.asarray(a)
    b = np.asarray(b)
    return np.allclose(a, b, rtol=tol, atol=tol), float(np.max(np.abs(a - b))) if a.size else 0.0


@pytest.mark.parametrize("alg", ALGS)
def test_site_draws_match_oracle(gpu, alg):
    # test/site_rng_gpu.jl:44-112: 257 sites, every draw kind, in one kernel
    key = (0x123456789ABCDEF0, 27, 3, 1, 2)
    counts = dict(words=19, uniform32=7, uniform64=7, normal32=8, normal64=8, bounded=11)
    lat = lmx.Lattice((4, 4, 4, 4))
    for first in (0, (1 << 40) + 5):  # the high counter / seed word of the site id is used
        print(f"draws {alg} first_site={first}", flush=True)
        g = lmx.site_rng_draw(lat, lmx.RNGStreamKey(*key), 257, algorithm=alg, first_site=first, upper=17, **counts)
        o = orc.site_rng_draw(key, alg, first, 257, upper=17, **counts)
        for k in ("words", "uniform32", "uniform64", "bounded"):
            assert np.array_equal(g[k], o[k]), k
        ok32, d32 = _close(g["normal32"], o["normal32"], 2e-6)
        ok64, d64 = _close(g["normal64"], o["normal64"], TOL64)
        print(f"  normals: max |GPU - oracle| Float32 {d32:.2e}, Float64 {d64:.2e}", flush=True)
        assert ok32 and ok64


def _expected_padded(core, dims, nw, phases):
    return halo_fill(core, dims, nw, phases)


@pytest.mark.parametrize("alg", ALGS)
def test_padded_fills_match_oracle(gpu, alg):
    dims, phases = (8, 3, 2, 4), (1, 1, 1, -1)
    lat = lmx.Lattice(dims, phases)
    tdt = {np.complex128: torch.complex128, np.complex64: torch.complex64, np.float64: torch.float64,
           np.float32: torch.float32}
    for dtype, (nc1, nc2), nw in ((np.complex128, (3, 2), 1), (np.complex64, (3, 4), 2), (np.float64, (5, 1), 1),
                                  (np.float32, (3, 3), 0)):
        for gaussian in (False, True):
            key = (0x51F15E, 11, 4, 0, 0x66) if gaussian else (0x8F3C2A19, 7, 2, 1, 0x55)
            print(f"fill {alg} {np.dtype(dtype).name} {nc1}x{nc2} nw={nw} gaussian={gaussian}", flush=True)
            P = tuple(d + 2 * nw for d in reversed(dims))
            C = torch.full(P + (nc2, nc1), 7.0, dtype=tdt[dtype], device=gpu)
            f = lmx.randomize_gaussian_matrix_ if gaussian else lmx.randomize_matrix_
            kw = dict(sigma=1.75) if gaussian else {}
            f(C, lmx.RNGStreamKey(*key), rng_algorithm=alg, lattice=lat, nw=nw, **kw)
            lat.synchronize()
            got = C.cpu().numpy()
            core = orc.randomize_fill(dtype, nc1, nc2, dims, key, gaussian=gaussian, sigma=1.75, algorithm=alg)
            want = _expected_padded(core, dims, nw, phases)
            if not gaussian:
                assert np.array_equal(got, want)
            else:
                ok, d = _close(got, want, TOL32 if dtype in (np.complex64, np.float32) else TOL64)
                print(f"  max |GPU - oracle| {d:.2e}", flush=True)
                assert ok


def test_spinor_fill_matches_padded_fill_and_oracle(gpu):
    dims = (4, 4, 2, 6)
    lat = lmx.Lattice(dims, (1, 1, 1, -1))
    keys = [lmx.RNGStreamKey(1235), lmx.RNGStreamKey(1235, 1)]
    psi = lat.spinor(2)
    lmx.randomize_gaussian_matrix_(psi, keys, sigma=math.sqrt(0.5))
    got = psi.to_numpy()
    for r, k in enumerate(keys):
        C = torch.zeros(tuple(d + 2 for d in reversed(dims)) + (4, 3), dtype=torch.complex128, device=gpu)
        lmx.randomize_gaussian_matrix_(C, k, sigma=math.sqrt(0.5), lattice=lat, nw=1)
        lat.synchronize()
        core = C.cpu().numpy()[1:-1, 1:-1, 1:-1, 1:-1].reshape(-1, 4, 3)
        assert np.array_equal(got[r], core)  # same device arithmetic in both layouts
        o = orc.randomize_fill(np.complex128, 3, 4, dims, (k.seed, k.sweep, 0, 0, 0), gaussian=True,
                               sigma=math.sqrt(0.5))
        ok, d = _close(got[r], o.reshape(-1, 4, 3), TOL64)
        assert ok, d
    u = lat.spinor(1)
    lmx.randomize_matrix_(u, seed=99, rng_algorithm="pcg32")
    o = orc.randomize_fill(np.complex128, 3, 4, dims, (99, 0, 0, 0, 0), algorithm="pcg32")
    assert np.array_equal(u.to_numpy(), o.reshape(-1, 4, 3))


def test_box_fills_are_slices_of_the_global_fill(gpu):
    G, box = (8, 4, 4, 8), (4, 4, 2, 4)
    key = lmx.RNGStreamKey(123, 9, 2)
    whole = lmx.Lattice(G).spinor(1)
    lmx.randomize_gaussian_matrix_(whole, key, sigma=0.5)
    W = whole.to_numpy().reshape(G[3], G[2], G[1], G[0], 4, 3)
    lb = lmx.Lattice(box)
    for off in ((0, 0, 0, 0), (4, 0, 2, 4), (4, 0, 0, 0)):
        part = lb.spinor(1)
        lmx.randomize_gaussian_matrix_(part, key, sigma=0.5, global_offset=off, global_dims=G)
        ref = W[off[3]:off[3] + box[3], off[2]:off[2] + box[2], off[1]:off[1] + box[1], off[0]:off[0] + box[0]]
        assert np.array_equal(part.to_numpy().reshape(ref.shape), ref)
        # a box of a larger lattice writes only its core: the halo is the exchange's
        C = torch.full(tuple(d + 2 for d in reversed(box)) + (4, 3), 3.0, dtype=torch.complex128, device=gpu)
        lmx.randomize_gaussian_matrix_(C, key, sigma=0.5, lattice=lb, nw=1, global_offset=off, global_dims=G)
        lb.synchronize()
        c = C.cpu().numpy()
        assert np.array_equal(c[1:-1, 1:-1, 1:-1, 1:-1], ref)
        assert np.all(c[0] == 3.0) and np.all(c[:, :, :, -1] == 3.0)


def test_benchmark_source_32x4(gpu):
    # the D^dagger D CG benchmark's source: randomize_gaussian_matrix!(source.f;
    # sigma = sqrt(1/2), seed = 1234 + 1), Philox (benchmark/ddagd_cg/ddagd_cg.jl:773-774)
    dims = (32, 32, 32, 32)
    lat = lmx.Lattice(dims, (1, 1, 1, -1))
    b = lat.spinor(1)
    lmx.randomize_gaussian_matrix_(b, sigma=math.sqrt(0.5), seed=1235)
    torch.cuda.synchronize()
    reps = 20
    t0 = torch.cuda.Event(enable_timing=True)
    t1 = torch.cuda.Event(enable_timing=True)
    t0.record()
    for _ in range(reps):
        lmx.randomize_gaussian_matrix_(b, sigma=math.sqrt(0.5), seed=1235)
    t1.record()
    torch.cuda.synchronize()
    ms = t0.elapsed_time(t1) / reps
    got = b.to_numpy()
    o = orc.randomize_fill(np.complex128, 3, 4, dims, (1235, 0, 0, 0, 0), gaussian=True, sigma=math.sqrt(0.5))
    ok, d = _close(got, o.reshape(-1, 4, 3), TOL64)
    exact = float(np.mean(got == o.reshape(-1, 4, 3)))
    sd = float(np.sqrt(np.mean(np.abs(got) ** 2) / 2))
    print(f"32^4 benchmark source: {ms * 1e3:.1f} us per fill ({lat.V * 192 / ms / 1e6:.0f} GB/s written), "
          f"max |GPU - oracle| {d:.2e}, {exact * 100:.2f}% of entries bit-identical, per-component sd {sd:.5f}",
          flush=True)
    assert ok
    assert abs(sd - math.sqrt(0.5)) < 1e-3


def test_fill_argument_errors(gpu):
    lat = lmx.Lattice((4, 4, 4, 4), (1, 1, 1, 1j))
    psi = lat.spinor(2)
    k = lmx.RNGStreamKey(1)
    with pytest.raises(lmx.ArgumentError):
        lmx.randomize_gaussian_matrix_(psi, [k, k], sigma=-1.0)
    with pytest.raises(lmx.ArgumentError):
        lmx.randomize_matrix_(psi, [k])  # one key per right-hand side
    with pytest.raises(lmx.ArgumentError):
        lmx.randomize_matrix_(psi, [k, k], rng_algorithm="mt19937")
    with pytest.raises(lmx.ArgumentError):
        lmx.randomize_matrix_(psi, [k, k], global_offset=(0, 0, 0, 0))
    with pytest.raises(lmx.ArgumentError):
        lmx.randomize_matrix_(psi, [k, k], global_offset=(2, 0, 0, 0), global_dims=(4, 4, 4, 4))
    with pytest.raises(lmx.ArgumentError):
        lmx.RNGStreamKey(-1)
    with pytest.raises(lmx.ArgumentError):
        lmx.RNGStreamKey(1, 0, 1 << 32)
    real = torch.zeros((6, 6, 6, 6, 1, 2), dtype=torch.float64, device=gpu)
    with pytest.raises(lmx.ArgumentError):  # a real field cannot take the complex phase i
        lmx.randomize_matrix_(real, k, lattice=lat, nw=1)
    lmx.randomize_matrix_(real[1:-1, 1:-1, 1:-1, 1:-1].contiguous(), k, lattice=lat, nw=0)  # no halo: no phase
    with pytest.raises(lmx.ArgumentError):
        lmx.site_rng_draw(lat, k, 4, normal64=3)
    with pytest.raises(lmx.ArgumentError):
        lmx.site_rng_draw(lat, k, 4, bounded=2, upper=0)
