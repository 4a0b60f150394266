"""lmx_cg_ddagd_noalloc (the allocation-free FP64 D^dagger D CG) against
lmx_cg_ddagd, through the C ABI.

Both entry points run the same kernels; the allocation-free one keeps its
reduction partials and CG state in the library's static per-GPU scratch
array instead of the context's buffers.  Two calls in a row share that
array, the second on a smaller lattice (it must not see the first call's
contents); each must equal lmx_cg_ddagd bit for bit (solution, iteration
count, reason, residuals).  A lattice above the scratch's 2^24-site limit is
rejected before any device work.
"""
import ctypes

import numpy as np
import pytest
import torch

import lmx
from lmx import _lib
from lmx import generators as gen

pytestmark = pytest.mark.gpu


def solve(entry, D, b, **kw):
    lat = b.lattice
    x, r, p, Ap, tmp = (lat.spinor() for _ in range(5))
    params = _lib.default_params(**kw)
    res = _lib.CGResultC()
    _lib.check(getattr(_lib.lib(), entry)(D._op, x.ptr, b.ptr, r.ptr, p.ptr, Ap.ptr, tmp.ptr, ctypes.byref(params),
                                          ctypes.byref(res)))
    return x.to_numpy(), res


def fields(res):
    return (res.converged, res.iterations, res.reason, res.residual_norm, res.relative_residual)


# (lattice, link ensemble, kappa, csw): a clover solve on the t-streaming
# stencil, then a smaller Wilson solve on the per-site kernel
CASES = [((8, 8, 8, 16), "weak", 0.124, 1.0), ((4, 4, 4, 4), "haar", 0.125, 0.0)]


def test_noalloc_cg_equals_cg_ddagd(gpu):
    for L, ens, kappa, csw in CASES:  # two calls in a row, the second on a smaller lattice
        print(f"noalloc CG on {L}, {ens} links, kappa {kappa}, csw {csw}", flush=True)
        lat = lmx.Lattice(L, (1, 1, 1, -1))
        U = gen.links(L, ens, 4242, eps=0.2)
        D = lmx.WilsonDiracOperator4D(lat.gauge_from_numpy(U), kappa, csw=csw)
        b = lat.spinor_from_numpy(gen.gaussian_spinor(L, 4343))
        x, r = solve("lmx_cg_ddagd_noalloc", D, b)
        x_ref, r_ref = solve("lmx_cg_ddagd", D, b)
        print(f"  iterations {r.iterations} (lmx_cg_ddagd {r_ref.iterations}), "
              f"relative residual {r.relative_residual:.3e}", flush=True)
        assert r_ref.converged
        assert fields(r) == fields(r_ref)
        assert np.array_equal(x, x_ref)
        # a maxiter stop: the state kept in the scratch carries the reason
        x_m, r_m = solve("lmx_cg_ddagd_noalloc", D, b, maxiter=7)
        x_mr, r_mr = solve("lmx_cg_ddagd", D, b, maxiter=7)
        assert fields(r_m) == fields(r_mr) and r_m.iterations == 7 and r_m.reason == _lib.LMX_CG_MAXITER
        assert np.array_equal(x_m, x_mr)


def test_noalloc_cg_rejects_volumes_beyond_its_scratch(gpu):
    # 64 x 64 x 64 x 65 sites > 2^24: rejected on the host before any launch
    # (the fields are real buffers of the full size all the same)
    L = (64, 64, 64, 65)
    lat = lmx.Lattice(L, (1, 1, 1, -1))
    V = int(np.prod(L))
    links = torch.empty(4 * 9 * V * 2, dtype=torch.float64, device=gpu)
    op = ctypes.c_void_p()
    _lib.check(_lib.lib().lmx_wilson_create(lat.ctx, ctypes.byref(op), ctypes.c_void_p(links.data_ptr()), 0.125,
                                            0.0))
    try:
        f = [torch.empty(12 * V * 2, dtype=torch.float64, device=gpu) for _ in range(6)]
        params = _lib.default_params()
        res = _lib.CGResultC()
        rc = _lib.lib().lmx_cg_ddagd_noalloc(op, *[ctypes.c_void_p(t.data_ptr()) for t in f], ctypes.byref(params),
                                             ctypes.byref(res))
        assert rc == _lib.LMX_ERR_ARG
        assert b"exceeds the allocation-free CG's scratch" in _lib.lib().lmx_last_error()
        del f
    finally:
        _lib.lib().lmx_wilson_destroy(op)
        del links
        torch.cuda.empty_cache()
