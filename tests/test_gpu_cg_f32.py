"""cg! on ComplexF32 fields (src/LinearAlgebras/cg.jl:73-137 is generic in
the element type) through lmx_cg_ddagd_f32: FP32 operator and vectors, FP64
site sums and scalars.  FP32 arithmetic cannot reproduce the FP64 oracle's
iterates bit for bit, so the bar is: the FP64 true residual of the FP32
solution (recomputed with the oracle's operator) near the requested rtol, the
iteration count near the FP64 oracle's at the same rtol, the solution near
the oracle's FP64 solution, and cg!'s statuses (converged at 0 for b = 0,
maximum_iterations, the alias error) as in the FP64 solver."""
import numpy as np
import pytest

import oracle as orc
import lmx
from lmx import generators as gen

pytestmark = pytest.mark.gpu


def _setup(L, kappa, csw, seed):
    U = gen.links(L, "haar", seed)
    b = gen.gaussian_spinor(L, seed + 1, nrhs=1)[0]
    lat = lmx.Lattice(L, (1, 1, 1, -1))
    D = lmx.WilsonDiracOperator4D(lat.gauge_from_numpy(U), kappa, csw)
    return U, b, lat, D


@pytest.mark.parametrize("csw", [0.0, 1.0])
def test_cg_f32_against_fp64_oracle(gpu, csw):
    L, kappa, rtol = (4, 4, 4, 8), 0.12, 1e-5
    U, b, lat, D = _setup(L, kappa, csw, 910)
    print(f"cg_ f32 L={L} kappa={kappa} csw={csw} rtol={rtol}", flush=True)
    x, bf, r, p, Ap = (lmx.spinor_f32(lat) for _ in range(5))
    bf.copy_(lmx.spinor_f32_from_numpy(lat, b))
    res = lmx.cg_(x, lmx.DdagDOp(D, lmx.spinor_f32(lat)), bf, r, p, Ap, rtol=rtol, maxiter=2000)
    F = orc.clover_field_strength(U, L) if csw else None
    op = orc.NormalOp(orc.OP_DDAGD_CLOVER if csw else orc.OP_DDAGD_WILSON, L, U=U, F=F, kappa=kappa, csw=csw)
    xo, ro = orc.cg(op, b, rtol=rtol, maxiter=2000)
    xs = lmx.spinor_f32_to_numpy(x)[0].astype(np.complex128)
    true_rel = np.linalg.norm(op.apply(xs) - b) / np.linalg.norm(b)
    err = np.linalg.norm(xs - xo) / np.linalg.norm(xo)
    print(f"  iterations {res.iterations} (FP64 oracle {ro.iterations}), reported rel {res.relative_residual:.3e}, "
          f"FP64 true rel {true_rel:.3e}, error vs FP64 {err:.3e}", flush=True)
    assert res.converged and res.reason == "converged"
    assert res.relative_residual <= rtol
    assert abs(res.iterations - ro.iterations) <= 3
    assert true_rel <= 3 * rtol
    assert err <= 1e-3


def test_cg_f32_statuses(gpu):
    L, kappa = (4, 4, 4, 4), 0.12
    U, b, lat, D = _setup(L, kappa, 0.0, 920)
    N = lmx.DdagDOp(D, lmx.spinor_f32(lat))
    x, bf, r, p, Ap = (lmx.spinor_f32(lat) for _ in range(5))
    # b = 0: converged at iteration 0 with x = 0 (cg.jl:95-103)
    res = lmx.cg_(x, N, bf, r, p, Ap, rtol=1e-6)
    assert res.converged and res.iterations == 0 and res.residual_norm == 0.0
    # maxiter reached: a status, not an exception (cg.jl:135-136)
    bf.copy_(lmx.spinor_f32_from_numpy(lat, b))
    res = lmx.cg_(x, N, bf, r, p, Ap, rtol=1e-12, maxiter=3, check_every=1)
    assert res.reason == "maximum_iterations" and res.iterations == 3 and not res.converged
    assert np.isfinite(res.residual_norm)
    # aliasing fields are rejected (_validate_cg_arguments, cg.jl:33-47)
    with pytest.raises(lmx.ArgumentError):
        lmx.cg_(x, N, bf, r, r, Ap)
    # fields of the wrong shape are rejected
    with pytest.raises(lmx.ArgumentError):
        lmx.cg_(x, N, bf, r, p, lmx.spinor_f32(lat, 2))
