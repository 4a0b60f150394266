"""LatticeMatrix(A, dim, PEs) construction and gather_matrix on the device
(src/LatticeMatrices_core.jl:405-461, 936-993) for any NC1 x NC2 and element
type, against the numpy restatement of set_halo! (tests/dense_reference.py).
Phases are +-1 and +-i, whose products are exact, so the bar is bit-identity."""
import numpy as np
import pytest
import torch

import lmx
from dense_reference import halo_fill

pytestmark = pytest.mark.gpu

CASES = [  # dtype, (nc1, nc2), nw, phases
    (np.complex128, (3, 4), 1, (1, 1, 1, -1)),
    (np.complex128, (3, 3), 2, (1, 1j, -1, -1j)),
    (np.complex64, (3, 3), 1, (1, -1, 1, -1)),
    (np.float64, (5, 1), 1, (1, 1, -1, -1)),
    (np.float32, (2, 2), 0, (1, 1, 1, 1)),
    (np.float32, (1, 1), 2, (-1, 1, 1, -1)),  # nw equal to the z extent
]


def _random(shape, dtype, seed):
    rng = np.random.default_rng(seed)
    a = rng.standard_normal(shape)
    if np.dtype(dtype).kind == "c":
        a = a + 1j * rng.standard_normal(shape)
    return a.astype(dtype)


@pytest.mark.parametrize("case", range(len(CASES)))
def test_pad_and_gather_match_set_halo(gpu, case):
    dtype, (nc1, nc2), nw, phases = CASES[case]
    dims = (4, 3, 2, 6)
    print(f"fields {np.dtype(dtype).name} {nc1}x{nc2} nw={nw} phases={phases}", flush=True)
    lat = lmx.Lattice(dims, phases)
    core = _random(tuple(reversed(dims)) + (nc2, nc1), dtype, 40 + case)
    A = torch.from_numpy(core).to(gpu)
    C = lmx.lattice_matrix(lat, A, nw)
    lat.synchronize()
    assert np.array_equal(C.cpu().numpy(), halo_fill(core, dims, nw, phases))
    B = lmx.gather_matrix(lat, C, nw)
    lat.synchronize()
    assert np.array_equal(B.cpu().numpy(), core)


def test_pad_agrees_with_spinor_export(gpu):
    # the generic construction of a (3,4) ComplexF64 field equals the spinor
    # path's import then export with its halo refill
    dims = (4, 4, 2, 4)
    lat = lmx.Lattice(dims, (1, 1, 1, -1))
    core = _random(tuple(reversed(dims)) + (4, 3), np.complex128, 7)
    psi = lat.spinor_from_numpy(core.reshape(-1, 4, 3))
    C = lmx.lattice_matrix(lat, torch.from_numpy(core).to(gpu), 1)
    lat.synchronize()
    assert np.array_equal(C.cpu().numpy().reshape(-1), psi.to_padded(1).reshape(-1))


def test_field_argument_errors(gpu):
    lat = lmx.Lattice((4, 4, 4, 4), (1, 1, 1, 1j))
    real = torch.zeros((4, 4, 4, 4, 1, 2), dtype=torch.float64, device=gpu)
    with pytest.raises(lmx.ArgumentError):  # a real field cannot take the phase i
        lmx.lattice_matrix(lat, real, 1)
    assert lmx.lattice_matrix(lat, real, 0).shape == real.shape  # no halo, no phase
    with pytest.raises(lmx.ArgumentError):  # halo wider than the lattice
        lmx.lattice_matrix(lat, real, 5)
    with pytest.raises(lmx.ArgumentError):
        lmx.lattice_matrix(lat, real.to(torch.int64), 1)
    with pytest.raises(lmx.ArgumentError):  # wrong padded extents
        lmx.gather_matrix(lat, real, 1)
    lat2 = lmx.Lattice((4, 4, 4, 4), (1, 1, 1, 2.0))
    c64 = torch.zeros((4, 4, 4, 4, 3, 3), dtype=torch.complex64, device=gpu)
    with pytest.raises(lmx.ArgumentError):  # single precision takes +-1 only
        lmx.lattice_matrix(lat2, c64, 1)
    c128 = torch.ones((4, 4, 4, 4, 3, 3), dtype=torch.complex128, device=gpu)
    C = lmx.lattice_matrix(lat2, c128, 1)  # ComplexF64 takes any nonzero phase
    lat2.synchronize()
    c = C.cpu().numpy()
    assert np.all(c[-1, 1:-1, 1:-1, 1:-1] == 2.0) and np.all(c[0, 1:-1, 1:-1, 1:-1] == 0.5)
