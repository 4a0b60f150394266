"""Host logic of mul_cached_clover!'s refresh decision (no device calls):
_clover_cache_is_current / _ensure_clover_cache_current!
(src/Operators/WilsonDiracCloverOperator.jl:603-623) compare the links object
and its core epoch with the ones the cache was built from.  Here the epoch is
(mark_halo_dirty_ count, torch's in-place version counter, device address),
and the refresh itself is replaced by a recorder."""
import types

import pytest
import torch

import lmx
from lmx.lattice import _ensure_clover_cache_current


def _stub_operator(U):
    lat = U.lattice
    op = object.__new__(lmx.WilsonDiracOperator4D)
    op.lattice, op.U, op.cache_epoch = lat, U, 0
    op._source = (U, U.core_epoch)
    op.refreshed = []

    def update_links(U2=None):  # what update_links records, without the library call
        if U2 is not None:
            op.U = U2
        op.refreshed.append(op.U)
        op.cache_epoch += 1
        op._source = (op.U, op.U.core_epoch)

    op.update_links = update_links
    return op


def test_refresh_decision():
    lat = types.SimpleNamespace(V=16, dims=(2, 2, 2, 2), device=torch.device("cpu"), nc=3, ncomp=12)
    U = lmx.GaugeField(lat)
    op = _stub_operator(U)
    _ensure_clover_cache_current(op, U)
    assert op.refreshed == []  # built from these links, unchanged
    U.data[1, 3, 5, 0] += 0.5  # in-place torch write (add_matrix!-like)
    _ensure_clover_cache_current(op, U)
    assert op.refreshed == [U]
    _ensure_clover_cache_current(op, U)
    assert len(op.refreshed) == 1  # current again
    U.data.data.mul_(2.0)  # a write torch does not count ...
    _ensure_clover_cache_current(op, U)
    assert len(op.refreshed) == 1
    U.mark_halo_dirty_()  # ... is declared as in the reference
    _ensure_clover_cache_current(op, U)
    assert len(op.refreshed) == 2
    U.data = U.data.clone()  # new storage
    _ensure_clover_cache_current(op, U)
    assert len(op.refreshed) == 3
    U2 = lmx.GaugeField(lat, U.data)  # same storage, another links object
    _ensure_clover_cache_current(op, U2)
    assert op.refreshed[-1] is U2 and op.U is U2
    with pytest.raises(TypeError):
        _ensure_clover_cache_current(op, U2.data)
    with pytest.raises(TypeError):
        _ensure_clover_cache_current(object(), U2)
