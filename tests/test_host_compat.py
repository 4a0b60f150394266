"""Host logic of the compatibility interfaces (no device calls): the work-field
pool behind DiracOp.temps / phitemps and the legacy cg (PreallocatedArray's
get_block / unused!), and DdagDOp's one-argument form, which exists for
DiracOp only (src/Operators/DiracOperators.jl:93-98)."""
import pytest

import lmx


def test_field_pool_hands_out_reuses_and_grows():
    made = []

    def make():
        made.append(object())
        return made[-1]

    pool = lmx.FieldPool(make, 4)
    assert made == []  # fields are made on first use
    a, ia = pool.get_block()
    blk, ib = pool.get_block(3)
    assert len(made) == 4 and ia == 0 and ib == [1, 2, 3] and a is made[0] and blk == made[1:4]
    more, im = pool.get_block(2)  # every field in use: the pool grows
    assert len(pool) == 6 and im == [4, 5] and more == made[4:6]
    pool.unused_(ib)
    again, ig = pool.get_block(2)
    assert ig == [1, 2] and again == made[1:3] and len(made) == 6  # returned fields are reused
    pool.unused_(ia)
    b, i0 = pool.get_block()
    assert i0 == 0 and b is made[0]


def test_fixed_pool_does_not_grow():
    pool = lmx.FieldPool.of(["r", "p", "Ap", "spare"])
    blk, idx = pool.get_block(3)
    assert blk == ["r", "p", "Ap"]
    with pytest.raises(lmx.ArgumentError):
        pool.get_block(2)
    pool.unused_(idx)
    assert pool.get_block(4)[0] == ["r", "p", "Ap", "spare"]


def test_one_argument_ddagd_needs_a_dirac_op():
    with pytest.raises(TypeError):
        lmx.DdagDOp(lmx.ScaledOperator(2.0))
