"""lmx: MI355X-native Wilson(-clover) D_W / D^dagger D CG path of LatticeMatrices.jl.

The compute path is liblmx.so (hand-written gfx950 HIP kernels behind the C
ABI of include/lmx.h); this package is its host-side mirror of the
reference's LatticeMatrix operator API.
"""
from . import _lib, generators
from ._lib import ArgumentError, LmxError
from .lattice import (Adjoint, add_matrix_, cg, CGResult, CloverFieldStrength4D, DdagDOp, DiracOp, FieldPool, GaugeField, Lattice, ScaledOperator, SpinorField,
                      WilsonDiracCloverOperator4D, WilsonDiracOperator4D, WilsonDiracOperator4D_Donly, adjoint,
                      axpby_, cg_, clear_matrix_, dot, dSFdU, mul_, mul_padded_, pseudofermion_action, solve_, substitute_,
                      ta_project_, wilson_pullback_, cg_batched_, autotune, inner_kernel_times, mul_cached_clover_,
                      mul_cached_clover_adjoint_, update_clover_, spinor_f32,
                      spinor_f32_from_numpy, spinor_f32_to_numpy, mul_f32_)
from .autodiff import (MulCachedClover, PseudofermionAction, mul_cached_clover_ad, pseudofermion_action_ad,
                       wilson_apply_ad)
from .fields import gather_matrix, lattice_matrix
from .rng import (DEFAULT_RANDOMIZE_SEED, RNGStreamKey, randomize_gaussian_matrix_, randomize_matrix_,
                  site_rng_draw)

__all__ = [
    "ArgumentError", "LmxError", "Adjoint", "cg", "CGResult", "CloverFieldStrength4D", "DdagDOp", "DiracOp", "FieldPool", "GaugeField", "Lattice", "ScaledOperator",
    "SpinorField", "WilsonDiracCloverOperator4D", "WilsonDiracOperator4D", "WilsonDiracOperator4D_Donly",
    "adjoint", "add_matrix_", "axpby_", "cg_", "clear_matrix_", "dot", "dSFdU", "mul_", "mul_padded_", "pseudofermion_action", "solve_",
    "substitute_", "ta_project_", "wilson_pullback_", "generators", "cg_batched_", "autotune", "inner_kernel_times",
    "mul_cached_clover_", "mul_cached_clover_adjoint_", "update_clover_", "spinor_f32",
    "spinor_f32_from_numpy", "spinor_f32_to_numpy", "mul_f32_", "MulCachedClover", "mul_cached_clover_ad", "wilson_apply_ad",
    "PseudofermionAction", "pseudofermion_action_ad", "gather_matrix", "lattice_matrix", "DEFAULT_RANDOMIZE_SEED", "RNGStreamKey", "randomize_gaussian_matrix_", "randomize_matrix_", "site_rng_draw",
]
