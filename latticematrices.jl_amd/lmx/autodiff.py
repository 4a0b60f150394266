"""The pseudofermion action as a differentiable function with a hand-written
reverse rule.

S(U, phi) = Re phi^dagger (D(U)^dagger D(U))^-1 phi for the Wilson or
Wilson-clover operator, as the reference differentiates it with its Enzyme
rules (dSFdU, ext/LatticeMatricesEnzymeExt.jl:387-460; the Wilson link
pullback, ext/AD/WilsonDiracOperator.jl:93-170; the clover pullback,
src/Operators/WilsonCloverPullback.jl:342-389).  The reverse rule does not
record the CG: the forward pass solves for eta = (D^dagger D)^-1 phi once
and, when the links need a gradient, applies the hand-written link pullback
to (eta, chi = D eta) on the device (lmx_wilson_force).  Gradients are with
respect to the real storage of the tensors, (re, im) pairs:

    dS/dU   = G = dS/dRe U + i dS/dIm U  (the Enzyme convention; its (re, im)
              pairs are exactly the gradient of the links tensor)
    dS/dphi = 2 eta                      (S is Hermitian-quadratic in phi)
"""
from __future__ import annotations

import torch

from .lattice import (GaugeField, SpinorField, WilsonDiracOperator4D, _ensure_clover_cache_current, adjoint,
                      dSFdU, dot, mul_, mul_cached_clover_, wilson_pullback_)


class PseudofermionAction(torch.autograd.Function):
    """S = PseudofermionAction.apply(links, phi, lattice, kappa, csw, rtol, maxiter)

    links: GaugeField.data (float64 [4, 9, V, 2]); phi: SpinorField.data of
    one field (float64 [1, 12, V, 2]).  Returns S as a 0-dim float64 tensor
    on the device.  Convergence failures of the solve raise RuntimeError."""

    @staticmethod
    def forward(ctx, links, phi, lattice, kappa, csw=0.0, rtol=1e-12, maxiter=20000):
        U = GaugeField(lattice, links.detach())
        D = WilsonDiracOperator4D(U, float(kappa), float(csw))
        p = SpinorField(lattice, 1, phi.detach())
        G = GaugeField(lattice)
        res, eta, _chi = dSFdU(G, D, p, rtol=rtol, maxiter=maxiter)
        if not res.converged:
            raise RuntimeError(f"pseudofermion action: CG did not converge ({res.reason}, {res.iterations} iterations)")
        S = dot(p, eta).real
        ctx.save_for_backward(G.data, eta.data)
        return torch.tensor(S, dtype=torch.float64, device=links.device)

    @staticmethod
    def backward(ctx, gS):
        G, eta = ctx.saved_tensors
        gl = gS * G if ctx.needs_input_grad[0] else None
        gp = 2.0 * gS * eta if ctx.needs_input_grad[1] else None
        return gl, gp, None, None, None, None, None


def pseudofermion_action_ad(U: GaugeField, phi: SpinorField, kappa: float, csw: float = 0.0, rtol: float = 1e-12,
                            maxiter: int = 20000):
    """S(U, phi) with autograd through the hand-written adjoint; U.data and
    phi.data may require grad."""
    return PseudofermionAction.apply(U.data, phi.data, U.lattice, kappa, csw, rtol, maxiter)


class MulCachedClover(torch.autograd.Function):
    """result = MulCachedClover.apply(links, psi, cache, U)

    mul_cached_clover!(result, cache, U1..U4, psi) as a differentiable
    function, with the reference's reverse rule
    (ext/AD/WilsonDiracCloverOperator.jl:348-444): the link cotangent is the
    analytic Wilson + clover pullback (wilson_clover_link_pullback!,
    src/Operators/WilsonCloverPullback.jl:328-389, coefficient -kappa on the
    device), the spinor cotangent D^dagger dresult.  links must be U.data
    (float64 [4, 9, V, 2]), psi a SpinorField's data (float64 [n, 12, V, 2]);
    the result is a new float64 [n, 12, V, 2] tensor.  Gradients are with
    respect to the (re, im) storage, the Enzyme convention dL/dRe + i dL/dIm."""

    @staticmethod
    def forward(ctx, links, psi, cache, U):
        if links.data_ptr() != U.data.data_ptr():
            raise ValueError("links must be the GaugeField's data tensor")
        L = cache.lattice
        x = SpinorField(L, psi.shape[0], psi.detach())
        out = SpinorField(L, psi.shape[0])
        mul_cached_clover_(out, cache, U, x)
        ctx.cache, ctx.U = cache, U
        ctx.save_for_backward(links, psi)
        return out.data

    @staticmethod
    def backward(ctx, g):
        links, psi = ctx.saved_tensors  # raises if either was written in place since the forward pass
        cache, U = ctx.cache, ctx.U
        _ensure_clover_cache_current(cache, U)
        L = cache.lattice
        n = psi.shape[0]
        dres = SpinorField(L, n, g.detach().to(torch.float64).contiguous())
        gl = gp = None
        if ctx.needs_input_grad[0]:
            dU = GaugeField(L)
            x = SpinorField(L, n, psi.detach())
            for i in range(n):
                wilson_pullback_(dU, cache, dres.rhs(i), x.rhs(i), -cache.kappa)
            gl = dU.data
        if ctx.needs_input_grad[1]:
            dpsi = SpinorField(L, n)
            mul_(dpsi, adjoint(cache), dres)
            gp = dpsi.data
        return gl, gp, None, None


def mul_cached_clover_ad(cache: WilsonDiracOperator4D, U: GaugeField, psi: SpinorField):
    """D(U) psi through mul_cached_clover_ with autograd; U.data and psi.data
    may require grad.  Returns the result tensor [n, 12, V, 2]."""
    return MulCachedClover.apply(U.data, psi.data, cache, U)


def wilson_apply_ad(D: WilsonDiracOperator4D, psi: SpinorField):
    """D psi (mul!(result, D, psi)) for a Wilson or Wilson-clover operator
    with autograd through its links (D.U.data) and psi.data: the reference's
    mul! reverse rules (ext/AD/WilsonDiracOperator.jl:93-170,
    ext/AD/WilsonDiracCloverOperator.jl:266-340) -- the analytic link pullback
    and dpsi = D^dagger dresult -- through MulCachedClover with the operator's
    own links (refreshing its derived copies first if they were written)."""
    return MulCachedClover.apply(D.U.data, psi.data, D, D.U)
