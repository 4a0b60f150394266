"""
    LatticeMatricesLmx

Drop-in MI355X backend for the Wilson(-clover) hot path of LatticeMatrices.jl:
methods on the reference's own types that call liblmx (include/lmx.h, gfx950
HIP kernels) through `ccall`, with the fields left in the reference's storage
(`LatticeMatrix.A` on `ROCArray`s, halo-padded, column-major).

Methods added (reference interface replaced, file:line in LatticeMatrices.jl):
  mul!(C, D::WilsonDiracOperator4D, ψ)            src/Operators/WilsonDiracOperator.jl:60-76
  mul!(C, D::Adjoint_WilsonDiracOperator4D, ψ)    :425-441
  mul!(C, D::WilsonDiracOperator4D_Donly, ψ) and its adjoint :535-551,646-662
  mul!(C, D::WilsonDiracCloverOperator4D, ψ) and its adjoint
                                                  src/Operators/WilsonDiracCloverOperator.jl:561-601
  cg!(x, DdagDOp(D, temp), b, r, p, Ap; rtol, atol, maxiter)
                                                  src/LinearAlgebras/cg.jl:73-137
  EnzymeRules.augmented_primal / reverse of mul!(result, D::WilsonDiracOperator4D, ψ)
                                                  ext/AD/WilsonDiracOperator.jl:93-170
  dSFdU(dfdU, D::WilsonDiracOperator4D or WilsonDiracCloverOperator4D, φ)
                                                  ext/LatticeMatricesEnzymeExt.jl:395-460 (for DiracOp)
ComplexF64 and ComplexF32 fields with NC = 3, and ComplexF64 fields with
NC = 2 (the reference's generic-NC kernels), on one rank.  The library keeps,
per operator object, its links converted to its own layout; they are
re-imported whenever a link field's core epoch (`halo_epoch.core`, advanced by
`mark_halo_dirty!` and by every public mutating operation) differs from the
one they were imported at, as the clover cache's `CloverCacheState` does
(src/Operators/WilsonDiracCloverOperator.jl:417-432).

`tests/test_gpu_julia_replay.py` in this repository replays the exact C-ABI
call sequence of every method here (ctypes, same arguments in the same order)
against the CPU oracle.
"""
module LatticeMatricesLmx

using LinearAlgebra
using AMDGPU
using LatticeMatrices
using LatticeMatrices: LatticeMatrix, WilsonDiracOperator4D, Adjoint_WilsonDiracOperator4D,
    WilsonDiracOperator4D_Donly, Adjoint_WilsonDiracOperator4D_Donly, WilsonDiracCloverOperator4D,
    Adjoint_WilsonDiracCloverOperator4D, DdagDOp, CGResult, mark_halo_dirty!, _mark_halo_clean!,
    _validate_cg_arguments, add_matrix!

const liblmx = get(ENV, "LMX_LIBRARY", joinpath(@__DIR__, "..", "latticematrices.jl_amd", "lmx", "liblmx.so"))

# ---------------------------------------------------------------------------
# Status codes and the C structs of include/lmx.h
# ---------------------------------------------------------------------------
const LMX_OK, LMX_ERR_ARG, LMX_ERR_ALIAS = Cint(0), Cint(1), Cint(2)
const LMX_F64 = Cint(0)
const LMX_DAGGER, LMX_DONLY = Cint(1), Cint(2)
const REASONS = (:converged, :maximum_iterations, :nonpositive_curvature, :numerical_breakdown)

struct LmxError <: Exception
    code::Cint
    msg::String
end

last_error() = unsafe_string(ccall((:lmx_last_error, liblmx), Cstring, ()))

# ArgumentError for what the reference throws as one (bad geometry,
# aliasing, tolerances); everything else as LmxError.
function check(rc::Cint)
    rc == LMX_OK && return nothing
    msg = last_error()
    throw(rc in (LMX_ERR_ARG, LMX_ERR_ALIAS) ? ArgumentError(msg) : LmxError(rc, msg))
end

struct LmxCGParams          # lmx_cg_params
    rtol::Float64
    atol::Float64
    maxiter::Int32
    precision::Int32
    reliable_delta::Float64
    reliable_maxiter::Int32
    check_every::Int32
    use_graph::Int32
    rhs_block::Int32
    reserved::NTuple{4,Int32}
end
LmxCGParams(rtol, atol, maxiter) =
    LmxCGParams(Float64(rtol), Float64(atol), Int32(maxiter), Int32(0), 0.1, Int32(100), Int32(0), Int32(0),
        Int32(0), (Int32(0), Int32(0), Int32(0), Int32(0)))

struct LmxCGResult          # lmx_cg_result
    converged::Int32
    iterations::Int32
    reason::Int32
    reliable_updates::Int32
    residual_norm::Float64
    relative_residual::Float64
    true_relative_residual::Float64
end

# ---------------------------------------------------------------------------
# Contexts: one per (lattice size, spinor phases, NC, HIP stream)
# ---------------------------------------------------------------------------
mutable struct LmxContext
    ptr::Ptr{Cvoid}
    nc::Int
end

const CONTEXTS = Dict{Any,LmxContext}()

function context(ψ::LatticeMatrix{4,T,AT,NC1}) where {T,AT,NC1}
    stream = AMDGPU.stream()
    phases = ComplexF64.(Tuple(ψ.phases))
    key = (ψ.gsize, phases, NC1, stream.stream)
    get!(CONTEXTS, key) do
        prod(ψ.dims) == 1 || throw(ArgumentError("liblmx runs a lattice on one rank (process grid $(ψ.dims))"))
        ph = Float64[x for p in phases for x in (real(p), imag(p))]
        h = Ref{Ptr{Cvoid}}(C_NULL)
        check(ccall((:lmx_context_create_nc, liblmx), Cint,
            (Ref{Ptr{Cvoid}}, Ptr{Int32}, Ptr{Float64}, Ptr{Cvoid}, Cint),
            h, Int32[ψ.gsize...], ph, stream.stream, NC1))
        ctx = LmxContext(h[], NC1)
        finalizer(c -> ccall((:lmx_context_destroy, liblmx), Cint, (Ptr{Cvoid},), c.ptr), ctx)
        ctx
    end
end

spinor_bytes(ctx::LmxContext, n) =
    ccall((:lmx_spinor_bytes, liblmx), Csize_t, (Ptr{Cvoid}, Cint, Cint), ctx.ptr, LMX_F64, n)

# ---------------------------------------------------------------------------
# Operator handles: the links in the library's layout plus lmx_wilson,
# cached per reference operator object and refreshed on a link write
# ---------------------------------------------------------------------------
mutable struct LmxOperator
    ptr::Ptr{Cvoid}
    ctx::LmxContext
    links::ROCVector{Float64}
    epochs::NTuple{4,UInt64}
    kappa::Float64
    work::ROCVector{UInt8}       # six library-layout spinors (cg!, the reverse rule, dSFdU)
end

const OPERATORS = WeakKeyDict{Any,LmxOperator}()

_links(D::WilsonDiracOperator4D) = D.U
_links(D::WilsonDiracOperator4D_Donly) = D.U
_links(D::WilsonDiracCloverOperator4D) = D.wilson.U
_kappa(D::WilsonDiracOperator4D) = D.κ
_kappa(D::WilsonDiracOperator4D_Donly) = 0.0
_kappa(D::WilsonDiracCloverOperator4D) = D.wilson.κ
_csw(D) = 0.0
_csw(D::WilsonDiracCloverOperator4D) = D.cSW

function import_links!(ctx, dst::ROCVector{Float64}, U)
    check(ccall((:lmx_links_import, liblmx), Cint, (Ptr{Cvoid}, Ptr{Cvoid}, Ptr{Ptr{Cvoid}}, Cint),
        ctx.ptr, pointer(dst), Ptr{Cvoid}[pointer(u.A) for u in U], U[1].nw))
end

function handle(D, ψ::LatticeMatrix)
    U = _links(D)
    epochs = ntuple(mu -> U[mu].halo_epoch.core, 4)
    op = get(OPERATORS, D, nothing)
    if op === nothing
        ctx = context(ψ)
        nc = ctx.nc
        links = AMDGPU.zeros(Float64, 2 * 4 * nc * nc * prod(ψ.gsize))
        import_links!(ctx, links, U)
        h = Ref{Ptr{Cvoid}}(C_NULL)
        check(ccall((:lmx_wilson_create, liblmx), Cint, (Ptr{Cvoid}, Ref{Ptr{Cvoid}}, Ptr{Cvoid}, Float64, Float64),
            ctx.ptr, h, pointer(links), _kappa(D), _csw(D)))
        work = AMDGPU.zeros(UInt8, Int(spinor_bytes(ctx, 6)))
        op = LmxOperator(h[], ctx, links, epochs, _kappa(D), work)
        finalizer(o -> ccall((:lmx_wilson_destroy, liblmx), Cint, (Ptr{Cvoid},), o.ptr), op)
        OPERATORS[D] = op
    elseif op.epochs != epochs            # links written since: re-import, rebuild the clover cache
        import_links!(op.ctx, op.links, U)
        check(ccall((:lmx_wilson_update_links, liblmx), Cint, (Ptr{Cvoid}, Ptr{Cvoid}), op.ptr, pointer(op.links)))
        op.epochs = epochs
    end
    if op.kappa != _kappa(D)              # the reference keeps kappa in a Ref that callers may set
        check(ccall((:lmx_wilson_set_kappa, liblmx), Cint, (Ptr{Cvoid}, Float64), op.ptr, _kappa(D)))
        op.kappa = _kappa(D)
    end
    return op
end

# ---------------------------------------------------------------------------
# mul! on the reference's storage
# ---------------------------------------------------------------------------
const LM64{NC,nw} = LatticeMatrix{4,ComplexF64,<:ROCArray,NC,4,nw}
const LM32{nw} = LatticeMatrix{4,ComplexF32,<:ROCArray,3,4,nw}

_parent_flags(D::WilsonDiracOperator4D) = (D, Cint(0))
_parent_flags(D::Adjoint_WilsonDiracOperator4D) = (D.parent, LMX_DAGGER)
_parent_flags(D::WilsonDiracOperator4D_Donly) = (D, LMX_DONLY)
_parent_flags(D::Adjoint_WilsonDiracOperator4D_Donly) = (D.parent, LMX_DONLY | LMX_DAGGER)
_parent_flags(D::WilsonDiracCloverOperator4D) = (D, Cint(0))
_parent_flags(D::Adjoint_WilsonDiracCloverOperator4D) = (D.parent, LMX_DAGGER)

const WilsonLike = Union{WilsonDiracOperator4D,Adjoint_WilsonDiracOperator4D,WilsonDiracOperator4D_Donly,
    Adjoint_WilsonDiracOperator4D_Donly,WilsonDiracCloverOperator4D,Adjoint_WilsonDiracCloverOperator4D}

function apply_padded!(C::LatticeMatrix, D, ψ::LatticeMatrix, fn::Symbol)
    C === ψ && throw(ArgumentError("output and input of the Wilson operator must not alias"))
    parent, flags = _parent_flags(D)
    op = handle(parent, ψ)
    if fn === :lmx_wilson_apply_padded
        check(ccall((:lmx_wilson_apply_padded, liblmx), Cint, (Ptr{Cvoid}, Ptr{Cvoid}, Ptr{Cvoid}, Cint, Cint, Cint),
            op.ptr, pointer(C.A), pointer(ψ.A), ψ.nw, flags, 1))
    else
        check(ccall((:lmx_wilson_apply_padded_f32, liblmx), Cint,
            (Ptr{Cvoid}, Ptr{Cvoid}, Ptr{Cvoid}, Cint, Cint, Cint),
            op.ptr, pointer(C.A), pointer(ψ.A), ψ.nw, flags, 1))
    end
    mark_halo_dirty!(C)                   # only C's core was written, as after the reference's mul!
    return C
end

LinearAlgebra.mul!(C::LM64{NC,nw}, D::WilsonLike, ψ::LM64{NC,nw}) where {NC,nw} =
    apply_padded!(C, D, ψ, :lmx_wilson_apply_padded)
LinearAlgebra.mul!(C::LM32{nw}, D::WilsonLike, ψ::LM32{nw}) where {nw} =
    apply_padded!(C, D, ψ, :lmx_wilson_apply_padded_f32)

# ---------------------------------------------------------------------------
# cg! of D^dagger D on LatticeMatrix fields
# ---------------------------------------------------------------------------
const WilsonOrClover = Union{WilsonDiracOperator4D,WilsonDiracCloverOperator4D}

function LatticeMatrices.cg!(x::LM64{NC,nw}, A::DdagDOp{<:WilsonOrClover}, b::LM64{NC,nw}, r::LM64{NC,nw},
    p::LM64{NC,nw}, Ap::LM64{NC,nw}; rtol=1e-10, atol=0, maxiter::Integer=5000) where {NC,nw}
    _validate_cg_arguments(x, A, b, r, p, Ap, rtol, atol, maxiter)
    op = handle(A.D, b)
    res = Ref{LmxCGResult}()
    check(ccall((:lmx_cg_ddagd_padded, liblmx), Cint,
        (Ptr{Cvoid}, Ptr{Cvoid}, Ptr{Cvoid}, Cint, Ptr{Cvoid}, Csize_t, Ref{LmxCGParams}, Ref{LmxCGResult}),
        op.ptr, pointer(x.A), pointer(b.A), nw, pointer(op.work), sizeof(op.work),
        LmxCGParams(rtol, atol, maxiter), res))
    mark_halo_dirty!(x)
    _mark_halo_clean!(x)                  # x's halo was refilled with its phases
    q = res[]
    return CGResult(q.converged != 0, Int(q.iterations), q.residual_norm, q.relative_residual, REASONS[q.reason + 1])
end

# cg! on LatticeMatrix{ComplexF32} (cg! is generic in the element type):
# FP32 operator and vectors, FP64 reductions, the same workspace
function LatticeMatrices.cg!(x::LM32{nw}, A::DdagDOp{<:WilsonOrClover}, b::LM32{nw}, r::LM32{nw},
    p::LM32{nw}, Ap::LM32{nw}; rtol=1e-10, atol=0, maxiter::Integer=5000) where {nw}
    _validate_cg_arguments(x, A, b, r, p, Ap, rtol, atol, maxiter)
    op = handle(A.D, b)
    res = Ref{LmxCGResult}()
    check(ccall((:lmx_cg_ddagd_padded_f32, liblmx), Cint,
        (Ptr{Cvoid}, Ptr{Cvoid}, Ptr{Cvoid}, Cint, Ptr{Cvoid}, Csize_t, Ref{LmxCGParams}, Ref{LmxCGResult}),
        op.ptr, pointer(x.A), pointer(b.A), nw, pointer(op.work), sizeof(op.work),
        LmxCGParams(rtol, atol, maxiter), res))
    mark_halo_dirty!(x)
    _mark_halo_clean!(x)                  # x's halo was refilled with its phases
    q = res[]
    return CGResult(q.converged != 0, Int(q.iterations), q.residual_norm, q.relative_residual, REASONS[q.reason + 1])
end

# ---------------------------------------------------------------------------
# Link cotangent and force: the library's own layout in op.work
# ---------------------------------------------------------------------------
function import_spinor!(op::LmxOperator, slot::Int, ψ::LatticeMatrix)
    dst = pointer(op.work) + slot * spinor_bytes(op.ctx, 1)
    check(ccall((:lmx_spinor_import, liblmx), Cint, (Ptr{Cvoid}, Ptr{Cvoid}, Ptr{Cvoid}, Cint, Cint),
        op.ctx.ptr, dst, pointer(ψ.A), ψ.nw, 1))
    return dst
end

# dU[mu] += link cotangent of mul!(result, D, psi) for the output cotangent dresult:
# the four padded link arrays go to the library layout, the pullback adds into
# them there, and they come back with the halo refilled (periodic).
function add_link_pullback!(dU, D, dresult::LatticeMatrix, psi::LatticeMatrix, coefficient)
    op = handle(D, psi)
    nc = op.ctx.nc
    G = AMDGPU.zeros(Float64, 2 * 4 * nc * nc * prod(psi.gsize))
    import_links!(op.ctx, G, dU)
    pd = import_spinor!(op, 0, dresult)
    pp = import_spinor!(op, 1, psi)
    check(ccall((:lmx_wilson_pullback, liblmx), Cint, (Ptr{Cvoid}, Ptr{Cvoid}, Ptr{Cvoid}, Ptr{Cvoid}, Float64),
        op.ptr, pointer(G), pd, pp, coefficient))
    check(ccall((:lmx_links_export, liblmx), Cint, (Ptr{Cvoid}, Ptr{Ptr{Cvoid}}, Cint, Ptr{Cvoid}),
        op.ctx.ptr, Ptr{Cvoid}[pointer(u.A) for u in dU], dU[1].nw, pointer(G)))
    mark_halo_dirty!.(dU)
    return dU
end

"""
    dSFdU(dfdU, D, φ; rtol=1e-10, maxiter=5000)

dfdU[mu] <- dS/dU_mu of S = φ† (D† D)⁻¹ φ in the Enzyme convention
(dS/dRe U + i dS/dIm U) for a Wilson or Wilson-clover operator: η = (D†D)⁻¹φ
by FP64 CG from zero, χ = D η, and the link pullback of dresult = -2χ with
coefficient -κ -- the computation the reference's dSFdU does by Enzyme for a
DiracOp (ext/LatticeMatricesEnzymeExt.jl:395-460), with the hand-written
adjoint and no stored Krylov history.  Returns the CGResult of the solve.
"""
function dSFdU(dfdU, D::WilsonOrClover, φ::LM64{NC,nw}; rtol=1e-10, maxiter::Integer=5000) where {NC,nw}
    op = handle(D, φ)
    nc = op.ctx.nc
    n1 = spinor_bytes(op.ctx, 1)
    G = AMDGPU.zeros(Float64, 2 * 4 * nc * nc * prod(φ.gsize))
    work = AMDGPU.zeros(UInt8, Int(4 * n1))
    pφ = import_spinor!(op, 0, φ)
    η = pointer(op.work) + n1
    χ = pointer(op.work) + 2 * n1
    res = Ref{LmxCGResult}()
    check(ccall((:lmx_wilson_force, liblmx), Cint,
        (Ptr{Cvoid}, Ptr{Cvoid}, Ptr{Cvoid}, Ptr{Cvoid}, Ptr{Cvoid}, Ptr{Cvoid}, Ptr{Cvoid}, Ref{LmxCGParams},
            Ref{LmxCGResult}),
        op.ptr, pointer(G), C_NULL, pφ, η, χ, pointer(work), LmxCGParams(rtol, 0.0, maxiter), res))
    check(ccall((:lmx_links_export, liblmx), Cint, (Ptr{Cvoid}, Ptr{Ptr{Cvoid}}, Cint, Ptr{Cvoid}),
        op.ctx.ptr, Ptr{Cvoid}[pointer(u.A) for u in dfdU], dfdU[1].nw, pointer(G)))
    mark_halo_dirty!.(dfdU)
    q = res[]
    return CGResult(q.converged != 0, Int(q.iterations), q.residual_norm, q.relative_residual, REASONS[q.reason + 1])
end

# ---------------------------------------------------------------------------
# Enzyme rules of mul!(result, D::WilsonDiracOperator4D, psi) on ROCArray
# fields (the reference's rule, ext/AD/WilsonDiracOperator.jl:93-170, with
# the pullback and D^dagger on the device): dU += pullback(dresult, psi; -kappa),
# dpsi += D^dagger dresult, then dresult is cleared.
# ---------------------------------------------------------------------------
import Enzyme
const ER = Enzyme.EnzymeRules

_shadow(x) = x isa Base.RefValue ? x[] : x

function ER.augmented_primal(cfg::ER.RevConfig, ::ER.Const{typeof(LinearAlgebra.mul!)}, ::Type{RT},
    result::ER.Annotation{<:LM64}, operator::ER.Annotation{<:WilsonDiracOperator4D},
    psi::ER.Annotation{<:LM64}) where {RT}
    primal_return = LinearAlgebra.mul!(result.val, operator.val, psi.val)
    primal = ER.needs_primal(cfg) ? primal_return : nothing
    shadow = ER.needs_shadow(cfg) ? _shadow(result.dval) : nothing
    return ER.augmented_rule_return_type(cfg, RT, nothing)(primal, shadow, nothing)
end

function ER.reverse(cfg::ER.RevConfig, ::ER.Const{typeof(LinearAlgebra.mul!)}, dresult_out, _tape,
    result::ER.Annotation{<:LM64}, operator::ER.Annotation{<:WilsonDiracOperator4D},
    psi::ER.Annotation{<:LM64})
    dresult = hasproperty(result, :dval) ? _shadow(result.dval) : nothing
    dresult isa LatticeMatrix || return (nothing, nothing, nothing)
    D = operator.val
    if hasproperty(operator, :dval) && _shadow(operator.dval) isa WilsonDiracOperator4D
        add_link_pullback!(_shadow(operator.dval).U, D, dresult, psi.val, -D.κ)
    end
    dpsi = hasproperty(psi, :dval) ? _shadow(psi.dval) : nothing
    if dpsi isa LatticeMatrix
        temporary = similar(psi.val)
        LinearAlgebra.mul!(temporary, D', dresult)    # D^dagger dresult on the device
        add_matrix!(dpsi, temporary)
    end
    fill!(dresult.A, zero(eltype(dresult.A)))
    mark_halo_dirty!(dresult)
    return (nothing, nothing, nothing)
end

export dSFdU

end # module
