// wilson.hip -- FP64 Wilson(-clover) apply (bit-exact with the reference's
// kernels) and the dispatch of the low-precision variants built in
// wilson_lo.hip.  The stencil itself is in wilson_kernels.h.
#include "wilson_kernels.h"

namespace lmx {

#define LMX_EXTERN(In, Out, Lk, Cl, NR) extern template int launch_lo_nr<In, Out, Lk, Cl, NR>(lmx_context*, const LoArgs&, Lk, Cl, bool);
LMX_LO_ALL(LMX_EXTERN)
#undef LMX_EXTERN

int launch_wilson_f64(lmx_context* ctx, const double2* U, const double2* clover, double kappa, double csw,
                      double2* out, const double2* in, int nrhs, int flags, const Epilogue& e) {
    if (ctx->V == 0 || nrhs == 0) return LMX_OK;
    const double ccoef = (-kappa) * csw;  // coefficient = -kappa * cSW (:572)
    const int V = (int)ctx->V;
    SpF64Out o{out, V};
    SpF64In i{in, V};
    LinkF64 l{U, V};
    F64Clover F{clover};
    EpiArgs epi{e.partials, 0, e.st, e.gate};
    const NoEx none{};
    if (e.kind == 1) {
        LMX_CHECK(nrhs == 1, LMX_ERR_UNSUPPORTED, "fused p.Ap needs a single right-hand side");
        SpF64In aux{e.aux, V};
        LMX_TRY((dispatch<SpF64In, LinkF64, SpF64Out, F64Clover, SpF64In, NoEx, 1, EPI_PAP, true, true>(
            ctx, o, i, l, F, aux, none, kappa, ccoef, 1, 1, flags, clover != nullptr, epi, 256)));
        return launch_finalize(ctx, nblocks(ctx->V), FIN_PAP, false, e.st, nullptr);
    }
    return dispatch<SpF64In, LinkF64, SpF64Out, F64Clover, SpF64In, NoEx, 1, EPI_NONE, true, true>(
        ctx, o, i, l, F, i, none, kappa, ccoef, nrhs, nrhs, flags, clover != nullptr, epi, 256);
}

// ---------------------------------------------------------------------------
// Low-precision variants for the mixed / multi-RHS solver: spinors FP32 or
// 16-bit (fixed point with a per-site FP32 scale), links FP32 or FP16, FP32
// arithmetic.
// ---------------------------------------------------------------------------
template <class In, class Out, class Lk, class Cl>
static int launch_lo_sp(lmx_context* ctx, const LoArgs& a, Lk U, Cl F, bool clover) {
    switch (a.rhs_block & 15) {
    case 0:
    case 1:
        return launch_lo_nr<In, Out, Lk, Cl, 1>(ctx, a, U, F, clover);
    case 2:
        return launch_lo_nr<In, Out, Lk, Cl, 2>(ctx, a, U, F, clover);
    case 4:
        return launch_lo_nr<In, Out, Lk, Cl, 4>(ctx, a, U, F, clover);
    default:
        set_error("unsupported rhs register block %d (built: 1, 2, 4)", a.rhs_block & 15);
        return LMX_ERR_UNSUPPORTED;
    }
}

int launch_wilson_lo(lmx_context* ctx, const LoArgs& a) {
    if (ctx->V == 0 || a.nrhs == 0) return LMX_OK;
    const bool cl = a.clover != nullptr;
    const int V = (int)ctx->V;
    if (a.link_prec == LMX_F32 && a.spinor_prec == LMX_F32)
        return launch_lo_sp<SpF32In, SpF32Out>(ctx, a, LinkF32{(const float2*)a.links, V},
                                               F32Clover{(const float2*)a.clover}, cl);
    if (a.link_prec == LMX_F16 && a.spinor_prec == LMX_F32)
        return launch_lo_sp<SpF32In, SpF32Out>(ctx, a, LinkF16{(const int4*)a.links, V},
                                               F16Clover{(const __half2*)a.clover}, cl);
    if (a.link_prec == LMX_F16 && a.spinor_prec == LMX_H16)
        return launch_lo_sp<SpH16In, SpH16Out>(ctx, a, LinkF16{(const int4*)a.links, V},
                                               F16Clover{(const __half2*)a.clover}, cl);
    set_error("unsupported precision pair (spinor %d, links %d)", a.spinor_prec, a.link_prec);
    return LMX_ERR_UNSUPPORTED;
}

}  // namespace lmx
