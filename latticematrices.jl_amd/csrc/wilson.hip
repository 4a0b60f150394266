// wilson.hip -- FP64 Wilson(-clover) apply (bit-exact with the reference's
// kernels) and the dispatch of the low-precision variants built in
// wilson_lo.hip.  The stencil itself is in wilson_kernels.h.
#include "wilson_kernels.h"

namespace lmx {

#define LMX_EXTERN(In, Out, Lk, Cl, NR) extern template int launch_lo_nr<In, Out, Lk, Cl, NR>(lmx_context*, const LoArgs&, Lk, Cl, bool);
LMX_LO_ALL(LMX_EXTERN)
#undef LMX_EXTERN
extern template int launch_lo_loop<SpF32In, SpF32Out, F16Clover>(lmx_context*, const LoArgs&, LinkF16, F16Clover, bool);
extern template int launch_lo_loop<SpH16In, SpH16Out, F16Clover>(lmx_context*, const LoArgs&, LinkF16, F16Clover, bool);
#define LMX_EXTERN_LDS(In, Out, Lk, Cl, NR) extern template int launch_lo_lds<In, Out, Lk, Cl>(lmx_context*, const LoArgs&, Lk, Cl, bool);
LMX_LO_COMBOS(LMX_EXTERN_LDS, 1)
#undef LMX_EXTERN_LDS

int launch_wilson_f64(lmx_context* ctx, const double2* U, const double2* clover, double kappa, double csw,
                      double2* out, const double2* in, int nrhs, int flags, const Epilogue& e) {
    if (ctx->V == 0 || nrhs == 0) return LMX_OK;
    if (ctx->nc != 3) {  // the reference's generic-NC kernels (wilson_nc.hip)
        LMX_CHECK(e.kind == 0, LMX_ERR_UNSUPPORTED, "fused epilogues are built for NC = 3");
        return launch_wilson_nc(ctx, U, clover, kappa, csw, out, in, nrhs, flags, -1, e.st, e.gate);
    }
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
    // several Wilson RHS: two per work item share each link load (the same
    // per-RHS arithmetic, so bit-identical to one per work item); the clover
    // variant would spill at two, so it keeps one
    if (nrhs >= 2 && clover == nullptr)
        return dispatch<SpF64In, LinkF64, SpF64Out, F64Clover, SpF64In, NoEx, 2, EPI_NONE, true, true>(
            ctx, o, i, l, F, i, none, kappa, ccoef, (nrhs + 1) / 2, nrhs, flags, false, epi, 256);
    return dispatch<SpF64In, LinkF64, SpF64Out, F64Clover, SpF64In, NoEx, 1, EPI_NONE, true, true>(
        ctx, o, i, l, F, i, none, kappa, ccoef, nrhs, nrhs, flags, clover != nullptr, epi, 256);
}

template <class SO>
static int wilson_f64_true(lmx_context* ctx, const double2* U, const double2* clover, double kappa, double csw,
                           const double2* in, const double2* b, SO rout, const int* done, const int* fold,
                           int dstride, int force, double* partials, long pstride, int nrhs, const int* gate) {
    const double ccoef = (-kappa) * csw;
    const int V = (int)ctx->V;
    SpF64Out o{nullptr, V};  // D^dagger in itself is not stored
    SpF64In i{in, V};
    LinkF64 l{U, V};
    F64Clover F{clover};
    EpiArgs epi{partials, pstride, nullptr, gate};
    TrueEx<SO> ex{SpF64In{b, V}, rout, done, fold, dstride, force};
    if (nrhs >= 2 && clover == nullptr)  // as launch_wilson_f64: two Wilson RHS per work item
        return dispatch<SpF64In, LinkF64, SpF64Out, F64Clover, SpF64In, TrueEx<SO>, 2, EPI_TRUE, true, true>(
            ctx, o, i, l, F, i, ex, kappa, ccoef, (nrhs + 1) / 2, nrhs, LMX_DAGGER, false, epi, 256);
    return dispatch<SpF64In, LinkF64, SpF64Out, F64Clover, SpF64In, TrueEx<SO>, 1, EPI_TRUE, true, true>(
        ctx, o, i, l, F, i, ex, kappa, ccoef, nrhs, nrhs, LMX_DAGGER, clover != nullptr, epi, 256);
}

int launch_wilson_f64_true(lmx_context* ctx, const double2* U, const double2* clover, double kappa, double csw,
                           const double2* in, const double2* b, void* r, float* r_norm, int spinor_prec,
                           const int* done, const int* fold, int dstride, int force, double* partials, long pstride,
                           int nrhs, const int* gate) {
    if (ctx->V == 0 || nrhs == 0) return LMX_OK;
    LMX_CHECK(ctx->nc == 3, LMX_ERR_UNSUPPORTED, "the fused true residual is built for NC = 3");
    const int V = (int)ctx->V;
    if (spinor_prec == LMX_H16)
        return wilson_f64_true(ctx, U, clover, kappa, csw, in, b, SpH16Out::at(r, r_norm, V, 0), done, fold, dstride,
                               force, partials, pstride, nrhs, gate);
    return wilson_f64_true(ctx, U, clover, kappa, csw, in, b, SpF32Out::at(r, r_norm, V, 0), done, fold, dstride,
                           force, partials, pstride, nrhs, gate);
}

// ---------------------------------------------------------------------------
// Low-precision variants for the mixed / multi-RHS solver: spinors FP32 or
// 16-bit (fixed point with a per-site FP32 scale), links FP32 or FP16, FP32
// arithmetic.
// ---------------------------------------------------------------------------
template <class In, class Out, class Lk, class Cl>
static int launch_lo_sp(lmx_context* ctx, const LoArgs& a, Lk U, Cl F, bool clover) {
    if constexpr (std::is_same<Lk, LinkF16>::value) {
        if ((a.rhs_block & kRhsLoop) && rhsloop_ok(ctx)) return launch_lo_loop<In, Out, Cl>(ctx, a, U, F, clover);
    }
    if ((a.rhs_block & kLdsStage) && lds_ok(ctx)) return launch_lo_lds<In, Out, Lk, Cl>(ctx, a, U, F, clover);
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

namespace lmx {

// ---------------------------------------------------------------------------
// mul! on the reference's own storage: psi and C are LatticeMatrix.A arrays
// (complex128, column-major [3,4,L1+2nw,...,L4+2nw], colour fastest, i.e.
// component spin*3 + colour of padded site p at p*12 + k).  psi's halo is not
// read: neighbours wrap in-kernel with the spinor phases (the halo fill of
// set_halo!, src/LatticeMatrices_core.jl:707-723, applied on the fly); only
// C's core sites are written.  Links and the clover cache are the operator's
// own (core-indexed).  Arithmetic as k_site with general phases.
// ---------------------------------------------------------------------------
struct PadGeom {
    int PS[4];  // padded strides (1, P1, P1 P2, P1 P2 P3)
    int nw;
    long Vp;    // padded volume
};

template <int MU, int PM, bool FWD, typename R, typename E, class Lk>
__device__ __forceinline__ void hop_pad(cx<R> (&acc)[12], const E* __restrict__ q, const Lk& U, const Geom& g,
                                        const PadGeom& pg, int s, long sp, const int x[4]) {
    bool w;
    const int t = FWD ? nbr_fwd(s, x, MU, g, w) : nbr_bwd(s, x, MU, g, w);
    const long step = (long)pg.PS[MU];
    const long tp = FWD ? (w ? sp - (long)(g.L[MU] - 1) * step : sp + step)
                        : (w ? sp + (long)(g.L[MU] - 1) * step : sp - step);
    cx<R> u[9], psi[12], h[6], o[6];
    U.template ld9<R>(MU, FWD ? s : t, u);
#pragma unroll
    for (int k = 0; k < 12; ++k) {
        const E v = q[tp * 12 + k];
        psi[k] = mk<R>(v.x, v.y);
    }
    if (w && !g.phase_one[MU]) {
        if (FWD)
            phase_spinor(psi, g.ph_re[MU], g.ph_im[MU]);
        else
            phase_spinor(psi, g.iph_re[MU], g.iph_im[MU]);
    }
    half_project<PM, MU>(psi, h);
    if (FWD)
        matvec_fwd(u, h, o);
    else
        matvec_bwd(u, h, o);
    reconstruct_add<PM, MU>(acc, o);
    pin_acc(acc);
}

// E: the field element (double2 for ComplexF64, float2 for ComplexF32); the
// arithmetic is in E's precision, links Lk and clover Cl of the same.
template <int FPM, int MODE, typename E, class Lk, class Cl>
__global__ void __launch_bounds__(256) k_wilson_pad(E* __restrict__ C, const E* __restrict__ psi, Lk U, Cl F, Geom g,
                                                    PadGeom pg, double kappa_d, double ccoef_d, int nrhs) {
    using R = typename std::conditional<std::is_same<E, float2>::value, float, double>::type;
    const long gid = (long)blockIdx.x * 256 + threadIdx.x;
    if (gid >= (long)g.V * nrhs) return;
    const int s = (int)(gid % g.V);
    const int r = (int)(gid / g.V);
    const R kappa = R(kappa_d), ccoef = R(ccoef_d);
    int x[4];
    site_coords(s, g, x);
    const long sp = (long)(x[0] + pg.nw) + (long)pg.PS[1] * (x[1] + pg.nw) + (long)pg.PS[2] * (x[2] + pg.nw) +
                    (long)pg.PS[3] * (x[3] + pg.nw);
    const E* q = psi + (long)r * pg.Vp * 12;
    cx<R> acc[12];
#pragma unroll
    for (int k = 0; k < 12; ++k) acc[k] = mk<R>(R(0), R(0));
    hop_pad<0, FPM, true>(acc, q, U, g, pg, s, sp, x);
    hop_pad<0, -FPM, false>(acc, q, U, g, pg, s, sp, x);
    hop_pad<1, FPM, true>(acc, q, U, g, pg, s, sp, x);
    hop_pad<1, -FPM, false>(acc, q, U, g, pg, s, sp, x);
    hop_pad<2, FPM, true>(acc, q, U, g, pg, s, sp, x);
    hop_pad<2, -FPM, false>(acc, q, U, g, pg, s, sp, x);
    hop_pad<3, FPM, true>(acc, q, U, g, pg, s, sp, x);
    hop_pad<3, -FPM, false>(acc, q, U, g, pg, s, sp, x);
    cx<R> res[12];
    if (MODE == MODE_DONLY) {
#pragma unroll
        for (int sp4 = 0; sp4 < 4; ++sp4)
#pragma unroll
            for (int c = 0; c < 3; ++c)
                res[sp4 * 3 + c] = cadd(mk<R>(R(0), R(0)), rmul(R(0.5), acc[c * 4 + sp4]));
    } else {
        cx<R> p[12];
#pragma unroll
        for (int k = 0; k < 12; ++k) {
            const E v = q[sp * 12 + k];
            p[k] = mk<R>(v.x, v.y);
        }
#pragma unroll
        for (int sp4 = 0; sp4 < 4; ++sp4)
#pragma unroll
            for (int c = 0; c < 3; ++c) res[sp4 * 3 + c] = cadd(p[sp4 * 3 + c], rmul(-kappa, acc[c * 4 + sp4]));
        if constexpr (MODE == MODE_CLOVER) clover_add<R>(F, s, g.V, p, ccoef, res);
    }
    E* o = C + (long)r * pg.Vp * 12 + sp * 12;
#pragma unroll
    for (int k = 0; k < 12; ++k) {
        E v;
        v.x = res[k].re;
        v.y = res[k].im;
        o[k] = v;
    }
}

template <typename E, class Lk, class Cl>
static int launch_padded(lmx_context* ctx, Lk l, Cl F, bool clover, double kappa, double csw, E* C, const E* psi,
                         int nw, int nrhs, int flags) {
    if (ctx->V == 0 || nrhs == 0) return LMX_OK;
    const Geom& g = ctx->g;
    PadGeom pg;
    pg.nw = nw;
    long st = 1;
    for (int d = 0; d < 4; ++d) {
        pg.PS[d] = (int)st;
        st *= g.L[d] + 2 * nw;
    }
    pg.Vp = st;
    LMX_CHECK(st * 12 < (1L << 40), LMX_ERR_ARG, "padded volume too large");
    const double ccoef = (-kappa) * csw;
    const long n = ctx->V * (long)nrhs;
    dim3 grid((unsigned)((n + 255) / 256)), blk(256);
    const bool dag = (flags & LMX_DAGGER) != 0, donly = (flags & LMX_DONLY) != 0;
#define LMX_PAD(FPM, MODE)                                                                                    \
    hipLaunchKernelGGL((k_wilson_pad<FPM, MODE, E, Lk, Cl>), grid, blk, 0, ctx->stream, C, psi, l, F, g, pg, kappa, \
                       ccoef, nrhs)
    if (donly) {
        if (dag)
            LMX_PAD(1, MODE_DONLY);
        else
            LMX_PAD(-1, MODE_DONLY);
    } else if (clover) {
        if (dag)
            LMX_PAD(1, MODE_CLOVER);
        else
            LMX_PAD(-1, MODE_CLOVER);
    } else {
        if (dag)
            LMX_PAD(1, MODE_WILSON);
        else
            LMX_PAD(-1, MODE_WILSON);
    }
#undef LMX_PAD
    LMX_HIP(hipGetLastError());
    return LMX_OK;
}

int launch_wilson_padded(lmx_context* ctx, const double2* U, const double2* clover, double kappa, double csw,
                         double2* C, const double2* psi, int nw, int nrhs, int flags) {
    return launch_padded<double2>(ctx, LinkF64{U, (int)ctx->V}, F64Clover{clover}, clover != nullptr, kappa, csw, C,
                                  psi, nw, nrhs, flags);
}

int launch_wilson_padded_f32(lmx_context* ctx, const float2* U, const float2* clover, double kappa, double csw,
                             float2* C, const float2* psi, int nw, int nrhs, int flags) {
    return launch_padded<float2>(ctx, LinkF32{U, (int)ctx->V}, F32Clover{clover}, clover != nullptr, kappa, csw, C,
                                 psi, nw, nrhs, flags);
}

}  // namespace lmx
