// wilson.hip -- the Wilson(-clover) hopping stencil on gfx950.
//
// One work item per lattice site and per register block of NR right-hand
// sides; SoA fields so every load of a wave is one contiguous 512 B/1 KiB
// segment; periodic/antiperiodic wrap inside the kernel (no halo); the
// workgroup->site mapping is XCD-contiguous (xcd_block) so that the x/y/z
// neighbour re-reads hit in the local L2.  The arithmetic follows the
// reference's half-spin kernels operation by operation:
//   kernel_WilsonDiracOperator4D! NC=3        src/Operators/WilsonDiracOperator.jl:413-419
//   kernel_adjoint_WilsonDiracOperator4D!     :514-520
//   kernel_WilsonDiracOperator4D_Donly!       :624-631, adjoint :736-743
//   kernel_WilsonDiracCloverOperator4D!       src/Operators/WilsonDiracCloverOperator.jl:537-559
#include "lmx_internal.h"

namespace lmx {

// ---------------------------------------------------------------------------
// Clover field storage: 6 anti-Hermitian planes, 9 reals each, as 27 complex
// slots per site: slots 0..8 carry the 18 imaginary diagonals (plane*3+i) in
// pairs, slots 9..26 the upper off-diagonals (plane*3 + {01,02,12}).
// ---------------------------------------------------------------------------
struct NoClover {
    using R = double;
};
template <class Acc>
struct CloverAcc {
    Acc a;
};

template <typename R, class Acc>
__device__ __forceinline__ void clover_add(const Acc& F, int s, int V, const cx<R> psi[12], R coefficient,
                                           cx<R> res[12]) {
    R d[18];
    cx<R> u[18];
#pragma unroll
    for (int q = 0; q < 9; ++q) {
        cx<R> v = F.ld((long)q * V + s);
        d[2 * q] = v.re;
        d[2 * q + 1] = v.im;
    }
#pragma unroll
    for (int q = 0; q < 18; ++q) u[q] = F.ld((long)(9 + q) * V + s);
    // kernel_add_clover_term!, src/Operators/WilsonDiracCloverOperator.jl:481-525
#pragma unroll
    for (int c = 0; c < 3; ++c) {
        cx<R> v1 = mk<R>(0, 0), v2 = mk<R>(0, 0), v3 = mk<R>(0, 0), v4 = mk<R>(0, 0);
#pragma unroll
        for (int ic = 0; ic < 3; ++ic) {
            cx<R> f[6];
#pragma unroll
            for (int p = 0; p < 6; ++p) {
                if (c == ic) {
                    f[p] = mk<R>(R(0), d[3 * p + c]);
                } else if (c < ic) {
                    const int j = (c == 0) ? (ic == 1 ? 0 : 1) : 2;
                    f[p] = u[3 * p + j];
                } else {
                    const int j = (ic == 0) ? (c == 1 ? 0 : 1) : 2;
                    f[p] = cneg(cconj(u[3 * p + j]));
                }
            }
            cx<R> f12 = f[0], f13 = f[1], f14 = f[2], f23 = f[3], f24 = f[4], f34 = f[5];
            cx<R> s1 = cadd(f12, f34), s2 = csub(f12, f34), s3 = cadd(f14, f23), s4 = csub(f23, f14);
            cx<R> dp = mk<R>(-s1.im, s1.re), dm = mk<R>(-s2.im, s2.re);
            cx<R> odp = mk<R>(-s3.im, s3.re), odm = mk<R>(-s4.im, s4.re);
            cx<R> p1 = psi[ic], p2 = psi[3 + ic], p3 = psi[6 + ic], p4 = psi[9 + ic];
            v1 = cadd(v1, cadd(cmul(dp, p1), cmul(cadd(cadd(cneg(f13), odp), f24), p2)));
            v2 = cadd(v2, cadd(cmul(cneg(dp), p2), cmul(csub(cadd(f13, odp), f24), p1)));
            v3 = cadd(v3, cadd(cmul(dm, p3), cmul(csub(cadd(cneg(f13), odm), f24), p4)));
            v4 = cadd(v4, cadd(cmul(cneg(dm), p4), cmul(cadd(cadd(f13, odm), f24), p3)));
        }
        res[c] = cadd(res[c], rmul(coefficient, v1));
        res[3 + c] = cadd(res[3 + c], rmul(coefficient, v2));
        res[6 + c] = cadd(res[6 + c], rmul(coefficient, v3));
        res[9 + c] = cadd(res[9 + c], rmul(coefficient, v4));
    }
}

// ---------------------------------------------------------------------------
// One direction of the hopping sum for NR right-hand sides.
// ---------------------------------------------------------------------------
template <int MU, int FPM, int NR, bool GP, class In, class Lk, typename R>
__device__ __forceinline__ void hop_mu(cx<R> (&acc)[NR][12], const In& in, const Lk& U, const Geom& g, int s,
                                       const int x[4], const long (&roff)[NR]) {
    bool w;
    cx<R> u[9];
    // forward: U_mu(x) (1 + FPM gamma_mu) psi(x+mu)
    const int sp = nbr_fwd(s, x, MU, g, w);
    const bool negf = w && g.phase_neg[MU];
#pragma unroll
    for (int k = 0; k < 9; ++k) {
        auto v = U.ld((long)(MU * 9 + k) * g.V + s);
        u[k] = mk<R>(R(v.re), R(v.im));
    }
#pragma unroll
    for (int r = 0; r < NR; ++r) {
        cx<R> psi[12], h[6], o[6];
#pragma unroll
        for (int k = 0; k < 12; ++k) psi[k] = in.ld(roff[r] + (long)k * g.V + sp);
        if (GP) {
            if (w && !g.phase_one[MU]) phase_spinor(psi, g.ph_re[MU], g.ph_im[MU]);
        }
        half_project<FPM, MU>(psi, h);
        if (!GP) neg_select(h, negf);
        matvec_fwd(u, h, o);
        reconstruct_add<FPM, MU>(acc[r], o);
    }
    // backward: U_mu(x-mu)^dagger (1 - FPM gamma_mu) psi(x-mu)
    const int sm = nbr_bwd(s, x, MU, g, w);
    const bool negb = w && g.phase_neg[MU];
#pragma unroll
    for (int k = 0; k < 9; ++k) {
        auto v = U.ld((long)(MU * 9 + k) * g.V + sm);
        u[k] = mk<R>(R(v.re), R(v.im));
    }
#pragma unroll
    for (int r = 0; r < NR; ++r) {
        cx<R> psi[12], h[6], o[6];
#pragma unroll
        for (int k = 0; k < 12; ++k) psi[k] = in.ld(roff[r] + (long)k * g.V + sm);
        if (GP) {
            if (w && !g.phase_one[MU]) phase_spinor(psi, g.iph_re[MU], g.iph_im[MU]);
        }
        half_project<-FPM, MU>(psi, h);
        if (!GP) neg_select(h, negb);
        matvec_bwd(u, h, o);
        reconstruct_add<-FPM, MU>(acc[r], o);
    }
}

struct EpiArgs {
    const double2* aux;
    CGState* st;
    double* partials;
    unsigned* counter;
    const int* gate;
};

enum { MODE_WILSON = 0, MODE_DONLY = 1, MODE_CLOVER = 2 };
enum { EPI_NONE = 0, EPI_PAP = 1 };

// out = psi - kappa*hop (+ clover) or 1/2 hop, for NR right-hand sides per item.
template <class In, class Lk, class Out, class Cl, int FPM, int MODE, int NR, int EPI, bool GP>
__global__ void __launch_bounds__(256) k_wilson(Out out, In in, Lk U, Cl F, Geom g, double kappa_d, double ccoef_d,
                                                long rstride, int ngroups, int nrhs, EpiArgs epi) {
    using R = typename In::R;
    if (epi.st != nullptr && epi.st->done) return;
    if (epi.gate != nullptr && *epi.gate == 0) return;
    __shared__ double lds4[4];
    const int lb = xcd_block(blockIdx.x, gridDim.x);
    const int sb = lb / ngroups, rg = lb - sb * ngroups;
    const int s = sb * 256 + threadIdx.x;
    const int r0 = rg * NR;
    const int nr = min(NR, nrhs - r0);
    const long base = (long)r0 * rstride;
    double red = 0.0;
    if (s < g.V) {
        int x[4];
        site_coords(s, g, x);
        cx<R> acc[NR][12];
        // right-hand sides beyond nr (a ragged last group) recompute the
        // group's first RHS and are not stored: no branch in the stencil
        long roff[NR];
#pragma unroll
        for (int r = 0; r < NR; ++r) {
            roff[r] = base + (long)(r < nr ? r : 0) * rstride;
#pragma unroll
            for (int k = 0; k < 12; ++k) acc[r][k] = mk<R>(R(0), R(0));
        }
        hop_mu<0, FPM, NR, GP>(acc, in, U, g, s, x, roff);
        hop_mu<1, FPM, NR, GP>(acc, in, U, g, s, x, roff);
        hop_mu<2, FPM, NR, GP>(acc, in, U, g, s, x, roff);
        hop_mu<3, FPM, NR, GP>(acc, in, U, g, s, x, roff);
        const R kappa = R(kappa_d);
#pragma unroll
        for (int r = 0; r < NR; ++r) {
            if (r < nr) {
                cx<R> res[12];
                if (MODE == MODE_DONLY) {
                    const R half = R(1) / R(2);
#pragma unroll
                    for (int sp = 0; sp < 4; ++sp)
#pragma unroll
                        for (int c = 0; c < 3; ++c)
                            res[sp * 3 + c] = cadd(mk<R>(R(0), R(0)), rmul(half, acc[r][c * 4 + sp]));
                } else {
                    cx<R> psi[12];
#pragma unroll
                    for (int k = 0; k < 12; ++k) psi[k] = in.ld(base + r * rstride + (long)k * g.V + s);
#pragma unroll
                    for (int sp = 0; sp < 4; ++sp)
#pragma unroll
                        for (int c = 0; c < 3; ++c)
                            res[sp * 3 + c] = cadd(psi[sp * 3 + c], rmul(-kappa, acc[r][c * 4 + sp]));
                    if constexpr (MODE == MODE_CLOVER) clover_add<R>(F.a, s, g.V, psi, R(ccoef_d), res);
                }
#pragma unroll
                for (int k = 0; k < 12; ++k) out.st(base + r * rstride + (long)k * g.V + s, res[k]);
                if constexpr (EPI == EPI_PAP) {
                    // Re sum conj(p) * Ap over the site, kernel_dot_D order (spin, colour).
                    double acc_re = 0.0;
#pragma unroll
                    for (int k = 0; k < 12; ++k) {
                        double2 pv = epi.aux[(long)k * g.V + s];
                        cx<double> t = cmul(cconj(mk<double>(pv.x, pv.y)), mk<double>(double(res[k].re), double(res[k].im)));
                        acc_re = acc_re + t.re;
                    }
                    red = acc_re;
                }
            }
        }
    }
    if constexpr (EPI == EPI_PAP) {
        // block partial of Re <p, Ap>; k_finalize (FIN_PAP) completes the sum
        double tot = block_tree(red, lds4);
        if (threadIdx.x == 0) epi.partials[lb] = tot;
    }
}

template <class Acc>
struct CloverWrap {
    Acc a;
};

template <class In, class Lk, class Out, class Cl, int NR, int EPI>
static int dispatch_fpm_mode(lmx_context* ctx, Out out, In in, Lk U, Cl F, double kappa, double ccoef, long rstride,
                             int ngroups, int nrhs, int flags, bool clover, const EpiArgs& epi) {
    const int nb = nblocks(ctx->V) * ngroups;
    dim3 grid(nb), block(256);
    const bool dag = (flags & LMX_DAGGER) != 0;
    const bool donly = (flags & LMX_DONLY) != 0;
#define LMX_LAUNCH(FPM, MODE)                                                                                   \
    do {                                                                                                         \
        if (ctx->g.general)                                                                                      \
            hipLaunchKernelGGL((k_wilson<In, Lk, Out, Cl, FPM, MODE, NR, EPI, true>), grid, block, 0, ctx->stream, \
                               out, in, U, F, ctx->g, kappa, ccoef, rstride, ngroups, nrhs, epi);                \
        else                                                                                                     \
            hipLaunchKernelGGL((k_wilson<In, Lk, Out, Cl, FPM, MODE, NR, EPI, false>), grid, block, 0,           \
                               ctx->stream, out, in, U, F, ctx->g, kappa, ccoef, rstride, ngroups, nrhs, epi);   \
    } while (0)
    if (donly) {
        if constexpr (EPI == EPI_NONE) {
            if (dag)
                LMX_LAUNCH(1, MODE_DONLY);
            else
                LMX_LAUNCH(-1, MODE_DONLY);
        } else {
            set_error("hopping-only operator has no fused reduction");
            return LMX_ERR_UNSUPPORTED;
        }
    } else if (clover) {
        if (dag)
            LMX_LAUNCH(1, MODE_CLOVER);
        else
            LMX_LAUNCH(-1, MODE_CLOVER);
    } else {
        if (dag)
            LMX_LAUNCH(1, MODE_WILSON);
        else
            LMX_LAUNCH(-1, MODE_WILSON);
    }
#undef LMX_LAUNCH
    LMX_HIP(hipGetLastError());
    return LMX_OK;
}

struct F64Clover {
    const double2* __restrict__ p;
    __device__ __forceinline__ cx<double> ld(long i) const {
        double2 v = p[i];
        return mk<double>(v.x, v.y);
    }
};

int launch_wilson_f64(lmx_context* ctx, const double2* U, const double2* clover, double kappa, double csw,
                      double2* out, const double2* in, int nrhs, int flags, const Epilogue& e) {
    if (ctx->V == 0 || nrhs == 0) return LMX_OK;
    EpiArgs epi{e.aux, e.st, e.partials, e.counter, e.gate};
    const long rstride = 12 * ctx->V;
    const double ccoef = (-kappa) * csw;  // coefficient = -kappa * cSW (:572)
    F64Out o{out};
    F64In i{in};
    F64In l{U};
    if (e.kind == 1) {
        LMX_CHECK(nrhs == 1, LMX_ERR_UNSUPPORTED, "fused p.Ap needs a single right-hand side");
        CloverWrap<F64Clover> F{F64Clover{clover}};
        LMX_TRY((dispatch_fpm_mode<F64In, F64In, F64Out, CloverWrap<F64Clover>, 1, EPI_PAP>(
            ctx, o, i, l, F, kappa, ccoef, rstride, 1, 1, flags, clover != nullptr, epi)));
        return launch_finalize(ctx, nblocks(ctx->V), FIN_PAP, false, e.st, nullptr);
    }
    CloverWrap<F64Clover> F{F64Clover{clover}};
    return dispatch_fpm_mode<F64In, F64In, F64Out, CloverWrap<F64Clover>, 1, EPI_NONE>(
        ctx, o, i, l, F, kappa, ccoef, rstride, nrhs, nrhs, flags, clover != nullptr, epi);
}

// ---------------------------------------------------------------------------
// Low-precision (FP32 spinor) variants for the mixed / multi-RHS solver.
// ---------------------------------------------------------------------------
struct F32Clover {
    const float2* __restrict__ p;
    __device__ __forceinline__ cx<float> ld(long i) const {
        float2 v = p[i];
        return mk<float>(v.x, v.y);
    }
};
struct F16Clover {
    const __half2* __restrict__ p;
    __device__ __forceinline__ cx<float> ld(long i) const {
        float2 v = __half22float2(p[i]);
        return mk<float>(v.x, v.y);
    }
};

template <class Lk, class Cl, int NR>
static int launch_lo_nr(lmx_context* ctx, Lk U, Cl F, double kappa, double ccoef, float2* out, const float2* in,
                        int nrhs, int flags, bool clover, int rhs_in_launch, const int* gate) {
    // One launch covers rhs_in_launch right-hand sides split into groups of NR.
    EpiArgs epi{nullptr, nullptr, nullptr, nullptr, gate};
    const long rstride = 12 * ctx->V;
    for (int r0 = 0; r0 < nrhs; r0 += rhs_in_launch) {
        const int n = std::min(rhs_in_launch, nrhs - r0);
        const int ngroups = (n + NR - 1) / NR;
        F32Out o{out + r0 * rstride};
        F32In i{in + r0 * rstride};
        LMX_TRY((dispatch_fpm_mode<F32In, Lk, F32Out, Cl, NR, EPI_NONE>(ctx, o, i, U, F, kappa, ccoef, rstride,
                                                                         ngroups, n, flags, clover, epi)));
    }
    return LMX_OK;
}

template <class Lk, class Cl>
static int launch_lo_any(lmx_context* ctx, Lk U, Cl F, double kappa, double ccoef, float2* out, const float2* in,
                         int nrhs, int flags, bool clover, int rhs_block, const int* gate) {
    // rhs_block encodes (rhs per launch) * 16 + (rhs per work item).
    int per_launch = rhs_block >> 4, nr = rhs_block & 15;
    if (per_launch <= 0) per_launch = 1;
    if (nr <= 0) nr = 1;
    switch (nr) {
    case 1:
        return launch_lo_nr<Lk, Cl, 1>(ctx, U, F, kappa, ccoef, out, in, nrhs, flags, clover, per_launch, gate);
    case 2:
        return launch_lo_nr<Lk, Cl, 2>(ctx, U, F, kappa, ccoef, out, in, nrhs, flags, clover, per_launch, gate);
    case 3:
        return launch_lo_nr<Lk, Cl, 3>(ctx, U, F, kappa, ccoef, out, in, nrhs, flags, clover, per_launch, gate);
    case 4:
        return launch_lo_nr<Lk, Cl, 4>(ctx, U, F, kappa, ccoef, out, in, nrhs, flags, clover, per_launch, gate);
    default:
        set_error("unsupported rhs register block %d", nr);
        return LMX_ERR_UNSUPPORTED;
    }
}

int launch_wilson_lo(lmx_context* ctx, int link_prec, const void* U, const void* clover, double kappa, double csw,
                     float2* out, const float2* in, int nrhs, int flags, int rhs_block, const int* gate) {
    if (ctx->V == 0 || nrhs == 0) return LMX_OK;
    const double ccoef = (-kappa) * csw;
    const bool cl = clover != nullptr;
    if (link_prec == LMX_F32) {
        return launch_lo_any(ctx, F32In{(const float2*)U}, CloverWrap<F32Clover>{F32Clover{(const float2*)clover}},
                             kappa, ccoef, out, in, nrhs, flags, cl, rhs_block, gate);
    }
    if (link_prec == LMX_F16) {
        return launch_lo_any(ctx, F16In{(const __half2*)U}, CloverWrap<F16Clover>{F16Clover{(const __half2*)clover}},
                             kappa, ccoef, out, in, nrhs, flags, cl, rhs_block, gate);
    }
    set_error("unsupported link precision %d", link_prec);
    return LMX_ERR_UNSUPPORTED;
}

}  // namespace lmx
