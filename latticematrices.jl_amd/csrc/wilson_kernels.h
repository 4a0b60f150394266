// wilson_kernels.h -- the Wilson(-clover) hopping stencil on gfx950 (kernel
// templates; wilson.hip instantiates FP64, wilson_lo.hip the low precisions).
//
// Two launch shapes share the per-direction arithmetic:
//  * k_stream: a workgroup owns 256 consecutive sites of one t-slice (a 3-D
//    block) and walks TL consecutive t-slices.  The backward t-hop
//    U_t(x-t)^dagger P psi(x-t) of the next slice is computed from the current
//    slice's own spinor and link and carried in registers, so the t-neighbours
//    (the re-reads that miss the 4 MB L2 in a site-parallel launch) are never
//    fetched twice; XCDs are mapped to (t-chunk, z-range) units so one step's
//    working set stays in the XCD's L2.
//  * k_site: one work item per site (lattices whose 3-D volume is not a
//    multiple of 256).
// SoA fields make every wave-wide load one contiguous segment; periodic /
// antiperiodic wrap happens in the kernel.  After every half-direction the
// accumulators are pinned (an empty asm with register operands), which
// keeps the compiler from sinking the arithmetic below later loads and
// running out of registers.  The arithmetic follows the reference's half-spin
// kernels operation by operation:
//   kernel_WilsonDiracOperator4D! NC=3        src/Operators/WilsonDiracOperator.jl:413-419
//   kernel_adjoint_WilsonDiracOperator4D!     :514-520
//   kernel_WilsonDiracOperator4D_Donly!       :624-631, adjoint :736-743
//   kernel_WilsonDiracCloverOperator4D!       src/Operators/WilsonDiracCloverOperator.jl:537-559
#pragma once

#include <algorithm>
#include <type_traits>

#include "lmx_internal.h"

namespace lmx {

// ---------------------------------------------------------------------------
// Clover field storage: 6 anti-Hermitian planes, 9 reals each, as 27 complex
// slots per site: slots 0..8 carry the 18 imaginary diagonals (plane*3+i) in
// pairs, slots 9..26 the upper off-diagonals (plane*3 + {01,02,12}).
// ---------------------------------------------------------------------------
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

struct F64Clover {
    const double2* __restrict__ p;
    __device__ __forceinline__ cx<double> ld(long i) const {
        double2 v = p[i];
        return mk<double>(v.x, v.y);
    }
};
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

// ---------------------------------------------------------------------------
// Per-half-direction arithmetic.  PM is the projector sign of this half.
// ---------------------------------------------------------------------------
template <int NR, typename R>
__device__ __forceinline__ void pin_all(cx<R> (&a)[NR][12]) {
#pragma unroll
    for (int r = 0; r < NR; ++r) pin_acc(a[r]);
}

template <typename R>
constexpr bool kSpatialOne = std::is_same<R, float>::value;

// negate a half-spinor on a workgroup-uniform condition (a t wrap in the
// streaming kernel): a scalar branch instead of per-lane selects
template <typename R>
__device__ __forceinline__ void neg_uniform(cx<R> h[6], bool neg) {
    if (neg) {
#pragma unroll
        for (int k = 0; k < 6; ++k) h[k] = cneg(h[k]);
    }
}

// keep a carried half-spinor in registers: loads after this are not hoisted above it
template <typename R>
__device__ __forceinline__ void pin6f(cx<R> (&a)[6]) {
#pragma unroll
    for (int k = 0; k < 6; ++k) asm volatile("" : "+v"(a[k].re), "+v"(a[k].im)::"memory");
}

template <int MU, int PM, bool FWD, int NR, bool GP, class In, class Lk, typename R>
__device__ __forceinline__ void hop_half(cx<R> (&acc)[NR][12], const In& in, const Lk& U, const Geom& g, int s,
                                         const int x[4], const int (&rr)[NR]) {
    bool w;
    const int t = FWD ? nbr_fwd(s, x, MU, g, w) : nbr_bwd(s, x, MU, g, w);
    const bool ng = w && g.phase_neg[MU];
    // general phases rotate the components, so the projected half-spinor can
    // leave the FP16 range: those launches keep the converted link entries
    LinkOp<Lk, R, kHalfProducts<In, Lk> && !GP> u;
    u.load(U, MU, FWD ? s : t);
#pragma unroll
    for (int r = 0; r < NR; ++r) {
        cx<R> psi[12], h[6], o[6];
        R c;
        in.ld12s(rr[r], t, psi, c);
        if (GP) {
            if (w && !g.phase_one[MU]) {
                if (FWD)
                    phase_spinor(psi, g.ph_re[MU], g.ph_im[MU]);
                else
                    phase_spinor(psi, g.iph_re[MU], g.iph_im[MU]);
            }
        }
        half_project<PM, MU>(psi, h);
        // low-precision launches run with x, y, z phases exactly 1 (others take
        // the general-phase kernel), so only t can wrap with a sign
        if (!GP && (MU == 3 || !kSpatialOne<R>)) neg_select(h, ng);
        if (FWD)
            u.fwd(h, o);
        else
            u.bwd(h, o);
        reconstruct_add_s<PM, MU>(acc[r], o, c);
    }
    pin_all(acc);
}

struct EpiArgs {
    double* partials;  // block partials of the fused reduction, row r at partials + r * pstride
    long pstride;
    CGState* st;       // early exit when st->done
    const int* gate;   // early exit while *gate == 0
};

enum { MODE_WILSON = 0, MODE_DONLY = 1, MODE_CLOVER = 2 };
#ifndef LMX_CARRY_NEXT
#define LMX_CARRY_NEXT 1
#endif
#ifndef LMX_NT_UT
#define LMX_NT_UT 1
#endif
#ifndef LMX_TWO_LANE
#define LMX_TWO_LANE 0
#endif

// Fused epilogues: EPI_PAP   Re <aux, out> (the CG's p.Ap on D^dagger, FP64 CG);
//                  EPI_XR    out = Ap is not stored: r -= alpha Ap, x += alpha p
//                            (p = aux) and |r|^2 (mixed solver, on D^dagger);
//                  EPI_TP    the stencil input is r: out = t = D r + beta t in
//                            place (= D p by linearity), p = r + beta p_prev
//                            written at the site, and |t|^2 (mixed solver, on D).
enum { EPI_NONE = 0, EPI_PAP = 1, EPI_XR = 3, EPI_TP = 4, EPI_TRUE = 5 };

struct NoEx {};
// Extra fields of EPI_XR, all offset to the launch's first RHS.  The x
// update is batched over groups of kXBatch iterations: with apply_x == 0 the
// kernel leaves x alone (alpha p is deferred, p kept in its ring slot); with
// apply_x == 1 it adds the deferred alpha_q[j] pq[j] of the marked slots,
// oldest first, and then alpha p.  APPLY fixes apply_x at compile time (0:
// defer, 1: apply), so the deferring kernel (three launches in four) is
// compiled without the registers of the x update.
template <class In, class Out, int APPLY>
struct XrEx {
    static constexpr int kApply = APPLY;
    In rin;        // r (read) ...
    Out rout;      // ... and written in place
    SpX4In xin;    // inner solution x (FP32, component pairs)
    SpX4Out xout;
    In pq[kXBatch - 1];        // ring slots of the deferred p (apply_x == 1)
    const double* alpha;       // alpha of RHS r at alpha[r * astride]
    const double* alpha_q;     // deferred alpha of slot j at alpha_q[r * astride + j]
    int astride;
    const int* done;           // retired RHS (no update) at done[r * dstride]
    const int* pmask;          // deferred slots (bit j) at pmask[r * dstride]
    int dstride;
    int apply_x;
};

// Extra fields of EPI_TP, offset to the launch's first RHS.  The search
// direction update p = r + beta p_prev of the conjugate-gradient iteration is
// folded into the D launch that stencils r: by linearity D p = D r + beta
// D p_prev, and D p_prev is the previous iteration's t, read and overwritten
// at the same site.  Both are read on every iteration: on the first
// (beta == 0) they must hold finite values (the solver's init writes t = 0
// and p_prev = r).
template <class In, class Out>
struct TpEx {
    In tin;               // t of the previous iteration (the launch's out buffer)
    In pin;               // p_prev
    Out pout;             // p, written for live RHS
    const double* beta;   // beta of RHS r at beta[r * astride]
    int astride;
    const int* done;      // retired RHS (p left as it is) at done[r * dstride]
    int dstride;
};

// EPI_TRUE (FP64 D^dagger of a reliable update, the operand D x): the true
// residual r = b - D^dagger D x at the site in FP64, written to the inner
// solver's r (16-bit or FP32) instead of storing D^dagger D x, with |r|^2 as
// the site's contribution (FP64, the unfused kernel's sum order).  An RHS
// takes part unless retired (done), or, on the forced final pass, retired
// with its fold outstanding; the batch-wide request is the launch's gate.
template <class Out>
struct TrueEx {
    SpF64In b;           // right-hand sides
    Out rout;            // r, in the inner format
    const int* done;     // done[r * dstride]
    const int* fold;     // fold[r * dstride]
    int dstride;
    int force;
};

// Finish one site from its own spinor psi: out = psi - kappa*hop (+ clover),
// or 1/2 hop; then the fused epilogue, returning the site's contribution to
// its reduction.
template <int MODE, int EPI, bool SPLIT = false, class Out, class Cl, class Aux, class Ex, typename R>
__device__ __forceinline__ double finish_site(const Out& out, const cx<R> (&psi)[12], const Cl& F, const Aux& aux,
                                              const Ex& ex, int r, int s, const Geom& g, R kappa, R ccoef,
                                              const cx<R> (&acc)[12], bool store) {
    cx<R> res[12];
    if (MODE == MODE_DONLY) {
        const R half = R(1) / R(2);
#pragma unroll
        for (int sp = 0; sp < 4; ++sp)
#pragma unroll
            for (int c = 0; c < 3; ++c) res[sp * 3 + c] = cadd(mk<R>(R(0), R(0)), rmul(half, acc[c * 4 + sp]));
    } else {
#pragma unroll
        for (int sp = 0; sp < 4; ++sp)
#pragma unroll
            for (int c = 0; c < 3; ++c) res[sp * 3 + c] = cadd(psi[sp * 3 + c], rmul(-kappa, acc[c * 4 + sp]));
        if constexpr (MODE == MODE_CLOVER) clover_add<R>(F, s, g.V, psi, ccoef, res);
    }
    double red = 0.0;
    if constexpr (EPI == EPI_TP) {
        // site scales (1 for unscaled formats) are folded into beta.  SPLIT:
        // p first and a compiler barrier, so that psi is dead before t is
        // read (the register allocation the LDS-staged kernel needs; the
        // streaming kernel allocates better with both loads together)
        const R be = R(ex.beta[(long)r * ex.astride]);
        if constexpr (SPLIT) {
            {
                cx<R> pv[12];
                R cp;
                ex.pin.ld12s_stream(r, s, pv, cp);
                const R bp = be * cp;
#pragma unroll
                for (int k = 0; k < 12; ++k)
                    pv[k] = mk<R>(fmaR(bp, pv[k].re, psi[k].re), fmaR(bp, pv[k].im, psi[k].im));
                if (store && ex.done[(long)r * ex.dstride] == 0) ex.pout.st12(r, s, pv);
            }
            asm volatile("" ::: "memory");
            cx<R> tv[12];
            R ct;
            ex.tin.ld12s_stream(r, s, tv, ct);
            const R bt = be * ct;
#pragma unroll
            for (int k = 0; k < 12; ++k)
                res[k] = mk<R>(fmaR(bt, tv[k].re, res[k].re), fmaR(bt, tv[k].im, res[k].im));
        } else {
            cx<R> pv[12], tv[12];
            R ct, cp;
            ex.tin.ld12s_stream(r, s, tv, ct);
            ex.pin.ld12s_stream(r, s, pv, cp);
            const R bt = be * ct, bp = be * cp;
#pragma unroll
            for (int k = 0; k < 12; ++k) {
                res[k] = mk<R>(fmaR(bt, tv[k].re, res[k].re), fmaR(bt, tv[k].im, res[k].im));
                pv[k] = mk<R>(fmaR(bp, pv[k].re, psi[k].re), fmaR(bp, pv[k].im, psi[k].im));
            }
            if (store && ex.done[(long)r * ex.dstride] == 0) ex.pout.st12(r, s, pv);
        }
    }
    if constexpr (EPI == EPI_XR) {
        const R a = R(ex.alpha[(long)r * ex.astride]);
        cx<R> rv[12];
        ex.rin.ld12_stream(r, s, rv);
        R site = R(0);  // per-site sum in working precision, accumulated over sites in FP64
#pragma unroll
        for (int k = 0; k < 12; ++k) {
            rv[k] = mk<R>(fmaR(-a, res[k].re, rv[k].re), fmaR(-a, res[k].im, rv[k].im));
            site = fmaR(rv[k].re, rv[k].re, site);
            site = fmaR(rv[k].im, rv[k].im, site);
        }
        red = double(site);
        const bool live = store && ex.done[(long)r * ex.dstride] == 0;
        if (live) ex.rout.st12(r, s, rv);
        if (Ex::kApply == 1 && live) {
            cx<R> pv[12], xv[12];
            aux.ld12_stream(r, s, pv);
            ex.xin.ld12_stream(r, s, xv);
            const int m = ex.pmask[(long)r * ex.dstride];
#pragma unroll
            for (int j = 0; j < kXBatch - 1; ++j) {
                if (m & (1 << j)) {
                    const R ap = R(ex.alpha_q[(long)r * ex.astride + j]);
                    cx<R> qv[12];
                    ex.pq[j].ld12_stream(r, s, qv);
#pragma unroll
                    for (int k = 0; k < 12; ++k)
                        xv[k] = mk<R>(fmaR(ap, qv[k].re, xv[k].re), fmaR(ap, qv[k].im, xv[k].im));
                }
            }
#pragma unroll
            for (int k = 0; k < 12; ++k) xv[k] = mk<R>(fmaR(a, pv[k].re, xv[k].re), fmaR(a, pv[k].im, xv[k].im));
            ex.xout.st12(r, s, xv);
        }
        return red;
    } else if constexpr (EPI == EPI_TRUE) {
        static_assert(std::is_same<R, double>::value, "the true residual is an FP64 epilogue");
        cx<double> bv[12];
        ex.b.ld12_stream(r, s, bv);
        cx<float> f[12];
        double v = 0.0;
#pragma unroll
        for (int k = 0; k < 12; ++k) {
            const double re = bv[k].re - res[k].re, im = bv[k].im - res[k].im;
            f[k] = mk<float>((float)re, (float)im);
            v += re * re + im * im;
        }
        const int d = ex.done[(long)r * ex.dstride];
        if (store && (d == 0 || (ex.force && ex.fold[(long)r * ex.dstride] != 0))) ex.rout.st12(r, s, f);
        return v;
    } else {
        if (store) out.st12(r, s, res);
    }
    if constexpr (EPI == EPI_PAP) {
        // Re sum conj(p) * Ap over the site, in kernel_dot_D order (spin, colour)
        cx<R> pv[12];
        aux.ld12(r, s, pv);
#pragma unroll
        for (int k = 0; k < 12; ++k) {
            cx<double> t = cmul(cconj(mk<double>(double(pv[k].re), double(pv[k].im))),
                                mk<double>(double(res[k].re), double(res[k].im)));
            red = red + t.re;
        }
    }
    if constexpr (EPI == EPI_TP) {
        R site = R(0);  // per-site sum in working precision, accumulated over sites in FP64
#pragma unroll
        for (int k = 0; k < 12; ++k) {
            site = fmaR(res[k].re, res[k].re, site);
            site = fmaR(res[k].im, res[k].im, site);
        }
        red = double(site);
    }
    return red;
}

template <int NR>
__device__ __forceinline__ void rhs_rows(int rg, int nrhs, int (&rr)[NR], int& nr) {
    const int r0 = rg * NR;
    nr = min(NR, nrhs - r0);
#pragma unroll
    for (int r = 0; r < NR; ++r) rr[r] = r0 + (r < nr ? r : 0);  // ragged group: recompute, do not store
}

// ---------------------------------------------------------------------------
// Per-site kernel
// ---------------------------------------------------------------------------
template <class In, class Lk, class Out, class Cl, class Aux, class Ex, int FPM, int MODE, int NR, int EPI, bool GP>
__global__ void __launch_bounds__(256) k_site(Out out, In in, Lk U, Cl F, Aux aux, Ex ex, Geom g, double kappa_d,
                                              double ccoef_d, int ngroups, int nrhs, EpiArgs epi) {
    using R = typename In::R;
    if (epi.st != nullptr && epi.st->done) return;
    if (epi.gate != nullptr && *epi.gate == 0) return;
    __shared__ double lds4[4];
    const int lb = xcd_block(blockIdx.x, gridDim.x);
    const int sb = lb / ngroups, rg = lb - sb * ngroups;
    const int s = sb * 256 + threadIdx.x;
    int rr[NR], nr;
    rhs_rows<NR>(rg, nrhs, rr, nr);
    double red[NR];
#pragma unroll
    for (int r = 0; r < NR; ++r) red[r] = 0.0;
    if (s < g.V) {
        int x[4];
        site_coords(s, g, x);
        cx<R> acc[NR][12];
#pragma unroll
        for (int r = 0; r < NR; ++r)
#pragma unroll
            for (int k = 0; k < 12; ++k) acc[r][k] = mk<R>(R(0), R(0));
        hop_half<0, FPM, true, NR, GP>(acc, in, U, g, s, x, rr);
        hop_half<0, -FPM, false, NR, GP>(acc, in, U, g, s, x, rr);
        hop_half<1, FPM, true, NR, GP>(acc, in, U, g, s, x, rr);
        hop_half<1, -FPM, false, NR, GP>(acc, in, U, g, s, x, rr);
        hop_half<2, FPM, true, NR, GP>(acc, in, U, g, s, x, rr);
        hop_half<2, -FPM, false, NR, GP>(acc, in, U, g, s, x, rr);
        hop_half<3, FPM, true, NR, GP>(acc, in, U, g, s, x, rr);
        hop_half<3, -FPM, false, NR, GP>(acc, in, U, g, s, x, rr);
#pragma unroll
        for (int r = 0; r < NR; ++r) {
            cx<R> psi[12];
            if (MODE != MODE_DONLY) in.ld12(rr[r], s, psi);
            red[r] = finish_site<MODE, EPI>(out, psi, F, aux, ex, rr[r], s, g, R(kappa_d), R(ccoef_d), acc[r],
                                            r < nr);
        }
    }
    if constexpr (EPI != EPI_NONE) {
#pragma unroll
        for (int r = 0; r < NR; ++r) {
            const double tot = block_tree(red[r], lds4);
            if (r < nr && threadIdx.x == 0) epi.partials[rr[r] * epi.pstride + sb] = tot;
        }
    }
}

// ---------------------------------------------------------------------------
// t-streaming kernel (3-D volume a multiple of 256)
// ---------------------------------------------------------------------------
template <class In, class Lk, class Out, class Cl, class Aux, class Ex, int FPM, int MODE, int NR, int EPI, bool GP>
__global__ void __launch_bounds__(256) k_stream(Out out, In in, Lk U, Cl F, Aux aux, Ex ex, Geom g, double kappa_d,
                                                double ccoef_d, int ngroups, int nrhs, int TL, int ZS, EpiArgs epi) {
    using R = typename In::R;
    if (epi.st != nullptr && epi.st->done) return;
    if (epi.gate != nullptr && *epi.gate == 0) return;
    __shared__ double wpart[EPI != EPI_NONE ? NR : 1][EPI != EPI_NONE ? kMaxTL : 1][4];
    const int V3 = g.stride[3];
    const int nblk3 = V3 >> 8;
    const int ntc = g.L[3] / TL;
    const int per = nblk3 / ZS;
    int lb = xcd_block(blockIdx.x, gridDim.x);
    const int rg = lb % ngroups;
    lb /= ngroups;
    const int b3lo = lb % per;
    lb /= per;
    const int tc = lb % ntc;
    const int b3 = (lb / ntc) * per + b3lo;
    const int s3 = (b3 << 8) + threadIdx.x;
    int rr[NR], nr;
    rhs_rows<NR>(rg, nrhs, rr, nr);
    int x[4];
    x[0] = s3 % g.L[0];
    x[1] = (s3 / g.L[0]) % g.L[1];
    x[2] = s3 / g.stride[2];
    const int t0 = tc * TL;
    const R kappa = R(kappa_d), ccoef = R(ccoef_d);
    // carried backward t-hop of the current slice: B = U_t(x-t)^dag P psi(x-t).
    // With FP16 pair products (kHP) it is formed from the integer components
    // like every other hop and carries the neighbour's site scale Bc.
    constexpr bool kHP = kHalfProducts<In, Lk> && !GP;
    using LOp = LinkOp<Lk, R, kHP>;
    cx<R> B[NR][6];
    R Bc[NR];
    {
        x[3] = t0;
        const int s = s3 + V3 * t0;
        bool w;
        const int tm = nbr_bwd(s, x, 3, g, w);
        const bool ng = w && g.phase_neg[3];
        LOp u;
        u.load(U, 3, tm);
#pragma unroll
        for (int r = 0; r < NR; ++r) {
            cx<R> psi[12], h[6];
            if constexpr (kHP)
                in.ld12s(rr[r], tm, psi, Bc[r]);
            else
                in.ld12(rr[r], tm, psi);
            if (GP) {
                if (w && !g.phase_one[3]) phase_spinor(psi, g.iph_re[3], g.iph_im[3]);
            }
            half_project<-FPM, 3>(psi, h);
            if (!GP) {
                if constexpr (MODE == MODE_CLOVER)
                    neg_select(h, ng);  // the clover variants sit at the register limit; keep their allocation
                else
                    neg_uniform(h, ng);
            }
            u.bwd(h, B[r]);
        }
    }
    // FP64 (memory-bound): the t+1 spinor loaded for the forward t-hop is
    // carried in registers as the next step's own spinor instead of being
    // re-read (by then it has usually left L2).  The low-precision kernels
    // are instruction-bound and keep their registers for occupancy.
    constexpr bool kCarry = LMX_CARRY_NEXT && sizeof(R) == 8;
    constexpr int NC = kCarry ? NR : 1;
    cx<R> cur[NC][12], nxt[NC][12];
    R curc[NC], nxtc[NC];
    for (int stp = 0; stp < TL; ++stp) {
        const int t = t0 + stp;
        x[3] = t;
        const int s = s3 + V3 * t;
        if constexpr (kCarry) {
#pragma unroll
            for (int r = 0; r < NR; ++r) {
#pragma unroll
                for (int k = 0; k < 12; ++k) cur[r][k] = nxt[r][k];
                curc[r] = nxtc[r];
            }
        }
        cx<R> acc[NR][12];
#pragma unroll
        for (int r = 0; r < NR; ++r)
#pragma unroll
            for (int k = 0; k < 12; ++k) acc[r][k] = mk<R>(R(0), R(0));
        hop_half<0, FPM, true, NR, GP>(acc, in, U, g, s, x, rr);
        hop_half<0, -FPM, false, NR, GP>(acc, in, U, g, s, x, rr);
        hop_half<1, FPM, true, NR, GP>(acc, in, U, g, s, x, rr);
        hop_half<1, -FPM, false, NR, GP>(acc, in, U, g, s, x, rr);
        hop_half<2, FPM, true, NR, GP>(acc, in, U, g, s, x, rr);
        hop_half<2, -FPM, false, NR, GP>(acc, in, U, g, s, x, rr);
        // t forward from memory, t backward from the carried B (reference order: forward, then backward)
        LOp ut;
        {
            bool w;
            const int tp = nbr_fwd(s, x, 3, g, w);
            const bool ng = w && g.phase_neg[3];
#if LMX_NT_UT
            if constexpr (std::is_same<Lk, LinkF64>::value)
                U.template ld9_nt<R>(3, s, ut.u);  // U_t(x) is used once: its backward use is the carried B
            else
#endif
                ut.load(U, 3, s);
#pragma unroll
            for (int r = 0; r < NR; ++r) {
                cx<R> psi[12], h[6], o[6];
                R c;
                in.ld12s(rr[r], tp, psi, c);
                if constexpr (kCarry) {  // slice t+1 is the next step's own slice: keep it
#pragma unroll
                    for (int k = 0; k < 12; ++k) nxt[r][k] = psi[k];
                    nxtc[r] = c;
                }
                if (GP) {
                    if (w && !g.phase_one[3]) phase_spinor(psi, g.ph_re[3], g.ph_im[3]);
                }
                half_project<FPM, 3>(psi, h);
                if (!GP) {
                if constexpr (MODE == MODE_CLOVER)
                    neg_select(h, ng);  // the clover variants sit at the register limit; keep their allocation
                else
                    neg_uniform(h, ng);
            }
                ut.fwd(h, o);
                reconstruct_add_s<FPM, 3>(acc[r], o, c);
                if constexpr (kHP)
                    reconstruct_add_s<-FPM, 3>(acc[r], B[r], Bc[r]);
                else
                    reconstruct_add<-FPM, 3>(acc[r], B[r]);
            }
        }
        double red[NR];
#pragma unroll
        for (int r = 0; r < NR; ++r) {
            // this slice's own spinor: the output's diagonal term and the next
            // slice's backward t-hop
            cx<R> p[12], h[6];
            if (kCarry && stp > 0) {
#pragma unroll
                for (int k = 0; k < 12; ++k) p[k] = mk<R>(cur[r][k].re * curc[r], cur[r][k].im * curc[r]);
            } else if constexpr (kHP) {
                R c;
                in.ld12s(rr[r], s, p, c);
                half_project<-FPM, 3>(p, h);
                Bc[r] = c;
#pragma unroll
                for (int k = 0; k < 12; ++k) p[k] = mk<R>(p[k].re * c, p[k].im * c);
            } else {
                in.ld12(rr[r], s, p);
            }
            red[r] = finish_site<MODE, EPI>(out, p, F, aux, ex, rr[r], s, g, kappa, ccoef, acc[r], r < nr);
            if constexpr (!kHP) half_project<-FPM, 3>(p, h);
            ut.bwd(h, B[r]);
        }
        if constexpr (EPI != EPI_NONE) {
            // wave sums of this step; the 4-wave combination waits for the end
#pragma unroll
            for (int r = 0; r < NR; ++r) {
                const double v = wave_halve(red[r]);
                if ((threadIdx.x & 63) == 0) wpart[r][stp][threadIdx.x >> 6] = v;
            }
        }
    }
    if constexpr (EPI != EPI_NONE) {
        // one partial per step and RHS: the 256 lexicographic sites b3 + nblk3*t,
        // combined in block_tree's order ((w0 + w2) + (w1 + w3))
        __syncthreads();
        const int tid = threadIdx.x;
        if (tid < NR * TL) {
            const int r = tid / TL, stp = tid - r * TL;
            const double* w = wpart[r][stp];
            const double tot = (w[0] + w[2]) + (w[1] + w[3]);
            int rsel = rr[0];
#pragma unroll
            for (int q = 1; q < NR; ++q)
                if (r == q) rsel = rr[q];
            if (r < nr) epi.partials[(long)rsel * epi.pstride + b3 + nblk3 * (t0 + stp)] = tot;
        }
    }
}

// ---------------------------------------------------------------------------
// LDS-staged streaming kernel (low precision, one RHS per work item).  The
// workgroup walks TL t-slices of its 256-site block like k_stream, but keeps
// the block's spinors at slices t and t+1 in LDS: each step stages slice t+1
// (one coalesced load per site), then every x, y or z neighbour that lies in
// the block, the t+1 neighbour and the site's own spinor are read from LDS;
// the other neighbours come from memory, as do the links.  The arithmetic and the
// hop order are k_stream's, so its results and block partials are
// bit-identical to k_stream at one RHS per work item.
// ---------------------------------------------------------------------------
template <int MU, int PM, bool FWD, typename R, class LOp>
__device__ __forceinline__ void hop_given(cx<R> (&acc)[12], cx<R> (&psi)[12], R c, const LOp& u) {
    cx<R> h[6], o[6];
    half_project<PM, MU>(psi, h);
    if (FWD)
        u.fwd(h, o);
    else
        u.bwd(h, o);
    reconstruct_add_s<PM, MU>(acc, o, c);
    pin_acc(acc);  // as hop_half: the next hop's loads are not hoisted above this one
}

// One spatial direction of the LDS-staged kernel: the forward and the
// backward hop, each neighbour read from the staged 256-site block (sites
// base .. base + 255 of slice t) when it lies in it, else from memory.
// WHERE: 0 test each neighbour, 1 always in the block, 2 never in it.
template <int MU, int FPM, int WHERE, class In, class Lk, typename R>
__device__ __forceinline__ void hop_pair_lds(cx<R> (&acc)[12], const In& in, const typename In::Slot& sc, const Lk& U,
                                             const Geom& g, int r, int s, const int x[4], int base) {
    LinkOp<Lk, R, kHalfProducts<In, Lk>> u;
    cx<R> psi[12];
    R c;
    bool w;
    u.load(U, MU, s);
    const int nf = nbr_fwd(s, x, MU, g, w);
    if (WHERE == 1 || (WHERE == 0 && unsigned(nf - base) < 256u))
        In::lds_ld12s(sc, nf - base, psi, c);
    else
        in.ld12s(r, nf, psi, c);
    hop_given<MU, FPM, true>(acc, psi, c, u);
    const int nb = nbr_bwd(s, x, MU, g, w);
    u.load(U, MU, nb);
    if (WHERE == 1 || (WHERE == 0 && unsigned(nb - base) < 256u))
        In::lds_ld12s(sc, nb - base, psi, c);
    else
        in.ld12s(r, nb, psi, c);
    hop_given<MU, -FPM, false>(acc, psi, c, u);
}

// ROWS: the blocks are whole x-rows of one z-plane (L0 divides 256, 256 / L0
// divides L1), so x neighbours are always in the block and z neighbours never.
template <class In, class Lk, class Out, class Cl, class Aux, class Ex, int FPM, int MODE, int EPI, bool ROWS>
__global__ void __launch_bounds__(256) k_lds(Out out, In in, Lk U, Cl F, Aux aux, Ex ex, Geom g, double kappa_d,
                                             double ccoef_d, int nrhs, int TL, int ZS, EpiArgs epi) {
    using R = float;
    if (epi.st != nullptr && epi.st->done) return;
    if (epi.gate != nullptr && *epi.gate == 0) return;
    __shared__ typename In::Slot slot[2];
    __shared__ double wpart[EPI != EPI_NONE ? kMaxTL : 1][4];
    const int V3 = g.stride[3];
    const int nblk3 = V3 >> 8;
    const int ntc = g.L[3] / TL;
    const int per = nblk3 / ZS;
    int lb = xcd_block(blockIdx.x, gridDim.x);
    const int r = lb % nrhs;
    lb /= nrhs;
    const int b3lo = lb % per;
    lb /= per;
    const int tc = lb % ntc;
    const int b3 = (lb / ntc) * per + b3lo;
    const int tid = threadIdx.x;
    const int s3 = (b3 << 8) + tid;
    const int L0 = g.L[0];
    int x[4];
    x[0] = s3 % L0;
    x[1] = (s3 / L0) % g.L[1];
    x[2] = s3 / g.stride[2];
    const int t0 = tc * TL;
    const R kappa = R(kappa_d), ccoef = R(ccoef_d);
    // carried backward t-hop of the current slice: B = U_t(x-t)^dag P psi(x-t)
    // (with FP16 pair products from the integer components, scale Bc, as k_stream)
    constexpr bool kHP = kHalfProducts<In, Lk>;
    using LOp = LinkOp<Lk, R, kHP>;
    cx<R> B[6];
    R Bc = R(1);
    {
        x[3] = t0;
        const int s = s3 + V3 * t0;
        bool w;
        const int tm = nbr_bwd(s, x, 3, g, w);
        const bool ng = w && g.phase_neg[3];
        LOp u;
        cx<R> psi[12], h[6];
        u.load(U, 3, tm);
        if constexpr (kHP)
            in.ld12s(r, tm, psi, Bc);
        else
            in.ld12(r, tm, psi);
        half_project<-FPM, 3>(psi, h);
        if constexpr (MODE == MODE_CLOVER)
            neg_select(h, ng);
        else
            neg_uniform(h, ng);
        u.bwd(h, B);
    }
    in.stage(r, s3 + V3 * t0, slot[0], tid);
    for (int stp = 0; stp < TL; ++stp) {
        const int t = t0 + stp;
        x[3] = t;
        const int s = s3 + V3 * t;
        const int cur = stp & 1;
        bool wt;
        const int tp = nbr_fwd(s, x, 3, g, wt);
        in.stage(r, tp, slot[cur ^ 1], tid);
        __syncthreads();
        const typename In::Slot& sc = slot[cur];
        cx<R> acc[12];
#pragma unroll
        for (int k = 0; k < 12; ++k) acc[k] = mk<R>(R(0), R(0));
        // x, y, z: each neighbour from the staged block when it lies in it
        const int base = s - tid;  // the block's first site in slice t
        hop_pair_lds<0, FPM, ROWS ? 1 : 0>(acc, in, sc, U, g, r, s, x, base);
        hop_pair_lds<1, FPM, 0>(acc, in, sc, U, g, r, s, x, base);
        hop_pair_lds<2, FPM, ROWS ? 2 : 0>(acc, in, sc, U, g, r, s, x, base);
        // t forward from the staged slice t+1, t backward from the carried B
        LOp ut;
        ut.load(U, 3, s);
        {
            const bool ng = wt && g.phase_neg[3];
            cx<R> psi[12], h[6], o[6];
            R c;
            In::lds_ld12s(slot[cur ^ 1], tid, psi, c);
            half_project<FPM, 3>(psi, h);
            if constexpr (MODE == MODE_CLOVER)
                neg_select(h, ng);
            else
                neg_uniform(h, ng);
            ut.fwd(h, o);
            reconstruct_add_s<FPM, 3>(acc, o, c);
            if constexpr (kHP)
                reconstruct_add_s<-FPM, 3>(acc, B, Bc);
            else
                reconstruct_add<-FPM, 3>(acc, B);
            pin_acc(acc);
        }
        double red;
        {
            cx<R> p[12], h[6];
            R c;
            In::lds_ld12s(sc, tid, p, c);
            // the next slice's carried hop first (U_t(x) is then dead in the epilogue)
            if constexpr (kHP) {
                half_project<-FPM, 3>(p, h);
                Bc = c;
            }
#pragma unroll
            for (int k = 0; k < 12; ++k) p[k] = mk<R>(p[k].re * c, p[k].im * c);
            if constexpr (!kHP) half_project<-FPM, 3>(p, h);
            ut.bwd(h, B);
            pin6f(B);
            red = finish_site<MODE, EPI, true>(out, p, F, aux, ex, r, s, g, kappa, ccoef, acc, true);
        }
        if constexpr (EPI != EPI_NONE) {
            const double v = wave_halve(red);
            if ((tid & 63) == 0) wpart[stp][tid >> 6] = v;
        }
        __syncthreads();  // slot cur is restaged at the next step
    }
    if constexpr (EPI != EPI_NONE) {
        // one partial per step: the 256 lexicographic sites b3 + nblk3*t,
        // combined in block_tree's order ((w0 + w2) + (w1 + w3))
        if (tid < TL) {
            const double* w = wpart[tid];
            epi.partials[(long)r * epi.pstride + b3 + nblk3 * (t0 + tid)] = (w[0] + w[2]) + (w[1] + w[3]);
        }
    }
}

// ---------------------------------------------------------------------------
// RHS-loop kernel (FP16 links): one work item per site of a 256-site block at
// one t-slice.  Its eight links (U_mu(x) and U_mu(x - mu)) are loaded once, as
// the packed FP16 words they are stored in (72 registers), and applied to
// every right-hand side of the launch in turn, so a launch over n RHS reads
// each link once instead of once per RHS group.  Each RHS's arithmetic is
// k_site's (the eight half-direction hops of hop_half in the reference order,
// then finish_site and a per-block partial at block b3 + nblk3 t), so the
// result equals k_site's bit for bit.  Workgroups are ordered t fastest
// after the XCD remap: the t-neighbour slices of a block run on the same XCD
// at the same time.
// ---------------------------------------------------------------------------
struct LinkWords {
    int w[9];
};
__device__ __forceinline__ LinkWords ld_link_words(const LinkF16& U, int mu, int s) {
    const int4 a = ldv(U.p + (size_t)(2 * mu) * U.V, s);
    const int4 b = ldv(U.p + (size_t)(2 * mu + 1) * U.V, s);
    const int c = ldv((const int*)(U.p + (size_t)8 * U.V) + (size_t)mu * U.V, s);
    return LinkWords{{a.x, a.y, a.z, a.w, b.x, b.y, b.z, b.w, c}};
}

template <int MU, int PM, bool FWD, class In>
__device__ __forceinline__ void hop_words(cx<float> (&acc)[12], const In& in, const LinkWords& lw, int r, int s0,
                                          const int x[4], const Geom& g) {
    // the site index "changes" here: the neighbour's address offsets are
    // recomputed per hop (a few integer operations) instead of 36 per-lane
    // offsets being kept live across the RHS loop
    int s = s0;
    asm volatile("" : "+v"(s));
    bool w;
    const int t = FWD ? nbr_fwd(s, x, MU, g, w) : nbr_bwd(s, x, MU, g, w);
    const bool ng = w && g.phase_neg[MU];
    cx<float> psi[12], h[6], o[6];
    // the link words "change" here, after the previous hop's register pin:
    // their use (FP16 -> FP32 conversions, or the pair products' operands)
    // stays with this hop instead of being scheduled at the top of the RHS loop
    int lw9[9];
#pragma unroll
    for (int k = 0; k < 9; ++k) {
        lw9[k] = lw.w[k];
        asm volatile("" : "+v"(lw9[k]));
    }
    LinkOp<LinkF16, float, kHalfProducts<In, LinkF16>> u;
    u.from_words(lw9);
    float c;
    in.ld12s(r, t, psi, c);
    half_project<PM, MU>(psi, h);
    if (MU == 3) neg_select(h, ng);  // low-precision launches: only t can wrap with a sign
    if (FWD)
        u.fwd(h, o);
    else
        u.bwd(h, o);
    reconstruct_add_s<PM, MU>(acc, o, c);
    pin_acc(acc);
}

template <class In, class Out, class Cl, class Aux, class Ex, int FPM, int MODE, int EPI>
__global__ void __launch_bounds__(256) k_rhsloop(Out out, In in, LinkF16 U, Cl F, Aux aux, Ex ex, Geom g,
                                                 double kappa_d, double ccoef_d, int nrhs, EpiArgs epi) {
    if (epi.gate != nullptr && *epi.gate == 0) return;
    __shared__ double lds4[4];
    const int lb = xcd_block(blockIdx.x, gridDim.x);
    const int nblk3 = g.stride[3] >> 8;
    const int t = lb % g.L[3];
    const int b3 = lb / g.L[3];
    const int s = (b3 << 8) + threadIdx.x + g.stride[3] * t;
    int x[4];
    site_coords(s, g, x);
    bool w;
    LinkWords lw[8];
#pragma unroll
    for (int mu = 0; mu < 4; ++mu) {
        lw[2 * mu] = ld_link_words(U, mu, s);
        lw[2 * mu + 1] = ld_link_words(U, mu, nbr_bwd(s, x, mu, g, w));
    }
    const float kappa = float(kappa_d), ccoef = float(ccoef_d);
    const int part = b3 + nblk3 * t;
    for (int r = 0; r < nrhs; ++r) {
        cx<float> acc[12];
#pragma unroll
        for (int k = 0; k < 12; ++k) acc[k] = mk<float>(0.f, 0.f);
        hop_words<0, FPM, true>(acc, in, lw[0], r, s, x, g);
        hop_words<0, -FPM, false>(acc, in, lw[1], r, s, x, g);
        hop_words<1, FPM, true>(acc, in, lw[2], r, s, x, g);
        hop_words<1, -FPM, false>(acc, in, lw[3], r, s, x, g);
        hop_words<2, FPM, true>(acc, in, lw[4], r, s, x, g);
        hop_words<2, -FPM, false>(acc, in, lw[5], r, s, x, g);
        hop_words<3, FPM, true>(acc, in, lw[6], r, s, x, g);
        hop_words<3, -FPM, false>(acc, in, lw[7], r, s, x, g);
        int so = s;  // the own-site offsets likewise, once per RHS
        asm volatile("" : "+v"(so));
        cx<float> psi[12];
        if (MODE != MODE_DONLY) in.ld12(r, so, psi);
        const double red = finish_site<MODE, EPI>(out, psi, F, aux, ex, r, so, g, kappa, ccoef, acc, true);
        if constexpr (EPI != EPI_NONE) {
            const double tot = block_tree(red, lds4);
            if (threadIdx.x == 0) epi.partials[(long)r * epi.pstride + part] = tot;
        }
    }
}

// ---------------------------------------------------------------------------
// FP64 t-streaming kernel with two lanes per site (512-thread workgroups over
// the same 256-site blocks).  Lane j of a pair (j = thread & 1) computes
// half-spinor component j of every projection: it loads spin j and the
// partner spin of the direction (3 - j for mu = x, y; 2 + j for mu = z, t)
// of each neighbour, applies the link to 3 colours, and accumulates spin j
// plus one of spins 2 / 3: spin 3 - j for mu = x, y and spin 2 + j for
// mu = z, t, so the pair swaps those two partial sums once, after mu = y.
// Every accumulator still receives its eight terms in the reference order
// (mu = x fwd, x bwd, ..., t fwd, t bwd) with the same operations, so the
// result is bit-identical to k_stream / k_site; the gain is twice the waves
// with fewer registers each for the latency-bound FP64 stencil.
// ---------------------------------------------------------------------------
template <int MU, int PM, typename R>
__device__ __forceinline__ void half2_project(const cx<R> (&pa)[3], const cx<R> (&pb)[3], int j, cx<R> (&h)[3]) {
    // (MU 0) p1 - PM i p4 | p2 - PM i p3   (MU 1) p1 - PM p4 | p2 + PM p3
    // (MU 2) p1 - PM i p3 | p2 + PM i p4   (MU 3) p1 - PM p3 | p2 - PM p4
    const bool plus = (MU == 1 || MU == 2) && j == 1;
#pragma unroll
    for (int c = 0; c < 3; ++c) {
        cx<R> x = (MU == 0 || MU == 2) ? pmi<PM>(pb[c]) : pmr<PM>(pb[c]);
        x.re = plus ? x.re : -x.re;
        x.im = plus ? x.im : -x.im;
        h[c] = cadd(pa[c], x);  // pa - x computed as pa + (-x): identical in IEEE arithmetic
    }
}

template <int MU, int PM, typename R>
__device__ __forceinline__ void half2_reconstruct(cx<R> (&a0)[3], cx<R> (&a2)[3], const cx<R> (&o)[3], int j) {
    // spin j += o; spin (MU < 2 ? 3 - j : 2 + j) += y with
    // (MU 0) +PM i o  (MU 1) j ? +PM o : -PM o  (MU 2) j ? -PM i o : +PM i o  (MU 3) -PM o
    const bool minus = MU == 3 || (MU == 1 && j == 0) || (MU == 2 && j == 1);
#pragma unroll
    for (int c = 0; c < 3; ++c) {
        a0[c] = cadd(a0[c], o[c]);
        cx<R> y = (MU == 0 || MU == 2) ? pmi<PM>(o[c]) : pmr<PM>(o[c]);
        y.re = minus ? -y.re : y.re;
        y.im = minus ? -y.im : y.im;
        a2[c] = cadd(a2[c], y);  // a - y computed as a + (-y)
    }
}

__device__ __forceinline__ void pin6(cx<double> (&a)[3]) {
#pragma unroll
    for (int k = 0; k < 3; ++k) asm volatile("" : "+v"(a[k].re), "+v"(a[k].im)::"memory");
}

// 3 colours of spin `spin` of field q (one RHS) at site s
__device__ __forceinline__ void ld_spin(const double2* q, int V, int spin, int s, cx<double> (&z)[3]) {
    const double2* b = q + (size_t)(spin * 3) * V;
#pragma unroll
    for (int c = 0; c < 3; ++c) {
        double2 v = b[(size_t)c * V + s];
        z[c] = mk<double>(v.x, v.y);
    }
}

template <int MU, int PM, bool FWD>
__device__ __forceinline__ void hop2(cx<double> (&a0)[3], cx<double> (&a2)[3], const double2* q, const LinkF64& U,
                                     const Geom& g, int s, const int x[4], int j) {
    bool w;
    const int t = FWD ? nbr_fwd(s, x, MU, g, w) : nbr_bwd(s, x, MU, g, w);
    const bool ng = w && g.phase_neg[MU];
    cx<double> u[9], pa[3], pb[3], h[3], o[3];
    U.template ld9<double>(MU, FWD ? s : t, u);
    ld_spin(q, g.V, j, t, pa);
    ld_spin(q, g.V, MU < 2 ? 3 - j : 2 + j, t, pb);
    half2_project<MU, PM>(pa, pb, j, h);
#pragma unroll
    for (int c = 0; c < 3; ++c) {
        h[c].re = ng ? -h[c].re : h[c].re;
        h[c].im = ng ? -h[c].im : h[c].im;
    }
#pragma unroll
    for (int a = 0; a < 3; ++a) {
        if (FWD)
            o[a] = muladdmulti(u[3 * a], h[0], u[3 * a + 1], h[1], u[3 * a + 2], h[2]);
        else
            o[a] = muladdmulti(cconj(u[a]), h[0], cconj(u[3 + a]), h[1], cconj(u[6 + a]), h[2]);
    }
    half2_reconstruct<MU, PM>(a0, a2, o, j);
    pin6(a0);
    pin6(a2);
}

__device__ __forceinline__ double shfl_pair(double v) { return __shfl_xor(v, 1, 64); }

template <int FPM, int MODE, int EPI>
__global__ void __launch_bounds__(512) k_stream2(SpF64Out out, SpF64In in, LinkF64 U, SpF64In aux, Geom g,
                                                 double kappa, int nrhs, int TL, int ZS, EpiArgs epi) {
    static_assert(MODE == MODE_WILSON || MODE == MODE_DONLY, "k_stream2: Wilson and hopping-only operators");
    if (epi.st != nullptr && epi.st->done) return;
    if (epi.gate != nullptr && *epi.gate == 0) return;
    __shared__ double spart[EPI == EPI_PAP ? kMaxTL : 1][EPI == EPI_PAP ? 256 : 1];
    const int V3 = g.stride[3];
    const int nblk3 = V3 >> 8;
    const int ntc = g.L[3] / TL;
    const int per = nblk3 / ZS;
    int lb = xcd_block(blockIdx.x, gridDim.x);
    const int r = lb % nrhs;
    lb /= nrhs;
    const int b3lo = lb % per;
    lb /= per;
    const int tc = lb % ntc;
    const int b3 = (lb / ntc) * per + b3lo;
    const int i = threadIdx.x >> 1, j = threadIdx.x & 1;
    const int s3 = (b3 << 8) + i;
    const int V = g.V;
    const double2* q = in.p + (size_t)r * 12 * V;
    double2* qo = out.p + (size_t)r * 12 * V;
    int x[4];
    x[0] = s3 % g.L[0];
    x[1] = (s3 / g.L[0]) % g.L[1];
    x[2] = s3 / g.stride[2];
    const int t0 = tc * TL;
    // carried backward t-hop of the current slice, component j
    cx<double> B[3];
    {
        x[3] = t0;
        const int s = s3 + V3 * t0;
        bool w;
        const int tm = nbr_bwd(s, x, 3, g, w);
        const bool ng = w && g.phase_neg[3];
        cx<double> u[9], pa[3], pb[3], h[3];
        U.template ld9<double>(3, tm, u);
        ld_spin(q, V, j, tm, pa);
        ld_spin(q, V, 2 + j, tm, pb);
        half2_project<3, -FPM>(pa, pb, j, h);
#pragma unroll
        for (int c = 0; c < 3; ++c) {
            h[c].re = ng ? -h[c].re : h[c].re;
            h[c].im = ng ? -h[c].im : h[c].im;
        }
#pragma unroll
        for (int a = 0; a < 3; ++a)
            B[a] = muladdmulti(cconj(u[a]), h[0], cconj(u[3 + a]), h[1], cconj(u[6 + a]), h[2]);
    }
    cx<double> ca[3], cb[3];  // this slice's spins j and 2 + j (carried from the t-forward hop)
    for (int stp = 0; stp < TL; ++stp) {
        const int t = t0 + stp;
        x[3] = t;
        const int s = s3 + V3 * t;
        if (stp == 0) {
            ld_spin(q, V, j, s, ca);
            ld_spin(q, V, 2 + j, s, cb);
        }
        cx<double> a0[3], a2[3];
#pragma unroll
        for (int c = 0; c < 3; ++c) a0[c] = a2[c] = mk<double>(0.0, 0.0);
        hop2<0, FPM, true>(a0, a2, q, U, g, s, x, j);
        hop2<0, -FPM, false>(a0, a2, q, U, g, s, x, j);
        hop2<1, FPM, true>(a0, a2, q, U, g, s, x, j);
        hop2<1, -FPM, false>(a0, a2, q, U, g, s, x, j);
        // spin 3 - j partial sums change hands: from here on lane j holds spin 2 + j
#pragma unroll
        for (int c = 0; c < 3; ++c) {
            a2[c].re = shfl_pair(a2[c].re);
            a2[c].im = shfl_pair(a2[c].im);
        }
        hop2<2, FPM, true>(a0, a2, q, U, g, s, x, j);
        hop2<2, -FPM, false>(a0, a2, q, U, g, s, x, j);
        cx<double> ut[9], na[3], nb[3];
        {
            bool w;
            const int tp = nbr_fwd(s, x, 3, g, w);
            const bool ng = w && g.phase_neg[3];
            cx<double> h[3], o[3];
            U.template ld9<double>(3, s, ut);
            ld_spin(q, V, j, tp, na);
            ld_spin(q, V, 2 + j, tp, nb);
            half2_project<3, FPM>(na, nb, j, h);
#pragma unroll
            for (int c = 0; c < 3; ++c) {
                h[c].re = ng ? -h[c].re : h[c].re;
                h[c].im = ng ? -h[c].im : h[c].im;
            }
#pragma unroll
            for (int a = 0; a < 3; ++a) o[a] = muladdmulti(ut[3 * a], h[0], ut[3 * a + 1], h[1], ut[3 * a + 2], h[2]);
            half2_reconstruct<3, FPM>(a0, a2, o, j);
            half2_reconstruct<3, -FPM>(a0, a2, B, j);
        }
        // output spins j and 2 + j
        cx<double> r0[3], r2[3];
        if (MODE == MODE_DONLY) {
#pragma unroll
            for (int c = 0; c < 3; ++c) {
                r0[c] = cadd(mk<double>(0.0, 0.0), rmul(0.5, a0[c]));
                r2[c] = cadd(mk<double>(0.0, 0.0), rmul(0.5, a2[c]));
            }
        } else {
#pragma unroll
            for (int c = 0; c < 3; ++c) {
                r0[c] = cadd(ca[c], rmul(-kappa, a0[c]));
                r2[c] = cadd(cb[c], rmul(-kappa, a2[c]));
            }
        }
#pragma unroll
        for (int c = 0; c < 3; ++c) {
            st_out(qo + (size_t)(j * 3 + c) * V + s, r0[c].re, r0[c].im);
            st_out(qo + (size_t)((2 + j) * 3 + c) * V + s, r2[c].re, r2[c].im);
        }
        if constexpr (EPI == EPI_PAP) {
            // Re sum_k conj(p_k) Ap_k in component order k = spin*3 + c: spins 0, 1, 2, 3
            // alternate between the lanes of the pair, the running sum passed along
            cx<double> pa[3], pb[3];
            const double2* qp = aux.p + (size_t)r * 12 * V;
            ld_spin(qp, V, j, s, pa);
            ld_spin(qp, V, 2 + j, s, pb);
            double red = 0.0;
#pragma unroll
            for (int half = 0; half < 2; ++half) {
                const cx<double>* pv = half ? pb : pa;
                const cx<double>* rv = half ? r2 : r0;
                // lane 0 adds spin 2*half, then lane 1 adds spin 2*half + 1
#pragma unroll
                for (int turn = 0; turn < 2; ++turn) {
                    if (j == turn) {
#pragma unroll
                        for (int c = 0; c < 3; ++c) {
                            cx<double> tt = cmul(cconj(pv[c]), rv[c]);
                            red = red + tt.re;
                        }
                    }
                    red = shfl_pair(red);
                }
            }
            // after the last pass lane 0 holds the site's sum
            if (j == 0) spart[stp][i] = red;
        }
        // next step: this slice's backward t-hop from its own spins, and the carry
        {
            cx<double> h[3];
            half2_project<3, -FPM>(ca, cb, j, h);
#pragma unroll
            for (int a = 0; a < 3; ++a)
                B[a] = muladdmulti(cconj(ut[a]), h[0], cconj(ut[3 + a]), h[1], cconj(ut[6 + a]), h[2]);
#pragma unroll
            for (int c = 0; c < 3; ++c) {
                ca[c] = na[c];
                cb[c] = nb[c];
            }
        }
    }
    if constexpr (EPI == EPI_PAP) {
        // the per-step block trees of block_tree(): wave halving over each 64-site
        // quarter, then (q0 + q2) + (q1 + q3); wave w takes steps w, w + 8, ...
        __syncthreads();
        const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
        for (int stp = wave; stp < TL; stp += 8) {
            double qs[4];
#pragma unroll
            for (int qq = 0; qq < 4; ++qq) qs[qq] = wave_halve(spart[stp][qq * 64 + lane]);
            if (lane == 0)
                epi.partials[(long)r * epi.pstride + b3 + nblk3 * (t0 + stp)] = (qs[0] + qs[2]) + (qs[1] + qs[3]);
        }
    }
}

// ---------------------------------------------------------------------------
// Launch shape and dispatch
// ---------------------------------------------------------------------------
struct Shape {
    bool stream;
    int TL, ZS;
};

// t-streaming when the 3-D volume splits into 256-site blocks; TL and ZS keep
// enough workgroups for the chip and keep one step's working set per XCD
// small.  A context override (tune_tl / tune_zs) wins when it divides, then
// the caller's (autotuned) hints.
inline Shape pick_shape(const lmx_context* ctx, int ngroups, int min_blocks, int tl_hint, int zs_hint = 0) {
    Shape sh{false, 1, 1};
    const Geom& g = ctx->g;
    const long V3 = (long)g.L[0] * g.L[1] * g.L[2];
    if (V3 % 256 != 0) return sh;
    const int nblk3 = (int)(V3 / 256);
    sh.stream = true;
    if (ctx->tune_tl > 0 && ctx->tune_tl <= kMaxTL && g.L[3] % ctx->tune_tl == 0) {
        sh.TL = ctx->tune_tl;
    } else if (tl_hint > 0 && tl_hint <= kMaxTL && g.L[3] % tl_hint == 0) {
        sh.TL = tl_hint;
    } else {
        const int cand[5] = {16, 8, 4, 2, 1};
        for (int c : cand) {
            if (g.L[3] % c) continue;
            if ((long)nblk3 * (g.L[3] / c) * ngroups >= min_blocks || c == 1) {
                sh.TL = c;
                break;
            }
        }
    }
    if (ctx->tune_zs > 0 && nblk3 % ctx->tune_zs == 0) {
        sh.ZS = ctx->tune_zs;
    } else if (zs_hint > 0 && nblk3 % zs_hint == 0) {
        sh.ZS = zs_hint;
    } else {
        const int zs_cand[3] = {4, 2, 1};
        for (int z : zs_cand)
            if (nblk3 % z == 0) {
                sh.ZS = z;
                break;
            }
    }
    return sh;
}

// GP (complex boundary phases) runs in the per-site kernel only; SITE_NG
// instantiates the per-site kernel for +-1 phases too (else the GP kernel
// covers lattices whose 3-D volume is not a multiple of 256).  The fused
// epilogues exist where D^dagger D uses them: EPI_PAP and EPI_XR on
// D^dagger, EPI_TP on D.
template <class In, class Lk, class Out, class Cl, class Aux, class Ex, int NR, int EPI, bool SITE_NG, bool DONLY_OK>
int dispatch(lmx_context* ctx, Out out, In in, Lk U, Cl F, Aux aux, const Ex& ex, double kappa, double ccoef,
             int ngroups, int nrhs, int flags, bool clover, const EpiArgs& epi, int min_blocks, int tl_hint = 0,
             int zs_hint = 0) {
    const bool dag = (flags & LMX_DAGGER) != 0;
    const bool donly = (flags & LMX_DONLY) != 0;
    const Shape sh = pick_shape(ctx, ngroups, min_blocks, tl_hint, zs_hint);
    const int nb_site = nblocks(ctx->V) * ngroups;
    const int nb_stream = (int)(ctx->V / 256 / sh.TL) * ngroups;
    const Geom& gg = ctx->g;
    const bool general = gg.general != 0 ||
                         (kSpatialOne<typename In::R> && !(gg.phase_one[0] && gg.phase_one[1] && gg.phase_one[2]));
    hipStream_t st = ctx->stream;
    // FP64 streaming Wilson / hopping-only launches use the two-lanes-per-site kernel
    constexpr bool kTwoLane = LMX_TWO_LANE && std::is_same<In, SpF64In>::value && NR == 1;
    constexpr bool kOnD = EPI == EPI_TP;
    if (EPI != EPI_NONE && (donly || dag == kOnD)) {
        set_error("fused reduction %d not built for this operator (p.Ap and x/r updates on D^dagger, |D p|^2 on D)",
                  EPI);
        return LMX_ERR_UNSUPPORTED;
    }
    if (donly && !DONLY_OK) {
        set_error("hopping-only operator is built for FP64 and FP32 fields only");
        return LMX_ERR_UNSUPPORTED;
    }
#define LMX_GO(FPM, MODE)                                                                                      \
    do {                                                                                                       \
        if constexpr (kTwoLane && MODE != MODE_CLOVER && (EPI == EPI_NONE || EPI == EPI_PAP)) {                \
            if (!general && sh.stream) {                                                                       \
                hipLaunchKernelGGL((k_stream2<FPM, MODE, EPI>), dim3(nb_stream), dim3(512), 0, st, out, in, U, aux, \
                                   ctx->g, kappa, nrhs, sh.TL, sh.ZS, epi);                                    \
                break;                                                                                         \
            }                                                                                                  \
        }                                                                                                      \
        if (general || (!sh.stream && !SITE_NG))                                                               \
            hipLaunchKernelGGL((k_site<In, Lk, Out, Cl, Aux, Ex, FPM, MODE, NR, EPI, true>), dim3(nb_site),    \
                               dim3(256), 0, st, out, in, U, F, aux, ex, ctx->g, kappa, ccoef, ngroups, nrhs, epi); \
        else if (sh.stream)                                                                                    \
            hipLaunchKernelGGL((k_stream<In, Lk, Out, Cl, Aux, Ex, FPM, MODE, NR, EPI, false>), dim3(nb_stream), \
                               dim3(256), 0, st, out, in, U, F, aux, ex, ctx->g, kappa, ccoef, ngroups, nrhs,  \
                               sh.TL, sh.ZS, epi);                                                             \
        else if constexpr (SITE_NG)                                                                            \
            hipLaunchKernelGGL((k_site<In, Lk, Out, Cl, Aux, Ex, FPM, MODE, NR, EPI, false>), dim3(nb_site),   \
                               dim3(256), 0, st, out, in, U, F, aux, ex, ctx->g, kappa, ccoef, ngroups, nrhs, epi); \
    } while (0)
    if constexpr (EPI == EPI_PAP || EPI == EPI_XR || EPI == EPI_TRUE) {
        if (clover)
            LMX_GO(1, MODE_CLOVER);
        else
            LMX_GO(1, MODE_WILSON);
    } else if constexpr (kOnD) {
        if (clover)
            LMX_GO(-1, MODE_CLOVER);
        else
            LMX_GO(-1, MODE_WILSON);
    } else {
        if (donly) {
            if constexpr (DONLY_OK) {
                if (dag)
                    LMX_GO(1, MODE_DONLY);
                else
                    LMX_GO(-1, MODE_DONLY);
            }
        } else if (clover) {
            if (dag)
                LMX_GO(1, MODE_CLOVER);
            else
                LMX_GO(-1, MODE_CLOVER);
        } else {
            if (dag)
                LMX_GO(1, MODE_WILSON);
            else
                LMX_GO(-1, MODE_WILSON);
        }
    }
#undef LMX_GO
    LMX_HIP(hipGetLastError());
    return LMX_OK;
}

// The EPI_XR fields of a launch starting at RHS r0.
template <class In, class Out, int APPLY>
XrEx<In, Out, APPLY> make_xr(const LoArgs& a, int V, int r0) {
    XrEx<In, Out, APPLY> ex{In::at(a.r, a.r_norm, V, r0),
                            Out::at(a.r, a.r_norm, V, r0),
                            SpX4In::at(a.x, nullptr, V, r0),
                            SpX4Out::at(a.x, nullptr, V, r0),
                            {},
                            a.alpha + (long)r0 * a.astride,
                            a.alpha_q + (long)r0 * a.astride,
                            a.astride,
                            a.done + (long)r0 * a.dstride,
                            a.pmask + (long)r0 * a.dstride,
                            a.dstride,
                            a.apply_x};
    for (int j = 0; j < kXBatch - 1; ++j)
        ex.pq[j] = In::at(a.pq[j] ? a.pq[j] : a.aux, a.pq[j] ? a.pq_norm[j] : a.aux_norm, V, r0);
    return ex;
}

// The EPI_TP fields of a launch starting at RHS r0 (t is the launch's out).
template <class In, class Out>
TpEx<In, Out> make_tp(const LoArgs& a, int V, int r0) {
    return TpEx<In, Out>{In::at(a.out, a.out_norm, V, r0),
                         In::at(a.pprev, a.pprev_norm, V, r0),
                         Out::at(a.pnext, a.pnext_norm, V, r0),
                         a.beta + (long)r0 * a.astride,
                         a.astride,
                         a.done + (long)r0 * a.dstride,
                         a.dstride};
}

// One low-precision (In, Out, Lk, Cl, NR) combination: (rhs_block >> 4) & 0xff
// right-hand sides per launch, NR per work item, (rhs_block >> 12) & 0xff
// t-slices per streaming workgroup, (rhs_block >> 20) & 0xf z-ranges (the
// 3-D share of one XCD's slab; 0 automatic); a.epi selects the fused
// epilogue.  Instantiated per combination in wilson_lo.hip.
template <class In, class Out, class Lk, class Cl, int NR>
int launch_lo_nr(lmx_context* ctx, const LoArgs& a, Lk U, Cl F, bool clover) {
    const double ccoef = (-a.kappa) * a.csw;
    int per_launch = (a.rhs_block >> 4) & 0xff;
    if (per_launch <= 0) per_launch = 1;
    const int tl = (a.rhs_block >> 12) & 0xff;
    const int zs = (a.rhs_block >> 20) & 0xf;
    const int V = (int)ctx->V;
    for (int r0 = 0; r0 < a.nrhs; r0 += per_launch) {
        const int n = std::min(per_launch, a.nrhs - r0);
        const int ngroups = (n + NR - 1) / NR;
        Out o = Out::at(a.out, a.out_norm, V, r0);
        In i = In::at(a.in, a.in_norm, V, r0);
        EpiArgs epi{a.partials ? a.partials + (long)r0 * a.pstride : nullptr, a.pstride, const_cast<CGState*>(a.st),
                    a.gate};
        const NoEx none{};
        if (a.epi == EPI_XR) {
            In p = In::at(a.aux, a.aux_norm, V, r0);
            if (a.apply_x) {
                const XrEx<In, Out, 1> ex = make_xr<In, Out, 1>(a, V, r0);
                LMX_TRY((dispatch<In, Lk, Out, Cl, In, XrEx<In, Out, 1>, NR, EPI_XR, false, false>(
                    ctx, o, i, U, F, p, ex, a.kappa, ccoef, ngroups, n, a.flags, clover, epi, 512, tl, zs)));
            } else {
                const XrEx<In, Out, 0> ex = make_xr<In, Out, 0>(a, V, r0);
                LMX_TRY((dispatch<In, Lk, Out, Cl, In, XrEx<In, Out, 0>, NR, EPI_XR, false, false>(
                    ctx, o, i, U, F, p, ex, a.kappa, ccoef, ngroups, n, a.flags, clover, epi, 512, tl, zs)));
            }
        } else if (a.epi == EPI_TP) {
            const TpEx<In, Out> ex = make_tp<In, Out>(a, V, r0);
            LMX_TRY((dispatch<In, Lk, Out, Cl, In, TpEx<In, Out>, NR, EPI_TP, false, false>(
                ctx, o, i, U, F, i, ex, a.kappa, ccoef, ngroups, n, a.flags, clover, epi, 512, tl, zs)));
        } else if (a.epi == EPI_NONE) {
            // the single-precision user apply (FP32 spinors and links) also
            // serves the hopping-only operator
            constexpr bool kDonly = std::is_same<In, SpF32In>::value && std::is_same<Lk, LinkF32>::value && NR == 1;
            LMX_TRY((dispatch<In, Lk, Out, Cl, In, NoEx, NR, EPI_NONE, false, kDonly>(
                ctx, o, i, U, F, i, none, a.kappa, ccoef, ngroups, n, a.flags, clover, epi, 512, tl, zs)));
        } else {
            set_error("fused epilogue %d is not built for low-precision spinors", a.epi);
            return LMX_ERR_UNSUPPORTED;
        }
    }
    return LMX_OK;
}

// The RHS-loop shape (rhs_block bit kRhsLoop): (rhs_block >> 4) & 0xff RHS
// per launch, all of them looped over inside each work item (FP16 links,
// lattices whose 3-D volume is a multiple of 256 and whose phases are +-1;
// others fall back to the per-RHS shapes).

template <class In, class Out, class Cl>
int launch_lo_loop(lmx_context* ctx, const LoArgs& a, LinkF16 U, Cl F, bool clover) {
    const double ccoef = (-a.kappa) * a.csw;
    int per_launch = (a.rhs_block >> 4) & 0xff;
    if (per_launch <= 0) per_launch = a.nrhs;
    const int V = (int)ctx->V;
    const bool dag = (a.flags & LMX_DAGGER) != 0;
    const int nb = (int)(ctx->V / 256);
    hipStream_t st = ctx->stream;
    for (int r0 = 0; r0 < a.nrhs; r0 += per_launch) {
        const int n = std::min(per_launch, a.nrhs - r0);
        Out o = Out::at(a.out, a.out_norm, V, r0);
        In i = In::at(a.in, a.in_norm, V, r0);
        EpiArgs epi{a.partials ? a.partials + (long)r0 * a.pstride : nullptr, a.pstride, nullptr, a.gate};
        if (a.epi == EPI_XR) {
            In p = In::at(a.aux, a.aux_norm, V, r0);
            if (!dag) {
                set_error("the fused r / x update runs on D^dagger");
                return LMX_ERR_UNSUPPORTED;
            }
#define LMX_LOOP_XR(AP)                                                                                         \
    do {                                                                                                        \
        const XrEx<In, Out, AP> ex = make_xr<In, Out, AP>(a, V, r0);                                            \
        if (clover)                                                                                             \
            hipLaunchKernelGGL((k_rhsloop<In, Out, Cl, In, XrEx<In, Out, AP>, 1, MODE_CLOVER, EPI_XR>), dim3(nb), \
                               dim3(256), 0, st, o, i, U, F, p, ex, ctx->g, a.kappa, ccoef, n, epi);            \
        else                                                                                                    \
            hipLaunchKernelGGL((k_rhsloop<In, Out, Cl, In, XrEx<In, Out, AP>, 1, MODE_WILSON, EPI_XR>), dim3(nb), \
                               dim3(256), 0, st, o, i, U, F, p, ex, ctx->g, a.kappa, ccoef, n, epi);            \
    } while (0)
            if (a.apply_x)
                LMX_LOOP_XR(1);
            else
                LMX_LOOP_XR(0);
#undef LMX_LOOP_XR
        } else if (a.epi == EPI_TP) {
            if (dag) {
                set_error("the direction update runs on D");
                return LMX_ERR_UNSUPPORTED;
            }
            const TpEx<In, Out> ex = make_tp<In, Out>(a, V, r0);
            if (clover)
                hipLaunchKernelGGL((k_rhsloop<In, Out, Cl, In, TpEx<In, Out>, -1, MODE_CLOVER, EPI_TP>), dim3(nb),
                                   dim3(256), 0, st, o, i, U, F, i, ex, ctx->g, a.kappa, ccoef, n, epi);
            else
                hipLaunchKernelGGL((k_rhsloop<In, Out, Cl, In, TpEx<In, Out>, -1, MODE_WILSON, EPI_TP>), dim3(nb),
                                   dim3(256), 0, st, o, i, U, F, i, ex, ctx->g, a.kappa, ccoef, n, epi);
        } else {
            set_error("the RHS-loop shape is built for the mixed solver's inner stencils");
            return LMX_ERR_UNSUPPORTED;
        }
        LMX_HIP(hipGetLastError());
    }
    return LMX_OK;
}

// The LDS-staged shape (rhs_block bit kLdsStage): one RHS per work item,
// (rhs_block >> 4) & 0xff RHS per launch, t-slices per workgroup and z-ranges
// as for launch_lo_nr; built for the mixed solver's two inner stencils (D with
// the direction update, D^dagger with the r / x update) on lattices that pass
// lds_ok.
template <class In, class Out, class Lk, class Cl>
int launch_lo_lds(lmx_context* ctx, const LoArgs& a, Lk U, Cl F, bool clover) {
    const double ccoef = (-a.kappa) * a.csw;
    int per_launch = (a.rhs_block >> 4) & 0xff;
    if (per_launch <= 0) per_launch = a.nrhs;
    const int tl = (a.rhs_block >> 12) & 0xff;
    const int zs = (a.rhs_block >> 20) & 0xf;
    const int V = (int)ctx->V;
    const bool dag = (a.flags & LMX_DAGGER) != 0;
    hipStream_t st = ctx->stream;
    const Geom& gg = ctx->g;
    const bool rows = gg.L[0] <= 256 && 256 % gg.L[0] == 0 && gg.L[1] % (256 / gg.L[0]) == 0;
    if (!((a.epi == EPI_TP && !dag) || (a.epi == EPI_XR && dag))) {
        set_error("the LDS-staged shape is built for the mixed solver's inner stencils");
        return LMX_ERR_UNSUPPORTED;
    }
    for (int r0 = 0; r0 < a.nrhs; r0 += per_launch) {
        const int n = std::min(per_launch, a.nrhs - r0);
        const Shape sh = pick_shape(ctx, n, 512, tl, zs);
        const int nb = (int)(ctx->V / 256 / sh.TL) * n;
        Out o = Out::at(a.out, a.out_norm, V, r0);
        In i = In::at(a.in, a.in_norm, V, r0);
        EpiArgs epi{a.partials ? a.partials + (long)r0 * a.pstride : nullptr, a.pstride, nullptr, a.gate};
#define LMX_LDS_ONE(EX, FPM, EPI, MODE, ROWS)                                                              \
    hipLaunchKernelGGL((k_lds<In, Lk, Out, Cl, In, EX, FPM, MODE, EPI, ROWS>), dim3(nb), dim3(256), 0, st, o, i, \
                       U, F, p, ex, ctx->g, a.kappa, ccoef, n, sh.TL, sh.ZS, epi)
#define LMX_LDS_GO(EX, FPM, EPI)                           \
    do {                                                  \
        if (clover && rows)                               \
            LMX_LDS_ONE(EX, FPM, EPI, MODE_CLOVER, true);  \
        else if (clover)                                  \
            LMX_LDS_ONE(EX, FPM, EPI, MODE_CLOVER, false); \
        else if (rows)                                    \
            LMX_LDS_ONE(EX, FPM, EPI, MODE_WILSON, true);  \
        else                                              \
            LMX_LDS_ONE(EX, FPM, EPI, MODE_WILSON, false); \
    } while (0)
        if (a.epi == EPI_TP) {
            using EX = TpEx<In, Out>;
            const EX ex = make_tp<In, Out>(a, V, r0);
            const In p = i;
            LMX_LDS_GO(EX, -1, EPI_TP);
        } else if (a.apply_x) {
            using EX = XrEx<In, Out, 1>;
            const EX ex = make_xr<In, Out, 1>(a, V, r0);
            const In p = In::at(a.aux, a.aux_norm, V, r0);
            LMX_LDS_GO(EX, 1, EPI_XR);
        } else {
            using EX = XrEx<In, Out, 0>;
            const EX ex = make_xr<In, Out, 0>(a, V, r0);
            const In p = In::at(a.aux, a.aux_norm, V, r0);
            LMX_LDS_GO(EX, 1, EPI_XR);
        }
#undef LMX_LDS_GO
#undef LMX_LDS_ONE
        LMX_HIP(hipGetLastError());
    }
    return LMX_OK;
}

// Low-precision combinations built: (spinor, links) = (F32, F32), (F32, F16), (H16, F16).
#define LMX_LO_COMBO_0(X, NR) X(SpF32In, SpF32Out, LinkF32, F32Clover, NR)
#define LMX_LO_COMBO_1(X, NR) X(SpF32In, SpF32Out, LinkF16, F16Clover, NR)
#define LMX_LO_COMBO_2(X, NR) X(SpH16In, SpH16Out, LinkF16, F16Clover, NR)
#define LMX_LO_COMBOS(X, NR) LMX_LO_COMBO_0(X, NR) LMX_LO_COMBO_1(X, NR) LMX_LO_COMBO_2(X, NR)
#define LMX_LO_ALL(X) LMX_LO_COMBOS(X, 1) LMX_LO_COMBOS(X, 2) LMX_LO_COMBOS(X, 4)

}  // namespace lmx
