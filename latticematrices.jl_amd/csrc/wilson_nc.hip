// wilson_nc.hip -- FP64 Wilson(-clover) operator, field strength and link
// pullback for NC != 3 (NC = 2 built), following the reference's generic
// kernels in their operation order, so that they are bit-identical to them:
//   kernel_WilsonDiracOperator4D!           src/Operators/WilsonDiracOperator.jl:80-147
//   kernel_adjoint_WilsonDiracOperator4D!   :444-511
//   kernel_WilsonDiracOperator4D_Donly!     :554-621 (and its adjoint :665-733)
//   mul_op(Oneγ{pm,mu}, ...)                src/Operators/Operators.jl:71-133
//   _clover_q_element_halo + store          src/Operators/WilsonDiracCloverOperator.jl:57-118
//   kernel_add_clover_term!                 :481-525
//   _kernel_wilson_link_pullback_direction! ext/AD/WilsonDiracOperator.jl:27-53
// (For NC = 3 the reference runs its half-spin kernels; so does this library,
// in wilson_kernels.h.)  One work item per site and right-hand side: these
// are small-NC fields, outside the NC = 3 production path, and their kernels
// keep the reference's loop order (for ic, for jc, forward hops mu = 1..4,
// backward hops mu = 1..4) rather than a register-tiled form.
//
// Layouts: spinor double2 [nrhs][4 NC][V], component spin*NC + colour (the
// reference's A[NC,4,...] element order), or the reference's padded storage
// itself (halo not read, wrap in-kernel); links double2 [4][NC NC][V],
// element row*NC + col; clover planes double2 [6][NC NC][V] full matrices.
#include "lmx_internal.h"

namespace lmx {

enum { NC_WILSON = 0, NC_DONLY = 1, NC_CLOVER = 2 };

// (1 + PM gamma_mu) x for the four spin components of one colour (mul_op).
template <int PM, int MU>
__device__ __forceinline__ void mul_op4(const cx<double> (&x)[4], cx<double> (&v)[4]) {
    auto im = [](cx<double> z) { return mk<double>(-z.im, z.re); };  // im * z
    const cx<double> x1 = x[0], x2 = x[1], x3 = x[2], x4 = x[3];
    if (PM > 0) {
        if (MU == 0) { v[0] = csub(x1, im(x4)); v[1] = csub(x2, im(x3)); v[2] = cadd(x3, im(x2)); v[3] = cadd(x4, im(x1)); }
        if (MU == 1) { v[0] = csub(x1, x4); v[1] = cadd(x2, x3); v[2] = cadd(x3, x2); v[3] = csub(x4, x1); }
        if (MU == 2) { v[0] = csub(x1, im(x3)); v[1] = cadd(x2, im(x4)); v[2] = cadd(x3, im(x1)); v[3] = csub(x4, im(x2)); }
        if (MU == 3) { v[0] = csub(x1, x3); v[1] = csub(x2, x4); v[2] = csub(x3, x1); v[3] = csub(x4, x2); }
    } else {
        if (MU == 0) { v[0] = cadd(x1, im(x4)); v[1] = cadd(x2, im(x3)); v[2] = csub(x3, im(x2)); v[3] = csub(x4, im(x1)); }
        if (MU == 1) { v[0] = cadd(x1, x4); v[1] = csub(x2, x3); v[2] = csub(x3, x2); v[3] = cadd(x4, x1); }
        if (MU == 2) { v[0] = cadd(x1, im(x3)); v[1] = csub(x2, im(x4)); v[2] = csub(x3, im(x1)); v[3] = cadd(x4, im(x2)); }
        if (MU == 3) { v[0] = cadd(x1, x3); v[1] = cadd(x2, x4); v[2] = cadd(x3, x1); v[3] = cadd(x4, x2); }
    }
}

// Spinor access: library SoA (component-major per right-hand side) or the
// reference's padded column-major storage (site-major, component fastest).
struct NcSoA {
    int V, ncomp;
    __device__ __forceinline__ long idx(int r, int s, const int*, int k) const {
        return ((long)r * ncomp + k) * V + s;
    }
};
struct NcPad {
    int P[4], nw, ncomp;
    long Vp;
    __device__ __forceinline__ long idx(int r, int, const int* x, int k) const {
        const long ps = (long)(x[0] + nw) + (long)P[0] * ((x[1] + nw) + (long)P[1] * ((x[2] + nw) + (long)P[2] * (x[3] + nw)));
        return ((long)r * Vp + ps) * ncomp + k;
    }
};

template <class A>
__device__ __forceinline__ cx<double> ld_sp(const double2* p, const A& a, int r, int s, const int* x, int k) {
    const double2 v = p[a.idx(r, s, x, k)];
    return mk<double>(v.x, v.y);
}

// the neighbour's value as the halo holds it: times the phase (forward wrap)
// or its inverse (backward wrap), skipped where the phase is exactly 1
__device__ __forceinline__ cx<double> phased(cx<double> z, bool wrap, int mu, bool fwd, const Geom& g) {
    if (!wrap || g.phase_one[mu]) return z;
    return fwd ? cmul(z, mk<double>(g.ph_re[mu], g.ph_im[mu])) : cmul(z, mk<double>(g.iph_re[mu], g.iph_im[mu]));
}

__device__ __forceinline__ void nbr_coords(const int x[4], int mu, int dir, const Geom& g, int y[4]) {
    for (int d = 0; d < 4; ++d) y[d] = x[d];
    y[mu] = dir > 0 ? (x[mu] + 1 == g.L[mu] ? 0 : x[mu] + 1) : (x[mu] == 0 ? g.L[mu] - 1 : x[mu] - 1);
}

__device__ __forceinline__ cx<double> ldz(const double2* p, long i) {
    const double2 v = p[i];
    return mk<double>(v.x, v.y);
}

// One hop of every (ic, jc) term: forward (FWD) adds (coef U_mu(x)[ic,jc]) v,
// backward adds (coef conj(U_mu(x-mu)[jc,ic])) v, v = mul_op of colour jc.
template <int NC, int PM, int MU, bool FWD, class A>
__device__ __forceinline__ void nc_hop(cx<double> (&res)[4 * NC], const double2* in, const A& a, const double2* U,
                                       const Geom& g, int r, int s, const int x[4], int ic, int jc, double coef) {
    int y[4];
    nbr_coords(x, MU, FWD ? 1 : -1, g, y);
    const bool wrap = FWD ? x[MU] == g.L[MU] - 1 : x[MU] == 0;
    const int t = y[0] + g.stride[1] * y[1] + g.stride[2] * y[2] + g.stride[3] * y[3];
    cx<double> xv[4], v[4];
#pragma unroll
    for (int sp = 0; sp < 4; ++sp) xv[sp] = phased(ld_sp(in, a, r, t, y, sp * NC + jc), wrap, MU, FWD, g);
    mul_op4<PM, MU>(xv, v);
    const cx<double> u = FWD ? rmul(coef, ldz(U, (long)(MU * NC * NC + ic * NC + jc) * g.V + s))
                             : rmul(coef, cconj(ldz(U, (long)(MU * NC * NC + jc * NC + ic) * g.V + t)));
#pragma unroll
    for (int ia = 0; ia < 4; ++ia) res[ia * NC + ic] = cadd(res[ia * NC + ic], cmul(u, v[ia]));
}

// kernel_add_clover_term! for NC colours (full planes F[pl][row*NC+col][V])
template <int NC, class A>
__device__ __forceinline__ void nc_clover_add(cx<double> (&res)[4 * NC], const double2* F, const double2* in,
                                              const A& a, int r, int s, const int x[4], int V, double coefficient) {
    for (int color = 0; color < NC; ++color) {
        cx<double> v1 = mk(0.0, 0.0), v2 = mk(0.0, 0.0), v3 = mk(0.0, 0.0), v4 = mk(0.0, 0.0);
        for (int ic = 0; ic < NC; ++ic) {
            cx<double> f[6];
#pragma unroll
            for (int pl = 0; pl < 6; ++pl) f[pl] = ldz(F, (long)(pl * NC * NC + color * NC + ic) * V + s);
            const cx<double> f12 = f[0], f13 = f[1], f14 = f[2], f23 = f[3], f24 = f[4], f34 = f[5];
            auto im = [](cx<double> z) { return mk<double>(-z.im, z.re); };
            const cx<double> dp = im(cadd(f12, f34)), dm = im(csub(f12, f34));
            const cx<double> odp = im(cadd(f14, f23)), odm = im(csub(f23, f14));
            const cx<double> p1 = ld_sp(in, a, r, s, x, 0 * NC + ic), p2 = ld_sp(in, a, r, s, x, 1 * NC + ic);
            const cx<double> p3 = ld_sp(in, a, r, s, x, 2 * NC + ic), p4 = ld_sp(in, a, r, s, x, 3 * NC + ic);
            v1 = cadd(v1, cadd(cmul(dp, p1), cmul(cadd(cadd(cneg(f13), odp), f24), p2)));
            v2 = cadd(v2, cadd(cmul(cneg(dp), p2), cmul(csub(cadd(f13, odp), f24), p1)));
            v3 = cadd(v3, cadd(cmul(dm, p3), cmul(csub(cadd(cneg(f13), odm), f24), p4)));
            v4 = cadd(v4, cadd(cmul(cneg(dm), p4), cmul(cadd(cadd(f13, odm), f24), p3)));
        }
        res[0 * NC + color] = cadd(res[0 * NC + color], rmul(coefficient, v1));
        res[1 * NC + color] = cadd(res[1 * NC + color], rmul(coefficient, v2));
        res[2 * NC + color] = cadd(res[2 * NC + color], rmul(coefficient, v3));
        res[3 * NC + color] = cadd(res[3 * NC + color], rmul(coefficient, v4));
    }
}

// out = D in for right-hand side blockIdx.y.  FPM: projector sign of the
// forward hops (-1 D, +1 D^dagger); NC_WILSON / NC_DONLY / NC_CLOVER.
template <int NC, int FPM, int MODE, class A>
__global__ void __launch_bounds__(256) k_wilson_nc(double2* __restrict__ out, const double2* __restrict__ in,
                                                   const double2* __restrict__ U, const double2* __restrict__ F,
                                                   Geom g, A a, double coef, double ccoef, const CGState* st,
                                                   const int* gate) {
    if (st != nullptr && st->done) return;
    if (gate != nullptr && *gate == 0) return;
    const int s = blockIdx.x * 256 + threadIdx.x;
    const int r = blockIdx.y;
    if (s >= g.V) return;
    int x[4];
    site_coords(s, g, x);
    cx<double> res[4 * NC];
#pragma unroll
    for (int k = 0; k < 4 * NC; ++k) res[k] = MODE == NC_DONLY ? mk(0.0, 0.0) : ld_sp(in, a, r, s, x, k);
    for (int ic = 0; ic < NC; ++ic) {
        for (int jc = 0; jc < NC; ++jc) {
            nc_hop<NC, FPM, 0, true>(res, in, a, U, g, r, s, x, ic, jc, coef);
            nc_hop<NC, FPM, 1, true>(res, in, a, U, g, r, s, x, ic, jc, coef);
            nc_hop<NC, FPM, 2, true>(res, in, a, U, g, r, s, x, ic, jc, coef);
            nc_hop<NC, FPM, 3, true>(res, in, a, U, g, r, s, x, ic, jc, coef);
            nc_hop<NC, -FPM, 0, false>(res, in, a, U, g, r, s, x, ic, jc, coef);
            nc_hop<NC, -FPM, 1, false>(res, in, a, U, g, r, s, x, ic, jc, coef);
            nc_hop<NC, -FPM, 2, false>(res, in, a, U, g, r, s, x, ic, jc, coef);
            nc_hop<NC, -FPM, 3, false>(res, in, a, U, g, r, s, x, ic, jc, coef);
        }
    }
    if constexpr (MODE == NC_CLOVER) nc_clover_add<NC>(res, F, in, a, r, s, x, g.V, ccoef);
#pragma unroll
    for (int k = 0; k < 4 * NC; ++k) out[a.idx(r, s, x, k)] = make_double2(res[k].re, res[k].im);
}

template <int NC, class A>
static int launch_nc_kernel(lmx_context* ctx, double2* out, const double2* in, const double2* U, const double2* F,
                            A a, double kappa, double csw, int nrhs, int flags, const CGState* st, const int* gate) {
    const bool dag = flags & LMX_DAGGER, donly = flags & LMX_DONLY, clover = F != nullptr && !donly;
    // coefficients: -kappa (Wilson), 1/2 (hopping only; its adjoint's local kappa = -1/2), -kappa * cSW (clover)
    const double coef = donly ? 0.5 : -kappa;
    const double ccoef = (-kappa) * csw;
    const dim3 grid((unsigned)((ctx->V + 255) / 256), (unsigned)nrhs);
#define LMX_NC_LAUNCH(FPM, MODE)                                                                             \
    hipLaunchKernelGGL((k_wilson_nc<NC, FPM, MODE, A>), grid, dim3(256), 0, ctx->stream, out, in, U, F, ctx->g, a, \
                       coef, ccoef, st, gate)
    if (donly) {
        if (dag) LMX_NC_LAUNCH(1, NC_DONLY); else LMX_NC_LAUNCH(-1, NC_DONLY);
    } else if (clover) {
        if (dag) LMX_NC_LAUNCH(1, NC_CLOVER); else LMX_NC_LAUNCH(-1, NC_CLOVER);
    } else {
        if (dag) LMX_NC_LAUNCH(1, NC_WILSON); else LMX_NC_LAUNCH(-1, NC_WILSON);
    }
#undef LMX_NC_LAUNCH
    LMX_HIP(hipGetLastError());
    return LMX_OK;
}

int launch_wilson_nc(lmx_context* ctx, const double2* U, const double2* clover, double kappa, double csw,
                     double2* out, const double2* in, int nrhs, int flags, int nw, const CGState* st,
                     const int* gate) {
    if (ctx->V == 0 || nrhs == 0) return LMX_OK;
    LMX_CHECK(ctx->nc == 2, LMX_ERR_UNSUPPORTED, "generic-NC kernels are built for NC = 2 (got %d)", ctx->nc);
    LMX_CHECK(nrhs <= 65535, LMX_ERR_ARG, "too many right-hand sides (%d)", nrhs);
    constexpr int NC = 2;
    if (nw < 0) {
        NcSoA a{(int)ctx->V, 4 * NC};
        return launch_nc_kernel<NC>(ctx, out, in, U, clover, a, kappa, csw, nrhs, flags, st, gate);
    }
    NcPad a;
    a.nw = nw;
    a.ncomp = 4 * NC;
    a.Vp = 1;
    for (int d = 0; d < 4; ++d) {
        a.P[d] = ctx->g.L[d] + 2 * nw;
        a.Vp *= a.P[d];
    }
    return launch_nc_kernel<NC>(ctx, out, in, U, clover, a, kappa, csw, nrhs, flags, st, gate);
}

// ---------------------------------------------------------------------------
// Field strength: F = (Q - Q^dagger) / 8 per plane, the four-leaf Q element
// by element (_clover_q_element_halo); one work item per site and plane.
// ---------------------------------------------------------------------------
__device__ __forceinline__ int shift_site(int s, const Geom& g, int mu, int dir) {
    int x[4], y[4];
    site_coords(s, g, x);
    nbr_coords(x, mu, dir, g, y);
    return y[0] + g.stride[1] * y[1] + g.stride[2] * y[2] + g.stride[3] * y[3];
}

template <int NC>
__device__ __forceinline__ cx<double> lk(const double2* U, int V, int mu, int site, int r, int c) {
    return ldz(U, (long)(mu * NC * NC + r * NC + c) * V + site);
}

template <int NC>
__device__ cx<double> q_element(const double2* U, const Geom& g, int mu, int nu, int x, int xpmu, int xpnu, int xmmu,
                                int xmnu, int xpnu_mmu, int xmmu_mnu, int xpmu_mnu, int row, int col) {
    const int V = g.V;
    cx<double> value = mk(0.0, 0.0);
    for (int a = 0; a < NC; ++a)
        for (int b = 0; b < NC; ++b)
            for (int c = 0; c < NC; ++c) {
                // +mu, +nu, -mu, -nu
                cx<double> t = cmul(cmul(cmul(lk<NC>(U, V, mu, x, row, a), lk<NC>(U, V, nu, xpmu, a, b)),
                                         cconj(lk<NC>(U, V, mu, xpnu, c, b))),
                                    cconj(lk<NC>(U, V, nu, x, col, c)));
                value = cadd(value, t);
                // +nu, -mu, -nu, +mu
                t = cmul(cmul(cmul(lk<NC>(U, V, nu, x, row, a), cconj(lk<NC>(U, V, mu, xpnu_mmu, b, a))),
                              cconj(lk<NC>(U, V, nu, xmmu, c, b))),
                         lk<NC>(U, V, mu, xmmu, c, col));
                value = cadd(value, t);
                // -mu, -nu, +mu, +nu
                t = cmul(cmul(cmul(cconj(lk<NC>(U, V, mu, xmmu, a, row)), cconj(lk<NC>(U, V, nu, xmmu_mnu, b, a))),
                              lk<NC>(U, V, mu, xmmu_mnu, b, c)),
                         lk<NC>(U, V, nu, xmnu, c, col));
                value = cadd(value, t);
                // -nu, +mu, +nu, -mu
                t = cmul(cmul(cmul(cconj(lk<NC>(U, V, nu, xmnu, a, row)), lk<NC>(U, V, mu, xmnu, a, b)),
                              lk<NC>(U, V, nu, xpmu_mnu, b, c)),
                         cconj(lk<NC>(U, V, mu, x, col, c)));
                value = cadd(value, t);
            }
    return value;
}

template <int NC>
__global__ void __launch_bounds__(256) k_clover_build_nc(double2* __restrict__ F, const double2* __restrict__ U,
                                                         Geom g) {
    const int s = blockIdx.x * 256 + threadIdx.x;
    const int pl = blockIdx.y;
    if (s >= g.V) return;
    const int mus[6] = {0, 0, 0, 1, 1, 2}, nus[6] = {1, 2, 3, 2, 3, 3};
    const int mu = mus[pl], nu = nus[pl];
    const int xpmu = shift_site(s, g, mu, 1), xpnu = shift_site(s, g, nu, 1), xmmu = shift_site(s, g, mu, -1),
              xmnu = shift_site(s, g, nu, -1);
    const int xpnu_mmu = shift_site(xpnu, g, mu, -1), xmmu_mnu = shift_site(xmmu, g, nu, -1),
              xpmu_mnu = shift_site(xmnu, g, mu, 1);
    for (int col = 0; col < NC; ++col)
        for (int row = 0; row <= col; ++row) {
            const cx<double> qrc = q_element<NC>(U, g, mu, nu, s, xpmu, xpnu, xmmu, xmnu, xpnu_mmu, xmmu_mnu,
                                                 xpmu_mnu, row, col);
            const cx<double> qcr = row == col ? qrc
                                              : q_element<NC>(U, g, mu, nu, s, xpmu, xpnu, xmmu, xmnu, xpnu_mmu,
                                                              xmmu_mnu, xpmu_mnu, col, row);
            const cx<double> value = rmul(0.125, csub(qrc, cconj(qcr)));
            F[(long)(pl * NC * NC + row * NC + col) * g.V + s] = make_double2(value.re, value.im);
            if (row != col) {
                const cx<double> m = cneg(cconj(value));
                F[(long)(pl * NC * NC + col * NC + row) * g.V + s] = make_double2(m.re, m.im);
            }
        }
}

int launch_clover_build_nc(lmx_context* ctx, const double2* U, double2* F) {
    if (ctx->V == 0) return LMX_OK;
    LMX_CHECK(ctx->nc == 2, LMX_ERR_UNSUPPORTED, "generic-NC kernels are built for NC = 2 (got %d)", ctx->nc);
    hipLaunchKernelGGL(k_clover_build_nc<2>, dim3((unsigned)((ctx->V + 255) / 256), 6), dim3(256), 0, ctx->stream,
                       F, U, ctx->g);
    LMX_HIP(hipGetLastError());
    return LMX_OK;
}

// Full planes [6][V][NC][NC] (lmx_clover_export) from the cache layout.
template <int NC>
__global__ void __launch_bounds__(256) k_clover_export_nc(double2* __restrict__ out, const double2* __restrict__ F,
                                                          int V) {
    const int s = blockIdx.x * 256 + threadIdx.x;
    const int pl = blockIdx.y;
    if (s >= V) return;
    for (int e = 0; e < NC * NC; ++e)
        out[((long)pl * V + s) * NC * NC + e] = F[(long)(pl * NC * NC + e) * V + s];
}

int launch_clover_export_nc(lmx_context* ctx, const double2* F, double2* out) {
    if (ctx->V == 0) return LMX_OK;
    LMX_CHECK(ctx->nc == 2, LMX_ERR_UNSUPPORTED, "generic-NC kernels are built for NC = 2 (got %d)", ctx->nc);
    hipLaunchKernelGGL(k_clover_export_nc<2>, dim3((unsigned)((ctx->V + 255) / 256), 6), dim3(256), 0, ctx->stream,
                       out, F, (int)ctx->V);
    LMX_HIP(hipGetLastError());
    return LMX_OK;
}

// ---------------------------------------------------------------------------
// Link pullback of the hopping term: dU_mu(x)[row, col] += coefficient *
// sum_spin (dres[row](x) conj(((1 - gamma_mu) psi[col])(x + mu))
//           + ((1 + gamma_mu) psi[row])(x) conj(dres[col](x + mu))),
// the neighbour values as the halo holds them.  One work item per site.
// ---------------------------------------------------------------------------
template <int NC, int MU>
__device__ __forceinline__ void nc_pullback_dir(double2* dU, const double2* dres, const double2* psi, const Geom& g,
                                                int s, const int x[4], double coefficient) {
    const NcSoA a{g.V, 4 * NC};
    int y[4];
    nbr_coords(x, MU, 1, g, y);
    const bool wrap = x[MU] == g.L[MU] - 1;
    const int t = y[0] + g.stride[1] * y[1] + g.stride[2] * y[2] + g.stride[3] * y[3];
    for (int row = 0; row < NC; ++row) {
        cx<double> xv[4], minus_psi[4];
#pragma unroll
        for (int sp = 0; sp < 4; ++sp) xv[sp] = ld_sp(psi, a, 0, s, x, sp * NC + row);
        mul_op4<1, MU>(xv, minus_psi);
        for (int col = 0; col < NC; ++col) {
            cx<double> pv[4], plus_psi[4];
#pragma unroll
            for (int sp = 0; sp < 4; ++sp) pv[sp] = phased(ld_sp(psi, a, 0, t, y, sp * NC + col), wrap, MU, true, g);
            mul_op4<-1, MU>(pv, plus_psi);
            cx<double> value = mk(0.0, 0.0);
#pragma unroll
            for (int sp = 0; sp < 4; ++sp) {
                value = cadd(value, cmul(ld_sp(dres, a, 0, s, x, sp * NC + row), cconj(plus_psi[sp])));
                const cx<double> d1 = phased(ld_sp(dres, a, 0, t, y, sp * NC + col), wrap, MU, true, g);
                value = cadd(value, cmul(minus_psi[sp], cconj(d1)));
            }
            const long o = (long)(MU * NC * NC + row * NC + col) * g.V + s;
            const cx<double> nv = cadd(ldz(dU, o), rmul(coefficient, value));
            dU[o] = make_double2(nv.re, nv.im);
        }
    }
}

template <int NC>
__global__ void __launch_bounds__(256) k_pullback_nc(double2* __restrict__ dU, const double2* __restrict__ dres,
                                                     const double2* __restrict__ psi, Geom g, double coefficient) {
    const int s = blockIdx.x * 256 + threadIdx.x;
    if (s >= g.V) return;
    int x[4];
    site_coords(s, g, x);
    nc_pullback_dir<NC, 0>(dU, dres, psi, g, s, x, coefficient);
    nc_pullback_dir<NC, 1>(dU, dres, psi, g, s, x, coefficient);
    nc_pullback_dir<NC, 2>(dU, dres, psi, g, s, x, coefficient);
    nc_pullback_dir<NC, 3>(dU, dres, psi, g, s, x, coefficient);
}

int launch_pullback_nc(lmx_context* ctx, double2* dU, const double2* dres, const double2* psi, double coefficient) {
    if (ctx->V == 0) return LMX_OK;
    LMX_CHECK(ctx->nc == 2, LMX_ERR_UNSUPPORTED, "generic-NC kernels are built for NC = 2 (got %d)", ctx->nc);
    hipLaunchKernelGGL(k_pullback_nc<2>, dim3((unsigned)((ctx->V + 255) / 256)), dim3(256), 0, ctx->stream, dU, dres,
                       psi, ctx->g, coefficient);
    LMX_HIP(hipGetLastError());
    return LMX_OK;
}

}  // namespace lmx
