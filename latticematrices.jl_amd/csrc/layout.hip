// layout.hip -- conversion between the reference's halo-padded column-major
// arrays (LatticeMatrix.A, src/LatticeMatrices_core.jl:361-363) and the
// library's structure-of-arrays layout.
//
// Reference spinor: A[3,4,L1+2nw,L2+2nw,L3+2nw,L4+2nw] complex128, colour
// fastest; links: A[3,3,...] per direction, row fastest.  Export refills the
// halo exactly as set_halo! would: ghost = core value times the boundary
// phase (forward wrap) or its inverse (backward wrap), direction by direction
// (exchange_dim_local!, :707-723), links periodic.
#include "lmx_internal.h"

namespace lmx {

struct Pad {
    int P[4];
    int nw;
    long Vp;
};

__device__ __forceinline__ long padded_site(const int x[4], const Pad& p) {
    return (long)(x[0] + p.nw) + (long)p.P[0] * ((x[1] + p.nw) + (long)p.P[1] * ((x[2] + p.nw) + (long)p.P[2] * (x[3] + p.nw)));
}

__global__ void __launch_bounds__(256) k_spinor_import(double2* __restrict__ dst, const double2* __restrict__ src,
                                                       Geom g, Pad p, int nrhs) {
    const long t = (long)blockIdx.x * 256 + threadIdx.x;
    if (t >= (long)g.V * nrhs) return;
    const int s = (int)(t % g.V);
    const int r = (int)(t / g.V);
    int x[4];
    site_coords(s, g, x);
    const double2* a = src + (long)r * p.Vp * 12 + padded_site(x, p) * 12;
    double2* o = dst + (long)r * 12 * g.V;
#pragma unroll
    for (int k = 0; k < 12; ++k) o[(long)k * g.V + s] = a[k];
}

__global__ void __launch_bounds__(256) k_spinor_export(double2* __restrict__ dst, const double2* __restrict__ src,
                                                       Geom g, Pad p, int nrhs) {
    const long t = (long)blockIdx.x * 256 + threadIdx.x;
    if (t >= p.Vp * nrhs) return;
    const long ps = t % p.Vp;
    const int r = (int)(t / p.Vp);
    int px[4];
    long rem = ps;
    for (int d = 0; d < 4; ++d) {
        px[d] = (int)(rem % p.P[d]);
        rem /= p.P[d];
    }
    int x[4], w[4];
    for (int d = 0; d < 4; ++d) {
        int c = px[d] - p.nw;
        w[d] = 0;
        if (c < 0) {
            c += g.L[d];
            w[d] = -1;
        } else if (c >= g.L[d]) {
            c -= g.L[d];
            w[d] = 1;
        }
        x[d] = c;
    }
    const int s = x[0] + g.stride[1] * x[1] + g.stride[2] * x[2] + g.stride[3] * x[3];
    const double2* in = src + (long)r * 12 * g.V;
    double2* o = dst + (long)r * p.Vp * 12 + ps * 12;
    for (int k = 0; k < 12; ++k) {
        double2 v = in[(long)k * g.V + s];
        cx<double> z = mk<double>(v.x, v.y);
        for (int d = 0; d < 4; ++d) {
            if (w[d] == 0 || g.phase_one[d]) continue;
            cx<double> ph = w[d] > 0 ? mk<double>(g.ph_re[d], g.ph_im[d]) : mk<double>(g.iph_re[d], g.iph_im[d]);
            z = cmul(z, ph);
        }
        o[k] = make_double2(z.re, z.im);
    }
}

struct LinkPtrs {
    const double2* p[4];
};
struct LinkOutPtrs {
    double2* p[4];
};

__global__ void __launch_bounds__(256) k_links_import(double2* __restrict__ dst, LinkPtrs src, Geom g, Pad p) {
    const int s = blockIdx.x * 256 + threadIdx.x;
    if (s >= g.V) return;
    int x[4];
    site_coords(s, g, x);
    const long ps = padded_site(x, p);
    for (int mu = 0; mu < 4; ++mu) {
        const double2* a = src.p[mu] + ps * 9;
        for (int row = 0; row < 3; ++row)
            for (int col = 0; col < 3; ++col) dst[(long)(mu * 9 + row * 3 + col) * g.V + s] = a[row + 3 * col];
    }
}

__global__ void __launch_bounds__(256) k_links_export(LinkOutPtrs dst, const double2* __restrict__ src, Geom g, Pad p) {
    const long ps = (long)blockIdx.x * 256 + threadIdx.x;
    if (ps >= p.Vp) return;
    int x[4];
    long rem = ps;
    for (int d = 0; d < 4; ++d) {
        int c = (int)(rem % p.P[d]) - p.nw;
        rem /= p.P[d];
        if (c < 0) c += g.L[d];
        if (c >= g.L[d]) c -= g.L[d];
        x[d] = c;
    }
    const int s = x[0] + g.stride[1] * x[1] + g.stride[2] * x[2] + g.stride[3] * x[3];
    for (int mu = 0; mu < 4; ++mu) {
        double2* a = dst.p[mu] + ps * 9;
        for (int row = 0; row < 3; ++row)
            for (int col = 0; col < 3; ++col) a[row + 3 * col] = src[(long)(mu * 9 + row * 3 + col) * g.V + s];
    }
}

static Pad make_pad(const lmx_context* ctx, int nw) {
    Pad p;
    p.nw = nw;
    p.Vp = 1;
    for (int d = 0; d < 4; ++d) {
        p.P[d] = ctx->g.L[d] + 2 * nw;
        p.Vp *= p.P[d];
    }
    return p;
}

static int check_nw(const lmx_context* ctx, int nw) {
    LMX_CHECK(nw >= 0, LMX_ERR_ARG, "nw must be non-negative, got %d", nw);
    for (int d = 0; d < 4; ++d)
        LMX_CHECK(nw <= ctx->g.L[d], LMX_ERR_ARG, "halo width nw=%d exceeds local lattice size %d in dimension %d", nw,
                  ctx->g.L[d], d + 1);
    return LMX_OK;
}

int spinor_import(lmx_context* ctx, void* dst, const void* src, int nw, int nrhs) {
    LMX_TRY(check_nw(ctx, nw));
    Pad p = make_pad(ctx, nw);
    const long n = ctx->V * (long)nrhs;
    if (n == 0) return LMX_OK;
    hipLaunchKernelGGL(k_spinor_import, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, ctx->stream, (double2*)dst,
                       (const double2*)src, ctx->g, p, nrhs);
    LMX_HIP(hipGetLastError());
    return LMX_OK;
}

int spinor_export(lmx_context* ctx, void* dst, int nw, const void* src, int nrhs) {
    LMX_TRY(check_nw(ctx, nw));
    Pad p = make_pad(ctx, nw);
    const long n = p.Vp * nrhs;
    if (n == 0) return LMX_OK;
    hipLaunchKernelGGL(k_spinor_export, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, ctx->stream, (double2*)dst,
                       (const double2*)src, ctx->g, p, nrhs);
    LMX_HIP(hipGetLastError());
    return LMX_OK;
}

int links_import(lmx_context* ctx, void* dst, const void* const src[4], int nw) {
    LMX_TRY(check_nw(ctx, nw));
    Pad p = make_pad(ctx, nw);
    LinkPtrs lp;
    for (int mu = 0; mu < 4; ++mu) {
        LMX_CHECK(src[mu] != nullptr, LMX_ERR_ARG, "link array %d is NULL", mu + 1);
        lp.p[mu] = (const double2*)src[mu];
    }
    if (ctx->V == 0) return LMX_OK;
    hipLaunchKernelGGL(k_links_import, dim3(nblocks(ctx->V)), dim3(256), 0, ctx->stream, (double2*)dst, lp, ctx->g, p);
    LMX_HIP(hipGetLastError());
    return LMX_OK;
}

int links_export(lmx_context* ctx, void* const dst[4], int nw, const void* src) {
    LMX_TRY(check_nw(ctx, nw));
    Pad p = make_pad(ctx, nw);
    LinkOutPtrs lp;
    for (int mu = 0; mu < 4; ++mu) {
        LMX_CHECK(dst[mu] != nullptr, LMX_ERR_ARG, "link array %d is NULL", mu + 1);
        lp.p[mu] = (double2*)dst[mu];
    }
    if (p.Vp == 0) return LMX_OK;
    hipLaunchKernelGGL(k_links_export, dim3((unsigned)((p.Vp + 255) / 256)), dim3(256), 0, ctx->stream, lp,
                       (const double2*)src, ctx->g, p);
    LMX_HIP(hipGetLastError());
    return LMX_OK;
}

}  // namespace lmx
