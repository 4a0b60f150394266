// layout.hip -- conversion between the reference's halo-padded column-major
// arrays (LatticeMatrix.A, src/LatticeMatrices_core.jl:361-363) and the
// library's structure-of-arrays layout.
//
// Reference spinor: A[NC,4,L1+2nw,L2+2nw,L3+2nw,L4+2nw] complex128, colour
// fastest; links: A[NC,NC,...] per direction, row fastest (NC the context's).  Export refills the
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

// T2: double2 (ComplexF64 fields) or float2 (ComplexF32), R its real type.
template <typename T2>
__global__ void __launch_bounds__(256) k_spinor_import(T2* __restrict__ dst, const T2* __restrict__ src,
                                                       Geom g, Pad p, int nrhs, int ncomp) {
    const long t = (long)blockIdx.x * 256 + threadIdx.x;
    if (t >= (long)g.V * nrhs) return;
    const int s = (int)(t % g.V);
    const int r = (int)(t / g.V);
    int x[4];
    site_coords(s, g, x);
    const T2* a = src + (long)r * p.Vp * ncomp + padded_site(x, p) * ncomp;
    T2* o = dst + (long)r * ncomp * g.V;
    for (int k = 0; k < ncomp; ++k) o[(long)k * g.V + s] = a[k];
}

template <typename T2, typename R>
__global__ void __launch_bounds__(256) k_spinor_export(T2* __restrict__ dst, const T2* __restrict__ src,
                                                       Geom g, Pad p, int nrhs, int ncomp) {
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
    const T2* in = src + (long)r * ncomp * g.V;
    T2* o = dst + (long)r * p.Vp * ncomp + ps * ncomp;
    for (int k = 0; k < ncomp; ++k) {
        const T2 v = in[(long)k * g.V + s];
        cx<R> z = mk<R>(v.x, v.y);
        for (int d = 0; d < 4; ++d) {
            if (w[d] == 0 || g.phase_one[d]) continue;
            cx<R> ph = w[d] > 0 ? mk<R>(R(g.ph_re[d]), R(g.ph_im[d])) : mk<R>(R(g.iph_re[d]), R(g.iph_im[d]));
            z = cmul(z, ph);
        }
        T2 out;
        out.x = z.re;
        out.y = z.im;
        o[k] = out;
    }
}

struct LinkPtrs {
    const double2* p[4];
};
struct LinkOutPtrs {
    double2* p[4];
};

__global__ void __launch_bounds__(256) k_links_import(double2* __restrict__ dst, LinkPtrs src, Geom g, Pad p, int nc) {
    const int s = blockIdx.x * 256 + threadIdx.x;
    if (s >= g.V) return;
    int x[4];
    site_coords(s, g, x);
    const long ps = padded_site(x, p);
    for (int mu = 0; mu < 4; ++mu) {
        const double2* a = src.p[mu] + ps * nc * nc;
        for (int row = 0; row < nc; ++row)
            for (int col = 0; col < nc; ++col) dst[(long)(mu * nc * nc + row * nc + col) * g.V + s] = a[row + nc * col];
    }
}

__global__ void __launch_bounds__(256) k_links_export(LinkOutPtrs dst, const double2* __restrict__ src, Geom g, Pad p,
                                                      int nc) {
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
        double2* a = dst.p[mu] + ps * nc * nc;
        for (int row = 0; row < nc; ++row)
            for (int col = 0; col < nc; ++col) a[row + nc * col] = src[(long)(mu * nc * nc + row * nc + col) * g.V + s];
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

template <typename T2>
static int spinor_import_t(lmx_context* ctx, void* dst, const void* src, int nw, int nrhs) {
    LMX_TRY(check_nw(ctx, nw));
    Pad p = make_pad(ctx, nw);
    const long n = ctx->V * (long)nrhs;
    if (n == 0) return LMX_OK;
    hipLaunchKernelGGL((k_spinor_import<T2>), dim3((unsigned)((n + 255) / 256)), dim3(256), 0, ctx->stream, (T2*)dst,
                       (const T2*)src, ctx->g, p, nrhs, spinor_ncomp(ctx));
    LMX_HIP(hipGetLastError());
    return LMX_OK;
}

template <typename T2, typename R>
static int spinor_export_t(lmx_context* ctx, void* dst, int nw, const void* src, int nrhs) {
    LMX_TRY(check_nw(ctx, nw));
    Pad p = make_pad(ctx, nw);
    const long n = p.Vp * nrhs;
    if (n == 0) return LMX_OK;
    hipLaunchKernelGGL((k_spinor_export<T2, R>), dim3((unsigned)((n + 255) / 256)), dim3(256), 0, ctx->stream,
                       (T2*)dst, (const T2*)src, ctx->g, p, nrhs, spinor_ncomp(ctx));
    LMX_HIP(hipGetLastError());
    return LMX_OK;
}

int spinor_import(lmx_context* ctx, void* dst, const void* src, int nw, int nrhs) {
    return spinor_import_t<double2>(ctx, dst, src, nw, nrhs);
}
int spinor_export(lmx_context* ctx, void* dst, int nw, const void* src, int nrhs) {
    return spinor_export_t<double2, double>(ctx, dst, nw, src, nrhs);
}
// ComplexF32 fields: LatticeMatrix{ComplexF32}.A <-> float2 [12][V]
int spinor_import_f32(lmx_context* ctx, void* dst, const void* src, int nw, int nrhs) {
    return spinor_import_t<float2>(ctx, dst, src, nw, nrhs);
}
int spinor_export_f32(lmx_context* ctx, void* dst, int nw, const void* src, int nrhs) {
    return spinor_export_t<float2, float>(ctx, dst, nw, src, nrhs);
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
    hipLaunchKernelGGL(k_links_import, dim3(nblocks(ctx->V)), dim3(256), 0, ctx->stream, (double2*)dst, lp, ctx->g, p,
                       ctx->nc);
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
                       (const double2*)src, ctx->g, p, ctx->nc);
    LMX_HIP(hipGetLastError());
    return LMX_OK;
}

// ---------------------------------------------------------------------------
// Generic fields: any NC1 x NC2 matrix per site, Float32 / Float64 /
// ComplexF32 / ComplexF64, element e = ic + NC1 * jc of a site contiguous.
// ---------------------------------------------------------------------------
int make_halo_phases(const lmx_context* ctx, int elem, HaloPhases* hp) {
    const bool real_el = elem == LMX_ELEM_F32 || elem == LMX_ELEM_F64;
    const bool single = elem == LMX_ELEM_F32 || elem == LMX_ELEM_C64;
    for (int d = 0; d < 4; ++d) {
        const double re = ctx->g.ph_re[d], im = ctx->g.ph_im[d];
        hp->skip[d] = ctx->g.phase_one[d];
        LMX_CHECK(!real_el || im == 0.0, LMX_ERR_ARG,
                  "a real-valued field cannot take the complex boundary phase of dimension %d", d + 1);
        LMX_CHECK(!single || hp->skip[d] || ctx->g.phase_neg[d], LMX_ERR_ARG,
                  "single-precision fields with a halo support boundary phases +-1 only (dimension %d)", d + 1);
        hp->plus_re[d] = re;
        hp->plus_im[d] = im;
        if (real_el) {
            hp->minus_re[d] = 1.0 / re;  // inv(x::Float64)
            hp->minus_im[d] = 0.0;
        } else {
            hp->minus_re[d] = ctx->g.iph_re[d];
            hp->minus_im[d] = ctx->g.iph_im[d];
        }
    }
    return LMX_OK;
}

template <class T>
struct C2 {
    T re, im;
};

// LatticeMatrix(A, dim, PEs) on one rank (src/LatticeMatrices_core.jl:405-461):
// padded[ps][e] = core[wrapped site][e] times the phases of the directions the
// ghost wraps across (exchange_dim_local!, :707-723).  One thread per element.
template <class T, bool CPLX>
__global__ void __launch_bounds__(256) k_field_pad(void* __restrict__ dst, const void* __restrict__ src, Geom g,
                                                   Pad p, int ne, HaloPhases ph) {
    const long t = (long)blockIdx.x * 256 + threadIdx.x;
    if (t >= p.Vp * ne) return;
    const long ps = t / ne;
    const int e = (int)(t - ps * ne);
    long rem = ps;
    int x[4], w[4];
    for (int d = 0; d < 4; ++d) {
        int c = (int)(rem % p.P[d]) - p.nw;
        rem /= p.P[d];
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
    const long s = x[0] + (long)g.stride[1] * x[1] + (long)g.stride[2] * x[2] + (long)g.stride[3] * x[3];
    if constexpr (CPLX) {
        C2<T> v = ((const C2<T>*)src)[s * ne + e];
        for (int d = 0; d < 4; ++d) {
            if (w[d] == 0 || ph.skip[d]) continue;
            const T pr = (T)(w[d] > 0 ? ph.plus_re[d] : ph.minus_re[d]);
            const T pi = (T)(w[d] > 0 ? ph.plus_im[d] : ph.minus_im[d]);
            const T nr = v.re * pr - v.im * pi;
            const T ni = v.re * pi + v.im * pr;
            v.re = nr;
            v.im = ni;
        }
        ((C2<T>*)dst)[t] = v;
    } else {
        T v = ((const T*)src)[s * ne + e];
        for (int d = 0; d < 4; ++d) {
            if (w[d] == 0 || ph.skip[d]) continue;
            v = v * (T)(w[d] > 0 ? ph.plus_re[d] : ph.minus_re[d]);
        }
        ((T*)dst)[t] = v;
    }
}

// gather_matrix on one rank (:936-993): the interior of the padded array, as
// raw elements of W bytes.  One thread per element.
template <class W>
__global__ void __launch_bounds__(256) k_field_unpad(W* __restrict__ dst, const W* __restrict__ src, Geom g, Pad p,
                                                     int ne) {
    const long t = (long)blockIdx.x * 256 + threadIdx.x;
    if (t >= (long)g.V * ne) return;
    const long s = t / ne;
    const int e = (int)(t - s * ne);
    int x[4];
    site_coords((int)s, g, x);
    dst[t] = src[padded_site(x, p) * ne + e];
}

static int check_field(const lmx_context* ctx, int elem, int nc1, int nc2, int nw) {
    LMX_CHECK(elem >= LMX_ELEM_F32 && elem <= LMX_ELEM_C128, LMX_ERR_ARG,
              "fields support Float32, Float64, ComplexF32 and ComplexF64 elements (got code %d)", elem);
    LMX_CHECK(nc1 >= 1 && nc2 >= 1 && nc1 * nc2 <= 4096, LMX_ERR_ARG, "bad matrix shape %d x %d", nc1, nc2);
    return check_nw(ctx, nw);
}

int launch_field_pad(lmx_context* ctx, void* dst, const void* src, int elem, int nc1, int nc2, int nw) {
    LMX_CHECK(dst && src, LMX_ERR_ARG, "NULL field");
    LMX_TRY(check_field(ctx, elem, nc1, nc2, nw));
    HaloPhases ph;
    if (nw > 0) {
        LMX_TRY(make_halo_phases(ctx, elem, &ph));
    } else {
        for (int d = 0; d < 4; ++d) {
            ph.skip[d] = 1;
            ph.plus_re[d] = ph.minus_re[d] = 1.0;
            ph.plus_im[d] = ph.minus_im[d] = 0.0;
        }
    }
    const Pad p = make_pad(ctx, nw);
    const int ne = nc1 * nc2;
    const long n = p.Vp * ne;
    if (n == 0) return LMX_OK;
    const dim3 grid((unsigned)((n + 255) / 256));
    switch (elem) {
        case LMX_ELEM_F32:
            hipLaunchKernelGGL((k_field_pad<float, false>), grid, dim3(256), 0, ctx->stream, dst, src, ctx->g, p, ne, ph);
            break;
        case LMX_ELEM_F64:
            hipLaunchKernelGGL((k_field_pad<double, false>), grid, dim3(256), 0, ctx->stream, dst, src, ctx->g, p, ne,
                               ph);
            break;
        case LMX_ELEM_C64:
            hipLaunchKernelGGL((k_field_pad<float, true>), grid, dim3(256), 0, ctx->stream, dst, src, ctx->g, p, ne, ph);
            break;
        default:
            hipLaunchKernelGGL((k_field_pad<double, true>), grid, dim3(256), 0, ctx->stream, dst, src, ctx->g, p, ne,
                               ph);
            break;
    }
    LMX_HIP(hipGetLastError());
    return LMX_OK;
}

int launch_field_unpad(lmx_context* ctx, void* dst, const void* src, int elem, int nc1, int nc2, int nw) {
    LMX_CHECK(dst && src, LMX_ERR_ARG, "NULL field");
    LMX_TRY(check_field(ctx, elem, nc1, nc2, nw));
    const Pad p = make_pad(ctx, nw);
    const int ne = nc1 * nc2;
    const long n = ctx->V * (long)ne;
    if (n == 0) return LMX_OK;
    const dim3 grid((unsigned)((n + 255) / 256));
    switch (elem) {
        case LMX_ELEM_F32:
            hipLaunchKernelGGL((k_field_unpad<unsigned>), grid, dim3(256), 0, ctx->stream, (unsigned*)dst,
                               (const unsigned*)src, ctx->g, p, ne);
            break;
        case LMX_ELEM_F64:
        case LMX_ELEM_C64:
            hipLaunchKernelGGL((k_field_unpad<unsigned long long>), grid, dim3(256), 0, ctx->stream,
                               (unsigned long long*)dst, (const unsigned long long*)src, ctx->g, p, ne);
            break;
        default:
            hipLaunchKernelGGL((k_field_unpad<uint4>), grid, dim3(256), 0, ctx->stream, (uint4*)dst, (const uint4*)src,
                               ctx->g, p, ne);
            break;
    }
    LMX_HIP(hipGetLastError());
    return LMX_OK;
}

}  // namespace lmx
