// lmx_device.h -- device-side building blocks of the Wilson(-clover) stencil.
//
// The whole library is compiled with -ffp-contract=off: a plain complex
// product or sum is evaluated exactly as Julia Base evaluates it (no fused
// multiply-add), and fma() appears only where the reference calls muladd.
// With that, the FP64 kernels reproduce the reference's rounding step by
// step (src/Operators/WilsonDiracOperator.jl:149-352).
#pragma once

#include <hip/hip_fp16.h>
#include <hip/hip_runtime.h>
#include <stdint.h>

#include <type_traits>

namespace lmx {

template <typename R>
struct cx {
    R re, im;
};

template <typename R>
__device__ __forceinline__ cx<R> mk(R r, R i) {
    cx<R> z;
    z.re = r;
    z.im = i;
    return z;
}
template <typename R>
__device__ __forceinline__ cx<R> cadd(cx<R> a, cx<R> b) { return mk<R>(a.re + b.re, a.im + b.im); }
template <typename R>
__device__ __forceinline__ cx<R> csub(cx<R> a, cx<R> b) { return mk<R>(a.re - b.re, a.im - b.im); }
template <typename R>
__device__ __forceinline__ cx<R> cneg(cx<R> a) { return mk<R>(-a.re, -a.im); }
template <typename R>
__device__ __forceinline__ cx<R> cconj(cx<R> a) { return mk<R>(a.re, -a.im); }
// Julia Base complex product, unfused.
template <typename R>
__device__ __forceinline__ cx<R> cmul(cx<R> a, cx<R> b) {
    return mk<R>(a.re * b.re - a.im * b.im, a.re * b.im + a.im * b.re);
}
template <typename R>
__device__ __forceinline__ cx<R> rmul(R a, cx<R> b) { return mk<R>(a * b.re, a * b.im); }

__device__ __forceinline__ double fmaR(double a, double b, double c) { return fma(a, b, c); }
__device__ __forceinline__ float fmaR(float a, float b, float c) { return fmaf(a, b, c); }

// Julia Base complex muladd with fused real muladds.
template <typename R>
__device__ __forceinline__ cx<R> cmuladd(cx<R> z, cx<R> w, cx<R> x) {
    return mk<R>(fmaR(z.re, w.re, -fmaR(z.im, w.im, -x.re)), fmaR(z.re, w.im, fmaR(z.im, w.re, x.im)));
}
// muladdmulti, src/Operators/WilsonDiracOperator.jl:149-155
template <typename R>
__device__ __forceinline__ cx<R> muladdmulti(cx<R> a1, cx<R> b1, cx<R> a2, cx<R> b2, cx<R> a3, cx<R> b3) {
    cx<R> acc = mk<R>(R(0), R(0));
    acc = cmuladd(a1, b1, acc);
    acc = cmuladd(a2, b2, acc);
    acc = cmuladd(a3, b3, acc);
    return acc;
}
// PM * im * z and PM * z for PM = +-1 (exact).
template <int PM, typename R>
__device__ __forceinline__ cx<R> pmi(cx<R> z) {
    return PM > 0 ? mk<R>(-z.im, z.re) : mk<R>(z.im, -z.re);
}
template <int PM, typename R>
__device__ __forceinline__ cx<R> pmr(cx<R> z) {
    return PM > 0 ? z : cneg(z);
}

// ---------------------------------------------------------------------------
// Geometry
// ---------------------------------------------------------------------------
struct Geom {
    int L[4];
    int stride[4];
    int V;
    int phase_one[4];  // phase == 1 exactly: the wrap multiply is skipped (as _mul_phase!)
    int phase_neg[4];  // phase == -1 exactly: the wrap is an exact sign flip
    int general;       // some phase is neither 1 nor -1 (complex multiply on wrap)
    double ph_re[4], ph_im[4];
    double iph_re[4], iph_im[4];
};

__device__ __forceinline__ void site_coords(int s, const Geom& g, int x[4]) {
    int r = s;
    x[0] = r % g.L[0];
    r /= g.L[0];
    x[1] = r % g.L[1];
    r /= g.L[1];
    x[2] = r % g.L[2];
    x[3] = r / g.L[2];
}

__device__ __forceinline__ int nbr_fwd(int s, const int x[4], int mu, const Geom& g, bool& wrap) {
    wrap = (x[mu] == g.L[mu] - 1);
    return wrap ? s - (g.L[mu] - 1) * g.stride[mu] : s + g.stride[mu];
}
__device__ __forceinline__ int nbr_bwd(int s, const int x[4], int mu, const Geom& g, bool& wrap) {
    wrap = (x[mu] == 0);
    return wrap ? s + (g.L[mu] - 1) * g.stride[mu] : s - g.stride[mu];
}

// Logical block index so that each XCD (workgroups are dispatched round-robin
// over the 8 XCDs) walks a contiguous range of the lattice: neighbour reuse
// then stays inside one XCD's L2.
__device__ __forceinline__ int xcd_block(int b, int nb) {
    const int nx = 8;
    int per = nb / nx, rem = nb % nx;
    int k = b % nx, j = b / nx;
    return (k < rem) ? k * (per + 1) + j : rem * (per + 1) + (k - rem) * per + j;
}

// ---------------------------------------------------------------------------
// Field accessors.  A spinor accessor moves the 12 colour-spin components
// (k = spin*3 + colour) of right-hand side r at site s; a link accessor the
// 9 entries (row*3 + col) of U_mu at site s.  Layouts, per right-hand side
// of V sites (components outermost, so a wave's load is one contiguous run):
//   F64  double2[12][V]      F32  float2[12][V]
//   H16  int4[3][V] + float[V]: components 4g..4g+3 as 16-bit fixed-point
//        (re, im) pairs, value = q * scale[s], scale = max |component| / 32752
//        (32752, not 32767: a sum or difference of two components, the
//        projected half-spinor, then stays within the FP16 range, 65504)
// Indices are 32-bit within one right-hand side or link direction (12 V < 2^31).
// ---------------------------------------------------------------------------
// Output stores bypass the caches' normal allocation (nontemporal): the
// stencil's output is not re-read by the same launch, and keeping it out of
// L2 leaves room for the neighbour spinors and links that are.
#ifndef LMX_NT_STORE
#define LMX_NT_STORE 1
#endif
typedef double lmx_dv2 __attribute__((ext_vector_type(2)));
typedef float lmx_fv2 __attribute__((ext_vector_type(2)));
typedef int lmx_iv4 __attribute__((ext_vector_type(4)));
typedef float lmx_fv4 __attribute__((ext_vector_type(4)));
__device__ __forceinline__ void st_out(double2* p, double re, double im) {
#if LMX_NT_STORE
    lmx_dv2 v = {re, im};
    __builtin_nontemporal_store(v, (lmx_dv2*)p);
#else
    *p = make_double2(re, im);
#endif
}
__device__ __forceinline__ void st_out(float2* p, float re, float im) {
#if LMX_NT_STORE
    lmx_fv2 v = {re, im};
    __builtin_nontemporal_store(v, (lmx_fv2*)p);
#else
    *p = make_float2(re, im);
#endif
}
__device__ __forceinline__ void st_out(int4* p, int4 w) {
#if LMX_NT_STORE
    lmx_iv4 v = {w.x, w.y, w.z, w.w};
    __builtin_nontemporal_store(v, (lmx_iv4*)p);
#else
    *p = w;
#endif
}
__device__ __forceinline__ void st_out(float4* p, float4 w) {
#if LMX_NT_STORE
    lmx_fv4 v = {w.x, w.y, w.z, w.w};
    __builtin_nontemporal_store(v, (lmx_fv4*)p);
#else
    *p = w;
#endif
}
__device__ __forceinline__ void st_out(float* p, float v) {
#if LMX_NT_STORE
    __builtin_nontemporal_store(v, p);
#else
    *p = v;
#endif
}

// Loads and stores address a field as (uniform base, 32-bit site index):
// the base (field + component offset) is the same for the whole wave, so
// the compiler keeps it in scalar registers and each access costs one
// 32-bit offset computation instead of a 64-bit address.
template <typename T>
__device__ __forceinline__ T ldv(const T* ubase, int s) {
    return *(const T*)((const char*)ubase + (unsigned)s * (unsigned)sizeof(T));
}

// The low-precision fields are read through global address-space pointers:
// the compiler then emits global loads, not flat ones (a flat load also
// counts against the LDS counter, so it serialises with the LDS-staged
// kernels' shared-memory traffic), and forms each address as a 32-bit lane
// offset from the uniform base.  The base must be uniform, as for ldv.
template <typename T>
using gptr_t = __attribute__((address_space(1))) const T*;
// element s of the field at base, read as the builtin type VT (a global
// address-space lvalue of a HIP vector type cannot be copied directly)
template <typename T, typename VT, bool NT>
__device__ __forceinline__ VT ldg_raw(const T* base, int s) {
    const gptr_t<T> g = (gptr_t<T>)base;
    using GV = __attribute__((address_space(1))) const VT*;
    const GV a = (GV)((__attribute__((address_space(1))) const char*)g + (unsigned)s * (unsigned)sizeof(T));
    if constexpr (NT) return __builtin_nontemporal_load(a);
    return *a;
}
__device__ __forceinline__ int ldg(const int* base, int s) { return ldg_raw<int, int, false>(base, s); }
__device__ __forceinline__ float ldg(const float* base, int s) { return ldg_raw<float, float, false>(base, s); }
__device__ __forceinline__ int4 ldg(const int4* base, int s) {
    const lmx_iv4 v = ldg_raw<int4, lmx_iv4, false>(base, s);
    return make_int4(v.x, v.y, v.z, v.w);
}
__device__ __forceinline__ float2 ldg(const float2* base, int s) {
    const lmx_fv2 v = ldg_raw<float2, lmx_fv2, false>(base, s);
    return make_float2(v.x, v.y);
}
__device__ __forceinline__ float4 ldg(const float4* base, int s) {
    const lmx_fv4 v = ldg_raw<float4, lmx_fv4, false>(base, s);
    return make_float4(v.x, v.y, v.z, v.w);
}

// Streamed loads (ldg with the nontemporal hint): for fields a launch reads
// once at the site itself (the fused epilogues' r, t, p, x_inner), so that
// they do not evict the neighbour spinors and links the stencil re-reads from
// L2.
#ifndef LMX_NT_STREAM
#define LMX_NT_STREAM 1
#endif
__device__ __forceinline__ int4 ldg_stream(const int4* base, int s) {
    const lmx_iv4 v = ldg_raw<int4, lmx_iv4, LMX_NT_STREAM != 0>(base, s);
    return make_int4(v.x, v.y, v.z, v.w);
}
__device__ __forceinline__ float ldg_stream(const float* base, int s) {
    return ldg_raw<float, float, LMX_NT_STREAM != 0>(base, s);
}
__device__ __forceinline__ float2 ldg_stream(const float2* base, int s) {
    const lmx_fv2 v = ldg_raw<float2, lmx_fv2, LMX_NT_STREAM != 0>(base, s);
    return make_float2(v.x, v.y);
}
__device__ __forceinline__ float4 ldg_stream(const float4* base, int s) {
    const lmx_fv4 v = ldg_raw<float4, lmx_fv4, LMX_NT_STREAM != 0>(base, s);
    return make_float4(v.x, v.y, v.z, v.w);
}
__device__ __forceinline__ double2 ldg_stream(const double2* base, int s) {
    const lmx_dv2 v = ldg_raw<double2, lmx_dv2, LMX_NT_STREAM != 0>(base, s);
    return make_double2(v.x, v.y);
}

struct SpF64In {
    using R = double;
    const double2* __restrict__ p;
    int V;
    static SpF64In at(const void* q, const float*, int V, int r0) {
        return SpF64In{(const double2*)q + (size_t)r0 * 12 * V, V};
    }
    // unscaled components and the site scale (1 here)
    __device__ __forceinline__ void ld12s(int r, int s, cx<double>* z, double& c) const {
        const double2* q = p + (size_t)r * 12 * V;
#pragma unroll
        for (int k = 0; k < 12; ++k) {
            double2 v = ldv(q + (size_t)k * V, s);
            z[k] = mk<double>(v.x, v.y);
        }
        c = 1.0;
    }
    __device__ __forceinline__ void ld12(int r, int s, cx<double>* z) const {
        double c;
        ld12s(r, s, z, c);
    }
    // streamed form (a field read once, at the site itself)
    __device__ __forceinline__ void ld12_stream(int r, int s, cx<double>* z) const {
        const double2* q = p + (size_t)r * 12 * V;
#pragma unroll
        for (int k = 0; k < 12; ++k) {
            double2 v = ldg_stream(q + (size_t)k * V, s);
            z[k] = mk<double>(v.x, v.y);
        }
    }
};
struct SpF64Out {
    using R = double;
    double2* __restrict__ p;
    int V;
    __device__ __forceinline__ void st12(int r, int s, const cx<double>* z) const {
        double2* q = p + (size_t)r * 12 * V;
#pragma unroll
        for (int k = 0; k < 12; ++k) st_out((double2*)((char*)(q + (size_t)k * V) + (unsigned)s * 16u), z[k].re, z[k].im);
    }
};
// A 256-site slice of one right-hand side staged in LDS by the LDS-staged
// streaming stencil: the site's stored words, component-major, so that
// consecutive lanes move consecutive words.
struct F32Slot {
    float2 w[12][256];
};
struct H16Slot {
    int4 w[3][256];
    float c[256];
};

struct SpF32In {
    using R = float;
    using Slot = F32Slot;
    const float2* __restrict__ p;
    int V;
    static SpF32In at(const void* q, const float*, int V, int r0) {
        return SpF32In{(const float2*)q + (size_t)r0 * 12 * V, V};
    }
    // site s of RHS r into lane `lane` of an LDS slot
    __device__ __forceinline__ void stage(int r, int s, Slot& sl, int lane) const {
        const float2* q = p + (size_t)r * 12 * V;
#pragma unroll
        for (int k = 0; k < 12; ++k) sl.w[k][lane] = ldg(q + (size_t)k * V, s);
    }
    __device__ static __forceinline__ void lds_ld12s(const Slot& sl, int lane, cx<float>* z, float& c) {
#pragma unroll
        for (int k = 0; k < 12; ++k) {
            const float2 v = sl.w[k][lane];
            z[k] = mk<float>(v.x, v.y);
        }
        c = 1.0f;
    }
    __device__ __forceinline__ void ld12s(int r, int s, cx<float>* z, float& c) const {
        const float2* q = p + (size_t)r * 12 * V;
#pragma unroll
        for (int k = 0; k < 12; ++k) {
            float2 v = ldg(q + (size_t)k * V, s);
            z[k] = mk<float>(v.x, v.y);
        }
        c = 1.0f;
    }
    __device__ __forceinline__ void ld12(int r, int s, cx<float>* z) const {
        float c;
        ld12s(r, s, z, c);
    }
    // streamed forms (fields read once, at the site itself)
    __device__ __forceinline__ void ld12s_stream(int r, int s, cx<float>* z, float& c) const {
        const float2* q = p + (size_t)r * 12 * V;
#pragma unroll
        for (int k = 0; k < 12; ++k) {
            float2 v = ldg_stream(q + (size_t)k * V, s);
            z[k] = mk<float>(v.x, v.y);
        }
        c = 1.0f;
    }
    __device__ __forceinline__ void ld12_stream(int r, int s, cx<float>* z) const {
        float c;
        ld12s_stream(r, s, z, c);
    }
};
struct SpF32Out {
    using R = float;
    float2* __restrict__ p;
    int V;
    static SpF32Out at(void* q, float*, int V, int r0) { return SpF32Out{(float2*)q + (size_t)r0 * 12 * V, V}; }
    __device__ __forceinline__ void st12(int r, int s, const cx<float>* z) const {
        float2* q = p + (size_t)r * 12 * V;
#pragma unroll
        for (int k = 0; k < 12; ++k) st_out((float2*)((char*)(q + (size_t)k * V) + (unsigned)s * 8u), z[k].re, z[k].im);
    }
};

// FP32 spinor in component pairs: float4 [nrhs][6][V] holding components
// (2m, 2m+1) -- the mixed solver's inner solution (half the instructions of
// the float2 layout for the same bytes).
struct SpX4In {
    using R = float;
    const float4* __restrict__ p;
    int V;
    static SpX4In at(const void* q, const float*, int V, int r0) {
        return SpX4In{(const float4*)q + (size_t)r0 * 6 * V, V};
    }
    __device__ __forceinline__ void ld12(int r, int s, cx<float>* z) const {
        const float4* q = p + (size_t)r * 6 * V;
#pragma unroll
        for (int m = 0; m < 6; ++m) {
            float4 v = ldg(q + (size_t)m * V, s);
            z[2 * m] = mk<float>(v.x, v.y);
            z[2 * m + 1] = mk<float>(v.z, v.w);
        }
    }
    __device__ __forceinline__ void ld12_stream(int r, int s, cx<float>* z) const {
        const float4* q = p + (size_t)r * 6 * V;
#pragma unroll
        for (int m = 0; m < 6; ++m) {
            float4 v = ldg_stream(q + (size_t)m * V, s);
            z[2 * m] = mk<float>(v.x, v.y);
            z[2 * m + 1] = mk<float>(v.z, v.w);
        }
    }
};
struct SpX4Out {
    using R = float;
    float4* __restrict__ p;
    int V;
    static SpX4Out at(void* q, float*, int V, int r0) { return SpX4Out{(float4*)q + (size_t)r0 * 6 * V, V}; }
    __device__ __forceinline__ void st12(int r, int s, const cx<float>* z) const {
        float4* q = p + (size_t)r * 6 * V;
#pragma unroll
        for (int m = 0; m < 6; ++m)
            st_out((float4*)((char*)(q + (size_t)m * V) + (unsigned)s * 16u),
                   make_float4(z[2 * m].re, z[2 * m].im, z[2 * m + 1].re, z[2 * m + 1].im));
    }
};

constexpr float kH16Max = 32752.0f;

__device__ __forceinline__ float h16_lo(int w) { return float(int(unsigned(w) << 16) >> 16); }
__device__ __forceinline__ float h16_hi(int w) { return float(w >> 16); }
// round(re * inv), round(im * inv) (|.| <= 32752, nearest-even, one rounding
// of the exact product) as a 16-bit pair: x + 1.5 * 2^23 lies in [2^23, 2^24),
// where the float spacing is 1, so the low 16 bits of its encoding are round(x)
// in two's complement; one FMA per component and one byte permute per pair.
__device__ __forceinline__ int h16_pack_scaled(float re, float im, float inv) {
    constexpr float kMagic = 12582912.0f;
    const float a = __builtin_fmaf(re, inv, kMagic), b = __builtin_fmaf(im, inv, kMagic);
    return int(__builtin_amdgcn_perm(__float_as_uint(b), __float_as_uint(a), 0x05040100u));
}

struct SpH16In {
    using R = float;
    using Slot = H16Slot;
    const int4* __restrict__ p;
    const float* __restrict__ scale;
    int V;
    static SpH16In at(const void* q, const float* n, int V, int r0) {
        return SpH16In{(const int4*)q + (size_t)r0 * 3 * V, n + (size_t)r0 * V, V};
    }
    // site s of RHS r into lane `lane` of an LDS slot
    __device__ __forceinline__ void stage(int r, int s, Slot& sl, int lane) const {
        const int4* q = p + (size_t)r * 3 * V;
#pragma unroll
        for (int g = 0; g < 3; ++g) sl.w[g][lane] = ldg(q + (size_t)g * V, s);
        sl.c[lane] = ldg(scale + (size_t)r * V, s);
    }
    // integer components (as floats) and the site scale, as ld12s
    __device__ static __forceinline__ void lds_ld12s(const Slot& sl, int lane, cx<float>* z, float& c) {
        c = sl.c[lane];
#pragma unroll
        for (int g = 0; g < 3; ++g) {
            const int4 v = sl.w[g][lane];
            z[4 * g + 0] = mk<float>(h16_lo(v.x), h16_hi(v.x));
            z[4 * g + 1] = mk<float>(h16_lo(v.y), h16_hi(v.y));
            z[4 * g + 2] = mk<float>(h16_lo(v.z), h16_hi(v.z));
            z[4 * g + 3] = mk<float>(h16_lo(v.w), h16_hi(v.w));
        }
    }
    // integer components (as floats) and the site scale: value = z * c
    __device__ __forceinline__ void ld12s(int r, int s, cx<float>* z, float& c) const {
        const int4* q = p + (size_t)r * 3 * V;
        c = ldg(scale + (size_t)r * V, s);
#pragma unroll
        for (int g = 0; g < 3; ++g) {
            int4 v = ldg(q + (size_t)g * V, s);
            z[4 * g + 0] = mk<float>(h16_lo(v.x), h16_hi(v.x));
            z[4 * g + 1] = mk<float>(h16_lo(v.y), h16_hi(v.y));
            z[4 * g + 2] = mk<float>(h16_lo(v.z), h16_hi(v.z));
            z[4 * g + 3] = mk<float>(h16_lo(v.w), h16_hi(v.w));
        }
    }
    __device__ __forceinline__ void ld12(int r, int s, cx<float>* z) const {
        float c;
        ld12s(r, s, z, c);
#pragma unroll
        for (int k = 0; k < 12; ++k) z[k] = mk<float>(z[k].re * c, z[k].im * c);
    }
    // streamed forms (fields read once, at the site itself)
    __device__ __forceinline__ void ld12s_stream(int r, int s, cx<float>* z, float& c) const {
        const int4* q = p + (size_t)r * 3 * V;
        c = ldg_stream(scale + (size_t)r * V, s);
#pragma unroll
        for (int g = 0; g < 3; ++g) {
            int4 v = ldg_stream(q + (size_t)g * V, s);
            z[4 * g + 0] = mk<float>(h16_lo(v.x), h16_hi(v.x));
            z[4 * g + 1] = mk<float>(h16_lo(v.y), h16_hi(v.y));
            z[4 * g + 2] = mk<float>(h16_lo(v.z), h16_hi(v.z));
            z[4 * g + 3] = mk<float>(h16_lo(v.w), h16_hi(v.w));
        }
    }
    __device__ __forceinline__ void ld12_stream(int r, int s, cx<float>* z) const {
        float c;
        ld12s_stream(r, s, z, c);
#pragma unroll
        for (int k = 0; k < 12; ++k) z[k] = mk<float>(z[k].re * c, z[k].im * c);
    }
};
struct SpH16Out {
    using R = float;
    int4* __restrict__ p;
    float* __restrict__ scale;
    int V;
    static SpH16Out at(void* q, float* n, int V, int r0) {
        return SpH16Out{(int4*)q + (size_t)r0 * 3 * V, n + (size_t)r0 * V, V};
    }
    __device__ __forceinline__ void st12(int r, int s, const cx<float>* z) const {
        float m = 0.0f;
#pragma unroll
        for (int k = 0; k < 12; ++k) m = fmaxf(m, fmaxf(fabsf(z[k].re), fabsf(z[k].im)));
        // 1 ulp reciprocal: m * inv stays within 32752 * (1 + 2^-22) < 32752.5
        const float inv = m > 0.0f ? kH16Max * __builtin_amdgcn_rcpf(m) : 0.0f;
        int4* q = p + (size_t)r * 3 * V;
#pragma unroll
        for (int g = 0; g < 3; ++g) {
            int4 v;
            v.x = h16_pack_scaled(z[4 * g + 0].re, z[4 * g + 0].im, inv);
            v.y = h16_pack_scaled(z[4 * g + 1].re, z[4 * g + 1].im, inv);
            v.z = h16_pack_scaled(z[4 * g + 2].re, z[4 * g + 2].im, inv);
            v.w = h16_pack_scaled(z[4 * g + 3].re, z[4 * g + 3].im, inv);
            st_out((int4*)((char*)(q + (size_t)g * V) + (unsigned)s * 16u), v);
        }
        st_out((float*)((char*)(scale + (size_t)r * V) + (unsigned)s * 4u), m * (1.0f / kH16Max));
    }
};

// Links: FP64 / FP32 as double2 / float2 [4][9][V]; FP16 packed per
// direction as two 16-byte words (entries 0-3, 4-7) and one 4-byte word
// (entry 8): int4 [4][2][V] followed by half2 [4][V], three loads a link.
#ifndef LMX_NT_LINKS
#define LMX_NT_LINKS 0
#endif
struct LinkF64 {
    const double2* __restrict__ p;
    int V;
    // nontemporal: a link used once, whose lines should not displace cached spinors
    template <typename R>
    __device__ __forceinline__ void ld9_nt(int mu, int s, cx<R>* u) const {
        const double2* q = p + (size_t)mu * 9 * V;
#pragma unroll
        for (int k = 0; k < 9; ++k) {
            const lmx_dv2* a = (const lmx_dv2*)((const char*)(q + (size_t)k * V) + (unsigned)s * 16u);
            lmx_dv2 w = __builtin_nontemporal_load(a);
            u[k] = mk<R>(R(w.x), R(w.y));
        }
    }
    template <typename R>
    __device__ __forceinline__ void ld9(int mu, int s, cx<R>* u) const {
        const double2* q = p + (size_t)mu * 9 * V;
#pragma unroll
        for (int k = 0; k < 9; ++k) {
#if LMX_NT_LINKS
            const lmx_dv2* a = (const lmx_dv2*)((const char*)(q + (size_t)k * V) + (unsigned)s * 16u);
            lmx_dv2 w = __builtin_nontemporal_load(a);
            u[k] = mk<R>(R(w.x), R(w.y));
#else
            double2 v = ldv(q + (size_t)k * V, s);
            u[k] = mk<R>(R(v.x), R(v.y));
#endif
        }
    }
};
struct LinkF32 {
    const float2* __restrict__ p;
    int V;
    template <typename R>
    __device__ __forceinline__ void ld9(int mu, int s, cx<R>* u) const {
        const float2* q = p + (size_t)mu * 9 * V;
#pragma unroll
        for (int k = 0; k < 9; ++k) {
            float2 v = ldg(q + (size_t)k * V, s);
            u[k] = mk<R>(R(v.x), R(v.y));
        }
    }
};
__device__ __forceinline__ float2 h2f(int w) {
    __half2 h = *reinterpret_cast<__half2*>(&w);
    return __half22float2(h);
}
struct LinkF16 {
    const int4* __restrict__ p;  // packed layout above
    int V;
    template <typename R>
    __device__ __forceinline__ void ld9(int mu, int s, cx<R>* u) const {
        const int4 a = ldg(p + (size_t)(2 * mu) * V, s);
        const int4 b = ldg(p + (size_t)(2 * mu + 1) * V, s);
        const int c = ldg((const int*)(p + (size_t)8 * V) + (size_t)mu * V, s);
        const int w[9] = {a.x, a.y, a.z, a.w, b.x, b.y, b.z, b.w, c};
#pragma unroll
        for (int k = 0; k < 9; ++k) {
            float2 v = h2f(w[k]);
            u[k] = mk<R>(R(v.x), R(v.y));
        }
    }
    // the nine entries as their (re, im) FP16 words
    __device__ __forceinline__ void ld9w(int mu, int s, unsigned* w) const {
        const int4 a = ldg(p + (size_t)(2 * mu) * V, s);
        const int4 b = ldg(p + (size_t)(2 * mu + 1) * V, s);
        const int c = ldg((const int*)(p + (size_t)8 * V) + (size_t)mu * V, s);
        w[0] = a.x; w[1] = a.y; w[2] = a.z; w[3] = a.w;
        w[4] = b.x; w[5] = b.y; w[6] = b.z; w[7] = b.w;
        w[8] = c;
    }
};

// Keep a 12-component accumulator in registers at this point: an empty asm
// that "modifies" the values stops the scheduler from sinking the arithmetic
// that produced them below the loads that follow (which otherwise keeps
// every loaded neighbour live at once and spills).
#ifndef LMX_PIN_MEM
#define LMX_PIN_MEM 1
#endif
template <typename R>
__device__ __forceinline__ void pin_acc(cx<R> (&a)[12]) {
#pragma unroll
    for (int k = 0; k < 12; k += 2)
#if LMX_PIN_MEM
        asm volatile("" : "+v"(a[k].re), "+v"(a[k].im), "+v"(a[k + 1].re), "+v"(a[k + 1].im)::"memory");
#else
        asm volatile("" : "+v"(a[k].re), "+v"(a[k].im), "+v"(a[k + 1].re), "+v"(a[k + 1].im));
#endif
}

// ---------------------------------------------------------------------------
// Spin projection / reconstruction (Appendix of the half-spin kernels)
//   _wilson_half_project3     src/Operators/WilsonDiracOperator.jl:157-175
//   _wilson_reconstruct_add3  src/Operators/WilsonDiracOperator.jl:231-263
// psi is indexed [spin*3 + colour]; h[2*colour + {0,1}]; acc[colour*4 + spin].
// ---------------------------------------------------------------------------
template <int PM, int MU, typename R>
__device__ __forceinline__ void half_project(const cx<R> psi[12], cx<R> h[6]) {
#pragma unroll
    for (int c = 0; c < 3; ++c) {
        cx<R> p1 = psi[c], p2 = psi[3 + c], p3 = psi[6 + c], p4 = psi[9 + c];
        if (MU == 0) {
            h[2 * c] = csub(p1, pmi<PM>(p4));
            h[2 * c + 1] = csub(p2, pmi<PM>(p3));
        } else if (MU == 1) {
            h[2 * c] = csub(p1, pmr<PM>(p4));
            h[2 * c + 1] = cadd(p2, pmr<PM>(p3));
        } else if (MU == 2) {
            h[2 * c] = csub(p1, pmi<PM>(p3));
            h[2 * c + 1] = cadd(p2, pmi<PM>(p4));
        } else {
            h[2 * c] = csub(p1, pmr<PM>(p3));
            h[2 * c + 1] = csub(p2, pmr<PM>(p4));
        }
    }
}

template <int PM, int MU, typename R>
__device__ __forceinline__ void reconstruct_add(cx<R> acc[12], const cx<R> r[6]) {
#pragma unroll
    for (int c = 0; c < 3; ++c) {
        cx<R> h1 = r[2 * c], h2 = r[2 * c + 1];
        cx<R>* a = acc + 4 * c;
        a[0] = cadd(a[0], h1);
        a[1] = cadd(a[1], h2);
        if (MU == 0) {
            a[2] = cadd(a[2], pmi<PM>(h2));
            a[3] = cadd(a[3], pmi<PM>(h1));
        } else if (MU == 1) {
            a[2] = cadd(a[2], pmr<PM>(h2));
            a[3] = csub(a[3], pmr<PM>(h1));
        } else if (MU == 2) {
            a[2] = cadd(a[2], pmi<PM>(h1));
            a[3] = csub(a[3], pmi<PM>(h2));
        } else {
            a[2] = csub(a[2], pmr<PM>(h1));
            a[3] = csub(a[3], pmr<PM>(h2));
        }
    }
}

// acc += c * (reconstructed r): the site scale of a 16-bit spinor applied
// once per half-spinor (c == 1 for FP32/FP64 fields, where fma(1, x, a) is
// exactly a + x).
template <typename R>
__device__ __forceinline__ cx<R> cfma(R c, cx<R> t, cx<R> a) {
    return mk<R>(fmaR(c, t.re, a.re), fmaR(c, t.im, a.im));
}
template <int PM, int MU, typename R>
__device__ __forceinline__ void reconstruct_add_s(cx<R> acc[12], const cx<R> r[6], R c) {
#pragma unroll
    for (int k = 0; k < 3; ++k) {
        cx<R> h1 = r[2 * k], h2 = r[2 * k + 1];
        cx<R>* a = acc + 4 * k;
        a[0] = cfma(c, h1, a[0]);
        a[1] = cfma(c, h2, a[1]);
        if (MU == 0) {
            a[2] = cfma(c, pmi<PM>(h2), a[2]);
            a[3] = cfma(c, pmi<PM>(h1), a[3]);
        } else if (MU == 1) {
            a[2] = cfma(c, pmr<PM>(h2), a[2]);
            a[3] = cfma(c, cneg(pmr<PM>(h1)), a[3]);
        } else if (MU == 2) {
            a[2] = cfma(c, pmi<PM>(h1), a[2]);
            a[3] = cfma(c, cneg(pmi<PM>(h2)), a[3]);
        } else {
            a[2] = cfma(c, cneg(pmr<PM>(h1)), a[2]);
            a[3] = cfma(c, cneg(pmr<PM>(h2)), a[3]);
        }
    }
}

// U h  (_wilson_half_matvec_forward3, :177-202)
template <typename R>
__device__ __forceinline__ void matvec_fwd(const cx<R> u[9], const cx<R> h[6], cx<R> r[6]) {
#pragma unroll
    for (int a = 0; a < 3; ++a) {
        r[2 * a] = muladdmulti(u[3 * a], h[0], u[3 * a + 1], h[2], u[3 * a + 2], h[4]);
        r[2 * a + 1] = muladdmulti(u[3 * a], h[1], u[3 * a + 1], h[3], u[3 * a + 2], h[5]);
    }
}
// U^dagger h  (_wilson_half_matvec_backward3, :204-229)
template <typename R>
__device__ __forceinline__ void matvec_bwd(const cx<R> u[9], const cx<R> h[6], cx<R> r[6]) {
#pragma unroll
    for (int a = 0; a < 3; ++a) {
        cx<R> w0 = cconj(u[a]), w1 = cconj(u[3 + a]), w2 = cconj(u[6 + a]);
        r[2 * a] = muladdmulti(w0, h[0], w1, h[2], w2, h[4]);
        r[2 * a + 1] = muladdmulti(w0, h[1], w1, h[3], w2, h[5]);
    }
}

// U h and U^dagger h from FP16 link entries, as FP16 pair products.
// v_dot2_f32_f16 computes a.lo * b.lo + a.hi * b.hi + c from FP16 pairs (the
// products are exact in FP32) into an FP32 accumulator, so one instruction
// gives the real or imaginary part of a complex product with a link entry
// w = (re, im) and sums it, where matvec_fwd with converted entries takes two
// mixed FMAs.  Per half-spinor element h the operands are hp = (h.re, h.im)
// and hs = (h.im, h.re); the minus sign of each sum is the VOP3P neg_hi
// modifier.  The instruction is written out: the compiler's builtin selects
// the modifier-free VOP2 form, which costs a negation per operand and a move
// per accumulator.  A VALU read (or a write) of a dot result needs three wait
// states after the dot, except as the next dot's accumulator: each block ends
// with s_nop 2, so its outputs are safe to use as soon as it returns (the
// compiler, which does not know the block, inserts nothing).  h is rounded to
// FP16 (nearest even, relative error at most 2^-12, as for the links
// themselves); the 16-bit stencils call these with h an integer sum of two
// fixed-point components, |h| <= 2 * 32752 <= 65504, the FP16 maximum.
typedef _Float16 lmx_hv2 __attribute__((ext_vector_type(2)));
typedef float lmx_fv2x __attribute__((ext_vector_type(2)));
__device__ __forceinline__ lmx_hv2 pk16(float a, float b) { return __builtin_convertvector((lmx_fv2x){a, b}, lmx_hv2); }

// one spin of U h (FWD: sum_j U[a][j] h_j) or U^dagger h (sum_j conj(U[j][a]) h_j)
// for the three rows a: outputs (re, im) of rows 0, 1, 2; u the nine words
// row-major; hp/hs the spin's three colours
template <bool FWD>
__device__ __forceinline__ void dot_spin(const unsigned u[9], const lmx_hv2 hp[3], const lmx_hv2 hs[3], cx<float> o[3]) {
    float r0, i0, r1, i1, r2, i2;
    if constexpr (FWD) {
        asm(
        "v_dot2_f32_f16 %0, %6, %15, 0 neg_hi:[0,1,0]\n"
        "v_dot2_f32_f16 %1, %6, %18, 0\n"
        "v_dot2_f32_f16 %2, %9, %15, 0 neg_hi:[0,1,0]\n"
        "v_dot2_f32_f16 %3, %9, %18, 0\n"
        "v_dot2_f32_f16 %4, %12, %15, 0 neg_hi:[0,1,0]\n"
        "v_dot2_f32_f16 %5, %12, %18, 0\n"
        "v_dot2_f32_f16 %0, %7, %16, %0 neg_hi:[0,1,0]\n"
        "v_dot2_f32_f16 %1, %7, %19, %1\n"
        "v_dot2_f32_f16 %2, %10, %16, %2 neg_hi:[0,1,0]\n"
        "v_dot2_f32_f16 %3, %10, %19, %3\n"
        "v_dot2_f32_f16 %4, %13, %16, %4 neg_hi:[0,1,0]\n"
        "v_dot2_f32_f16 %5, %13, %19, %5\n"
        "v_dot2_f32_f16 %0, %8, %17, %0 neg_hi:[0,1,0]\n"
        "v_dot2_f32_f16 %1, %8, %20, %1\n"
        "v_dot2_f32_f16 %2, %11, %17, %2 neg_hi:[0,1,0]\n"
        "v_dot2_f32_f16 %3, %11, %20, %3\n"
        "v_dot2_f32_f16 %4, %14, %17, %4 neg_hi:[0,1,0]\n"
        "v_dot2_f32_f16 %5, %14, %20, %5\n"
        "s_nop 2\n"
            : "=&v"(r0), "=&v"(i0), "=&v"(r1), "=&v"(i1), "=&v"(r2), "=&v"(i2)
            : "v"(u[0]), "v"(u[1]), "v"(u[2]), "v"(u[3]), "v"(u[4]), "v"(u[5]), "v"(u[6]), "v"(u[7]), "v"(u[8]),
              "v"(hp[0]), "v"(hp[1]), "v"(hp[2]), "v"(hs[0]), "v"(hs[1]), "v"(hs[2]));
    } else {
        asm(
        "v_dot2_f32_f16 %0, %6, %15, 0\n"
        "v_dot2_f32_f16 %1, %6, %18, 0 neg_hi:[0,1,0]\n"
        "v_dot2_f32_f16 %2, %7, %15, 0\n"
        "v_dot2_f32_f16 %3, %7, %18, 0 neg_hi:[0,1,0]\n"
        "v_dot2_f32_f16 %4, %8, %15, 0\n"
        "v_dot2_f32_f16 %5, %8, %18, 0 neg_hi:[0,1,0]\n"
        "v_dot2_f32_f16 %0, %9, %16, %0\n"
        "v_dot2_f32_f16 %1, %9, %19, %1 neg_hi:[0,1,0]\n"
        "v_dot2_f32_f16 %2, %10, %16, %2\n"
        "v_dot2_f32_f16 %3, %10, %19, %3 neg_hi:[0,1,0]\n"
        "v_dot2_f32_f16 %4, %11, %16, %4\n"
        "v_dot2_f32_f16 %5, %11, %19, %5 neg_hi:[0,1,0]\n"
        "v_dot2_f32_f16 %0, %12, %17, %0\n"
        "v_dot2_f32_f16 %1, %12, %20, %1 neg_hi:[0,1,0]\n"
        "v_dot2_f32_f16 %2, %13, %17, %2\n"
        "v_dot2_f32_f16 %3, %13, %20, %3 neg_hi:[0,1,0]\n"
        "v_dot2_f32_f16 %4, %14, %17, %4\n"
        "v_dot2_f32_f16 %5, %14, %20, %5 neg_hi:[0,1,0]\n"
        "s_nop 2\n"
            : "=&v"(r0), "=&v"(i0), "=&v"(r1), "=&v"(i1), "=&v"(r2), "=&v"(i2)
            : "v"(u[0]), "v"(u[1]), "v"(u[2]), "v"(u[3]), "v"(u[4]), "v"(u[5]), "v"(u[6]), "v"(u[7]), "v"(u[8]),
              "v"(hp[0]), "v"(hp[1]), "v"(hp[2]), "v"(hs[0]), "v"(hs[1]), "v"(hs[2]));
    }
    o[0] = mk<float>(r0, i0);
    o[1] = mk<float>(r1, i1);
    o[2] = mk<float>(r2, i2);
}
// r[2a + sp] = (U h)[a] of spin sp (FWD) or (U^dagger h)[a]; matvec_fwd / matvec_bwd's sums
template <bool FWD>
__device__ __forceinline__ void matvec_h(const unsigned w[9], const cx<float> h[6], cx<float> r[6]) {
#pragma unroll
    for (int sp = 0; sp < 2; ++sp) {
        lmx_hv2 hp[3], hs[3];
        cx<float> o[3];
#pragma unroll
        for (int j = 0; j < 3; ++j) {
            hp[j] = pk16(h[2 * j + sp].re, h[2 * j + sp].im);
            hs[j] = pk16(h[2 * j + sp].im, h[2 * j + sp].re);
        }
        dot_spin<FWD>(w, hp, hs, o);
#pragma unroll
        for (int a = 0; a < 3; ++a) r[2 * a + sp] = o[a];
    }
}

// The link operand of one hop.  With 16-bit spinors and FP16 links (HP) it
// keeps the nine FP16 words and multiplies with the pair products above;
// otherwise it holds the entries converted to R.
template <class Lk, typename R, bool HP>
struct LinkOp {
    cx<R> u[9];
    __device__ __forceinline__ void load(const Lk& U, int mu, int s) { U.template ld9<R>(mu, s, u); }
    template <class W>
    __device__ __forceinline__ void from_words(const W& lw) {
#pragma unroll
        for (int k = 0; k < 9; ++k) {
            const float2 v = h2f(lw[k]);
            u[k] = mk<R>(R(v.x), R(v.y));
        }
    }
    __device__ __forceinline__ void fwd(const cx<R> h[6], cx<R> o[6]) const { matvec_fwd(u, h, o); }
    __device__ __forceinline__ void bwd(const cx<R> h[6], cx<R> o[6]) const { matvec_bwd(u, h, o); }
};
template <class Lk>
struct LinkOp<Lk, float, true> {
    unsigned w[9];
    __device__ __forceinline__ void load(const Lk& U, int mu, int s) { U.ld9w(mu, s, w); }
    template <class W>
    __device__ __forceinline__ void from_words(const W& lw) {
#pragma unroll
        for (int k = 0; k < 9; ++k) w[k] = unsigned(lw[k]);
    }
    __device__ __forceinline__ void fwd(const cx<float> h[6], cx<float> o[6]) const { matvec_h<true>(w, h, o); }
    __device__ __forceinline__ void bwd(const cx<float> h[6], cx<float> o[6]) const { matvec_h<false>(w, h, o); }
};
template <class In, class Lk>
constexpr bool kHalfProducts = std::is_same<In, SpH16In>::value && std::is_same<Lk, LinkF16>::value;

// Negate a half-spinor where `neg` holds (a boundary phase of -1).  Negation
// commutes exactly with projection and the link product, so flipping the
// projected half-spinor is bit-identical to flipping the neighbour spinor as
// the halo fill does; selects keep the stencil free of divergent branches.
template <typename R>
__device__ __forceinline__ void neg_select(cx<R> h[6], bool neg) {
#pragma unroll
    for (int k = 0; k < 6; ++k) {
        h[k].re = neg ? -h[k].re : h[k].re;
        h[k].im = neg ? -h[k].im : h[k].im;
    }
}

// Multiply a full spinor by a boundary phase (the halo fill's _mul_phase!).
template <typename R>
__device__ __forceinline__ void phase_spinor(cx<R> psi[12], double pre, double pim) {
    cx<R> ph = mk<R>(R(pre), R(pim));
#pragma unroll
    for (int k = 0; k < 12; ++k) psi[k] = cmul(psi[k], ph);
}

// ---------------------------------------------------------------------------
// Fixed-order reduction tree (block of 256 = 4 waves of 64).  The CPU oracle
// restates exactly this tree, so FP64 sums agree bit for bit.
// ---------------------------------------------------------------------------
constexpr int kRedBlock = 256;

__device__ __forceinline__ double wave_halve(double v) {
#pragma unroll
    for (int off = 32; off >= 1; off >>= 1) v = v + __shfl_down(v, off, 64);
    return v;
}

// Returns the block total in thread 0 (other threads: unspecified).
__device__ __forceinline__ double block_tree(double v, double* lds4) {
    v = wave_halve(v);
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    if (lane == 0) lds4[wave] = v;
    __syncthreads();
    double t = 0.0;
    if (threadIdx.x == 0) {
        double w0 = lds4[0], w1 = lds4[1], w2 = lds4[2], w3 = lds4[3];
        w0 = w0 + w2;
        w1 = w1 + w3;
        t = w0 + w1;
    }
    __syncthreads();
    return t;
}

// Second stage (its own one-block launch, after the producing kernel has
// finished, so every partial is visible without a per-block fence): 256
// lanes sum the block partials with stride 256 in order, then the block tree.
__device__ __forceinline__ double finish_partials(const double* partials, int nb, double* lds4) {
    double acc = 0.0;
    for (int b = threadIdx.x; b < nb; b += kRedBlock) acc = acc + partials[b];
    return block_tree(acc, lds4);
}

}  // namespace lmx
