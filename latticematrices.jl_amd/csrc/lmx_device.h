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
// Field accessors.  idx is a linear element index (component*V + site).
// ---------------------------------------------------------------------------
struct F64In {
    using R = double;
    const double2* __restrict__ p;
    __device__ __forceinline__ cx<double> ld(long i) const {
        double2 v = p[i];
        return mk<double>(v.x, v.y);
    }
};
struct F32In {
    using R = float;
    const float2* __restrict__ p;
    __device__ __forceinline__ cx<float> ld(long i) const {
        float2 v = p[i];
        return mk<float>(v.x, v.y);
    }
};
struct F16In {
    using R = float;
    const __half2* __restrict__ p;
    __device__ __forceinline__ cx<float> ld(long i) const {
        float2 v = __half22float2(p[i]);
        return mk<float>(v.x, v.y);
    }
};
struct F64Out {
    using R = double;
    double2* __restrict__ p;
    __device__ __forceinline__ void st(long i, cx<double> z) const { p[i] = make_double2(z.re, z.im); }
};
struct F32Out {
    using R = float;
    float2* __restrict__ p;
    __device__ __forceinline__ void st(long i, cx<float> z) const { p[i] = make_float2(z.re, z.im); }
};

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
