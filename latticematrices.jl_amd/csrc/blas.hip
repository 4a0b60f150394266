// blas.hip -- BLAS-1 and the fused FP64 CG vector kernels.
//
// Each reduction is two launches: every block of the producing kernel reduces
// its 256 sites with the fixed tree of lmx_device.h and stores one partial;
// a one-block k_finalize launch then sums the partials in a fixed order and
// updates the device-resident CG state (alpha, beta, convergence) without a
// host round trip.  The
// element-wise updates use the reference's unfused forms
//   axpby!(a, X, b, Y): Y = a*X + b*Y   src/LinearAlgebras/linearalgebra_D.jl:15-23
// and the per-site dot kernel_dot_D (:3331-3341).
#include "lmx_internal.h"

namespace lmx {


__device__ __forceinline__ void finalize_scalar(int kind, CGState* st, double re, double im, double* out2) {
    switch (kind) {
    case FIN_OUT:
        out2[0] = re;
        out2[1] = im;
        break;
    case FIN_BB: {  // cg.jl:78-86
        st->bb = re;
        if (!isfinite(re) || re < 0.0) {
            st->done = 1;
            st->reason = LMX_CG_BREAKDOWN;
            st->rnorm = NAN;
            st->bnorm = NAN;
        } else {
            st->bnorm = sqrt(re);
            st->tol = fmax(st->atol, st->rtol * st->bnorm);
        }
        break;
    }
    case FIN_RR0: {  // cg.jl:95-103
        st->rr = re;
        if (!isfinite(re) || re < 0.0) {
            st->done = 1;
            st->reason = LMX_CG_BREAKDOWN;
            st->rnorm = NAN;
        } else {
            st->rnorm = sqrt(re);
            if (st->rnorm <= st->tol) {
                st->done = 1;
                st->converged = 1;
                st->reason = LMX_CG_CONVERGED;
            } else if (st->maxiter <= 0) {
                st->done = 1;
                st->reason = LMX_CG_MAXITER;
            }
        }
        break;
    }
    case FIN_PAP: {  // cg.jl:107-116
        st->pAp = re;
        if (!isfinite(re)) {
            st->done = 1;
            st->reason = LMX_CG_BREAKDOWN;
        } else if (re <= 0.0) {
            st->done = 1;
            st->reason = LMX_CG_NONPOSITIVE_CURVATURE;
        } else {
            st->alpha = st->rr / re;
        }
        break;
    }
    case FIN_RRNEW: {  // cg.jl:120-132
        st->rr_new = re;
        st->iteration += 1;
        if (!isfinite(re) || re < 0.0) {
            st->done = 1;
            st->reason = LMX_CG_BREAKDOWN;
            st->rnorm = NAN;
        } else {
            st->rnorm = sqrt(re);
            if (st->rnorm <= st->tol) {
                st->done = 1;
                st->converged = 1;
                st->reason = LMX_CG_CONVERGED;
            } else if (st->iteration >= st->maxiter) {
                st->done = 1;
                st->reason = LMX_CG_MAXITER;
            } else {
                st->beta = re / st->rr;
                st->rr = re;
            }
        }
        break;
    }
    }
}

// Block partials of (re, im) over the block's 256 sites (fixed tree).
__device__ __forceinline__ void store_partials(double re, double im, bool want_im, int lb, double* partials) {
    __shared__ double lds4[4];
    const int nb = gridDim.x;
    double tre = block_tree(re, lds4);
    double tim = want_im ? block_tree(im, lds4) : 0.0;
    if (threadIdx.x == 0) {
        partials[lb] = tre;
        if (want_im) partials[nb + lb] = tim;
    }
}

// Second stage of every reduction: one block sums the partials in a fixed
// order and updates the device CG state (or writes the scalar out).
__global__ void __launch_bounds__(256) k_finalize(const double* __restrict__ partials, int nb, int kind, int want_im,
                                                  CGState* st, double* out2) {
    if (st != nullptr && st->done) return;
    __shared__ double lds4[4];
    double fre = finish_partials(partials, nb, lds4);
    double fim = want_im ? finish_partials(partials + nb, nb, lds4) : 0.0;
    if (threadIdx.x == 0) finalize_scalar(kind, st, fre, fim, out2);
}

int launch_finalize(lmx_context* ctx, int nb, int kind, bool want_im, CGState* st, double* out2) {
    hipLaunchKernelGGL(k_finalize, dim3(1), dim3(256), 0, ctx->stream, ctx->partials, nb, kind, want_im ? 1 : 0, st,
                       out2);
    LMX_HIP(hipGetLastError());
    return LMX_OK;
}

// The CG vector kernels take fields of element type ComplexF64 (R = double,
// double2 storage: the bit-exact FP64 CG) or ComplexF32 (R = float, float2:
// cg! on single-precision fields, which the reference runs in the fields'
// element type); per-site arithmetic is in R, the site-to-site sums of the
// reductions in FP64.
template <typename R>
struct Vec2;
template <>
struct Vec2<double> {
    using T = double2;
};
template <>
struct Vec2<float> {
    using T = float2;
};

// dot(A, B) = sum conj(A) B
template <int NCOMP, typename R = double>
__global__ void __launch_bounds__(256) k_dot(const typename Vec2<R>::T* __restrict__ A,
                                             const typename Vec2<R>::T* __restrict__ B, int V, double* partials,
                                             int want_im, CGState* st) {
    if (st != nullptr && st->done) return;
    const int lb = blockIdx.x;
    const int s = lb * 256 + threadIdx.x;
    double re = 0.0, im = 0.0;
    if (s < V) {
        cx<R> acc = mk<R>(R(0), R(0));
#pragma unroll
        for (int k = 0; k < NCOMP; ++k) {
            const auto a = A[(long)k * V + s], b = B[(long)k * V + s];
            acc = cadd(acc, cmul(cconj(mk<R>(a.x, a.y)), mk<R>(b.x, b.y)));
        }
        re = double(acc.re);
        im = double(acc.im);
    }
    store_partials(re, im, want_im != 0, lb, partials);
}

__global__ void __launch_bounds__(256) k_axpby(double a, const double* __restrict__ X, double b, double* __restrict__ Y,
                                               int64_t n) {
    int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x;
    if (i < n) Y[i] = a * X[i] + b * Y[i];
}

__global__ void __launch_bounds__(256) k_scale(double2* __restrict__ out, const double2* __restrict__ in, double sr,
                                               double si, int64_t n) {
    int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x;
    if (i < n) {
        double2 v = in[i];
        cx<double> z = cmul(mk<double>(sr, si), mk<double>(v.x, v.y));
        out[i] = make_double2(z.re, z.im);
    }
}

// r = b - Ap (axpby!(1,b,0,r); axpby!(-1,Ap,1,r)); p = r; rr = |r|^2  (cg.jl:91-95)
template <int NCOMP, typename R = double>
__global__ void __launch_bounds__(256) k_cg_setup_r(const typename Vec2<R>::T* __restrict__ b,
                                                    const typename Vec2<R>::T* __restrict__ Ap,
                                                    typename Vec2<R>::T* __restrict__ r,
                                                    typename Vec2<R>::T* __restrict__ p, int V, double* partials,
                                                    CGState* st) {
    if (st->done) return;
    const int lb = blockIdx.x;
    const int s = lb * 256 + threadIdx.x;
    double re = 0.0;
    if (s < V) {
        cx<R> acc = mk<R>(R(0), R(0));
#pragma unroll
        for (int k = 0; k < NCOMP; ++k) {
            const long i = (long)k * V + s;
            const auto bv = b[i], av = Ap[i];
            typename Vec2<R>::T rv;
            rv.x = R(-1) * av.x + bv.x;
            rv.y = R(-1) * av.y + bv.y;
            r[i] = rv;
            p[i] = rv;
            cx<R> rc = mk<R>(rv.x, rv.y);
            acc = cadd(acc, cmul(cconj(rc), rc));
        }
        re = double(acc.re);
    }
    store_partials(re, 0.0, false, lb, partials);
}

// x = alpha*p + x; r = (-alpha)*Ap + r; rr_new = |r|^2  (cg.jl:116-120)
template <int NCOMP, typename R = double>
__global__ void __launch_bounds__(256) k_cg_update_xr(typename Vec2<R>::T* __restrict__ x,
                                                      typename Vec2<R>::T* __restrict__ r,
                                                      const typename Vec2<R>::T* __restrict__ p,
                                                      const typename Vec2<R>::T* __restrict__ Ap, int V,
                                                      double* partials, CGState* st) {
    if (st->done) return;
    const R alpha = R(st->alpha);
    const R malpha = -alpha;
    const int lb = blockIdx.x;
    const int s = lb * 256 + threadIdx.x;
    double re = 0.0;
    if (s < V) {
        cx<R> acc = mk<R>(R(0), R(0));
#pragma unroll
        for (int k = 0; k < NCOMP; ++k) {
            const long i = (long)k * V + s;
            auto pv = p[i], xv = x[i], av = Ap[i], rv = r[i];
            xv.x = alpha * pv.x + R(1) * xv.x;
            xv.y = alpha * pv.y + R(1) * xv.y;
            rv.x = malpha * av.x + R(1) * rv.x;
            rv.y = malpha * av.y + R(1) * rv.y;
            x[i] = xv;
            r[i] = rv;
            cx<R> rc = mk<R>(rv.x, rv.y);
            acc = cadd(acc, cmul(cconj(rc), rc));
        }
        re = double(acc.re);
    }
    store_partials(re, 0.0, false, lb, partials);
}

// p = 1*r + beta*p  (cg.jl:130-131)
template <int NCOMP, typename R = double>
__global__ void __launch_bounds__(256) k_cg_update_p(const typename Vec2<R>::T* __restrict__ r,
                                                     typename Vec2<R>::T* __restrict__ p, int V, CGState* st) {
    if (st->done) return;
    const R beta = R(st->beta);
    const int s = blockIdx.x * 256 + threadIdx.x;
    if (s < V) {
#pragma unroll
        for (int k = 0; k < NCOMP; ++k) {
            const long i = (long)k * V + s;
            auto rv = r[i], pv = p[i];
            pv.x = R(1) * rv.x + beta * pv.x;
            pv.y = R(1) * rv.y + beta * pv.y;
            p[i] = pv;
        }
    }
}

__global__ void k_cg_init_state(CGState* st, double rtol, double atol, int maxiter) {
    if (threadIdx.x == 0) {
        CGState s;
        s.bb = s.bnorm = s.tol = s.rr = s.pAp = s.alpha = s.beta = s.rr_new = s.rnorm = 0.0;
        s.rtol = rtol;
        s.atol = atol;
        s.iteration = 0;
        s.maxiter = maxiter;
        s.done = 0;
        s.reason = LMX_CG_MAXITER;
        s.converged = 0;
        s.pad[0] = s.pad[1] = s.pad[2] = 0;
        *st = s;
    }
}

static inline dim3 grid_for(int64_t n) { return dim3((unsigned)((n + 255) / 256)); }

int launch_dot(lmx_context* ctx, const double2* A, const double2* B, int fin_kind, CGState* st, double* out2) {
    const int nb = nblocks(ctx->V);
    const bool want_im = fin_kind == FIN_OUT;
    if (ctx->nc == 3)
        hipLaunchKernelGGL(k_dot<12>, dim3(nb), dim3(256), 0, ctx->stream, A, B, (int)ctx->V, ctx->partials,
                           want_im ? 1 : 0, st);
    else
        hipLaunchKernelGGL(k_dot<8>, dim3(nb), dim3(256), 0, ctx->stream, A, B, (int)ctx->V, ctx->partials,
                           want_im ? 1 : 0, st);
    LMX_HIP(hipGetLastError());
    return launch_finalize(ctx, nb, fin_kind, want_im, st, out2);
}

int launch_axpby(lmx_context* ctx, double a, const double2* X, double b, double2* Y, int64_t n_complex) {
    const int64_t n = 2 * n_complex;
    if (n == 0) return LMX_OK;
    hipLaunchKernelGGL(k_axpby, grid_for(n), dim3(256), 0, ctx->stream, a, (const double*)X, b, (double*)Y, n);
    LMX_HIP(hipGetLastError());
    return LMX_OK;
}

// Zero fill as a kernel rather than hipMemsetAsync: the entry points that clear
// buffers may be captured into a caller's HIP graph, where a kernel node
// replays like every other launch.
__global__ void __launch_bounds__(256) k_zero(double2* __restrict__ p, int64_t n) {
    const int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x;
    if (i < n) p[i] = make_double2(0.0, 0.0);
}

int launch_zero(lmx_context* ctx, double2* p, int64_t n) {
    if (n == 0) return LMX_OK;
    hipLaunchKernelGGL(k_zero, grid_for(n), dim3(256), 0, ctx->stream, p, n);
    LMX_HIP(hipGetLastError());
    return LMX_OK;
}

int launch_scale(lmx_context* ctx, double2* out, const double2* in, double sr, double si, int64_t n) {
    if (n == 0) return LMX_OK;
    hipLaunchKernelGGL(k_scale, grid_for(n), dim3(256), 0, ctx->stream, out, in, sr, si, n);
    LMX_HIP(hipGetLastError());
    return LMX_OK;
}

int launch_cg_setup_r(lmx_context* ctx, const double2* b, const double2* Ap, double2* r, double2* p, CGState* st) {
    if (ctx->nc == 3)
        hipLaunchKernelGGL(k_cg_setup_r<12>, dim3(nblocks(ctx->V)), dim3(256), 0, ctx->stream, b, Ap, r, p,
                           (int)ctx->V, ctx->partials, st);
    else
        hipLaunchKernelGGL(k_cg_setup_r<8>, dim3(nblocks(ctx->V)), dim3(256), 0, ctx->stream, b, Ap, r, p,
                           (int)ctx->V, ctx->partials, st);
    LMX_HIP(hipGetLastError());
    return launch_finalize(ctx, nblocks(ctx->V), FIN_RR0, false, st, nullptr);
}

int launch_cg_update_xr(lmx_context* ctx, double2* x, double2* r, const double2* p, const double2* Ap, CGState* st) {
    if (ctx->nc == 3)
        hipLaunchKernelGGL(k_cg_update_xr<12>, dim3(nblocks(ctx->V)), dim3(256), 0, ctx->stream, x, r, p, Ap,
                           (int)ctx->V, ctx->partials, st);
    else
        hipLaunchKernelGGL(k_cg_update_xr<8>, dim3(nblocks(ctx->V)), dim3(256), 0, ctx->stream, x, r, p, Ap,
                           (int)ctx->V, ctx->partials, st);
    LMX_HIP(hipGetLastError());
    return launch_finalize(ctx, nblocks(ctx->V), FIN_RRNEW, false, st, nullptr);
}

int launch_cg_update_p(lmx_context* ctx, const double2* r, double2* p, CGState* st) {
    if (ctx->nc == 3)
        hipLaunchKernelGGL(k_cg_update_p<12>, dim3(nblocks(ctx->V)), dim3(256), 0, ctx->stream, r, p, (int)ctx->V, st);
    else
        hipLaunchKernelGGL(k_cg_update_p<8>, dim3(nblocks(ctx->V)), dim3(256), 0, ctx->stream, r, p, (int)ctx->V, st);
    LMX_HIP(hipGetLastError());
    return LMX_OK;
}

// The same CG vector kernels on ComplexF32 fields (NC = 3).
int launch_dot_f32(lmx_context* ctx, const float2* A, const float2* B, int fin_kind, CGState* st, double* out2) {
    const int nb = nblocks(ctx->V);
    const bool want_im = fin_kind == FIN_OUT;
    hipLaunchKernelGGL((k_dot<12, float>), dim3(nb), dim3(256), 0, ctx->stream, A, B, (int)ctx->V, ctx->partials,
                       want_im ? 1 : 0, st);
    LMX_HIP(hipGetLastError());
    return launch_finalize(ctx, nb, fin_kind, want_im, st, out2);
}

int launch_cg_setup_r_f32(lmx_context* ctx, const float2* b, const float2* Ap, float2* r, float2* p, CGState* st) {
    hipLaunchKernelGGL((k_cg_setup_r<12, float>), dim3(nblocks(ctx->V)), dim3(256), 0, ctx->stream, b, Ap, r, p,
                       (int)ctx->V, ctx->partials, st);
    LMX_HIP(hipGetLastError());
    return launch_finalize(ctx, nblocks(ctx->V), FIN_RR0, false, st, nullptr);
}

int launch_cg_update_xr_f32(lmx_context* ctx, float2* x, float2* r, const float2* p, const float2* Ap, CGState* st) {
    hipLaunchKernelGGL((k_cg_update_xr<12, float>), dim3(nblocks(ctx->V)), dim3(256), 0, ctx->stream, x, r, p, Ap,
                       (int)ctx->V, ctx->partials, st);
    LMX_HIP(hipGetLastError());
    return launch_finalize(ctx, nblocks(ctx->V), FIN_RRNEW, false, st, nullptr);
}

int launch_cg_update_p_f32(lmx_context* ctx, const float2* r, float2* p, CGState* st) {
    hipLaunchKernelGGL((k_cg_update_p<12, float>), dim3(nblocks(ctx->V)), dim3(256), 0, ctx->stream, r, p,
                       (int)ctx->V, st);
    LMX_HIP(hipGetLastError());
    return LMX_OK;
}

int launch_cg_init_state(lmx_context* ctx, CGState* st, double rtol, double atol, int maxiter) {
    hipLaunchKernelGGL(k_cg_init_state, dim3(1), dim3(64), 0, ctx->stream, st, rtol, atol, maxiter);
    LMX_HIP(hipGetLastError());
    return LMX_OK;
}

}  // namespace lmx
