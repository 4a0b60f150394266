// clover.hip -- clover field strength cache and precision conversions.
//
// F_mu,nu = (Q - Q^dagger)/8 with Q the four-leaf clover, planes
// (12,13,14,23,24,34): kernel_clover_field_strength_halo! NC=3 and
// _clover_store_antihermitian3! (src/Operators/WilsonDiracCloverOperator.jl:164-226).
// The cache holds only the 9 independent reals of each anti-Hermitian plane
// (27 complex slots per site, 432 B FP64 instead of 864 B for full matrices);
// the full matrices are recovered exactly (diagonal real parts are exactly 0,
// lower entries are -conj of upper ones, as the reference stores them).
#include "lmx_internal.h"

namespace lmx {

__device__ __forceinline__ void ld_link(const double2* __restrict__ U, int mu, int s, int V, cx<double> u[9]) {
#pragma unroll
    for (int k = 0; k < 9; ++k) {
        double2 v = U[(long)(mu * 9 + k) * V + s];
        u[k] = mk<double>(v.x, v.y);
    }
}
__device__ __forceinline__ void dag3(cx<double> u[9]) {
    cx<double> t[9];
#pragma unroll
    for (int i = 0; i < 3; ++i)
#pragma unroll
        for (int j = 0; j < 3; ++j) t[3 * i + j] = cconj(u[3 * j + i]);
#pragma unroll
    for (int k = 0; k < 9; ++k) u[k] = t[k];
}
// _clover_mul_matrix3 (:137-149): C = A*B with muladdmulti rows
__device__ __forceinline__ void mul3(const cx<double> A[9], const cx<double> B[9], cx<double> out[9]) {
    cx<double> r[9];
#pragma unroll
    for (int i = 0; i < 3; ++i)
#pragma unroll
        for (int j = 0; j < 3; ++j)
            r[3 * i + j] = muladdmulti(A[3 * i], B[j], A[3 * i + 1], B[3 + j], A[3 * i + 2], B[6 + j]);
#pragma unroll
    for (int k = 0; k < 9; ++k) out[k] = r[k];
}

__device__ __forceinline__ int shift(int s, int mu, int dir, const Geom& g) {
    int x[4];
    site_coords(s, g, x);
    bool w;
    return dir > 0 ? nbr_fwd(s, x, mu, g, w) : nbr_bwd(s, x, mu, g, w);
}

// leaf = ((A*B)*C)*D, with optional daggers on each factor
__device__ __forceinline__ void leaf(const double2* __restrict__ U, int V, int ma, int sa, bool da, int mb, int sb,
                                     bool db, int mc, int sc, bool dc, int md, int sd, bool dd, cx<double> out[9]) {
    cx<double> a[9], b[9];
    ld_link(U, ma, sa, V, a);
    if (da) dag3(a);
    ld_link(U, mb, sb, V, b);
    if (db) dag3(b);
    mul3(a, b, a);
    ld_link(U, mc, sc, V, b);
    if (dc) dag3(b);
    mul3(a, b, a);
    ld_link(U, md, sd, V, b);
    if (dd) dag3(b);
    mul3(a, b, out);
}

__global__ void __launch_bounds__(256) k_clover_build(const double2* __restrict__ U, double2* __restrict__ F, Geom g) {
    const int s = blockIdx.x * 256 + threadIdx.x;
    if (s >= g.V) return;
    const int V = g.V;
    const int pmu[6] = {0, 0, 0, 1, 1, 2};
    const int pnu[6] = {1, 2, 3, 2, 3, 3};
    double dg[18];
    cx<double> up[18];
#pragma unroll 1
    for (int pl = 0; pl < 6; ++pl) {
        const int mu = pmu[pl], nu = pnu[pl];
        const int xpmu = shift(s, mu, 1, g), xpnu = shift(s, nu, 1, g);
        const int xmmu = shift(s, mu, -1, g), xmnu = shift(s, nu, -1, g);
        const int xpnu_mmu = shift(xpnu, mu, -1, g);
        const int xmmu_mnu = shift(xmmu, nu, -1, g);
        const int xpmu_mnu = shift(xmnu, mu, 1, g);
        cx<double> q[9], t[9];
        leaf(U, V, mu, s, false, nu, xpmu, false, mu, xpnu, true, nu, s, true, q);
        leaf(U, V, nu, s, false, mu, xpnu_mmu, true, nu, xmmu, true, mu, xmmu, false, t);
#pragma unroll
        for (int k = 0; k < 9; ++k) q[k] = cadd(q[k], t[k]);
        leaf(U, V, mu, xmmu, true, nu, xmmu_mnu, true, mu, xmmu_mnu, false, nu, xmnu, false, t);
#pragma unroll
        for (int k = 0; k < 9; ++k) q[k] = cadd(q[k], t[k]);
        leaf(U, V, nu, xmnu, true, mu, xmnu, false, nu, xpmu_mnu, false, mu, s, true, t);
#pragma unroll
        for (int k = 0; k < 9; ++k) q[k] = cadd(q[k], t[k]);
        const double f = 1.0 / 8.0;
        dg[3 * pl + 0] = rmul(f, csub(q[0], cconj(q[0]))).im;
        dg[3 * pl + 1] = rmul(f, csub(q[4], cconj(q[4]))).im;
        dg[3 * pl + 2] = rmul(f, csub(q[8], cconj(q[8]))).im;
        up[3 * pl + 0] = rmul(f, csub(q[1], cconj(q[3])));
        up[3 * pl + 1] = rmul(f, csub(q[2], cconj(q[6])));
        up[3 * pl + 2] = rmul(f, csub(q[5], cconj(q[7])));
    }
#pragma unroll
    for (int qd = 0; qd < 9; ++qd) F[(long)qd * V + s] = make_double2(dg[2 * qd], dg[2 * qd + 1]);
#pragma unroll
    for (int qd = 0; qd < 18; ++qd) F[(long)(9 + qd) * V + s] = make_double2(up[qd].re, up[qd].im);
}

// compact cache -> full planes [6][V][3][3] complex
__global__ void __launch_bounds__(256) k_clover_export(const double2* __restrict__ F, double2* __restrict__ out, int V) {
    const int s = blockIdx.x * 256 + threadIdx.x;
    if (s >= V) return;
    for (int pl = 0; pl < 6; ++pl) {
        double d[3];
        for (int i = 0; i < 3; ++i) {
            const int q = 3 * pl + i;
            double2 v = F[(long)(q / 2) * V + s];
            d[i] = (q % 2) ? v.y : v.x;
        }
        double2 u01 = F[(long)(9 + 3 * pl + 0) * V + s];
        double2 u02 = F[(long)(9 + 3 * pl + 1) * V + s];
        double2 u12 = F[(long)(9 + 3 * pl + 2) * V + s];
        double2* o = out + ((long)pl * V + s) * 9;
        o[0] = make_double2(0.0, d[0]);
        o[1] = u01;
        o[2] = u02;
        o[3] = make_double2(-u01.x, u01.y);
        o[4] = make_double2(0.0, d[1]);
        o[5] = u12;
        o[6] = make_double2(-u02.x, u02.y);
        o[7] = make_double2(-u12.x, u12.y);
        o[8] = make_double2(0.0, d[2]);
    }
}

__global__ void __launch_bounds__(256) k_convert(const double2* __restrict__ in, float2* __restrict__ o32,
                                                 __half2* __restrict__ o16, int64_t n) {
    int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x;
    if (i >= n) return;
    double2 v = in[i];
    float2 f = make_float2((float)v.x, (float)v.y);
    if (o32) o32[i] = f;
    if (o16) o16[i] = __float22half2_rn(f);
}

int launch_clover_build_f64(lmx_context* ctx, const double2* U, double2* clover) {
    if (ctx->V == 0) return LMX_OK;
    hipLaunchKernelGGL(k_clover_build, dim3(nblocks(ctx->V)), dim3(256), 0, ctx->stream, U, clover, ctx->g);
    LMX_HIP(hipGetLastError());
    return LMX_OK;
}

int launch_clover_export(lmx_context* ctx, const double2* clover, double2* F_out) {
    if (ctx->V == 0) return LMX_OK;
    hipLaunchKernelGGL(k_clover_export, dim3(nblocks(ctx->V)), dim3(256), 0, ctx->stream, clover, F_out, (int)ctx->V);
    LMX_HIP(hipGetLastError());
    return LMX_OK;
}

static int convert(lmx_context* ctx, const double2* in, float2* o32, __half2* o16, int64_t n) {
    if (n == 0) return LMX_OK;
    hipLaunchKernelGGL(k_convert, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, ctx->stream, in, o32, o16, n);
    LMX_HIP(hipGetLastError());
    return LMX_OK;
}

// FP16 links packed per direction: entries 0-3 and 4-7 as two 16-byte words
// (int4 [4][2][V]), entry 8 as one 4-byte word (half2 [4][V]) after them.
__global__ void __launch_bounds__(256) k_pack_links16(const double2* __restrict__ U, int4* __restrict__ P, int V) {
    const int s = blockIdx.x * 256 + threadIdx.x;
    const int mu = blockIdx.y;
    if (s >= V) return;
    int w[9];
#pragma unroll
    for (int k = 0; k < 9; ++k) {
        double2 v = U[((size_t)mu * 9 + k) * V + s];
        __half2 h = __float22half2_rn(make_float2((float)v.x, (float)v.y));
        w[k] = *reinterpret_cast<int*>(&h);
    }
    P[((size_t)2 * mu) * V + s] = make_int4(w[0], w[1], w[2], w[3]);
    P[((size_t)2 * mu + 1) * V + s] = make_int4(w[4], w[5], w[6], w[7]);
    reinterpret_cast<int*>(P + (size_t)8 * V)[(size_t)mu * V + s] = w[8];
}

int launch_convert_links(lmx_context* ctx, const double2* U, float2* U32, __half2* U16) {
    LMX_TRY(convert(ctx, U, U32, nullptr, 36 * ctx->V));
    if (U16 && ctx->V > 0) {
        hipLaunchKernelGGL(k_pack_links16, dim3(nblocks(ctx->V), 4), dim3(256), 0, ctx->stream, U, (int4*)U16,
                           (int)ctx->V);
        LMX_HIP(hipGetLastError());
    }
    return LMX_OK;
}

int launch_convert_clover(lmx_context* ctx, const double2* F, float2* F32, __half2* F16) {
    return convert(ctx, F, F32, F16, 27 * ctx->V);
}

}  // namespace lmx
