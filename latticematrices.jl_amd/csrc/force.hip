// force.hip -- hand-written adjoint of the Wilson hopping term (the fermion
// force) and the traceless anti-Hermitian projection.
//
// Link cotangent, per site x and direction mu (row r, column c):
//   dU_mu(x)[r,c] += coefficient * ( sum_k P-(dres(x))_r,k conj(P-(psi(x+mu))_c,k)
//                                  + sum_k P+(psi(x))_r,k  conj(P+(dres(x+mu))_c,k) )
// with P-/P+ the half-spin projections of (1 -+ gamma_mu): the Enzyme reverse
// rule kernel _kernel_wilson_link_pullback! (ext/AD/WilsonDiracOperator.jl:62-91)
// and _wilson_halfspin_link_pullback_direction3!/_row3!
// (src/Operators/WilsonDiracOperator.jl:354-410), in its operation order.
// Values at x+mu carry the spinor boundary phase, as the reference's
// set_halo! of both fields does.
#include "lmx_internal.h"

namespace lmx {

template <int MU>
__device__ __forceinline__ void pullback_mu(double2* __restrict__ dU, const double2* __restrict__ dres,
                                            const double2* __restrict__ psi, const Geom& g, int s, const int x[4],
                                            const cx<double> d0[12], const cx<double> p0[12], double coefficient) {
    bool w;
    const int sp = nbr_fwd(s, x, MU, g, w);
    cx<double> d1[12], p1[12];
#pragma unroll
    for (int k = 0; k < 12; ++k) {
        double2 a = dres[(long)k * g.V + sp], b = psi[(long)k * g.V + sp];
        d1[k] = mk<double>(a.x, a.y);
        p1[k] = mk<double>(b.x, b.y);
    }
    if (w && !g.phase_one[MU]) {
        phase_spinor(d1, g.ph_re[MU], g.ph_im[MU]);
        phase_spinor(p1, g.ph_re[MU], g.ph_im[MU]);
    }
    cx<double> plus_psi[6], minus_dres[6], dplus[6], mpsi[6];
    half_project<-1, MU>(p1, plus_psi);
    half_project<1, MU>(d1, minus_dres);
    half_project<-1, MU>(d0, dplus);
    half_project<1, MU>(p0, mpsi);
#pragma unroll
    for (int row = 0; row < 3; ++row) {
#pragma unroll
        for (int col = 0; col < 3; ++col) {
            cx<double> v = cmul(dplus[2 * row], cconj(plus_psi[2 * col]));
            v = cadd(v, cmul(dplus[2 * row + 1], cconj(plus_psi[2 * col + 1])));
            v = cadd(v, cmul(mpsi[2 * row], cconj(minus_dres[2 * col])));
            v = cadd(v, cmul(mpsi[2 * row + 1], cconj(minus_dres[2 * col + 1])));
            const long i = (long)(MU * 9 + row * 3 + col) * g.V + s;
            double2 o = dU[i];
            cx<double> nv = cadd(mk<double>(o.x, o.y), rmul(coefficient, v));
            dU[i] = make_double2(nv.re, nv.im);
        }
    }
}

__global__ void __launch_bounds__(256) k_pullback(double2* __restrict__ dU, const double2* __restrict__ dres,
                                                  const double2* __restrict__ psi, Geom g, double coefficient) {
    const int s = xcd_block(blockIdx.x, gridDim.x) * 256 + threadIdx.x;
    if (s >= g.V) return;
    int x[4];
    site_coords(s, g, x);
    cx<double> d0[12], p0[12];
#pragma unroll
    for (int k = 0; k < 12; ++k) {
        double2 a = dres[(long)k * g.V + s], b = psi[(long)k * g.V + s];
        d0[k] = mk<double>(a.x, a.y);
        p0[k] = mk<double>(b.x, b.y);
    }
    pullback_mu<0>(dU, dres, psi, g, s, x, d0, p0, coefficient);
    pullback_mu<1>(dU, dres, psi, g, s, x, d0, p0, coefficient);
    pullback_mu<2>(dU, dres, psi, g, s, x, d0, p0, coefficient);
    pullback_mu<3>(dU, dres, psi, g, s, x, d0, p0, coefficient);
}

// F = TA(U G^dagger): M[i][j] = muladdmulti over k of U[i][k] conj(G[j][k]),
// then kernel_traceless_antihermitian! NC=3 (src/TA/TA.jl:68-116).
__global__ void __launch_bounds__(256) k_ta_project(double2* __restrict__ F, const double2* __restrict__ U,
                                                    const double2* __restrict__ G, int V) {
    const int s = blockIdx.x * 256 + threadIdx.x;
    if (s >= V) return;
    for (int mu = 0; mu < 4; ++mu) {
        cx<double> u[9], gg[9], m[9];
#pragma unroll
        for (int k = 0; k < 9; ++k) {
            double2 a = U[(long)(mu * 9 + k) * V + s], b = G[(long)(mu * 9 + k) * V + s];
            u[k] = mk<double>(a.x, a.y);
            gg[k] = mk<double>(b.x, b.y);
        }
#pragma unroll
        for (int i = 0; i < 3; ++i)
#pragma unroll
            for (int j = 0; j < 3; ++j)
                m[3 * i + j] = muladdmulti(u[3 * i], cconj(gg[3 * j]), u[3 * i + 1], cconj(gg[3 * j + 1]),
                                           u[3 * i + 2], cconj(gg[3 * j + 2]));
        const double fac13 = 1.0 / 3.0;
        const double tri = fac13 * ((m[0].im + m[4].im) + m[8].im);
        cx<double> o[9];
        o[0] = mk<double>(0.0, m[0].im - tri);
        o[4] = mk<double>(0.0, m[4].im - tri);
        o[8] = mk<double>(0.0, m[8].im - tri);
        cx<double> x12 = csub(m[1], cconj(m[3]));
        cx<double> x13 = csub(m[2], cconj(m[6]));
        cx<double> x23 = csub(m[5], cconj(m[7]));
        o[1] = rmul(0.5, x12);
        o[2] = rmul(0.5, x13);
        o[5] = rmul(0.5, x23);
        o[3] = rmul(0.5, cneg(cconj(x12)));
        o[6] = rmul(0.5, cneg(cconj(x13)));
        o[7] = rmul(0.5, cneg(cconj(x23)));
#pragma unroll
        for (int k = 0; k < 9; ++k) F[(long)(mu * 9 + k) * V + s] = make_double2(o[k].re, o[k].im);
    }
}

int launch_pullback_f64(lmx_context* ctx, double2* dU, const double2* dres, const double2* psi, double coefficient) {
    if (ctx->V == 0) return LMX_OK;
    hipLaunchKernelGGL(k_pullback, dim3(nblocks(ctx->V)), dim3(256), 0, ctx->stream, dU, dres, psi, ctx->g,
                       coefficient);
    LMX_HIP(hipGetLastError());
    return LMX_OK;
}

int launch_ta_project(lmx_context* ctx, double2* F, const double2* U, const double2* G) {
    if (ctx->V == 0) return LMX_OK;
    hipLaunchKernelGGL(k_ta_project, dim3(nblocks(ctx->V)), dim3(256), 0, ctx->stream, F, U, G, (int)ctx->V);
    LMX_HIP(hipGetLastError());
    return LMX_OK;
}

}  // namespace lmx
