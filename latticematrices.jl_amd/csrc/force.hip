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

// ---------------------------------------------------------------------------
// Wilson-clover link pullback (wilson_clover_link_pullback!,
// src/Operators/WilsonCloverPullback.jl:349-389): the hopping part above plus
//   1. the field-strength cotangent per plane
//      dF[c][ic] += coefficient * sum_out dres[c,out] conj(sum_in g[out,in] psi[ic,in])
//      with g = gamma_mu gamma_nu (_kernel_clover_field_cotangent_plane!, :102-119);
//   2. its chain rule through F = (Q - Q^dagger)/8 and the four leaves of Q
//      onto the links (_kernel_clover_links_pullback!, :236-280): for each link
//      and each occurrence K in a leaf, with dQ = (dF - dF^dagger)/8 at the
//      leaf's anchor, left^dagger dQ right^dagger, or right dQ^dagger left for
//      an adjointed factor.
// Matrix products are plain complex arithmetic summed left to right; the
// operation order follows the reference's loops (planes 12..34, leaves,
// occurrences 1..4).
// ---------------------------------------------------------------------------
struct GammaProducts {
    double2 g[6][4][4];
};

__device__ __forceinline__ void mat_mul_plain(const cx<double>* A, const cx<double>* B, cx<double>* out) {
    cx<double> r[9];
#pragma unroll
    for (int i = 0; i < 3; ++i)
#pragma unroll
        for (int j = 0; j < 3; ++j)
            r[3 * i + j] =
                cadd(cadd(cmul(A[3 * i], B[j]), cmul(A[3 * i + 1], B[3 + j])), cmul(A[3 * i + 2], B[6 + j]));
#pragma unroll
    for (int k = 0; k < 9; ++k) out[k] = r[k];
}
__device__ __forceinline__ void mat_dag(const cx<double>* A, cx<double>* out) {
    cx<double> r[9];
#pragma unroll
    for (int i = 0; i < 3; ++i)
#pragma unroll
        for (int j = 0; j < 3; ++j) r[3 * i + j] = cconj(A[3 * j + i]);
#pragma unroll
    for (int k = 0; k < 9; ++k) out[k] = r[k];
}

__global__ void __launch_bounds__(256) k_clover_cotangent(double2* __restrict__ dF, const double2* __restrict__ dres,
                                                          const double2* __restrict__ psi, int V, double coefficient,
                                                          GammaProducts gp) {
    const int s = blockIdx.x * 256 + threadIdx.x;
    if (s >= V) return;
    cx<double> d[12], p[12];
#pragma unroll
    for (int k = 0; k < 12; ++k) {
        double2 a = dres[(long)k * V + s], b = psi[(long)k * V + s];
        d[k] = mk<double>(a.x, a.y);
        p[k] = mk<double>(b.x, b.y);
    }
    for (int pl = 0; pl < 6; ++pl)
        for (int ic = 0; ic < 3; ++ic)
            for (int c = 0; c < 3; ++c) {
                cx<double> value = mk<double>(0.0, 0.0);
                for (int o = 0; o < 4; ++o) {
                    cx<double> tp = mk<double>(0.0, 0.0);
                    for (int i = 0; i < 4; ++i) {
                        const double2 gg = gp.g[pl][o][i];
                        tp = cadd(tp, cmul(mk<double>(gg.x, gg.y), p[i * 3 + ic]));
                    }
                    value = cadd(value, cmul(d[o * 3 + c], cconj(tp)));
                }
                const long e = ((long)pl * 9 + c * 3 + ic) * V + s;
                double2 cur = dF[e];
                cx<double> nv = cadd(mk<double>(cur.x, cur.y), rmul(coefficient, value));
                dF[e] = make_double2(nv.re, nv.im);
            }
}

// leaf p of plane (mu, nu): (mu,nu,-mu,-nu), (nu,-mu,-nu,mu), (-mu,-nu,mu,nu), (-nu,mu,nu,-mu)
// (_clover_path_descriptor, :140-159): factor origins relative to the anchor,
// directions and adjoint flags.
struct LeafPath {
    int origin[4][4];
    int dir[4];
    int dag[4];
};
__device__ __forceinline__ void make_leaf(int mu, int nu, int p, LeafPath& lp) {
    const int m = mu + 1, n = nu + 1;
    int sd[4];
    if (p == 0) { sd[0] = m; sd[1] = n; sd[2] = -m; sd[3] = -n; }
    else if (p == 1) { sd[0] = n; sd[1] = -m; sd[2] = -n; sd[3] = m; }
    else if (p == 2) { sd[0] = -m; sd[1] = -n; sd[2] = m; sd[3] = n; }
    else { sd[0] = -n; sd[1] = m; sd[2] = n; sd[3] = -m; }
    int pos[4] = {0, 0, 0, 0};
    for (int k = 0; k < 4; ++k) {
        const int d = (sd[k] > 0 ? sd[k] : -sd[k]) - 1;
        if (sd[k] < 0) pos[d] -= 1;
        for (int e = 0; e < 4; ++e) lp.origin[k][e] = pos[e];
        lp.dir[k] = d;
        lp.dag[k] = sd[k] < 0;
        if (sd[k] > 0) pos[d] += 1;
    }
}

__device__ __forceinline__ int site_of(const Geom& g, const int x[4], const int* off, int sign) {
    int s = 0;
    for (int d = 0; d < 4; ++d) {
        int c = x[d] + sign * off[d];
        c = c < 0 ? c + g.L[d] : (c >= g.L[d] ? c - g.L[d] : c);
        s += c * g.stride[d];
    }
    return s;
}

__device__ __forceinline__ void ld_link(const double2* __restrict__ U, int V, int mu, int s, cx<double>* u) {
#pragma unroll
    for (int k = 0; k < 9; ++k) {
        double2 v = U[(long)(mu * 9 + k) * V + s];
        u[k] = mk<double>(v.x, v.y);
    }
}

__global__ void __launch_bounds__(256) k_clover_links_pullback(double2* __restrict__ dU, const double2* __restrict__ U,
                                                               const double2* __restrict__ dF, Geom g) {
    const int y = blockIdx.x * 256 + threadIdx.x;
    if (y >= g.V) return;
    const int V = g.V;
    int xy[4];
    site_coords(y, g, xy);
    const int pmu[6] = {0, 0, 0, 1, 1, 2}, pnu[6] = {1, 2, 3, 2, 3, 3};
    for (int pl = 0; pl < 6; ++pl) {
        const int mu = pmu[pl], nu = pnu[pl];
        cx<double> dmu[9], dnu[9];
        for (int k = 0; k < 9; ++k) dmu[k] = dnu[k] = mk<double>(0.0, 0.0);
        for (int p = 0; p < 4; ++p) {
            LeafPath lp;
            make_leaf(mu, nu, p, lp);
            cx<double> hmu[9], hnu[9];
            for (int k = 0; k < 9; ++k) hmu[k] = hnu[k] = mk<double>(0.0, 0.0);
            for (int K = 0; K < 4; ++K) {
                // anchor = y - origin_K; factor j at anchor + origin_j
                const int anchor = site_of(g, xy, lp.origin[K], -1);
                int xa[4];
                site_coords(anchor, g, xa);
                cx<double> left[9], right[9], t[9], m[9];
                const bool has_left = K > 0, has_right = K < 3;
                // left = M1 ... M_{K-1}
                for (int j = 0; j < K; ++j) {
                    ld_link(U, V, lp.dir[j], site_of(g, xa, lp.origin[j], 1), m);
                    if (lp.dag[j]) mat_dag(m, m);
                    if (j == 0) {
                        for (int k = 0; k < 9; ++k) left[k] = m[k];
                    } else {
                        mat_mul_plain(left, m, left);
                    }
                }
                // right = M_{K+1} ... M4
                for (int j = K + 1; j < 4; ++j) {
                    ld_link(U, V, lp.dir[j], site_of(g, xa, lp.origin[j], 1), m);
                    if (lp.dag[j]) mat_dag(m, m);
                    if (j == K + 1) {
                        for (int k = 0; k < 9; ++k) right[k] = m[k];
                    } else {
                        mat_mul_plain(right, m, right);
                    }
                }
                cx<double> f[9], dq[9];
#pragma unroll
                for (int k = 0; k < 9; ++k) {
                    double2 v = dF[((long)pl * 9 + k) * V + anchor];
                    f[k] = mk<double>(v.x, v.y);
                }
#pragma unroll
                for (int r = 0; r < 3; ++r)
#pragma unroll
                    for (int c = 0; c < 3; ++c) dq[3 * r + c] = rmul(0.125, csub(f[3 * r + c], cconj(f[3 * c + r])));
                cx<double> h[9];
                if (lp.dag[K]) {  // right * dQ^dagger * left
                    mat_dag(dq, dq);
                    if (has_right)
                        mat_mul_plain(right, dq, t);
                    else
                        for (int k = 0; k < 9; ++k) t[k] = dq[k];
                    if (has_left)
                        mat_mul_plain(t, left, h);
                    else
                        for (int k = 0; k < 9; ++k) h[k] = t[k];
                } else {  // left^dagger * dQ * right^dagger
                    if (has_left) {
                        mat_dag(left, left);
                        mat_mul_plain(left, dq, t);
                    } else {
                        for (int k = 0; k < 9; ++k) t[k] = dq[k];
                    }
                    if (has_right) {
                        mat_dag(right, right);
                        mat_mul_plain(t, right, h);
                    } else {
                        for (int k = 0; k < 9; ++k) h[k] = t[k];
                    }
                }
                cx<double>* acc = lp.dir[K] == mu ? hmu : hnu;
                for (int k = 0; k < 9; ++k) acc[k] = cadd(acc[k], h[k]);
            }
            for (int k = 0; k < 9; ++k) {
                dmu[k] = cadd(dmu[k], hmu[k]);
                dnu[k] = cadd(dnu[k], hnu[k]);
            }
        }
        for (int k = 0; k < 9; ++k) {
            const long em = (long)(mu * 9 + k) * V + y, en = (long)(nu * 9 + k) * V + y;
            double2 a = dU[em];
            cx<double> na = cadd(mk<double>(a.x, a.y), dmu[k]);
            dU[em] = make_double2(na.re, na.im);
            double2 b = dU[en];
            cx<double> nb = cadd(mk<double>(b.x, b.y), dnu[k]);
            dU[en] = make_double2(nb.re, nb.im);
        }
    }
}

// gamma_mu gamma_nu of the chiral gamma matrices (src/Operators/Operators.jl:44-67),
// exact (entries 0, +-1, +-i), summed as the oracle does
static GammaProducts gamma_products() {
    struct C2 {
        double re, im;
    };
    C2 gm[4][4][4] = {};
    gm[0][0][3] = {0, -1}; gm[0][1][2] = {0, -1}; gm[0][2][1] = {0, 1}; gm[0][3][0] = {0, 1};
    gm[1][0][3] = {-1, 0}; gm[1][1][2] = {1, 0}; gm[1][2][1] = {1, 0}; gm[1][3][0] = {-1, 0};
    gm[2][0][2] = {0, -1}; gm[2][1][3] = {0, 1}; gm[2][2][0] = {0, 1}; gm[2][3][1] = {0, -1};
    gm[3][0][2] = {-1, 0}; gm[3][1][3] = {-1, 0}; gm[3][2][0] = {-1, 0}; gm[3][3][1] = {-1, 0};
    const int pmu[6] = {0, 0, 0, 1, 1, 2}, pnu[6] = {1, 2, 3, 2, 3, 3};
    GammaProducts out;
    for (int pl = 0; pl < 6; ++pl)
        for (int i = 0; i < 4; ++i)
            for (int j = 0; j < 4; ++j) {
                double re = 0.0, im = 0.0;
                for (int k = 0; k < 4; ++k) {
                    const C2 a = gm[pmu[pl]][i][k], b = gm[pnu[pl]][k][j];
                    const double pr = a.re * b.re - a.im * b.im, pi = a.re * b.im + a.im * b.re;
                    re = re + pr;
                    im = im + pi;
                }
                out.g[pl][i][j] = make_double2(re, im);
            }
    return out;
}

int launch_clover_pullback(lmx_context* ctx, double2* dU, const double2* U, double2* dF, const double2* dres,
                           const double2* psi, double coefficient) {
    if (ctx->V == 0) return LMX_OK;
    static const GammaProducts gp = gamma_products();
    LMX_TRY(launch_zero(ctx, dF, 6 * 9 * (int64_t)ctx->V));
    hipLaunchKernelGGL(k_clover_cotangent, dim3(nblocks(ctx->V)), dim3(256), 0, ctx->stream, dF, dres, psi,
                       (int)ctx->V, coefficient, gp);
    hipLaunchKernelGGL(k_clover_links_pullback, dim3(nblocks(ctx->V)), dim3(256), 0, ctx->stream, dU, U,
                       (const double2*)dF, ctx->g);
    LMX_HIP(hipGetLastError());
    return LMX_OK;
}

}  // namespace lmx
