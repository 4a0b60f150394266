/*
 * lmx_oracle.c -- CPU oracle for the Wilson(-clover) D_W / D_W^dagger path and
 * the D^dagger D conjugate-gradient solve of LatticeMatrices.jl.
 *
 * TEST INFRASTRUCTURE ONLY.  Only tests/, __graft_entry__.smoke() and the
 * cpu_baseline leg of bench.py may load this library, and only as the checker
 * (or, for cpu_baseline, as the timed CPU restatement).  The product library
 * (liblmx.so) never links, loads or calls it.
 *
 * Every routine restates the reference's arithmetic in the reference's
 * operation order (file:line cited per function).  Conventions:
 *   - Complex numbers are pairs of doubles (re, im).
 *   - Julia's `muladd(a::Float64, b, c)` is taken to be one fused multiply-add
 *     (LLVM contracts it on every FMA-capable host); the complex `muladd` is
 *     Julia Base's decomposition
 *         re = muladd(re z, re w, -muladd(im z, im w, -re x))
 *         im = muladd(re z, im w,  muladd(im z, re w,  im x)).
 *   - Plain complex `*`, `+`, `-` are not contracted (Julia does not contract
 *     without muladd/@fastmath); this file is compiled with -ffp-contract=off.
 *   - Spinor layout: psi[(site*4 + spin)*3 + colour] complex, site index
 *     x-fastest (src/Latticeindices.jl:88-95).  This is the reference's
 *     (3,4,X,Y,Z,T) column-major array without halo.
 *   - Link layout: U[((mu*V + site)*3 + row)*3 + col] complex (row-major 3x3).
 *   - Boundary phases act on spinors only: a hop that wraps forward in
 *     direction mu multiplies the neighbour spinor by phase[mu], a backward
 *     wrap by inv(phase[mu]) (src/LatticeMatrices_core.jl:707-723,1066-1070).
 *     Applying the phase while reading the neighbour is bit-identical to the
 *     reference's halo fill (the same complex product of the same values).
 *
 * Parity pinning (tests/test_oracle_*.py):
 *   wilson       : dense gamma-matrix restatement of test/wilson_halfspin.jl:8-42
 *                  (numpy), free-field plane-wave closed form, adjoint identity,
 *                  gamma5-hermiticity, gauge covariance, tiny-lattice brute force.
 *   dot/axpby/cg : test/cg.jl known answers (scaled operator, zero rhs,
 *                  maxiter 0, indefinite operator, D^dagger D of (2+i)I),
 *                  S = eta^dagger eta for phi = D^dagger eta, true residual.
 *   clover       : dense restatement of test/wilson_clover.jl:47-112, csw=0
 *                  equals Wilson, unit gauge gives F=0, F anti-Hermitian,
 *                  gamma5-hermiticity, gauge covariance.
 *   pullback/TA  : central finite differences of S along su(3) directions,
 *                  linear-functional finite differences of <chi, D(U) psi>.
 *   clover pullback / clover dSFdU : exact central differences of the
 *                  functional Re <w, D_clover(U) psi> in single link entries
 *                  (real-affine in any one entry, F recomputed from U),
 *                  finite differences of S along su(3) directions, S = |eta|^2,
 *                  planted mistakes (cSW sign, dropped clover part).  The
 *                  SMatrix product order of the reference is not pinned: the
 *                  products are restated as plain sums left to right.
 */
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

typedef struct {
    double re, im;
} cx;

static inline cx C(double r, double i) {
    cx z;
    z.re = r;
    z.im = i;
    return z;
}
static inline cx cadd(cx a, cx b) { return C(a.re + b.re, a.im + b.im); }
static inline cx csub(cx a, cx b) { return C(a.re - b.re, a.im - b.im); }
static inline cx cneg(cx a) { return C(-a.re, -a.im); }
static inline cx cconj(cx a) { return C(a.re, -a.im); }
/* Julia Base complex product (no contraction). */
static inline cx cmul(cx a, cx b) {
    return C(a.re * b.re - a.im * b.im, a.re * b.im + a.im * b.re);
}
/* Float64 * Complex */
static inline cx rmul(double a, cx b) { return C(a * b.re, a * b.im); }
/* Julia Base complex muladd with fused real muladds. */
static inline cx cmuladd(cx z, cx w, cx x) {
    return C(fma(z.re, w.re, -fma(z.im, w.im, -x.re)), fma(z.re, w.im, fma(z.im, w.re, x.im)));
}
/* src/Operators/WilsonDiracOperator.jl:149-155 */
static inline cx muladdmulti(cx a1, cx b1, cx a2, cx b2, cx a3, cx b3) {
    cx acc = C(0.0, 0.0);
    acc = cmuladd(a1, b1, acc);
    acc = cmuladd(a2, b2, acc);
    acc = cmuladd(a3, b3, acc);
    return acc;
}
/* PM * im * z with PM = +-1 (exact). */
static inline cx pmi(int pm, cx z) { return C(-(double)pm * z.im, (double)pm * z.re); }
static inline cx pmr(int pm, cx z) { return C((double)pm * z.re, (double)pm * z.im); }

typedef struct {
    int L[4];
    long V;
    long stride[4];
    cx ph[4];
    cx iph[4];
} geom;

/* Complex reciprocal used for backward boundary phases (inv(phase)). */
static cx cinv(cx z) {
    double d = z.re * z.re + z.im * z.im;
    return C(z.re / d, -z.im / d);
}

static void geom_init(geom* g, const int* L, const double* phases) {
    g->V = 1;
    for (int d = 0; d < 4; ++d) {
        g->L[d] = L[d];
        g->stride[d] = g->V;
        g->V *= L[d];
        g->ph[d] = C(phases ? phases[2 * d] : 1.0, phases ? phases[2 * d + 1] : 0.0);
        g->iph[d] = cinv(g->ph[d]);
    }
}

static inline void site_coords(const geom* g, long s, int x[4]) {
    for (int d = 0; d < 4; ++d) {
        x[d] = (int)(s % g->L[d]);
        s /= g->L[d];
    }
}

/* neighbour site in direction mu, dir = +1/-1; *wrap = 1 when crossing the boundary */
static inline long nbr(const geom* g, long s, const int x[4], int mu, int dir, int* wrap) {
    if (dir > 0) {
        if (x[mu] == g->L[mu] - 1) {
            *wrap = 1;
            return s - (long)(g->L[mu] - 1) * g->stride[mu];
        }
        *wrap = 0;
        return s + g->stride[mu];
    }
    if (x[mu] == 0) {
        *wrap = 1;
        return s + (long)(g->L[mu] - 1) * g->stride[mu];
    }
    *wrap = 0;
    return s - g->stride[mu];
}

static inline void load_spinor(const double* psi, long s, cx out[12]) {
    const double* p = psi + s * 24;
    for (int k = 0; k < 12; ++k) out[k] = C(p[2 * k], p[2 * k + 1]);
}
/* psi(spin, colour) at k = spin*3 + colour */
#define SP(a, spin, col) (a)[(spin) * 3 + (col)]

static inline void load_link(const double* U, const geom* g, int mu, long s, cx u[9]) {
    const double* p = U + ((long)mu * g->V + s) * 18;
    for (int k = 0; k < 9; ++k) u[k] = C(p[2 * k], p[2 * k + 1]);
}

/* _wilson_half_project3, src/Operators/WilsonDiracOperator.jl:157-175 */
static inline void half_project(const cx psi[12], int col, int pm, int mu, cx* h1, cx* h2) {
    cx p1 = SP(psi, 0, col), p2 = SP(psi, 1, col), p3 = SP(psi, 2, col), p4 = SP(psi, 3, col);
    switch (mu) {
    case 0:
        *h1 = csub(p1, pmi(pm, p4));
        *h2 = csub(p2, pmi(pm, p3));
        break;
    case 1:
        *h1 = csub(p1, pmr(pm, p4));
        *h2 = cadd(p2, pmr(pm, p3));
        break;
    case 2:
        *h1 = csub(p1, pmi(pm, p3));
        *h2 = cadd(p2, pmi(pm, p4));
        break;
    default:
        *h1 = csub(p1, pmr(pm, p3));
        *h2 = csub(p2, pmr(pm, p4));
        break;
    }
}

/* _wilson_reconstruct_add3, src/Operators/WilsonDiracOperator.jl:231-263.
 * acc is indexed [colour*4 + spin] (the reference accumulator order c11,c12,...). */
static inline void reconstruct_add(cx acc[12], const cx h[6], int pm, int mu) {
    for (int c = 0; c < 3; ++c) {
        cx h1 = h[2 * c], h2 = h[2 * c + 1];
        cx* a = acc + 4 * c;
        a[0] = cadd(a[0], h1);
        a[1] = cadd(a[1], h2);
        switch (mu) {
        case 0:
            a[2] = cadd(a[2], pmi(pm, h2));
            a[3] = cadd(a[3], pmi(pm, h1));
            break;
        case 1:
            a[2] = cadd(a[2], pmr(pm, h2));
            a[3] = csub(a[3], pmr(pm, h1));
            break;
        case 2:
            a[2] = cadd(a[2], pmi(pm, h1));
            a[3] = csub(a[3], pmi(pm, h2));
            break;
        default:
            a[2] = csub(a[2], pmr(pm, h1));
            a[3] = csub(a[3], pmr(pm, h2));
            break;
        }
    }
}

static inline void scale_spinor(cx psi[12], cx ph) {
    for (int k = 0; k < 12; ++k) psi[k] = cmul(psi[k], ph);
}

/* _wilson_hopping_accumulator3, src/Operators/WilsonDiracOperator.jl:265-320 */
static void hopping_site(const geom* g, const double* U, const double* psi, long s, int fpm, cx acc[12]) {
    int x[4];
    site_coords(g, s, x);
    for (int k = 0; k < 12; ++k) acc[k] = C(0.0, 0.0);
    for (int mu = 0; mu < 4; ++mu) {
        cx sp[12], u[9], h[6], r[6];
        int wrap;
        /* forward: U_mu(x) (1 + fpm gamma_mu... ) psi(x+mu)  (_wilson_half_matvec_forward3 :177-202) */
        long sp_site = nbr(g, s, x, mu, +1, &wrap);
        load_spinor(psi, sp_site, sp);
        if (wrap) scale_spinor(sp, g->ph[mu]);
        for (int c = 0; c < 3; ++c) half_project(sp, c, fpm, mu, &h[2 * c], &h[2 * c + 1]);
        load_link(U, g, mu, s, u);
        for (int a = 0; a < 3; ++a) {
            r[2 * a] = muladdmulti(u[3 * a], h[0], u[3 * a + 1], h[2], u[3 * a + 2], h[4]);
            r[2 * a + 1] = muladdmulti(u[3 * a], h[1], u[3 * a + 1], h[3], u[3 * a + 2], h[5]);
        }
        reconstruct_add(acc, r, fpm, mu);
        /* backward: U_mu(x-mu)^dagger psi(x-mu)  (_wilson_half_matvec_backward3 :204-229) */
        long sm_site = nbr(g, s, x, mu, -1, &wrap);
        load_spinor(psi, sm_site, sp);
        if (wrap) scale_spinor(sp, g->iph[mu]);
        for (int c = 0; c < 3; ++c) half_project(sp, c, -fpm, mu, &h[2 * c], &h[2 * c + 1]);
        load_link(U, g, mu, sm_site, u);
        for (int a = 0; a < 3; ++a) {
            cx w0 = cconj(u[a]), w1 = cconj(u[3 + a]), w2 = cconj(u[6 + a]);
            r[2 * a] = muladdmulti(w0, h[0], w1, h[2], w2, h[4]);
            r[2 * a + 1] = muladdmulti(w0, h[1], w1, h[3], w2, h[5]);
        }
        reconstruct_add(acc, r, -fpm, mu);
    }
}

/*
 * D_W, D_W^dagger and the hopping-only variants.
 *   kernel_WilsonDiracOperator4D! NC=3            src/Operators/WilsonDiracOperator.jl:413-419
 *   kernel_adjoint_WilsonDiracOperator4D! NC=3    :514-520
 *   kernel_WilsonDiracOperator4D_Donly! NC=3      :624-631 (and adjoint :736-743)
 *   _write_wilson_result3!                        :322-352
 */
void orc_wilson(double* out, const double* in, const double* U, const int* L, const double* phases,
                double kappa, int dagger, int donly) {
    geom g;
    geom_init(&g, L, phases);
    const int fpm = dagger ? 1 : -1;
    const double coef = donly ? 0.5 : -kappa;
#pragma omp parallel for schedule(static)
    for (long s = 0; s < g.V; ++s) {
        cx acc[12];
        hopping_site(&g, U, in, s, fpm, acc);
        double* o = out + s * 24;
        const double* p = in + s * 24;
        for (int spin = 0; spin < 4; ++spin)
            for (int c = 0; c < 3; ++c) {
                cx h = acc[4 * c + spin];
                cx base = donly ? C(0.0, 0.0) : C(p[2 * (spin * 3 + c)], p[2 * (spin * 3 + c) + 1]);
                cx v = cadd(base, rmul(coef, h));
                o[2 * (spin * 3 + c)] = v.re;
                o[2 * (spin * 3 + c) + 1] = v.im;
            }
    }
}

/* ------------------------------------------------------------------------ */
/* BLAS-1 and reductions                                                    */
/* ------------------------------------------------------------------------ */

/* kernel_dot_D per-site value, src/LinearAlgebras/linearalgebra_D.jl:3331-3341:
 * s = 0; for spin, for colour: s += conj(A)*B */
static inline cx site_dot(const double* a, const double* b, long s, int ncomp) {
    cx acc = C(0.0, 0.0);
    const double* pa = a + s * 2 * ncomp;
    const double* pb = b + s * 2 * ncomp;
    for (int k = 0; k < ncomp; ++k)
        acc = cadd(acc, cmul(cconj(C(pa[2 * k], pa[2 * k + 1])), C(pb[2 * k], pb[2 * k + 1])));
    return acc;
}

/* The fixed reduction tree of the HIP library: blocks of 256 sites, each wave
 * of 64 halves its values (offsets 32..1), the 4 wave sums halve again; the
 * block partials are summed by 256 strided sequential lanes which are then
 * reduced by the same block tree.  Restated here so that the oracle's CG can
 * reproduce the device's rounding exactly (order 1); order 0 is the plain
 * sequential sum of a single-threaded reference run. */
#define RBLK 256
static void tree_halve(double* w, int n) {
    for (int off = n / 2; off >= 1; off /= 2)
        for (int i = 0; i < off; ++i) w[i] = w[i] + w[i + off];
}
static double block_tree(const double* v /*256*/) {
    double waves[4];
    double w[64];
    for (int q = 0; q < 4; ++q) {
        memcpy(w, v + 64 * q, sizeof(w));
        tree_halve(w, 64);
        waves[q] = w[0];
    }
    tree_halve(waves, 4);
    return waves[0];
}
static double tree_finish(const double* partials, long nb) {
    double lanes[RBLK];
    for (int t = 0; t < RBLK; ++t) {
        double acc = 0.0;
        for (long b = t; b < nb; b += RBLK) acc = acc + partials[b];
        lanes[t] = acc;
    }
    return block_tree(lanes);
}

typedef cx (*site_fn)(const void* ctx, long s);

static cx reduce_sites(long V, int order, site_fn f, const void* ctx) {
    if (order == 0) {
        cx acc = C(0.0, 0.0);
        for (long s = 0; s < V; ++s) acc = cadd(acc, f(ctx, s));
        return acc;
    }
    long nb = (V + RBLK - 1) / RBLK;
    double* pre = (double*)malloc(sizeof(double) * nb);
    double* pim = (double*)malloc(sizeof(double) * nb);
#pragma omp parallel for schedule(static)
    for (long b = 0; b < nb; ++b) {
        double vr[RBLK], vi[RBLK];
        for (int t = 0; t < RBLK; ++t) {
            long s = b * RBLK + t;
            cx z = s < V ? f(ctx, s) : C(0.0, 0.0);
            vr[t] = z.re;
            vi[t] = z.im;
        }
        pre[b] = block_tree(vr);
        pim[b] = block_tree(vi);
    }
    cx out = C(tree_finish(pre, nb), tree_finish(pim, nb));
    free(pre);
    free(pim);
    return out;
}

typedef struct {
    const double *a, *b;
    int ncomp;
} dot_ctx;
static cx dot_site_fn(const void* c, long s) {
    const dot_ctx* d = (const dot_ctx*)c;
    return site_dot(d->a, d->b, s, d->ncomp);
}

/* LinearAlgebra.dot, src/LinearAlgebras/linearalgebra_D.jl:3325-3329 (one
 * rank), over sites of ncomp = 4 NC complex components each */
void orc_dot_n(const double* a, const double* b, long V, int ncomp, int order, double* out2) {
    dot_ctx d = {a, b, ncomp};
    cx z = reduce_sites(V, order, dot_site_fn, &d);
    out2[0] = z.re;
    out2[1] = z.im;
}
void orc_dot(const double* a, const double* b, long V, int order, double* out2) { orc_dot_n(a, b, V, 12, order, out2); }

/* axpby!(a, X, b, Y): Y = a*X + b*Y, src/LinearAlgebras/linearalgebra_D.jl:2-23.
 * a and b are real here (CG only passes reals). */
static void axpby_n(double a, const double* X, double b, double* Y, long n) {
#pragma omp parallel for schedule(static)
    for (long i = 0; i < n; ++i) Y[i] = a * X[i] + b * Y[i];
}
void orc_axpby(double a, const double* X, double b, double* Y, long V) { axpby_n(a, X, b, Y, V * 24); }

/* ------------------------------------------------------------------------ */
/* Clover                                                                   */
/* ------------------------------------------------------------------------ */

static const int plane_mu[6] = {0, 0, 0, 1, 1, 2};
static const int plane_nu[6] = {1, 2, 3, 2, 3, 3};

/* _clover_mul_matrix3, src/Operators/WilsonDiracCloverOperator.jl:137-149 (row-major 9-tuples) */
static void mat_mul(const cx* A, const cx* B, cx* out) {
    cx r[9];
    for (int i = 0; i < 3; ++i)
        for (int j = 0; j < 3; ++j)
            r[3 * i + j] = muladdmulti(A[3 * i], B[j], A[3 * i + 1], B[3 + j], A[3 * i + 2], B[6 + j]);
    memcpy(out, r, sizeof(r));
}
static void mat_dag(const cx* A, cx* out) {
    cx r[9];
    for (int i = 0; i < 3; ++i)
        for (int j = 0; j < 3; ++j) r[3 * i + j] = cconj(A[3 * j + i]);
    memcpy(out, r, sizeof(r));
}
/* _clover_path_matrix3 :159-162 = ((A*B)*C)*D */
static void path4(const cx* A, const cx* B, const cx* Cm, const cx* D, cx* out) {
    cx t[9];
    mat_mul(A, B, t);
    mat_mul(t, Cm, t);
    mat_mul(t, D, out);
}

static long shift_site(const geom* g, long s, int mu, int dir) {
    int x[4], w;
    site_coords(g, s, x);
    return nbr(g, s, x, mu, dir, &w);
}

/* kernel_clover_field_strength_halo! NC=3 + _clover_store_antihermitian3!,
 * src/Operators/WilsonDiracCloverOperator.jl:164-226; planes (12,13,14,23,24,34) :1.
 * Links are periodic (their phases are one).  F layout: F[((plane*V + site)*3+row)*3+col]. */
void orc_clover_field_strength(double* F, const double* U, const int* L) {
    geom g;
    geom_init(&g, L, NULL);
#pragma omp parallel for schedule(static)
    for (long s = 0; s < g.V; ++s) {
        for (int pl = 0; pl < 6; ++pl) {
            int mu = plane_mu[pl], nu = plane_nu[pl];
            long xpmu = shift_site(&g, s, mu, +1), xpnu = shift_site(&g, s, nu, +1);
            long xmmu = shift_site(&g, s, mu, -1), xmnu = shift_site(&g, s, nu, -1);
            long xpnu_mmu = shift_site(&g, xpnu, mu, -1);
            long xmmu_mnu = shift_site(&g, xmmu, nu, -1);
            long xpmu_mnu = shift_site(&g, xmnu, mu, +1);
            cx a[9], b[9], c[9], d[9], q[9], t[9];
            /* +mu, +nu, -mu, -nu */
            load_link(U, &g, mu, s, a);
            load_link(U, &g, nu, xpmu, b);
            load_link(U, &g, mu, xpnu, c);
            mat_dag(c, c);
            load_link(U, &g, nu, s, d);
            mat_dag(d, d);
            path4(a, b, c, d, q);
            /* +nu, -mu, -nu, +mu */
            load_link(U, &g, nu, s, a);
            load_link(U, &g, mu, xpnu_mmu, b);
            mat_dag(b, b);
            load_link(U, &g, nu, xmmu, c);
            mat_dag(c, c);
            load_link(U, &g, mu, xmmu, d);
            path4(a, b, c, d, t);
            for (int k = 0; k < 9; ++k) q[k] = cadd(q[k], t[k]);
            /* -mu, -nu, +mu, +nu */
            load_link(U, &g, mu, xmmu, a);
            mat_dag(a, a);
            load_link(U, &g, nu, xmmu_mnu, b);
            mat_dag(b, b);
            load_link(U, &g, mu, xmmu_mnu, c);
            load_link(U, &g, nu, xmnu, d);
            path4(a, b, c, d, t);
            for (int k = 0; k < 9; ++k) q[k] = cadd(q[k], t[k]);
            /* -nu, +mu, +nu, -mu */
            load_link(U, &g, nu, xmnu, a);
            mat_dag(a, a);
            load_link(U, &g, mu, xmnu, b);
            load_link(U, &g, nu, xpmu_mnu, c);
            load_link(U, &g, mu, s, d);
            mat_dag(d, d);
            path4(a, b, c, d, t);
            for (int k = 0; k < 9; ++k) q[k] = cadd(q[k], t[k]);
            /* store (q - q^dagger)/8 */
            const double f = 1.0 / 8.0;
            cx v11 = rmul(f, csub(q[0], cconj(q[0])));
            cx v12 = rmul(f, csub(q[1], cconj(q[3])));
            cx v13 = rmul(f, csub(q[2], cconj(q[6])));
            cx v22 = rmul(f, csub(q[4], cconj(q[4])));
            cx v23 = rmul(f, csub(q[5], cconj(q[7])));
            cx v33 = rmul(f, csub(q[8], cconj(q[8])));
            cx out[9] = {v11, v12, v13, cneg(cconj(v12)), v22, v23, cneg(cconj(v13)), cneg(cconj(v23)), v33};
            double* o = F + ((long)pl * g.V + s) * 18;
            for (int k = 0; k < 9; ++k) {
                o[2 * k] = out[k].re;
                o[2 * k + 1] = out[k].im;
            }
        }
    }
}

static inline cx iunit(cx z) { return C(-z.im, z.re); } /* (0,1) * z */

/* kernel_add_clover_term!, src/Operators/WilsonDiracCloverOperator.jl:481-525 */
static void clover_site_add(const geom* g, const double* F, const double* psi, long s, double coefficient,
                            cx res[12] /* [spin*3+colour] */) {
    cx p[12];
    load_spinor(psi, s, p);
    for (int color = 0; color < 3; ++color) {
        cx v1 = C(0, 0), v2 = C(0, 0), v3 = C(0, 0), v4 = C(0, 0);
        for (int ic = 0; ic < 3; ++ic) {
            cx f[6];
            for (int pl = 0; pl < 6; ++pl) {
                const double* q = F + (((long)pl * g->V + s) * 9 + color * 3 + ic) * 2;
                f[pl] = C(q[0], q[1]);
            }
            cx f12 = f[0], f13 = f[1], f14 = f[2], f23 = f[3], f24 = f[4], f34 = f[5];
            cx dp = iunit(cadd(f12, f34));
            cx dm = iunit(csub(f12, f34));
            cx odp = iunit(cadd(f14, f23));
            cx odm = iunit(csub(f23, f14));
            cx p1 = SP(p, 0, ic), p2 = SP(p, 1, ic), p3 = SP(p, 2, ic), p4 = SP(p, 3, ic);
            v1 = cadd(v1, cadd(cmul(dp, p1), cmul(cadd(cadd(cneg(f13), odp), f24), p2)));
            v2 = cadd(v2, cadd(cmul(cneg(dp), p2), cmul(csub(cadd(f13, odp), f24), p1)));
            v3 = cadd(v3, cadd(cmul(dm, p3), cmul(csub(cadd(cneg(f13), odm), f24), p4)));
            v4 = cadd(v4, cadd(cmul(cneg(dm), p4), cmul(cadd(cadd(f13, odm), f24), p3)));
        }
        SP(res, 0, color) = cadd(SP(res, 0, color), rmul(coefficient, v1));
        SP(res, 1, color) = cadd(SP(res, 1, color), rmul(coefficient, v2));
        SP(res, 2, color) = cadd(SP(res, 2, color), rmul(coefficient, v3));
        SP(res, 3, color) = cadd(SP(res, 3, color), rmul(coefficient, v4));
    }
}

/* Fused Wilson + clover, kernel_WilsonDiracCloverOperator4D! (:537-546) and its adjoint (:548-559):
 * the Wilson result C = psi - kappa*hop is formed first, then the clover term is added with
 * coefficient = -kappa * cSW. */
void orc_wilson_clover(double* out, const double* in, const double* U, const double* F, const int* L,
                       const double* phases, double kappa, double csw, int dagger) {
    geom g;
    geom_init(&g, L, phases);
    const int fpm = dagger ? 1 : -1;
    const double coefficient = (-kappa) * csw;
#pragma omp parallel for schedule(static)
    for (long s = 0; s < g.V; ++s) {
        cx acc[12], res[12];
        hopping_site(&g, U, in, s, fpm, acc);
        const double* p = in + s * 24;
        for (int spin = 0; spin < 4; ++spin)
            for (int c = 0; c < 3; ++c) {
                cx base = C(p[2 * (spin * 3 + c)], p[2 * (spin * 3 + c) + 1]);
                SP(res, spin, c) = cadd(base, rmul(-kappa, acc[4 * c + spin]));
            }
        clover_site_add(&g, F, in, s, coefficient, res);
        double* o = out + s * 24;
        for (int k = 0; k < 12; ++k) {
            o[2 * k] = res[k].re;
            o[2 * k + 1] = res[k].im;
        }
    }
}

/* ------------------------------------------------------------------------ */
/* CG                                                                       */
/* ------------------------------------------------------------------------ */

enum { ORC_CONVERGED = 0, ORC_MAXITER = 1, ORC_NONPOSITIVE_CURVATURE = 2, ORC_BREAKDOWN = 3 };

enum { ORC_OP_DDAGD_WILSON = 0, ORC_OP_DDAGD_CLOVER = 1, ORC_OP_SCALE = 2, ORC_OP_DDAGD_SCALE = 3 };

typedef struct {
    int kind;
    const double* U;
    const double* F;
    int L[4];
    double phases[8];
    double kappa, csw;
    double scale_re, scale_im;
    double* tmp; /* D x temporary (DdagDOp temp) */
    int nc;      /* colours (0 means 3); NC != 3 runs the generic kernels of lmx_nc_oracle.c */
} orc_op;

void orc_wilson_nc(double* out, const double* in, const double* U, const int* L, const double* phases, double kappa,
                   int nc, int dagger, int donly);
void orc_wilson_clover_nc(double* out, const double* in, const double* U, const double* F, const int* L,
                          const double* phases, double kappa, double csw, int nc, int dagger);

static int op_nc(const orc_op* op) { return op->nc > 0 ? op->nc : 3; }

static void scale_field(double* out, const double* in, cx a, long n) {
#pragma omp parallel for schedule(static)
    for (long i = 0; i < n; ++i) {
        cx v = cmul(a, C(in[2 * i], in[2 * i + 1]));
        out[2 * i] = v.re;
        out[2 * i + 1] = v.im;
    }
}

/* DdagDOp.mul!: temp = D x; y = D' temp (src/Operators/DiracOperators.jl:104-121) */
static void op_apply(const orc_op* op, double* y, const double* x, long V) {
    const int nc = op_nc(op);
    const long n = V * 4 * nc;
    switch (op->kind) {
    case ORC_OP_DDAGD_WILSON:
        if (nc == 3) {
            orc_wilson(op->tmp, x, op->U, op->L, op->phases, op->kappa, 0, 0);
            orc_wilson(y, op->tmp, op->U, op->L, op->phases, op->kappa, 1, 0);
        } else {
            orc_wilson_nc(op->tmp, x, op->U, op->L, op->phases, op->kappa, nc, 0, 0);
            orc_wilson_nc(y, op->tmp, op->U, op->L, op->phases, op->kappa, nc, 1, 0);
        }
        break;
    case ORC_OP_DDAGD_CLOVER:
        if (nc == 3) {
            orc_wilson_clover(op->tmp, x, op->U, op->F, op->L, op->phases, op->kappa, op->csw, 0);
            orc_wilson_clover(y, op->tmp, op->U, op->F, op->L, op->phases, op->kappa, op->csw, 1);
        } else {
            orc_wilson_clover_nc(op->tmp, x, op->U, op->F, op->L, op->phases, op->kappa, op->csw, nc, 0);
            orc_wilson_clover_nc(y, op->tmp, op->U, op->F, op->L, op->phases, op->kappa, op->csw, nc, 1);
        }
        break;
    case ORC_OP_SCALE: /* mul!(y, scale, x), src/LinearAlgebras/mul.jl:2-9 */
        scale_field(y, x, C(op->scale_re, op->scale_im), n);
        break;
    default: /* D = scale, D' = conj(scale) */
        scale_field(op->tmp, x, C(op->scale_re, op->scale_im), n);
        scale_field(y, op->tmp, C(op->scale_re, -op->scale_im), n);
        break;
    }
}

typedef struct {
    int converged;
    int iterations;
    double residual_norm;
    double relative_residual;
    int reason;
} orc_cg_result;

static double rel_res(double rn, double bn) {
    if (bn == 0.0) return rn == 0.0 ? 0.0 : INFINITY;
    return rn / bn;
}
static void set_result(orc_cg_result* r, int conv, int it, double rn, double bn, int reason) {
    r->converged = conv;
    r->iterations = it;
    r->residual_norm = rn;
    r->relative_residual = rel_res(rn, bn);
    r->reason = reason;
}

/* cg!, src/LinearAlgebras/cg.jl:73-137, in its operation order. */
int orc_cg(double* x, const double* b, double* r, double* p, double* Ap, const orc_op* op, long V, double rtol,
           double atol, int maxiter, int dot_order, orc_cg_result* res) {
    const int ncomp = 4 * op_nc(op);
    const long n = V * 2 * ncomp; /* doubles per field */
    double d2[2];
    orc_dot_n(b, b, V, ncomp, dot_order, d2);
    double bb = d2[0];
    if (!isfinite(bb) || bb < 0.0) {
        set_result(res, 0, 0, NAN, NAN, ORC_BREAKDOWN);
        return 0;
    }
    double bnorm = sqrt(bb);
    double tol = fmax(atol, rtol * bnorm);
    op_apply(op, Ap, x, V);
    memcpy(r, b, sizeof(double) * n);     /* axpby!(1, b, 0, r) */
    axpby_n(-1.0, Ap, 1.0, r, n);           /* axpby!(-1, Ap, 1, r) */
    memcpy(p, r, sizeof(double) * n);     /* axpby!(1, r, 0, p) */
    orc_dot_n(r, r, V, ncomp, dot_order, d2);
    double rr = d2[0];
    if (!isfinite(rr) || rr < 0.0) {
        set_result(res, 0, 0, NAN, bnorm, ORC_BREAKDOWN);
        return 0;
    }
    double rn = sqrt(rr);
    if (rn <= tol) {
        set_result(res, 1, 0, rn, bnorm, ORC_CONVERGED);
        return 0;
    }
    for (int it = 1; it <= maxiter; ++it) {
        op_apply(op, Ap, p, V);
        orc_dot_n(p, Ap, V, ncomp, dot_order, d2);
        double pAp = d2[0];
        if (!isfinite(pAp)) {
            set_result(res, 0, it - 1, rn, bnorm, ORC_BREAKDOWN);
            return 0;
        } else if (pAp <= 0.0) {
            set_result(res, 0, it - 1, rn, bnorm, ORC_NONPOSITIVE_CURVATURE);
            return 0;
        }
        double alpha = rr / pAp;
        axpby_n(alpha, p, 1.0, x, n);
        axpby_n(-alpha, Ap, 1.0, r, n);
        orc_dot_n(r, r, V, ncomp, dot_order, d2);
        double rr_new = d2[0];
        if (!isfinite(rr_new) || rr_new < 0.0) {
            set_result(res, 0, it, NAN, bnorm, ORC_BREAKDOWN);
            return 0;
        }
        rn = sqrt(rr_new);
        if (rn <= tol) {
            set_result(res, 1, it, rn, bnorm, ORC_CONVERGED);
            return 0;
        }
        double beta = rr_new / rr;
        axpby_n(1.0, r, beta, p, n);
        rr = rr_new;
    }
    set_result(res, 0, maxiter, rn, bnorm, ORC_MAXITER);
    return 0;
}

/* Apply the normal operator once (y = A x), for true-residual checks. */
void orc_op_apply(const orc_op* op, double* y, const double* x, long V) { op_apply(op, y, x, V); }

/* ------------------------------------------------------------------------ */
/* Force: link pullback and traceless anti-Hermitian projection             */
/* ------------------------------------------------------------------------ */

/* _wilson_halfspin_link_pullback_direction3! / _row3!,
 * src/Operators/WilsonDiracOperator.jl:354-410, driven per site by
 * _kernel_wilson_link_pullback! (ext/AD/WilsonDiracOperator.jl:62-91).
 * dU[mu] += coefficient * value; dresult and psi are spinors (their x+mu
 * values carry the spinor boundary phase, as after set_halo!). */
void orc_wilson_pullback(double* dU, const double* dres, const double* psi, const int* L, const double* phases,
                         double coefficient) {
    geom g;
    geom_init(&g, L, phases);
#pragma omp parallel for schedule(static)
    for (long s = 0; s < g.V; ++s) {
        int x[4];
        site_coords(&g, s, x);
        cx d0[12], p0[12];
        load_spinor(dres, s, d0);
        load_spinor(psi, s, p0);
        for (int mu = 0; mu < 4; ++mu) {
            int wrap;
            long sp = nbr(&g, s, x, mu, +1, &wrap);
            cx d1[12], p1[12];
            load_spinor(dres, sp, d1);
            load_spinor(psi, sp, p1);
            if (wrap) {
                scale_spinor(d1, g.ph[mu]);
                scale_spinor(p1, g.ph[mu]);
            }
            cx plus_psi[6], minus_dres[6];
            for (int c = 0; c < 3; ++c) {
                half_project(p1, c, -1, mu, &plus_psi[2 * c], &plus_psi[2 * c + 1]);
                half_project(d1, c, +1, mu, &minus_dres[2 * c], &minus_dres[2 * c + 1]);
            }
            for (int row = 0; row < 3; ++row) {
                cx dplus1, dplus2, mpsi1, mpsi2;
                half_project(d0, row, -1, mu, &dplus1, &dplus2);
                half_project(p0, row, +1, mu, &mpsi1, &mpsi2);
                for (int col = 0; col < 3; ++col) {
                    cx v = cmul(dplus1, cconj(plus_psi[2 * col]));
                    v = cadd(v, cmul(dplus2, cconj(plus_psi[2 * col + 1])));
                    v = cadd(v, cmul(mpsi1, cconj(minus_dres[2 * col])));
                    v = cadd(v, cmul(mpsi2, cconj(minus_dres[2 * col + 1])));
                    double* o = dU + (((long)mu * g.V + s) * 9 + row * 3 + col) * 2;
                    cx nv = cadd(C(o[0], o[1]), rmul(coefficient, v));
                    o[0] = nv.re;
                    o[1] = nv.im;
                }
            }
        }
    }
}

/* M = U G^dagger per link (row i of U times column j of G^dagger, as one
 * muladd chain), then kernel_traceless_antihermitian! NC=3 (src/TA/TA.jl:68-116). */
void orc_ta_force(double* Fout, const double* U, const double* G, long nlinks) {
#pragma omp parallel for schedule(static)
    for (long l = 0; l < nlinks; ++l) {
        cx u[9], gg[9], m[9];
        for (int k = 0; k < 9; ++k) {
            u[k] = C(U[l * 18 + 2 * k], U[l * 18 + 2 * k + 1]);
            gg[k] = C(G[l * 18 + 2 * k], G[l * 18 + 2 * k + 1]);
        }
        for (int i = 0; i < 3; ++i)
            for (int j = 0; j < 3; ++j)
                m[3 * i + j] = muladdmulti(u[3 * i], cconj(gg[3 * j]), u[3 * i + 1], cconj(gg[3 * j + 1]), u[3 * i + 2],
                                           cconj(gg[3 * j + 2]));
        const double fac13 = 1.0 / 3.0;
        double tri = fac13 * ((m[0].im + m[4].im) + m[8].im);
        cx y11 = C(0.0, m[0].im - tri), y22 = C(0.0, m[4].im - tri), y33 = C(0.0, m[8].im - tri);
        cx x12 = csub(m[1], cconj(m[3]));
        cx x13 = csub(m[2], cconj(m[6]));
        cx x23 = csub(m[5], cconj(m[7]));
        cx x21 = cneg(cconj(x12)), x31 = cneg(cconj(x13)), x32 = cneg(cconj(x23));
        cx o[9] = {y11, rmul(0.5, x12), rmul(0.5, x13), rmul(0.5, x21), y22, rmul(0.5, x23),
                   rmul(0.5, x31), rmul(0.5, x32), y33};
        for (int k = 0; k < 9; ++k) {
            Fout[l * 18 + 2 * k] = o[k].re;
            Fout[l * 18 + 2 * k + 1] = o[k].im;
        }
    }
}

/* dSFdU for the Wilson operator (ext/LatticeMatricesEnzymeExt.jl:387-460):
 * eta = (D^dagger D)^-1 phi (zero start), chi = D eta, and the Enzyme reverse
 * rule of mul!(., D, eta) seeded with dresult = -2 chi (the pullback of
 * g = -2 Re dot(chi, D eta), ext/AD/AD.jl:4014-4032) with coefficient -kappa.
 * G is overwritten; eta/chi are outputs; work needs 4 spinor fields. */
int orc_dsfdu(double* G, double* eta, double* chi, const double* U, const double* phi, const int* L,
              const double* phases, double kappa, double rtol, int maxiter, int dot_order, double* work,
              orc_cg_result* res) {
    geom g;
    geom_init(&g, L, phases);
    long V = g.V;
    double *r = work, *p = work + 24 * V, *Ap = work + 48 * V, *tmp = work + 72 * V;
    orc_op op;
    memset(&op, 0, sizeof(op));
    op.kind = ORC_OP_DDAGD_WILSON;
    op.U = U;
    memcpy(op.L, L, sizeof(op.L));
    for (int k = 0; k < 8; ++k) op.phases[k] = phases ? phases[k] : (k % 2 ? 0.0 : 1.0);
    op.kappa = kappa;
    op.tmp = tmp;
    memset(eta, 0, sizeof(double) * 24 * V);
    orc_cg(eta, phi, r, p, Ap, &op, V, rtol, 0.0, maxiter, dot_order, res);
    orc_wilson(chi, eta, U, L, phases, kappa, 0, 0);
    double* dres = tmp;
    for (long i = 0; i < 24 * V; ++i) dres[i] = -2.0 * chi[i];
    memset(G, 0, sizeof(double) * 4 * V * 18);
    orc_wilson_pullback(G, dres, eta, L, phases, -kappa);
    return 0;
}

/* ------------------------------------------------------------------------ */
/* Wilson-clover link pullback and force                                    */
/* ------------------------------------------------------------------------ */

/* gamma_mu gamma_nu per plane (clover_gamma_products,
 * src/Operators/WilsonDiracCloverOperator.jl:10-13) from the chiral gamma
 * matrices of src/Operators/Operators.jl:44-67; entries 0, +-1, +-i. */
static cx gamma_products[6][4][4];
static int gamma_ready = 0;
static void init_gamma_products(void) {
    if (gamma_ready) return;
    cx gm[4][4][4];
    memset(gm, 0, sizeof(gm));
    /* gamma1 */
    gm[0][0][3] = C(0, -1); gm[0][1][2] = C(0, -1); gm[0][2][1] = C(0, 1); gm[0][3][0] = C(0, 1);
    /* gamma2 */
    gm[1][0][3] = C(-1, 0); gm[1][1][2] = C(1, 0); gm[1][2][1] = C(1, 0); gm[1][3][0] = C(-1, 0);
    /* gamma3 */
    gm[2][0][2] = C(0, -1); gm[2][1][3] = C(0, 1); gm[2][2][0] = C(0, 1); gm[2][3][1] = C(0, -1);
    /* gamma4 */
    gm[3][0][2] = C(-1, 0); gm[3][1][3] = C(-1, 0); gm[3][2][0] = C(-1, 0); gm[3][3][1] = C(-1, 0);
    for (int pl = 0; pl < 6; ++pl)
        for (int i = 0; i < 4; ++i)
            for (int j = 0; j < 4; ++j) {
                cx v = C(0, 0);
                for (int k = 0; k < 4; ++k) v = cadd(v, cmul(gm[plane_mu[pl]][i][k], gm[plane_nu[pl]][k][j]));
                gamma_products[pl][i][j] = v;
            }
    gamma_ready = 1;
}

/* _kernel_clover_field_cotangent_plane!, src/Operators/WilsonCloverPullback.jl:102-119:
 * dF[plane][c][ic] += conj(coefficient) * sum_out dres[c,out] conj(sum_in g[out,in] psi[ic,in])
 * (coefficient real).  dF layout as F: dF[((plane*V + site)*3+row)*3+col] complex. */
static void clover_field_cotangent(double* dF, const double* dres, const double* psi, long V, double coefficient) {
#pragma omp parallel for schedule(static)
    for (long s = 0; s < V; ++s) {
        cx d[12], p[12];
        load_spinor(dres, s, d);
        load_spinor(psi, s, p);
        for (int pl = 0; pl < 6; ++pl)
            for (int ic = 0; ic < 3; ++ic)
                for (int c = 0; c < 3; ++c) {
                    cx value = C(0, 0);
                    for (int o = 0; o < 4; ++o) {
                        cx tp = C(0, 0);
                        for (int i = 0; i < 4; ++i) tp = cadd(tp, cmul(gamma_products[pl][o][i], p[i * 3 + ic]));
                        value = cadd(value, cmul(d[o * 3 + c], cconj(tp)));
                    }
                    double* e = dF + (((long)pl * V + s) * 9 + c * 3 + ic) * 2;
                    cx nv = cadd(C(e[0], e[1]), rmul(coefficient, value));
                    e[0] = nv.re;
                    e[1] = nv.im;
                }
    }
}

/* 3x3 product with plain complex arithmetic, sums left to right (the
 * StaticArrays SMatrix products of WilsonCloverPullback.jl; their exact
 * evaluation order is not pinned here). */
static void mat_mul_plain(const cx* A, const cx* B, cx* out) {
    cx r[9];
    for (int i = 0; i < 3; ++i)
        for (int j = 0; j < 3; ++j)
            r[3 * i + j] = cadd(cadd(cmul(A[3 * i], B[j]), cmul(A[3 * i + 1], B[3 + j])), cmul(A[3 * i + 2], B[6 + j]));
    memcpy(out, r, sizeof(r));
}

/* Path descriptors of the four leaves of plane (mu, nu)
 * (_clover_path_descriptor, WilsonCloverPullback.jl:140-159): per factor the
 * origin relative to the anchor, the direction, and whether it is adjointed. */
typedef struct {
    int origin[4][4];
    int dir[4];
    int dag[4];
} leaf_path;
static void make_path(const int* signed_dirs, leaf_path* lp) {
    int pos[4] = {0, 0, 0, 0};
    for (int k = 0; k < 4; ++k) {
        int sd = signed_dirs[k], d = (sd > 0 ? sd : -sd) - 1;
        if (sd < 0) pos[d] -= 1;
        memcpy(lp->origin[k], pos, sizeof(pos));
        lp->dir[k] = d;
        lp->dag[k] = sd < 0;
        if (sd > 0) pos[d] += 1;
    }
}
static long offset_site(const geom* g, long s, const int* off) {
    for (int d = 0; d < 4; ++d) {
        int n = off[d];
        while (n > 0) { s = shift_site(g, s, d, +1); --n; }
        while (n < 0) { s = shift_site(g, s, d, -1); ++n; }
    }
    return s;
}

/* _clover_path_occurrence_pullback (WilsonCloverPullback.jl:185-211) for
 * occurrence K of the link at y: dQ = (dF - dF^dagger)/8 at the anchor;
 * h = left^dagger dQ right^dagger, or right dQ^dagger left for an adjointed
 * factor (left = M1..M_{K-1}, right = M_{K+1}..M4). */
static void occurrence_pullback(const geom* g, const double* U, const double* dFp, long y, const leaf_path* lp, int K,
                                cx h[9]) {
    int neg[4];
    for (int d = 0; d < 4; ++d) neg[d] = -lp->origin[K][d];
    long anchor = offset_site(g, y, neg);
    cx M[4][9];
    for (int j = 0; j < 4; ++j) {
        long site = offset_site(g, anchor, lp->origin[j]);
        load_link(U, g, lp->dir[j], site, M[j]);
        if (lp->dag[j]) mat_dag(M[j], M[j]);
    }
    cx dq[9], f[9];
    const double* e = dFp + anchor * 18;
    for (int k = 0; k < 9; ++k) f[k] = C(e[2 * k], e[2 * k + 1]);
    for (int r = 0; r < 3; ++r)
        for (int c = 0; c < 3; ++c) dq[3 * r + c] = rmul(0.125, csub(f[3 * r + c], cconj(f[3 * c + r])));
    cx left[9], right[9], t[9];
    int has_left = K > 0, has_right = K < 3;
    if (K == 1) memcpy(left, M[0], sizeof(left));
    if (K == 2) mat_mul_plain(M[0], M[1], left);
    if (K == 3) {
        mat_mul_plain(M[0], M[1], t);
        mat_mul_plain(t, M[2], left);
    }
    if (K == 0) {
        mat_mul_plain(M[1], M[2], t);
        mat_mul_plain(t, M[3], right);
    }
    if (K == 1) mat_mul_plain(M[2], M[3], right);
    if (K == 2) memcpy(right, M[3], sizeof(right));
    if (lp->dag[K]) {
        /* right * dQ^dagger * left */
        cx dqd[9];
        mat_dag(dq, dqd);
        if (has_right) mat_mul_plain(right, dqd, t); else memcpy(t, dqd, sizeof(t));
        if (has_left) mat_mul_plain(t, left, h); else memcpy(h, t, sizeof(t));
    } else {
        /* left^dagger * dQ * right^dagger */
        cx ld[9], rd[9];
        if (has_left) {
            mat_dag(left, ld);
            mat_mul_plain(ld, dq, t);
        } else {
            memcpy(t, dq, sizeof(t));
        }
        if (has_right) {
            mat_dag(right, rd);
            mat_mul_plain(t, rd, h);
        } else {
            memcpy(h, t, sizeof(t));
        }
    }
}

/* _kernel_clover_links_pullback! (WilsonCloverPullback.jl:236-280), planes in
 * order (_clover_links_pullback!, :282-293): for each link site and plane
 * (mu, nu), the four leaves (mu,nu,-mu,-nu), (nu,-mu,-nu,mu), (-mu,-nu,mu,nu),
 * (-nu,mu,nu,-mu) and their four factors; dU[mu] += dmu, dU[nu] += dnu. */
static void clover_links_pullback(double* dU, const double* U, const double* dF, const geom* g) {
    leaf_path paths[6][4];
    for (int pl = 0; pl < 6; ++pl) {
        int m = plane_mu[pl] + 1, n = plane_nu[pl] + 1;
        int p1[4] = {m, n, -m, -n}, p2[4] = {n, -m, -n, m}, p3[4] = {-m, -n, m, n}, p4[4] = {-n, m, n, -m};
        make_path(p1, &paths[pl][0]);
        make_path(p2, &paths[pl][1]);
        make_path(p3, &paths[pl][2]);
        make_path(p4, &paths[pl][3]);
    }
#pragma omp parallel for schedule(static)
    for (long y = 0; y < g->V; ++y) {
        for (int pl = 0; pl < 6; ++pl) {
            const int mu = plane_mu[pl], nu = plane_nu[pl];
            const double* dFp = dF + (long)pl * g->V * 18;
            cx dmu[9], dnu[9];
            for (int k = 0; k < 9; ++k) dmu[k] = dnu[k] = C(0, 0);
            for (int p = 0; p < 4; ++p) {
                cx hmu[9], hnu[9];
                for (int k = 0; k < 9; ++k) hmu[k] = hnu[k] = C(0, 0);
                for (int K = 0; K < 4; ++K) {
                    cx h[9];
                    occurrence_pullback(g, U, dFp, y, &paths[pl][p], K, h);
                    cx* acc = paths[pl][p].dir[K] == mu ? hmu : hnu;
                    for (int k = 0; k < 9; ++k) acc[k] = cadd(acc[k], h[k]);
                }
                for (int k = 0; k < 9; ++k) {
                    dmu[k] = cadd(dmu[k], hmu[k]);
                    dnu[k] = cadd(dnu[k], hnu[k]);
                }
            }
            double* om = dU + ((long)mu * g->V + y) * 18;
            double* on = dU + ((long)nu * g->V + y) * 18;
            for (int k = 0; k < 9; ++k) {
                cx a = cadd(C(om[2 * k], om[2 * k + 1]), dmu[k]);
                om[2 * k] = a.re;
                om[2 * k + 1] = a.im;
                cx b = cadd(C(on[2 * k], on[2 * k + 1]), dnu[k]);
                on[2 * k] = b.re;
                on[2 * k + 1] = b.im;
            }
        }
    }
}

/* wilson_clover_link_pullback! (WilsonCloverPullback.jl:349-389) with the
 * coefficients of this library's pullback: dU += coefficient * (hopping
 * pullback) + the clover part with coefficient * csw (the reference passes
 * -kappa*scale and -kappa*cSW*scale).  dU: [4][V][3][3] complex. */
void orc_clover_pullback(double* dU, const double* U, const double* dres, const double* psi, const int* L,
                         const double* phases, double coefficient, double csw) {
    geom g;
    geom_init(&g, L, NULL);
    init_gamma_products();
    orc_wilson_pullback(dU, dres, psi, L, phases, coefficient);
    double* dF = (double*)calloc((size_t)6 * g.V * 18, sizeof(double));
    clover_field_cotangent(dF, dres, psi, g.V, coefficient * csw);
    clover_links_pullback(dU, U, dF, &g);
    free(dF);
}

/* dSFdU for the Wilson-clover operator: as orc_dsfdu with D the clover
 * operator (F the field strength of U) and the clover link pullback. */
int orc_dsfdu_clover(double* G, double* eta, double* chi, const double* U, const double* F, const double* phi,
                     const int* L, const double* phases, double kappa, double csw, double rtol, int maxiter,
                     int dot_order, double* work, orc_cg_result* res) {
    geom g;
    geom_init(&g, L, phases);
    long V = g.V;
    double *r = work, *p = work + 24 * V, *Ap = work + 48 * V, *tmp = work + 72 * V;
    orc_op op;
    memset(&op, 0, sizeof(op));
    op.kind = ORC_OP_DDAGD_CLOVER;
    op.U = U;
    op.F = F;
    memcpy(op.L, L, sizeof(op.L));
    for (int k = 0; k < 8; ++k) op.phases[k] = phases ? phases[k] : (k % 2 ? 0.0 : 1.0);
    op.kappa = kappa;
    op.csw = csw;
    op.tmp = tmp;
    memset(eta, 0, sizeof(double) * 24 * V);
    orc_cg(eta, phi, r, p, Ap, &op, V, rtol, 0.0, maxiter, dot_order, res);
    orc_wilson_clover(chi, eta, U, F, L, phases, kappa, csw, 0);
    double* dres = tmp;
    for (long i = 0; i < 24 * V; ++i) dres[i] = -2.0 * chi[i];
    memset(G, 0, sizeof(double) * 4 * V * 18);
    orc_clover_pullback(G, U, dres, eta, L, phases, -kappa, csw);
    return 0;
}

/* ------------------------------------------------------------------------ */
/* The reference's CPU path of the D^dagger D CG, as cg! runs it on the      */
/* JACC threads backend: halo-padded AoS fields (LatticeMatrix.A, nw = 1),   */
/* set_halo! before every operator application, the per-site half-spin       */
/* kernel reading its neighbours from the halo, unfused BLAS-1 over the core  */
/* sites, reductions per thread then summed.  CPU baseline only: bench.py     */
/* times it beside the GPU solve.                                             */
/* ------------------------------------------------------------------------ */
typedef struct {
    int L[4], P[4];
    long V, Vp, pstride[4];
    cx ph[4], iph[4];
} pgeom;

static void pgeom_init(pgeom* g, const int* L, const double* phases) {
    g->V = 1;
    g->Vp = 1;
    for (int d = 0; d < 4; ++d) {
        g->L[d] = L[d];
        g->P[d] = L[d] + 2;
        g->pstride[d] = g->Vp;
        g->V *= L[d];
        g->Vp *= g->P[d];
        g->ph[d] = C(phases ? phases[2 * d] : 1.0, phases ? phases[2 * d + 1] : 0.0);
        g->iph[d] = cinv(g->ph[d]);
    }
}

/* padded index of core site s (x fastest) */
static inline long pcore(const pgeom* g, long s) {
    long ps = 0;
    for (int d = 0; d < 4; ++d) {
        ps += (long)(s % g->L[d] + 1) * g->pstride[d];
        s /= g->L[d];
    }
    return ps;
}

/* set_halo! / exchange_dim_local! on one rank (src/LatticeMatrices_core.jl:555-569,707-723):
 * direction by direction over the whole padded extent of the others, the minus
 * ghost = the last core slice times inv(phase), the plus ghost = the first
 * times phase (skipped for phase 1); ne complex entries per site. */
static void set_halo(const pgeom* g, double* a, int ne, int with_phases) {
    for (int d = 0; d < 4; ++d) {
        const long sd = g->pstride[d];
        const long nslab = g->Vp / g->P[d];
        const int skip = !with_phases || (g->ph[d].re == 1.0 && g->ph[d].im == 0.0);
#pragma omp parallel for schedule(static)
        for (long q = 0; q < nslab; ++q) {
            /* q enumerates the padded sites with x_d = 0 */
            const long lo = q % sd, hi = q / sd;
            const long base = lo + hi * sd * g->P[d];
            double* gm = a + 2 * ne * base;
            double* gp = a + 2 * ne * (base + (long)(g->L[d] + 1) * sd);
            const double* cm = a + 2 * ne * (base + (long)g->L[d] * sd);
            const double* cp = a + 2 * ne * (base + sd);
            for (int k = 0; k < ne; ++k) {
                cx vm = C(cm[2 * k], cm[2 * k + 1]), vp = C(cp[2 * k], cp[2 * k + 1]);
                if (!skip) {
                    vm = cmul(vm, g->iph[d]);
                    vp = cmul(vp, g->ph[d]);
                }
                gm[2 * k] = vm.re;
                gm[2 * k + 1] = vm.im;
                gp[2 * k] = vp.re;
                gp[2 * k + 1] = vp.im;
            }
        }
    }
}

/* link U_mu at padded site ps from the column-major A[3,3,...] arrays */
static inline void load_link_pad(const double* const* U, int mu, long ps, cx u[9]) {
    const double* p = U[mu] + ps * 18;
    for (int row = 0; row < 3; ++row)
        for (int col = 0; col < 3; ++col) u[row * 3 + col] = C(p[2 * (row + 3 * col)], p[2 * (row + 3 * col) + 1]);
}

/* mul!(C, D, psi) or mul!(C, D', psi): set_halo!(psi), then the half-spin kernel per core site */
static void refcpu_wilson(const pgeom* g, double* out, double* in, const double* const* U, double kappa, int dagger) {
    set_halo(g, in, 12, 1);
    const int fpm = dagger ? 1 : -1;
#pragma omp parallel for schedule(static)
    for (long s = 0; s < g->V; ++s) {
        const long ps = pcore(g, s);
        cx acc[12], sp[12], u[9], h[6], r[6];
        for (int k = 0; k < 12; ++k) acc[k] = C(0.0, 0.0);
        for (int mu = 0; mu < 4; ++mu) {
            load_spinor(in, ps + g->pstride[mu], sp);
            for (int c = 0; c < 3; ++c) half_project(sp, c, fpm, mu, &h[2 * c], &h[2 * c + 1]);
            load_link_pad(U, mu, ps, u);
            for (int a = 0; a < 3; ++a) {
                r[2 * a] = muladdmulti(u[3 * a], h[0], u[3 * a + 1], h[2], u[3 * a + 2], h[4]);
                r[2 * a + 1] = muladdmulti(u[3 * a], h[1], u[3 * a + 1], h[3], u[3 * a + 2], h[5]);
            }
            reconstruct_add(acc, r, fpm, mu);
            const long sm = ps - g->pstride[mu];
            load_spinor(in, sm, sp);
            for (int c = 0; c < 3; ++c) half_project(sp, c, -fpm, mu, &h[2 * c], &h[2 * c + 1]);
            load_link_pad(U, mu, sm, u);
            for (int a = 0; a < 3; ++a) {
                cx w0 = cconj(u[a]), w1 = cconj(u[3 + a]), w2 = cconj(u[6 + a]);
                r[2 * a] = muladdmulti(w0, h[0], w1, h[2], w2, h[4]);
                r[2 * a + 1] = muladdmulti(w0, h[1], w1, h[3], w2, h[5]);
            }
            reconstruct_add(acc, r, -fpm, mu);
        }
        double* o = out + ps * 24;
        const double* p = in + ps * 24;
        for (int spin = 0; spin < 4; ++spin)
            for (int c = 0; c < 3; ++c) {
                cx v = cadd(C(p[2 * (spin * 3 + c)], p[2 * (spin * 3 + c) + 1]), rmul(-kappa, acc[4 * c + spin]));
                o[2 * (spin * 3 + c)] = v.re;
                o[2 * (spin * 3 + c) + 1] = v.im;
            }
    }
}

/* dot over the core sites: per-site sums, per-thread partial sums, then summed */
static double refcpu_dot_re(const pgeom* g, const double* a, const double* b) {
    double acc = 0.0;
#pragma omp parallel for schedule(static) reduction(+ : acc)
    for (long s = 0; s < g->V; ++s) acc += site_dot(a, b, pcore(g, s), 12).re;
    return acc;
}

/* axpby!(a, X, b, Y) over the core sites */
static void refcpu_axpby(const pgeom* g, double a, const double* X, double b, double* Y) {
#pragma omp parallel for schedule(static)
    for (long s = 0; s < g->V; ++s) {
        const long o = pcore(g, s) * 24;
        for (int k = 0; k < 24; ++k) Y[o + k] = a * X[o + k] + b * Y[o + k];
    }
}

/* cg!(x, DdagDOp(D, temp), b, r, p, Ap) (src/LinearAlgebras/cg.jl:73-137) on
 * padded fields: spinors [Vp][4][3] complex (LatticeMatrix.A, nw = 1), links
 * four column-major [Vp][3][3] arrays with their periodic halo filled. */
int orc_refcpu_cg_ddagd(double* x, const double* b, double* r, double* p, double* Ap, double* tmp,
                        const double* const* U, const int* L, const double* phases, double kappa, double rtol,
                        int maxiter, orc_cg_result* res) {
    pgeom g;
    pgeom_init(&g, L, phases);
    double bb = refcpu_dot_re(&g, b, b);
    double bnorm = sqrt(bb), tol = rtol * bnorm;
    refcpu_wilson(&g, tmp, x, U, kappa, 0);
    refcpu_wilson(&g, Ap, tmp, U, kappa, 1);
    memcpy(r, b, sizeof(double) * 24 * g.Vp);
    refcpu_axpby(&g, -1.0, Ap, 1.0, r);
    memcpy(p, r, sizeof(double) * 24 * g.Vp);
    double rr = refcpu_dot_re(&g, r, r);
    double rn = sqrt(rr);
    if (rn <= tol) {
        set_result(res, 1, 0, rn, bnorm, ORC_CONVERGED);
        return 0;
    }
    for (int it = 1; it <= maxiter; ++it) {
        refcpu_wilson(&g, tmp, p, U, kappa, 0);
        refcpu_wilson(&g, Ap, tmp, U, kappa, 1);
        const double pAp = refcpu_dot_re(&g, p, Ap);
        if (!(pAp > 0.0)) {
            set_result(res, 0, it - 1, rn, bnorm, isfinite(pAp) ? ORC_NONPOSITIVE_CURVATURE : ORC_BREAKDOWN);
            return 0;
        }
        const double alpha = rr / pAp;
        refcpu_axpby(&g, alpha, p, 1.0, x);
        refcpu_axpby(&g, -alpha, Ap, 1.0, r);
        const double rr_new = refcpu_dot_re(&g, r, r);
        rn = sqrt(rr_new);
        if (rn <= tol) {
            set_result(res, 1, it, rn, bnorm, ORC_CONVERGED);
            return 0;
        }
        refcpu_axpby(&g, 1.0, r, rr_new / rr, p);
        rr = rr_new;
    }
    set_result(res, 0, maxiter, rn, bnorm, ORC_MAXITER);
    return 0;
}

/* LatticeMatrix(A, dim, PEs; nw = 1) of a core spinor [V][4][3] or of four
 * core link fields [4][V][3][3] (row-major in, column-major padded out). */
void orc_refcpu_pad_spinor(double* dst, const double* src, const int* L, const double* phases) {
    pgeom g;
    pgeom_init(&g, L, phases);
    memset(dst, 0, sizeof(double) * 24 * g.Vp);
    for (long s = 0; s < g.V; ++s) memcpy(dst + pcore(&g, s) * 24, src + s * 24, sizeof(double) * 24);
    set_halo(&g, dst, 12, 1);
}
void orc_refcpu_pad_links(double* const* dst, const double* src, const int* L) {
    pgeom g;
    pgeom_init(&g, L, NULL);
    for (int mu = 0; mu < 4; ++mu) {
        memset(dst[mu], 0, sizeof(double) * 18 * g.Vp);
        for (long s = 0; s < g.V; ++s) {
            const double* a = src + ((long)mu * g.V + s) * 18;
            double* o = dst[mu] + pcore(&g, s) * 18;
            for (int row = 0; row < 3; ++row)
                for (int col = 0; col < 3; ++col) {
                    o[2 * (row + 3 * col)] = a[2 * (row * 3 + col)];
                    o[2 * (row + 3 * col) + 1] = a[2 * (row * 3 + col) + 1];
                }
        }
        set_halo(&g, dst[mu], 9, 0);
    }
}
void orc_refcpu_unpad_spinor(double* dst, const double* src, const int* L) {
    pgeom g;
    pgeom_init(&g, L, NULL);
    for (long s = 0; s < g.V; ++s) memcpy(dst + s * 24, src + pcore(&g, s) * 24, sizeof(double) * 24);
}

int orc_num_threads(void) {
#ifdef _OPENMP
    return omp_get_max_threads();
#else
    return 1;
#endif
}

void orc_set_num_threads(int n) {
#ifdef _OPENMP
    if (n > 0) omp_set_num_threads(n);
#else
    (void)n;
#endif
}
