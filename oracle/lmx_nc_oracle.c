/*
 * lmx_nc_oracle.c -- CPU oracle of the reference's generic-NC Wilson and
 * Wilson-clover kernels (NC != 3, e.g. the SU(2) fields of the reference's own
 * tests).
 *
 * TEST INFRASTRUCTURE ONLY, like lmx_oracle.c: only tests/, smoke() and
 * bench.py's cpu_baseline leg load it, as the checker.
 *
 * For NC = 3 the reference runs its half-spin kernels (restated in
 * lmx_oracle.c); for any other NC it runs the plain kernels restated here in
 * their operation order:
 *   kernel_WilsonDiracOperator4D!           src/Operators/WilsonDiracOperator.jl:80-147
 *   kernel_adjoint_WilsonDiracOperator4D!   :444-511
 *   kernel_WilsonDiracOperator4D_Donly!     :554-621
 *   kernel_adjoint_..._Donly!               :665-733 (local kappa = -0.5)
 *   mul_op(Oneγ{pm,mu}, x, ic, idx)         src/Operators/Operators.jl:71-133
 *   _clover_q_element_halo + store          src/Operators/WilsonDiracCloverOperator.jl:57-118
 *   kernel_add_clover_term!                 :481-525
 *   _kernel_wilson_link_pullback_direction! ext/AD/WilsonDiracOperator.jl:27-53
 * Layouts (no halo, site x-fastest):
 *   spinor psi[(s*4 + spin)*NC + colour]  (the reference's A[NC,4,...] core)
 *   links  U[((mu*V + s)*NC + row)*NC + col]
 *   planes F[((plane*V + s)*NC + row)*NC + col], planes (12,13,14,23,24,34)
 * A neighbour read across the boundary carries the spinor phase (forward
 * wrap: phase, backward wrap: inv(phase)), the value set_halo! puts in the
 * halo; links are periodic.  `i * z` is evaluated as (-im z, re z), equal to
 * Julia's Complex{Bool} product up to the sign of exact zeros.
 *
 * Pinned by (tests/test_oracle_nc.py): the reference's dense restatement of
 * test/wilson_clover.jl:47-112 (numpy, any NC) on its closed-form SU(2)
 * fields (:12-45); F anti-Hermitian and zero for unit links; cSW = 0 equal
 * to Wilson; adjoint identity and gamma5-Hermiticity; gauge covariance under
 * the test's SU(2) transformation (:125-150); the NC = 3 generic order against
 * the half-spin oracle to rounding; the free-field closed form; exact central
 * differences of Re <w, D(U) psi> for the link pullback; planted mistakes.
 */
#include <math.h>
#include <stdlib.h>
#include <string.h>

typedef struct {
    double re, im;
} ncx;

static inline ncx Z(double r, double i) {
    ncx z;
    z.re = r;
    z.im = i;
    return z;
}
static inline ncx zadd(ncx a, ncx b) { return Z(a.re + b.re, a.im + b.im); }
static inline ncx zsub(ncx a, ncx b) { return Z(a.re - b.re, a.im - b.im); }
static inline ncx zneg(ncx a) { return Z(-a.re, -a.im); }
static inline ncx zconj(ncx a) { return Z(a.re, -a.im); }
static inline ncx zmul(ncx a, ncx b) { return Z(a.re * b.re - a.im * b.im, a.re * b.im + a.im * b.re); }
static inline ncx zrmul(double a, ncx b) { return Z(a * b.re, a * b.im); }
static inline ncx zi(ncx z) { return Z(-z.im, z.re); } /* im * z */

typedef struct {
    int L[4];
    long V, stride[4];
    ncx ph[4], iph[4];
} ngeom;

static void ngeom_init(ngeom* g, const int* L, const double* phases) {
    g->V = 1;
    for (int d = 0; d < 4; ++d) {
        g->L[d] = L[d];
        g->stride[d] = g->V;
        g->V *= L[d];
        g->ph[d] = Z(phases ? phases[2 * d] : 1.0, phases ? phases[2 * d + 1] : 0.0);
        const double q = g->ph[d].re * g->ph[d].re + g->ph[d].im * g->ph[d].im;
        g->iph[d] = Z(g->ph[d].re / q, -g->ph[d].im / q);
    }
}

static inline void ncoords(const ngeom* g, long s, int x[4]) {
    for (int d = 0; d < 4; ++d) {
        x[d] = (int)(s % g->L[d]);
        s /= g->L[d];
    }
}

/* neighbour of s in direction mu (dir +1 / -1); *wrap set when it crosses the boundary */
static inline long nnbr(const ngeom* g, long s, const int x[4], int mu, int dir, int* wrap) {
    if (dir > 0) {
        *wrap = x[mu] == g->L[mu] - 1;
        return *wrap ? s - (long)(g->L[mu] - 1) * g->stride[mu] : s + g->stride[mu];
    }
    *wrap = x[mu] == 0;
    return *wrap ? s + (long)(g->L[mu] - 1) * g->stride[mu] : s - g->stride[mu];
}

static inline ncx ld(const double* a, long i) { return Z(a[2 * i], a[2 * i + 1]); }
static inline void st(double* a, long i, ncx z) {
    a[2 * i] = z.re;
    a[2 * i + 1] = z.im;
}

/* spinor value (colour c, spin sp) of site s as the halo holds it */
static inline ncx psi_at(const double* psi, long s, int nc, int sp, int c, int wrap, ncx ph) {
    ncx v = ld(psi, (s * 4 + sp) * nc + c);
    return wrap ? zmul(v, ph) : v;
}

/* mul_op(Oneγ{pm,mu}, x, c, idx): the four spin components of (1 + pm gamma_mu) x */
static void mul_op(int pm, int mu, const ncx x[4], ncx v[4]) {
    const ncx x1 = x[0], x2 = x[1], x3 = x[2], x4 = x[3];
    if (pm > 0) {
        switch (mu) {
        case 0: v[0] = zsub(x1, zi(x4)); v[1] = zsub(x2, zi(x3)); v[2] = zadd(x3, zi(x2)); v[3] = zadd(x4, zi(x1)); break;
        case 1: v[0] = zsub(x1, x4); v[1] = zadd(x2, x3); v[2] = zadd(x3, x2); v[3] = zsub(x4, x1); break;
        case 2: v[0] = zsub(x1, zi(x3)); v[1] = zadd(x2, zi(x4)); v[2] = zadd(x3, zi(x1)); v[3] = zsub(x4, zi(x2)); break;
        default: v[0] = zsub(x1, x3); v[1] = zsub(x2, x4); v[2] = zsub(x3, x1); v[3] = zsub(x4, x2); break;
        }
    } else {
        switch (mu) {
        case 0: v[0] = zadd(x1, zi(x4)); v[1] = zadd(x2, zi(x3)); v[2] = zsub(x3, zi(x2)); v[3] = zsub(x4, zi(x1)); break;
        case 1: v[0] = zadd(x1, x4); v[1] = zsub(x2, x3); v[2] = zsub(x3, x2); v[3] = zadd(x4, x1); break;
        case 2: v[0] = zadd(x1, zi(x3)); v[1] = zsub(x2, zi(x4)); v[2] = zsub(x3, zi(x1)); v[3] = zadd(x4, zi(x2)); break;
        default: v[0] = zadd(x1, x3); v[1] = zadd(x2, x4); v[2] = zadd(x3, x1); v[3] = zadd(x4, x2); break;
        }
    }
}

/* Generic-NC hopping operator at site s: res = psi (or 0 for Donly), then for
 * ic, jc: forward hops mu = 1..4, backward hops mu = 1..4, each adding
 * (coef * U) * v[ia] (backward: (coef * conj(U[jc, ic](x - mu))) * v[ia]).
 * fpm: projector sign of the forward hops (-1: D, +1: D^dagger). */
static void wilson_nc_site(const ngeom* g, const double* U, const double* psi, long s, int nc, int fpm, double coef,
                           int donly, ncx* res /* [spin*nc + colour] */) {
    int x[4];
    ncoords(g, s, x);
    long sp[4], sm[4];
    int wp[4], wm[4];
    for (int mu = 0; mu < 4; ++mu) {
        sp[mu] = nnbr(g, s, x, mu, +1, &wp[mu]);
        sm[mu] = nnbr(g, s, x, mu, -1, &wm[mu]);
    }
    for (int ic = 0; ic < nc; ++ic)
        for (int ia = 0; ia < 4; ++ia) res[ia * nc + ic] = donly ? Z(0.0, 0.0) : ld(psi, (s * 4 + ia) * nc + ic);
    for (int ic = 0; ic < nc; ++ic) {
        for (int jc = 0; jc < nc; ++jc) {
            for (int mu = 0; mu < 4; ++mu) {
                ncx xv[4], v[4];
                for (int a = 0; a < 4; ++a) xv[a] = psi_at(psi, sp[mu], nc, a, jc, wp[mu], g->ph[mu]);
                mul_op(fpm, mu, xv, v);
                const ncx u = zrmul(coef, ld(U, ((mu * g->V + s) * nc + ic) * nc + jc));
                for (int ia = 0; ia < 4; ++ia) res[ia * nc + ic] = zadd(res[ia * nc + ic], zmul(u, v[ia]));
            }
            for (int mu = 0; mu < 4; ++mu) {
                ncx xv[4], v[4];
                for (int a = 0; a < 4; ++a) xv[a] = psi_at(psi, sm[mu], nc, a, jc, wm[mu], g->iph[mu]);
                mul_op(-fpm, mu, xv, v);
                const ncx u = zrmul(coef, zconj(ld(U, ((mu * g->V + sm[mu]) * nc + jc) * nc + ic)));
                for (int ia = 0; ia < 4; ++ia) res[ia * nc + ic] = zadd(res[ia * nc + ic], zmul(u, v[ia]));
            }
        }
    }
}

/* mul!(C, D, psi) for NC != 3: D (dagger 0) or D^dagger (dagger 1), the
 * hopping-only operator with donly (coefficient 1/2, no identity term). */
void orc_wilson_nc(double* out, const double* in, const double* U, const int* L, const double* phases, double kappa,
                   int nc, int dagger, int donly) {
    ngeom g;
    ngeom_init(&g, L, phases);
    const int fpm = dagger ? 1 : -1;
    const double coef = donly ? 0.5 : -kappa;
#pragma omp parallel for schedule(static)
    for (long s = 0; s < g.V; ++s) {
        ncx res[64];
        wilson_nc_site(&g, U, in, s, nc, fpm, coef, donly, res);
        for (int k = 0; k < 4 * nc; ++k) st(out, s * 4 * nc + k, res[k]);
    }
}

static const int nplane_mu[6] = {0, 0, 0, 1, 1, 2};
static const int nplane_nu[6] = {1, 2, 3, 2, 3, 3};

static inline long nshift(const ngeom* g, long s, int mu, int dir) {
    int x[4], w;
    ncoords(g, s, x);
    return nnbr(g, s, x, mu, dir, &w);
}

/* U_mu(site)[r, c] */
static inline ncx lk(const double* U, const ngeom* g, int nc, int mu, long site, int r, int c) {
    return ld(U, ((mu * g->V + site) * nc + r) * nc + c);
}

/* _clover_q_element_halo, src/Operators/WilsonDiracCloverOperator.jl:57-84 */
static ncx q_element(const double* U, const ngeom* g, int nc, int mu, int nu, long x, int row, int col) {
    const long xpmu = nshift(g, x, mu, 1), xpnu = nshift(g, x, nu, 1), xmmu = nshift(g, x, mu, -1),
               xmnu = nshift(g, x, nu, -1);
    const long xpnu_mmu = nshift(g, xpnu, mu, -1), xmmu_mnu = nshift(g, xmmu, nu, -1),
               xpmu_mnu = nshift(g, xmnu, mu, 1);
    ncx value = Z(0.0, 0.0);
    for (int a = 0; a < nc; ++a)
        for (int b = 0; b < nc; ++b)
            for (int c = 0; c < nc; ++c) {
                ncx t = zmul(zmul(zmul(lk(U, g, nc, mu, x, row, a), lk(U, g, nc, nu, xpmu, a, b)),
                                  zconj(lk(U, g, nc, mu, xpnu, c, b))),
                             zconj(lk(U, g, nc, nu, x, col, c)));
                value = zadd(value, t);
                t = zmul(zmul(zmul(lk(U, g, nc, nu, x, row, a), zconj(lk(U, g, nc, mu, xpnu_mmu, b, a))),
                              zconj(lk(U, g, nc, nu, xmmu, c, b))),
                         lk(U, g, nc, mu, xmmu, c, col));
                value = zadd(value, t);
                t = zmul(zmul(zmul(zconj(lk(U, g, nc, mu, xmmu, a, row)), zconj(lk(U, g, nc, nu, xmmu_mnu, b, a))),
                              lk(U, g, nc, mu, xmmu_mnu, b, c)),
                         lk(U, g, nc, nu, xmnu, c, col));
                value = zadd(value, t);
                t = zmul(zmul(zmul(zconj(lk(U, g, nc, nu, xmnu, a, row)), lk(U, g, nc, mu, xmnu, a, b)),
                              lk(U, g, nc, nu, xpmu_mnu, b, c)),
                         zconj(lk(U, g, nc, mu, x, col, c)));
                value = zadd(value, t);
            }
    return value;
}

/* CloverFieldStrength4D(U) for NC != 3: F_plane = (Q - Q^dagger) / 8, element
 * by element as kernel_clover_field_strength_halo! (:86-118) stores it. */
void orc_clover_field_strength_nc(double* F, const double* U, const int* L, int nc) {
    ngeom g;
    ngeom_init(&g, L, NULL);
    for (int pl = 0; pl < 6; ++pl) {
        const int mu = nplane_mu[pl], nu = nplane_nu[pl];
#pragma omp parallel for schedule(static)
        for (long s = 0; s < g.V; ++s) {
            for (int col = 0; col < nc; ++col)
                for (int row = 0; row <= col; ++row) {
                    const ncx qrc = q_element(U, &g, nc, mu, nu, s, row, col);
                    const ncx qcr = row == col ? qrc : q_element(U, &g, nc, mu, nu, s, col, row);
                    const ncx value = zrmul(0.125, zsub(qrc, zconj(qcr)));
                    st(F, ((pl * g.V + s) * nc + row) * nc + col, value);
                    if (row != col) st(F, ((pl * g.V + s) * nc + col) * nc + row, zneg(zconj(value)));
                }
        }
    }
}

/* kernel_add_clover_term!, src/Operators/WilsonDiracCloverOperator.jl:481-525 */
static void clover_add_nc(const ngeom* g, const double* F, const double* psi, long s, int nc, double coefficient,
                          ncx* res) {
    for (int color = 0; color < nc; ++color) {
        ncx v1 = Z(0, 0), v2 = Z(0, 0), v3 = Z(0, 0), v4 = Z(0, 0);
        for (int ic = 0; ic < nc; ++ic) {
            ncx f[6];
            for (int pl = 0; pl < 6; ++pl) f[pl] = ld(F, ((pl * g->V + s) * nc + color) * nc + ic);
            const ncx f12 = f[0], f13 = f[1], f14 = f[2], f23 = f[3], f24 = f[4], f34 = f[5];
            const ncx dp = zi(zadd(f12, f34)), dm = zi(zsub(f12, f34));
            const ncx odp = zi(zadd(f14, f23)), odm = zi(zsub(f23, f14));
            const ncx p1 = ld(psi, (s * 4 + 0) * nc + ic), p2 = ld(psi, (s * 4 + 1) * nc + ic);
            const ncx p3 = ld(psi, (s * 4 + 2) * nc + ic), p4 = ld(psi, (s * 4 + 3) * nc + ic);
            v1 = zadd(v1, zadd(zmul(dp, p1), zmul(zadd(zadd(zneg(f13), odp), f24), p2)));
            v2 = zadd(v2, zadd(zmul(zneg(dp), p2), zmul(zsub(zadd(f13, odp), f24), p1)));
            v3 = zadd(v3, zadd(zmul(dm, p3), zmul(zsub(zadd(zneg(f13), odm), f24), p4)));
            v4 = zadd(v4, zadd(zmul(zneg(dm), p4), zmul(zadd(zadd(f13, odm), f24), p3)));
        }
        res[0 * nc + color] = zadd(res[0 * nc + color], zrmul(coefficient, v1));
        res[1 * nc + color] = zadd(res[1 * nc + color], zrmul(coefficient, v2));
        res[2 * nc + color] = zadd(res[2 * nc + color], zrmul(coefficient, v3));
        res[3 * nc + color] = zadd(res[3 * nc + color], zrmul(coefficient, v4));
    }
}

/* Wilson-clover mul! for NC != 3: the generic Wilson kernel, then the clover
 * term with coefficient (-kappa) * cSW (kernel_WilsonDiracCloverOperator4D!
 * and its adjoint, :537-559). */
void orc_wilson_clover_nc(double* out, const double* in, const double* U, const double* F, const int* L,
                          const double* phases, double kappa, double csw, int nc, int dagger) {
    ngeom g;
    ngeom_init(&g, L, phases);
    const int fpm = dagger ? 1 : -1;
    const double coefficient = (-kappa) * csw;
#pragma omp parallel for schedule(static)
    for (long s = 0; s < g.V; ++s) {
        ncx res[64];
        wilson_nc_site(&g, U, in, s, nc, fpm, -kappa, 0, res);
        clover_add_nc(&g, F, in, s, nc, coefficient, res);
        for (int k = 0; k < 4 * nc; ++k) st(out, s * 4 * nc + k, res[k]);
    }
}

/* _kernel_wilson_link_pullback_direction! for any NC
 * (ext/AD/WilsonDiracOperator.jl:27-53), driven per site by
 * _kernel_wilson_link_pullback! (:62-91): dU[mu](x)[row, col] += coefficient *
 * sum_spin (dres[row](x) conj(((1 - gamma_mu) psi[col])(x + mu)) +
 * ((1 + gamma_mu) psi[row])(x) conj(dres[col](x + mu))). */
void orc_wilson_pullback_nc(double* dU, const double* dres, const double* psi, const int* L, const double* phases,
                            double coefficient, int nc) {
    ngeom g;
    ngeom_init(&g, L, phases);
#pragma omp parallel for schedule(static)
    for (long s = 0; s < g.V; ++s) {
        int x[4];
        ncoords(&g, s, x);
        for (int mu = 0; mu < 4; ++mu) {
            int w;
            const long sp = nnbr(&g, s, x, mu, +1, &w);
            for (int row = 0; row < nc; ++row) {
                ncx xv[4], minus_psi[4];
                for (int a = 0; a < 4; ++a) xv[a] = ld(psi, (s * 4 + a) * nc + row);
                mul_op(+1, mu, xv, minus_psi);
                for (int col = 0; col < nc; ++col) {
                    ncx pv[4], plus_psi[4];
                    for (int a = 0; a < 4; ++a) pv[a] = psi_at(psi, sp, nc, a, col, w, g.ph[mu]);
                    mul_op(-1, mu, pv, plus_psi);
                    ncx value = Z(0.0, 0.0);
                    for (int spin = 0; spin < 4; ++spin) {
                        value = zadd(value, zmul(ld(dres, (s * 4 + spin) * nc + row), zconj(plus_psi[spin])));
                        value = zadd(value, zmul(minus_psi[spin],
                                                 zconj(psi_at(dres, sp, nc, spin, col, w, g.ph[mu]))));
                    }
                    const long o = ((mu * g.V + s) * nc + row) * nc + col;
                    st(dU, o, zadd(ld(dU, o), zrmul(coefficient, value)));
                }
            }
        }
    }
}
