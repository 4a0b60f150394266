/*
 * lmx_rng_oracle.c -- CPU oracle for the site random-number streams and the
 * global-site lattice fills of LatticeMatrices.jl.
 *
 * TEST INFRASTRUCTURE ONLY.  Only tests/, __graft_entry__.smoke() and the
 * cpu_baseline leg of bench.py may load this code (it is linked into
 * liblmx_oracle.so), and only as the checker.  The product library never
 * links, loads or calls it.
 *
 * Restated from src/utilities/randomgenerator.jl and
 * src/LinearAlgebras/linearalgebra_D.jl:2550-2760 (file:line per function):
 *   - a stream key (seed, sweep, direction, color, subgroup) is folded into a
 *     64-bit stream seed with SplitMix64 finalisers and fixed salts;
 *   - each zero-based global site id (x fastest, column-major) owns one
 *     stream: PCG32 (XSH-RR), xoshiro256++ or Philox4x32-10 (counter =
 *     (block lo, block hi, site lo, site hi), one key per field, four words
 *     cached and handed out in order);
 *   - uniforms: Float32 = (u32 >> 8) 2^-24, Float64 = (u64 >> 11) 2^-53 with
 *     u64 = (first u32 << 32) | second u32 for the 32-bit generators, and the
 *     high half of the 64-bit output for xoshiro's u32;
 *   - normals: Box-Muller from an open uniform (radius) then a half-open one
 *     (angle), with the C library's log / sqrt / cos / sin standing in for
 *     Julia's (both are faithfully rounded: last-bit differences possible);
 *   - fills: per site, elements in (ic fastest, jc) order; complex elements
 *     draw (re, im); Gaussian real fields use both normals of a pair.
 *
 * Parity pinning (tests/test_oracle_rng.py):
 *   philox4x32_10 : Random123's published known-answer vectors (the
 *                   zero-counter/zero-key one is also test/site_rng.jl:136-145);
 *   pcg32_step    : the PCG reference demo sequence (seed 42, sequence 54);
 *   site streams  : test/site_rng.jl:23-37 (PCG32 words), :120-134
 *                   (xoshiro256++ u64 words), :146-160 (Philox words) for the
 *                   key (0x123456789abcdef0, 17, 3, 1, 2) at site 41, which pin
 *                   SplitMix64, the key folding, the per-site seeding of all
 *                   three generators and the Philox lane order;
 *   uniforms      : exact identities with the golden words, range [0, 1) and
 *                   the open variant never 0 over long runs;
 *   normals       : Box-Muller identities recomputed with numpy from the same
 *                   uniforms, and test/site_rng.jl:70-83's moments;
 *   fills         : test/random_fill.jl:1-43's site-loop restatements,
 *                   decomposition independence (a sub-box with an offset
 *                   equals the slice of the whole fill), random_fill.jl's
 *                   statistical test (:117-138).
 */
#include <math.h>
#include <stdint.h>
#include <string.h>

enum { ORC_RNG_PCG32 = 0, ORC_RNG_XOSHIRO = 1, ORC_RNG_PHILOX = 2 };
enum { ORC_ELEM_F32 = 0, ORC_ELEM_F64 = 1, ORC_ELEM_C64 = 2, ORC_ELEM_C128 = 3 };

/* SplitMix64 finaliser -- randomgenerator.jl:122-127 */
uint64_t orc_splitmix64(uint64_t x) {
    uint64_t z = x + 0x9e3779b97f4a7c15ULL;
    z = (z ^ (z >> 30)) * 0xbf58476d1ce4e5b9ULL;
    z = (z ^ (z >> 27)) * 0x94d049bb133111ebULL;
    return z ^ (z >> 31);
}

/* _mix_site_rng_tag -- :129-131 */
static uint64_t mix_tag(uint64_t h, uint64_t tag, uint64_t salt) {
    return orc_splitmix64(h ^ orc_splitmix64(tag + salt));
}

/* _rng_stream_seed -- :133-140 */
uint64_t orc_rng_stream_seed(uint64_t seed, uint64_t sweep, uint32_t direction, uint32_t color, uint32_t subgroup) {
    uint64_t h = orc_splitmix64(seed ^ 0x243f6a8885a308d3ULL);
    h = mix_tag(h, sweep, 0xa4093822299f31d0ULL);
    h = mix_tag(h, (uint64_t)direction, 0x082efa98ec4e6c89ULL);
    h = mix_tag(h, (uint64_t)color, 0x452821e638d01377ULL);
    h = mix_tag(h, (uint64_t)subgroup, 0xbe5466cf34e90c6cULL);
    return h;
}

/* PCG32 step (XSH-RR output of the old state) -- :7-22 */
uint32_t orc_pcg32_step(uint64_t* state, uint64_t inc) {
    const uint64_t old = *state;
    *state = old * 0x5851F42D4C957F2DULL + (inc | 1ULL);
    const uint32_t xs = (uint32_t)(((old >> 18) ^ old) >> 27);
    const unsigned rot = (unsigned)(old >> 59);
    return (xs >> rot) | (xs << ((32u - rot) & 31u));
}

/* Philox4x32-10 block transform -- :267-281 (round :247-265) */
void orc_philox4x32_10(const uint32_t ctr[4], const uint32_t key[2], uint32_t out[4]) {
    uint32_t c0 = ctr[0], c1 = ctr[1], c2 = ctr[2], c3 = ctr[3], k0 = key[0], k1 = key[1];
    for (int r = 0; r < 10; ++r) {
        const uint64_t p0 = (uint64_t)0xd2511f53u * c0;
        const uint64_t p1 = (uint64_t)0xcd9e8d57u * c2;
        const uint32_t n0 = (uint32_t)(p1 >> 32) ^ c1 ^ k0;
        const uint32_t n1 = (uint32_t)p1;
        const uint32_t n2 = (uint32_t)(p0 >> 32) ^ c3 ^ k1;
        const uint32_t n3 = (uint32_t)p0;
        c0 = n0;
        c1 = n1;
        c2 = n2;
        c3 = n3;
        k0 += 0x9e3779b9u;
        k1 += 0xbb67ae85u;
    }
    out[0] = c0;
    out[1] = c1;
    out[2] = c2;
    out[3] = c3;
}

typedef struct {
    int alg;
    uint64_t a, b, c, d; /* PCG: state, inc; xoshiro: s0..s3; Philox: block, site, key */
    uint32_t w[4];
    int lane;
} orc_stream;

/* site_rng(key, global_site, algorithm) -- :152-200 */
static void stream_init(orc_stream* s, uint64_t stream_seed, uint64_t site, int alg) {
    memset(s, 0, sizeof(*s));
    s->alg = alg;
    if (alg == ORC_RNG_PHILOX) { /* :184-200: the key does not depend on the site */
        s->a = 0;
        s->b = site;
        s->c = orc_splitmix64(stream_seed ^ 0x9216d5d98979fb1bULL);
        s->lane = 4;
        return;
    }
    const uint64_t h = mix_tag(stream_seed, site, 0x13198a2e03707344ULL); /* _site_rng_seed :152-156 */
    if (alg == ORC_RNG_PCG32) {                                           /* :158-167 */
        s->a = orc_splitmix64(h ^ 0xc0ac29b7c97c50ddULL);
        s->b = orc_splitmix64(h ^ 0x3f84d5b5b5470917ULL) | 1ULL;
    } else { /* :169-182 */
        const uint64_t g = 0x9e3779b97f4a7c15ULL;
        s->a = orc_splitmix64(h);
        s->b = orc_splitmix64(h + g);
        s->c = orc_splitmix64(h + g + g);
        s->d = orc_splitmix64(h + g + g + g);
    }
}

static inline uint64_t rotl64(uint64_t v, int k) { return (v << k) | (v >> (64 - k)); }

/* xoshiro256++ -- :224-240 */
static uint64_t xoshiro_next(orc_stream* s) {
    const uint64_t result = rotl64(s->a + s->d, 23) + s->a;
    const uint64_t t = s->b << 17;
    uint64_t s2 = s->c ^ s->a;
    uint64_t s3 = s->d ^ s->b;
    const uint64_t s1 = s->b ^ s2;
    const uint64_t s0 = s->a ^ s3;
    s2 ^= t;
    s3 = rotl64(s3, 45);
    s->a = s0;
    s->b = s1;
    s->c = s2;
    s->d = s3;
    return result;
}

/* rand_u32 -- :219-222 (PCG32), :242-245 (xoshiro: high half), :283-319 (Philox lanes) */
static uint32_t stream_u32(orc_stream* s) {
    if (s->alg == ORC_RNG_PCG32) return orc_pcg32_step(&s->a, s->b);
    if (s->alg == ORC_RNG_XOSHIRO) return (uint32_t)(xoshiro_next(s) >> 32);
    if (s->lane == 4) {
        const uint32_t ctr[4] = {(uint32_t)s->a, (uint32_t)(s->a >> 32), (uint32_t)s->b, (uint32_t)(s->b >> 32)};
        const uint32_t key[2] = {(uint32_t)s->c, (uint32_t)(s->c >> 32)};
        orc_philox4x32_10(ctr, key, s->w);
        s->a += 1;
        s->lane = 1;
        return s->w[0];
    }
    return s->w[s->lane++];
}

/* rand_u64 -- :322-333 (32-bit generators: high word first), :227 (xoshiro) */
static uint64_t stream_u64(orc_stream* s) {
    if (s->alg == ORC_RNG_XOSHIRO) return xoshiro_next(s);
    const uint64_t hi = stream_u32(s);
    const uint64_t lo = stream_u32(s);
    return (hi << 32) | lo;
}

/* rand_uniform -- :341-352 */
static float uniform_f32(orc_stream* s) { return (float)(stream_u32(s) >> 8) * 5.9604645e-8f; }
static double uniform_f64(orc_stream* s) { return (double)(stream_u64(s) >> 11) * 1.1102230246251565e-16; }

/* rand_uniform_open -- :360-368 */
static float uniform_open_f32(orc_stream* s) {
    const float v = uniform_f32(s);
    return v == 0.0f ? 2.9802322e-8f : v;
}
static double uniform_open_f64(orc_stream* s) {
    const double v = uniform_f64(s);
    return v == 0.0 ? 5.551115123125783e-17 : v;
}

/* rand_normal_pair -- :380-386 (T(2pi) is 2pi rounded to T) */
static void normal_pair_f32(orc_stream* s, float* z0, float* z1) {
    const float u = uniform_open_f32(s);
    const float v = uniform_f32(s);
    const float radius = sqrtf(-2.0f * logf(u));
    const float angle = (float)6.283185307179586 * v;
    *z0 = radius * cosf(angle);
    *z1 = radius * sinf(angle);
}
static void normal_pair_f64(orc_stream* s, double* z0, double* z1) {
    const double u = uniform_open_f64(s);
    const double v = uniform_f64(s);
    const double radius = sqrt(-2.0 * log(u));
    const double angle = 6.283185307179586 * v;
    *z0 = radius * cos(angle);
    *z1 = radius * sin(angle);
}

/* rand_bounded -- :406-414 (upper != 0) */
static uint32_t bounded_u32(orc_stream* s, uint32_t upper) {
    const uint32_t threshold = (0u - upper) % upper;
    for (;;) {
        const uint32_t v = stream_u32(s);
        if (v >= threshold) return v % upper;
    }
}

/* The draw sequence of test/site_rng_gpu.jl:7-42 for sites first_site ..
 * first_site+nsites-1: counts[0..5] = words, uniform32, uniform64, normal32,
 * normal64 (even), bounded; outputs are [count, nsites] column-major. */
void orc_site_rng_draw(uint64_t stream_seed, int alg, uint64_t first_site, int nsites, const int counts[6],
                       uint32_t upper, uint32_t* words, float* u32, double* u64, float* n32, double* n64,
                       uint32_t* bounded) {
    for (int i = 0; i < nsites; ++i) {
        orc_stream s;
        stream_init(&s, stream_seed, first_site + (uint64_t)i, alg);
        for (int k = 0; k < counts[0]; ++k) words[(long)i * counts[0] + k] = stream_u32(&s);
        for (int k = 0; k < counts[1]; ++k) u32[(long)i * counts[1] + k] = uniform_f32(&s);
        for (int k = 0; k < counts[2]; ++k) u64[(long)i * counts[2] + k] = uniform_f64(&s);
        for (int k = 0; k < counts[3]; k += 2) normal_pair_f32(&s, &n32[(long)i * counts[3] + k], &n32[(long)i * counts[3] + k + 1]);
        for (int k = 0; k < counts[4]; k += 2) normal_pair_f64(&s, &n64[(long)i * counts[4] + k], &n64[(long)i * counts[4] + k + 1]);
        for (int k = 0; k < counts[5]; ++k) bounded[(long)i * counts[5] + k] = bounded_u32(&s, upper);
    }
}

/* randomize_matrix! (gaussian = 0) and randomize_gaussian_matrix! (gaussian =
 * 1) over the core of a D-dimensional field, no halo: out is [nc1, nc2, L1,
 * ..., LD] column-major of element type elem, the local box starts at
 * global coordinate offset[] of a lattice of extents gdims[] (global site ids
 * column-major, global_site_id, src/LatticeMatrices.jl:334-345).  Per site:
 * kernel_randomize_global_sites! / _uniform_site_fill! (linearalgebra_D.jl:
 * 2673-2716) and kernel_randomize_gaussian_global_sites! /
 * _gaussian_site_fill! (:2718-2775). */
void orc_randomize_fill(void* out, int elem, int nc1, int nc2, int D, const int* L, const int* offset,
                        const int* gdims, int gaussian, double sigma, uint64_t stream_seed, int alg) {
    long V = 1;
    for (int d = 0; d < D; ++d) V *= L[d];
    const int ne = nc1 * nc2;
    const float sig32 = (float)sigma;
    for (long site = 0; site < V; ++site) {
        long rem = site;
        uint64_t gid = 0, stride = 1;
        for (int d = 0; d < D; ++d) {
            const long x = rem % L[d];
            rem /= L[d];
            gid += (uint64_t)(x + offset[d]) * stride;
            stride *= (uint64_t)gdims[d];
        }
        orc_stream s;
        stream_init(&s, stream_seed, gid, alg);
        int use_spare = 0;
        float spare32 = 0.0f;
        double spare64 = 0.0;
        for (int e = 0; e < ne; ++e) {
            const long o = site * ne + e;
            switch (elem) {
                case ORC_ELEM_F32:
                    if (!gaussian) {
                        ((float*)out)[o] = uniform_f32(&s) - 0.5f;
                    } else {
                        float z;
                        if (use_spare) z = spare32;
                        else normal_pair_f32(&s, &z, &spare32);
                        ((float*)out)[o] = sig32 * z;
                        use_spare = !use_spare;
                    }
                    break;
                case ORC_ELEM_F64:
                    if (!gaussian) {
                        ((double*)out)[o] = uniform_f64(&s) - 0.5;
                    } else {
                        double z;
                        if (use_spare) z = spare64;
                        else normal_pair_f64(&s, &z, &spare64);
                        ((double*)out)[o] = sigma * z;
                        use_spare = !use_spare;
                    }
                    break;
                case ORC_ELEM_C64: {
                    float re, im;
                    if (!gaussian) {
                        re = uniform_f32(&s);
                        im = uniform_f32(&s);
                        re -= 0.5f;
                        im -= 0.5f;
                    } else {
                        normal_pair_f32(&s, &re, &im);
                        re *= sig32;
                        im *= sig32;
                    }
                    ((float*)out)[2 * o] = re;
                    ((float*)out)[2 * o + 1] = im;
                    break;
                }
                default: {
                    double re, im;
                    if (!gaussian) {
                        re = uniform_f64(&s);
                        im = uniform_f64(&s);
                        re -= 0.5;
                        im -= 0.5;
                    } else {
                        normal_pair_f64(&s, &re, &im);
                        re *= sigma;
                        im *= sigma;
                    }
                    ((double*)out)[2 * o] = re;
                    ((double*)out)[2 * o + 1] = im;
                    break;
                }
            }
        }
    }
}
