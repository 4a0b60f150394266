// rng.hip -- site random-number streams and global-site lattice fills
// (src/utilities/randomgenerator.jl, src/LinearAlgebras/linearalgebra_D.jl:2550-2775).
//
// Every zero-based global site id (x fastest) owns one stream, so a field is
// the same for any split of the lattice into boxes.  One thread per site draws
// that site's elements in order: the streams are sequential within a site and
// independent across sites, and a site's 12 complex spinor entries cost 12
// Philox blocks plus 12 Box-Muller pairs, so the fill is ALU bound (FP64 log /
// sincos) with a single coalesced write of each element.
//
// Streams (randomgenerator.jl):
//   PCG32 (XSH-RR of the old state, :7-22), seeded from the site seed :158-167
//   xoshiro256++ (:224-245), four SplitMix64 words of the site seed :169-182
//   Philox4x32-10 (:247-319): counter (block lo, block hi, site lo, site hi),
//     key = SplitMix64(stream seed ^ salt) shared by all sites :184-200, four
//     output words handed out in order before the next block.
// Uniforms (:341-368), Box-Muller normals (:380-386), bounded integers by
// rejection (:406-414).  Integer and uniform results are bit-identical to the
// reference on every backend; normals use the device log / sqrt / sincos,
// which may differ from the host library's in the last bit (the reference
// documents the same for its own GPU backends, :371-376).
#include "lmx_internal.h"

namespace lmx {

__host__ __device__ __forceinline__ uint64_t splitmix64(uint64_t x) {
    uint64_t z = x + 0x9e3779b97f4a7c15ULL;
    z = (z ^ (z >> 30)) * 0xbf58476d1ce4e5b9ULL;
    z = (z ^ (z >> 27)) * 0x94d049bb133111ebULL;
    return z ^ (z >> 31);
}

// _mix_site_rng_tag, randomgenerator.jl:129-131
__host__ __device__ __forceinline__ uint64_t mix_tag(uint64_t h, uint64_t tag, uint64_t salt) {
    return splitmix64(h ^ splitmix64(tag + salt));
}

// _rng_stream_seed, :133-140 (once per field, on the host)
static uint64_t stream_seed(const lmx_rng_key& k) {
    uint64_t h = splitmix64(k.seed ^ 0x243f6a8885a308d3ULL);
    h = mix_tag(h, k.sweep, 0xa4093822299f31d0ULL);
    h = mix_tag(h, (uint64_t)k.direction, 0x082efa98ec4e6c89ULL);
    h = mix_tag(h, (uint64_t)k.color, 0x452821e638d01377ULL);
    h = mix_tag(h, (uint64_t)k.subgroup, 0xbe5466cf34e90c6cULL);
    return h;
}

// Per-field stream parameters: the stream seed (PCG32 / xoshiro derive each
// site's state from it) and the Philox key.
struct FieldSeed {
    uint64_t h;
    uint64_t philox_key;
};

static FieldSeed field_seed(const lmx_rng_key& k) {
    FieldSeed f;
    f.h = stream_seed(k);
    f.philox_key = splitmix64(f.h ^ 0x9216d5d98979fb1bULL);
    return f;
}

struct PcgStream {
    uint64_t state, inc;
    __device__ __forceinline__ PcgStream(const FieldSeed& f, uint64_t site) {
        const uint64_t h = mix_tag(f.h, site, 0x13198a2e03707344ULL);
        state = splitmix64(h ^ 0xc0ac29b7c97c50ddULL);
        inc = splitmix64(h ^ 0x3f84d5b5b5470917ULL) | 1ULL;
    }
    __device__ __forceinline__ uint32_t u32() {
        const uint64_t old = state;
        state = old * 0x5851F42D4C957F2DULL + (inc | 1ULL);
        const uint32_t xs = (uint32_t)(((old >> 18) ^ old) >> 27);
        const uint32_t rot = (uint32_t)(old >> 59);
        return (xs >> rot) | (xs << ((32u - rot) & 31u));
    }
    __device__ __forceinline__ uint64_t u64() {
        const uint64_t hi = u32();
        return (hi << 32) | u32();
    }
};

struct XoshiroStream {
    uint64_t s0, s1, s2, s3;
    __device__ __forceinline__ XoshiroStream(const FieldSeed& f, uint64_t site) {
        const uint64_t h = mix_tag(f.h, site, 0x13198a2e03707344ULL);
        const uint64_t g = 0x9e3779b97f4a7c15ULL;
        s0 = splitmix64(h);
        s1 = splitmix64(h + g);
        s2 = splitmix64(h + g + g);
        s3 = splitmix64(h + g + g + g);
    }
    static __device__ __forceinline__ uint64_t rotl(uint64_t v, int k) { return (v << k) | (v >> (64 - k)); }
    __device__ __forceinline__ uint64_t u64() {
        const uint64_t result = rotl(s0 + s3, 23) + s0;
        const uint64_t t = s1 << 17;
        s2 ^= s0;
        s3 ^= s1;
        s1 ^= s2;
        s0 ^= s3;
        s2 ^= t;
        s3 = rotl(s3, 45);
        return result;
    }
    __device__ __forceinline__ uint32_t u32() { return (uint32_t)(u64() >> 32); }
};

// The four words of the current block are kept in w0..w3 and shifted down as
// they are consumed (no indexed register access); left counts the words not
// yet handed out.
struct PhiloxStream {
    uint64_t block;
    uint32_t c2, c3, k0, k1;
    uint32_t w0, w1, w2, w3;
    int left;
    __device__ __forceinline__ PhiloxStream(const FieldSeed& f, uint64_t site) {
        block = 0;
        c2 = (uint32_t)site;
        c3 = (uint32_t)(site >> 32);
        k0 = (uint32_t)f.philox_key;
        k1 = (uint32_t)(f.philox_key >> 32);
        w0 = w1 = w2 = w3 = 0;
        left = 0;
    }
    __device__ __forceinline__ void next_block() {
        uint32_t x0 = (uint32_t)block, x1 = (uint32_t)(block >> 32), x2 = c2, x3 = c3, a = k0, b = k1;
#pragma unroll
        for (int r = 0; r < 10; ++r) {
            const uint64_t p0 = (uint64_t)0xd2511f53u * x0;
            const uint64_t p1 = (uint64_t)0xcd9e8d57u * x2;
            const uint32_t n0 = (uint32_t)(p1 >> 32) ^ x1 ^ a;
            const uint32_t n2 = (uint32_t)(p0 >> 32) ^ x3 ^ b;
            x1 = (uint32_t)p1;
            x3 = (uint32_t)p0;
            x0 = n0;
            x2 = n2;
            a += 0x9e3779b9u;
            b += 0xbb67ae85u;
        }
        w0 = x0;
        w1 = x1;
        w2 = x2;
        w3 = x3;
        block += 1;
        left = 4;
    }
    __device__ __forceinline__ uint32_t u32() {
        if (left == 0) next_block();
        const uint32_t v = w0;
        w0 = w1;
        w1 = w2;
        w2 = w3;
        --left;
        return v;
    }
    __device__ __forceinline__ uint64_t u64() {
        const uint64_t hi = u32();
        return (hi << 32) | u32();
    }
};

template <class S>
__device__ __forceinline__ float uniform_f32(S& s) {
    return (float)(s.u32() >> 8) * 5.9604645e-8f;
}
template <class S>
__device__ __forceinline__ double uniform_f64(S& s) {
    return (double)(s.u64() >> 11) * 1.1102230246251565e-16;
}
template <class T, class S>
__device__ __forceinline__ T uniform(S& s) {
    if constexpr (sizeof(T) == 4) return uniform_f32(s);
    else return uniform_f64(s);
}

template <class T, class S>
__device__ __forceinline__ void normal_pair(S& s, T& z0, T& z1) {
    if constexpr (sizeof(T) == 4) {
        float u = uniform_f32(s);
        u = u == 0.0f ? 2.9802322e-8f : u;
        const float v = uniform_f32(s);
        const float radius = sqrtf(-2.0f * logf(u));
        const float angle = (float)6.283185307179586 * v;
        float sn, cs;
        sincosf(angle, &sn, &cs);
        z0 = radius * cs;
        z1 = radius * sn;
    } else {
        double u = uniform_f64(s);
        u = u == 0.0 ? 5.551115123125783e-17 : u;
        const double v = uniform_f64(s);
        const double radius = sqrt(-2.0 * log(u));
        const double angle = 6.283185307179586 * v;
        double sn, cs;
        sincos(angle, &sn, &cs);
        z0 = radius * cs;
        z1 = radius * sn;
    }
}

template <class S>
__device__ __forceinline__ uint32_t bounded(S& s, uint32_t upper) {
    const uint32_t threshold = (0u - upper) % upper;
    for (;;) {
        const uint32_t v = s.u32();
        if (v >= threshold) return v % upper;
    }
}

// ---------------------------------------------------------------------------
// The per-site draw sequence of test/site_rng_gpu.jl:7-42
// ---------------------------------------------------------------------------
struct DrawOut {
    uint32_t* words;
    float* u32;
    double* u64;
    float* n32;
    double* n64;
    uint32_t* bnd;
    int n[6];
    uint32_t upper;
};

template <class S>
__global__ void __launch_bounds__(256) k_site_rng_draw(FieldSeed f, uint64_t first_site, int nsites, DrawOut o) {
    const int i = blockIdx.x * 256 + threadIdx.x;
    if (i >= nsites) return;
    S s(f, first_site + (uint64_t)i);
    const long b = i;
    for (int k = 0; k < o.n[0]; ++k) o.words[b * o.n[0] + k] = s.u32();
    for (int k = 0; k < o.n[1]; ++k) o.u32[b * o.n[1] + k] = uniform_f32(s);
    for (int k = 0; k < o.n[2]; ++k) o.u64[b * o.n[2] + k] = uniform_f64(s);
    for (int k = 0; k < o.n[3]; k += 2) normal_pair<float>(s, o.n32[b * o.n[3] + k], o.n32[b * o.n[3] + k + 1]);
    for (int k = 0; k < o.n[4]; k += 2) normal_pair<double>(s, o.n64[b * o.n[4] + k], o.n64[b * o.n[4] + k + 1]);
    for (int k = 0; k < o.n[5]; ++k) o.bnd[b * o.n[5] + k] = bounded(s, o.upper);
}

// ---------------------------------------------------------------------------
// Fills
// ---------------------------------------------------------------------------
// Site fill in the reference's element order (ic fastest, then jc):
// _uniform_site_fill! (linearalgebra_D.jl:2689-2716) and _gaussian_site_fill!
// (:2737-2775).  emit(e, re, im) receives element e (im = 0 for real fields).
template <class T, bool CPLX, bool GAUSS, class S, class F>
__device__ __forceinline__ void site_fill(S& s, int ne, T sigma, F&& emit) {
    if constexpr (!GAUSS) {
        for (int e = 0; e < ne; ++e) {
            if constexpr (CPLX) {
                const T re = uniform<T>(s);
                const T im = uniform<T>(s);
                emit(e, re - T(0.5), im - T(0.5));
            } else {
                emit(e, uniform<T>(s) - T(0.5), T(0));
            }
        }
    } else if constexpr (CPLX) {
        for (int e = 0; e < ne; ++e) {
            T z0, z1;
            normal_pair<T>(s, z0, z1);
            emit(e, sigma * z0, sigma * z1);
        }
    } else {
        for (int e = 0; e < ne; e += 2) {
            T z0, z1;
            normal_pair<T>(s, z0, z1);
            emit(e, sigma * z0, T(0));
            if (e + 1 < ne) emit(e + 1, sigma * z1, T(0));
        }
    }
}

// Placement of the local box in the global lattice and of the core in the
// padded array.  With halo != 0 the grid covers the padded volume and each
// ghost site regenerates its wrapped core site and applies the boundary
// phases the way exchange_dim_local! does (src/LatticeMatrices_core.jl:707-723):
// plus ghost = core * phase, minus ghost = core * inv(phase), direction by
// direction, skipped where the phase is exactly one.
struct FillGeom {
    int L[4];
    int P[4];     // padded extents L + 2 nw
    int nw;
    int halo;     // fill the halo too (the local box is the whole lattice)
    long Vp;      // threads: padded volume (halo) or core volume
    uint64_t goff[4];
    uint64_t gstride[4];
    HaloPhases ph;
};

template <class T>
struct Elem2 {
    T re, im;
};

template <class S, class T, bool CPLX, bool GAUSS>
__global__ void __launch_bounds__(256) k_rng_fill_padded(void* __restrict__ out, FieldSeed f, FillGeom g, int nc1,
                                                          int nc2, double sigma) {
    const long t = (long)blockIdx.x * 256 + threadIdx.x;
    if (t >= g.Vp) return;
    int x[4], w[4];
    long ps = 0;
    if (g.halo) {
        long rem = t;
        for (int d = 0; d < 4; ++d) {
            int c = (int)(rem % g.P[d]) - g.nw;
            rem /= g.P[d];
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
        ps = t;
    } else {
        long rem = t;
        for (int d = 0; d < 4; ++d) {
            x[d] = (int)(rem % g.L[d]);
            rem /= g.L[d];
            w[d] = 0;
        }
        ps = (long)(x[0] + g.nw) +
             (long)g.P[0] * ((x[1] + g.nw) + (long)g.P[1] * ((x[2] + g.nw) + (long)g.P[2] * (x[3] + g.nw)));
    }
    uint64_t gid = 0;
    for (int d = 0; d < 4; ++d) gid += ((uint64_t)x[d] + g.goff[d]) * g.gstride[d];
    S s(f, gid);
    const int ne = nc1 * nc2;
    const T sig = (T)sigma;
    site_fill<T, CPLX, GAUSS>(s, ne, sig, [&](int e, T re, T im) {
        for (int d = 0; d < 4; ++d) {
            if (w[d] == 0 || g.ph.skip[d]) continue;
            const T pr = (T)(w[d] > 0 ? g.ph.plus_re[d] : g.ph.minus_re[d]);
            if constexpr (CPLX) {
                const T pi = (T)(w[d] > 0 ? g.ph.plus_im[d] : g.ph.minus_im[d]);
                const T nr = re * pr - im * pi;
                const T ni = re * pi + im * pr;
                re = nr;
                im = ni;
            } else {
                re = re * pr;
            }
        }
        if constexpr (CPLX) {
            Elem2<T> v;
            v.re = re;
            v.im = im;
            ((Elem2<T>*)out)[ps * ne + e] = v;
        } else {
            ((T*)out)[ps * ne + e] = re;
        }
    });
}

// Straight into the library's F64 spinor layout (component k = colour + 3
// spin, site-major planes): the Gaussian or uniform (3,4) ComplexF64 fill of
// one right-hand side.
template <class S, bool GAUSS>
__global__ void __launch_bounds__(256) k_rng_fill_spinor(double2* __restrict__ out, FieldSeed f, Geom g,
                                                          FillGeom fg, double sigma) {
    const int site = blockIdx.x * 256 + threadIdx.x;
    if (site >= g.V) return;
    int x[4];
    site_coords(site, g, x);
    uint64_t gid = 0;
    for (int d = 0; d < 4; ++d) gid += ((uint64_t)x[d] + fg.goff[d]) * fg.gstride[d];
    S s(f, gid);
    site_fill<double, true, GAUSS>(s, 12, sigma, [&](int e, double re, double im) {
        st_out(out + (long)e * g.V + site, re, im);
    });
}

// ---------------------------------------------------------------------------
// Launchers
// ---------------------------------------------------------------------------
static int check_key_alg(const lmx_rng_key* key, int algorithm) {
    LMX_CHECK(key != nullptr, LMX_ERR_ARG, "RNG stream key is NULL");
    LMX_CHECK(algorithm == LMX_RNG_PCG32 || algorithm == LMX_RNG_XOSHIRO256PP || algorithm == LMX_RNG_PHILOX4X32,
              LMX_ERR_ARG, "unknown site RNG algorithm %d", algorithm);
    return LMX_OK;
}

int launch_site_rng_draw(lmx_context* ctx, const lmx_rng_key* key, int algorithm, uint64_t first_site, int nsites,
                         const lmx_rng_draws* n, uint32_t* words, float* u32, double* u64, float* n32, double* n64,
                         uint32_t* bnd) {
    LMX_TRY(check_key_alg(key, algorithm));
    LMX_CHECK(n != nullptr && nsites >= 0, LMX_ERR_ARG, "bad site_rng_draw arguments");
    const int32_t c[6] = {n->words, n->uniform32, n->uniform64, n->normal32, n->normal64, n->bounded};
    void* const ptrs[6] = {words, u32, u64, n32, n64, bnd};
    for (int k = 0; k < 6; ++k) {
        LMX_CHECK(c[k] >= 0, LMX_ERR_ARG, "draw counts must be non-negative");
        LMX_CHECK(c[k] == 0 || ptrs[k] != nullptr, LMX_ERR_ARG, "output %d is NULL with a nonzero count", k);
    }
    LMX_CHECK(c[3] % 2 == 0 && c[4] % 2 == 0, LMX_ERR_ARG, "normal draws come in Box-Muller pairs (even counts)");
    LMX_CHECK(c[5] == 0 || n->upper != 0, LMX_ERR_ARG, "rand_bounded needs a nonzero upper bound");
    if (nsites == 0) return LMX_OK;
    DrawOut o{words, u32, u64, n32, n64, bnd, {c[0], c[1], c[2], c[3], c[4], c[5]}, n->upper};
    const FieldSeed f = field_seed(*key);
    const dim3 grid((unsigned)((nsites + 255) / 256));
    if (algorithm == LMX_RNG_PCG32)
        hipLaunchKernelGGL(k_site_rng_draw<PcgStream>, grid, dim3(256), 0, ctx->stream, f, first_site, nsites, o);
    else if (algorithm == LMX_RNG_XOSHIRO256PP)
        hipLaunchKernelGGL(k_site_rng_draw<XoshiroStream>, grid, dim3(256), 0, ctx->stream, f, first_site, nsites, o);
    else
        hipLaunchKernelGGL(k_site_rng_draw<PhiloxStream>, grid, dim3(256), 0, ctx->stream, f, first_site, nsites, o);
    LMX_HIP(hipGetLastError());
    return LMX_OK;
}

// Global placement of the context's lattice; NULL offset / dims = the whole lattice.
static int make_fill_geom(const lmx_context* ctx, int nw, const int32_t* goff, const int32_t* gdims, FillGeom& g) {
    LMX_CHECK(nw >= 0, LMX_ERR_ARG, "nw must be non-negative, got %d", nw);
    LMX_CHECK((goff == nullptr) == (gdims == nullptr), LMX_ERR_ARG,
              "global_offset and global_dims are given together or not at all");
    bool whole = true;
    uint64_t st = 1;
    for (int d = 0; d < 4; ++d) {
        const int L = ctx->g.L[d];
        LMX_CHECK(nw <= L, LMX_ERR_ARG, "halo width nw=%d exceeds local lattice size %d in dimension %d", nw, L, d + 1);
        const int64_t o = goff ? goff[d] : 0, G = gdims ? gdims[d] : L;
        LMX_CHECK(o >= 0 && G >= 1 && o + L <= G, LMX_ERR_ARG,
                  "local box [%lld, %lld) does not fit the global extent %lld in dimension %d", (long long)o,
                  (long long)(o + L), (long long)G, d + 1);
        if (o != 0 || G != L) whole = false;
        g.L[d] = L;
        g.P[d] = L + 2 * nw;
        g.goff[d] = (uint64_t)o;
        g.gstride[d] = st;
        st *= (uint64_t)G;
    }
    g.nw = nw;
    g.halo = (whole && nw > 0) ? 1 : 0;
    long v = 1;
    for (int d = 0; d < 4; ++d) v *= g.halo ? g.P[d] : g.L[d];
    g.Vp = v;
    return LMX_OK;
}

template <class S, class T, bool CPLX>
static void launch_padded_t(lmx_context* ctx, void* out, const FieldSeed& f, const FillGeom& g, int nc1, int nc2,
                            int gauss, double sigma) {
    const dim3 grid((unsigned)((g.Vp + 255) / 256));
    if (gauss)
        hipLaunchKernelGGL((k_rng_fill_padded<S, T, CPLX, true>), grid, dim3(256), 0, ctx->stream, out, f, g, nc1, nc2,
                           sigma);
    else
        hipLaunchKernelGGL((k_rng_fill_padded<S, T, CPLX, false>), grid, dim3(256), 0, ctx->stream, out, f, g, nc1,
                           nc2, sigma);
}

template <class S>
static void launch_padded_s(lmx_context* ctx, void* out, const FieldSeed& f, const FillGeom& g, int elem, int nc1,
                            int nc2, int gauss, double sigma) {
    switch (elem) {
        case LMX_ELEM_F32: launch_padded_t<S, float, false>(ctx, out, f, g, nc1, nc2, gauss, sigma); break;
        case LMX_ELEM_F64: launch_padded_t<S, double, false>(ctx, out, f, g, nc1, nc2, gauss, sigma); break;
        case LMX_ELEM_C64: launch_padded_t<S, float, true>(ctx, out, f, g, nc1, nc2, gauss, sigma); break;
        default: launch_padded_t<S, double, true>(ctx, out, f, g, nc1, nc2, gauss, sigma); break;
    }
}

static int check_fill(int dist, double sigma) {
    LMX_CHECK(dist == LMX_FILL_UNIFORM || dist == LMX_FILL_GAUSSIAN, LMX_ERR_ARG, "unknown fill distribution %d", dist);
    LMX_CHECK(dist == LMX_FILL_UNIFORM || sigma >= 0.0, LMX_ERR_ARG, "sigma must be non-negative");
    return LMX_OK;
}

int launch_randomize_padded(lmx_context* ctx, void* C, int elem, int nc1, int nc2, int nw, int dist,
                            const lmx_rng_key* key, double sigma, int algorithm, const int32_t* goff,
                            const int32_t* gdims) {
    LMX_TRY(check_key_alg(key, algorithm));
    LMX_TRY(check_fill(dist, sigma));
    LMX_CHECK(elem >= LMX_ELEM_F32 && elem <= LMX_ELEM_C128, LMX_ERR_ARG,
              "random lattice fills support Float32, Float64, ComplexF32 and ComplexF64 elements (got code %d)", elem);
    LMX_CHECK(nc1 >= 1 && nc2 >= 1 && nc1 * nc2 <= 4096, LMX_ERR_ARG, "bad matrix shape %d x %d", nc1, nc2);
    LMX_CHECK(C != nullptr, LMX_ERR_ARG, "field array is NULL");
    FillGeom g;
    LMX_TRY(make_fill_geom(ctx, nw, goff, gdims, g));
    if (g.halo) {
        LMX_TRY(make_halo_phases(ctx, elem, &g.ph));
    } else {
        for (int d = 0; d < 4; ++d) {
            g.ph.skip[d] = 1;
            g.ph.plus_re[d] = g.ph.minus_re[d] = 1.0;
            g.ph.plus_im[d] = g.ph.minus_im[d] = 0.0;
        }
    }
    if (g.Vp == 0) return LMX_OK;
    const FieldSeed f = field_seed(*key);
    const int gauss = dist == LMX_FILL_GAUSSIAN;
    if (algorithm == LMX_RNG_PCG32) launch_padded_s<PcgStream>(ctx, C, f, g, elem, nc1, nc2, gauss, sigma);
    else if (algorithm == LMX_RNG_XOSHIRO256PP) launch_padded_s<XoshiroStream>(ctx, C, f, g, elem, nc1, nc2, gauss, sigma);
    else launch_padded_s<PhiloxStream>(ctx, C, f, g, elem, nc1, nc2, gauss, sigma);
    LMX_HIP(hipGetLastError());
    return LMX_OK;
}

template <class S>
static void launch_spinor_s(lmx_context* ctx, double2* out, const FieldSeed& f, const FillGeom& g, int gauss,
                            double sigma) {
    const dim3 grid(nblocks(ctx->V));
    if (gauss)
        hipLaunchKernelGGL((k_rng_fill_spinor<S, true>), grid, dim3(256), 0, ctx->stream, out, f, ctx->g, g, sigma);
    else
        hipLaunchKernelGGL((k_rng_fill_spinor<S, false>), grid, dim3(256), 0, ctx->stream, out, f, ctx->g, g, sigma);
}

int launch_spinor_randomize(lmx_context* ctx, void* dst, int nrhs, int dist, const lmx_rng_key* keys, double sigma,
                            int algorithm, const int32_t* goff, const int32_t* gdims) {
    LMX_CHECK(dst != nullptr && keys != nullptr && nrhs >= 0, LMX_ERR_ARG, "bad spinor_randomize arguments");
    LMX_TRY(check_fill(dist, sigma));
    for (int r = 0; r < nrhs; ++r) LMX_TRY(check_key_alg(&keys[r], algorithm));
    FillGeom g;
    LMX_TRY(make_fill_geom(ctx, 0, goff, gdims, g));
    if (ctx->V == 0) return LMX_OK;
    const int gauss = dist == LMX_FILL_GAUSSIAN;
    for (int r = 0; r < nrhs; ++r) {
        const FieldSeed f = field_seed(keys[r]);
        double2* out = (double2*)dst + (long)r * 12 * ctx->V;
        if (algorithm == LMX_RNG_PCG32) launch_spinor_s<PcgStream>(ctx, out, f, g, gauss, sigma);
        else if (algorithm == LMX_RNG_XOSHIRO256PP) launch_spinor_s<XoshiroStream>(ctx, out, f, g, gauss, sigma);
        else launch_spinor_s<PhiloxStream>(ctx, out, f, g, gauss, sigma);
        LMX_HIP(hipGetLastError());
    }
    return LMX_OK;
}

}  // namespace lmx
