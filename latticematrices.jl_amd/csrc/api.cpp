// api.cpp -- the extern "C" boundary (include/lmx.h) and the host drivers of
// the FP64 CG, the force and the operator objects.
#include <algorithm>
#include <cmath>
#include <cstring>
#include <mutex>
#include <new>

#include <cstdlib>

#include "lmx_internal.h"

namespace lmx {

static thread_local std::string g_last_error;

void set_error(const char* fmt, ...) {
    char buf[1024];
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(buf, sizeof(buf), fmt, ap);
    va_end(ap);
    g_last_error = buf;
}
void clear_error() { g_last_error.clear(); }

int spinor_import(lmx_context* ctx, void* dst, const void* src, int nw, int nrhs);
int spinor_export(lmx_context* ctx, void* dst, int nw, const void* src, int nrhs);
int spinor_import_f32(lmx_context* ctx, void* dst, const void* src, int nw, int nrhs);
int spinor_export_f32(lmx_context* ctx, void* dst, int nw, const void* src, int nrhs);
int links_import(lmx_context* ctx, void* dst, const void* const src[4], int nw);
int links_export(lmx_context* ctx, void* const dst[4], int nw, const void* src);

static int check_ctx(const lmx_context* ctx) {
    LMX_CHECK(ctx != nullptr, LMX_ERR_ARG, "context is NULL");
    return LMX_OK;
}
static int check_op(const lmx_wilson* op) {
    LMX_CHECK(op != nullptr, LMX_ERR_ARG, "operator is NULL");
    LMX_CHECK(op->ctx != nullptr, LMX_ERR_ARG, "operator has no context");
    return LMX_OK;
}

// Entry points built for NC = 3 fields only (the low-precision and mixed paths,
// the TA projection, the spinor fills).
static int check_nc3(const lmx_context* ctx, const char* what) {
    LMX_CHECK(ctx->nc == 3, LMX_ERR_UNSUPPORTED, "%s is built for NC = 3 fields (this context has NC = %d)", what,
              ctx->nc);
    return LMX_OK;
}
static int check_nw_fits(const lmx_context* ctx, int nw) {
    for (int d = 0; d < 4; ++d)
        LMX_CHECK(nw <= ctx->g.L[d], LMX_ERR_ARG, "halo width nw=%d exceeds local lattice size %d in dimension %d", nw,
                  ctx->g.L[d], d + 1);
    return LMX_OK;
}

// _validate_cg_arguments, src/LinearAlgebras/cg.jl:33-47
static int validate_cg(const void* const* fields, const char* const* names, int n, const lmx_cg_params* p) {
    for (int i = 0; i < n; ++i) LMX_CHECK(fields[i] != nullptr, LMX_ERR_ARG, "CG field %s is NULL", names[i]);
    for (int i = 0; i < n; ++i)
        for (int j = i + 1; j < n; ++j)
            LMX_CHECK(fields[i] != fields[j], LMX_ERR_ALIAS, "CG fields %s and %s must not alias", names[i], names[j]);
    LMX_CHECK(p != nullptr, LMX_ERR_ARG, "params is NULL");
    LMX_CHECK(p->rtol >= 0.0, LMX_ERR_ARG, "rtol must be nonnegative");
    LMX_CHECK(p->atol >= 0.0, LMX_ERR_ARG, "atol must be nonnegative");
    LMX_CHECK(std::isfinite(p->rtol), LMX_ERR_ARG, "rtol must be finite");
    LMX_CHECK(std::isfinite(p->atol), LMX_ERR_ARG, "atol must be finite");
    LMX_CHECK(p->maxiter >= 0, LMX_ERR_ARG, "maxiter must be nonnegative");
    return LMX_OK;
}

static double rel_res(double rn, double bn) {  // _cg_relative_residual, cg.jl:17-23
    if (bn == 0.0) return rn == 0.0 ? 0.0 : INFINITY;
    return rn / bn;
}

static int read_state(lmx_context* ctx, CGState* dev, CGState* host) {
    LMX_HIP(hipMemcpyAsync(host, dev, sizeof(CGState), hipMemcpyDeviceToHost, ctx->stream));
    LMX_HIP(hipStreamSynchronize(ctx->stream));
    return LMX_OK;
}

static void fill_result(const CGState& s, lmx_cg_result* r) {
    r->converged = s.converged;
    r->iterations = s.iteration;
    r->reason = s.reason;
    r->reliable_updates = 0;
    r->residual_norm = s.rnorm;
    if (std::isnan(s.bnorm))
        r->relative_residual = NAN;
    else
        r->relative_residual = rel_res(s.rnorm, s.bnorm);
    r->true_relative_residual = -1.0;
}

static int chunk_of(const lmx_cg_params* p) { return p->check_every > 0 ? p->check_every : 16; }

int ensure_lo_links(lmx_wilson* op) {
    if (op->lo_valid) return LMX_OK;
    lmx_context* ctx = op->ctx;
    const size_t nl = 36 * (size_t)ctx->V;
    if (!op->links32) LMX_HIP(hipMalloc(&op->links32, nl * sizeof(float2)));
    if (!op->links16) LMX_HIP(hipMalloc(&op->links16, nl * sizeof(__half2)));
    LMX_TRY(launch_convert_links(ctx, op->links, op->links32, op->links16));
    if (op->clover64) {
        const size_t nc = 27 * (size_t)ctx->V;
        if (!op->clover32) LMX_HIP(hipMalloc(&op->clover32, nc * sizeof(float2)));
        if (!op->clover16) LMX_HIP(hipMalloc(&op->clover16, nc * sizeof(__half2)));
        LMX_TRY(launch_convert_clover(ctx, op->clover64, op->clover32, op->clover16));
    }
    op->lo_valid = true;
    return LMX_OK;
}

// ---------------------------------------------------------------------------
// Allocation-free FP64 CG (lmx_cg_ddagd_noalloc).  Its reduction partials and
// CG state live in `scratch`, a static device array: every GPU the library
// runs on holds its own copy until the process ends, and every call on that
// GPU shares it, so calls that overlap on one GPU race.  The host side never
// uses the array's name as an address: lmx_context_create resolves the
// device address once per device (with that device current) and keeps it.
// ---------------------------------------------------------------------------
constexpr size_t kNoAllocStateBytes = 256;                 // one CGState, padded
constexpr int64_t kNoAllocMaxVolume = int64_t(1) << 24;    // 64^4 sites
constexpr size_t kNoAllocScratchBytes =
    kNoAllocStateBytes + sizeof(double) * 2 * (size_t)(kNoAllocMaxVolume / kRedBlock);  // 1 MiB + 256 B
static_assert(sizeof(CGState) <= kNoAllocStateBytes, "CGState must fit its scratch slot");

}  // namespace lmx

static __device__ __align__(256) unsigned char scratch[lmx::kNoAllocScratchBytes];

namespace lmx {

constexpr int kMaxDevices = 64;
static std::mutex g_scratch_mutex;
static void* g_scratch_ptr[kMaxDevices];

// Device address of this device's `scratch`; `device` must be current.
static int noalloc_scratch_address(int device, void** out) {
    LMX_CHECK(device >= 0 && device < kMaxDevices, LMX_ERR_UNSUPPORTED, "device ordinal %d out of range", device);
    std::lock_guard<std::mutex> lock(g_scratch_mutex);
    if (g_scratch_ptr[device] == nullptr) {
        void* ptr = nullptr;
        hipError_t e = hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch));
        LMX_CHECK(e == hipSuccess && ptr != nullptr, LMX_ERR_HIP, "hipGetSymbolAddress(scratch) failed on device %d: %s",
                  device, hipGetErrorString(e));
        g_scratch_ptr[device] = ptr;
    }
    *out = g_scratch_ptr[device];
    return LMX_OK;
}

// cg!(x, DdagDOp(D, tmp), b, r, p, Ap) in FP64 on the context c, whose
// partials and state buffers the kernels use as their scratch.
static int cg_ddagd_run(lmx_wilson* op, lmx_context* c, void* x, const void* b, void* r, void* p, void* Ap,
                        void* tmp, const lmx_cg_params* params, lmx_cg_result* result) {
    CGState* st = c->state;
    const double2* U = op->links;
    const double2* F = op->clover64;
    double2 *X = (double2*)x, *R = (double2*)r, *P = (double2*)p, *AP = (double2*)Ap, *T = (double2*)tmp;
    const double2* B = (const double2*)b;
    Epilogue gate{0, nullptr, st, nullptr, nullptr};
    Epilogue pap{1, P, st, c->partials, nullptr};
    LMX_TRY(launch_cg_init_state(c, st, params->rtol, params->atol, params->maxiter));
    LMX_TRY(launch_dot(c, B, B, 1, st, nullptr));
    LMX_TRY(launch_wilson_f64(c, U, F, op->kappa, op->csw, T, X, 1, 0, gate));
    LMX_TRY(launch_wilson_f64(c, U, F, op->kappa, op->csw, AP, T, 1, LMX_DAGGER, gate));
    LMX_TRY(launch_cg_setup_r(c, B, AP, R, P, st));
    LMX_TRY(read_state(c, st, c->host_state));
    const int chunk = chunk_of(params);
    const long max_polls = (long)params->maxiter / chunk + 3;
    long polls = 0;
    while (!c->host_state->done) {
        LMX_CHECK(++polls <= max_polls, LMX_ERR_HIP, "CG state did not advance (iteration %d)",
                  c->host_state->iteration);
        for (int k = 0; k < chunk; ++k) {
            LMX_TRY(launch_wilson_f64(c, U, F, op->kappa, op->csw, T, P, 1, 0, gate));
            if (c->nc == 3) {  // p.Ap fused into D^dagger's epilogue
                LMX_TRY(launch_wilson_f64(c, U, F, op->kappa, op->csw, AP, T, 1, LMX_DAGGER, pap));
            } else {
                LMX_TRY(launch_wilson_f64(c, U, F, op->kappa, op->csw, AP, T, 1, LMX_DAGGER, gate));
                LMX_TRY(launch_dot(c, P, AP, FIN_PAP, st, nullptr));
            }
            LMX_TRY(launch_cg_update_xr(c, X, R, P, AP, st));
            LMX_TRY(launch_cg_update_p(c, R, P, st));
        }
        LMX_TRY(read_state(c, st, c->host_state));
    }
    fill_result(*c->host_state, result);
    return LMX_OK;
}

// cg! on ComplexF32 fields (float2 [12][V]): the same step order, stopping
// test and failure reasons as cg_ddagd_run, the vectors and the operator (FP32
// links, and the FP32 clover cache) in FP32 arithmetic, the reductions summed
// over sites in FP64 and the CG scalars held in FP64 on the device.
static int apply_f32_gated(lmx_wilson* op, float2* out, const float2* in, int flags, CGState* st) {
    LoArgs a;
    memset(&a, 0, sizeof(a));
    a.link_prec = LMX_F32;
    a.spinor_prec = LMX_F32;
    a.links = op->links32;
    a.clover = op->clover32;
    a.kappa = op->kappa;
    a.csw = op->csw;
    a.out = out;
    a.in = in;
    a.nrhs = 1;
    a.flags = flags;
    a.epi = 0;
    a.rhs_block = (1 << 4) | 1;
    a.st = st;
    return launch_wilson_lo(op->ctx, a);
}

static int cg_ddagd_run_f32(lmx_wilson* op, lmx_context* c, float2* X, const float2* B, float2* R, float2* P,
                            float2* AP, float2* T, const lmx_cg_params* params, lmx_cg_result* result) {
    CGState* st = c->state;
    LMX_TRY(launch_cg_init_state(c, st, params->rtol, params->atol, params->maxiter));
    LMX_TRY(launch_dot_f32(c, B, B, FIN_BB, st, nullptr));
    LMX_TRY(apply_f32_gated(op, T, X, 0, st));
    LMX_TRY(apply_f32_gated(op, AP, T, LMX_DAGGER, st));
    LMX_TRY(launch_cg_setup_r_f32(c, B, AP, R, P, st));
    LMX_TRY(read_state(c, st, c->host_state));
    const int chunk = chunk_of(params);
    const long max_polls = (long)params->maxiter / chunk + 3;
    long polls = 0;
    while (!c->host_state->done) {
        LMX_CHECK(++polls <= max_polls, LMX_ERR_HIP, "CG state did not advance (iteration %d)",
                  c->host_state->iteration);
        for (int k = 0; k < chunk; ++k) {
            LMX_TRY(apply_f32_gated(op, T, P, 0, st));
            LMX_TRY(apply_f32_gated(op, AP, T, LMX_DAGGER, st));
            LMX_TRY(launch_dot_f32(c, P, AP, FIN_PAP, st, nullptr));
            LMX_TRY(launch_cg_update_xr_f32(c, X, R, P, AP, st));
            LMX_TRY(launch_cg_update_p_f32(c, R, P, st));
        }
        LMX_TRY(read_state(c, st, c->host_state));
    }
    fill_result(*c->host_state, result);
    return LMX_OK;
}

static int refresh_derived(lmx_wilson* op) {
    lmx_context* ctx = op->ctx;
    if (op->csw != 0.0 && ctx->nc != 3) {
        if (!op->clover64) LMX_HIP(hipMalloc(&op->clover64, 6 * (size_t)link_nelem(ctx) * ctx->V * sizeof(double2)));
        LMX_TRY(launch_clover_build_nc(ctx, op->links, op->clover64));
    } else if (op->csw != 0.0) {
        if (!op->clover64) LMX_HIP(hipMalloc(&op->clover64, 27 * (size_t)ctx->V * sizeof(double2)));
        LMX_TRY(launch_clover_build_f64(ctx, op->links, op->clover64));
    }
    op->lo_valid = false;
    return LMX_OK;
}

}  // namespace lmx

using namespace lmx;

extern "C" {

const char* lmx_last_error(void) { return g_last_error.c_str(); }
int lmx_abi_version(void) { return LMX_ABI_VERSION; }

void lmx_cg_params_default(lmx_cg_params* p) {
    if (!p) return;
    std::memset(p, 0, sizeof(*p));
    p->rtol = 1e-10;
    p->atol = 0.0;
    p->maxiter = 5000;
    p->precision = LMX_F64;
    p->reliable_delta = 0.1;
    p->reliable_maxiter = 100;
    p->check_every = 0;
    p->use_graph = 0;
    p->rhs_block = 0;
}

int lmx_context_create(lmx_context** out, const int32_t dims[4], const double* phases, void* stream) {
    return lmx_context_create_nc(out, dims, phases, stream, 3);
}

int lmx_context_create_nc(lmx_context** out, const int32_t dims[4], const double* phases, void* stream, int nc) {
    clear_error();
    LMX_CHECK(out != nullptr && dims != nullptr, LMX_ERR_ARG, "NULL argument");
    *out = nullptr;
    LMX_CHECK(nc >= 1, LMX_ERR_ARG, "NC must be positive, got %d", nc);
    LMX_CHECK(nc == 2 || nc == 3, LMX_ERR_UNSUPPORTED, "fields of NC = %d colours are not built (NC = 2 and 3 are)",
              nc);
    int64_t V = 1;
    for (int d = 0; d < 4; ++d) {
        LMX_CHECK(dims[d] > 0, LMX_ERR_ARG, "global lattice sizes must be positive, got %d in dimension %d", dims[d],
                  d + 1);
        V *= dims[d];
    }
    LMX_CHECK(V < (int64_t(1) << 31) / 16, LMX_ERR_ARG, "lattice volume %lld too large", (long long)V);
    lmx_context* c = new (std::nothrow) lmx_context();
    LMX_CHECK(c != nullptr, LMX_ERR_NOMEM, "host allocation failed");
    Geom& g = c->g;
    int64_t st = 1;
    for (int d = 0; d < 4; ++d) {
        g.L[d] = dims[d];
        g.stride[d] = (int)st;
        st *= dims[d];
        double re = phases ? phases[2 * d] : 1.0, im = phases ? phases[2 * d + 1] : 0.0;
        if (re == 0.0 && im == 0.0) {
            delete c;
            set_error("boundary phases must be nonzero because negative wraps use inv(phase)");
            return LMX_ERR_ARG;
        }
        g.ph_re[d] = re;
        g.ph_im[d] = im;
        const double dd = re * re + im * im;
        g.iph_re[d] = re / dd;
        g.iph_im[d] = -im / dd;
        g.phase_one[d] = (re == 1.0 && im == 0.0) ? 1 : 0;
        g.phase_neg[d] = (re == -1.0 && im == 0.0) ? 1 : 0;
    }
    g.general = 0;
    for (int d = 0; d < 4; ++d)
        if (!g.phase_one[d] && !g.phase_neg[d]) g.general = 1;
    g.V = (int)V;
    c->V = V;
    c->nc = nc;
    c->stream = (hipStream_t)stream;
    const char* tl = getenv("LMX_STREAM_TL");
    const char* zs = getenv("LMX_STREAM_ZS");
    c->tune_tl = tl ? atoi(tl) : 0;
    c->tune_zs = zs ? atoi(zs) : 0;
    hipError_t e = hipGetDevice(&c->device);
    if (e != hipSuccess) {
        delete c;
        set_error("hipGetDevice failed: %s", hipGetErrorString(e));
        return LMX_ERR_HIP;
    }
    {
        int rc = noalloc_scratch_address(c->device, &c->noalloc_scratch);
        if (rc != LMX_OK) {
            delete c;
            return rc;
        }
    }
    c->partials_cap = 2 * (size_t)nblocks(V) * 16 + 1024;
    if (hipMalloc(&c->partials, c->partials_cap * sizeof(double)) != hipSuccess ||
        hipMalloc(&c->counters, kCounters * sizeof(unsigned)) != hipSuccess ||
        hipMalloc(&c->state, kMaxStates * sizeof(CGState)) != hipSuccess ||
        hipMalloc(&c->dot_out, 2 * kMaxStates * sizeof(double)) != hipSuccess ||
        hipHostMalloc(&c->host_state, kMaxStates * sizeof(CGState), hipHostMallocDefault) != hipSuccess ||
        hipHostMalloc(&c->host_dot, 2 * kMaxStates * sizeof(double), hipHostMallocDefault) != hipSuccess) {
        set_error("device allocation of the context scratch failed");
        lmx_context_destroy(c);
        return LMX_ERR_NOMEM;
    }
    if (hipMemset(c->counters, 0, kCounters * sizeof(unsigned)) != hipSuccess ||
        hipMemset(c->state, 0, kMaxStates * sizeof(CGState)) != hipSuccess ||
        hipDeviceSynchronize() != hipSuccess) {
        set_error("context scratch initialisation failed");
        lmx_context_destroy(c);
        return LMX_ERR_HIP;
    }
    *out = c;
    return LMX_OK;
}

int lmx_context_destroy(lmx_context* c) {
    if (!c) return LMX_OK;
    if (c->stream) (void)hipStreamSynchronize(c->stream);
    else (void)hipDeviceSynchronize();
    if (c->partials) (void)hipFree(c->partials);
    if (c->counters) (void)hipFree(c->counters);
    if (c->state) (void)hipFree(c->state);
    if (c->dot_out) (void)hipFree(c->dot_out);
    if (c->host_state) (void)hipHostFree(c->host_state);
    if (c->host_dot) (void)hipHostFree(c->host_dot);
    delete c;
    return LMX_OK;
}

int lmx_context_set_stream(lmx_context* c, void* stream) {
    LMX_TRY(check_ctx(c));
    c->stream = (hipStream_t)stream;
    return LMX_OK;
}

int lmx_context_synchronize(lmx_context* c) {
    LMX_TRY(check_ctx(c));
    LMX_HIP(hipStreamSynchronize(c->stream));
    return LMX_OK;
}

int64_t lmx_context_volume(const lmx_context* c) { return c ? c->V : -1; }

size_t lmx_spinor_bytes(const lmx_context* c, int precision, int nrhs) {
    if (!c) return 0;
    // LMX_H16: 12 16-bit (re, im) pairs plus one float scale per site
    const size_t n = (size_t)spinor_ncomp(c);
    size_t e = precision == LMX_F64 ? n * sizeof(double2) : precision == LMX_F32 ? n * sizeof(float2)
             : precision == LMX_H16 ? n * sizeof(int) + sizeof(float) : n * sizeof(__half2);
    return e * (size_t)c->V * (size_t)std::max(nrhs, 0);
}
size_t lmx_links_bytes(const lmx_context* c, int precision) {
    if (!c) return 0;
    size_t e = precision == LMX_F64 ? sizeof(double2) : precision == LMX_F32 ? sizeof(float2) : sizeof(__half2);  // H16: FP16 links
    return e * 4 * (size_t)link_nelem(c) * (size_t)c->V;
}
size_t lmx_clover_bytes(const lmx_context* c, int precision) {
    if (!c) return 0;
    size_t e = precision == LMX_F64 ? sizeof(double2) : precision == LMX_F32 ? sizeof(float2) : sizeof(__half2);
    // NC = 3: the 27 complex slots of the compact planes; other NC: 6 full planes
    return e * (c->nc == 3 ? 27 : 6 * (size_t)link_nelem(c)) * (size_t)c->V;
}

int lmx_spinor_import(lmx_context* c, void* dst, const void* src, int nw, int nrhs) {
    clear_error();
    LMX_TRY(check_ctx(c));
    LMX_CHECK(dst && src && nrhs >= 0, LMX_ERR_ARG, "bad spinor_import arguments");
    return spinor_import(c, dst, src, nw, nrhs);
}
int lmx_spinor_export(lmx_context* c, void* dst, int nw, const void* src, int nrhs) {
    clear_error();
    LMX_TRY(check_ctx(c));
    LMX_CHECK(dst && src && nrhs >= 0, LMX_ERR_ARG, "bad spinor_export arguments");
    return spinor_export(c, dst, nw, src, nrhs);
}
int lmx_site_rng_draw(lmx_context* c, const lmx_rng_key* key, int algorithm, uint64_t first_site, int32_t nsites,
                      const lmx_rng_draws* n, uint32_t* words, float* uniform32, double* uniform64, float* normal32,
                      double* normal64, uint32_t* bounded) {
    clear_error();
    LMX_TRY(check_ctx(c));
    return launch_site_rng_draw(c, key, algorithm, first_site, nsites, n, words, uniform32, uniform64, normal32,
                                normal64, bounded);
}
int lmx_randomize_padded(lmx_context* c, void* C, int elem, int nc1, int nc2, int nw, int dist,
                         const lmx_rng_key* key, double sigma, int algorithm, const int32_t* global_offset,
                         const int32_t* global_dims) {
    clear_error();
    LMX_TRY(check_ctx(c));
    return launch_randomize_padded(c, C, elem, nc1, nc2, nw, dist, key, sigma, algorithm, global_offset, global_dims);
}
int lmx_spinor_randomize(lmx_context* c, void* dst, int nrhs, int dist, const lmx_rng_key* keys, double sigma,
                         int algorithm, const int32_t* global_offset, const int32_t* global_dims) {
    clear_error();
    LMX_TRY(check_ctx(c));
    LMX_TRY(check_nc3(c, "the spinor fill"));
    return launch_spinor_randomize(c, dst, nrhs, dist, keys, sigma, algorithm, global_offset, global_dims);
}
int lmx_field_pad(lmx_context* c, void* dst, const void* src, int elem, int nc1, int nc2, int nw) {
    clear_error();
    LMX_TRY(check_ctx(c));
    return launch_field_pad(c, dst, src, elem, nc1, nc2, nw);
}
int lmx_field_gather(lmx_context* c, void* dst, const void* src, int elem, int nc1, int nc2, int nw) {
    clear_error();
    LMX_TRY(check_ctx(c));
    return launch_field_unpad(c, dst, src, elem, nc1, nc2, nw);
}
int lmx_links_import(lmx_context* c, void* dst, const void* const src[4], int nw) {
    clear_error();
    LMX_TRY(check_ctx(c));
    LMX_CHECK(dst && src, LMX_ERR_ARG, "bad links_import arguments");
    return links_import(c, dst, src, nw);
}
int lmx_links_export(lmx_context* c, void* const dst[4], int nw, const void* src) {
    clear_error();
    LMX_TRY(check_ctx(c));
    LMX_CHECK(dst && src, LMX_ERR_ARG, "bad links_export arguments");
    return links_export(c, dst, nw, src);
}

int lmx_axpby(lmx_context* c, double a, const void* X, double b, void* Y, int nrhs) {
    clear_error();
    LMX_TRY(check_ctx(c));
    LMX_CHECK(X && Y && nrhs >= 0, LMX_ERR_ARG, "bad axpby arguments");
    return launch_axpby(c, a, (const double2*)X, b, (double2*)Y, spinor_ncomp(c) * c->V * (int64_t)nrhs);
}

int lmx_dot(lmx_context* c, const void* A, const void* B, double out[2]) {
    clear_error();
    LMX_TRY(check_ctx(c));
    LMX_CHECK(A && B && out, LMX_ERR_ARG, "bad dot arguments");
    if (c->V == 0) {
        out[0] = out[1] = 0.0;
        return LMX_OK;
    }
    LMX_TRY(launch_dot(c, (const double2*)A, (const double2*)B, 0, nullptr, c->dot_out));
    LMX_HIP(hipMemcpyAsync(c->host_dot, c->dot_out, 2 * sizeof(double), hipMemcpyDeviceToHost, c->stream));
    LMX_HIP(hipStreamSynchronize(c->stream));
    out[0] = c->host_dot[0];
    out[1] = c->host_dot[1];
    return LMX_OK;
}

int lmx_scale(lmx_context* c, void* out, const void* in, double sr, double si, int nrhs) {
    clear_error();
    LMX_TRY(check_ctx(c));
    LMX_CHECK(out && in && nrhs >= 0, LMX_ERR_ARG, "bad scale arguments");
    return launch_scale(c, (double2*)out, (const double2*)in, sr, si, spinor_ncomp(c) * c->V * (int64_t)nrhs);
}

int lmx_wilson_create(lmx_context* c, lmx_wilson** out, const void* links, double kappa, double csw) {
    clear_error();
    LMX_TRY(check_ctx(c));
    LMX_CHECK(out && links, LMX_ERR_ARG, "bad wilson_create arguments");
    LMX_CHECK(std::isfinite(kappa) && std::isfinite(csw), LMX_ERR_ARG, "kappa and csw must be finite");
    lmx_wilson* op = new (std::nothrow) lmx_wilson();
    LMX_CHECK(op != nullptr, LMX_ERR_NOMEM, "host allocation failed");
    std::memset(op, 0, sizeof(*op));
    op->ctx = c;
    op->links = (const double2*)links;
    op->kappa = kappa;
    op->csw = csw;
    int rc = refresh_derived(op);
    if (rc != LMX_OK) {
        lmx_wilson_destroy(op);
        return rc;
    }
    *out = op;
    return LMX_OK;
}

int lmx_wilson_destroy(lmx_wilson* op) {
    if (!op) return LMX_OK;
    if (op->ctx) (void)hipStreamSynchronize(op->ctx->stream);
    if (op->links32) (void)hipFree(op->links32);
    if (op->links16) (void)hipFree(op->links16);
    if (op->clover64) (void)hipFree(op->clover64);
    if (op->clover32) (void)hipFree(op->clover32);
    if (op->clover16) (void)hipFree(op->clover16);
    if (op->clover_dF) (void)hipFree(op->clover_dF);
    delete op;
    return LMX_OK;
}

int lmx_wilson_set_kappa(lmx_wilson* op, double kappa) {
    LMX_TRY(check_op(op));
    LMX_CHECK(std::isfinite(kappa), LMX_ERR_ARG, "kappa must be finite");
    op->kappa = kappa;
    return LMX_OK;
}

int lmx_wilson_update_links(lmx_wilson* op, const void* links) {
    clear_error();
    LMX_TRY(check_op(op));
    LMX_CHECK(links != nullptr, LMX_ERR_ARG, "links is NULL");
    op->links = (const double2*)links;
    return refresh_derived(op);
}

int lmx_clover_export(lmx_wilson* op, void* F_out) {
    clear_error();
    LMX_TRY(check_op(op));
    LMX_CHECK(F_out != nullptr, LMX_ERR_ARG, "F_out is NULL");
    LMX_CHECK(op->clover64 != nullptr, LMX_ERR_ARG, "operator has no clover term (csw == 0)");
    if (op->ctx->nc != 3) return launch_clover_export_nc(op->ctx, op->clover64, (double2*)F_out);
    return launch_clover_export(op->ctx, op->clover64, (double2*)F_out);
}

int lmx_wilson_apply(lmx_wilson* op, void* out, const void* in, int flags, int nrhs) {
    clear_error();
    LMX_TRY(check_op(op));
    LMX_CHECK(out && in && nrhs >= 0, LMX_ERR_ARG, "bad wilson_apply arguments");
    LMX_CHECK(out != in, LMX_ERR_ALIAS, "output and input of the Wilson operator must not alias");
    LMX_CHECK((flags & ~(LMX_DAGGER | LMX_DONLY)) == 0, LMX_ERR_ARG, "unknown flags %d", flags);
    Epilogue e{0, nullptr, nullptr, nullptr, nullptr};
    return launch_wilson_f64(op->ctx, op->links, (flags & LMX_DONLY) ? nullptr : op->clover64, op->kappa, op->csw,
                             (double2*)out, (const double2*)in, nrhs, flags, e);
}

int lmx_wilson_apply_f32(lmx_wilson* op, void* out, const void* in, int flags, int nrhs) {
    clear_error();
    LMX_TRY(check_op(op));
    LMX_CHECK(out && in && nrhs >= 0, LMX_ERR_ARG, "bad wilson_apply_f32 arguments");
    LMX_TRY(check_nc3(op->ctx, "the single-precision apply"));
    LMX_CHECK(out != in, LMX_ERR_ALIAS, "output and input of the Wilson operator must not alias");
    LMX_CHECK((flags & ~(LMX_DAGGER | LMX_DONLY)) == 0, LMX_ERR_ARG, "unknown flags %d", flags);
    if (nrhs == 0 || op->ctx->V == 0) return LMX_OK;
    LMX_TRY(ensure_lo_links(op));
    LoArgs a;
    memset(&a, 0, sizeof(a));
    a.link_prec = LMX_F32;
    a.spinor_prec = LMX_F32;
    a.links = op->links32;
    a.clover = (flags & LMX_DONLY) ? nullptr : op->clover32;
    a.kappa = op->kappa;
    a.csw = op->csw;
    a.out = out;
    a.in = in;
    a.nrhs = nrhs;
    a.flags = flags;
    a.epi = 0;
    a.rhs_block = (std::min(nrhs, 12) << 4) | 1;  // one RHS per work item, up to 12 per launch
    return launch_wilson_lo(op->ctx, a);
}

int lmx_wilson_apply_padded(lmx_wilson* op, void* C, const void* psi, int nw, int flags, int nrhs) {
    clear_error();
    LMX_TRY(check_op(op));
    LMX_CHECK(C && psi && nrhs >= 0 && nw >= 0, LMX_ERR_ARG, "bad wilson_apply_padded arguments");
    LMX_CHECK(C != psi, LMX_ERR_ALIAS, "output and input of the Wilson operator must not alias");
    LMX_CHECK((flags & ~(LMX_DAGGER | LMX_DONLY)) == 0, LMX_ERR_ARG, "unknown flags %d", flags);
    if (op->ctx->nc != 3) {
        LMX_TRY(check_nw_fits(op->ctx, nw));
        return launch_wilson_nc(op->ctx, op->links, (flags & LMX_DONLY) ? nullptr : op->clover64, op->kappa, op->csw,
                                (double2*)C, (const double2*)psi, nrhs, flags, nw, nullptr, nullptr);
    }
    return launch_wilson_padded(op->ctx, op->links, (flags & LMX_DONLY) ? nullptr : op->clover64, op->kappa, op->csw,
                                (double2*)C, (const double2*)psi, nw, nrhs, flags);
}

int lmx_wilson_apply_padded_f32(lmx_wilson* op, void* C, const void* psi, int nw, int flags, int nrhs) {
    clear_error();
    LMX_TRY(check_op(op));
    LMX_TRY(check_nc3(op->ctx, "the single-precision apply"));
    LMX_TRY(check_op(op));
    LMX_CHECK(C && psi && nrhs >= 0 && nw >= 0, LMX_ERR_ARG, "bad wilson_apply_padded_f32 arguments");
    LMX_CHECK(C != psi, LMX_ERR_ALIAS, "output and input of the Wilson operator must not alias");
    LMX_CHECK((flags & ~(LMX_DAGGER | LMX_DONLY)) == 0, LMX_ERR_ARG, "unknown flags %d", flags);
    if (nrhs == 0 || op->ctx->V == 0) return LMX_OK;
    LMX_TRY(ensure_lo_links(op));
    return launch_wilson_padded_f32(op->ctx, op->links32, (flags & LMX_DONLY) ? nullptr : op->clover32, op->kappa,
                                    op->csw, (float2*)C, (const float2*)psi, nw, nrhs, flags);
}

int lmx_ddagd_apply(lmx_wilson* op, void* out, const void* in, void* tmp, int nrhs) {
    clear_error();
    LMX_TRY(check_op(op));
    LMX_CHECK(out && in && tmp && nrhs >= 0, LMX_ERR_ARG, "bad ddagd_apply arguments");
    LMX_CHECK(tmp != in && tmp != out, LMX_ERR_ALIAS, "DdagDOp temporary field must not alias its input or output");
    LMX_CHECK(out != in, LMX_ERR_ALIAS, "output and input of D^dagger D must not alias");
    Epilogue e{0, nullptr, nullptr, nullptr, nullptr};
    LMX_TRY(launch_wilson_f64(op->ctx, op->links, op->clover64, op->kappa, op->csw, (double2*)tmp,
                              (const double2*)in, nrhs, 0, e));
    return launch_wilson_f64(op->ctx, op->links, op->clover64, op->kappa, op->csw, (double2*)out,
                             (const double2*)tmp, nrhs, LMX_DAGGER, e);
}

int lmx_cg_ddagd(lmx_wilson* op, void* x, const void* b, void* r, void* p, void* Ap, void* tmp,
                 const lmx_cg_params* params, lmx_cg_result* result) {
    clear_error();
    LMX_TRY(check_op(op));
    const void* f[6] = {x, b, r, p, Ap, tmp};
    const char* names[6] = {"x", "b", "r", "p", "Ap", "tmp"};
    LMX_TRY(validate_cg(f, names, 6, params));
    LMX_CHECK(result != nullptr, LMX_ERR_ARG, "result is NULL");
    return cg_ddagd_run(op, op->ctx, x, b, r, p, Ap, tmp, params, result);
}

int lmx_cg_ddagd_f32(lmx_wilson* op, void* x, const void* b, void* r, void* p, void* Ap, void* tmp,
                     const lmx_cg_params* params, lmx_cg_result* result) {
    clear_error();
    LMX_TRY(check_op(op));
    const void* f[6] = {x, b, r, p, Ap, tmp};
    const char* names[6] = {"x", "b", "r", "p", "Ap", "tmp"};
    LMX_TRY(validate_cg(f, names, 6, params));
    LMX_CHECK(result != nullptr, LMX_ERR_ARG, "result is NULL");
    LMX_TRY(check_nc3(op->ctx, "the single-precision CG"));
    LMX_TRY(ensure_lo_links(op));
    return cg_ddagd_run_f32(op, op->ctx, (float2*)x, (const float2*)b, (float2*)r, (float2*)p, (float2*)Ap,
                            (float2*)tmp, params, result);
}

int lmx_cg_ddagd_padded_f32(lmx_wilson* op, void* x, const void* b, int nw, void* work, size_t work_bytes,
                            const lmx_cg_params* params, lmx_cg_result* result) {
    clear_error();
    LMX_TRY(check_op(op));
    const void* f[2] = {x, b};
    const char* names[2] = {"x", "b"};
    LMX_TRY(validate_cg(f, names, 2, params));
    LMX_CHECK(result != nullptr && work != nullptr, LMX_ERR_ARG, "bad cg_ddagd_padded_f32 arguments");
    lmx_context* c = op->ctx;
    LMX_TRY(check_nc3(c, "the single-precision CG"));
    const size_t field = lmx_spinor_bytes(c, LMX_F32, 1);
    LMX_CHECK(work_bytes >= 6 * field, LMX_ERR_WORKSPACE,
              "workspace of %zu bytes is below the %zu of six single-precision spinor fields", work_bytes, 6 * field);
    LMX_TRY(ensure_lo_links(op));
    unsigned char* w = (unsigned char*)work;
    float2 *X = (float2*)w, *B = (float2*)(w + field), *R = (float2*)(w + 2 * field), *P = (float2*)(w + 3 * field),
           *AP = (float2*)(w + 4 * field), *T = (float2*)(w + 5 * field);
    LMX_TRY(spinor_import_f32(c, X, x, nw, 1));
    LMX_TRY(spinor_import_f32(c, B, b, nw, 1));
    LMX_TRY(cg_ddagd_run_f32(op, c, X, B, R, P, AP, T, params, result));
    LMX_TRY(spinor_export_f32(c, x, nw, X, 1));
    LMX_HIP(hipStreamSynchronize(c->stream));
    return LMX_OK;
}

int lmx_cg_ddagd_noalloc(lmx_wilson* op, void* x, const void* b, void* r, void* p, void* Ap, void* tmp,
                         const lmx_cg_params* params, lmx_cg_result* result) {
    clear_error();
    LMX_TRY(check_op(op));
    const void* f[6] = {x, b, r, p, Ap, tmp};
    const char* names[6] = {"x", "b", "r", "p", "Ap", "tmp"};
    LMX_TRY(validate_cg(f, names, 6, params));
    LMX_CHECK(result != nullptr, LMX_ERR_ARG, "result is NULL");
    const lmx_context* c = op->ctx;
    LMX_CHECK(c->V <= kNoAllocMaxVolume, LMX_ERR_ARG,
              "lattice volume %lld exceeds the allocation-free CG's scratch (at most %lld sites)", (long long)c->V,
              (long long)kNoAllocMaxVolume);
    LMX_CHECK(c->noalloc_scratch != nullptr, LMX_ERR_HIP, "no scratch address for device %d", c->device);
    // the same drivers and kernels as lmx_cg_ddagd, with the context's
    // partials and state replaced by this GPU's static scratch
    lmx_context local = *c;
    local.state = (CGState*)c->noalloc_scratch;
    local.partials = (double*)((unsigned char*)c->noalloc_scratch + kNoAllocStateBytes);
    local.partials_cap = (kNoAllocScratchBytes - kNoAllocStateBytes) / sizeof(double);
    local.counters = nullptr;
    local.dot_out = nullptr;
    return cg_ddagd_run(op, &local, x, b, r, p, Ap, tmp, params, result);
}

int lmx_cg_ddagd_padded(lmx_wilson* op, void* x, const void* b, int nw, void* work, size_t work_bytes,
                        const lmx_cg_params* params, lmx_cg_result* result) {
    clear_error();
    LMX_TRY(check_op(op));
    const void* f[2] = {x, b};
    const char* names[2] = {"x", "b"};
    LMX_TRY(validate_cg(f, names, 2, params));
    LMX_CHECK(result != nullptr && work != nullptr, LMX_ERR_ARG, "bad cg_ddagd_padded arguments");
    lmx_context* c = op->ctx;
    const size_t field = lmx_spinor_bytes(c, LMX_F64, 1);
    LMX_CHECK(work_bytes >= 6 * field, LMX_ERR_WORKSPACE, "workspace of %zu bytes is below the %zu of six spinor fields",
              work_bytes, 6 * field);
    unsigned char* w = (unsigned char*)work;
    void *X = w, *B = w + field, *R = w + 2 * field, *P = w + 3 * field, *AP = w + 4 * field, *T = w + 5 * field;
    // the LatticeMatrix arrays -> the library layout; the solve; x back with
    // its halo refilled as set_halo! would (b is only read)
    LMX_TRY(spinor_import(c, X, x, nw, 1));
    LMX_TRY(spinor_import(c, B, b, nw, 1));
    LMX_TRY(cg_ddagd_run(op, c, X, B, R, P, AP, T, params, result));
    LMX_TRY(spinor_export(c, x, nw, X, 1));
    LMX_HIP(hipStreamSynchronize(c->stream));
    return LMX_OK;
}

int lmx_cg_generic(lmx_context* c, lmx_apply_fn A, void* A_user, void* x, const void* b, void* r, void* p, void* Ap,
                   const lmx_cg_params* params, lmx_cg_result* result) {
    clear_error();
    LMX_TRY(check_ctx(c));
    LMX_CHECK(A != nullptr, LMX_ERR_ARG, "operator callback is NULL");
    const void* f[5] = {x, b, r, p, Ap};
    const char* names[5] = {"x", "b", "r", "p", "Ap"};
    LMX_TRY(validate_cg(f, names, 5, params));
    LMX_CHECK(result != nullptr, LMX_ERR_ARG, "result is NULL");
    CGState* st = c->state;
    double2 *X = (double2*)x, *R = (double2*)r, *P = (double2*)p, *AP = (double2*)Ap;
    const double2* B = (const double2*)b;
    LMX_TRY(launch_cg_init_state(c, st, params->rtol, params->atol, params->maxiter));
    LMX_TRY(launch_dot(c, B, B, 1, st, nullptr));
    LMX_TRY(read_state(c, st, c->host_state));
    if (!c->host_state->done) {
        LMX_CHECK(A(A_user, AP, X, c->stream) == 0, LMX_ERR_ARG, "operator callback failed");
        LMX_TRY(launch_cg_setup_r(c, B, AP, R, P, st));
        LMX_TRY(read_state(c, st, c->host_state));
    }
    long polls = 0;
    while (!c->host_state->done) {
        LMX_CHECK(++polls <= (long)params->maxiter + 2, LMX_ERR_HIP, "CG state did not advance (iteration %d)",
                  c->host_state->iteration);
        LMX_CHECK(A(A_user, AP, P, c->stream) == 0, LMX_ERR_ARG, "operator callback failed");
        LMX_TRY(launch_dot(c, P, AP, 4, st, nullptr));
        LMX_TRY(launch_cg_update_xr(c, X, R, P, AP, st));
        LMX_TRY(launch_cg_update_p(c, R, P, st));
        LMX_TRY(read_state(c, st, c->host_state));
    }
    fill_result(*c->host_state, result);
    return LMX_OK;
}

int lmx_cg_batched_workspace(lmx_wilson* op, int nrhs, const lmx_cg_params* params, size_t* bytes) {
    clear_error();
    LMX_TRY(check_op(op));
    LMX_CHECK(bytes && params && nrhs >= 0, LMX_ERR_ARG, "bad workspace query");
    LMX_TRY(check_nc3(op->ctx, "the batched solver"));
    *bytes = cg_batched_workspace_impl(op, nrhs, params);
    return LMX_OK;
}

int lmx_cg_batched(lmx_wilson* op, void* x, const void* b, int nrhs, void* work, size_t bytes,
                   const lmx_cg_params* params, lmx_cg_result* results) {
    clear_error();
    LMX_TRY(check_op(op));
    LMX_CHECK(x && b && results && params, LMX_ERR_ARG, "bad cg_batched arguments");
    LMX_TRY(check_nc3(op->ctx, "the batched solver"));
    LMX_CHECK(x != b, LMX_ERR_ALIAS, "CG fields x and b must not alias");
    LMX_CHECK(nrhs >= 0 && nrhs <= kMaxStates, LMX_ERR_ARG, "nrhs must be in [0, %d]", kMaxStates);
    const void* f[2] = {x, b};
    const char* names[2] = {"x", "b"};
    LMX_TRY(validate_cg(f, names, 2, params));
    return cg_batched_impl(op, (double2*)x, (const double2*)b, nrhs, work, bytes, params, results);
}

namespace lmx {
// Hopping pullback, plus the clover part (coefficient * csw) for a clover operator.
static int operator_pullback(lmx_wilson* op, double2* dU, const double2* dres, const double2* psi,
                             double coefficient) {
    if (op->ctx->nc != 3) {
        LMX_CHECK(op->csw == 0.0, LMX_ERR_UNSUPPORTED, "the clover link pullback is built for NC = 3 fields");
        return launch_pullback_nc(op->ctx, dU, dres, psi, coefficient);
    }
    LMX_TRY(launch_pullback_f64(op->ctx, dU, dres, psi, coefficient));
    if (op->csw == 0.0) return LMX_OK;
    if (!op->clover_dF) LMX_HIP(hipMalloc(&op->clover_dF, 6 * 9 * (size_t)op->ctx->V * sizeof(double2)));
    return launch_clover_pullback(op->ctx, dU, op->links, op->clover_dF, dres, psi, coefficient * op->csw);
}
}  // namespace lmx

int lmx_wilson_pullback(lmx_wilson* op, void* dU, const void* dres, const void* psi, double coefficient) {
    clear_error();
    LMX_TRY(check_op(op));
    LMX_CHECK(dU && dres && psi, LMX_ERR_ARG, "bad pullback arguments");
    return operator_pullback(op, (double2*)dU, (const double2*)dres, (const double2*)psi, coefficient);
}

int lmx_ta_project(lmx_context* c, void* F, const void* links, const void* G) {
    clear_error();
    LMX_TRY(check_ctx(c));
    LMX_CHECK(F && links && G, LMX_ERR_ARG, "bad ta_project arguments");
    LMX_TRY(check_nc3(c, "the traceless anti-Hermitian projection"));
    return launch_ta_project(c, (double2*)F, (const double2*)links, (const double2*)G);
}

int lmx_wilson_force(lmx_wilson* op, void* G, void* F, const void* phi, void* eta, void* chi, void* work,
                     const lmx_cg_params* params, lmx_cg_result* result) {
    clear_error();
    LMX_TRY(check_op(op));
    LMX_CHECK(G && phi && eta && chi && work && params && result, LMX_ERR_ARG, "bad force arguments");
    lmx_context* c = op->ctx;
    if (F) LMX_TRY(check_nc3(c, "the traceless anti-Hermitian projection"));
    const size_t fb = (size_t)spinor_ncomp(c) * (size_t)c->V;
    double2* w = (double2*)work;
    double2 *r = w, *p = w + fb, *Ap = w + 2 * fb, *tmp = w + 3 * fb;
    LMX_TRY(launch_zero(c, (double2*)eta, (int64_t)fb));
    LMX_TRY(lmx_cg_ddagd(op, eta, phi, r, p, Ap, tmp, params, result));
    Epilogue e{0, nullptr, nullptr, nullptr, nullptr};
    LMX_TRY(launch_wilson_f64(c, op->links, op->clover64, op->kappa, op->csw, (double2*)chi, (const double2*)eta, 1, 0,
                              e));
    // dresult = -2 chi (the dot pullback of g = -2 Re dot(chi, D eta))
    LMX_TRY(launch_scale(c, tmp, (const double2*)chi, -2.0, 0.0, (int64_t)fb));
    LMX_TRY(launch_zero(c, (double2*)G, 4 * (int64_t)link_nelem(c) * c->V));
    LMX_TRY(operator_pullback(op, (double2*)G, tmp, (const double2*)eta, -op->kappa));
    if (F) LMX_TRY(launch_ta_project(c, (double2*)F, op->links, (const double2*)G));
    LMX_HIP(hipStreamSynchronize(c->stream));
    return LMX_OK;
}

}  // extern "C"

namespace lmx {
int lmx_cg_ddagd_internal(lmx_wilson* op, double2* x, const double2* b, double2* r, double2* p, double2* Ap,
                          double2* tmp, const lmx_cg_params* params, lmx_cg_result* result) {
    return lmx_cg_ddagd(op, x, b, r, p, Ap, tmp, params, result);
}
}  // namespace lmx
