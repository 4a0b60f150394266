// lmx_internal.h -- host-side structures shared by the library's translation units.
#pragma once

#include <hip/hip_fp16.h>
#include <hip/hip_runtime.h>

#include <cstdarg>
#include <cstdio>
#include <string>

#include "../../include/lmx.h"
#include "lmx_device.h"

namespace lmx {

// Thread-local last error (lmx_last_error).
void set_error(const char* fmt, ...);
void clear_error();

#define LMX_HIP(call)                                                                              \
    do {                                                                                           \
        hipError_t e_ = (call);                                                                    \
        if (e_ != hipSuccess) {                                                                    \
            ::lmx::set_error("%s failed: %s (%s:%d)", #call, hipGetErrorString(e_), __FILE__, __LINE__); \
            return LMX_ERR_HIP;                                                                    \
        }                                                                                          \
    } while (0)

#define LMX_CHECK(cond, code, ...)         \
    do {                                   \
        if (!(cond)) {                     \
            ::lmx::set_error(__VA_ARGS__); \
            return (code);                 \
        }                                  \
    } while (0)

#define LMX_TRY(call)              \
    do {                           \
        int rc_ = (call);          \
        if (rc_ != LMX_OK) return rc_; \
    } while (0)

// Device-resident state of one CG solve (single RHS).  Written only by the
// finishing block of each reduction; read by every kernel of the iteration
// for its early exit.
struct CGState {
    double bb, bnorm, tol, rr, pAp, alpha, beta, rr_new, rnorm;
    double rtol, atol;
    int iteration, maxiter, done, reason, converged;
    int pad[3];
};

}  // namespace lmx

struct lmx_context {
    lmx::Geom g;
    int64_t V;
    int nc;  // colours of the fields (3: the half-spin kernels; 2: the generic-NC kernels)
    hipStream_t stream;
    int device;
    // Reduction scratch: per-block partials and self-resetting completion counters.
    double* partials;
    size_t partials_cap;  // doubles
    unsigned* counters;   // kCounters entries
    lmx::CGState* state;  // device CG states (kMaxStates)
    lmx::CGState* host_state;  // pinned mirror
    double* dot_out;      // device scalars for lmx_dot
    double* host_dot;     // pinned mirror
    // Stencil launch-shape overrides (0 = automatic): t-slices per workgroup
    // and 3-D sub-ranges, from LMX_STREAM_TL / LMX_STREAM_ZS at creation.
    int tune_tl, tune_zs;
    // Device address of this GPU's static scratch array (the allocation-free
    // CG's partials and state), resolved once per device at creation.
    void* noalloc_scratch;
};

struct lmx_wilson {
    lmx_context* ctx;
    const double2* links;  // borrowed F64 links
    double kappa;
    double csw;
    // Owned derived copies.
    float2* links32;
    __half2* links16;
    double2* clover64;
    float2* clover32;
    __half2* clover16;
    bool lo_valid;
    double2* clover_dF;  // scratch of the clover link pullback (6 planes), on first use
    // Autotuned launch shape per (precision, nrhs); 0 = untuned.
    int tuned_block[4][65];
};

namespace lmx {

constexpr int kCounters = 64;
constexpr int kMaxTL = 32;  // t-slices per streaming workgroup at most (LDS of the fused reductions)
constexpr int kMaxStates = 64;

inline int nblocks(int64_t V) { return int((V + kRedBlock - 1) / kRedBlock); }

// ---- launchers (wilson.hip) ----
struct Epilogue {
    int kind;           // 0 none, 1 pAp (p = aux, finalize alpha)
    const double2* aux; // p field for kind 1
    CGState* st;        // early-exit state (may be null)
    double* partials;
    unsigned* counter;
    const int* gate = nullptr;  // if non-null, the launch does nothing while *gate == 0
};

int launch_wilson_f64(lmx_context* ctx, const double2* U, const double2* clover, double kappa, double csw,
                      double2* out, const double2* in, int nrhs, int flags, const Epilogue& epi);
// FP64 D^dagger of the mixed solver's reliable update with the true residual
// fused: r = b - D^dagger in, written in the inner format (spinor_prec LMX_H16
// with its scales r_norm, or LMX_F32) for the RHS not retired (done) -- or, if
// force, retired with fold set -- and |r|^2 as block partials, row r at
// partials + r * pstride; gate as for Epilogue
int launch_wilson_f64_true(lmx_context* ctx, const double2* U, const double2* clover, double kappa, double csw,
                           const double2* in, const double2* b, void* r, float* r_norm, int spinor_prec,
                           const int* done, const int* fold, int dstride, int force, double* partials, long pstride,
                           int nrhs, const int* gate);
// mul! on the reference's halo-padded AoS spinor arrays (halo ignored, core written)
int launch_wilson_padded(lmx_context* ctx, const double2* U, const double2* clover, double kappa, double csw,
                         double2* C, const double2* psi, int nw, int nrhs, int flags);
int launch_wilson_padded_f32(lmx_context* ctx, const float2* U, const float2* clover, double kappa, double csw,
                             float2* C, const float2* psi, int nw, int nrhs, int flags);

// Low-precision apply: spinors LMX_F32 or LMX_H16 (data + per-site scale),
// links LMX_F32 or LMX_F16, FP32 arithmetic.  epi selects a fused epilogue
// whose block partials (one per 256 consecutive sites) go to row r of
// partials (partials + r * pstride): 2 = |out|^2 on D; 3 = on D^dagger,
// out = Ap is not stored but r -= alpha Ap, x += alpha p (p = aux; r in the
// spinor format, x FP32) with |r|^2, skipping RHS whose done flag is set;
// the x update can be deferred over a group of iterations (apply_x, pmask).
//
// The mixed solver's x_inner update is batched over groups of kXBatch
// iterations: phase j < kXBatch - 1 of a group defers alpha_j p_j (p_j stays
// in ring slot j, bit j of the RHS's mask marks it), phase kXBatch - 1 adds
// the marked terms and its own alpha p -- one x_inner read and write per
// group instead of one per iteration.
constexpr int kXBatch = 4;
// rhs_block bit of the RHS-loop shape: each work item loads its site's eight
// FP16 links once and applies them to every RHS of the launch in turn.
constexpr int kRhsLoop = 1 << 28;
// Whether a launch can take the RHS-loop shape on this context (3-D volume a
// multiple of 256, boundary phases 1 in x, y, z and +-1 in t).
inline bool rhsloop_ok(const lmx_context* ctx) {
    const Geom& g = ctx->g;
    const long V3 = (long)g.L[0] * g.L[1] * g.L[2];
    return V3 % 256 == 0 && !g.general && g.phase_one[0] && g.phase_one[1] && g.phase_one[2];
}
// rhs_block bit of the LDS-staged streaming shape: one RHS per work item,
// the workgroup's 256-site block at slices t and t+1 held in LDS, so that the
// x, y, z neighbours inside the block, the t+1 neighbour and the site's own
// spinor are read from LDS.
constexpr int kLdsStage = 1 << 29;
// Whether a launch can take it: the RHS-loop conditions.
inline bool lds_ok(const lmx_context* ctx) { return rhsloop_ok(ctx); }
struct LoArgs {
    int link_prec, spinor_prec;
    const void* links;
    const void* clover;
    double kappa, csw;
    void* out;
    float* out_norm;
    const void* in;
    const float* in_norm;
    const void* aux;
    const float* aux_norm;
    double* partials;
    long pstride;
    int nrhs, flags, rhs_block;
    const int* gate;
    int epi;
    void* r;
    float* r_norm;
    void* x;  // FP32 component pairs, float4 [nrhs][6][V]
    const double* alpha;
    const double* alpha_q;  // deferred alphas: slot j of RHS r at alpha_q[r * astride + j]
    int astride;
    const int* done;
    const int* pmask;       // deferred slots of RHS r (bit j) at pmask[r * dstride]
    int dstride;
    // apply_x == 0 defers alpha p; apply_x == 1 adds alpha_q[j] pq[j] for the
    // marked slots j (oldest first) and then alpha p
    int apply_x;
    const void* pq[kXBatch - 1];
    const float* pq_norm[kXBatch - 1];
    // EPI_TP (in = r, out = t updated in place): p = r + beta p_prev written
    // to pnext for the RHS not marked done; beta of RHS r at beta[r * astride]
    const double* beta;
    const void* pprev;
    const float* pprev_norm;
    void* pnext;
    float* pnext_norm;
    // a device CG state whose done flag ends the launch early (the FP32 cg!)
    const CGState* st;
};
int launch_wilson_lo(lmx_context* ctx, const LoArgs& a);

int launch_clover_build_f64(lmx_context* ctx, const double2* U, double2* clover);
int launch_clover_export(lmx_context* ctx, const double2* clover, double2* F_out);
int launch_convert_links(lmx_context* ctx, const double2* U, float2* U32, __half2* U16);
int launch_convert_clover(lmx_context* ctx, const double2* F, float2* F32, __half2* F16);

// ---- generic-NC FP64 kernels (wilson_nc.hip; ctx->nc != 3) ----
// nw < 0: library SoA spinors; nw >= 0: the reference's padded storage.
int launch_wilson_nc(lmx_context* ctx, const double2* U, const double2* clover, double kappa, double csw,
                     double2* out, const double2* in, int nrhs, int flags, int nw, const CGState* st,
                     const int* gate);
int launch_clover_build_nc(lmx_context* ctx, const double2* U, double2* F);
int launch_clover_export_nc(lmx_context* ctx, const double2* F, double2* out);
int launch_pullback_nc(lmx_context* ctx, double2* dU, const double2* dres, const double2* psi, double coefficient);
// complex components per spinor site and entries per link
inline int spinor_ncomp(const lmx_context* ctx) { return 4 * ctx->nc; }
inline int link_nelem(const lmx_context* ctx) { return ctx->nc * ctx->nc; }

// ---- BLAS / CG kernels (blas.hip) ----
enum { FIN_OUT = 0, FIN_BB = 1, FIN_RR0 = 2, FIN_RRNEW = 3, FIN_PAP = 4 };
int launch_finalize(lmx_context* ctx, int nb, int kind, bool want_im, CGState* st, double* out2);
int launch_dot(lmx_context* ctx, const double2* A, const double2* B, int fin_kind, CGState* st, double* out2);
int launch_axpby(lmx_context* ctx, double a, const double2* X, double b, double2* Y, int64_t n);
int launch_zero(lmx_context* ctx, double2* p, int64_t n);
int launch_scale(lmx_context* ctx, double2* out, const double2* in, double sr, double si, int64_t n);
int launch_cg_setup_r(lmx_context* ctx, const double2* b, const double2* Ap, double2* r, double2* p, CGState* st);
int launch_cg_update_xr(lmx_context* ctx, double2* x, double2* r, const double2* p, const double2* Ap, CGState* st);
int launch_cg_update_p(lmx_context* ctx, const double2* r, double2* p, CGState* st);
int launch_cg_init_state(lmx_context* ctx, CGState* st, double rtol, double atol, int maxiter);
// the same on ComplexF32 fields (float2 [12][V]; NC = 3)
int launch_dot_f32(lmx_context* ctx, const float2* A, const float2* B, int fin_kind, CGState* st, double* out2);
int launch_cg_setup_r_f32(lmx_context* ctx, const float2* b, const float2* Ap, float2* r, float2* p, CGState* st);
int launch_cg_update_xr_f32(lmx_context* ctx, float2* x, float2* r, const float2* p, const float2* Ap, CGState* st);
int launch_cg_update_p_f32(lmx_context* ctx, const float2* r, float2* p, CGState* st);

// ---- pullback / TA (force.hip) ----
int launch_pullback_f64(lmx_context* ctx, double2* dU, const double2* dres, const double2* psi, double coefficient);
int launch_ta_project(lmx_context* ctx, double2* F, const double2* U, const double2* G);
// clover part of the Wilson-clover link pullback (dF: 6*9*V scratch)
// Boundary phases of a halo refill for fields of element type elem
// (LMX_ELEM_*), exchange_dim_local! (src/LatticeMatrices_core.jl:707-723):
// a plus ghost is the wrapped core value times the phase, a minus ghost times
// its inverse, direction by direction, skipped where the phase is exactly 1.
// Real fields take real phases (inverse 1/x, as inv(x::Float64)), complex
// ones the context's complex inverse; single-precision fields take +-1 only,
// so that the phase and its inverse are exact in any precision.
struct HaloPhases {
    int skip[4];
    double plus_re[4], plus_im[4], minus_re[4], minus_im[4];
};
int make_halo_phases(const lmx_context* ctx, int elem, HaloPhases* hp);
// Generic padded-field construction and interior gather (any NC1 x NC2, any
// element type): LatticeMatrix(A, dim, PEs) and gather_matrix on one rank.
int launch_field_pad(lmx_context* ctx, void* dst, const void* src, int elem, int nc1, int nc2, int nw);
int launch_field_unpad(lmx_context* ctx, void* dst, const void* src, int elem, int nc1, int nc2, int nw);
int launch_site_rng_draw(lmx_context* ctx, const lmx_rng_key* key, int algorithm, uint64_t first_site, int nsites,
                         const lmx_rng_draws* n, uint32_t* words, float* u32, double* u64, float* n32, double* n64,
                         uint32_t* bnd);
int launch_randomize_padded(lmx_context* ctx, void* C, int elem, int nc1, int nc2, int nw, int dist,
                            const lmx_rng_key* key, double sigma, int algorithm, const int32_t* goff,
                            const int32_t* gdims);
int launch_spinor_randomize(lmx_context* ctx, void* dst, int nrhs, int dist, const lmx_rng_key* keys, double sigma,
                            int algorithm, const int32_t* goff, const int32_t* gdims);
int launch_clover_pullback(lmx_context* ctx, double2* dU, const double2* U, double2* dF, const double2* dres,
                           const double2* psi, double coefficient);

// ---- batched / mixed CG (mixed.hip) ----
int cg_batched_impl(lmx_wilson* op, double2* x, const double2* b, int nrhs, void* work, size_t bytes,
                    const lmx_cg_params* params, lmx_cg_result* results);
size_t cg_batched_workspace_impl(lmx_wilson* op, int nrhs, const lmx_cg_params* params);

int ensure_lo_links(lmx_wilson* op);

}  // namespace lmx
