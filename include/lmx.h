/*
 * lmx.h -- C ABI of the MI355X-native Wilson(-clover) D_W / D^dagger D library.
 *
 * One context = one lattice geometry + one HIP stream.  Every entry point
 * returns an int status (LMX_OK on success); lmx_last_error() gives the
 * message of the last failure on the calling thread.  No C++ exception
 * crosses the ABI.  All field pointers are DEVICE pointers; sizes are in
 * bytes or counts as named.  Calls are asynchronous on the context stream
 * unless documented otherwise (solvers and reductions return host values and
 * are therefore synchronous).  A context is not thread-safe: use one per host
 * thread.
 *
 * Internal (device) layouts, sites x-fastest (site = x + Lx*(y + Ly*(z + Lz*t))):
 *   spinor  F64: double2 [nrhs][12][V]    component k = spin*3 + colour
 *           F32: float2  [nrhs][12][V]
 *   links   F64: double2 [4][9][V]        element k = row*3 + col
 *           F32: float2  [4][9][V]
 *           F16: per direction two 16-byte words (elements 0-3, 4-7) and one
 *                4-byte word (element 8): int4 [4][2][V] then half2 [4][V]
 *   clover  F64: double2 [27][V]          6 anti-Hermitian planes (12,13,14,23,24,34):
 *           slots 0..8 hold the 18 imaginary diagonals (plane*3+i) in pairs,
 *           slots 9..26 the upper off-diagonals (plane*3 + {01,02,12}).
 * The reference layout (Julia, column-major, halo width nw) is
 *   spinor  A[3,4,L1+2nw,...,L4+2nw]   complex, colour fastest
 *   link    A[3,3,L1+2nw,...,L4+2nw]   per direction, row fastest
 * and is converted by lmx_spinor_import/export and lmx_links_import.
 *
 * Reference citations are relative to the LatticeMatrices.jl source tree.
 */
#ifndef LMX_H
#define LMX_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define LMX_ABI_VERSION 1

/* ---- status codes ------------------------------------------------------ */
enum {
    LMX_OK = 0,
    LMX_ERR_ARG = 1,         /* bad argument (ArgumentError in the reference) */
    LMX_ERR_ALIAS = 2,       /* fields that must be distinct alias each other */
    LMX_ERR_HIP = 3,         /* HIP runtime failure */
    LMX_ERR_NOMEM = 4,       /* device allocation failed */
    LMX_ERR_UNSUPPORTED = 5, /* combination not built */
    LMX_ERR_WORKSPACE = 6    /* caller workspace too small */
};

/* ---- precisions of stored fields --------------------------------------- */
/* LMX_H16: spinors as 16-bit fixed point with one FP32 scale per site
 * (value = q * scale, scale = largest |component| / 32752, so that the sum
 * of two components stays within the FP16 range), links FP16; the
 * mixed solver's inner vectors.  Spinor layout int4[nrhs][3][V] followed by
 * float[nrhs][V]. */
enum { LMX_F64 = 0, LMX_F32 = 1, LMX_F16 = 2, LMX_H16 = 3 };

/* ---- operator flags ------------------------------------------------------ */
enum {
    LMX_DAGGER = 1, /* apply D^dagger (Adjoint_WilsonDiracOperator4D) */
    LMX_DONLY = 2   /* hopping term only, coefficient +1/2 (WilsonDiracOperator4D_Donly) */
};

/* ---- CG reasons: CGResult.reason, src/LinearAlgebras/cg.jl:1-13 --------- */
enum {
    LMX_CG_CONVERGED = 0,
    LMX_CG_MAXITER = 1,
    LMX_CG_NONPOSITIVE_CURVATURE = 2,
    LMX_CG_BREAKDOWN = 3
};

typedef struct lmx_context lmx_context;
typedef struct lmx_wilson lmx_wilson;

/* Solver parameters.  Defaults (lmx_cg_params_default) are cg!'s keyword
 * defaults rtol=1e-10, atol=0, maxiter=5000 (src/LinearAlgebras/cg.jl:73-74). */
typedef struct {
    double rtol;
    double atol;
    int32_t maxiter;
    int32_t precision;        /* inner precision of lmx_cg_batched: LMX_F64 (plain), LMX_F32, LMX_F16, LMX_H16 */
    double reliable_delta;    /* reliable update when |r_inner| < delta * |r at last update| */
    int32_t reliable_maxiter; /* ... or after this many inner iterations */
    int32_t check_every;      /* iterations between host convergence polls */
    int32_t use_graph;        /* capture check_every iterations in a hipGraph */
    int32_t rhs_block;        /* inner dslash shape (0: autotuned): (ZS << 20) | (TL << 12) | (RHS per launch << 4)
                                 | RHS per work item (1, 2, 4); TL = t-slices per streaming workgroup, ZS = 3-D
                                 sub-ranges of the lattice (one XCD's slab spans one), 0 automatic for both;
                                 bit 28 set: the RHS-loop shape (FP16 links), (1 << 28) | (RHS per launch << 4) | 1,
                                 each work item applying its site's links to every RHS of the launch;
                                 bit 29 set: the LDS-staged shape, (1 << 29) | (ZS << 20) | (TL << 12) | (RHS per
                                 launch << 4) | 1, each workgroup holding its block's spinors at slices t and t+1
                                 in LDS (lattices whose x extent divides 256 and whose 256-site blocks are whole
                                 x-rows of one z-plane; elsewhere the streaming shape runs) */
    int32_t reserved[4];
} lmx_cg_params;

/* CGResult (src/LinearAlgebras/cg.jl:7-13) plus the true residual and the
 * reliable-update count of the mixed solver. */
typedef struct {
    int32_t converged;
    int32_t iterations;
    int32_t reason;
    int32_t reliable_updates;
    double residual_norm;
    double relative_residual;
    double true_relative_residual; /* |b - A x| / |b| in FP64, or -1 if not computed */
} lmx_cg_result;

/* Caller-side operator for lmx_cg_generic: out = A in, enqueued on `stream`.
 * Returns 0 on success. */
typedef int (*lmx_apply_fn)(void* user, void* out, const void* in, void* stream);

/* ======================================================================= */
/* Context                                                                  */
/* ======================================================================= */

/* Last error message of the calling thread ("" if none). */
const char* lmx_last_error(void);
int lmx_abi_version(void);

/* Create a context for a lattice of extents dims (x,y,z,t).  phases holds the
 * spinor boundary phases as (re, im) pairs per direction; NULL means all ones.
 * They play the role of LatticeMatrix(...; phases) on the spinor lattice
 * (src/LatticeMatrices_core.jl:297-307,161).  Zero phases are rejected, as in
 * the reference (:329-330).  stream: hipStream_t to run on (NULL: the null
 * stream).  Replaces the single-rank construction of LatticeMatrix /
 * LatticeMatrix_standard (src/LatticeMatrices_core.jl:297-307). */
int lmx_context_create(lmx_context** ctx, const int32_t dims[4], const double* phases, void* stream);
/* The same for fields of nc colours: LatticeMatrix{4,T,AT,NC1,...} with
 * NC1 = nc (src/LatticeMatrices_core.jl:297-307).  nc = 3 is
 * lmx_context_create; nc = 2 runs the reference's generic-NC kernels
 * (kernel_WilsonDiracOperator4D! and its adjoint / Donly variants,
 * src/Operators/WilsonDiracOperator.jl:80-147,444-511,554-621,665-733; the
 * generic clover field strength and term,
 * src/Operators/WilsonDiracCloverOperator.jl:57-118,481-525; the generic link
 * pullback, ext/AD/WilsonDiracOperator.jl:27-53) in their operation order.
 * On such a context spinors hold 4 nc complex components per site (component
 * spin*nc + colour), links nc*nc entries (row*nc + col), the clover cache six
 * full nc x nc planes; import / export, BLAS-1, lmx_wilson_apply(_padded),
 * lmx_ddagd_apply, lmx_cg_ddagd(_noalloc), lmx_cg_generic, lmx_clover_export,
 * lmx_wilson_pullback (Wilson operator) and lmx_wilson_force (F = NULL) work
 * on it; the single-precision, mixed and batched paths, the autotuner, the
 * TA projection, the clover link pullback and lmx_spinor_randomize return
 * LMX_ERR_UNSUPPORTED.  Other nc: LMX_ERR_UNSUPPORTED. */
int lmx_context_create_nc(lmx_context** ctx, const int32_t dims[4], const double* phases, void* stream, int nc);
int lmx_context_destroy(lmx_context* ctx);
int lmx_context_set_stream(lmx_context* ctx, void* stream);
int lmx_context_synchronize(lmx_context* ctx);
int64_t lmx_context_volume(const lmx_context* ctx);
size_t lmx_spinor_bytes(const lmx_context* ctx, int precision, int nrhs);
size_t lmx_links_bytes(const lmx_context* ctx, int precision);
size_t lmx_clover_bytes(const lmx_context* ctx, int precision);

/* ======================================================================= */
/* Layout boundary                                                          */
/* ======================================================================= */

/* Reference padded AoS complex128 spinor(s) -> internal F64 SoA.  src holds
 * nrhs consecutive arrays A[3,4,L1+2nw,...] (LatticeMatrix.A,
 * src/LatticeMatrices_core.jl:361-363); only the core is read (halo ignored). */
int lmx_spinor_import(lmx_context* ctx, void* dst, const void* src_padded, int nw, int nrhs);
/* Internal F64 SoA -> reference padded AoS; the halo is filled with the
 * phase-wrapped core values exactly as set_halo! would (:707-723). */
int lmx_spinor_export(lmx_context* ctx, void* dst_padded, int nw, const void* src, int nrhs);
/* LatticeMatrix(A, dim, PEs; nw, phases) on one rank
 * (src/LatticeMatrices_core.jl:405-461) for any field: the core array
 * A[nc1, nc2, L1, ..., L4] (column-major, element type elem = LMX_ELEM_*, no
 * halo) -> the padded storage dst[nc1, nc2, L1+2nw, ..., L4+2nw] with the halo
 * filled as set_halo! does (the context's boundary phases: real for real
 * fields, +-1 for single precision). */
int lmx_field_pad(lmx_context* ctx, void* dst_padded, const void* src_core, int elem, int nc1, int nc2, int nw);
/* gather_matrix(ls) on one rank (src/LatticeMatrices_core.jl:936-993): the
 * interior of padded storage as the core array A[nc1, nc2, L1, ..., L4]. */
int lmx_field_gather(lmx_context* ctx, void* dst_core, const void* src_padded, int elem, int nc1, int nc2, int nw);
/* Four reference link arrays A[3,3,L+2nw...] (complex128) -> internal F64 links. */
int lmx_links_import(lmx_context* ctx, void* dst, const void* const src_padded[4], int nw);
/* Internal F64 links -> four padded link arrays (halo periodic). */
int lmx_links_export(lmx_context* ctx, void* const dst_padded[4], int nw, const void* src);

/* ======================================================================= */
/* BLAS-1 on spinor fields (F64, nrhs fields stacked)                        */
/* ======================================================================= */

/* Y = a*X + b*Y   -- axpby!, src/LinearAlgebras/linearalgebra_D.jl:2-23 */
int lmx_axpby(lmx_context* ctx, double a, const void* X, double b, void* Y, int nrhs);
/* out = sum conj(A) B (complex) -- dot, src/LinearAlgebras/linearalgebra_D.jl:3325-3341.
 * Synchronous; out[0]=re, out[1]=im. */
int lmx_dot(lmx_context* ctx, const void* A, const void* B, double out[2]);
/* out = s * in (complex scalar) -- mul!(C, a::Number, x), src/LinearAlgebras/mul.jl:2-9 */
int lmx_scale(lmx_context* ctx, void* out, const void* in, double s_re, double s_im, int nrhs);

/* ======================================================================= */
/* Wilson and Wilson-clover operators                                       */
/* ======================================================================= */

/* Create a Wilson operator over internal F64 links (borrowed, not copied;
 * the caller keeps them alive) with hopping parameter kappa.  csw != 0 makes
 * it the Wilson-clover operator and builds the field-strength cache.
 * Replaces WilsonDiracOperator4D(U, kappa) (src/Operators/WilsonDiracOperator.jl:2-35)
 * and WilsonDiracCloverOperator4D(U, kappa, cSW)
 * (src/Operators/WilsonDiracCloverOperator.jl:447-456). */
int lmx_wilson_create(lmx_context* ctx, lmx_wilson** op, const void* links, double kappa, double csw);
int lmx_wilson_destroy(lmx_wilson* op);
int lmx_wilson_set_kappa(lmx_wilson* op, double kappa);
/* Point the operator at (possibly new) links and refresh every derived copy
 * (low-precision links, clover cache) -- update_clover!(operator, U),
 * src/Operators/WilsonDiracCloverOperator.jl:468-479. */
int lmx_wilson_update_links(lmx_wilson* op, const void* links);
/* Export the six clover planes as full F64 3x3 matrices, [6][V][3][3]
 * complex, planes (12,13,14,23,24,34) -- CloverFieldStrength4D components. */
int lmx_clover_export(lmx_wilson* op, void* F_out);

/* out = D in  (flags: LMX_DAGGER, LMX_DONLY) on nrhs stacked F64 fields.
 * Replaces mul!(C, D, psi), mul!(C, D', psi) and the _Donly variants
 * (src/Operators/WilsonDiracOperator.jl:60-76,425-441,535-551,646-662) and the
 * clover mul! (src/Operators/WilsonDiracCloverOperator.jl:561-601). */
int lmx_wilson_apply(lmx_wilson* op, void* out, const void* in, int flags, int nrhs);
/* out = D in in single precision: mul!(C, D, psi) on ComplexF32 fields
 * (LatticeMatrix{...,ComplexF32}, test/wilson_halfspin.jl:105,124-130).  out
 * and in are float2 [nrhs][12][V] (component spin*3+colour); the operator's
 * links and clover cache are used rounded to FP32 (copies the operator
 * builds, with a device allocation, on first use and after
 * lmx_wilson_update_links) and the arithmetic is FP32.  Flags as
 * lmx_wilson_apply (LMX_DAGGER, LMX_DONLY). */
int lmx_wilson_apply_f32(lmx_wilson* op, void* out, const void* in, int flags, int nrhs);
/* The same mul! directly on the reference's storage, without conversion:
 * psi and C are LatticeMatrix.A arrays (complex128, column-major
 * [3,4,L1+2nw,L2+2nw,L3+2nw,L4+2nw], nrhs of them stacked).  psi's halo is
 * not read (neighbours wrap in-kernel with the context's spinor phases,
 * which is what set_halo! would have put there); only C's core sites are
 * written, its halo is left as it was (stale, as after the reference's
 * mul!).  Replaces mul!(C::LatticeMatrix, D, psi) / mul!(C, D', psi) with
 * the LatticeMatrix arrays passed in place
 * (src/Operators/WilsonDiracOperator.jl:60-76,425-441; clover
 * src/Operators/WilsonDiracCloverOperator.jl:561-601). */
int lmx_wilson_apply_padded(lmx_wilson* op, void* C, const void* psi, int nw, int flags, int nrhs);
/* The same on ComplexF32 storage (LatticeMatrix{...,ComplexF32}.A, complex64,
 * column-major [3,4,L1+2nw,...]): FP32 links (the operator's, rounded) and
 * FP32 arithmetic, as lmx_wilson_apply_f32. */
int lmx_wilson_apply_padded_f32(lmx_wilson* op, void* C, const void* psi, int nw, int flags, int nrhs);
/* out = D^dagger D in with caller temp -- DdagDOp(D, temp) mul!,
 * src/Operators/DiracOperators.jl:104-121 (temp must not alias in/out). */
int lmx_ddagd_apply(lmx_wilson* op, void* out, const void* in, void* tmp, int nrhs);

/* ======================================================================= */
/* Solvers                                                                  */
/* ======================================================================= */

void lmx_cg_params_default(lmx_cg_params* p);

/* cg!(x, DdagDOp(D, tmp), b, r, p, Ap; rtol, atol, maxiter) in FP64 with the
 * reference's operation order (src/LinearAlgebras/cg.jl:73-137).  x is the
 * initial guess and the solution; r, p, Ap, tmp are caller-owned work fields
 * whose contents are unspecified on return.  Fields must not alias
 * (LMX_ERR_ALIAS, as _validate_cg_arguments :33-47). */
int lmx_cg_ddagd(lmx_wilson* op, void* x, const void* b, void* r, void* p, void* Ap, void* tmp,
                 const lmx_cg_params* params, lmx_cg_result* result);

/* cg!(x, DdagDOp(D, tmp), b, r, p, Ap; rtol, atol, maxiter) on ComplexF32
 * fields (src/LinearAlgebras/cg.jl:73-137 is generic in the element type;
 * single-precision operator tests test/wilson_halfspin.jl:124-130): x, b, r,
 * p, Ap, tmp are float2 [12][V] fields (lmx_wilson_apply_f32's layout, NC =
 * 3); the operator runs with FP32 links (and the FP32 clover cache) in FP32
 * arithmetic, the vector updates in FP32, the reductions' site sums and the
 * CG scalars in FP64.  The same step order, stopping test and failure
 * reasons as lmx_cg_ddagd.  Fields must not alias (LMX_ERR_ALIAS). */
int lmx_cg_ddagd_f32(lmx_wilson* op, void* x, const void* b, void* r, void* p, void* Ap, void* tmp,
                     const lmx_cg_params* params, lmx_cg_result* result);

/* cg!(x, DdagDOp(D, temp), b, r, p, Ap; rtol, atol, maxiter)
 * (src/LinearAlgebras/cg.jl:73-137) with x and b the LatticeMatrix.A arrays
 * themselves (complex128, column-major [nc,4,L1+2nw,...,L4+2nw]): the solve
 * runs on the library layout in `work` (at least lmx_spinor_bytes(ctx,
 * LMX_F64, 6) bytes, caller-owned; LMX_ERR_WORKSPACE if smaller), x is the
 * initial guess and the solution, and on return x's halo holds what set_halo!
 * would put there.  The reference's r, p, Ap and temp have no role here (the
 * work fields are in `work`).  Synchronous. */
int lmx_cg_ddagd_padded(lmx_wilson* op, void* x, const void* b, int nw, void* work, size_t work_bytes,
                        const lmx_cg_params* params, lmx_cg_result* result);

/* lmx_cg_ddagd_padded for LatticeMatrix{ComplexF32} fields (complex64
 * arrays, NC = 3): the solve is lmx_cg_ddagd_f32 on the library layout in
 * `work` (at least lmx_spinor_bytes(ctx, LMX_F32, 6) bytes).  Synchronous. */
int lmx_cg_ddagd_padded_f32(lmx_wilson* op, void* x, const void* b, int nw, void* work, size_t work_bytes,
                            const lmx_cg_params* params, lmx_cg_result* result);

/* No counterpart in the reference: lmx_cg_ddagd for callers that cannot
 * allocate device memory (code between graph captures, tight serving or HMC
 * loops, harnesses that forbid hipMalloc).  Same arguments, kernels and
 * results (bit for bit) as lmx_cg_ddagd; it calls no hipMalloc / hipFree.
 * The reduction partials and the CG state live in a static device array of
 * 1 MiB + 256 B that each GPU holds until the process ends, so the lattice
 * volume is at most 2^24 sites (64^4; LMX_ERR_ARG beyond).  Every call on
 * one GPU shares that array: calls that overlap on one GPU (several host
 * threads or streams at once) race and must be serialised by the caller.
 * The operator must be a Wilson or Wilson-clover operator whose clover
 * cache was built at creation or by lmx_wilson_update_links. */
int lmx_cg_ddagd_noalloc(lmx_wilson* op, void* x, const void* b, void* r, void* p, void* Ap, void* tmp,
                         const lmx_cg_params* params, lmx_cg_result* result);

/* cg! with a caller operator A (callback on the context stream), FP64,
 * single field -- cg!(x, A, b, r, p, Ap) for any mul!-implementing A. */
int lmx_cg_generic(lmx_context* ctx, lmx_apply_fn A, void* A_user, void* x, const void* b, void* r, void* p,
                   void* Ap, const lmx_cg_params* params, lmx_cg_result* result);

/* Batched D^dagger D solve of nrhs right-hand sides (x, b: nrhs stacked F64
 * fields, x is the initial guess).  params->precision selects plain FP64 CG
 * per RHS (LMX_F64) or the mixed solver: FP32 (LMX_F32), FP16-stored links
 * with FP32 spinors (LMX_F16), or FP16 links with 16-bit spinors (LMX_H16)
 * inner CG (FP32 arithmetic; with LMX_H16 the link products multiply FP16
 * link entries by the projected half-spinor rounded to FP16, accumulating in
 * FP32) with FP64 reliable updates, FP64
 * reductions and FP64 solution accumulation.  Convergence is on the FP64
 * true residual |b - A x| <= max(atol, rtol |b|) per RHS.  The workspace is
 * caller-owned (size from lmx_cg_batched_workspace).  results[nrhs]. */
int lmx_cg_batched_workspace(lmx_wilson* op, int nrhs, const lmx_cg_params* params, size_t* bytes);
int lmx_cg_batched(lmx_wilson* op, void* x, const void* b, int nrhs, void* workspace, size_t workspace_bytes,
                   const lmx_cg_params* params, lmx_cg_result* results);

/* ======================================================================= */
/* Force                                                                    */
/* ======================================================================= */

/* Link cotangent of the Wilson mul!: dU[mu](x) += coefficient * (...) for
 * dresult and psi (F64 spinors) -- the Enzyme reverse rule kernel
 * _kernel_wilson_link_pullback! (ext/AD/WilsonDiracOperator.jl:62-91,137-143)
 * with _wilson_halfspin_link_pullback_direction3!
 * (src/Operators/WilsonDiracOperator.jl:354-410).  For a Wilson-clover
 * operator (csw != 0) the clover part is added with coefficient * csw, as
 * wilson_clover_link_pullback!(dU, D, U, dresult, psi; coefficient = scale)
 * does with -kappa*scale and -kappa*cSW*scale
 * (src/Operators/WilsonCloverPullback.jl:349-389).  dU: internal F64 links layout. */
int lmx_wilson_pullback(lmx_wilson* op, void* dU, const void* dresult, const void* psi, double coefficient);

/* F[mu](x) = TA(U[mu](x) G[mu](x)^dagger) (kernel_traceless_antihermitian!,
 * src/TA/TA.jl:68-116); all in internal F64 links layout. */
int lmx_ta_project(lmx_context* ctx, void* F, const void* links, const void* G);

/* dSFdU(dfdU, D, phi) (ext/LatticeMatricesEnzymeExt.jl:395-460) for the
 * Wilson or Wilson-clover operator: eta = (D^dagger D)^-1 phi by FP64 CG
 * from zero, chi = D eta,
 * G = pullback(dresult = -2 chi, eta, coefficient -kappa) (the Enzyme
 * convention G = dS/dReU + i dS/dImU), and, if F is non-NULL, the su(3) force
 * F = TA(U G^dagger).  eta/chi: caller F64 spinors (outputs).  work: 4 spinors. */
int lmx_wilson_force(lmx_wilson* op, void* G, void* F, const void* phi, void* eta, void* chi, void* work,
                     const lmx_cg_params* params, lmx_cg_result* result);

/* ======================================================================= */
/* Autotuning                                                               */
/* ======================================================================= */

typedef struct {
    int32_t rhs_block;   /* chosen shape, encoded as lmx_cg_params.rhs_block */
    int32_t block_size;  /* threads per workgroup */
    double ms_per_apply; /* measured D apply time of the chosen shape, per RHS */
} lmx_tune_result;

/* Time one group of the mixed solver's inner D / D^dagger iterations of
 * `precision` over the launch shapes (RHS per launch {1,2,4,6,12}, RHS per
 * work item {1,2,4}, t-slices per workgroup, z-ranges), cache the winner per
 * (volume, precision, nrhs, clover term or not, arch) in cache_path (may be
 * NULL: always measure; a cached shape this build cannot launch on this
 * lattice is ignored and retuned) and use it for later lmx_cg_batched calls
 * with rhs_block = 0. */
int lmx_autotune(lmx_wilson* op, int precision, int nrhs, const char* cache_path, lmx_tune_result* out);

/* Average duration in microseconds of the mixed solver's two inner stencil
 * kernels at launch shape rhs_block (0: the shape lmx_autotune chose, or the
 * default), over `groups` groups of four inner iterations as the solver runs
 * them, each launch bracketed by HIP events on the context stream:
 * us[0] = D with |D p|^2 (one launch over nrhs RHS), us[1] = D^dagger with
 * the r update, the x update deferred three times in four and applied the
 * fourth.  Measurement only (used for the roofline of these kernels); it
 * allocates its buffers. */
int lmx_inner_kernel_times(lmx_wilson* op, int precision, int nrhs, int rhs_block, int groups, double us[2]);

/* ======================================================================= */
/* Site random numbers and lattice fills                                    */
/* ======================================================================= */

/* RNGStreamKey(seed, sweep, direction, color, subgroup)
 * (src/utilities/randomgenerator.jl:98-118): the tags of one stream family;
 * each zero-based global site id (x fastest, global_site_id,
 * src/LatticeMatrices.jl:334-345) owns one stream of the family. */
typedef struct {
    uint64_t seed;
    uint64_t sweep;
    uint32_t direction;
    uint32_t color;
    uint32_t subgroup;
    uint32_t reserved;
} lmx_rng_key;

/* Site RNG algorithms: PCG32(), Xoshiro256PlusPlus(), Philox4x32()
 * (randomgenerator.jl:52-59). */
enum { LMX_RNG_PCG32 = 0, LMX_RNG_XOSHIRO256PP = 1, LMX_RNG_PHILOX4X32 = 2 };
/* Element types of a random fill: Float32, Float64, ComplexF32, ComplexF64
 * (_check_randomize_eltype, src/LinearAlgebras/linearalgebra_D.jl:2546-2549). */
enum { LMX_ELEM_F32 = 0, LMX_ELEM_F64 = 1, LMX_ELEM_C64 = 2, LMX_ELEM_C128 = 3 };
/* randomize_matrix! (uniform in [-0.5, 0.5)) or randomize_gaussian_matrix!
 * (normal, standard deviation sigma). */
enum { LMX_FILL_UNIFORM = 0, LMX_FILL_GAUSSIAN = 1 };

/* Draws per site for lmx_site_rng_draw, taken in this order. */
typedef struct {
    int32_t words;     /* rand_u32 */
    int32_t uniform32; /* rand_uniform(rng, Float32) */
    int32_t uniform64; /* rand_uniform(rng, Float64) */
    int32_t normal32;  /* rand_normal_pair(rng, Float32), two values per pair (even count) */
    int32_t normal64;  /* rand_normal_pair(rng, Float64), two values per pair (even count) */
    int32_t bounded;   /* rand_bounded(rng, upper) */
    uint32_t upper;    /* nonzero when bounded > 0 */
    int32_t reserved;
} lmx_rng_draws;

/* For each site i in [0, nsites): stream = site_rng(key, first_site + i,
 * algorithm) (randomgenerator.jl:152-200), then the draws of *n in order
 * (rand_u32 :219-319, rand_uniform :341-352, rand_normal_pair :380-386,
 * rand_bounded :406-414).  Outputs are device arrays [count, nsites]
 * column-major (the draws of one site are contiguous).  Integer and uniform
 * outputs are bit-identical to the reference; normals agree to the last bits
 * of the device log / sqrt / sincos. */
int lmx_site_rng_draw(lmx_context* ctx, const lmx_rng_key* key, int algorithm, uint64_t first_site, int32_t nsites,
                      const lmx_rng_draws* n, uint32_t* words, float* uniform32, double* uniform64, float* normal32,
                      double* normal64, uint32_t* bounded);

/* randomize_matrix!(C, key; rng_algorithm) (dist = LMX_FILL_UNIFORM,
 * src/LinearAlgebras/linearalgebra_D.jl:2560-2582) and
 * randomize_gaussian_matrix!(C, key; sigma, rng_algorithm) (LMX_FILL_GAUSSIAN,
 * :2610-2637) on a LatticeMatrix array in the reference's own storage: C is
 * A[nc1, nc2, L1+2nw, ..., L4+2nw] column-major of element type elem over the
 * context's lattice.  global_offset / global_dims place that lattice as a box
 * of a larger global lattice (both NULL: it is the whole lattice).  When it is
 * the whole lattice the halo is then refilled as set_halo! does
 * (src/LatticeMatrices_core.jl:707-723, the context's boundary phases: real
 * for real fields, +-1 for single precision); a box of a larger lattice has
 * only its core written.  sigma >= 0 (ArgumentError otherwise). */
int lmx_randomize_padded(lmx_context* ctx, void* C, int elem, int nc1, int nc2, int nw, int dist,
                         const lmx_rng_key* key, double sigma, int algorithm, const int32_t* global_offset,
                         const int32_t* global_dims);

/* The same fill of nrhs (3,4) ComplexF64 spinor fields written straight into
 * the library's F64 spinor layout (field r from keys[r]): identical to
 * randomize_*_matrix!(psi, keys[r]) followed by lmx_spinor_import.  The
 * D^dagger D CG benchmark's source is keys = {RNGStreamKey(seed + 1)},
 * LMX_FILL_GAUSSIAN, sigma = sqrt(1/2), Philox
 * (benchmark/ddagd_cg/ddagd_cg.jl:773-774). */
int lmx_spinor_randomize(lmx_context* ctx, void* dst, int nrhs, int dist, const lmx_rng_key* keys, double sigma,
                         int algorithm, const int32_t* global_offset, const int32_t* global_dims);

#ifdef __cplusplus
}
#endif

#endif /* LMX_H */
