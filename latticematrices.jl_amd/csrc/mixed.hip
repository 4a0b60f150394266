// mixed.hip -- batched D^dagger D CG over many right-hand sides, plain FP64
// or mixed precision with FP64 reliable updates.
//
// Mixed solver (per right-hand side i, all RHS advanced in lockstep by
// batched kernels; every scalar lives on the device):
//   inner CG in FP32 arithmetic on FP32 or 16-bit stored vectors with FP32
//   or FP16 links; p.Ap = |D p|^2 and |r|^2 accumulated in FP64 in the
//   stencils' epilogues (the direction update p = r + beta p rides on the D
//   launch, which stencils r and forms D p = D r + beta D p_prev; the x/r
//   update rides on the D^dagger launch), alpha/beta per RHS on the device;
//   a reliable update replaces the inner residual by the FP64 true residual
//   r = b - A x (A = D^dagger D in FP64) and folds the inner solution into the
//   FP64 solution x, when |r| drops below delta * (largest |r| since the
//   last update), after reliable_maxiter iterations, or when the inner
//   residual reaches the tolerance.  Convergence is declared only on the
//   FP64 true residual.  The update decision is a device flag that gates the
//   FP64 kernels, so one iteration is a fixed sequence of launches that can be
//   captured once in a hipGraph and replayed.
// The reference solver this replaces is cg! (src/LinearAlgebras/cg.jl:73-137)
// on DdagDOp (src/Operators/DiracOperators.jl:104-121).
#include <algorithm>
#include <cmath>
#include <cstddef>
#include <cstring>
#include <vector>

#include "lmx_internal.h"

namespace lmx {

struct MixRhs {
    double bnorm, tol, rr, rr_prev, pAp, alpha, beta, rref, rtrue;
    double alpha_q[kXBatch - 1];  // alphas of the deferred x_inner += alpha p, by ring slot
    int done, converged, reason, since, iterations, updates;
    int pmask;  // ring slots whose alpha p awaits the group's last iteration (or a flush)
    // retired (maximum iterations, breakdown, nonpositive curvature) with its
    // inner solution and deferred alpha p not yet in x: the solve's final
    // pass folds them in and computes the true residual
    int fold;
};
struct MixBatch {
    int do_update;  // reliable update requested this iteration
    int n_active;   // RHS still iterating
    double delta;
    int rmax;
    int maxiter;    // an RHS stops once its inner iteration count reaches it
};

// An inner (low-precision) vector of nrhs fields: FP32 (n == null) or
// 16-bit fixed point with per-site scales n.
struct InnerVec {
    void* p;
    float* n;
};

struct MixPtrs {
    float4* xlo;  // inner solution: FP32 component pairs, float4 [nrhs][6][V]
    InnerVec r, p[kXBatch], t;  // p ring: iteration k writes p[k % kXBatch] from p[(k - 1) % kXBatch]
    double2* r64;  // A x of a reliable update (FP64); the FP64 residual is formed on the fly from it
    double2* t64;  // D x, the intermediate of that A x
    double* partials;
    MixRhs* rhs;
    MixBatch* batch;
};

static inline size_t align256(size_t v) { return (v + 255) & ~size_t(255); }

static size_t layout(lmx_context* ctx, int nrhs, int sp, MixPtrs* P, char* base) {
    const size_t V = (size_t)ctx->V;
    const size_t f32 = 12 * V * sizeof(float2) * nrhs, f64 = 12 * V * sizeof(double2) * nrhs;
    const size_t nb = (size_t)nblocks(ctx->V);
    size_t off = 0;
    auto take = [&](size_t bytes) {
        char* p = base ? base + off : nullptr;
        off += align256(bytes);
        return p;
    };
    auto inner = [&]() {
        InnerVec v;
        if (sp == LMX_H16) {
            v.p = take(3 * V * sizeof(int4) * nrhs);
            v.n = (float*)take(V * sizeof(float) * nrhs);
        } else {
            v.p = take(f32);
            v.n = nullptr;
        }
        return v;
    };
    MixPtrs q;
    q.xlo = (float4*)take(f32);
    q.r = inner();
    for (int j = 0; j < kXBatch; ++j) q.p[j] = inner();
    q.t = inner();
    q.r64 = (double2*)take(f64);
    q.t64 = (double2*)take(f64);
    q.partials = (double*)take(2 * nb * nrhs * sizeof(double));
    q.rhs = (MixRhs*)take(nrhs * sizeof(MixRhs));
    q.batch = (MixBatch*)take(sizeof(MixBatch));
    if (P) *P = q;
    return off;
}

// ---------------------------------------------------------------------------
// Reductions: every block of row i (blockIdx.y = RHS) stores the block sums
// of (a, b) over its 256 sites; a separate one-block-per-RHS finalize launch
// sums them in a fixed order and updates that RHS's state.
// ---------------------------------------------------------------------------
__device__ __forceinline__ void row_partials(double a, double b, double* partials, int nb, int bx) {
    __shared__ double lds4[4];
    const int i = blockIdx.y;
    double ta = block_tree(a, lds4);
    double tb = block_tree(b, lds4);
    if (threadIdx.x == 0) {
        partials[(size_t)(2 * i) * nb + bx] = ta;
        partials[(size_t)(2 * i + 1) * nb + bx] = tb;
    }
}

__device__ __forceinline__ void retire(MixRhs* st, MixBatch* batch, int reason, int converged, int fold) {
    st->done = 1;
    st->reason = reason;
    st->converged = converged;
    st->fold = fold;
    atomicSub(&batch->n_active, 1);
}

enum { MX_INIT = 0, MX_PAP = 1, MX_RR = 2, MX_TRUE = 3 };

// flag: MX_TRUE -- run even without a reliable-update request (the final
//                  pass: RHS still iterating, and retired RHS awaiting their fold);
//       MX_RR   -- the iteration's phase j in its group: j < kXBatch - 1
//                  deferred its x_inner += alpha p into ring slot j.
// An RHS stops with LMX_CG_MAXITER when its iteration count reaches maxiter
// (cg!, src/LinearAlgebras/cg.jl:105,135-136), in the iteration's own
// finalize, so a graph replay that runs past maxiter leaves it untouched.
template <int KIND>
__global__ void __launch_bounds__(256) k_mx_fin(const double* __restrict__ partials, int nb, MixRhs* rhs,
                                                MixBatch* batch, double rtol, double atol, int flag) {
    const int i = blockIdx.x;
    MixRhs* st = rhs + i;
    if (KIND == MX_PAP && i == 0 && threadIdx.x == 0) batch->do_update = 0;
    if (KIND == MX_TRUE && !flag && !batch->do_update) return;
    const bool final_fold = KIND == MX_TRUE && flag && st->done && st->fold;
    if (st->done && !final_fold) return;
    if (threadIdx.x == 0) {
        if (KIND == MX_RR) {
            if (flag < kXBatch - 1) {
                st->alpha_q[flag] = st->alpha;
                st->pmask |= 1 << flag;
            } else {
                st->pmask = 0;  // the group's last iteration applied them
            }
        }
        if (KIND == MX_TRUE) st->pmask = 0;  // k_mx_accum folded the deferred terms
    }
    __shared__ double lds4[4];
    const double A = finish_partials(partials + (size_t)(2 * i) * nb, nb, lds4);
    const double B = (KIND == MX_INIT) ? finish_partials(partials + (size_t)(2 * i + 1) * nb, nb, lds4) : 0.0;
    if (threadIdx.x != 0) return;
    if (KIND == MX_INIT) {  // A = |r|^2, B = |b|^2
        st->bnorm = sqrt(B);
        st->tol = fmax(atol, rtol * st->bnorm);
        st->rr = A;
        st->rr_prev = A;
        st->rtrue = sqrt(A);
        st->rref = st->rtrue;
        st->since = 0;
        st->iterations = 0;
        st->updates = 0;
        st->beta = 0.0;
        if (!isfinite(A) || !isfinite(B))
            retire(st, batch, LMX_CG_BREAKDOWN, 0, 0);
        else if (st->rtrue <= st->tol)
            retire(st, batch, LMX_CG_CONVERGED, 1, 0);
        else if (batch->maxiter <= 0)
            retire(st, batch, LMX_CG_MAXITER, 0, 0);
    } else if (KIND == MX_PAP) {  // A = p.Ap = |D p|^2
        st->pAp = A;
        if (!isfinite(A))
            retire(st, batch, LMX_CG_BREAKDOWN, 0, 1);
        else if (A <= 0.0)
            retire(st, batch, LMX_CG_NONPOSITIVE_CURVATURE, 0, 1);
        else
            st->alpha = st->rr / A;
    } else if (KIND == MX_RR) {  // A = |r_inner|^2
        st->iterations += 1;
        st->since += 1;
        if (!isfinite(A)) {
            retire(st, batch, LMX_CG_BREAKDOWN, 0, 1);
            return;
        }
        const double rn = sqrt(A);
        st->rref = fmax(st->rref, rn);
        st->rr_prev = st->rr;
        st->beta = A / st->rr;
        st->rr = A;
        if (st->iterations >= batch->maxiter) {
            // the final pass folds x_inner in; a true residual within the
            // tolerance there still counts as converged
            retire(st, batch, LMX_CG_MAXITER, 0, 1);
            return;
        }
        if (rn < batch->delta * st->rref || st->since >= batch->rmax || rn <= st->tol) atomicOr(&batch->do_update, 1);
    } else if (final_fold) {  // MX_TRUE on a retired RHS: A = |b - A x|^2 after its fold
        st->rtrue = sqrt(A);
        st->updates += 1;
        st->fold = 0;
        if (st->reason == LMX_CG_MAXITER && isfinite(A) && st->rtrue <= st->tol) {
            st->reason = LMX_CG_CONVERGED;
            st->converged = 1;
        }
    } else {  // MX_TRUE: A = |b - A x|^2 in FP64
        st->rtrue = sqrt(A);
        st->updates += 1;
        if (!isfinite(A)) {
            retire(st, batch, LMX_CG_BREAKDOWN, 0, 0);
        } else if (st->rtrue <= st->tol) {
            retire(st, batch, LMX_CG_CONVERGED, 1, 0);
        } else {
            st->beta = A / st->rr_prev;
            st->rr = A;
            st->rref = st->rtrue;
            st->since = 0;
        }
    }
}

// r = b - A x with A x in ax64 (FP64; r itself is not kept in FP64: the next
// reliable update recomputes it); r_lo = r rounded to the inner format; the
// first iteration's D launch (beta = 0) forms t = D r_lo + 0 t and
// p = r_lo + 0 p_prev from t = 0 and p_prev = r_lo written here; x_lo = 0;
// |r|^2 and |b|^2 per RHS.
template <class SO>
__global__ void __launch_bounds__(256) k_mx_init(const double2* __restrict__ b, const double2* __restrict__ ax64, SO rlo,
                                                 SO t, SO pprev, float4* __restrict__ xlo, int V, double* partials) {
    const int i = blockIdx.y;
    const int s = blockIdx.x * 256 + threadIdx.x;
    double rr = 0.0, bb = 0.0;
    if (s < V) {
        const size_t o = (size_t)i * 12 * V;
        cx<float> f[12];
#pragma unroll
        for (int k = 0; k < 12; ++k) {
            const size_t e = o + (size_t)k * V + s;
            double2 bv = b[e], av = ax64[e];
            double2 rv = make_double2(bv.x - av.x, bv.y - av.y);
            f[k] = mk<float>((float)rv.x, (float)rv.y);
            rr += rv.x * rv.x + rv.y * rv.y;
            bb += bv.x * bv.x + bv.y * bv.y;
        }
        rlo.st12(i, s, f);
        pprev.st12(i, s, f);
#pragma unroll
        for (int k = 0; k < 12; ++k) f[k] = mk<float>(0.f, 0.f);
        t.st12(i, s, f);
#pragma unroll
        for (int m = 0; m < 6; ++m) xlo[(size_t)i * 6 * V + (size_t)m * V + s] = make_float4(0.f, 0.f, 0.f, 0.f);
    }
    row_partials(rr, bb, partials, gridDim.x, blockIdx.x);
}

__global__ void k_mx_reset(MixRhs* rhs, MixBatch* batch, int nrhs, double delta, int rmax, int maxiter) {
    const int t = threadIdx.x;
    if (t < nrhs) {
        MixRhs z;
        memset(&z, 0, sizeof(z));
        z.reason = LMX_CG_MAXITER;
        rhs[t] = z;
    }
    if (t == 0) {
        batch->do_update = 0;
        batch->n_active = nrhs;
        batch->delta = delta;
        batch->rmax = rmax;
        batch->maxiter = maxiter;
    }
}

// reliable update, part 1: x += x_lo (+ the deferred alpha p of the marked
// ring slots, oldest first) in FP64, x_lo = 0
template <class SI>
struct RingSlots {
    SI p[kXBatch - 1];
};
template <class SI>
__device__ __forceinline__ void accum_site(double2* __restrict__ x, float4* __restrict__ xlo, const RingSlots<SI>& ring,
                                           int V, const MixRhs& st, int mask, size_t o, size_t o4, int i, int s) {
    double d[24];
#pragma unroll
    for (int k = 0; k < 24; ++k) d[k] = 0.0;
#pragma unroll
    for (int j = 0; j < kXBatch - 1; ++j) {
        if (mask & (1 << j)) {
            const double ap = st.alpha_q[j];
            cx<float> q[12];
            ring.p[j].ld12(i, s, q);
#pragma unroll
            for (int k = 0; k < 12; ++k) {
                d[2 * k] += ap * (double)q[k].re;
                d[2 * k + 1] += ap * (double)q[k].im;
            }
        }
    }
#pragma unroll
    for (int m = 0; m < 6; ++m) {
        const float4 l = xlo[o4 + (size_t)m * V + s];
        const size_t e0 = o + (size_t)(2 * m) * V + s, e1 = e0 + V;
        double a0 = (double)l.x, b0 = (double)l.y, a1 = (double)l.z, b1 = (double)l.w;
        if (mask) {
            a0 += d[4 * m];
            b0 += d[4 * m + 1];
            a1 += d[4 * m + 2];
            b1 += d[4 * m + 3];
        }
        const double2 h0 = x[e0], h1 = x[e1];
        x[e0] = make_double2(h0.x + a0, h0.y + b0);
        x[e1] = make_double2(h1.x + a1, h1.y + b1);
        xlo[o4 + (size_t)m * V + s] = make_float4(0.f, 0.f, 0.f, 0.f);
    }
}

// A reliable-update kernel works on RHS i when the iteration requested an
// update and i still iterates, or, in the final (forced) pass, when i still
// iterates or retired with its fold pending.
__device__ __forceinline__ bool update_runs(const MixRhs& st, const MixBatch* batch, int force) {
    return force ? (!st.done || st.fold) : (batch->do_update && !st.done);
}

// (grid-stride over the sites: a small grid, so that the launch costs little
// on the iterations where the update is gated off)
template <class SI>
__global__ void __launch_bounds__(256) k_mx_accum(double2* __restrict__ x, float4* __restrict__ xlo, RingSlots<SI> ring,
                                                  int V, const MixRhs* rhs, const MixBatch* batch, int force) {
    const int i = blockIdx.y;
    if (!update_runs(rhs[i], batch, force)) return;
    const size_t o = (size_t)i * 12 * V, o4 = (size_t)i * 6 * V;
    const int mask = rhs[i].pmask;  // slots not yet overwritten by a later p update
    for (int s = blockIdx.x * 256 + threadIdx.x; s < V; s += gridDim.x * 256)
        accum_site(x, xlo, ring, V, rhs[i], mask, o, o4, i, s);
}

// ---------------------------------------------------------------------------
// Host driver
// ---------------------------------------------------------------------------
struct MixRun {
    lmx_wilson* op;
    lmx_context* ctx;
    MixPtrs w;
    double2* x;
    const double2* b;
    int nrhs;
    int link_prec;
    int spinor_prec;
    int rhs_block;
};

template <class SI>
static SI in_of(const InnerVec& v, int V) {
    return SI::at(v.p, v.n, V, 0);
}
template <class SO>
static SO out_of(const InnerVec& v, int V) {
    return SO::at(v.p, v.n, V, 0);
}

static int apply_ddagd64(MixRun& R, double2* out, const double2* in, double2* tmp, const int* gate) {
    Epilogue e{0, nullptr, nullptr, nullptr, nullptr, gate};
    lmx_wilson* op = R.op;
    LMX_TRY(launch_wilson_f64(R.ctx, op->links, op->clover64, op->kappa, op->csw, tmp, in, R.nrhs, 0, e));
    return launch_wilson_f64(R.ctx, op->links, op->clover64, op->kappa, op->csw, out, tmp, R.nrhs, LMX_DAGGER, e);
}

static LoArgs lo_args(lmx_wilson* op, int link_prec, int spinor_prec) {
    LoArgs a;
    memset(&a, 0, sizeof(a));
    a.link_prec = link_prec;
    a.spinor_prec = spinor_prec;
    a.links = link_prec == LMX_F16 ? (const void*)op->links16 : (const void*)op->links32;
    a.clover = op->clover64 ? (link_prec == LMX_F16 ? (const void*)op->clover16 : (const void*)op->clover32) : nullptr;
    a.kappa = op->kappa;
    a.csw = op->csw;
    return a;
}

// One inner iteration's operator work, in two launches with fused epilogues:
//   t = D r + beta t = D p (beta from the previous iteration, or from the
//   reliable update that replaced r; 0 on the first iteration), writing
//   p = r + beta p_prev at each site, with the block partials of
//   |t|^2 = p.Ap (FP64 sums), then
//   (after the finalize has set alpha) Ap = D^dagger t, kept in registers:
//   r -= alpha Ap with the block partials of |r|^2, and the x_inner update
//   batched over groups of kXBatch iterations: phase j < kXBatch - 1 defers
//   alpha p (its p stays in ring slot j, the finalize records alpha and
//   marks the slot), the group's last phase adds the marked terms and its own
//   alpha p -- the x_inner traffic of one iteration for kXBatch.  A reliable
//   update inside a group folds the marked terms into x in k_mx_accum.
static_assert(sizeof(MixRhs) % sizeof(double) == 0, "MixRhs is addressed in doubles and ints");

static int apply_d_tp(MixRun& R, int e) {
    LoArgs a = lo_args(R.op, R.link_prec, R.spinor_prec);
    a.nrhs = R.nrhs;
    a.rhs_block = R.rhs_block;
    a.gate = &R.w.batch->n_active;
    a.out = R.w.t.p;
    a.out_norm = R.w.t.n;
    a.in = R.w.r.p;
    a.in_norm = R.w.r.n;
    const InnerVec& pp = R.w.p[(e + kXBatch - 1) % kXBatch];
    a.pprev = pp.p;
    a.pprev_norm = pp.n;
    a.pnext = R.w.p[e].p;
    a.pnext_norm = R.w.p[e].n;
    a.beta = (const double*)((const char*)R.w.rhs + offsetof(MixRhs, beta));
    a.astride = (int)(sizeof(MixRhs) / sizeof(double));
    a.done = (const int*)((const char*)R.w.rhs + offsetof(MixRhs, done));
    a.dstride = (int)(sizeof(MixRhs) / sizeof(int));
    a.flags = 0;
    a.epi = 4;
    a.partials = R.w.partials;
    a.pstride = 2L * nblocks(R.ctx->V);
    return launch_wilson_lo(R.ctx, a);
}

static int apply_ddag_xr(MixRun& R, int e) {
    LoArgs a = lo_args(R.op, R.link_prec, R.spinor_prec);
    a.nrhs = R.nrhs;
    a.rhs_block = R.rhs_block;
    a.gate = &R.w.batch->n_active;
    a.out = nullptr;  // Ap is consumed in registers
    a.in = R.w.t.p;
    a.in_norm = R.w.t.n;
    a.aux = R.w.p[e].p;
    a.aux_norm = R.w.p[e].n;
    for (int j = 0; j < kXBatch - 1; ++j) {
        a.pq[j] = R.w.p[j].p;
        a.pq_norm[j] = R.w.p[j].n;
    }
    a.apply_x = e == kXBatch - 1 ? 1 : 0;
    a.flags = LMX_DAGGER;
    a.epi = 3;
    a.r = R.w.r.p;
    a.r_norm = R.w.r.n;
    a.x = R.w.xlo;
    a.alpha = (const double*)((const char*)R.w.rhs + offsetof(MixRhs, alpha));
    a.alpha_q = (const double*)((const char*)R.w.rhs + offsetof(MixRhs, alpha_q));
    a.astride = (int)(sizeof(MixRhs) / sizeof(double));
    a.done = (const int*)((const char*)R.w.rhs + offsetof(MixRhs, done));
    a.pmask = (const int*)((const char*)R.w.rhs + offsetof(MixRhs, pmask));
    a.dstride = (int)(sizeof(MixRhs) / sizeof(int));
    a.partials = R.w.partials;
    a.pstride = 2L * nblocks(R.ctx->V);
    return launch_wilson_lo(R.ctx, a);
}

// The reliable-update kernels run on most iterations only to find their gate
// off; they stride over the lattice with at most this many workgroups per RHS.
constexpr unsigned kUpdateBlocks = 256;

// reliable update, parts 2 and 3: t64 = D x, then the FP64 D^dagger with the
// true residual fused (r = b - D^dagger D x in the inner format, |r|^2 in
// FP64 block partials); gated like part 1, or forced on the final pass
static int enqueue_true_ddagd(MixRun& R, const double2* x, const int* gate, int force) {
    lmx_wilson* op = R.op;
    const Epilogue e{0, nullptr, nullptr, nullptr, nullptr, gate};
    LMX_TRY(launch_wilson_f64(R.ctx, op->links, op->clover64, op->kappa, op->csw, R.w.t64, x, R.nrhs, 0, e));
    const int* done = (const int*)((const char*)R.w.rhs + offsetof(MixRhs, done));
    const int* fold = (const int*)((const char*)R.w.rhs + offsetof(MixRhs, fold));
    return launch_wilson_f64_true(R.ctx, op->links, op->clover64, op->kappa, op->csw, R.w.t64, R.b, R.w.r.p, R.w.r.n,
                                  R.spinor_prec, done, fold, (int)(sizeof(MixRhs) / sizeof(int)), force,
                                  R.w.partials, 2L * nblocks(R.ctx->V), R.nrhs, gate);
}
template <class SO>
static void enqueue_init(MixRun& R, dim3 grid) {
    const int V = (int)R.ctx->V;
    hipLaunchKernelGGL((k_mx_init<SO>), grid, dim3(256), 0, R.ctx->stream, R.b, R.w.r64, out_of<SO>(R.w.r, V),
                       out_of<SO>(R.w.t, V), out_of<SO>(R.w.p[kXBatch - 1], V), R.w.xlo, V, R.w.partials);
}
// reliable-update part 1 (gated, or forced at the end): x += x_inner (+ the marked alpha p)
template <class SI>
static RingSlots<SI> ring_of(MixRun& R) {
    RingSlots<SI> g;
    for (int j = 0; j < kXBatch - 1; ++j) g.p[j] = in_of<SI>(R.w.p[j], (int)R.ctx->V);
    return g;
}
static void enqueue_accum(MixRun& R, dim3 grid, int force) {
    const int V = (int)R.ctx->V;
    grid.x = std::min(grid.x, kUpdateBlocks);
    if (R.spinor_prec == LMX_H16)
        hipLaunchKernelGGL((k_mx_accum<SpH16In>), grid, dim3(256), 0, R.ctx->stream, R.x, R.w.xlo, ring_of<SpH16In>(R),
                           V, R.w.rhs, R.w.batch, force);
    else
        hipLaunchKernelGGL((k_mx_accum<SpF32In>), grid, dim3(256), 0, R.ctx->stream, R.x, R.w.xlo, ring_of<SpF32In>(R),
                           V, R.w.rhs, R.w.batch, force);
}

// One iteration with phase e (iteration k has e = k % kXBatch).
static int enqueue_iteration(MixRun& R, int e) {
    lmx_context* c = R.ctx;
    dim3 grid(nblocks(c->V), R.nrhs), blk(256);
    const int nb = nblocks(c->V);
    MixRhs* rhs = R.w.rhs;
    MixBatch* bt = R.w.batch;
    LMX_TRY(apply_d_tp(R, e));
    hipLaunchKernelGGL(k_mx_fin<MX_PAP>, dim3(R.nrhs), blk, 0, c->stream, R.w.partials, nb, rhs, bt, 0.0, 0.0, 0);
    LMX_TRY(apply_ddag_xr(R, e));
    hipLaunchKernelGGL(k_mx_fin<MX_RR>, dim3(R.nrhs), blk, 0, c->stream, R.w.partials, nb, rhs, bt, 0.0, 0.0, e);
    enqueue_accum(R, grid, 0);
    LMX_TRY(enqueue_true_ddagd(R, R.x, &bt->do_update, 0));
    hipLaunchKernelGGL(k_mx_fin<MX_TRUE>, dim3(R.nrhs), blk, 0, c->stream, R.w.partials, nb, rhs, bt, 0.0, 0.0, 0);
    LMX_HIP(hipGetLastError());
    return LMX_OK;
}

static int spinor_prec_of(int precision) { return precision == LMX_H16 ? LMX_H16 : LMX_F32; }
static int link_prec_of(int precision) { return precision == LMX_F32 ? LMX_F32 : LMX_F16; }

size_t cg_batched_workspace_impl(lmx_wilson* op, int nrhs, const lmx_cg_params* params) {
    if (params->precision == LMX_F64) return align256(4 * 12 * (size_t)op->ctx->V * sizeof(double2));
    return layout(op->ctx, std::max(nrhs, 1), spinor_prec_of(params->precision), nullptr, nullptr);
}

int lmx_cg_ddagd_internal(lmx_wilson* op, double2* x, const double2* b, double2* r, double2* p, double2* Ap,
                          double2* tmp, const lmx_cg_params* params, lmx_cg_result* result);

static int default_rhs_block(lmx_wilson* op, int prec, int nrhs) {
    const int t = (prec >= 0 && prec < 4 && nrhs >= 0 && nrhs <= 64) ? op->tuned_block[prec][nrhs] : 0;
    if (t > 0) return t;
    // all RHS in one launch, two per work item when that divides evenly
    const int per = std::min(nrhs, 12);
    const int nr = (per % 2 == 0) ? 2 : 1;
    return (per << 4) | nr;
}

int cg_batched_impl(lmx_wilson* op, double2* x, const double2* b, int nrhs, void* work, size_t bytes,
                    const lmx_cg_params* params, lmx_cg_result* results) {
    lmx_context* c = op->ctx;
    const size_t need = cg_batched_workspace_impl(op, nrhs, params);
    LMX_CHECK(work != nullptr || need == 0, LMX_ERR_ARG, "workspace is NULL");
    LMX_CHECK(bytes >= need, LMX_ERR_WORKSPACE, "workspace of %zu bytes is smaller than the %zu required", bytes, need);
    if (nrhs == 0) return LMX_OK;
    const size_t fs = 12 * (size_t)c->V;
    if (params->precision == LMX_F64) {
        // plain FP64 cg! per right-hand side, the reference algorithm
        double2* w = (double2*)work;
        for (int i = 0; i < nrhs; ++i)
            LMX_TRY(lmx_cg_ddagd_internal(op, x + i * fs, b + i * fs, w, w + fs, w + 2 * fs, w + 3 * fs, params,
                                          results + i));
        return LMX_OK;
    }
    LMX_CHECK(params->precision == LMX_F32 || params->precision == LMX_F16 || params->precision == LMX_H16,
              LMX_ERR_ARG, "unknown precision %d", params->precision);
    LMX_CHECK(nrhs <= kMaxStates, LMX_ERR_ARG, "at most %d right-hand sides per batch", kMaxStates);
    LMX_CHECK(params->reliable_delta > 0.0 && params->reliable_delta < 1.0, LMX_ERR_ARG,
              "reliable_delta must be in (0, 1)");
    LMX_TRY(ensure_lo_links(op));
    MixRun R;
    R.op = op;
    R.ctx = c;
    R.spinor_prec = spinor_prec_of(params->precision);
    R.link_prec = link_prec_of(params->precision);
    layout(c, nrhs, R.spinor_prec, &R.w, (char*)work);
    R.x = x;
    R.b = b;
    R.nrhs = nrhs;
    R.rhs_block = params->rhs_block > 0 ? params->rhs_block : default_rhs_block(op, params->precision, nrhs);
    dim3 grid(nblocks(c->V), nrhs), blk(256);
    const int rmax = params->reliable_maxiter > 0 ? params->reliable_maxiter : 100;
    hipLaunchKernelGGL(k_mx_reset, dim3(1), dim3(64), 0, c->stream, R.w.rhs, R.w.batch, nrhs, params->reliable_delta,
                       rmax, params->maxiter);
    LMX_TRY(apply_ddagd64(R, R.w.r64, x, R.w.t64, nullptr));
    if (R.spinor_prec == LMX_H16)
        enqueue_init<SpH16Out>(R, grid);
    else
        enqueue_init<SpF32Out>(R, grid);
    hipLaunchKernelGGL(k_mx_fin<MX_INIT>, dim3(nrhs), blk, 0, c->stream, R.w.partials, nblocks(c->V), R.w.rhs,
                       R.w.batch, params->rtol, params->atol, 0);
    LMX_HIP(hipGetLastError());

    // iterations between host polls; a multiple of kXBatch with a graph, so
    // every replay starts a group (the x_inner batching groups iterations)
    int chunk = params->check_every > 0 ? params->check_every : 10;
    if (params->use_graph) chunk = (chunk + kXBatch - 1) / kXBatch * kXBatch;
    hipGraphExec_t exec = nullptr;
    hipGraph_t graph = nullptr;
    if (params->use_graph) {
        hipStream_t cap;
        LMX_HIP(hipStreamCreateWithFlags(&cap, hipStreamNonBlocking));
        hipStream_t saved = c->stream;
        c->stream = cap;
        hipError_t e = hipStreamBeginCapture(cap, hipStreamCaptureModeRelaxed);
        int rc = LMX_OK;
        if (e == hipSuccess) {
            for (int k = 0; k < chunk && rc == LMX_OK; ++k) rc = enqueue_iteration(R, k % kXBatch);
            e = hipStreamEndCapture(cap, &graph);
        }
        c->stream = saved;
        (void)hipStreamDestroy(cap);
        if ((rc != LMX_OK || e != hipSuccess) && graph) (void)hipGraphDestroy(graph);
        LMX_TRY(rc);
        LMX_HIP(e);
        LMX_HIP(hipGraphInstantiate(&exec, graph, nullptr, nullptr, 0));
    }
    std::vector<MixRhs> hs(nrhs);
    MixBatch hb;
    int launched = 0;
    int rc = LMX_OK;
    while (true) {
        LMX_HIP(hipMemcpyAsync(&hb, R.w.batch, sizeof(MixBatch), hipMemcpyDeviceToHost, c->stream));
        LMX_HIP(hipStreamSynchronize(c->stream));
        if (hb.n_active <= 0 || launched >= params->maxiter) break;
        if (exec) {
            if (hipGraphLaunch(exec, c->stream) != hipSuccess) {
                rc = LMX_ERR_HIP;
                set_error("hipGraphLaunch failed");
                break;
            }
            launched += chunk;
        } else {
            const int n = std::min(chunk, params->maxiter - launched);
            for (int k = 0; k < n && rc == LMX_OK; ++k) rc = enqueue_iteration(R, (launched + k) % kXBatch);
            if (rc != LMX_OK) break;
            launched += n;
        }
    }
    if (exec) (void)hipGraphExecDestroy(exec);
    if (graph) (void)hipGraphDestroy(graph);
    LMX_TRY(rc);
    // RHS stopped by maxiter (or retired by a breakdown): fold the inner
    // solution in and report the true residual; the ring slots marked in
    // pmask hold the still deferred alpha p
    enqueue_accum(R, grid, 1);
    LMX_TRY(enqueue_true_ddagd(R, x, nullptr, 1));
    hipLaunchKernelGGL(k_mx_fin<MX_TRUE>, dim3(nrhs), blk, 0, c->stream, R.w.partials, nblocks(c->V), R.w.rhs,
                       R.w.batch, 0.0, 0.0, 1);
    LMX_HIP(hipGetLastError());
    LMX_HIP(hipMemcpyAsync(hs.data(), R.w.rhs, nrhs * sizeof(MixRhs), hipMemcpyDeviceToHost, c->stream));
    LMX_HIP(hipStreamSynchronize(c->stream));
    for (int i = 0; i < nrhs; ++i) {
        const MixRhs& s = hs[i];
        lmx_cg_result& r = results[i];
        r.converged = s.converged;
        r.iterations = s.iterations;
        r.reason = s.done ? s.reason : LMX_CG_MAXITER;
        r.reliable_updates = s.updates;
        r.residual_norm = s.rtrue;
        r.relative_residual = s.bnorm == 0.0 ? (s.rtrue == 0.0 ? 0.0 : INFINITY) : s.rtrue / s.bnorm;
        r.true_relative_residual = r.relative_residual;
    }
    return LMX_OK;
}

// ---------------------------------------------------------------------------
// Autotuner: time one group of kXBatch inner iterations' operator pairs as
// the solver runs them (D of r with the direction update and |D p|^2,
// D^dagger with the fused r update, the x update deferred on all but the
// last, which adds the deferred terms from their own buffers) for every RHS
// blocking (RHS per launch in {1,2,4,6,12}, RHS per work item in {1,2,4})
// and streaming depth (t-slices per workgroup: automatic, 4, 8) and keep
// the fastest per (volume, precision, nrhs, clover term or not, arch);
// optionally persisted as text lines (cache_lookup).
// ---------------------------------------------------------------------------
static std::string arch_name(int dev) {
    hipDeviceProp_t prop;
    if (hipGetDeviceProperties(&prop, dev) != hipSuccess) return "unknown";
    std::string a = prop.gcnArchName;
    size_t c = a.find(':');
    return c == std::string::npos ? a : a.substr(0, c);
}

// A cached launch shape is used only if it is one this build can launch on
// this lattice: RHS per work item 1, 2 or 4 dividing the RHS per launch,
// t-slices per workgroup dividing Lt (at most kMaxTL), z-ranges 1, 2 or 4
// dividing the 256-site blocks of a t-slice, streaming fields only where the
// lattice streams.  Anything else (a foreign or damaged line) is retuned.
static bool block_launchable(const lmx_context* c, int blk) {
    if ((blk & kRhsLoop) != 0) {  // the RHS-loop shape: RHS per launch only
        const int rest = blk & ~kRhsLoop;
        return rhsloop_ok(c) && (rest & ~(0xff << 4)) == 1 && ((rest >> 4) & 0xff) >= 1;
    }
    if ((blk & kLdsStage) != 0) {  // the LDS-staged shape: one RHS per work item
        if (!lds_ok(c) || (blk & 15) != 1) return false;
        blk &= ~kLdsStage;
    }
    if (blk <= 0 || (blk >> 24) != 0) return false;
    const int nr = blk & 15, per = (blk >> 4) & 0xff, tl = (blk >> 12) & 0xff, zs = (blk >> 20) & 0xf;
    if (nr != 1 && nr != 2 && nr != 4) return false;
    if (per < nr || per % nr != 0) return false;
    const long V3 = (long)c->g.L[0] * c->g.L[1] * c->g.L[2];
    const bool streams = V3 % 256 == 0;
    if (tl != 0 && (!streams || tl > kMaxTL || c->g.L[3] % tl != 0)) return false;
    if (zs != 0 && (!streams || (zs != 1 && zs != 2 && zs != 4) || (V3 / 256) % zs != 0)) return false;
    return true;
}

// Cache lines: "lmx1 V precision nrhs clover arch rhs_block us_per_rhs"; the
// key is (V, precision, nrhs, clover or not, arch), the last matching
// launchable line wins.
static constexpr const char* kTuneTag = "lmx1";
static bool cache_lookup(const char* path, const lmx_context* c, int prec, int nrhs, int clover,
                         const std::string& arch, int* block, double* us) {
    if (!path) return false;
    FILE* f = fopen(path, "r");
    if (!f) return false;
    char line[256];
    bool found = false;
    while (fgets(line, sizeof(line), f)) {
        char tag[16], a[64];
        long v;
        int p, n, cl, blk;
        double t;
        if (sscanf(line, "%15s %ld %d %d %d %63s %d %lf", tag, &v, &p, &n, &cl, a, &blk, &t) != 8) continue;
        if (strcmp(tag, kTuneTag) != 0 || v != (long)c->V || p != prec || n != nrhs || cl != clover || arch != a)
            continue;
        if (!block_launchable(c, blk) || !(t > 0.0)) continue;
        *block = blk;
        *us = t;
        found = true;
    }
    fclose(f);
    return found;
}

}  // namespace lmx

namespace lmx {
// The buffers and launch arguments of one group of the mixed solver's inner
// iterations (D of r with the direction update and |D p|^2, then D^dagger
// with the r update and the x update deferred or applied), for timing launch
// shapes (lmx_autotune) and single kernels (lmx_inner_kernel_times).  Freed by
// the destructor.
struct InnerGroup {
    char* buf = nullptr;
    LoArgs ld, lx;
    ~InnerGroup() {
        if (buf) (void)hipFree(buf);
    }
};

static int inner_group_setup(lmx_wilson* op, int precision, int nrhs, InnerGroup& G) {
    lmx_context* c = op->ctx;
    LMX_TRY(ensure_lo_links(op));
    const int sp = spinor_prec_of(precision);
    // buffers of one group of inner iterations: p, t = D p, r, x_inner,
    // partials, alpha (one per RHS and deferred slot), done (zeros), the
    // deferred-slot masks (all slots marked), the deferred p slots (the first
    // also serves as p_prev) and beta (nonzero, so that the direction update
    // reads t and p_prev as in every iteration after the first)
    const size_t fb = align256(lmx_spinor_bytes(c, sp, nrhs));
    const size_t xb = align256(12 * (size_t)c->V * nrhs * sizeof(float2));
    const size_t nbk = (size_t)nblocks(c->V);
    const size_t pb = align256(2 * nbk * nrhs * sizeof(double));
    const size_t ab = align256((nrhs + kXBatch) * sizeof(double)), db = align256(nrhs * sizeof(int));
    const size_t bb = align256(nrhs * sizeof(double));
    const size_t total = 3 * fb + xb + pb + ab + 2 * db + (kXBatch - 1) * fb + bb;
    if (hipMalloc(&G.buf, total) != hipSuccess) {
        G.buf = nullptr;
        set_error("inner-iteration buffer allocation of %zu bytes failed", total);
        return LMX_ERR_NOMEM;
    }
    char* buf = G.buf;
    LMX_HIP(hipMemsetAsync(buf, 0, total, c->stream));
    char* pbuf = buf;
    char* tbuf = buf + fb;
    char* rbuf = buf + 2 * fb;
    char* xbuf = buf + 3 * fb;
    double* parts = (double*)(xbuf + xb);
    const double* alpha = (const double*)((char*)parts + pb);
    const int* done = (const int*)((const char*)alpha + ab);
    int* masks = (int*)((char*)done + db);
    char* qbuf = (char*)masks + db;  // kXBatch - 1 deferred p slots
    const double* beta = (const double*)(qbuf + (kXBatch - 1) * fb);
    static_assert(kXBatch >= 2 && kXBatch <= 8, "deferred-slot masks are filled bytewise");
    LMX_HIP(hipMemsetAsync(masks, (1 << (kXBatch - 1)) - 1, db, c->stream));  // slot bits 0..kXBatch-2 set
    LMX_HIP(hipMemsetAsync((void*)beta, 0x3f, bb, c->stream));  // bytes 0x3f: beta = 4.8e-4 per RHS
    auto norm_of = [&](char* q) -> float* {  // per-site scales follow the 16-bit data
        return sp == LMX_H16 ? (float*)(q + (size_t)nrhs * c->V * 3 * sizeof(int4)) : nullptr;
    };
    LoArgs& ld = G.ld;  // t = D r + beta t, p = r + beta p_prev, |t|^2
    ld = lo_args(op, link_prec_of(precision), sp);
    ld.nrhs = nrhs;
    ld.in = rbuf;
    ld.in_norm = norm_of(rbuf);
    ld.out = tbuf;
    ld.out_norm = norm_of(tbuf);
    ld.pprev = qbuf;
    ld.pprev_norm = norm_of(qbuf);
    ld.pnext = pbuf;
    ld.pnext_norm = norm_of(pbuf);
    ld.beta = beta;
    ld.astride = 1;
    ld.done = done;
    ld.dstride = 1;
    ld.epi = 4;
    ld.partials = parts;
    ld.pstride = 2L * (long)nbk;
    LoArgs& lx = G.lx;  // D^dagger t with the x / r update
    lx = lo_args(op, link_prec_of(precision), sp);
    lx.nrhs = nrhs;
    lx.in = tbuf;
    lx.in_norm = norm_of(tbuf);
    lx.aux = pbuf;
    lx.aux_norm = norm_of(pbuf);
    lx.flags = LMX_DAGGER;
    lx.epi = 3;
    lx.r = rbuf;
    lx.r_norm = norm_of(rbuf);
    lx.x = xbuf;
    lx.alpha = alpha;
    lx.alpha_q = alpha;
    lx.astride = 1;
    lx.done = done;
    lx.pmask = masks;  // every deferred slot marked
    lx.dstride = 1;
    for (int j = 0; j < kXBatch - 1; ++j) {
        lx.pq[j] = qbuf + (size_t)j * fb;
        lx.pq_norm[j] = norm_of(qbuf + (size_t)j * fb);
    }
    lx.partials = parts;
    lx.pstride = 2L * (long)nbk;
    return LMX_OK;
}
}  // namespace lmx

extern "C" int lmx_autotune(lmx_wilson* op, int precision, int nrhs, const char* cache_path, lmx_tune_result* out) {
    using namespace lmx;
    clear_error();
    LMX_CHECK(op != nullptr && op->ctx != nullptr && out != nullptr, LMX_ERR_ARG, "bad autotune arguments");
    LMX_CHECK(precision == LMX_F32 || precision == LMX_F16 || precision == LMX_H16, LMX_ERR_ARG,
              "autotune precision must be F32, F16 or H16");
    LMX_CHECK(nrhs >= 1 && nrhs <= 64, LMX_ERR_ARG, "nrhs must be in [1, 64]");
    lmx_context* c = op->ctx;
    LMX_CHECK(c->nc == 3, LMX_ERR_UNSUPPORTED, "the autotuner is built for NC = 3 fields (this context has NC = %d)",
              c->nc);
    const std::string arch = arch_name(c->device);
    int best = 0;
    double best_us = 0.0;
    const int clover = op->csw != 0.0 ? 1 : 0;
    if (cache_lookup(cache_path, c, precision, nrhs, clover, arch, &best, &best_us)) {
        op->tuned_block[precision][nrhs] = best;
        out->rhs_block = best;
        out->block_size = 256;
        out->ms_per_apply = best_us * 1e-3;
        return LMX_OK;
    }
    InnerGroup G;
    LMX_TRY(inner_group_setup(op, precision, nrhs, G));
    LoArgs& ld = G.ld;
    LoArgs& lx = G.lx;
    hipEvent_t e0, e1;
    (void)hipEventCreate(&e0);
    (void)hipEventCreate(&e1);
    auto group = [&]() {
        int rc2 = LMX_OK;
        for (int j = 0; j < kXBatch && rc2 == LMX_OK; ++j) {
            lx.apply_x = j == kXBatch - 1 ? 1 : 0;
            rc2 = launch_wilson_lo(c, ld);
            if (rc2 == LMX_OK) rc2 = launch_wilson_lo(c, lx);
        }
        return rc2;
    };
    const int per_opts[5] = {1, 2, 4, 6, 12};
    int rc = LMX_OK;
    best = 0;
    best_us = 1e30;
    // t-slices per streaming workgroup: automatic, 4 or 8; z-ranges per
    // XCD slab: automatic, 1 or 2 (streaming lattices only)
    const long V3 = (long)c->g.L[0] * c->g.L[1] * c->g.L[2];
    const bool streams = V3 % 256 == 0;
    const int tl_opts[3] = {0, 4, 8};
    const int zs_opts[3] = {0, 1, 2};
    for (int pi = 0; pi < 5 && rc == LMX_OK; ++pi) {
        const int per = per_opts[pi];
        if (per > nrhs) break;
        for (int ti = 0; ti < 3 && rc == LMX_OK; ++ti) {
        const int tl = tl_opts[ti];
        if (tl > 0 && (!streams || c->g.L[3] % tl != 0)) continue;
        for (int zi = 0; zi < 3 && rc == LMX_OK; ++zi) {
        const int zs = zs_opts[zi];
        if (zs > 0 && (!streams || (V3 / 256) % zs != 0)) continue;
        for (int nr = 1; nr <= 4 && rc == LMX_OK; ++nr) {
            if (nr == 3 || nr > per || per % nr != 0) continue;  // built: 1, 2, 4 per work item
            const int blk = (zs << 20) | (tl << 12) | (per << 4) | nr;
            ld.rhs_block = lx.rhs_block = blk;
            rc = group();
            const int reps = 2;
            (void)hipEventRecord(e0, c->stream);
            for (int k = 0; k < reps && rc == LMX_OK; ++k) rc = group();
            (void)hipEventRecord(e1, c->stream);
            if (hipEventSynchronize(e1) != hipSuccess) rc = LMX_ERR_HIP;
            float ms = 0.f;
            (void)hipEventElapsedTime(&ms, e0, e1);
            const double us = 1e3 * ms / reps / kXBatch / nrhs / 2;  // per operator application and RHS
            if (rc == LMX_OK && us < best_us) {
                best_us = us;
                best = blk;
            }
        }
        }
        }
    }
    // the LDS-staged shape: RHS per launch, t-slices and z-ranges as above
    for (int pi = 0; pi < 5 && rc == LMX_OK && lds_ok(c); ++pi) {
        const int per = per_opts[pi];
        if (per > nrhs) break;
        for (int ti = 0; ti < 3 && rc == LMX_OK; ++ti) {
            const int tl = tl_opts[ti];
            if (tl > 0 && c->g.L[3] % tl != 0) continue;
            for (int zi = 0; zi < 3 && rc == LMX_OK; ++zi) {
                const int zs = zs_opts[zi];
                if (zs > 0 && (V3 / 256) % zs != 0) continue;
                const int blk = kLdsStage | (zs << 20) | (tl << 12) | (per << 4) | 1;
                ld.rhs_block = lx.rhs_block = blk;
                rc = group();
                const int reps = 2;
                (void)hipEventRecord(e0, c->stream);
                for (int k = 0; k < reps && rc == LMX_OK; ++k) rc = group();
                (void)hipEventRecord(e1, c->stream);
                if (hipEventSynchronize(e1) != hipSuccess) rc = LMX_ERR_HIP;
                float ms = 0.f;
                (void)hipEventElapsedTime(&ms, e0, e1);
                const double us = 1e3 * ms / reps / kXBatch / nrhs / 2;
                if (rc == LMX_OK && us < best_us) {
                    best_us = us;
                    best = blk;
                }
            }
        }
    }
    // the RHS-loop shape (FP16 links): RHS per launch only
    for (int pi = 0; pi < 5 && rc == LMX_OK && precision != LMX_F32 && rhsloop_ok(c); ++pi) {
        const int per = per_opts[pi];
        if (per > nrhs) break;
        const int blk = kRhsLoop | (per << 4) | 1;
        ld.rhs_block = lx.rhs_block = blk;
        rc = group();
        const int reps = 2;
        (void)hipEventRecord(e0, c->stream);
        for (int k = 0; k < reps && rc == LMX_OK; ++k) rc = group();
        (void)hipEventRecord(e1, c->stream);
        if (hipEventSynchronize(e1) != hipSuccess) rc = LMX_ERR_HIP;
        float ms = 0.f;
        (void)hipEventElapsedTime(&ms, e0, e1);
        const double us = 1e3 * ms / reps / kXBatch / nrhs / 2;
        if (rc == LMX_OK && us < best_us) {
            best_us = us;
            best = blk;
        }
    }
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    LMX_TRY(rc);
    op->tuned_block[precision][nrhs] = best;
    if (cache_path) {
        FILE* f = fopen(cache_path, "a");
        if (f) {
            fprintf(f, "%s %ld %d %d %d %s %d %.3f\n", kTuneTag, (long)c->V, precision, nrhs, clover, arch.c_str(), best,
                    best_us);
            fclose(f);
        }
    }
    out->rhs_block = best;
    out->block_size = 256;
    out->ms_per_apply = best_us * 1e-3;
    return LMX_OK;
}

extern "C" int lmx_inner_kernel_times(lmx_wilson* op, int precision, int nrhs, int rhs_block, int groups,
                                      double us[2]) {
    using namespace lmx;
    clear_error();
    LMX_CHECK(op != nullptr && op->ctx != nullptr && us != nullptr, LMX_ERR_ARG, "bad inner_kernel_times arguments");
    LMX_CHECK(precision == LMX_F32 || precision == LMX_F16 || precision == LMX_H16, LMX_ERR_ARG,
              "precision must be F32, F16 or H16");
    LMX_CHECK(nrhs >= 1 && nrhs <= 64 && groups >= 1, LMX_ERR_ARG, "nrhs must be in [1, 64], groups positive");
    lmx_context* c = op->ctx;
    LMX_CHECK(c->nc == 3, LMX_ERR_UNSUPPORTED, "the mixed solver's kernels are built for NC = 3 fields");
    if (rhs_block == 0) rhs_block = op->tuned_block[precision][nrhs];
    InnerGroup G;
    LMX_TRY(inner_group_setup(op, precision, nrhs, G));
    G.ld.rhs_block = G.lx.rhs_block = rhs_block;
    // one event before the first launch and one after each launch: kernel k's
    // duration is the interval that ends at its own event
    const int nl = 2 * kXBatch * (groups + 1);
    std::vector<hipEvent_t> ev(nl + 1);
    for (auto& e : ev) LMX_HIP(hipEventCreate(&e));
    int rc = LMX_OK;
    int k = 0;
    LMX_HIP(hipEventRecord(ev[0], c->stream));
    for (int g = 0; g <= groups && rc == LMX_OK; ++g) {  // group 0 warms up
        for (int j = 0; j < kXBatch && rc == LMX_OK; ++j) {
            G.lx.apply_x = j == kXBatch - 1 ? 1 : 0;
            rc = launch_wilson_lo(c, G.ld);
            if (hipEventRecord(ev[++k], c->stream) != hipSuccess) rc = LMX_ERR_HIP;
            if (rc == LMX_OK) rc = launch_wilson_lo(c, G.lx);
            if (rc == LMX_OK && hipEventRecord(ev[++k], c->stream) != hipSuccess) rc = LMX_ERR_HIP;
        }
    }
    if (rc == LMX_OK && hipEventSynchronize(ev[nl]) != hipSuccess) rc = LMX_ERR_HIP;
    double tot[2] = {0.0, 0.0};
    for (int i = 2 * kXBatch; i < nl && rc == LMX_OK; ++i) {  // skip the warm-up group
        float ms = 0.f;
        if (hipEventElapsedTime(&ms, ev[i], ev[i + 1]) != hipSuccess) rc = LMX_ERR_HIP;
        tot[i % 2] += ms;
    }
    for (auto& e : ev) (void)hipEventDestroy(e);
    LMX_TRY(rc);
    const double per = (double)groups * kXBatch;
    us[0] = 1e3 * tot[0] / per;  // D of r with the direction update and |D p|^2
    us[1] = 1e3 * tot[1] / per;  // D^dagger with the r update, x deferred 3 of 4 times
    return LMX_OK;
}
