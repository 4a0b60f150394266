"""Benchmark of the Wilson D^dagger D CG path on one MI355X.

Workload (N=1): 32^4 Wilson D^dagger D, weak-field SU(3) links
U = exp(i 0.2 H) (seeded), Gaussian source, kappa 0.124, antiperiodic in t,
FP64 true relative residual 1e-10, zero start.  One step = one full solve.
Inputs are generated on the host (numpy, seeded) and uploaded before the
timed region; the timed region holds only device work.

Prints one JSON line (rank 0).  Run: python bench.py [--steps K] [--warmup W]
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(ROOT, "latticematrices.jl_amd"))

HBM_PEAK_GBS = 8000.0  # MI355X HBM3E, 8 TB/s


def log(msg):
    print(msg, file=sys.stderr, flush=True)


def parse():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=3)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--dims", type=str, default="32,32,32,32")
    ap.add_argument("--kappa", type=float, default=0.124)
    ap.add_argument("--eps", type=float, default=0.2)
    ap.add_argument("--rtol", type=float, default=1e-10)
    ap.add_argument("--dslash-reps", type=int, default=50)
    ap.add_argument("--cpu-iters", type=int, default=12)
    ap.add_argument("--no-cpu", action="store_true")
    return ap.parse_args()


def time_dslash(lat, D, x, y, reps, flags_dagger=False):
    """Average duration of one D_W launch, measured with HIP events on the
    stream the library launches on (the context stream = torch's current)."""
    import torch
    import lmx
    A = lmx.adjoint(D) if flags_dagger else D
    for _ in range(3):
        lmx.mul_(y, A, x)
    s = torch.cuda.Event(enable_timing=True)
    e = torch.cuda.Event(enable_timing=True)
    s.record(lat.stream)
    for _ in range(reps):
        lmx.mul_(y, A, x)
    e.record(lat.stream)
    e.synchronize()
    return s.elapsed_time(e) / reps


def cpu_baseline(U, b, L, kappa, iters):
    """Oracle (C/OpenMP restatement of the reference CPU algorithm) on the
    same 32^4 inputs: a bounded sample of `iters` CG iterations."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import oracle as orc
    threads = orc.num_threads()
    op = orc.NormalOp(orc.OP_DDAGD_WILSON, L, U=U, kappa=kappa)
    t0 = time.perf_counter()
    _, res = orc.cg(op, b, rtol=1e-30, maxiter=iters, order=1)
    dt = time.perf_counter() - t0
    return dt, res.iterations, threads


def main():
    args = parse()
    if args.gpus != 1:
        print(json.dumps({"error": "this build runs on one GPU; --gpus must be 1 (multi-GPU sharding is a design "
                                   "note only)"}))
        sys.exit(2)
    import numpy as np
    import torch

    import lmx
    from lmx import generators as gen

    L = tuple(int(v) for v in args.dims.split(","))
    V = int(np.prod(L))
    log(f"bench: generating inputs {L} weak-field eps={args.eps}")
    U = gen.links(L, "weak", 20240601, eps=args.eps)
    b = gen.gaussian_spinor(L, 20240602)
    lat = lmx.Lattice(L, (1, 1, 1, -1))
    Ug = lat.gauge_from_numpy(U)
    D = lmx.WilsonDiracOperator4D(Ug, args.kappa)
    bg = lat.spinor_from_numpy(b)
    x, r, p, Ap, tmp = (lat.spinor() for _ in range(5))
    N = lmx.DdagDOp(D, tmp)

    # ---- dslash roofline (FP64 D_W, 960 B/site algorithmic) -------------
    t_ms = time_dslash(lat, D, bg, tmp, args.dslash_reps)
    bytes_per_launch = 960.0 * V
    achieved = bytes_per_launch / (t_ms * 1e-3) / 1e9
    log(f"bench: FP64 D_W {t_ms * 1e3:.1f} us/launch, {achieved:.0f} GB/s algorithmic")

    # ---- FP64 CG time-to-solution ---------------------------------------
    def solve():
        lmx.clear_matrix_(x)
        return lmx.cg_(x, N, bg, r, p, Ap, rtol=args.rtol, maxiter=20000)

    for w in range(args.warmup):
        res = solve()
        log(f"bench: warmup {w}: {res.iterations} iterations, converged={res.converged}")
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for k in range(args.steps):
        res = solve()
        log(f"bench: step {k}: {res.iterations} iterations")
    torch.cuda.synchronize()
    t_solve = (time.perf_counter() - t0) / args.steps
    # true residual in FP64
    lmx.mul_(Ap, N, x)
    lmx.axpby_(-1.0, bg, 1.0, Ap)
    true_rel = float(np.sqrt(lmx.dot(Ap, Ap).real / lmx.dot(bg, bg).real))
    log(f"bench: FP64 CG {t_solve:.3f} s/solve, {res.iterations} iterations, true rel residual {true_rel:.3e}")

    cpu = None
    if not args.no_cpu:
        dt, it, threads = cpu_baseline(U, b, L, args.kappa, args.cpu_iters)
        per_iter = dt / max(it, 1)
        cpu = {"value": per_iter * res.iterations, "unit": "s/RHS", "cores": threads, "kind": "port",
               "sample": f"{it} CG iterations of the 32^4 FP64 solve (oracle, {threads} OpenMP threads), "
                         f"{per_iter:.3f} s/iteration x {res.iterations} iterations"}
        log(f"bench: cpu baseline {per_iter:.3f} s/iteration on {threads} threads")

    out = {
        "metric": "D^dagger D CG time-to-solution/RHS @1e-10 (mixed+multiRHS vs FP64); dslash HBM GB/s",
        "value": t_solve,
        "unit": "s/RHS",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": t_solve * 1e3,
        "higher_is_better": False,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic (seeded weak-field SU(3) links eps=0.2, Gaussian source)",
        "config": {"workload": f"Wilson D^dagger D CG {'x'.join(map(str, L))}, kappa {args.kappa}, "
                               f"rtol {args.rtol}, FP64", "lattice": list(L), "kappa": args.kappa,
                   "rhs": 1, "precision": "fp64"},
        "iterations": res.iterations,
        "true_relative_residual": true_rel,
        "fp64_time_per_rhs_s": t_solve,
        "dslash_gbs": achieved,
        "roofline": {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": achieved / HBM_PEAK_GBS, "traffic": None,
                     "kernel": "k_wilson FP64 D_W", "bytes_per_launch": bytes_per_launch,
                     "us_per_launch": t_ms * 1e3},
        "cpu_baseline": cpu,
    }
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
