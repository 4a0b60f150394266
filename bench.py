"""Benchmark of the Wilson D^dagger D CG path on one MI355X.

Workload (N=1): 32^4 Wilson D^dagger D, weak-field SU(3) links
U = exp(i 0.2 H) (seeded), kappa 0.124, antiperiodic in t, the reference
benchmark's Gaussian sources (source r: randomize_gaussian_matrix! with
sigma = sqrt(1/2), seed 1235 + r, Philox4x32, generated on the device),
FP64 true relative residual 1e-10, zero start.
  - plain FP64 GPU CG (the reference algorithm cg!, one RHS), and
  - the mixed-precision multi-RHS CG (FP16 links, inner vectors as 16-bit
    fixed point with a per-site scale, FP32 math; FP64 reliable updates /
    reductions / solution) on `--nrhs` sources.
One step = one full solve of the mixed multi-RHS batch; `value` is its
time-to-solution per RHS.  Links are generated on the host (numpy, seeded)
and uploaded, sources on the device, all before the timed region; the timed
region holds device work only (plus the solver's host convergence polls).

Also measured, each with HIP events on the library stream: the mixed solve's
two inner stencil kernels at the autotuned shape (`roofline`: D^dagger with
the fused r / x update, the dominant kernel of the timed solve; and D of r
with the direction update and |Dp|^2), the FP64 D_W launch (960 B/site algorithmic: links 576 + psi 192 +
out 192), and a plain 1 GiB copy kernel used to calibrate the rocprof byte
counters.  The CPU baseline is one full solve of the same FP64 system by the
reference's own CPU algorithm (oracle/: padded storage, halo refresh per
operator application, unfused BLAS-1, OpenMP threads).

Prints one JSON line.  Run: python bench.py [--steps K] [--warmup W]
"""
from __future__ import annotations

import argparse
import json
import math
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
    ap.add_argument("--steps", type=int, default=2)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--dims", type=str, default="32,32,32,32")
    ap.add_argument("--kappa", type=float, default=0.124)
    ap.add_argument("--eps", type=float, default=0.2)
    ap.add_argument("--rtol", type=float, default=1e-10)
    ap.add_argument("--nrhs", type=int, default=12)
    ap.add_argument("--precision", type=str, default="h16", choices=["h16", "f16", "f32"])
    ap.add_argument("--alt-precision", type=str, default="f32,f16",
                    help="comma-separated inner precisions (f32, f16, h16) timed once each after the headline solve "
                         "(reported, not the value); 'none' for none")
    ap.add_argument("--rhs-block", type=int, default=0)
    ap.add_argument("--graph", type=int, default=1)
    ap.add_argument("--dslash-reps", type=int, default=50)
    ap.add_argument("--cpu-maxiter", type=int, default=20000)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--fp64-steps", type=int, default=1)
    ap.add_argument("--output", type=str, default=None,
                    help="append the benchmark's result rows (the reference's CSV columns) to this CSV file")
    ap.add_argument("--mode", type=str, default="full", choices=["full", "kernels"],
                    help="kernels: only the dslash / copy kernel timings (for profiler counter runs)")
    return ap.parse_args()


def timed(lat, fn, reps):
    """Average duration of fn() over reps launches, HIP events on the library stream."""
    import torch
    for _ in range(3):
        fn()
    s = torch.cuda.Event(enable_timing=True)
    e = torch.cuda.Event(enable_timing=True)
    s.record(lat.stream)
    for _ in range(reps):
        fn()
    e.record(lat.stream)
    e.synchronize()
    return s.elapsed_time(e) / reps


def measured_traffic(name, L, key="traffic_bytes_per_launch", nrhs=None):
    """HBM bytes per launch of a kernel from a committed rocprofv3 FETCH_SIZE /
    WRITE_SIZE summary (profiles/<name>.json, calibrated by a 1 GiB copy);
    None if absent or measured on another workload."""
    try:
        with open(os.path.join(ROOT, "profiles", name + ".json")) as f:
            t = json.load(f)
    except (OSError, ValueError):
        return None, None
    if tuple(L) != (32, 32, 32, 32) or (nrhs is not None and t.get("nrhs") != nrhs) or key not in t:
        return None, None
    return float(t[key]), f"profiles/{name}.md"


def inner_kernel_bytes(precision, n):
    """Algorithmic bytes per site of one launch of each inner stencil over n
    RHS (links once per launch): D with the direction update stencils r, reads
    the previous t and p, writes t = D p and p; D^dagger with the fused update
    reads t, reads and writes r, and once in four (averaged here) reads the
    four p of the group and reads and writes the FP32 x_inner (192 B)."""
    links = 288.0 if precision == "f32" else 144.0
    spinor = 52.0 if precision == "h16" else 96.0
    return links + n * 5 * spinor, links + n * (3 * spinor + (4 * spinor + 192.0) / 4)


def lo_dslash_bytes(precision, n):
    """Algorithmic bytes per site per RHS of one operator application,
    averaged as the autotuner times them: a group of four inner iterations,
    each D with the direction update (stencils r, reads t and p_prev, writes t
    and p) and D^dagger with the r update (reads t and r, writes r), the last
    D^dagger also adding the four alpha p (reads four p) to x_inner (FP32,
    read and written); links once per launch of n RHS.  Per application:
    links / n + 4 spinors + (4 spinors + 192) / 8."""
    links = 288.0 if precision == "f32" else 144.0
    spinor = 52.0 if precision == "h16" else 96.0
    group = 8 * links / n + 32 * spinor + 4 * spinor + 2 * 96.0
    return group / 8


def copy_calibration(lat):
    """A plain device copy of 1 GiB (read 1 GiB + write 1 GiB): the reference
    kernel for scaling the rocprof FETCH_SIZE / WRITE_SIZE counters."""
    import torch
    n = (1 << 30) // 8
    a = torch.ones(n, dtype=torch.float64, device=lat.device)
    b = torch.empty_like(a)
    ms = timed(lat, lambda: b.copy_(a), 10)
    del a, b
    return ms


def result_rows(L, n, args, solves):
    """The reference benchmark's VERIFY / RESULT lines (benchmark/ddagd_cg/ddagd_cg.jl:883-892) for each
    timed solve, printed on stderr (stdout keeps the one JSON line), and its CSV row (:835-880, the columns
    that apply to this build) appended to --output.  solves: (precision label, iterations, [seconds per
    solve], true relative residual, converged)."""
    import csv
    import datetime
    import torch
    dev = torch.cuda.get_device_name(0).replace(" ", "_")
    lattice = "x".join(map(str, L))
    verify_rtol = 10 * args.rtol
    rows = []
    for prec, iters, secs, true_res, conv in solves:
        med = sorted(secs)[len(secs) // 2]
        passed = bool(conv) and math.isfinite(true_res) and true_res <= verify_rtol
        ms_it = 0.0 if iters == 0 else 1e3 * med / iters
        log(f"VERIFY passed={str(passed).lower()} converged={str(bool(conv)).lower()} "
            f"true_relative_residual={true_res:.9e} threshold={verify_rtol:.9e}")
        log(f"RESULT operator=wilson backend=hip device={dev} device_ordinal=0 precision={prec} ranks=1 "
            f"threads=1 lattice={lattice} grid=1x1x1x1 iterations={iters} median_seconds={med:.9f} "
            f"ms_per_iteration={ms_it:.9f} true_residual={true_res:.9e} passed={str(passed).lower()}")
        rows.append({"timestamp": datetime.datetime.now(datetime.timezone.utc).strftime("%Y-%m-%dT%H:%M:%S.%f")[:-3] + "Z",
                     "operator": "wilson", "backend": "hip", "device": dev, "precision": prec, "mpi_ranks": 1,
                     "threads_per_rank": 1, "lattice": lattice, "process_grid": "1x1x1x1", "local_lattice": lattice,
                     "halo": 0, "rtol": args.rtol, "atol": 0.0, "verify_rtol": verify_rtol, "maxiter": 20000,
                     "repeats": len(secs), "nrhs": n if prec != "Float64" else 1, "iterations": iters,
                     "converged": bool(conv), "true_relative_residual": true_res, "passed": passed,
                     "min_solve_seconds": min(secs), "median_solve_seconds": med,
                     "max_solve_seconds": max(secs), "milliseconds_per_iteration": ms_it,
                     "torch_version": torch.__version__, "hip_version": str(torch.version.hip)})
    if args.output:
        new = not os.path.exists(args.output)
        with open(args.output, "a", newline="") as f:
            w = csv.DictWriter(f, fieldnames=list(rows[0].keys()))
            if new:
                w.writeheader()
            w.writerows(rows)
        log(f"CSV appended: {os.path.abspath(args.output)}")


def host_cpu():
    """CPU model, sockets and physical cores of the host (lscpu)."""
    import subprocess
    info = {}
    try:
        out = subprocess.run(["lscpu"], capture_output=True, text=True, timeout=20).stdout
        for line in out.splitlines():
            k, _, v = line.partition(":")
            info[k.strip()] = v.strip()
    except (OSError, subprocess.SubprocessError):
        pass
    try:
        cores = int(info["Core(s) per socket"]) * int(info["Socket(s)"])
    except (KeyError, ValueError):
        cores = None
    return {"model": info.get("Model name", "unknown"), "physical_cores": cores,
            "logical_cpus": os.cpu_count(), "affinity_cpus": len(os.sched_getaffinity(0))}


def cpu_baseline(U, b, L, kappa, rtol, maxiter, gpu_iterations):
    """The same FP64 solve (zero start, rtol on the recursive residual, as
    cg!) by the reference's own CPU algorithm, restated in C with OpenMP:
    LatticeMatrix padded storage (nw = 1), set_halo! before every operator
    application, the half-spin kernel reading neighbours from the halo,
    unfused BLAS-1 and per-thread reductions (the JACC threads backend).  One
    whole solve is timed (setup excluded); threads = OMP_NUM_THREADS."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import oracle as orc
    threads = orc.num_threads()
    host = host_cpu()
    x, res, dt = orc.refcpu_cg_ddagd(U, b, L, kappa, rtol=rtol, maxiter=maxiter)
    return {"value": dt, "unit": "s/RHS", "cores": threads, "kind": "port",
            "sample": f"one whole FP64 D^dagger D CG solve of the benchmark's RHS 0 (32^4, kappa {kappa}, rtol "
                      f"{rtol:g}, zero start): {res.iterations} iterations ({'converged' if res.converged else res.reason}; "
                      f"the GPU FP64 CG took {gpu_iterations}) on the reference's CPU algorithm restated in C/OpenMP "
                      f"(padded storage, halo refresh per application, unfused BLAS-1), {threads} threads on "
                      f"{host['model']} ({host['physical_cores']} physical cores, {host['affinity_cpus']} CPUs "
                      f"available to this process)",
            "host": host, "iterations": res.iterations}, x


def require_converged(name, results, rtol):
    """A time-to-solution counts only for solves that reached the tolerance:
    every RHS converged with its FP64 true relative residual within rtol
    (the plain FP64 CG reports its recursive residual, as cg! does)."""
    bad = [i for i, q in enumerate(results) if not q.converged]
    worst = max((q.true_relative_residual if q.true_relative_residual >= 0 else q.relative_residual)
                for q in results)
    if bad or not worst <= rtol:
        print(json.dumps({"error": f"{name} solve did not converge: RHS {bad} not converged, worst relative "
                                   f"residual {worst:.3e} (tolerance {rtol:g})"}), flush=True)
        sys.exit(1)


def main():
    args = parse()
    if args.gpus != 1:
        print(json.dumps({"error": "this build runs on one GPU; --gpus must be 1 (the multi-GPU sharding of the "
                                   "lattice is a design note only)"}))
        sys.exit(2)
    import numpy as np
    import torch

    import lmx
    from lmx import generators as gen

    L = tuple(int(v) for v in args.dims.split(","))
    V = int(np.prod(L))
    n = args.nrhs
    log(f"bench: generating inputs {L} weak-field eps={args.eps}, {n} Gaussian sources")
    U = gen.links(L, "weak", 20240601, eps=args.eps)
    lat = lmx.Lattice(L, (1, 1, 1, -1))
    Ug = lat.gauge_from_numpy(U)
    D = lmx.WilsonDiracOperator4D(Ug, args.kappa)
    # benchmark/ddagd_cg/ddagd_cg.jl:773-774 with --seed = 1234 + r for source r
    bg = lat.spinor(n)
    lmx.randomize_gaussian_matrix_(bg, [lmx.RNGStreamKey(1235 + r) for r in range(n)], sigma=0.5 ** 0.5)
    b0 = bg.rhs(0)
    x1, r, p, Ap, tmp = (lat.spinor() for _ in range(5))
    N = lmx.DdagDOp(D, tmp)

    # ---- dslash roofline (FP64 D_W) and calibration copy ------------------
    t_d64 = timed(lat, lambda: lmx.mul_(tmp, D, b0), args.dslash_reps)
    bytes_d64 = 960.0 * V
    ach_d64 = bytes_d64 / (t_d64 * 1e-3) / 1e9
    log(f"bench: FP64 D_W {t_d64 * 1e3:.1f} us/launch, {ach_d64:.0f} GB/s algorithmic")
    t_copy = copy_calibration(lat)
    log(f"bench: 1 GiB copy {t_copy * 1e3:.1f} us ({2 * (1 << 30) / (t_copy * 1e-3) / 1e9:.0f} GB/s)")

    if args.mode == "kernels":
        tune = lmx.autotune(D, args.precision, n, None)
        bytes_dm = lo_dslash_bytes(args.precision, n)
        log(f"bench: {args.precision} fused D / D^dagger pair x{n}: {tune['ms_per_apply'] * 1e3:.1f} us per application "
            f"per RHS, {bytes_dm * V / (tune['ms_per_apply'] * 1e-3) / 1e9:.0f} GB/s algorithmic")
        # the two inner stencils launched on their own at the tuned shape, five
        # groups of four inner iterations (the first warms up): the dispatches
        # a profile of this run attributes to them
        kt = lmx.inner_kernel_times(D, args.precision, n, tune["rhs_block"], groups=4)
        bD, bX = inner_kernel_bytes(args.precision, n)
        log(f"bench: {args.precision} inner D {kt['D']:.1f} us ({bD * V / (kt['D'] * 1e-6) / 1e9:.0f} GB/s), "
            f"D^dagger with the update {kt['Ddag']:.1f} us ({bX * V / (kt['Ddag'] * 1e-6) / 1e9:.0f} GB/s) per launch")
        torch.cuda.synchronize()
        print(json.dumps({"mode": "kernels", "fp64_dslash_us": t_d64 * 1e3, "fp64_dslash_gbs": ach_d64,
                          "copy_1gib_us": t_copy * 1e3, "lo_dslash_us_per_rhs": tune["ms_per_apply"] * 1e3,
                          "lo_rhs_block": tune["rhs_block"], "inner_D_us": kt["D"], "inner_Ddag_us": kt["Ddag"]}),
              flush=True)
        return

    # ---- plain FP64 CG (the reference algorithm), RHS 0 -------------------
    def solve64():
        lmx.clear_matrix_(x1)
        return lmx.cg_(x1, N, b0, r, p, Ap, rtol=args.rtol, maxiter=20000)

    res64 = solve64()
    torch.cuda.synchronize()
    t64s = []
    for _ in range(args.fp64_steps):
        t0 = time.perf_counter()
        res64 = solve64()
        torch.cuda.synchronize()
        t64s.append(time.perf_counter() - t0)
    t64 = sum(t64s) / len(t64s)
    log(f"bench: FP64 CG {t64:.3f} s/RHS, {res64.iterations} iterations, converged={res64.converged}")
    require_converged("FP64", [res64], args.rtol)
    x64 = x1.to_numpy()
    lmx.mul_(Ap, N, x1)  # the FP64 true residual |b - A x| / |b| (true_relative_residual!, ddagd_cg.jl:825)
    b0n = b0.to_numpy()
    true64 = float(np.linalg.norm(Ap.to_numpy() - b0n) / np.linalg.norm(b0n))

    # ---- mixed multi-RHS CG ------------------------------------------------
    tune = lmx.autotune(D, args.precision, n, None)  # measured in this run, never read from a cache
    bytes_dm = lo_dslash_bytes(args.precision, n)
    ach_dm = bytes_dm * V / (tune["ms_per_apply"] * 1e-3) / 1e9
    blk = tune["rhs_block"]
    kind = "LDS-staged" if blk & (1 << 29) else ("RHS-loop" if blk & (1 << 28) else "streaming")
    log(f"bench: autotuned {args.precision} fused D / D^dagger pair x{n}: {kind}, TL {(blk >> 12) & 0xff}, z-ranges "
        f"{(blk >> 20) & 0xf} (0: automatic), {(blk >> 4) & 0xff} RHS per launch, {blk & 15} per work item: "
        f"{tune['ms_per_apply'] * 1e3:.1f} us per application per RHS, {ach_dm:.0f} GB/s algorithmic")
    kt = lmx.inner_kernel_times(D, args.precision, n, tune["rhs_block"], groups=4)
    bD, bX = inner_kernel_bytes(args.precision, n)
    log(f"bench: {args.precision} inner D {kt['D']:.1f} us, D^dagger with the update {kt['Ddag']:.1f} us per launch")
    xm = lat.spinor(n)

    def solve_mixed():
        lmx.clear_matrix_(xm)
        return lmx.cg_batched_(xm, D, bg, precision=args.precision, rtol=args.rtol, use_graph=bool(args.graph),
                               rhs_block=args.rhs_block)

    for w in range(args.warmup):
        resm = solve_mixed()
        log(f"bench: warmup {w}: iterations {[q.iterations for q in resm]}")
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    tsteps = []
    for k in range(args.steps):
        tk = time.perf_counter()
        resm = solve_mixed()
        torch.cuda.synchronize()
        tsteps.append(time.perf_counter() - tk)
        log(f"bench: step {k}: iterations {[q.iterations for q in resm]} "
            f"updates {[q.reliable_updates for q in resm]}")
    torch.cuda.synchronize()
    tm = (time.perf_counter() - t0) / args.steps
    per_rhs = tm / n
    xm0 = xm.to_numpy()[0]
    err0 = float(np.linalg.norm(xm0 - x64) / np.linalg.norm(x64))
    true_res = [q.true_relative_residual for q in resm]
    require_converged(f"mixed {args.precision}", resm, args.rtol)
    log(f"bench: mixed {args.precision} x{n}: {tm:.3f} s/batch = {per_rhs:.4f} s/RHS; speedup vs FP64 "
        f"{t64 / per_rhs:.2f}x; max true rel residual {max(true_res):.3e}; RHS0 error vs FP64 {err0:.3e}")
    # the reference benchmark's result rows: the FP64 solve, and the mixed batch per RHS
    result_rows(L, n, args, [("Float64", res64.iterations, t64s, true64, res64.converged),
                             (f"{args.precision}-mixed-x{n}", max(q.iterations for q in resm),
                              [t / n for t in tsteps], max(true_res), all(q.converged for q in resm))])

    # the configuration's own inner format (FP32 spinors and links) and FP16
    # links with FP32 spinors, each timed once after the headline solve
    alt = []
    for prec in [q for q in args.alt_precision.split(",") if q and q != "none" and q != args.precision]:
        if prec not in ("f32", "f16", "h16"):
            raise SystemExit(f"unknown precision {prec}")
        lmx.autotune(D, prec, n, None)

        def solve_alt():
            lmx.clear_matrix_(xm)
            return lmx.cg_batched_(xm, D, bg, precision=prec, rtol=args.rtol, use_graph=bool(args.graph))

        solve_alt()
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        resa = solve_alt()
        torch.cuda.synchronize()
        ta = (time.perf_counter() - t0) / n
        require_converged(f"mixed {prec}", resa, args.rtol)
        erra = float(np.linalg.norm(xm.to_numpy()[0] - x64) / np.linalg.norm(x64))
        a = {"precision": prec, "s_per_rhs": ta, "speedup_vs_fp64": t64 / ta,
             "iterations": [q.iterations for q in resa], "reliable_updates": [q.reliable_updates for q in resa],
             "true_relative_residual_max": max(q.true_relative_residual for q in resa), "rhs0_error_vs_fp64": erra}
        alt.append(a)
        log(f"bench: mixed {prec} x{n} (one timed solve): {ta:.4f} s/RHS; speedup vs FP64 {t64 / ta:.2f}x; "
            f"iterations {min(a['iterations'])}-{max(a['iterations'])}; max true rel residual "
            f"{a['true_relative_residual_max']:.3e}; RHS0 error vs FP64 {erra:.3e}")

    traffic, traffic_src = measured_traffic("k_stream_fp64_dw_32x32x32x32", L)
    pair_name = f"k_lds_{args.precision}_pair_32x32x32x32_{n}rhs"
    tr_D, tr_src = measured_traffic(pair_name, L, "D_traffic_bytes_per_launch", n)
    tr_X, _ = measured_traffic(pair_name, L, "Ddag_traffic_bytes_per_launch", n)
    cpu = None
    if not args.no_cpu:
        log("bench: cpu baseline: one FP64 solve on the reference's CPU algorithm ...")
        cpu, xc = cpu_baseline(U, bg.rhs(0).to_numpy(), L, args.kappa, args.rtol, args.cpu_maxiter, res64.iterations)
        cpu["rhs0_error_vs_gpu_fp64"] = float(np.linalg.norm(xc - x64) / np.linalg.norm(x64))
        log(f"bench: cpu baseline {cpu['value']:.1f} s: {cpu['sample']}; error vs the GPU FP64 solution "
            f"{cpu['rhs0_error_vs_gpu_fp64']:.2e}")

    def roofline(kernel, bytes_per_launch, us, tr, src):
        ach = bytes_per_launch / (us * 1e-6) / 1e9
        return {"bound": "hbm", "achieved": ach, "peak": HBM_PEAK_GBS, "unit": "GB/s", "frac": ach / HBM_PEAK_GBS,
                "traffic": tr, "kernel": kernel, "bytes_per_launch": bytes_per_launch, "us_per_launch": us,
                "traffic_source": src}

    lo_name = {"h16": "SpH16In, LinkF16, SpH16Out", "f16": "SpF32In, LinkF16, SpF32Out",
               "f32": "SpF32In, LinkF32, SpF32Out"}[args.precision]
    fam = "k_lds" if blk & (1 << 29) else ("k_rhsloop" if blk & (1 << 28) else "k_stream")
    roof_X = roofline(f"{fam}<{lo_name}, ..., XrEx, +1, ...> (D^dagger with the r / x update, {n} RHS, the mean "
                      f"of the deferring and x-applying launches as the solver runs them; the dominant kernel of "
                      f"the timed solve)", bX * V, kt["Ddag"], tr_X, tr_src)
    roof_D = roofline(f"{fam}<{lo_name}, ..., TpEx, -1, ...> (D of r with the direction update and |Dp|^2, "
                      f"{n} RHS)", bD * V, kt["D"], tr_D, tr_src)
    roof_64 = roofline("k_stream<SpF64In, LinkF64, SpF64Out, F64Clover, SpF64In, NoEx, -1, 0, 1, 0, false> "
                       "(FP64 D_W, 32^4; the reliable updates and the FP64 CG)", bytes_d64, t_d64 * 1e3, traffic,
                       traffic_src)

    out = {
        "metric": "D^dagger D CG time-to-solution/RHS @1e-10 (mixed+multiRHS vs FP64); dslash HBM GB/s",
        "value": per_rhs,
        "unit": "s/RHS",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": tm * 1e3,
        "higher_is_better": False,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": {"h16": "16-bit fixed-point inner spinors, f16 links, link products on f16 operands with f32 accumulation, f32 math otherwise, f64 reliable updates",
                  "f16": "f16 links, f32 inner spinors and math, f64 reliable updates",
                  "f32": "f32 inner links, spinors and math, f64 reliable updates"}[args.precision],
        "data": "synthetic (seeded weak-field SU(3) links eps=0.2; the reference benchmark's Gaussian sources, "
                "Philox seed 1235 + r, sigma sqrt(1/2), generated on the device)",
        "config": {"workload": f"Wilson D^dagger D CG {'x'.join(map(str, L))}, kappa {args.kappa}, rtol "
                               f"{args.rtol}, {n} RHS mixed vs 1 RHS FP64",
                   "lattice": list(L), "kappa": args.kappa, "rhs": n, "precision": args.precision},
        "fp64_time_per_rhs_s": t64,
        "fp64_iterations": res64.iterations,
        "speedup_vs_fp64": t64 / per_rhs,
        "mixed_iterations": [q.iterations for q in resm],
        "mixed_reliable_updates": [q.reliable_updates for q in resm],
        "mixed_true_relative_residual_max": max(true_res),
        "mixed_rhs0_error_vs_fp64": err0,
        "dslash_gbs": ach_d64,
        "copy_1gib_us": t_copy * 1e3,
        "mixed_dslash": {"kernels": "D with the direction update and |Dp|^2 + D^dagger with the fused x/r "
                                    "update, per application",
                         "rhs_block": tune["rhs_block"], "us_per_rhs": tune["ms_per_apply"] * 1e3,
                         "algorithmic_bytes_per_site_per_rhs": bytes_dm, "achieved_gbs": ach_dm,
                         "frac": ach_dm / HBM_PEAK_GBS},
        "roofline": roof_X,
        "roofline_kernels": {"inner_Ddag": roof_X, "inner_D": roof_D, "fp64_dw": roof_64},
        "alt_precision": alt,
        "cpu_baseline": cpu,
    }
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
