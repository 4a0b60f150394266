"""Build liblmx.so (gfx950) from the HIP/C++ sources in this directory.

Each source is compiled separately (in parallel) with hipcc and linked into a
shared library next to the Python package.  Flags:
  -ffp-contract=off   plain complex arithmetic unfused, fma() only where the
                      reference's muladd is (bit-exact FP64 parity);
  -Xclang -target-feature -Xclang -packed-fp32-ops   no packed FP32 ops.
"""
from __future__ import annotations

import concurrent.futures as cf
import os
import shutil
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
PKG = os.path.join(os.path.dirname(HERE), "lmx")
OBJ = os.path.join(HERE, "_build")
# (source, object name, extra defines): the low-precision stencil is one
# object per (precision pair, RHS per work item) so those build in parallel.
SOURCES = [(s, os.path.splitext(s)[0], []) for s in
           ["wilson.hip", "api.cpp", "blas.hip", "clover.hip", "force.hip", "layout.hip", "mixed.hip", "rng.hip",
            "wilson_nc.hip"]]
SOURCES[1:1] = [("wilson_lo.hip", f"wilson_lo_c{c}_nr{n}", [f"-DLMX_LO_COMBO={c}", f"-DLMX_LO_NR={n}"])
                for c in (0, 1, 2) for n in (1, 2, 4)]
ARCH = os.environ.get("LMX_OFFLOAD_ARCH", "gfx950")


def hipcc() -> str:
    for c in (os.environ.get("HIPCC"), "/opt/rocm/bin/hipcc", shutil.which("hipcc")):
        if c and os.path.exists(c):
            return c
    raise RuntimeError("hipcc not found")


def flags() -> list[str]:
    return ["-O3", "-std=c++17", "-fPIC", f"--offload-arch={ARCH}", "-ffp-contract=off",
            "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
            "-Wall", "-Wno-unused-function", "-Wno-unused-variable", "-Wno-unknown-pragmas",
            "-I", os.path.join(os.path.dirname(os.path.dirname(HERE)), "include")]


def _compile(job) -> str:
    src, name, defs = job
    out = os.path.join(OBJ, name + ".o")
    cmd = [hipcc()] + flags() + defs + ["-c", os.path.join(HERE, src), "-o", out]
    if src.endswith(".cpp"):
        cmd[1:1] = ["-x", "hip"]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"compile failed: {' '.join(cmd)}\n{r.stdout}\n{r.stderr}")
    return out


def build(verbose: bool = False) -> str:
    os.makedirs(OBJ, exist_ok=True)
    workers = max(1, min(8, os.cpu_count() or 1, len(SOURCES)))
    with cf.ThreadPoolExecutor(max_workers=workers) as ex:
        objs = list(ex.map(_compile, SOURCES))
    lib = os.path.join(PKG, "liblmx.so")
    tmp = lib + ".tmp"
    cmd = [hipcc(), "-shared", f"--offload-arch={ARCH}", "-o", tmp] + objs
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"link failed: {' '.join(cmd)}\n{r.stdout}\n{r.stderr}")
    os.replace(tmp, lib)
    if verbose:
        print(f"built {lib}")
    return lib


if __name__ == "__main__":
    build(verbose=True)
    sys.exit(0)
