"""Builds libsupernormal_amd.so (the gfx950 kernels behind include/sn_api.h).

Every .hip file under csrc/ is compiled on its own with hipcc for gfx950 and
the objects are linked into supernormal_amd/_lib/libsupernormal_amd.so, in the
source tree so that the library travels with the repository snapshot.

Files whose results must be bit-identical to the C oracle (marching,
occupancy, patch weights) are compiled with -ffp-contract=off; every file is
compiled without packed-fp32 instructions.

Usage:  python -m supernormal_amd.build [--jobs N] [--verbose]
"""
from __future__ import annotations

import argparse
import concurrent.futures as cf
import os
import shutil
import subprocess
import sys
from pathlib import Path

PKG = Path(__file__).resolve().parent
CSRC = PKG / "csrc"
OUT_DIR = PKG / "_lib"
LIB_NAME = "libsupernormal_amd.so"
ARCH = "gfx950"

# Files that restate integer/ordering-sensitive reference semantics.
NO_CONTRACT = {"ray_marching.hip", "occgrid.hip", "render_weight.hip", "render_step.hip", "sampler.hip",
               "step_kernels.hip"}
# Files whose fp32 divisions and square roots are the hardware reciprocal / square root
# (a few ulp) rather than the correctly rounded sequences: the fused rendering step,
# compared with its fp64 oracle at 2e-3 (its ~280 divisions per lane were a third of
# its VALU instructions; 202 -> 169 us per late-iteration launch).
FAST_DIV = {"render_step.hip"}

COMMON_FLAGS = [
    f"--offload-arch={ARCH}",
    "-O3",
    "-fPIC",
    "-std=c++17",
    "-munsafe-fp-atomics",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wno-unused-command-line-argument",
]


def hipcc() -> str:
    for cand in (os.environ.get("HIPCC"), "/opt/rocm/bin/hipcc", shutil.which("hipcc")):
        if cand and Path(cand).exists():
            return cand
    raise RuntimeError("hipcc not found (ROCm is required to build the HIP kernels)")


def _compile(src: Path, obj: Path, verbose: bool) -> None:
    flags = list(COMMON_FLAGS)
    if src.name in NO_CONTRACT:
        flags.append("-ffp-contract=off")
    if src.name in FAST_DIV:
        flags.append("-fno-hip-fp32-correctly-rounded-divide-sqrt")
    cmd = [hipcc(), *flags, "-I", str(CSRC), "-c", str(src), "-o", str(obj)]
    if verbose:
        print(" ".join(cmd), flush=True)
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"hipcc failed on {src.name}:\n{r.stdout}\n{r.stderr}")
    warn = "\n".join(l for l in r.stderr.splitlines()
                     if "packed-fp32-ops" not in l and l.strip())
    if verbose and warn:
        print(warn, flush=True)


def build(jobs: int = 8, verbose: bool = False) -> Path:
    OUT_DIR.mkdir(exist_ok=True)
    obj_dir = PKG / "_build"
    obj_dir.mkdir(exist_ok=True)
    srcs = sorted(CSRC.glob("*.hip"))
    objs = [obj_dir / (s.stem + ".o") for s in srcs]
    with cf.ThreadPoolExecutor(max_workers=max(1, jobs)) as ex:
        futs = [ex.submit(_compile, s, o, verbose) for s, o in zip(srcs, objs)]
        for f in futs:
            f.result()
    lib = OUT_DIR / LIB_NAME
    tmp = OUT_DIR / (LIB_NAME + ".tmp")
    cmd = [hipcc(), f"--offload-arch={ARCH}", "-shared", "-fPIC", "-o", str(tmp),
           *map(str, objs)]
    if verbose:
        print(" ".join(cmd), flush=True)
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"link failed:\n{r.stdout}\n{r.stderr}")
    os.replace(tmp, lib)
    return lib


def main(argv=None) -> int:
    ap = argparse.ArgumentParser(description=__doc__.splitlines()[0])
    ap.add_argument("--jobs", type=int, default=min(8, os.cpu_count() or 1))
    ap.add_argument("--verbose", action="store_true")
    a = ap.parse_args(argv)
    lib = build(a.jobs, a.verbose)
    print(lib)
    return 0


if __name__ == "__main__":
    sys.exit(main())
