"""Builds libngp_hip.so (all HIP kernels + the C ABI) for gfx950 with hipcc.

Usage: python -m jaxngp_amd.build  [--jobs N]

Each source compiles to an object with its own flags, then one shared library is
linked in-tree next to this file, so it travels with the source tree.
"""
from __future__ import annotations

import argparse
import os
import shutil
import subprocess
import sys
from concurrent.futures import ThreadPoolExecutor
from pathlib import Path

PKG_DIR = Path(__file__).resolve().parent
REPO = PKG_DIR.parent
CSRC = PKG_DIR / "csrc"
INCLUDE = REPO / "include"
BUILD_DIR = PKG_DIR / "_build"
LIB_PATH = PKG_DIR / "libngp_hip.so"
ARCH = "gfx950"

COMMON_FLAGS = [
    f"--offload-arch={ARCH}",
    "-O3",
    "-fPIC",
    "-std=c++17",
    "-fno-gpu-rdc",
    # keep hipcc from forming packed-FP32 instructions
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    f"-I{INCLUDE}",
    f"-I{CSRC}",
    "-Wall",
    "-Wno-unused-function",
    "-Wno-unknown-pragmas",
]

# The marcher, integrator and encoder follow the oracle's rounding convention:
# every f32 operation rounded on its own.
SOURCES = {
    "status.hip": [],
    "packbits_morton.hip": ["-ffp-contract=off"],
    "march.hip": ["-ffp-contract=off"],
    "integrate.hip": ["-ffp-contract=off"],
    "hashgrid.hip": ["-ffp-contract=off"],
    "nerf_mlp.hip": ["-ffp-contract=off"],
    "harness.hip": ["-ffp-contract=off"],
    "proposal.hip": ["-ffp-contract=off"],
    "tcnn_hashgrid.hip": ["-ffp-contract=off"],
    "skysphere.hip": [],  # tolerance-judged against float64: FMAs allowed
}


def hipcc() -> str:
    for cand in (shutil.which("hipcc"), "/opt/rocm/bin/hipcc"):
        if cand and os.path.exists(cand):
            return cand
    raise RuntimeError("hipcc not found: the HIP toolchain is required to build libngp_hip.so")


def _compile(src: str, extra: list[str]) -> Path:
    obj = BUILD_DIR / (Path(src).stem + ".o")
    cmd = [hipcc(), *COMMON_FLAGS, *extra, "-c", str(CSRC / src), "-o", str(obj)]
    res = subprocess.run(cmd, capture_output=True, text=True)
    if res.returncode != 0:
        raise RuntimeError(f"hipcc failed for {src}:\n{' '.join(cmd)}\n{res.stdout}\n{res.stderr}")
    return obj


def build(jobs: int = 8, verbose: bool = True) -> Path:
    """Compiles every source and links libngp_hip.so; returns its path."""
    BUILD_DIR.mkdir(exist_ok=True)
    srcs = [s for s in SOURCES if (CSRC / s).exists()]
    with ThreadPoolExecutor(max_workers=max(1, jobs)) as ex:
        objs = list(ex.map(lambda s: _compile(s, SOURCES[s]), srcs))
    tmp = LIB_PATH.with_suffix(".so.tmp")
    cmd = [hipcc(), f"--offload-arch={ARCH}", "-shared", "-fPIC", "-o", str(tmp), *map(str, objs)]
    res = subprocess.run(cmd, capture_output=True, text=True)
    if res.returncode != 0:
        raise RuntimeError(f"link failed:\n{' '.join(cmd)}\n{res.stdout}\n{res.stderr}")
    os.replace(tmp, LIB_PATH)
    if verbose:
        print(f"built {LIB_PATH} from {len(objs)} sources", file=sys.stderr)
    return LIB_PATH


def main() -> None:
    ap = argparse.ArgumentParser()
    ap.add_argument("--jobs", type=int, default=8)
    args = ap.parse_args()
    build(jobs=args.jobs)


if __name__ == "__main__":
    main()
