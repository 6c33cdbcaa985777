"""Compile the gfx950 HIP library (libpdq.so) in-tree with hipcc.

The translation units (C ABI, cooperative dense/isotropic kernel,
lane-per-instance isotropic kernel, block-diagonal kernel) compile in
parallel and link into one shared library.
"""

from __future__ import annotations

import concurrent.futures
import os
import subprocess

HERE = os.path.dirname(os.path.abspath(__file__))
SRC = os.path.join(HERE, "csrc")
OUT_DIR = os.path.join(HERE, "_lib")
LIB = os.path.join(OUT_DIR, "libpdq.so")
UNITS = [
    "pdq_capi.hip",
    "pdq_coop.hip",
    "pdq_lane.hip",
    "pdq_lane_filter.hip",
    "pdq_lane_fixedpoint.hip",
    "pdq_lane_fixedinterval.hip",
    "pdq_lane_filter_opts.hip",
    "pdq_lane_fixedpoint_opts.hip",
    "pdq_lane_fixedinterval_opts.hip",
    "pdq_lane_lml.hip",
    "pdq_lane_offgrid.hip",
    "pdq_post.hip",
    "pdq_lane_sample.hip",
    "pdq_bd.hip",
    "pdq_bd_filter.hip",
    "pdq_bd_fixedpoint.hip",
    "pdq_bd_fixedinterval.hip",
]

HIPCC_FLAGS = [
    "-O3",
    "-std=c++17",
    "-fPIC",
    "--offload-arch=gfx950",
    # never emit packed FP32 instructions (all arithmetic here is FP64 anyway)
    "-Xclang",
    "-target-feature",
    "-Xclang",
    "-packed-fp32-ops",
    "-Wall",
    "-Wno-unused-function",
    "-Wno-unused-variable",
    "-Wno-unused-result",
]


def _hipcc() -> str:
    return os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")


def _compile(unit: str, verbose: bool) -> str:
    obj = os.path.join(OUT_DIR, unit.replace(".hip", ".o"))
    cmd = [_hipcc(), *HIPCC_FLAGS, "-I", os.path.join(HERE, "..", "include"), "-c", os.path.join(SRC, unit), "-o", obj]
    if verbose:
        print(" ".join(cmd), flush=True)
    subprocess.run(cmd, check=True)
    return obj


def build(verbose: bool = False) -> str:
    """Always recompiles (the caller decides when a build is needed)."""
    os.makedirs(OUT_DIR, exist_ok=True)
    with concurrent.futures.ThreadPoolExecutor(max_workers=len(UNITS)) as ex:
        objs = list(ex.map(lambda u: _compile(u, verbose), UNITS))
    tmp = LIB + ".tmp"
    subprocess.run([_hipcc(), "-shared", "-fPIC", "--offload-arch=gfx950", *objs, "-o", tmp], check=True)
    os.replace(tmp, LIB)
    return LIB


if __name__ == "__main__":
    print(build(verbose=True))
