"""Builds libvxl.so (gfx950) in-tree from csrc/*.hip with hipcc.

Flags that matter for parity and for these GPUs:
  -ffp-contract=off                 no FMA contraction: SAT and DDA arithmetic is bit-exact
                                    with the reference's binary32 operation order
  -fno-gpu-flush-denormals-to-zero  IEEE denormals, like the reference's host f32
  -Xclang -target-feature -Xclang -packed-fp32-ops
                                    no packed FP32 instructions are generated
"""
from __future__ import annotations

import concurrent.futures as cf
import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
BUILD = os.path.join(HERE, "_build")
LIB = os.path.join(HERE, "libvxl.so")
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
ARCH = os.environ.get("VXL_OFFLOAD_ARCH", "gfx950")

CFLAGS = [
    f"--offload-arch={ARCH}", "-O3", "-fPIC", "-std=c++17",
    "-ffp-contract=off", "-fno-fast-math", "-fno-gpu-flush-denormals-to-zero",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wall", "-Wno-unused-function", "-Wno-unused-variable",
]


def sources():
    return sorted(os.path.join(CSRC, f) for f in os.listdir(CSRC) if f.endswith(".hip"))


def _compile(src: str) -> str:
    obj = os.path.join(BUILD, os.path.basename(src) + ".o")
    cmd = [HIPCC, *CFLAGS, "-c", src, "-o", obj]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"hipcc failed for {src}:\n{r.stdout}\n{r.stderr}")
    return obj


PROBE_SRC = os.path.join(HERE, "probe", "valu_probe.hip")
PROBE_LIB = os.path.join(HERE, "libvxl_probe.so")


def build_probe(verbose: bool = False) -> str:
    """libvxl_probe.so: the VALU issue-rate probe bench.py --valu-probe runs (measurement
    infrastructure, not part of the libvxl C ABI)."""
    tmp = PROBE_LIB + ".tmp"
    r = subprocess.run([HIPCC, *CFLAGS, "-shared", "-o", tmp, PROBE_SRC], capture_output=True,
                       text=True)
    if r.returncode != 0:
        raise RuntimeError(f"hipcc failed for {PROBE_SRC}:\n{r.stdout}\n{r.stderr}")
    os.replace(tmp, PROBE_LIB)
    if verbose:
        print(f"built {PROBE_LIB}")
    return PROBE_LIB


def build(verbose: bool = False) -> str:
    os.makedirs(BUILD, exist_ok=True)
    srcs = sources()
    with cf.ThreadPoolExecutor(max_workers=min(8, len(srcs))) as ex:
        objs = list(ex.map(_compile, srcs))
    tmp = LIB + ".tmp"
    cmd = [HIPCC, f"--offload-arch={ARCH}", "-shared", "-fPIC", "-o", tmp, *objs]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"link failed:\n{r.stdout}\n{r.stderr}")
    os.replace(tmp, LIB)
    if verbose:
        print(f"built {LIB}")
    build_probe(verbose)
    return LIB


if __name__ == "__main__":
    build(verbose=True)
    sys.exit(0)
