"""Builds libgsr.so (HIP, gfx950) in-tree with hipcc.

    python home-measure-privacy_amd/gsr/build.py

Every HIP source is compiled without FMA contraction (the backward replays the
forward's blend arithmetic bit for bit) and without packed FP32 instructions.
"""
from __future__ import annotations

import concurrent.futures as cf
import glob
import os
import shutil
import subprocess
import sys

PKG = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(PKG))
CSRC = os.path.join(os.path.dirname(PKG), "csrc")
INCLUDE = os.path.join(ROOT, "include")
LIB = os.path.join(PKG, "libgsr.so")
OBJ = os.path.join(os.path.dirname(PKG), "_build")
ARCH = os.environ.get("GSR_OFFLOAD_ARCH", "gfx950")

HIPCC_FLAGS = [
    f"--offload-arch={ARCH}", "-O3", "-std=c++17", "-fPIC", "-ffp-contract=off", "-munsafe-fp-atomics",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wall", "-Wno-unused-function", "-Wno-unused-variable",
    f"-I{INCLUDE}", f"-I{CSRC}",
]


# Per-file additions (later flags win).  The blend kernels keep -ffp-contract=off:
# the backward replays the forward's arithmetic bit for bit.  The projection keeps
# it too, with correctly rounded division and square root: its splat records feed
# the blend's discrete rules (the 1/255 skip, the 0.99 clamp, the T stop, the
# square cull), and a centre one ulp off the reference's operation order flips
# such a rule at some pixel of a 10k-Gaussian 256x256 frame, which moves that
# frame's mean/scale/rotation gradients by 1.5e-4 rel-L2 against the FP64
# oracle (5e-6 in the reference's order).  Only its SH colour sum, which feeds
# no rule, contracts (a pragma in sh_colour).  The chain feeds no rule: fused
# multiply-adds cut its instruction count by a quarter.
FILE_FLAGS = {
    "gsr_chain.hip": ["-ffp-contract=fast"],
}


def flags_for(src: str) -> list:
    return HIPCC_FLAGS + FILE_FLAGS.get(os.path.basename(src), [])


def hipcc() -> str:
    for cand in (shutil.which("hipcc"), "/opt/rocm/bin/hipcc"):
        if cand and os.path.exists(cand):
            return cand
    raise RuntimeError("hipcc not found")


def _compile(src: str) -> str:
    obj = os.path.join(OBJ, os.path.basename(src) + ".o")
    cmd = [hipcc(), *flags_for(src), "-c", src, "-o", obj]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"hipcc failed for {src}:\n{r.stdout}\n{r.stderr}")
    return obj


def build(verbose: bool = True) -> str:
    """Compiles every csrc/*.hip and links libgsr.so.  Always rebuilds."""
    os.makedirs(OBJ, exist_ok=True)
    srcs = sorted(glob.glob(os.path.join(CSRC, "*.hip")))
    with cf.ThreadPoolExecutor(max_workers=min(8, len(srcs))) as ex:
        objs = list(ex.map(_compile, srcs))
    tmp = LIB + ".tmp"
    cmd = [hipcc(), f"--offload-arch={ARCH}", "-shared", "-fPIC", "-o", tmp, *objs]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"link failed:\n{r.stdout}\n{r.stderr}")
    os.replace(tmp, LIB)
    if verbose:
        print(f"built {LIB}")
    build_calib(verbose)
    return LIB


CALIB = os.path.join(PKG, "libgsr_calib.so")


def build_calib(verbose: bool = True) -> str:
    """The bench's measurement kernels (csrc/calib), a library of their own."""
    src = os.path.join(CSRC, "calib", "gsr_calib.hip")
    tmp = CALIB + ".tmp"
    cmd = [hipcc(), *HIPCC_FLAGS, "-shared", src, "-o", tmp]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"hipcc failed for {src}:\n{r.stdout}\n{r.stderr}")
    os.replace(tmp, CALIB)
    if verbose:
        print(f"built {CALIB}")
    return CALIB


if __name__ == "__main__":
    build()
    sys.exit(0)
