"""Build liboct_hip.so (gfx950) in-tree with hipcc, and the host tile generator
liboct_hostgen.so with gcc.

    python -m oct_recon_af_v1_amd.build
"""
from __future__ import annotations

import os
import subprocess
import sys
from pathlib import Path

PKG = Path(__file__).resolve().parent
ROOT = PKG.parent
SRC = PKG / "csrc" / "oct_hip.hip"
LIB = PKG / "liboct_hip.so"
ARCH = os.environ.get("OCT_HIP_ARCH", "gfx950")

HIPCC_FLAGS = [
    f"--offload-arch={ARCH}", "-O3", "-std=c++17", "-shared", "-fPIC",
    # packed FP32 (v_pk_*_f32) is disabled for every HIP translation unit of this library
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-I", str(ROOT / "include"),
]


def hipcc() -> str:
    for c in (os.environ.get("HIPCC"), "/opt/rocm/bin/hipcc", "hipcc"):
        if c and (Path(c).exists() or c == "hipcc"):
            return c
    raise RuntimeError("hipcc not found")


HOSTGEN_SRC = PKG / "csrc" / "hostgen.c"
HOSTGEN_LIB = PKG / "liboct_hostgen.so"
# host generator: AVX2/FMA (present on every x86-64 host of an MI355X system), OpenMP threads
GCC_FLAGS = ["-O3", "-mavx2", "-mfma", "-fno-math-errno", "-fopenmp", "-fPIC", "-shared", "-std=c11"]


def build_hostgen(verbose: bool = False) -> Path:
    tmp = HOSTGEN_LIB.with_suffix(".so.tmp")
    cmd = [os.environ.get("CC", "gcc"), *GCC_FLAGS, str(HOSTGEN_SRC), "-o", str(tmp), "-lm"]
    if verbose:
        print(" ".join(cmd), flush=True)
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"gcc failed ({r.returncode}):\n{r.stderr[-6000:]}")
    os.replace(tmp, HOSTGEN_LIB)
    return HOSTGEN_LIB


def build(verbose: bool = False) -> Path:
    srcs = sorted((PKG / "csrc").glob("*.hpp")) + [SRC, ROOT / "include" / "oct_hip.h"]
    tmp = LIB.with_suffix(".so.tmp")
    cmd = [hipcc(), *HIPCC_FLAGS, str(SRC), "-o", str(tmp)]
    if verbose:
        print(" ".join(cmd), flush=True)
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"hipcc failed ({r.returncode}):\n{r.stderr[-6000:]}")
    os.replace(tmp, LIB)
    del srcs
    build_hostgen(verbose)
    return LIB


if __name__ == "__main__":
    print(build(verbose="-v" in sys.argv))
