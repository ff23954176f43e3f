"""Build libfcap.so (HIP, gfx950) in-tree.

    python -m freecappuccino_amd.build          # or __graft_entry__.build()

Every .hip translation unit under csrc/ is compiled with hipcc for gfx950
only, every .cpp (host-only code) with the host C++ compiler, and all are
linked into freecappuccino_amd/libfcap.so.  HIP flags:
  -ffp-contract=off   the FP64 parity kernels reproduce the reference's
                      rounding sequence and must not be contracted into FMAs
                      (kernels that want FMAs call fma() explicitly);
  -Xclang -target-feature -Xclang -packed-fp32-ops   no packed FP32 ops.
"""
from __future__ import annotations

import concurrent.futures as cf
import glob
import os
import shutil
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
CSRC = os.path.join(HERE, "csrc")
OUT = os.path.join(HERE, "libfcap.so")
BUILD_DIR = os.path.join(HERE, "_build")

HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
ARCH = "gfx950"
FLAGS = [
    f"--offload-arch={ARCH}", "-O3", "-fPIC", "-std=c++17", "-ffp-contract=off",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-I", os.path.join(ROOT, "include"), "-I", CSRC, "-Wall", "-Wno-unused-result",
]


# rocSOLVER (dense Cholesky of the coarse operator) and the rocBLAS it runs on
ROCM_LIB = os.path.join(os.path.dirname(os.path.dirname(HIPCC)), "lib")
LINK_LIBS = ["-L" + ROCM_LIB, "-Wl,-rpath," + ROCM_LIB, "-lrocsolver", "-lrocblas"]

CXX = os.environ.get("CXX", "g++")
HOST_FLAGS = ["-O3", "-fPIC", "-std=c++17", "-pthread", "-I", os.path.join(ROOT, "include"), "-I", CSRC, "-Wall"]


def _compile(src: str) -> str:
    obj = os.path.join(BUILD_DIR, os.path.basename(src) + ".o")
    if src.endswith(".cpp"):        # host-only code (the polyMesh reader)
        cmd = [CXX, *HOST_FLAGS, "-c", src, "-o", obj]
    else:
        cmd = [HIPCC, *FLAGS, "-c", src, "-o", obj]
    p = subprocess.run(cmd, capture_output=True, text=True)
    if p.returncode != 0:
        raise RuntimeError(f"{cmd[0]} failed for {src}:\n{p.stderr}")
    return obj


def build(verbose: bool = False) -> str:
    os.makedirs(BUILD_DIR, exist_ok=True)
    srcs = sorted(glob.glob(os.path.join(CSRC, "*.hip")) + glob.glob(os.path.join(CSRC, "*.cpp")))
    jobs = min(len(srcs), max(1, min(8, os.cpu_count() or 1)))
    with cf.ThreadPoolExecutor(jobs) as ex:
        objs = list(ex.map(_compile, srcs))
    tmp = OUT + ".tmp"
    cmd = [HIPCC, f"--offload-arch={ARCH}", "-shared", "-fPIC", "-pthread", *objs, "-o", tmp, *LINK_LIBS]
    p = subprocess.run(cmd, capture_output=True, text=True)
    if p.returncode != 0:
        raise RuntimeError(f"link failed:\n{p.stderr}")
    os.replace(tmp, OUT)
    if verbose:
        print(f"built {OUT}")
    return OUT


def clean() -> None:
    shutil.rmtree(BUILD_DIR, ignore_errors=True)


if __name__ == "__main__":
    build(verbose=True)
    sys.exit(0)
