"""Build recipe of libhsrq.so (HIP, gfx950) and the drop-in CLI driver.

    python -m hsrq3in_amd.build            # library + driver
"""
from __future__ import annotations

import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
CSRC = os.path.join(HERE, "csrc")
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")

HIP_FLAGS = [
    "--offload-arch=gfx950", "-O3", "-std=c++17", "-fPIC",
    # no implicit contraction: every code copy rounds alike (fma() is explicit)
    "-ffp-contract=off",
    # no packed-FP32 instructions in generated code
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wall",
]

SOURCES = ["hsrq_kernels.hip", "hsrq_gen.hip", "hsrq_cplx.hip", "hsrq_gemm.hip", "hsrq_sched.cpp",
           "hsrq_api.cpp", "hsrq_server.cpp", "hsrq_post.cpp", "hsrq_noalloc.hip"]
LINK = []


def _run(cmd):
    r = subprocess.run(cmd, cwd=CSRC, stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True)
    out = "\n".join(line for line in r.stdout.splitlines()
                    if "packed-fp32-ops" not in line and "not a recognized feature" not in line)
    if r.returncode != 0:
        raise RuntimeError(f"build failed: {' '.join(cmd)}\n{out}")
    return out


def build_library() -> str:
    out = os.path.join(HERE, "libhsrq.so")
    tmp = out + ".tmp"
    _run([HIPCC, *HIP_FLAGS, "-shared", *SOURCES, *LINK, "-o", tmp])
    os.replace(tmp, out)
    return out


def build_driver() -> str:
    """eigvecs: the reference's CLI (main.c) linked against libhsrq."""
    src = os.path.join(HERE, "driver", "main.c")
    out = os.path.join(HERE, "driver", "eigvecs")
    if not os.path.exists(src):
        return ""
    cmd = ["gcc", "-std=gnu99", "-O2", "-march=x86-64-v3", "-fopenmp", "-Wall",
           "-I", os.path.join(ROOT, "include"), src, "-o", out,
           "-L", HERE, "-lhsrq", f"-Wl,-rpath,{HERE}", "-Wl,-rpath,$ORIGIN/..", "-lm"]
    r = subprocess.run(cmd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"driver build failed:\n{r.stdout}")
    return out


def build_all():
    lib = build_library()
    drv = build_driver()
    return lib, drv


if __name__ == "__main__":
    for p in build_all():
        if p:
            print(p)
    sys.exit(0)
