"""Build the gfx950 shared library ``libtt_amd.so`` in-tree with hipcc.

``python -m tracetransform_amd.build`` (or ``__graft_entry__.build()``)
compiles every HIP/C++ source under ``csrc/`` into one shared library next
to this file.  The library exports exactly the C ABI of ``include/tt.h``.
"""
from __future__ import annotations

import os
import shutil
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
CSRC = os.path.join(HERE, "csrc")
LIB = os.path.join(HERE, "libtt_amd.so")
DEMO = os.path.join(HERE, "tt_demo")
ARCH = os.environ.get("TT_OFFLOAD_ARCH", "gfx950")

# Device code: no FMA contraction (median indices depend on exact float
# rounding), IEEE denormals, correctly rounded sqrt/div, and no packed-FP32
# instructions.
DEVICE_FLAGS = [
    f"--offload-arch={ARCH}",
    "-O3",
    "-std=c++17",
    "-ffp-contract=off",
    "-fno-fast-math",
    "-fno-gpu-flush-denormals-to-zero",
    "-fhip-fp32-correctly-rounded-divide-sqrt",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-fvisibility=hidden",
    "-fPIC",
    "-Wall",
    "-Wno-unused-command-line-argument",
]


def sources():
    out = []
    for name in sorted(os.listdir(CSRC)):
        if name.endswith((".hip", ".cpp")):
            out.append(os.path.join(CSRC, name))
    return out


def hipcc() -> str:
    path = shutil.which("hipcc") or "/opt/rocm/bin/hipcc"
    if not os.path.exists(path):
        raise RuntimeError("hipcc not found; the HIP toolchain is required")
    return path


def build(verbose: bool = False) -> str:
    objs = []
    objdir = os.path.join(HERE, "_build")
    os.makedirs(objdir, exist_ok=True)
    procs = []
    for src in sources():  # one hipcc per source, run side by side
        obj = os.path.join(objdir, os.path.basename(src) + ".o")
        lang = ["-x", "hip"] if src.endswith((".hip", ".cpp")) else []
        cmd = [hipcc(), *lang, *DEVICE_FLAGS, "-I", os.path.join(ROOT, "include"),
               "-c", src, "-o", obj]
        if verbose:
            print(" ".join(cmd), flush=True)
        procs.append((cmd, subprocess.Popen(cmd)))
        objs.append(obj)
    failed = [cmd for cmd, pr in procs if pr.wait() != 0]
    if failed:
        raise subprocess.CalledProcessError(1, failed[0])
    tmp = LIB + ".tmp"
    cmd = [hipcc(), f"--offload-arch={ARCH}", "-shared", "-fPIC", "-o", tmp, *objs,
           "-L/opt/rocm/lib", "-lrocsolver", "-lrocblas", "-Wl,-rpath,/opt/rocm/lib"]
    if verbose:
        print(" ".join(cmd), flush=True)
    subprocess.run(cmd, check=True)
    os.replace(tmp, LIB)
    build_demo(verbose)
    return LIB


def build_demo(verbose: bool = False) -> str:
    """Host-only C++ front end ``tt_demo`` over the library (cpp/demo.cpp),
    built with g++ against the HIP runtime API; it finds libtt_amd.so next to
    itself through an $ORIGIN rpath."""
    tmp = DEMO + ".tmp"
    cmd = ["g++", "-std=c++17", "-O2", "-Wall", "-D__HIP_PLATFORM_AMD__",
           "-I", os.path.join(ROOT, "include"), "-I", os.path.join(HERE, "cpp"),
           "-I", "/opt/rocm/include", os.path.join(HERE, "cpp", "demo.cpp"),
           "-L", HERE, "-ltt_amd", "-L/opt/rocm/lib", "-lamdhip64",
           "-Wl,-rpath,$ORIGIN", "-Wl,-rpath,/opt/rocm/lib", "-o", tmp]
    if verbose:
        print(" ".join(cmd), flush=True)
    subprocess.run(cmd, check=True)
    os.replace(tmp, DEMO)
    return DEMO


if __name__ == "__main__":
    print(build(verbose="-v" in sys.argv))
