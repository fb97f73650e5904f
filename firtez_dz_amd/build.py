"""Build the HIP library (gfx950) in-tree: firtez_dz_amd/libfirtez_gpu.so.

  python -m firtez_dz_amd.build
"""
from __future__ import annotations

import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
LIB = os.path.join(HERE, "libfirtez_gpu.so")
SOURCES = [os.path.join(HERE, "csrc", "fz_gpu.hip")]
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")

FLAGS = [
    "--offload-arch=gfx950", "-O3", "-std=c++17", "-fPIC", "-shared",
    # code objects loadable by both the system ROCm runtime and the one bundled with PyTorch
    "-mcode-object-version=5",
    # no packed FP32 instructions anywhere in the device code
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    # FP32 division and sqrt through the hardware reciprocal / sqrt (a few ulp);
    # FP64 arithmetic is unaffected
    "-fno-hip-fp32-correctly-rounded-divide-sqrt",
    "-I", os.path.join(ROOT, "include"),
]


def build(verbose: bool = False) -> str:
    tmp = LIB + ".tmp"
    cmd = [HIPCC, *FLAGS, *SOURCES, "-o", tmp]
    if verbose:
        print(" ".join(cmd))
    res = subprocess.run(cmd, capture_output=True, text=True)
    if res.returncode != 0:
        sys.stderr.write(res.stdout + res.stderr)
        raise RuntimeError("hipcc failed")
    os.replace(tmp, LIB)
    return LIB


FLANG = "/opt/rocm/bin/amdflang"
CALLER_SRC = os.path.join(ROOT, "tests", "fortran", "fz_caller.f90")
CALLER = os.path.join(ROOT, "tests", "fortran", "fz_caller")


def build_fortran(verbose: bool = False) -> str:
    """Compile the iso_c_binding module and link the test caller
    (tests/fortran/fz_caller.f90) against it and libfirtez_gpu.so."""
    mdir = os.path.join(HERE, "fortran")
    src = os.path.join(mdir, "firtez_gpu_mod.f90")
    out = os.path.join(mdir, "firtez_gpu_mod.o")
    cmds = [[FLANG, "-c", "-O2", "-fPIC", "-J", mdir, src, "-o", out]]
    if os.path.exists(CALLER_SRC):
        cmds.append([FLANG, "-O2", "-J", mdir, CALLER_SRC, out, "-L", HERE, "-lfirtez_gpu",
                     "-Wl,-rpath,$ORIGIN/../../firtez_dz_amd", "-o", CALLER])
    for cmd in cmds:
        if verbose:
            print(" ".join(cmd))
        subprocess.check_call(cmd)
    return out


if __name__ == "__main__":
    print(build(verbose=True))
