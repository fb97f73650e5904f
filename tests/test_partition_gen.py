"""The generated partition-function code (firtez_dz_amd/csrc/fz_partition_gen.h)
is up to date with its generator and, compiled for the host, reproduces the
oracle's interpreter of the same programs (PARTITION_FUNCTION,
atom_database.f90:121-1036) bit for bit for all 92 elements."""
import ctypes
import os
import runpy
import subprocess
import tempfile

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
CSRC = os.path.join(ROOT, "firtez_dz_amd", "csrc")


def test_generated_header_is_current(tmp_path):
    gen = runpy.run_path(os.path.join(CSRC, "gen_partition.py"))
    out = tmp_path / "gen.h"
    gen["OUT"] = str(out)
    g = dict(gen)
    g["main"].__globals__["OUT"] = str(out)
    g["main"]()
    assert out.read_text() == open(os.path.join(CSRC, "fz_partition_gen.h")).read()


def test_generated_code_matches_interpreter(cpu, tmp_path):
    src = tmp_path / "pf.cpp"
    src.write_text(
        "#include <cmath>\n#define __device__\n#define __noinline__ __attribute__((noinline))\n"
        "#define __forceinline__ inline\nusing std::exp; using std::log;\n"
        '#include "%s"\n'
        'extern "C" void pf(int e, double T, double* u, double* du) { fz::partition_gen(e, T, u, du); }\n'
        % os.path.join(CSRC, "fz_partition_gen.h"))
    lib = tmp_path / "libpf.so"
    subprocess.run(["g++", "-O2", "-ffp-contract=off", "-Wno-unknown-pragmas", "-shared", "-fPIC", str(src), "-o",
                    str(lib)], check=True)
    L = ctypes.CDLL(str(lib))
    dp = ctypes.POINTER(ctypes.c_double)
    L.pf.argtypes = [ctypes.c_int, ctypes.c_double, dp, dp]
    u, du = np.zeros(3), np.zeros(3)
    temps = np.concatenate([np.linspace(1500.0, 60000.0, 700), [5040.0, 13000.0, 16200.0, 30000.0]])
    bad = 0
    for e in range(92):
        for T in temps:
            L.pf(e, float(T), u.ctypes.data_as(dp), du.ctypes.data_as(dp))
            ru, rdu = cpu.partition(e + 1, float(T))   # the oracle numbers elements from 1
            if not (np.array_equal(u, ru) and np.array_equal(du, rdu)):
                bad += 1
    assert bad == 0, bad
