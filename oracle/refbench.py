"""CPU baseline on the host cores: the reference's own Fortran column path
(oracle/_ref/libfzref.so: chemical_equilibrium, cont_opacity, line_opacity,
absorption_matrix, rtesolver, ... compiled from the reference's sources by
oracle/ref/Makefile, driven by the harness's FORWARD1D column loop), one
process per core like the reference's MPI workers.

TEST INFRASTRUCTURE: only bench.py's cpu_baseline leg runs it, as a child
process (this module never touches the GPU).

  python -m oracle.refbench --nz 96 --seconds 10
prints one JSON object: column-wavelengths per second over all processes.
The process count defaults to the cores this process may run on
(os.sched_getaffinity), at most 16 (the host share of one GPU).
"""
from __future__ import annotations

import argparse
import json
import multiprocessing as mp
import os
import sys
import time

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)

ATMPAR = (1, 1, 1, 1, 1, 1, 1, 0)


def _setup(nz):
    import numpy as np
    import oracle
    from firtez_dz_amd import lines
    ref = oracle.Ref()
    ls = lines.fe6301_6302(nwave=112, start=-624.0, step=20.0)
    numw = ref.setup(ls, resp=True, atmpar=ATMPAR)
    return ref, numw, np


def _worker(args):
    nz, seconds, seed, gain = args
    from firtez_dz_amd import atmos
    ref, numw, np = _setup(nz)
    ref.set_gain(gain)
    atm, z = atmos.seeded(4, 4, nz, seed=seed)
    models = []
    for c in range(atm.shape[2]):
        m = atmos.column_model(atm, z, c)
        pg, rho, pel, mw, _ = ref.he(m[0], float(m[1, -1]), z)
        m[1], m[2], m[3], m[4] = pg, rho, pel, mw
        models.append(m)
    done = 0
    t0 = time.perf_counter()
    while time.perf_counter() - t0 < seconds:
        ref.forward1d(models[done % len(models)])
        done += 1
    return done, time.perf_counter() - t0, numw


def host_cores() -> tuple:
    """(cores to use, cores in this process's affinity mask)."""
    n = len(os.sched_getaffinity(0))
    return min(n, 16), n


def run(nz: int, seconds: float, procs: int = 0) -> dict:
    use, avail = host_cores()
    procs = procs or use
    ref, numw, np = _setup(nz)
    gain = ref.gain()
    del ref
    ctx = mp.get_context("fork")
    with ctx.Pool(procs) as pool:
        res = pool.map(_worker, [(nz, seconds, 3000 + i, gain) for i in range(procs)])
    rate = sum(d * nw / t for d, t, nw in res)
    cols = sum(d for d, _, _ in res)
    return {"value": rate, "unit": "column-λ/s", "cores": procs, "kind": "reference",
            "sample": "%d columns (nz=%d, %d wavelengths, Jacobians of 7 parameters) of seeded 4x4 cubes, the "
                      "reference's Fortran modules (amdflang -O3) driven by the harness's FORWARD1D loop, %d "
                      "processes x %.0f s on %d of the %d cores in the affinity mask"
                      % (cols, nz, numw, procs, seconds, procs, avail)}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--nz", type=int, default=96)
    ap.add_argument("--seconds", type=float, default=10.0)
    ap.add_argument("--procs", type=int, default=0, help="0: the cores in the affinity mask, at most 16")
    a = ap.parse_args()
    print(json.dumps(run(a.nz, a.seconds, a.procs)), flush=True)


if __name__ == "__main__":
    main()
