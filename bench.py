"""Benchmark: forward synthesis + per-height Jacobians of a seeded cube.

Workload (one "step"): FORWARD3D with response functions over a 288x288
cube of columns, nz = 96 heights (dz = 15 km), the Fe I 6301.5/6302.5 pair
(one region, 112 wavelengths at 20 mA), Jacobians of Stokes IQUV w.r.t.
T, |B|, inclination, azimuth and v_los at every height.  Inputs (the seeded
atmosphere, Pg in hydrostatic equilibrium) are resident in HBM before timing;
each step runs the microphysics and radiative-transfer kernels and writes
the 71 GB Jacobian to HBM.

Prints one JSON line (rank 0).  Only N = 1 GPU is supported.
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "column-λ/s fwd+Jacobian"
PEAK_FP32_TFLOPS = 157.3
PEAK_HBM_GBS = 8000.0


def cpu_baseline(ls, nz, gain, seconds=12.0, threads=16):
    """The oracle (FP64 C restatement, OpenMP over columns) on a bounded sample
    of the same workload: columns of the seeded cube, full Jacobians."""
    import oracle
    from firtez_dz_amd import atmos
    cpu = oracle.Cpu()
    numw = cpu.setup(ls, resp=True, atmpar=(1, 1, 1, 1, 1, 1, 1, 0))
    cpu.set_gain(gain)
    atm, z = atmos.seeded(16, 16, nz, seed=2024)
    # hydrostatic Pg for the sample columns (not timed)
    models = []
    for c in range(atm.shape[2]):
        m = atmos.column_model(atm, z, c)
        pg, rho, pel, mw, _ = cpu.he(m[0], float(m[1, -1]), z)
        m[1], m[2], m[3], m[4] = pg, rho, pel, mw
        models.append(m)
    models = np.array(models)
    done = 0
    t0 = time.perf_counter()
    while time.perf_counter() - t0 < seconds:
        batch = models[: max(threads, 16)]
        cpu.forward_batch(batch, nthreads=threads)
        done += len(batch)
    dt = time.perf_counter() - t0
    return {"value": done * numw / dt, "unit": "column-λ/s", "cores": threads, "kind": "port",
            "sample": "%d columns (nz=%d, %d wavelengths, Jacobians of 7 parameters) of a seeded 16x16 cube, "
                      "FP64 oracle, %d OpenMP threads, %.1f s" % (done, nz, numw, threads, dt)}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--nx", type=int, default=288)
    ap.add_argument("--ny", type=int, default=288)
    ap.add_argument("--nz", type=int, default=96)
    ap.add_argument("--tile", type=int, default=8192)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--cpu-seconds", type=float, default=12.0)
    args = ap.parse_args()
    if args.gpus != 1:
        print("bench.py measures one GPU only: this build has no multi-GPU path (columns are independent; "
              "see the sharding design).", file=sys.stderr)
        sys.exit(2)

    import torch
    from firtez_dz_amd import atmos, flops, gpu, lines

    ls = lines.fe6301_6302(nwave=112, start=-624.0, step=20.0)
    print("[bench] seeding %dx%dx%d atmosphere" % (args.nx, args.ny, args.nz), file=sys.stderr, flush=True)
    atm, z = atmos.seeded(args.nx, args.ny, args.nz, dz_km=15.0, seed=2)
    ncol = args.nx * args.ny
    ctx = gpu.Context(0)
    numw = ctx.configure(ls, z)
    gain = ctx.gain_hsra()
    ctx.set_tile(args.tile, 256)
    d_atm = torch.from_numpy(atm).cuda()
    st = ctx.hydrostatic_torch(d_atm)
    torch.cuda.synchronize()
    assert int(st.max().item()) == 0, "hydrostatic integration failed"
    npar = gpu.NPARAM[gpu.RF_SPHERICAL]
    stokes = torch.empty((ncol, 4, numw), dtype=torch.float32, device="cuda")
    jac = torch.empty((ncol, npar, args.nz, 4, numw), dtype=torch.float32, device="cuda")
    status = torch.zeros(ncol, dtype=torch.int32, device="cuda")
    print("[bench] buffers: jacobian %.1f GB" % (jac.numel() * 4 / 1e9), file=sys.stderr, flush=True)
    ctx.set_profiling(True)
    for _ in range(args.warmup):
        ctx.forward_torch(d_atm, gpu.RF_SPHERICAL, stokes, jac, status)
    torch.cuda.synchronize()
    assert int(status.max().item()) == 0, "a column failed"
    micro_ms, rt_ms = [], []
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(args.steps):
        ctx.forward_torch(d_atm, gpu.RF_SPHERICAL, stokes, jac, status)
        m, r, nl = ctx.last_timing()
        micro_ms.append(m)
        rt_ms.append(r)
    torch.cuda.synchronize()
    elapsed = time.perf_counter() - t0
    ms_step = elapsed / args.steps * 1e3
    value = ncol * numw * args.steps / elapsed
    # roofline of the dominant kernel (radiative transfer), HIP-event timed
    rt_avg_ms = float(np.mean(rt_ms))
    nlaunch = (ncol + args.tile - 1) // args.tile
    ncomp = 16
    fl = flops.per_column_lambda(args.nz, 2, ncomp, npar)
    # per launch: the tile's columns x wavelengths x flop, over the average launch time
    launch_ms = rt_avg_ms / nlaunch
    achieved = fl * ncol * numw / (rt_avg_ms * 1e-3) / 1e12
    print("[bench] step %.1f ms; micro %.1f ms; rt %.1f ms (%d launches); %.0f kflop/column-lambda"
          % (ms_step, np.mean(micro_ms), rt_avg_ms, nlaunch, fl / 1e3), file=sys.stderr, flush=True)
    line = {
        "metric": METRIC, "value": value, "unit": "column-λ/s", "n_gpus": 1, "steps": args.steps,
        "warmup": args.warmup, "ms_per_step": ms_step, "higher_is_better": True, "scaling": "weak",
        "vs_baseline": None, "dtype": "f32 (f64 EOS/opacities)", "data": "synthetic (seeded atmosphere)",
        "config": {"workload": "forward+Jacobian %dx%d columns x nz=%d, Fe I 6301.5+6302.5, %d wavelengths, "
                               "d/d(T,|B|,incl,azim,v_los) per height" % (args.nx, args.ny, args.nz, numw),
                   "columns": ncol, "nz": args.nz, "wavelengths": numw, "parameters": npar,
                   "kernel_ms": {"micro": float(np.mean(micro_ms)), "rt": rt_avg_ms}},
        # compute-bound: FP32 arithmetic on the vector ALUs (no matrix work in
        # this kernel); priced against MI355X's 157.3 TFLOP/s FP32 peak
        "roofline": {"bound": "mfma", "pipe": "fp32 valu", "achieved": achieved, "peak": PEAK_FP32_TFLOPS,
                     "unit": "TFLOP/s", "frac": achieved / PEAK_FP32_TFLOPS, "traffic": None,
                     "kernel": "rt_kernel<1>", "launch_ms": launch_ms, "launches_per_step": nlaunch,
                     "flop_per_column_lambda": fl},
    }
    if not args.no_cpu:
        print("[bench] CPU baseline (oracle) ...", file=sys.stderr, flush=True)
        line["cpu_baseline"] = cpu_baseline(ls, args.nz, gain, seconds=args.cpu_seconds)
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
