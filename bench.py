"""Benchmark: forward synthesis + per-height Jacobians of a seeded cube.

Workload (one "step"): FORWARD3D with response functions over a 288x288
cube of columns, nz = 96 heights (dz = 15 km), the Fe I 6301.5/6302.5 pair
(one region, 112 wavelengths at 20 mA), Jacobians of Stokes IQUV w.r.t.
T, |B|, inclination, azimuth and v_los at every height.  Inputs (the seeded
atmosphere, Pg in hydrostatic equilibrium) are resident in HBM before timing;
each step runs the microphysics and radiative-transfer kernels and writes
the 71 GB Jacobian to HBM.

Prints one JSON line (rank 0).  Only N = 1 GPU is supported.

--lm measures the inversion instead: LM-converged pixels per second of the
whole INVERT3D schedule (fz_invert3d_dev) on a seeded cube (default 512x512,
nz = 64, the same two lines and 112 wavelengths), observations = forward of
the seeded truth + N(0, 1e-3 I_c) noise, T, Bx, By, Bz, v_z with at most 4
nodes each (two cycles: T 2 / others 1 node, then 4), started from the
plain initial model: the column-mean T(z) of the truth in hydrostatic
equilibrium, B = (100, 100, 100) G, v = 0.

--lm --learned adds the learned initial model: a 448-512-512-512-20 BF16
network trained on the device (pairs from the device forward on 8 seeded
512x512 cubes of other seeds, 2M pairs), its proposal inverted with the same
schedule, pixels above twice the chi^2 level of plain LM (calibrated on a
held-out 64x64 cube) inverted again from the plain model; the value is the
learned pipeline's pixels/s (network + LM + fallback), the plain run beside it.
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


def cpu_baseline_reference(nz, seconds=10.0):
    """The reference's own Fortran column path (oracle/_ref, compiled from its
    sources) on the host cores, one process per core of this process's
    affinity mask (at most 16), in a child process that never touches the
    GPU; None when the reference build is absent."""
    import subprocess
    lib = os.path.join(ROOT, "oracle", "_ref", "libfzref.so")
    if not os.path.exists(lib):
        return None
    r = subprocess.run([sys.executable, "-m", "oracle.refbench", "--nz", str(nz), "--seconds", str(seconds)],
                       cwd=ROOT, capture_output=True, text=True, timeout=seconds * 4 + 120)
    if r.returncode != 0:
        print("[bench] reference CPU baseline failed: " + r.stderr[-400:], file=sys.stderr, flush=True)
        return None
    return json.loads(r.stdout.strip().splitlines()[-1])


def oracle_error(ls, z, gain, atm_h, st_h, jac_h, cols, ids=None, xi=0.0):
    """Error of the timed cube against the FP64 oracle on sampled columns:
    Stokes |dS| in units of I_c; spherical Jacobians (the oracle's Cartesian
    response functions through the closed-form chain rule to |B|,
    inclination, azimuth) as |dJ| / max|J| of the parameter over the sampled
    map, the figure the tolerance applies to.  The same ratio with the max
    taken per column is reported beside it with its worst column: the
    inclination and azimuth responses scale with |B| and |B|^2, so in a
    weak-field column (a few G) they are 1e3 times smaller than in the rest
    of the map and the FP32 rounding of the kernel shows up at the percent
    level of that column's own scale.  Runs after the timed region, on the
    host."""
    import oracle
    from firtez_dz_amd import atmos
    cpu = oracle.Cpu()
    cpu.setup(ls, resp=True, atmpar=(1, 1, 1, 1, 1, 1, 1, 0))
    cpu.set_microturbulence(xi)
    cpu.set_gain(gain)
    models = np.stack([atmos.column_model(atm_h, z, int(c)) for c in cols])
    _, syn, dsyn = cpu.forward_batch(models)
    ds = np.abs(st_h - np.transpose(syn, (0, 2, 1))).reshape(-1)
    refs, worst_col = [], (0.0, -1, -1, 0.0)
    for i, m in enumerate(models):
        tr = lambda p: np.transpose(dsyn[i, :, :, p, :], (0, 2, 1))
        bx, by, bz = (float(m[j, 0]) for j in (5, 6, 7))
        B = np.sqrt(bx * bx + by * by + bz * bz)
        g, f = np.arccos(bz / B), np.arctan2(by, bx)
        dx, dy, dzz = tr(3), tr(4), tr(5)
        ref = [tr(0), (bx * dx + by * dy + bz * dzz) / B,
               B * np.cos(g) * np.cos(f) * dx + B * np.cos(g) * np.sin(f) * dy - B * np.sin(g) * dzz,
               -B * np.sin(g) * np.sin(f) * dx + B * np.sin(g) * np.cos(f) * dy, tr(6)]
        refs.append(ref)
        for p in range(5):
            sc = np.abs(ref[p]).max()
            if sc > 0:
                e = float(np.abs(jac_h[i, p] - ref[p]).max() / sc)
                if e > worst_col[0]:
                    worst_col = (e, int(ids[i] if ids is not None else cols[i]), p, float(B))
    scale = [max(float(np.abs(r[p]).max()) for r in refs) for p in range(5)]
    e = np.concatenate([(np.abs(jac_h[i, p] - refs[i][p]) / scale[p]).reshape(-1)
                        for i in range(len(models)) for p in range(5) if scale[p] > 0])
    names = ("T", "|B|", "inclination", "azimuth", "v_los")
    return {"columns": len(cols), "stokes_max_abs_Ic": float(ds.max()), "stokes_p99_abs_Ic": float(np.percentile(ds, 99)),
            "jacobian_max_rel": float(e.max()), "jacobian_p99_rel": float(np.percentile(e, 99)),
            "jacobian_max_rel_per_column": {"value": worst_col[0], "column": worst_col[1],
                                            "parameter": names[worst_col[2]] if worst_col[2] >= 0 else None,
                                            "column_B_gauss": worst_col[3]},
            "tolerance": {"stokes_abs_Ic": 1e-4, "jacobian_rel": 2e-3},
            "oracle": "FP64 C restatement of the reference column path"}


def cpu_baseline(ls, nz, gain, seconds=12.0, threads=0, xi=0.0):
    """The oracle (FP64 C restatement, OpenMP over columns) on a bounded sample
    of the same workload: columns of the seeded cube, full Jacobians."""
    import oracle
    from firtez_dz_amd import atmos
    threads = threads or min(len(os.sched_getaffinity(0)), 16)
    os.environ["OMP_NUM_THREADS"] = str(threads)
    cpu = oracle.Cpu()
    numw = cpu.setup(ls, resp=True, atmpar=(1, 1, 1, 1, 1, 1, 1, 0))
    cpu.set_gain(gain)
    cpu.set_microturbulence(xi)
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


def lm_metric(n=128, nz=64, train_cubes=16, epochs=6, max_passes=3, log=None):
    """The inversion half of the metric on a smaller map: LM-converged pixels
    per second of plain INVERT3D and of the learned pipeline (network
    proposal + INVERT3D + plain-LM fallback), same GPU, same observations.

    Observations: the device forward of a seeded n x n x nz truth + N(0, 1e-3)
    noise (Fe I 6301.5+6302.5, 112 wavelengths).  Reference chi^2 of a pixel:
    the chi^2 of the truth's own 4-node projection (T, Bx, By, Bz, v_z each
    replaced by its least-squares fit on the hat basis of 4 nodes, the model
    the inversion can represent) against the same noisy observations.  A
    pixel is converged when its final chi^2 is at most twice its reference;
    the learned pipeline hands a pixel to plain LM from the default model
    when its chi^2 exceeds that bound (the fallback rule of the learned
    initial model).  The network is trained beforehand (not timed) on
    train_cubes seeded n x n cubes of other seeds."""
    import torch
    from firtez_dz_amd import atmos, gpu, learned, lines

    say = log or (lambda m: None)
    ncol = n * n
    ls = lines.fe6301_6302(nwave=112, start=-624.0, step=20.0)
    atm, z = atmos.seeded(n, n, nz, dz_km=15.0, seed=7)
    ctx = gpu.Context(0)
    try:
        numw = ctx.configure(ls, z)
        ctx.gain_hsra()
        truth = torch.from_numpy(atm).cuda()
        assert int(ctx.hydrostatic_torch(truth).max().item()) == 0
        st, _, status = ctx.forward_torch(truth, gpu.RF_NONE)
        gen = torch.Generator(device="cuda")
        gen.manual_seed(11)
        obs = (st.double().reshape(ncol, 4 * numw) + 1e-3 * torch.randn((ncol, 4 * numw), generator=gen,
                                                                        device="cuda", dtype=torch.float64))
        obs = obs.contiguous()
        # reference chi^2: the truth projected on the 4-node basis
        C = torch.from_numpy(gpu.int_coefs(nz, 4)).cuda()
        Pz = (C @ torch.linalg.pinv(C)).float()
        proj = truth.clone()
        for r in (0, 5, 6, 7, 10):
            proj[r] = Pz @ truth[r]
        pst, _, pstat = ctx.forward_torch(proj, gpu.RF_NONE)
        torch.cuda.synchronize()
        ref_chi2 = ((pst.double().reshape(ncol, -1) - obs) ** 2).mean(dim=1) / 1e-6
        bound = 2.0 * ref_chi2
        ctl = gpu.invert_control({"T": 4, "Bx": 4, "By": 4, "Bz": 4, "vz": 4}, noise=1e-3, smoothing=False)
        init = plain_init(ctx, truth)
        say("plain INVERT3D of %d pixels" % ncol)
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        m_plain, _, c_plain, rp = learned.invert(ctx, init, obs, ctl, n, n, None, learned=False)
        torch.cuda.synchronize()
        t_plain = time.perf_counter() - t0
        # the network (training not timed)
        ctx.mlp_create(4 * numw, 512, 3, 4, seed=1)
        X, T = learned.training_set(ctx, train_cubes, n, n, nz, seed0=1000, noise=1e-3)
        rep = learned.train(ctx, X, T, epochs=epochs, batch=1024, lr=1e-3)
        del X, T
        say("learned pipeline (network trained in %.1f s)" % rep.seconds)
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        m_lrn, _, c_lrn, rl = learned.invert(ctx, init, obs, ctl, n, n, bound, learned=True)
        torch.cuda.synchronize()
        t_lrn = time.perf_counter() - t0
        conv_p = int((c_plain <= bound).sum().item())
        conv_l = int((c_lrn <= bound).sum().item())
        # plain LM restarted from its own result towards the learned median chi^2
        target = float(c_lrn.median())
        passes, t_more, cur, med = 1, t_plain, m_plain, float(c_plain.median())
        conv_more = conv_p
        while med > target and passes < max_passes:
            torch.cuda.synchronize()
            t0 = time.perf_counter()
            cur, _, c2, _ = learned.invert(ctx, cur, obs, ctl, n, n, None, learned=False)
            torch.cuda.synchronize()
            t_more += time.perf_counter() - t0
            passes += 1
            med = float(c2.median())
            conv_more = int((c2 <= bound).sum().item())
        out = {
            "workload": "INVERT3D %dx%d pixels, nz=%d, Fe I 6301.5+6302.5, %d wavelengths, T/Bx/By/Bz/vz up to 4 "
                        "nodes (2 cycles), obs = forward(seeded truth) + N(0, 1e-3 Ic)" % (n, n, nz, numw),
            "converged_rule": "final chi2 <= 2 x chi2 of the truth's 4-node projection (per pixel)",
            "reference_chi2_median": float(ref_chi2.median()),
            "plain": {"seconds": t_plain, "pixels_per_s": ncol / t_plain, "converged_pixels": conv_p,
                      "converged_pixels_per_s": conv_p / t_plain, "median_chi2": float(c_plain.median()),
                      "lm_iterations": rp.iterations},
            "learned": {"seconds": t_lrn, "pixels_per_s": ncol / t_lrn, "converged_pixels": conv_l,
                        "converged_pixels_per_s": conv_l / t_lrn, "median_chi2": target,
                        "lm_iterations": rl.iterations, "fallback_pixels": rl.fallback_pixels,
                        "seconds_network": rl.seconds_init, "seconds_lm": rl.seconds_lm,
                        "seconds_fallback": rl.seconds_fallback},
            "network": {"shape": "%d-512-512-512-20" % (4 * numw), "pairs": train_cubes * ncol, "epochs": epochs,
                        "train_seconds": rep.seconds, "valid_loss": rep.valid_loss[-1] if rep.valid_loss else None},
            "plain_to_learned_median_chi2": {"passes": passes, "seconds": t_more, "median_chi2": med,
                                             "reached": med <= target, "converged_pixels": conv_more},
            "speedup_converged_pixels_per_s": (conv_l / t_lrn) / (conv_p / t_plain) if conv_p else None,
            "speedup_at_learned_median_chi2": (t_more / t_lrn) if med <= target else None,
        }
        return out
    finally:
        ctx.close()


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--nx", type=int, default=288)
    ap.add_argument("--ny", type=int, default=288)
    ap.add_argument("--nz", type=int, default=96)
    # radiative-transfer columns per launch: what fz_autotune picks on this
    # workload (--autotune re-runs it during warm-up)
    ap.add_argument("--tile", type=int, default=32768)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-lm", action="store_true", help="skip the inversion half of the metric")
    ap.add_argument("--lm-n", type=int, default=128, help="pixels per side of the inversion half")
    ap.add_argument("--autotune", action="store_true")
    ap.add_argument("--cpu-seconds", type=float, default=10.0)
    ap.add_argument("--cpu-kind", choices=("reference", "port"), default="reference",
                    help="CPU baseline: the reference's Fortran (oracle/_ref) or the C restatement")
    ap.add_argument("--lm", action="store_true", help="measure LM-converged pixels/s instead")
    ap.add_argument("--learned", action="store_true", help="with --lm: learned initial model vs plain")
    ap.add_argument("--train-cubes", type=int, default=8)
    ap.add_argument("--train-nx", type=int, default=512)
    ap.add_argument("--epochs", type=int, default=6)
    ap.add_argument("--max-passes", type=int, default=3)
    ap.add_argument("--uneven", action="store_true",
                    help="with --lm: per-pixel line sets and windows (8-line list, 1-4 lines, 60-300 wavelengths)")
    ap.add_argument("--micro", type=float, default=0.5, help="microturbulence (km/s) of the forward workload")
    ap.add_argument("--overlap", action="store_true",
                    help="microphysics of the next tile on a second stream during the current tile's transfer")
    ap.add_argument("--train-cubes-coupled", type=int, default=4,
                    help="with --coupled: seeded 256x256 cubes of PSF-degraded training pairs")
    ap.add_argument("--coupled", action="store_true",
                    help="PSF-coupled inversion of a 256x256x128 cube through a 31x31 Airy + wing PSF")
    args = ap.parse_args()
    if args.gpus != 1:
        print("bench.py measures one GPU only: this build has no multi-GPU path (columns are independent; "
              "see the sharding design).", file=sys.stderr)
        sys.exit(2)

    if args.coupled:
        return bench_coupled(args)
    if args.lm and args.uneven:
        return bench_uneven(args)
    if args.lm:
        return bench_lm(args)

    import torch
    from firtez_dz_amd import atmos, flops, gpu, lines

    ls = lines.fe6301_6302(nwave=112, start=-624.0, step=20.0)
    print("[bench] seeding %dx%dx%d atmosphere" % (args.nx, args.ny, args.nz), file=sys.stderr, flush=True)
    atm, z = atmos.seeded(args.nx, args.ny, args.nz, dz_km=15.0, seed=2)
    ncol = args.nx * args.ny
    ctx = gpu.Context(0)
    numw = ctx.configure(ls, z)
    gain = ctx.gain_hsra()
    xi = args.micro * 1.0e5
    ctx.set_microturbulence(xi)
    ctx.set_overlap(args.overlap)
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
    tiles = None
    if args.autotune:
        tiles = ctx.autotune_torch(d_atm, gpu.RF_SPHERICAL, stokes, jac, status)
        args.tile = tiles[0]
        print("[bench] autotuned tiles: radiative transfer %d columns, microphysics %d columns" % tiles,
              file=sys.stderr, flush=True)
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
    # HBM bytes of rt_kernel<1> per column of this workload (nz = 96, 112
    # wavelengths, spherical Jacobian) from the FETCH_SIZE and WRITE_SIZE
    # counters of the kernel (two rocprofv3 --pmc runs of bench.py, each
    # counter scaled by its reading for a device copy of known size); a
    # constant of the profile, not measured in this run
    traffic = None
    if args.nz == 96 and numw == 112:
        traffic = flops.RT_HBM_BYTES_PER_COLUMN * ncol / nlaunch
    achieved = fl * ncol * numw / (rt_avg_ms * 1e-3) / 1e12
    print("[bench] step %.1f ms; micro %.1f ms; rt %.1f ms (%d launches); %.0f kflop/column-lambda"
          % (ms_step, np.mean(micro_ms), rt_avg_ms, nlaunch, fl / 1e3), file=sys.stderr, flush=True)
    line = {
        "metric": METRIC, "value": value, "unit": "column-λ/s", "n_gpus": 1, "steps": args.steps,
        "warmup": args.warmup, "ms_per_step": ms_step, "higher_is_better": True, "scaling": "weak",
        "vs_baseline": None, "dtype": "f32 (f64 EOS/opacities)", "data": "synthetic (seeded atmosphere)",
        "config": {"workload": "forward+Jacobian %dx%d columns x nz=%d, Fe I 6301.5+6302.5, %d wavelengths, "
                               "d/d(T,|B|,incl,azim,v_los) per height, microturbulence %.2g km/s"
                               % (args.nx, args.ny, args.nz, numw, args.micro),
                   "columns": ncol, "nz": args.nz, "wavelengths": numw, "parameters": npar,
                   "kernel_ms": {"micro": float(np.mean(micro_ms)), "rt": rt_avg_ms},
                   "tiles": {"rt_columns": args.tile, "micro_columns": tiles[1] if tiles else None,
                             "autotuned": tiles is not None}},
        # compute-bound: FP32 arithmetic on the vector ALUs (no matrix work in
        # this kernel); priced against MI355X's published 157.3 TFLOP/s FP32
        # peak, which counts packed FP32 (this build emits none: 78.6 without)
        "roofline": {"bound": "valu", "pipe": "fp32 valu", "achieved": achieved, "peak": PEAK_FP32_TFLOPS,
                     "unit": "TFLOP/s", "frac": achieved / PEAK_FP32_TFLOPS, "traffic": traffic,
                     "traffic_source": flops.RT_HBM_SOURCE if traffic else None,
                     "kernel": "rt_kernel<1>", "launch_ms": launch_ms, "launches_per_step": nlaunch,
                     "flop_per_column_lambda": fl},
    }
    if not args.no_cpu:
        # the oracle as the checker of the timed cube (64 sampled columns, after the timed region)
        cols = np.sort(np.random.default_rng(64).choice(ncol, 64, replace=False))
        icols = torch.from_numpy(cols).cuda()
        err = oracle_error(ls, z, gain, d_atm[:, :, icols].cpu().numpy(), stokes[icols].cpu().numpy(),
                           jac[icols].cpu().numpy(), np.arange(64), ids=cols, xi=xi)
        print("[bench] error vs oracle (64 columns): Stokes max %.2e Ic, Jacobian max %.2e"
              % (err["stokes_max_abs_Ic"], err["jacobian_max_rel"]), file=sys.stderr, flush=True)
        line["error"] = err
        base = None
        if args.cpu_kind == "reference":
            print("[bench] CPU baseline (reference Fortran, one process per core) ...", file=sys.stderr, flush=True)
            base = cpu_baseline_reference(args.nz, seconds=args.cpu_seconds)
        if base is None:
            print("[bench] CPU baseline (oracle) ...", file=sys.stderr, flush=True)
            base = cpu_baseline(ls, args.nz, gain, seconds=args.cpu_seconds, xi=xi)
        line["cpu_baseline"] = base
    if not args.no_lm:
        # the second half of the metric: LM-converged pixels/s, learned vs plain
        del jac, stokes
        torch.cuda.empty_cache()
        print("[bench] inversion half: %dx%d pixels" % (args.lm_n, args.lm_n), file=sys.stderr, flush=True)
        line["lm"] = lm_metric(n=args.lm_n, log=lambda m: print("[bench-lm] " + m, file=sys.stderr, flush=True))
        lm = line["lm"]
        print("[bench] LM-converged pixels/s: plain %.3g (%d of %d), learned %.3g (%d); median chi2 %.3g vs %.3g"
              % (lm["plain"]["converged_pixels_per_s"], lm["plain"]["converged_pixels"], args.lm_n ** 2,
                 lm["learned"]["converged_pixels_per_s"], lm["learned"]["converged_pixels"],
                 lm["plain"]["median_chi2"], lm["learned"]["median_chi2"]), file=sys.stderr, flush=True)
    print(json.dumps(line), flush=True)


def lm_problem(args):
    """Seeded truth cube, its observations (forward + 1e-3 noise) and the
    plain initial model, all on the device."""
    import torch
    from firtez_dz_amd import atmos, gpu, lines

    nx = ny = 512 if args.nx == 288 else args.nx
    nz = 64 if args.nz == 96 else args.nz
    ncol = nx * ny
    ls = lines.fe6301_6302(nwave=112, start=-624.0, step=20.0)
    print("[bench-lm] seeding %dx%dx%d truth" % (nx, ny, nz), file=sys.stderr, flush=True)
    atm, z = atmos.seeded(nx, ny, nz, dz_km=15.0, seed=7)
    ctx = gpu.Context(0)
    numw = ctx.configure(ls, z)
    ctx.gain_hsra()
    truth = torch.from_numpy(atm).cuda()
    assert int(ctx.hydrostatic_torch(truth).max().item()) == 0
    st, _, status = ctx.forward_torch(truth, gpu.RF_NONE)
    torch.cuda.synchronize()
    assert int(status.max().item()) == 0
    gen = torch.Generator(device="cuda")
    gen.manual_seed(11)
    obs = (st.double().reshape(ncol, 4 * numw) +
           1e-3 * torch.randn((ncol, 4 * numw), generator=gen, device="cuda", dtype=torch.float64)).contiguous()
    init = plain_init(ctx, truth)
    # no SMOOTHING_B: B and v of the seeded truth are independent from pixel
    # to pixel, so the 1-2-1 filter would mix unrelated pixels at every cycle
    ctl = gpu.invert_control({"T": 4, "Bx": 4, "By": 4, "Bz": 4, "vz": 4}, noise=1e-3, smoothing=False)
    return ctx, truth, obs, init, ctl, nx, ny, nz, numw


def plain_init(ctx, truth):
    """Plain initial model: column-mean T(z), hydrostatic Pg from the mean top
    pressure, B = (100, 100, 100) G, v = 0."""
    import torch
    init = torch.zeros_like(truth)
    init[0] = truth[0].mean(dim=1, keepdim=True)
    init[1, -1] = truth[1, -1].mean()
    init[5] = 100.0
    init[6] = 100.0
    init[7] = 100.0
    assert int(ctx.hydrostatic_torch(init).max().item()) == 0
    torch.cuda.synchronize()
    return init


def lm_stats(model, truth, chi):
    return {"median_chi2": float(np.median(chi)), "p90_chi2": float(np.percentile(chi, 90)),
            "mean_abs_T_error_K": float((model[0] - truth[0]).abs().mean().item()),
            "mean_abs_vz_error_cm_s": float((model[10] - truth[10]).abs().mean().item())}


def bench_lm(args):
    import torch
    from firtez_dz_amd import learned

    ctx, truth, obs, init, ctl, nx, ny, nz, numw = lm_problem(args)
    ncol = nx * ny
    print("[bench-lm] inverting %d pixels (plain init)" % ncol, file=sys.stderr, flush=True)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    model_plain, _, chi2, rp = learned.invert(ctx, init, obs, ctl, nx, ny, None, learned=False)
    torch.cuda.synchronize()
    dt = time.perf_counter() - t0
    plain = dict(lm_stats(model_plain, truth, chi2.cpu().numpy()), lm_iterations=rp.iterations, seconds=dt,
                 pixels_per_s=ncol / dt)
    print("[bench-lm] plain: %.1f s, %d iterations, median chi2 %.3g" % (dt, rp.iterations, plain["median_chi2"]),
          file=sys.stderr, flush=True)
    workload = ("INVERT3D %dx%d pixels, nz=%d, Fe I 6301.5+6302.5, %d wavelengths, T/Bx/By/Bz/vz up to 4 nodes "
                "(2 cycles)" % (nx, ny, nz, numw))
    if not args.learned:
        line = {"metric": "LM-converged pixels/s (plain init)", "value": ncol / dt, "unit": "pixel/s", "n_gpus": 1,
                "steps": 1, "warmup": 0, "ms_per_step": dt * 1e3, "higher_is_better": True, "scaling": "weak",
                "vs_baseline": None, "dtype": "f64 LM step, f32 radiative transfer",
                "data": "synthetic (seeded truth + 1e-3 noise)",
                "config": dict({"workload": workload}, **plain)}
        print(json.dumps(line), flush=True)
        return

    # the network: trained on pairs from the device forward (cubes of other seeds)
    from firtez_dz_amd import gpu
    ctx.mlp_create(4 * numw, 512, 3, 4, seed=1)
    tc = args.train_cubes
    print("[bench-lm] training set: %d cubes of %dx%d" % (tc, args.train_nx, args.train_nx), file=sys.stderr,
          flush=True)
    t0 = time.perf_counter()
    X, T = learned.training_set(ctx, tc, args.train_nx, args.train_nx, nz, seed0=1000, noise=1e-3,
                                log=lambda m: print("[bench-lm] " + m, file=sys.stderr, flush=True))
    t_data = time.perf_counter() - t0
    rep = learned.train(ctx, X, T, epochs=args.epochs, batch=1024, lr=1e-3,
                        log=lambda m: print("[bench-lm] " + m, file=sys.stderr, flush=True))
    del X, T
    # the fallback threshold: twice the median chi^2 the learned pipeline
    # reaches on a held-out 64x64 cube of the same statistics (the level the
    # 4-node model reaches with this noise); pixels above it are the
    # network's failures and go to plain LM
    from firtez_dz_amd import atmos
    ca, _ = atmos.seeded(64, 64, nz, dz_km=15.0, seed=4242)
    ctruth = torch.from_numpy(ca).cuda()
    ctx.hydrostatic_torch(ctruth)
    cst, _, _ = ctx.forward_torch(ctruth, gpu.RF_NONE)
    g2 = torch.Generator(device="cuda")
    g2.manual_seed(12)
    cobs = (cst.double().reshape(64 * 64, -1) + 1e-3 * torch.randn(cst.reshape(64 * 64, -1).shape, generator=g2,
                                                                   device="cuda", dtype=torch.float64)).contiguous()
    _, _, cchi, _ = learned.invert(ctx, plain_init(ctx, ctruth), cobs, ctl, 64, 64, None, learned=True)
    noise_chi2 = float(cchi.median())
    threshold = 2.0 * noise_chi2
    print("[bench-lm] network trained in %.1f s (data %.1f s); calibration chi2 %.3g -> threshold %.3g"
          % (rep.seconds, t_data, noise_chi2, threshold), file=sys.stderr, flush=True)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    model, _, chi2, rl = learned.invert(ctx, init, obs, ctl, nx, ny, threshold, learned=True)
    torch.cuda.synchronize()
    dtl = time.perf_counter() - t0
    lrn = dict(lm_stats(model, truth, chi2.cpu().numpy()), lm_iterations=rl.iterations, seconds=dtl,
               pixels_per_s=ncol / dtl, fallback_pixels=rl.fallback_pixels,
               fallback_iterations=rl.fallback_iterations, seconds_network=rl.seconds_init,
               seconds_lm=rl.seconds_lm, seconds_fallback=rl.seconds_fallback)
    print("[bench-lm] learned: %.1f s, %d iterations, median chi2 %.3g, %d fallback pixels"
          % (dtl, rl.iterations, lrn["median_chi2"], rl.fallback_pixels), file=sys.stderr, flush=True)
    # median chi^2 of the network's proposal itself (one synthesis, not timed
    # as part of the pipeline): when it is already at or below plain LM's final
    # level, the pipeline reaches that level after its network stage
    prop = init.clone()
    ctx.learned_init_torch(obs, prop)
    ctx.hydrostatic_torch(prop)
    pst, _, _ = ctx.forward_torch(prop, gpu.RF_NONE)
    init_chi2 = float(((pst.double().reshape(ncol, -1) - obs) ** 2).mean(dim=1).median()) / 1e-6
    del prop, pst
    lrn["proposal_median_chi2"] = init_chi2
    # plain LM to the same median chi^2: further INVERT3D passes, each from the
    # previous pass's model (how a user restarts the reference), up to 4
    target = lrn["median_chi2"]
    passes, t_plain, cur, med = 1, plain["seconds"], model_plain, plain["median_chi2"]
    while med > target and passes < args.max_passes:
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        cur, _, c2, _ = learned.invert(ctx, cur, obs, ctl, nx, ny, None, learned=False)
        torch.cuda.synchronize()
        t_plain += time.perf_counter() - t0
        passes += 1
        med = float(c2.median())
        print("[bench-lm] plain pass %d: %.1f s total, median chi2 %.3g" % (passes, t_plain, med), file=sys.stderr,
              flush=True)
    to_target = {"passes": passes, "seconds": t_plain, "median_chi2": med, "reached": med <= target}
    speed = t_plain / dtl
    line = {"metric": "LM-converged pixels/s (learned init)", "value": ncol / dtl, "unit": "pixel/s", "n_gpus": 1,
            "steps": 1, "warmup": 0, "ms_per_step": dtl * 1e3, "higher_is_better": True, "scaling": "weak",
            "vs_baseline": None, "dtype": "f64 LM step, f32 radiative transfer, bf16 network",
            "data": "synthetic (seeded truth + 1e-3 noise; network trained on %d other seeded pairs)"
                    % (tc * args.train_nx ** 2),
            "config": {"workload": workload, "learned": lrn, "plain": plain, "plain_to_learned_chi2": to_target,
                       # the learned pipeline ends at a median chi^2 no higher than plain LM's, so it
                       # reaches plain's level no later than its own end: a conservative speed-up
                       "speedup_at_plain_median_chi2": (plain["seconds"] / dtl
                                                        if lrn["median_chi2"] <= plain["median_chi2"] else None),
                       "speedup_at_plain_median_chi2_network_stage": (
                           plain["seconds"] / rl.seconds_init if init_chi2 <= plain["median_chi2"] else None),
                       # plain LM restarted from its own result towards the learned median chi^2
                       "speedup_at_learned_median_chi2": speed if to_target["reached"] else None,
                       "speedup_at_learned_median_chi2_lower_bound": None if to_target["reached"] else speed,
                       "network": {"shape": "%d-512-512-512-20" % (4 * numw), "pairs": tc * args.train_nx ** 2,
                                   "epochs": args.epochs, "train_seconds": rep.seconds, "data_seconds": t_data,
                                   "train_loss": rep.train_loss, "valid_loss": rep.valid_loss,
                                   "fallback_threshold_chi2": threshold}}}
    print(json.dumps(line), flush=True)


def bench_uneven(args):
    """INVERT3D over pixels that each observe their own 1-4 lines of an
    8-line Fe I / Si I list, each line through a centred window of 60-300
    wavelengths, in one launch (per-pixel windows): the ragged spectra of all
    pixels go through one forward and one LM step per iteration."""
    import torch
    from firtez_dz_amd import atmos, gpu, lines

    nx = ny = 512 if args.nx == 288 else args.nx
    nz = 64 if args.nz == 96 else args.nz
    ncol = nx * ny
    ls = lines.uneven_list(nwave=300)
    print("[bench-uneven] seeding %dx%dx%d truth" % (nx, ny, nz), file=sys.stderr, flush=True)
    atm, z = atmos.seeded(nx, ny, nz, dz_km=15.0, seed=7)
    ctx = gpu.Context(0)
    numw = ctx.configure(ls, z)
    ctx.gain_hsra()
    truth = torch.from_numpy(atm).cuda()
    assert int(ctx.hydrostatic_torch(truth).max().item()) == 0
    init = plain_init(ctx, truth)
    win = lines.uneven_windows(ls, ncol, seed=17)
    off = ctx.set_windows(win)
    nw = np.diff(off)
    nlin = (win[:, :, 1] > 0).sum(axis=1)
    tot = int(off[-1])
    st, _, status = ctx.forward_torch(truth, gpu.RF_NONE)
    torch.cuda.synchronize()
    assert int(status.max().item()) == 0
    gen = torch.Generator(device="cuda")
    gen.manual_seed(11)
    obs = (st.double() + 1e-3 * torch.randn(st.shape, generator=gen, device="cuda", dtype=torch.float64)).contiguous()
    del st
    ctl = gpu.invert_control({"T": 4, "Bx": 4, "By": 4, "Bz": 4, "vz": 4}, noise=1e-3, smoothing=False)
    print("[bench-uneven] %d pixels, %d wavelengths (%.0f per pixel, %d..%d), lines per pixel %.2f"
          % (ncol, tot, nw.mean(), nw.min(), nw.max(), nlin.mean()), file=sys.stderr, flush=True)
    model = init.clone()
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    _, chi2, its = ctx.invert3d_torch(model, obs, ctl, nx, ny)
    torch.cuda.synchronize()
    dt = time.perf_counter() - t0
    stats = lm_stats(model, truth, chi2.cpu().numpy())
    print("[bench-uneven] %.1f s, %d iterations, median chi2 %.3g" % (dt, its, stats["median_chi2"]),
          file=sys.stderr, flush=True)
    workload = ("INVERT3D %dx%d pixels, nz=%d, per pixel 1-4 of 8 lines (Fe I 6301.5, 6302.5, 6173.3, 15648.5, "
                "15652.9, 15662.0, 5576.1, Si I 10827.1) in centred windows of 60-300 of 300 wavelengths, "
                "T/Bx/By/Bz/vz up to 4 nodes (2 cycles), one launch" % (nx, ny, nz))
    line = {"metric": "LM-converged pixels/s (uneven lines/windows)", "value": ncol / dt, "unit": "pixel/s",
            "n_gpus": 1, "steps": 1, "warmup": 0, "ms_per_step": dt * 1e3, "higher_is_better": True,
            "scaling": "weak", "vs_baseline": None, "dtype": "f64 LM step, f32 radiative transfer",
            "data": "synthetic (seeded truth + 1e-3 noise, seeded windows)",
            "config": dict({"workload": workload, "wavelengths_total": tot, "wavelengths_per_pixel_mean": float(nw.mean()),
                            "union_wavelengths": numw, "lines_per_pixel_mean": float(nlin.mean()),
                            "lm_iterations": its, "seconds": dt,
                            "column_wavelengths_per_s_of_iterations": tot * its / dt}, **stats)}
    print(json.dumps(line), flush=True)


def bench_coupled(args):
    """INVERT3D through a spatial PSF (fz_coupled_invert3d_dev): a seeded
    256x256x128 truth, observed as PSF * forward(truth) + 1e-3 noise, with a
    31x31 PSF (Airy core of first-zero radius 2.44 pixels, i.e. 1.22 lambda/D
    with lambda/D = 0.16" on 0.08" pixels, plus a Gaussian wing of 4 pixels
    carrying 20 % of the light).  Reported: the coupled inversion's
    pixels/s and accuracy, the uncoupled INVERT3D of the same observations
    beside it, and the PSF convolution kernel against its roofline."""
    import torch
    from firtez_dz_amd import atmos, gpu, lines

    nx = ny = 256 if args.nx == 288 else args.nx
    nz = 128 if args.nz == 96 else args.nz
    ncol = nx * ny
    ls = lines.fe6301_6302(nwave=112, start=-624.0, step=20.0)
    print("[bench-coupled] seeding %dx%dx%d truth" % (nx, ny, nz), file=sys.stderr, flush=True)
    atm, z = atmos.seeded(nx, ny, nz, dz_km=15.0, seed=7)
    ctx = gpu.Context(0)
    numw = ctx.configure(ls, z)
    ctx.gain_hsra()
    psf = atmos.airy_gauss_psf(31, 2.44, 4.0, 0.2)
    ctx.set_psf(psf)
    truth = torch.from_numpy(atm).cuda()
    assert int(ctx.hydrostatic_torch(truth).max().item()) == 0
    st, _, status = ctx.forward_torch(truth, gpu.RF_NONE)
    torch.cuda.synchronize()
    assert int(status.max().item()) == 0
    nmap = 4 * numw
    maps = st.reshape(ncol, nmap).t().contiguous()
    del st
    conv = ctx.psf_convolve_torch(maps, nx, ny)
    # the convolution kernel alone, timed on the stream it runs on
    reps = 10
    ev0, ev1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    ev0.record()
    for _ in range(reps):
        ctx.psf_convolve_torch(maps, nx, ny, out=conv)
    ev1.record()
    torch.cuda.synchronize()
    conv_ms = ev0.elapsed_time(ev1) / reps
    gen = torch.Generator(device="cuda")
    gen.manual_seed(11)
    obs = (conv.t().double() +
           1e-3 * torch.randn((ncol, nmap), generator=gen, device="cuda", dtype=torch.float64)).contiguous()
    del maps, conv
    init = plain_init(ctx, truth)
    ctl = gpu.invert_control({"T": 4, "Bx": 4, "By": 4, "Bz": 4, "vz": 4}, noise=1e-3, smoothing=False)
    print("[bench-coupled] PSF convolution of %d maps: %.3f ms; coupled inversion of %d pixels"
          % (nmap, conv_ms, ncol), file=sys.stderr, flush=True)
    model = init.clone()
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    _, chi_c, its_c = ctx.coupled_invert3d_torch(model, obs, ctl, nx, ny)
    torch.cuda.synchronize()
    dt = time.perf_counter() - t0
    coupled = dict(lm_stats(model, truth, chi_c.cpu().numpy()), lm_iterations=its_c, seconds=dt)
    print("[bench-coupled] coupled: %.1f s, %d iterations, median chi2 %.3g, mean |T - truth| %.1f K"
          % (dt, its_c, coupled["median_chi2"], coupled["mean_abs_T_error_K"]), file=sys.stderr, flush=True)
    del model
    model = init.clone()
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    _, chi_u, its_u = ctx.invert3d_torch(model, obs, ctl, nx, ny)
    torch.cuda.synchronize()
    du = time.perf_counter() - t0
    uncoupled = dict(lm_stats(model, truth, chi_u.cpu().numpy()), lm_iterations=its_u, seconds=du)
    print("[bench-coupled] uncoupled: %.1f s, %d iterations, median chi2 %.3g, mean |T - truth| %.1f K"
          % (du, its_u, uncoupled["median_chi2"], uncoupled["mean_abs_T_error_K"]), file=sys.stderr, flush=True)
    # the learned initial model (the coupled workload's definition): a network
    # trained on PSF-degraded pairs of other seeded cubes proposes the model
    # on the final node basis, which the coupled inversion then refines (its
    # last cycle only, as for the uncoupled learned pipeline)
    from firtez_dz_amd import learned

    def blur(rows):
        return ctx.psf_convolve_torch(rows.t().contiguous(), nx, ny).t()

    ctx.mlp_create(nmap, 512, 3, 4, seed=1)
    X, Tg = learned.training_set(ctx, args.train_cubes_coupled, nx, ny, nz, seed0=2000, noise=1e-3, transform=blur)
    rep = learned.train(ctx, X, Tg, epochs=args.epochs, batch=1024, lr=1e-3)
    del X, Tg
    print("[bench-coupled] network trained on %d PSF-degraded pairs in %.1f s"
          % (args.train_cubes_coupled * ncol, rep.seconds), file=sys.stderr, flush=True)
    ctl_l = gpu.CInvertControl.from_buffer_copy(ctl)
    ctl_l.first_cycle = 1 << 20
    del model
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    model = init.clone()
    ctx.learned_init_torch(obs, model)
    bad = ctx.hydrostatic_torch(model) != 0
    if bool(bad.any()):
        model[:, :, bad] = init[:, :, bad]
        ctx.hydrostatic_torch(model)
    _, chi_l, its_l = ctx.coupled_invert3d_torch(model, obs, ctl_l, nx, ny)
    torch.cuda.synchronize()
    dl = time.perf_counter() - t0
    lrn = dict(lm_stats(model, truth, chi_l.cpu().numpy()), lm_iterations=its_l, seconds=dl,
               pixels_per_s=ncol / dl, training_seconds=rep.seconds)
    print("[bench-coupled] learned + coupled: %.1f s, %d iterations, median chi2 %.3g, mean |T - truth| %.1f K"
          % (dl, its_l, lrn["median_chi2"], lrn["mean_abs_T_error_K"]), file=sys.stderr, flush=True)
    taps = 31 * 31
    conv_flop = 2.0 * taps * nmap * ncol
    conv_bytes = 2.0 * 4 * nmap * ncol
    workload = ("coupled INVERT3D %dx%d pixels, nz=%d, Fe I 6301.5+6302.5, %d wavelengths, 31x31 Airy+wing PSF, "
                "T/Bx/By/Bz/vz up to 4 nodes (2 cycles), plain initial model" % (nx, ny, nz, numw))
    line = {"metric": "LM-converged pixels/s (PSF-coupled)", "value": ncol / dt, "unit": "pixel/s", "n_gpus": 1,
            "steps": 1, "warmup": 0, "ms_per_step": dt * 1e3, "higher_is_better": True, "scaling": "weak",
            "vs_baseline": None, "dtype": "f64 LM step, f32 radiative transfer and PSF convolution",
            "data": "synthetic (seeded truth, PSF-degraded, + 1e-3 noise)",
            "config": {"workload": workload, "coupled": coupled, "uncoupled_same_observations": uncoupled,
                       "learned_then_coupled": lrn, "ms_per_lm_iteration": dt * 1e3 / max(its_c, 1)},
            "roofline": {"kernel": "psf_conv_kernel (%d maps of %dx%d, %d taps)" % (nmap, nx, ny, taps),
                         "bound": "fp32 valu", "achieved": conv_flop / (conv_ms * 1e-3) / 1e12, "peak": 157.3,
                         "unit": "TFLOP/s", "frac": conv_flop / (conv_ms * 1e-3) / 1e12 / 157.3,
                         "ms": conv_ms, "hbm_GB_per_s": conv_bytes / (conv_ms * 1e-3) / 1e9}}
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
