"""The forward + Jacobian workload at its own shape: nz = 96 heights at
dz = 15 km with the top pressure of that grid (the optically thinnest top
layers of any test), the Fe I 6301.5/6302.5 pair in one region of 112
wavelengths at 20 mA, Jacobians w.r.t. (T, |B|, inclination, azimuth, v_los)
at every height.

  * a seeded 16x16x96 cube: Stokes of all 256 columns against the FP64
    oracle; spherical Jacobians of 16 columns against the oracle's Cartesian
    response functions through the closed-form chain rule (max and p99);
  * the full 288x288x96 cube: every column succeeds, Stokes and Jacobians
    are finite, a slice re-run as its own cube is bit-identical to the same
    columns of the whole-cube run (the launch boundary at 32768 columns
    falls inside the slice), and 64 sampled columns match the oracle.

Tolerances as in test_gpu_parity: Stokes 1e-4 I_c; Jacobians 2e-3 of the
largest response of the spectrum to the same parameter.
"""
import numpy as np
import pytest

from firtez_dz_amd import atmos, gpu, lines

pytestmark = pytest.mark.gpu
STOKES_TOL = 1e-4
RF_TOL = 2e-3
CART = (1, 1, 1, 1, 1, 1, 1, 0)
NZ = 96
DZ_KM = 15.0


def report(name, err, tol, p99=None):
    extra = "" if p99 is None else "  p99 %.3e" % p99
    print("[config2] %-50s max %.3e%s (tol %.1e)" % (name, err, extra, tol), flush=True)
    assert err <= tol, name


def cart_to_sph(jc, bx, by, bz):
    """d/d(Bx,By,Bz) -> d/d(|B|, gamma, phi): the chain rule with
    Bx = B sin(g) cos(f), By = B sin(g) sin(f), Bz = B cos(g)."""
    B = np.sqrt(bx * bx + by * by + bz * bz)
    g = np.arccos(bz / B)
    f = np.arctan2(by, bx)
    dx, dy, dz = jc
    dB = (bx * dx + by * dy + bz * dz) / B
    dg = B * np.cos(g) * np.cos(f) * dx + B * np.cos(g) * np.sin(f) * dy - B * np.sin(g) * dz
    df = -B * np.sin(g) * np.sin(f) * dx + B * np.sin(g) * np.cos(f) * dy
    return dB, dg, df


def oracle_spherical(dsyn, model):
    """The oracle's DSYN1D (nz, numw, 8, 4) -> [5][nz][4][numw] in the device's
    spherical order (T, |B|, gamma, phi, v_los)."""
    tr = lambda p: np.transpose(dsyn[:, :, p, :], (0, 2, 1))
    bx, by, bz = (float(model[i, 0]) for i in (5, 6, 7))
    dB, dg, df = cart_to_sph([tr(3), tr(4), tr(5)], bx, by, bz)
    return np.stack([tr(0), dB, dg, df, tr(6)])


def jacobian_errors(ours, ref):
    """Per parameter: |dJ| / max|J| over heights, Stokes and wavelengths."""
    out = []
    for p in range(ref.shape[0]):
        sc = np.abs(ref[p]).max()
        if sc > 0:
            out.append(np.abs(ours[p] - ref[p]).reshape(-1) / sc)
    return np.concatenate(out)


def test_config2_cube_vs_oracle(gpu_ctx, cpu):
    import torch
    ls = lines.fe6301_6302(nwave=112, start=-624.0, step=20.0)
    atm, z = atmos.seeded(16, 16, NZ, dz_km=DZ_KM, seed=31)
    assert atm[1, -1, 0] == np.float32(atmos.pg_top_for(NZ))
    gpu_ctx.configure(ls, z)
    gain = gpu_ctx.gain_hsra()
    t = torch.from_numpy(atm.copy()).cuda()
    assert np.all(gpu_ctx.hydrostatic_torch(t).cpu().numpy() == 0)
    atm_in = t.cpu().numpy()
    st, jac, status = gpu_ctx.forward_torch(t, gpu.RF_SPHERICAL)
    torch.cuda.synchronize()
    st, jac, status = st.cpu().numpy(), jac.cpu().numpy(), status.cpu().numpy()
    assert np.all(status == 0)
    models = np.stack([atmos.column_model(atm_in, z, c) for c in range(256)])
    cpu.setup(ls, resp=False, atmpar=CART)
    cpu.set_gain(gain)
    _, syn, _ = cpu.forward_batch(models)
    ds = np.abs(st - np.transpose(syn, (0, 2, 1))).reshape(-1)
    report("16x16x96 Stokes vs oracle (256 columns)", ds.max(), STOKES_TOL, np.percentile(ds, 99))
    cpu.setup(ls, resp=True, atmpar=CART)
    cpu.set_gain(gain)
    cols = np.random.default_rng(31).choice(256, 16, replace=False)
    _, syn_s, dsyn = cpu.forward_batch(models[cols])
    errs = []
    for i, c in enumerate(cols):
        ref = oracle_spherical(dsyn[i], models[c])
        e = jacobian_errors(jac[c], ref)
        assert e.max() <= RF_TOL, "column %d: %.3e" % (c, e.max())
        errs.append(e)
    e = np.concatenate(errs)
    report("16x16x96 spherical Jacobians vs oracle (16 columns)", e.max(), RF_TOL, np.percentile(e, 99))


def test_config2_full_cube(gpu_ctx, cpu):
    """288x288x96: status, finiteness, slice-vs-whole identity and a sampled
    oracle check of the benchmark cube itself."""
    import torch
    nx = ny = 288
    ls = lines.fe6301_6302(nwave=112, start=-624.0, step=20.0)
    atm, z = atmos.seeded(nx, ny, NZ, dz_km=DZ_KM, seed=2)
    ncol = nx * ny
    gpu_ctx.configure(ls, z)
    gain = gpu_ctx.gain_hsra()
    numw = gpu_ctx.numw
    d_atm = torch.from_numpy(atm).cuda()
    assert int(gpu_ctx.hydrostatic_torch(d_atm).max()) == 0
    atm_in = d_atm.cpu().numpy()
    stokes = torch.empty((ncol, 4, numw), dtype=torch.float32, device="cuda")
    jac = torch.empty((ncol, 5, NZ, 4, numw), dtype=torch.float32, device="cuda")
    status = torch.zeros(ncol, dtype=torch.int32, device="cuda")
    gpu_ctx.set_tile(32768, 256)
    try:
        gpu_ctx.forward_torch(d_atm, gpu.RF_SPHERICAL, stokes, jac, status)
        torch.cuda.synchronize()
    finally:
        gpu_ctx.set_tile(8192, 256)
    assert int(status.max()) == 0 and int(status.min()) == 0
    assert bool(torch.isfinite(stokes).all())
    for c0 in range(0, ncol, 8192):
        assert bool(torch.isfinite(jac[c0:c0 + 8192]).all()), c0
    print("[config2] 288x288x96: 82944 columns, status 0, Stokes and Jacobians finite", flush=True)
    # columns 32000..33535 as a cube of their own (the whole-cube run splits
    # its radiative transfer at column 32768)
    s0, s1 = 32000, 33536
    sub = torch.from_numpy(np.ascontiguousarray(atm_in[:, :, s0:s1])).cuda()
    st2, jac2, status2 = gpu_ctx.forward_torch(sub, gpu.RF_SPHERICAL)
    torch.cuda.synchronize()
    assert int(status2.max()) == 0
    assert torch.equal(st2, stokes[s0:s1]) and torch.equal(jac2, jac[s0:s1])
    assert torch.equal(sub, d_atm[:, :, s0:s1])
    print("[config2] slice of 1536 columns across the launch boundary: bit-identical", flush=True)
    cols = np.sort(np.random.default_rng(2).choice(ncol, 64, replace=False))
    h_st = stokes[cols].cpu().numpy()
    h_jac = jac[torch.from_numpy(cols).cuda()].cpu().numpy()
    models = np.stack([atmos.column_model(atm_in, z, int(c)) for c in cols])
    cpu.setup(ls, resp=True, atmpar=CART)
    cpu.set_gain(gain)
    _, syn, dsyn = cpu.forward_batch(models)
    ds = np.abs(h_st - np.transpose(syn, (0, 2, 1))).reshape(-1)
    e = np.concatenate([jacobian_errors(h_jac[i], oracle_spherical(dsyn[i], models[i])) for i in range(len(cols))])
    report("288x288x96 Stokes vs oracle (64 sampled columns)", ds.max(), STOKES_TOL, np.percentile(ds, 99))
    report("288x288x96 spherical Jacobians vs oracle (64 columns)", e.max(), RF_TOL, np.percentile(e, 99))
    del jac, jac2
    torch.cuda.empty_cache()
