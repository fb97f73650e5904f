"""Microturbulence on the device (fz_set_microturbulence) against the
oracle with the same xi: the benchmark's line set and grid (Fe I 6301.5 +
6302.5, 112 wavelengths, nz = 96, dz 15 km) with xi = 0.5 km/s, Stokes
vectors and spherical per-height Jacobians of 16 columns; and xi back at 0
reproduces the thermal-width forward bit for bit."""
import numpy as np
import pytest

from firtez_dz_amd import atmos, gpu, lines

pytestmark = pytest.mark.gpu
XI = 5.0e4
RF_TOL = 2e-3
STOKES_TOL = 1e-4
CART = (1, 1, 1, 1, 1, 1, 1, 0)


def test_microturbulence_vs_oracle(gpu_ctx, cpu):
    import torch
    ls = lines.fe6301_6302(nwave=112, start=-624.0, step=20.0)
    atm, z = atmos.seeded(4, 4, 96, dz_km=15.0, seed=12)
    ncol = 16
    cpu.setup(ls, resp=True, atmpar=CART)
    for i in range(ncol):
        pg, rho, pel, mw, _ = cpu.he(atm[0, :, i], float(atm[1, -1, i]), z)
        atm[1, :, i], atm[2, :, i], atm[3, :, i], atm[4, :, i] = pg, rho, pel, mw
    gpu_ctx.configure(ls, z)
    cpu.set_gain(gpu_ctx.gain_hsra())
    t0 = torch.from_numpy(atm.copy()).cuda()
    st0, jac0, _ = gpu_ctx.forward_torch(t0, gpu.RF_SPHERICAL)
    try:
        gpu_ctx.set_microturbulence(XI)
        cpu.set_microturbulence(XI)
        t = torch.from_numpy(atm.copy()).cuda()
        st, jac, status = gpu_ctx.forward_torch(t, gpu.RF_SPHERICAL)
        torch.cuda.synchronize()
        st, jac, out = st.cpu().numpy(), jac.cpu().numpy(), t.cpu().numpy()
        assert np.all(status.cpu().numpy() == 0)
        worst_s, worst_j = 0.0, [0.0] * 5
        for i in range(ncol):
            m = atmos.column_model(out, z, i)
            _, syn, dsyn, _ = cpu.forward1d(m)
            worst_s = max(worst_s, float(np.max(np.abs(st[i] - syn))))
            tr = lambda p: np.transpose(dsyn[:, :, p, :], (0, 2, 1))
            bx, by, bz = (float(m[j, 0]) for j in (5, 6, 7))
            B = np.sqrt(bx * bx + by * by + bz * bz)
            g, f = np.arccos(bz / B), np.arctan2(by, bx)
            dx, dy, dzz = tr(3), tr(4), tr(5)
            ref = [tr(0), (bx * dx + by * dy + bz * dzz) / B,
                   B * np.cos(g) * np.cos(f) * dx + B * np.cos(g) * np.sin(f) * dy - B * np.sin(g) * dzz,
                   -B * np.sin(g) * np.sin(f) * dx + B * np.sin(g) * np.cos(f) * dy, tr(6)]
            for p in range(5):
                worst_j[p] = max(worst_j[p], float(np.max(np.abs(jac[i, p] - ref[p])) / np.abs(ref[p]).max()))
        print("[micro] xi %.1f km/s, 16 columns nz=96: Stokes max err %.2e; Jacobians (T, |B|, incl, azim, v) "
              "max err %s of each column's largest" % (XI / 1e5, worst_s, " ".join("%.1e" % w for w in worst_j)),
              flush=True)
        assert worst_s <= STOKES_TOL
        assert max(worst_j) <= RF_TOL
        assert float(np.max(np.abs(st - st0.cpu().numpy()))) > 1e-3          # the width matters
    finally:
        gpu_ctx.set_microturbulence(0.0)
        cpu.set_microturbulence(0.0)
    t1 = torch.from_numpy(atm.copy()).cuda()
    st1, jac1, _ = gpu_ctx.forward_torch(t1, gpu.RF_SPHERICAL)
    torch.cuda.synchronize()
    assert torch.equal(st1, st0) and torch.equal(jac1, jac0)
