"""Near-degenerate polarisation eigenvalues in the evolution operator.

Column 2849 of the benchmark cube (288x288x96, dz 15 km, seed 2, Fe I
6301.5 + 6302.5 over 112 wavelengths) has a layer (height 33, near the
6302.5 core) where th = sqrt(d^2 + 4 (e.r)^2) is small against
(e.e + r.r)/2: the closed form of the operator's directional derivatives
(rtesolver.f90:476-1002) cancels terms that grow like 1/th^1.5 there, and in
FP32 the v_los response of that point came out 6 % of the column's largest
v_los response away from the FP64 oracle.  The kernel evaluates the
operator through the series of its coefficient functions below th = 2; the
column's spherical Jacobians must now meet the same tolerance as every
other column.
"""
import numpy as np
import pytest

from firtez_dz_amd import atmos, gpu, lines

pytestmark = pytest.mark.gpu
RF_TOL = 2e-3
STOKES_TOL = 1e-4
CART = (1, 1, 1, 1, 1, 1, 1, 0)


def test_degenerate_layer_column(gpu_ctx, cpu):
    import torch
    ls = lines.fe6301_6302(nwave=112, start=-624.0, step=20.0)
    atm, z = atmos.seeded(288, 288, 96, dz_km=15.0, seed=2)
    cols = [2849, 100]
    cpu.setup(ls, resp=True, atmpar=CART)
    sub = np.ascontiguousarray(atm[:, :, cols])
    for i in range(len(cols)):
        pg, rho, pel, mw, _ = cpu.he(sub[0, :, i], float(sub[1, -1, i]), z)
        sub[1, :, i], sub[2, :, i], sub[3, :, i], sub[4, :, i] = pg, rho, pel, mw
    gpu_ctx.configure(ls, z)
    gain = gpu_ctx.gain_hsra()
    cpu.set_gain(gain)
    t = torch.from_numpy(sub.copy()).cuda()
    st, jac, status = gpu_ctx.forward_torch(t, gpu.RF_SPHERICAL)
    torch.cuda.synchronize()
    st, jac, status, out = st.cpu().numpy(), jac.cpu().numpy(), status.cpu().numpy(), t.cpu().numpy()
    assert np.all(status == 0)
    names = ("T", "|B|", "inclination", "azimuth", "v_los")
    for i, c in enumerate(cols):
        m = atmos.column_model(out, z, i)
        _, syn, dsyn, _ = cpu.forward1d(m)
        assert np.max(np.abs(st[i] - syn)) <= STOKES_TOL
        tr = lambda p: np.transpose(dsyn[:, :, p, :], (0, 2, 1))
        bx, by, bz = (float(m[j, 0]) for j in (5, 6, 7))
        B = np.sqrt(bx * bx + by * by + bz * bz)
        g, f = np.arccos(bz / B), np.arctan2(by, bx)
        dx, dy, dzz = tr(3), tr(4), tr(5)
        ref = [tr(0), (bx * dx + by * dy + bz * dzz) / B,
               B * np.cos(g) * np.cos(f) * dx + B * np.cos(g) * np.sin(f) * dy - B * np.sin(g) * dzz,
               -B * np.sin(g) * np.sin(f) * dx + B * np.sin(g) * np.cos(f) * dy, tr(6)]
        for p in range(5):
            sc = np.abs(ref[p]).max()
            err = float(np.max(np.abs(jac[i, p] - ref[p])) / sc)
            print("[degenerate] column %d %-11s max err %.3e of the column's largest response (tol %.0e)"
                  % (c, names[p], err, RF_TOL), flush=True)
            assert err <= RF_TOL, (c, names[p])
