"""FULL_STOKES on the device (fz_forward_full_dev): the Stokes vector after
every height against the oracle's (itself pinned by truncated columns), for
the Stokes-only and the response-function kernels; the top height is the
emergent Stokes output bit for bit."""
import numpy as np
import pytest

from firtez_dz_amd import atmos, gpu, lines

pytestmark = pytest.mark.gpu
TOL = 1e-4


@pytest.mark.parametrize("rf", [gpu.RF_NONE, gpu.RF_SPHERICAL], ids=["stokes", "jacobian"])
def test_full_stokes_vs_oracle(gpu_ctx, cpu, rf):
    import torch
    ls = lines.fe6301_6302(nwave=30, start=-300.0, step=20.0)
    atm, z = atmos.seeded(4, 2, 48, dz_km=15.0, seed=13)
    ncol = 8
    cpu.setup(ls, resp=False, atmpar=(1, 1, 1, 1, 1, 1, 1, 0))
    for i in range(ncol):
        pg, rho, pel, mw, _ = cpu.he(atm[0, :, i], float(atm[1, -1, i]), z)
        atm[1, :, i], atm[2, :, i], atm[3, :, i], atm[4, :, i] = pg, rho, pel, mw
    gpu_ctx.configure(ls, z)
    cpu.set_gain(gpu_ctx.gain_hsra())
    t = torch.from_numpy(atm.copy()).cuda()
    st, _, status, full = gpu_ctx.forward_full_torch(t, rf)
    torch.cuda.synchronize()
    st, full, out = st.cpu().numpy(), full.cpu().numpy(), t.cpu().numpy()
    assert np.all(status.cpu().numpy() == 0)
    assert np.array_equal(full[:, -1], st)
    worst = 0.0
    for i in range(ncol):
        _, _, _, _, fs = cpu.forward1d_full(atmos.column_model(out, z, i))
        worst = max(worst, float(np.max(np.abs(full[i] - fs))))
    print("[fullstokes] rf %d: max |dS| over all heights %.2e Ic" % (rf, worst), flush=True)
    assert worst <= TOL
