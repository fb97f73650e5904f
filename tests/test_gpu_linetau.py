"""LINE OPTICAL DEPTH on the device (fz_line_tau_dev) against the oracle's
(pinned by the continuum trapezoid of an absent line): log10 tau of every
wavelength at every height of 8 seeded columns (Fe I 6301.5 + 6302.5), on
the heights the opacities reach (the deepest, opaque heights contribute 0
on both sides)."""
import numpy as np
import pytest

from firtez_dz_amd import atmos, gpu, lines

pytestmark = pytest.mark.gpu


def test_line_tau_vs_oracle(gpu_ctx, cpu):
    import torch
    ls = lines.fe6301_6302(nwave=30, start=-300.0, step=20.0)
    atm, z = atmos.seeded(4, 2, 48, dz_km=15.0, seed=17)
    ncol = 8
    cpu.setup(ls, resp=False, atmpar=(1, 1, 1, 1, 1, 1, 1, 0))
    for i in range(ncol):
        pg, rho, pel, mw, _ = cpu.he(atm[0, :, i], float(atm[1, -1, i]), z)
        atm[1, :, i], atm[2, :, i], atm[3, :, i], atm[4, :, i] = pg, rho, pel, mw
    gpu_ctx.configure(ls, z)
    cpu.set_gain(gpu_ctx.gain_hsra())
    t = torch.from_numpy(atm.copy()).cuda()
    tl, status = gpu_ctx.line_tau_torch(t)
    torch.cuda.synchronize()
    tl, out = tl.cpu().numpy(), t.cpu().numpy()
    assert np.all(status.cpu().numpy() == 0)
    worst = 0.0
    for i in range(ncol):
        _, _, _, _, ref = cpu.forward1d_line_tau(atmos.column_model(out, z, i))
        worst = max(worst, float(np.max(np.abs(tl[i] - ref))))
    print("[linetau] 8 columns nz=48: max |d log10 tau| %.2e" % worst, flush=True)
    # FP32 propagation matrix against the oracle's FP64 ETAI: ~2e-5 relative in tau
    assert worst <= 5e-5
