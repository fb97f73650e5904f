"""fz_set_overlap: the microphysics of the next radiative-transfer tile on the
library's second stream while the current tile's transfer runs must give
the forward of the one-stream order bit for bit (Stokes, per-height
Jacobians, the updated cube and the status), here over five tiles of a
320-column cube with one failing column."""
import numpy as np
import pytest

from firtez_dz_amd import atmos, gpu, lines

pytestmark = pytest.mark.gpu


def test_overlap_matches_sequential(gpu_ctx):
    import torch
    ls = lines.fe6301_6302(nwave=40, start=-400.0, step=20.0)
    atm, z = atmos.seeded(16, 20, 48, seed=31)
    atm[0, :, 77] = -5.0                      # T <= 0: this column's EOS fails
    gpu_ctx.configure(ls, z)
    gpu_ctx.gain_hsra()
    t = torch.from_numpy(atm.copy()).cuda()
    ok = torch.ones(t.shape[2], dtype=torch.bool, device="cuda")
    ok[77] = False
    sub = t[:, :, ok].contiguous()
    assert int(gpu_ctx.hydrostatic_torch(sub).max()) == 0
    t[:, :, ok] = sub
    outs = []
    try:
        gpu_ctx.set_tile(64, 256)
        for on in (False, True):
            gpu_ctx.set_overlap(on)
            a = t.clone()
            st, jac, status = gpu_ctx.forward_torch(a, gpu.RF_SPHERICAL)
            torch.cuda.synchronize()
            outs.append((st.cpu().numpy(), jac.cpu().numpy(), status.cpu().numpy(), a.cpu().numpy()))
    finally:
        gpu_ctx.set_overlap(False)
        gpu_ctx.set_tile(32768, 256)
    (s0, j0, e0, a0), (s1, j1, e1, a1) = outs
    assert e0[77] != 0 and np.array_equal(e0, e1)
    good = e0 == 0
    assert np.array_equal(s0[good], s1[good]) and np.array_equal(j0[good], j1[good])
    assert np.array_equal(a0[:, :, good], a1[:, :, good])
    print("[overlap] 5 tiles of 64 columns: second-stream microphysics identical to one stream", flush=True)
