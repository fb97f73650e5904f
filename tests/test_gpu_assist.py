"""SET_TAU's assist initialisers on the GPU (fz_assist_dev and the first
cycle of fz_invert3d_dev) against the oracle's restatement
(oracle/invert.py assist, itself pinned by closed forms in
tests/test_oracle_assist.py)."""
import numpy as np
import pytest

from firtez_dz_amd import atmos, gpu, lines
from oracle import invert as oinv

pytestmark = pytest.mark.gpu


def report(name, err, tol):
    print("[assist] %-48s max err %.3e (tol %.1e)" % (name, err, tol), flush=True)
    assert err <= tol, name


def test_assists_vs_oracle(gpu_ctx):
    import torch
    ls = lines.fe6301_6302()
    atm, z = atmos.seeded(6, 5, 48, seed=8)
    gpu_ctx.configure(ls, z)
    gpu_ctx.gain_hsra()
    t = torch.from_numpy(atm.copy()).cuda()
    assert int(gpu_ctx.hydrostatic_torch(t).max()) == 0
    st, _, _ = gpu_ctx.forward_torch(t, gpu.RF_NONE)
    obs = st.double().reshape(30, -1).contiguous()
    gpu_ctx.set_inversion(gpu.inversion({"T": 2, "Bx": 1, "By": 1, "Bz": 1, "vz": 1}))
    start = t.clone()
    numw = gpu_ctx.numw
    for flags in (oinv.ASSIST_B | oinv.ASSIST_V, oinv.ASSIST_T, oinv.ASSIST_P,
                  oinv.ASSIST_B | oinv.ASSIST_V | oinv.ASSIST_T | oinv.ASSIST_P):
        d = start.clone()
        gpu_ctx.assist_torch(obs, d, flags)
        torch.cuda.synchronize()
        ref = oinv.assist(obs.cpu().numpy(), start.cpu().numpy(), flags, numw, numw, 4, ls.lines[0].lambda0)
        got = d.cpu().numpy()
        for row in (0, 1, 5, 6, 7, 10, 11):
            err = np.max(np.abs(got[row] - ref[row]) / (np.abs(ref[row]) + 1e-3))
            report("flags %d row %s" % (flags, gpu.ATM[row]), err, 1e-5)


def test_invert3d_with_assists_runs(gpu_ctx):
    """INVERT3D starting from the assisted model: the first cycle applies the
    assists, the inversion completes with every pixel's status clean."""
    import torch
    ls = lines.fe6301_6302()
    atm, z = atmos.seeded(4, 4, 48, seed=10)
    gpu_ctx.configure(ls, z)
    gpu_ctx.gain_hsra()
    t = torch.from_numpy(atm.copy()).cuda()
    assert int(gpu_ctx.hydrostatic_torch(t).max()) == 0
    st, _, _ = gpu_ctx.forward_torch(t, gpu.RF_NONE)
    obs = st.double().reshape(16, -1).contiguous()
    init = t.clone()
    init[5:8] = 0.0
    init[10] = 0.0
    ctl = gpu.invert_control({"T": 2, "Bx": 1, "By": 1, "Bz": 1, "vz": 1}, noise=1e-3, smoothing=False,
                             max_iterations=4, assist=oinv.ASSIST_B | oinv.ASSIST_V)
    model = init.clone()
    _, chi2, its = gpu_ctx.invert3d_torch(model, obs, ctl, 4, 4)
    torch.cuda.synchronize()
    c = chi2.cpu().numpy()
    assert its > 0 and np.all(np.isfinite(c))
    assert float(model[5:8].abs().max()) > 0.0            # the field rows moved off zero
