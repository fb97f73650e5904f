"""STOKES SETUP AUTO on the device: fz_auto_weights_dev against the oracle's
SET_WEIGHTS3D restatement, and INVERT3D with the per-pixel weights
(fz_set_auto_weights) against the oracle's inversion with the same weights."""
import numpy as np
import pytest

from firtez_dz_amd import atmos, gpu, lines

pytestmark = pytest.mark.gpu

CART = (1, 1, 1, 1, 1, 1, 1, 0)
NOISE = 1e-3


def test_auto_weights_vs_oracle(gpu_ctx):
    import torch
    from oracle import invert
    ls = lines.fe6301_6302(nwave=24, start=-200.0, step=30.0)
    gpu_ctx.configure(ls, np.linspace(0.0, 6e7, 16).astype(np.float32))
    numw = gpu_ctx.numw
    rng = np.random.default_rng(11)
    for invstk in ((1, 1, 1, 1), (1, 0, 0, 1), (0, 1, 1, 0)):
        nstk = sum(invstk)
        ncol = 300
        scale = rng.uniform(1e-4, 0.3, (ncol, nstk, 1))
        obs = (rng.normal(0.0, 1.0, (ncol, nstk, numw)) * scale).reshape(ncol, nstk * numw)
        obs[7, :numw] = 0.5                             # a constant block: the cap
        print("[autow] invstk %s ncol %d numw %d" % (invstk, ncol, numw), flush=True)
        got = gpu_ctx.auto_weights_torch(torch.from_numpy(obs).cuda(), invstk, NOISE).cpu().numpy()
        ref = invert.auto_weights(obs, invstk, NOISE, numw)
        err = float(np.max(np.abs(got - ref) / ref))
        print("[autow] max relative difference %.3g" % err, flush=True)
        assert err <= 3e-7                               # one REAL(4) rounding of WSTK


def test_invert3d_auto_weights_vs_oracle(gpu_ctx, cpu):
    import torch
    from oracle import invert
    ls = lines.fe6301_6302(nwave=48, start=-200.0, step=30.0)
    nx, ny = 2, 2
    atm, z = atmos.seeded(nx, ny, 64, seed=21)
    cpu.setup(ls, resp=True, atmpar=CART)
    ncol = nx * ny
    for c in range(ncol):
        pg, rho, pel, mw, _ = cpu.he(atm[0, :, c], float(atm[1, -1, c]), z)
        atm[1, :, c], atm[2, :, c], atm[3, :, c], atm[4, :, c] = pg, rho, pel, mw
    rng = np.random.default_rng(21)
    gpu_ctx.configure(ls, z)
    cpu.set_gain(gpu_ctx.gain_hsra())
    obs = []
    for c in range(ncol):
        mo, syn, _, _ = cpu.forward1d(atmos.column_model(atm, z, c))
        atm[:, :, c] = mo[:12]
        obs.append(syn.reshape(-1) + rng.normal(0, NOISE, syn.size))
    obs = np.array(obs)
    guess = atm.copy()
    guess[0] += np.float32(150.0)
    guess[5:8] *= np.float32(0.8)
    guess[10] += np.float32(5e4)
    nodes = {"T": 1, "Bx": 1, "By": 1, "Bz": 1, "vz": 1}
    ctl = gpu.invert_control(nodes, noise=NOISE, smoothing=False)
    gpu_ctx.set_auto_weights(True)
    try:
        d_atm = torch.from_numpy(guess.copy()).cuda()
        print("[autow] invert3d 2x2 one node, AUTO weights", flush=True)
        st, chi2, its = gpu_ctx.invert3d_torch(d_atm, torch.from_numpy(obs.copy()).cuda(), ctl, nx, ny)
        torch.cuda.synchronize()
    finally:
        gpu_ctx.set_auto_weights(False)
    g_model, g_chi2 = d_atm.cpu().numpy(), chi2.cpu().numpy()
    idx = {gpu.INV_PARAMS.index(k): v for k, v in nodes.items()}
    o_model, o_syn, o_chi2, o_its = invert.invert3d(cpu, guess, z, obs, nx, ny, idx, noise=NOISE, smoothing=False,
                                                   auto=True)
    print("[autow] iterations device %d oracle %d; median chi2 device %.4g oracle %.4g"
          % (its, o_its, np.median(g_chi2), np.median(o_chi2)), flush=True)
    assert abs(its - o_its) <= 2
    assert abs(np.median(g_chi2) - np.median(o_chi2)) / np.median(o_chi2) <= 0.1
    np.testing.assert_allclose(g_chi2, o_chi2, rtol=0.1)
    for row in (0, 5, 6, 7, 10):
        scale = max(1.0, float(np.abs(atm[row]).max()))
        err = float(np.max(np.abs(g_model[row] - o_model[row]))) / scale
        print("[autow] model row %d vs oracle (relative to max) %.3g" % (row, err), flush=True)
        assert err <= 2e-3
    # the AUTO weights differ from the map-wide ones: chi2 of the same model changes
    o_model2, _, o_chi2_map, _ = invert.invert3d(cpu, guess, z, obs, nx, ny, idx, noise=NOISE, smoothing=False)
    assert not np.allclose(o_chi2_map, o_chi2, rtol=1e-3)


def test_auto_weights_refused_by_coupled_inversion(gpu_ctx):
    import torch
    ls = lines.fe6301_6302(nwave=24, start=-200.0, step=30.0)
    atm, z = atmos.seeded(4, 4, 32, seed=2)
    gpu_ctx.configure(ls, z)
    gpu_ctx.set_psf(np.full((3, 3), 1.0 / 9.0))
    gpu_ctx.set_auto_weights(True)
    try:
        ctl = gpu.invert_control({"T": 1}, noise=NOISE, smoothing=False)
        obs = torch.zeros((16, 4 * gpu_ctx.numw), dtype=torch.float64, device="cuda")
        with pytest.raises(gpu.FzError):
            gpu_ctx.coupled_invert3d_torch(torch.from_numpy(atm).cuda(), obs, ctl, 4, 4)
    finally:
        gpu_ctx.set_auto_weights(False)
        gpu_ctx.set_psf(None)
