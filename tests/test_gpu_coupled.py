"""The PSF-coupled inversion on the device (fz_set_psf,
fz_coupled_invert3d_dev) against the oracle's restatement of the same
design (oracle/coupled.py: the map's chi^2 through the PSF of oracle/psf.py,
acceptance on the map, each pixel stepped by the oracle's GET_PERT_1D with
the adjoint-convolved residual), and what the coupling buys: on observations
degraded by the PSF, the coupled inversion fits them to the noise while the
uncoupled INVERT3D of the same observations cannot.
The reference wires none of this (GET_DMODEL3DC is a stub,
get_dmodel.f90:1707-1711), so the oracle and the properties pin it."""
import numpy as np
import pytest

from firtez_dz_amd import atmos, gpu, lines
from oracle import coupled as OC
from oracle import psf as P

pytestmark = pytest.mark.gpu
CART = (1, 1, 1, 1, 1, 1, 1, 0)
NOISE = 1e-3


def _problem(cpu, nx, ny, nz, ls, psf, seed):
    atm, z = atmos.seeded(nx, ny, nz, seed=seed)
    cpu.setup(ls, resp=True, atmpar=CART)
    ncol = nx * ny
    for c in range(ncol):
        pg, rho, pel, mw, _ = cpu.he(atm[0, :, c], float(atm[1, -1, c]), z)
        atm[1, :, c], atm[2, :, c], atm[3, :, c], atm[4, :, c] = pg, rho, pel, mw
    return atm, z


def _observe(cpu, atm, z, nx, ny, psf, rng):
    """PSF * (the oracle's synthesis of every pixel) + noise, rows I, Q, U, V."""
    ncol = nx * ny
    syn = []
    for c in range(ncol):
        mo, s, _, _ = cpu.forward1d(atmos.column_model(atm, z, c))
        atm[:, :, c] = mo[:12]
        syn.append(s.reshape(-1))
    syn = np.array(syn)                                   # [ncol][4 numw]
    nf = syn.shape[1]
    maps = syn.T.reshape(nf, nx, ny)
    conv = P.convolve(maps, np.asarray(psf, np.float32).astype(np.float64)).reshape(nf, ncol).T
    return conv + rng.normal(0, NOISE, conv.shape)


def test_coupled_invert3d_vs_oracle(gpu_ctx, cpu):
    import torch
    ls = lines.fe6301_6302(nwave=24, start=-180.0, step=15.0)
    nx, ny, nz = 5, 4, 64
    rng = np.random.default_rng(5)
    psf = rng.random((3, 5)) + np.pad(np.array([[4.0]]), ((1, 1), (2, 2)))
    psf = psf / psf.sum()
    atm, z = _problem(cpu, nx, ny, nz, ls, psf, seed=17)
    gpu_ctx.configure(ls, z)
    gain = gpu_ctx.gain_hsra()
    cpu.set_gain(gain)
    obs = _observe(cpu, atm, z, nx, ny, psf, rng)
    guess = atm.copy()
    guess[0] += np.float32(150.0)
    guess[5:8] *= np.float32(0.8)
    guess[10] += np.float32(5e4)
    nodes = {"T": 1, "Bx": 1, "By": 1, "Bz": 1, "vz": 1}
    ctl = gpu.invert_control(nodes, noise=NOISE, smoothing=False, max_iterations=30)
    gpu_ctx.set_psf(psf)
    try:
        d_atm = torch.from_numpy(guess.copy()).cuda()
        st, chi2, its = gpu_ctx.coupled_invert3d_torch(d_atm, torch.from_numpy(obs.copy()).cuda(), ctl, nx, ny)
        torch.cuda.synchronize()
    finally:
        gpu_ctx.set_psf(None)
    g_chi, g_model = chi2.cpu().numpy(), d_atm.cpu().numpy()
    idx = {gpu.INV_PARAMS.index(k): v for k, v in nodes.items()}
    o_model, o_chi, o_its = OC.coupled_invert3d(cpu, guess, z, obs, nx, ny, idx, psf, noise=NOISE, smoothing=False,
                                                max_iterations=30)
    print("[coupled] iterations device %d oracle %d; map chi2 device %.4g oracle %.4g (start offsets T+150 K, "
          "B x0.8, v+0.5 km/s)" % (its, o_its, g_chi.mean(), o_chi.mean()), flush=True)
    assert abs(its - o_its) <= 2
    assert abs(g_chi.mean() - o_chi.mean()) <= 0.1 * o_chi.mean()
    assert o_chi.mean() < 1.5 and g_chi.mean() < 1.5
    for name, row, tol in (("T", 0, 10.0), ("v_z", 10, 5e3)):
        eg = float(np.abs(g_model[row] - atm[row]).mean())
        eo = float(np.abs(o_model[row] - atm[row]).mean())
        print("[coupled] mean |%s - truth|: device %.3g, oracle %.3g" % (name, eg, eo), flush=True)
        assert eg < tol and eo < tol


def test_coupling_fits_psf_degraded_data(gpu_ctx, cpu):
    """A 12x12 map observed through a 7x7 Airy + wing PSF: the coupled
    inversion reaches the noise (one node per parameter represents the
    start offsets exactly); INVERT3D without the PSF on the same
    observations is left with the PSF's mixing in its residual."""
    import torch
    ls = lines.fe6301_6302(nwave=32, start=-240.0, step=15.0)
    nx = ny = 12
    rng = np.random.default_rng(8)
    psf = atmos.airy_gauss_psf(7, 1.5, 2.0, 0.2)
    atm, z = _problem(cpu, nx, ny, 64, ls, psf, seed=23)
    gpu_ctx.configure(ls, z)
    cpu.set_gain(gpu_ctx.gain_hsra())
    obs = _observe(cpu, atm, z, nx, ny, psf, rng)
    guess = atm.copy()
    guess[0] += np.float32(120.0)
    guess[10] += np.float32(4e4)
    ctl = gpu.invert_control({"T": 1, "Bx": 1, "By": 1, "Bz": 1, "vz": 1}, noise=NOISE, smoothing=False)
    d_obs = torch.from_numpy(obs.copy()).cuda()
    gpu_ctx.set_psf(psf)
    try:
        d1 = torch.from_numpy(guess.copy()).cuda()
        _, chi_c, its_c = gpu_ctx.coupled_invert3d_torch(d1, d_obs, ctl, nx, ny)
    finally:
        gpu_ctx.set_psf(None)
    d2 = torch.from_numpy(guess.copy()).cuda()
    _, chi_u, its_u = gpu_ctx.invert3d_torch(d2, d_obs, ctl, nx, ny)
    torch.cuda.synchronize()
    cc, cu = float(chi_c.mean()), float(chi_u.mean())
    t_err_c = float(np.abs(d1.cpu().numpy()[0] - atm[0]).mean())
    t_err_u = float(np.abs(d2.cpu().numpy()[0] - atm[0]).mean())
    print("[coupled] PSF-degraded 12x12: chi2 coupled %.3g (%d iterations), uncoupled %.3g (%d); mean |T - truth| "
          "%.1f K vs %.1f K" % (cc, its_c, cu, its_u, t_err_c, t_err_u), flush=True)
    assert cc < 1.5
    assert cu > 2.0 * cc
    assert t_err_c < t_err_u
