"""The oracle's PSF-coupled inversion (oracle/coupled.py) pinned by what
the mathematics fixes, on the CPU:
  * the gradient it steps with, J_q^T W (PSF^T r)_q, is the derivative of
    the map's chi^2 (central finite differences through the oracle's
    forward, an asymmetric PSF so that the convolution in place of its
    adjoint fails);
  * on a one-pixel map with the unit PSF it is the uncoupled INVERT3D of
    oracle/invert.py (itself checked against the device and the
    reference's test_inv1): same iterations, same model, same chi^2;
  * a one-pixel PSF a scales the problem: the coupled inversion of obs
    equals the uncoupled inversion of obs / a with the noise / a (the
    sum PSF^2 scaling of the step)."""
import numpy as np
import pytest

from firtez_dz_amd import atmos, lines
from oracle import coupled as OC
from oracle import psf as P
from oracle.invert import invert3d, weights

CART = (1, 1, 1, 1, 1, 1, 1, 0)
NOISE = 1e-3


def _truth(cpu, nx, ny, nz, ls, seed):
    atm, z = atmos.seeded(nx, ny, nz, seed=seed)
    cpu.setup(ls, resp=True, atmpar=CART)
    cpu.set_gain(cpu.gain())
    for c in range(nx * ny):
        pg, rho, pel, mw, _ = cpu.he(atm[0, :, c], float(atm[1, -1, c]), z)
        atm[1, :, c], atm[2, :, c], atm[3, :, c], atm[4, :, c] = pg, rho, pel, mw
    return atm, z


def _synth(cpu, atm, z):
    out = []
    for c in range(atm.shape[2]):
        mo, s, _, _ = cpu.forward1d(atmos.column_model(atm, z, c))
        atm[:, :, c] = mo[:12]
        out.append(s.reshape(-1))
    return np.array(out)


def test_coupled_gradient_is_the_map_chi2_derivative(cpu):
    from oracle import lm_cfg
    ls = lines.fe6301_6302(nwave=20, start=-150.0, step=15.0)
    nx, ny, nz = 4, 3, 32
    atm, z = _truth(cpu, nx, ny, nz, ls, seed=17)
    rng = np.random.default_rng(5)
    psf = rng.random((3, 5)) + np.pad(np.array([[3.0]]), ((1, 1), (2, 2)))
    psf = (psf / psf.sum()).astype(np.float32).astype(np.float64)
    ncol = nx * ny
    isig = weights((1, 1, 1, 1), (1, 1, 1, 1), NOISE, cpu.numw)
    nfreq = len(isig)
    obs = P.convolve(_synth(cpu, atm.copy(), z).T.reshape(nfreq, nx, ny), psf).reshape(nfreq, ncol).T
    guess = atm.copy()
    # v_z offsets that differ from pixel to pixel, so that the residual maps
    # vary across the map and the PSF's asymmetry shows
    guess[10] += rng.uniform(-8e4, 8e4, ncol).astype(np.float32)[None, :]

    def chi2(model):
        syn = _synth(cpu, model.copy(), z)
        r = obs - P.convolve(syn.T.reshape(nfreq, nx, ny), psf).reshape(nfreq, ncol).T
        return float(np.sum(r * r * isig)), r

    c0, r0 = chi2(guess)
    gres = P.convolve(r0.T.reshape(nfreq, nx, ny), psf, adjoint=True).reshape(nfreq, ncol).T
    gwrong = P.convolve(r0.T.reshape(nfreq, nx, ny), psf).reshape(nfreq, ncol).T
    # one v_z node (its response has no part in the bottom boundary, which the
    # reference's T response leaves out)
    sel = [0, 0, 0, 0, 0, 0, 1, 0]
    cfg = lm_cfg(nz, cpu.numw, sel, sel, inu=nfreq)
    coefs = cpu.coefs_block(cfg, {6: cpu.int_coefs(nz, 1)})
    h = np.float32(500.0)
    for q in (0, 5, 10):
        m = guess.copy()
        _, _, dsyn, _ = cpu.forward1d(atmos.column_model(m, z, q))
        J = cpu.eqv_dsyn(cfg, coefs, dsyn)[:, 0].astype(np.float64)
        g = -2.0 * np.sum(J * isig * gres[q])
        g_wrong = -2.0 * np.sum(J * isig * gwrong[q])
        mp, mm = guess.copy(), guess.copy()
        mp[10, :, q] += h
        mm[10, :, q] -= h
        fd = (chi2(mp)[0] - chi2(mm)[0]) / (2.0 * float(h))
        print("pixel %d: J^T W PSF^T r %.6e, finite difference %.6e, J^T W PSF r %.6e" % (q, g, fd, g_wrong))
        assert abs(g - fd) <= 2e-3 * abs(fd)
        assert abs(g_wrong - fd) > 2e-2 * abs(fd)


def _one_pixel(cpu, seed):
    ls = lines.fe6301_6302(nwave=24, start=-180.0, step=15.0)
    atm, z = _truth(cpu, 1, 1, 48, ls, seed=seed)
    syn = _synth(cpu, atm.copy(), z)
    guess = atm.copy()
    guess[0] += np.float32(120.0)
    guess[5:8] *= np.float32(0.8)
    guess[10] += np.float32(4e4)
    return atm, z, syn, guess


# one node per parameter: one cycle that converges to the noise, through
# accepted and rejected steps; T up to 4 nodes: two cycles, whose different
# stop thresholds MAX(1, MAXSTEPS - (MAXITER - cycle)) end them
NODES_1 = {0: 1, 3: 1, 4: 1, 5: 1, 6: 1}
NODES_2CYC = {0: 4, 3: 1, 4: 1, 5: 1, 6: 1}


@pytest.mark.parametrize("nodes", [NODES_1, NODES_2CYC], ids=["one-cycle", "two-cycles"])
def test_one_pixel_unit_psf_is_the_uncoupled_inversion(cpu, nodes):
    atm, z, syn, guess = _one_pixel(cpu, seed=29)
    obs = syn + np.random.default_rng(3).normal(0, NOISE, syn.shape)
    mc, chic, itc = OC.coupled_invert3d(cpu, guess, z, obs, 1, 1, nodes, np.ones((1, 1)), noise=NOISE,
                                        smoothing=False)
    mu, _, chiu, itu = invert3d(cpu, guess, z, obs, 1, 1, nodes, noise=NOISE, smoothing=False)
    print("coupled %d iterations chi2 %.6g; uncoupled %d iterations chi2 %.6g" % (itc, chic[0], itu, chiu[0]))
    assert itc == itu
    assert np.array_equal(mc, mu)
    assert abs(chic[0] - chiu[0]) <= 1e-12 * chiu[0]


def test_one_pixel_psf_scale(cpu):
    atm, z, syn, guess = _one_pixel(cpu, seed=31)
    a = 0.5
    obs = a * syn + np.random.default_rng(4).normal(0, NOISE, syn.shape)
    mc, chic, itc = OC.coupled_invert3d(cpu, guess, z, obs, 1, 1, NODES_1, np.full((1, 1), a), noise=NOISE,
                                        smoothing=False)
    mu, _, chiu, itu = invert3d(cpu, guess, z, obs / a, 1, 1, NODES_1, noise=NOISE / a, smoothing=False)
    print("PSF %.2f: coupled %d iterations chi2 %.6g; uncoupled of obs/a %d iterations chi2 %.6g"
          % (a, itc, chic[0], itu, chiu[0]))
    assert itc == itu
    assert np.allclose(mc, mu, rtol=1e-5, atol=1e-3)
    assert abs(chic[0] - chiu[0]) <= 1e-9 * chiu[0]
