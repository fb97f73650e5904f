"""SET_TAU's assist initialisers (invert3d.f90:528-700) in the oracle
(oracle/invert.py), pinned by what they are meant to recover:

  * T: the brightness temperature of the mean intensity, i.e. the Planck
    function ratio B(T)/B(6390 K) equals the intensity (Planck's law
    evaluated directly), and 6390 K for unit intensity;
  * P: Pg scaled to 1.333e5 at height ceiling(nz/3), ratios kept;
  * V: symmetric absorption lines shifted by small known amounts give v_z
    proportional to the shift, redshift positive, 2 km/s for the largest;
  * B: Q and U proportional to cos 2phi and sin 2phi give the azimuth phi of
    the transverse field (Bx, By) = 1000 G (-sin phi, cos phi) scaled by the
    relative amplitude; Bz follows -V projected on the profile derivative.
"""
import math

import numpy as np

from oracle import invert

NUMW, ITZ, NZ, NCOL = 80, 80, 30, 40
LAM0 = 6301.5


def base_model():
    atm = np.zeros((12, NZ, NCOL), np.float32)
    atm[0] = np.linspace(9000, 4000, NZ, dtype=np.float32)[:, None]
    atm[1] = np.geomspace(1e5, 10.0, NZ).astype(np.float32)[:, None]
    return atm


def profile(shift=0.0, depth=0.6, width=6.0):
    w = np.arange(NUMW) - NUMW / 2 + 0.5
    return 1.0 - depth * np.exp(-((w - shift) / width) ** 2)


def test_assist_t_is_brightness_temperature():
    rng = np.random.default_rng(1)
    mean = rng.uniform(0.4, 1.3, NCOL)
    mean[0] = 1.0
    obs = np.repeat(mean[:, None], 4 * NUMW, axis=1)
    out = invert.assist(obs, base_model(), invert.ASSIST_T, NUMW, ITZ, 4, LAM0)
    kz = math.ceil(NZ / 3) - 1
    lam = LAM0 * 1e-8
    x = lambda T: invert.HPLA * invert.LIGHT / (lam * invert.KBOL * T)
    planck = lambda T: 1.0 / np.expm1(x(T))
    tb = out[0, kz].astype(np.float64)
    assert abs(tb[0] - 6390.0) < 1e-3
    assert np.max(np.abs(planck(tb) / planck(6390.0) - mean)) < 1e-5
    # every height shifted by the same amount, log tau5 reset around kz
    shift = out[0].astype(np.float64) - base_model()[0].astype(np.float64)
    assert np.allclose(shift, shift[kz][None, :], atol=2e-3)
    assert np.all(out[11, kz] == 0) and np.all(np.delete(out[11], kz, axis=0) == 1)


def test_assist_p_scales_to_reference_pressure():
    out = invert.assist(np.ones((NCOL, 4 * NUMW)), base_model(), invert.ASSIST_P, NUMW, ITZ, 4, LAM0)
    kz = math.ceil(NZ / 3) - 1
    assert np.allclose(out[1, kz], 1.333e5, rtol=1e-6)
    ratio = out[1] / base_model()[1]
    assert np.allclose(ratio, ratio[0][None, :], rtol=1e-6)


def test_assist_v_follows_line_shift():
    shifts = np.linspace(-0.3, 0.3, NCOL)
    obs = np.ones((NCOL, 4 * NUMW))
    for p, s in enumerate(shifts):
        obs[p, :NUMW] = profile(s)
    out = invert.assist(obs, base_model(), invert.ASSIST_V, NUMW, ITZ, 4, LAM0)
    vz = out[10, 0].astype(np.float64)
    assert np.all(out[10] == out[10, 0][None, :])
    assert abs(np.max(np.abs(vz)) - 2.0e5) < 1.0
    # redshift (positive shift) -> positive v_z, linear to second order in the shift
    assert np.max(np.abs(vz / 2.0e5 - shifts / np.max(np.abs(shifts)))) < 2e-2


def test_assist_b_recovers_azimuth():
    rng = np.random.default_rng(3)
    phi = rng.uniform(-math.pi / 2, math.pi / 2, NCOL)
    amp = rng.uniform(0.2, 1.0, NCOL)
    bz = rng.uniform(-1.0, 1.0, NCOL)
    i0 = profile()
    d1 = np.gradient(i0)
    d2 = np.gradient(d1)
    obs = np.zeros((NCOL, 4 * NUMW))
    obs[:, :NUMW] = i0
    obs[:, NUMW:2 * NUMW] = (amp * np.cos(2 * phi))[:, None] * d2
    obs[:, 2 * NUMW:3 * NUMW] = (amp * np.sin(2 * phi))[:, None] * d2
    obs[:, 3 * NUMW:] = -bz[:, None] * d1
    out = invert.assist(obs, base_model(), invert.ASSIST_B, NUMW, ITZ, 4, LAM0)
    bx, by, bzo = (out[i, 0].astype(np.float64) for i in (5, 6, 7))
    # the transverse direction: (Bx, By) along (-sin phi, cos phi), modulo 180 degrees
    az = np.arctan2(-bx, by)
    dphi = np.angle(np.exp(2j * (az - phi))) / 2
    assert np.max(np.abs(dphi)) < 1e-5
    bt = np.hypot(bx, by)
    assert np.allclose(bt / bt.max(), amp / amp.max(), rtol=1e-5)
    assert np.allclose(bzo / np.abs(bzo).max(), bz / np.abs(bz).max(), atol=1e-5)
