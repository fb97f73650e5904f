"""Microturbulence xi in the oracle (fzc_set_microturbulence), pinned by
closed forms and finite differences.  The reference's Doppler width is
thermal only (absorption_matrix.f90:1794-1795); with xi every line's width
is sqrt(2kT/m + xi^2):
  * the radiative damping parameter Gamma_rad lambda / (4 pi v_D) with that
    v_D, and the line opacity's 1/v_D factor: kl(T, xi) v_D(T, xi) equals
    kl(T, 0) v_D(T, 0);
  * xi = 0 is the thermal width bit for bit;
  * the per-height T response of a column (whose Voigt argument, damping
    and opacity all carry d v_D/dT = k / (m v_D)) against central
    differences of the forward in the temperature of one interior height."""
import math

import numpy as np
import pytest

from firtez_dz_amd import atmos, lines

from test_oracle_damping import KBOL, MAMU, MATOM

DPI = math.pi
XI = 8.0e4
CART = (1, 1, 1, 1, 1, 1, 1, 0)


@pytest.fixture
def xi_cpu(cpu):
    yield cpu
    cpu.set_microturbulence(0.0)


def _matom(zn):
    return MATOM[zn - 1]


def test_doppler_width_in_damping_and_opacity(xi_cpu):
    cpu = xi_cpu
    line = lines.line_from_db(4)                     # Fe I 6301.5
    cpu.setup(lines.LineSetup([line], [lines.Region([0], 4, 0.0, 10.0)]), resp=True, atmpar=CART)
    lam = line.lambda0 * 1e-8
    for T in (4000.0, 6500.25):
        vth2 = 2.0 * KBOL * T / (MAMU * _matom(line.zn))
        a0, _ = cpu.damping(0, T, 0.0, np.zeros(25))
        e, k0, _ = cpu.opac_line(0, T, 1e13, 1e17, np.zeros(25))
        cpu.set_microturbulence(XI)
        a1, _ = cpu.damping(0, T, 0.0, np.zeros(25))
        e1, k1, _ = cpu.opac_line(0, T, 1e13, 1e17, np.zeros(25))
        cpu.set_microturbulence(0.0)
        a2, _ = cpu.damping(0, T, 0.0, np.zeros(25))
        ref = (0.2223 / lam ** 2) * lam / (4.0 * DPI * math.sqrt(vth2 + XI * XI))
        assert e == 0 and e1 == 0
        assert abs(a1 - ref) <= 1e-12 * ref
        assert a2 == a0
        assert abs(k1 * math.sqrt(vth2 + XI * XI) - k0 * math.sqrt(vth2)) <= 1e-12 * k0 * math.sqrt(vth2)


def test_temperature_response_by_central_differences(xi_cpu):
    cpu = xi_cpu
    ls = lines.fe6301_6302(nwave=40, start=-400.0, step=20.0)
    atm, z = atmos.seeded(1, 1, 64, seed=5)
    cpu.setup(ls, resp=True, atmpar=CART)
    cpu.set_gain(cpu.gain())
    pg, rho, pel, mw, _ = cpu.he(atm[0, :, 0], float(atm[1, -1, 0]), z)
    atm[1, :, 0], atm[2, :, 0], atm[3, :, 0], atm[4, :, 0] = pg, rho, pel, mw
    cpu.set_microturbulence(XI)
    m = atmos.column_model(atm, z, 0)
    _, s0, dsyn, _ = cpu.forward1d(m.copy())
    worst = 0.0
    for k in (20, 35, 45):
        h = np.float32(2.0)
        mp, mm = m.copy(), m.copy()
        mp[0, k] += h
        mm[0, k] -= h
        _, sp, _, _ = cpu.forward1d(mp)
        _, sm, _, _ = cpu.forward1d(mm)
        fd = (sp - sm) / (float(mp[0, k]) - float(mm[0, k]))     # (4, numw)
        rf = dsyn[k, :, 0, :].T                                    # (4, numw)
        err = float(np.max(np.abs(rf - fd)) / np.max(np.abs(fd)))
        worst = max(worst, err)
        print("height %d: T response vs central difference, max err %.2e of max" % (k, err))
    assert worst < 2e-3
    # the width changes the profile: the xi = 0 synthesis differs
    cpu.set_microturbulence(0.0)
    _, s1, _, _ = cpu.forward1d(m.copy())
    assert np.max(np.abs(s1 - s0)) > 1e-3
