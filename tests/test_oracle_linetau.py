"""LINE OPTICAL DEPTH in the oracle (fzc_set_line_tau): log10 of
tau_lambda(z) = trapezoid from the top (1e-9) of KLINTAU * rho with
KLINTAU = ETAI / N_H / m_u / MW (forward.f90:966-968, 1008-1015).  Pinned
where it has a closed form: a line too weak to absorb (log gf = -30) at
5000 A leaves ETAI the 5000 A continuum opacity, so KLINTAU = kc5 / MW with
the kc5 the forward returns, and the line optical depth must equal that
trapezoid (REAL(4) steps, as the reference writes them) evaluated here; the
top height is log10(1e-9)."""
import dataclasses

import numpy as np

from firtez_dz_amd import atmos, lines

CART = (1, 1, 1, 1, 1, 1, 1, 0)


def test_line_tau_of_an_absent_line_is_the_continuum_trapezoid(cpu):
    base = lines.line_from_db(4)
    ghost = dataclasses.replace(base, lambda0=5000.0, loggf=-30.0)
    ls = lines.LineSetup([ghost], [lines.Region([0], 3, -50.0, 50.0)])
    atm, z = atmos.seeded(1, 1, 40, dz_km=15.0, seed=3)
    cpu.setup(ls, resp=False, atmpar=CART)
    cpu.set_gain(cpu.gain())
    pg, rho, pel, mw, _ = cpu.he(atm[0, :, 0], float(atm[1, -1, 0]), z)
    atm[1, :, 0], atm[2, :, 0], atm[3, :, 0], atm[4, :, 0] = pg, rho, pel, mw
    mo, _, _, kc5, tl = cpu.forward1d_line_tau(atmos.column_model(atm, z, 0))
    nz = mo.shape[1]
    thr = 2.0e10
    kl = np.where(kc5 < thr, (kc5 / mo[4].astype(np.float64)), 0.0).astype(np.float32)
    rho_f = mo[2].astype(np.float32)
    zz = np.asarray(mo[12], np.float32).astype(np.float64)
    tau = np.float32(1e-9)
    ref = np.zeros(nz)
    ref[-1] = np.log10(np.float32(1e-9), dtype=np.float32)
    for k in range(nz - 2, -1, -1):
        x = (kl[k] * rho_f[k] + kl[k + 1] * rho_f[k + 1]) / np.float32(2.0)
        tau = np.float32(np.float64(tau) + np.float64(x) * ((zz[k + 1] - zz[k]) * 1000.0 * 100.0))
        ref[k] = np.log10(tau, dtype=np.float32)
    err = float(np.max(np.abs(tl[:, 1] - ref)))
    print("line optical depth of an absent line vs the continuum trapezoid: max |d log10 tau| %.2e" % err)
    assert tl[-1, 1] == ref[-1]
    assert err <= 2e-6
    assert np.all(np.diff(tl[:, 1]) <= 0)         # tau grows downward
