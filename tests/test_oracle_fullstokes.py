"""FULL_STOKES in the oracle (fzc_set_full_stokes), pinned by what fixes it:
the Stokes vector after height k is the emergent Stokes vector of the same
column cut at height k (each layer's solve is local and, on a uniform grid,
the cut column's top layer has the same thickness), so every height of the
output must match a forward of the truncated column; the top height is the
emergent vector itself."""
import numpy as np

from firtez_dz_amd import atmos, lines

CART = (1, 1, 1, 1, 1, 1, 1, 0)


def test_full_stokes_is_the_truncated_column(cpu):
    ls = lines.fe6301_6302(nwave=30, start=-300.0, step=20.0)
    atm, z = atmos.seeded(1, 1, 48, dz_km=15.0, seed=9)
    cpu.setup(ls, resp=False, atmpar=CART)
    cpu.set_gain(cpu.gain())
    pg, rho, pel, mw, _ = cpu.he(atm[0, :, 0], float(atm[1, -1, 0]), z)
    atm[1, :, 0], atm[2, :, 0], atm[3, :, 0], atm[4, :, 0] = pg, rho, pel, mw
    m = atmos.column_model(atm, z, 0)
    _, syn, _, _, fs = cpu.forward1d_full(m.copy())
    scale = float(np.max(np.abs(syn)))
    assert np.max(np.abs(fs[-1] - syn)) <= 1e-6 * scale
    worst = 0.0
    for k in (5, 20, 33, 46):
        _, syn_k, _, _ = cpu.forward1d(np.ascontiguousarray(m[:, :k + 1]))
        worst = max(worst, float(np.max(np.abs(fs[k] - syn_k))) / scale)
    print("full Stokes vs truncated columns: max diff %.2e of max |I|" % worst)
    assert worst <= 1e-6
    # the bottom is the boundary: Q, U, V zero
    assert np.all(fs[0, 1:] == 0.0)
