"""NLTE departure coefficients in the oracle (POPL/POPU per region and
height: OPAC_LINE line_opacity.f90:34-217, SOURCE_FUNCTION(_DER)
rtesolver.f90:1008-1066), pinned to the reference's own code:

  * tests/golden/nlte.npz was written by the reference's compiled modules
    with non-unit POPL/POPU handed to FORWARD_PARAM (tests/golden/
    make_golden.py): Stokes and the seven Cartesian response functions;
  * unit coefficients set explicitly reproduce LTE bit for bit;
  * the LTE region of the fixture is unaffected by the other region's
    coefficients (region-local, as POPL(L,K) is indexed by region).
"""
import numpy as np

from firtez_dz_amd import lines

CART = (1, 1, 1, 1, 1, 1, 1, 0)


def _setup(cpu, g):
    ls = lines.setup([((4,), int(g["nwave"]), float(g["wave_start"]), float(g["wave_step"])),
                      ((5,), int(g["nwave"]), float(g["wave_start"]), float(g["wave_step"]))])
    cpu.setup(ls, resp=True, atmpar=CART)
    cpu.set_gain(g["gain"])


def test_nlte_vs_reference_fixture(cpu, golden):
    g = golden["nlte"]
    _setup(cpu, g)
    cpu.set_departure(g["popl"], g["popu"])
    try:
        mo, syn, dsyn, _ = cpu.forward1d(g["model_in"])
    finally:
        cpu.set_departure(None)
    scale = np.abs(g["stokes"]).max()
    assert np.max(np.abs(syn - g["stokes"])) <= 1e-12 * scale
    rf = dsyn[g["rf_heights"]][:, :, :7, :]
    for p in range(7):
        sc = np.abs(g["rf"][:, :, p, :]).max()
        assert np.max(np.abs(rf[:, :, p, :] - g["rf"][:, :, p, :])) <= 1e-6 * sc, p
    # the departures change the first region's profiles and leave the second's
    _setup(cpu, g)
    _, lte, _, _ = cpu.forward1d(g["model_in"])
    n = int(g["nwave"])
    assert np.max(np.abs(lte[:, :n] - syn[:, :n])) > 1e-3
    assert np.max(np.abs(lte[:, n:] - syn[:, n:])) <= 1e-13 * scale


def test_unit_coefficients_are_lte(cpu, golden):
    g = golden["nlte"]
    _setup(cpu, g)
    _, lte, dl, _ = cpu.forward1d(g["model_in"])
    nz = g["model_in"].shape[1]
    cpu.set_departure(np.ones((2, nz)), np.ones((2, nz)))
    try:
        _, one, d1, _ = cpu.forward1d(g["model_in"])
    finally:
        cpu.set_departure(None)
    assert np.array_equal(lte, one) and np.array_equal(dl, d1)
