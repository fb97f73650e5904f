"""Defined failure paths of the forward on the GPU.  The reference prints and
STOPs on each of these conditions; here each becomes an FZ_E* code in the
affected column's status (or the entry point's return code), and every other
column is computed exactly as without the bad one:

  * T = 0 or T < 0 at one height, Pg < 0 at one height: the equation of
    state fails (chemical_equilibrium.f90:259-268);
  * a NaN in B or v_los: non-finite profile / propagation matrix /
    evolution operator (voigt.f:91-95, absorption_matrix.f90:1882-1889,
    rtesolver.f90:346-352);
  * a NaN temperature;
  * a line below 1800 A: the continuum opacity refuses it
    (cont_opacity.f90:83-86), for the HSRA normalisation and every column.

The expected code of each bad column is the oracle's (its FORWARD1D returns
the same codes); the conditions are set on the host and every kernel still
runs to its normal end (the fixed-point iterations are bounded).
"""
import numpy as np
import pytest

from firtez_dz_amd import atmos, gpu, lines

pytestmark = pytest.mark.gpu
CART = (1, 1, 1, 1, 1, 1, 1, 0)
BAD = {3: ("T", 0.0), 5: ("T", -100.0), 6: ("Pg", -1.0), 9: ("Bx", np.nan), 12: ("vz", np.nan), 14: ("T", np.nan)}


def test_bad_columns_report_their_code_only(gpu_ctx, cpu):
    import torch
    ls = lines.fe6301_6302()
    atm, z = atmos.seeded(4, 4, 32, seed=3)
    gpu_ctx.configure(ls, z)
    gain = gpu_ctx.gain_hsra()
    t = torch.from_numpy(atm.copy()).cuda()
    assert int(gpu_ctx.hydrostatic_torch(t).max()) == 0
    good = t.cpu().numpy()
    bad = good.copy()
    for c, (name, v) in BAD.items():
        bad[gpu.ATM.index(name), 10, c] = v
    outs = []
    for cube in (good, bad):
        d = torch.from_numpy(cube.copy()).cuda()
        st, jac, status = gpu_ctx.forward_torch(d, gpu.RF_SPHERICAL)
        torch.cuda.synchronize()
        outs.append((st.cpu().numpy(), jac.cpu().numpy(), status.cpu().numpy()))
    (st0, jac0, s0), (st1, jac1, s1) = outs
    assert np.all(s0 == 0)
    cpu.setup(ls, resp=True, atmpar=CART)
    cpu.set_gain(gain)
    for c in range(16):
        if c in BAD:
            try:
                cpu.forward1d(atmos.column_model(bad, z, c))
                code = 0
            except RuntimeError as e:
                code = int(str(e).split()[-1])
            print("[errors] column %2d %s = %s: device status %d, oracle %d"
                  % (c, BAD[c][0], BAD[c][1], s1[c], code), flush=True)
            assert code != 0 and s1[c] == code, c
        else:
            assert s1[c] == 0
            assert np.array_equal(st1[c], st0[c]) and np.array_equal(jac1[c], jac0[c]), c


def test_line_below_1800_angstrom(gpu_ctx):
    import torch
    ln = lines.Line(**dict(lines.line_from_db(4).__dict__, lambda0=1500.0))
    ls = lines.LineSetup([ln], [lines.Region([0], 32, -300.0, 20.0)])
    atm, z = atmos.seeded(2, 2, 32, seed=4)
    gpu_ctx.configure(ls, z)
    with pytest.raises(gpu.FzError) as ei:
        gpu_ctx.gain_hsra()
    assert "(4)" in str(ei.value)
    gpu_ctx.set_gain(np.ones(4 * gpu_ctx.numw, np.float32))
    t = torch.from_numpy(atm.copy()).cuda()
    assert int(gpu_ctx.hydrostatic_torch(t).max()) == 0
    st, _, status = gpu_ctx.forward_torch(t, gpu.RF_NONE)
    torch.cuda.synchronize()
    s = status.cpu().numpy()
    print("[errors] line at 1500 A: column status %s" % np.unique(s), flush=True)
    assert np.all(s == 4)
