"""fz_autotune and the launch tiles change how the columns are cut into
launches, never a column's arithmetic: the tiles it picks, and a few fixed
ones, reproduce the default forward bit for bit (Stokes, Jacobians, status
and the updated model cube).  The cube is larger than the smallest
radiative-transfer and microphysics candidates, so the trials really split
it.  A context of its own keeps the chosen tiles out of the other tests."""
import numpy as np
import pytest

from firtez_dz_amd import atmos, gpu, lines

pytestmark = pytest.mark.gpu


def test_autotuned_tiles_are_bit_identical():
    import torch
    ls = lines.fe6301_6302(nwave=24, start=-150.0, step=15.0)
    nx = ny = 72                                    # 5184 columns
    atm, z = atmos.seeded(nx, ny, 32, seed=13)
    ctx = gpu.Context(0)
    try:
        ctx.configure(ls, z)
        ctx.gain_hsra()
        base = torch.from_numpy(atm.copy()).cuda()
        assert int(ctx.hydrostatic_torch(base).max()) == 0

        def run():
            t = base.clone()
            st, jac, status = ctx.forward_torch(t, gpu.RF_SPHERICAL)
            torch.cuda.synchronize()
            return [x.cpu().numpy() for x in (st, jac, status, t)]

        ref = run()
        assert np.all(ref[2] == 0)
        ncol = nx * ny
        stokes = torch.empty(ctx._stokes_shape(ncol), dtype=torch.float32, device="cuda")
        jac = torch.empty((ncol, gpu.NPARAM[gpu.RF_SPHERICAL], 32, 4, ctx.numw), dtype=torch.float32,
                          device="cuda")
        status = torch.zeros(ncol, dtype=torch.int32, device="cuda")
        trial = base.clone()
        tiles = ctx.autotune_torch(trial, gpu.RF_SPHERICAL, stokes, jac, status)
        torch.cuda.synchronize()
        print("[autotune] chosen tiles: radiative transfer %d, microphysics %d" % tiles, flush=True)
        cases = [("autotuned", None), ("rt 2048 / block 128", (2048, 128)), ("rt 3000 / block 512", (3000, 512)),
                 ("rt 4096 / block 256", (4096, 256))]
        for name, tb in cases:
            if tb:
                ctx.set_tile(*tb)
            got = run()
            for a, b, what in zip(got, ref, ("stokes", "jacobian", "status", "model")):
                assert np.array_equal(a, b, equal_nan=True), (name, what)
            print("[autotune] %s: bit-identical" % name, flush=True)
    finally:
        ctx.close()
