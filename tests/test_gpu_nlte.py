"""NLTE departure coefficients on the GPU (fz_set_departure: POPL/POPU per
region and height in OPAC_LINE and SOURCE_FUNCTION, line_opacity.f90:101-106,
rtesolver.f90:1030-1064):

  * the reference's own output for two unblended regions with
    height-dependent coefficients in one of them (tests/golden/nlte.npz):
    Stokes and the seven Cartesian response functions;
  * a seeded 4x4x64 cube against the oracle with departures in both
    regions: Stokes and the spherical Jacobians;
  * unit coefficients give the LTE output bit for bit, and switching them
    off restores it.
"""
import numpy as np
import pytest

from firtez_dz_amd import atmos, gpu, lines

pytestmark = pytest.mark.gpu
STOKES_TOL = 1e-4
RF_TOL = 2e-3
CART = (1, 1, 1, 1, 1, 1, 1, 0)


def report(name, err, tol):
    print("[nlte] %-50s max err %.3e (tol %.1e)" % (name, err, tol), flush=True)
    assert err <= tol, name


def _run(ctx, atm, rf):
    import torch
    t = torch.from_numpy(np.ascontiguousarray(atm)).cuda()
    st, jac, status = ctx.forward_torch(t, rf)
    torch.cuda.synchronize()
    return st.cpu().numpy(), (jac.cpu().numpy() if jac is not None else None), status.cpu().numpy()


def _setup(g):
    n, w0, dw = int(g["nwave"]), float(g["wave_start"]), float(g["wave_step"])
    return lines.setup([((4,), n, w0, dw), ((5,), n, w0, dw)])


def test_nlte_vs_reference_fixture(gpu_ctx, golden):
    g = golden["nlte"]
    m = g["model_in"]
    gpu_ctx.configure(_setup(g), m[12].astype(np.float64))
    gpu_ctx.set_gain(g["gain"])
    gpu_ctx.set_departure(g["popl"], g["popu"])
    atm = np.ascontiguousarray(m[:12, :, None])
    st, jac, status = _run(gpu_ctx, atm, gpu.RF_CARTESIAN)
    gpu_ctx.set_departure(None)
    assert status[0] == 0
    report("NLTE Stokes vs reference", np.max(np.abs(st[0] - g["stokes"])), STOKES_TOL)
    names = ("T", "Pg", "rho", "Bx", "By", "Bz", "vz")
    for p in range(7):
        ours = jac[0, p][g["rf_heights"]]
        ref = np.transpose(g["rf"][:, :, p, :], (0, 2, 1))
        sc = np.abs(ref).max()
        if sc > 0:
            report("NLTE RF d/d%s vs reference" % names[p], np.max(np.abs(ours - ref)) / sc, RF_TOL)


def test_nlte_seeded_cube_vs_oracle(gpu_ctx, cpu):
    import torch
    ls = lines.setup([((4,), 64, -320.0, 10.0), ((5,), 48, -240.0, 10.0)])
    atm, z = atmos.seeded(4, 4, 64, seed=12)
    nz = 64
    x = np.linspace(0.0, 1.0, nz)
    popl = np.vstack([1.0 + 0.3 * x ** 2, 0.9 + 0.2 * x])
    popu = np.vstack([1.0 - 0.4 * x ** 1.5, 1.1 - 0.3 * x ** 2])
    gpu_ctx.configure(ls, z)
    gain = gpu_ctx.gain_hsra()
    t = torch.from_numpy(atm.copy()).cuda()
    assert int(gpu_ctx.hydrostatic_torch(t).max()) == 0
    atm_in = t.cpu().numpy()
    gpu_ctx.set_departure(popl, popu)
    st, jac, status = _run(gpu_ctx, atm_in, gpu.RF_SPHERICAL)
    gpu_ctx.set_departure(None)
    st_lte, _, _ = _run(gpu_ctx, atm_in, gpu.RF_NONE)
    assert np.all(status == 0)
    cpu.setup(ls, resp=True, atmpar=CART)
    cpu.set_gain(gain)
    cpu.set_departure(popl, popu)
    try:
        models = np.stack([atmos.column_model(atm_in, z, c) for c in range(16)])
        _, syn, dsyn = cpu.forward_batch(models)
    finally:
        cpu.set_departure(None)
    report("NLTE seeded Stokes vs oracle (16 columns)", np.max(np.abs(st - np.transpose(syn, (0, 2, 1)))),
           STOKES_TOL)
    assert np.max(np.abs(st - st_lte)) > 1e-3          # the departures matter
    worst = 0.0
    for c in range(16):
        tr = lambda p: np.transpose(dsyn[c][:, :, p, :], (0, 2, 1))
        bx, by, bz = (float(models[c][i, 0]) for i in (5, 6, 7))
        B = np.sqrt(bx * bx + by * by + bz * bz)
        gm, f = np.arccos(bz / B), np.arctan2(by, bx)
        dx, dy, dzz = tr(3), tr(4), tr(5)
        ref = [tr(0), (bx * dx + by * dy + bz * dzz) / B,
               B * np.cos(gm) * np.cos(f) * dx + B * np.cos(gm) * np.sin(f) * dy - B * np.sin(gm) * dzz,
               -B * np.sin(gm) * np.sin(f) * dx + B * np.sin(gm) * np.cos(f) * dy, tr(6)]
        for p in range(5):
            sc = np.abs(ref[p]).max()
            if sc > 0:
                worst = max(worst, np.max(np.abs(jac[c, p] - ref[p])) / sc)
    report("NLTE seeded spherical Jacobians vs oracle", worst, RF_TOL)


def test_unit_departure_is_lte(gpu_ctx):
    import torch
    ls = lines.fe6301_6302(nwave=56, start=-600.0, step=40.0)
    atm, z = atmos.seeded(2, 2, 32, seed=9)
    gpu_ctx.configure(ls, z)
    gpu_ctx.gain_hsra()
    t = torch.from_numpy(atm.copy()).cuda()
    assert int(gpu_ctx.hydrostatic_torch(t).max()) == 0
    atm = t.cpu().numpy()
    st0, j0, s0 = _run(gpu_ctx, atm, gpu.RF_SPHERICAL)
    assert np.all(s0 == 0) and np.all(np.isfinite(st0))
    gpu_ctx.set_departure(np.ones((1, 32)), np.ones((1, 32)))
    st1, j1, _ = _run(gpu_ctx, atm, gpu.RF_SPHERICAL)
    gpu_ctx.set_departure(None)
    st2, j2, _ = _run(gpu_ctx, atm, gpu.RF_SPHERICAL)
    assert np.array_equal(st0, st1) and np.array_equal(j0, j1)
    assert np.array_equal(st0, st2) and np.array_equal(j0, j2)
    with pytest.raises(gpu.FzError):
        gpu_ctx.set_departure(np.zeros((1, 32)), np.ones((1, 32)))


def test_departure_survives_lsf_and_resets_with_configure(gpu_ctx):
    """fz_set_lsf keeps the departure coefficients (the order of the two
    set-up calls does not matter); fz_configure resets them to LTE."""
    import torch
    ls = lines.fe6301_6302(nwave=56, start=-600.0, step=40.0)
    atm, z = atmos.seeded(2, 2, 32, seed=9)
    gpu_ctx.configure(ls, z)
    gpu_ctx.gain_hsra()
    t = torch.from_numpy(atm.copy()).cuda()
    assert int(gpu_ctx.hydrostatic_torch(t).max()) == 0
    atm = t.cpu().numpy()
    lte, _, _ = _run(gpu_ctx, atm, gpu.RF_NONE)
    x = np.linspace(0.0, 1.0, 32)
    popl, popu = (1.0 + 0.3 * x)[None, :], (1.0 - 0.3 * x)[None, :]
    kern, valid = lines.lsf_kernel(ls, [40.0], [0.0])
    try:
        gpu_ctx.set_departure(popl, popu)
        gpu_ctx.set_lsf(kern, valid)
        a, _, _ = _run(gpu_ctx, atm, gpu.RF_NONE)
        gpu_ctx.set_lsf(None)
        gpu_ctx.set_departure(None)
        gpu_ctx.set_lsf(kern, valid)
        gpu_ctx.set_departure(popl, popu)
        b, _, _ = _run(gpu_ctx, atm, gpu.RF_NONE)
        gpu_ctx.set_departure(None)
        c, _, _ = _run(gpu_ctx, atm, gpu.RF_NONE)
    finally:
        gpu_ctx.set_lsf(None)
        gpu_ctx.set_departure(None)
    assert np.array_equal(a, b)
    assert np.max(np.abs(a - c)) > 1e-3
    gpu_ctx.set_departure(popl, popu)
    gpu_ctx.configure(ls, z)
    gpu_ctx.gain_hsra()
    d, _, _ = _run(gpu_ctx, atm, gpu.RF_NONE)
    assert np.array_equal(d, lte)
