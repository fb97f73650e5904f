"""GPU parity: the HIP path (through the C ABI) against the reference's golden
outputs and the FP64 oracle on the same seeded inputs.

Tolerances (FP32 bulk arithmetic on the device, FP64 in the oracle):
  Stokes          max |dS| <= 1e-4 I_c   (I_c: HSRA continuum = 1 after GAIN1D)
  Jacobians       max |dJ| <= RF_TOL * max|J| per parameter: the error of
                  d(I,Q,U,V)/dx over all heights and wavelengths, relative to
                  the largest response of the continuum-normalised spectrum
                  to that parameter (so a Stokes parameter with almost no
                  signal, e.g. Q for a 3 G transverse field, is judged on the
                  scale of the parameter's effect, not on its own tiny scale)
  EOS / HE (FP64) relative 1e-9
"""
import numpy as np
import pytest

from firtez_dz_amd import atmos, lines

pytestmark = pytest.mark.gpu
STOKES_TOL = 1e-4
RF_TOL = 2e-3
CART = (1, 1, 1, 1, 1, 1, 1, 0)


def torch_dev():
    import torch
    assert torch.cuda.is_available(), "GPU test without a GPU"
    return torch


def run_gpu(ctx, atm_np, rf_mode):
    torch = torch_dev()
    atm = torch.from_numpy(np.ascontiguousarray(atm_np)).cuda()
    st, jac, status = ctx.forward_torch(atm, rf_mode)
    torch.cuda.synchronize()
    return (atm.cpu().numpy(), st.cpu().numpy(), None if jac is None else jac.cpu().numpy(),
            status.cpu().numpy())


def model_rows(models):
    """(ncol,13,nz) models -> device rows (12,nz,ncol)."""
    return np.ascontiguousarray(np.transpose(models[:, :12, :], (1, 2, 0)))


def cart_to_sph(jc, bx, by, bz):
    """d/d(Bx,By,Bz) -> d/d(|B|, gamma, phi) by the closed-form chain rule."""
    B = np.sqrt(bx * bx + by * by + bz * bz)
    g = np.arccos(bz / B)
    f = np.arctan2(by, bx)
    dx, dy, dz = jc
    dB = (bx * dx + by * dy + bz * dz) / B
    dg = B * np.cos(g) * np.cos(f) * dx + B * np.cos(g) * np.sin(f) * dy - B * np.sin(g) * dz
    df = -B * np.sin(g) * np.sin(f) * dx + B * np.sin(g) * np.cos(f) * dy
    return dB, dg, df


def report(name, err, tol):
    print("[parity] %-40s max err %.3e (tol %.1e)" % (name, err, tol), flush=True)
    assert err <= tol, name


def test_gain_hsra(gpu_ctx, golden):
    g = golden["syn1"]
    ls = lines.setup([((4, 5), 500, -700.0, 5.0)])
    gpu_ctx.configure(ls, g["model_in"][12].astype(np.float64))
    gain = gpu_ctx.gain_hsra()
    report("GAIN1D vs reference (relative)", np.max(np.abs(gain - g["gain"]) / np.abs(g["gain"])), 1e-6)


def test_syn1_stokes_vs_reference(gpu_ctx, golden):
    g = golden["syn1"]
    ls = lines.setup([((4, 5), 500, -700.0, 5.0)])
    gpu_ctx.configure(ls, g["model_in"][12].astype(np.float64))
    gpu_ctx.set_gain(g["gain"])
    atm = model_rows(g["model_in"][None])
    atm_out, st, _, status = run_gpu(gpu_ctx, atm, 0)
    assert status[0] == 0
    report("test_syn1 Stokes vs reference", np.max(np.abs(st[0] - g["stokes"])), STOKES_TOL)
    # rho, Pe, mu and log tau5 written back like the reference updates the model
    for row, name in ((2, "rho"), (3, "Pe"), (4, "mu")):
        report("test_syn1 model %s (relative)" % name,
               np.max(np.abs(atm_out[row, :, 0] - g["model_out"][row]) / np.abs(g["model_out"][row])), 1e-6)
    report("test_syn1 log tau5", np.max(np.abs(atm_out[11, :, 0] - g["model_out"][11])), 1e-5)


def test_rf_cartesian_vs_reference(gpu_ctx, golden):
    g = golden["rf"]
    ls = lines.setup([((4, 5), 500, -700.0, 5.0)])
    gpu_ctx.configure(ls, g["model_in"][12].astype(np.float64))
    gpu_ctx.set_gain(g["gain"])
    _, st, jac, status = run_gpu(gpu_ctx, model_rows(g["model_in"][None]), 2)
    assert status[0] == 0
    report("test_rf Stokes vs reference", np.max(np.abs(st[0] - g["stokes"])), STOKES_TOL)
    # device [np][nz][4][numw] vs fixture [nz_sub][numw][8][4]
    names = ("T", "Pg", "rho", "Bx", "By", "Bz", "vz")
    for p in range(7):
        ours = jac[0, p][g["rf_heights"]]                  # (nzs, 4, numw)
        ref = np.transpose(g["rf"][:, :, p, :], (0, 2, 1))  # (nzs, 4, numw)
        for s in range(4):
            sc = np.abs(ref).max()
            if sc == 0:
                continue
            report("test_rf RF d/d%s Stokes %d" % (names[p], s), np.max(np.abs(ours[:, s] - ref[:, s])) / sc, RF_TOL)


def test_seeded_patch_vs_reference(gpu_ctx, golden):
    g = golden["seeded_c1"]
    gpu_ctx.configure(lines.fe6301_6302(), g["z"])
    gpu_ctx.set_gain(g["gain"])
    atm = model_rows(g["models"])
    _, st, jac, status = run_gpu(gpu_ctx, atm, 2)
    assert np.all(status == 0)
    report("seeded C1 Stokes vs reference (16 cols)", np.max(np.abs(st - g["stokes"])), STOKES_TOL)
    for i in range(len(g["rf"])):
        for p in range(7):
            ref = np.transpose(g["rf"][i][:, :, p, :], (0, 2, 1))
            for s in range(4):
                sc = np.abs(ref).max()
                if sc > 0:
                    report("seeded col %d RF p%d s%d" % (i, p, s), np.max(np.abs(jac[i, p][:, s] - ref[:, s])) / sc,
                           RF_TOL)


def test_hydrostatic_vs_oracle(gpu_ctx, cpu):
    import torch
    nx = ny = 4
    atm, z = atmos.seeded(nx, ny, 64, seed=3)
    gpu_ctx.configure(lines.fe6301_6302(), z)
    t = torch.from_numpy(atm.copy()).cuda()
    status = gpu_ctx.hydrostatic_torch(t)
    torch.cuda.synchronize()
    out = t.cpu().numpy()
    assert np.all(status.cpu().numpy() == 0)
    worst = 0.0
    for col in range(nx * ny):
        pg, rho, pel, mw, _ = cpu.he(atm[0, :, col], float(atm[1, -1, col]), z)
        worst = max(worst, np.max(np.abs(out[1, :, col] - pg) / pg), np.max(np.abs(out[2, :, col] - rho) / rho))
    report("hydrostatic Pg/rho vs oracle (relative)", worst, 1e-6)


def test_seeded_cube_spherical_vs_oracle(gpu_ctx, cpu):
    """16x16x64 seeded cube: Stokes for all columns, spherical Jacobians for a
    sample, against the oracle's Cartesian RFs and the closed-form chain rule."""
    import torch
    ls = lines.fe6301_6302()
    atm, z = atmos.seeded(16, 16, 64, seed=5)
    gpu_ctx.configure(ls, z)
    gain = gpu_ctx.gain_hsra()
    t = torch.from_numpy(atm.copy()).cuda()
    assert np.all(gpu_ctx.hydrostatic_torch(t).cpu().numpy() == 0)
    st, jac, status = gpu_ctx.forward_torch(t, 1)
    torch.cuda.synchronize()
    st, jac, status, atm_h = st.cpu().numpy(), jac.cpu().numpy(), status.cpu().numpy(), t.cpu().numpy()
    assert np.all(status == 0)
    cpu.setup(ls, resp=True, atmpar=CART)
    cpu.set_gain(gain)
    models = np.stack([atmos.column_model(atm_h, z, c) for c in range(256)])
    _, syn, _ = cpu.forward_batch(models)
    report("seeded 16x16x64 Stokes vs oracle", np.max(np.abs(st - np.transpose(syn, (0, 2, 1)))), STOKES_TOL)
    for c in (0, 37, 101, 255):
        _, s1, d1, _ = cpu.forward1d(models[c])
        dT = np.transpose(d1[:, :, 0, :], (0, 2, 1))
        dv = np.transpose(d1[:, :, 6, :], (0, 2, 1))
        jc = [np.transpose(d1[:, :, p, :], (0, 2, 1)) for p in (3, 4, 5)]
        bx, by, bz = (float(models[c][i, 0]) for i in (5, 6, 7))
        dB, dg, df = cart_to_sph(jc, bx, by, bz)
        for p, ref in ((0, dT), (1, dB), (2, dg), (3, df), (4, dv)):
            for s in range(4):
                sc = np.abs(ref).max()
                if sc > 0:
                    report("cube col %d spherical p%d s%d" % (c, p, s),
                           np.max(np.abs(jac[c, p][:, s] - ref[:, s])) / sc, RF_TOL)


def test_forward3d_host_mirror(gpu_ctx, golden):
    """FORWARD3D on Fortran-order host cubes with an AM_I_DONE mask."""
    from firtez_dz_amd import gpu
    g = golden["seeded_c1"]
    gpu_ctx.configure(lines.fe6301_6302(), g["z"])
    gpu_ctx.set_gain(g["gain"])
    nz = len(g["z"])
    nx, ny = 4, 4
    cube = {}
    names = gpu.ATM
    for i, nm in enumerate(names):
        c = np.zeros((nz, ny, nx), np.float32, order="F")
        for k, col in enumerate(g["cols"]):
            x, y = divmod(int(col), 16)
            c[:, y, x] = g["models"][k][i]
        cube[nm] = c
    done = np.zeros((ny, nx), np.int32, order="F")
    done[1, 2] = 1
    syn, dsyn = gpu_ctx.forward3d(cube, am_i_done=done, rf_mode=gpu.RF_CARTESIAN)
    numw = gpu_ctx.numw
    worst = 0.0
    for k, col in enumerate(g["cols"]):
        x, y = divmod(int(col), 16)
        if done[y, x]:
            assert np.all(syn[:, y, x] == 0)
            continue
        ours = syn[:, y, x].reshape(4, numw)
        worst = max(worst, np.max(np.abs(ours - g["stokes"][k])))
    report("FORWARD3D host mirror Stokes vs reference", worst, STOKES_TOL)
    assert dsyn.shape == (4 * numw, 7 * nz, ny, nx)


def test_hydrostatic_mode_vs_reference(gpu_ctx, golden):
    """HYDROSTATIC: yes against the reference's own output (tests/test_he,
    test_syn2 set-up with a field and a flow): the re-integrated model, Stokes
    and all eight response functions of DSYN1D, including the hydrostatic T
    term, the outgoing-vector quirk of the rho, B and v responses and the P0
    (top pressure) response."""
    g = golden["he"]
    ls = lines.setup([((4, 5), int(g["nwave"]), float(g["wave_start"]), float(g["wave_step"]))])
    gpu_ctx.configure(ls, g["model_in"][12].astype(np.float64))
    gpu_ctx.set_gain(g["gain"])
    gpu_ctx.set_hydrostatic(True)
    try:
        atm_out, st, jac, status = run_gpu(gpu_ctx, model_rows(g["model_in"][None]), 3)
    finally:
        gpu_ctx.set_hydrostatic(False)
    assert status[0] == 0
    assert jac.shape[1] == 8
    for row, name in ((1, "Pg"), (2, "rho"), (3, "Pe"), (4, "mu")):
        report("HE model %s (relative)" % name,
               np.max(np.abs(atm_out[row, :, 0] - g["model_out"][row]) / np.abs(g["model_out"][row])), 1e-6)
    report("HE Stokes vs reference", np.max(np.abs(st[0] - g["stokes"])), STOKES_TOL)
    names = ("T", "Pg", "rho", "Bx", "By", "Bz", "vz", "P0")
    for p in range(8):
        ours = jac[0, p][g["rf_heights"]]
        ref = np.transpose(g["rf"][:, :, p, :], (0, 2, 1))
        sc = np.abs(ref).max()
        for s in range(4):
            report("HE RF d/d%s Stokes %d" % (names[p], s), np.max(np.abs(ours[:, s] - ref[:, s])) / sc, RF_TOL)


def test_hydrostatic_mode_seeded_vs_oracle(gpu_ctx, cpu):
    """HYDROSTATIC: yes on seeded columns (2 lines, 112 wavelengths), Stokes
    and Cartesian response functions against the oracle in the same mode."""
    import torch
    ls = lines.fe6301_6302()
    atm, z = atmos.seeded(4, 4, 64, seed=9)
    gpu_ctx.configure(ls, z)
    gain = gpu_ctx.gain_hsra()
    gpu_ctx.set_hydrostatic(True)
    try:
        t = torch.from_numpy(atm.copy()).cuda()
        st, jac, status = gpu_ctx.forward_torch(t, 2)
        torch.cuda.synchronize()
    finally:
        gpu_ctx.set_hydrostatic(False)
    st, jac, status, atm_h = st.cpu().numpy(), jac.cpu().numpy(), status.cpu().numpy(), t.cpu().numpy()
    assert np.all(status == 0)
    cpu.setup(ls, resp=True, atmpar=CART)
    cpu.set_gain(gain)
    cpu.set_hydrostatic(True)
    try:
        worst = 0.0
        for c in (0, 5, 15):
            m = atmos.column_model(atm, z, c)       # Pg only at the top, as the GPU got it
            mo, syn, dsyn, _ = cpu.forward1d(m)
            report("HE seeded col %d Pg vs oracle (relative)" % c,
                   np.max(np.abs(atm_h[1, :, c] - mo[1]) / mo[1]), 1e-6)
            worst = max(worst, np.max(np.abs(st[c] - syn)))
            for p in range(7):
                ref = np.transpose(dsyn[:, :, p, :], (0, 2, 1))
                sc = np.abs(ref).max()
                report("HE seeded col %d RF p%d" % (c, p), np.max(np.abs(jac[c, p] - ref)) / sc, RF_TOL)
        report("HE seeded Stokes vs oracle", worst, STOKES_TOL)
    finally:
        cpu.set_hydrostatic(False)


def test_lsf_forward_vs_oracle(gpu_ctx, cpu):
    """MLSF on: two regions (odd and even pixel counts, the second with an
    off-centre Gaussian), Stokes and Cartesian response functions against the
    oracle with the same LSF_KERNEL; then the second region switched off
    (LSF_VALID = 0) must leave its pixels as without the LSF."""
    import torch
    ls = lines.setup([((4,), 61, -600.0, 20.0), ((5,), 52, -520.0, 20.0)])
    atm, z = atmos.seeded(3, 3, 64, seed=13)
    gpu_ctx.configure(ls, z)
    gain = gpu_ctx.gain_hsra()
    t = torch.from_numpy(atm.copy()).cuda()
    assert np.all(gpu_ctx.hydrostatic_torch(t).cpu().numpy() == 0)
    atm = t.cpu().numpy()
    kern, valid = lines.lsf_kernel(ls, [45.0, 35.0], [0.0, -15.0])
    cpu.setup(ls, resp=True, atmpar=CART)
    cpu.set_gain(gain)
    try:
        gpu_ctx.set_lsf(kern, valid)
        cpu.set_lsf(kern, valid)
        _, st, jac, status = run_gpu(gpu_ctx, atm, 2)
        assert np.all(status == 0)
        worst = 0.0
        for c in (0, 4, 8):
            _, syn, dsyn, _ = cpu.forward1d(atmos.column_model(atm, z, c))
            worst = max(worst, np.max(np.abs(st[c] - syn)))
            for p in range(7):
                ref = np.transpose(dsyn[:, :, p, :], (0, 2, 1))
                report("LSF col %d RF p%d" % (c, p), np.max(np.abs(jac[c, p] - ref)) / np.abs(ref).max(), RF_TOL)
        report("LSF Stokes vs oracle", worst, STOKES_TOL)
        gpu_ctx.set_lsf(kern, [1, 0])
        _, st1, _, _ = run_gpu(gpu_ctx, atm, 0)
        gpu_ctx.set_lsf(None)
        _, st0, _, _ = run_gpu(gpu_ctx, atm, 0)
        report("LSF_VALID=0 region unchanged", np.max(np.abs(st1[:, :, 61:] - st0[:, :, 61:])), 0.0)
        assert np.max(np.abs(st1[:, 0, :61] - st0[:, 0, :61])) > 1e-3
    finally:
        gpu_ctx.set_lsf(None)
        cpu.set_lsf(None)


def test_eos_outputs_vs_oracle(gpu_ctx, cpu):
    """The forward's equation-of-state outputs written back to the model (rho,
    Pe, mu: the electron-density fixed point with its warm start from the
    height below) against the oracle on seeded columns.  The fixed point runs
    in the reference's operation order, so agreement is at FP64 rounding."""
    import torch
    ls = lines.fe6301_6302()
    atm, z = atmos.seeded(6, 5, 64, seed=17)
    gpu_ctx.configure(ls, z)
    gain = gpu_ctx.gain_hsra()
    t = torch.from_numpy(atm.copy()).cuda()
    assert np.all(gpu_ctx.hydrostatic_torch(t).cpu().numpy() == 0)
    atm_in = t.cpu().numpy()
    atm_out, _, _, status = run_gpu(gpu_ctx, atm_in, 0)
    assert np.all(status == 0)
    cpu.setup(ls, resp=False, atmpar=CART)
    cpu.set_gain(gain)
    worst = 0.0
    for c in range(atm.shape[2]):
        mo, _, _, _ = cpu.forward1d(atmos.column_model(atm_in, z, c))
        for row in (2, 3, 4):
            worst = max(worst, float(np.max(np.abs(atm_out[row, :, c] - mo[row]) / np.abs(mo[row]))))
    report("EOS rho/Pe/mu vs oracle (relative, 30 columns)", worst, 1e-12)


def test_p0_response_vs_oracle(gpu_ctx, cpu):
    """P0 (slot 8 of DSYN1D, forward.f90:543-545) without HYDROSTATIC on
    seeded columns: the propagated unit response times PG(K)/PG(NZ) and the
    propagated Pg|T response, the top layer 1, against the oracle."""
    import torch
    ls = lines.fe6301_6302()
    atm, z = atmos.seeded(3, 3, 64, seed=23)
    gpu_ctx.configure(ls, z)
    gain = gpu_ctx.gain_hsra()
    t = torch.from_numpy(atm.copy()).cuda()
    assert np.all(gpu_ctx.hydrostatic_torch(t).cpu().numpy() == 0)
    atm_in = t.cpu().numpy()
    _, st, jac, status = run_gpu(gpu_ctx, atm_in, 3)
    assert np.all(status == 0) and jac.shape[1] == 8
    cpu.setup(ls, resp=True, atmpar=(1, 1, 1, 1, 1, 1, 1, 1))
    cpu.set_gain(gain)
    for c in (0, 4, 8):
        _, syn, dsyn, _ = cpu.forward1d(atmos.column_model(atm_in, z, c))
        for p in (1, 7):
            ref = np.transpose(dsyn[:, :, p, :], (0, 2, 1))
            report("P0 mode col %d RF p%d" % (c, p), np.max(np.abs(jac[c, p] - ref)) / np.abs(ref).max(), RF_TOL)
        report("P0 mode col %d Stokes" % c, np.max(np.abs(st[c] - syn)), STOKES_TOL)


@pytest.mark.parametrize("hydro", [False, True])
def test_tiled_forward_identical(gpu_ctx, hydro):
    """Six radiative-transfer tiles (the last one ragged) against one: each
    column's arithmetic does not depend on the tiling, so Stokes, Jacobians,
    status and the model outputs (Pg under HYDROSTATIC) must be bit-identical
    to the single-tile forward, also with profiling on."""
    import torch
    ls = lines.fe6301_6302()
    atm, z = atmos.seeded(16, 16, 64, seed=11)
    gpu_ctx.configure(ls, z)
    gpu_ctx.gain_hsra()
    t0 = torch.from_numpy(atm.copy()).cuda()
    assert np.all(gpu_ctx.hydrostatic_torch(t0).cpu().numpy() == 0)
    gpu_ctx.set_hydrostatic(hydro)
    try:
        outs = []
        for tile, prof in ((8192, False), (48, False), (48, True)):
            gpu_ctx.set_tile(tile, 256)
            gpu_ctx.set_profiling(prof)
            t = t0.clone()
            st, jac, status = gpu_ctx.forward_torch(t, 1)
            torch.cuda.synchronize()
            if prof:
                micro_ms, rt_ms, launches = gpu_ctx.last_timing()
                assert launches == 6 and micro_ms > 0 and rt_ms > 0
            outs.append([x.cpu().numpy() for x in (st, jac, status, t)])
    finally:
        gpu_ctx.set_tile(8192, 256)
        gpu_ctx.set_profiling(False)
        gpu_ctx.set_hydrostatic(False)
    assert np.all(outs[0][2] == 0)
    for o in outs[1:]:
        for a, b in zip(outs[0], o):
            assert np.array_equal(a, b)
    print("[parity] tiled forward (6 tiles, HYDROSTATIC %s): bit-identical" % hydro, flush=True)
