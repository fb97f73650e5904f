"""Per-pixel spectral windows (fz_set_windows) on the GPU.

  * the ragged forward equals the dense forward restricted to each pixel's
    windows, bit for bit (the same per-wavelength arithmetic, only the
    work-item mapping differs), including single-wavelength windows, windows
    of exactly one and of one-plus-one 64-lane chunks, and unobserved regions;
  * the ragged node-contracted response functions equal the dense ones
    restricted, bit for bit;
  * one LM step on ragged rows against the oracle run per pixel with that
    pixel's own NFREQ (its chi^2 normalisation included);
  * INVERT3D with windows against the oracle's INVERT3D run per pixel on the
    line setup whose regions are that pixel's windows;
  * the uneven-lines workload itself: the eight-line list (Fe I 6301.5,
    6302.5, 6173.3, 15648.5, 15652.9, 15662.0, 5576.1, Si I 10827.1), 1-4
    lines per pixel in windows of 60-300 wavelengths, Stokes and node
    response functions against the oracle per pixel, and the LM step and
    INVERT3D tests above on it.
"""
import numpy as np
import pytest

import oracle
from firtez_dz_amd import atmos, gpu, lines

pytestmark = pytest.mark.gpu
NOISE = 1e-3
CART = (1, 1, 1, 1, 1, 1, 1, 0)
NODES = {"T": 4, "Bx": 4, "By": 4, "Bz": 4, "vz": 4}


def report(name, err, tol):
    print("[windows] %-52s max err %.3e (tol %.1e)" % (name, err, tol), flush=True)
    assert err <= tol, name


def three_regions():
    return lines.setup([((4, 5), 112, -624.0, 20.0), ((2,), 80, -400.0, 10.0), ((14,), 96, -1200.0, 25.0)])


def windows_for(ls, ncol, seed):
    w = lines.uneven_windows(ls, ncol, seed, nlines=(1, 3), width=(1, 112))
    # edge cases first: every region whole; one wavelength; exactly one
    # 64-lane chunk; 65 = a chunk plus one, with a one-point window elsewhere
    w[0] = [(0, 112), (0, 80), (0, 96)]
    w[1] = [(50, 1), (0, 0), (0, 0)]
    w[2] = [(0, 0), (8, 64), (0, 0)]
    w[3] = [(10, 65), (0, 0), (95, 1)]
    return w


def config3_windows(ls, ncol, seed):
    """1-4 of the eight lines per pixel, windows of 60-300 wavelengths; the
    first pixels are given every line between them, so that each line of the
    list is observed whatever the draw."""
    w = lines.uneven_windows(ls, ncol, seed, nlines=(1, 4), width=(60, 300))
    nreg = len(ls.regions)
    for c in range(min(ncol, 2)):
        w[c] = 0
        for r in range(c * nreg // 2, (c + 1) * nreg // 2):
            cnt = 60 + 60 * (r % 5)
            w[c, r] = ((ls.regions[r].nwave - cnt) // 2, cnt)
    return w


LINESETS = {"three": (three_regions, windows_for), "config3": (lines.uneven_list, config3_windows)}


def setup_cube(gpu_ctx, nx, ny, nz, seed, lineset="three"):
    import torch
    assert torch.cuda.is_available()
    ls = LINESETS[lineset][0]()
    atm, z = atmos.seeded(nx, ny, nz, seed=seed)
    gpu_ctx.configure(ls, z)
    gain = gpu_ctx.gain_hsra()
    t = torch.from_numpy(atm).cuda()
    assert int(gpu_ctx.hydrostatic_torch(t).max()) == 0
    return ls, z, gain, t


def test_windows_forward_matches_dense(gpu_ctx):
    import torch
    nx, ny = 4, 4
    ls, z, gain, truth = setup_cube(gpu_ctx, nx, ny, 64, 11)
    ncol = nx * ny
    w = windows_for(ls, ncol, 5)
    a_dense, a_win = truth.clone(), truth.clone()
    st_d, _, status_d = gpu_ctx.forward_torch(a_dense, gpu.RF_NONE)
    off = gpu_ctx.set_windows(w)
    try:
        print("[windows] forward: %d pixels, %d..%d wavelengths (union %d)"
              % (ncol, np.diff(off).min(), np.diff(off).max(), gpu_ctx.numw), flush=True)
        st_w, _, status_w = gpu_ctx.forward_torch(a_win, gpu.RF_NONE)
        torch.cuda.synchronize()
        assert st_w.shape == (4 * int(off[-1]),)
        assert torch.equal(status_w, status_d) and int(status_d.max()) == 0
        assert torch.equal(a_win, a_dense)          # the model updates are per column
        sd, sw = st_d.cpu().numpy(), st_w.cpu().numpy()
        for c in range(ncol):
            idx = lines.window_index(ls, w[c])
            assert off[c + 1] - off[c] == idx.size
            got = sw[4 * off[c]:4 * off[c + 1]].reshape(4, idx.size)
            assert np.array_equal(got, sd[c][:, idx]), c
        with pytest.raises(gpu.FzError):
            gpu_ctx.forward_torch(truth.clone(), gpu.RF_SPHERICAL)
        with pytest.raises(gpu.FzError):
            gpu_ctx.forward_torch(truth[:, :, :ncol - 1].contiguous(), gpu.RF_NONE)
    finally:
        gpu_ctx.set_windows(None)


def test_windows_reject_bad_input(gpu_ctx):
    ls, z, gain, truth = setup_cube(gpu_ctx, 2, 1, 32, 3)
    w = np.zeros((2, 3, 2), np.int32)
    w[0, 0] = (0, 10)
    with pytest.raises(gpu.FzError):
        gpu_ctx.set_windows(w)                         # pixel 1 observes nothing
    w[1, 2] = (90, 7)
    with pytest.raises(gpu.FzError):
        gpu_ctx.set_windows(w)                         # 90 + 7 > 96
    w[1, 2] = (-1, 3)
    with pytest.raises(gpu.FzError):
        gpu_ctx.set_windows(w)
    w[1, 2] = (0, 3)
    off = gpu_ctx.set_windows(w)
    assert list(off) == [0, 10, 13]
    gpu_ctx.configure(ls, z)                           # a new configuration drops the windows
    assert gpu_ctx.win_offsets is None


def test_windows_eqv_matches_dense(gpu_ctx):
    import torch
    nx, ny = 4, 4
    ls, z, gain, truth = setup_cube(gpu_ctx, nx, ny, 64, 12)
    ncol = nx * ny
    w = windows_for(ls, ncol, 6)
    guess = truth.clone()
    guess[0] += 200.0
    guess[10] += 1e5
    ifreep, nfreq = gpu_ctx.set_inversion(gpu.inversion(NODES))
    numw = gpu_ctx.numw
    st_d, eqv_d, status_d = gpu_ctx.forward_nodes_torch(guess.clone())
    off = gpu_ctx.set_windows(w)
    try:
        st_w, eqv_w, status_w = gpu_ctx.forward_nodes_torch(guess.clone())
        torch.cuda.synchronize()
        assert eqv_w.shape == (ifreep * 4 * int(off[-1]),)
        sd, sw, ed, ew = st_d.cpu().numpy(), st_w.cpu().numpy(), eqv_d.cpu().numpy(), eqv_w.cpu().numpy()
        for c in range(ncol):
            idx = lines.window_index(ls, w[c])
            nw = idx.size
            assert np.array_equal(sw[4 * off[c]:4 * off[c + 1]].reshape(4, nw), sd[c][:, idx])
            blk = ew[ifreep * 4 * off[c]:ifreep * 4 * off[c + 1]].reshape(ifreep, 4, nw)
            ref = ed[c].reshape(ifreep, 4, numw)[:, :, idx]
            assert np.array_equal(blk, ref), c
        print("[windows] EQVDSYN ragged == dense restricted (%d pixels, %d free parameters)" % (ncol, ifreep),
              flush=True)
    finally:
        gpu_ctx.set_windows(None)


@pytest.mark.parametrize("lineset", ["three", "config3"])
def test_windows_lm_step_vs_oracle(gpu_ctx, cpu, lineset):
    """Device and oracle fed the same ragged inputs; the oracle runs each
    pixel as its own problem of NFREQ = 4*numw_pixel."""
    import torch
    nx, ny, nz = 4, 2, 64
    ls, z, gain, truth = setup_cube(gpu_ctx, nx, ny, nz, 13, lineset)
    ncol = nx * ny
    w = LINESETS[lineset][1](ls, ncol, 7)
    st_true, _, _ = gpu_ctx.forward_torch(truth.clone(), gpu.RF_NONE)
    rng = np.random.default_rng(3)
    guess = truth.clone()
    guess[0] += torch.from_numpy(250.0 * np.linspace(1.0, 0.2, nz, dtype=np.float32)[:, None]).cuda()
    guess[5:8] *= 0.7
    guess[10] += 1.0e5
    ifreep, nfreq = gpu_ctx.set_inversion(gpu.inversion(NODES))
    atmpar, nd = [0] * 8, [0] * 8
    for k, v in NODES.items():
        i = gpu.INV_PARAMS.index(k)
        atmpar[i], nd[i] = 1, v
    cfg0 = oracle.lm_cfg(nz, gpu_ctx.numw, atmpar, nd)
    coefs = cpu.coefs_block(cfg0, {i: cpu.int_coefs(nz, nd[i]) for i in range(8) if atmpar[i]})
    h_true = st_true.cpu().numpy()
    off = gpu_ctx.set_windows(w)
    try:
        tot = int(off[-1])
        obs = np.zeros(4 * tot)
        for c in range(ncol):
            idx = lines.window_index(ls, w[c])
            obs[4 * off[c]:4 * off[c + 1]] = (h_true[c][:, idx] + rng.normal(0, NOISE, (4, idx.size))).reshape(-1)
        atm = guess.clone()
        st, eqv, status = gpu_ctx.forward_nodes_torch(atm)
        assert int(status.max()) == 0
        d = lambda a, dt=torch.float64: torch.from_numpy(np.ascontiguousarray(a)).to(dt).cuda()
        sla0 = np.tile(np.array([0, 1e29, 1e29, 0, 0, 0, 0, 0, 10.0]), (ncol, 1))
        d_obs, sla, bsyn = d(obs), d(sla0), d(np.zeros(4 * tot))
        isig = d(np.full(nfreq, 1.0 / NOISE ** 2))
        ebest = torch.zeros_like(eqv)
        best = atm.clone()
        done = torch.zeros(ncol, dtype=torch.int32, device=atm.device)
        delta = torch.zeros((ncol, ifreep), dtype=torch.float64, device=atm.device)
        torch.cuda.synchronize()
        h_st, h_eqv, h_cur = st.cpu().numpy().astype(np.float64), eqv.cpu().numpy(), atm.cpu().numpy()
        gpu_ctx.lm_step_torch(d_obs, st, bsyn, sla, eqv, ebest, isig, atm, best, done, delta)
        torch.cuda.synchronize()
        g_sla, g_cur, g_del = sla.cpu().numpy(), atm.cpu().numpy(), delta.cpu().numpy()
        g_eb, g_bs = ebest.cpu().numpy(), bsyn.cpu().numpy()
        werr = dict(chi=0.0, delta=0.0, model=0.0)
        for c in range(ncol):
            p0, p1 = int(off[c]), int(off[c + 1])
            nw = p1 - p0
            cfg = oracle.lm_cfg(nz, nw, atmpar, nd)
            e_c = h_eqv[ifreep * 4 * p0:ifreep * 4 * p1].reshape(ifreep, 4 * nw)
            mc = atmos.column_model(h_cur, z, c)
            mc_o, mb_o, sl, bs, jb, dl = cpu.lm_step(cfg, coefs, obs[4 * p0:4 * p1], h_st[4 * p0:4 * p1],
                                                     np.zeros(4 * nw), sla0[c], e_c.T, np.zeros_like(e_c.T),
                                                     np.full(4 * nw, 1.0 / NOISE ** 2), mc, mc)
            assert g_sla[c, 0] == sl[0] and g_sla[c, 8] == pytest.approx(sl[8], rel=1e-14)
            werr["chi"] = max(werr["chi"], abs(g_sla[c, 1] - sl[1]) / sl[1])
            assert np.array_equal(g_bs[4 * p0:4 * p1], h_st[4 * p0:4 * p1])
            assert np.array_equal(g_eb[ifreep * 4 * p0:ifreep * 4 * p1], h_eqv[ifreep * 4 * p0:ifreep * 4 * p1])
            werr["delta"] = max(werr["delta"], np.max(np.abs(g_del[c] - dl) / (np.abs(dl) + 1e-6 * np.abs(dl).max())))
            for row in (0, 5, 6, 7, 10):
                werr["model"] = max(werr["model"], np.max(np.abs(g_cur[row, :, c] - mc_o[row]) /
                                                          (np.abs(mc_o[row]) + 1e-3)))
        report("%s: ragged LM step chi2 vs oracle per pixel (relative)" % lineset, werr["chi"], 1e-11)
        report("%s: ragged LM step node perturbation vs oracle" % lineset, werr["delta"], 1e-6)
        report("%s: ragged LM step new model vs oracle (relative)" % lineset, werr["model"], 2e-6)
    finally:
        gpu_ctx.set_windows(None)


@pytest.mark.parametrize("lineset", ["three", "config3"])
def test_windows_invert3d_vs_oracle(gpu_ctx, cpu, lineset):
    """INVERT3D (one node per parameter, no smoothing) on 4 pixels of
    different windows, against the oracle's INVERT3D of each pixel alone on
    the line setup made of its windows (same GAIN1D values)."""
    import torch
    from oracle import invert
    nx, ny, nz = 2, 2, 64
    ls, z, gain, truth = setup_cube(gpu_ctx, nx, ny, nz, 21, lineset)
    ncol = nx * ny
    if lineset == "three":
        w = np.zeros((ncol, 3, 2), np.int32)
        w[0] = [(20, 72), (0, 0), (0, 0)]
        w[1] = [(0, 0), (10, 60), (20, 56)]
        w[2] = [(0, 112), (0, 0), (30, 40)]
        w[3] = [(40, 30), (5, 70), (0, 0)]
    else:
        w = config3_windows(ls, ncol, 21)
    true = truth.cpu().numpy()
    rng = np.random.default_rng(21)
    obs_parts, per_pixel = [], []
    for c in range(ncol):
        idx = lines.window_index(ls, w[c])
        lsc = lines.window_setup(ls, w[c])
        cpu.setup(lsc, resp=True, atmpar=CART)
        cpu.set_gain(gain[:, idx])
        mo, syn, _, _ = cpu.forward1d(atmos.column_model(true, z, c))
        o = syn.reshape(-1) + rng.normal(0, NOISE, syn.size)
        obs_parts.append(o)
        per_pixel.append((lsc, idx, o))
    guess = true.copy()
    guess[0] += np.float32(150.0)
    guess[5:8] *= np.float32(0.8)
    guess[10] += np.float32(5e4)
    nodes = {"T": 1, "Bx": 1, "By": 1, "Bz": 1, "vz": 1}
    ctl = gpu.invert_control(nodes, noise=NOISE, smoothing=False)
    off = gpu_ctx.set_windows(w)
    try:
        d_atm = torch.from_numpy(guess.copy()).cuda()
        d_obs = torch.from_numpy(np.concatenate(obs_parts)).cuda()
        st, chi2, its = gpu_ctx.invert3d_torch(d_atm, d_obs, ctl, nx, ny)
        torch.cuda.synchronize()
    finally:
        gpu_ctx.set_windows(None)
    g_model, g_chi2, g_st = d_atm.cpu().numpy(), chi2.cpu().numpy(), st.cpu().numpy()
    idxn = {gpu.INV_PARAMS.index(k): v for k, v in nodes.items()}
    werr = dict(chi=0.0, model=0.0, syn=0.0)
    for c, (lsc, idx, o) in enumerate(per_pixel):
        cpu.setup(lsc, resp=True, atmpar=CART)
        cpu.set_gain(gain[:, idx])
        o_model, o_syn, o_chi2, o_its = invert.invert3d(cpu, guess[:, :, c:c + 1], z, o[None, :], 1, 1, idxn,
                                                       noise=NOISE, smoothing=False)
        print("[windows] pixel %d (%d wavelengths): chi2 device %.4g oracle %.4g (oracle %d iterations)"
              % (c, idx.size, g_chi2[c], o_chi2[0], o_its), flush=True)
        assert g_chi2[c] < 2.0 and o_chi2[0] < 2.0
        werr["chi"] = max(werr["chi"], abs(g_chi2[c] - o_chi2[0]) / o_chi2[0])
        werr["syn"] = max(werr["syn"], float(np.max(np.abs(g_st[4 * off[c]:4 * off[c + 1]] - o_syn[0].reshape(-1)))))
        for row in (0, 5, 6, 7, 10):
            scale = max(1.0, float(np.abs(true[row]).max()))
            werr["model"] = max(werr["model"], float(np.max(np.abs(g_model[row, :, c] - o_model[row, :, 0]))) / scale)
    report("%s: windowed INVERT3D chi2 vs oracle per pixel (relative)" % lineset, werr["chi"], 0.1)
    report("%s: windowed INVERT3D synthesis vs oracle" % lineset, werr["syn"], 5e-3)
    report("%s: windowed INVERT3D model vs oracle (relative to max)" % lineset, werr["model"], 2e-3)


def test_config3_eight_lines_vs_oracle(gpu_ctx, cpu):
    """The eight-line list with 60-300-point windows: ragged Stokes and node
    response functions (T, B, v_z at 4 nodes) of every pixel against the
    oracle run on that pixel's own windows."""
    import torch
    nx, ny, nz = 3, 3, 64
    ls, z, gain, truth = setup_cube(gpu_ctx, nx, ny, nz, 41, "config3")
    ncol = nx * ny
    w = config3_windows(ls, ncol, 41)
    assert np.all((w[:, :, 1] == 0) | ((w[:, :, 1] >= 60) & (w[:, :, 1] <= 300)))
    assert np.all(np.any(w[:, :, 1] > 0, axis=0)), "every line observed"
    guess = truth.clone()
    guess[0] += 150.0
    guess[5:8] *= 0.8
    guess[10] += 5e4
    ifreep, _ = gpu_ctx.set_inversion(gpu.inversion(NODES))
    atmpar, nd = [0] * 8, [0] * 8
    for k, v in NODES.items():
        i = gpu.INV_PARAMS.index(k)
        atmpar[i], nd[i] = 1, v
    off = gpu_ctx.set_windows(w)
    try:
        a1 = guess.clone()
        st, _, status = gpu_ctx.forward_torch(a1, gpu.RF_NONE)
        a2 = guess.clone()
        st2, eqv, status2 = gpu_ctx.forward_nodes_torch(a2)
        torch.cuda.synchronize()
        assert int(status.max()) == 0 and int(status2.max()) == 0
        h_st, h_st2, h_eqv, h_atm = st.cpu().numpy(), st2.cpu().numpy(), eqv.cpu().numpy(), a1.cpu().numpy()
    finally:
        gpu_ctx.set_windows(None)
    assert np.array_equal(h_st, h_st2)
    worst_s = worst_j = 0.0
    for c in range(ncol):
        idx = lines.window_index(ls, w[c])
        nw = idx.size
        lsc = lines.window_setup(ls, w[c])
        cpu.setup(lsc, resp=True, atmpar=CART)
        cpu.set_gain(gain[:, idx])
        _, syn, dsyn, _ = cpu.forward1d(atmos.column_model(h_atm, z, c))
        worst_s = max(worst_s, float(np.max(np.abs(h_st[4 * off[c]:4 * off[c + 1]].reshape(4, nw) - syn))))
        cfg = oracle.lm_cfg(nz, nw, atmpar, nd)
        coefs = cpu.coefs_block(cfg, {i: cpu.int_coefs(nz, nd[i]) for i in range(8) if atmpar[i]})
        ref = cpu.eqv_dsyn(cfg, coefs, dsyn)                       # (4*nw, ifreep)
        blk = h_eqv[ifreep * 4 * off[c]:ifreep * 4 * off[c + 1]].reshape(ifreep, 4 * nw)
        for j in range(ifreep):
            sc = np.abs(ref[:, j]).max()
            if sc > 0:
                worst_j = max(worst_j, float(np.max(np.abs(blk[j] - ref[:, j])) / sc))
        print("[windows] config3 pixel %d: regions %s, %d wavelengths"
              % (c, list(np.nonzero(w[c, :, 1])[0]), nw), flush=True)
    report("config3: eight-line ragged Stokes vs oracle", worst_s, 1e-4)
    report("config3: eight-line ragged EQVDSYN vs oracle", worst_j, 2e-3)
