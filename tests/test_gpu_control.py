"""The reference's own test control files run through the device path
(firtez_dz_amd/control.py) in the reference's directory layout
(tests/test_*/control_file.dat, default/*.bin, relative paths unchanged),
their output files read back with the reference's formats and compared with
the reference's outputs held in tests/golden (syn1.npz, rf.npz, he.npz,
written by the reference's Fortran)."""
import os
import shutil

import numpy as np
import pytest

from firtez_dz_amd import control, modelio

pytestmark = pytest.mark.gpu
HERE = os.path.dirname(os.path.abspath(__file__))
GOLD = os.path.join(HERE, "golden")
STOKES_TOL = 1e-4
RF_TOL = 2e-3


def report(name, err, tol):
    print("[control] %-44s max err %.3e (tol %.1e)" % (name, err, tol), flush=True)
    assert err <= tol, name


@pytest.fixture
def tree(tmp_path):
    (tmp_path / "default").mkdir()
    for f in ("test_atmos.bin", "rf_test_atmos.bin"):
        shutil.copy(os.path.join(GOLD, f), tmp_path / "default" / f)
    for t in ("test_syn1", "test_syn2", "test_rf", "test_he", "test_tau1", "test_tau2", "test_inv1"):
        d = tmp_path / "tests" / t
        (d / "out_dir").mkdir(parents=True)
        shutil.copy(os.path.join(GOLD, "control", t + ".dat"), d / "control_file.dat")
    return tmp_path


def run(tree, name, ctx, edit=None):
    """Runs tests/<name>/control_file.dat; returns (control, result, output
    directory): OUTPATH relative to the test directory, the directory itself
    when the file sets none."""
    wd = str(tree / "tests" / name)
    c = control.parse_file(os.path.join(wd, "control_file.dat"))
    if edit:
        edit(c)
    print("[control] running %s (%s)" % (name, c.mode), flush=True)
    return c, control.run(c, wd, ctx), os.path.join(wd, c.outpath)


def model_rel(cube, ref, rows):
    names = {1: "Pg", 2: "rho", 3: "Pe", 4: "mu"}
    return max(float(np.max(np.abs(cube[names[r]][:, 0, 0] - ref[r]) / np.abs(ref[r]))) for r in rows)


def test_syn1_and_tau1(gpu_ctx, golden, tree):
    g = golden["syn1"]
    _, out, wd = run(tree, "test_syn1", gpu_ctx)
    st, wave, index = modelio.read_profiles(os.path.join(wd, "syn_profile.bin"))
    assert st.shape == (4, 500, 1, 1)
    assert wave[0] == -700.0 and np.allclose(index, 6301.5010, atol=1e-3)
    report("test_syn1 profiles vs reference", float(np.max(np.abs(st[:, :, 0, 0] - g["stokes"]))), STOKES_TOL)
    m = modelio.read_model(os.path.join(wd, "out_test_atmos.bin"))
    report("test_syn1 out model rho/Pe/mu (relative)", model_rel(m, g["model_out"], (2, 3, 4)), 1e-6)
    report("test_syn1 out model log tau5", float(np.max(np.abs(m["tau5"][:, 0, 0] - g["model_out"][11]))), 1e-5)
    # MODE tau: the same model outputs without lines
    _, _, wd = run(tree, "test_tau1", gpu_ctx)
    m = modelio.read_model(os.path.join(wd, "out_test_atmos.bin"))
    report("test_tau1 out model rho/Pe/mu (relative)", model_rel(m, g["model_out"], (2, 3, 4)), 1e-6)
    report("test_tau1 out model log tau5", float(np.max(np.abs(m["tau5"][:, 0, 0] - g["model_out"][11]))), 1e-5)


def test_rf(gpu_ctx, golden, tree):
    g = golden["rf"]
    _, out, wd = run(tree, "test_rf", gpu_ctx)
    st, _, _ = modelio.read_profiles(os.path.join(wd, "syn_profile.bin"))
    report("test_rf profiles vs reference", float(np.max(np.abs(st[:, :, 0, 0] - g["stokes"]))), STOKES_TOL)
    d, z, wave, logv = modelio.read_rfs(os.path.join(wd, "rfout_rf_test_atmos.bin"))
    assert list(logv) == [True, False, False, False, False, True, False, False]
    assert d.shape == (8, 4, 500, 256, 1, 1)
    for p, name in ((0, "T"), (5, "Bz")):
        ours = d[p][:, :, g["rf_heights"], 0, 0]                 # (4, nw, nzs)
        ref = np.transpose(g["rf"][:, :, p, :], (2, 1, 0))       # (4, nw, nzs)
        report("test_rf rfout d/d%s vs reference" % name, float(np.max(np.abs(ours - ref)) / np.abs(ref).max()),
               RF_TOL)
    assert not np.any(d[[1, 2, 3, 4, 6, 7]])


def test_hydrostatic_modes(gpu_ctx, golden, tree):
    g = golden["he"]
    for name in ("test_he", "test_tau2", "test_syn2"):
        _, _, wd = run(tree, name, gpu_ctx)
        m = modelio.read_model(os.path.join(wd, "out_test_atmos.bin"))
        report("%s out model Pg/rho/Pe/mu (relative)" % name, model_rel(m, g["model_out"], (1, 2, 3, 4)), 1e-6)
        if name != "test_he":
            report("%s out model log tau5" % name, float(np.max(np.abs(m["tau5"][:, 0, 0] - g["model_out"][11]))),
                   1e-5)
    st, _, _ = modelio.read_profiles(str(tree / "tests" / "test_syn2" / "out_dir" / "syn_profile.bin"))
    assert st.shape == (4, 500, 1, 1) and np.all(np.isfinite(st))


def test_inversion_after_synthesis(gpu_ctx, tree):
    """test_inv1 unchanged: it reads test_syn1's output and inverts T and
    v_z with HYDROSTATIC (Pg re-integrated from the top), vertical
    regularisation, the CG solver and PERTURBATION_TYPE 2, its nodes growing
    to NZ = 256 per parameter (512 free parameters: the LM step with its work
    arrays in device memory).  The start model is the one that made the
    observations, but the hydrostatic integration replaces its Pg, so the
    start misfit is that of the hydrostatic synthesis of the same model; the
    inversion must lower it."""
    run(tree, "test_syn1", gpu_ctx)

    def hydro_start(c):
        c.hydrostatic = True
        c.fileprofile = "he_start.bin"

    run(tree, "test_syn1", gpu_ctx, hydro_start)
    out_syn1 = tree / "tests" / "test_syn1" / "out_dir"
    obs, _, _ = modelio.read_profiles(str(out_syn1 / "syn_profile.bin"))
    start, _, _ = modelio.read_profiles(str(out_syn1 / "he_start.bin"))
    c, out, wd = run(tree, "test_inv1", gpu_ctx)
    st, _, _ = modelio.read_profiles(os.path.join(wd, "out_syn_profile.bin"))
    chi = float(np.asarray(out["chi2"]).reshape(-1)[0])
    e0 = float(np.sqrt(np.mean((start - obs) ** 2)))
    e1 = float(np.sqrt(np.mean((st - obs) ** 2)))
    print("[control] test_inv1: %d LM iterations, chi2 %.4g; rms misfit start %.3e -> %.3e"
          % (out["iterations"], chi, e0, e1), flush=True)
    assert np.isfinite(chi) and e1 < 0.5 * e0
    m = modelio.read_model(os.path.join(wd, "out_test_atmos.bin"))
    assert np.all(np.isfinite(m["T"])) and np.all(m["T"] >= 2500.0)


def test_inversion_auto_weights(gpu_ctx, tree):
    """test_inv1 with STOKES SETUP AUTO: every pixel weighted by the spread
    of its own observed Stokes blocks (SET_WEIGHTS3D); the inversion must
    still lower the misfit."""
    run(tree, "test_syn1", gpu_ctx)

    def hydro_start(c):
        c.hydrostatic = True
        c.fileprofile = "he_start.bin"

    run(tree, "test_syn1", gpu_ctx, hydro_start)
    out_syn1 = tree / "tests" / "test_syn1" / "out_dir"
    obs, _, _ = modelio.read_profiles(str(out_syn1 / "syn_profile.bin"))
    start, _, _ = modelio.read_profiles(str(out_syn1 / "he_start.bin"))

    def auto(c):
        c.auto_weights = True

    c, out, wd = run(tree, "test_inv1", gpu_ctx, auto)
    st, _, _ = modelio.read_profiles(os.path.join(wd, "out_syn_profile.bin"))
    chi = float(np.asarray(out["chi2"]).reshape(-1)[0])
    e0 = float(np.sqrt(np.mean((start - obs) ** 2)))
    e1 = float(np.sqrt(np.mean((st - obs) ** 2)))
    print("[control] test_inv1 AUTO: %d LM iterations, chi2 %.4g; rms misfit start %.3e -> %.3e"
          % (out["iterations"], chi, e0, e1), flush=True)
    assert np.isfinite(chi) and e1 < 0.5 * e0


def test_syn1_without_continuum_normalisation(gpu_ctx, tree):
    """test_syn1 with CONTINUUM NORMALIZATION: F (GET_NORM without HSRA,
    normalization.f90:746-751: GAIN = 1, V with the opposite sign): the
    written profiles are the normalised ones divided by the HSRA gain."""
    _, _, wd = run(tree, "test_syn1", gpu_ctx)
    st_t, _, _ = modelio.read_profiles(os.path.join(wd, "syn_profile.bin"))
    gain = gpu_ctx.gain_hsra()                                  # (4, numw) of the same lines

    def no_norm(c):
        c.hsra_norm = False
        c.fileprofile = "raw_profile.bin"

    run(tree, "test_syn1", gpu_ctx, no_norm)
    st_f, _, _ = modelio.read_profiles(os.path.join(wd, "raw_profile.bin"))
    sign = np.array([1.0, 1.0, 1.0, -1.0])[:, None, None, None]
    ref = st_t / (gain[:, :, None, None] * sign)
    err = float(np.max(np.abs(st_f - ref)) / np.max(np.abs(ref)))
    report("test_syn1, CONTINUUM NORMALIZATION: F", err, 1e-6)


def test_syn1_full_stokes(gpu_ctx, tree):
    """test_syn1 with MISC SETUP FULL_STOKES: fs_<FILEPROFILE> (WRITE_BIN,
    SYN5D(4, NUMW, NZ, NY, NX)) holds the Stokes vector after every height;
    its top height is the written profile."""
    def full(c):
        c.full_stokes = True
        c.fileprofile = "fs_run.bin"

    c, out, wd = run(tree, "test_syn1", gpu_ctx, full)
    st, _, _ = modelio.read_profiles(os.path.join(wd, "fs_run.bin"))
    fs = modelio.read_bin(os.path.join(wd, "fs_fs_run.bin"))
    assert fs.shape == (4, st.shape[1], fs.shape[2], 1, 1) and fs.shape[2] > 1
    rows = [s for s in range(4) if c.invstk[s]]      # the profile file holds the set-up Stokes rows only
    report("test_syn1 FULL_STOKES top height vs profile", float(np.max(np.abs(fs[rows, :, -1] - st[rows]))), 0.0)


def test_syn1_line_optical_depth(gpu_ctx, tree):
    """test_syn1 with LINE OPTICAL DEPTH: y: tl_<FILEMODEL> (WRITE_LOGTAUS)
    holds log10 tau of every wavelength and height, -9 at the top, growing
    downward, identical to fz_line_tau_dev on the written model."""
    import torch

    def tl(c):
        c.line_tau = True

    c, out, wd = run(tree, "test_syn1", gpu_ctx, tl)
    taus, z, ind, wave = modelio.read_logtaus(os.path.join(wd, "tl_" + c.filemodel))
    nw, nz = taus.shape[:2]
    assert taus.shape[2:] == (1, 1) and nz == len(z) and nw == len(wave)
    assert np.all(taus[:, -1] == np.float32(np.log10(np.float32(1e-9))))
    assert np.all(np.diff(taus[:, :, 0, 0], axis=1) <= 0)
    m = modelio.read_model(os.path.join(wd, "out_" + c.filemodel))
    atm = torch.from_numpy(modelio.to_device_rows(m)).cuda()
    t2, st = gpu_ctx.line_tau_torch(atm)
    assert int(st.max()) == 0
    report("test_syn1 LINE OPTICAL DEPTH vs fz_line_tau_dev", float(np.max(np.abs(taus[:, :, 0, 0] -
                                                                                t2.cpu().numpy()[0].T))), 0.0)


def test_syn1_with_line_spread_function(gpu_ctx, golden, tree):
    """test_syn1 with a LINE SPREAD FUNCTION: section (a Gaussian of sigma
    30 mA centred on 0, NSET_LSF / LSF_INIT): the written profiles equal the
    profiles without it circularly convolved with the LSF taps, the FFT
    product of forward.f90:810-844 written with numpy's DFT."""
    from firtez_dz_amd import lines
    wd = tree / "tests" / "test_syn1"
    _, _, out = run(tree, "test_syn1", gpu_ctx)
    st0, _, _ = modelio.read_profiles(os.path.join(out, "syn_profile.bin"))
    txt = (wd / "control_file.dat").read_text().replace("END:", "LINE SPREAD FUNCTION:\n30.0 0.0\nEND:")
    (wd / "control_file.dat").write_text(txt)
    c, _, out = run(tree, "test_syn1", gpu_ctx)
    assert c.lsf == [(30.0, 0.0)]
    st1, _, _ = modelio.read_profiles(os.path.join(out, "syn_profile.bin"))
    kern, valid = lines.lsf_kernel(control.line_setup(c), [30.0], [0.0])
    assert valid[0] == 1
    n = kern.size
    taps = kern * n
    ref = np.stack([np.fft.irfft(np.fft.rfft(st0[s, :, 0, 0]) * np.fft.rfft(taps), n) for s in range(4)])
    report("test_syn1 + LSF vs convolved profiles", float(np.max(np.abs(st1[:, :, 0, 0] - ref))), 1e-5)
    assert float(np.max(np.abs(st1 - st0))) > 1e-3          # the LSF changed the profiles
    gpu_ctx.set_lsf(None)


def test_coupled_inversion_control_file(gpu_ctx, tmp_path):
    """MODE inversion with "COUPLED INVERSION:" yes / FILENAME: the runner
    reads the PSF file (READ_PSF_2D's format under DATAPATH), inverts the
    PSF-degraded profiles through it and writes the reference's outputs; they
    equal the library's coupled inversion called directly on the same files,
    bit for bit, and the context is left without a PSF."""
    import torch
    from firtez_dz_amd import atmos, gpu
    from oracle import psf as P
    nx, ny, nz = 4, 3, 48
    text = ("LINES:\n4,5 24 -240.0 20.0\nFILEPROFILE:\nobs.bin\nFILEMODEL:\nstart.bin\nMODE:\ninversion\n"
            "STOKES SETUP:\nSTKI 1\nSTKQ 1\nSTKU 1\nSTKV 1\n"
            "INVERSION SETUP:\nTEM 1\nBX 1\nBY 1\nBZ 1\nVLOS 1\n"
            "MISC SETUP:\nDATAPATH ./\nMODPATH ./\nOUTPATH out/\nNOISE 1e-3\n"
            "COUPLED INVERSION:\nyes\nFILENAME psf.bin\n"
            "BOX:\n4 16.0\n3 16.0\n48 15.0\n")
    (tmp_path / "control_file.dat").write_text(text)
    (tmp_path / "out").mkdir()
    c = control.parse(text)
    ls = control.line_setup(c)
    atm, z = atmos.seeded(nx, ny, nz, seed=41)
    gpu_ctx.configure(ls, z)
    gpu_ctx.gain_hsra()
    t = torch.from_numpy(atm.copy()).cuda()
    assert int(gpu_ctx.hydrostatic_torch(t).max()) == 0
    a = t.cpu().numpy()
    names = {"T": 0, "Pg": 1, "rho": 2, "Pe": 3, "mu": 4, "Bx": 5, "By": 6, "Bz": 7, "vx": 8, "vy": 9, "vz": 10,
             "tau5": 11}
    cube = {n: np.asfortranarray(a[i].reshape(nz, nx, ny).transpose(0, 2, 1)) for n, i in names.items()}
    syn = gpu_ctx.forward3d(cube)
    syn = syn[0] if isinstance(syn, tuple) else syn                  # SYN3D (4 numw, ny, nx)
    numw = gpu_ctx.numw
    nf = 4 * numw
    psf = np.random.default_rng(2).random((3, 3)) + np.pad([[3.0]], 1)
    psf /= psf.sum()
    modelio.write_psf(str(tmp_path / "psf.bin"), psf)
    psf = modelio.read_psf(str(tmp_path / "psf.bin"))
    maps = np.asarray(syn).reshape(nf, ny * nx, order="F").reshape(nf, nx, ny)
    obs = P.convolve(maps, psf).reshape(nf, nx, ny).transpose(0, 2, 1)             # (nf, ny, nx)
    obs = obs + np.random.default_rng(3).normal(0, 1e-3, obs.shape)
    st = obs.reshape(numw, 4, ny, nx, order="F").transpose(1, 0, 2, 3)
    wave, index = control._wave_index(ls)
    modelio.write_profiles(str(tmp_path / "obs.bin"), st, wave, index)
    start = {n: np.asfortranarray(v.copy()) for n, v in cube.items()}
    start["T"] += np.float32(120.0)
    start["vz"] += np.float32(4e4)
    start["Bx"] *= np.float32(0.8)
    start["z_km"] = np.asarray(z, np.float64)
    modelio.write_model(str(tmp_path / "start.bin"), start)
    print("[control] coupled inversion control file: %dx%dx%d, PSF 3x3" % (nx, ny, nz), flush=True)
    out = control.run(c, str(tmp_path), gpu_ctx)
    with pytest.raises(gpu.FzError):
        gpu_ctx.psf_convolve_torch(torch.zeros((1, nx * ny), device="cuda"), nx, ny)
    # the same inversion called directly on the files
    cube2 = modelio.read_model(str(tmp_path / "start.bin"))
    for n in ("vx", "vy"):
        cube2.setdefault(n, np.zeros_like(cube2["T"], order="F"))
    prof, _, _ = modelio.read_profiles(str(tmp_path / "obs.bin"))
    obs3d = np.asfortranarray(np.concatenate([prof[s] for s in range(4)], axis=0).astype(np.float64))
    gpu_ctx.set_psf(modelio.read_psf(str(tmp_path / "psf.bin")))
    try:
        syn_d, chi_d, its_d = gpu_ctx.invert3d(cube2, obs3d, control.control_for_inversion(c, nz), coupled=True)
    finally:
        gpu_ctx.set_psf(None)
    chi = np.asarray(out["chi2"])
    print("[control] coupled: %d iterations, mean chi2 %.4g" % (out["iterations"], float(chi.mean())), flush=True)
    assert out["iterations"] == its_d and np.array_equal(chi, chi_d)
    m = modelio.read_model(str(tmp_path / "out" / "out_start.bin"))
    for n in ("T", "Bx", "By", "Bz", "vz"):
        assert np.array_equal(m[n], cube2[n]), n
    # cout_: the PSF-convolved synthesis of the inverted model; out_: its own
    # synthesis (DEALLOCATE_AND_WRITE, pre_post_duties.f90:195-245)
    w, _, _ = modelio.read_profiles(str(tmp_path / "out" / "cout_obs.bin"))
    assert np.array_equal(w, _st4(syn_d, numw).astype(np.float32))
    res = gpu_ctx.forward3d(cube2)
    syn_u = res[0] if isinstance(res, tuple) else res
    w, _, _ = modelio.read_profiles(str(tmp_path / "out" / "out_obs.bin"))
    assert np.array_equal(w, _st4(syn_u, numw).astype(np.float32))
    assert float(chi.mean()) < 2.0


def _st4(syn3d, numw):
    ny, nx = syn3d.shape[1:]
    return np.asarray(syn3d).reshape((numw, 4, ny, nx), order="F").transpose(1, 0, 2, 3)
