"""The Fortran side of the boundary on the GPU: tests/fortran/fz_caller.f90,
compiled against the FIRTEZ_GPU module and the library by
firtez_dz_amd.build.build_fortran(), calls FZ_FORWARD3D (spherical response
functions), FZ_INVERT3D, FZ_GET_PERTURBATION (including a rejected DTYPE),
FZ_HESSIAN and, between FZ_INIT_COUPLED_INVERSION and FZ_END_COUPLED_INVERSION,
FZ_COUPLED_INVERT3D (refused once the PSF is gone) on inputs written here;
every output must equal the same entry point called through ctypes on the
same inputs, bit for bit."""
import os
import subprocess

import numpy as np
import pytest

from firtez_dz_amd import atmos, gpu, lines

pytestmark = pytest.mark.gpu
ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
CALLER = os.path.join(ROOT, "tests", "fortran", "fz_caller")


def test_fortran_caller_matches_python(gpu_ctx, tmp_path):
    import torch
    if not os.path.exists(CALLER):
        pytest.fail("tests/fortran/fz_caller is not built (firtez_dz_amd.build.build_fortran)")
    nx, ny, nz = 3, 2, 32
    ls = lines.fe6301_6302(nwave=40, start=-400.0, step=20.0)
    atm, z = atmos.seeded(nx, ny, nz, seed=21)
    gpu_ctx.configure(ls, z)
    gpu_ctx.gain_hsra()
    t = torch.from_numpy(atm.copy()).cuda()
    assert int(gpu_ctx.hydrostatic_torch(t).max()) == 0
    dev = t.cpu().numpy()                                   # [12][nz][col], col = y + ny * x
    cube4 = np.stack([dev[p].T.reshape(nx, ny, nz) for p in range(12)])      # C (12,nx,ny,nz) = F (nz,ny,nx,12)
    inv4 = cube4.copy()
    inv4[0] += 150.0
    inv4[5:8] *= 0.8
    ctl = gpu.invert_control({"T": 2, "Bx": 1, "By": 1, "Bz": 1, "vz": 1}, noise=1e-3, smoothing=False,
                             max_iterations=6)
    rng = np.random.default_rng(3)
    np_, nf = 10, 120
    ijac = rng.normal(size=(nf, np_)) * np.exp(rng.uniform(-1, 1, np_))
    diff = rng.normal(size=nf) * 0.1
    pen = rng.normal(size=np_) * 0.01
    L = np.eye(np_) - np.eye(np_, k=-1)
    hssr = 0.05 * (L.T @ L)
    lam, tol = 0.3, 1e-4
    jac = rng.normal(size=(np_, nf))
    # a 1 x 3 PSF (the 3 x 2 map allows half-widths 1 along x, 0 along y)
    psf = np.array([[0.2, 0.7, 0.1]])
    cl, cr = lines.c_arrays(ls)
    fin, fout = tmp_path / "in.bin", tmp_path / "out.bin"
    with open(fin, "wb") as fh:
        fh.write(np.array([len(ls.lines), len(ls.regions), nz, nx, ny, np_, nf, 1], np.int32).tobytes())
        fh.write(bytes(cl) + bytes(cr))
        fh.write(np.asarray(z, np.float64).tobytes())
        fh.write(cube4.astype(np.float32).tobytes() + inv4.astype(np.float32).tobytes())
        # an unformatted READ of a derived type transfers its components one
        # after another, without the C struct's alignment padding (4 bytes
        # between PEN_TYP and PEN_ALP)
        raw, pad = bytes(ctl), gpu.CInvertControl.pen_alp.offset
        assert pad - (gpu.CInvertControl.pen_typ.offset + 32) == 4
        fh.write(raw[:pad - 4] + raw[pad:])
        fh.write(np.asfortranarray(ijac).tobytes(order="F") + diff.tobytes() + pen.tobytes() +
                 np.asfortranarray(hssr).tobytes(order="F") + np.array([lam, tol]).tobytes() +
                 np.asfortranarray(jac).tobytes(order="F"))
        fh.write(np.array(psf.shape, np.int32).tobytes() + np.asfortranarray(psf).tobytes(order="F"))
    print("[fortran] running fz_caller", flush=True)
    r = subprocess.run([CALLER, str(fin), str(fout)], capture_output=True, text=True, timeout=300)
    print(r.stdout[-500:], r.stderr[-500:], flush=True)
    assert r.returncode == 0 and "FZ_CALLER OK" in r.stdout
    # the same calls through ctypes
    cube = {n: np.asfortranarray(cube4[p].transpose(2, 1, 0)) for p, n in enumerate(gpu.ATM)}
    syn, dsyn = gpu_ctx.forward3d(cube, rf_mode=gpu.RF_SPHERICAL)
    inv_start = inv4.copy()        # the views below let invert3d overwrite inv4
    inv = {n: np.asfortranarray(inv4[p].transpose(2, 1, 0)) for p, n in enumerate(gpu.ATM)}
    syni, chi, its = gpu_ctx.invert3d(inv, syn, ctl)
    delta = gpu_ctx.get_perturbation(ijac, diff, lam, pen, hssr, dtype=1, svdtol=tol)
    hss = gpu_ctx.hessian(jac)
    invc = {n: np.asfortranarray(inv_start[p].transpose(2, 1, 0)) for p, n in enumerate(gpu.ATM)}
    gpu_ctx.set_psf(psf)
    try:
        sync, chic, itsc = gpu_ctx.invert3d(invc, syn, ctl, coupled=True)
    finally:
        gpu_ctx.set_psf(None)
    numw = gpu_ctx.numw
    with open(fout, "rb") as fh:
        f_numw, f_its = np.frombuffer(fh.read(8), np.int32)
        rd = lambda n, dt: np.frombuffer(fh.read(n * np.dtype(dt).itemsize), dt)
        f_cube = rd(12 * nx * ny * nz, np.float32).reshape(12, nx, ny, nz)
        f_syn = rd(4 * numw * ny * nx, np.float64).reshape(nx, ny, 4 * numw)
        f_dsyn = rd(4 * numw * 5 * nz * ny * nx, np.float32).reshape(nx, ny, 5 * nz, 4 * numw)
        f_inv = rd(12 * nx * ny * nz, np.float32).reshape(12, nx, ny, nz)
        f_syni = rd(4 * numw * ny * nx, np.float64).reshape(nx, ny, 4 * numw)
        f_chi = rd(ny * nx, np.float64).reshape(nx, ny)
        f_delta = rd(np_, np.float64)
        f_hss = rd(np_ * np_, np.float64).reshape(np_, np_).T
        f_itsc = int(rd(1, np.int32)[0])
        f_invc = rd(12 * nx * ny * nz, np.float32).reshape(12, nx, ny, nz)
        f_sync = rd(4 * numw * ny * nx, np.float64).reshape(nx, ny, 4 * numw)
        f_chic = rd(ny * nx, np.float64).reshape(nx, ny)
    assert f_numw == numw and f_its == its
    F = lambda a: np.ascontiguousarray(np.asarray(a).transpose(tuple(range(a.ndim))[::-1]))
    assert np.array_equal(f_syn, F(syn)) and np.array_equal(f_dsyn, F(dsyn))
    for p, n in enumerate(gpu.ATM):
        assert np.array_equal(f_cube[p], F(cube[n])), n
        assert np.array_equal(f_inv[p], F(inv[n])), n
    assert np.array_equal(f_syni, F(syni)) and np.array_equal(f_chi, F(chi))
    assert np.array_equal(f_delta, delta) and np.array_equal(f_hss, hss)
    assert f_itsc == itsc and np.array_equal(f_sync, F(sync)) and np.array_equal(f_chic, F(chic))
    for p, n in enumerate(gpu.ATM):
        assert np.array_equal(f_invc[p], F(invc[n])), n
    print("[fortran] FZ_FORWARD3D, FZ_INVERT3D (%d iterations), FZ_GET_PERTURBATION, FZ_HESSIAN, "
          "FZ_COUPLED_INVERT3D (%d iterations): identical to the ctypes calls" % (its, itsc), flush=True)
