"""The spatial PSF convolution on the device (fz_set_psf,
fz_psf_convolve_dev) against the oracle's restatement of the reference's FFT
route (oracle/psf.py): forward and adjoint, square and non-square maps, odd
and even extents, a 31 x 31 Airy + Gaussian-wing PSF as in the coupled
configuration and small asymmetric ones; the adjoint identity on the
device's own outputs; the argument errors (PSF wider than the map, no PSF)."""
import numpy as np
import pytest

from firtez_dz_amd import atmos, gpu
from oracle import psf as P

pytestmark = pytest.mark.gpu


def _dev(ctx, maps, nx, ny, adjoint=False):
    import torch
    t = torch.from_numpy(np.ascontiguousarray(maps.reshape(maps.shape[0], nx * ny), np.float32)).cuda()
    out = ctx.psf_convolve_torch(t, nx, ny, adjoint=adjoint)
    torch.cuda.synchronize()
    return out.cpu().numpy().reshape(maps.shape)


@pytest.mark.parametrize("nx,ny,psf_kind", [(64, 64, "airy31"), (45, 38, "airy31"), (20, 17, "rand7x5"),
                                            (33, 16, "rand3x9")])
def test_convolution_vs_oracle(gpu_ctx, nx, ny, psf_kind):
    rng = np.random.default_rng(nx * ny)
    if psf_kind == "airy31":
        psf = atmos.airy_gauss_psf(31, 2.5, 6.0)
    else:
        a, b = (int(v) for v in psf_kind[4:].split("x"))
        psf = rng.random((b, a))
    gpu_ctx.set_psf(psf)
    try:
        maps = rng.normal(size=(6, nx, ny)).astype(np.float32)
        for adj in (False, True):
            got = _dev(gpu_ctx, maps, nx, ny, adj)
            ref = P.convolve(maps.astype(np.float64), psf.astype(np.float32).astype(np.float64), adjoint=adj)
            err = np.abs(got - ref).max() / np.abs(ref).max()
            print("[psf] %dx%d %s adjoint=%d: max err %.2e of max |out|" % (nx, ny, psf_kind, adj, err), flush=True)
            assert err < 2e-6
        x = maps[:1]
        y = rng.normal(size=(1, nx, ny)).astype(np.float32)
        lhs = np.sum(_dev(gpu_ctx, x, nx, ny).astype(np.float64) * y)
        rhs = np.sum(x.astype(np.float64) * _dev(gpu_ctx, y, nx, ny, True))
        assert abs(lhs - rhs) <= 1e-5 * (abs(lhs) + abs(rhs)) / 2
    finally:
        gpu_ctx.set_psf(None)


def test_bench_shape_and_widest_psf(gpu_ctx):
    """The coupled benchmark's shape (448 maps of 256 x 256, the 31 x 31
    Airy + wing PSF) and the widest PSF the kernel takes (63 x 63, the
    64-tap instantiation) on a non-square map, against the oracle."""
    rng = np.random.default_rng(4)
    for nx, ny, nm, psf in ((256, 256, 448, atmos.airy_gauss_psf(31, 2.44, 4.0, 0.2)),
                            (70, 128, 3, rng.random((63, 63)))):
        gpu_ctx.set_psf(psf)
        try:
            maps = rng.normal(size=(nm, nx, ny)).astype(np.float32)
            for adj in (False, True):
                got = _dev(gpu_ctx, maps, nx, ny, adj)
                ref = P.convolve(maps.astype(np.float64), psf.astype(np.float32).astype(np.float64), adjoint=adj)
                err = np.abs(got - ref).max() / np.abs(ref).max()
                # FP32 accumulation over the taps: the rounding grows like sqrt(taps)
                tol = 1e-7 * np.sqrt(psf.size)
                print("[psf] %d maps %dx%d, PSF %dx%d adjoint=%d: max err %.2e of max |out| (tol %.1e)"
                      % (nm, nx, ny, psf.shape[0], psf.shape[1], adj, err, tol), flush=True)
                assert err < tol
        finally:
            gpu_ctx.set_psf(None)


def test_psf_errors(gpu_ctx):
    import torch
    t = torch.zeros((1, 20 * 20), dtype=torch.float32, device="cuda")
    with pytest.raises(gpu.FzError, match="odd"):
        gpu_ctx.set_psf(np.ones((4, 5)))
    with pytest.raises(gpu.FzError, match="no PSF"):
        gpu_ctx.psf_convolve_torch(t, 20, 20)
    gpu_ctx.set_psf(atmos.airy_gauss_psf(31, 2.5, 6.0))
    try:
        with pytest.raises(gpu.FzError, match="wider than the map"):
            gpu_ctx.psf_convolve_torch(t, 20, 20)
    finally:
        gpu_ctx.set_psf(None)
