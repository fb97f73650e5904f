"""The oracle's PSF convolution (oracle/psf.py, the reference's READ_PSF +
SPATIAL_CONV_INIT + SMOOTH_MODEL steps with numpy.fft) pinned by what the
mathematics fixes: scipy.signal.convolve2d with periodic boundaries (a
direct sum, independent of the FFT route), the adjoint identity
<A x, y> = <x, A^T y>, a delta PSF as the identity and a shifted delta as a
circular shift."""
import numpy as np
import pytest
from scipy.signal import convolve2d

from firtez_dz_amd import atmos
from oracle import psf as P


@pytest.mark.parametrize("nx,ny,npx,npy", [(24, 20, 7, 5), (33, 40, 31, 31), (16, 16, 9, 15)])
def test_convolution_matches_direct_periodic_sum(nx, ny, npx, npy):
    rng = np.random.default_rng(nx + ny)
    psf = rng.random((npy, npx))
    maps = rng.normal(size=(3, nx, ny))
    out = P.convolve(maps, psf)
    # the kernel as an (x, y) array centred on its middle element
    kxy = psf.T
    for m in range(3):
        ref = convolve2d(maps[m], kxy, mode="same", boundary="wrap")
        assert np.allclose(out[m], ref, rtol=1e-12, atol=1e-12 * np.abs(ref).max())


def test_adjoint_identity_and_special_kernels():
    rng = np.random.default_rng(3)
    nx, ny = 30, 26
    psf = atmos.airy_gauss_psf(31, 2.5, 6.0)
    assert abs(psf.sum() - 1.0) < 1e-12
    x, y = rng.normal(size=(nx, ny)), rng.normal(size=(nx, ny))
    lhs = np.sum(P.convolve(x, psf) * y)
    rhs = np.sum(x * P.convolve(y, psf, adjoint=True))
    assert abs(lhs - rhs) < 1e-12 * abs(lhs)
    delta = np.zeros((5, 5))
    delta[2, 2] = 1.0
    assert np.allclose(P.convolve(x, delta), x, atol=1e-14)
    shifted = np.zeros((5, 5))
    shifted[2, 3] = 1.0                    # PSF(Y0, X0 + 1): one pixel along +x
    assert np.allclose(P.convolve(x, shifted), np.roll(x, 1, axis=0), atol=1e-14)
    assert np.allclose(P.convolve(x, shifted, adjoint=True), np.roll(x, -1, axis=0), atol=1e-14)
