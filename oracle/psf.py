"""The reference's spatial PSF convolution, restated in numpy (FP64).

TEST INFRASTRUCTURE (the oracle): only tests/ use it, as the checker.

READ_PSF (pre_post_duties.f90:488-555) embeds the central part of the PSF,
half-widths XS = MIN(NPX/2, NX/2) and YS = MIN(NPY/2, NY/2) around
(NPY/2+1, NPX/2+1), into an NY x NX array centred at (NY/2+1, NX/2+1);
SPATIAL_CONV_INIT (:559-588) shifts that centre to (1, 1) with CSHIFT, divides
by NX*NY (FFTW's unnormalised inverse) and keeps the r2c transform;
SMOOTH_MODEL (invert3d.f90:412-488) multiplies a map's transform by it and
transforms back: a circular convolution.  Here the same steps with
numpy.fft; the adjoint multiplies by the conjugate transform.  Maps are
(nx, ny) arrays (x first, the cube's column order col = y + ny * x).
"""
from __future__ import annotations

import numpy as np


def embedded_kernel(psf: np.ndarray, nx: int, ny: int) -> np.ndarray:
    """DPSF after SPATIAL_CONV_INIT's shift, as an (nx, ny) array with the
    PSF centre at [0, 0] (before the 1/(NX*NY) factor, which cancels in the
    forward-inverse pair).  psf: (npy, npx) like PSF(NPY, NPX)."""
    npy, npx = psf.shape
    xs, ys = min(npx // 2, nx // 2), min(npy // 2, ny // 2)
    x0, y0 = npx // 2, npy // 2            # 0-based PSF(NPY/2+1, NPX/2+1)
    k = np.zeros((nx, ny))
    for dx in range(-xs, xs + 1):
        for dy in range(-ys, ys + 1):
            k[dx % nx, dy % ny] += psf[y0 + dy, x0 + dx]
    return k


def convolve(maps: np.ndarray, psf: np.ndarray, adjoint: bool = False) -> np.ndarray:
    """maps (..., nx, ny) -> the PSF's circular convolution (or its adjoint)."""
    nx, ny = maps.shape[-2:]
    fk = np.fft.rfft2(embedded_kernel(psf, nx, ny))
    if adjoint:
        fk = np.conj(fk)
    return np.fft.irfft2(np.fft.rfft2(maps) * fk, s=(nx, ny))

