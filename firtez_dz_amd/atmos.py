"""Seeded synthetic atmospheres for the benchmark and parity workloads.

Per column (x, y) and height z (z = 0 at the bottom, uniform spacing dz):
  T(z)   = 4000 K + 5000 K * exp(-z / 200 km) + dT,   dT = white N(0,1) noise
           smoothed by a 3-D Gaussian (sigma = 3 cells; periodic in x, y,
           reflecting in z) and rescaled to a standard deviation of 200 K;
  |B| ~ U(0, 2000) G, cos(inclination) ~ U(-1, 1), azimuth ~ U(0, pi),
  v_z ~ N(0, 2 km/s), constant in z; v_x = v_y = 0;
  Pg from hydrostatic equilibrium integrated from the top (RK4 with the
  equation of state at every stage, hydrostatic_tools.f90:150-222), with a
  top pressure chosen so the bottom sits near log tau_5000 = +1.5.
Random numbers: numpy PCG64 with the given seed.  The rows follow the
device layout atm[12][nz][ncol] (T, Pg, rho, Pe, mu, Bx, By, Bz, vx, vy, vz,
tau5) with col = y + ny * x.
"""
from __future__ import annotations

import numpy as np

PG_TOP = {64: 40.0, 96: 0.3}


def pg_top_for(nz: int) -> float:
    return PG_TOP.get(nz, 40.0 if nz <= 64 else 0.3)


def seeded(nx: int, ny: int, nz: int, dz_km: float = 15.0, seed: int = 0):
    """Returns (atm[12, nz, ncol] float32 with Pg set only at the top, z_km[nz])."""
    from scipy.ndimage import gaussian_filter
    rng = np.random.Generator(np.random.PCG64(seed))
    z = np.arange(nz, dtype=np.float64) * dz_km
    base = 4000.0 + 5000.0 * np.exp(-z / 200.0)
    noise = rng.standard_normal((nz, ny, nx))
    sm = gaussian_filter(noise, sigma=3.0, mode=("reflect", "wrap", "wrap"))
    sm = sm / max(sm.std(), 1e-30) * 200.0
    T = base[:, None, None] + sm                       # (nz, ny, nx)
    ncol = nx * ny
    B = rng.uniform(0.0, 2000.0, (ny, nx))
    cg = rng.uniform(-1.0, 1.0, (ny, nx))
    phi = rng.uniform(0.0, np.pi, (ny, nx))
    v = rng.normal(0.0, 2.0e5, (ny, nx))
    sg = np.sqrt(1.0 - cg * cg)
    atm = np.zeros((12, nz, ncol), np.float32)
    # col = y + ny * x  -> transpose (ny, nx) to x-major flattening
    def cols(a2):
        return np.ascontiguousarray(a2.T).reshape(-1)
    atm[0] = np.ascontiguousarray(np.transpose(T, (0, 2, 1))).reshape(nz, ncol)
    atm[1, -1, :] = pg_top_for(nz)
    atm[5] = cols(B * sg * np.cos(phi))[None, :]
    atm[6] = cols(B * sg * np.sin(phi))[None, :]
    atm[7] = cols(B * cg)[None, :]
    atm[10] = cols(v)[None, :]
    return atm, z


def spherical(atm):
    """|B|, inclination, azimuth (mod pi) of each column from the Cartesian rows."""
    bx, by, bz = atm[5].astype(np.float64), atm[6].astype(np.float64), atm[7].astype(np.float64)
    B = np.sqrt(bx * bx + by * by + bz * bz)
    gam = np.arccos(np.clip(bz / np.maximum(B, 1e-30), -1, 1))
    phi = np.mod(np.arctan2(by, bx), np.pi)
    return B, gam, phi


def column_model(atm, z_km, col):
    """MODEL(NZ,13) of one column as a (13, nz) array [.., tau5, z]."""
    nz = atm.shape[1]
    m = np.zeros((13, nz), np.float32)
    m[:12] = atm[:, :, col]
    m[12] = np.asarray(z_km, np.float32)
    return m


def airy_gauss_psf(n: int, core: float, wing: float, wing_frac: float = 0.1) -> np.ndarray:
    """A synthetic n x n spatial PSF, PSF(NPY, NPX) with the centre at
    (n//2, n//2): an Airy pattern whose first zero lies `core` pixels from
    the centre (1.22 lambda/D over the pixel size) plus a Gaussian wing of
    standard deviation `wing` pixels carrying `wing_frac` of the light; unit
    sum."""
    from scipy.special import j1
    r0 = n // 2
    y, x = np.mgrid[-r0:n - r0, -r0:n - r0].astype(np.float64)
    r = np.hypot(x, y)
    v = 3.8317 * r / core             # first zero of J1 at v = 3.8317
    with np.errstate(invalid="ignore", divide="ignore"):
        airy = np.where(v > 0, (2.0 * j1(v) / v) ** 2, 1.0)
    airy /= airy.sum()
    g = np.exp(-0.5 * (r / wing) ** 2)
    g /= g.sum()
    return (1.0 - wing_frac) * airy + wing_frac * g
