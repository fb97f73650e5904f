"""Readers/writers for the reference's binary model format (version 3).

A model file is a Fortran unformatted sequential file (4-byte record
markers) of REAL(4): record 1 = [a, b, 130904, nrec, 4, 13, nx, ny, nz]; the
following records hold the 13 parameter cubes (NZ,NY,NX) in the order
T, Pg, rho, Bx, By, Bz, vz, Pe, mu, log tau5, x, y, z (misc.f90:78-503,
801-929).  z is in km, B in G, v in cm/s.
"""
from __future__ import annotations

import numpy as np

MODEL_MAGIC = 130904
ORDER = ("T", "Pg", "rho", "Bx", "By", "Bz", "vz", "Pe", "mu", "tau5", "x", "y", "z")


def _records(path):
    raw = open(path, "rb").read()
    out, pos = [], 0
    while pos < len(raw):
        n = int(np.frombuffer(raw, np.int32, 1, pos)[0])
        out.append(np.frombuffer(raw, np.float32, n // 4, pos + 4).copy())
        pos += 4 + n + 4
    return out


def read_model(path):
    """Returns dict of (nz, ny, nx) float32 Fortran-order cubes + 'z' (nz,) in km."""
    recs = _records(path)
    hdr = recs[0]
    if abs(int(hdr[2]) - MODEL_MAGIC) > 0:
        raise ValueError("%s is not a model atmosphere file" % path)
    npar, nx, ny, nz = (int(v) for v in hdr[5:9])
    data = np.concatenate(recs[1:])
    n = nx * ny * nz
    if npar != 13 or data.size != 13 * n:
        raise ValueError("unexpected model layout")
    cube = {}
    for i, name in enumerate(ORDER):
        cube[name] = np.asfortranarray(data[i * n:(i + 1) * n].reshape((nx, ny, nz)).transpose(2, 1, 0))
    cube["z_km"] = cube["z"][:, 0, 0].astype(np.float64)
    return cube


def to_device_rows(cube):
    """Model cube -> device rows atm[12][nz][ncol] (col = y + ny*x)."""
    nz, ny, nx = cube["T"].shape
    names = ("T", "Pg", "rho", "Pe", "mu", "Bx", "By", "Bz", None, None, "vz", "tau5")
    atm = np.zeros((12, nz, nx * ny), np.float32)
    for i, nm in enumerate(names):
        if nm is None:
            continue
        atm[i] = np.ascontiguousarray(cube[nm].transpose(0, 2, 1)).reshape(nz, nx * ny)
    return atm
