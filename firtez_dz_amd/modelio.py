"""Readers/writers for the reference's binary files (version 3).

All are Fortran unformatted sequential files (4-byte record markers) of
REAL(4), with a header record [SS+VV, SS-VV, magic, nrec, ndims, sizes...]
and data records of at most 268435446 values (misc.f90:15-17):
  * model 130904 (misc.f90:78-503, 801-929): header [.., 130904, nrec, 4, 13,
    nx, ny, nz]; the 13 cubes (NZ,NY,NX) T, Pg, rho, Bx, By, Bz, vz, Pe, mu,
    log tau5, x, y, z.  z in km, B in G, v in cm/s.
  * profiles 160904 (misc.f90:1147-1249, 1338-1642): header [.., 160904,
    nrec, 4, nx, ny, nw, 4]; a record of line indices, one of wavelengths (mA),
    then STOKES(4, nw, ny, nx).
  * response functions 18060904 (misc.f90:1862-1999): header [.., 18060904,
    nrec, 6, nx, ny, nz, nw, 4, npar]; records z, line index, wavelength,
    parameter mask (8), then DSTOKES(p, s, w, k, y, x) for the stored p.
  * line-spread function 3330904 (READ_LSF_FILE, misc.f90:2422-2555):
    header [.., 3330904, nrec, 1, nw]; then the nw samples of one region.
"""
from __future__ import annotations

import numpy as np

MODEL_MAGIC = 130904
PROFILE_MAGIC = 160904
LSF_MAGIC = 3330904
RF_MAGIC = 18060904
PSF_MAGIC = 161906
BIN_MAGIC = 160206
LOGTAU_MAGIC = 12200904
VERSION = (3000, 3)      # SS, VV (format 3000, version 3) written as SS+VV, SS-VV (misc.f90:944-953)
NMAXREC = 1073741784 // 4
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


def _write_records(path, header, blocks):
    """header and blocks: float32 arrays, each written as records; the data
    block is split into records of at most NMAXREC values."""
    with open(path, "wb") as f:
        def rec(a):
            a = np.ascontiguousarray(a, dtype=np.float32)
            n = np.array([a.nbytes], np.int32)
            f.write(n.tobytes())
            f.write(a.tobytes())
            f.write(n.tobytes())
        rec(header)
        for b in blocks[:-1]:
            rec(b)
        data = np.ascontiguousarray(blocks[-1], dtype=np.float32).reshape(-1)
        for i in range(0, data.size, NMAXREC):
            rec(data[i:i + NMAXREC])


def _nrecs(n):
    return int(np.ceil(np.float32(n) / np.float32(NMAXREC)))


def write_model(path, cube, x=None, y=None):
    """WRITE_MODEL3D_V3: cube dict of (nz, ny, nx) arrays (names of ORDER;
    'z_km' or 'z'), x (nx,) and y (ny,) in km (default 0)."""
    nz, ny, nx = cube["T"].shape
    ss, vv = VERSION
    n = nx * ny * nz
    hdr = np.array([ss + vv, ss - vv, MODEL_MAGIC, _nrecs(13 * n) + 1.0, 4.0, 13, nx, ny, nz], np.float32)
    z = np.asarray(cube["z_km"] if "z_km" in cube else cube["z"][:, 0, 0], np.float64)
    xs = np.zeros(nx) if x is None else np.asarray(x, np.float64)
    ys = np.zeros(ny) if y is None else np.asarray(y, np.float64)
    parts = []
    for name in ORDER:
        if name == "x":
            c = np.broadcast_to(xs.astype(np.float32)[None, None, :], (nz, ny, nx))
        elif name == "y":
            c = np.broadcast_to(ys.astype(np.float32)[None, :, None], (nz, ny, nx))
        elif name == "z":
            c = np.broadcast_to(z.astype(np.float32)[:, None, None], (nz, ny, nx))
        else:
            c = np.asarray(cube[name], np.float32)
        parts.append(np.asarray(c).reshape(-1, order="F"))
    _write_records(path, hdr, [np.concatenate(parts)])


def write_profiles(path, stokes, wave_ma, index=None):
    """WRITE_PROFILES3D_V3: stokes (4, nw, ny, nx) (Fortran STOKES(4,NW,NY,NX)),
    wave_ma (nw,) relative wavelengths, index (nw,) line index of each point."""
    st = np.asarray(stokes, np.float32)
    ns, nw, ny, nx = st.shape
    ss, vv = VERSION
    hdr = np.array([ss + vv, ss - vv, PROFILE_MAGIC, _nrecs(st.size) + 3.0, 4, nx, ny, nw, ns], np.float32)
    ind = np.zeros(nw) if index is None else np.asarray(index, np.float64)
    _write_records(path, hdr, [ind.astype(np.float32), np.asarray(wave_ma, np.float64).astype(np.float32),
                               st.reshape(-1, order="F")])


def read_profiles(path):
    """READ_PROFILE_3D (v3): returns (stokes (4, nw, ny, nx), wave_ma, index)."""
    recs = _records(path)
    hdr = recs[0]
    if int(round(float(hdr[2]))) != PROFILE_MAGIC:
        raise ValueError("%s is not a profile file" % path)
    nx, ny, nw, ns = (int(v) for v in hdr[5:9])
    ind, wave = recs[1], recs[2]
    data = np.concatenate(recs[3:])
    if data.size != ns * nw * ny * nx:
        raise ValueError("unexpected profile layout")
    return data.reshape((ns, nw, ny, nx), order="F"), wave, ind


def write_rfs(path, dstokes, z_km, wave_ma, logv, index=None):
    """WRITE_RFS3D_V3: dstokes (8, 4, nw, nz, ny, nx) in Fortran order
    DSTOKES(P,S,W,K,J,I); logv (8,) which parameters are stored (T|Pg, Pg|T,
    rho|T, Bx, By, Bz, vz, P0)."""
    d = np.asarray(dstokes, np.float32)
    npar, ns, nw, nz, ny, nx = d.shape
    logv = np.asarray(logv, bool)
    ss, vv = VERSION
    keep = d[logv]
    hdr = np.array([ss + vv, ss - vv, RF_MAGIC, 1 + 1 + 1 + 1 + 1 + _nrecs(keep.size), 6, nx, ny, nz, nw, ns, npar],
                   np.float32)
    ind = np.zeros(nw) if index is None else np.asarray(index, np.float64)
    _write_records(path, hdr, [np.asarray(z_km, np.float64).astype(np.float32), ind.astype(np.float32),
                               np.asarray(wave_ma, np.float64).astype(np.float32), logv.astype(np.float32),
                               keep.reshape(-1, order="F")])


def read_rfs(path):
    """Reads a WRITE_RFS3D_V3 file: returns (dstokes (8, 4, nw, nz, ny, nx) with
    zeros for the parameters not stored, z_km, wave_ma, logv)."""
    recs = _records(path)
    hdr = recs[0]
    if int(round(float(hdr[2]))) != RF_MAGIC:
        raise ValueError("%s is not a response-function file" % path)
    nx, ny, nz, nw, ns, npar = (int(v) for v in hdr[5:11])
    z, ind, wave, logv = recs[1], recs[2], recs[3], recs[4] > 0.5
    data = np.concatenate(recs[5:])
    keep = data.reshape((int(logv.sum()), ns, nw, nz, ny, nx), order="F")
    out = np.zeros((npar, ns, nw, nz, ny, nx), np.float32)
    out[logv] = keep
    return out, z, wave, logv


def write_lsf(path, samples):
    """Line-spread-function file of one spectral region (the layout
    READ_LSF_FILE reads; the record count in the header is the data records
    plus one, as the reference's own writer sets it)."""
    data = np.asarray(samples, np.float32).reshape(-1)
    nw = data.size
    header = np.array([VERSION[0] + VERSION[1], VERSION[0] - VERSION[1], LSF_MAGIC, _nrecs(nw) + 1, 1, nw],
                      np.float32)
    _write_records(path, header, [data])


def read_lsf(path, nwave=None):
    """Samples (float64) of a line-spread-function file; nwave, when given,
    must equal the file's size, as READ_LSF_FILE requires of the region."""
    recs = _records(path)
    hdr = recs[0]
    if abs(int(round(float(hdr[2]))) - LSF_MAGIC) > 0:
        raise ValueError("%s is not a line spread function file" % path)
    if int(hdr[4]) != 1:
        raise ValueError("%s: a line spread function file has one dimension" % path)
    nw = int(hdr[5])
    if nwave is not None and nw != nwave:
        raise ValueError("%s holds %d wavelengths, the region has %d" % (path, nw, nwave))
    data = np.concatenate(recs[1:]) if len(recs) > 1 else np.zeros(0, np.float32)
    if data.size != nw:
        raise ValueError("%s: %d samples, header says %d" % (path, data.size, nw))
    return data.astype(np.float64)


def write_departure(path, popl, popu):
    """NLTE departure-coefficient file of one line: one Fortran unformatted
    record (BETAL, BETAU) of two REAL*8 per height, bottom first, as
    READ_DEPARTURE_COEFFICIENTS reads it (input_file.f90:2146-2166)."""
    popl = np.asarray(popl, np.float64).reshape(-1)
    popu = np.asarray(popu, np.float64).reshape(-1)
    with open(path, "wb") as f:
        n = np.array([16], np.int32).tobytes()
        for bl, bu in zip(popl, popu):
            f.write(n + np.array([bl, bu], np.float64).tobytes() + n)


def read_departure(path, nz):
    """(BETAL[nz], BETAU[nz]) of a departure-coefficient file; the file must
    hold at least nz records of two REAL*8 (the reference reads NZ of them)."""
    raw = open(path, "rb").read()
    bl, bu, pos = [], [], 0
    for k in range(nz):
        if pos + 4 > len(raw):
            raise ValueError("%s: %d records, the model has %d heights" % (path, k, nz))
        n = int(np.frombuffer(raw, np.int32, 1, pos)[0])
        if n < 16:
            raise ValueError("%s: record %d holds %d bytes, two REAL*8 expected" % (path, k + 1, n))
        v = np.frombuffer(raw, np.float64, 2, pos + 4)
        bl.append(float(v[0]))
        bu.append(float(v[1]))
        pos += 4 + n + 4
    return np.array(bl), np.array(bu)


def write_psf(path, psf):
    """Spatial PSF file, psf (npy, npx) like PSF(NPY, NPX): the header record
    (VV, -VV, 161906, records, 2, NPX, NPY) with version VV = 2, then the
    values row by row (x fastest), as READ_PSF_2D (misc.f90:2206-2348)
    reads them and the reference's Python writer writes them."""
    a = np.asarray(psf, np.float32)
    if a.ndim != 2:
        raise ValueError("a PSF is two-dimensional")
    npy, npx = a.shape
    header = np.array([2, -2, PSF_MAGIC, _nrecs(a.size) + 1, 2, npx, npy], np.float32)
    _write_records(path, header, [a.reshape(-1)])


def read_psf(path):
    """PSF(NPY, NPX) (float64 of the file's float32 values) of a spatial PSF
    file (READ_PSF_2D, misc.f90:2206-2348: version 2 only, two dimensions,
    PSF(iy, ix) = value iy * NPX + ix of the data records)."""
    recs = _records(path)
    hdr = recs[0]
    pos, neg = int(hdr[0]), int(hdr[1])
    med = (pos + neg) // 2
    if (pos - med) + (neg - med) > 0.01:
        raise ValueError("%s: not a recognised format" % path)
    if abs(int(round(float(hdr[2]))) - PSF_MAGIC) > 0:
        raise ValueError("%s is not a spatial PSF file" % path)
    if int(hdr[4]) != 2:
        raise ValueError("%s: a spatial PSF file has two dimensions" % path)
    vv = ((pos - med) - (neg - med)) // 2
    if vv != 2:
        raise ValueError("%s: PSF file version %d is not supported (the reference reads version 2)" % (path, vv))
    npx, npy = int(hdr[5]), int(hdr[6])
    data = np.concatenate(recs[1:]) if len(recs) > 1 else np.zeros(0, np.float32)
    if data.size != npx * npy:
        raise ValueError("%s: %d values, header says %d x %d" % (path, data.size, npx, npy))
    return data.reshape(npy, npx).astype(np.float64)


def write_bin(path, array):
    """WRITE_BIN (misc.f90:2352-2418): the header record (SS+VV, SS-VV,
    160206, records, RK, SH(1..RK)) with format 3000 version 3, then the
    array's values in Fortran order; array holds the Fortran shape SH."""
    a = np.asarray(array, np.float32)
    ss, vv = VERSION
    header = np.array([ss + vv, ss - vv, BIN_MAGIC, _nrecs(a.size) + 1, a.ndim] + list(a.shape), np.float32)
    _write_records(path, header, [a.reshape(-1, order="F")])


def read_bin(path):
    """The array of a WRITE_BIN file, with its Fortran shape."""
    recs = _records(path)
    hdr = recs[0]
    if abs(int(round(float(hdr[2]))) - BIN_MAGIC) > 0:
        raise ValueError("%s is not a WRITE_BIN file" % path)
    rk = int(hdr[4])
    shape = tuple(int(v) for v in hdr[5:5 + rk])
    data = np.concatenate(recs[1:]) if len(recs) > 1 else np.zeros(0, np.float32)
    if data.size != int(np.prod(shape)):
        raise ValueError("%s: %d values, header says %s" % (path, data.size, shape))
    return data.reshape(shape, order="F")


def write_logtaus(path, taus, z_km, index, wave_ma):
    """WRITE_LOGTAUS_V3 (misc.f90:2088-2199): taus (NW, NZ, NY, NX) of
    log10 tau per wavelength and height; the header record (SS+VV, SS-VV,
    12200904, records, 4, NX, NY, NZ, NW), then Z, IND and WAVE records and
    the values with the wavelength fastest."""
    t = np.asarray(taus, np.float32)
    nw, nz, ny, nx = t.shape
    ss, vv = VERSION
    header = np.array([ss + vv, ss - vv, LOGTAU_MAGIC, _nrecs(t.size) + 4, 4, nx, ny, nz, nw], np.float32)
    _write_records(path, header, [np.asarray(z_km, np.float64).astype(np.float32),
                                  np.asarray(index, np.float64).astype(np.float32),
                                  np.asarray(wave_ma, np.float64).astype(np.float32), t.reshape(-1, order="F")])


def read_logtaus(path):
    """(taus (NW, NZ, NY, NX), z, index, wave) of a WRITE_LOGTAUS file."""
    recs = _records(path)
    hdr = recs[0]
    if abs(int(round(float(hdr[2]))) - LOGTAU_MAGIC) > 0:
        raise ValueError("%s is not an optical-depth file" % path)
    nx, ny, nz, nw = (int(v) for v in hdr[5:9])
    data = np.concatenate(recs[4:]) if len(recs) > 4 else np.zeros(0, np.float32)
    if data.size != nw * nz * ny * nx:
        raise ValueError("%s: %d values, header says %d" % (path, data.size, nw * nz * ny * nx))
    return data.reshape((nw, nz, ny, nx), order="F"), recs[1], recs[2], recs[3]
