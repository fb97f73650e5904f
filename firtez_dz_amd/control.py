"""Reference control files through the device path.

`parse` reads a FIRTEZ-dz control file (READ_INPUT_FILE,
input_file.f90:371-707, with the line reader of :60-98 and the field
readers SPLIT_READ_LINE_WEIGHTS / _STRING / GET_NUMBER, :1310-1420,
1868-1916) into a `Control`; `run` executes it the way the reference's main
program does -- one FORWARD3D for MODE synthesis / tau / hydrostatic
(invert3d.f90:158-177), INVERT3D for MODE inversion -- and writes the
reference's output files (DEALLOCATE_AND_WRITE, pre_post_duties.f90:108-355):

  synthesis    OUTPATH/<FILEPROFILE>, OUTPATH/out_<FILEMODEL>, and with
               WRITE RESPONSE FUNCTION: yes OUTPATH/rfout_<FILEMODEL>
  inversion    OUTPATH/out_<FILEPROFILE> (inverted Stokes parameters only),
               OUTPATH/out_<FILEMODEL>; observations from DATAPATH/<FILEPROFILE>;
               with COUPLED INVERSION also OUTPATH/cout_<FILEPROFILE>, the
               synthesis convolved with the PSF
  tau          OUTPATH/out_<FILEMODEL> (log tau5 from the forward's EOS)
  hydrostatic  OUTPATH/out_<FILEMODEL> (Pg, rho, Pe, mu from RK4 from the top)

Supported fields: LINES, FILEPROFILE, FILEMODEL, MODE, HYDROSTATIC,
HYDROSTATIC DERIVATIVES, STOKES SETUP, INVERSION SETUP (parameters with their
maximum nodes, POW, MAXITER), VERTICAL REGULARIZATION TYPE / NORM, WRITE
RESPONSE FUNCTION, MISC SETUP (DATAPATH, MODPATH, OUTPATH, SLDPATH, NOISE,
SVDTOL, LINEAR_SOLVER, PERTURBATION_TYPE).  The line data come from the
package's table of the reference's lines_database.dat (SLDPATH is read but
not needed).  STOKES SETUP: AUTO weights every pixel by the spread of its own
observed Stokes blocks (SET_WEIGHTS3D) in the uncoupled inversion.
LINE OPTICAL DEPTH: y writes OUTPATH/tl_<FILEMODEL> (WRITE_LOGTAUS) with
log10 tau of every wavelength and height in synthesis and inversion modes.  MISC SETUP FULL_STOKES writes, in synthesis mode,
OUTPATH/fs_<FILEPROFILE> (WRITE_BIN) with the Stokes vector after every
height, (4, NUMW, NZ, NY, NX).  BOX (three rows "N D" for x,
y, z) declares the grid instead of the model header (BOX_AVAILABLE,
input_file.f90:711-731); the runner requires the model file to hold that
grid.  CONTINUUM NORMALIZATION
(T: the HSRA continuum, the default; F: no normalisation, GAIN = 1 with the
V sign of GET_NORM) chooses GAIN1D.  COUPLED
INVERSION (yes and FILENAME, the PSF file in READ_PSF_2D's format under
DATAPATH) runs MODE inversion through that spatial PSF.  MISC SETUP TEMP / PGAS / BVEC / VLOS switch on SET_TAU's assist
initialisers.  LINE SPREAD FUNCTION
(Gaussian rows or LSF files) sets the library's LSF; NLTE DEPARTURE
COEFFICIENTS (per-line files of BETAL, BETAU per height) sets POPL/POPU.
"""
from __future__ import annotations

import os
from dataclasses import dataclass, field
from typing import Dict, List, Optional, Tuple

import numpy as np

from . import gpu as _gpu
from . import lines as _lines
from . import modelio

# INVERSION SETUP names (NSET_VAR_INVERSION, input_file.f90:1100-1115) -> INV_PARAMS
PARAM_NAMES = {"TEM": "T", "PGAS": "Pg", "RHO": "rho", "BX": "Bx", "BY": "By", "BZ": "Bz", "VLOS": "vz", "P0": "P0"}
# default weights of VERTICAL REGULARIZATION TYPE / NORM (input_file.f90:1139-1262)
_REG_DEF = {"TEM": 1, "P0": 0, "PGAS": 1, "RHO": 1, "BX": 0, "BY": 0, "BZ": 1, "VLOS": 1}
_UNSUPPORTED = ()


@dataclass
class Control:
    lines: List[Tuple[Tuple[int, ...], int, float, float]] = field(default_factory=list)
    fileprofile: Optional[str] = None
    filemodel: Optional[str] = None
    mode: Optional[str] = None              # synthesis, inversion, tau, hydrostatic
    hydrostatic: bool = False
    hydro_ders: bool = True
    invstk: List[bool] = field(default_factory=lambda: [False] * 4)
    wstk: List[float] = field(default_factory=lambda: [0.0] * 4)
    params: Dict[str, Optional[int]] = field(default_factory=dict)   # reference name -> NSLB_MAX (None: NZ)
    cycpow: int = 2
    maxsteps: int = 5
    vregul: bool = False
    pen_typ: Dict[str, int] = field(default_factory=dict)
    pen_alp: Dict[str, float] = field(default_factory=dict)
    write_rf: bool = False
    datapath: str = ""
    modpath: str = ""
    outpath: str = ""
    sldpath: str = ""
    noise: float = 4.0e-3
    svdtol: float = 0.0                     # 0: the library's default (RSVDTOL 1e-4)
    solver: int = 1
    node_expansion: int = 1
    # LINE SPREAD FUNCTION: per region (sigma, w0) in mA, or an LSF file name
    # (DATAPATH-relative), or None for EMPTY; empty list: no LSF
    lsf: List[object] = field(default_factory=list)
    # NLTE DEPARTURE COEFFICIENTS: (line index, file name or LTE) rows
    nlte: List[Tuple[int, str]] = field(default_factory=list)
    # MISC SETUP TEMP / PGAS / BVEC / VLOS: ASSIST_* bits
    assist: int = 0
    # COUPLED INVERSION: yes/no and the PSF file (DATAPATH-relative)
    coupled: bool = False
    psf_file: Optional[str] = None
    # CONTINUUM NORMALIZATION: T (HSRA, the default) or F (GAIN = 1)
    hsra_norm: bool = True
    # MISC SETUP FULL_STOKES: the Stokes vector after every height
    full_stokes: bool = False
    # LINE OPTICAL DEPTH: log10 tau of every wavelength (MTAULIN)
    line_tau: bool = False
    # STOKES SETUP: AUTO -- every pixel weighted by its own observations
    auto_weights: bool = False
    # BOX: ((NX, DX), (NY, DY), (NZ, DZ)) or None (grid from the model header)
    box: Optional[Tuple[Tuple[int, float], ...]] = None


def _number(tok: str) -> float:
    """GET_NUMBER (input_file.f90:1310-1352): integers, reals and the
    E/D exponent forms."""
    t = tok.replace("D", "E").replace("d", "E")
    return float(t)


def _nint(v: float) -> int:
    """Fortran NINT: halves away from zero."""
    return int(np.sign(v) * np.floor(abs(v) + 0.5))


def _field(line: str, match: str, default):
    """SPLIT_READ_LINE_WEIGHTS: (present, value) where value is the second
    column when given, else default."""
    cols = line.split()
    if not cols or cols[0] != match:
        return False, None
    if len(cols) > 1:
        return True, _number(cols[1])
    return True, default


def _string(line: str, match: str):
    cols = line.split()
    if len(cols) >= 2 and cols[0] == match:
        return cols[1]
    return None


def read_lines(text: str) -> List[str]:
    """The reader of input_file.f90:60-98: adjusted lines, no empty lines, no
    comment lines, '#' comments cut, everything after END: dropped."""
    out = []
    for raw in text.splitlines():
        ln = raw.strip()
        if not ln or ln.startswith("#"):
            continue
        if "#" in ln:
            ln = ln[:ln.index("#")]
        if ln.strip() == "END:":
            break
        out.append(ln)
    return out


def parse(text: str) -> Control:
    """Control from the text of a control file."""
    ls = read_lines(text)
    c = Control()
    nstk = 0
    i = 0

    def section(start):
        """lines of the section that starts after index start (up to the next
        line with a colon)"""
        j = start + 1
        while j < len(ls) and ":" not in ls[j]:
            j += 1
        return ls[start + 1:j], j - 1

    while i < len(ls):
        key = ls[i].strip()
        if key in _UNSUPPORTED:
            raise NotImplementedError("control field %s is not supported by this path" % key)
        if key == "LINES:":
            body, i = section(i)
            for row in body:
                cols = row.split()
                ids = tuple(int(v) for v in cols[0].split(",") if v)
                c.lines.append((ids, int(_number(cols[1])), _number(cols[2]), _number(cols[3])))
        elif key in ("FILEPROFILE:", "FILEMODEL:", "MODE:", "HYDROSTATIC:", "HYDROSTATIC DERIVATIVES:",
                     "WRITE RESPONSE FUNCTION:", "LINE OPTICAL DEPTH:"):
            i += 1
            val = ls[i].strip() if i < len(ls) else ""
            if key == "FILEPROFILE:":
                c.fileprofile = val
            elif key == "FILEMODEL:":
                c.filemodel = val
            elif key == "MODE:":
                m = val[:1].lower()
                c.mode = {"s": "synthesis", "i": "inversion", "t": "tau", "h": "hydrostatic"}.get(m)
                if c.mode is None:
                    raise ValueError('"MODE:" must be synthesis, inversion, tau or hydrostatic')
            elif key == "HYDROSTATIC:":
                c.hydrostatic = val[:1] in ("y", "Y")
            elif key == "HYDROSTATIC DERIVATIVES:":
                c.hydro_ders = val[:1] not in ("n", "N")
            elif key == "LINE OPTICAL DEPTH:":
                # input_file.f90:559-565: y / Y switches MTAULIN on
                c.line_tau = val[:1] in ("y", "Y")
            else:
                if val[:1] in ("y", "Y"):
                    c.write_rf = True
        elif key == "STOKES SETUP:":
            body, i = section(i)
            for row in body:
                for s, name in enumerate(("STKI", "STKQ", "STKU", "STKV")):
                    ok, w = _field(row, name, 1)
                    if ok:
                        c.invstk[s] = True
                        c.wstk[s] = float(w)
                        nstk += 1
                if _field(row, "AUTO", 1)[0]:
                    c.auto_weights = True
        elif key == "INVERSION SETUP:":
            body, i = section(i)
            for row in body:
                for name in PARAM_NAMES:
                    ok, v = _field(row, name, None)
                    if ok:
                        c.params[name] = None if v is None else _nint(v)
                ok, v = _field(row, "POW", 1)
                if ok:
                    c.cycpow = _nint(v)
                ok, v = _field(row, "MAXITER", 1)
                if ok:
                    c.maxsteps = _nint(v)
        elif key in ("VERTICAL REGULARIZATION TYPE:", "VERTICAL REGULARIZATION NORM:"):
            body, i = section(i)
            c.vregul = True
            for row in body:
                for name in PARAM_NAMES:
                    ok, v = _field(row, name, _REG_DEF[name])
                    if ok:
                        if key.endswith("TYPE:"):
                            c.pen_typ[name] = _nint(v)
                        else:
                            c.pen_alp[name] = float(np.float32(v))   # PEN_ALP is REAL(SP) (invert_param.f90:66)
        elif key == "NLTE DEPARTURE COEFFICIENTS:":
            # NSET_NLTE (input_file.f90:1979-2144): "<line index> <file | LTE>" rows
            body, i = section(i)
            for row in body:
                cols = row.split()
                if len(cols) != 2:
                    raise ValueError('wrong input format in "NLTE DEPARTURE COEFFICIENTS": %r' % row)
                c.nlte.append((int(_number(cols[0])), cols[1]))
        elif key == "LINE SPREAD FUNCTION:":
            # NSET_LSF (input_file.f90:1719-1865): one row per region, either
            # "sigma w0" (Gaussian, sigma /= 0) or an LSF file name, or EMPTY
            body, i = section(i)
            for row in body:
                cols = row.split()
                if len(cols) == 2:
                    sig, w0 = _number(cols[0]), _number(cols[1])
                    if sig == 0:
                        raise ValueError("LSF standard deviation of spectral region %d is 0" % (len(c.lsf) + 1))
                    c.lsf.append((sig, w0))
                elif len(cols) == 1:
                    c.lsf.append(None if cols[0].upper() == "EMPTY" else cols[0])
                else:
                    raise ValueError('wrong input format in "LINE SPREAD FUNCTION": %r' % row)
        elif key == "BOX:":
            # NREAD_BOX (input_file.f90:219-250): the next three rows, "N D"
            # for x, y and z (SPLIT_READ_BOX_LINE, :771-829)
            rows = []
            for _ in range(3):
                i += 1
                cols = ls[i].split() if i < len(ls) else []
                if len(cols) < 2:
                    raise ValueError('"BOX:" needs three rows of "N D"')
                rows.append((int(_number(cols[0])), float(_number(cols[1]))))
            c.box = tuple(rows)
        elif key == "CONTINUUM NORMALIZATION:":
            # NSET_CONTINUUM_NORMALIZATION (input_file.f90:1480-1520): the
            # first character of the section's last row, T or F
            body, i = section(i)
            first = body[-1].strip()[:1] if body else ""
            if first not in ("T", "F"):
                raise ValueError('"CONTINUUM NORMALIZATION:" must be either TRUE or FALSE, not %r' % first)
            c.hsra_norm = first == "T"
        elif key == "COUPLED INVERSION:":
            # NSET_COUPLED_INVERSION (input_file.f90:1524-1607): a one-column
            # row y/n, "FILENAME <psf file>"; BLCKSZ / NMTHRD / PSFRAD set up
            # the reference's unwired block solver and are read and not used
            body, i = section(i)
            for row in body:
                cols = row.split()
                if len(cols) == 1:
                    if cols[0][:1] in ("y", "Y"):
                        c.coupled = True
                elif len(cols) == 2:
                    if cols[0] == "FILENAME":
                        c.psf_file = cols[1]
                    elif cols[0] in ("BLCKSZ", "NMTHRD", "PSFRAD"):
                        _number(cols[1])
                else:
                    raise ValueError('"COUPLED INVERSION:" rows are yes/no and FILENAME <psf file>: %r' % row)
            if c.coupled and c.psf_file is None:
                raise ValueError('"COUPLED INVERSION:" with yes needs a FILENAME row')
        elif key == "MISC SETUP:":
            body, i = section(i)
            for row in body:
                # ASSIST_T/P/B/V (input_file.f90:170-176): SET_TAU's initialisers
                for name, bit in (("TEMP", _gpu.ASSIST_T), ("PGAS", _gpu.ASSIST_P), ("BVEC", _gpu.ASSIST_B),
                                  ("VLOS", _gpu.ASSIST_V)):
                    if _field(row, name, 0)[0]:
                        c.assist |= bit
                if _field(row, "FULL_STOKES", 0)[0]:
                    c.full_stokes = True
                ok, v = _field(row, "SVDTOL", 0)
                if ok:
                    c.svdtol = float(v)
                ok, v = _field(row, "NOISE", 0)
                if ok:
                    c.noise = float(np.float32(v))
                ok, v = _field(row, "LINEAR_SOLVER", 1)
                if ok:
                    c.solver = _nint(v)
                ok, v = _field(row, "PERTURBATION_TYPE", 1)
                if ok:
                    c.node_expansion = _nint(v)
                for name, attr in (("DATAPATH", "datapath"), ("MODPATH", "modpath"), ("OUTPATH", "outpath"),
                                   ("SLDPATH", "sldpath")):
                    s = _string(row, name)
                    if s is not None:
                        setattr(c, attr, s)
        i += 1
    # READ_INPUT_FILE's checks and defaults (input_file.f90:580-705)
    if c.filemodel is None:
        raise ValueError('the control file must contain a non-empty "FILEMODEL:" field')
    if c.mode is None:
        raise ValueError('the control file must contain a "MODE:" field')
    if c.mode == "synthesis" and nstk == 0:
        c.invstk = [True] * 4
        c.wstk = [1.0] * 4
    if c.mode == "inversion":
        if not any(c.invstk):
            raise ValueError('"STOKES SETUP:" is mandatory in inversion mode')
        if not c.params:
            raise ValueError('"INVERSION SETUP:" must name at least one parameter in inversion mode')
    if c.lsf and len(c.lsf) != len(c.lines):
        raise ValueError("the number of LSF (%d) and LINE (%d) do not match" % (len(c.lsf), len(c.lines)))
    if c.mode in ("synthesis", "inversion"):
        if not c.lines:
            raise ValueError('missing spectral lines ("LINES:")')
        if c.fileprofile is None:
            raise ValueError('"FILEPROFILE:" is required in synthesis and inversion modes')
    return c


def parse_file(path: str) -> Control:
    with open(path) as fh:
        return parse(fh.read())


def lsf_arrays(c: Control, ls: _lines.LineSetup, base: str = "."):
    """LSF_KERNEL and LSF_VALID of LSF_INIT (pre_post_duties.f90:770-898)
    from the LINE SPREAD FUNCTION rows: Gaussians where (sigma, w0) is
    given, the samples of DATAPATH/<file> (READ_LSF_FILE) where a file is
    named, nothing (an invalid region) for EMPTY."""
    sig = [e[0] if isinstance(e, tuple) else -1.0 for e in c.lsf]
    w0 = [e[1] if isinstance(e, tuple) else -1.0 for e in c.lsf]
    taps = {}
    for r, e in enumerate(c.lsf):
        if isinstance(e, str):
            path = os.path.join(base, c.datapath, e)
            if not os.path.exists(path):
                raise FileNotFoundError("cannot find LSF file %s in directory %s" % (e, c.datapath))
            taps[r] = modelio.read_lsf(path, ls.regions[r].nwave)
    return _lines.lsf_kernel(ls, sig, w0, taps)


def departure_arrays(c: Control, nz: int, base: str = "."):
    """POPL, POPU (nregions, nz) from the NLTE DEPARTURE COEFFICIENTS rows,
    with NSET_NLTE's rules (input_file.f90:1979-2144): a line with a file
    must be the main line of a LINES row and must not share its region with
    another line; its region takes the file's coefficients
    (READ_DEPARTURE_COEFFICIENTS, :2146-2166, DATAPATH-relative); every
    other region stays at 1 (LTE)."""
    popl = np.ones((len(c.lines), nz))
    popu = np.ones((len(c.lines), nz))
    for lid, fname in c.nlte:
        if fname.upper() == "LTE":
            continue
        for ids, *_ in c.lines:
            if lid in ids and len(ids) > 1:
                raise ValueError("NLTE lines cannot be blended (line %d)" % lid)
        regs = [r for r, (ids, *_) in enumerate(c.lines) if ids[0] == lid]
        if not regs:
            raise ValueError("departure coefficients for line %d, which is not under LINES" % lid)
        path = os.path.join(base, c.datapath, fname)
        if not os.path.exists(path):
            raise FileNotFoundError("cannot find NLTE file %s in directory %s" % (fname, c.datapath))
        bl, bu = modelio.read_departure(path, nz)
        popl[regs[0]], popu[regs[0]] = bl, bu
    return popl, popu


def line_setup(c: Control) -> _lines.LineSetup:
    """The configured regions; a mode without LINES (tau, hydrostatic) gets
    one wavelength of Fe I 6301.5, which leaves the model outputs unchanged
    (the reference sets up placeholder lines the same way, SETUP_FAKE_LINES)."""
    if c.lines:
        return _lines.setup(c.lines)
    return _lines.setup([((4,), 1, 0.0, 1.0)])


def _wave_index(ls: _lines.LineSetup):
    """WAVE and INDEX of WAVE_INIT (pre_post_duties.f90:592-621)."""
    wave = ls.wave()
    index = np.concatenate([np.full(r.nwave, ls.lines[r.lines[0]].lambda0) for r in ls.regions])
    return wave, index


def control_for_inversion(c: Control, nz: int) -> _gpu.CInvertControl:
    nodes = {PARAM_NAMES[k]: (nz if v is None else v) for k, v in c.params.items()}
    return _gpu.invert_control(nodes, invstk=[int(b) for b in c.invstk],
                               wstk=[c.wstk[s] if c.invstk[s] else 0.0 for s in range(4)], noise=c.noise,
                               cycpow=c.cycpow, maxsteps=c.maxsteps, solver=c.solver,
                               node_expansion=c.node_expansion, svdtol=c.svdtol, smoothing=True,
                               regul=c.vregul, pen_typ={PARAM_NAMES[k]: v for k, v in c.pen_typ.items()},
                               pen_alp={PARAM_NAMES[k]: v for k, v in c.pen_alp.items()}, assist=c.assist)


def run(c: Control, workdir: str = ".", ctx: Optional[_gpu.Context] = None) -> dict:
    """Executes the control file in workdir (its relative paths are taken from
    there, as the reference takes them from its working directory).  Returns
    the paths written and, for an inversion, chi^2 and the LM iteration count."""
    own = ctx is None
    if own:
        ctx = _gpu.Context(0)
    hydro_before = ctx.hydrostatic
    try:
        def path(p, name):
            return os.path.join(workdir, p + name)

        cube = modelio.read_model(path(c.modpath, c.filemodel))
        z = cube["z_km"]
        nz, ny, nx = cube["T"].shape
        if c.box is not None and (c.box[0][0], c.box[1][0], c.box[2][0]) != (nx, ny, nz):
            raise ValueError("BOX declares %dx%dx%d, the model file holds %dx%dx%d"
                             % (c.box[0][0], c.box[1][0], c.box[2][0], nx, ny, nz))
        for name in ("vx", "vy"):
            cube.setdefault(name, np.zeros_like(cube["T"], order="F"))
        ls = line_setup(c)
        ctx.configure(ls, z)
        if c.hsra_norm:
            ctx.gain_hsra()
        else:
            # GET_NORM without HSRA (normalization.f90:746-751): GAIN = 1,
            # V with the opposite sign
            g = np.ones((4, ctx.numw), np.float32)
            g[3] = -1.0
            ctx.set_gain(g)
        if c.lsf:
            ctx.set_lsf(*lsf_arrays(c, ls, workdir))
        if any(f.upper() != "LTE" for _, f in c.nlte):
            ctx.set_departure(*departure_arrays(c, nz, workdir))
        ctx.set_hydrostatic(c.hydrostatic)
        ctx.set_auto_weights(c.auto_weights)
        if c.hydrostatic and not c.hydro_ders and c.mode == "inversion":
            raise NotImplementedError("HYDROSTATIC DERIVATIVES: no is not supported by this path")
        out = {"written": []}
        wave, index = _wave_index(ls)
        numw = ctx.numw
        if c.mode == "hydrostatic":
            _hydrostatic_cube(ctx, cube)
        elif c.mode in ("synthesis", "tau"):
            want_rf = c.mode == "synthesis" and c.write_rf
            rf_mode = _gpu.RF_NONE
            if want_rf:
                rf_mode = _gpu.RF_CARTESIAN_P0 if "P0" in c.params else _gpu.RF_CARTESIAN
            res = ctx.forward3d(cube, rf_mode=rf_mode)
            syn3d, dsyn = (res if isinstance(res, tuple) else (res, None))
            if c.mode == "synthesis":
                st = _stokes4(syn3d, numw, ny, nx)
                for s in range(4):
                    if not c.invstk[s]:
                        st[s] = 0.0
                p = path(c.outpath, c.fileprofile)
                modelio.write_profiles(p, st, wave, index)
                out["written"].append(p)
                if c.full_stokes:
                    # DEALLOCATE_AND_WRITE (pre_post_duties.f90:265-271):
                    # fs_<FILEPROFILE>, SYN5D(4, NUMW, NZ, NY, NX)
                    p = path(c.outpath, "fs_" + c.fileprofile)
                    modelio.write_bin(p, _full_stokes(ctx, cube))
                    out["written"].append(p)
            if want_rf:
                npar = _gpu.NPARAM[rf_mode]
                # DSYN(NFREQ, IFREEP, NY, NX): NFREQ = Stokes-major blocks of
                # NUMW, IFREEP = parameter-major blocks of NZ
                d = np.asarray(dsyn).reshape((numw, 4, nz, npar, ny, nx), order="F")
                w = np.zeros((8, 4, numw, nz, ny, nx), np.float32)
                logv = np.zeros(8, bool)
                for k in c.params:
                    slot = _gpu.INV_PARAMS.index(PARAM_NAMES[k])
                    logv[slot] = True
                    w[slot] = d[:, :, :, slot].transpose(1, 0, 2, 3, 4)
                for s in range(4):
                    if not c.invstk[s]:
                        w[:, s] = 0.0
                p = path(c.outpath, "rfout_" + c.filemodel)
                modelio.write_rfs(p, w, z, wave, logv, index)
                out["written"].append(p)
        else:
            prof, pwave, _ = modelio.read_profiles(path(c.datapath, c.fileprofile))
            if prof.shape[1] != numw or prof.shape[2] != ny or prof.shape[3] != nx:
                raise ValueError("observed profiles (%s) do not match the lines and the model" % (prof.shape,))
            rows = [s for s in range(4) if c.invstk[s]]
            obs3d = np.asfortranarray(np.concatenate([prof[s] for s in rows], axis=0).astype(np.float64))
            if c.coupled:
                # READ_PSF (pre_post_duties.f90:488-555) reads DATAPATH//FILENAME
                ppath = path(c.datapath, c.psf_file)
                if not os.path.exists(ppath):
                    raise FileNotFoundError("cannot find PSF file %s in directory %s" % (c.psf_file, c.datapath))
                ctx.set_psf(modelio.read_psf(ppath))
                try:
                    syn_c, chi2, its = ctx.invert3d(cube, obs3d, control_for_inversion(c, nz), coupled=True)
                finally:
                    ctx.set_psf(None)         # END_COUPLED_INVERSION_*
                # DEALLOCATE_AND_WRITE (pre_post_duties.f90:195-245) writes the
                # inverted model's own synthesis to out_<FILEPROFILE> and its
                # PSF convolution to cout_<FILEPROFILE>
                res = ctx.forward3d(cube)
                syn3d = res[0] if isinstance(res, tuple) else res
            else:
                syn3d, chi2, its = ctx.invert3d(cube, obs3d, control_for_inversion(c, nz))
            outs = [("out_", syn3d)] + ([("cout_", syn_c)] if c.coupled else [])
            for prefix, sy in outs:
                st = _stokes4(sy, numw, ny, nx)
                for s in range(4):
                    if not c.invstk[s]:
                        st[s] = 0.0
                p = path(c.outpath, prefix + c.fileprofile)
                modelio.write_profiles(p, st, wave, index)
                out["written"].append(p)
            out["chi2"] = chi2
            out["iterations"] = its
        if c.line_tau and c.mode in ("synthesis", "inversion"):
            # WRITE_LINTAU (pre_post_duties.f90:168-180): tl_<FILEMODEL>
            p = path(c.outpath, "tl_" + c.filemodel)
            modelio.write_logtaus(p, _line_tau(ctx, cube), z, index, wave)
            out["written"].append(p)
        p = path(c.outpath, "out_" + c.filemodel)
        modelio.write_model(p, cube, x=cube["x"][0, 0, :], y=cube["y"][0, :, 0])
        out["written"].append(p)
        return out
    finally:
        # a caller's context leaves as it came (HYDROSTATIC is a context
        # setting that outlives the run)
        if own:
            ctx.close()
        else:
            ctx.set_hydrostatic(hydro_before)
            ctx.set_auto_weights(False)


def _full_stokes(ctx, cube):
    """SYN5D(4, NUMW, NZ, NY, NX) of the cube (FULL_STOKES, forward.f90:980-983)
    from a device forward; the cube itself is left as it is."""
    import torch
    nz, ny, nx = cube["T"].shape
    atm = torch.from_numpy(modelio.to_device_rows(cube)).cuda()
    _, _, status, full = ctx.forward_full_torch(atm)
    if int(status.max()) != 0:
        raise _gpu.FzError("full Stokes synthesis failed in %d columns" % int((status != 0).sum()))
    f = full.cpu().numpy().reshape(nx, ny, nz, 4, ctx.numw)        # [x][y][z][s][l], column = y + ny x
    return np.asfortranarray(f.transpose(3, 4, 2, 1, 0))


def _line_tau(ctx, cube):
    """TAU3DLIN(NUMW, NZ, NY, NX) of the cube (LINE OPTICAL DEPTH) from the
    device; the cube itself is left as it is."""
    import torch
    nz, ny, nx = cube["T"].shape
    atm = torch.from_numpy(modelio.to_device_rows(cube)).cuda()
    tl, status = ctx.line_tau_torch(atm)
    if int(status.max()) != 0:
        raise _gpu.FzError("line optical depth failed in %d columns" % int((status != 0).sum()))
    t = tl.cpu().numpy().reshape(nx, ny, nz, ctx.numw)              # [x][y][z][l], column = y + ny x
    return np.asfortranarray(t.transpose(3, 2, 1, 0))


def _stokes4(syn3d, numw, ny, nx):
    """SYN3D(4*NUMW, NY, NX) (Stokes-major blocks) -> STOKES(4, NUMW, NY, NX)."""
    return np.asarray(syn3d, np.float64).reshape((numw, 4, ny, nx), order="F").transpose(1, 0, 2, 3).copy()


def _hydrostatic_cube(ctx, cube):
    """RK4_INTEGRATION_LOG from the top on every column (MODE hydrostatic:
    FORWARD1D stops after the integration, forward.f90:891-909)."""
    import torch
    nz, ny, nx = cube["T"].shape
    atm = torch.from_numpy(modelio.to_device_rows(cube)).cuda()
    st = ctx.hydrostatic_torch(atm)
    if int(st.max()) != 0:
        raise _gpu.FzError("hydrostatic integration failed in %d columns" % int((st != 0).sum()))
    a = atm.cpu().numpy()
    for i, name in ((1, "Pg"), (2, "rho"), (3, "Pe"), (4, "mu")):
        cube[name][...] = np.asfortranarray(a[i].reshape(nz, nx, ny).transpose(0, 2, 1))
