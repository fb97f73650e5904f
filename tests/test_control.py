"""Control-file reader (firtez_dz_amd/control.py) on the reference's own
control files (tests/golden/control/*.dat are the files of the reference's
tests/test_*/control_file.dat, kept as data) and on hand-made cases of the
reader rules of input_file.f90:60-98, 1310-1420."""
import os

import numpy as np
import pytest

from firtez_dz_amd import control, lines

HERE = os.path.dirname(os.path.abspath(__file__))
CTL = os.path.join(HERE, "golden", "control")


def load(name):
    return control.parse_file(os.path.join(CTL, name + ".dat"))


def test_syn1():
    c = load("test_syn1")
    assert c.mode == "synthesis" and not c.hydrostatic
    assert c.lines == [((4, 5), 500, -700.0, 5.0)]
    assert c.fileprofile == "syn_profile.bin" and c.filemodel == "test_atmos.bin"
    assert c.invstk == [True] * 4 and c.wstk == [1.0] * 4
    assert (c.modpath, c.sldpath, c.outpath) == ("../../default/", "../../default/", "./out_dir/")
    assert not c.write_rf and not c.params
    ls = control.line_setup(c)
    assert ls.numw == 500 and [l.id for l in ls.lines] == [4, 5]
    # WAVE / INDEX of WAVE_INIT: start + step*j, the main line's lambda0
    wave, index = control._wave_index(ls)
    assert wave[0] == -700.0 and wave[-1] == -700.0 + 5.0 * 499
    assert np.all(index == ls.lines[0].lambda0)


def test_syn2_stops_at_end():
    c = load("test_syn2")
    # the row after END: is not read (input_file.f90:86-89)
    assert c.lines == [((72,), 500, -700.0, 5.0)]
    assert c.hydrostatic and c.mode == "synthesis"


def test_rf():
    c = load("test_rf")
    assert c.write_rf and c.params == {"TEM": None, "BZ": None}
    assert c.mode == "synthesis"


def test_he_tau():
    he, t1, t2 = load("test_he"), load("test_tau1"), load("test_tau2")
    assert he.mode == "hydrostatic" and he.hydrostatic and not he.lines
    assert t1.mode == "tau" and not t1.hydrostatic
    assert t2.mode == "tau" and t2.hydrostatic
    # modes without LINES get one placeholder wavelength
    assert control.line_setup(t1).numw == 1


def test_inv1():
    c = load("test_inv1")
    assert c.mode == "inversion" and c.hydrostatic and c.hydro_ders
    assert c.params == {"TEM": None, "VLOS": None}
    assert c.vregul and c.pen_typ == {"TEM": 1, "VLOS": 1}
    assert c.pen_alp == {"TEM": pytest.approx(0.01), "VLOS": pytest.approx(0.1)}
    assert c.solver == 2 and c.node_expansion == 2
    assert c.datapath == "../test_syn1/out_dir/"
    ctl = control.control_for_inversion(c, 256)
    assert list(ctl.atmpar) == [1, 0, 0, 0, 0, 0, 1, 0]
    assert list(ctl.nslb_max)[0] == 256 and list(ctl.nslb_max)[6] == 256
    assert ctl.regul == 1 and ctl.solver == 2 and ctl.node_expansion == 2 and ctl.smoothing == 1
    assert list(ctl.pen_typ)[0] == 1 and ctl.pen_alp[6] == pytest.approx(0.1)


def test_reader_rules():
    text = """
# a comment line
LINES:
4,5 112 -624.0 20.0   # trailing comment
  6 80 -400 10
MODE:
inversion
FILEMODEL:
m.bin
FILEPROFILE:
p.bin
STOKES SETUP:
STKI
STKV 2.0
INVERSION SETUP:
TEM 4
BZ 2
VLOS
POW 3
MAXITER 7
MISC SETUP:
NOISE 1.0e-3
SVDTOL 1D-5
OUTPATH ./o/
END:
MODE:
synthesis
"""
    c = control.parse(text)
    assert c.lines == [((4, 5), 112, -624.0, 20.0), ((6,), 80, -400.0, 10.0)]
    assert c.mode == "inversion"
    assert c.invstk == [True, False, False, True] and c.wstk[3] == 2.0
    assert c.params == {"TEM": 4, "BZ": 2, "VLOS": None}
    assert c.cycpow == 3 and c.maxsteps == 7
    assert c.noise == pytest.approx(1e-3) and c.svdtol == pytest.approx(1e-5)
    assert c.outpath == "./o/"
    ctl = control.control_for_inversion(c, 64)
    assert list(ctl.nslb_max)[:8] == [4, 0, 0, 0, 0, 2, 64, 0]
    assert list(ctl.wstk) == [1.0, 0.0, 0.0, 2.0]
    # halves round away from zero like NINT
    assert control._nint(2.5) == 3 and control._nint(-2.5) == -3


def test_reader_errors():
    with pytest.raises(ValueError):
        control.parse("MODE:\nsynthesis\n")                       # no FILEMODEL
    with pytest.raises(ValueError):
        control.parse("FILEMODEL:\nm.bin\n")                      # no MODE
    with pytest.raises(ValueError):
        control.parse("FILEMODEL:\nm.bin\nMODE:\ninversion\nLINES:\n4 10 0 1\nFILEPROFILE:\np\n")  # no Stokes
    with pytest.raises(ValueError):
        control.parse("FILEMODEL:\nm.bin\nMODE:\nsynthesis\nFILEPROFILE:\np\n")   # no lines
    # BOX (NREAD_BOX, input_file.f90:219-250): three rows "N D"
    c = control.parse("BOX:\n4 16.0\n3 16.0\n64 15.0\nFILEMODEL:\nm\nMODE:\ntau\n")
    assert c.box == ((4, 16.0), (3, 16.0), (64, 15.0))
    with pytest.raises(ValueError):
        control.parse("BOX:\n4 16.0\n3\nFILEMODEL:\nm\nMODE:\ntau\n")
    # LINE OPTICAL DEPTH (input_file.f90:559-565): y switches it on
    assert control.parse("LINE OPTICAL DEPTH:\ny\nFILEMODEL:\nm\nMODE:\ntau\n").line_tau
    assert not control.parse("LINE OPTICAL DEPTH:\nno\nFILEMODEL:\nm\nMODE:\ntau\n").line_tau
    # CONTINUUM NORMALIZATION (input_file.f90:1480-1520): T or F on the
    # section's last row, anything else is an error
    assert control.parse("FILEMODEL:\nm\nMODE:\ntau\n").hsra_norm
    assert not control.parse("CONTINUUM NORMALIZATION:\nF\nFILEMODEL:\nm\nMODE:\ntau\n").hsra_norm
    assert control.parse("CONTINUUM NORMALIZATION:\nF\nTRUE\nFILEMODEL:\nm\nMODE:\ntau\n").hsra_norm
    with pytest.raises(ValueError):
        control.parse("CONTINUUM NORMALIZATION:\nyes\nFILEMODEL:\nm\nMODE:\ntau\n")
    assert control.parse("FILEMODEL:\nm\nMODE:\ntau\nMISC SETUP:\nFULL_STOKES\n").full_stokes


def test_line_spread_function_section(tmp_path):
    """NSET_LSF (input_file.f90:1719-1865): one row per region -- "sigma w0",
    an LSF file name or EMPTY; sigma 0 and a row count different from the
    LINES count are errors."""
    from firtez_dz_amd import control, lines, modelio
    base = ("LINES:\n4,5 100 -300.0 10.0\n4 50 -200.0 10.0\n5 40 -100.0 10.0\nFILEPROFILE:\np.bin\n"
            "FILEMODEL:\nm.bin\nMODE:\nsynthesis\nMISC SETUP:\nDATAPATH ./\n")
    c = control.parse(base + "LINE SPREAD FUNCTION:\n25.0 3.0\nlsf_b.bin\nEMPTY\nEND:\n")
    assert c.lsf == [(25.0, 3.0), "lsf_b.bin", None]
    taps = np.exp(-0.5 * ((np.arange(50) - 25) / 3.0) ** 2)
    taps = np.roll(taps / taps.sum(), -25)
    modelio.write_lsf(str(tmp_path / "lsf_b.bin"), taps)
    ls = control.line_setup(c)
    kern, valid = control.lsf_arrays(c, ls, str(tmp_path))
    assert list(valid) == [1, 1, 0]
    k0, _ = lines.lsf_kernel(ls, [25.0, -1, -1], [3.0, -1, -1])
    assert np.array_equal(kern[:100], k0[:100])
    assert np.allclose(kern[100:150] * 50, taps / taps.sum(), rtol=1e-6) and np.all(kern[150:] == 0)
    with pytest.raises(ValueError):
        control.parse(base + "LINE SPREAD FUNCTION:\n0.0 3.0\n1 1\n1 1\nEND:\n")
    with pytest.raises(ValueError):
        control.parse(base + "LINE SPREAD FUNCTION:\n20.0 3.0\nEND:\n")


def test_nlte_departure_section(tmp_path):
    """NSET_NLTE / READ_DEPARTURE_COEFFICIENTS (input_file.f90:1979-2166):
    "<line> <file|LTE>" rows; the file's (BETAL, BETAU) per height go to the
    region whose LINES row names that line alone; a blended NLTE line and a
    line not under LINES are errors."""
    from firtez_dz_amd import control, modelio
    base = ("LINES:\n4 60 -300.0 10.0\n5 40 -200.0 10.0\nFILEPROFILE:\np.bin\nFILEMODEL:\nm.bin\n"
            "MODE:\nsynthesis\nMISC SETUP:\nDATAPATH ./\n")
    nz = 16
    bl, bu = np.linspace(0.8, 1.3, nz), np.linspace(1.1, 0.7, nz)
    modelio.write_departure(str(tmp_path / "b5.bin"), bl, bu)
    c = control.parse(base + "NLTE DEPARTURE COEFFICIENTS:\n4 LTE\n5 b5.bin\nEND:\n")
    assert c.nlte == [(4, "LTE"), (5, "b5.bin")]
    popl, popu = control.departure_arrays(c, nz, str(tmp_path))
    assert np.all(popl[0] == 1) and np.all(popu[0] == 1)
    assert np.array_equal(popl[1], bl) and np.array_equal(popu[1], bu)
    blended = base.replace("4 60 -300.0 10.0\n5 40", "4,5 60 -300.0 10.0\n5 40")
    with pytest.raises(ValueError, match="blended"):
        control.departure_arrays(control.parse(blended + "NLTE DEPARTURE COEFFICIENTS:\n5 b5.bin\nEND:\n"),
                                 nz, str(tmp_path))
    with pytest.raises(ValueError, match="not under LINES"):
        control.departure_arrays(control.parse(base + "NLTE DEPARTURE COEFFICIENTS:\n2 b5.bin\nEND:\n"),
                                 nz, str(tmp_path))
    with pytest.raises(ValueError):
        modelio.read_departure(str(tmp_path / "b5.bin"), nz + 1)


def test_misc_setup_assists():
    """MISC SETUP TEMP / PGAS / BVEC / VLOS set ASSIST_T/P/B/V
    (input_file.f90:170-176); FULL_STOKES (:177) sets the full-Stokes output."""
    from firtez_dz_amd import control, gpu
    base = ("LINES:\n4,5 112 -624.0 20.0\nFILEPROFILE:\np.bin\nFILEMODEL:\nm.bin\nMODE:\ninversion\n"
            "STOKES SETUP:\nSTKI\nSTKQ\nSTKU\nSTKV\nINVERSION SETUP:\nTEM 4\nBZ 1\n")
    c = control.parse(base + "MISC SETUP:\nBVEC\nVLOS\nTEMP\nEND:\n")
    assert c.assist == gpu.ASSIST_B | gpu.ASSIST_V | gpu.ASSIST_T
    assert control.control_for_inversion(c, 64).assist == c.assist
    assert control.parse(base + "MISC SETUP:\nPGAS\nEND:\n").assist == gpu.ASSIST_P
    assert control.parse(base + "MISC SETUP:\nFULL_STOKES\nEND:\n").full_stokes
    assert not c.full_stokes
    assert not c.auto_weights
    assert control.parse(base.replace("STKV\n", "STKV\nAUTO\n")).auto_weights


def test_coupled_inversion_section(tmp_path):
    """NSET_COUPLED_INVERSION (input_file.f90:1524-1607): a yes/no row and
    FILENAME <psf file> (BLCKSZ / NMTHRD / PSFRAD accepted); yes without a
    FILENAME and a row of three columns are errors.  READ_PSF_2D's file
    (misc.f90:2206-2348): the header record (2, -2, 161906, records, 2, NPX,
    NPY) and the values row by row, x fastest, written and read back."""
    from firtez_dz_amd import modelio
    base = ("LINES:\n4 20 -200.0 20.0\nFILEPROFILE:\np.bin\nFILEMODEL:\nm.bin\nMODE:\ninversion\n"
            "STOKES SETUP:\nSTKI 1\nINVERSION SETUP:\nTEM 2\n")
    c = control.parse(base + "COUPLED INVERSION:\nyes\nFILENAME psf.bin\nPSFRAD 5\nEND:\n")
    assert c.coupled and c.psf_file == "psf.bin"
    c = control.parse(base + "COUPLED INVERSION:\nno\nFILENAME psf.bin\n")
    assert not c.coupled
    with pytest.raises(ValueError):
        control.parse(base + "COUPLED INVERSION:\nyes\n")
    with pytest.raises(ValueError):
        control.parse(base + "COUPLED INVERSION:\nyes\nFILENAME a b\n")
    psf = np.arange(15, dtype=np.float64).reshape(3, 5) / 7.0        # PSF(NPY=3, NPX=5)
    f = str(tmp_path / "psf.bin")
    modelio.write_psf(f, psf)
    raw = open(f, "rb").read()
    hdr = np.frombuffer(raw, np.float32, 7, 4)
    assert np.frombuffer(raw, np.int32, 1, 0)[0] == 28
    assert list(hdr) == [2, -2, 161906, 2, 2, 5, 3]
    data = np.frombuffer(raw, np.float32, 15, 4 + 28 + 4 + 4)
    assert data[1] == np.float32(psf[0, 1]) and data[5] == np.float32(psf[1, 0])      # x fastest
    back = modelio.read_psf(f)
    assert back.shape == (3, 5) and np.array_equal(back, psf.astype(np.float32).astype(np.float64))


def test_write_bin_layout(tmp_path):
    """WRITE_BIN (misc.f90:2352-2418): header (3003, 2997, 160206, records,
    rank, shape...) and the values in Fortran order, read back unchanged."""
    from firtez_dz_amd import modelio
    a = np.asfortranarray(np.arange(2 * 3 * 4 * 5, dtype=np.float32).reshape(2, 3, 4, 5))
    f = str(tmp_path / "x.bin")
    modelio.write_bin(f, a)
    raw = open(f, "rb").read()
    hdr = np.frombuffer(raw, np.float32, 9, 4)
    assert list(hdr) == [3003, 2997, 160206, 2, 4, 2, 3, 4, 5]
    data = np.frombuffer(raw, np.float32, a.size, 4 + 36 + 4 + 4)
    assert data[1] == a[1, 0, 0, 0] and data[2] == a[0, 1, 0, 0]         # Fortran order
    assert np.array_equal(modelio.read_bin(f), a)


def test_write_logtaus_layout(tmp_path):
    """WRITE_LOGTAUS_V3 (misc.f90:2088-2199): header (3003, 2997, 12200904,
    4 + records, 4, NX, NY, NZ, NW), the Z, IND and WAVE records, then the
    values with the wavelength fastest; read back unchanged."""
    from firtez_dz_amd import modelio
    t = np.asfortranarray(np.arange(5 * 4 * 3 * 2, dtype=np.float32).reshape(5, 4, 3, 2))
    f = str(tmp_path / "tl.bin")
    modelio.write_logtaus(f, t, np.arange(4.0), np.full(5, 6301.5), np.linspace(-100, 100, 5))
    raw = open(f, "rb").read()
    assert list(np.frombuffer(raw, np.float32, 9, 4)) == [3003, 2997, 12200904, 5, 4, 2, 3, 4, 5]
    back, z, ind, wave = modelio.read_logtaus(f)
    assert np.array_equal(back, t) and np.array_equal(z, np.arange(4.0, dtype=np.float32))
