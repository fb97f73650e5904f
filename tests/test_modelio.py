"""Binary file formats of the reference (model 130904, profiles 160904,
response functions 18060904, LSF 3330904): the model writer reproduces the reference's
own model file byte for byte, and profiles / RF files round-trip."""
import os

import numpy as np
import pytest

from firtez_dz_amd import modelio

GOLD = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


def test_model_round_trip_is_byte_identical(tmp_path):
    src = os.path.join(GOLD, "test_atmos.bin")
    cube = modelio.read_model(src)
    assert cube["T"].shape == (256, 1, 1)
    out = tmp_path / "m.bin"
    modelio.write_model(str(out), cube, x=cube["x"][0, 0, :], y=cube["y"][0, :, 0])
    assert open(src, "rb").read() == out.read_bytes()


def test_profiles_round_trip(tmp_path):
    rng = np.random.default_rng(0)
    st = rng.normal(size=(4, 30, 3, 5)).astype(np.float32)
    wave = np.linspace(-700, 800, 30)
    modelio.write_profiles(str(tmp_path / "p.bin"), st, wave, index=np.full(30, 4.0))
    s2, w2, i2 = modelio.read_profiles(str(tmp_path / "p.bin"))
    assert np.array_equal(s2, st) and np.allclose(w2, wave.astype(np.float32)) and np.all(i2 == 4)
    raw = (tmp_path / "p.bin").read_bytes()
    hdr = np.frombuffer(raw, np.float32, 9, 4)
    assert list(hdr[2:]) == [160904, 4, 4, 5, 3, 30, 4]


def test_rf_round_trip(tmp_path):
    rng = np.random.default_rng(1)
    d = rng.normal(size=(8, 4, 6, 7, 2, 3)).astype(np.float32)
    logv = np.array([1, 0, 0, 1, 1, 1, 1, 0], bool)
    z = np.arange(7) * 15.0
    modelio.write_rfs(str(tmp_path / "r.bin"), d, z, np.arange(6) * 20.0, logv)
    d2, z2, w2, lv2 = modelio.read_rfs(str(tmp_path / "r.bin"))
    assert np.array_equal(lv2, logv)
    assert np.array_equal(d2[logv], d[logv]) and np.all(d2[~logv] == 0)
    assert np.allclose(z2, z)


def test_lsf_file_round_trip_and_layout(tmp_path):
    """LSF file (READ_LSF_FILE, misc.f90:2422-2555): header [SS+VV, SS-VV,
    3330904, nrec, 1, nw] then the samples; the bytes are checked against a
    file assembled by hand from that layout, and lsf_kernel takes the samples
    of a region that has no Gaussian width."""
    import struct

    from firtez_dz_amd import lines
    y = np.exp(-0.5 * ((np.arange(61) - 30) / 3.0) ** 2)
    y = np.roll(y / y.sum(), -30)
    p = tmp_path / "lsf.bin"
    modelio.write_lsf(str(p), y)
    hdr = np.array([3003, 2997, 3330904, 2, 1, 61], np.float32)
    body = y.astype(np.float32)
    ref = (struct.pack("<i", 24) + hdr.tobytes() + struct.pack("<i", 24) +
           struct.pack("<i", 244) + body.tobytes() + struct.pack("<i", 244))
    assert p.read_bytes() == ref
    back = modelio.read_lsf(str(p), nwave=61)
    assert np.array_equal(back, body.astype(np.float64))
    with pytest.raises(ValueError):
        modelio.read_lsf(str(p), nwave=60)
    ls = lines.setup([((4,), 61, -600.0, 20.0), ((5,), 52, -520.0, 20.0)])
    k, v = lines.lsf_kernel(ls, [-1, 30.0], [-1, 0.0], taps={0: back})
    assert list(v) == [1, 1]
    assert np.allclose(k[:61] * 61, back / back.sum(), rtol=1e-12)
