"""Spectral line-spread function (MLSF) in the oracle and the host-side
kernel builder.

The reference convolves through FFTW (forward.f90:621-657, 814-844), which is
not present here; the oracle's direct circular convolution is pinned against
numpy's DFT with the same unnormalised r2c -> product -> c2r sequence, and
against closed forms (a delta kernel is the identity, a constant row stays
constant, an invalid region is untouched).  The Gaussian kernel of LSF_INIT
(pre_post_duties.f90:823-876) is pinned by its symmetry, centring and sum.
"""
import numpy as np
import pytest

from firtez_dz_amd import atmos, lines

CART = (1, 1, 1, 1, 1, 1, 1, 0)


def two_regions():
    # odd and even pixel counts, Fe I 6301.5 and 6302.5 in regions of their own
    return lines.setup([((4,), 61, -600.0, 20.0), ((5,), 52, -520.0, 20.0)])


def fftw_like(x, k):
    """FFTW's r2c(x) * r2c(k) -> unnormalised c2r, in numpy."""
    n = x.size
    return np.fft.irfft(np.fft.rfft(x) * np.fft.rfft(k), n) * n


def apply_reference_lsf(row, ls, kern, valid):
    out = row.copy()
    p0 = 0
    for r, reg in enumerate(ls.regions):
        n = reg.nwave
        if valid[r]:
            out[p0:p0 + n] = fftw_like(row[p0:p0 + n], kern[p0:p0 + n]).astype(np.float32)
        p0 += n
    return out


def test_gaussian_kernel_closed_form():
    ls = two_regions()
    k, v = lines.lsf_kernel(ls, [40.0, 30.0], [0.0, 0.0])
    assert list(v) == [1, 1]
    a, b = k[:61] * 61, k[61:] * 52
    assert abs(a.sum() - 1) < 1e-14 and abs(b.sum() - 1) < 1e-14
    # odd count: centred on tap 0 and symmetric t <-> n - t
    assert np.argmax(a) == 0 and np.allclose(a[1:], a[1:][::-1], rtol=1e-13)
    # even count: the centre is half a step to the red, folded onto tap 0
    assert np.argmax(b) == 0 and np.allclose(b[1:], b[1:][::-1], rtol=1e-13)
    x = np.arange(-30, 31) * 20.0
    g = np.exp(-x ** 2 / (2 * 40.0 ** 2))
    assert np.allclose(a, np.roll(g / g.sum(), -30), rtol=1e-13)
    # a centre offset of one pixel moves the peak to tap 1
    k1, _ = lines.lsf_kernel(ls, [40.0, 30.0], [20.0, 0.0])
    assert np.argmax(k1[:61]) == 1
    assert np.allclose(k1[:61], np.roll(k[:61], 1), atol=1e-12)
    # neither a width nor samples: region not valid, no taps
    k2, v2 = lines.lsf_kernel(ls, [40.0, -1], [0.0, -1])
    assert list(v2) == [1, 0] and not np.any(k2[61:])
    # samples that do not sum to 1 within 0.01 are rejected
    k3, v3 = lines.lsf_kernel(ls, [-1, -1], [-1, -1], taps={0: np.full(61, 1.0 / 61), 1: np.full(52, 0.9 / 52)})
    assert list(v3) == [1, 0] and np.allclose(k3[:61], 1.0 / 61 / 61)


@pytest.fixture()
def setup_cpu(cpu):
    ls = two_regions()
    cpu.setup(ls, resp=True, atmpar=CART)
    cpu.set_gain(cpu.gain())
    yield cpu, ls
    cpu.set_lsf(None)


def test_oracle_lsf_matches_dft(setup_cpu):
    cpu, ls = setup_cpu
    atm, z = atmos.seeded(2, 1, 48, seed=21)
    m = atmos.column_model(atm, z, 1)
    pg, rho, pel, mw, _ = cpu.he(m[0], float(m[1, -1]), z)
    m[1], m[2], m[3], m[4] = pg, rho, pel, mw
    _, syn0, dsyn0, _ = cpu.forward1d(m)
    kern, valid = lines.lsf_kernel(ls, [45.0, 35.0], [10.0, -15.0])
    cpu.set_lsf(kern, valid)
    _, syn1, dsyn1, _ = cpu.forward1d(m)
    for s in range(4):
        ref = apply_reference_lsf(syn0[s], ls, kern, valid)
        err = np.max(np.abs(syn1[s] - ref)) / np.abs(syn0[0]).max()
        assert err < 2e-7, (s, err)
    worst = 0.0
    for p in (0, 1, 3, 6):
        for k in (5, 30):
            for s in range(4):
                row = dsyn0[k, :, p, s]
                ref = apply_reference_lsf(row, ls, kern, valid)
                sc = max(np.abs(row).max(), 1e-300)
                worst = max(worst, np.max(np.abs(dsyn1[k, :, p, s] - ref)) / sc)
    assert worst < 2e-7, worst
    # the convolution smooths: Stokes I's line core rises
    assert syn1[0, :61].min() > syn0[0, :61].min()


def test_oracle_lsf_delta_constant_invalid(setup_cpu):
    cpu, ls = setup_cpu
    atm, z = atmos.seeded(1, 1, 40, seed=3)
    m = atmos.column_model(atm, z, 0)
    pg, rho, pel, mw, _ = cpu.he(m[0], float(m[1, -1]), z)
    m[1], m[2], m[3], m[4] = pg, rho, pel, mw
    _, syn0, dsyn0, _ = cpu.forward1d(m)
    # delta taps: the identity up to the REAL(4) store
    kern = np.zeros(ls.numw)
    kern[0], kern[61] = 1.0 / 61, 1.0 / 52
    cpu.set_lsf(kern, [1, 1])
    _, syn1, dsyn1, _ = cpu.forward1d(m)
    assert np.array_equal(syn1, syn0.astype(np.float32).astype(np.float64))
    assert np.array_equal(dsyn1, dsyn0.astype(np.float32).astype(np.float64))
    # region 2 not valid: its pixels are bitwise those without the LSF
    kern, _ = lines.lsf_kernel(ls, [45.0, 35.0], [0.0, 0.0])
    cpu.set_lsf(kern, [1, 0])
    _, syn2, dsyn2, _ = cpu.forward1d(m)
    assert np.array_equal(syn2[:, 61:], syn0[:, 61:]) and np.array_equal(dsyn2[:, 61:], dsyn0[:, 61:])
    assert not np.array_equal(syn2[:, :61], syn0[:, :61])
    # a unit-sum kernel keeps a constant row constant
    flat = np.zeros(ls.numw)
    flat[:61] = 0.75
    y = apply_reference_lsf(flat, ls, kern, [1, 0])
    assert np.allclose(y[:61], 0.75, rtol=1e-7)
