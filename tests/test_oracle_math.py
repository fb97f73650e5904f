"""Closed-form and invariant checks of the oracle's pieces, chosen so that a
dropped term, a wrong sign or a transposed operand fails one of them:
  * Voigt/Faraday-Voigt against the Faddeeva function (scipy.special.wofz)
    and the a = 0 limit (Gaussian, Dawson);
  * Zeeman patterns: LS Lande factors, per-type normalisation, the
    sigma-component centroid = g_eff * 4.66685e-10 * lambda0^2;
  * the evolution operator against expm(-K dz), its derivatives against the
    Frechet derivative of expm, the semigroup law and det O = exp(-4 eta_I dz);
  * the response functions against central finite differences of the forward
    synthesis (magnetic field and velocity: no equation-of-state iteration in
    the loop, so the differences are clean).
"""
import numpy as np
import pytest
from scipy.linalg import expm, expm_frechet
from scipy.special import dawsn, wofz

from firtez_dz_amd import lines


def kmat(x):
    """Propagation matrix whose exponential the reference's closed form
    evaluates (rtesolver.f90:374-473): first row/column eta, and the
    magneto-optical block K(2,3) = -rho_V, K(2,4) = rho_U, K(3,4) = -rho_Q."""
    ei, eq, eu, ev, rq, ru, rv = x
    return np.array([[ei, eq, eu, ev],
                     [eq, ei, -rv, ru],
                     [eu, rv, ei, -rq],
                     [ev, -ru, rq, ei]])


def test_voigt_vs_faddeeva(cpu):
    v = np.linspace(-30, 30, 6001)
    for a in (0.01, 0.05, 0.2, 1.0, 3.0):
        h, f = cpu.voigt(a, v)
        w = wofz(v + 1j * a)
        assert np.max(np.abs(h - w.real)) < 2e-5, a
        assert np.max(np.abs(f - w.imag)) < 2e-5, a
        # L is odd, H is even
        assert np.allclose(f, -f[::-1], atol=1e-14) and np.allclose(h, h[::-1], atol=1e-14)


def test_voigt_zero_damping_limit(cpu):
    v = np.linspace(-25, 25, 2001)
    h, f = cpu.voigt(0.0, v)
    assert np.max(np.abs(h - np.exp(-v * v))) < 1e-14
    # the reference interpolates a 28-point Dawson table with 3-point Lagrange
    # polynomials (voigt.f:47-74); near v = 2.4 that is good to ~1e-2
    assert np.max(np.abs(f - 2.0 / np.sqrt(np.pi) * dawsn(v))) < 1.2e-2


def test_zeeman_patterns(cpu):
    ls = lines.fe6301_6302()
    cpu.setup(ls, resp=False)
    for j, (geff, n) in enumerate(((5.0 / 3.0, 13), (2.5, 3))):
        s, st, ty = cpu.zeeman(j)
        assert len(s) == n
        for t in (-1, 0, 1):
            assert abs(st[ty == t].sum() - 1.0) < 1e-14
        l0 = ls.lines[j].lambda0
        unit = 4.66685e-10 * l0 * l0
        for t in (-1, 1):
            cen = np.sum(st[ty == t] * s[ty == t])
            assert abs(abs(cen) - geff * unit) < 1e-6 * unit
        assert abs(np.sum(st[ty == 0] * s[ty == 0])) < 1e-6 * unit


def random_x(rng, scale):
    pol = rng.normal(0, 1, 6) * scale * rng.uniform(0.05, 0.8)
    return np.concatenate([[scale], pol])


def test_evolution_operator_vs_expm(cpu):
    rng = np.random.default_rng(3)
    for _ in range(200):
        dz_km = rng.uniform(1, 30)
        x = random_x(rng, 10 ** rng.uniform(-9, -6))
        O, _ = cpu.evol(x, dz_km)
        ref = expm(-kmat(x) * dz_km * 1e5)
        assert np.max(np.abs(O - ref)) < 1e-10 * max(1.0, np.abs(ref).max())


def test_evolution_operator_invariants(cpu):
    rng = np.random.default_rng(4)
    for _ in range(50):
        x = random_x(rng, 10 ** rng.uniform(-8, -6.5))
        a, _ = cpu.evol(x, 7.0)
        b, _ = cpu.evol(x, 5.0)
        c, _ = cpu.evol(x, 12.0)
        assert np.max(np.abs(a @ b - c)) < 1e-12
        assert abs(np.linalg.det(c) - np.exp(-4 * x[0] * 12e5)) < 1e-10 * np.exp(-4 * x[0] * 12e5) + 1e-300
    # unpolarised limit: O = exp(-eta_I dz) * 1
    O, _ = cpu.evol(np.array([2e-7, 0, 0, 0, 0, 0, 0]), 10.0)
    assert np.allclose(O, np.exp(-2e-7 * 1e6) * np.eye(4), rtol=1e-14, atol=0)


def test_evolution_operator_derivatives_vs_frechet(cpu):
    rng = np.random.default_rng(5)
    for _ in range(100):
        dz_km = rng.uniform(2, 20)
        x = random_x(rng, 10 ** rng.uniform(-8, -6))
        dx = rng.normal(0, 1, (8, 7)) * x[0]
        _, dO = cpu.evol(x, dz_km, dx)
        for p in range(8):
            _, ref = expm_frechet(-kmat(x) * dz_km * 1e5, -kmat(dx[p]) * dz_km * 1e5)
            assert np.max(np.abs(dO[p] - ref)) < 1e-8 * np.abs(ref).max() + 1e-300


@pytest.mark.parametrize("slot,row,step", [(3, 5, 1.0), (4, 6, 1.0), (5, 7, 1.0), (6, 10, 200.0)])
def test_response_functions_vs_finite_differences(cpu, golden, slot, row, step):
    """d Stokes / d(Bx, By, Bz, v_z) at three heights by central differences."""
    g = golden["seeded_c1"]
    cpu.setup(lines.fe6301_6302(), resp=True, atmpar=(1, 1, 1, 1, 1, 1, 1, 0))
    cpu.set_gain(g["gain"])
    m = g["models"][0].copy()
    _, syn, dsyn, _ = cpu.forward1d(m)
    for k in (20, 35, 50):
        mp, mm = m.copy(), m.copy()
        mp[row, k] += step
        mm[row, k] -= step
        h = float(mp[row, k]) - float(mm[row, k])
        sp = cpu.forward1d(mp)[1]
        sm = cpu.forward1d(mm)[1]
        fd = (sp - sm) / h
        an = dsyn[k, :, slot, :].T
        scale = np.abs(an).max()
        assert np.max(np.abs(fd - an)) < 2e-3 * scale, (slot, k)
