"""Pins the oracle's Levenberg-Marquardt step (oracle/fzlm.c) with checks that
do not restate its formulas:
  * node basis: partition of unity, and expansion = linear interpolation of
    the node values on their positions (numpy.interp);
  * cycle schedule: the node counts the reference's two-cycle default gives;
  * damped solve: numpy.linalg.pinv of H + lambda diag(H) with the same
    relative cutoff; the conjugate-gradient solver meets its own stopping test;
  * the full step: with one node per parameter the applied perturbation is a
    constant shift limited to 15 %; a rejected step restores the best model
    and raises lambda tenfold; on a linear forward model repeated steps reach
    the least-squares solution.
The reference's lm.f90 needs LAPACK (absent here) and get_dmodel.f90 needs MPI,
so neither is built; these checks are what pins the restatement.
"""
import numpy as np
import pytest

import oracle

NORMS = np.array([6000.0, 1e5, 1e-5, 1e3, 1e3, 1e3, 1e6, 600.0])


def test_node_basis_is_linear_interpolation(cpu):
    rng = np.random.default_rng(0)
    for nz, nj in [(64, 2), (64, 3), (64, 4), (96, 5), (64, 8), (64, 63)]:
        c = cpu.int_coefs(nz, nj)
        assert np.allclose(c.sum(axis=1), 1.0, atol=1e-14)
        pos = np.linspace(1.0, nz, nj)
        v = rng.normal(size=nj)
        assert np.allclose(c @ v, np.interp(np.arange(1, nz + 1), pos, v), atol=1e-13)
    assert np.array_equal(cpu.int_coefs(64, 1), np.ones((64, 1)))
    assert np.array_equal(cpu.int_coefs(16, 16), np.eye(16))


def test_cycle_schedule_default():
    """TEM/BX/BY/BZ/VLOS with 4 nodes max, POW 2: MAXITER = floor(log 4/log 2) = 2;
    cycle 1 gives T two nodes and the others one; cycle 2 all four (capped by
    NSLB_MAX)."""
    cpu = oracle.Cpu()
    maxiter = int(np.floor(np.log(4.0) / np.log(2.0)))
    assert maxiter == 2
    got = [[min(cpu.njevals(p, ci, maxiter, 2, 1, 64), 4) for p in (1, 4, 5, 6, 7)] for ci in (1, 2)]
    assert got == [[2, 1, 1, 1, 1], [4, 4, 4, 4, 4]]


def _pinv_delta(jac, diff, lam, svdtol, reg_pen=None, reg_hss=None):
    h = jac.T @ jac
    if reg_hss is not None:
        h = h + reg_hss
    a = h + np.diag(np.diag(h)) * lam
    g = jac.T @ diff - (0 if reg_pen is None else reg_pen)
    s = np.linalg.svd(a, compute_uv=False)
    # the reference keeps singular values >= svdtol * max
    return np.linalg.pinv(a, rcond=svdtol * (1 - 1e-12)) @ g, s


@pytest.mark.parametrize("seed", range(6))
def test_damped_solve_matches_pinv(cpu, seed):
    rng = np.random.default_rng(seed)
    nf, np_ = 448, 20
    jac = rng.normal(size=(nf, np_)) * np.logspace(0, -6, np_)   # ill-conditioned columns
    diff = rng.normal(size=nf)
    lam = 10 ** rng.uniform(-2, 1)
    ref, s = _pinv_delta(jac, diff, lam, 1e-4)
    got = cpu.get_perturbation(jac, diff, lam, dtype=1, svdtol=1e-4)
    assert np.sum(s < 1e-4 * s.max()) > 0          # the cutoff is exercised
    assert np.allclose(got, ref, rtol=1e-9, atol=1e-9 * np.abs(ref).max())
    # with a regularisation term
    rp = rng.normal(size=np_) * 0.1
    l = np.eye(np_) - np.eye(np_, k=-1)
    rh = l.T @ l * 0.01
    ref2, _ = _pinv_delta(jac, diff, lam, 1e-4, rp, rh)
    got2 = cpu.get_perturbation(jac, diff, lam, reg_pen=rp, reg_hss=rh, dtype=1, svdtol=1e-4)
    assert np.allclose(got2, ref2, rtol=1e-9, atol=1e-9 * np.abs(ref2).max())


def test_cg_solver_meets_its_stopping_rule(cpu):
    rng = np.random.default_rng(11)
    nf, np_ = 200, 12
    jac = rng.normal(size=(nf, np_))
    diff = rng.normal(size=nf)
    lam = 0.1
    h = jac.T @ jac
    a = h + np.diag(np.diag(h)) * lam
    b = jac.T @ diff
    x = cpu.get_perturbation(jac, diff, lam, dtype=2)
    cond = np.mean(((a @ x - b) / (np.abs(b) + 1e-5)) ** 2)
    assert cond < 1e-3
    assert np.allclose(x, np.linalg.solve(a, b), rtol=0.1, atol=0.05 * np.abs(np.linalg.solve(a, b)).max())


def _toy(nz=32, numw=20):
    z = np.linspace(0, 31 * 15, nz)
    m = np.zeros((13, nz), np.float32)
    m[0] = 5000.0
    m[1] = 1e4
    m[2] = 1e-7
    m[5], m[6], m[7] = 100.0, -50.0, 300.0
    m[10] = 1e5
    m[11] = np.linspace(1.0, -5.0, nz)        # log tau5, bottom first
    m[12] = z
    return m


def _cfg(nz, numw, nodes=(1, 0, 0, 1, 1, 1, 1, 0), **kw):
    atm = tuple(int(n > 0) for n in nodes)
    return oracle.lm_cfg(nz, numw, atm, nodes, **kw)


def test_one_node_step_is_limited_constant_shift(cpu):
    nz, numw = 32, 20
    cfg = _cfg(nz, numw)
    coefs = np.ones((5, nz))
    nf, np_ = 4 * numw, 5
    rng = np.random.default_rng(2)
    jac = (rng.normal(size=(nf, np_)) / NORMS[[0, 3, 4, 5, 6]]).astype(np.float32)
    m = _toy(nz, numw)
    syn = rng.normal(size=nf)
    # observed = a large linear offset (T +8000 K, B +/-, v +2 km/s) plus noise
    obs = syn + jac.astype(np.float64) @ np.array([8000.0, 40.0, -60.0, 90.0, 2e5]) + rng.normal(size=nf) * 1e-3
    isig = np.full(nf, 1.0 / 1e-3 ** 2)
    sla = np.array([0, 1e29, 1e29, 0, 0, 0, 0, 0, 10.0])
    mc, mb, sl, bs, jb, delta = cpu.lm_step(cfg, coefs, obs, syn, syn.copy(), sla, jac, jac.copy(), isig, m, m)
    # accepted: best = input model, lambda 10/sqrt(10), steps max(0,-1)=0
    assert np.array_equal(mb, m)
    assert sl[8] == pytest.approx(10 / np.sqrt(10)) and sl[0] == 0 and sl[2] == sl[1]
    # one node per parameter: constant shift per parameter, clamped to
    # max(0.15 * norm, 0.15 * |mean|)
    nclamped = 0
    for p, col in zip((0, 3, 4, 5, 6), (0, 5, 6, 7, 10)):
        shift = mc[col].astype(np.float64) - m[col]
        assert np.ptp(shift) < 1e-4 * max(1.0, abs(float(m[col][0]))), p
        lim = max(0.15 * NORMS[p], 0.15 * abs(float(m[col].mean())))
        assert abs(shift[0]) <= lim * (1 + 1e-6), p
        nclamped += abs(abs(shift[0]) - lim) < 1e-6 * lim
    assert nclamped >= 1
    # delta returns the applied (limited) node perturbation in physical units
    got = np.array([mc[c][5] - m[c][5] for c in (0, 5, 6, 7, 10)], np.float64)
    assert np.allclose(got, delta, rtol=1e-5, atol=1e-2)
    # before the limit, the step is the damped least-squares solution
    jn = jac.astype(np.float64) * (np.sqrt(isig) / np.sqrt(nf))[:, None] * NORMS[[0, 3, 4, 5, 6]]
    r = (obs - syn) * np.sqrt(isig) / np.sqrt(nf)
    raw, _ = _pinv_delta(jn, r, 10 / np.sqrt(10), 1e-4)
    raw = raw * NORMS[[0, 3, 4, 5, 6]]
    for i, p in enumerate((0, 3, 4, 5, 6)):
        col = (0, 5, 6, 7, 10)[i]
        lim = max(0.15 * NORMS[p], 0.15 * abs(float(m[col].mean())))
        assert got[i] == pytest.approx(np.clip(raw[i], -lim, lim), rel=1e-5, abs=1e-2)


def test_rejected_step_restores_best(cpu):
    nz, numw = 32, 20
    cfg = _cfg(nz, numw)
    coefs = np.ones((5, nz))
    nf = 4 * numw
    rng = np.random.default_rng(3)
    jc = rng.normal(size=(nf, 5)).astype(np.float32)
    jbest = rng.normal(size=(nf, 5)).astype(np.float32)
    best = _toy(nz, numw)
    cur = best.copy()
    cur[0] += 300.0
    obs = rng.normal(size=nf)
    syn = obs + 5.0                      # far worse than the stored best chi2
    bsyn = obs + 0.01
    isig = np.ones(nf)
    chibest = np.sum((obs - bsyn) ** 2) / nf
    sla = np.array([1.0, 0.0, chibest, 0.0, 0, 0, 0, 0, 0.5])
    mc, mb, sl, bs, jb, delta = cpu.lm_step(cfg, coefs, obs, syn, bsyn, sla, jc, jbest, isig, cur, best)
    assert np.array_equal(mb, best)             # best kept
    assert np.array_equal(jb, jbest)            # best Jacobian kept
    assert np.array_equal(bs, bsyn)
    assert sl[0] == 2.0 and sl[8] == pytest.approx(5.0) and sl[2] == chibest
    # the new trial starts from the best model (T offset of the rejected one is gone)
    assert abs(float(mc[0][10] - best[0][10])) <= 0.15 * 6000 + 1e-3


def test_linear_forward_converges_to_least_squares(cpu):
    """syn(x) = s0 + J x exactly, with x the (constant) T, Bz and v_z of the
    model and one node each: repeated steps reach the truth.  (With several
    nodes per parameter the reference's SIR-type re-projection
    expand(contract(.)) moves the model even at zero step, so its fixed point
    is not the least-squares node vector; one node keeps the check exact.)"""
    nz, numw = 32, 30
    nodes = (1, 0, 0, 0, 0, 1, 1, 0)
    cfg = _cfg(nz, numw, nodes=nodes)
    coefs = np.ones((3, nz))
    nf, np_ = 4 * numw, 3
    rng = np.random.default_rng(5)
    scale = np.array([1 / 6000.0, 1 / 1e3, 1 / 1e6])
    jac = (rng.normal(size=(nf, np_)) * scale).astype(np.float32)
    m0 = _toy(nz, numw)
    m0[11] = np.linspace(0.0, -4.0, nz)           # keep the T tail rule out of the way
    true = m0.copy()
    true[0] += np.float32(700.0)
    true[7] += np.float32(450.0)
    true[10] += np.float32(3e5)

    def params(m):
        return np.array([m[0].mean(), m[7].mean(), m[10].mean()], np.float64)

    s0 = rng.normal(size=nf)

    def synth(m):
        return s0 + jac.astype(np.float64) @ (params(m) - params(m0))

    obs = synth(true)
    isig = np.full(nf, 1e4)
    cur, best = m0.copy(), m0.copy()
    sla = np.array([0, 1e29, 1e29, 0, 0, 0, 0, 0, 10.0])
    bsyn = synth(cur)
    jbest = jac.copy()
    chis = []
    for it in range(30):
        syn = synth(cur)
        cur, best, sla, bsyn, jbest, _ = cpu.lm_step(cfg, coefs, obs, syn, bsyn, sla, jac, jbest, isig, cur,
                                                      best)
        chis.append(sla[2])
    assert all(b <= a for a, b in zip(chis, chis[1:]))   # the best chi2 never increases
    assert chis[-1] < 1e-6 * chis[0]
    assert np.allclose(params(best), params(true), rtol=1e-4)


@pytest.mark.parametrize("expansion", [1, 2])
def test_zero_step_reprojection(cpu, expansion):
    """CONTRACT_VECTOR (get_dmodel.f90:861-880) is a basis-weighted mean,
    not the pseudo-inverse of EXPAND_VECTOR (:839-857), so with several
    nodes a step of zero still moves the model (SORT_DELTA_TO_ATMOSPHERE,
    :1715-1836): NODE EXPANSION 2 replaces a parameter by E(C x); NODE
    EXPANSION 1 adds E(C(E(C x) - x)).  Checked with the node basis built as
    numpy.interp hat functions, on v_z (no tail rule, no bounds in reach)."""
    nz, numw, nj = 32, 20, 4
    nodes = (0, 0, 0, 0, 0, 0, nj, 0)
    cfg = _cfg(nz, numw, nodes=nodes, node_expansion=expansion)
    pos = np.linspace(1.0, nz, nj)
    E = np.stack([np.interp(np.arange(1, nz + 1), pos, np.eye(nj)[j]) for j in range(nj)], axis=1)
    C = (E / E.sum(axis=0)).T
    coefs = cpu.coefs_block(cfg, {6: cpu.int_coefs(nz, nj)})
    nf = 4 * numw
    rng = np.random.default_rng(9)
    jac = rng.normal(size=(nf, nj)).astype(np.float32)
    m = _toy(nz, numw)
    m[10] = (1e5 + 4e4 * np.sin(np.linspace(0.0, 5.0, nz))).astype(np.float32)
    syn = rng.normal(size=nf)
    sla = np.array([0, 1e29, 1e29, 0, 0, 0, 0, 0, 10.0])
    mc, mb, sl, bs, jb, delta = cpu.lm_step(cfg, coefs, syn.copy(), syn, syn.copy(), sla, jac, jac.copy(),
                                            np.ones(nf), m, m)
    x = m[10].astype(np.float64)
    assert np.max(np.abs(C @ E - np.eye(nj))) > 0.1          # not a projection
    if expansion == 2:
        want = E @ (C @ x)
    else:
        want = x + E @ (C @ (E @ (C @ x) - x))
    assert np.allclose(mc[10], want, rtol=2e-6, atol=0.05), np.max(np.abs(mc[10] - want))
    assert np.max(np.abs(mc[10] - x)) > 100.0                # the model moved
