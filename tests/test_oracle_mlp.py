"""The learned initial model's restatement (oracle/mlp.py), pinned by the
mathematics: central finite differences of the FP64 loss fix its gradient,
an independent BF16 conversion (torch) fixes the rounding, and the target /
decode pair is checked as a round trip on models the node basis represents."""
import numpy as np
import pytest

import oracle
from oracle import mlp


def rand_params(rng, nin, nh, nl, nout):
    dims = [nin] + [nh] * nl + [nout]
    parts = []
    for l in range(nl + 1):
        parts.append(rng.normal(0, np.sqrt(2.0 / dims[l]), dims[l] * dims[l + 1]))
        parts.append(rng.normal(0, 0.1, dims[l + 1]))
    norms = np.concatenate([rng.normal(0, 0.1, nin), rng.uniform(0.5, 2.0, nin), rng.normal(0, 0.1, nout),
                            rng.uniform(0.5, 2.0, nout)])
    return np.concatenate(parts + [norms])


def test_bf16_rounding():
    torch = pytest.importorskip("torch")
    x = np.array([1.0, 1 + 2.0 ** -8, 1 + 3 * 2.0 ** -8, -2.5e-3, 3.14159, 65504.0, 1e-30, -7.0e12], np.float32)
    ref = torch.from_numpy(x).to(torch.bfloat16).to(torch.float64).numpy()
    assert np.array_equal(mlp.bf16(x), ref)
    # ties go to the even mantissa
    assert mlp.bf16(np.float32(1 + 2.0 ** -8)) == 1.0
    assert mlp.bf16(np.float32(1 + 3 * 2.0 ** -8)) == 1 + 2.0 ** -6
    rng = np.random.default_rng(0)
    y = rng.normal(size=2000).astype(np.float32) * np.float32(10.0) ** rng.integers(-6, 6, 2000).astype(np.float32)
    assert np.array_equal(mlp.bf16(y), torch.from_numpy(y).to(torch.bfloat16).to(torch.float64).numpy())


def test_gradient_finite_differences():
    rng = np.random.default_rng(1)
    nin, nh, nl, nout, n = 6, 5, 2, 3, 7
    p = rand_params(rng, nin, nh, nl, nout)
    x = rng.normal(size=(n, nin))
    t = rng.normal(size=(n, nout))
    g = mlp.grad_fp64(p, x, t, nin, nh, nl, nout)
    npar = len(p) - 2 * (nin + nout)
    h = 1e-6
    fd = np.zeros(npar)
    for i in range(npar):
        a, b = p.copy(), p.copy()
        a[i] += h
        b[i] -= h
        fd[i] = (mlp.loss_fp64(a, x, t, nin, nh, nl, nout) - mlp.loss_fp64(b, x, t, nin, nh, nl, nout)) / (2 * h)
    assert np.max(np.abs(fd - g[:npar])) < 1e-6 * max(1.0, np.abs(g).max())
    assert not np.any(g[npar:])
    # the check has teeth: a sign mistake in the first layer's weight gradient
    # is far outside the bound
    bad = g.copy()
    bad[:nin * nh] *= -1.0
    assert np.max(np.abs(fd - bad[:npar])) > 1e-3 * np.abs(g).max()


def test_emulated_gradient_close_to_fp64():
    rng = np.random.default_rng(2)
    nin, nh, nl, nout, n = 40, 32, 3, 20, 50
    p = rand_params(rng, nin, nh, nl, nout)
    x = rng.normal(size=(n, nin))
    t = rng.normal(size=(n, nout))
    g64 = mlp.grad_fp64(p, x, t, nin, nh, nl, nout)
    gbf = mlp.loss_and_grad(p, x, t, nin, nh, nl, nout, emulate=True)[1]
    # BF16 operands (8-bit mantissas) through three layers with cancelling
    # sums: a few per cent of the largest entry, per layer; a sign or index
    # mistake would be of order one
    dims = [nin] + [nh] * nl + [nout]
    o = 0
    for l in range(nl + 1):
        sl = slice(o, o + dims[l] * dims[l + 1] + dims[l + 1])
        assert np.max(np.abs(gbf[sl] - g64[sl])) < 0.15 * np.abs(g64[sl]).max(), l
        o = sl.stop


def test_targets_decode_round_trip():
    cpu = oracle.Cpu()
    nz, nj, ncol = 64, 4, 9
    coefs = np.array(cpu.int_coefs(nz, nj))
    rng = np.random.default_rng(3)
    nodes_T = rng.uniform(4000, 9000, (ncol, nj))
    atm = np.zeros((12, nz, ncol), np.float32)
    atm[0] = (coefs @ nodes_T.T).astype(np.float32)
    B = rng.uniform(50, 2000, ncol)
    g = np.arccos(rng.uniform(-1, 1, ncol))
    f = rng.uniform(0, np.pi, ncol)
    atm[5] = (B * np.sin(g) * np.cos(f))[None, :]
    atm[6] = (B * np.sin(g) * np.sin(f))[None, :]
    atm[7] = (B * np.cos(g))[None, :]
    atm[10] = rng.normal(0, 2e5, ncol)[None, :]
    t = mlp.targets(atm, coefs)
    assert t.shape == (ncol, 5 * nj)
    assert np.allclose(t[:, :nj], nodes_T, rtol=1e-6)
    tm, ts = t.mean(axis=0), t.std(axis=0) + 1.0
    out = mlp.decode((t - tm) / ts, tm, ts, coefs)
    assert np.allclose(out["T"], atm[0].T, rtol=1e-6)
    assert np.allclose(out["Bz"], atm[7].T, atol=1e-3)
    assert np.allclose(out["vz"], atm[10].T, atol=1e-2)
    # B_perp and the azimuth modulo pi (the 180-degree ambiguity of Q and U)
    bp = np.hypot(out["Bx"], out["By"])
    assert np.allclose(bp, np.hypot(atm[5], atm[6]).T, atol=1e-3)
    same = np.isclose(out["Bx"], atm[5].T, atol=1e-3) & np.isclose(out["By"], atm[6].T, atol=1e-3)
    flip = np.isclose(out["Bx"], -atm[5].T, atol=1e-3) & np.isclose(out["By"], -atm[6].T, atol=1e-3)
    assert np.all(same | flip)
