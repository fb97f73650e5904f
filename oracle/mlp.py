"""Restatement of the learned initial model's network (no counterpart in the
reference: FIRTEZ-dz starts every inversion from the input model file,
invert3d.f90:57-178).

TEST INFRASTRUCTURE (the oracle): only tests/ use it, as the checker.

The device network (firtez_dz_amd/csrc/fz_mlp.hpp) is a ReLU multilayer
perceptron with BF16 weights, BF16 activations and FP32 accumulation.  Here
the same arithmetic runs in FP64 with the BF16 roundings placed where the
device rounds (inputs after standardisation, weights, hidden activations, the
output gradient and the back-propagated gradients), so that the device and
this restatement differ only by accumulation order and by the rare BF16
rounding that the order flips.  `grad_fp64` is the plain FP64 gradient; the
tests pin it by central finite differences of `loss_fp64` (the mathematics
fixes it), and pin `grad` against it at BF16 resolution.

Parameter vector layout (as fz_mlp_get_params): W_0 [out][in], b_0, W_1, b_1,
..., followed by xmean[nin], xstd[nin], tmean[nout], tstd[nout].
"""
from __future__ import annotations

import numpy as np


def bf16(x):
    """Round to the nearest BF16 (ties to even), returned as float64."""
    a = np.asarray(x, np.float32)
    u = a.view(np.uint32).astype(np.uint64)
    r = ((u + 0x7FFF + ((u >> 16) & 1)) >> 16) << 16
    return r.astype(np.uint32).view(np.float32).astype(np.float64)


def split(params, nin, nh, nl, nout):
    p = np.asarray(params, np.float64)
    dims = [nin] + [nh] * nl + [nout]
    ws, bs, o = [], [], 0
    for l in range(nl + 1):
        di, do = dims[l], dims[l + 1]
        ws.append(p[o:o + di * do].reshape(do, di))
        o += di * do
        bs.append(p[o:o + do])
        o += do
    norms = p[o:]
    xm, xs = norms[:nin], norms[nin:2 * nin]
    tm, ts = norms[2 * nin:2 * nin + nout], norms[2 * nin + nout:2 * nin + 2 * nout]
    return ws, bs, (xm, xs, tm, ts), o


def forward(params, x, nin, nh, nl, nout, emulate=True):
    """Standardised network output Y [n][nout]; with emulate the BF16
    roundings of the device."""
    rd = bf16 if emulate else (lambda v: np.asarray(v, np.float64))
    f32 = (lambda v: np.asarray(v, np.float32)) if emulate else (lambda v: np.asarray(v, np.float64))
    ws, bs, (xm, xs, _, _), _ = split(params, nin, nh, nl, nout)
    a = rd(f32((f32(x) - f32(xm)) / f32(xs)))
    acts = [a]
    for l in range(nl + 1):
        z = a @ rd(ws[l]).T + f32(bs[l])
        if l < nl:
            a = rd(np.maximum(z, 0.0))
            acts.append(a)
        else:
            return z, acts
    raise AssertionError


def loss_and_grad(params, x, t, nin, nh, nl, nout, emulate=True):
    """Mean squared error on standardised targets and its gradient with
    respect to the parameter vector (zeros on the normalisation entries)."""
    rd = bf16 if emulate else (lambda v: np.asarray(v, np.float64))
    f32 = (lambda v: np.asarray(v, np.float32)) if emulate else (lambda v: np.asarray(v, np.float64))
    ws, bs, (_, _, tm, ts), npar = split(params, nin, nh, nl, nout)
    y, acts = forward(params, x, nin, nh, nl, nout, emulate)
    tn = np.asarray(f32((f32(t) - f32(tm)) / f32(ts)), np.float64)
    n = y.shape[0]
    d = y - tn
    loss = float(np.mean(d * d))
    gz = 2.0 * d / (n * nout)
    g = np.zeros(len(params))
    dims = [nin] + [nh] * nl + [nout]
    offs = []
    o = 0
    for l in range(nl + 1):
        offs.append(o)
        o += dims[l] * dims[l + 1] + dims[l + 1]
    gzb = rd(gz)
    gzf = gz
    for l in range(nl, -1, -1):
        di, do = dims[l], dims[l + 1]
        a_in = acts[l]
        gw = gzb.T @ a_in
        g[offs[l]:offs[l] + di * do] = gw.reshape(-1)
        g[offs[l] + di * do:offs[l] + di * do + do] = gzf.sum(axis=0)
        if l == 0:
            break
        da = gzb @ rd(ws[l])
        da = da * (a_in > 0)
        gzf = da
        gzb = rd(da)
    assert o == npar
    return loss, g


def loss_fp64(params, x, t, nin, nh, nl, nout):
    return loss_and_grad(params, x, t, nin, nh, nl, nout, emulate=False)[0]


def grad_fp64(params, x, t, nin, nh, nl, nout):
    return loss_and_grad(params, x, t, nin, nh, nl, nout, emulate=False)[1]


def hat_pinv(coefs):
    """Least-squares node fit through the hat basis coefs [nz][nj]."""
    c = np.asarray(coefs, np.float64)
    return np.linalg.solve(c.T @ c, c.T)


def targets(atm, coefs):
    """Training targets of a model cube atm[12][nz][ncol]: per column the
    node values of T, Bz, B_perp cos 2phi, B_perp sin 2phi, v_z."""
    p = hat_pinv(coefs)
    T = p @ atm[0].astype(np.float64)
    bx = p @ atm[5].astype(np.float64)
    by = p @ atm[6].astype(np.float64)
    bz = p @ atm[7].astype(np.float64)
    vz = p @ atm[10].astype(np.float64)
    bp = np.hypot(bx, by)
    ph2 = 2.0 * np.arctan2(by, bx)
    return np.concatenate([T, bz, bp * np.cos(ph2), bp * np.sin(ph2), vz], axis=0).T   # [ncol][5*nj]


def decode(y, tm, ts, coefs):
    """Network output (standardised, [ncol][5*nj]) -> T, Bx, By, Bz, vz at the
    heights of the hat basis coefs [nz][nj] (T kept in 2500-18000 K)."""
    c = np.asarray(coefs, np.float64)
    nj = c.shape[1]
    q = np.asarray(y, np.float64) * ts + tm
    T, bz, b1, b2, vz = (q[:, i * nj:(i + 1) * nj] for i in range(5))
    bp = np.hypot(b1, b2)
    ph = 0.5 * np.arctan2(b2, b1)
    bx, by = bp * np.cos(ph), bp * np.sin(ph)
    out = {}
    for name, v in (("T", T), ("Bx", bx), ("By", by), ("Bz", bz), ("vz", vz)):
        out[name] = v @ c.T        # [ncol][nz]
    out["T"] = np.clip(out["T"], 2500.0, 18000.0)
    return out
