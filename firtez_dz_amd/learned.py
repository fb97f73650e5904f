"""Learned initial model for the inversion.

The reference starts INVERT3D from the input model file (invert3d.f90:57-178)
and has no learned component.  Here a small network (fz_mlp_* entry points,
BF16 MFMA kernels) maps observed Stokes profiles to node values of T, Bz,
B_perp cos 2phi, B_perp sin 2phi and v_z; the proposal replaces those rows of
the default initial model, Pg is re-integrated hydrostatically, and the
unchanged INVERT3D runs from there.  Pixels whose final chi^2 exceeds a
threshold are inverted again from the default model (plain LM) and keep the
better of the two fits, so plain LM judges every proposal it does not accept.

  training_set   atmosphere/Stokes pairs from the device forward (seeded
                 cubes of the same statistics as the observations, Gaussian
                 noise added to the profiles)
  train          Adam over shuffled mini-batches, cosine learning-rate decay
  invert         learned initial model + INVERT3D + plain-LM fallback
"""
from __future__ import annotations

import math
import time
from dataclasses import dataclass, field
from typing import List, Optional

import numpy as np

from . import atmos
from . import gpu as _gpu


def training_set(ctx, ncubes: int, nx: int, ny: int, nz: int, dz_km: float = 15.0, seed0: int = 1000,
                 noise: float = 1e-3, log=None, transform=None):
    """Returns (X [n][4*numw] float32, T [n][5*nodes] float32) on the device,
    n = ncubes*nx*ny.  Each cube is atmos.seeded(seed0 + i) put in
    hydrostatic equilibrium and synthesised by the device forward;
    transform (optional) maps a cube's Stokes rows [ncol][4*numw] to the
    observed ones before the noise (e.g. the spatial PSF of a coupled
    inversion)."""
    import torch
    ncol = nx * ny
    nin = 4 * ctx.numw
    X = torch.empty((ncubes * ncol, nin), dtype=torch.float32, device="cuda")
    T = torch.empty((ncubes * ncol, ctx.mlp_shape[3]), dtype=torch.float32, device="cuda")
    gen = torch.Generator(device="cuda")
    gen.manual_seed(seed0)
    for i in range(ncubes):
        atm, _ = atmos.seeded(nx, ny, nz, dz_km=dz_km, seed=seed0 + i)
        a = torch.from_numpy(atm).cuda()
        st = ctx.hydrostatic_torch(a)
        stokes, _, status = ctx.forward_torch(a, _gpu.RF_NONE)
        ok = (st == 0) & (status == 0)
        rows = slice(i * ncol, (i + 1) * ncol)
        obs = stokes.reshape(ncol, nin)
        if transform is not None:
            obs = transform(obs)
        X[rows] = obs + noise * torch.randn((ncol, nin), generator=gen, device="cuda")
        T[rows] = ctx.mlp_targets_torch(a)
        if not bool(ok.all()):
            # a column the EOS rejected is replaced by the first good one
            bad = torch.nonzero(~ok).flatten() + i * ncol
            good = int(torch.nonzero(ok).flatten()[0]) + i * ncol
            X[bad] = X[good]
            T[bad] = T[good]
        if log:
            log("training set: cube %d/%d (%d columns)" % (i + 1, ncubes, ncol))
    torch.cuda.synchronize()
    return X, T


@dataclass
class TrainReport:
    epochs: int
    steps: int
    seconds: float
    train_loss: List[float] = field(default_factory=list)
    valid_loss: List[float] = field(default_factory=list)
    valid_loss_initial: float = float("nan")


def train(ctx, X, T, epochs: int = 8, batch: int = 1024, lr: float = 1e-3, lr_min: float = 2e-5,
          valid_frac: float = 0.02, seed: int = 0, log=None) -> TrainReport:
    """Adam on the device network over (X, T); the last valid_frac of the rows
    is held out for the validation loss (mean squared error on standardised
    targets)."""
    import torch
    n = X.shape[0]
    nv = max(1, int(n * valid_frac))
    nt = n - nv
    ctx.mlp_fit_norms_torch(X[:nt], T[:nt])
    gen = torch.Generator(device="cuda")
    gen.manual_seed(seed)
    loss = torch.zeros(1, dtype=torch.float32, device="cuda")
    steps_per_epoch = nt // batch
    total = steps_per_epoch * epochs
    rep = TrainReport(epochs, total, 0.0)
    vidx = torch.arange(nt, n, dtype=torch.int32, device="cuda")
    rep.valid_loss_initial = _valid_loss(ctx, X, T, vidx)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    step = 0
    for ep in range(epochs):
        perm = torch.randperm(nt, generator=gen, device="cuda").to(torch.int32)
        acc = torch.zeros(1, dtype=torch.float64, device="cuda")
        for s in range(steps_per_epoch):
            lr_t = lr_min + 0.5 * (lr - lr_min) * (1.0 + math.cos(math.pi * step / max(1, total)))
            ctx.mlp_train_step_torch(X, T, perm[s * batch:(s + 1) * batch], lr_t, loss)
            acc += loss
            step += 1
        vl = _valid_loss(ctx, X, T, vidx)
        rep.train_loss.append(float(acc.item()) / max(1, steps_per_epoch))
        rep.valid_loss.append(vl)
        if log:
            log("epoch %d: train %.4f valid %.4f" % (ep + 1, rep.train_loss[-1], vl))
    torch.cuda.synchronize()
    rep.seconds = time.perf_counter() - t0
    return rep


def _valid_loss(ctx, X, T, idx, chunk: int = 8192) -> float:
    tot, cnt = 0.0, 0
    for i in range(0, idx.shape[0], chunk):
        sl = idx[i:i + chunk]
        loss, _ = ctx.mlp_gradient_torch(X, T, sl)
        tot += float(loss.item()) * sl.shape[0]
        cnt += sl.shape[0]
    return tot / max(1, cnt)


@dataclass
class InvertReport:
    iterations: int
    fallback_pixels: int
    fallback_iterations: int
    seconds_init: float
    seconds_lm: float
    seconds_fallback: float


def invert(ctx, init, obs, control, nx: int, ny: int, threshold: Optional[float], learned: bool = True,
           final_cycle_only: bool = True):
    """INVERT3D from the learned initial model (learned=True) or from init.

    init: [12][nz][ncol] default initial model (not modified); obs [ncol][NFREQ]
    float64.  The proposal is already on the final node basis, so by default
    its inversion runs only the last cycle (the coarse cycles would first
    contract it to fewer nodes).  Pixels with a final chi^2 above threshold
    are inverted again from init with the full schedule, without the spatial
    smoothing (they are no longer neighbours), and keep the lower chi^2.
    Returns (model, stokes, chi2, InvertReport)."""
    import torch
    ncol = nx * ny
    t0 = time.perf_counter()
    atm = init.clone()
    ctl = control
    if learned and final_cycle_only:
        ctl = _gpu.CInvertControl.from_buffer_copy(control)
        ctl.first_cycle = 1 << 20
    if learned:
        ctx.learned_init_torch(obs, atm)
        st = ctx.hydrostatic_torch(atm)
        bad = st != 0
        if bool(bad.any()):
            atm[:, :, bad] = init[:, :, bad]
            ctx.hydrostatic_torch(atm)
    torch.cuda.synchronize()
    t1 = time.perf_counter()
    stokes, chi2, its = ctx.invert3d_torch(atm, obs, ctl, nx, ny)
    torch.cuda.synchronize()
    t2 = time.perf_counter()
    nfb, its_fb = 0, 0
    if learned and threshold is not None:
        sel = torch.nonzero(chi2 > threshold).flatten()
        nfb = int(sel.numel())
        if nfb:
            sub = init[:, :, sel].contiguous()
            ctx.hydrostatic_torch(sub)
            ctl = _gpu.CInvertControl.from_buffer_copy(control)
            ctl.smoothing = 0
            s2, c2, its_fb = ctx.invert3d_torch(sub, obs[sel].contiguous(), ctl, nfb, 1)
            better = c2 < chi2[sel]
            keep = sel[better]
            atm[:, :, keep] = sub[:, :, better]
            stokes[keep] = s2[better]
            chi2[keep] = c2[better]
    torch.cuda.synchronize()
    t3 = time.perf_counter()
    return atm, stokes, chi2, InvertReport(its, nfb, its_fb, t1 - t0, t2 - t1, t3 - t2)
