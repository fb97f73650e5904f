"""Restatement of the reference's inversion cycle driver (INVERT3D,
invert3d.f90:57-403) over the oracle's per-column forward and LM step.

TEST INFRASTRUCTURE (the oracle): only tests/ use it, as the checker.  The
arithmetic lives in fzcpu.c / fzlm.c; this module keeps the schedule:
  * MAXITER from the parameters' maximum nodes (pre_post_duties.f90:635-646),
    nodes per cycle min(GET_NJEVALS, NSLB_MAX) (get_dmodel.f90:375-452);
  * INITIALIZE_CYCLE (invert3d.f90:228-325): INV_MAS reset, SET_TAU
    synthesis, best := current, SMOOTHING_B (model_spatial_extension.f90:
    351-491: 1-2-1 filter along x, y, z of every inverted parameter, FP64);
  * iterations: forward with response functions, node contraction, LM step
    per pixel, MAKE_DECISION (invert3d.f90:820-859);
  * END_CYCLE (:328-403): current := best and a synthesis.
Columns are numbered col = y + ny * x; atm is [12][nz][ncol] float32.
"""
from __future__ import annotations

import math

import numpy as np

from . import lm_cfg

ROWOF = (0, 1, 2, 5, 6, 7, 10, 1)


def njevals(apar, curic, maxiter, cycpow, nz, toffset=1):
    if apar == 1:
        if curic == 0 or curic == maxiter:
            return nz
        return min(nz, int(cycpow ** (curic + toffset - 1)))
    if curic in (0,):
        return nz
    if curic == 1:
        return 1
    if curic == maxiter:
        return nz
    return min(nz, int(cycpow ** (curic - 1)))


def weights(invstk, wstk, noise, numw):
    """ISIGMAP: REAL(4) normalised weights over REAL(8) ISIGMA**2."""
    w = np.array([wstk[s] if invstk[s] else 0.0 for s in range(4)], np.float32)
    nstk = int(sum(1 for s in range(4) if invstk[s]))
    minw = np.sqrt(np.float32(np.sum(w * w, dtype=np.float32))) / np.sqrt(np.float32(nstk))
    isigma = float(np.float32(noise))
    out = []
    for s in range(4):
        if invstk[s]:
            ws = np.float32(w[s] / minw)
            out.append(np.full(numw, float(np.float32(ws * ws)) / (isigma * isigma)))
    return np.concatenate(out)


def auto_weights(obs, invstk, noise, numw):
    """SET_WEIGHTS3D (get_dmodel.f90:2051-2170) on obs [ncol, NFREQ]: per pixel,
    SET_WEIGHTS1D (:2191-2224) -- weight of inverted block s = SQRT(1/STD**2),
    STD from STDDEV (:2228-2251: mean as the running sum of x_i/n, n - 1
    variance); weights above 10 x the smallest cut to it; all divided by
    SQRT(SUM w**2)/SQRT(NSTKINV) -- then WSTK = REAL(4) of them and SET_ISIGMAP
    (:356-373): WSTK**2 (REAL(4)) / ISIGMA**2.  Returns [ncol, NFREQ]."""
    obs = np.asarray(obs, np.float64)
    isigma = float(np.float32(noise))
    nstk = int(sum(1 for s in range(4) if invstk[s]))
    out = np.zeros_like(obs)
    for c in range(obs.shape[0]):
        w = [0.0] * 4
        minw = 1.0e60
        b = 0
        for s in range(4):
            if not invstk[s]:
                continue
            x = obs[c, b * numw:(b + 1) * numw]
            mn = 0.0
            for v in x:
                mn += float(v) / float(numw)
            var = 0.0
            for v in x:
                var += (float(v) - mn) * (float(v) - mn)
            sd = math.sqrt(var / (numw - 1.0))
            w[s] = math.sqrt(1.0 / (sd * sd)) if sd > 0 else math.inf
            minw = min(minw, w[s])
            b += 1
        minw *= 10.0
        w = [min(v, minw) for v in w]
        nrm = math.sqrt(sum(v * v for v in w)) / math.sqrt(nstk)
        b = 0
        for s in range(4):
            if invstk[s]:
                ws = np.float32(w[s] / nrm)
                out[c, b * numw:(b + 1) * numw] = float(np.float32(ws * ws)) / (isigma * isigma)
                b += 1
    return out


def smooth(a, nx, ny):
    """SMOOTHING_B on one parameter: a [nz][ncol] float32 -> float32."""
    nz = a.shape[0]
    v = a.astype(np.float64).reshape(nz, nx, ny)       # col = y + ny*x
    if nx > 4:
        w = v.copy()
        w[:, 1:-1, :] = (v[:, 2:, :] + v[:, :-2, :] + v[:, 1:-1, :] * 2.0) / 4.0
        v = w
    if ny > 4:
        w = v.copy()
        w[:, :, 1:-1] = (v[:, :, 2:] + v[:, :, :-2] + v[:, :, 1:-1] * 2.0) / 4.0
        v = w
    if nz > 4:
        w = v.copy()
        w[1:-1] = (v[2:] + v[:-2] + v[1:-1] * 2.0) / 4.0
        v = w
    return v.reshape(nz, nx * ny).astype(np.float32)


def invert3d(cpu, atm, z, obs, nx, ny, max_nodes, invstk=(1, 1, 1, 1), wstk=(1, 1, 1, 1), noise=4e-3, cycpow=2,
             maxsteps=5, solver=1, node_expansion=1, svdtol=1e-4, smoothing=True, max_iterations=1000,
             first_cycle=1, regul=False, pen_typ=None, pen_alp=None, auto=False):
    """cpu: oracle.Cpu set up with resp=True and all Cartesian slots (and
    set_hydrostatic for HYDROSTATIC: yes).  max_nodes: {physical index 0..6:
    NSLB_MAX}; pen_typ / pen_alp: {physical index: PEN_TYP / PEN_ALP} of the
    Tikhonov terms when regul; auto: STOKES SETUP AUTO (per-pixel weights of
    auto_weights in place of wstk).  Returns (model, stokes [ncol,4,numw], chi2
    [ncol], iterations)."""
    cur = np.array(atm, np.float32, copy=True)
    nz, ncol = cur.shape[1], cur.shape[2]
    numw = cpu.numw
    atmpar = [1 if p in max_nodes else 0 for p in range(8)]
    vmax = {p: min(max_nodes[p], nz) for p in max_nodes}
    nfreev, nfreep = len(vmax), sum(vmax.values())
    maxiter = 1 if nfreep == nfreev else int(math.floor(math.log(max(vmax.values())) / math.log(cycpow)))
    isig = weights(invstk, wstk, noise, numw)
    nfreq = len(isig)
    isig_pix = auto_weights(obs, invstk, noise, numw) if auto else [isig] * ncol
    rows = [s for s in range(4) if invstk[s]]
    total = 0

    def forward_all(model, resp):
        syn_all = np.zeros((ncol, 4, numw))
        dsyn_all = []
        for c in range(ncol):
            m = np.zeros((13, nz), np.float32)
            m[:12] = model[:, :, c]
            m[12] = np.asarray(z, np.float32)
            mo, syn, dsyn, _ = cpu.forward1d(m)
            model[:, :, c] = mo[:12]
            syn_all[c] = syn
            dsyn_all.append(dsyn if resp else None)
        return syn_all, dsyn_all

    sla = None
    syn = None
    for cyc in range(min(max(1, first_cycle), maxiter), maxiter + 1):
        nodes = [0] * 8
        for p in vmax:
            nodes[p] = min(njevals(p + 1, cyc, maxiter, cycpow, nz), vmax[p])
        ptyp, palp = [0] * 8, [0.0] * 8
        for p, v in (pen_typ or {}).items():
            ptyp[p] = int(v)
        for p, v in (pen_alp or {}).items():
            palp[p] = float(v)
        cfg = lm_cfg(nz, numw, atmpar, nodes, invstk=invstk, solver=solver, node_expansion=node_expansion,
                     svdtol=svdtol, inu=nfreq, regul=regul, pen_typ=ptyp, pen_alp=palp, curic=cyc)
        coefs = cpu.coefs_block(cfg, {p: cpu.int_coefs(nz, nodes[p]) for p in vmax})
        np_ = coefs.shape[0]
        sla = np.tile(np.array([0, 1e29, 1e29, 0, 0, 0, 0, 0, 10.0]), (ncol, 1))
        done = np.zeros(ncol, bool)
        if not atmpar[1] and not atmpar[7]:
            forward_all(cur, False)                       # SET_TAU
        best = cur.copy()
        if smoothing:
            for p in vmax:
                r = ROWOF[p]
                best[r] = smooth(best[r], nx, ny)
                cur[r] = best[r]
        bsyn = np.zeros((ncol, nfreq))
        ebest = np.zeros((ncol, nfreq, np_), np.float32)
        need = max(1, maxsteps - (maxiter - cyc))
        for it in range(max_iterations):
            syn, dsyn = forward_all(cur, True)
            for c in range(ncol):
                if done[c]:
                    continue
                eq = cpu.eqv_dsyn(cfg, coefs, dsyn[c])
                mc = np.zeros((13, nz), np.float32)
                mc[:12] = cur[:, :, c]
                mc[12] = z
                mb = mc.copy()
                mb[:12] = best[:, :, c]
                s_rows = syn[c][rows].reshape(-1)
                mc2, mb2, sl, bs, jb, _ = cpu.lm_step(cfg, coefs, obs[c], s_rows, bsyn[c], sla[c], eq, ebest[c],
                                                      isig_pix[c], mc, mb)
                cur[:, :, c] = mc2[:12]
                best[:, :, c] = mb2[:12]
                sla[c], bsyn[c], ebest[c] = sl, bs, jb
                if sl[0] > maxsteps:
                    done[c] = True
            total += 1
            if ncol - done.sum() <= 0 or sla[:, 0].sum() / ncol >= need:
                break
        cur = best.copy()
        syn, _ = forward_all(cur, False)
    return cur, syn, sla[:, 2].copy(), total


# SET_TAU's assist initialisers (invert3d.f90:528-700)
ASSIST_B, ASSIST_V, ASSIST_T, ASSIST_P = 1, 2, 4, 8
KBOL, HPLA, LIGHT = 1.38056503e-16, 6.62606957e-27, 2.99792458e10   # cons_param.f90:18-25


def assist(obs, atm, flags, numw, itz, nstk, lambda0):
    """The assists on a start model: obs (ncol, nfreq) float64 (the inverted
    Stokes blocks of NUMW each), atm (12, nz, ncol) float32 (a modified copy
    is returned); itz wavelengths of the first region, its main line at
    lambda0 (A).  Steps and REAL(4) roundings of invert3d.f90:528-700."""
    atm = np.array(atm, dtype=np.float32, copy=True)
    nz, ncol = atm.shape[1], atm.shape[2]
    f32 = np.float32
    first = obs[:, :itz]

    def norm(v):
        return v / np.sum(np.abs(v))

    def cdiff(v):
        d = np.empty_like(v)
        d[1:-1] = (v[2:] - v[:-2]) / 2.0
        d[0], d[-1] = d[1], d[-2]
        return d

    if flags & (ASSIST_B | ASSIST_V):
        v0 = norm(first.sum(axis=0))
        v1 = norm(cdiff(v0))
        v2 = norm(cdiff(v1))
    kz = int(math.ceil(float(f32(nz)) / 3.0)) - 1          # CEILING(REAL(NZ)/3.)
    if flags & ASSIST_B and nstk == 4:
        t3 = obs[:, 3 * numw:3 * numw + itz] @ v1
        t1 = obs[:, numw:numw + itz] @ v2
        t2 = obs[:, 2 * numw:2 * numw + itz] @ v2
        t3 = t3 / np.max(np.abs(t3))
        t4 = np.sqrt(t1 ** 2 + t2 ** 2)
        t4 = t4 / np.max(t4)
        atn2 = np.arctan2(t2, t1) / 2.0
        bx = ((-np.sin(atn2).astype(f32) * f32(1000.0)).astype(np.float64) * t4).astype(f32)
        by = ((np.cos(atn2).astype(f32) * f32(1000.0)).astype(np.float64) * t4).astype(f32)
        bz = t3.astype(f32) * f32(-1000.0)
        atm[5], atm[6], atm[7] = bx[None, :], by[None, :], bz[None, :]
    if flags & ASSIST_V:
        kol = float(v1 @ v0)
        tv = first @ v1 - (first @ v0) * kol
        tv = tv / np.max(np.abs(tv))
        atm[10] = (tv.astype(f32) * f32(-2.0e5))[None, :]
    if flags & (ASSIST_T | ASSIST_P):
        atm[11] = 1.0
        atm[11, kz] = 0.0
    if flags & ASSIST_T:
        lk = (lambda0 * 1.0e-8 * KBOL) / (HPLA * LIGHT)
        mean = first.sum(axis=1) / float(itz)
        tb = 1.0 / (lk * np.log(1.0 + 1.0 / mean * (np.exp(1.0 / 6.39e3 / lk) - 1.0)))
        dt = tb - atm[0, kz].astype(np.float64)
        atm[0] = (atm[0].astype(np.float64) + dt[None, :]).astype(f32)
    if flags & ASSIST_P:
        pz = atm[1, kz].copy()
        atm[1] = atm[1] * f32(1.333e5) / pz[None, :]
    return atm
