"""The PSF-coupled inversion cycle, restated over the oracle's forward and LM
step (numpy driver, FP64).

TEST INFRASTRUCTURE (the oracle): only tests/ use it, as the checker.

The reference allocates the coupled inversion's buffers
(coupled_inversion.f90:10-207) but never wires its mathematics
(GET_DMODEL3DC is a stub, get_dmodel.f90:1707-1711), so this restates the
design the device implements, from its definition:
  * chi^2 of the map = mean over pixels of sum_f W_f (obs - PSF * syn)^2 / NFREQ,
    the convolution that of oracle/psf.py on every (Stokes row, wavelength) map;
  * acceptance by the map's chi^2 and UPDATE_LAMBDA's rules
    (get_dmodel.f90:917-956) on the map (lambda / sqrt(10) or x 10, x 3
    never taken here, floor 1e-2, the 1 % step-count rule);
  * the step of pixel q from the gradient J_q^T W (PSF^T r)_q and the matrix
    (sum PSF^2) J_q^T W J_q: the oracle's GET_PERT_1D (fzc_lm_step) fed
    J' = sqrt(h) J and r' = (PSF^T r) / sqrt(h) with h = sum PSF^2, so that
    J'^T J' = h J^T J and J'^T r' = J^T PSF^T r, with acceptance forced and
    INV_SLA(9) = lambda sqrt(10) (the accepted branch divides it back);
  * the cycle schedule, SET_TAU and the smoothing of oracle/invert.py, and
    MAKE_DECISION's stop rule (invert3d.f90:820-859) on the map's step
    count: a cycle ends when it reaches MAX(1, MAXSTEPS - (MAXITER - cycle)).
Columns are numbered col = y + ny * x; atm is [12][nz][ncol] float32.
"""
from __future__ import annotations

import math

import numpy as np

from . import lm_cfg
from . import psf as P
from .invert import ROWOF, njevals, smooth, weights

_TRACE = False


def coupled_invert3d(cpu, atm, z, obs, nx, ny, max_nodes, psf, invstk=(1, 1, 1, 1), wstk=(1, 1, 1, 1), noise=4e-3,
                     cycpow=2, maxsteps=5, solver=1, node_expansion=1, svdtol=1e-4, smoothing=True,
                     max_iterations=1000, first_cycle=1):
    cur = np.array(atm, np.float32, copy=True)
    nz, ncol = cur.shape[1], cur.shape[2]
    numw = cpu.numw
    psf32 = np.asarray(psf, np.float32).astype(np.float64)
    h = float(np.sum(psf32 * psf32))
    atmpar = [1 if p in max_nodes else 0 for p in range(8)]
    vmax = {p: min(max_nodes[p], nz) for p in max_nodes}
    nfreev, nfreep = len(vmax), sum(vmax.values())
    maxiter = 1 if nfreep == nfreev else int(math.floor(math.log(max(vmax.values())) / math.log(cycpow)))
    isig = weights(invstk, wstk, noise, numw)
    nfreq = len(isig)
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

    def conv_rows(syn_all, adjoint=False):
        """[ncol][nfreq] rows of the inverted Stokes parameters -> PSF (or its
        adjoint) applied to every row's map."""
        maps = syn_all.T.reshape(nfreq, nx, ny)          # row f of pixel col = y + ny x
        return P.convolve(maps, psf32, adjoint).reshape(nfreq, ncol).T

    def chi_map(model):
        syn_all, dsyn = forward_all(model, True)
        rows_syn = np.concatenate([syn_all[:, s, :] for s in rows], axis=1)
        r = obs - conv_rows(rows_syn)
        chi_pix = np.sum(r * r * isig, axis=1) / nfreq
        return chi_pix.mean(), chi_pix, r, syn_all, dsyn

    best = cur.copy()
    chi_pix = None
    for cyc in range(min(max(1, first_cycle), maxiter), maxiter + 1):
        nodes = [0] * 8
        for p in vmax:
            nodes[p] = min(njevals(p + 1, cyc, maxiter, cycpow, nz), vmax[p])
        cfg = lm_cfg(nz, numw, atmpar, nodes, invstk=invstk, solver=solver, node_expansion=node_expansion,
                     svdtol=svdtol, inu=nfreq)
        coefs = cpu.coefs_block(cfg, {p: cpu.int_coefs(nz, nodes[p]) for p in vmax})
        if not atmpar[1] and not atmpar[7]:
            forward_all(cur, False)                       # SET_TAU
        if smoothing:
            for p in vmax:
                r_ = ROWOF[p]
                cur[r_] = smooth(cur[r_], nx, ny)
        chibest, lam, steps, precyc = 1.0e29, 10.0, 0.0, 0.0
        eqvb = resb = None
        for it in range(max_iterations):
            chi, chi_pix, r, syn_all, dsyn = chi_map(cur)
            total += 1
            if _TRACE:
                print("  it %d chi %.4f best %.4f lam %.3g steps %d" % (it, chi, chibest, lam, steps))
            if chi < chibest:
                if (chibest - chi) > 0.01 * chibest:
                    steps = max(0.0, steps - 1.0)
                else:
                    steps = steps + 1.0
                chibest = chi
                lam = lam / math.sqrt(10.0)
                best = cur.copy()
                eqvb = [cpu.eqv_dsyn(cfg, coefs, dsyn[c]) for c in range(ncol)]
                resb = r
            else:
                if abs(precyc - 1.0) < 1e-2:
                    precyc = 0.0
                    lam = lam * 3.0
                else:
                    lam = lam * 10.0
                steps = steps + 1.0
                cur = best.copy()
            lam = max(1e-2, lam)
            if steps >= max(1, maxsteps - (maxiter - cyc)):    # MAKE_DECISION on the map
                break
            gres = conv_rows(resb, adjoint=True)
            sq = math.sqrt(h)
            for c in range(ncol):
                mc = np.zeros((13, nz), np.float32)
                mc[:12] = cur[:, :, c]
                mc[12] = z
                sla = np.array([0, 1e29, 1e30, 0, 0, 0, 0, 0, lam * math.sqrt(10.0)])
                jac = (eqvb[c].astype(np.float64) * sq).astype(np.float32)
                mc2, _, _, _, _, _ = cpu.lm_step(cfg, coefs, gres[c] / sq, np.zeros(nfreq), np.zeros(nfreq), sla, jac,
                                                 jac.copy(), isig, mc, mc.copy())
                cur[:, :, c] = mc2[:12]
        cur = best.copy()
    chi, chi_pix, r, syn_all, _ = chi_map(cur)
    return cur, chi_pix, total
