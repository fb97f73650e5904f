"""The oracle: a CPU restatement of the FIRTEZ-dz column path, and (where it
was built) the reference's own Fortran modules behind a small harness.

TEST INFRASTRUCTURE.  Only tests/, __graft_entry__.smoke() and the
CPU-baseline leg of bench.py import this package, always as the checker.
The product (firtez_dz_amd) never imports it.

  Cpu   -> oracle/libfzcpu.so  (fzcpu.c, FP64, OpenMP over columns)
  Ref   -> oracle/_ref/libfzref.so (reference Fortran + harness; built only
           where the reference tree exists: make -C oracle/ref)
"""
from __future__ import annotations

import ctypes
import os
import subprocess

import numpy as np

from firtez_dz_amd import lines as _lines

HERE = os.path.dirname(os.path.abspath(__file__))
CPU_LIB = os.path.join(HERE, "libfzcpu.so")
REF_LIB = os.path.join(HERE, "_ref", "libfzref.so")

_dp = ctypes.POINTER(ctypes.c_double)
_fp = ctypes.POINTER(ctypes.c_float)
_ip = ctypes.POINTER(ctypes.c_int32)


def _ptr(a, t):
    return a.ctypes.data_as(t)


def build_cpu():
    subprocess.check_call(["make", "-s", "-C", HERE])


def build_ref():
    """Build oracle/_ref from the reference sources (only where they exist)."""
    src = os.environ.get("FZ_REFERENCE", "/root/reference") + "/src"
    if not os.path.isdir(src):
        return False
    subprocess.check_call(["make", "-s", "-C", os.path.join(HERE, "ref"), "REF=" + src])
    return True


def ref_available() -> bool:
    return os.path.exists(REF_LIB)


class LmCfg(ctypes.Structure):
    """fzc_lmcfg (oracle/fzcpu.h): per-pixel inversion configuration."""
    _fields_ = [("nz", ctypes.c_int32), ("numw", ctypes.c_int32),
                ("atmpar", ctypes.c_int32 * 8), ("invstk", ctypes.c_int32 * 4),
                ("nodes", ctypes.c_int32 * 8), ("pen_typ", ctypes.c_int32 * 8),
                ("pen_alp", ctypes.c_double * 8),
                ("regul", ctypes.c_int32), ("solver", ctypes.c_int32),
                ("node_expansion", ctypes.c_int32), ("curic", ctypes.c_int32),
                ("svdtol", ctypes.c_double), ("inu", ctypes.c_double), ("mpert", ctypes.c_double)]


def lm_cfg(nz, numw, atmpar, nodes, invstk=(1, 1, 1, 1), solver=1, node_expansion=1, svdtol=1e-4,
           regul=False, pen_typ=(0,) * 8, pen_alp=(0.0,) * 8, inu=None, mpert=0.15, curic=1):
    c = LmCfg()
    c.nz, c.numw = nz, numw
    c.atmpar[:] = list(atmpar)
    c.invstk[:] = list(invstk)
    c.nodes[:] = [n if a else 0 for n, a in zip(nodes, atmpar)]
    c.pen_typ[:] = list(pen_typ)
    c.pen_alp[:] = [float(np.float32(a)) for a in pen_alp]   # PEN_ALP is REAL(SP)
    c.regul, c.solver, c.node_expansion, c.curic = int(regul), solver, node_expansion, curic
    c.svdtol = svdtol
    c.inu = float(sum(invstk) * numw) if inu is None else float(inu)
    c.mpert = mpert
    return c


class Cpu:
    """ctypes view of fzcpu.c."""

    def __init__(self):
        if not os.path.exists(CPU_LIB):
            build_cpu()
        self.lib = ctypes.CDLL(CPU_LIB)
        L = self.lib
        L.fzc_setup.restype = ctypes.c_int
        L.fzc_partition.argtypes = [ctypes.c_int, ctypes.c_double, _dp, _dp]
        L.fzc_eos.argtypes = [ctypes.c_float, ctypes.c_float, _dp, _dp, _dp]
        L.fzc_opac_cont.argtypes = [ctypes.c_float, ctypes.c_double, ctypes.c_double, ctypes.c_double, _dp, _dp]
        L.fzc_opac_line.argtypes = [ctypes.c_int, ctypes.c_float, ctypes.c_double, ctypes.c_double, _dp, _dp]
        L.fzc_damping.argtypes = [ctypes.c_int, ctypes.c_float, ctypes.c_double, _dp, _dp]
        L.fzc_voigt.argtypes = [ctypes.c_int, ctypes.c_double, _dp, _dp, _dp]
        L.fzc_zeeman.argtypes = [ctypes.c_int, _dp, _dp, _ip]
        L.fzc_gain.argtypes = [_fp]
        L.fzc_set_gain.argtypes = [_fp]
        L.fzc_forward1d.argtypes = [ctypes.c_int, _fp, _dp, _dp, _dp]
        L.fzc_forward_batch.argtypes = [ctypes.c_int, ctypes.c_int, _fp, _dp, _dp, ctypes.c_int]
        L.fzc_he.argtypes = [ctypes.c_int, _fp, _fp, _fp, _fp, _fp, _dp, _dp]
        L.fzc_set_hydrostatic.argtypes = [ctypes.c_int]
        L.fzc_set_microturbulence.argtypes = [ctypes.c_double]
        L.fzc_set_full_stokes.argtypes = [ctypes.c_void_p]
        L.fzc_set_line_tau.argtypes = [ctypes.c_void_p]
        L.fzc_set_lsf.argtypes = [ctypes.c_void_p, ctypes.c_void_p]
        L.fzc_set_departure.argtypes = [ctypes.c_int, ctypes.c_int, ctypes.c_void_p, ctypes.c_void_p]
        L.fzc_evol.argtypes = [_dp, ctypes.c_double, _dp, _dp, _dp]
        L.fzc_int_coefs.argtypes = [ctypes.c_int, ctypes.c_int, _dp]
        L.fzc_njevals.argtypes = [ctypes.c_int] * 6
        L.fzc_eqv_dsyn.argtypes = [ctypes.POINTER(LmCfg), _dp, _dp, _fp]
        L.fzc_get_perturbation.argtypes = [ctypes.c_int, ctypes.c_int, _dp, _dp, ctypes.c_double, _dp, _dp, _dp,
                                           ctypes.c_int, ctypes.c_double]
        L.fzc_lm_step.argtypes = [ctypes.POINTER(LmCfg), _dp, _dp, _dp, _dp, _dp, _fp, _fp, _dp, _fp, _dp]
        self.numw = 0
        self.resp = False

    def setup(self, ls: _lines.LineSetup, resp=True, atmpar=(1, 1, 0, 1, 1, 1, 1, 0)):
        cl, cr = _lines.c_arrays(ls)
        ap = (ctypes.c_int32 * 8)(*atmpar)
        n = self.lib.fzc_setup(cl, len(ls.lines), cr, len(ls.regions), int(resp), ap)
        if n < 0:
            raise RuntimeError("fzc_setup failed %d" % n)
        self.numw, self.resp, self.atmpar = n, bool(resp), tuple(atmpar)
        self.nreg = len(ls.regions)
        return n

    def partition(self, elem, T):
        u = np.zeros(3)
        du = np.zeros(3)
        self.lib.fzc_partition(elem, T, _ptr(u, _dp), _ptr(du, _dp))
        return u, du

    def eos(self, T, P, ne=0.0):
        ne_ = ctypes.c_double(ne)
        out = np.zeros(5)
        d = np.zeros(25)
        e = self.lib.fzc_eos(T, P, ctypes.byref(ne_), _ptr(out, _dp), _ptr(d, _dp))
        return e, out, d

    def opac_cont(self, T, lam, ne, nh, d):
        kc = ctypes.c_double()
        d = np.array(d, dtype=np.float64)
        e = self.lib.fzc_opac_cont(T, lam, ne, nh, ctypes.byref(kc), _ptr(d, _dp))
        return e, kc.value, d

    def opac_line(self, line, T, ne, nh, d):
        kl = ctypes.c_double()
        d = np.array(d, dtype=np.float64)
        e = self.lib.fzc_opac_line(line, T, ne, nh, ctypes.byref(kl), _ptr(d, _dp))
        return e, kl.value, d

    def damping(self, line, T, nh, d):
        dm = ctypes.c_double()
        d = np.array(d, dtype=np.float64)
        self.lib.fzc_damping(line, T, nh, ctypes.byref(dm), _ptr(d, _dp))
        return dm.value, d

    def voigt(self, a, v):
        v = np.ascontiguousarray(v, dtype=np.float64)
        h = np.zeros_like(v)
        f = np.zeros_like(v)
        self.lib.fzc_voigt(len(v), a, _ptr(v, _dp), _ptr(h, _dp), _ptr(f, _dp))
        return h, f

    def zeeman(self, line):
        s = np.zeros(64)
        st = np.zeros(64)
        ty = np.zeros(64, dtype=np.int32)
        n = self.lib.fzc_zeeman(line, _ptr(s, _dp), _ptr(st, _dp), _ptr(ty, _ip))
        return s[:n], st[:n], ty[:n]

    def evol(self, x7, dz_km, dx=None):
        """Evolution operator (4x4) and its derivatives along dx (8 x 7)."""
        x7 = np.ascontiguousarray(x7, dtype=np.float64)
        dx = np.zeros((8, 7)) if dx is None else np.ascontiguousarray(dx, dtype=np.float64)
        ev = np.zeros(16)
        dev = np.zeros(8 * 16)
        e = self.lib.fzc_evol(_ptr(x7, _dp), dz_km, _ptr(dx, _dp), _ptr(ev, _dp), _ptr(dev, _dp))
        if e:
            raise RuntimeError("fzc_evol error %d" % e)
        return ev.reshape(4, 4), dev.reshape(8, 4, 4)

    def gain(self):
        g = np.zeros((self.numw, 4), dtype=np.float32)
        e = self.lib.fzc_gain(_ptr(g, _fp))
        if e:
            raise RuntimeError("fzc_gain error %d" % e)
        return g.T.copy()            # (4, numw)

    def set_gain(self, g):
        g = np.ascontiguousarray(np.asarray(g, dtype=np.float32).T)
        self.lib.fzc_set_gain(_ptr(g, _fp))

    def set_departure(self, popl=None, popu=None):
        """NLTE departure coefficients POPL, POPU as arrays (nreg, nz)
        (input_file.f90:2146-2166); None: LTE."""
        if popl is None:
            self.lib.fzc_set_departure(0, 0, None, None)
            return
        pl = np.asfortranarray(popl, dtype=np.float64)
        pu = np.asfortranarray(popu, dtype=np.float64)
        self.lib.fzc_set_departure(pl.shape[0], pl.shape[1], pl.ctypes.data, pu.ctypes.data)

    def set_hydrostatic(self, on=True):
        """HYDROSTATIC: yes -- Pg re-integrated from the top inside forward1d
        and the hydrostatic term in dI/dT."""
        self.lib.fzc_set_hydrostatic(int(on))

    def set_microturbulence(self, xi_cm_s=0.0):
        """Doppler width sqrt(2kT/m + xi^2) for every line (0: thermal only)."""
        self.lib.fzc_set_microturbulence(float(xi_cm_s))

    def forward1d_full(self, model):
        """forward1d plus FULL_STOKES: (model, syn, dsyn, kc5, fs [nz][4][numw])."""
        nz = model.shape[1]
        fs = np.zeros((nz, 4, self.numw))
        self.lib.fzc_set_full_stokes(fs.ctypes.data)
        try:
            out = self.forward1d(model)
        finally:
            self.lib.fzc_set_full_stokes(None)
        return out + (fs,)

    def forward1d_line_tau(self, model):
        """forward1d plus LINE OPTICAL DEPTH: (model, syn, dsyn, kc5, log10 tau [nz][numw])."""
        nz = model.shape[1]
        tl = np.zeros((nz, self.numw))
        self.lib.fzc_set_line_tau(tl.ctypes.data)
        try:
            out = self.forward1d(model)
        finally:
            self.lib.fzc_set_line_tau(None)
        return out + (tl,)

    def set_lsf(self, lsf_kernel=None, valid=None):
        """MLSF: LSF_KERNEL(NUMW) (region taps YK/SKI, tap 0 first) and
        LSF_VALID(NUML); None switches the convolution off."""
        if lsf_kernel is None:
            self._lsf = None
            return self.lib.fzc_set_lsf(None, None)
        k = np.ascontiguousarray(lsf_kernel, dtype=np.float64)
        v = np.ascontiguousarray(np.ones(self.nreg, np.int32) if valid is None else valid, dtype=np.int32)
        self._lsf = (k, v)
        r = self.lib.fzc_set_lsf(k.ctypes.data, v.ctypes.data)
        if r:
            raise ValueError("fzc_set_lsf: region longer than 4096 points")
        return r

    def forward1d(self, model):
        """model: (13, nz) float32 [T,Pg,rho,Pe,mu,Bx,By,Bz,vx,vy,vz,tau5,z];
        returns (model_out, syn (4,numw), dsyn (nz,numw,8,4) or None, kc5)."""
        m = np.ascontiguousarray(model, dtype=np.float32).copy()
        nz = m.shape[1]
        syn = np.zeros((self.numw, 4))
        dsyn = np.zeros((nz, self.numw, 8, 4)) if self.resp else np.zeros(1)
        kc5 = np.zeros(nz)
        e = self.lib.fzc_forward1d(nz, _ptr(m, _fp), _ptr(syn, _dp), _ptr(dsyn, _dp), _ptr(kc5, _dp))
        if e:
            raise RuntimeError("fzc_forward1d error %d" % e)
        return m, syn.T.copy(), (dsyn if self.resp else None), kc5

    def forward_batch(self, models, nthreads=0):
        """models: (ncol, 13, nz) float32 (updated copy returned)."""
        m = np.ascontiguousarray(models, dtype=np.float32).copy()
        ncol, _, nz = m.shape
        syn = np.zeros((ncol, self.numw, 4))
        dsyn = np.zeros((ncol, nz, self.numw, 8, 4)) if self.resp else None
        e = self.lib.fzc_forward_batch(ncol, nz, _ptr(m, _fp), _ptr(syn, _dp),
                                       _ptr(dsyn, _dp) if dsyn is not None else None, nthreads)
        if e:
            raise RuntimeError("fzc_forward_batch error %d" % e)
        return m, syn, dsyn

    # ---------------------------------------------------------------- LM step
    def int_coefs(self, nz, nj):
        """COEFS (nz, nj) of GET_INT_COEFS."""
        c = np.zeros((nj, nz))
        e = self.lib.fzc_int_coefs(nz, nj, _ptr(c, _dp))
        if e:
            raise RuntimeError("fzc_int_coefs error %d" % e)
        return c.T.copy()

    def njevals(self, apar, curic, maxiter, cycpow=2, toffset=1, nz=64):
        return self.lib.fzc_njevals(apar, curic, maxiter, cycpow, toffset, nz)

    @staticmethod
    def coefs_block(cfg, coefs_by_param):
        """Concatenate per-parameter COEFS (nz, nj) in inverted-parameter order
        into the (ifreep, nz) block the LM routines take."""
        cols = []
        for p in range(8):
            if cfg.atmpar[p]:
                c = np.asarray(coefs_by_param[p], dtype=np.float64)
                assert c.shape == (cfg.nz, cfg.nodes[p])
                cols.append(c.T)
        return np.ascontiguousarray(np.concatenate(cols, axis=0))

    def eqv_dsyn(self, cfg, coefs, dsyn):
        """dsyn (nz, numw, 8, 4) from forward1d -> EQVDSYN (nfreq, ifreep) float32."""
        co = np.ascontiguousarray(coefs, dtype=np.float64)
        ifreep = co.shape[0]
        nfreq = sum(cfg.invstk) * cfg.numw
        out = np.zeros((ifreep, nfreq), np.float32)
        d = np.ascontiguousarray(dsyn, dtype=np.float64)
        self.lib.fzc_eqv_dsyn(ctypes.byref(cfg), _ptr(co, _dp), _ptr(d, _dp), _ptr(out, _fp))
        return out.T.copy()

    def get_perturbation(self, jac, diff, lam, reg_pen=None, reg_hss=None, dtype=1, svdtol=1e-4):
        """jac (nf, np) -> delta (np) of GET_PERTURBATION."""
        nf, np_ = jac.shape
        jt = np.ascontiguousarray(np.asarray(jac, dtype=np.float64).T)
        df = np.ascontiguousarray(diff, dtype=np.float64)
        delta = np.zeros(np_)
        rp = None if reg_pen is None else _ptr(np.ascontiguousarray(reg_pen, dtype=np.float64), _dp)
        rh = None if reg_hss is None else _ptr(np.ascontiguousarray(np.asarray(reg_hss, np.float64).T), _dp)
        e = self.lib.fzc_get_perturbation(np_, nf, _ptr(jt, _dp), _ptr(df, _dp), lam, _ptr(delta, _dp), rp, rh,
                                          dtype, svdtol)
        if e:
            raise RuntimeError("fzc_get_perturbation error %d" % e)
        return delta

    def lm_step(self, cfg, coefs, obs, syn, best_syn, sla, jac_cur, jac_best, isigmap, model_cur, model_best):
        """One GET_PERT_1D.  jac_* (nfreq, ifreep) float32, models (13, nz)
        float32.  Returns (model_cur, model_best, sla, best_syn, jac_best, delta)."""
        co = np.ascontiguousarray(coefs, dtype=np.float64)
        ifreep = co.shape[0]
        m = np.ascontiguousarray(np.stack([model_cur, model_best]), dtype=np.float32).copy()
        bs = np.ascontiguousarray(best_syn, dtype=np.float64).copy()
        sl = np.ascontiguousarray(sla, dtype=np.float64).copy()
        jc = np.ascontiguousarray(np.asarray(jac_cur, np.float32).T)
        jb = np.ascontiguousarray(np.asarray(jac_best, np.float32).T).copy()
        ob = np.ascontiguousarray(obs, dtype=np.float64)
        sy = np.ascontiguousarray(syn, dtype=np.float64)
        isg = np.ascontiguousarray(isigmap, dtype=np.float64)
        delta = np.zeros(ifreep)
        e = self.lib.fzc_lm_step(ctypes.byref(cfg), _ptr(co, _dp), _ptr(ob, _dp), _ptr(sy, _dp), _ptr(bs, _dp),
                                 _ptr(sl, _dp), _ptr(jc, _fp), _ptr(jb, _fp), _ptr(isg, _dp), _ptr(m, _fp),
                                 _ptr(delta, _dp))
        if e:
            raise RuntimeError("fzc_lm_step error %d" % e)
        return m[0], m[1], sl, bs, jb.T.copy(), delta

    def he(self, tem, pg_top, z_km, pel_top=0.0):
        nz = len(tem)
        t = np.ascontiguousarray(tem, dtype=np.float32).copy()
        pg = np.zeros(nz, np.float32)
        pg[-1] = pg_top
        rho = np.zeros(nz, np.float32)
        pel = np.zeros(nz, np.float32)
        pel[-1] = pel_top
        mw = np.zeros(nz, np.float32)
        nh = np.zeros(nz)
        zz = np.ascontiguousarray(z_km, dtype=np.float64)
        e = self.lib.fzc_he(nz, _ptr(t, _fp), _ptr(pg, _fp), _ptr(rho, _fp), _ptr(pel, _fp), _ptr(mw, _fp),
                            _ptr(nh, _dp), _ptr(zz, _dp))
        if e:
            raise RuntimeError("fzc_he error %d" % e)
        return pg, rho, pel, mw, nh


class Ref:
    """ctypes view of the reference Fortran modules (oracle/_ref/libfzref.so)."""

    def __init__(self):
        if not ref_available():
            raise RuntimeError("oracle/_ref not built (needs the reference tree)")
        self.lib = ctypes.CDLL(REF_LIB)
        self.numw = 0

    def setup(self, ls: _lines.LineSetup, resp=True, atmpar=(1, 1, 0, 1, 1, 1, 1, 0)):
        n = len(ls.lines)
        zn = np.array([l.zn for l in ls.lines], np.int32)
        ion = np.array([l.ion for l in ls.lines], np.int32)
        l0 = np.array([l.lambda0 for l in ls.lines])
        lg = np.array([l.loggf for l in ls.lines])
        ep = np.array([l.eplow for l in ls.lines])
        sg = np.array([l.sigma for l in ls.lines])
        al = np.array([l.alpha for l in ls.lines])
        slj = np.array([[l.sl, l.ll, l.jl, l.su, l.lu, l.ju] for l in ls.lines])
        nr = len(ls.regions)
        nbl = np.array([len(r.lines) for r in ls.regions], np.int32)
        blid = np.zeros((nr, 4), np.int32)
        for i, r in enumerate(ls.regions):
            blid[i, :len(r.lines)] = r.lines
        nw = np.array([r.nwave for r in ls.regions], np.int32)
        wi = np.array([r.wave_start for r in ls.regions])
        dw = np.array([r.wave_step for r in ls.regions])
        ap = np.array(atmpar, np.int32)
        self._keep = (zn, ion, l0, lg, ep, sg, al, slj, nbl, blid, nw, wi, dw, ap)
        f = self.lib.fzref_setup
        f.argtypes = [ctypes.c_int, _ip, _ip, _dp, _dp, _dp, _dp, _dp, _dp, ctypes.c_int, _ip, _ip, _ip,
                      _dp, _dp, ctypes.c_int, _ip]
        f(n, _ptr(zn, _ip), _ptr(ion, _ip), _ptr(l0, _dp), _ptr(lg, _dp), _ptr(ep, _dp), _ptr(sg, _dp),
          _ptr(al, _dp), _ptr(slj, _dp), nr, _ptr(nbl, _ip), _ptr(blid, _ip), _ptr(nw, _ip), _ptr(wi, _dp),
          _ptr(dw, _dp), int(resp), _ptr(ap, _ip))
        self.numw = int(nw.sum())
        self.resp = bool(resp)
        return self.numw

    def partition(self, elem, T):
        u = np.zeros(3)
        du = np.zeros(3)
        f = self.lib.fzref_partition
        f.argtypes = [ctypes.c_int, ctypes.c_double, _dp, _dp]
        f(elem, T, _ptr(u, _dp), _ptr(du, _dp))
        return u, du

    def atomdata(self):
        xi, xii, ab, m = (np.zeros(92) for _ in range(4))
        den = np.zeros(2)
        self.lib.fzref_atomdata(*[_ptr(a, _dp) for a in (xi, xii, ab, m, den)])
        return xi, xii, ab, m, den

    def eos(self, T, P, ne=0.0):
        t = ctypes.c_float(T)
        p = ctypes.c_float(P)
        ne_ = ctypes.c_double(ne)
        out = np.zeros(5)
        d = np.zeros(25)
        self.lib.fzref_eos(ctypes.byref(t), ctypes.byref(p), ctypes.byref(ne_), _ptr(out, _dp), _ptr(d, _dp))
        return 0, out, d

    def opac_cont(self, T, lam, ne, nh):
        t = ctypes.c_float(T)
        kc = ctypes.c_double()
        d = np.zeros(25)
        f = self.lib.fzref_opac_cont
        f.argtypes = [ctypes.POINTER(ctypes.c_float), ctypes.c_double, ctypes.c_double, ctypes.c_double,
                      _dp, _dp]
        f(ctypes.byref(t), lam, ne, nh, ctypes.byref(kc), _ptr(d, _dp))
        return 0, kc.value, d

    def opac_line(self, line, T, ne, nh):
        t = ctypes.c_float(T)
        kl = ctypes.c_double()
        d = np.zeros(25)
        f = self.lib.fzref_opac_line
        f.argtypes = [ctypes.c_int, ctypes.POINTER(ctypes.c_float), ctypes.c_double, ctypes.c_double, _dp, _dp]
        f(line, ctypes.byref(t), ne, nh, ctypes.byref(kl), _ptr(d, _dp))
        return 0, kl.value, d

    def damping(self, line, T, nh):
        t = ctypes.c_float(T)
        dm = ctypes.c_double()
        d = np.zeros(25)
        f = self.lib.fzref_damping
        f.argtypes = [ctypes.c_int, ctypes.POINTER(ctypes.c_float), ctypes.c_double, _dp, _dp]
        f(line, ctypes.byref(t), nh, ctypes.byref(dm), _ptr(d, _dp))
        return dm.value, d

    def voigt(self, a, v):
        v = np.ascontiguousarray(v, dtype=np.float64)
        h = np.zeros_like(v)
        fo = np.zeros_like(v)
        f = self.lib.fzref_voigt
        f.argtypes = [ctypes.c_int, ctypes.c_double, _dp, _dp, _dp]
        f(len(v), a, _ptr(v, _dp), _ptr(h, _dp), _ptr(fo, _dp))
        return h, fo

    def zeeman(self, line):
        s = np.zeros(64)
        st = np.zeros(64)
        ty = np.zeros(64, dtype=np.int32)
        n = ctypes.c_int32()
        f = self.lib.fzref_zeeman
        f.argtypes = [ctypes.c_int, ctypes.POINTER(ctypes.c_int32), _dp, _dp, _ip]
        f(line, ctypes.byref(n), _ptr(s, _dp), _ptr(st, _dp), _ptr(ty, _ip))
        return s[:n.value], st[:n.value], ty[:n.value]

    def gain(self):
        g = np.zeros((self.numw, 4), dtype=np.float32)
        self.lib.fzref_gain(_ptr(g, _fp))
        return g.T.copy()

    def set_gain(self, g):
        g = np.ascontiguousarray(np.asarray(g, dtype=np.float32).T)
        self.lib.fzref_set_gain(_ptr(g, _fp))

    def set_hydrostatic(self, on=True):
        f = self.lib.fzref_set_hydro
        f.argtypes = [ctypes.c_int]
        f(int(on))

    def set_departure(self, popl=None, popu=None):
        """POPL, POPU (nreg, nz) for the reference's FORWARD_PARAM arrays; None: LTE."""
        f = self.lib.fzref_set_departure
        f.argtypes = [ctypes.c_int, ctypes.c_int, ctypes.c_void_p, ctypes.c_void_p]
        if popl is None:
            one = np.ones(1)
            f(0, 0, one.ctypes.data, one.ctypes.data)
            return
        self._pl = np.asfortranarray(popl, dtype=np.float64)
        self._pu = np.asfortranarray(popu, dtype=np.float64)
        f(self._pl.shape[0], self._pl.shape[1], self._pl.ctypes.data, self._pu.ctypes.data)

    def forward1d(self, model):
        m = np.asfortranarray(np.asarray(model, dtype=np.float32).T.copy())   # (nz,13) Fortran
        nz = m.shape[0]
        syn = np.zeros((self.numw, 4))
        dsyn = np.zeros((nz, self.numw, 8, 4))
        kc5 = np.zeros(nz)
        f = self.lib.fzref_forward1d
        f.argtypes = [ctypes.c_int, _fp, _dp, _dp, _dp]
        f(nz, _ptr(m, _fp), _ptr(syn, _dp), _ptr(dsyn, _dp), _ptr(kc5, _dp))
        return np.ascontiguousarray(m.T), syn.T.copy(), (dsyn if self.resp else None), kc5

    def he(self, tem, pg_top, z_km, pel_top=0.0):
        nz = len(tem)
        t = np.ascontiguousarray(tem, dtype=np.float32).copy()
        pg = np.zeros(nz, np.float32)
        pg[-1] = pg_top
        rho = np.zeros(nz, np.float32)
        pel = np.zeros(nz, np.float32)
        pel[-1] = pel_top
        mw = np.zeros(nz, np.float32)
        nh = np.zeros(nz)
        zz = np.ascontiguousarray(z_km, dtype=np.float64).copy()
        f = self.lib.fzref_he
        f.argtypes = [ctypes.c_int, _fp, _fp, _fp, _fp, _fp, _dp, _dp]
        f(nz, _ptr(t, _fp), _ptr(pg, _fp), _ptr(rho, _fp), _ptr(pel, _fp), _ptr(mw, _fp), _ptr(nh, _dp),
          _ptr(zz, _dp))
        return pg, rho, pel, mw, nh
