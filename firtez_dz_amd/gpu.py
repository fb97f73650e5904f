"""Python host binding of the C ABI (include/firtez_gpu.h).

Mirrors the reference's FORWARD3D operator (forward.f90:43-306): a cube of
columns goes in, Stokes profiles (SYN3D) and response functions (DSYN) come
out, and the cube's rho, Pe, mu and log tau5 are updated in place.  The HIP
library must be present: there is no CPU fallback, and every entry point
raises on a nonzero return code.
"""
from __future__ import annotations

import ctypes
import os
from typing import Optional

import numpy as np

from . import lines as _lines

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libfirtez_gpu.so")

RF_NONE, RF_SPHERICAL, RF_CARTESIAN, RF_CARTESIAN_P0 = 0, 1, 2, 3
NPARAM = {RF_NONE: 0, RF_SPHERICAL: 5, RF_CARTESIAN: 7, RF_CARTESIAN_P0: 8}
PARAM_NAMES = {RF_SPHERICAL: ("T", "B", "gamma", "phi", "vlos"),
               RF_CARTESIAN: ("T", "Pg", "rho", "Bx", "By", "Bz", "vz"),
               RF_CARTESIAN_P0: ("T", "Pg", "rho", "Bx", "By", "Bz", "vz", "P0")}
# device atmosphere rows (d_atm[12][nz][ncol])
ATM = ("T", "Pg", "rho", "Pe", "mu", "Bx", "By", "Bz", "vx", "vy", "vz", "tau5")

ERRORS = {1: "equation of state", 2: "non-finite value", 3: "ionisation stage", 4: "wavelength < 1800 A",
          5: "configuration", 6: "HIP runtime", 7: "device memory", 8: "argument"}

_lib = None


class FzError(RuntimeError):
    pass


class CInversion(ctypes.Structure):
    """fz_inversion (include/firtez_gpu.h)."""
    _fields_ = [("atmpar", ctypes.c_int32 * 8), ("invstk", ctypes.c_int32 * 4), ("nodes", ctypes.c_int32 * 8),
                ("pen_typ", ctypes.c_int32 * 8), ("pen_alp", ctypes.c_double * 8),
                ("regul", ctypes.c_int32), ("solver", ctypes.c_int32), ("node_expansion", ctypes.c_int32),
                ("maxsteps", ctypes.c_int32), ("svdtol", ctypes.c_double), ("inu", ctypes.c_double),
                ("mpert", ctypes.c_double)]


class CInvertControl(ctypes.Structure):
    """fz_invert_control (include/firtez_gpu.h)."""
    _fields_ = [("atmpar", ctypes.c_int32 * 8), ("nslb_max", ctypes.c_int32 * 8), ("invstk", ctypes.c_int32 * 4),
                ("wstk", ctypes.c_float * 4), ("noise", ctypes.c_double), ("cycpow", ctypes.c_int32),
                ("maxsteps", ctypes.c_int32), ("solver", ctypes.c_int32), ("node_expansion", ctypes.c_int32),
                ("regul", ctypes.c_int32), ("pen_typ", ctypes.c_int32 * 8), ("pen_alp", ctypes.c_double * 8),
                ("svdtol", ctypes.c_double), ("max_iterations", ctypes.c_int32), ("smoothing", ctypes.c_int32),
                ("first_cycle", ctypes.c_int32), ("assist", ctypes.c_int32)]


# SET_TAU assist initialisers (fz_invert_control.assist, fz_assist_dev)
ASSIST_B, ASSIST_V, ASSIST_T, ASSIST_P = 1, 2, 4, 8

# inverted physical parameters, in the reference's INV_ATMPAR order
INV_PARAMS = ("T", "Pg", "rho", "Bx", "By", "Bz", "vz", "P0")


def library():
    """Load libfirtez_gpu.so (built by firtez_dz_amd.build); raise if absent."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise FzError("libfirtez_gpu.so is not built: run python -m firtez_dz_amd.build")
    # One HIP runtime per process: when PyTorch is installed its bundled
    # libamdhip64 (same SONAME) is loaded first and the library binds to it,
    # so device pointers and streams can be shared with torch tensors.
    try:
        import torch  # noqa: F401
    except ImportError:
        pass
    L = ctypes.CDLL(LIB_PATH)
    vp, ip, dp, fp = ctypes.c_void_p, ctypes.c_int, ctypes.POINTER(ctypes.c_double), ctypes.POINTER(ctypes.c_float)
    L.fz_create.argtypes = [ctypes.POINTER(vp), ip]
    L.fz_destroy.argtypes = [vp]
    L.fz_last_error.argtypes = [vp]
    L.fz_last_error.restype = ctypes.c_char_p
    L.fz_configure.argtypes = [vp, ctypes.POINTER(_lines.CLine), ip, ctypes.POINTER(_lines.CRegion), ip, ip, dp]
    L.fz_numw.argtypes = [vp]
    L.fz_gain_hsra.argtypes = [vp, fp]
    L.fz_set_gain.argtypes = [vp, fp]
    L.fz_forward_dev.argtypes = [vp, ip, vp, ip, vp, vp, vp, vp]
    L.fz_hydrostatic_dev.argtypes = [vp, ip, vp, vp, vp]
    L.fz_forward3d.argtypes = [vp, ip, ip] + [fp] * 12 + [ctypes.POINTER(ctypes.c_int32), ip, dp, fp]
    L.fz_set_tile.argtypes = [vp, ip, ip]
    L.fz_forward_full_dev.argtypes = [vp, ip, vp, ip, vp, vp, vp, vp, vp]
    L.fz_line_tau_dev.argtypes = [vp, ip, vp, vp, vp, vp]
    L.fz_set_overlap.argtypes = [vp, ip]
    L.fz_set_auto_weights.argtypes = [vp, ip]
    L.fz_auto_weights_dev.argtypes = [vp, ip, vp, vp, ctypes.c_float, vp, vp]
    L.fz_set_hydrostatic.argtypes = [vp, ip]
    L.fz_set_microturbulence.argtypes = [vp, ctypes.c_double]
    L.fz_set_lsf.argtypes = [vp, vp, vp]
    L.fz_set_departure.argtypes = [vp, vp, vp]
    L.fz_assist_dev.argtypes = [vp, ip, vp, ip, vp, vp]
    L.fz_set_psf.argtypes = [vp, vp, ip, ip]
    L.fz_coupled_invert3d_dev.argtypes = [vp, ip, ip, vp, vp, vp, vp, vp, ctypes.POINTER(ctypes.c_int), vp]
    L.fz_psf_convolve_dev.argtypes = [vp, ip, ip, ip, vp, vp, ip, vp]
    L.fz_set_windows.argtypes = [vp, ip, vp]
    L.fz_window_offsets.argtypes = [vp, vp]
    L.fz_synchronize.argtypes = [vp]
    L.fz_set_profiling.argtypes = [vp, ip]
    L.fz_last_timing.argtypes = [vp, ctypes.POINTER(ctypes.c_float), ctypes.POINTER(ctypes.c_float),
                                 ctypes.POINTER(ctypes.c_int)]
    L.fz_int_coefs.argtypes = [ip, ip, dp]
    L.fz_set_inversion.argtypes = [vp, ctypes.POINTER(CInversion), ctypes.POINTER(ctypes.c_int),
                                   ctypes.POINTER(ctypes.c_int)]
    L.fz_forward_nodes_dev.argtypes = [vp, ip, vp, vp, vp, vp, vp]
    L.fz_lm_step_dev.argtypes = [vp, ip] + [vp] * 12
    L.fz_invert3d_dev.argtypes = [vp, ip, ip, vp, vp, ctypes.POINTER(CInvertControl), vp, vp,
                                  ctypes.POINTER(ctypes.c_int), vp]
    L.fz_invert3d.argtypes = [vp, ip, ip] + [fp] * 12 + [dp, ctypes.POINTER(CInvertControl), dp, dp,
                                                       ctypes.POINTER(ctypes.c_int)]
    L.fz_coupled_invert3d.argtypes = L.fz_invert3d.argtypes
    L.fz_autotune.argtypes = [vp, ip, vp, ip, vp, vp, vp, vp, ctypes.POINTER(ctypes.c_int), ctypes.POINTER(ctypes.c_int)]
    L.fz_mlp_create.argtypes = [vp, ip, ip, ip, ip, ctypes.c_uint64]
    L.fz_mlp_size.argtypes = [vp]
    L.fz_mlp_size.restype = ctypes.c_int64
    L.fz_mlp_get_params.argtypes = [vp, vp]
    L.fz_mlp_set_params.argtypes = [vp, vp]
    L.fz_mlp_fit_norms_dev.argtypes = [vp, ip, vp, vp, vp]
    L.fz_mlp_targets_dev.argtypes = [vp, ip, vp, vp, vp]
    L.fz_mlp_forward_dev.argtypes = [vp, ip, vp, vp, vp, vp]
    L.fz_mlp_gradient_dev.argtypes = [vp, ip, vp, vp, vp, vp, vp, vp]
    L.fz_mlp_train_step_dev.argtypes = [vp, ip, vp, vp, vp, ctypes.c_float, vp, vp]
    L.fz_learned_init_dev.argtypes = [vp, ip, vp, vp, vp]
    L.fz_get_perturbation.argtypes = [vp, ip, ip, dp, dp, ctypes.c_double, dp, dp, dp, ip, ctypes.c_double]
    L.fz_get_perturbation_dev.argtypes = [vp, ip, ip, vp, vp, ctypes.c_double, vp, vp, vp, ip, ctypes.c_double, vp]
    L.fz_hessian.argtypes = [vp, ip, ip, dp, dp]
    L.fz_get_dmodel3ds.argtypes = [vp, ip, ip, dp, dp, dp, dp, ip, fp, fp, dp, ctypes.POINTER(fp),
                                   ctypes.POINTER(fp), ctypes.POINTER(ctypes.c_int32)]
    L.fz_get_perturbation_dev_noalloc.argtypes = [vp, ip, ip, vp, vp, ctypes.c_double, vp, vp, vp, ip,
                                                  ctypes.c_double, vp]
    _lib = L
    return L


def exported_symbols():
    """Names declared in include/firtez_gpu.h."""
    return ["fz_create", "fz_destroy", "fz_last_error", "fz_configure", "fz_numw", "fz_gain_hsra", "fz_set_gain",
            "fz_forward_dev", "fz_forward_full_dev", "fz_line_tau_dev", "fz_forward3d", "fz_hydrostatic_dev", "fz_malloc", "fz_free", "fz_memcpy_h2d",
            "fz_memcpy_d2h", "fz_synchronize", "fz_set_tile", "fz_set_profiling", "fz_set_overlap", "fz_last_timing",
            "fz_set_auto_weights", "fz_auto_weights_dev",
            "fz_set_hydrostatic", "fz_set_microturbulence", "fz_set_lsf", "fz_set_departure", "fz_set_psf", "fz_psf_convolve_dev", "fz_coupled_invert3d_dev", "fz_assist_dev", "fz_set_windows", "fz_window_offsets",
            "fz_int_coefs", "fz_set_inversion", "fz_forward_nodes_dev", "fz_lm_step_dev", "fz_invert3d_dev",
            "fz_invert3d", "fz_coupled_invert3d",
            "fz_autotune", "fz_mlp_create", "fz_mlp_size", "fz_mlp_get_params", "fz_mlp_set_params", "fz_mlp_fit_norms_dev",
            "fz_mlp_targets_dev", "fz_mlp_forward_dev", "fz_mlp_gradient_dev", "fz_mlp_train_step_dev",
            "fz_learned_init_dev", "fz_get_perturbation", "fz_get_perturbation_dev", "fz_hessian", "fz_get_dmodel3ds",
            "fz_get_perturbation_dev_noalloc"]


def invert_control(max_nodes: dict, invstk=(1, 1, 1, 1), wstk=(1.0, 1.0, 1.0, 1.0), noise=4e-3, cycpow=2,
                   maxsteps=5, solver=1, node_expansion=1, svdtol=0.0, max_iterations=0, smoothing=True,
                   regul=False, pen_typ: Optional[dict] = None, pen_alp: Optional[dict] = None,
                   first_cycle=1, assist=0) -> CInvertControl:
    """fz_invert_control from {parameter name: maximum nodes}, e.g. {"T": 4, "vz": 4};
    assist: ASSIST_* bits (SET_TAU's initialisers of the first cycle)."""
    c = CInvertControl()
    for name, n in max_nodes.items():
        i = INV_PARAMS.index(name)
        c.atmpar[i] = 1
        c.nslb_max[i] = int(n)
    c.invstk[:] = [int(x) for x in invstk]
    c.wstk[:] = [float(x) for x in wstk]
    c.noise, c.cycpow, c.maxsteps, c.solver, c.node_expansion = noise, cycpow, maxsteps, solver, node_expansion
    c.regul, c.svdtol, c.max_iterations, c.smoothing = int(regul), svdtol, max_iterations, int(smoothing)
    c.first_cycle = int(first_cycle)
    c.assist = int(assist)
    for name, v in (pen_typ or {}).items():
        c.pen_typ[INV_PARAMS.index(name)] = int(v)
    for name, v in (pen_alp or {}).items():
        c.pen_alp[INV_PARAMS.index(name)] = float(np.float32(v))   # REAL(SP) PEN_ALP
    return c


def int_coefs(nz: int, nj: int) -> np.ndarray:
    """Node basis COEFS (nz, nj) of GET_INT_COEFS (get_dmodel.f90:563-644)."""
    c = np.zeros((nj, nz))
    rc = library().fz_int_coefs(nz, nj, c.ctypes.data_as(ctypes.POINTER(ctypes.c_double)))
    if rc:
        raise FzError("fz_int_coefs: bad sizes nz=%d nj=%d" % (nz, nj))
    return c.T.copy()


def inversion(nodes: dict, invstk=(1, 1, 1, 1), solver=1, node_expansion=1, maxsteps=5, svdtol=0.0, inu=0.0,
              mpert=0.0, regul=False, pen_typ: Optional[dict] = None, pen_alp: Optional[dict] = None) -> CInversion:
    """fz_inversion from {parameter name: nodes}, e.g. {"T": 2, "Bz": 1, "vz": 1}."""
    c = CInversion()
    for name, n in nodes.items():
        i = INV_PARAMS.index(name)
        c.atmpar[i] = 1
        c.nodes[i] = int(n)
    c.invstk[:] = [int(x) for x in invstk]
    for name, v in (pen_typ or {}).items():
        c.pen_typ[INV_PARAMS.index(name)] = int(v)
    for name, v in (pen_alp or {}).items():
        c.pen_alp[INV_PARAMS.index(name)] = float(np.float32(v))   # REAL(SP) PEN_ALP
    c.regul, c.solver, c.node_expansion, c.maxsteps = int(regul), solver, node_expansion, maxsteps
    c.svdtol, c.inu, c.mpert = svdtol, inu, mpert
    return c


def _f(a):
    return a.ctypes.data_as(ctypes.POINTER(ctypes.c_float))


class Context:
    """One device context: line setup, height grid, GAIN1D."""

    def __init__(self, device: int = 0):
        self.lib = library()
        self.h = ctypes.c_void_p()
        rc = self.lib.fz_create(ctypes.byref(self.h), device)
        if rc:
            raise FzError("fz_create failed (%d): no usable GPU %d" % (rc, device))
        self.numw = 0
        self.nz = 0
        self.lineset = None
        self.win_offsets = None
        self.hydrostatic = False

    def close(self):
        if self.h:
            self.lib.fz_destroy(self.h)
            self.h = ctypes.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def _check(self, rc, what):
        if rc:
            msg = self.lib.fz_last_error(self.h).decode()
            raise FzError("%s: %s error (%d) %s" % (what, ERRORS.get(rc, "?"), rc, msg))

    def configure(self, ls: _lines.LineSetup, z_km):
        cl, cr = _lines.c_arrays(ls)
        z = np.ascontiguousarray(z_km, dtype=np.float64)
        self._check(self.lib.fz_configure(self.h, cl, len(ls.lines), cr, len(ls.regions), len(z),
                                          z.ctypes.data_as(ctypes.POINTER(ctypes.c_double))), "fz_configure")
        self.numw = self.lib.fz_numw(self.h)
        self.nz = len(z)
        self.lineset = ls
        self.win_offsets = None
        return self.numw

    def gain_hsra(self) -> np.ndarray:
        """GAIN1D(4, numw) computed on the device from the HSRA model."""
        g = np.zeros((self.numw, 4), np.float32)
        self._check(self.lib.fz_gain_hsra(self.h, _f(g)), "fz_gain_hsra")
        return g.T.copy()

    def set_gain(self, gain):
        g = np.ascontiguousarray(np.asarray(gain, np.float32).T)
        self._check(self.lib.fz_set_gain(self.h, _f(g)), "fz_set_gain")

    def set_microturbulence(self, xi_cm_s: float = 0.0):
        """Microturbulence xi (cm/s) in every line's Doppler width
        sqrt(2kT/m + xi^2); 0 is the reference's thermal width."""
        self._check(self.lib.fz_set_microturbulence(self.h, float(xi_cm_s)), "fz_set_microturbulence")

    def set_overlap(self, on: bool = True):
        """Microphysics of the next radiative-transfer tile on a second stream
        while the current tile's transfer runs (the library's default is off)."""
        self._check(self.lib.fz_set_overlap(self.h, int(on)), "fz_set_overlap")

    def set_auto_weights(self, on: bool = True):
        """STOKES SETUP: AUTO -- invert3d weights every pixel by the spread of
        its own observed Stokes blocks (SET_WEIGHTS3D) instead of wstk."""
        self._check(self.lib.fz_set_auto_weights(self.h, int(on)), "fz_set_auto_weights")

    def auto_weights_torch(self, obs, invstk, noise: float):
        """SET_WEIGHTS3D on device: obs [ncol, NFREQ] float64 -> the per-pixel
        ISIGMAP3D [ncol, NFREQ] float64 (WSTK^2 / ISIGMA^2)."""
        import torch
        assert obs.is_cuda and obs.dtype == torch.float64 and obs.is_contiguous() and obs.dim() == 2
        out = torch.empty_like(obs)
        inv = (ctypes.c_int32 * 4)(*[int(bool(v)) for v in invstk])
        stream = torch.cuda.current_stream(obs.device).cuda_stream
        self._check(self.lib.fz_auto_weights_dev(self.h, obs.shape[0], obs.data_ptr(), inv, float(noise),
                                                 out.data_ptr(), stream), "fz_auto_weights_dev")
        return out

    def set_hydrostatic(self, on: bool = True):
        """HYDROSTATIC: yes -- Pg re-integrated from the top inside every
        forward, and the hydrostatic term in dI/dT."""
        self._check(self.lib.fz_set_hydrostatic(self.h, int(on)), "fz_set_hydrostatic")
        self.hydrostatic = bool(on)

    def set_psf(self, psf=None):
        """Spatial PSF of the coupled inversion, PSF(NPY, NPX) as READ_PSF reads
        it (a numpy array shaped (npy, npx)); None clears it."""
        if psf is None:
            self._check(self.lib.fz_set_psf(self.h, None, 0, 0), "fz_set_psf")
            return
        p = np.asfortranarray(psf, dtype=np.float64)
        self._psf = p
        self._check(self.lib.fz_set_psf(self.h, p.ctypes.data, p.shape[0], p.shape[1]), "fz_set_psf")

    def psf_convolve_torch(self, maps, nx, ny, adjoint=False, out=None):
        """maps: float32 cuda tensor [nmaps][nx*ny] (col = y + ny*x); returns the
        PSF's periodic convolution (or its adjoint)."""
        import torch
        assert maps.is_cuda and maps.dtype == torch.float32 and maps.is_contiguous()
        nmaps = maps.shape[0]
        if out is None:
            out = torch.empty_like(maps)
        self._check(self.lib.fz_psf_convolve_dev(self.h, nx, ny, nmaps, maps.data_ptr(), out.data_ptr(), int(adjoint),
                                                 torch.cuda.current_stream(maps.device).cuda_stream),
                    "fz_psf_convolve_dev")
        return out

    def set_lsf(self, lsf_kernel=None, valid=None):
        """MLSF: LSF_KERNEL(NUMW) and LSF_VALID(NUML) as LSF_INIT leaves them
        (lines.lsf_kernel builds both); None switches the convolution off."""
        if lsf_kernel is None:
            self._check(self.lib.fz_set_lsf(self.h, None, None), "fz_set_lsf")
            return
        k = np.ascontiguousarray(lsf_kernel, dtype=np.float64)
        if k.size != self.numw:
            raise ValueError("LSF kernel has %d entries, NUMW is %d" % (k.size, self.numw))
        v = None if valid is None else np.ascontiguousarray(valid, dtype=np.int32)
        self._check(self.lib.fz_set_lsf(self.h, k.ctypes.data, None if v is None else v.ctypes.data), "fz_set_lsf")
        self._lsf = (k, v)

    def set_departure(self, popl=None, popu=None):
        """NLTE departure coefficients POPL, POPU as arrays (nreg, nz)
        (input_file.f90:2146-2166); None: LTE."""
        if popl is None:
            self._check(self.lib.fz_set_departure(self.h, None, None), "fz_set_departure")
            return
        self._pop = (np.asfortranarray(popl, dtype=np.float64), np.asfortranarray(popu, dtype=np.float64))
        if self._pop[0].shape != (len(self.lineset.regions), self.nz) or self._pop[1].shape != self._pop[0].shape:
            raise FzError("departure coefficients must be (nregions, nz) = (%d, %d)"
                          % (len(self.lineset.regions), self.nz))
        self._check(self.lib.fz_set_departure(self.h, self._pop[0].ctypes.data, self._pop[1].ctypes.data),
                    "fz_set_departure")

    def set_windows(self, windows=None):
        """Per-pixel spectral windows: windows int [ncol][nreg][2] = (first
        wavelength index in the region, count); None switches them off.
        Afterwards spectra are ragged (see fz_set_windows) and
        self.win_offsets [ncol+1] holds each pixel's wavelength offset."""
        self.win_offsets = None          # the library drops the old windows first
        if windows is None:
            self._check(self.lib.fz_set_windows(self.h, 0, None), "fz_set_windows")
            return None
        w = np.ascontiguousarray(windows, dtype=np.int32)
        if w.ndim != 3 or w.shape[2] != 2:
            raise ValueError("windows must be [ncol][nreg][2]")
        ncol = w.shape[0]
        self._check(self.lib.fz_set_windows(self.h, ncol, w.ctypes.data), "fz_set_windows")
        off = np.zeros(ncol + 1, dtype=np.int64)
        self._check(self.lib.fz_window_offsets(self.h, off.ctypes.data), "fz_window_offsets")
        self.win_offsets = off
        return off

    def autotune_torch(self, atm, rf_mode, stokes, jac, status):
        """Pick the microphysics and radiative-transfer tiles by timing forwards
        on these buffers; returns (rt tile, microphysics tile)."""
        import torch
        t, m = ctypes.c_int(), ctypes.c_int()
        self._check(self.lib.fz_autotune(self.h, atm.shape[2], atm.data_ptr(), rf_mode, stokes.data_ptr(),
                                         None if jac is None else jac.data_ptr(), status.data_ptr(),
                                         torch.cuda.current_stream(atm.device).cuda_stream, ctypes.byref(t),
                                         ctypes.byref(m)), "fz_autotune")
        return t.value, m.value

    def set_tile(self, columns: int, block: int = 256):
        self._check(self.lib.fz_set_tile(self.h, columns, block), "fz_set_tile")

    # ---------------------------------------------------------------- device path
    def forward_dev(self, ncol: int, atm_ptr: int, rf_mode: int, stokes_ptr: int, jac_ptr: Optional[int],
                    status_ptr: Optional[int] = None, stream: Optional[int] = None):
        """Asynchronous launch on device buffers (see fz_forward_dev)."""
        self._check(self.lib.fz_forward_dev(self.h, ncol, atm_ptr, rf_mode, stokes_ptr, jac_ptr or None,
                                            status_ptr or None, stream or None), "fz_forward_dev")

    def hydrostatic_dev(self, ncol: int, atm_ptr: int, status_ptr: Optional[int] = None,
                        stream: Optional[int] = None):
        self._check(self.lib.fz_hydrostatic_dev(self.h, ncol, atm_ptr, status_ptr or None, stream or None),
                    "fz_hydrostatic_dev")

    def set_profiling(self, on=True):
        self._check(self.lib.fz_set_profiling(self.h, int(on)), "fz_set_profiling")

    def last_timing(self):
        """(micro_ms, rt_ms, rt_launches) of the last forward call (HIP events)."""
        a, b, n = ctypes.c_float(), ctypes.c_float(), ctypes.c_int()
        self._check(self.lib.fz_last_timing(self.h, ctypes.byref(a), ctypes.byref(b), ctypes.byref(n)),
                    "fz_last_timing")
        return a.value, b.value, n.value

    def synchronize(self):
        self._check(self.lib.fz_synchronize(self.h), "fz_synchronize")

    # ---------------------------------------------------------------- torch convenience
    def _stokes_shape(self, ncol):
        # ragged [4*sum numw_col] with per-pixel windows, else [ncol][4][numw]
        if self.win_offsets is not None:
            return (4 * int(self.win_offsets[-1]),)
        return (ncol, 4, self.numw)

    def forward_torch(self, atm, rf_mode=RF_SPHERICAL, stokes=None, jac=None, status=None):
        """atm: torch float32 cuda tensor [12, nz, ncol] (updated in place).
        Returns (stokes [ncol,4,numw], jac [ncol,np,nz,4,numw] or None, status [ncol])."""
        import torch
        assert atm.is_cuda and atm.dtype == torch.float32 and atm.is_contiguous()
        assert atm.shape[0] == 12 and atm.shape[1] == self.nz
        ncol = atm.shape[2]
        dev = atm.device
        if stokes is None:
            stokes = torch.empty(self._stokes_shape(ncol), dtype=torch.float32, device=dev)
        npar = NPARAM[rf_mode]
        if npar and jac is None:
            jac = torch.empty((ncol, npar, self.nz, 4, self.numw), dtype=torch.float32, device=dev)
        if status is None:
            status = torch.zeros(ncol, dtype=torch.int32, device=dev)
        stream = torch.cuda.current_stream(dev).cuda_stream
        self.forward_dev(ncol, atm.data_ptr(), rf_mode, stokes.data_ptr(), jac.data_ptr() if npar else None,
                         status.data_ptr(), stream)
        return stokes, (jac if npar else None), status

    def forward_full_torch(self, atm, rf_mode=RF_NONE, stokes=None, jac=None, status=None, full=None):
        """forward_torch plus FULL_STOKES (fz_forward_full_dev): also returns
        full [ncol, nz, 4, numw], the Stokes vector after every height (bottom
        first) times GAIN1D."""
        import torch
        assert atm.is_cuda and atm.dtype == torch.float32 and atm.is_contiguous()
        assert atm.shape[0] == 12 and atm.shape[1] == self.nz
        ncol = atm.shape[2]
        dev = atm.device
        if stokes is None:
            stokes = torch.empty(self._stokes_shape(ncol), dtype=torch.float32, device=dev)
        npar = NPARAM[rf_mode]
        if npar and jac is None:
            jac = torch.empty((ncol, npar, self.nz, 4, self.numw), dtype=torch.float32, device=dev)
        if status is None:
            status = torch.zeros(ncol, dtype=torch.int32, device=dev)
        if full is None:
            full = torch.empty((ncol, self.nz, 4, self.numw), dtype=torch.float32, device=dev)
        stream = torch.cuda.current_stream(dev).cuda_stream
        self._check(self.lib.fz_forward_full_dev(self.h, ncol, atm.data_ptr(), rf_mode, stokes.data_ptr(),
                                                 jac.data_ptr() if npar else None, full.data_ptr(),
                                                 status.data_ptr(), stream), "fz_forward_full_dev")
        return stokes, (jac if npar else None), status, full

    def line_tau_torch(self, atm, status=None):
        """LINE OPTICAL DEPTH (fz_line_tau_dev): log10 tau [ncol, nz, numw] of
        every wavelength; atm [12, nz, ncol] float32 cuda (EOS-updated)."""
        import torch
        ncol = atm.shape[2]
        out = torch.empty((ncol, self.nz, self.numw), dtype=torch.float32, device=atm.device)
        if status is None:
            status = torch.zeros(ncol, dtype=torch.int32, device=atm.device)
        self._check(self.lib.fz_line_tau_dev(self.h, ncol, atm.data_ptr(), out.data_ptr(), status.data_ptr(),
                                             torch.cuda.current_stream(atm.device).cuda_stream), "fz_line_tau_dev")
        return out, status

    def hydrostatic_torch(self, atm, status=None):
        import torch
        ncol = atm.shape[2]
        if status is None:
            status = torch.zeros(ncol, dtype=torch.int32, device=atm.device)
        self.hydrostatic_dev(ncol, atm.data_ptr(), status.data_ptr(), torch.cuda.current_stream(atm.device).cuda_stream)
        return status

    # ---------------------------------------------------------------- inversion
    def set_inversion(self, inv: CInversion):
        """Configure the inversion cycle; returns (IFREEP, NFREQ)."""
        npr, nf = ctypes.c_int(), ctypes.c_int()
        self._check(self.lib.fz_set_inversion(self.h, ctypes.byref(inv), ctypes.byref(npr), ctypes.byref(nf)),
                    "fz_set_inversion")
        self.ifreep, self.nfreq = npr.value, nf.value
        return self.ifreep, self.nfreq

    def forward_nodes_torch(self, atm, stokes=None, eqv=None, status=None):
        """FORWARD3D in inversion mode: atm [12, nz, ncol] float32 cuda (updated
        in place) -> (stokes [ncol,4,numw], eqv [ncol,IFREEP,NFREQ], status)."""
        import torch
        assert atm.is_cuda and atm.dtype == torch.float32 and atm.is_contiguous()
        ncol = atm.shape[2]
        dev = atm.device
        if stokes is None:
            stokes = torch.empty(self._stokes_shape(ncol), dtype=torch.float32, device=dev)
        if eqv is None:
            if self.win_offsets is not None:
                eqv = torch.empty(self.ifreep * self.nfreq // self.numw * int(self.win_offsets[-1]),
                                  dtype=torch.float32, device=dev)
            else:
                eqv = torch.empty((ncol, self.ifreep, self.nfreq), dtype=torch.float32, device=dev)
        if status is None:
            status = torch.zeros(ncol, dtype=torch.int32, device=dev)
        stream = torch.cuda.current_stream(dev).cuda_stream
        self._check(self.lib.fz_forward_nodes_dev(self.h, ncol, atm.data_ptr(), stokes.data_ptr(), eqv.data_ptr(),
                                                  status.data_ptr(), stream), "fz_forward_nodes_dev")
        return stokes, eqv, status

    def lm_step_torch(self, obs, stokes, best_syn, sla, eqv, eqv_best, isigmap, atm, atm_best, done=None,
                      delta=None):
        """One GET_PERT_1D per pixel (see fz_lm_step_dev); all arguments torch
        cuda tensors, updated in place."""
        import torch
        ncol = atm.shape[2]
        stream = torch.cuda.current_stream(atm.device).cuda_stream
        for name, t, dt in (("obs", obs, torch.float64), ("best_syn", best_syn, torch.float64),
                            ("sla", sla, torch.float64), ("isigmap", isigmap, torch.float64),
                            ("stokes", stokes, torch.float32), ("eqv", eqv, torch.float32),
                            ("eqv_best", eqv_best, torch.float32), ("atm", atm, torch.float32),
                            ("atm_best", atm_best, torch.float32)):
            if not (t.is_cuda and t.dtype == dt and t.is_contiguous()):
                raise FzError("%s must be a contiguous %s cuda tensor" % (name, dt))
        self._check(self.lib.fz_lm_step_dev(self.h, ncol, obs.data_ptr(), stokes.data_ptr(), best_syn.data_ptr(),
                                            sla.data_ptr(), eqv.data_ptr(), eqv_best.data_ptr(), isigmap.data_ptr(),
                                            atm.data_ptr(), atm_best.data_ptr(),
                                            done.data_ptr() if done is not None else None,
                                            delta.data_ptr() if delta is not None else None, stream),
                    "fz_lm_step_dev")

    # ---------------------------------------------------------------- learned initial model
    def mlp_create(self, nin: int, nhidden: int = 512, nlayers: int = 3, nodes: int = 4, seed: int = 1):
        """Network nin -> nhidden (x nlayers, ReLU) -> 5*nodes, BF16 weights."""
        self._check(self.lib.fz_mlp_create(self.h, nin, nhidden, nlayers, nodes, seed), "fz_mlp_create")
        self.mlp_shape = (nin, nhidden, nlayers, 5 * nodes)

    def mlp_params(self) -> np.ndarray:
        n = self.lib.fz_mlp_size(self.h)
        if n < 0:
            raise FzError("no network: call mlp_create first")
        out = np.empty(n, np.float32)
        self._check(self.lib.fz_mlp_get_params(self.h, out.ctypes.data), "fz_mlp_get_params")
        return out

    def mlp_set_params(self, params):
        p = np.ascontiguousarray(params, dtype=np.float32)
        if p.size != self.lib.fz_mlp_size(self.h):
            raise ValueError("parameter vector of %d, network has %d" % (p.size, self.lib.fz_mlp_size(self.h)))
        self._check(self.lib.fz_mlp_set_params(self.h, p.ctypes.data), "fz_mlp_set_params")

    @staticmethod
    def _stream(t):
        import torch
        return torch.cuda.current_stream(t.device).cuda_stream

    def mlp_fit_norms_torch(self, x, t):
        self._check(self.lib.fz_mlp_fit_norms_dev(self.h, x.shape[0], x.data_ptr(), t.data_ptr(), self._stream(x)),
                    "fz_mlp_fit_norms_dev")

    def mlp_targets_torch(self, atm):
        import torch
        ncol = atm.shape[2]
        t = torch.empty((ncol, self.mlp_shape[3]), dtype=torch.float32, device=atm.device)
        self._check(self.lib.fz_mlp_targets_dev(self.h, ncol, atm.data_ptr(), t.data_ptr(), self._stream(atm)),
                    "fz_mlp_targets_dev")
        return t

    def mlp_forward_torch(self, x):
        """Standardised network output for inputs x [n][nin] (float32 or float64)."""
        import torch
        y = torch.empty((x.shape[0], self.mlp_shape[3]), dtype=torch.float32, device=x.device)
        xf = x.data_ptr() if x.dtype == torch.float32 else None
        xd = x.data_ptr() if x.dtype == torch.float64 else None
        self._check(self.lib.fz_mlp_forward_dev(self.h, x.shape[0], xf, xd, y.data_ptr(), self._stream(x)),
                    "fz_mlp_forward_dev")
        return y

    def mlp_gradient_torch(self, x, t, idx=None):
        """(loss, gradient of the parameter vector) on rows x, targets t."""
        import torch
        n = x.shape[0] if idx is None else idx.shape[0]
        loss = torch.empty(1, dtype=torch.float32, device=x.device)
        g = torch.empty(self.lib.fz_mlp_size(self.h), dtype=torch.float32, device=x.device)
        self._check(self.lib.fz_mlp_gradient_dev(self.h, n, x.data_ptr(), t.data_ptr(),
                                                 None if idx is None else idx.data_ptr(), loss.data_ptr(),
                                                 g.data_ptr(), self._stream(x)), "fz_mlp_gradient_dev")
        return loss, g

    def mlp_train_step_torch(self, x, t, idx, lr, loss):
        self._check(self.lib.fz_mlp_train_step_dev(self.h, idx.shape[0], x.data_ptr(), t.data_ptr(), idx.data_ptr(),
                                                   float(lr), loss.data_ptr(), self._stream(x)),
                    "fz_mlp_train_step_dev")

    def assist_torch(self, obs, atm, flags):
        """SET_TAU's assist initialisers (fz_assist_dev) on device tensors:
        obs [ncol][NFREQ] float64 (after set_inversion), atm [12][nz][ncol]."""
        self._check(self.lib.fz_assist_dev(self.h, obs.shape[0], obs.data_ptr(), int(flags), atm.data_ptr(),
                                           self._stream(atm)), "fz_assist_dev")

    def learned_init_torch(self, obs, atm):
        """Write the network's T, Bx, By, Bz, vz for observations obs
        [ncol][4*numw] (float64) into atm [12][nz][ncol]."""
        self._check(self.lib.fz_learned_init_dev(self.h, obs.shape[0], obs.data_ptr(), atm.data_ptr(),
                                                 self._stream(atm)), "fz_learned_init_dev")

    def get_perturbation(self, jac, diff, lam, reg_pen=None, reg_hss=None, dtype=1, svdtol=1e-4):
        """GET_PERTURBATION (lm.f90:137-204) on host arrays: jac (nf, np)
        = IJACOBIAN, diff (nf), reg_pen (np), reg_hss (np, np); returns
        DELTA (np)."""
        jac = np.asarray(jac, dtype=np.float64)
        nf, np_ = jac.shape
        jt = np.ascontiguousarray(jac.T)                  # IJACOBIAN(NF,NP) column-major
        df = np.ascontiguousarray(diff, dtype=np.float64)
        delta = np.zeros(np_)
        dp = ctypes.POINTER(ctypes.c_double)
        rp = None if reg_pen is None else np.ascontiguousarray(reg_pen, dtype=np.float64)
        rh = None if reg_hss is None else np.ascontiguousarray(np.asarray(reg_hss, np.float64).T)
        self._check(self.lib.fz_get_perturbation(self.h, np_, nf, jt.ctypes.data_as(dp), df.ctypes.data_as(dp),
                                                 float(lam), delta.ctypes.data_as(dp),
                                                 None if rp is None else rp.ctypes.data_as(dp),
                                                 None if rh is None else rh.ctypes.data_as(dp), int(dtype),
                                                 float(svdtol)), "fz_get_perturbation")
        return delta

    def get_perturbation_torch(self, jt, diff, lam, delta, reg_pen=None, reg_hss=None, dtype=1, svdtol=1e-4):
        """fz_get_perturbation_dev on device tensors: jt (np, nf) (the
        IJACOBIAN(NF,NP) layout), diff (nf), delta (np) out, reg_hss (np, np)
        with REG_HSS(i, j) at [j, i]."""
        np_, nf = jt.shape
        ptr = lambda t: None if t is None else t.data_ptr()
        self._check(self.lib.fz_get_perturbation_dev(self.h, np_, nf, jt.data_ptr(), diff.data_ptr(), float(lam),
                                                     delta.data_ptr(), ptr(reg_pen), ptr(reg_hss), int(dtype),
                                                     float(svdtol), self._stream(jt)), "fz_get_perturbation_dev")

    def get_perturbation_noalloc_torch(self, jt, diff, lam, delta, reg_pen=None, reg_hss=None, dtype=1,
                                       svdtol=1e-4):
        """fz_get_perturbation_dev_noalloc (the library's fixed scratch, no
        device allocation); arguments as get_perturbation_torch."""
        np_, nf = jt.shape
        ptr = lambda t: None if t is None else t.data_ptr()
        self._check(self.lib.fz_get_perturbation_dev_noalloc(self.h, np_, nf, jt.data_ptr(), diff.data_ptr(),
                                                             float(lam), delta.data_ptr(), ptr(reg_pen),
                                                             ptr(reg_hss), int(dtype), float(svdtol),
                                                             self._stream(jt)), "fz_get_perturbation_dev_noalloc")

    def get_dmodel3ds(self, obs3d, syn3d, best_syn, inv_mas, dsyn, best_dsyn, isigmap3d, model: dict,
                      best_model: dict, am_i_done):
        """GET_DMODEL3DS (get_dmodel.f90:1327-1577) on host arrays in the
        reference's column-major layout, given here as numpy arrays indexed
        [x, y, ...] (C order of the reversed Fortran shape): obs3d, syn3d,
        best_syn, isigmap3d (nx, ny, NFREQ); inv_mas (nx, ny, ld >= 9);
        dsyn, best_dsyn (nx, ny, IFREEP, NFREQ) float32; model / best_model
        {name: (nx, ny, nz) float32} with the names of ATM; am_i_done
        (nx, ny) int32.  best_syn, inv_mas, best_dsyn, the models and
        am_i_done are updated in place."""
        nx, ny = obs3d.shape[:2]
        dp, fp = ctypes.POINTER(ctypes.c_double), ctypes.POINTER(ctypes.c_float)
        for arr, dt in ((obs3d, np.float64), (syn3d, np.float64), (best_syn, np.float64), (inv_mas, np.float64),
                        (dsyn, np.float32), (best_dsyn, np.float32), (isigmap3d, np.float64), (am_i_done, np.int32)):
            if arr.dtype != dt or not arr.flags.c_contiguous:
                raise ValueError("get_dmodel3ds: contiguous %s arrays expected" % np.dtype(dt).name)
        cubes = [model[n] for n in ATM]
        bcubes = [best_model[n] for n in ATM]
        for cb in cubes + bcubes:
            if cb.dtype != np.float32 or not cb.flags.c_contiguous:
                raise ValueError("get_dmodel3ds: contiguous float32 model cubes expected")
        mp = (fp * 12)(*[cb.ctypes.data_as(fp) for cb in cubes])
        bp = (fp * 12)(*[cb.ctypes.data_as(fp) for cb in bcubes])
        self._check(self.lib.fz_get_dmodel3ds(self.h, nx, ny, obs3d.ctypes.data_as(dp), syn3d.ctypes.data_as(dp),
                                              best_syn.ctypes.data_as(dp), inv_mas.ctypes.data_as(dp),
                                              inv_mas.shape[2], dsyn.ctypes.data_as(fp), best_dsyn.ctypes.data_as(fp),
                                              isigmap3d.ctypes.data_as(dp), mp, bp,
                                              am_i_done.ctypes.data_as(ctypes.POINTER(ctypes.c_int32))),
                    "fz_get_dmodel3ds")

    def hessian(self, jacobian):
        """HESSIAN (lm.f90:30-39): jacobian (np, nf) -> HSS (np, np)."""
        jacobian = np.asarray(jacobian, dtype=np.float64)
        np_, nf = jacobian.shape
        jf = np.asfortranarray(jacobian)                   # JACOBIAN(NP,NF) column-major
        hss = np.zeros((np_, np_), order="F")
        dp = ctypes.POINTER(ctypes.c_double)
        self._check(self.lib.fz_hessian(self.h, np_, nf, jf.ctypes.data_as(dp), hss.ctypes.data_as(dp)),
                    "fz_hessian")
        return np.ascontiguousarray(hss)

    def invert3d_torch(self, atm, obs, control: CInvertControl, nx: int, ny: int, stokes=None, chi2=None):
        """INVERT3D on device tensors: atm [12, nz, nx*ny] float32 (initial
        model in, inverted model out), obs [nx*ny, NFREQ] float64.  Returns
        (stokes [ncol,4,numw], chi2 [ncol], LM iterations)."""
        import torch
        ncol = nx * ny
        assert atm.is_cuda and atm.dtype == torch.float32 and atm.is_contiguous() and atm.shape[2] == ncol
        assert obs.is_cuda and obs.dtype == torch.float64 and obs.is_contiguous()
        assert obs.shape[0] == ncol or self.win_offsets is not None
        if stokes is None:
            stokes = torch.empty(self._stokes_shape(ncol), dtype=torch.float32, device=atm.device)
        if chi2 is None:
            chi2 = torch.empty(ncol, dtype=torch.float64, device=atm.device)
        its = ctypes.c_int()
        stream = torch.cuda.current_stream(atm.device).cuda_stream
        self._check(self.lib.fz_invert3d_dev(self.h, nx, ny, atm.data_ptr(), obs.data_ptr(), ctypes.byref(control),
                                             stokes.data_ptr(), chi2.data_ptr(), ctypes.byref(its), stream),
                    "fz_invert3d_dev")
        return stokes, chi2, its.value

    def coupled_invert3d_torch(self, atm, obs, control: CInvertControl, nx: int, ny: int, stokes=None, chi2=None):
        """INVERT3D through the spatial PSF of set_psf (fz_coupled_invert3d_dev):
        as invert3d_torch, obs being the PSF-degraded observations; stokes
        receives the PSF-convolved synthesis of the inverted model."""
        import torch
        ncol = nx * ny
        assert atm.is_cuda and atm.dtype == torch.float32 and atm.is_contiguous() and atm.shape[2] == ncol
        assert obs.is_cuda and obs.dtype == torch.float64 and obs.is_contiguous() and obs.shape[0] == ncol
        if stokes is None:
            stokes = torch.empty(self._stokes_shape(ncol), dtype=torch.float32, device=atm.device)
        if chi2 is None:
            chi2 = torch.empty(ncol, dtype=torch.float64, device=atm.device)
        its = ctypes.c_int()
        stream = torch.cuda.current_stream(atm.device).cuda_stream
        self._check(self.lib.fz_coupled_invert3d_dev(self.h, nx, ny, atm.data_ptr(), obs.data_ptr(),
                                                     ctypes.byref(control), stokes.data_ptr(), chi2.data_ptr(),
                                                     ctypes.byref(its), stream), "fz_coupled_invert3d_dev")
        return stokes, chi2, its.value

    def invert3d(self, cube: dict, obs3d, control: CInvertControl, coupled: bool = False):
        """INVERT3D on host cubes (invert3d.f90:57-178): cube as for forward3d
        (updated in place with the inverted model), obs3d (NFREQ, ny, nx) float64
        Fortran order.  Returns (SYN3D (4*numw, ny, nx), chi2 (ny, nx), iterations).
        coupled: the PSF-coupled inversion (fz_coupled_invert3d) with the PSF
        of set_psf; obs3d then holds PSF-degraded profiles."""
        nz, ny, nx = cube["T"].shape
        arrs = []
        for name in ATM:
            a = cube[name]
            if not (a.dtype == np.float32 and a.flags.f_contiguous):
                raise FzError("cube['%s'] must be float32 Fortran-contiguous" % name)
            arrs.append(a)
        obs = np.asfortranarray(obs3d, dtype=np.float64)
        syn = np.zeros((4 * self.numw, ny, nx), np.float64, order="F")
        chi = np.zeros((ny, nx), np.float64, order="F")
        its = ctypes.c_int()
        dp = ctypes.POINTER(ctypes.c_double)
        fn = self.lib.fz_coupled_invert3d if coupled else self.lib.fz_invert3d
        self._check(fn(self.h, nx, ny, *[_f(a) for a in arrs], obs.ctypes.data_as(dp), ctypes.byref(control),
                       syn.ctypes.data_as(dp), chi.ctypes.data_as(dp), ctypes.byref(its)), fn.__name__)
        return syn, chi, its.value

    # ---------------------------------------------------------------- FORWARD3D mirror
    def forward3d(self, cube: dict, am_i_done=None, rf_mode=RF_NONE):
        """FORWARD3D on host cubes (forward.f90:43-306).

        cube: dict of Fortran-order arrays shaped (nz, ny, nx) float32 for
        'T','Pg','rho','Pe','mu','Bx','By','Bz','vx','vy','vz','tau5'; rho, Pe,
        mu and tau5 are updated in place.  Returns SYN3D (4*numw, ny, nx)
        float64 and, for rf_mode != RF_NONE, DSYN (4*numw, np*nz, ny, nx) float32
        (both Fortran order)."""
        nz, ny, nx = cube["T"].shape
        if nz != self.nz:
            raise FzError("cube has nz=%d, context nz=%d" % (nz, self.nz))
        arrs = []
        for name in ATM:
            a = cube[name]
            if not (a.dtype == np.float32 and a.flags.f_contiguous):
                raise FzError("cube['%s'] must be float32 Fortran-contiguous" % name)
            arrs.append(a)
        nfreq = 4 * self.numw
        syn = np.zeros((nfreq, ny, nx), dtype=np.float64, order="F")
        npar = NPARAM[rf_mode]
        dsyn = np.zeros((nfreq, npar * nz, ny, nx), dtype=np.float32, order="F") if npar else None
        done = None
        if am_i_done is not None:
            done = np.asfortranarray(am_i_done, dtype=np.int32)
        rc = self.lib.fz_forward3d(self.h, nx, ny, *[_f(a) for a in arrs],
                                   done.ctypes.data_as(ctypes.POINTER(ctypes.c_int32)) if done is not None else None,
                                   rf_mode, syn.ctypes.data_as(ctypes.POINTER(ctypes.c_double)),
                                   _f(dsyn) if dsyn is not None else None)
        self._check(rc, "fz_forward3d")
        return syn, dsyn

