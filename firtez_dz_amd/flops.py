"""Algorithmic work of the radiative-transfer kernel (rt_kernel in
csrc/fz_gpu.hip), counted from its source per (height, wavelength) and summed
to a per-(column, wavelength) figure used for the roofline.

Convention: add, sub, mul = 1; FMA = 2; div, sqrt, exp, log, sin, cos and a
hardware reciprocal = 1 each.  Loads/stores, selects, comparisons and integer
index arithmetic are not counted.  The counts are per non-boundary layer.
They were tallied by hand from the source of each device function (the
names below follow csrc/fz_rt.hpp and rt_kernel); change them together with
the kernel.  Response-function directions that a mode does not store are
not formed by the kernel (the compiler drops them), so the per-line
derivative work depends on the parameter set.
"""
from __future__ import annotations

# per Zeeman component
VOIGT = 66           # Hui 7/7 rational: r, s (4); 5 + 6 two-FMA rounds (44); real/imag parts (6);
                     # |den|^2 (3), reciprocal (1), H and L (8)
COMP_FWD = 2 + VOIGT + 4              # u, profile, 2 accumulations (absorb)
COMP_DER = 2 + VOIGT + 4 + 5 + 20     # u, profile, dH/du, dH/da, 8 accumulations (absorb_der)
# per line
LINE_SETUP = 5                        # inverse Doppler width (from the stored 1/sqrt(T)), shift, field scale
LINE_FWD = LINE_SETUP + 2 * 14 + 14   # 2 lin() combinations, 7 accumulations
LIN_ALL = 8 * 14 + 2 * 15 + 2 * 8     # 8 lin(), 2 lin_dgamma(), 2 lin_dphi() combinations
# accumulation into the 7 propagation-matrix elements per direction formed
ACC = {"value": 2, "T": 7, "Pg": 5, "rho": 5, "v": 2, "B": 2, "gamma": 2, "phi": 2}
DIRS = {5: ("T", "v", "B", "gamma", "phi"), 7: ("T", "Pg", "rho", "v", "B", "gamma", "phi")}
# per layer
GEOMETRY = 0          # |B|, inclination and azimuth terms: formed by opac_kernel, read from the record
SOURCE = 6            # 1/T, c2/T, exp, -1, reciprocal, x c1
SOURCE_DER = 11       # the same and dS/dT = S (ex/(ex-1)) (c2/T) / T
EVOL_BUILD = 155                      # evol_build: scaling, eigen-parameters, exponentials, sin/cos, O
FORMAL = 40                           # I_k = (1-O) S + O I_{k-1}
M4V = 46                              # v = I_in - S, M4raw v
EVOL_DIR = 257                        # evol_dir: one directional derivative (dO) v
SOURCE_TERM = 12                      # (1-O) e0 dS/dT
PROPAGATE_PER_PARAM = 28              # ACC . local response
ACC_UPDATE = 112                      # ACC <- ACC . O_k
STORE_SCALE = 4                       # x GAIN1D per stored parameter


def line_der(nparam: int) -> int:
    acc = 7 * (ACC["value"] + sum(ACC[d] for d in DIRS[nparam]))
    return LINE_SETUP + LIN_ALL + 2 + acc


def per_layer(nlines: int, ncomp: int, nparam: int, chain: int = 0) -> dict:
    fwd = GEOMETRY + SOURCE + nlines * LINE_FWD + ncomp * COMP_FWD + EVOL_BUILD + FORMAL
    der = 0
    if nparam:
        der = (GEOMETRY + SOURCE_DER + nlines * line_der(nparam) + ncomp * COMP_DER + EVOL_BUILD + M4V
               + nparam * (EVOL_DIR + PROPAGATE_PER_PARAM + STORE_SCALE) + SOURCE_TERM + ACC_UPDATE + chain)
    return {"pass1": fwd, "pass2": der, "total": fwd + der}


def per_column_lambda(nz: int, nlines: int, ncomp: int, nparam: int) -> float:
    """FLOPs per (column, wavelength) over nz layers (boundary layer excluded)."""
    chain = 3 * 4 * 5 if nparam == 7 else 0
    return (nz - 1) * per_layer(nlines, ncomp, nparam, chain)["total"]


# bytes that must cross HBM per (column, wavelength): the Jacobian and Stokes
# written once, plus the per-height inputs amortised over the wavelengths
def bytes_per_column_lambda(nz: int, numw: int, nparam: int, nfield: int) -> float:
    out = 4 * 4 + nparam * nz * 4 * 4
    inp = (12 + nfield) * nz * 4 / numw
    return out + inp


# HBM traffic of rt_kernel<1> per column of the forward + Jacobian workload
# (nz = 96, 112 wavelengths, spherical Jacobian), from the FETCH_SIZE and
# WRITE_SIZE counters of one 32768-column launch (each scaled by its reading
# for a device copy of known size: FETCH_SIZE x2, WRITE_SIZE x1):
# (4.736e6 KiB x 2 + 3.309e7 KiB) x 1024 / 32768 columns
RT_HBM_BYTES_PER_COLUMN = 1.3300e6
RT_HBM_SOURCE = ("rocprofv3 --pmc FETCH_SIZE / WRITE_SIZE of rt_kernel<1>, a 32768-column launch of this "
                 "workload, scaled per column (profiles/rt_kernel_fwd_jacobian_288x288x96.md)")
