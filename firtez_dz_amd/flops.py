"""Algorithmic work of the radiative-transfer kernel (rt_kernel in
csrc/fz_gpu.hip), counted from its source per (height, wavelength) and summed
to a per-(column, wavelength) figure used for the roofline.

Convention: add, sub, mul = 1; FMA = 2; div, sqrt, exp, log, sin, cos, sinh,
cosh = 1 each.  Loads/stores and integer index arithmetic are not counted.
The counts are per non-boundary layer.  The numbers were tallied by hand
from the source of each device function and are kept next to it in this
file; change them together with the kernel.
"""
from __future__ import annotations

# per Zeeman component
VOIGT = 95           # Hui 7/7 rational in complex float arithmetic (fz_rt.hpp voigt)
COMP_FWD = 2 + VOIGT + 4              # u, profile, 2 accumulations (absorb)
COMP_DER = 2 + VOIGT + 9 + 24         # u, profile, dH/du, dH/da, 8 accumulations (absorb_der)
# per line
LINE_FWD = 7 + 24 + 14                # Doppler width/shift, 2 lin() combinations, 7 accumulations
LINE_DER = 15 + 12 * 12 + 7 * 30      # setup, 12 lin() combinations, 7 directions x 7 elements
# per layer
GEOMETRY = 20
SOURCE = 4
SOURCE_DER = 10
EVOL_BUILD = 160                      # evol_build: scaling, eigenvalues, exp/cosh/sinh/cos/sin, O assembly
                                      # (cosh, sinh from one exp, a reciprocal and a 4-term series for small
                                      # arguments: kept at 160, an undercount of about 15)
FORMAL = 40                           # I_k = (1-O) S + O I_{k-1}
M4V = 28                              # v = I_in - S, M4raw v
EVOL_DIR = 257                        # evol_dir: one directional derivative (dO) v
SOURCE_TERM = 12                      # (1-O) e0 dS/dT
PROPAGATE_PER_PARAM = 28              # ACC . local response
ACC_UPDATE = 112                      # ACC <- ACC . O_k


def per_layer(nlines: int, ncomp: int, nparam: int, chain: int = 0) -> dict:
    fwd = GEOMETRY + SOURCE + nlines * LINE_FWD + ncomp * COMP_FWD + EVOL_BUILD + FORMAL
    der = 0
    if nparam:
        der = (GEOMETRY + SOURCE_DER + nlines * LINE_DER + ncomp * COMP_DER + EVOL_BUILD + M4V
               + nparam * (EVOL_DIR + PROPAGATE_PER_PARAM) + SOURCE_TERM + ACC_UPDATE + chain)
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
