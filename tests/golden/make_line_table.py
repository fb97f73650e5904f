"""Derive the per-line constants of the reference's line database.

Runs only where the reference tree is present.  It reads, as text:
  * default/lines_database.dat, parsed like LINESDATABASE_GET13/GET8
    (lines_database.f90:178-300): numeric columns go through a REAL(4) read
    and are then rounded to 1e-6 (lines_database.f90:235-242);
  * the Anstee-Barklem-O'Mara cross-section tables CS_/AL_{SP,PD,DF}
    (damping.f90:24-835), used when a row gives no sigma/alpha.
It restates NREAD_LINES_DATABASE's choice of collisional parameters
(lines_database.f90:482-545): effective quantum numbers (GET_NEFF,
damping.f90:1016-1090) and the bilinear table interpolation
(TRANSITION_SP/PD/DF, damping.f90:1126-1305, including the SP row-index test
that compares against the s-grid, damping.f90:1155).

Output: firtez_dz_amd/data/lines_database.json (data only).
"""
import json
import os
import re
import sys

import numpy as np

REF = os.environ.get("FZ_REFERENCE", "/root/reference")
HERE = os.path.dirname(os.path.abspath(__file__))
OUT = os.path.join(HERE, "..", "..", "firtez_dz_amd", "data", "lines_database.json")

f32 = np.float32
EVOLT = 1.60217656e-12
HPLA = 6.62606957e-27
LIGHT = 2.99792458e10
XII_H = float(f32(13.595))          # atom_database.f90:71, default-real literal


def read_real(tok):
    """REAL(4) read, promoted, then rounded to 1e-6 (lines_database.f90:235-242)."""
    if "." in tok:
        d = float(f32(float(tok)))
    else:
        d = float(int(tok))
    return float(np.rint(d * 1.0e6)) * 1.0e-6


def elecconf(tok):
    """S, L, J from e.g. '5.0P2.0' (lines_database.f90:692-738)."""
    s = float(f32(float(tok[0:3])))
    ell = "SPDFGHI".index(tok[3])
    j = float(f32(float(tok[4:7])))
    return (s - 1.0) / 2.0, float(ell), j


def shell_code(sh):
    m = {"s": 0, "p": 1, "d": 2, "f": 3}
    codes, flag = [], 1
    for ch in sh[:2]:
        flag = 1
        c = m.get(ch.lower())
        if c is not None:
            flag = 0
        codes.append(c if c is not None else -1)
    return codes, flag


def parse_tables(text):
    tabs = {}
    for m in re.finditer(r"PARAMETER :: (\w\w_\w\w)\((\d+),(\d+)\) = RESHAPE\(.*?SOURCE=\s*&?\s*\(/(.*?)/\)(.*?)\)\s*\n",
                         text, re.S):
        name, n1, n2, body, tail = m.group(1), int(m.group(2)), int(m.group(3)), m.group(4), m.group(5)
        body = body.replace("&", " ")
        vals = [float(v) for v in re.findall(r"[-+]?\d*\.?\d+(?:[eEdD][-+]?\d+)?", body)]
        assert len(vals) == n1 * n2, (name, len(vals), n1 * n2)
        if re.search(r"order\s*=\s*\(/\s*2\s*,\s*1\s*/\)", tail, re.I):
            # RESHAPE(..., ORDER=(/2,1/)): the second index varies fastest
            tabs[name] = np.array(vals, dtype=np.float64).reshape((n1, n2))
        else:
            tabs[name] = np.array(vals, dtype=np.float64).reshape((n2, n1)).T   # Fortran order
    return tabs


def neff(zn, ion, eplow, l0, xi, xii, elow=None, eupp=None):
    xlow = eplow * EVOLT
    xupp = xlow + HPLA * LIGHT / (l0 * float(f32(1e-8)))
    xh = XII_H * EVOLT
    if elow is not None:
        lim_lo, lim_up = elow, eupp
    else:
        if zn > 1:
            lim = (xi[zn - 1] if ion == 1 else xii[zn - 1]) * EVOLT
        else:
            lim = xii[0] * EVOLT
        lim_lo = lim_up = lim / (HPLA * LIGHT)
    xlow /= HPLA * LIGHT
    xupp /= HPLA * LIGHT
    xh /= HPLA * LIGHT
    return np.sqrt(xh / (lim_lo - xlow)), np.sqrt(xh / (lim_up - xupp))


def bilinear(tab_s, tab_a, gx, gy, x, y, iy_grid):
    ix = int(np.argmin(np.abs(x - gx)))
    iy = int(np.argmin(np.abs(y - gy)))
    if x < gx[ix]:
        ix -= 1
    if y < iy_grid[iy]:
        iy -= 1
    a1 = abs((gx[ix + 1] - x) * (gy[iy + 1] - y))
    a2 = abs((gx[ix] - x) * (gy[iy + 1] - y))
    a3 = abs((gx[ix] - x) * (gy[iy] - y))
    a4 = abs((gx[ix + 1] - x) * (gy[iy] - y))
    at = a1 + a2 + a3 + a4
    def interp(t):
        return (t[ix, iy] * a1 + t[ix + 1, iy] * a2 + t[ix + 1, iy + 1] * a3 + t[ix, iy + 1] * a4) / at
    return interp(tab_s), interp(tab_a)


def grid(start, n):
    return np.array([float(f32(start)) + float(i) * float(f32(0.1)) for i in range(n)])


def collisional(nlo, nup, tr, tabs):
    """Returns (sigma, alpha, ierr) like GET_COLLISIONAL_PARAM (damping.f90:1094-1122)."""
    kind = sum(tr)
    if kind == 1:
        nss = nlo if tr[0] == 0 else nup
        nsp = nlo if tr[0] == 1 else nup
        if nss > 5.5 or nss < 1.0 or nsp > 5.5 or nsp < float(f32(1.3)):
            return None, None, 3
        gs, gp = grid(1.0, 46), grid(1.3, 43)
        s, a = bilinear(tabs["CS_SP"], tabs["AL_SP"], gs, gp, nss, nsp, gs)
        return s, a, 0
    if kind == 3:
        nsp = nlo if tr[0] == 1 else nup
        nsd = nlo if tr[0] == 2 else nup
        if nsp > 5.5 or nsp < float(f32(1.3)) or nsd > 5.5 or nsd < float(f32(2.3)):
            return None, None, 3
        gp, gd = grid(1.3, 43), grid(2.3, 33)
        s, a = bilinear(tabs["CS_PD"], tabs["AL_PD"], gp, gd, nsp, nsd, gd)
        return s, a, 0
    if kind == 5:
        nsd = nlo if tr[0] == 2 else nup
        nsf = nlo if tr[0] == 3 else nup
        if nsd > 5.5 or nsd < float(f32(2.3)) or nsf > 6.0 or nsf < float(f32(3.3)):
            return None, None, 3
        gd, gf = grid(2.3, 33), grid(3.3, 28)
        s, a = bilinear(tabs["CS_DF"], tabs["AL_DF"], gd, gf, nsd, nsf, gf)
        return s, a, 0
    return None, None, 1


def atom_potentials(text):
    def arr(name):
        m = re.search(name + r"\(:\)=\(/(.*?)/\)", text, re.S)
        toks = re.findall(r"[-+]?\d*\.?\d+(?:[eEdD][-+]?\d+)?", m.group(1).replace("&", " "))
        out = []
        for t in toks:
            if "D" in t or "d" in t:
                out.append(float(t.replace("D", "e").replace("d", "e")))
            else:
                out.append(float(f32(float(t))))
        return out
    return arr("XI"), arr("XII")


def main():
    src = os.path.join(REF, "src")
    tabs = parse_tables(open(os.path.join(src, "damping.f90")).read())
    xi, xii = atom_potentials(open(os.path.join(src, "atom_database.f90")).read())
    rows = []
    for raw in open(os.path.join(REF, "default", "lines_database.dat")):
        toks = raw.split()
        if not toks or toks[0] == "END":
            break
        rec = {"id": int(toks[0]), "zn": int(toks[1]), "ion": int(toks[2]),
               "lambda0": read_real(toks[3]), "loggf": read_real(toks[6]),
               "eplow": read_real(toks[7])}
        lo, up = elecconf(toks[4]), elecconf(toks[5])
        rec["sl"], rec["ll"], rec["jl"] = lo
        rec["su"], rec["lu"], rec["ju"] = up
        alpha = sigma = 0.0
        elow = eupp = 0.0
        shell = ""
        if len(toks) >= 13:
            alpha, sigma = read_real(toks[8]), read_real(toks[9])
            shell = toks[10]
            elow, eupp = read_real(toks[11]), read_real(toks[12])
        source = "database"
        if (alpha == 0.0 or sigma == 0.0) and shell:
            codes, flag = shell_code(shell)
            if flag == 0:
                if elow != 0.0 and eupp != 0.0:
                    nlo, nup = neff(rec["zn"], rec["ion"], rec["eplow"], rec["lambda0"], xi, xii, elow, eupp)
                else:
                    nlo, nup = neff(rec["zn"], rec["ion"], rec["eplow"], rec["lambda0"], xi, xii)
                s, a, ierr = collisional(nlo, nup, codes, tabs)
                if ierr == 0:
                    sigma, alpha, source = s, a, "abo_interpolated"
                else:
                    source = "unsold (abo ierr=%d)" % ierr
            else:
                source = "unsold"
        elif alpha == 0.0 or sigma == 0.0:
            source = "unsold"
        rec["sigma"], rec["alpha"], rec["collisional_source"] = float(sigma), float(alpha), source
        rows.append(rec)
    os.makedirs(os.path.dirname(OUT), exist_ok=True)
    with open(OUT, "w") as fh:
        json.dump({"source": "default/lines_database.dat + ABO interpolation (see header of "
                             "tests/golden/make_line_table.py)", "lines": rows}, fh, indent=1)
    for r in rows[:6]:
        print(r["id"], r["lambda0"], r["sigma"], r["alpha"], r["collisional_source"])


if __name__ == "__main__":
    sys.exit(main())
