"""Extract the reference's atomic and HSRA data tables into a C data header.

Checker/build-time tool: runs only where the reference tree is present and
reads its source files as text.  It emits firtez_dz_amd/csrc/fz_atomdata.h,
which holds data only:

  * 92-element first/second ionisation potentials, abundances (Thevenin 1989)
    and atomic masses (atom_database.f90:47-119).  Default-real literals in
    the reference are REAL(4); they are stored here already rounded through
    float32 so the double-precision arithmetic sees the same values.
  * The polynomial partition-function fits U_I,II,III(T) and dlnU/dT of
    PARTITION_FUNCTION (atom_database.f90:121-1036).  Each element's fit is
    encoded as a predicated, branch-free stack program (opcodes below): an
    assignment under IF(cond) becomes  var = cond ? expr : var.  Constant
    sub-expressions are folded with Fortran kind rules (INTEGER, REAL(4),
    REAL(8)), so e.g. 3*1.326861 is a REAL(4) product as in the reference.
  * The HSRA model used for the continuum normalisation
    (normalization.f90:146-345, SET_HSRA_MODEL): T, Pg, rho, z at 387 heights.

The program format and its interpreter (fz_pf_eval in fz_physics.h) belong
to this project; only the numbers come from the reference.
"""
import os
import re
import sys

import numpy as np

REF = os.environ.get("FZ_REFERENCE", "/root/reference")
HERE = os.path.dirname(os.path.abspath(__file__))
OUT = os.path.join(HERE, "..", "firtez_dz_amd", "csrc", "fz_atomdata.h")

f32 = np.float32

# variables of the partition-function programs
VARS = ["T", "X", "Y", "DX", "UI", "UII", "UIII", "DUI", "DUII", "DUIII"]
V = {n: i for i, n in enumerate(VARS)}
# opcodes
OPS = ["END", "CONST", "VAR", "ADD", "SUB", "MUL", "DIV", "NEG", "EXP", "MAX", "AMAX1",
       "STORE", "GT", "LT", "LE", "GE", "NE", "EQ", "SEL", "AND"]
OP = {n: i for i, n in enumerate(OPS)}

# ---------------------------------------------------------------- tokenizer
TOK = re.compile(r"""
    (?P<num>(\d+\.\d*|\.\d+|\d+)([EeDd][-+]?\d+)?) |
    (?P<rel>\.(GT|LT|LE|GE|NE|EQ|AND|OR)\.) |
    (?P<id>[A-Za-z_][A-Za-z_0-9]*) |
    (?P<op>\*\*|[-+*/(),=])
""", re.X | re.I)


def tokenize(s):
    out, pos = [], 0
    s = s.strip()
    while pos < len(s):
        if s[pos].isspace():
            pos += 1
            continue
        m = TOK.match(s, pos)
        if not m:
            raise ValueError("cannot tokenize: %r at %d" % (s, pos))
        if m.group("num"):
            txt = m.group("num")
            if re.search(r"[Dd]", txt):
                out.append(("num", ("r8", float(txt.replace("D", "e").replace("d", "e")))))
            elif "." in txt or re.search(r"[Ee]", txt):
                out.append(("num", ("r4", f32(float(txt)))))
            else:
                out.append(("num", ("int", int(txt))))
        elif m.group("rel"):
            out.append(("rel", m.group("rel").upper().strip(".")))
        elif m.group("id"):
            out.append(("id", m.group("id").upper()))
        else:
            out.append(("op", m.group("op")))
        pos = m.end()
    return out


# ----------------------------------------------------- expression nodes
# ("const", kind, value) | ("var", name) | ("bin", op, a, b) | ("neg", a)
# | ("call", fname, [args]) | ("rel", op, a, b) | ("and", a, b)
RANK = {"int": 0, "r4": 1, "r8": 2}


def conv(kind, val, to):
    if to == "int":
        return int(val)
    if to == "r4":
        return f32(val)
    return float(val)


def fold_bin(op, a, b):
    """Fold a binary op of two constants with Fortran kind promotion."""
    k = max(a[1], b[1], key=lambda q: RANK[q])
    x, y = conv(a[1], a[2], k), conv(b[1], b[2], k)
    if k == "int":
        if op == "/":
            r = abs(x) // abs(y) * (1 if (x >= 0) == (y >= 0) else -1)
        else:
            r = {"+": x + y, "-": x - y, "*": x * y}[op]
    else:
        with np.errstate(all="ignore"):
            r = {"+": x + y, "-": x - y, "*": x * y, "/": x / y}[op]
        r = conv(k, r, k)
    return ("const", k, r)


class Parser:
    def __init__(self, toks):
        self.t, self.i = toks, 0

    def peek(self):
        return self.t[self.i] if self.i < len(self.t) else (None, None)

    def take(self, kind=None, val=None):
        tk = self.peek()
        if kind and tk[0] != kind or val and tk[1] != val:
            raise ValueError("expected %s %s, got %s" % (kind, val, tk))
        self.i += 1
        return tk

    def logical(self):
        a = self.relation()
        while self.peek() == ("rel", "AND"):
            self.take()
            a = ("and", a, self.relation())
        return a

    def relation(self):
        a = self.expr()
        tk = self.peek()
        if tk[0] == "rel" and tk[1] not in ("AND", "OR"):
            self.take()
            return ("rel", tk[1], a, self.expr())
        return a

    def expr(self):
        neg = False
        if self.peek() in (("op", "-"), ("op", "+")):
            neg = self.take()[1] == "-"
        a = self.term()
        if neg:
            a = ("const", a[1], -a[2]) if a[0] == "const" else ("neg", a)
        while self.peek() in (("op", "+"), ("op", "-")):
            op = self.take()[1]
            a = self.binop(op, a, self.term())
        return a

    def term(self):
        a = self.factor()
        while self.peek() in (("op", "*"), ("op", "/")):
            op = self.take()[1]
            a = self.binop(op, a, self.factor())
        return a

    def factor(self):
        a = self.primary()
        if self.peek() == ("op", "**"):
            raise ValueError("** not used by the partition-function fits")
        return a

    def primary(self):
        tk = self.take()
        if tk[0] == "num":
            return ("const", tk[1][0], tk[1][1])
        if tk == ("op", "("):
            a = self.expr()
            self.take("op", ")")
            return a
        if tk[0] == "id":
            if self.peek() == ("op", "("):
                self.take()
                args = [self.expr()]
                while self.peek() == ("op", ","):
                    self.take()
                    args.append(self.expr())
                self.take("op", ")")
                return ("call", tk[1], args)
            return ("var", tk[1])
        raise ValueError("bad primary %s" % (tk,))

    @staticmethod
    def binop(op, a, b):
        if a[0] == "const" and b[0] == "const":
            return fold_bin(op, a, b)
        return ("bin", op, a, b)


# ------------------------------------------------------------- codegen
class Prog:
    def __init__(self):
        self.code, self.consts = [], []

    def const(self, node):
        k, v = node[1], node[2]
        val = float(f32(v)) if k == "r4" else float(v)
        self.consts.append(val)
        self.code.append((OP["CONST"], len(self.consts) - 1))

    def emit_expr(self, n):
        kind = n[0]
        if kind == "const":
            self.const(n)
        elif kind == "var":
            self.code.append((OP["VAR"], V[n[1]]))
        elif kind == "neg":
            self.emit_expr(n[1])
            self.code.append((OP["NEG"], 0))
        elif kind == "bin":
            self.emit_expr(n[2])
            self.emit_expr(n[3])
            self.code.append((OP[{"+": "ADD", "-": "SUB", "*": "MUL", "/": "DIV"}[n[1]]], 0))
        elif kind == "call":
            f, args = n[1], n[2]
            for a in args:
                self.emit_expr(a)
            if f == "EXP":
                self.code.append((OP["EXP"], 0))
            elif f in ("MAX", "AMAX1"):
                for _ in range(len(args) - 1):
                    self.code.append((OP[f], 0))
            else:
                raise ValueError("unknown function " + f)
        elif kind == "rel":
            self.emit_expr(n[2])
            self.emit_expr(n[3])
            self.code.append((OP[n[1]], 0))
        elif kind == "and":
            self.emit_expr(n[1])
            self.emit_expr(n[2])
            self.code.append((OP["AND"], 0))
        else:
            raise ValueError(n)

    def assign(self, var, expr, conds):
        self.emit_expr(expr)
        if conds:
            self.code.append((OP["VAR"], V[var]))
            c = conds[0]
            for extra in conds[1:]:
                c = ("and", c, extra)
            self.emit_expr(c)
            self.code.append((OP["SEL"], 0))
        self.code.append((OP["STORE"], V[var]))


def parse_statements(lines, prog):
    conds = []
    for ln in lines:
        up = ln.upper().strip()
        if not up:
            continue
        m = re.match(r"IF\s*\((.*)\)\s*THEN$", up)
        if m:
            conds.append(Parser(tokenize(m.group(1))).logical())
            continue
        if up in ("ENDIF", "END IF"):
            conds.pop()
            continue
        m = re.match(r"IF\s*\((.*?)\)\s*([A-Z]+\s*=.*)$", up)
        if m and m.group(2).split("=")[0].strip() in V:
            # single-line IF: find the matching parenthesis of the condition
            depth, k = 0, up.index("(")
            for k in range(up.index("("), len(up)):
                depth += {"(": 1, ")": -1}.get(up[k], 0)
                if depth == 0:
                    break
            cond = Parser(tokenize(up[up.index("(") + 1:k])).logical()
            lhs, rhs = up[k + 1:].split("=", 1)
            prog.assign(lhs.strip(), Parser(tokenize(rhs)).expr(), conds + [cond])
            continue
        lhs, rhs = up.split("=", 1)
        p = Parser(tokenize(rhs))
        e = p.expr()
        if p.i != len(p.t):
            raise ValueError("trailing tokens in " + up)
        prog.assign(lhs.strip(), e, list(conds))
    assert not conds


def partition_programs(text):
    body = text[text.index("PURE SUBROUTINE PARTITION_FUNCTION"):text.index("END SUBROUTINE PARTITION_FUNCTION")]
    body = body[body.index("SELECT CASE"):]
    lines = []
    for raw in body.split("\n"):
        s = raw.split("!")[0].rstrip()
        if lines and lines[-1].endswith("&"):
            lines[-1] = lines[-1][:-1] + " " + s.strip()
        else:
            lines.append(s)
    cases, cur = {}, None
    for ln in lines:
        m = re.match(r"\s*CASE\s*\((\d+)\)\s*$", ln, re.I)
        if m:
            cur = int(m.group(1))
            cases[cur] = []
            continue
        if re.match(r"\s*(END SELECT|SELECT CASE)", ln, re.I):
            continue
        if cur is not None:
            cases[cur].append(ln)
    assert sorted(cases) == list(range(1, 93)), sorted(cases)
    progs = []
    for z in range(1, 93):
        p = Prog()
        parse_statements(cases[z], p)
        p.code.append((OP["END"], 0))
        progs.append(p)
    return progs


def real_array(text, name, sp_default=True):
    m = re.search(name + r"\(:\)\s*=\s*\(/(.*?)/\)", text, re.S)
    toks = re.findall(r"[-+]?(?:\d+\.\d*|\.\d+|\d+)(?:[EeDd][-+]?\d+)?", m.group(1).replace("&", " "))
    out = []
    for t in toks:
        if re.search(r"[Dd]", t):
            out.append(float(t.replace("D", "e").replace("d", "e")))
        else:
            out.append(float(f32(float(t))) if sp_default else float(t))
    return out


def c_doubles(vals, per=4):
    rows = []
    for i in range(0, len(vals), per):
        rows.append("  " + ", ".join(repr(float(v)) if np.isfinite(v) else "0.0" for v in vals[i:i + per]))
    return ",\n".join(rows)


def main():
    src = os.path.join(REF, "src")
    atom = open(os.path.join(src, "atom_database.f90")).read()
    init = atom[atom.index("SUBROUTINE ATOM_INIT"):atom.index("END SUBROUTINE ATOM_INIT")]
    xi = real_array(init, "XI")
    xii = real_array(init, "XII")
    ab = real_array(init, r"\bABUND")
    ma = real_array(init, "MATOM")
    assert len(xi) == len(xii) == len(ab) == len(ma) == 92
    progs = partition_programs(atom)

    norm = open(os.path.join(src, "normalization.f90")).read()
    hs = norm[norm.index("SUBROUTINE SET_HSRA_MODEL\n"):norm.index("END SUBROUTINE SET_HSRA_MODEL\n")]
    h_t, h_p, h_r, h_z = (real_array(hs, n) for n in ("HSRA_TEM", "HSRA_PG", "HSRA_RHO", "HSRA_Z"))
    assert len(h_t) == len(h_p) == len(h_r) == len(h_z) == 387

    code, consts, start = [], [], [0]
    for p in progs:
        base = len(consts)
        for op, arg in p.code:
            code.append((op, arg + base if op == OP["CONST"] else arg))
        consts.extend(p.consts)
        start.append(len(code))

    with open(OUT, "w") as fh:
        fh.write("/* Generated by oracle/extract_tables.py from the reference's data tables\n"
                 " * (atom_database.f90:47-119, 121-1036; normalization.f90:146-345).\n"
                 " * Data only: element constants, partition-function programs, HSRA model.\n"
                 " * Do not edit by hand. */\n#pragma once\n\n")
        fh.write("#ifndef FZ_DATA_QUAL\n#define FZ_DATA_QUAL static\n#endif\n")
        fh.write("#define FZ_NELEM 92\n#define FZ_HSRA_LEN 387\n")
        fh.write("#define FZ_PF_NCODE %d\n#define FZ_PF_NCONST %d\n\n" % (len(code), len(consts)))
        for i, n in enumerate(OPS):
            fh.write("#define FZ_PF_%s %d\n" % (n, i))
        for i, n in enumerate(VARS):
            fh.write("#define FZ_PV_%s %d\n" % (n, i))
        fh.write("\n")
        for name, arr in (("fz_xi", xi), ("fz_xii", xii), ("fz_abund", ab), ("fz_matom", ma)):
            fh.write("FZ_DATA_QUAL const double %s[FZ_NELEM] = {\n%s};\n" % (name, c_doubles(arr)))
        fh.write("FZ_DATA_QUAL const unsigned short fz_pf_start[FZ_NELEM + 1] = {\n  %s};\n"
                 % ", ".join(str(s) for s in start))
        fh.write("FZ_DATA_QUAL const unsigned char fz_pf_op[FZ_PF_NCODE] = {\n  %s};\n"
                 % ", ".join(str(op) for op, _ in code))
        fh.write("FZ_DATA_QUAL const unsigned short fz_pf_arg[FZ_PF_NCODE] = {\n  %s};\n"
                 % ", ".join(str(a) for _, a in code))
        fh.write("FZ_DATA_QUAL const double fz_pf_const[FZ_PF_NCONST] = {\n%s};\n" % c_doubles(consts))
        for name, arr in (("fz_hsra_t", h_t), ("fz_hsra_pg", h_p), ("fz_hsra_rho", h_r), ("fz_hsra_z", h_z)):
            fh.write("FZ_DATA_QUAL const float %s[FZ_HSRA_LEN] = {\n%s};\n"
                     % (name, c_doubles([float(f32(v)) for v in arr], 6)))
    print("wrote", OUT, "ops", len(code), "consts", len(consts))


if __name__ == "__main__":
    sys.exit(main())
