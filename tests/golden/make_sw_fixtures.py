"""Writes tests/golden/sw_blocks.npz: SW derivative blocks at seeded geometries.

Run in a container where the reference is mounted read-only at /root/reference:

    python tests/golden/make_sw_fixtures.py

Two independent sources, stored side by side:

1. `ref_*`: the reference's own AutoDiffCalculator blocks.  The FastDifferentiation
   output under src/autodiff/generated_derivatives/SW/ is straight-line arithmetic
   (`var"##n" = expr`, `result[CartesianIndex(a,b,c)] = var"##n"`); this script
   reads those files as text, evaluates them line by line in FP64 with Python
   floats and stores only the resulting numbers.  That reproduces the baked
   constants, including cos(theta0) = -0.333333333333 (SW/H3_exec_ijk.jl:39).
   The iik block is built at run time by Symbolics (SW/H3_exec_iik.jl:1-30) and
   has no straight-line form; it comes from source 2 with cos(theta0) = -1/3.
2. `sym_*`: sympy derivatives of the SW energy terms (SW.jl:76-83) evaluated with
   mpmath at 40 significant digits, with the per-block cos(theta0) constants.

No reference source text is stored; the npz holds inputs and outputs only.
"""
from __future__ import annotations

import os
import re
import sys

import numpy as np

REF = "/root/reference/src/autodiff/generated_derivatives/SW"
OUT = os.path.join(os.path.dirname(os.path.abspath(__file__)), "sw_blocks.npz")

EPS, SIG, A_, LAM, GAM = 2.1683, 2.0951, 1.80, 21.0, 1.20
AA, BB, P, Q = 7.049556277, 0.6022245584, 4.0, 0.0
C0_FD = -0.333333333333
C0_IIK = -1.0 / 3.0
RC = A_ * SIG

BLOCKS3 = ["iij", "ijj", "iik", "ikk", "jjk", "jkk", "ijk"]
GROUPS3 = {"iij": (0, 0, 1), "ijj": (0, 1, 1), "iik": (0, 0, 2), "ikk": (0, 2, 2),
           "jjk": (1, 1, 2), "jkk": (1, 2, 2), "ijk": (0, 1, 2)}
BLOCKS2 = {"ij": (0, 1), "ik": (0, 2), "jk": (1, 2)}


def load_generated(path):
    """Evaluate a FastDifferentiation-generated Julia function as text."""
    src = open(path).read()
    stmts = []
    for line in src.splitlines():
        s = line.strip()
        if not s or s.startswith("#") or s.startswith("function") or s in ("end", "begin"):
            continue
        if s.startswith("result_element_type") or s.startswith("result = Array"):
            continue
        if s.startswith("return") or s.startswith("@inbounds") or s.startswith("#="):
            continue
        m = re.match(r'^(var"##\d+"|result\[CartesianIndex\(([\d, ]+)\)\])\s*=\s*(.+)$', s)
        if not m:
            continue
        lhs, idx, rhs = m.group(1), m.group(2), m.group(3)
        rhs = re.sub(r'var"##(\d+)"', r"v\1", rhs)
        rhs = re.sub(r"input_variables\[(\d+)\]", lambda mm: f"x[{int(mm.group(1)) - 1}]", rhs)
        rhs = re.sub(r"(\d)(v\d+|x\[)", r"\1*\2", rhs)      # 2var -> 2*var
        rhs = rhs.replace("^", "**")
        if idx is not None:
            key = tuple(int(t) - 1 for t in idx.split(","))
            stmts.append(("res", key, rhs))
        else:
            stmts.append(("var", lhs and re.sub(r'var"##(\d+)"', r"v\1", lhs), rhs))
    code = ["def f(x):", "    from math import sqrt, exp", "    res = {}"]
    for kind, tgt, rhs in stmts:
        if kind == "var":
            code.append(f"    {tgt} = {rhs}")
        else:
            code.append(f"    res[{tgt}] = {rhs}")
    code.append("    return res")
    ns = {}
    exec(compile("\n".join(code), path, "exec"), ns)
    return ns["f"]


def as_array(res, ndim):
    out = np.zeros((3,) * ndim)
    for k, v in res.items():
        out[k] = v
    return out


def sympy_blocks(r9_list, r3_list):
    import mpmath
    import sympy as sp

    mpmath.mp.dps = 40
    xs = sp.symbols("x0:9", real=True)
    ri, rj, rk = xs[0:3], xs[3:6], xs[6:9]
    rij = [ri[c] - rj[c] for c in range(3)]
    rik = [ri[c] - rk[c] for c in range(3)]
    dij = sp.sqrt(sum(t ** 2 for t in rij))
    dik = sp.sqrt(sum(t ** 2 for t in rik))
    c0 = sp.Symbol("c0")
    cs = sum(rij[c] * rik[c] for c in range(3)) / (dij * dik)
    le, gs, rc = sp.Float(repr(LAM * EPS), 30), sp.Float(repr(SIG * GAM), 30), sp.Float(repr(RC), 30)
    h = le * (cs - c0) ** 2 * sp.exp(gs / (dij - rc)) * sp.exp(gs / (dik - rc))
    r = sp.symbols("r0:3", real=True)
    d = sp.sqrt(sum(t ** 2 for t in r))
    sig = sp.Float(repr(SIG), 30)
    phi2 = sp.Float(repr(AA * EPS), 30) * (sp.Float(repr(BB), 30) * (sig / d) ** 4 - 1) * \
        sp.exp(sig / (d - rc))
    out3 = np.zeros((len(r9_list), 7, 3, 3, 3))
    out2 = np.zeros((len(r9_list), 3, 3, 3))
    for b, name in enumerate(BLOCKS3):
        g = GROUPS3[name]
        c0v = C0_IIK if name == "iik" else C0_FD
        hb = h.subs(c0, sp.Float(repr(c0v), 30))
        for a in range(3):
            for bb in range(3):
                da = sp.diff(hb, xs[3 * g[0] + a], xs[3 * g[1] + bb])
                for c in range(3):
                    e = sp.diff(da, xs[3 * g[2] + c])
                    f = sp.lambdify(xs, e, modules="mpmath")
                    for t, r9 in enumerate(r9_list):
                        out3[t, b, a, bb, c] = float(f(*[mpmath.mpf(repr(float(v))) for v in r9]))
        print("sym block", name, flush=True)
    for b, (name, g) in enumerate(BLOCKS2.items()):
        hb = h.subs(c0, sp.Float(repr(C0_FD), 30))
        for a in range(3):
            for bb in range(3):
                e = sp.diff(hb, xs[3 * g[0] + a], xs[3 * g[1] + bb])
                f = sp.lambdify(xs, e, modules="mpmath")
                for t, r9 in enumerate(r9_list):
                    out2[t, b, a, bb] = float(f(*[mpmath.mpf(repr(float(v))) for v in r9]))
    p3 = np.zeros((len(r3_list), 3, 3, 3))
    p2 = np.zeros((len(r3_list), 3, 3))
    for a in range(3):
        for bb in range(3):
            e2 = sp.diff(phi2, r[a], r[bb])
            f2 = sp.lambdify(r, e2, modules="mpmath")
            for t, r3 in enumerate(r3_list):
                p2[t, a, bb] = float(f2(*[mpmath.mpf(repr(float(v))) for v in r3]))
            for c in range(3):
                f3 = sp.lambdify(r, sp.diff(e2, r[c]), modules="mpmath")
                for t, r3 in enumerate(r3_list):
                    p3[t, a, bb, c] = float(f3(*[mpmath.mpf(repr(float(v))) for v in r3]))
    return out3, out2, p3, p2


def geometries(seed=20240):
    """Triplets (r_i, r_j, r_k) around diamond-silicon bond geometry, both
    neighbours inside r_cut, plus pair vectors."""
    rng = np.random.default_rng(seed)
    a = 5.43
    r9s, r3s = [], []
    bonds = a / 4 * np.array([[1, 1, 1], [1, -1, -1], [-1, 1, -1], [-1, -1, 1]], dtype=float)
    while len(r9s) < 20:
        ri = rng.uniform(0, 16.29, size=3)
        j, k = rng.choice(4, size=2, replace=False)
        rj = ri + bonds[j] + rng.normal(0, 0.12, 3)
        rk = ri + bonds[k] + rng.normal(0, 0.12, 3)
        if len(r9s) >= 16:   # a few stretched bonds close to the cutoff
            rj = ri + bonds[j] * rng.uniform(1.35, 1.55)
        uij, uik = ri - rj, ri - rk
        if np.linalg.norm(uij) < RC - 1e-3 and np.linalg.norm(uik) < RC - 1e-3:
            # the reference evaluates at [r_i; r_i - r_ij; r_i - r_ik]
            r9s.append(np.concatenate([ri, ri - uij, ri - uik]))
    while len(r3s) < 20:
        v = bonds[rng.integers(4)] * rng.uniform(0.9, 1.6) + rng.normal(0, 0.1, 3)
        if np.linalg.norm(v) < RC - 1e-3:
            r3s.append(v)
    return np.array(r9s), np.array(r3s)


def main():
    r9, r3 = geometries()
    ref3 = np.full((len(r9), 7, 3, 3, 3), np.nan)
    ref2 = np.zeros((len(r9), 3, 3, 3))
    for b, name in enumerate(BLOCKS3):
        if name == "iik":
            continue
        f = load_generated(os.path.join(REF, f"H3_exec_{name}.jl"))
        for t in range(len(r9)):
            ref3[t, b] = as_array(f(list(r9[t])), 3)
    for b, name in enumerate(BLOCKS2):
        f = load_generated(os.path.join(REF, f"H2_exec_{name}.jl"))
        for t in range(len(r9)):
            ref2[t, b] = as_array(f(list(r9[t])), 2)
    fpair3 = load_generated(os.path.join(REF, "H3_exec_ij.jl"))
    fpair2 = load_generated(os.path.join(REF, "H2_exec_pair.jl"))
    refp3 = np.array([as_array(fpair3(list(v)), 3) for v in r3])
    refp2 = np.array([as_array(fpair2(list(v)), 2) for v in r3])
    sym3, sym2, symp3, symp2 = sympy_blocks(r9, r3)
    np.savez_compressed(OUT, r9=r9, r3=r3, ref3=ref3, ref2=ref2, refp3=refp3, refp2=refp2,
                        sym3=sym3, sym2=sym2, symp3=symp3, symp2=symp2,
                        c0_fd=C0_FD, c0_iik=C0_IIK)
    # agreement report between the two sources
    m = ~np.isnan(ref3)
    print("max |ref3 - sym3| =", np.abs(ref3[m] - sym3[m]).max())
    print("max |ref2 - sym2| =", np.abs(ref2 - sym2).max())
    print("max |refp3 - symp3| =", np.abs(refp3 - symp3).max())
    print("max |refp2 - symp2| =", np.abs(refp2 - symp2).max())
    print("wrote", OUT)


if __name__ == "__main__":
    sys.exit(main())
