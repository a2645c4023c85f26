"""The symbolic adjoint point body (formura_amd/symadj.py) against the reverse
sweep of the instruction list, on the CPU.

Both bodies are compiled into one host library (g++) that applies them at every
point of a small periodic grid with random state, random lam and parameters
away from the program's own.  The gradient fields must agree to 1e-13 relative
(the bodies round differently: grouped coefficients, contraction), and the
parameter gradients summed over the grid to 1e-12: the symbolic body sums an
affine parameter derivative by parts onto the point's own state, so only the
sums over the periodic grid are equal, not the per-point contributions.  A
mistake planted in the symbolic derivation (one stencil coefficient, one
Jacobian entry, one parameter term) must fail the comparison.

The sweep body itself is checked against torch CPU autograd of the
Gray-Scott step and against central differences by the GPU tests of the
adjoint libraries (tests/test_adjoint.py); here it is the reference.
"""
import ctypes
import os
import re
import subprocess

import numpy as np
import pytest
import sympy as sp

from formura_amd import build as B
from formura_amd import symadj
from formura_amd.emit_adjoint import AdjointEmitter
from formura_amd.emit_hip import generate

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
WORK = os.path.join(ROOT, "formura_amd", "_build", "symadj_check")

# (program, config, parameters, grid) -- small grids of the same programs
CASES = {
    "grayscott2d": ("grayscott2d.fmr", "gs_64.yaml", ["f", "k", "ru", "rv"], (12, 10)),
    "diffusion2d": ("diffusion2d.fmr", "diffusion2d_64.yaml", [], (9, 11)),
    "heat3d": ("heat3d.fmr", "heat3d_16.yaml", [], (6, 7, 5)),
}


def _emitter(case, symbolic=True):
    fmr, yml, params, _ = CASES[case]
    _, _, mm = generate(os.path.join(B.PROG_DIR, fmr), os.path.join(B.PROG_DIR, yml), "symadj_" + case, "flat")
    return AdjointEmitter(mm, "symadj_" + case, params, "double", symbolic=symbolic)


def _harness(case, sym_lines, sweep_lines, nf, npar):
    grid = CASES[case][3]
    dim = len(grid)
    ivs = ["i%d" % a for a in range(dim)]

    def idx(a):
        terms = ["((%s + %d + %d) %% %d)" % (ivs[k], a[k], 4 * grid[k], grid[k]) for k in range(dim)]
        e = terms[0]
        for k in range(1, dim):
            e = "(%s)*%d + %s" % (e, grid[k], terms[k])
        return e

    def body(lines, fn):
        loops = "".join("for (int %s = 0; %s < %d; ++%s) " % (v, v, n, v) for v, n in zip(ivs, grid))
        out = ["extern \"C\" void %s(const double* X, const double* L, const double* P, double* A, double* PG) {" % fn,
               "  const long long n = %d;" % int(np.prod(grid)),
               "  for (int p = 0; p < %d; ++p) PG[p] = 0;" % max(1, npar),
               "  " + loops + "{",
               "    const long long q = %s;" % idx([0] * dim)]
        out += ["    " + l for l in lines]
        out += ["    A[%d * n + q] = acc%d;" % (g, g) for g in range(nf)]
        out += ["    PG[%d] += pg%d;" % (p, p) for p in range(max(1, npar))]
        out += ["  }", "}"]
        return "\n".join(out)

    src = ["#include <cmath>", "using std::sqrt; using std::exp; using std::log; using std::pow;",
           "#define XS(f) (X + (long long)(f) * %d)" % int(np.prod(grid)),
           "#define LS(f) (L + (long long)(f) * %d)" % int(np.prod(grid))]
    src.append(body(sym_lines, "adj_sym"))
    src.append(body(sweep_lines, "adj_sweep"))
    return "\n".join(src), idx


def _compile(case, sym_lines=None):
    em = _emitter(case)
    grid = CASES[case][3]
    dim = len(grid)
    nf, npar = len(em.fields), len(em.params)
    holder = {}

    def xload(f, a):
        return "XS(%d)[%s]" % (f, holder["idx"](a))

    def lload(f, a):
        return "LS(%d)[%s]" % (f, holder["idx"](a))

    _, holder["idx"] = _harness(case, [], [], nf, npar)
    if sym_lines is None:
        sym_lines, _ = em.point_body(xload, lload)
    sweep_lines, _ = em.sweep_point_body(xload, lload)
    src, _ = _harness(case, sym_lines, sweep_lines, nf, npar)
    os.makedirs(WORK, exist_ok=True)
    tag = "%s_%d" % (case, abs(hash("\n".join(sym_lines))) % 10 ** 8)
    cpp = os.path.join(WORK, tag + ".cpp")
    so = os.path.join(WORK, "lib" + tag + ".so")
    with open(cpp, "w") as f:
        f.write(src)
    subprocess.run(["g++", "-O1", "-shared", "-fPIC", "-o", so, cpp], check=True)
    lib = ctypes.CDLL(so)
    return lib, em, nf, npar, int(np.prod(grid))


def _run(lib, nf, npar, n, seed=5):
    rng = np.random.default_rng(seed)
    X = rng.uniform(0.05, 1.0, nf * n)
    L = rng.standard_normal(nf * n)
    P = np.array([0.035, 0.062, 1.7e-5, 0.9e-5][:max(1, npar)])
    out = {}
    for fn in ("adj_sym", "adj_sweep"):
        A = np.zeros(nf * n)
        PG = np.zeros(max(1, npar))
        dp = ctypes.POINTER(ctypes.c_double)
        getattr(lib, fn)(X.ctypes.data_as(dp), L.ctypes.data_as(dp), P.ctypes.data_as(dp),
                         A.ctypes.data_as(dp), PG.ctypes.data_as(dp))
        out[fn] = (A, PG)
    return out


def _agree(out, npar):
    (a1, p1), (a2, p2) = out["adj_sym"], out["adj_sweep"]
    ok = np.abs(a1 - a2).max() <= 1e-13 * np.abs(a2).max()
    if npar:
        ok = ok and bool(np.all(np.abs(p1 - p2) <= 1e-12 * np.abs(p2)))
    return ok, np.abs(a1 - a2).max() / np.abs(a2).max(), (np.abs(p1 - p2) / np.abs(p2) if npar else None)


@pytest.mark.parametrize("case", sorted(CASES))
def test_symbolic_body_matches_reverse_sweep(case):
    lib, em, nf, npar, n = _compile(case)
    ok, ea, ep = _agree(_run(lib, nf, npar, n), npar)
    assert ok, (case, ea, ep)


def test_symbolic_body_is_used_and_compact():
    """The Gray-Scott body is the grouped form: one transposed stencil per field
    (four lam loads summed, one coefficient), the Jacobian from the point's own
    (u, v), parameter gradients by parts; the HLLC step (branches) keeps the sweep."""
    em = _emitter("grayscott2d")
    sa = symadj.SymbolicAdjoint(em.plan, em.dim, em.fields, em.params)
    lines, uses = sa.point_body(lambda f, a: "X%d%s" % (f, a), lambda f, a: "L%d%s" % (f, a),
                                ["acc0", "acc1"], ["pg%d" % p for p in range(4)])
    text = "\n".join(lines)
    assert uses
    # the state enters only at the point itself; lam at the five stencil points
    assert set(re.findall(r"X\d\[[^\]]*\]", text)) == {"X0[0, 0]", "X1[0, 0]"}
    assert len(set(re.findall(r"L\d\[[^\]]*\]", text))) == 2 * 5 + 1
    assert len(lines) <= 20
    fmr, yml, params, real = B.ADJOINTS["euler3d_16_adj"]
    _, _, mm = generate(os.path.join(B.PROG_DIR, fmr), os.path.join(B.PROG_DIR, yml), "euler3d_16_adj", "flat")
    e3 = AdjointEmitter(mm, "euler3d_16_adj", params, real)
    with pytest.raises(symadj.Unsupported):
        symadj.SymbolicAdjoint(e3.plan, e3.dim, e3.fields, e3.params)


@pytest.mark.parametrize("plant", ["stencil", "jacobian", "param"])
def test_planted_mistakes_fail(plant):
    """A wrong stencil coefficient, a wrong Jacobian entry or a wrong parameter
    term in the symbolic body is caught by the comparison."""
    case = "grayscott2d"
    em = _emitter(case)
    grid = CASES[case][3]
    _, idx = _harness(case, [], [], 2, 4)
    sa = symadj.SymbolicAdjoint(em.plan, em.dim, em.fields, em.params)
    outs = sa.outputs()
    if plant == "stencil":
        y = sa._lsym(0, [0, 1])  # one neighbour's lam in lam_prev_u
        outs[0] = outs[0].xreplace({y: sp.Rational(101, 100) * y})
    elif plant == "jacobian":
        x = sa._xsym(1, [0, 0])  # v at the point: v^2 -> v^2 (1 + 1e-6)
        outs[1] = outs[1] + sp.Rational(1, 10 ** 6) * x * x * sa._lsym(1, [1, 1])
    else:
        outs[2 + 2] = outs[2 + 2] * sp.Rational(1000001, 1000000)
    sa.outputs = lambda: outs
    lines, _ = sa.point_body(lambda f, a: "XS(%d)[%s]" % (f, idx(a)), lambda f, a: "LS(%d)[%s]" % (f, idx(a)),
                             ["acc0", "acc1"], ["pg%d" % p for p in range(4)])
    lib, em, nf, npar, n = _compile(case, sym_lines=lines)
    ok, ea, ep = _agree(_run(lib, nf, npar, n), npar)
    assert not ok, (plant, ea, ep)
