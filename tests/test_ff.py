"""The float-float increment of the ``f32c`` ensemble form, checked on the CPU.

``emit_batch.ff_increment`` turns the increment dq of each update q + dq into
FP32 pairs (hi, lo) built with TwoProduct (fused multiply-add) and TwoSum; the
stencil part (the Laplacian) stays plain FP32 and its rounded coefficient
multiplies it as a pair.  The generated C lines are evaluated here by a
small interpreter with numpy float32 arithmetic (every operation rounded to
FP32, as on the device) and an exact fused multiply-add (rational arithmetic,
one rounding).  For random Gray-Scott states near the steady balance -- where
-u v^2 and F (1 - u) nearly cancel -- hi + lo must equal the exact increment
of the same FP32 inputs and FP64 parameters to about 2^-40 of the terms'
magnitude, while the plain FP32 increment (``f32k``) is off by about 2^-24.
"""
import os
import re
from fractions import Fraction

import numpy as np
import pytest

from formura_amd import build as B
from formura_amd.emit import Body
from formura_amd.emit_batch import ff_increment, BatchEmitter
from formura_amd.emit_hip import generate

f32 = np.float32


def fma(a, b, c):
    return f32(float(Fraction(float(a)) * Fraction(float(b)) + Fraction(float(c))))


class Lds:
    def __init__(self, vals):
        self.vals = vals

    def __getitem__(self, i):
        return self.vals[i]


def run_lines(lines, env):
    for line in lines:
        m = re.match(r"const float (\w+) = (.*);$", line.strip())
        if not m:
            continue
        expr = re.sub(r"(\d\.\d+e[+-]\d+)f", r"f32(\1)", m.group(2))
        expr = re.sub(r"(\d+\.\d*)f\b", r"f32(\1)", expr).replace("fmaf(", "fma(")
        env[m.group(1)] = f32(eval(expr, env))
    return env


def nodes():
    fmr, yml, _ = B.PROGRAMS["gs_64"]
    _, _, mm = generate(os.path.join(B.PROG_DIR, fmr), os.path.join(B.PROG_DIR, yml), "gs_64", "flat")
    em = BatchEmitter(mm, "t", "f32c", ["f", "k"], 16, 32, 4)
    return em, mm


def test_ff_increment_accuracy():
    em, mm = nodes()
    rng = np.random.default_rng(3)
    RY = 100
    s = em.plan.sleeve
    worst_ff, worst_plain = 0.0, 0.0
    # the diffusion term ru * (d_xx + d_yy) (rv for v) of each node: the product
    # whose left operand is the literal ru / rv
    diff_id = {}
    for kk in em.plan.stores:
        node = em.plan.mm[kk]
        diff_id[node.micro[-1].inst[1]] = next(
            j for j, m in enumerate(node.micro)
            if m.inst[0] == "Binop" and m.inst[1] == "*" and node.micro[m.inst[2]].src in ("ru", "rv"))
    for trial in range(40):
        F, k = 0.01 + 0.05 * rng.random(), 0.045 + 0.025 * rng.random()
        # a state near the balance -u v^2 + F (1 - u) = 0, small stencil differences
        u0 = 0.3 + 0.5 * rng.random()
        v0 = float(np.sqrt(F * (1 - u0) / u0))
        vals = {}
        grid = {}
        for di in range(-3, 4):
            for dj in range(-3, 4):
                grid[(di, dj)] = (f32(u0 * (1 + 1e-3 * rng.standard_normal())),
                                  f32(v0 * (1 + 1e-3 * rng.standard_normal())))
        l0 = {di * RY + dj: grid[(di, dj)][0] for (di, dj) in grid}
        l1 = {di * RY + dj: grid[(di, dj)][1] for (di, dj) in grid}
        P = [F, k]
        Ph = [f32(x) for x in P]
        Pl = [f32(x - float(h)) for x, h in zip(P, Ph)]
        for kk in em.plan.stores:
            node = em.plan.mm[kk]
            base = em.plan.rng[kk].offset - 2 * s
            fname = node.micro[-1].inst[1]

            def load(field, c, base=base):
                return "lds%d[o + (%d)*RY + (%d)]" % (em.fi[field], c[0] + base, c[1] + base)

            body = Body(node, load, lambda *a: None, lambda ax: None, lambda n, v: "", "s%d_" % kk,
                        params={"f": "P[0]", "k": "P[1]"}, real="float")
            q, dq = em._update_ids(node, fname, base)
            fl, hi, lo = ff_increment(node, body, dq, {"f": 0, "k": 1}, "c%d_" % kk)
            env = {"f32": f32, "fma": fma, "lds0": Lds(l0), "lds1": Lds(l1), "o": 0, "RY": RY,
                   "P": [f32(x) for x in P], "Ph": Ph, "Pl": Pl}
            run_lines(body.lines, env)
            run_lines(fl, env)
            got = Fraction(float(env[hi])) + Fraction(float(env[lo] if lo != "0.0f" else 0.0))
            plain = Fraction(float(env["s%d_%d" % (kk, dq)]))
            # exact increment of the same FP32 inputs (exact constants, FP64 parameters)
            c = (base, base)  # the cell itself: cursor (0, 0)

            def U(di, dj, fi):
                return Fraction(float(grid[(c[0] + di, c[1] + dj)][fi]))
            u, v = U(0, 0, 0), U(0, 0, 1)
            fi = em.fi[fname]
            qq = U(0, 0, fi)
            lap = ((U(1, 0, fi) + U(-1, 0, fi) - 2 * qq) + (U(0, 1, fi) + U(0, -1, fi) - 2 * qq)) / Fraction(float(1e-4))
            Ff, kf = Fraction(F), Fraction(k)
            # the stencil part (the Laplacian) as the kernel has it in plain FP32,
            # times the exact (FP64) coefficient; the reaction terms exact
            dm = node.micro[diff_id[fname]]
            diff = Fraction(float(body.vals[dm.inst[2]])) * Fraction(float(env["s%d_%d" % (kk, dm.inst[3])]))
            if fname == "u":
                exact = diff - u * v * v + Ff * (1 - u)
                full = Fraction(2e-5) * lap - u * v * v + Ff * (1 - u)
            else:
                exact = diff + u * v * v - (Ff + kf) * v
                full = Fraction(1e-5) * lap + u * v * v - (Ff + kf) * v
            scale = abs(u * v * v) + abs(Ff * (1 - u)) + abs((Ff + kf) * v)
            worst_ff = max(worst_ff, float(abs(got - exact) / scale))
            worst_plain = max(worst_plain, float(abs(plain - full) / scale))
            # all told, the pair is within the plain stencil part's own rounding
            assert abs(got - full) <= 2.0 ** -21 * abs(diff) + 2.0 ** -40 * scale
    assert worst_ff < 2.0 ** -40, worst_ff
    assert worst_plain > 2.0 ** -30, worst_plain   # the plain FP32 increment is really worse
    print("increment error / scale: float-float %.2e, plain FP32 %.2e" % (worst_ff, worst_plain))


def test_stencil_parts_stay_plain():
    """The Laplacian stays plain FP32 (no pair for a value that reads two or more
    cursors through + - and FP32-exact constant factors); the diffusivity, not
    an FP32 number, multiplies it as a pair."""
    em, mm = nodes()
    kk = em.plan.stores[0]
    node = em.plan.mm[kk]
    base = em.plan.rng[kk].offset - 2 * em.plan.sleeve
    body = Body(node, lambda f, c: "x", lambda *a: None, lambda ax: None, lambda n, v: "", "s_",
                params={"f": "P[0]", "k": "P[1]"}, real="float")
    q, dq = em._update_ids(node, node.micro[-1].inst[1], base)
    fl, hi, lo = ff_increment(node, body, dq, {"f": 0, "k": 1}, "c_")
    text = "\n".join(fl)
    # the d_xx / d_yy quotients (x 10000) are read as plain values, never paired
    assert "10000" not in text and "1.000000000e+04f" not in text
    assert "fmaf(" in text and "Ph[0]" in text and "Pl[0]" in text
    ru = np.float32(2e-5)
    assert ("%.9ef" % float(ru)) in text and ("%.9ef" % float(np.float32(2e-5 - float(ru)))) in text
