"""Compact adjoint point bodies, derived symbolically from a step's instruction lists.

For a step ``out_f(i) = E_f(s_g(i + delta), p)`` (loads at offsets delta of the
drifting frame, named literals p), input point q receives

    lam_prev_g(q) = sum over delta, f of  lam_f(q - delta) * dE_f/ds_g(i + delta) |_(i = q - delta)

and the parameter gradients are  sum over i, f of lam_f(i) * dE_f/dp(i).

``SymbolicAdjoint`` writes each E_f as a sympy expression (literals as the
exact values of their doubles, divisions by literals as products with exact
reciprocals), differentiates it, and emits one point body in which

* stencil entries whose partial derivative reads no state (a diffusion term's
  neighbours: a coefficient times lam) are grouped by coefficient, so the
  transposed stencil is one sum of lam values times one coefficient;
* state-dependent partials (the reaction's Jacobian) are evaluated from the
  state around q, renamed to offsets relative to q;
* a parameter whose derivative is affine in the state, c0 + sum c_L s_h(i + d'),
  is summed by parts over the periodic grid:
  sum_i lam_f(i) c_L s_h(i + d') = sum_q s_h(q) c_L lam_f(q - d'), so its
  gradient needs the point's own state and lam values the transposed stencil
  already reads.  Other parameters are evaluated at output i = q;
* common subexpressions are shared (sympy.cse).

Summing by parts moves a parameter's contributions between tiles, not their
total: the kernels sum each point's contribution once over the whole grid.
The body is derivative code: it is compiled with reassociation and
contraction allowed, and agrees with the instruction-by-instruction reverse
sweep to rounding.  Steps with branches, comparisons, index reads, reads of
manifest temporaries or non-floating values are not handled here
(``Unsupported``); their adjoints keep the reverse sweep of the instruction
list.
"""
from __future__ import annotations

import math
from fractions import Fraction
from typing import Callable, Dict, List, Tuple

import sympy as sp

from .emit import CTYPE, _const_literal, _f32, show_double


class Unsupported(Exception):
    """The step uses an operation the symbolic derivation does not handle."""


_UNARY = {"external-call/sqrt": sp.sqrt, "external-call/exp": sp.exp, "external-call/log": sp.log,
          "external-call/sin": sp.sin, "external-call/cos": sp.cos}
_LOADS = ("LoadCursorStatic", "LoadCursor", "LoadIndex")


def _tag(d) -> str:
    return "_".join(("m%d" % -x) if x < 0 else ("p%d" % x) for x in d)


class SymbolicAdjoint:
    """Symbolic adjoint of one step (``plan``: the step's GraphPlan, every
    manifest temporary inlined; ``fields``: state fields in layout order)."""

    def __init__(self, plan, dim: int, fields: List[Tuple[str, object]], params: List[str], real: str = "double",
                 max_ops: int = 4000):
        self.plan = plan
        self.dim = int(dim)
        self.fields = [n for n, _ in fields]
        self.fi = {n: i for i, n in enumerate(self.fields)}
        self.params = list(params)
        self.real = real
        self.psym = [sp.Symbol("P_%d" % i) for i in range(len(self.params))]
        self.loads: Dict[Tuple[int, tuple], sp.Symbol] = {}
        self.cnames: Dict[float, sp.Symbol] = {}
        self.cvals: Dict[sp.Symbol, sp.Rational] = {}
        self.exprs: List[Tuple[int, sp.Expr]] = []  # (output field index, E_f)
        s = plan.sleeve
        for k in plan.stores:
            node = plan.mm[k]
            base = plan.rng[k].offset - 2 * s
            self.exprs.append(self._node_expr(node, base))
        if sum(sp.count_ops(e) for _, e in self.exprs) > max_ops:
            raise Unsupported("step too large for the symbolic adjoint")

    # -- the step as expressions ----------------------------------------------
    def _const(self, v: float) -> sp.Expr:
        """A literal: small integers and powers of two as exact numbers (so that
        x*1, x - x and 2*x simplify); any other value as a symbol C_k that
        stands for that double, folded into one literal per constant
        subexpression when printed."""
        m, _ = math.frexp(abs(v)) if v else (0.5, 0)
        if v == int(v) and abs(v) <= 2 ** 20 or m == 0.5:
            return sp.Rational(v)
        if v not in self.cnames:
            self.cnames[v] = sp.Symbol("C_%d" % len(self.cnames))
            self.cvals[self.cnames[v]] = sp.Rational(v)
        return self.cnames[v]

    def _load(self, g: int, d: tuple) -> sp.Symbol:
        key = (g, tuple(d))
        if key not in self.loads:
            self.loads[key] = sp.Symbol("L_%d_%s" % (g, _tag(d)))
        return self.loads[key]

    def _node_expr(self, node, base):
        vals: List[sp.Expr] = []
        pidx = {p: i for i, p in enumerate(self.params)}
        for m in node.micro:
            k = m.inst[0]
            if k != "Store" and CTYPE.get(m.typ, m.typ) not in ("double", "float"):
                raise Unsupported("non-floating value (%s)" % m.typ)
            if m.src in pidx and k not in _LOADS:
                vals.append(self.psym[pidx[m.src]])
            elif k == "Imm":
                vals.append(self._const(float(Fraction(m.inst[1]))))
            elif k == "LoadCursorStatic":
                vals.append(self._load(self.fi[m.inst[2]], tuple(c + base for c in m.inst[1])))
            elif k == "Uniop":
                op, a = m.inst[1], vals[m.inst[2]]
                if op == "-":
                    vals.append(-a)
                elif op == "+":
                    vals.append(a)
                elif op in _UNARY:
                    vals.append(_UNARY[op](a))
                else:
                    raise Unsupported("unary %s" % op)
            elif k == "Binop":
                op, a, b = m.inst[1], vals[m.inst[2]], vals[m.inst[3]]
                if op == "+":
                    vals.append(a + b)
                elif op == "-":
                    vals.append(a - b)
                elif op == "*":
                    vals.append(a * b)
                elif op == "/":
                    vals.append(a / b)
                elif op == "**":
                    vals.append(a ** b)
                else:
                    raise Unsupported("binary %s" % op)
            elif k == "Store":
                return self.fi[m.inst[1]], vals[m.inst[2]]
            else:
                raise Unsupported(k)
        raise Unsupported("node without a store")

    # -- derivation -------------------------------------------------------------
    def _xsym(self, h: int, rho) -> sp.Symbol:
        return sp.Symbol("X_%d_%s" % (h, _tag(rho)))

    def _lsym(self, f: int, rho) -> sp.Symbol:
        return sp.Symbol("Y_%d_%s" % (f, _tag(rho)))

    def _at(self, e: sp.Expr, shift) -> sp.Expr:
        """``e`` with every load s_h(i + d') written as the state at q + d' + shift."""
        sub = {sym: self._xsym(h, [a + b for a, b in zip(d, shift)])
               for (h, d), sym in self.loads.items() if sym in e.free_symbols}
        return e.xreplace(sub)

    @staticmethod
    def _grouped(terms: List[Tuple[sp.Expr, sp.Expr]]) -> sp.Expr:
        """sum of coef * y over (coef, y) with equal coefficients grouped."""
        groups: Dict[sp.Expr, List[sp.Expr]] = {}
        for c, y in terms:
            if c != 0:
                groups.setdefault(c, []).append(y)
        return sp.Add(*[c * sp.Add(*ys) for c, ys in groups.items()])

    def outputs(self) -> List[sp.Expr]:
        """[lam_prev_g for every field g] + [the parameter gradient terms]."""
        nf = len(self.fields)
        lsyms = set(self.loads.values())
        acc: List[List[Tuple[sp.Expr, sp.Expr]]] = [[] for _ in range(nf)]
        for f, e in self.exprs:
            for (g, d), sym in self.loads.items():
                if sym not in e.free_symbols:
                    continue
                part = sp.diff(e, sym)
                neg = [-x for x in d]
                acc[g].append((self._at(part, neg), self._lsym(f, neg)))
        out = [self._grouped(a) for a in acc]
        for pi, P in enumerate(self.psym):
            terms = []
            for f, e in self.exprs:
                G = sp.diff(e, P)
                if G == 0:
                    continue
                used = [s for s in G.free_symbols if s in lsyms]
                firsts = {s: sp.diff(G, s) for s in used}
                if all(not (c.free_symbols & lsyms) for c in firsts.values()):
                    # affine in the state: summation by parts onto the point q
                    c0 = G.xreplace({s: 0 for s in used})
                    terms.append((c0, self._lsym(f, [0] * self.dim)))
                    inv = {sym: key for key, sym in self.loads.items()}
                    for s, c in firsts.items():
                        h, d = inv[s]
                        terms.append((c * self._xsym(h, [0] * self.dim), self._lsym(f, [-x for x in d])))
                else:
                    terms.append((self._at(G, [0] * self.dim), self._lsym(f, [0] * self.dim)))
            out.append(self._grouped(terms) if terms else sp.Integer(0))
        return out

    def _hoist(self, e: sp.Expr, defs: Dict[sp.Expr, sp.Symbol]) -> sp.Expr:
        """``e`` with each maximal sum or product of parameters and literals
        that is not a plain literal replaced by a symbol U_k (``defs``: its
        definition), so that a point evaluates it as one value."""
        uset = set(self.cvals) | set(self.psym)

        def uniform(x):
            return x.free_symbols <= uset

        def const(x):
            return x.free_symbols <= set(self.cvals)

        def name(g):
            if g not in defs:
                defs[g] = sp.Symbol("U_%d" % len(defs))
            return defs[g]

        def walk(x):
            if x.is_Atom:
                return x
            if uniform(x):
                return x if const(x) else name(x)
            if isinstance(x, (sp.Add, sp.Mul)):
                us = [a for a in x.args if uniform(a)]
                rest = [walk(a) for a in x.args if not uniform(a)]
                if us and not all(const(a) for a in us):
                    return x.func(name(x.func(*us)), *rest)
                return x.func(*(us + rest))
            return x.func(*[walk(a) for a in x.args])

        return walk(e)

    def state_offsets(self) -> set:
        """Offsets (relative to q) at which the derived body reads the state:
        {(0, ..., 0)} for a step whose state-dependent partials and parameter
        derivatives read only the cell itself (a pointwise reaction), empty for
        a step linear in the state without parameters."""
        offs = set()
        for e in self.outputs():
            for sym in e.free_symbols:
                if sym.name.startswith("X_"):
                    offs.add(tuple(_untag(sym.name.split("_", 2)[2])))
        return offs

    # -- C ------------------------------------------------------------------------
    def point_body(self, xload: Callable, lload: Callable, acc_names, pg_names) -> Tuple[List[str], bool]:
        """Statements that set ``acc_names[g]`` (in the library's real type) and
        ``pg_names[p]`` (double) for one point; ``xload(h, rho)``/``lload(f,
        rho)`` give the C expressions of the state and of lam at q + rho.  The
        parameter terms are evaluated in double also in an FP32 library (they
        cancel: dL/dF sums lam_u (1 - u) - lam_v v).  Returns (lines, uses_state)."""
        outs = self.outputs()
        nf = len(self.fields)
        used = set()
        lines = []
        # coefficients of the parameters (the same at every point) become one
        # value each, evaluated in double and rounded once to the library's type
        defs: Dict[sp.Expr, sp.Symbol] = {}
        outs = [self._hoist(e, defs) for e in outs]
        wide = self._printer(xload, lload, "double", used)
        for g, u in defs.items():
            if self.real == "double":
                lines.append("const double %s = %s;" % (u.name, wide(g)))
            else:
                lines.append("const %s %s = (%s)%s;" % (self.real, u.name, self.real, wide(g)))
        names = list(acc_names) + list(pg_names)
        if self.real == "double":  # one pass: the two share their lam sums
            passes = ((outs, names, "double", "sa_"),)
        else:
            passes = ((outs[:nf], names[:nf], self.real, "sa_"), (outs[nf:], names[nf:], "double", "sp_"))
        for exprs, nms, typ, pre in passes:
            if not nms:
                continue
            repl, red = sp.cse(exprs, symbols=sp.numbered_symbols(pre), optimizations="basic")
            c = self._printer(xload, lload, typ, used)
            for sym, e in repl:
                lines.append("const %s %s = %s;" % (typ, sym.name, c(e)))
            for i, (name, e) in enumerate(zip(nms, red)):
                t = typ if exprs is not outs or i < nf else "double"
                lines.append("%s %s = %s;" % (t, name, c(e)))
        return lines, "X" in used

    def _printer(self, xload: Callable, lload: Callable, T: str, used: set) -> Callable:
        """C printer of the derived expressions in type T ("double"/"float"):
        state and lam loads cast to T where the library's type differs,
        constant subexpressions folded into one literal of type T."""
        R = self.real
        cast = "(%s)" % T if T != R else ""

        def sym_name(s: sp.Symbol) -> str:
            n = s.name
            if n.startswith("X_"):
                used.add("X")
                _, h, tag = n.split("_", 2)
                return cast + xload(int(h), _untag(tag))
            if n.startswith("Y_"):
                _, f, tag = n.split("_", 2)
                return cast + lload(int(f), _untag(tag))
            if n.startswith("P_"):
                return "(%s)P[%s]" % (T, n[2:]) if T != "double" else "P[%s]" % n[2:]
            return n

        def lit(v) -> str:
            x = float(v)
            return _const_literal(_f32(x), "float") if T == "float" else show_double(x)

        cset = set(self.cvals)

        def const(e) -> bool:
            return e.free_symbols <= cset

        def value(e):
            """A constant subexpression evaluated exactly (or to 40 digits), rounded once."""
            x = e.xreplace(self.cvals)
            return x if x.is_Rational else sp.N(x, 40)

        def c(e) -> str:
            if isinstance(e, sp.Symbol) and e not in cset:
                return sym_name(e)
            if const(e):
                return lit(value(e))
            if isinstance(e, sp.Add):
                cs = [t for t in e.args if const(t)]
                if cs:
                    rest = sp.Add(*[t for t in e.args if not const(t)], evaluate=False)
                    k = value(sp.Add(*cs))
                    if k == 0:
                        return c(rest)
                    return "(%s %s %s)" % (c(rest), "-" if k < 0 else "+", lit(abs(k)))
                pos, neg = [], []
                for t in e.args:
                    co, rest = t.as_coeff_Mul()
                    if co.is_Number and co < 0:
                        neg.append(c(-t))
                    else:
                        pos.append(c(t))
                if pos:
                    s = " + ".join(pos) + "".join(" - " + n for n in neg)
                else:
                    s = "-" + " - ".join(neg)
                return "(" + s + ")"
            if isinstance(e, sp.Mul):
                cs = [t for t in e.args if const(t)]
                if cs:
                    k = value(sp.Mul(*cs))
                    rest = sp.Mul(*[t for t in e.args if not const(t)], evaluate=False)
                    if k == 1:
                        return c(rest)
                    if k == -1:
                        return "(-" + c(rest) + ")"
                    return "(%s*%s)" % (lit(k), c(rest))
                num, den = [], []
                for t in e.args:
                    if isinstance(t, sp.Pow) and t.exp.is_Integer and t.exp < 0:
                        den.append(c(t.base) if t.exp == -1 else c(sp.Pow(t.base, -t.exp)))
                    else:
                        num.append(c(t))
                s = "*".join(num) if num else lit(1)
                if den:
                    s = "(%s/(%s))" % (s, "*".join(den))
                return "(" + s + ")"
            if isinstance(e, sp.Pow):
                b, x = e.base, e.exp
                if x.is_Integer and 1 < x <= 4:
                    return "(" + "*".join([c(b)] * int(x)) + ")"
                if x == -1:
                    return "(%s/%s)" % (lit(1), c(b))
                if x == sp.Rational(1, 2):
                    return "sqrt(%s)" % c(b)
                if x == -sp.Rational(1, 2):
                    return "(%s/sqrt(%s))" % (lit(1), c(b))
                return "pow(%s, %s)" % (c(b), c(x))
            if isinstance(e, (sp.exp, sp.log, sp.sin, sp.cos)):
                return "%s(%s)" % (type(e).__name__, c(e.args[0]))
            raise Unsupported("cannot print %s" % type(e).__name__)

        return c


def _untag(tag: str) -> List[int]:
    return [(-int(t[1:]) if t[0] == "m" else int(t[1:])) for t in tag.split("_")]
