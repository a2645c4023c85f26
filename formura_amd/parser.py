"""Parser for ``.fmr`` programs.

A hand-written recursive-descent parser that accepts the grammar of
``src/Formura/Parser.hs``:

* whitespace never contains a newline; ``#`` comments and ``\\``-newline
  continuations are whitespace (:52-58, :88-99), and statements are separated
  by ``;`` or newlines (:304-307);
* identifiers are alphabetic (letter, then alnum/``_``/``'``) or symbolic runs
  (:113-145); keywords (:77-85) are not identifiers, and a keyword must not be
  followed by a character of its own class (:67-74);
* the operator table (:199-211), highest first: ``.`` (right), ``**``,
  ``* /``, prefix ``+ -``, ``+ -``, ``&&``, ``||``, ``<? >? <%`` (n-ary),
  comparisons (non-associative);
* application is juxtaposition and binds tighter than every operator; a
  postfix ``[..]`` is grid access (:249-263);
* a one-element tuple is a parenthesised expression (:162-172);
* ``begin function out = f(in) ... end function`` and the ``returns`` form
  become ``f = fun(in) let ... in out`` (:314-346);
* ``dx``/``total_grid_x`` are injected as exact rationals (:493-511).
"""
from __future__ import annotations

from fractions import Fraction
from typing import List

from .config import NumericalConfig
from .syntax import (Apply, Binding, Binop, ElemType, FunType, Grid, GridType,
                     Ident, Imm, Lambda, Let, NPlusK, Naryop, Program,
                     SpecialDecl, Subst, TopType, Triop, Tuple, TupleType,
                     TypeDecl, Uniop, Vector, VectorType, ELEMENT_TYPENAMES)

COMPARISON_OPS = ["<=", "==", "!=", ">=", "<", ">"]
MINMAX_OPS = ["<?", ">?", "<%"]
KEYWORDS = set(["begin", "end", "function", "returns", "let", "in", "fun",
                "dimension", "axes", "grid_struct_type_name", "grid_struct_instance_name",
                "if", "then", "else", "const", "extern", "manifest",
                "+", "-", "*", "/", ".", "**", "::", "=", ",", "&&", "||"]
               + COMPARISON_OPS + MINMAX_OPS)


class ParseError(Exception):
    pass


def _is_alpha0(c):
    return c.isalpha()


def _is_alpha1(c):
    return c.isalnum() or c == "_" or c == "'"


def _is_symbol(c):
    return c.isprintable() and not (_is_alpha1(c) or c.isspace() or c in "\"#();[\\]{}")


class _Fail(Exception):
    """Backtrackable failure."""


class Parser:
    def __init__(self, text: str, filename: str = "<fmr>"):
        self.s = text
        self.i = 0
        self.fn = filename

    # -- low level ------------------------------------------------------------
    def peek(self, k=0):
        j = self.i + k
        return self.s[j] if j < len(self.s) else ""

    def err(self, msg):
        line = self.s.count("\n", 0, self.i) + 1
        col = self.i - (self.s.rfind("\n", 0, self.i) + 1) + 1
        raise ParseError("%s:%d:%d: %s" % (self.fn, line, col, msg))

    def ws(self):
        """``someSpace`` without newlines (Parser.hs:52-58)."""
        while True:
            c = self.peek()
            if c in (" ", "\t", "\f", "\v"):
                self.i += 1
            elif c == "#":
                while self.peek() not in ("", "\n"):
                    self.i += 1
            elif c == "\\":
                j = self.i + 1
                while j < len(self.s) and self.s[j] in " \t":
                    j += 1
                if j < len(self.s) and self.s[j] == "\n":
                    self.i = j + 1
                elif j + 1 < len(self.s) and self.s[j] == "\r" and self.s[j + 1] == "\n":
                    self.i = j + 2
                else:
                    return
            else:
                return

    def try_keyword(self, k):
        if not self.s.startswith(k, self.i):
            return False
        nxt = self.peek(len(k))
        cls = _is_alpha1 if _is_alpha1(k[-1]) else _is_symbol
        if nxt and cls(nxt):
            return False
        self.i += len(k)
        self.ws()
        return True

    def keyword(self, k):
        if not self.try_keyword(k):
            self.err("expected keyword %r" % k)

    def try_symbolic(self, c):
        if self.peek() == c:
            self.i += 1
            self.ws()
            return True
        return False

    def symbolic(self, c):
        if not self.try_symbolic(c):
            self.err("expected %r" % c)

    def try_ident_name(self, mode="SA"):
        j = self.i
        if "S" in mode and _is_symbol(self.peek()):
            while j < len(self.s) and _is_symbol(self.s[j]):
                j += 1
        elif "A" in mode and _is_alpha0(self.peek()):
            j += 1
            while j < len(self.s) and _is_alpha1(self.s[j]):
                j += 1
        else:
            return None
        name = self.s[self.i:j]
        if name in KEYWORDS:
            return None
        self.i = j
        self.ws()
        return name

    def try_number(self):
        """``naturalOrScientific`` as an exact rational (Parser.hs:459-462)."""
        j = self.i
        if not self.peek().isdigit():
            return None
        while j < len(self.s) and self.s[j].isdigit():
            j += 1
        is_sci = False
        if j < len(self.s) and self.s[j] == "." and j + 1 < len(self.s) and self.s[j + 1].isdigit():
            j += 1
            while j < len(self.s) and self.s[j].isdigit():
                j += 1
            is_sci = True
        if j < len(self.s) and self.s[j] in "eE":
            k = j + 1
            if k < len(self.s) and self.s[k] in "+-":
                k += 1
            if k < len(self.s) and self.s[k].isdigit():
                while k < len(self.s) and self.s[k].isdigit():
                    k += 1
                j = k
                is_sci = True
        text = self.s[self.i:j]
        self.i = j
        self.ws()
        return Fraction(text) if is_sci else Fraction(int(text))

    def at_statement_end(self):
        c = self.peek()
        return c in ("", "\n", "\r", ";")

    # -- expressions ----------------------------------------------------------
    def expr(self, term):
        """``exprOf`` (Parser.hs:199-211), lowest level first."""
        return self._cmp(term)

    def _cmp(self, term):
        a = self._minmax(term)
        for op in COMPARISON_OPS:
            if self.try_keyword(op):
                b = self._minmax(term)
                return Binop(op, a, b)
        return a

    def _minmax(self, term):
        a = self._or(term)
        while True:
            for op in MINMAX_OPS:
                if self.try_keyword(op):
                    b = self._or(term)
                    ap = a.xs if isinstance(a, Naryop) and a.op == op else [a]
                    bp = b.xs if isinstance(b, Naryop) and b.op == op else [b]
                    a = Naryop(op, ap + bp)
                    break
            else:
                return a

    def _or(self, term):
        a = self._and(term)
        while self.try_keyword("||"):
            a = Binop("||", a, self._and(term))
        return a

    def _and(self, term):
        a = self._add(term)
        while self.try_keyword("&&"):
            a = Binop("&&", a, self._add(term))
        return a

    def _add(self, term):
        a = self._prefix(term)
        while True:
            if self.try_keyword("+"):
                a = Binop("+", a, self._prefix(term))
            elif self.try_keyword("-"):
                a = Binop("-", a, self._prefix(term))
            else:
                return a

    def _prefix(self, term):
        if self.try_keyword("+"):
            return Uniop("+", self._mul(term))
        if self.try_keyword("-"):
            return Uniop("-", self._mul(term))
        return self._mul(term)

    def _mul(self, term):
        a = self._pow(term)
        while True:
            if self.try_keyword("*"):
                a = Binop("*", a, self._pow(term))
            elif self.try_keyword("/"):
                a = Binop("/", a, self._pow(term))
            else:
                return a

    def _pow(self, term):
        a = self._dot(term)
        while self.try_keyword("**"):
            a = Binop("**", a, self._dot(term))
        return a

    def _dot(self, term):
        a = term()
        if self.try_keyword("."):
            return Binop(".", a, self._dot(term))
        return a

    def rexpr(self):
        return self.expr(self.fexpr)

    def fexpr(self):
        """Application chain (Parser.hs:249-263)."""
        f = self.aexpr()
        if f is None:
            self.err("expected an expression")
        while True:
            if self.peek() == "[":
                f = Grid(self.grid_indices(self.npk), f)
                continue
            save = self.i
            x = self.aexpr()
            if x is None:
                self.i = save
                return f
            f = Apply(f, x)

    def aexpr(self):
        """``aexpr`` (Parser.hs:266-267); returns None without consuming on failure."""
        if self.peek() == "(":
            return self.tuple_of(self.rexpr)
        if self.try_keyword("let"):
            b = self.binding()
            self.keyword("in")
            return Let(b, self.rexpr())
        if self.try_keyword("fun"):
            lhs = self.tuple_of(self.lexpr)
            return Lambda(lhs, self.rexpr())
        if self.try_keyword("if"):
            c = self.rexpr()
            self.keyword("then")
            x = self.rexpr()
            self.keyword("else")
            y = self.rexpr()
            return Triop("ite", c, x, y)
        name = self.try_ident_name("SA")
        if name is not None:
            return Ident(name)
        n = self.try_number()
        if n is not None:
            return Imm(n)
        return None

    def tuple_of(self, p):
        self.symbolic("(")
        xs = []
        if not self.try_symbolic(")"):
            xs.append(p())
            while self.try_symbolic(","):
                xs.append(p())
            self.symbolic(")")
        if len(xs) == 1:
            return xs[0]
        return Tuple(xs)

    def grid_indices(self, p):
        self.symbolic("[")
        xs = []
        if not self.try_symbolic("]"):
            xs.append(p())
            while self.try_symbolic(","):
                xs.append(p())
            self.symbolic("]")
        return xs

    def npk(self):
        """``nPlusK`` (Parser.hs:183-191)."""
        name = self.try_ident_name("A") or ""
        k = Fraction(0)
        if self.try_symbolic("+"):
            k = self.const_rational_expr()
        elif self.try_symbolic("-"):
            k = -self.const_rational_expr()
        if self.peek() not in (",", "]"):
            self.err("malformed grid index")
        return NPlusK(name, k)

    def const_rational_expr(self):
        """``constRationalExpr`` evaluated with ``evalCRE`` (Parser.hs:444-457)."""
        e = self.expr(self._imm_term)
        return _eval_cre(e, self)

    def _imm_term(self):
        n = self.try_number()
        if n is None:
            if self.peek() == "(":
                return self.tuple_of(lambda: self.expr(self._imm_term))
            self.err("expected a rational literal")
        return Imm(n)

    # -- l-expressions and types ---------------------------------------------
    def lexpr(self):
        if self.peek() == "(":
            f = self.tuple_of(self.lexpr)
        else:
            name = self.try_ident_name("SA")
            if name is None:
                self.err("expected an l-expression")
            f = Ident(name)
        while True:
            if self.peek() == "[":
                f = Grid(self.grid_indices(self.npk), f)
            elif self.peek() == "(":
                save = self.i
                self.symbolic("(")
                name = self.try_ident_name("SA")
                if name is not None and self.try_symbolic(")"):
                    f = Vector(name, f)
                else:
                    self.i = save
                    return f
            else:
                return f

    def try_type_fexpr(self):
        save = self.i
        t = None
        if self.peek() == "(":
            try:
                t = self.tuple_of(self._type_fexpr_strict)
                if isinstance(t, Tuple):
                    t = TupleType(t.items)
            except ParseError:
                self.i = save
                return None
        elif self.try_keyword("function"):
            t = FunType()
        else:
            name = self.try_ident_name("SA")
            if name is None or name not in ELEMENT_TYPENAMES:
                self.i = save
                return None
            t = ElemType(name)
        while True:
            if self.peek() == "[":
                offs = self.grid_indices(self.const_rational_expr)
                t = GridType(offs, t)
            elif self.peek() == "(":
                s2 = self.i
                self.symbolic("(")
                n = self.try_number()
                if n is not None and n.denominator == 1 and self.try_symbolic(")"):
                    t = VectorType(int(n), t)
                else:
                    self.i = s2
                    return t
            else:
                return t

    def _type_fexpr_strict(self):
        t = self.try_type_fexpr()
        if t is None:
            self.err("expected a type")
        return t

    def try_mod_type(self):
        """``modTypeExpr <* keyword "::"`` under ``try`` (Parser.hs:351,409-419)."""
        save = self.i
        mods = []

        def modifiers():
            while True:
                for kw, m in (("const", "const"), ("extern", "extern"), ("manifest", "manifest")):
                    if self.try_keyword(kw):
                        mods.append(m)
                        break
                else:
                    return

        modifiers()
        t = self.try_type_fexpr()
        modifiers()
        if self.try_keyword("::"):
            return mods, (t if t is not None else TopType())
        self.i = save
        return None

    # -- statements -------------------------------------------------------------
    def statement_delimiter(self):
        n = 0
        while True:
            if self.try_symbolic(";"):
                n += 1
            elif self.peek() == "\n":
                self.i += 1
                self.ws()
                n += 1
            elif self.peek() == "\r" and self.peek(1) == "\n":
                self.i += 2
                self.ws()
                n += 1
            else:
                return n > 0

    def binding(self, terminators=("end", "in")):
        stmts = []
        while True:
            save = self.i
            if any(self._lookahead_keyword(t) for t in terminators) or self.peek() == "":
                break
            got = self.statement_compound()
            stmts.extend(got)
            if not self.statement_delimiter():
                break
        return Binding(stmts)

    def _lookahead_keyword(self, k):
        save = self.i
        ok = self.try_keyword(k)
        self.i = save
        return ok

    def statement_compound(self):
        if self.try_keyword("begin"):
            self.keyword("function")
            save = self.i
            fn = inx = outx = None
            # returns-form: f(x) returns y
            name = self.try_ident_name("SA")
            if name is not None and self.peek() == "(":
                try:
                    inx_ = self.tuple_of(self.lexpr)
                    if self.try_keyword("returns"):
                        fn, inx, outx = Ident(name), inx_, self.rexpr()
                except ParseError:
                    pass
            if fn is None:
                self.i = save
                outx = self.rexpr()
                self.keyword("=")
                name = self.try_ident_name("SA")
                if name is None:
                    self.err("Malformed Function Syntax")
                fn = Ident(name)
                inx = self.tuple_of(self.lexpr)
            if not self.statement_delimiter():
                self.err("expected a newline after the function header")
            b = self.binding(terminators=("end",))
            self.keyword("end")
            self.keyword("function")
            return [Subst(fn, Lambda(inx, Let(b, outx)))]
        mt = self.try_mod_type()
        if mt is not None:
            mods, typ = mt
            lamrs = []
            while True:
                lhs = self.lexpr()
                rhs = self.rexpr() if self.try_keyword("=") else None
                lamrs.append((lhs, rhs))
                if not self.try_symbolic(","):
                    break
            return ([TypeDecl(mods, typ, l) for l, _ in lamrs]
                    + [Subst(l, r) for l, r in lamrs if r is not None])
        lhs = self.lexpr()
        self.keyword("=")
        rhs = self.rexpr()
        return [Subst(lhs, rhs)]

    def special_declaration(self):
        if self.try_keyword("dimension"):
            self.keyword("::")
            n = self.try_number()
            if n is None or n.denominator != 1:
                self.err("dimension must be a natural number")
            return SpecialDecl("dimension", int(n))
        if self.try_keyword("axes"):
            self.keyword("::")
            xs = [self.try_ident_name("SA")]
            while self.try_symbolic(","):
                xs.append(self.try_ident_name("SA"))
            if any(x is None for x in xs):
                self.err("malformed axes declaration")
            return SpecialDecl("axes", xs)
        for kw in ("grid_struct_type_name", "grid_struct_instance_name"):
            if self.try_keyword(kw):
                self.keyword("::")
                name = self.try_ident_name("SA")
                if name is None:
                    self.err("expected a name")
                return SpecialDecl(kw, name)
        return None

    def program(self, nc: NumericalConfig) -> Program:
        """``program`` (Parser.hs:493-511)."""
        # leading blank lines and comments
        while True:
            self.ws()
            if not self.statement_delimiter():
                break
        decls, stmts = [], []
        while self.peek() != "":
            d = self.special_declaration()
            if d is not None:
                decls.append(d)
            else:
                stmts.extend(self.statement_compound())
            if not self.statement_delimiter():
                break
        self.ws()
        if self.peek() != "":
            self.err("unexpected input")
        axes = [d.value for d in decls if d.kind == "axes"]
        if not axes:
            raise ParseError("no axes declaration found.")
        ivars = axes[0]
        mpi = nc.mpi_shape if nc.mpi_shape is not None else [1] * len(nc.grid_per_node)
        extents = []
        for a, l, n in zip(ivars, nc.length_per_node, nc.grid_per_node):
            extents.append(Subst(Ident("d" + a), Imm(Fraction(l) / n)))
        for a, n, m in zip(ivars, nc.grid_per_node, mpi):
            extents.append(Subst(Ident("total_grid_" + a), Imm(Fraction(n * m))))
        return Program(decls, Binding(extents + stmts), nc)


def _eval_cre(e, p):
    if isinstance(e, Imm):
        return e.value
    if isinstance(e, Uniop) and e.op in "+-":
        v = _eval_cre(e.a, p)
        return v if e.op == "+" else -v
    if isinstance(e, Binop) and e.op in ("+", "-", "*", "/"):
        a, b = _eval_cre(e.a, p), _eval_cre(e.b, p)
        return {"+": a + b, "-": a - b, "*": a * b, "/": a / b if b else p.err("division by zero")}[e.op]
    p.err("unsupported operator in const rational expression")


def parse_program(text: str, nc: NumericalConfig, filename="<fmr>") -> Program:
    return Parser(text, filename).program(nc)


def desugar(prog: Program) -> Program:
    """Pad every grid index and grid type to the declared dimension (Desugar.hs:34-56)."""
    dims = [d.value for d in prog.decls if d.kind == "dimension"]
    if not dims:
        raise ParseError("no dimension declaration found.")
    dim = dims[0]

    def pad(xs, zero):
        return (list(xs) + [zero] * dim)[:dim]

    def walk(x):
        if isinstance(x, Grid):
            return Grid(pad(x.npks, NPlusK("", 0)), walk(x.x))
        if isinstance(x, GridType):
            return GridType(pad(x.offs, Fraction(0)), walk(x.elem))
        if isinstance(x, Tuple):
            return Tuple([walk(i) for i in x.items])
        if isinstance(x, TupleType):
            return TupleType([walk(i) for i in x.items])
        if isinstance(x, Uniop):
            return Uniop(x.op, walk(x.a))
        if isinstance(x, Binop):
            return Binop(x.op, walk(x.a), walk(x.b))
        if isinstance(x, Triop):
            return Triop(x.op, walk(x.a), walk(x.b), walk(x.c))
        if isinstance(x, Naryop):
            return Naryop(x.op, [walk(i) for i in x.xs])
        if isinstance(x, Apply):
            return Apply(walk(x.f), walk(x.x))
        if isinstance(x, Vector):
            return Vector(x.name, walk(x.x))
        if isinstance(x, Let):
            return Let(walk_binding(x.binding), walk(x.x))
        if isinstance(x, Lambda):
            return Lambda(walk(x.lhs), walk(x.body))
        if isinstance(x, VectorType):
            return VectorType(x.n, walk(x.elem))
        return x

    def walk_binding(b):
        out = []
        for s in b.stmts:
            if isinstance(s, Subst):
                out.append(Subst(walk(s.lhs), walk(s.rhs)))
            else:
                out.append(TypeDecl(s.mods, walk(s.typ), walk(s.lhs)))
        return Binding(out)

    return Program(prog.decls, walk_binding(prog.binding), prog.config)
