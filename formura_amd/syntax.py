"""Syntax tree of ``.fmr`` programs and the value/type terms of the evaluator.

The constructors follow ``src/Formura/Syntax.hs`` (expression functors
:46-266, statements :268-281, program :326-338) and the value forms of
``src/Formura/OrthotopeMachine/Graph.hs`` (``NodeValue``, ``FunValue``,
``RXExpr``/``ValueExpr``). Expressions are immutable plain objects; a value
produced by evaluation (``NodeValue``/``FunValue``/``Tuple`` of values) can sit
inside an expression, as the reference's ``RXExpr`` allows after lexical
resolution (Translate.hs:353-383).
"""
from __future__ import annotations

from fractions import Fraction
from typing import List, Optional, Tuple as PyTuple


class Expr:
    __slots__ = ()


class Ident(Expr):
    __slots__ = ("name",)

    def __init__(self, name: str):
        self.name = name

    def __repr__(self):
        return self.name


class Imm(Expr):
    __slots__ = ("value",)

    def __init__(self, value: Fraction):
        self.value = Fraction(value)

    def __repr__(self):
        return str(self.value)


class Tuple(Expr):
    __slots__ = ("items",)

    def __init__(self, items):
        self.items = list(items)

    def __repr__(self):
        return "(" + ", ".join(map(repr, self.items)) + ")"


class Uniop(Expr):
    __slots__ = ("op", "a")

    def __init__(self, op, a):
        self.op, self.a = op, a

    def __repr__(self):
        return "(%s%r)" % (self.op, self.a)


class Binop(Expr):
    __slots__ = ("op", "a", "b")

    def __init__(self, op, a, b):
        self.op, self.a, self.b = op, a, b

    def __repr__(self):
        return "(%r %s %r)" % (self.a, self.op, self.b)


class Triop(Expr):
    __slots__ = ("op", "a", "b", "c")

    def __init__(self, op, a, b, c):
        self.op, self.a, self.b, self.c = op, a, b, c

    def __repr__(self):
        return "(if %r then %r else %r)" % (self.a, self.b, self.c)


class Naryop(Expr):
    __slots__ = ("op", "xs")

    def __init__(self, op, xs):
        self.op, self.xs = op, list(xs)


class NPlusK:
    """``NPlusK`` (Syntax.hs:288-305): an index name plus a rational offset."""

    __slots__ = ("name", "k")

    def __init__(self, name: str, k: Fraction):
        self.name, self.k = name, Fraction(k)

    def __repr__(self):
        if not self.k:
            return self.name or "0"
        return "%s%+s" % (self.name, self.k)


class Grid(Expr):
    __slots__ = ("npks", "x")

    def __init__(self, npks: List[NPlusK], x):
        self.npks, self.x = list(npks), x

    def __repr__(self):
        return "%r%r" % (self.x, self.npks)


class Vector(Expr):
    __slots__ = ("name", "x")

    def __init__(self, name, x):
        self.name, self.x = name, x


class Apply(Expr):
    __slots__ = ("f", "x")

    def __init__(self, f, x):
        self.f, self.x = f, x

    def __repr__(self):
        return "%r(%r)" % (self.f, self.x)


class Let(Expr):
    __slots__ = ("binding", "x")

    def __init__(self, binding: "Binding", x):
        self.binding, self.x = binding, x


class Lambda(Expr):
    __slots__ = ("lhs", "body")

    def __init__(self, lhs, body):
        self.lhs, self.body = lhs, body

    def __repr__(self):
        return "fun%r %r" % (self.lhs, self.body)


# --- values (Graph.hs NodeValueF / FunValueF) -------------------------------

class NodeValue(Expr):
    __slots__ = ("nid", "typ")

    def __init__(self, nid: int, typ):
        self.nid, self.typ = nid, typ

    def __repr__(self):
        return "#%d::%r" % (self.nid, self.typ)


class FunValue(Expr):
    __slots__ = ("lhs", "body")

    def __init__(self, lhs, body):
        self.lhs, self.body = lhs, body

    def __repr__(self):
        return "<fun %r>" % (self.lhs,)


# --- statements (Syntax.hs:240-281) -------------------------------------------

class Subst:
    __slots__ = ("lhs", "rhs")

    def __init__(self, lhs, rhs):
        self.lhs, self.rhs = lhs, rhs


class TypeDecl:
    __slots__ = ("mods", "typ", "lhs")

    def __init__(self, mods, typ, lhs):
        self.mods, self.typ, self.lhs = list(mods), typ, lhs


class Binding:
    __slots__ = ("stmts",)

    def __init__(self, stmts):
        self.stmts = list(stmts)

    def type_decls(self):
        """``typeDecls`` (Syntax.hs:249-253): typed declarations only."""
        return [(s.lhs, s.typ) for s in self.stmts
                if isinstance(s, TypeDecl) and not isinstance(s.typ, TopType)]

    def type_modifiers(self):
        """``typeModifiers`` (Syntax.hs:256-259)."""
        return [(s.lhs, m) for s in self.stmts if isinstance(s, TypeDecl) for m in s.mods]

    def substs(self):
        """``substs`` (Syntax.hs:262-265)."""
        return [(s.lhs, s.rhs) for s in self.stmts if isinstance(s, Subst)]


# --- types (Syntax.hs:44-78, 189-212, 307-354) --------------------------------

class Type:
    __slots__ = ()


class TopType(Type):
    __slots__ = ()

    def __eq__(self, o):
        return isinstance(o, TopType)

    def __hash__(self):
        return 1

    def __repr__(self):
        return "top"


class FunType(Type):
    __slots__ = ()

    def __eq__(self, o):
        return isinstance(o, FunType)

    def __hash__(self):
        return 2

    def __repr__(self):
        return "function"


class ElemType(Type):
    __slots__ = ("name",)

    def __init__(self, name):
        self.name = name

    def __eq__(self, o):
        return isinstance(o, ElemType) and o.name == self.name

    def __hash__(self):
        return hash(("E", self.name))

    def __repr__(self):
        return self.name


def vec_eq(a, b) -> bool:
    """``Vec`` equality zips and truncates (Vec.hs:162-163,189-194)."""
    return all(x == y for x, y in zip(a, b))


class GridType(Type):
    __slots__ = ("offs", "elem")

    def __init__(self, offs, elem):
        self.offs, self.elem = tuple(Fraction(o) for o in offs), elem

    def __eq__(self, o):
        return isinstance(o, GridType) and vec_eq(self.offs, o.offs) and self.elem == o.elem

    def __hash__(self):
        return hash(("G", self.elem))

    def __repr__(self):
        return "%r[%s]" % (self.elem, ",".join(str(x) for x in self.offs))


class TupleType(Type):
    __slots__ = ("items",)

    def __init__(self, items):
        self.items = list(items)

    def __eq__(self, o):
        return isinstance(o, TupleType) and len(o.items) == len(self.items) and all(
            a == b for a, b in zip(self.items, o.items))

    def __hash__(self):
        return hash(("T", len(self.items)))

    def __repr__(self):
        return "(" + ", ".join(map(repr, self.items)) + ")"


class VectorType(Type):
    __slots__ = ("n", "elem")

    def __init__(self, n, elem):
        self.n, self.elem = n, elem

    def __eq__(self, o):
        return isinstance(o, VectorType) and o.n == self.n and o.elem == self.elem

    def __hash__(self):
        return hash(("V", self.n))


# element-type lattice void < bool < int < Rational < float < double < real < Real
# (Type.hs:31-46); the meet of two element types is the larger one.
ELEMENT_TYPENAMES = ["void", "bool", "int", "Rational", "float", "double", "real", "Real"]
_ELEM_RANK = {n: i for i, n in enumerate(ELEMENT_TYPENAMES)}


def elem_meet(a: str, b: str) -> Optional[str]:
    r = max(_ELEM_RANK.get(a, 1 << 30), _ELEM_RANK.get(b, 1 << 30))
    return ELEMENT_TYPENAMES[r] if r < len(ELEMENT_TYPENAMES) else None


def type_meet(a: Type, b: Type) -> Type:
    """Meet of two node types (``semiLatticeOfOMNodeType``, Graph.hs:151-166)."""

    def go(x, y):
        if x == y:
            return x
        if isinstance(x, ElemType) and isinstance(y, ElemType):
            m = elem_meet(x.name, y.name)
            return TopType() if m is None else ElemType(m)
        if isinstance(x, ElemType) and isinstance(y, GridType):
            d = go(x, y.elem)
            return TopType() if isinstance(d, TopType) else GridType(y.offs, d)
        if isinstance(x, GridType) and isinstance(y, GridType):
            if vec_eq(x.offs, y.offs):
                return GridType(x.offs, go(x.elem, y.elem))
            return TopType()
        return TopType()

    c = go(a, b)
    if isinstance(c, TopType):
        return go(b, a)
    return c


class SpecialDecl:
    __slots__ = ("kind", "value")

    def __init__(self, kind, value):
        self.kind, self.value = kind, value


class Program:
    __slots__ = ("decls", "binding", "config")

    def __init__(self, decls, binding, config):
        self.decls, self.binding, self.config = decls, binding, config
