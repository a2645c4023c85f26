"""Evaluation of a ``.fmr`` program into the orthotope-machine (OM) graph.

Follows ``src/Formura/OrthotopeMachine/Translate.hs``:

* every literal becomes an ``Imm`` node when evaluated (:175-177): there is no
  constant folding in the graph;
* operators on tuples act element-wise, on functions they build new functions
  (:218-252), ``if`` zips tuples through ``TExpr`` (:187-190, TExpr.hs:22-31);
* grid access ``x[i+k]`` with a rational offset becomes an integer ``Shift`` by
  ``-floor(off - k)`` plus a residual fractional type offset (:287-313);
* ``fun`` resolves free names lexically but leaves ``let`` bindings to be
  evaluated where the function is applied (:353-383), and application binds
  the parameters on top of the caller's bindings (:346-351);
* ``withBindings`` binds statements one by one, with index names on a grid
  LHS bound to ``LoadIndex`` nodes, casts to the declared type and marks
  ``manifest`` names (:453-527);
* ``genGlobalFunction`` adds ``Load`` nodes for the parameters and one
  manifest ``Store`` per state field (:559-584);
* ``calc_sleeve`` is ``calcSleeve`` with the lexicographic ``Vec`` maximum
  (:613-623, Vec.hs:165-166).
"""
from __future__ import annotations

import dataclasses
import math
from fractions import Fraction
from typing import Dict, List, Optional

from . import syntax as S
from .config import ConfigException, InternalConfig, convert_config
from .parser import desugar


class TranslateError(Exception):
    pass


@dataclasses.dataclass
class OMNode:
    inst: tuple          # ("Imm", r) ("Load", name) ("LoadIndex", axis) ("Store", name, a)
                         # ("Uniop", op, a) ("Binop", op, a, b) ("Triop", op, a, b, c)
                         # ("Naryop", op, [xs]) ("Shift", vec, a)
    typ: S.Type
    manifest: bool = False
    name: Optional[str] = None   # source name annotation


class Graph:
    def __init__(self):
        self.nodes: List[OMNode] = []

    def insert(self, inst, typ) -> S.NodeValue:
        nid = len(self.nodes)
        self.nodes.append(OMNode(inst, typ))
        return S.NodeValue(nid, typ)


def _type_of_val(v):
    """``typeOfVal`` (Translate.hs:157-162)."""
    if isinstance(v, S.Imm):
        return S.ElemType("Rational")
    if isinstance(v, S.NodeValue):
        return v.typ
    if isinstance(v, S.FunValue):
        return S.FunType()
    if isinstance(v, S.Tuple):
        return S.TupleType([_type_of_val(x) for x in v.items])
    raise TranslateError("no match (typeOfVal): %r" % (v,))


def _names_of_lhs(l):
    """``namesOfLhs`` (Translate.hs:389-394)."""
    if isinstance(l, S.Ident):
        return l
    if isinstance(l, (S.Grid, S.Vector)):
        return _names_of_lhs(l.x)
    if isinstance(l, S.Tuple):
        return S.Tuple([_names_of_lhs(x) for x in l.items])
    raise TranslateError("malformed left-hand side: %r" % (l,))


def _tuple_contents(x):
    if isinstance(x, S.Tuple):
        out = []
        for i in x.items:
            out.extend(_tuple_contents(i))
        return out
    return [x]


def _match_to_idents(pat, val):
    """``matchToIdents`` (Translate.hs:412-421); ``val`` may be a value or a type."""
    if isinstance(pat, S.Tuple):
        items = val.items if isinstance(val, (S.Tuple, S.TupleType)) else None
        if items is None:
            raise TranslateError("the LHS expects a tuple, but RHS is not a tuple.")
        if len(items) != len(pat.items):
            raise TranslateError("tuple length mismatch.")
        out = []
        for p, v in zip(pat.items, items):
            out.extend(_match_to_idents(p, v))
        return out
    if isinstance(pat, S.Ident):
        return [(pat.name, val)]
    raise TranslateError("no match (matchToIdents)")


def match_to_lhs(l, val):
    return _match_to_idents(_names_of_lhs(l), val)


def _index_names_of_lhs(l):
    if isinstance(l, S.Grid):
        return [n.name if n.name else None for n in l.npks]
    return []


def _scope_holder(l):
    """``lexicalScopeHolder`` (Translate.hs:433-440)."""
    return {x.name: x for x in _tuple_contents(_names_of_lhs(l))}


COMPARISON_OPS = {"<=", "==", "!=", ">=", "<", ">"}


def _map_elem_type(f, t):
    if isinstance(t, S.ElemType):
        return S.ElemType(f(t.name))
    if isinstance(t, S.GridType):
        return S.GridType(t.offs, _map_elem_type(f, t.elem))
    return t


def _is_boolish(t):
    if isinstance(t, S.ElemType):
        return t.name == "bool"
    if isinstance(t, S.GridType):
        return _is_boolish(t.elem)
    return False


class _TAtom:
    __slots__ = ("x",)

    def __init__(self, x):
        self.x = x


def _texpr(v):
    if isinstance(v, S.Tuple):
        return [_texpr(i) for i in v.items]
    return _TAtom(v)


def _tzip3(a, b, c):
    """Applicative zip of three ``TExpr`` trees with atom broadcast (TExpr.hs:22-31)."""
    def zip2(x, y, combine):
        if isinstance(x, _TAtom) and isinstance(y, _TAtom):
            return combine(x, y)
        if isinstance(x, list) and isinstance(y, _TAtom):
            return [zip2(i, y, combine) for i in x]
        if isinstance(x, _TAtom) and isinstance(y, list):
            return [zip2(x, j, combine) for j in y]
        if isinstance(x, list) and isinstance(y, list):
            if len(x) != len(y):
                raise TranslateError("Tuple length mismatch")
            return [zip2(i, j, combine) for i, j in zip(x, y)]
        raise TranslateError("Tuple length mismatch")

    ab = zip2(a, b, lambda x, y: _TAtom((x.x, y.x)))
    return zip2(ab, c, lambda x, y: _TAtom((x.x[0], x.x[1], y.x)))


class Translator:
    """Builds one OM graph (one of init / first_step / filter / step)."""

    def __init__(self):
        self.g = Graph()

    def insert(self, inst, typ):
        return self.g.insert(inst, typ)

    # -- gen -------------------------------------------------------------------
    def gen(self, e, env):
        if isinstance(e, (S.NodeValue, S.FunValue)):
            return e
        if isinstance(e, S.Imm):
            return self.insert(("Imm", e.value), S.ElemType("Rational"))
        if isinstance(e, S.Ident):
            if e.name not in env:
                raise TranslateError("undefined variable: " + e.name)
            return env[e.name]
        if isinstance(e, S.Tuple):
            return S.Tuple([self.gen(x, env) for x in e.items])
        if isinstance(e, S.Uniop):
            return self.go_uniop(e.op, self.gen(e.a, env))
        if isinstance(e, S.Binop):
            a = self.gen(e.a, env)
            b = self.gen(e.b, env)
            return self.go_binop(e.op, a, b)
        if isinstance(e, S.Triop):
            a = self.gen(e.a, env)
            b = self.gen(e.b, env)
            c = self.gen(e.c, env)
            tree = _tzip3(_texpr(a), _texpr(b), _texpr(c))

            def run(t):
                if isinstance(t, list):
                    return S.Tuple([run(i) for i in t])
                return self.go_triop(e.op, *t.x)
            return run(tree)
        if isinstance(e, S.Naryop):
            raise TranslateError("n-ary operator %s is not supported by the code generator" % e.op)
        if isinstance(e, S.Grid):
            return self.grid(e.npks, self.gen(e.x, env))
        if isinstance(e, S.Apply):
            f0 = self.gen(e.f, env)
            a0 = self.gen(e.x, env)
            return self.go_apply(f0, a0, env)
        if isinstance(e, S.Let):
            return self.with_bindings(e.binding, env, lambda env2: self.gen(e.x, env2))
        if isinstance(e, S.Lambda):
            lex = dict(env)
            lex.update(_scope_holder(e.lhs))
            return S.FunValue(e.lhs, self.resolve_lex(e.body, lex))
        raise TranslateError("cannot evaluate %r" % (e,))

    def resolve_lex(self, e, lex):
        """``resolveLex`` (Translate.hs:365-383)."""
        if isinstance(e, S.Ident):
            if e.name not in lex:
                raise TranslateError("undefined variable: %s\nwhen resolving lexical binding." % e.name)
            return lex[e.name]
        if isinstance(e, S.Lambda):
            l2 = dict(lex)
            l2.update(_scope_holder(e.lhs))
            return S.FunValue(e.lhs, self.resolve_lex(e.body, l2))
        if isinstance(e, S.FunValue):
            l2 = dict(lex)
            l2.update(_scope_holder(e.lhs))
            return S.FunValue(e.lhs, self.resolve_lex(e.body, l2))
        if isinstance(e, S.Let):
            l2 = dict(lex)
            for l, _ in e.binding.substs():
                l2.update(_scope_holder(l))
            return S.Let(e.binding, self.resolve_lex(e.x, l2))
        if isinstance(e, S.Tuple):
            return S.Tuple([self.resolve_lex(x, lex) for x in e.items])
        if isinstance(e, S.Uniop):
            return S.Uniop(e.op, self.resolve_lex(e.a, lex))
        if isinstance(e, S.Binop):
            return S.Binop(e.op, self.resolve_lex(e.a, lex), self.resolve_lex(e.b, lex))
        if isinstance(e, S.Triop):
            return S.Triop(e.op, self.resolve_lex(e.a, lex), self.resolve_lex(e.b, lex),
                           self.resolve_lex(e.c, lex))
        if isinstance(e, S.Naryop):
            return S.Naryop(e.op, [self.resolve_lex(x, lex) for x in e.xs])
        if isinstance(e, S.Grid):
            return S.Grid(e.npks, self.resolve_lex(e.x, lex))
        if isinstance(e, S.Apply):
            return S.Apply(self.resolve_lex(e.f, lex), self.resolve_lex(e.x, lex))
        return e  # Imm, NodeValue

    # -- operators -------------------------------------------------------------
    def go_uniop(self, op, a):
        """``goUniop`` (Translate.hs:218-225)."""
        if isinstance(a, S.NodeValue):
            return self.insert(("Uniop", op, a.nid), a.typ)
        if isinstance(a, S.FunValue):
            return S.FunValue(S.Ident("x"), S.Uniop(op, S.Apply(a, S.Ident("x"))))
        if isinstance(a, S.Tuple):
            return S.Tuple([self.go_uniop(op, x) for x in a.items])
        raise TranslateError("unexpected path in unary operator")

    def cast_val(self, t1, vx):
        """``castVal`` (Translate.hs:166-172)."""
        t0 = _type_of_val(vx)
        if t1 == t0:
            return vx
        if isinstance(t1, S.ElemType) and isinstance(t0, S.ElemType):
            return vx
        if (isinstance(t1, S.GridType) and isinstance(t1.elem, S.ElemType)
                and isinstance(t0, S.ElemType) and isinstance(vx, S.NodeValue)):
            return S.NodeValue(vx.nid, S.GridType(t1.offs, t1.elem))
        if isinstance(t1, S.GridType) and isinstance(t0, S.GridType) and S.vec_eq(t1.offs, t0.offs):
            return vx
        raise TranslateError("cannot convert type %r to %r" % (t0, t1))

    def go_binop(self, op, a, b):
        """``goBinop`` (Translate.hs:227-252)."""
        if op == ".":
            return S.FunValue(S.Ident("x"), S.Apply(a, S.Apply(b, S.Ident("x"))))
        if isinstance(a, S.NodeValue) and isinstance(b, S.NodeValue):
            ct = S.type_meet(a.typ, b.typ)
            if isinstance(ct, S.TopType):
                raise TranslateError("there is no common type that can accomodate both hand side: %r %s %r"
                                     % (a.typ, op, b.typ))
            a2 = self.cast_val(ct, a)
            b2 = self.cast_val(ct, b)
            typ = _map_elem_type(lambda _: "bool", ct) if op in COMPARISON_OPS else ct
            return self.insert(("Binop", op, a2.nid, b2.nid), typ)
        fa, fb = isinstance(a, S.FunValue), isinstance(b, S.FunValue)
        if fa and fb:
            return S.FunValue(S.Ident("x"), S.Binop(op, S.Apply(a, S.Ident("x")), S.Apply(b, S.Ident("x"))))
        if fa and isinstance(b, S.NodeValue):
            return S.FunValue(S.Ident("x"), S.Binop(op, S.Apply(a, S.Ident("x")), b))
        if isinstance(a, S.NodeValue) and fb:
            return S.FunValue(S.Ident("x"), S.Binop(op, a, S.Apply(b, S.Ident("x"))))
        if isinstance(a, S.Tuple) and isinstance(b, S.Tuple):
            if len(a.items) != len(b.items):
                raise TranslateError("tuple length mismatch.")
            return S.Tuple([self.go_binop(op, x, y) for x, y in zip(a.items, b.items)])
        if isinstance(a, S.NodeValue) and isinstance(b, S.Tuple):
            return S.Tuple([self.go_binop(op, a, y) for y in b.items])
        if isinstance(a, S.Tuple) and isinstance(b, S.NodeValue):
            return S.Tuple([self.go_binop(op, x, b) for x in a.items])
        raise TranslateError("unexpected path in binary operator: %s %r %r" % (op, a, b))

    def go_triop(self, op, a, b, c):
        """``goTriop`` (Translate.hs:259-269)."""
        if op == "ite" and all(isinstance(x, S.NodeValue) for x in (a, b, c)):
            if not _is_boolish(a.typ):
                raise TranslateError("the first argument of if-expr must be of type bool")
            bct = S.type_meet(b.typ, c.typ)
            if isinstance(S.type_meet(a.typ, bct), S.TopType):
                raise TranslateError("Type mismatch in if-then-else expr: %r %r %r" % (a.typ, b.typ, c.typ))
            return self.insert(("Triop", op, a.nid, b.nid, c.nid), bct)
        raise TranslateError("unexpected path in trinary operator" + op)

    def grid(self, npks, v):
        """``gen`` of ``GridF`` (Translate.hs:287-313)."""
        if isinstance(v, S.NodeValue):
            t = v.typ
            if isinstance(t, S.ElemType):
                return v
            if isinstance(t, S.GridType):
                pat = [n.k for n in npks]
                new_pos = [o - k for o, k in zip(t.offs, pat)]
                int_off = [math.floor(p) for p in new_pos]
                new_off = [p - i for p, i in zip(new_pos, int_off)]
                typ1 = S.GridType(new_off, t.elem)
                if all(i == 0 for i in int_off):
                    return S.NodeValue(v.nid, typ1)
                return self.insert(("Shift", tuple(-i for i in int_off), v.nid), typ1)
            raise TranslateError("no match in grid access")
        if isinstance(v, S.Tuple):
            return S.Tuple([self.grid(npks, x) for x in v.items])
        raise TranslateError("unexpected pattern in gen of grid: %r" % (v,))

    def eval_to_imm(self, x):
        if isinstance(x, S.Imm):
            return x.value
        if isinstance(x, S.NodeValue):
            n = self.g.nodes[x.nid]
            if n.inst[0] == "Imm":
                return n.inst[1]
            raise TranslateError("tuple applied to non-constant integer")
        return None

    def go_apply(self, f, x, env):
        """``goApply`` (Translate.hs:332-351)."""
        if isinstance(f, S.Tuple):
            r = self.eval_to_imm(x)
            if r is None:
                raise TranslateError("tuple applied to non-constant integer: %r" % (x,))
            if r.denominator != 1:
                raise TranslateError("non-integer indexing in tuple access")
            n = int(r)
            if n < 0 or n >= len(f.items):
                raise TranslateError("tuple access out of bounds")
            return f.items[n]
        if isinstance(f, S.FunValue):
            env2 = dict(env)
            env2.update(dict(match_to_lhs(f.lhs, x)))
            return self.gen(f.body, env2)
        raise TranslateError("unexpected combination of application: %r applied to %r" % (f, x))

    # -- bindings --------------------------------------------------------------
    def with_bindings(self, b: S.Binding, env, cont):
        """``withBindings`` (Translate.hs:453-527)."""
        type_dict = {}
        for l, t in b.type_decls():
            type_dict.update(dict(match_to_lhs(l, t)))
        mod_dict: Dict[str, list] = {}
        for l, m in b.type_modifiers():
            for x in _tuple_contents(_names_of_lhs(l)):
                mod_dict.setdefault(x.name, []).append(m)
        ext_funs = [l.name for l, t in b.type_decls()
                    if isinstance(l, S.Ident) and isinstance(t, S.FunType)
                    and "extern" in mod_dict.get(l.name, [])]
        env1 = dict(env)
        for f in ext_funs:
            env1[f] = S.FunValue(S.Ident("q"), S.Uniop("external-call/" + f, S.Ident("q")))
        cur = env1
        bound = {}
        for l0, rhs in b.substs():
            idx_env = dict(cur)
            for ax, name in enumerate(_index_names_of_lhs(l0)):
                if name is not None:
                    idx_env[name] = self.insert(("LoadIndex", ax), S.ElemType("Rational"))
            v0 = self.gen(rhs, idx_env)
            if isinstance(l0, S.Grid):
                neg = [S.NPlusK(n.name, -n.k) for n in l0.npks]
                lvs = [(nm, self.grid(neg, v)) for nm, v in match_to_lhs(l0.x, v0)]
            else:
                lvs = match_to_lhs(l0, v0)
            nvs = {}
            for name0, v1 in lvs:
                v = self.cast_val(type_dict[name0], v1) if name0 in type_dict else v1
                self._annotate(v, name0, "manifest" in mod_dict.get(name0, []))
                nvs[name0] = v
            cur = dict(cur)
            cur.update(nvs)
            bound.update(nvs)
        final = dict(env1)
        final.update(bound)
        return cont(final)

    def _annotate(self, v, name, is_manifest):
        if isinstance(v, S.NodeValue):
            n = self.g.nodes[v.nid]
            if n.name is None:
                n.name = name
            if is_manifest:
                n.manifest = True
                n.name = name
        elif isinstance(v, S.Tuple):
            for x in v.items:
                self._annotate(x, name, is_manifest)

    def gen_global_function(self, gbinds: S.Binding, input_type, output_pattern, fundef):
        """``genGlobalFunction`` (Translate.hs:559-584); returns the type of the result."""
        if not isinstance(fundef, S.Lambda):
            raise TranslateError("Identifier specified for function generation is not of function type.")

        def body(env):
            init_binds = {}
            for name1, t1 in match_to_lhs(fundef.lhs, input_type):
                init_binds[name1] = self.insert(("Load", name1), _to_node_type(t1))
            env2 = dict(env)
            env2.update(init_binds)
            ret = self.gen(fundef.body, env2)
            for name1, rv1 in match_to_lhs(output_pattern, ret):
                if not isinstance(rv1, S.NodeValue):
                    raise TranslateError("The return type of a global function must be a tuple of grids.")
                st = self.insert(("Store", name1, rv1.nid), S.ElemType("void"))
                node = self.g.nodes[st.nid]
                node.manifest = True
                node.name = name1 + "_next"
            return _type_of_val(ret)

        return self.with_bindings(gbinds, {}, body)


def _to_node_type(t):
    if isinstance(t, (S.ElemType, S.GridType)):
        return t
    raise TranslateError("incompatible type `%r` encountered in conversion to node type." % (t,))


def calc_sleeve(g: Graph) -> int:
    """``calcSleeve`` (Translate.hs:613-623) with lexicographic ``max`` on vectors."""
    tbl = []

    def lexmax(*vs):
        best = vs[0]
        for v in vs[1:]:
            # default Ord max: max x y = if x <= y then y else x
            best = v if tuple(best) <= tuple(v) else best
        return best

    dim = None
    for n in g.nodes:
        if n.inst[0] == "Shift":
            dim = len(n.inst[1])
            break
    zero = (0,) * (dim or 1)
    for n in g.nodes:
        k = n.inst[0]
        if k == "Shift":
            base = tbl[n.inst[2]]
            d = n.inst[1]
            tbl.append(tuple(a + b for a, b in zip(d, base)) if len(base) == len(d) else tuple(d))
        elif k == "Uniop":
            tbl.append(tbl[n.inst[2]])
        elif k == "Binop":
            tbl.append(lexmax(tbl[n.inst[2]], tbl[n.inst[3]]))
        elif k == "Triop":
            tbl.append(lexmax(tbl[n.inst[2]], tbl[n.inst[3]], tbl[n.inst[4]]))
        else:
            tbl.append(zero)
    return max((max(v) for v in tbl), default=0)


@dataclasses.dataclass
class OMProgram:
    """``MachineProgram`` over OM graphs (Graph.hs:238-251)."""

    dimension: int
    axes: List[str]
    grid_struct_type_name: str
    grid_struct_instance_name: str
    config: InternalConfig
    init_graph: Graph
    first_step_graph: Optional[Graph]
    filter_graph: Optional[Graph]
    step_graph: Graph
    state_signature: Dict[str, S.Type]   # name -> type, iterated in key order

    @property
    def comm_bases(self):
        return {1: [[1]], 2: [[1, 0], [0, 1], [1, 1]],
                3: [[1, 0, 0], [0, 1, 0], [0, 0, 1], [1, 1, 0], [1, 0, 1], [0, 1, 1], [1, 1, 1]]}[self.dimension]


def _lookup_toplevel(prog: S.Program, name, required=True):
    hits = [rhs for lhs, rhs in prog.binding.substs() if isinstance(lhs, S.Ident) and lhs.name == name]
    if not hits:
        if required:
            raise TranslateError("Identifier `%s` not found." % name)
        return None
    if len(hits) > 1:
        raise TranslateError("Multiple declaration of identifier `%s` found." % name)
    return hits[0]


def _setup_env(prog: S.Program):
    """``setupGlobalEnvironment`` (Translate.hs:81-130)."""
    dims = [d.value for d in prog.decls if d.kind == "dimension"]
    if len(dims) != 1:
        raise TranslateError("no dimension declaration found." if not dims else "multiple dimension declaration found.")
    axes = [d.value for d in prog.decls if d.kind == "axes"]
    if len(axes) != 1:
        raise TranslateError("no axes declaration found." if not axes else "multiple axes declaration found.")
    if len(axes[0]) != dims[0]:
        raise TranslateError("number of declared axes does not match the declared dimension.")
    tn = [d.value for d in prog.decls if d.kind == "grid_struct_type_name"]
    inn = [d.value for d in prog.decls if d.kind == "grid_struct_instance_name"]
    if len(tn) > 1 or len(inn) > 1:
        raise TranslateError("multiple grid_struct names found.")
    if dims[0] not in (1, 2, 3):
        raise TranslateError("Not support")
    return dims[0], axes[0], (tn[0] if tn else "Formura_Grid_Struct"), (inn[0] if inn else "formura_data")


def gen_om_program(prog: S.Program) -> OMProgram:
    """``genOMProgram`` (Translate.hs:625-686)."""
    prog = desugar(prog)
    dim, axes, tname, iname = _setup_env(prog)
    step_def = _lookup_toplevel(prog, "step")
    if not isinstance(step_def, S.Lambda):
        raise TranslateError("step is not a function")
    lhs_of_step = step_def.lhs
    gb = prog.binding

    t_init = Translator()
    init_type = t_init.gen_global_function(gb, S.TupleType([]), lhs_of_step, _lookup_toplevel(prog, "init"))

    def optional(name):
        d = _lookup_toplevel(prog, name, required=False)
        if d is None:
            return None
        tr = Translator()
        ty = tr.gen_global_function(gb, init_type, lhs_of_step, d)
        if ty != init_type:
            raise TranslateError("the return type of %s : %r\nmust match the return type of init : %r" % (name, ty, init_type))
        return tr.g

    first = optional("first_step")
    filt = optional("filter")
    t_step = Translator()
    step_type = t_step.gen_global_function(gb, init_type, lhs_of_step, step_def)
    if step_type != init_type:
        raise TranslateError("the return type of step : %r\nmust match the return type of init : %r" % (step_type, init_type))
    sig = dict(sorted(match_to_lhs(lhs_of_step, step_type), key=lambda kv: kv[0]))

    sleeve = calc_sleeve(t_step.g)
    s0 = calc_sleeve(first) if first is not None else None
    sf = calc_sleeve(filt) if filt is not None else None
    try:
        cfg = convert_config(sleeve, s0, sf, prog.config)
    except ConfigException as e:
        raise ConfigException("Error: " + str(e))
    return OMProgram(dim, list(axes), tname, iname, cfg, t_init.g, first, filt, t_step.g, sig)
