"""Manifestation: OM graph -> one micro-instruction DAG per manifest node.

Follows ``src/Formura/OrthotopeMachine/Manifestation.hs``:

* each manifest node (``Store`` nodes and ``manifest ::`` names) becomes one
  loop nest, visited in ascending OM id (:170-207);
* delayed nodes are inlined at every cursor; ``Shift`` moves the cursor,
  ``Load`` becomes ``LoadCursorStatic``, a reference to another manifest node
  becomes ``LoadCursor`` (:116-152);
* micro-node ids are allocated in post-order and shared per
  ``(cursor, OM node)`` (:85-93, :104-113), so ascending id is a valid
  evaluation order;

and the range analysis of ``src/Formura/Generator/Templates.hs``:

* ``calc_range`` (:388-424) is a scalar ``[lo, hi]`` over every cursor
  component, merged through ``LoadCursor`` chains.
"""
from __future__ import annotations

import dataclasses
from typing import Dict, List, Tuple

from . import syntax as S
from .translate import Graph, OMProgram


@dataclasses.dataclass
class Micro:
    inst: tuple      # ("Imm", r) ("LoadCursorStatic", cursor, name) ("LoadCursor", cursor, omid)
                     # ("LoadIndex", axis) ("Uniop", op, a) ("Binop", op, a, b) ("Triop", op, a, b, c)
                     # ("Store", name, a)
    typ: str         # element type name (micro types are element types only)
    src: str = None  # source name of the OM node (set for literals bound to a name)
    prec: str = None  # C type of the value in the mixed-precision form ("float"/"double")


@dataclasses.dataclass
class MMNode:
    omid: int
    micro: List[Micro]        # index == micro id, topological order
    typ: S.Type               # OM node type (void for Store)
    name: str


def _micro_type(t: S.Type) -> str:
    if isinstance(t, S.GridType):
        return _micro_type(t.elem)
    if isinstance(t, S.ElemType):
        return t.name
    raise ValueError("Top type encountered while manifestation")


def _add(c, v):
    return tuple(a + b for a, b in zip(c, v))


def manifest_graph(g: Graph, dim: int) -> Dict[int, MMNode]:
    """``manifestG`` (Manifestation.hs:170-207)."""
    manifest_ids = [i for i, n in enumerate(g.nodes) if n.manifest]
    mset = set(manifest_ids)
    out = {}
    for root in manifest_ids:
        idmap: Dict[Tuple[tuple, int], int] = {}
        micro: Dict[int, Micro] = {}

        def insert(c, i, inst):
            key = (c, i)
            if key not in idmap:
                idmap[key] = len(idmap)
            j = idmap[key]
            micro[j] = Micro(inst, _micro_type(g.nodes[i].typ),
                             g.nodes[i].name if inst[0] not in ("LoadCursorStatic", "LoadCursor", "Store") else None)
            return j

        def code_at(c, nid):
            if nid in mset:
                return insert(c, nid, ("LoadCursor", c, nid))
            return delayed_at(c, nid)

        def delayed_at(c, nid):
            inst = g.nodes[nid].inst
            k = inst[0]
            if k == "Imm":
                return insert(c, nid, ("Imm", inst[1]))
            if k == "Uniop":
                ja = code_at(c, inst[2])
                return insert(c, nid, ("Uniop", inst[1], ja))
            if k == "Binop":
                ja = code_at(c, inst[2])
                jb = code_at(c, inst[3])
                return insert(c, nid, ("Binop", inst[1], ja, jb))
            if k == "Triop":
                ja = code_at(c, inst[2])
                jb = code_at(c, inst[3])
                jc = code_at(c, inst[4])
                return insert(c, nid, ("Triop", inst[1], ja, jb, jc))
            if k == "Shift":
                return code_at(_add(c, inst[1]), inst[2])
            if k == "Load":
                return insert(c, nid, ("LoadCursorStatic", c, inst[1]))
            if k == "LoadIndex":
                return insert(c, nid, ("LoadIndex", inst[1]))
            if k == "Store":
                ja = code_at(c, inst[2])
                return insert(c, nid, ("Store", inst[1], ja))
            raise ValueError("manifestation path unimplemented for keyword: %s" % k)

        delayed_at((0,) * dim, root)
        node = g.nodes[root]
        out[root] = MMNode(root, [micro[j] for j in range(len(micro))], node.typ,
                           node.name or ("mn%d" % root))
    return out


@dataclasses.dataclass(frozen=True)
class MMRange:
    lower: int = 0
    upper: int = 0

    def union(self, o: "MMRange") -> "MMRange":
        return MMRange(min(self.lower, o.lower), max(self.upper, o.upper))

    def merge(self, o: "MMRange") -> "MMRange":
        return MMRange(self.lower + o.lower, self.upper + o.upper)

    @property
    def size(self) -> int:
        """``toSize`` (Templates.hs:411-412)."""
        return abs(self.lower) + abs(self.upper)

    @property
    def offset(self) -> int:
        """``toOffset`` (Templates.hs:414-415)."""
        return abs(self.lower)


def _from_cursor(c) -> MMRange:
    return MMRange(min(c), max(c))


def calc_range(mm: Dict[int, MMNode]) -> Dict[int, MMRange]:
    """``calcRange`` (Templates.hs:417-424)."""
    tbl: Dict[int, MMRange] = {}
    for omid in sorted(mm):
        rng = MMRange()
        for m in mm[omid].micro:
            if m.inst[0] == "LoadCursorStatic":
                rng = rng.union(_from_cursor(m.inst[1]))
            elif m.inst[0] == "LoadCursor":
                rng = rng.union(_from_cursor(m.inst[1]).merge(tbl[m.inst[2]]))
        tbl[omid] = rng
    return tbl


@dataclasses.dataclass
class MMProgram:
    om: OMProgram
    init: Dict[int, MMNode]
    first_step: Dict[int, MMNode] = None
    filter: Dict[int, MMNode] = None
    step: Dict[int, MMNode] = None


def gen_mm_program(om: OMProgram) -> MMProgram:
    """``genMMProgram`` (Manifestation.hs:209-258)."""
    d = om.dimension
    return MMProgram(
        om=om,
        init=manifest_graph(om.init_graph, d),
        first_step=manifest_graph(om.first_step_graph, d) if om.first_step_graph else None,
        filter=manifest_graph(om.filter_graph, d) if om.filter_graph else None,
        step=manifest_graph(om.step_graph, d),
    )


def inner_temporaries(mm: Dict[int, MMNode]) -> set:
    """The manifest temporaries that no Store node reads (the wave speeds a
    flux temporary reads, for instance): ``inline_temporaries(mm, keep=...)``
    with the complement keeps only the temporaries the update reads."""
    read_by_store = set()
    for node in mm.values():
        if node.micro[-1].inst[0] == "Store":
            read_by_store |= {m.inst[2] for m in node.micro if m.inst[0] == "LoadCursor"}
    return {k for k, n in mm.items() if n.micro[-1].inst[0] != "Store" and k not in read_by_store}


def cheap_temporaries(mm: Dict[int, MMNode]) -> set:
    """Temporaries worth recomputing where read: those no Store node reads that
    either evaluate no division, square root or other external call (limited
    reconstructions: a few comparisons and products) or that every reader reads
    at one and the same cursor (wave speeds of one face: a merged kernel
    evaluates them once)."""
    read_by_store, cursors = set(), {}
    for node in mm.values():
        for m in node.micro:
            if m.inst[0] == "LoadCursor":
                cursors.setdefault(m.inst[2], set()).add(tuple(m.inst[1]))
                if node.micro[-1].inst[0] == "Store":
                    read_by_store.add(m.inst[2])
    out = set()
    for k, n in mm.items():
        if n.micro[-1].inst[0] == "Store" or k in read_by_store:
            continue
        costly = any((m.inst[0] == "Binop" and m.inst[1] in ("/", "**")) or
                     (m.inst[0] == "Uniop" and m.inst[1].startswith("external-call/")) for m in n.micro)
        if not costly or len(cursors.get(k, set())) <= 1:
            out.add(k)
    return out


def inline_temporaries(mm: Dict[int, MMNode], prec: Dict[int, Dict[int, str]] = None,
                       keep=frozenset()) -> Dict[int, MMNode]:
    """Replace every ``LoadCursor(c, X)`` of the Store nodes by X's own micro DAG
    with its cursors shifted by ``c`` (recursively), sharing identical
    instructions.  A manifest temporary holds, at each input position, the
    value of its DAG at that position (Templates.hs:442-474 index it so), so
    recomputing it where it is read gives the same numbers; the Store nodes of
    the result read only the state fields.  Returns the Store nodes only.

    ``keep``: temporaries that stay materialized.  They are returned too, reads
    of them stay ``LoadCursor`` (their cursor shifted like any other), and the
    other temporaries are inlined into them as into the Store nodes.

    Index reads keep the reference's values: a nest evaluates ``LoadIndex k``
    as its own loop index plus the offsets (Templates.hs:481), and a consumer
    at index i reads temporary X at X's index i + c + off(consumer) - off(X)
    (Templates.hs:474), so an index read of X inlined through an accumulated
    cursor c reads i + c_k + off(root) - off(X): ``("LoadIndex", k, delta)``.

    ``prec`` (node -> micro -> C type, from emit.mixed_precision on the nodes
    before inlining) is carried into the copies and is part of the sharing key,
    so a value keeps the precision of the expression it came from wherever it
    is recomputed: the face flux that cell i recomputes as its left face is
    the same FP32 computation as the one cell i-1 does as its right face.
    """
    out = {}
    rng = calc_range(mm)
    for k in sorted(mm):
        node = mm[k]
        if node.micro[-1].inst[0] != "Store" and k not in keep:
            continue
        new: List[Micro] = []
        memo: Dict[tuple, int] = {}

        def emit(inst, typ, p=None):
            key = (inst, typ, p)
            if key not in memo:
                memo[key] = len(new)
                new.append(Micro(inst, typ, prec=p))
            return memo[key]

        def copy(src: MMNode, shift, j, cache):
            if j in cache:
                return cache[j]
            m = src.micro[j]
            k0 = m.inst[0]
            p = prec[src.omid].get(j) if prec and src.omid in prec else None
            emit_ = lambda inst, typ: emit(inst, typ, p)
            if k0 == "LoadCursorStatic":
                r = emit_(("LoadCursorStatic", _add(m.inst[1], shift), m.inst[2]), m.typ)
            elif k0 == "LoadCursor" and m.inst[2] in keep:
                r = emit_(("LoadCursor", _add(m.inst[1], shift), m.inst[2]), m.typ)
            elif k0 == "LoadCursor":
                tgt = mm[m.inst[2]]
                r = copy(tgt, _add(m.inst[1], shift), len(tgt.micro) - 1, {})
            elif k0 == "LoadIndex":
                ax = m.inst[1]
                delta = (m.inst[2] if len(m.inst) > 2 else 0) + shift[ax] + rng[k].offset - rng[src.omid].offset
                r = emit_(("LoadIndex", ax, delta) if delta else ("LoadIndex", ax), m.typ)
            elif k0 == "Imm":
                r = emit_(m.inst, m.typ)
                if m.src:
                    new[r].src = m.src
            elif k0 == "Uniop":
                r = emit_(("Uniop", m.inst[1], copy(src, shift, m.inst[2], cache)), m.typ)
                new[r].src = new[r].src or m.src
            elif k0 == "Binop":
                a = copy(src, shift, m.inst[2], cache)
                b = copy(src, shift, m.inst[3], cache)
                r = emit_(("Binop", m.inst[1], a, b), m.typ)
                new[r].src = new[r].src or m.src
            elif k0 == "Triop":
                a = copy(src, shift, m.inst[2], cache)
                b = copy(src, shift, m.inst[3], cache)
                c = copy(src, shift, m.inst[4], cache)
                r = emit_(("Triop", m.inst[1], a, b, c), m.typ)
            elif k0 == "Store":
                r = emit_(("Store", m.inst[1], copy(src, shift, m.inst[2], cache)), m.typ)
            else:
                raise ValueError(k0)
            cache[j] = r
            return r

        dim = None
        for m in node.micro:
            if m.inst[0] in ("LoadCursorStatic", "LoadCursor"):
                dim = len(m.inst[1])
                break
        root = copy(node, (0,) * (dim or 3), len(node.micro) - 1, {})
        if root != len(new) - 1:  # the node's value must be its last instruction
            new.append(dataclasses.replace(new[root]))
        out[k] = MMNode(k, new, node.typ, node.name)
    return out
