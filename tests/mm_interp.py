"""Test helper: evaluate a manifest-machine graph with numpy (float64, IEEE, no FMA).

One Formura step evaluated straight from the front-end's micro-instruction
DAGs, with the placement of the reference's generated C: output index i of a
Store node with range r reads input index i + d + off(r) - 2s (periodic), and a
manifest temporary is indexed by input position.  Used to check the front-end
(parser -> OM graph -> manifestation) against the oracle on the CPU, without the
HIP emitter.
"""
import math

import numpy as np

from formura_amd.manifest import calc_range


def _shift(a, offs):
    # value at index i is a[(i + off) mod N]  ->  np.roll by -off
    return np.roll(a, shift=[-o for o in offs], axis=tuple(range(a.ndim)))


def _eval_node(node, fields, temps, base, params=None, shape=None, offsets=None):
    vals = []
    params = params or {}
    ext = {"sqrt": np.sqrt, "exp": np.exp, "fabs": np.fabs, "sin": np.sin, "cos": np.cos}
    for m in node.micro:
        k = m.inst[0]
        if k == "Imm":
            v = float(params[m.src]) if m.src in params else float(m.inst[1])
        elif k == "LoadCursorStatic":
            v = _shift(fields[m.inst[2]], [c + b for c, b in zip(m.inst[1], base)])
        elif k == "LoadCursor":
            v = _shift(temps[m.inst[2]], [c + b for c, b in zip(m.inst[1], base)])
        elif k == "LoadIndex":
            # i_k + offset_k (Templates.hs:481, block offset 0), as a double
            ax = m.inst[1]
            ix = np.arange(shape[ax], dtype=np.float64) + float((offsets or [0] * len(shape))[ax])
            v = ix.reshape([-1 if a == ax else 1 for a in range(len(shape))])
        elif k == "Uniop":
            a = vals[m.inst[2]]
            op = m.inst[1]
            if op == "-":
                v = -a
            elif op == "+":
                v = a
            elif np.isscalar(a):
                v = getattr(math, op.split("/", 1)[1])(a)
            else:
                v = ext[op.split("/", 1)[1]](a)
        elif k == "Binop":
            a, b = vals[m.inst[2]], vals[m.inst[3]]
            op = m.inst[1]
            if op == "**" and np.isscalar(a) and np.isscalar(b):
                vals.append(math.pow(a, b))  # host libm pow, as the reference's C calls it
                continue
            with np.errstate(all="ignore"):
                v = {"+": lambda: a + b, "-": lambda: a - b, "*": lambda: a * b, "/": lambda: np.divide(a, b),
                     "**": lambda: np.power(a, b), "<": lambda: a < b, ">": lambda: a > b,
                     "<=": lambda: a <= b, ">=": lambda: a >= b, "==": lambda: a == b,
                     "!=": lambda: a != b, "&&": lambda: np.logical_and(a, b),
                     "||": lambda: np.logical_or(a, b)}[op]()
        elif k == "Triop":
            v = np.where(vals[m.inst[2]], vals[m.inst[3]], vals[m.inst[4]])
        elif k == "Store":
            v = vals[m.inst[2]]
        else:
            raise NotImplementedError(k)
        vals.append(v)
    return vals[-1]


def step(mm_graph, sleeve, fields, params=None, offsets=None):
    """One application of a step graph to a dict of periodic numpy fields;
    ``offsets``: the Navi offsets the kernel is called with (index reads)."""
    rng = calc_range(mm_graph)
    shape = next(iter(fields.values())).shape
    temps, out = {}, {}
    for k in sorted(mm_graph):
        node = mm_graph[k]
        is_store = node.micro[-1].inst[0] == "Store"
        base = [rng[k].offset - 2 * sleeve] * len(shape) if is_store else [0] * len(shape)
        v = np.broadcast_to(np.asarray(_eval_node(node, fields, temps, base, params, shape, offsets),
                                       dtype=np.float64), shape)
        if is_store:
            out[node.micro[-1].inst[1]] = np.array(v)
        else:
            temps[k] = np.array(v)
    return out


def init(mm_init, shape):
    """The init graph at offset 0 (Formura_Setup, Templates.hs:272-280): every node
    over the whole grid, temporaries indexed by position, no halo."""
    temps, out = {}, {}
    for k in sorted(mm_init):
        node = mm_init[k]
        is_store = node.micro[-1].inst[0] == "Store"
        v = np.broadcast_to(np.asarray(_eval_node(node, {}, temps, [0] * len(shape), None, shape),
                                       dtype=np.float64), shape)
        if is_store:
            out[node.micro[-1].inst[1]] = np.array(v)
        else:
            temps[k] = np.array(v)
    return out
