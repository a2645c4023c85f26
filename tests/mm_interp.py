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
            # i_k + offset_k (Templates.hs:481, block offset 0), as a double; an
            # inlined temporary's index read carries its nest's constant delta
            ax = m.inst[1]
            ix = (np.arange(shape[ax], dtype=np.float64) + float((offsets or [0] * len(shape))[ax])
                  + float(m.inst[2] if len(m.inst) > 2 else 0))
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


def step_nests(mm_graph, sleeve, fields, offsets=None, params=None):
    """One step evaluated literally as the reference's generated C does it
    (Templates.hs:336-346, 417-487): the state copied into buff with the
    low-side halo (buff[j] = data[(j - 2s) mod N]); every manifest node, in
    ascending id, a loop nest over its own extent N + 2s - size(range); a
    state load at cursor d reads buff[i + d + off(range)], a temporary X at
    X[i + d + off(range) - off(range X)] (X held over its own extent, not
    periodically), an index read is the nest's loop index i + offset; Store
    nodes write output i."""
    rng = calc_range(mm_graph)
    shape = next(iter(fields.values())).shape
    dim = len(shape)
    offsets = offsets or [0] * dim
    s2 = 2 * sleeve
    buff = {f: np.pad(a, [(s2, 0)] * dim, mode="wrap") for f, a in fields.items()}
    temps, out = {}, {}
    ext = {"sqrt": np.sqrt, "exp": np.exp, "fabs": np.fabs, "sin": np.sin, "cos": np.cos}
    for k in sorted(mm_graph):
        node = mm_graph[k]
        r = rng[k]
        E = [n + s2 - r.size for n in shape]
        vals = []
        for m in node.micro:
            kk = m.inst[0]
            if m.src in (params or {}) and kk not in ("LoadCursorStatic", "LoadCursor", "LoadIndex", "Store"):
                v = float(params[m.src])  # a literal bound to another value
            elif kk == "Imm":
                v = float(m.inst[1])
            elif kk == "LoadCursorStatic":
                o = [c + r.offset for c in m.inst[1]]
                v = buff[m.inst[2]][tuple(slice(a, a + e) for a, e in zip(o, E))]
            elif kk == "LoadCursor":
                o = [c + r.offset - rng[m.inst[2]].offset for c in m.inst[1]]
                v = temps[m.inst[2]][tuple(slice(a, a + e) for a, e in zip(o, E))]
            elif kk == "LoadIndex":
                ax = m.inst[1]
                ix = np.arange(E[ax], dtype=np.float64) + float(offsets[ax])
                v = ix.reshape([-1 if a == ax else 1 for a in range(dim)])
            elif kk == "Uniop":
                a = vals[m.inst[2]]
                op = m.inst[1]
                v = -a if op == "-" else (a if op == "+" else ext[op.split("/", 1)[1]](a))
            elif kk == "Binop":
                a, b = vals[m.inst[2]], vals[m.inst[3]]
                with np.errstate(all="ignore"):
                    v = {"+": lambda: a + b, "-": lambda: a - b, "*": lambda: a * b, "/": lambda: np.divide(a, b),
                         "<": lambda: a < b, ">": lambda: a > b, "<=": lambda: a <= b, ">=": lambda: a >= b,
                         "==": lambda: a == b, "!=": lambda: a != b, "&&": lambda: np.logical_and(a, b),
                         "||": lambda: np.logical_or(a, b)}[m.inst[1]]()
            elif kk == "Triop":
                v = np.where(vals[m.inst[2]], vals[m.inst[3]], vals[m.inst[4]])
            elif kk == "Store":
                v = vals[m.inst[2]]
            else:
                raise NotImplementedError(kk)
            vals.append(v)
        v = np.array(np.broadcast_to(np.asarray(vals[-1], dtype=np.float64), E))
        if node.micro[-1].inst[0] == "Store":
            assert list(E) == list(shape)
            out[node.micro[-1].inst[1]] = v
        else:
            temps[k] = v
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


def _eval_block(node, buff, base, out_shape, params, index):
    """A Store node over one block of the reference's temporally blocked update:
    output i reads buff[i + d + off(range)] (Templates.hs:473, no wrap: buff
    carries the block's high-side walls); ``index[k]`` is the value of the
    index expression along axis k for every output i (Templates.hs:481)."""
    vals = []
    params = params or {}
    ext = {"sqrt": np.sqrt, "exp": np.exp, "fabs": np.fabs, "sin": np.sin, "cos": np.cos}
    for m in node.micro:
        k = m.inst[0]
        if m.src in params and k not in ("LoadCursorStatic", "LoadCursor", "LoadIndex", "Store"):
            v = float(params[m.src])  # a literal bound to a value (the emitter's ``params``)
        elif k == "Imm":
            v = float(m.inst[1])
        elif k == "LoadCursorStatic":
            o = [c + b for c, b in zip(m.inst[1], base)]
            v = buff[m.inst[2]][tuple(slice(a, a + n) for a, n in zip(o, out_shape))]
        elif k == "LoadIndex":
            ax = m.inst[1]
            v = index[ax].reshape([-1 if a == ax else 1 for a in range(len(out_shape))])
        elif k == "Uniop":
            a = vals[m.inst[2]]
            op = m.inst[1]
            v = -a if op == "-" else (a if op == "+" else ext[op.split("/", 1)[1]](a))
        elif k == "Binop":
            a, b = vals[m.inst[2]], vals[m.inst[3]]
            with np.errstate(all="ignore"):
                v = {"+": lambda: a + b, "-": lambda: a - b, "*": lambda: a * b, "/": lambda: np.divide(a, b),
                     "**": lambda: np.power(a, b), "<": lambda: a < b, ">": lambda: a > b,
                     "<=": lambda: a <= b, ">=": lambda: a >= b, "==": lambda: a == b,
                     "!=": lambda: a != b}[m.inst[1]]()
        elif k == "Triop":
            v = np.where(vals[m.inst[2]], vals[m.inst[3]], vals[m.inst[4]])
        elif k == "Store":
            v = vals[m.inst[2]]
        else:
            raise NotImplementedError(k)
        vals.append(v)
    return vals[-1]


def forward_tb(mm_graph, sleeve, fields, nt, gpb, offsets, params=None):
    """One Formura_Forward of a temporally blocked program, restating
    Templates.hs:311-381 step by step (no MPI, one rank).

    tmp_floor[f] = data[(f - 2 s nt) mod N] for f in [0, N + 2 s nt) per axis
    (copy with offset 2 s nt, then the periodic low-side halo).  Blocks of gpb
    cells run in loop order (the first axis outermost, block index j ascending,
    so positions descend), first the region b0, then the regions bs (:316-321);
    block j starts at floor offset gpb (m - 1) - gpb j.  A block copies its
    floor into rslt, then for each of the nt sub-steps: rslt -> buff, the
    walls of the blocks above (written by them at the same sub-step) -> the
    high side of buff, one Formura_Step with the Navi of the Forward's start
    and the block's floor offset as block offsets (an index reads
    i + offset + block_offset), and the low side of buff -> the walls.  rslt
    goes back to the floor, and the floor's first N cells become the state.
    The walls are static arrays: zero before their first write.  Programs
    without manifest temporaries."""
    s = sleeve
    names = sorted(fields)
    shape = fields[names[0]].shape
    dim = len(shape)
    rng = calc_range(mm_graph)
    for k in mm_graph:
        if mm_graph[k].micro[-1].inst[0] != "Store":
            raise NotImplementedError("temporaries under temporal blocking")
    h = 2 * s * nt
    fsize = [n + h for n in shape]
    for f in fsize:
        assert f % gpb == 0
    m = [f // gpb for f in fsize]
    floor = {f: fields[f][np.ix_(*[(np.arange(fs) - h) % n for fs, n in zip(fsize, shape)])] for f in names}
    walls = [dict() for _ in range(dim)]
    dd = [h // gpb for _ in range(dim)]
    b0 = [(0, m[a] - 1 - dd[a]) for a in range(dim)]
    bs = [[((m[a] - 1 - dd[a]) if a == j else 0, (m[a] - 1 - dd[a]) if a > j else m[a]) for a in range(dim)]
          for j in range(dim)]

    def blocks(region):
        import itertools
        return itertools.product(*[range(l, u) for l, u in region])

    def update(J):
        fo = [gpb * (m[a] - 1) - gpb * J[a] for a in range(dim)]
        sl = tuple(slice(o, o + gpb) for o in fo)
        rslt = {f: floor[f][sl].copy() for f in names}
        for it in range(nt):
            buff = {f: np.zeros([gpb + 2 * s] * dim) for f in names}
            for f in names:
                buff[f][tuple(slice(0, gpb) for _ in range(dim))] = rslt[f]
            for a in range(dim):
                key = tuple(J[b] for b in range(dim) if b != a) + (it,)
                region = tuple(slice(gpb, gpb + 2 * s) if b == a else slice(0, gpb + 2 * s) for b in range(dim))
                for f in names:
                    buff[f][region] = walls[a].get(key, {}).get(f, 0.0)
            index = [np.arange(gpb, dtype=np.float64) + float(offsets[a] + fo[a]) for a in range(dim)]
            new = {}
            for k in sorted(mm_graph):
                node = mm_graph[k]
                base = [rng[k].offset] * dim
                new[node.micro[-1].inst[1]] = np.array(np.broadcast_to(
                    np.asarray(_eval_block(node, buff, base, [gpb] * dim, params, index), dtype=np.float64),
                    [gpb] * dim))
            rslt = new
            for a in range(dim):
                key = tuple(J[b] for b in range(dim) if b != a) + (it,)
                region = tuple(slice(0, 2 * s) if b == a else slice(0, gpb + 2 * s) for b in range(dim))
                walls[a][key] = {f: buff[f][region].copy() for f in names}
        for f in names:
            floor[f][sl] = rslt[f]

    for J in blocks(b0):
        update(J)
    for region in bs:
        for J in blocks(region):
            update(J)
    return {f: floor[f][tuple(slice(0, n) for n in shape)].copy() for f in names}


_NP_TYPE = {"float": np.float32, "double": np.float64, "Rational": np.float64}


def step_typed(mm_graph, sleeve, fields, offsets=None):
    """One step of a program with FP32 (``float``) values, in the C types of the
    generated code (Templates.hs:459-467): every micro value is held in its
    declared type; an operation on two floats is done in float, one with a
    double operand in double (the usual arithmetic conversions), and the result
    is converted to the declared type on assignment.  Stores only (no
    temporaries); index reads as in ``step``."""
    rng = calc_range(mm_graph)
    shape = next(iter(fields.values())).shape
    out = {}
    for k in sorted(mm_graph):
        node = mm_graph[k]
        if node.micro[-1].inst[0] != "Store":
            raise NotImplementedError("temporaries")
        base = [rng[k].offset - 2 * sleeve] * len(shape)
        vals, types = [], []
        for m in node.micro:
            kk = m.inst[0]
            t = _NP_TYPE.get(str(m.typ), np.float64)
            if kk == "Imm":
                v = np.float64(float(m.inst[1]))
            elif kk == "LoadCursorStatic":
                v = _shift(fields[m.inst[2]], [c + b for c, b in zip(m.inst[1], base)])
            elif kk == "LoadIndex":
                ax = m.inst[1]
                ix = np.arange(shape[ax], dtype=np.float64) + float((offsets or [0] * len(shape))[ax])
                v = ix.reshape([-1 if a == ax else 1 for a in range(len(shape))])
            elif kk == "Uniop" and m.inst[1] == "-":
                v = -vals[m.inst[2]]
            elif kk == "Binop":
                a, b = vals[m.inst[2]], vals[m.inst[3]]
                wide = np.float64 if np.float64 in (types[m.inst[2]], types[m.inst[3]]) else np.float32
                a, b = np.asarray(a, dtype=wide), np.asarray(b, dtype=wide)
                op = m.inst[1]
                v = a + b if op == "+" else a - b if op == "-" else a * b if op == "*" else a / b
            elif kk == "Store":
                out[m.inst[1]] = np.array(np.broadcast_to(vals[m.inst[2]], shape)).astype(fields[m.inst[1]].dtype)
                vals.append(None)
                types.append(None)
                continue
            else:
                raise NotImplementedError(kk)
            vals.append(np.asarray(v, dtype=t))
            types.append(t)
    return out
