"""Adjoint (reverse-mode) kernels generated from a program's step graph.

For a step ``out = S(in; p)`` the adjoint maps ``lam_out`` (dL/d out) to
``lam_in`` (dL/d in) and accumulates dL/dp for the lifted literals ``p``.

The kernel is written in gather form, one work-item per input point q.
Output i reads input i + d + off - 2s, so input q is read by outputs
i = q - delta for every distinct load offset delta = d + off - 2s.  For each
delta, the kernel does the following:

* evaluate the step's forward instruction list at output q - delta;
* run the reverse sweep of each Store node, restricted to the instructions on
  a path from a load at offset delta to the Store (a transposed stencil entry
  times the linearised update);
* add it to lam_in[q].

Paths that are linear in literals fold to a few multiplications of lam_out.
The nonlinear part (the reaction term, or the flux Jacobian) appears only on
the paths through the cell's own value.  Literal parameters get their own
restricted sweep at output i = q and are reduced per workgroup in FP64.
Contributions are summed in a fixed order, so results are deterministic (no
atomics).

The library also contains the forward step of the same program with the
parameters as run-time values.  Its instruction list is the plain path's,
so at the program's own parameters it is bit-identical to the plain path.
The library's C ABI is ``include/formura_adjoint.h``.
"""
from __future__ import annotations

import json
import math
from typing import Dict, List

from . import syntax as S
from .emit import Body, EmitError, LDS_LOAD_LINE, tile_load, GraphPlan, _const_literal, _f32, show_double, state_fields
from .manifest import calc_range, inline_temporaries
from .symadj import SymbolicAdjoint, Unsupported


def _operands(inst):
    k = inst[0]
    if k == "Uniop":
        return [inst[2]]
    if k == "Binop":
        return [inst[2], inst[3]]
    if k == "Triop":
        return [inst[2], inst[3], inst[4]]
    if k == "Store":
        return [inst[2]]
    return []


class ReverseSweep:
    """Reverse-mode derivative of one Store node's instruction list, restricted
    to paths that end in ``targets`` (micro ids).  Emits C statements; the
    forward values are the variables ``fwd + str(j)``."""

    def __init__(self, node, targets, seed: str, fwd: str, out: str, consts: Dict[int, float],
                 real: str, leaves=()):
        micro = node.micro
        leaves = set(leaves)
        on = [False] * len(micro)
        for j, m in enumerate(micro):
            if j in targets:
                on[j] = True
            elif j not in leaves:
                on[j] = any(on[o] for o in _operands(m.inst))
        self.lines: List[str] = []
        self.result: Dict[int, str] = {}
        root = len(micro) - 1
        if not on[root]:
            return
        contrib: Dict[int, List[str]] = {root: [seed]}
        R = real

        def v(j):
            return "%s%d" % (fwd, j)

        def const_inv(j):
            if j in consts and consts[j] not in (0.0,) and isinstance(consts[j], float):
                c = 1.0 / consts[j]
                return _const_literal(_f32(c), "float") if R == "float" else show_double(c)
            return None

        for j in range(root, -1, -1):
            if not on[j] or j not in contrib:
                continue
            g = "%s%d" % (out, j)
            self.lines.append("const %s %s = %s;" % (R, g, " + ".join(contrib[j])))
            if j in targets:
                self.result[j] = g
            if j in leaves:
                continue
            m = micro[j]
            k = m.inst[0]
            ops = _operands(m.inst)

            def add(o, expr):
                if on[o]:
                    contrib.setdefault(o, []).append("(%s)" % expr)

            if k == "Store":
                add(ops[0], g)
            elif k == "Binop":
                op, a, b = m.inst[1], m.inst[2], m.inst[3]
                if op == "+":
                    add(a, g)
                    add(b, g)
                elif op == "-":
                    add(a, g)
                    add(b, "-" + g)
                elif op == "*":
                    add(a, "%s*%s" % (g, v(b)))
                    add(b, "%s*%s" % (g, v(a)))
                elif op == "/":
                    inv = const_inv(b)
                    add(a, "%s*%s" % (g, inv) if inv else "%s/%s" % (g, v(b)))
                    add(b, "-%s*%s/%s" % (g, v(j), v(b)))
                elif op == "**":
                    add(a, "%s*%s*pow(%s,%s-1)" % (g, v(b), v(a), v(b)))
                    add(b, "%s*%s*log(%s)" % (g, v(j), v(a)))
                # comparisons and logic carry no derivative
            elif k == "Uniop":
                op, a = m.inst[1], m.inst[2]
                if op == "-":
                    add(a, "-" + g)
                elif op == "+":
                    add(a, g)
                elif op == "external-call/sqrt":
                    add(a, "%s/(2*%s)" % (g, v(j)))
                elif op == "external-call/exp":
                    add(a, "%s*%s" % (g, v(j)))
                elif op == "external-call/fabs":
                    add(a, "(%s >= 0 ? %s : -%s)" % (v(a), g, g))
                elif op == "external-call/log":
                    add(a, "%s/%s" % (g, v(a)))
                elif op == "external-call/sin":
                    add(a, "%s*cos(%s)" % (g, v(a)))
                elif op == "external-call/cos":
                    add(a, "-%s*sin(%s)" % (g, v(a)))
                else:
                    raise EmitError("no derivative for %s" % op)
            elif k == "Triop":
                c, a, b = m.inst[2], m.inst[3], m.inst[4]
                add(a, "(%s ? %s : (%s)0)" % (v(c), g, R))
                add(b, "(%s ? (%s)0 : %s)" % (v(c), R, g))


class AdjointEmitter:
    def __init__(self, mm, name: str, params: List[str], real: str = "double", block: int = 256,
                 tile=(32, 32, 4), adj_tile=(32, 32), levels=4, row_block=1, fwd_threads=None, adj_threads=None,
                 symbolic: bool = True, fused: bool = True, adj_prefetch: bool = True, adj_index_arrays: bool = True,
                 unit_inplace=None, unit_next: bool = True, long_jobs_first: bool = True):
        self.mm, self.om, self.name = mm, mm.om, name
        self.symbolic = bool(symbolic)
        # units of adjoint steps through k_adjoint_unit (odd levels recomputed in the tile)
        self.fused_ok = bool(fused)
        # k_adjoint_unit: lam updated in place in one LDS buffer per field (else
        # ping-pong).  Default: in place for double, whose two lam buffers would
        # hold a CU to two workgroups (in float four fit either way, and the extra
        # barrier per level costs more than it gains)
        self.unit_inplace = (real == "double") if unit_inplace is None else bool(unit_inplace)
        # k_adjoint_unit_next: a unit's adjoint and, in the same launch, the
        # level-2 state of the unit reversed next (when the executor has it)
        self.unit_next = bool(unit_next)
        # the executor places a launch's 4-step forward job before its 2-step one
        # (the longer workgroups start first and the shorter ones pack the end:
        # FP64 reversal 4.9 % faster than in program order, FP32 0.5 %)
        self.long_jobs_first = bool(long_jobs_first)
        # pointwise adjoint kernel: fetch the next level's state while a level
        # computes (registers), and keep element indices in registers
        self.adj_prefetch, self.adj_index_arrays = bool(adj_prefetch), bool(adj_index_arrays)
        self.real = real
        self.tile = tuple(tile) if tile else None
        # tile of the adjoint kernel (defaults to the forward kernel's)
        self.adj_tile = tuple(adj_tile) if adj_tile else (self.tile[:2] if self.tile else None)
        # most steps one fused adjoint launch reverses (2-D tiled kernels)
        self.levels = max(1, int(levels))
        # points per work-item along the first axis in the adjoint levels kernel
        self.row_block = int(row_block)
        # work-items per workgroup of the tiled forward and adjoint kernels (the
        # forward: 1024 for FP64 state, 512 for FP32; the adjoint unit with the
        # forward jobs it carries: 512 for FP64, 256 for FP32 -- the faster on a
        # 1024^2 Gray-Scott reversal for each)
        if fwd_threads is None:
            fwd_threads = 1024 if real == "double" else 512
        if adj_threads is None:
            adj_threads = 512 if real == "double" else 256
        self.fwd_threads, self.adj_threads = int(fwd_threads), int(adj_threads)
        self.dim = self.om.dimension
        self.N = self.om.config.grid_per_node
        self.params = list(params)
        step = mm.step
        if any(n.micro[-1].inst[0] != "Store" for n in step.values()):
            step = inline_temporaries(step)
        self.plan = GraphPlan(step, self.om.config.sleeve, self.dim, "step")
        self.fields = state_fields(self.om)
        self.fi = {n: i for i, (n, _) in enumerate(self.fields)}
        self.block = block
        names = {m.src for n in self.plan.mm.values() for m in n.micro if m.src}
        missing = [p for p in self.params if p not in names]
        if missing:
            raise EmitError("parameters %s are not named values of the step" % missing)
        for n in self.plan.mm.values():
            if any(m.inst[0] == "LoadIndex" for m in n.micro):
                raise EmitError("index expressions are not supported by the adjoint generator")

    # -- helpers --------------------------------------------------------------
    def _ivs(self):
        return ["i%d" % (k + 1) for k in range(self.dim)]

    def _idx(self, names):
        e = names[0]
        for k in range(1, self.dim):
            e = "(%s)*N%d+%s" % (e, k + 1, names[k])
        return e

    def _pos(self, shift):
        ivs = self._ivs()
        return [("fmr_mod<%d>(%s%+d)" % (self.N[k], ivs[k], shift[k])) if shift[k] else ivs[k]
                for k in range(self.dim)]

    def _decompose(self):
        ivs = self._ivs()
        lines = ["const long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x;",
                 "const bool live = t < NTOT;",
                 "const long long tt = live ? t : 0;"]
        rem = "tt"
        for k in range(self.dim - 1, -1, -1):
            if k == 0:
                lines.append("const int %s = (int)(%s);" % (ivs[k], rem))
            else:
                lines.append("const int %s = (int)(%s %% N%d);" % (ivs[k], rem, k + 1))
                rem = "(%s / N%d)" % (rem, k + 1)
        return lines

    def _forward_body(self, k, shift, prefix, xload, exact=True):
        """Forward instructions of Store node k at output point (q + shift); the
        state value of field f at q + a is ``xload(f, a)``.  With ``exact`` off
        (values that only feed derivatives) constant divisors become reciprocal
        multiplications."""
        s = self.plan.sleeve
        base = self.plan.rng[k].offset - 2 * s

        def load(field, c):
            return xload(self.fi[field], [shift[a] + c[a] + base for a in range(self.dim)])

        pmap = {p: "P[%d]" % i for i, p in enumerate(self.params)}
        # exact FP64 bodies divide by constants without a branch (the fused
        # sequence; a wave holding a dividend out of its range redoes that
        # division, a wave-uniform branch), as the plain path's tile kernels do
        return Body(self.plan.mm[k], load, lambda *a: None, lambda ax: None, lambda n, v: "", prefix,
                    params=pmap, real=("float" if self.real == "float" else None), recip=not exact,
                    divfix=exact and self.real == "double")

    def deltas(self):
        """Distinct load offsets delta = c + base over all stores and fields."""
        s = self.plan.sleeve
        out = []
        for k in self.plan.stores:
            base = self.plan.rng[k].offset - 2 * s
            for m in self.plan.mm[k].micro:
                if m.inst[0] == "LoadCursorStatic":
                    d = tuple(c + base for c in m.inst[1])
                    if d not in out:
                        out.append(d)
        return sorted(out)

    @staticmethod
    def _state_dependent(node):
        on = []
        for m in node.micro:
            on.append(m.inst[0] in ("LoadCursorStatic", "LoadCursor", "LoadIndex")
                      or any(on[o] for o in _operands(m.inst)))
        return on

    def point_body(self, xload, lload, pguard=None):
        """Statements computing acc<g> (dL/d input field g at q) and pg<p> (the
        parameter-gradient contribution of point q) for one point q; returns
        (lines, uses_state): uses_state is False when no derivative reads a
        state-dependent forward value (a step linear in the state with no
        parameters: its adjoint does not need the forward states).

        With ``symbolic`` (the default) the body comes from
        ``symadj.SymbolicAdjoint`` where the step allows it (grouped transposed
        stencil, parameter gradients summed by parts); otherwise, and with
        ``symbolic`` off, from the reverse sweep of the instruction list per
        load offset (``sweep_point_body``)."""
        sa = self.symbolic_adjoint()
        if sa is not None:
            try:
                nf, np_ = len(self.fields), max(1, len(self.params))
                lines, uses = sa.point_body(xload, lload, ["acc%d" % g for g in range(nf)],
                                            ["pg%d" % p for p in range(len(self.params))])
                if not self.params:
                    lines.append("double pg0 = 0;")
                return lines, uses
            except Unsupported:
                pass
        return self.sweep_point_body(xload, lload, pguard)

    def symbolic_adjoint(self):
        """The step's ``SymbolicAdjoint``, or None (``symbolic`` off, or a step
        it does not handle)."""
        if not self.symbolic:
            return None
        if not hasattr(self, "_sym"):
            try:
                self._sym = SymbolicAdjoint(self.plan, self.dim, self.fields, self.params, self.real)
            except Unsupported:
                self._sym = None
        return self._sym

    def pointwise_state(self) -> bool:
        """True when the adjoint body reads the state at the point only (or not
        at all): the tiled adjoint then keeps each level's state in registers
        instead of staging a state tile with halo in LDS."""
        sa = self.symbolic_adjoint()
        if sa is None:
            return False
        try:
            return sa.state_offsets() <= {tuple([0] * self.dim)}
        except Unsupported:
            return False

    def fused_unit(self) -> bool:
        """True when a unit of up to 4 adjoint steps runs as a two-step forward
        launch (the unit's level-2 state) and one ``k_adjoint_unit`` launch that
        recomputes each odd level from the even level below it in its tile:
        tiled 2-D programs whose adjoint reads the state at the point only."""
        return bool(self.fused_ok and self.tiled() and self.pointwise_state() and self.levels <= 4)

    def sweep_point_body(self, xload, lload, pguard=None):
        """``point_body`` by the reverse sweep of the step's instruction list:
        for each load offset delta, the forward instructions at output q - delta
        and the sweep restricted to the paths from the loads at delta; the
        parameter gradients at output q (under ``pguard``)."""
        import re
        s = self.plan.sleeve
        dim = self.dim
        R = self.real
        nf = len(self.fields)
        np_ = max(1, len(self.params))
        uses_state = bool(self.params)
        body = ["%s acc%d = 0;" % (R, g) for g in range(nf)]
        for di, delta in enumerate(self.deltas()):
            shift = [-x for x in delta]  # output point i = q - delta
            body.append("{  // outputs that read this point at offset %s" % (delta,))
            for si, k in enumerate(self.plan.stores):
                node = self.plan.mm[k]
                base = self.plan.rng[k].offset - 2 * s
                fb = self._forward_body(k, shift, "f%d_%d_" % (di, si), xload, exact=False)
                dep = self._state_dependent(node)
                body += ["  " + l for l in fb.lines]
                consts = getattr(fb, "vals", {})
                body.append("  const %s lam%d_%d = %s;" % (R, di, si, lload(self.fi[node.micro[-1].inst[1]], shift)))
                for g, (gname, _) in enumerate(self.fields):
                    targets = {j for j, m in enumerate(node.micro)
                               if m.inst[0] == "LoadCursorStatic" and m.inst[2] == gname
                               and tuple(c + base for c in m.inst[1]) == delta}
                    if not targets:
                        continue
                    pleaves = {j for j, m in enumerate(node.micro) if m.src in self.params}
                    rs = ReverseSweep(node, targets, "lam%d_%d" % (di, si), "f%d_%d_" % (di, si),
                                      "b%d_%d_%d_" % (di, si, g), consts, R, leaves=pleaves)
                    if rs.result:
                        for l in rs.lines:
                            for j in re.findall(r"\bf%d_%d_(\d+)\b" % (di, si), l):
                                if dep[int(j)] and int(j) not in consts:
                                    uses_state = True
                        body += ["  " + l for l in rs.lines]
                        body.append("  acc%d += %s;" % (g, " + ".join(rs.result[j] for j in sorted(rs.result))))
            body.append("}")
        # parameter gradients at output i = q
        body += ["double pg%d = 0;" % p for p in range(np_)]
        if self.params:
            body.append("if (%s) {" % pguard if pguard else "{")
            for si, k in enumerate(self.plan.stores):
                node = self.plan.mm[k]
                fb = self._forward_body(k, [0] * dim, "p%d_" % si, xload, exact=False)
                body += ["  " + l for l in fb.lines]
                body.append("  const %s plam%d = %s;" % (R, si, lload(self.fi[node.micro[-1].inst[1]], [0] * dim)))
                for pi, p in enumerate(self.params):
                    leaves = {j for j, m in enumerate(node.micro) if m.src in self.params}
                    targets = {j for j, m in enumerate(node.micro) if m.src == p}
                    rs = ReverseSweep(node, targets, "plam%d" % si, "p%d_" % si, "pb%d_%d_" % (si, pi),
                                      getattr(fb, "vals", {}), R, leaves=leaves)
                    if rs.result:
                        body += ["  " + l for l in rs.lines]
                        body.append("  pg%d += (double)(%s);" % (pi, " + ".join(rs.result[j] for j in sorted(rs.result))))
            body.append("}")
        return body, uses_state

    def tiled(self):
        """Tile shape for the LDS-tiled 2-D kernels, or None (3-D programs and
        grids the tile does not divide use the one-point-per-thread kernels)."""
        if self.dim != 2 or self.tile is None:
            return None
        tx, ty, k = self.tile
        s = self.plan.sleeve
        ax, ay = self.adj_tile
        if (self.N[0] % tx or self.N[1] % ty or self.N[0] % ax or self.N[1] % ay
                or 2 * (self.levels + 1) * s >= min(self.N) or 2 * k * s >= min(self.N)):
            return None
        return self.tile

    def emit(self) -> str:
        s = self.plan.sleeve
        dim = self.dim
        R = self.real
        nf = len(self.fields)
        np_ = max(1, len(self.params))
        deltas = self.deltas()

        def gidx(a):
            return self._idx(self._pos(a))

        flat, uses_state = self.point_body(lambda f, a: "FLD%d(st)[%s]" % (f, gidx(a)),
                                           lambda f, a: "LAM%d(lam_next)[%s]" % (f, gidx(a)))
        body = self._decompose() + flat
        body += ["if (live) LAM%d(lam_prev)[t] = acc%d;" % (g, g) for g in range(nf)]
        body += ["pg%d = live ? pg%d : 0.0;" % (p, p) for p in range(np_)]
        body.append("{ const double pv[NP] = {%s}; block_reduce_params<%d>(pgrad + (long long)blockIdx.x * NP, pv); }"
                    % (", ".join("pg%d" % p for p in range(np_)), self.block))
        # forward kernel (parameters at run time)
        fwd = self._decompose()
        fwd.append("if (!live) return;")
        for si, k in enumerate(self.plan.stores):
            fb = self._forward_body(k, [0] * dim, "s%d_" % si, lambda f, a: "FLD%d(st)[%s]" % (f, gidx(a)))
            fwd.append("{")
            fwd += ["  " + l for l in fb.lines]
            fwd.append("  FLD%d(out)[t] = s%d_%d;" % (self.fi[self.plan.mm[k].micro[-1].inst[1]], si,
                                                    self.plan.mm[k].micro[-1].inst[2]))
            fwd.append("}")
        tile = self.tiled()
        tiled_src = self._emit_tiles(tile, uses_state) if tile else ""
        desc = {"name": self.name, "real": R, "fields": [n for n, _ in self.fields], "params": self.params,
                "grid": self.N, "sleeve": s, "block": self.block, "deltas": [list(d) for d in deltas],
                "state_free": not uses_state, "tile": list(tile) if tile else None,
                "adjoint_tile": list(self.adj_tile) if tile else None}
        flds = "\n".join("#define FLD%d(p) ((%s*)((p) + (size_t)%d * NTOT * sizeof(%s)))" % (i, R, i, R)
                         for i in range(nf))
        lams = "\n".join("#define LAM%d(p) ((%s*)((p) + (size_t)%d * NTOT * sizeof(%s)))" % (i, R, i, R)
                         for i in range(nf))
        if tile:
            nt = (self.N[0] // self.adj_tile[0]) * (self.N[1] // self.adj_tile[1])
            fwd_call = "  fwd_tiles(src, dst_of, nsteps, params_dev, s);"
            adj_call = "  adj_levels(1, (const char*)state_prev, nullptr, (const char*)lam_next, (char*)lam_prev, params_dev, pgrad, s);"
            steps_call = ("  if (nsteps > 1) fwd_levels(nsteps - 1, (const char*)state0, (char*)scratch, params_dev, s);\n"
                          "  adj_levels(nsteps, (const char*)state0, (const char*)scratch, (const char*)lam_next, "
                          "(char*)lam_prev, params_dev, pgrad, s);")
            if self.fused_unit():
                steps_call = ("  if (nsteps > 2) fwd_even((const char*)state0, (char*)scratch, params_dev, s);\n"
                              "  adj_unit(nsteps, (const char*)state0, (const char*)scratch, (const char*)lam_next, "
                              "(char*)lam_prev, params_dev, pgrad, s);")
            inner_call = "  if (nlevels > 0) fwd_levels(nlevels, (const char*)state0, (char*)scratch, params_dev, s);"
            levels_call = ("  adj_levels(nsteps, (const char*)state0, (const char*)scratch, (const char*)lam_next, "
                           "(char*)lam_prev, params_dev, pgrad, s);")
            slots = nt
            levels = self.levels
        else:
            fwd_call = ("  for (int t = 0; t < nsteps; ++t) {\n"
                        "    char* dst = dst_of(nsteps - 1 - t);\n"
                        "    hipLaunchKernelGGL(k_forward, dim3(nb), dim3(BLOCK), 0, s, src, dst, params_dev);\n"
                        "    src = dst;\n  }")
            adj_call = ("  hipLaunchKernelGGL(k_adjoint, dim3(nb), dim3(BLOCK), 0, s, (const char*)state_prev, "
                        "(const char*)lam_next, (char*)lam_prev, params_dev, pgrad);")
            steps_call = ("  (void)scratch;\n"
                          "  hipLaunchKernelGGL(k_adjoint, dim3(nb), dim3(BLOCK), 0, s, (const char*)state0, "
                          "(const char*)lam_next, (char*)lam_prev, params_dev, pgrad);")
            inner_call = "  (void)state0; (void)scratch; (void)params_dev; (void)s;"
            levels_call = steps_call
            slots = None
            levels = 1
        desc["levels"] = levels
        desc["fused_unit"] = bool(tile and self.fused_unit())
        return ADJ_TEMPLATE % dict(
            name=self.name, R=R, nf=nf, np=np_, has_params="true" if self.params else "false",
            block=self.block, sleeve=s,
            consts="\n".join("constexpr int N%d = %d;" % (k + 1, n) for k, n in enumerate(self.N)),
            ntot="*".join("(long long)%d" % n for n in self.N), flds=flds, lams=lams,
            desc=json.dumps(desc, sort_keys=True), fwd="\n".join("  " + l for l in fwd), adj="\n".join("  " + l for l in body),
            fused="true" if (tile and self.fused_unit()) else "false",
            unit_stub="" if (tile and self.fused_unit()) else UNIT_STUB,
            tiled=tiled_src, fwd_call=fwd_call, adj_call=adj_call, steps_call=steps_call, levels=levels,
            inner_call=inner_call, levels_call=levels_call,
            slots=("%d" % slots) if slots else "(int)((NTOT + BLOCK - 1) / BLOCK)")

    def _emit_tiles(self, tile, uses_state):
        """LDS-tiled 2-D kernels.

        k_forward_tile<K, ALL>: K steps per launch (tile plus a low-side halo of
        2*K*s, ping-pong LDS levels), parameters at run time.  With ALL every
        level's tile is also written out: level j to buffer j-1 (the inner
        states a fused adjoint needs).

        k_adjoint_levels<K>: K adjoint steps per launch.  Input point q is read
        by outputs q - delta with delta in [-2s, 0]; they read inputs in
        [q - 2s, q + 2s].  Reversing K steps for a TX x TY tile therefore needs
        lam_K on the tile plus a high-side halo of 2*K*s, and the state X_m of
        level m on [-2s, T + 2(m+1)s) around the tile.  lam lives in two LDS
        buffers (ping-pong); X_m is loaded into LDS from HBM level by level.
        The parameter gradients of level m are summed over the tile's own
        outputs only, so each output counts once."""
        s = self.plan.sleeve
        R = self.real
        nf = len(self.fields)
        np_ = max(1, len(self.params))
        tx, ty, kmax = tile
        ax, ay = self.adj_tile
        # forward point function on LDS levels
        pbody, ploads = [], []
        for si, k in enumerate(self.plan.stores):
            node = self.plan.mm[k]
            fb = self._forward_body(k, [0, 0], "s%d_" % si,
                                    lambda f, a: "lds%d[o + (%d)*RY + (%d)]" % (f, a[0], a[1]))
            # the node's LDS loads (names unique to the node) before every node's arithmetic
            ploads += [l for l in fb.lines if LDS_LOAD_LINE.search(l)]
            pbody.append("{")
            pbody += ["  " + l for l in fb.lines if not LDS_LOAD_LINE.search(l)]
            pbody.append("  o%d = s%d_%d;" % (self.fi[node.micro[-1].inst[1]], si, node.micro[-1].inst[2]))
            pbody.append("}")
        pbody = ploads + pbody
        lparams = ", ".join("const real_t* __restrict__ lds%d" % i for i in range(nf))
        outs = ", ".join("real_t& o%d" % i for i in range(nf))
        src = []
        src.append("template <int RY>\n__device__ __forceinline__ void pt_step(%s, const int o, const double* __restrict__ P, %s) {\n%s\n}\n"
                   % (lparams, outs, "\n".join("  " + l for l in pbody)))
        lds = "\n".join("  __shared__ real_t B%d[2][RX * RY];" % i for i in range(nf))
        load_tile = tile_load("RX * RY", ["const int r = idx / RY, c = idx - r * RY;",
                                          "const long long g = (long long)wrapn(x0 - H + r, N1) * N2 + wrapn(y0 - H + c, N2);"],
                              [("B%d[0][idx]" % i, "FLD%d(st_in)[g]" % i) for i in range(nf)], "real_t", "ld", threads=self.fwd_threads)
        srcs = ", ".join("B%d[sb]" % i for i in range(nf))
        decl = " ".join("real_t o%d;" % i for i in range(nf))
        refs = ", ".join("o%d" % i for i in range(nf))
        st_lds = " ".join("B%d[db][o] = o%d;" % (i, i) for i in range(nf))
        st_lvl = " ".join("FLD%d(lvl)[g] = o%d;" % (i, i) for i in range(nf))
        st_glb = " ".join("FLD%d(dst)[g] = o%d;" % (i, i) for i in range(nf))
        fwd_args = dict(TX=tx, TY=ty, lds=lds, srcs=srcs, decl=decl, refs=refs, st_lds=st_lds, st_lvl=st_lvl,
                        st_glb=st_glb)
        src.append(TILE_FWD % dict(fwd_args, FTH=self.fwd_threads, load_tile=load_tile))
        if self.pointwise_state():
            src.append(self._emit_adj_pointwise(uses_state))
        else:
            src.append(self._emit_adj_staged(uses_state))
        ucases = ""
        if self.fused_unit():
            src.append(self._emit_adj_unit(uses_state))
            ucases = "".join("    case %d: hipLaunchKernelGGL(k_adjoint_unit<%d>, dim3(NTA), dim3(%d), 0, s, st0, ev, lam_next, "
                             "lam_prev, P, pgrad); break;\n" % (k, k, self.adj_threads)
                             for k in range(1, self.levels + 1))
        fcases = "".join("    case %d: hipLaunchKernelGGL((k_forward_tile<%d, false>), dim3(NT), dim3(%d), 0, s, in, out, P); break;\n"
                         % (k, k, self.fwd_threads) for k in range(1, kmax + 1))
        lcases = "".join("    case %d: hipLaunchKernelGGL((k_forward_tile<%d, true>), dim3(NT), dim3(%d), 0, s, in, out, P); break;\n"
                         % (k, k, self.fwd_threads) for k in range(1, self.levels))
        acases = "".join("    case %d: hipLaunchKernelGGL(k_adjoint_levels<%d>, dim3(NTA), dim3(%d), 0, s, st0, lv, lam_next, "
                         "lam_prev, P, pgrad); break;\n" % (k, k, self.adj_threads) for k in range(1, self.levels + 1))
        src.append(TILE_HOST % dict(TX=tx, TY=ty, AX=ax, AY=ay, K=kmax, fcases=fcases, lcases=lcases, acases=acases))
        if ucases:
            src.append(FWD_JOB)
            if self.unit_next:
                src.append(self._emit_unit_next(uses_state, fwd_args))
            src.append(UNIT_HOST % dict(ucases=ucases, ncases=self._unit_next_cases(),
                                        has_next="true" if self.unit_next else "false",
                                        long_first="true" if self.long_jobs_first else "false"))
        return "\n".join(src)

    def _unit_next_cases(self):
        if not self.unit_next:
            return ""
        return "".join("    case %d: hipLaunchKernelGGL(k_adjoint_unit_fwd<%d>, dim3(NTA + nj * NT), dim3(%d), 0, s, st0, ev, "
                       "lam_next, lam_prev, P, pgrad, nj, j0, j1); break;\n" % (k, k, self.adj_threads)
                       for k in range(1, self.levels + 1))

    def _emit_unit_next(self, uses_state, fwd_args):
        """k_adjoint_unit_fwd<K>: the workgroups [0, NTA) run k_adjoint_unit<K>'s
        tiles, the workgroups after them the forward tiles of up to two forward
        jobs (2 or 4 steps: a level-2 state, a recomputation launch) that the
        executor found independent of the adjoint.  The block dispatcher starts
        the forward tiles as the adjoint's last round leaves slots free.  Both
        parts are the kernels' own bodies as device functions over one LDS pool
        (the larger of the two footprints)."""
        nf = len(self.fields)
        s = self.plan.sleeve
        tx, ty, _ = self.tile
        ax, ay = self.adj_tile
        th = self.adj_threads
        if (tx, ty) != (ax, ay) or self.tile[2] < 4:
            raise EmitError("k_adjoint_unit_fwd needs the same forward and adjoint tile and 4-step forward launches")
        lt = tile_load("RX * RY", ["const int r = idx / RY, c = idx - r * RY;",
                                   "const long long g = (long long)wrapn(x0 - H + r, N1) * N2 + wrapn(y0 - H + c, N2);"],
                       [("B%d[0][idx]" % i, "FLD%d(st_in)[g]" % i) for i in range(nf)], "real_t", "ld", threads=th)
        fwd = TILE_FWD % dict(fwd_args, FTH=th, load_tile=lt)
        fwd = fwd[fwd.index("template <int K, bool ALL>"):]
        fwd = _device_part(fwd, "k_forward_tile", "fwd_part",
                           ["  __shared__ real_t B%d[2][RX * RY];" % i for i in range(nf)],
                           ["  real_t (*B%d)[RX * RY] = reinterpret_cast<real_t (*)[RX * RY]>(pool + %d * RX * RY);"
                            % (i, 2 * i) for i in range(nf)])
        adj = self._emit_adj_unit(uses_state)
        nl = 1 if self.unit_inplace else 2
        lpool = ((["  real_t* L%d = pool + %d * LH * LW;" % (i, nf + i) for i in range(nf)]) if self.unit_inplace else
                 (["  real_t (*L%d)[LH * LW] = reinterpret_cast<real_t (*)[LH * LW]>(pool + %d * LH * LW);"
                   % (i, nf + 2 * i) for i in range(nf)]))
        adj = _device_part(adj, "k_adjoint_unit", "adj_part",
                           ["  __shared__ real_t F%d[LH * LW];" % i for i in range(nf)]
                           + [("  __shared__ real_t L%d[LH * LW];" if self.unit_inplace else
                               "  __shared__ real_t L%d[2][LH * LW];") % i for i in range(nf)],
                           ["  real_t* F%d = pool + %d * LH * LW;" % (i, i) for i in range(nf)] + lpool)
        lh, lw = ax + 2 * self.levels * s, ay + 2 * self.levels * s
        pool = max(nf * (1 + nl) * lh * lw, nf * 2 * (tx + 8 * s) * (ty + 8 * s))
        # as many workgroups per CU as the unit kernel's LDS admits (the register
        # budget of the two parts together would otherwise allow fewer)
        # (HIP's second launch bound is the minimum number of waves per SIMD)
        esize = 8 if self.real == "double" else 4
        minb = max(1, min(2048 // th, (160 * 1024) // (pool * esize + 1024)))
        # waves per SIMD for minb workgroups of th / 64 waves on 4 SIMDs, rounded up
        return "\n".join([fwd, adj, UNIT_NEXT % dict(ATH=th, POOL=pool, MINW=max(1, -(-minb * th // 256)))])

    def _emit_adj_unit(self, uses_state):
        """k_adjoint_unit<K>: K (<= 4) adjoint levels over the adjoint tile; the
        state of an odd level m is recomputed at the points from level m-1
        (st0 or the level-2 state ev) staged in LDS with its low-side halo (the
        step's exact instruction list, as the forward kernels evaluate it), an
        even level's is read from that frame (or from HBM when no odd level
        above it staged it)."""
        nf = len(self.fields)
        np_ = max(1, len(self.params))
        ax, ay = self.adj_tile
        th = self.adj_threads
        fl = "\n".join("  __shared__ real_t F%d[LH * LW];" % i for i in range(nf))
        inpl = self.unit_inplace
        ll = "\n".join(("  __shared__ real_t L%d[LH * LW];" if inpl else "  __shared__ real_t L%d[2][LH * LW];") % i
                       for i in range(nf))
        lbuf = (lambda i: "L%d" % i) if inpl else (lambda i: "L%d[sb]" % i)
        stage = tile_load("FR * LW", ["const int r = idx / LW, c = idx - r * LW;",
                                      "const long long g = (long long)wrapn(x0 - 2 * SLEEVE + r, N1) * N2 + "
                                      "wrapn(y0 - 2 * SLEEVE + c, N2);"],
                          [("F%d[idx]" % i, "FLD%d(xs)[g]" % i) for i in range(nf)], "real_t", "fs",
                          threads=th, indent="        ", max_count="LH * LW")
        return TILE_ADJ_UNIT % dict(
            TX=ax, TY=ay, ATH=th, fl=fl, ll=ll, stage=stage,
            lload=tile_load("LH * LW", ["const int r = idx / LW, c = idx - r * LW;",
                                        "const long long g = (long long)wrapn(x0 + r, N1) * N2 + wrapn(y0 + c, N2);"],
                            [(("L%d[idx]" if inpl else "L%d[0][idx]") % i, "LAM%d(lam_next)[g]" % i)
                             for i in range(nf)], "real_t", "ul", threads=th),
            inplace="true" if inpl else "false",
            rdecl=(" ".join("real_t res%d[NIT];" % i for i in range(nf)) if inpl else
                   "const int sb = mm & 1, db = sb ^ 1;"),
            wback=" ".join("L%d[lo] = res%d[it];" % (i, i) for i in range(nf)) if inpl else "",
            srcs=", ".join("F%d" % i for i in range(nf)),
            xdecl=" ".join("real_t sv%d;" % i for i in range(nf)),
            xstep=", ".join("sv%d" % i for i in range(nf)),
            xframe=" ".join("sv%d = F%d[fo];" % (i, i) for i in range(nf)),
            xglob=" ".join("sv%d = FLD%d(xs)[g];" % (i, i) for i in range(nf)),
            xargs=", ".join("sv%d" % i for i in range(nf)), largs=", ".join(lbuf(i) for i in range(nf)),
            adecl=" ".join("real_t a%d;" % i for i in range(nf)) + " " + " ".join("double q%d;" % p for p in range(np_)),
            aargs=", ".join(["a%d" % i for i in range(nf)] + ["q%d" % p for p in range(np_)]),
            st_lds=(" ".join("res%d[it] = a%d;" % (i, i) for i in range(nf)) if inpl else
                    " ".join("L%d[db][lo] = a%d;" % (i, i) for i in range(nf))),
            st_glb=" ".join("LAM%d(lam_prev)[go] = a%d;" % (i, i) for i in range(nf)),
            pdecl=" ".join("double pgs%d = 0;" % p for p in range(np_)),
            pacc=" ".join("pgs%d += q%d;" % (p, p) for p in range(np_)),
            pgs=", ".join("pgs%d" % p for p in range(np_)))

    def _emit_adj_pointwise(self, uses_state):
        """k_adjoint_levels<K> for a body that reads the state at the point only:
        lam in LDS (updated in place per level), each work-item's points fixed
        for the launch, the state of level m at those points in registers,
        level m-1's fetched while level m computes."""
        R = self.real
        nf = len(self.fields)
        np_ = max(1, len(self.params))
        ax, ay = self.adj_tile
        pb, _ = self.point_body(lambda f, a: "x%d" % f, lambda f, a: "L%d[lo + (%d)*LW + (%d)]" % (f, a[0], a[1]))
        xs = ", ".join("const real_t x%d" % i for i in range(nf))
        ls = ", ".join("const real_t* __restrict__ L%d" % i for i in range(nf))
        aouts = ", ".join(["real_t& oacc%d" % i for i in range(nf)] + ["double& opg%d" % p for p in range(np_)])
        fin = ["oacc%d = acc%d;" % (i, i) for i in range(nf)] + ["opg%d = pg%d;" % (p, p) for p in range(np_)]
        out = ["template <int LW>\n__device__ __forceinline__ void adj_point(%s, %s, const int lo, "
               "const double* __restrict__ P, %s) {\n"
               "#pragma clang fp contract(fast)\n%s\n}\n"
               % (xs, ls, aouts, "\n".join("  " + l for l in pb + fin))]
        ll = "\n".join("  __shared__ real_t L%d[LH * LW];" % i for i in range(nf))
        out.append(TILE_ADJ_PW % dict(
            TX=ax, TY=ay, ATH=self.adj_threads, ll=ll, uses_state="true" if uses_state else "false",
            lload=tile_load("LH * LW", ["const int r = idx / LW, c = idx - r * LW;",
                                        "const long long g = (long long)wrapn(x0 + r, N1) * N2 + wrapn(y0 + c, N2);"],
                            [("L%d[idx]" % i, "LAM%d(lam_next)[g]" % i) for i in range(nf)], "real_t", "ll",
                            threads=self.adj_threads),
            prefetch="true" if self.adj_prefetch else "false", idxarr="true" if self.adj_index_arrays else "false",
            xdecl=" ".join("real_t xv%d[NL_IT], nx%d[PREFETCH ? NL_IT : 1];" % (i, i) for i in range(nf)),
            xfetch_nx="\n".join("          nx%d[it] = FLD%d(xs)[g];" % (i, i) for i in range(nf)),
            xfetch_xv="\n".join("          xv%d[it] = FLD%d(xs)[g];" % (i, i) for i in range(nf)),
            xswap=" ".join("xv%d[it] = nx%d[it];" % (i, i) for i in range(nf)),
            xargs=", ".join("xv%d[it]" % i for i in range(nf)), largs=", ".join("L%d" % i for i in range(nf)),
            rdecl=" ".join("real_t res%d[NL_IT];" % i for i in range(nf)),
            rset=" ".join("res%d[it] = a%d;" % (i, i) for i in range(nf)),
            adecl=" ".join("real_t a%d;" % i for i in range(nf)) + " " + " ".join("double q%d;" % p for p in range(np_)),
            aargs=", ".join(["a%d" % i for i in range(nf)] + ["q%d" % p for p in range(np_)]),
            st_lds=" ".join("L%d[lo] = res%d[it];" % (i, i) for i in range(nf)),
            st_glb=" ".join("LAM%d(lam_prev)[go] = a%d;" % (i, i) for i in range(nf)),
            pdecl=" ".join("double pgs%d = 0;" % p for p in range(np_)),
            pacc=" ".join("pgs%d += q%d;" % (p, p) for p in range(np_)),
            pgs=", ".join("pgs%d" % p for p in range(np_))))
        return "\n".join(out)

    def _emit_adj_staged(self, uses_state):
        """k_adjoint_levels<K> with each level's state tile (and halo) staged in LDS."""
        R = self.real
        nf = len(self.fields)
        np_ = max(1, len(self.params))
        ax, ay = self.adj_tile
        src = []
        # adjoint point function on LDS tiles
        pb, _ = self.point_body(lambda f, a: "X%d[xo + (%d)*XW + (%d)]" % (f, a[0], a[1]),
                                lambda f, a: "L%d[lo + (%d)*LW + (%d)]" % (f, a[0], a[1]), pguard="own")
        xs = ", ".join("const real_t* __restrict__ X%d" % i for i in range(nf))
        ls = ", ".join("const real_t* __restrict__ L%d" % i for i in range(nf))
        aouts = ", ".join(["real_t& oacc%d" % i for i in range(nf)] + ["double& opg%d" % p for p in range(np_)])
        fin = ["oacc%d = acc%d;" % (i, i) for i in range(nf)] + ["opg%d = pg%d;" % (p, p) for p in range(np_)]
        # derivative code only: the compiler may reassociate and contract (hoisting
        # the products of uniform coefficients); the step's own values stay exact
        src.append("template <int XW, int LW>\n__device__ __forceinline__ void adj_point(%s, %s, const int xo, const int lo, "
                   "const bool own, const double* __restrict__ P, %s) {\n"
                   "#pragma clang fp reassociate(on)\n#pragma clang fp contract(fast)\n%s\n}\n"
                   % (xs, ls, aouts, "\n".join("  " + l for l in pb + fin)))
        xl = "\n".join("  __shared__ real_t X%d[XH * XW];" % i for i in range(nf))
        ll = "\n".join("  __shared__ real_t L%d[LH * LW];" % i for i in range(nf))
        src.append(TILE_ADJ % dict(
            TX=ax, TY=ay, ATH=self.adj_threads, xl=xl, ll=ll, uses_state="true" if uses_state else "false",
            lload=tile_load("LH * LW", ["const int r = idx / LW, c = idx - r * LW;",
                                        "const long long g = (long long)wrapn(x0 + r, N1) * N2 + wrapn(y0 + c, N2);"],
                            [("L%d[idx]" % i, "LAM%d(lam_next)[g]" % i) for i in range(nf)], "real_t", "ll",
                            threads=self.adj_threads),
            xdecl=" ".join("real_t nx%d[NX_IT];" % i for i in range(nf)),
            xfetch="\n".join("          nx%d[it] = FLD%d(xs)[g];" % (i, i) for i in range(nf)),
            xput="\n".join("        X%d[idx] = nx%d[it];" % (i, i) for i in range(nf)),
            xargs=", ".join("X%d" % i for i in range(nf)), largs=", ".join("L%d" % i for i in range(nf)),
            rdecl=" ".join("real_t res%d[NL_IT * RB];" % i for i in range(nf)),
            rset=" ".join("res%d[it * RB + q] = a%d;" % (i, i) for i in range(nf)),
            rb=self.row_block,
            adecl=" ".join("real_t a%d;" % i for i in range(nf)) + " " + " ".join("double q%d;" % p for p in range(np_)),
            aargs=", ".join(["a%d" % i for i in range(nf)] + ["q%d" % p for p in range(np_)]),
            st_lds=" ".join("L%d[lo] = res%d[it * RB + q];" % (i, i) for i in range(nf)),
            st_glb=" ".join("LAM%d(lam_prev)[g] = a%d;" % (i, i) for i in range(nf)),
            pdecl=" ".join("double pgs%d = 0;" % p for p in range(np_)),
            pacc=" ".join("pgs%d += q%d;" % (p, p) for p in range(np_)),
            pgs=", ".join("pgs%d" % p for p in range(np_))))
        return "\n".join(src)


def _device_part(text: str, kernel: str, name: str, lds_lines: List[str], pool_lines: List[str]) -> str:
    """A generated kernel as a device function ``name`` over an LDS pool: its
    signature gains ``const int bid`` (the workgroup's tile number) and
    ``real_t* pool``; its LDS arrays become pointers into the pool."""
    head = text.index("__global__ void __launch_bounds__(")
    open_brace = text.index(") {", head)
    params = text[text.index(kernel + "(", head) + len(kernel) + 1:open_brace]
    text = (text[:head] + "__device__ __forceinline__ void %s(const int bid, real_t* __restrict__ pool, %s) {"
            % (name, " ".join(params.split())) + text[open_brace + 3:])
    for old, new in zip(lds_lines, pool_lines):
        assert text.count(old + "\n") == 1, old
        text = text.replace(old + "\n", new + "\n")
    text = text.replace("xcd_tile(blockIdx.x, NT)", "xcd_tile(bid, NT)")
    text = text.replace("(long long)blockIdx.x * NP", "(long long)bid * NP")
    if "blockIdx" in text or "__shared__ real_t" in text:
        raise EmitError("%s: workgroup index or LDS array left in the device part" % name)
    return text


FWD_JOB = r"""
// A forward job riding in an adjoint launch: k (2 or 4) steps from `in` into `out`.
struct FwdJob {
  const char* in;
  char* out;
  int k;
};
"""

UNIT_NEXT = r"""
// The adjoint of K steps (workgroups [0, NTA), k_adjoint_unit's tiles) and nj
// (0..2) forward jobs (NT workgroups each, the forward tiles) in one launch: the
// forward tiles take the slots the adjoint's last round of workgroups leaves
// free.  The executor only pairs work that touches disjoint buffers.
template <int K>
__global__ void __launch_bounds__(%(ATH)d, %(MINW)d) k_adjoint_unit_fwd(const char* __restrict__ st0, const char* __restrict__ ev,
                                                          const char* __restrict__ lam_next, char* __restrict__ lam_prev,
                                                          const double* __restrict__ P, double* __restrict__ pgrad,
                                                          const int nj, const FwdJob j0, const FwdJob j1) {
  __shared__ real_t pool[%(POOL)d];
  // role 0: adjoint tile t; role q + 1: tile t of job q.  Every adjoint tile
  // comes first, so the forward tiles fill the adjoint's last round (measured:
  // interleaving the roles within each XCD made the reversal 13 %% slower)
  const int b = (int)blockIdx.x;
  const int role = b < NTA ? 0 : 1 + (b - NTA) / NT, t = b < NTA ? b : (b - NTA) %% NT;
  if (role > nj) return;
  if (role == 0) {
    adj_part<K>(t, pool, st0, ev, lam_next, lam_prev, P, pgrad);
  } else {
    const FwdJob& J = role == 1 ? j0 : j1;
    if (J.k == 2)
      fwd_part<2, false>(t, pool, J.in, J.out, P);
    else
      fwd_part<4, false>(t, pool, J.in, J.out, P);
  }
}
"""

TILE_FWD = r"""
// XCD-aware tile order: workgroup ids go round-robin to the 8 XCDs, so each XCD
// gets a contiguous band of tiles (shared halos stay in its L2).
__device__ __forceinline__ int xcd_tile(int b, int nt) { return (nt %% 8 == 0) ? ((b %% 8) * (nt / 8) + b / 8) : b; }
__device__ __forceinline__ int wrapn(int i, int n) { return i < 0 ? i + n : (i >= n ? i - n : i); }

template <int K, bool ALL>
__global__ void __launch_bounds__(%(FTH)d) k_forward_tile(const char* __restrict__ st_in, char* __restrict__ st_out,
                                                      const double* __restrict__ P) {
  constexpr int TX = %(TX)d, TY = %(TY)d, H = 2 * K * SLEEVE, RX = TX + H, RY = TY + H;
  constexpr int NTY = N2 / TY, NT = (N1 / TX) * NTY;
%(lds)s
  const int tile = xcd_tile(blockIdx.x, NT);
  const int x0 = (tile / NTY) * TX, y0 = (tile %% NTY) * TY;
  // run-time parameters in registers for the whole launch (a scalar load per
  // point would also drain the LDS reads in flight: both count in lgkmcnt)
  double Pl[NP];
#pragma unroll
  for (int p = 0; p < NP; ++p) Pl[p] = HAS_PARAMS ? P[p] : 0.0;
%(load_tile)s
  __syncthreads();
#pragma unroll
  for (int j = 1; j <= K; ++j) {
    const int sb = (j - 1) & 1, db = j & 1;
    const int lo = 2 * j * SLEEVE, W = RY - lo, HH = RX - lo;
    char* __restrict__ dst = ALL ? st_out + (size_t)(j - 1) * STATE_BYTES : st_out;
    for (int idx = threadIdx.x; idx < HH * W; idx += %(FTH)d) {
      const int rr = idx / W;
      const int r = lo + rr, c = lo + (idx - rr * W);
      const int o = r * RY + c;
      %(decl)s
      pt_step<RY>(%(srcs)s, o, Pl, %(refs)s);
      if (j < K) {
        %(st_lds)s
        if (ALL && r >= H && c >= H) {
          char* __restrict__ lvl = dst;
          const long long g = (long long)(x0 + r - H) * N2 + (y0 + c - H);
          %(st_lvl)s
        }
      } else {
        const long long g = (long long)(x0 + r - H) * N2 + (y0 + c - H);
        %(st_glb)s
      }
    }
    if (j < K) __syncthreads();
  }
}
"""

TILE_ADJ = r"""
// K adjoint steps: lam_K (HBM) -> lam_0 (HBM); X_0 = st0, X_m = level buffer m-1 of lv
template <int K>
__global__ void __launch_bounds__(%(ATH)d) k_adjoint_levels(const char* __restrict__ st0, const char* __restrict__ lv,
                                                        const char* __restrict__ lam_next, char* __restrict__ lam_prev,
                                                        const double* __restrict__ P, double* __restrict__ pgrad) {
  constexpr int TX = %(TX)d, TY = %(TY)d;
  constexpr int XH = TX + 2 * (K + 1) * SLEEVE, XW = TY + 2 * (K + 1) * SLEEVE;
  constexpr int LH = TX + 2 * K * SLEEVE, LW = TY + 2 * K * SLEEVE;
  constexpr int NTY = N2 / TY, NT = (N1 / TX) * NTY;
  constexpr bool USES_STATE = %(uses_state)s;
%(xl)s
%(ll)s
  const int tile = xcd_tile(blockIdx.x, NT);
  const int x0 = (tile / NTY) * TX, y0 = (tile %% NTY) * TY;
  // run-time parameters in registers for the whole launch (a scalar load per
  // point would also drain the LDS reads in flight: both count in lgkmcnt)
  double Pl[NP];
#pragma unroll
  for (int p = 0; p < NP; ++p) Pl[p] = HAS_PARAMS ? P[p] : 0.0;
  // the state tile of the next level is fetched into registers (nx*) while the
  // current level computes; it enters LDS once every read of the current one ended.
  // Element it of this thread is (r, c) = divmod(idx, XW) of the largest region;
  // its periodic global offset xg and the lowest level that reads it (level m
  // reads rows < TX + 2(m+2)s and columns < TY + 2(m+2)s) are computed once.
  constexpr int NX_IT = (XH * XW + %(ATH)d - 1) / %(ATH)d;
  static_assert((long long)N1 * N2 < (1LL << 31), "element offsets of a field are ints");
  %(xdecl)s
  int xg[NX_IT], xm[NX_IT];
#pragma unroll
  for (int it = 0; it < NX_IT; ++it) {
    const int idx = threadIdx.x + it * %(ATH)d;
    const int r = idx / XW, c = idx - r * XW;
    xg[it] = wrapn(x0 - 2 * SLEEVE + r, N1) * N2 + wrapn(y0 - 2 * SLEEVE + c, N2);
    const int er = r - TX, ec = c - TY, e = er > ec ? er : ec;  // beyond the tile by e cells
    xm[it] = idx < XH * XW ? (e < 4 * SLEEVE ? 0 : (e - 4 * SLEEVE) / (2 * SLEEVE) + 1) : K;
  }
  auto xfetch = [&](int m) {
    const char* __restrict__ xs = m == 0 ? st0 : lv + (size_t)(m - 1) * STATE_BYTES;
#pragma unroll
    for (int it = 0; it < NX_IT; ++it) {
      if (xm[it] <= m) {
        const int g = xg[it];
%(xfetch)s
      }
    }
  };
  auto xput = [&](int m) {
#pragma unroll
    for (int it = 0; it < NX_IT; ++it) {
      const int idx = threadIdx.x + it * %(ATH)d;
      if (xm[it] <= m) {
%(xput)s
      }
    }
  };
  if (USES_STATE) xfetch(K - 1);
%(lload)s
  if (USES_STATE) {
    xput(K - 1);
    if (K > 1) xfetch(K - 2);
  }
  __syncthreads();
  %(pdecl)s
  // lam is updated in place: a level's values are computed into registers from
  // lam_{m+1} in LDS, and written over it after every thread has read it
  // each work-item evaluates RB vertically adjacent points of a column, so the
  // LDS reads that their stencils share are issued once
  constexpr int RB = %(rb)d;
  static_assert(TX %% RB == 0 && (2 * SLEEVE) %% RB == 0, "row blocks tile every level");
  constexpr int NL_IT = ((LH / RB) * LW + %(ATH)d - 1) / %(ATH)d;
#pragma unroll
  for (int mm = 0; mm < K; ++mm) {
    const int m = K - 1 - mm;  // level: reverse step m, lam_{m+1} -> lam_m at state X_m
    const int HH = TX + 2 * m * SLEEVE, W = TY + 2 * m * SLEEVE;
    %(rdecl)s
#pragma unroll
    for (int it = 0; it < NL_IT; ++it) {
      // whole rows of the adjoint buffer, columns beyond the level masked: a
      // wavefront's LDS accesses are contiguous (no bank conflicts at row wraps)
      const int idx = threadIdx.x + it * %(ATH)d;
      const int rq = idx / LW, c = idx - rq * LW;
      if (idx < (HH / RB) * LW && c < W) {
#pragma unroll
        for (int q = 0; q < RB; ++q) {
          const int r = rq * RB + q;
          const bool own = r < TX && c < TY;
          %(adecl)s
          adj_point<XW, LW>(%(xargs)s, %(largs)s, (r + 2 * SLEEVE) * XW + (c + 2 * SLEEVE), r * LW + c, own, Pl, %(aargs)s);
          if (m > 0) {
            %(rset)s
          } else {
            const long long g = (long long)(x0 + r) * N2 + (y0 + c);
            %(st_glb)s
          }
          if (own) { %(pacc)s }
        }
      }
    }
    if (m > 0) {
      __syncthreads();  // every read of lam_{m+1} and X_m is done
#pragma unroll
      for (int it = 0; it < NL_IT; ++it) {
        const int idx = threadIdx.x + it * %(ATH)d;
        const int rq = idx / LW, c = idx - rq * LW;
        if (idx < (HH / RB) * LW && c < W) {
#pragma unroll
          for (int q = 0; q < RB; ++q) {
            const int lo = (rq * RB + q) * LW + c;
            %(st_lds)s
          }
        }
      }
      if (USES_STATE) {
        xput(m - 1);
        if (m > 1) xfetch(m - 2);
      }
      __syncthreads();
    }
  }
  { const double pv[NP] = {%(pgs)s}; block_reduce_params<%(ATH)d>(pgrad + (long long)blockIdx.x * NP, pv); }
}
"""

TILE_ADJ_PW = r"""
// K adjoint steps: lam_K (HBM) -> lam_0 (HBM); X_0 = st0, X_m = level buffer m-1 of lv.
// The adjoint body reads the state at the point only: work-item points are fixed
// for the launch (element it of this thread is (r, c) = divmod(idx, LW) of the
// lam buffer) and the state of level m at those points is in registers (xv*);
// with PREFETCH level m-1's is fetched (nx*) while level m computes.  With
// IDXARR the elements' rows, columns and global offsets stay in registers,
// else they are recomputed where used (fewer registers, more waves per CU).
template <int K>
__global__ void __launch_bounds__(%(ATH)d) k_adjoint_levels(const char* __restrict__ st0, const char* __restrict__ lv,
                                                        const char* __restrict__ lam_next, char* __restrict__ lam_prev,
                                                        const double* __restrict__ P, double* __restrict__ pgrad) {
  constexpr int TX = %(TX)d, TY = %(TY)d;
  constexpr int LH = TX + 2 * K * SLEEVE, LW = TY + 2 * K * SLEEVE;
  constexpr int NTY = N2 / TY, NT = (N1 / TX) * NTY;
  constexpr bool USES_STATE = %(uses_state)s, PREFETCH = %(prefetch)s, IDXARR = %(idxarr)s;
  static_assert((long long)N1 * N2 < (1LL << 31), "element offsets of a field are ints");
%(ll)s
  const int tile = xcd_tile(blockIdx.x, NT);
  const int x0 = (tile / NTY) * TX, y0 = (tile %% NTY) * TY;
  double Pl[NP];
#pragma unroll
  for (int p = 0; p < NP; ++p) Pl[p] = HAS_PARAMS ? P[p] : 0.0;
  constexpr int NL_IT = (LH * LW + %(ATH)d - 1) / %(ATH)d;
  // element it: row (LH: none) and column of the lam buffer, periodic global offset
  int rq[IDXARR ? NL_IT : 1], cq[IDXARR ? NL_IT : 1], gq[IDXARR ? NL_IT : 1];
  auto element = [&](int it, int& r, int& c, int& g) {
    if constexpr (IDXARR) {
      r = rq[it]; c = cq[it]; g = gq[it];
    } else {
      const int idx = threadIdx.x + it * %(ATH)d;
      const int rr = idx / LW;
      c = idx - rr * LW;
      r = idx < LH * LW ? rr : LH;
      g = wrapn(x0 + (idx < LH * LW ? rr : 0), N1) * N2 + wrapn(y0 + c, N2);
    }
  };
  if constexpr (IDXARR) {
#pragma unroll
    for (int it = 0; it < NL_IT; ++it) {
      const int idx = threadIdx.x + it * %(ATH)d;
      const int r = idx / LW, c = idx - r * LW;
      rq[it] = idx < LH * LW ? r : LH;
      cq[it] = c;
      gq[it] = wrapn(x0 + (idx < LH * LW ? r : 0), N1) * N2 + wrapn(y0 + c, N2);
    }
  }
  %(xdecl)s
  // the state of level m at the elements level m computes (rows < TX + 2ms,
  // columns < TY + 2ms), into nx (PREFETCH) or xv
  auto xfetch = [&](int m) {
    const char* __restrict__ xs = m == 0 ? st0 : lv + (size_t)(m - 1) * STATE_BYTES;
#pragma unroll
    for (int it = 0; it < NL_IT; ++it) {
      int r, c, g;
      element(it, r, c, g);
      if (r < TX + 2 * m * SLEEVE && c < TY + 2 * m * SLEEVE) {
        if constexpr (PREFETCH) {
%(xfetch_nx)s
        } else {
%(xfetch_xv)s
        }
      }
    }
  };
  if (USES_STATE && PREFETCH) xfetch(K - 1);
%(lload)s
  __syncthreads();
  %(pdecl)s
#pragma unroll
  for (int mm = 0; mm < K; ++mm) {
    const int m = K - 1 - mm;  // level: reverse step m, lam_{m+1} -> lam_m at state X_m
    const int HH = TX + 2 * m * SLEEVE, W = TY + 2 * m * SLEEVE;
    if (USES_STATE) {
      if constexpr (PREFETCH) {
#pragma unroll
        for (int it = 0; it < NL_IT; ++it) { %(xswap)s }
        if (m > 0) xfetch(m - 1);
      } else {
        xfetch(m);
      }
    }
    %(rdecl)s
#pragma unroll
    for (int it = 0; it < NL_IT; ++it) {
      int r, c, g;
      element(it, r, c, g);
      if (r < HH && c < W) {
        const int lo = r * LW + c;
        %(adecl)s
        adj_point<LW>(%(xargs)s, %(largs)s, lo, Pl, %(aargs)s);
        if (m > 0) {
          %(rset)s
        } else {
          const long long go = (long long)(x0 + r) * N2 + (y0 + c);
          %(st_glb)s
        }
        if (r < TX && c < TY) { %(pacc)s }
      }
    }
    if (m > 0) {
      __syncthreads();  // every read of lam_{m+1} is done
#pragma unroll
      for (int it = 0; it < NL_IT; ++it) {
        int r, c, g;
        element(it, r, c, g);
        if (r < HH && c < W) { const int lo = r * LW + c; %(st_lds)s }
      }
      __syncthreads();
    }
  }
  { const double pv[NP] = {%(pgs)s}; block_reduce_params<%(ATH)d>(pgrad + (long long)blockIdx.x * NP, pv); }
}
"""

TILE_ADJ_UNIT = r"""
// K (<= 4) adjoint steps from the state at their start (st0) and the state two
// steps later (ev, for K >= 3): lam_K (HBM) -> lam_0 (HBM).  Levels come in
// pairs (odd m, even m-1): level m-1's state is staged in the frame F (tile rows
// [-2s, T + 2ms), the low-side halo the step reads), level m's state is
// recomputed at each point of the adjoint region from F with the step's exact
// instruction list, and level m-1's is read from F.  The adjoint body reads the
// state at the point only.  lam either ping-pongs between two LDS buffers (one
// barrier per level) or is updated in place (LAM_INPLACE: a level's results are
// held in registers until every work-item has read lam_{m+1}, then written
// back; one LDS buffer, so more workgroups fit on a CU); level m's work-items
// cover exactly its region [0, TX + 2ms) x [0, TY + 2ms).
template <int K>
__global__ void __launch_bounds__(%(ATH)d) k_adjoint_unit(const char* __restrict__ st0, const char* __restrict__ ev,
                                                      const char* __restrict__ lam_next, char* __restrict__ lam_prev,
                                                      const double* __restrict__ P, double* __restrict__ pgrad) {
  constexpr int TX = %(TX)d, TY = %(TY)d;
  constexpr int LH = TX + 2 * K * SLEEVE, LW = TY + 2 * K * SLEEVE;
  constexpr int NTY = N2 / TY, NT = (N1 / TX) * NTY;
  static_assert(K <= 4, "levels 0 and 2 are given, 1 and 3 recomputed");
  static_assert((long long)N1 * N2 < (1LL << 31), "element offsets of a field are ints");
  constexpr bool LAM_INPLACE = %(inplace)s;
  // most points a work-item holds in one level (the largest region, m = K - 1)
  constexpr int NIT = ((TX + 2 * (K - 1) * SLEEVE) * (TY + 2 * (K - 1) * SLEEVE) + %(ATH)d - 1) / %(ATH)d;
%(fl)s
%(ll)s
  const int tile = xcd_tile(blockIdx.x, NT);
  const int x0 = (tile / NTY) * TX, y0 = (tile %% NTY) * TY;
  double Pl[NP];
#pragma unroll
  for (int p = 0; p < NP; ++p) Pl[p] = HAS_PARAMS ? P[p] : 0.0;
%(lload)s
  %(pdecl)s
#pragma unroll
  for (int mm = 0; mm < K; ++mm) {
    const int m = K - 1 - mm;  // level: reverse step m, lam_{m+1} -> lam_m (ping-pong: buffer mm & 1)
    const int HH = TX + 2 * m * SLEEVE, W = TY + 2 * m * SLEEVE;
    // state m: recomputed (odd m) from the frame, read from the frame (even m
    // below an odd level), or read from HBM (even m = K-1)
    const bool recompute = (m & 1) != 0, framed = !recompute && m + 1 < K;
    if (recompute) {
      // the level above has read the frame (in place: its write-back barrier)
      if (!LAM_INPLACE && mm > 0) __syncthreads();
      const char* __restrict__ xs = m == 1 ? st0 : ev;  // level m-1
      const int FR = TX + 2 * (m + 1) * SLEEVE;  // frame rows staged: [-2s, T + 2ms) of the tile
%(stage)s
    }
    __syncthreads();  // lam_{m+1} and the frame are in LDS
    %(rdecl)s
#pragma unroll
    for (int it = 0; it < (HH * W + %(ATH)d - 1) / %(ATH)d; ++it) {
      const int idx = threadIdx.x + it * %(ATH)d;
      if (idx < HH * W) {
        const int r = idx / W, c = idx - r * W;
        const int lo = r * LW + c, fo = (r + 2 * SLEEVE) * LW + (c + 2 * SLEEVE);
        %(xdecl)s
        if (recompute) {
          pt_step<LW>(%(srcs)s, fo, Pl, %(xstep)s);
        } else if (framed) {
          %(xframe)s
        } else {
          const char* __restrict__ xs = m == 0 ? st0 : ev;
          const int g = wrapn(x0 + r, N1) * N2 + wrapn(y0 + c, N2);
          %(xglob)s
        }
        %(adecl)s
        adj_point<LW>(%(xargs)s, %(largs)s, lo, Pl, %(aargs)s);
        if (m > 0) {
          %(st_lds)s
        } else {
          const long long go = (long long)(x0 + r) * N2 + (y0 + c);
          %(st_glb)s
        }
        if (r < TX && c < TY) { %(pacc)s }
      }
    }
    if (LAM_INPLACE && m > 0) {
      __syncthreads();  // every read of lam_{m+1} and of the frame is done
#pragma unroll
      for (int it = 0; it < (HH * W + %(ATH)d - 1) / %(ATH)d; ++it) {
        const int idx = threadIdx.x + it * %(ATH)d;
        if (idx < HH * W) {
          const int r = idx / W, c = idx - r * W;
          const int lo = r * LW + c;
          %(wback)s
        }
      }
    }
  }
  { const double pv[NP] = {%(pgs)s}; block_reduce_params<%(ATH)d>(pgrad + (long long)blockIdx.x * NP, pv); }
}
"""

UNIT_STUB = r"""
// no fused unit in this library (FUSED_UNIT is false): never called
static void fwd_even(const char*, char*, const double*, hipStream_t) { fmr_rt::fatal("fwd_even: no fused unit"); }
static void adj_unit(int, const char*, const char*, const char*, char*, const double*, double*, hipStream_t) {
  fmr_rt::fatal("adj_unit: no fused unit");
}
struct FwdJob {
  const char* in;
  char* out;
  int k;
};
constexpr bool HAS_UNIT_FWD = false;
constexpr bool LONG_JOBS_FIRST = false;
static void adj_unit_fwd(int, const char*, const char*, const char*, char*, const double*, double*, int, FwdJob, FwdJob,
                         hipStream_t) {
  fmr_rt::fatal("adj_unit_fwd: no fused unit");
}
"""

UNIT_HOST = r"""
// the state two steps after `in` into `out` (one launch)
static void fwd_even(const char* in, char* out, const double* P, hipStream_t s) {
  fwd_tiles(in, [&](int) { return out; }, 2, P, s);
}

// the adjoint of k (1..KA <= 4) steps from st0 in one launch (ev: the state two steps later, for k >= 3)
static void adj_unit(int k, const char* st0, const char* ev, const char* lam_next, char* lam_prev, const double* P,
                     double* pgrad, hipStream_t s) {
  switch (k) {
%(ucases)s    default: fmr_rt::fatal("bad adjoint step count");
  }
}

// adj_unit and, in the same launch, nj (0..2) forward jobs (k_adjoint_unit_fwd)
constexpr bool HAS_UNIT_FWD = %(has_next)s;
constexpr bool LONG_JOBS_FIRST = %(long_first)s;
static void adj_unit_fwd(int k, const char* st0, const char* ev, const char* lam_next, char* lam_prev,
                         const double* P, double* pgrad, int nj, FwdJob j0, FwdJob j1, hipStream_t s) {
  if (nj < 0 || nj > 2) fmr_rt::fatal("adj_unit_fwd: bad job count");
  for (int q = 0; q < nj; ++q) {
    const FwdJob& J = q == 0 ? j0 : j1;
    if (J.k != 2 && J.k != 4) fmr_rt::fatal("adj_unit_fwd: a forward job is 2 or 4 steps");
  }
  switch (k) {
%(ncases)s    default: fmr_rt::fatal("bad adjoint step count");
  }
}
"""

TILE_HOST = r"""
constexpr int NT = (N1 / %(TX)d) * (N2 / %(TY)d);
constexpr int NTA = (N1 / %(AX)d) * (N2 / %(AY)d);

// nsteps forward steps, K = %(K)d per launch, ping-pong so that the last lands in dst_of(0)
template <class DstOf>
static void fwd_tiles(const char* src, DstOf dst_of, int nsteps, const double* P, hipStream_t s) {
  int launches = (nsteps + %(K)d - 1) / %(K)d, left = nsteps;
  for (int l = 0; l < launches; ++l) {
    const int k = left >= %(K)d ? %(K)d : left;
    char* out = dst_of(launches - 1 - l);
    const char* in = src;
    switch (k) {
%(fcases)s      default: fmr_rt::fatal("bad step count");
    }
    src = out;
    left -= k;
  }
}

// j forward steps from `in`, level i (1..j) written to state buffer i-1 of `out`
static void fwd_levels(int j, const char* in, char* out, const double* P, hipStream_t s) {
  switch (j) {
%(lcases)s    default: fmr_rt::fatal("bad level count");
  }
}

static void adj_levels(int k, const char* st0, const char* lv, const char* lam_next, char* lam_prev, const double* P,
                       double* pgrad, hipStream_t s) {
  switch (k) {
%(acases)s    default: fmr_rt::fatal("bad adjoint step count");
  }
}
"""


ADJ_TEMPLATE = r'''// Generated by formura_amd: forward and adjoint steps of %(name)s. Do not edit.
#include "formura_hip_rt.h"
#include "formura_adjoint.h"
#include <vector>

namespace {
typedef %(R)s real_t;
%(consts)s
constexpr long long NTOT = %(ntot)s;
constexpr int NF = %(nf)d;
constexpr int NP = %(np)d;
constexpr bool HAS_PARAMS = %(has_params)s;
constexpr int BLOCK = %(block)d;
constexpr int SLEEVE = %(sleeve)d;
constexpr int KA = %(levels)d;  // most adjoint steps per Formura_AD_Adjoint_Steps call
constexpr size_t STATE_BYTES = (size_t)NF * NTOT * sizeof(%(R)s);
// a unit of adjoint steps is one kernel that recomputes its inner states (no scratch used)
constexpr bool FUSED_UNIT = %(fused)s;
%(flds)s
%(lams)s
const char* const kDescription = R"FMRJSON(%(desc)s)FMRJSON";

// out[p] += the sum of v[p] over the workgroup: wave sums by shuffles, then the
// waves' sums in wave order (a fixed order: deterministic, no atomics)
template <int BS>
__device__ __forceinline__ void block_reduce_params(double* __restrict__ out, const double* v) {
  __shared__ double red[NP][BS / 64];
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
#pragma unroll
  for (int p = 0; p < NP; ++p) {
    double x = v[p];
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) x += __shfl_down(x, off, 64);
    if (lane == 0) red[p][w] = x;
  }
  __syncthreads();
  if ((int)threadIdx.x < NP) {
    double t = 0;
    for (int q = 0; q < BS / 64; ++q) t += red[threadIdx.x][q];
    out[threadIdx.x] += t;
  }
}

__global__ void __launch_bounds__(BLOCK) k_forward(const char* __restrict__ st, char* __restrict__ out,
                                                   const double* __restrict__ P) {
%(fwd)s
}

__global__ void __launch_bounds__(BLOCK) k_adjoint(const char* __restrict__ st, const char* __restrict__ lam_next,
                                                   char* __restrict__ lam_prev, const double* __restrict__ P,
                                                   double* __restrict__ pgrad) {
%(adj)s
}
%(tiled)s
%(unit_stub)s
}  // namespace

extern "C" const char* Formura_AD_Describe(void) { return kDescription; }
extern "C" long long Formura_AD_State_Bytes(void) { return (long long)NF * NTOT * sizeof(real_t); }
extern "C" int Formura_AD_Param_Slots(void) { return %(slots)s; }
extern "C" int Formura_AD_Levels(void) { return KA; }

// nsteps forward steps from `in`; ping-pong between `out` and `tmp` so that the
// result lands in `out` (in may alias neither).
extern "C" void Formura_AD_Forward(const void* in, void* out, void* tmp, int nsteps, const double* params_dev,
                                   void* stream) {
  hipStream_t s = (hipStream_t)stream;
  const unsigned nb = (unsigned)((NTOT + BLOCK - 1) / BLOCK);
  (void)nb;
  const char* src = (const char*)in;
  // destination of the launch that has `left` launches after it
  auto dst_of = [&](int left) { return (left %% 2 == 0) ? (char*)out : (char*)tmp; };
%(fwd_call)s
  if (nsteps == 0 && in != out)
    FMR_CHECK(hipMemcpyAsync(out, in, (size_t)NF * NTOT * sizeof(real_t), hipMemcpyDeviceToDevice, s));
  FMR_CHECK(hipGetLastError());
}

// one adjoint step: state_prev is the input of the step, lam_next = dL/d(output);
// writes lam_prev = dL/d(input) and adds dL/dparams into pgrad[slot][param].
extern "C" void Formura_AD_Adjoint(const void* state_prev, const void* lam_next, void* lam_prev,
                                   const double* params_dev, double* pgrad, void* stream) {
  hipStream_t s = (hipStream_t)stream;
  const unsigned nb = (unsigned)((NTOT + BLOCK - 1) / BLOCK);
  (void)nb;
%(adj_call)s
  FMR_CHECK(hipGetLastError());
}

// nsteps (1..Formura_AD_Levels()) adjoint steps in one call: state0 is the input
// of the first of them, lam_next = dL/d(output of the last); the inner states
// are recomputed into scratch (nsteps-1 state buffers).  Writes lam_prev =
// dL/d(state0) and adds dL/dparams of all nsteps steps into pgrad.
extern "C" void Formura_AD_Adjoint_Steps(const void* state0, void* scratch, const void* lam_next, void* lam_prev,
                                         int nsteps, const double* params_dev, double* pgrad, void* stream) {
  hipStream_t s = (hipStream_t)stream;
  const unsigned nb = (unsigned)((NTOT + BLOCK - 1) / BLOCK);
  (void)nb;
  if (nsteps < 1 || nsteps > KA) fmr_rt::fatal("Formura_AD_Adjoint_Steps: step count out of range");
%(steps_call)s
  FMR_CHECK(hipGetLastError());
}

// The two halves of Formura_AD_Adjoint_Steps, for callers that overlap the
// forward work of one unit with the adjoint work of another on two streams:
// the nlevels (0..Formura_AD_Levels()-1) steps after state0, level i into
// state buffer i of scratch ...
extern "C" void Formura_AD_Inner_States(const void* state0, void* scratch, int nlevels, const double* params_dev,
                                        void* stream) {
  hipStream_t s = (hipStream_t)stream;
  if (nlevels < 0 || nlevels > KA - 1) fmr_rt::fatal("Formura_AD_Inner_States: level count out of range");
%(inner_call)s
  FMR_CHECK(hipGetLastError());
}

// Executor of a reversal schedule (see include/formura_adjoint.h): forward ops on
// stream F, adjoint ops on stream A, ordered by events on the buffers they share.
// Events are re-recorded in place (a wait takes the event as last recorded).
namespace {
struct EventSet {
  std::vector<hipEvent_t> e;
  explicit EventSet(size_t n) : e(n) {
    for (auto& x : e) FMR_CHECK(hipEventCreateWithFlags(&x, hipEventDisableTiming));
  }
  ~EventSet() {
    for (auto x : e) (void)hipEventDestroy(x);
  }
};
}  // namespace

// ... and the nsteps adjoint steps over them (scratch as filled above).
extern "C" void Formura_AD_Adjoint_Levels(const void* state0, const void* scratch, const void* lam_next, void* lam_prev,
                                          int nsteps, const double* params_dev, double* pgrad, void* stream) {
  hipStream_t s = (hipStream_t)stream;
  const unsigned nb = (unsigned)((NTOT + BLOCK - 1) / BLOCK);
  (void)nb;
  if (nsteps < 1 || nsteps > KA) fmr_rt::fatal("Formura_AD_Adjoint_Levels: step count out of range");
%(levels_call)s
  FMR_CHECK(hipGetLastError());
}

// The reversal on one stream with a fused unit: the operations in program order,
// each unit's adjoint launch also carrying up to two forward jobs of later
// operations -- a unit's level-2 state, or a 2- or 4-step recomputation --
// that are ready and touch nothing pending.  A job at position j may run now if
// every operation before j that touches its output buffer, and every one that
// writes its input buffer, was issued in an earlier launch (so, in stream
// order, has finished); the adjoint it rides with is such an operation, so it
// never shares a buffer with the job.  Buffers: the states 0..nbufs-1 and the
// two scratch halves nbufs, nbufs+1 (a unit's level-2 state).  Returns as
// Formura_AD_Reverse.
static int reverse_fused(const int* ops, int nops, void* const* bufs, int nbufs, void* tmp, char* const* sc,
                         char* lam, char* lam2, const double* P, double* pgrad, hipStream_t s) {
  const int nb = nbufs + 2;
  std::vector<int> half(nops, 0);
  std::vector<char> adj_done(nops, 0), lvl_done(nops, 0);  // adjoint issued; level-2 state issued (or none needed)
  std::vector<std::vector<int>> touch(nb), writer(nb);
  int turn = 0;
  for (int i = 0; i < nops; ++i) {
    const int kind = ops[4 * i], a = ops[4 * i + 1], b = ops[4 * i + 2], k = ops[4 * i + 3];
    if (a < 0 || a >= nbufs || (kind == 0 && (b < 0 || b >= nbufs || b == a)) || k < 1 || (kind == 1 && k > KA))
      fmr_rt::fatal("Formura_AD_Reverse: bad operation");
    if (kind == 0) {
      touch[a].push_back(i);
      touch[b].push_back(i);
      writer[b].push_back(i);
      adj_done[i] = lvl_done[i] = 0;
    } else {
      half[i] = KA > 1 ? (turn++ & 1) : 0;
      touch[a].push_back(i);
      touch[nbufs + half[i]].push_back(i);
      lvl_done[i] = k <= 2;  // no level-2 state for a unit of one or two steps
    }
  }
  // an operation is issued when all its launches are: a forward one when its
  // launch is (adj_done doubles as its flag), a unit when its adjoint and its
  // level-2 state are
  auto issued = [&](int q) { return ops[4 * q] == 0 ? adj_done[q] != 0 : (adj_done[q] && lvl_done[q]); };
  std::vector<size_t> tcur(nb, 0), wcur(nb, 0);
  auto first_open = [&](std::vector<int>& list, size_t& cur) {
    while (cur < list.size() && issued(list[cur])) ++cur;
    return cur < list.size() ? list[cur] : nops;
  };
  auto ready = [&](int j, int in, int out) {
    // the cursors skip issued operations; an unissued one before j blocks
    return first_open(touch[out], tcur[out]) >= j && first_open(writer[in], wcur[in]) >= j;
  };
  char* lcur = lam;
  char* lnxt = lam2;
  const int WINDOW = 256;
  for (int i = 0; i < nops; ++i) {
    const int kind = ops[4 * i], a = ops[4 * i + 1], b = ops[4 * i + 2], k = ops[4 * i + 3];
    if (kind == 0) {
      if (!adj_done[i]) {
        Formura_AD_Forward(bufs[a], bufs[b], tmp, k, P, (void*)s);
        adj_done[i] = 1;
      }
      continue;
    }
    if (!lvl_done[i]) {
      fwd_even((const char*)bufs[a], sc[half[i]], P, s);
      lvl_done[i] = 1;
    }
    FwdJob jb[2] = {{nullptr, nullptr, 0}, {nullptr, nullptr, 0}};
    int pick[2] = {-1, -1}, nj = 0;
    for (int j = i + 1; HAS_UNIT_FWD && j < nops && j <= i + WINDOW && nj < 2; ++j) {
      const int kj = ops[4 * j + 3], aj = ops[4 * j + 1];
      if (ops[4 * j] == 1) {
        if (!lvl_done[j] && ready(j, aj, nbufs + half[j])) {
          jb[nj] = FwdJob{(const char*)bufs[aj], sc[half[j]], 2};
          pick[nj++] = j;
        }
      } else if (!adj_done[j] && (kj == 2 || kj == 4) && ready(j, aj, ops[4 * j + 2])) {
        jb[nj] = FwdJob{(const char*)bufs[aj], (char*)bufs[ops[4 * j + 2]], kj};
        pick[nj++] = j;
      }
    }
    if (LONG_JOBS_FIRST && nj == 2 && jb[1].k > jb[0].k) {
      const FwdJob t = jb[0];
      jb[0] = jb[1];
      jb[1] = t;
    }
    if (nj > 0)
      adj_unit_fwd(k, (const char*)bufs[a], sc[half[i]], lcur, lnxt, P, pgrad, nj, jb[0], jb[1], s);
    else
      adj_unit(k, (const char*)bufs[a], sc[half[i]], lcur, lnxt, P, pgrad, s);
    adj_done[i] = 1;
    for (int q = 0; q < nj; ++q) {
      if (ops[4 * pick[q]] == 1)
        lvl_done[pick[q]] = 1;
      else
        adj_done[pick[q]] = 1;
    }
    char* t = lcur;
    lcur = lnxt;
    lnxt = t;
  }
  return lcur == lam ? 0 : 1;
}

extern "C" int Formura_AD_Reverse(const int* ops, int nops, void* const* bufs, int nbufs, void* tmp, void* scratch,
                                  void* lam, void* lam2, const double* params_dev, double* pgrad, void* stream_a,
                                  void* stream_f) {
  hipStream_t A = (hipStream_t)stream_a, F = (hipStream_t)stream_f;
  const bool two = F != A;
  // events: [0, nbufs) last adjoint read of a buffer; nbufs+h last use of scratch
  // half h; nbufs+2+h inner states of half h ready; nbufs+4 stream joins
  EventSet ev(two ? (size_t)nbufs + 5 : 0);
  std::vector<char> was_read(nbufs, 0);
  char half_used[2] = {0, 0};
  char* lcur = (char*)lam;
  char* lnxt = (char*)lam2;
  char* sc[2] = {(char*)scratch, (char*)scratch + (size_t)(KA > 1 ? KA - 1 : 0) * STATE_BYTES};
  int turn = 0;
  if (two) {  // F starts after everything A has queued (the saved states)
    FMR_CHECK(hipEventRecord(ev.e[nbufs + 4], A));
    FMR_CHECK(hipStreamWaitEvent(F, ev.e[nbufs + 4], 0));
  }
  if (FUSED_UNIT && !two) {
    const int r = reverse_fused(ops, nops, bufs, nbufs, tmp, sc, (char*)lam, (char*)lam2, params_dev, pgrad, A);
    FMR_CHECK(hipGetLastError());
    return r;
  }
  for (int i = 0; i < nops; ++i) {
    const int kind = ops[4 * i], a = ops[4 * i + 1], b = ops[4 * i + 2], k = ops[4 * i + 3];
    if (a < 0 || a >= nbufs || (kind == 0 && (b < 0 || b >= nbufs || b == a)) || k < 1)
      fmr_rt::fatal("Formura_AD_Reverse: bad operation");
    if (kind == 0) {  // bufs[b] = k forward steps from bufs[a], on F
      if (two && was_read[b]) FMR_CHECK(hipStreamWaitEvent(F, ev.e[b], 0));
      Formura_AD_Forward(bufs[a], bufs[b], tmp, k, params_dev, (void*)F);
    } else {  // the adjoint of the k steps from bufs[a], on A (inner states on F, or in the fused kernel)
      if (k > KA) fmr_rt::fatal("Formura_AD_Reverse: unit longer than Formura_AD_Levels()");
      const int h = KA > 1 ? (turn++ & 1) : 0;
      if (FUSED_UNIT && !two) {
        fmr_rt::fatal("Formura_AD_Reverse: one-stream fused reversal runs in reverse_fused");
      } else if (FUSED_UNIT) {  // the level-2 state on F, the unit's adjoint (odd levels recomputed) on A
        if (two && half_used[h]) FMR_CHECK(hipStreamWaitEvent(F, ev.e[nbufs + h], 0));
        if (k > 2) fwd_even((const char*)bufs[a], sc[h], params_dev, F);
        if (two) {
          FMR_CHECK(hipEventRecord(ev.e[nbufs + 2 + h], F));
          FMR_CHECK(hipStreamWaitEvent(A, ev.e[nbufs + 2 + h], 0));
        }
        adj_unit(k, (const char*)bufs[a], sc[h], lcur, lnxt, params_dev, pgrad, A);
      } else {
        if (two && half_used[h]) FMR_CHECK(hipStreamWaitEvent(F, ev.e[nbufs + h], 0));
        if (k > 1) Formura_AD_Inner_States(bufs[a], sc[h], k - 1, params_dev, (void*)F);
        if (two) {
          FMR_CHECK(hipEventRecord(ev.e[nbufs + 2 + h], F));
          FMR_CHECK(hipStreamWaitEvent(A, ev.e[nbufs + 2 + h], 0));
        }
        Formura_AD_Adjoint_Levels(bufs[a], sc[h], lcur, lnxt, k, params_dev, pgrad, (void*)A);
      }
      char* t = lcur;
      lcur = lnxt;
      lnxt = t;
      if (two) {
        FMR_CHECK(hipEventRecord(ev.e[a], A));
        was_read[a] = 1;
        FMR_CHECK(hipEventRecord(ev.e[nbufs + h], A));
        half_used[h] = 1;
      }
    }
  }
  if (two) {  // A's later work sees everything F did; the events die on return
    FMR_CHECK(hipEventRecord(ev.e[nbufs + 4], F));
    FMR_CHECK(hipStreamWaitEvent(A, ev.e[nbufs + 4], 0));
    FMR_CHECK(hipStreamSynchronize(A));
  }
  FMR_CHECK(hipGetLastError());
  return lcur == (char*)lam ? 0 : 1;
}
'''
