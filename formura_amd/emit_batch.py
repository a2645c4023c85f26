"""Batched ensembles of one 2-D program: many instances, uneven grids, per-instance
parameters, one persistent launch per K steps, hipGraph-captured step loops.

The generated library implements ``include/formura_batch.h``.  Every instance
is an independent periodic run of the program with its own grid (n1 x n2, both
multiples of the tile), its own values of the lifted literals (e.g. ``f`` and
``k`` of Gray-Scott), and its own drifting frame (offset = -s*steps mod n).

Two precisions:

``f64``   the program's own arithmetic, bit-identical to single-instance runs
          of the same program (and to the reference's C) when the parameters
          equal the program's literals;
``f32k``  FP32 state and FP32 arithmetic, with Kahan-compensated accumulation
          of the time update ``q' = q + dq`` (Kahan 1965): each field carries a
          compensation term c, updated as y = dq - c, t = q + y,
          c' = (t - q) - y, q' = t.  The compensation moves with the cell in the
          drifting frame and stays in LDS across the K fused steps.
``f32c``  as ``f32k``, and the increment dq evaluated in float-float FP32
          arithmetic (``ff_increment``): every value of dq is a pair (hi, lo)
          of FP32 numbers built with error-free transformations -- TwoProduct
          through the fused multiply-add, TwoSum (Knuth) -- and constants and
          lifted parameters enter as their FP32 pairs; only the stencil parts
          (linear combinations of loads at two or more cursors with constant
          factors that are FP32 numbers: the Laplacian) stay plain FP32, and a
          rounded coefficient (a diffusivity) multiplies them as a pair.  The pair's
          low part is folded into the Kahan step (y = dq_hi - (c - dq_lo)).
          Near a steady pattern the point-local terms of a reaction nearly
          cancel (-u v^2 + F (1 - u) in Gray-Scott), so their plain FP32
          rounding is the same bias at every step and accumulates; the pairs
          remove it.

Kernel: the ``tile2d`` scheme (tile + low-side halo of 2*K*s in LDS, K steps,
ping-pong buffers) over a table of tiles of all instances; a fixed-size grid of
workgroups loops over the tile table (persistent launch).  The host captures G
launches (an even number, so the ping-pong buffers return to their start) in
a hipGraph and replays it.
"""
from __future__ import annotations

import json
from typing import Dict, List

from . import syntax as S
from .emit import LDS_LOAD_LINE, Body, tile_load, EmitError, GraphPlan, state_fields
from .manifest import calc_range

PRECISIONS = {"f64": ("double", False), "f32k": ("float", True), "f32c": ("float", True),
              "f32x": ("float", True)}


def _f32(x: float) -> float:
    import numpy as np
    return float(np.float32(x))


def _f32_exact(c, reciprocal: bool) -> bool:
    """Whether a folded constant (or its reciprocal) is an FP32 number to 2^-40."""
    kind, v = c
    if kind != "value" or v == 0.0:
        return False
    x = 1.0 / v if reciprocal else v
    return abs(_f32(x) - x) <= 2.0 ** -40 * abs(x)


def _pow2(x: float) -> bool:
    import math
    return x != 0.0 and math.isfinite(x) and math.frexp(abs(x))[0] == 0.5


def _flit(x: float) -> str:
    """An FP32 literal that parses to exactly the float x."""
    if x != x or x in (float("inf"), float("-inf")):
        raise EmitError("non-finite constant in the float-float increment")
    return "%.9ef" % x


def ff_increment(node, body: Body, root: int, pidx: Dict[str, int], pre: str, stencil_plain: bool = True,
                 comp_load=None):
    """Float-float evaluation of the sub-DAG of ``node`` below micro ``root`` (the
    increment dq of an update q + dq).  ``body`` is the node's plain FP32 Body
    (its values ``<prefix><j>`` and folded constants ``body.vals``); ``pidx``
    maps lifted literals to their index in the parameter array P (FP64).
    Returns (lines, hi, lo) with C expressions of the pair's parts.

    Stencil-linear values that read two or more cursors, through + - and
    constant factors that are FP32 numbers (to 2^-40), keep their plain FP32
    value (lo = 0); all other values are pairs.  Operations on pairs:
      x + y: s = xh + yh, e = TwoSum error + xl + yl   (Knuth's TwoSum)
      x * y: p = xh * yh, e = fma(xh, yh, -p) + xh*yl + xl*yh   (TwoProduct)
      x / c: x * (hi, lo) of 1/c (c a constant, 1/c rounded in FP64)
    A pair is left unnormalized between operations (|lo| stays near ulp(hi));
    the result is normalized once (Fast2Sum).  Operations outside + - * / and
    unary minus (square roots, comparisons, selections) take the hi parts.

    ``comp_load(field, cursor)``: a C expression of the Kahan compensation c of
    a state load; the load is then the pair (q, -c) -- the compensated value
    the accumulation carries -- and every value is a pair (no plain stencil
    part), so the increment is computed from the state to about 2^-48."""
    micro = node.micro
    # the plain FP32 value of micro j is named like the Body's: prefix + j
    prefix = body.result.rsplit("_", 1)[0] + "_"
    need = set()
    stack = [root]
    while stack:
        j = stack.pop()
        if j in need:
            continue
        need.add(j)
        m = micro[j]
        k = m.inst[0]
        if k == "Uniop":
            stack.append(m.inst[2])
        elif k == "Binop":
            stack += [m.inst[2], m.inst[3]]
        elif k == "Triop":
            stack += [m.inst[2], m.inst[3], m.inst[4]]
    const, lin, curs = {}, {}, {}
    for j in sorted(need):
        m = micro[j]
        k = m.inst[0]
        if m.src in pidx and k not in ("LoadCursorStatic", "LoadCursor", "LoadIndex"):
            const[j] = ("param", pidx[m.src])
        elif j in body.vals and isinstance(body.vals[j], float):
            const[j] = ("value", body.vals[j])
        elif j in body.vals:
            const[j] = ("value", float(body.vals[j]))
        if j in const:
            lin[j], curs[j] = True, set()
        elif k == "LoadCursorStatic":
            lin[j], curs[j] = True, {(m.inst[2], tuple(m.inst[1]))}
        elif k == "Uniop" and m.inst[1] == "-":
            lin[j], curs[j] = lin[m.inst[2]], set(curs[m.inst[2]])
        elif k == "Binop" and m.inst[1] in ("+", "-"):
            a, b = m.inst[2], m.inst[3]
            lin[j], curs[j] = lin[a] and lin[b], curs[a] | curs[b]
        elif k == "Binop" and m.inst[1] in ("*", "/"):
            # a constant factor keeps the product in the plain stencil part only if
            # its FP32 value is exact (to 2^-40): a rounded diffusion coefficient is
            # the same relative bias at every step
            a, b = m.inst[2], m.inst[3]
            div = m.inst[1] == "/"
            lin[j] = ((a in const and lin[b] and not div and _f32_exact(const[a], False))
                      or (b in const and lin[a] and _f32_exact(const[b], div)))
            curs[j] = curs[a] | curs[b]
        else:
            lin[j], curs[j] = False, set().union(*[curs.get(x, set()) for x in m.inst[2:] if isinstance(x, int)])
    plain = ({j for j in need if j not in const and lin[j] and len(curs[j]) >= 2}
             if stencil_plain and comp_load is None else set())
    lines = []
    pair = {}

    def cpair(c):
        kind, v = const[c]
        if kind == "param":
            return "Ph[%d]" % v, "Pl[%d]" % v
        h = _f32(v)
        lo = _f32(v - h)
        return _flit(h), (_flit(lo) if lo != 0.0 else None)

    def get(x):
        if x in const:
            return cpair(x)
        if micro[x].inst[0] == "LoadCursorStatic" and comp_load is not None:
            return "%s%d" % (prefix, x), "(-%s)" % comp_load(micro[x].inst[2], micro[x].inst[1])
        if x in plain or micro[x].inst[0] == "LoadCursorStatic":
            return "%s%d" % (prefix, x), None
        return pair[x]

    for j in sorted(need):
        if j in const or j in plain or micro[j].inst[0] == "LoadCursorStatic":
            continue
        m = micro[j]
        k = m.inst[0]
        h, l = "%sh%d" % (pre, j), "%sl%d" % (pre, j)
        if k == "Uniop" and m.inst[1] == "-":
            xh, xl = get(m.inst[2])
            lines.append("const float %s = -%s;" % (h, xh))
            pair[j] = (h, None) if xl is None else (h, "(-%s)" % xl)
            if xl is not None:
                lines.append("const float %s = -%s;" % (l, xl))
                pair[j] = (h, l)
            continue
        if k == "Binop" and m.inst[1] in ("+", "-"):
            (xh, xl), (yh, yl) = get(m.inst[2]), get(m.inst[3])
            if m.inst[1] == "-":
                yh = "(-%s)" % yh
                yl = None if yl is None else "(-%s)" % yl
            lines.append("const float %s = %s + %s;" % (h, xh, yh))
            lines.append("const float %sb%d = %s - %s;" % (pre, j, h, xh))
            err = "((%s - (%s - %sb%d)) + (%s - %sb%d))" % (xh, h, pre, j, yh, pre, j)
            extra = "".join(" + %s" % t for t in (xl, yl) if t is not None)
            lines.append("const float %s = %s%s;" % (l, err, extra))
            pair[j] = (h, l)
            continue
        if k == "Binop" and m.inst[1] in ("*", "/"):
            a, b = m.inst[2], m.inst[3]
            if m.inst[1] == "/":
                if b not in const or const[b][0] != "value":
                    raise EmitError("float-float increment: division by a non-constant")
                r = 1.0 / const[b][1]
                yh, yl = _flit(_f32(r)), _f32(r - _f32(r))
                yl = _flit(yl) if yl != 0.0 else None
            else:
                yh, yl = get(b)
            xh, xl = get(a)
            if a in const and b not in const:
                (xh, xl), (yh, yl), a, b = (yh, yl), (xh, xl), b, a
            if b in const and const[b][0] == "value" and yl is None and _pow2(const[b][1] if m.inst[1] == "*" else 1.0 / const[b][1]):
                # a power of two: the product is exact
                lines.append("const float %s = %s * %s;" % (h, xh, yh))
                if xl is None:
                    pair[j] = (h, None)
                else:
                    lines.append("const float %s = %s * %s;" % (l, xl, yh))
                    pair[j] = (h, l)
                continue
            lines.append("const float %s = %s * %s;" % (h, xh, yh))
            e = "fmaf(%s, %s, -%s)" % (xh, yh, h)
            if yl is not None:
                e = "fmaf(%s, %s, %s)" % (xh, yl, e)
            if xl is not None:
                e = "fmaf(%s, %s, %s)" % (xl, yh, e)
            lines.append("const float %s = %s;" % (l, e))
            pair[j] = (h, l)
            continue
        # anything else: on the hi parts, in FP32 (the Body's expression with hi operands)
        lines.append("const float %s = %s%d;" % (h, prefix, j))
        pair[j] = (h, None)
    rh, rl = get(root)
    if rl is None:
        return lines, rh, "0.0f"
    lines.append("const float %sH = %s + %s;" % (pre, rh, rl))
    lines.append("const float %sL = %s - (%sH - %s);" % (pre, rl, pre, rh))
    return lines, "%sH" % pre, "%sL" % pre


class BatchEmitter:
    def __init__(self, mm, name: str, precision: str, params: List[str], tx: int, ty: int, k: int,
                 graph_launches: int = 16, grid_wgs: int = 2048, threads: int = 256):
        self.mm, self.om, self.name = mm, mm.om, name
        if self.om.dimension != 2:
            raise EmitError("batched ensembles are built for 2-D programs")
        if precision not in PRECISIONS:
            raise EmitError("unknown precision %r" % precision)
        self.real, self.kahan = PRECISIONS[precision]
        self.precision = precision
        self.compensated = precision in ("f32c", "f32x")
        # f32x: the increment reads the compensated state (q, -c) as pairs
        self.comp_loads = precision == "f32x"
        self.params = list(params)
        self.tx, self.ty, self.k = tx, ty, k
        self.G = graph_launches + (graph_launches % 2)
        self.grid_wgs = grid_wgs
        # work-items per workgroup
        self.threads = int(threads)
        self.plan = GraphPlan(mm.step, self.om.config.sleeve, 2, "step")
        if self.plan.temps:
            raise EmitError("batched ensembles need a program without manifest temporaries")
        self.fields = state_fields(self.om)
        self.fi = {n: i for i, (n, _) in enumerate(self.fields)}
        names = set()
        for node in self.plan.mm.values():
            for m in node.micro:
                if m.src:
                    names.add(m.src)
                if m.inst[0] == "LoadIndex":
                    raise EmitError("index expressions are not supported in batched ensembles")
        missing = [p for p in self.params if p not in names]
        if missing:
            raise EmitError("parameters %s are not literals of the step" % missing)

    def emit(self) -> str:
        s = self.plan.sleeve
        nf = len(self.fields)
        real = self.real
        pmap = {p: "P[%d]" % i for i, p in enumerate(self.params)}
        lines = []
        if self.compensated:
            # FP32 pairs of the FP64 parameters: P = Ph + Pl to about 2^-48
            lines += ["float Ph[NP], Pl[NP];",
                      "#pragma unroll",
                      "for (int q = 0; q < NP; ++q) { Ph[q] = (float)P[q]; Pl[q] = (float)(P[q] - (double)Ph[q]); }"]
        hoisted = []
        for k in self.plan.stores:
            node = self.plan.mm[k]
            base = self.plan.rng[k].offset - 2 * s
            fname = node.micro[-1].inst[1]
            fi = self.fi[fname]

            def load(field, c, base=base):
                return "lds%d[o + (%d)*RY + (%d)]" % (self.fi[field], c[0] + base, c[1] + base)

            # FP64: divisions by constants as the fused sequence, redone with the
            # guarded form only by a wave holding a dividend outside its range
            body = Body(node, load, lambda *a: None, lambda ax: None, lambda n, v: "", "s%d_" % k,
                        params=pmap, real=("float" if real == "float" else None), divfix=real != "float")
            # the point's loads first (names unique to the node), then its arithmetic
            hoisted += [l for l in body.lines if LDS_LOAD_LINE.search(l)]
            lines.append("{")
            lines += ["  " + l for l in body.lines if not LDS_LOAD_LINE.search(l)]
            if self.compensated:
                q, dq = self._update_ids(node, fname, base)

                def cload(field, c, base=base):
                    return "clds%d[o + (%d)*RY + (%d)]" % (self.fi[field], c[0] + base, c[1] + base)
                fl, hi, lo = ff_increment(node, body, dq, {p: i for i, p in enumerate(self.params)}, "c%d_" % k,
                                          comp_load=cload if self.comp_loads else None)
                lines += ["  " + l for l in fl]
                lines.append("  old%d = s%d_%d; inc%d = %s; incl%d = %s;" % (fi, k, q, fi, hi, fi, lo))
            elif self.kahan:
                q, dq = self._update_ids(node, fname, base)
                lines.append("  old%d = s%d_%d; inc%d = s%d_%d;" % (fi, k, q, fi, k, dq))
            else:
                lines.append("  old%d = 0; inc%d = s%d_%d;" % (fi, fi, k, node.micro[-1].inst[2]))
            lines.append("}")
        params = ", ".join("const %s* __restrict__ lds%d" % (real, i) for i in range(nf))
        if self.comp_loads:
            params += "".join(", const %s* __restrict__ clds%d" % (real, i) for i in range(nf))
        outs = ", ".join("%s& old%d, %s& inc%d" % (real, i, real, i) for i in range(nf))
        if self.compensated:
            outs += "".join(", float& incl%d" % i for i in range(nf))
        cbase = "(%d)*RY + (%d)" % (self.plan.rng[self.plan.stores[0]].offset - 2 * s,
                                    self.plan.rng[self.plan.stores[0]].offset - 2 * s)
        lds = "\n".join("  __shared__ %s B%d[2][RX * RY];" % (real, i) for i in range(nf))
        if self.kahan:
            lds += "\n" + "\n".join("  __shared__ %s C%d[2][RX * RY];" % (real, i) for i in range(nf))
        pairs = [("B%d[0][idx]" % i, "src[%dLL * plane + g]" % i) for i in range(nf)]
        if self.kahan:
            pairs += [("C%d[0][idx]" % i, "src[%dLL * plane + g]" % (nf + i)) for i in range(nf)]
        load_tile = tile_load("RX * RY", ["const int r = idx / RY, c = idx - r * RY;",
                                          "const long long g = (long long)wrapn(tx0 - H + r, I.n1) * I.n2 "
                                          "+ wrapn(ty0 - H + c, I.n2);"], pairs, real, "ld", indent="    ",
                              threads=self.threads)
        srcs = ", ".join("B%d[sb]" % i for i in range(nf))
        if self.comp_loads:
            srcs += "".join(", C%d[sb]" % i for i in range(nf))
        decl = " ".join("%s old%d, inc%d;" % (real, i, i) for i in range(nf))
        refs = ", ".join("old%d, inc%d" % (i, i) for i in range(nf))
        if self.compensated:
            decl += " " + " ".join("float incl%d;" % i for i in range(nf))
            refs += "".join(", incl%d" % i for i in range(nf))
        if self.comp_loads:
            # f32x: the state is the float-float pair (q, -c) and the update a
            # float-float sum of (q, -c) and the increment (dq_hi, dq_lo):
            # TwoSum of the high parts, the low parts added, then Fast2Sum
            # (kt - kc is the new value to about 2^-48)
            upd = "\n".join(
                ("        const float ks%(i)d = old%(i)d + inc%(i)d;\n"
                 "        const float kb%(i)d = ks%(i)d - old%(i)d;\n"
                 "        const float kl%(i)d = ((old%(i)d - (ks%(i)d - kb%(i)d)) + (inc%(i)d - kb%(i)d))"
                 " + (incl%(i)d - C%(i)d[sb][o + CB]);\n"
                 "        const float kt%(i)d = ks%(i)d + kl%(i)d;\n"
                 "        const float kc%(i)d = (kt%(i)d - ks%(i)d) - kl%(i)d;") % {"i": i} for i in range(nf))
            st_lds = " ".join("B%d[db][o] = kt%d; C%d[db][o] = kc%d;" % (i, i, i, i) for i in range(nf))
            st_glb = " ".join("dst[%dLL * plane + g] = kt%d; dst[%dLL * plane + g] = kc%d;" % (i, i, nf + i, i)
                              for i in range(nf))
        elif self.kahan:
            # Kahan: ky = dq - c, kt = q + ky, kc = (kt - q) - ky (names that shadow nothing);
            # f32c adds the increment's low part: ky = dq_hi - (c - dq_lo)
            cterm = ("(C%(i)d[sb][o + CB] - incl%(i)d)" if self.compensated else "C%(i)d[sb][o + CB]")
            upd = "\n".join(
                ("        const %(r)s ky%(i)d = inc%(i)d - " + cterm + ";\n"
                "        const %(r)s kt%(i)d = old%(i)d + ky%(i)d;\n"
                "        const %(r)s kc%(i)d = (kt%(i)d - old%(i)d) - ky%(i)d;") % {"r": real, "i": i} for i in range(nf))
            st_lds = " ".join("B%d[db][o] = kt%d; C%d[db][o] = kc%d;" % (i, i, i, i) for i in range(nf))
            st_glb = " ".join("dst[%dLL * plane + g] = kt%d; dst[%dLL * plane + g] = kc%d;" % (i, i, nf + i, i)
                              for i in range(nf))
        else:
            upd = "\n".join("        const %s kt%d = inc%d;" % (real, i, i) for i in range(nf))
            st_lds = " ".join("B%d[db][o] = kt%d;" % (i, i) for i in range(nf))
            st_glb = " ".join("dst[%dLL * plane + g] = kt%d;" % (i, i) for i in range(nf))
        nslots = nf * (2 if self.kahan else 1)
        desc = {"name": self.name, "precision": self.precision, "real": real, "kahan": self.kahan,
                "fields": [n for n, _ in self.fields], "params": self.params, "sleeve": s,
                "tile": [self.tx, self.ty], "steps_per_launch": self.k, "graph_launches": self.G,
                "grid_wgs": self.grid_wgs, "threads": self.threads, "slots_per_cell": nslots}
        return TEMPLATE % dict(
            name=self.name, real=real, nf=nf, nslots=nslots, np=max(1, len(self.params)), sleeve=s,
            TX=self.tx, TY=self.ty, NTH=self.threads, KMAX=self.k, G=self.G, GRID=self.grid_wgs, desc=json.dumps(desc),
            pt_params=params, pt_outs=outs, pt_body="\n".join("  " + l for l in hoisted + lines),
            lds=lds, load_tile=load_tile, srcs=srcs, decl=decl, refs=refs, upd=upd, st_lds=st_lds,
            st_glb=st_glb, CB=cbase, kahan=1 if self.kahan else 0,
            preload="\n".join("  { hipFuncAttributes fa; FMR_CHECK(hipFuncGetAttributes(&fa, (const void*)k_batch<%d>)); }"
                              % kk for kk in range(1, self.k + 1)),
            launch_cases="".join(
                "    case %d: hipLaunchKernelGGL(k_batch<%d>, dim3(grid), dim3(%d), 0, b->stream, b->inst_d, "
                "b->tile_inst_d, b->ntiles, in, out); break;\n" % (kk, kk, self.threads) for kk in range(1, self.k + 1)))

    def _update_ids(self, node, fname, base):
        st = node.micro[-1]
        a = node.micro[st.inst[2]]
        if a.inst[0] != "Binop" or a.inst[1] != "+":
            raise EmitError("the f32k form needs an update of the form q + dq (node %s)" % node.name)
        for x, y in ((a.inst[2], a.inst[3]), (a.inst[3], a.inst[2])):
            m = node.micro[x]
            if (m.inst[0] == "LoadCursorStatic" and m.inst[2] == fname
                    and all(c + base == base for c in m.inst[1])):
                return x, y
        raise EmitError("the f32k form needs an update q + dq with q the cell's own value (node %s)" % node.name)


TEMPLATE = r'''// Generated by formura_amd: batched ensemble of %(name)s. Do not edit.
#include "formura_hip_rt.h"
#include "formura_batch.h"
#include <vector>

namespace {
typedef %(real)s real_t;
constexpr int NF = %(nf)d;        // state fields
constexpr int NSLOT = %(nslots)d; // arrays per instance (fields, then compensations)
constexpr int NP = %(np)d;        // lifted parameters
constexpr int SLEEVE = %(sleeve)d;
constexpr int KMAX = %(KMAX)d;
constexpr int GRAPH_LAUNCHES = %(G)d;
const char* const kDescription = R"FMRJSON(%(desc)s)FMRJSON";

struct Inst {
  long long base;  // element offset of the instance's arrays
  int n1, n2, nty, tile0;
  double p[NP];
};

template <int RY>
__device__ __forceinline__ void pt(%(pt_params)s, const int o, const double* P, %(pt_outs)s) {
%(pt_body)s
}

__device__ __forceinline__ int wrapn(int i, int n) { return i < 0 ? i + n : (i >= n ? i - n : i); }

template <int K>
__global__ void __launch_bounds__(%(NTH)d) k_batch(const Inst* __restrict__ inst, const int* __restrict__ tile_inst,
                                               int ntiles, const real_t* __restrict__ in, real_t* __restrict__ out) {
  constexpr int TX = %(TX)d, TY = %(TY)d, H = 2 * K * SLEEVE, RX = TX + H, RY = TY + H;
  constexpr int CB = %(CB)s;  // the cell's own value relative to the output point
%(lds)s
  for (int tile = blockIdx.x; tile < ntiles; tile += gridDim.x) {  // persistent loop over all instances' tiles
    const Inst I = inst[tile_inst[tile]];
    const int lt = tile - I.tile0;
    const int tx0 = (lt / I.nty) * TX, ty0 = (lt %% I.nty) * TY;
    const long long plane = (long long)I.n1 * I.n2;
    const real_t* __restrict__ src = in + I.base;
    real_t* __restrict__ dst = out + I.base;
    double P[NP];
#pragma unroll
    for (int q = 0; q < NP; ++q) P[q] = I.p[q];
%(load_tile)s
    __syncthreads();
#pragma unroll
    for (int j = 1; j <= K; ++j) {
      const int sb = (j - 1) & 1, db = j & 1;
      const int lo = 2 * j * SLEEVE, W = RY - lo, HH = RX - lo;
      for (int idx = threadIdx.x; idx < HH * W; idx += %(NTH)d) {
        const int rr = idx / W;
        const int r = lo + rr, c = lo + (idx - rr * W);
        const int o = r * RY + c;
        %(decl)s
        pt<RY>(%(srcs)s, o, P, %(refs)s);
%(upd)s
        if (j < K) {
          %(st_lds)s
        } else {
          const long long g = (long long)(tx0 + r - H) * I.n2 + (ty0 + c - H);
          %(st_glb)s
        }
      }
      __syncthreads();
    }
  }
}
}  // namespace

struct FormuraBatch {
  int ninst = 0, ntiles = 0;
  std::vector<Inst> inst;
  std::vector<long long> steps_done;
  Inst* inst_d = nullptr;
  int* tile_inst_d = nullptr;
  real_t *cur = nullptr, *nxt = nullptr;
  long long total = 0;
  hipStream_t stream = nullptr;
  hipEvent_t ev0 = nullptr, ev1 = nullptr;
  // one captured graph per ping-pong parity: graph p reads buf[p] (and, with an
  // even number of launches, leaves the state there)
  real_t* buf[2] = {nullptr, nullptr};
  hipGraphExec_t exec[2] = {nullptr, nullptr};
  hipGraph_t graph[2] = {nullptr, nullptr};
  double last_ms = 0.0;
  long long launches = 0;
};

static void launch_k(FormuraBatch* b, int k, const real_t* in, real_t* out) {
  const int grid = b->ntiles < %(GRID)d ? b->ntiles : %(GRID)d;
  switch (k) {
%(launch_cases)s    default: fmr_rt::fatal("bad step count");
  }
  ++b->launches;
}

extern "C" const char* Formura_Batch_Describe(void) { return kDescription; }

extern "C" FormuraBatch* Formura_Batch_Create(int ninst, const int* n1, const int* n2, const double* params) {
  constexpr int TX = %(TX)d, TY = %(TY)d;
  if (ninst <= 0) fmr_rt::fatal("Formura_Batch_Create: no instances");
  FormuraBatch* b = new FormuraBatch();
  b->ninst = ninst;
  b->inst.resize(ninst);
  b->steps_done.assign(ninst, 0);
  std::vector<int> tile_inst;
  long long base = 0;
  for (int i = 0; i < ninst; ++i) {
    if (n1[i] %% TX || n2[i] %% TY || n1[i] <= 2 * KMAX * SLEEVE || n2[i] <= 2 * KMAX * SLEEVE)
      fmr_rt::fatal("Formura_Batch_Create: every grid extent must be a multiple of the tile and exceed the halo");
    Inst& I = b->inst[i];
    I.base = base;
    I.n1 = n1[i];
    I.n2 = n2[i];
    I.nty = n2[i] / TY;
    I.tile0 = (int)tile_inst.size();
    for (int q = 0; q < NP; ++q) I.p[q] = params ? params[(long long)i * NP + q] : 0.0;
    const int nt = (n1[i] / TX) * (n2[i] / TY);
    for (int t = 0; t < nt; ++t) tile_inst.push_back(i);
    base += (long long)NSLOT * n1[i] * n2[i];
  }
  b->ntiles = (int)tile_inst.size();
  b->total = base;
  int ndev = 0;
  if (hipGetDeviceCount(&ndev) != hipSuccess || ndev < 1) fmr_rt::fatal("no HIP device is visible");
  b->inst_d = (Inst*)fmr_rt::device_alloc(sizeof(Inst) * ninst);
  b->tile_inst_d = (int*)fmr_rt::device_alloc(sizeof(int) * tile_inst.size());
  b->cur = (real_t*)fmr_rt::device_alloc(sizeof(real_t) * base);
  b->nxt = (real_t*)fmr_rt::device_alloc(sizeof(real_t) * base);
  b->buf[0] = b->cur;
  b->buf[1] = b->nxt;
  // load every kernel's code object now: the first launch of a kernel must not
  // happen inside a stream capture (Formura_Batch_Step captures launches)
%(preload)s
  FMR_CHECK(hipStreamCreateWithFlags(&b->stream, hipStreamNonBlocking));
  FMR_CHECK(hipEventCreate(&b->ev0));
  FMR_CHECK(hipEventCreate(&b->ev1));
  FMR_CHECK(hipMemcpyAsync(b->inst_d, b->inst.data(), sizeof(Inst) * ninst, hipMemcpyHostToDevice, b->stream));
  FMR_CHECK(hipMemcpyAsync(b->tile_inst_d, tile_inst.data(), sizeof(int) * tile_inst.size(), hipMemcpyHostToDevice,
                           b->stream));
  FMR_CHECK(hipStreamSynchronize(b->stream));
  return b;
}

// state of one instance as NF fields of n1*n2 doubles (row-major).  With Kahan
// compensation a value x is stored as the pair (q, c) = (RN(x), RN(q - x)),
// which carries x to about 2^-48 (q - c is the compensated value)
extern "C" void Formura_Batch_Set_State(FormuraBatch* b, int i, const double* fields) {
  if (i < 0 || i >= b->ninst) fmr_rt::fatal("Formura_Batch_Set_State: bad instance");
  const Inst& I = b->inst[i];
  const long long plane = (long long)I.n1 * I.n2;
  std::vector<real_t> h((size_t)NSLOT * plane, (real_t)0);
  for (long long e = 0; e < NF * plane; ++e) {
    h[e] = (real_t)fields[e];
    if (NSLOT > NF) h[NF * plane + e] = (real_t)((double)h[e] - fields[e]);
  }
  FMR_CHECK(hipMemcpyAsync(b->cur + I.base, h.data(), sizeof(real_t) * h.size(), hipMemcpyHostToDevice, b->stream));
  FMR_CHECK(hipStreamSynchronize(b->stream));
  b->steps_done[i] = 0;
}

extern "C" void Formura_Batch_Get_State(FormuraBatch* b, int i, double* fields) {
  if (i < 0 || i >= b->ninst) fmr_rt::fatal("Formura_Batch_Get_State: bad instance");
  const Inst& I = b->inst[i];
  const long long plane = (long long)I.n1 * I.n2;
  std::vector<real_t> h((size_t)NSLOT * plane);
  FMR_CHECK(hipMemcpyAsync(h.data(), b->cur + I.base, sizeof(real_t) * h.size(), hipMemcpyDeviceToHost, b->stream));
  FMR_CHECK(hipStreamSynchronize(b->stream));
  // with Kahan compensation the value is q - c, formed in FP64
  for (long long e = 0; e < NF * plane; ++e)
    fields[e] = NSLOT > NF ? (double)h[e] - (double)h[NF * plane + e] : (double)h[e];
}

// drifting-frame offsets of instance i after the steps run so far
extern "C" void Formura_Batch_Offsets(FormuraBatch* b, int i, int* off) {
  if (i < 0 || i >= b->ninst) fmr_rt::fatal("Formura_Batch_Offsets: bad instance");
  const Inst& I = b->inst[i];
  const long long sh = (long long)SLEEVE * b->steps_done[i];
  off[0] = (int)(((-sh) %% I.n1 + I.n1) %% I.n1);
  off[1] = (int)(((-sh) %% I.n2 + I.n2) %% I.n2);
}

// nsteps time steps of every instance; returns the device time in ms (HIP events
// on the batch's stream).  Full K-step launches are replayed from a hipGraph of
// GRAPH_LAUNCHES launches.
extern "C" double Formura_Batch_Step(FormuraBatch* b, int nsteps) {
  if (nsteps < 0) fmr_rt::fatal("Formura_Batch_Step: negative count");
  b->launches = 0;
  FMR_CHECK(hipEventRecord(b->ev0, b->stream));
  long long full = nsteps / KMAX;
  const int rem = nsteps %% KMAX;
  if (full >= GRAPH_LAUNCHES) {
    const int par = (b->cur == b->buf[0]) ? 0 : 1;
    if (!b->exec[par]) {
      FMR_CHECK(hipStreamBeginCapture(b->stream, hipStreamCaptureModeThreadLocal));
      for (int g = 0; g < GRAPH_LAUNCHES; ++g) {
        launch_k(b, KMAX, b->cur, b->nxt);
        real_t* t = b->cur; b->cur = b->nxt; b->nxt = t;
      }
      FMR_CHECK(hipStreamEndCapture(b->stream, &b->graph[par]));
      FMR_CHECK(hipGraphInstantiate(&b->exec[par], b->graph[par], nullptr, nullptr, 0));
      b->launches -= GRAPH_LAUNCHES;
    }
    // graph `par` reads buf[par] and, after an even number of launches, leaves the
    // state in buf[par] == cur again
    while (full >= GRAPH_LAUNCHES) {
      FMR_CHECK(hipGraphLaunch(b->exec[par], b->stream));
      b->launches += GRAPH_LAUNCHES;
      full -= GRAPH_LAUNCHES;
    }
  }
  for (long long l = 0; l < full; ++l) {
    launch_k(b, KMAX, b->cur, b->nxt);
    real_t* t = b->cur; b->cur = b->nxt; b->nxt = t;
  }
  if (rem) {
    launch_k(b, rem, b->cur, b->nxt);
    real_t* t = b->cur; b->cur = b->nxt; b->nxt = t;
  }
  FMR_CHECK(hipGetLastError());
  FMR_CHECK(hipEventRecord(b->ev1, b->stream));
  FMR_CHECK(hipEventSynchronize(b->ev1));
  float ms = 0.f;
  FMR_CHECK(hipEventElapsedTime(&ms, b->ev0, b->ev1));
  b->last_ms = ms;
  for (int i = 0; i < b->ninst; ++i) b->steps_done[i] += nsteps;
  return ms;
}

extern "C" long long Formura_Batch_Last_Launches(FormuraBatch* b) { return b->launches; }
extern "C" long long Formura_Batch_Cells(FormuraBatch* b) { return b->total / NSLOT; }

extern "C" void Formura_Batch_Destroy(FormuraBatch* b) {
  if (!b) return;
  FMR_CHECK(hipStreamSynchronize(b->stream));
  for (int p = 0; p < 2; ++p) {
    if (b->exec[p]) FMR_CHECK(hipGraphExecDestroy(b->exec[p]));
    if (b->graph[p]) FMR_CHECK(hipGraphDestroy(b->graph[p]));
  }
  fmr_rt::device_free(b->inst_d);
  fmr_rt::device_free(b->tile_inst_d);
  fmr_rt::device_free(b->cur);
  fmr_rt::device_free(b->nxt);
  FMR_CHECK(hipEventDestroy(b->ev0));
  FMR_CHECK(hipEventDestroy(b->ev1));
  FMR_CHECK(hipStreamDestroy(b->stream));
  delete b;
}
'''
