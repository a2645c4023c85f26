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
from .emit import Body, EmitError, GraphPlan, _const_literal, _f32, show_double, state_fields
from .manifest import calc_range, inline_temporaries


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
    def __init__(self, mm, name: str, params: List[str], real: str = "double", block: int = 256):
        self.mm, self.om, self.name = mm, mm.om, name
        self.real = real
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

    def _forward_body(self, k, shift, prefix):
        """Forward instructions of Store node k at output point (q + shift)."""
        s = self.plan.sleeve
        base = self.plan.rng[k].offset - 2 * s

        def load(field, c):
            return "FLD%d(st)[%s]" % (self.fi[field], self._idx(self._pos([shift[a] + c[a] + base for a in range(self.dim)])))

        pmap = {p: "P[%d]" % i for i, p in enumerate(self.params)}
        b = Body(self.plan.mm[k], load, lambda *a: None, lambda ax: None, lambda n, v: "", prefix,
                 params=pmap, real=("float" if self.real == "float" else None))
        return b

    def emit(self) -> str:
        s = self.plan.sleeve
        dim = self.dim
        R = self.real
        nf = len(self.fields)
        np_ = max(1, len(self.params))
        # distinct load offsets delta = c + base, per (store, field)
        deltas = []
        for k in self.plan.stores:
            base = self.plan.rng[k].offset - 2 * s
            for m in self.plan.mm[k].micro:
                if m.inst[0] == "LoadCursorStatic":
                    d = tuple(c + base for c in m.inst[1])
                    if d not in deltas:
                        deltas.append(d)
        deltas.sort()
        body = self._decompose()
        body += ["%s acc%d = 0;" % (R, g) for g in range(nf)]
        for di, delta in enumerate(deltas):
            shift = [-x for x in delta]  # output point i = q - delta
            body.append("{  // outputs that read this point at offset %s" % (delta,))
            lam = ["LAM%d(lam_next)[%s]" % (self.fi[self.plan.mm[k].micro[-1].inst[1]], self._idx(self._pos(shift)))
                   for k in self.plan.stores]
            for si, k in enumerate(self.plan.stores):
                node = self.plan.mm[k]
                base = self.plan.rng[k].offset - 2 * s
                fb = self._forward_body(k, shift, "f%d_%d_" % (di, si))
                body += ["  " + l for l in fb.lines]
                consts = getattr(fb, "vals", {})
                body.append("  const %s lam%d_%d = %s;" % (R, di, si, lam[si]))
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
                        body += ["  " + l for l in rs.lines]
                        body.append("  acc%d += %s;" % (g, " + ".join(rs.result[j] for j in sorted(rs.result))))
            body.append("}")
        # parameter gradients at output i = q
        pg = ["double pg%d = 0;" % p for p in range(np_)]
        if self.params:
            pg.append("{")
            for si, k in enumerate(self.plan.stores):
                node = self.plan.mm[k]
                fb = self._forward_body(k, [0] * dim, "p%d_" % si)
                pg += ["  " + l for l in fb.lines]
                pg.append("  const %s plam%d = LAM%d(lam_next)[%s];"
                          % (R, si, self.fi[node.micro[-1].inst[1]], self._idx(self._ivs())))
                for pi, p in enumerate(self.params):
                    leaves = {j for j, m in enumerate(node.micro) if m.src in self.params}
                    targets = {j for j, m in enumerate(node.micro) if m.src == p}
                    rs = ReverseSweep(node, targets, "plam%d" % si, "p%d_" % si, "pb%d_%d_" % (si, pi),
                                      getattr(fb, "vals", {}), R, leaves=leaves)
                    if rs.result:
                        pg += ["  " + l for l in rs.lines]
                        pg.append("  pg%d += (double)(%s);" % (pi, " + ".join(rs.result[j] for j in sorted(rs.result))))
            pg.append("}")
        body += pg
        body += ["if (live) LAM%d(lam_prev)[t] = acc%d;" % (g, g) for g in range(nf)]
        # block reduction of the parameter gradients (FP64), one slot per block
        body.append("__shared__ double red[%d][%d];" % (np_, self.block))
        body += ["red[%d][threadIdx.x] = live ? pg%d : 0.0;" % (p, p) for p in range(np_)]
        body.append("__syncthreads();")
        body.append("for (int w = %d; w > 0; w >>= 1) {" % (self.block // 2))
        body += ["  if ((int)threadIdx.x < w) red[%d][threadIdx.x] += red[%d][threadIdx.x + w];" % (p, p) for p in range(np_)]
        body.append("  __syncthreads();")
        body.append("}")
        body.append("if (threadIdx.x == 0) {")
        body += ["  pgrad[(long long)blockIdx.x * NP + %d] += red[%d][0];" % (p, p) for p in range(np_)]
        body.append("}")
        # forward kernel (parameters at run time)
        fwd = self._decompose()
        fwd.append("if (!live) return;")
        for si, k in enumerate(self.plan.stores):
            fb = self._forward_body(k, [0] * dim, "s%d_" % si)
            fwd.append("{")
            fwd += ["  " + l for l in fb.lines]
            fwd.append("  FLD%d(out)[t] = s%d_%d;" % (self.fi[self.plan.mm[k].micro[-1].inst[1]], si,
                                                    self.plan.mm[k].micro[-1].inst[2]))
            fwd.append("}")
        desc = {"name": self.name, "real": R, "fields": [n for n, _ in self.fields], "params": self.params,
                "grid": self.N, "sleeve": s, "block": self.block, "deltas": [list(d) for d in deltas]}
        flds = "\n".join("#define FLD%d(p) ((%s*)((p) + (size_t)%d * NTOT * sizeof(%s)))" % (i, R, i, R)
                         for i in range(nf))
        lams = "\n".join("#define LAM%d(p) ((%s*)((p) + (size_t)%d * NTOT * sizeof(%s)))" % (i, R, i, R)
                         for i in range(nf))
        return ADJ_TEMPLATE % dict(
            name=self.name, R=R, nf=nf, np=np_, block=self.block,
            consts="\n".join("constexpr int N%d = %d;" % (k + 1, n) for k, n in enumerate(self.N)),
            ntot="*".join("(long long)%d" % n for n in self.N), flds=flds, lams=lams,
            desc=json.dumps(desc), fwd="\n".join("  " + l for l in fwd), adj="\n".join("  " + l for l in body))


ADJ_TEMPLATE = r'''// Generated by formura_amd: forward and adjoint steps of %(name)s. Do not edit.
#include "formura_hip_rt.h"
#include "formura_adjoint.h"

namespace {
typedef %(R)s real_t;
%(consts)s
constexpr long long NTOT = %(ntot)s;
constexpr int NF = %(nf)d;
constexpr int NP = %(np)d;
constexpr int BLOCK = %(block)d;
%(flds)s
%(lams)s
const char* const kDescription = R"FMRJSON(%(desc)s)FMRJSON";

__global__ void __launch_bounds__(BLOCK) k_forward(const char* __restrict__ st, char* __restrict__ out,
                                                   const double* __restrict__ P) {
%(fwd)s
}

__global__ void __launch_bounds__(BLOCK) k_adjoint(const char* __restrict__ st, const char* __restrict__ lam_next,
                                                   char* __restrict__ lam_prev, const double* __restrict__ P,
                                                   double* __restrict__ pgrad) {
%(adj)s
}
}  // namespace

extern "C" const char* Formura_AD_Describe(void) { return kDescription; }
extern "C" long long Formura_AD_State_Bytes(void) { return (long long)NF * NTOT * sizeof(real_t); }
extern "C" int Formura_AD_Param_Slots(void) { return (int)((NTOT + BLOCK - 1) / BLOCK); }

// nsteps forward steps from `in`; ping-pong between `out` and `tmp` so that the
// result lands in `out` (in may alias neither).
extern "C" void Formura_AD_Forward(const void* in, void* out, void* tmp, int nsteps, const double* params_dev,
                                   void* stream) {
  hipStream_t s = (hipStream_t)stream;
  const unsigned nb = (unsigned)((NTOT + BLOCK - 1) / BLOCK);
  const char* src = (const char*)in;
  for (int t = 0; t < nsteps; ++t) {
    char* dst = ((nsteps - 1 - t) %% 2 == 0) ? (char*)out : (char*)tmp;
    hipLaunchKernelGGL(k_forward, dim3(nb), dim3(BLOCK), 0, s, src, dst, params_dev);
    src = dst;
  }
  if (nsteps == 0 && in != out)
    FMR_CHECK(hipMemcpyAsync(out, in, (size_t)NF * NTOT * sizeof(real_t), hipMemcpyDeviceToDevice, s));
  FMR_CHECK(hipGetLastError());
}

// one adjoint step: state_prev is the input of the step, lam_next = dL/d(output);
// writes lam_prev = dL/d(input) and adds dL/dparams into pgrad[block][param].
extern "C" void Formura_AD_Adjoint(const void* state_prev, const void* lam_next, void* lam_prev,
                                   const double* params_dev, double* pgrad, void* stream) {
  hipStream_t s = (hipStream_t)stream;
  const unsigned nb = (unsigned)((NTOT + BLOCK - 1) / BLOCK);
  hipLaunchKernelGGL(k_adjoint, dim3(nb), dim3(BLOCK), 0, s, (const char*)state_prev, (const char*)lam_next,
                     (char*)lam_prev, params_dev, pgrad);
  FMR_CHECK(hipGetLastError());
}
'''
