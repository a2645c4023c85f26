"""Code emission: the drop-in C header and the HIP library of one program.

The header reproduces the layout the reference emits (``Encode.hs:11-27``,
``Templates.hs:23-116``): the ``#define``s, ``Formura_Grid_Struct`` (fields in
``Data.Map`` key order, Templates.hs:53-73), ``Formura_Navi`` (field order of
``mkNavi``, Templates.hs:81-115), ``extern Formura_Grid_Struct formura_data``
and the prototypes of ``to_pos_*``, ``Formura_Init``, ``Formura_Forward`` and
``Formura_Finalize``. A user ``main.c`` written against the reference compiles
and links against the HIP library unchanged.

The library replaces the generated OpenMP/MPI C:

* ``Formura_Setup`` (the init graph) runs on the host, as a loop nest built the
  way ``mkKernel`` builds it (Templates.hs:426-487), and the state is uploaded;
* each ``Formura_Forward`` advances ``temporal_blocking_interval`` (or one) steps
  on the GPU.  The halo of the reference (``buff[j] = data[(j-2s) mod N]``,
  Templates.hs:336-346 / Functions.hs:265-290) becomes periodic wrap-around
  addressing in HBM, and the output keeps the reference's drifting frame:
  output index ``i`` of a step holds position ``i + off(range) - 2s`` of the
  input, and ``offset_a`` decreases by ``s`` per step (Templates.hs:328,346).
* a manifest temporary ``formura_mnX`` (Templates.hs:442-457) is evaluated over
  the whole periodic domain and indexed by input position, which gives every
  consumer (``LoadCursor``, :474) the same value the reference computes.

The per-point arithmetic is the micro-instruction DAG of the manifest node in
micro-id order, with the C types of Templates.hs:459-467, and no reassociation:
HIP code is compiled with ``-ffp-contract=off``.  Sub-expressions whose inputs
are all literals are folded on the host in IEEE double (the reference's C
evaluates them at run time with the same correctly rounded operations; ``pow``
and ``exp`` use the host libm, as the reference's C does).
"""
from __future__ import annotations

import json
import math
import re
import struct
from fractions import Fraction
from typing import Callable, Dict, List, Optional

from . import syntax as S
from .manifest import MMNode, MMProgram, MMRange, calc_range


class EmitError(Exception):
    pass


CTYPE = {"Rational": "double", "double": "double", "float": "float", "int": "int",
         "bool": "bool", "void": "void", "real": "double", "Real": "double"}


def c_elem_type(t: S.Type) -> str:
    """``mapType`` of the state struct (Templates.hs:63-70)."""
    if isinstance(t, S.GridType):
        return c_elem_type(t.elem)
    if isinstance(t, S.ElemType):
        if t.name == "Rational":
            return "double"
        if t.name in ("int", "float", "double"):
            return t.name
        return t.name
    raise EmitError("Unable to translate type to C: %r" % (t,))


def show_double(x: float) -> str:
    """A C literal that denotes exactly the double ``x``."""
    if math.isnan(x):
        return "(0.0/0.0)"
    if math.isinf(x):
        return "(1.0/0.0)" if x > 0 else "(-1.0/0.0)"
    r = repr(float(x))
    if "e" not in r and "." not in r and "n" not in r:
        r += ".0"
    return r


def _f32(x: float) -> float:
    return struct.unpack("f", struct.pack("f", x))[0] if math.isfinite(x) else x


# --- constant folding in IEEE double -------------------------------------------

def _div(a, b):
    try:
        return a / b
    except ZeroDivisionError:
        if a != a or a == 0:
            return float("nan")
        return math.copysign(float("inf"), a) * math.copysign(1.0, b)


def _pow(a, b):
    try:
        return math.pow(a, b)
    except (OverflowError, ValueError):
        return float("nan") if a < 0 else float("inf")


_EXTERN_FOLD = {"sqrt": math.sqrt, "exp": math.exp, "fabs": math.fabs, "sin": math.sin,
                "cos": math.cos, "tan": math.tan, "log": math.log, "atan": math.atan,
                "floor": math.floor, "ceil": math.ceil, "tanh": math.tanh}


def _fold(m, vals, typ):
    """Fold one micro instruction whose operands are constants; None if not foldable."""
    k = m.inst[0]
    try:
        if k == "Imm":
            v = float(Fraction(m.inst[1]))
        elif k == "Uniop":
            a = vals[m.inst[2]]
            op = m.inst[1]
            if op == "-":
                v = -a
            elif op == "+":
                v = a
            elif op.startswith("external-call/"):
                f = _EXTERN_FOLD.get(op.split("/", 1)[1])
                if f is None:
                    return None
                v = float(f(float(a)))
            else:
                return None
        elif k == "Binop":
            a, b = vals[m.inst[2]], vals[m.inst[3]]
            op = m.inst[1]
            if isinstance(a, bool) or isinstance(b, bool):
                if op not in ("&&", "||", "==", "!="):
                    a, b = float(a), float(b)
            if op == "+":
                v = a + b
            elif op == "-":
                v = a - b
            elif op == "*":
                v = a * b
            elif op == "/":
                v = _div(a, b)
            elif op == "**":
                v = _pow(float(a), float(b))
            elif op == "<":
                v = a < b
            elif op == "<=":
                v = a <= b
            elif op == ">":
                v = a > b
            elif op == ">=":
                v = a >= b
            elif op == "==":
                v = a == b
            elif op == "!=":
                v = a != b
            elif op == "&&":
                v = bool(a) and bool(b)
            elif op == "||":
                v = bool(a) or bool(b)
            else:
                return None
        elif k == "Triop":
            c, a, b = vals[m.inst[2]], vals[m.inst[3]], vals[m.inst[4]]
            v = a if c else b
        else:
            return None
    except (ValueError, OverflowError):
        return None
    if typ == "bool":
        return bool(v)
    if typ == "float":
        return _f32(float(v))
    if typ in ("double", "Rational", "real", "Real"):
        return float(v)
    return None


def _divc_one_step(b: float) -> bool:
    """``fmr_divc1`` applies: |b*RN(1/b) - 1| <= 2^-54, exactly (b > 0)."""
    y = 1.0 / b
    return abs(Fraction(b) * Fraction(y) - 1) <= Fraction(1, 2 ** 54)


def _divc_ok(b) -> bool:
    """A constant divisor that ``fmr_divc`` handles: finite, 2^-100 <= |b| <= 2^100."""
    if isinstance(b, bool):
        return False
    b = float(b)
    return math.isfinite(b) and 2.0 ** -100 <= abs(b) <= 2.0 ** 100


# values whose every integer combination lands in fmr_divc's range or on zero
LEVEL_LO = 2.0 ** -848


def level_division_bound(node) -> "Optional[float]":
    """For the per-level range test of the tile kernels (``div`` "level"): the
    largest XMAX such that, if every value the node loads is 0 or a finite x
    with 2^-848 <= |x| <= XMAX, every dividend of its FP64 divisions by
    constants (fmr_divc / fmr_divc1) is 0 or in [2^-900, 2^900] -- or None if
    a dividend is not an integer combination of loads and literals.

    Why: a finite double with |x| >= 2^-848 is an integer multiple of 2^-900
    (its ulp is at least 2^(-848-52)); so are literals of that size, their
    integer multiples, and every rounded sum or difference of such multiples
    (the exact result is a multiple, and rounding it to a coarser grid keeps
    it one); a nonzero one has magnitude >= 2^-900.  A dividend sum_i c_i x_i
    + sum_j k_j with integer c_i is bounded by sum |c_i| XMAX + sum |k_j| (to
    rounding of the partial sums, covered by halving XMAX)."""
    vals: Dict[int, object] = {}
    for j, m in enumerate(node.micro):
        ops = _operands(m)
        if m.inst[0] == "Imm" or (ops and all(o in vals for o in ops)):
            v = _fold(m, vals, m.typ)
            if v is not None:
                vals[j] = v
    memo: Dict[int, object] = {}

    def walk(j):
        """(sum |c_i| over loads, sum |k_j| over literals) of micro j, or None."""
        if j in memo:
            return memo[j]
        m = node.micro[j]
        k = m.inst[0]
        r = None
        if j in vals:
            v = vals[j]
            if isinstance(v, float) and math.isfinite(v) and (v == 0.0 or abs(v) >= LEVEL_LO):
                r = (0.0, abs(v))
        elif k == "LoadCursorStatic":
            r = (1.0, 0.0)
        elif k == "Uniop" and m.inst[1] in ("-", "+"):
            r = walk(m.inst[2])
        elif k == "Binop" and m.inst[1] in ("+", "-"):
            a, b = walk(m.inst[2]), walk(m.inst[3])
            r = (a[0] + b[0], a[1] + b[1]) if a and b else None
        elif k == "Binop" and m.inst[1] == "*":
            for c, x in ((m.inst[2], m.inst[3]), (m.inst[3], m.inst[2])):
                cv = vals.get(c)
                if isinstance(cv, float) and cv == int(cv) and cv != 0.0 and abs(cv) <= 2.0 ** 20 and x not in vals:
                    t = walk(x)
                    r = (abs(cv) * t[0], abs(cv) * t[1]) if t else None
                    break
        memo[j] = r
        return r

    bound = None
    for j, m in enumerate(node.micro):
        if (m.inst[0] == "Binop" and m.inst[1] == "/" and m.inst[3] in vals and m.inst[2] not in vals
                and CTYPE.get(m.typ, m.typ) == "double" and _divc_ok(vals[m.inst[3]])):
            mant, _ = math.frexp(abs(float(vals[m.inst[3]])))
            if mant == 0.5:
                continue  # a power-of-two divisor is an exact product
            w = walk(m.inst[2])
            if w is None or w[1] >= 2.0 ** 899:
                return None
            if w[0] == 0.0:
                continue
            xmax = 2.0 ** math.floor(math.log2((2.0 ** 899 - w[1]) / w[0]))
            bound = xmax if bound is None else min(bound, xmax)
    return bound if bound is not None else 2.0 ** 896


def redo_lower_bound(b: float) -> float:
    """Smallest |a| the scaled redo (fmr_divc_redo) divides with the fused
    sequence for the divisor magnitude b: 2^ceil(log2 b) * 2^-1021 >= b 2^-1021,
    so that |a/b| >= 2^-1021 (a normal quotient)."""
    m, e = math.frexp(abs(b))
    return math.ldexp(1.0, (e - 1 if m == 0.5 else e) - 1021)


def _divfix_lines(ct, name, src, sg, fn, b, y, cast):
    """A division by a constant as the branch-free fused sequence, computed
    again where a lane of the wave holds a dividend outside its range (a
    wave-uniform branch).  Two corrections (fmr_divc): tiny dividends through
    the sequence scaled by 2^600, the rest with the division (fmr_divc_redo);
    one correction (fmr_divc1, remainder at scale 2^120, exact down to the
    subnormals): only dividends above 2^900, Inf and NaN, with the division."""
    if fn == "fmr_divc1":
        return ["%s %s = %s(%sfmr_divc1_fast(%s, %s, %s));" % (ct, name, cast, sg, src, b, y),
                "if (__builtin_expect(fmr_divc1_out_of_range(%s) != 0ull, 0)) %s = %s(%sfmr_divc1(%s, %s, %s));"
                % (src, name, cast, sg, src, b, y)]
    lo = show_double(redo_lower_bound(float(b)))
    return ["%s %s = %s(%s%s_fast(%s, %s, %s));" % (ct, name, cast, sg, fn, src, b, y),
            "if (__builtin_expect(fmr_divc_out_of_range(%s) != 0ull, 0)) %s = %s(%s%s_redo(%s, %s, %s, %s));"
            % (src, name, cast, sg, fn, src, b, y, lo)]


def _const_literal(v, typ):
    if typ == "bool":
        return "true" if v else "false"
    if typ == "float":
        return show_double(v) + "f" if math.isfinite(v) else "((float)%s)" % show_double(v)
    return show_double(v)


class Body:
    """C statements for one manifest node's micro DAG."""

    def __init__(self, node: MMNode, load: Callable, tmp_load: Callable, index: Callable,
                 store: Callable, prefix: str, fold: bool = True, const_kw="const ",
                 params: Dict[str, str] = None, real: str = None, recip: bool = False,
                 prec: Dict[int, str] = None, divc: bool = True, spec_ok: str = None, divfix: bool = False,
                 divfast: bool = False, fastdiv: bool = False):
        """``params``: literal names bound to run-time expressions (not folded);
        ``real``: if set ("float"), every floating micro value is computed in
        that type (the FP32 form; constants are folded in double and rounded,
        and a division by a constant becomes a multiplication by its reciprocal);
        ``recip``: divide by constants through their reciprocal in double too
        (not bit-exact: for derivative code, never for a step's own values);
        ``prec``: per-micro C type of the floating values ("double"/"float",
        from ``mixed_precision``); operands of another precision are converted,
        FP32 values divide by constants through the reciprocal, and every
        FP64 square-root argument is checked to be positive (FMR_POSITIVE);
        ``divc``: an FP64 division by a constant is emitted as ``fmr_divc``
        (device code: the correctly rounded quotient from the host-rounded
        reciprocal and two fused corrections) or, for a power-of-two
        divisor, as the exact product with its reciprocal; host code keeps
        the division (with one correction instead of two where the divisor
        allows it: ``fmr_divc1``);
        ``spec_ok``: name of a wave mask the caller declared zero: such
        divisions use the branch-free forms (``fmr_divcs``/``fmr_divc1s``),
        which set the bits of lanes whose dividend lies outside their range --
        the caller then evaluates the node again without ``spec_ok``;
        ``divfix``: such divisions use the branch-free sequence and, where a
        lane of the wave holds a dividend outside its range, are computed
        again with the guarded form (a branch on a wave-uniform value, the
        division alone redone);
        ``divfast``: such divisions use the branch-free sequence with no test
        (the caller has established that every dividend is in range:
        ``level_division_bound``);
        ``fastdiv``: an FP64 division by a computed value is
        ``fmr_div_contract`` (reciprocal approximation and Newton steps; the
        opt-in contract form only, not IEEE)."""
        self.lines: List[str] = []
        sfx = (lambda fn: fn + "s") if spec_ok else (lambda fn: fn)
        sarg = (", " + spec_ok) if spec_ok else ""
        self.spec_used = False
        vals: Dict[int, object] = {}
        params = params or {}
        ctypes: Dict[int, str] = {}
        for j, m in enumerate(node.micro):
            typ = m.typ
            ct = CTYPE.get(typ, typ)
            if real and ct == "double":
                ct = real
            if prec and ct == "double" and j in prec:
                ct = prec[j]
            ctypes[j] = ct
            name = "%s%d" % (prefix, j)
            k = m.inst[0]
            if k == "Store":
                self.lines.append(store(m.inst[1], "%s%d" % (prefix, m.inst[2])))
                continue
            if m.src in params and k not in ("LoadCursorStatic", "LoadCursor", "LoadIndex"):
                self.lines.append("%s%s %s = %s;" % (const_kw, ct, name, params[m.src]))
                continue
            if (real == "float" and k == "Binop" and m.inst[1] == "-" and m.inst[2] not in vals
                    and node.micro[m.inst[2]].inst[0] == "Binop" and node.micro[m.inst[2]].inst[1] == "+"
                    and node.micro[m.inst[3]].inst[0] == "Binop" and node.micro[m.inst[3]].inst[1] == "*"
                    and vals.get(node.micro[m.inst[3]].inst[2]) == 2.0):
                # FP32 form: (a + b) - 2*c evaluated as (a - c) + (b - c).  For a, b
                # close to c the differences are exact (Sterbenz), so a discrete
                # second difference loses no digits to cancellation.
                pa, pb = node.micro[m.inst[2]].inst[2], node.micro[m.inst[2]].inst[3]
                pc = node.micro[m.inst[3]].inst[3]
                self.lines.append("%s%s %s = (%s%d - %s%d) + (%s%d - %s%d);"
                                  % (const_kw, ct, name, prefix, pa, prefix, pc, prefix, pb, prefix, pc))
                continue
            if (ct == "float" and not (real or recip or prec) and k == "Binop" and m.inst[1] == "/"
                    and m.inst[3] in vals and m.inst[2] not in vals and divc and _divc_ok(vals[m.inst[3]])):
                # a float divided by a constant: the reference's C divides in
                # double (the literal is a double, Templates.hs:475) and rounds
                # to float on assignment; that double quotient (fmr_divc, the
                # correctly rounded one) rounded to float is the correctly
                # rounded float quotient (53 >= 2*24 + 2: no double rounding)
                b = float(vals[m.inst[3]])
                if ctypes.get(m.inst[3]) == "float":
                    b = float(_f32(b))  # a float constant holds its value rounded to float
                src = "(double)%s%d" % (prefix, m.inst[2])
                mant, _ = math.frexp(abs(b))
                if mant == 0.5:
                    rhs = "%s*%s" % (src, show_double(1.0 / b))
                else:
                    fn = sfx("fmr_divc1" if _divc_one_step(abs(b)) else "fmr_divc")
                    self.spec_used |= bool(spec_ok)
                    sg = "" if b > 0 else "-"
                    rhs = "%s%s(%s, %s, %s%s)" % (sg, fn, src, show_double(abs(b)), show_double(1.0 / abs(b)), sarg)
                    if divfix:
                        self.lines += _divfix_lines(ct, name, src, sg, fn, show_double(abs(b)),
                                                    show_double(1.0 / abs(b)), "(float)")
                        continue
                self.lines.append("%s%s %s = (float)(%s);" % (const_kw, ct, name, rhs))
                continue
            if ((real or recip or (ct == "float" and prec)) and k == "Binop" and m.inst[1] == "/" and m.inst[3] in vals
                    and m.inst[2] not in vals):
                rc = 1.0 / vals[m.inst[3]] if vals[m.inst[3]] else float("inf")
                lit = _const_literal(_f32(rc), "float") if ct == "float" else show_double(rc)
                src = "%s%d" % (prefix, m.inst[2])
                if prec and ctypes.get(m.inst[2]) in ("float", "double") and ctypes[m.inst[2]] != ct:
                    src = "(%s)%s" % (ct, src)
                self.lines.append("%s%s %s = %s*%s;" % (const_kw, ct, name, src, lit))
                continue
            if (divc and ct == "double" and k == "Binop" and m.inst[1] == "/" and m.inst[3] in vals
                    and m.inst[2] not in vals and _divc_ok(vals[m.inst[3]])):
                b = float(vals[m.inst[3]])
                src = "%s%d" % (prefix, m.inst[2])
                if prec and ctypes.get(m.inst[2]) == "float":
                    src = "(double)%s" % src
                mant, _ = math.frexp(abs(b))
                if mant == 0.5:
                    # 1/b is exact: a*(1/b) and a/b are the same real number, rounded once
                    rhs = "%s*%s" % (src, show_double(1.0 / b))
                else:
                    fn = sfx("fmr_divc1" if _divc_one_step(abs(b)) else "fmr_divc")
                    if divfast:
                        fn += "_fast"
                    self.spec_used |= bool(spec_ok)
                    if b > 0:
                        rhs = "%s(%s, %s, %s%s)" % (fn, src, show_double(b), show_double(1.0 / b), sarg)
                    else:
                        rhs = "-%s(%s, %s, %s%s)" % (fn, src, show_double(-b), show_double(-1.0 / b), sarg)
                    if divfix and not divfast:
                        self.lines += _divfix_lines(ct, name, src, "" if b > 0 else "-", fn, show_double(abs(b)),
                                                    show_double(1.0 / abs(b)), "")
                        continue
                self.lines.append("%s%s %s = %s;" % (const_kw, ct, name, rhs))
                continue
            if fold:
                ops = _operands(m)
                if all(o in vals for o in ops):
                    v = _fold(m, vals, typ)
                    if v is not None:
                        vals[j] = v
                        f32 = (real == "float" or ct == "float") and isinstance(v, float)
                        lit = _const_literal(_f32(v) if f32 else v, "float" if f32 else typ)
                        self.lines.append("%s%s %s = %s;" % (const_kw, ct, name, lit))
                        continue
            def a(i, ct=ct, k=k):
                v = "%s%d" % (prefix, i)
                if prec and ct in ("float", "double") and ctypes.get(i) in ("float", "double") and ctypes[i] != ct:
                    return "(%s)%s" % (ct, v)
                return v
            if prec and k == "Binop" and m.inst[1] in ("<", "<=", ">", ">=", "==", "!="):
                # a comparison is evaluated in the wider type of its operands
                wide = "double" if "double" in (ctypes.get(m.inst[2]), ctypes.get(m.inst[3])) else ct
                a = lambda i, wide=wide: ("(%s)%s%d" % (wide, prefix, i)
                                          if ctypes.get(i) in ("float", "double") and ctypes[i] != wide
                                          else "%s%d" % (prefix, i))
            if k == "Imm":
                rhs = show_double(float(Fraction(m.inst[1])))
                if real == "float":
                    rhs = _const_literal(_f32(float(Fraction(m.inst[1]))), "float")
            elif k == "LoadCursorStatic":
                rhs = load(m.inst[2], m.inst[1])
            elif k == "LoadCursor":
                rhs = tmp_load(m.inst[2], m.inst[1])
            elif k == "LoadIndex":
                rhs = index(m.inst[1])
                if len(m.inst) > 2 and m.inst[2]:
                    # an index read of an inlined temporary: its nest's loop index
                    # differs from this one's by a constant (manifest.inline_temporaries)
                    rhs = "(%s%+d)" % (rhs, m.inst[2])
            elif k == "Uniop":
                op = m.inst[1]
                if op.startswith("external-call/"):
                    rhs = "%s((%s)%s)" % (op.split("/", 1)[1], ct if prec else (real or "double"), a(m.inst[2]))
                    if prec and op == "external-call/sqrt" and ct == "double":
                        self.lines.append("FMR_POSITIVE(%s);" % a(m.inst[2]))
                else:
                    rhs = "%s%s" % (op, a(m.inst[2]))
            elif k == "Binop":
                op = m.inst[1]
                if (fastdiv and op == "/" and ct == "double" and m.inst[3] not in vals
                        and ctypes.get(m.inst[2]) == "double" and ctypes.get(m.inst[3]) == "double"):
                    self.lines.append("%s%s %s = fmr_div_contract(%s, %s);" % (const_kw, ct, name, a(m.inst[2]),
                                                                             a(m.inst[3])))
                    continue
                if op == "**":
                    pt = ct if prec else (real or "double")
                    rhs = "pow((%s)%s,(%s)%s)" % (pt, a(m.inst[2]), pt, a(m.inst[3]))
                else:
                    rhs = "%s%s%s" % (a(m.inst[2]), op, a(m.inst[3]))
            elif k == "Triop":
                rhs = "%s%d?%s:%s" % (prefix, m.inst[2], a(m.inst[3]), a(m.inst[4]))
            else:
                raise EmitError("Unimplemented for keyword: %s" % k)
            self.lines.append("%s%s %s = %s;" % (const_kw, ct, name, rhs))
        self.result = "%s%d" % (prefix, len(node.micro) - 1)
        self.vals = vals


COMPARISONS = ("<", "<=", ">", ">=", "==", "!=")


def mixed_precision(node: MMNode) -> Dict[int, str]:
    """Precision of each floating micro value of a node in the mixed form.

    FP64: (1) every value that decides a branch or feeds a square root, with
    all its operands -- for an approximate Riemann solver the wave-speed
    estimates and the pressures and densities they come from; (2) the loads of
    the (FP64) state; (3) the update chain: from the stored value down through
    + - * / and unary minus, so that q' = q - dt*(F_i - F_{i-1})/dx is summed in
    FP64 (flux differences telescope, totals are conserved to FP64 rounding).
    FP32: everything else (the flux arithmetic)."""
    micro = node.micro
    fp64 = [False] * len(micro)

    def mark(i):
        stack = [i]
        while stack:
            j = stack.pop()
            if j is None or fp64[j]:
                continue
            fp64[j] = True
            stack += [o for o in _operands(micro[j]) if o is not None]

    for j, m in enumerate(micro):
        k = m.inst[0]
        if k == "Binop" and m.inst[1] in COMPARISONS:
            mark(j)
        elif k == "Triop":
            mark(m.inst[2])
        elif k == "Uniop" and m.inst[1] == "external-call/sqrt":
            mark(j)
        elif k in ("LoadCursorStatic", "LoadCursor"):
            fp64[j] = True
    # constants (and values computed from constants only) keep their exact FP64
    # value; FP32 consumers convert them (at compile time)
    const = [False] * len(micro)
    for j, m in enumerate(micro):
        ops = _operands(m)
        const[j] = m.inst[0] == "Imm" or (m.inst[0] in ("Uniop", "Binop", "Triop") and all(
            o is not None and const[o] for o in ops))
        if const[j]:
            fp64[j] = True
    st = micro[-1]
    if st.inst[0] == "Store":
        stack = [st.inst[2]]
        while stack:
            j = stack.pop()
            m = micro[j]
            if (m.inst[0] == "Binop" and m.inst[1] in ("+", "-", "*", "/")) or (m.inst[0] == "Uniop" and m.inst[1] == "-"):
                fp64[j] = True
                stack += _operands(m)
    return {j: ("double" if fp64[j] else "float") for j in range(len(micro))}


def mixed_precision_graph(mm: Dict[int, MMNode]):
    """The mixed form over a step graph whose manifest temporaries stay
    materialized (one buffer per temporary, as the reference's
    ``formura_mnX``).  Returns (precision of every micro value per node,
    storage type per temporary).

    The rules of ``mixed_precision`` hold across nodes: a value that decides a
    branch or feeds a square root is FP64 with all its operands, and when such
    an operand is a temporary read (LoadCursor) the temporary's own result is
    FP64 with all its operands, in its node, and so on; state loads are FP64;
    the update chain of a stored value (+ - * / and unary minus) is FP64 down
    to the reads it reaches, which are converted on load.  Every other value,
    and every temporary whose result is not FP64, is FP32: an approximate
    Riemann solver keeps its wave speeds and the densities and pressures they
    come from in FP64 and computes and stores its fluxes in FP32.  A face flux
    stays one stored value that both neighbouring cells subtract, so totals are
    still conserved to FP64 rounding."""
    fp64 = {k: [False] * len(n.micro) for k, n in mm.items()}
    closed = {k: [False] * len(n.micro) for k, n in mm.items()}
    work = []

    def run():
        while work:
            k, j = work.pop()
            if j is None or closed[k][j]:
                continue
            closed[k][j] = True
            fp64[k][j] = True
            m = mm[k].micro[j]
            if m.inst[0] == "LoadCursor" and m.inst[2] in mm:
                t = m.inst[2]
                work.append((t, len(mm[t].micro) - 1))
            work.extend((k, o) for o in _operands(m) if o is not None)

    for k, node in mm.items():
        micro = node.micro
        const = [False] * len(micro)
        for j, m in enumerate(micro):
            kind = m.inst[0]
            if kind == "Binop" and m.inst[1] in COMPARISONS:
                work.append((k, j))
            elif kind == "Triop":
                work.append((k, m.inst[2]))
            elif kind == "Uniop" and m.inst[1] == "external-call/sqrt":
                work.append((k, j))
            elif kind == "LoadCursorStatic":
                fp64[k][j] = True
            ops = _operands(m)
            const[j] = kind == "Imm" or (kind in ("Uniop", "Binop", "Triop") and all(
                o is not None and const[o] for o in ops))
            if const[j]:
                fp64[k][j] = True
        st = micro[-1]
        if st.inst[0] == "Store":
            stack = [st.inst[2]]
            while stack:
                j = stack.pop()
                m = micro[j]
                if (m.inst[0] == "Binop" and m.inst[1] in ("+", "-", "*", "/")) or (m.inst[0] == "Uniop" and m.inst[1] == "-"):
                    fp64[k][j] = True
                    stack += _operands(m)
                elif m.inst[0] == "LoadCursor":
                    fp64[k][j] = True
    run()
    prec = {k: {j: ("double" if v else "float") for j, v in enumerate(f)} for k, f in fp64.items()}
    storage = {k: ("double" if fp64[k][-1] else "float") for k, n in mm.items() if n.micro[-1].inst[0] != "Store"}
    return prec, storage


def _operands(m):
    k = m.inst[0]
    if k == "Uniop":
        return [m.inst[2]]
    if k == "Binop":
        return [m.inst[2], m.inst[3]]
    if k == "Triop":
        return [m.inst[2], m.inst[3], m.inst[4]]
    if k == "Imm":
        return []
    return [None]  # loads are never constant


# --- the drop-in header ---------------------------------------------------------

def navi_fields(om) -> List[tuple]:
    """``mkNavi`` without MPI (Templates.hs:81-115): (name, ctype, init expression)."""
    cfg = om.config
    axes = om.axes
    out = [("time_step", "int", "0")]
    out += [("lower_" + a, "int", "0") for a in axes]
    out += [("upper_" + a, "int", str(n)) for a, n in zip(axes, cfg.grid_per_node)]
    out += [("space_interval_" + a, "double", show_double(d)) for a, d in zip(axes, cfg.space_interval)]
    out += [("my_rank", "int", "0")]
    out += [("offset_" + a, "int", "0") for a in axes]
    out += [("length_" + a, "double", show_double(float(l))) for a, l in zip(axes, cfg.length_per_node)]
    out += [("total_grid_" + a, "int", str(n)) for a, n in zip(axes, cfg.grid_per_node)]
    return out


def state_fields(om) -> List[tuple]:
    return [(name, c_elem_type(t)) for name, t in om.state_signature.items()]


def emit_header(om, with_ext: bool = True) -> str:
    cfg = om.config
    if cfg.mpi_shape is not None:
        raise EmitError("mpi_shape is set: the HIP backend runs one rank on one GPU; omit mpi_shape "
                        "(the periodic halo is wrapped in HBM)")
    dims = "".join("[%d]" % n for n in cfg.grid_per_node)
    L = ["#pragma once",
         "#include <stdio.h>", "#include <stdlib.h>", "#include <stdbool.h>", "#include <math.h>",
         "#define Ns %d" % cfg.sleeve]
    L += ["#define L%d %d" % (i + 1, n) for i, n in enumerate(cfg.grid_per_node)]
    L.append("typedef struct {")
    L += ["%s %s%s;" % (ct, name, dims) for name, ct in state_fields(om)]
    L.append("} %s;" % om.grid_struct_type_name)
    L.append("typedef struct {")
    L += ["%s %s;" % (ct, name) for name, ct, _ in navi_fields(om)]
    L.append("} Formura_Navi;")
    L.append("extern %s %s;" % (om.grid_struct_type_name, om.grid_struct_instance_name))
    L += ["double to_pos_%s(int,Formura_Navi);" % a for a in om.axes]
    L.append("void Formura_Init(int *,char ***,Formura_Navi *);")
    L.append("void Formura_Forward(Formura_Navi *);")
    L.append("void Formura_Finalize();")
    if with_ext:
        L.append("#include \"formura_hip.h\"")
    return "\n".join(L) + "\n"


# --- the HIP library -------------------------------------------------------------

class GraphPlan:
    """What one MM graph (step, filter or first_step) needs on the GPU."""

    def __init__(self, mm: Dict[int, MMNode], sleeve: int, dim: int, tag: str, storage: Dict[int, str] = None):
        self.mm = mm
        # storage type of each manifest temporary (mixed form), else its C type
        self.storage = storage
        self.sleeve = sleeve
        self.tag = tag
        self.rng = calc_range(mm)
        self.temps = [k for k in sorted(mm) if not _is_void(mm[k].typ)]
        self.stores = [k for k in sorted(mm) if _is_void(mm[k].typ)]
        for k in self.stores:
            if self.rng[k].size != 2 * sleeve:
                raise EmitError("manifest node %s of %s has stencil range [%d,%d], which the reference only "
                                "evaluates consistently when it spans 2*sleeve=%d"
                                % (mm[k].name, tag, self.rng[k].lower, self.rng[k].upper, 2 * sleeve))
        for k in self.temps:
            if any(m.inst[0] == "LoadIndex" for m in mm[k].micro):
                raise EmitError("index access inside the manifest temporary %s is not supported" % mm[k].name)
            if not isinstance(mm[k].typ, S.GridType):
                raise EmitError("manifest temporary %s is not a grid" % mm[k].name)
        self.max_reach = 0
        for k, node in mm.items():
            base = self.rng[k].offset - 2 * sleeve if k in self.stores else 0
            for m in node.micro:
                if m.inst[0] in ("LoadCursorStatic", "LoadCursor"):
                    for c in m.inst[1]:
                        self.max_reach = max(self.max_reach, abs(c + base))

    def temp_ctype(self, k):
        if self.storage and k in self.storage and c_elem_type(self.mm[k].typ) == "double":
            return self.storage[k]
        return c_elem_type(self.mm[k].typ)


def _is_void(t):
    return isinstance(t, S.ElemType) and t.name == "void"


def lib_description(om, name: str, schedule: str) -> dict:
    cfg = om.config
    return {
        "name": name,
        "dimension": om.dimension,
        "axes": om.axes,
        "grid": cfg.grid_per_node,
        "sleeve": cfg.sleeve,
        "steps_per_forward": cfg.steps_per_forward,
        "fields": [{"name": n, "ctype": t} for n, t in state_fields(om)],
        "navi": [{"name": n, "ctype": t} for n, t, _ in navi_fields(om)],
        "filter_interval": cfg.filter_interval,
        "schedule": schedule,
    }


# a point function's LDS load statement as Body emits it ("const T name = ldsI[...];")
LDS_LOAD_LINE = re.compile(r"^const \w+ \w+ = lds\d+\[[^\]]*\];$")


def tile_load(count: str, index_lines: List[str], pairs: List[tuple], ctype: str, tag: str = "ld",
              threads: int = 256, indent: str = "  ", max_count: Optional[str] = None, check: str = None) -> str:
    """C statements loading a tile of ``count`` elements (a constant expression)
    into LDS with ``threads`` work-items: every global load of the work-item is
    issued (unrolled, into registers) before the first LDS write, so the tile
    costs one memory latency, not one per loop trip.  ``index_lines`` compute
    ``g`` from ``idx``; ``pairs`` are (lds_element, global_element) in terms of
    ``idx`` and ``g``; all are of C type ``ctype``.  ``max_count`` (a constant
    expression) bounds ``count`` when ``count`` is not itself constant.
    ``check``: a statement applied to each loaded value (``%s`` is the value)
    as it is written to LDS."""
    it = "%s_IT" % tag
    decl = ["%sconstexpr int %s = (%s + %d) / %d;" % (indent, it, max_count or count, threads - 1, threads)]
    decl.append("%s%s %s;" % (indent, ctype, ", ".join("%s_v%d[%s]" % (tag, i, it) for i in range(len(pairs)))))
    out = []
    out.append("%s#pragma unroll" % indent)
    out.append("%sfor (int it = 0; it < %s; ++it) {" % (indent, it))
    out.append("%s  const int idx = threadIdx.x + it * %d;" % (indent, threads))
    out.append("%s  if (idx < (%s)) {" % (indent, count))
    out += ["%s    %s" % (indent, l) for l in index_lines]
    out += ["%s    %s_v%d[it] = %s;" % (indent, tag, i, src) for i, (dst, src) in enumerate(pairs)]
    out.append("%s  }" % indent)
    out.append("%s}" % indent)
    issue, out = out, []
    out.append("%s#pragma unroll" % indent)
    out.append("%sfor (int it = 0; it < %s; ++it) {" % (indent, it))
    out.append("%s  const int idx = threadIdx.x + it * %d;" % (indent, threads))
    out.append("%s  if (idx < (%s)) {" % (indent, count))
    out += ["%s    %s = %s_v%d[it];" % (indent, dst, tag, i) for i, (dst, src) in enumerate(pairs)]
    if check:
        out += ["%s    %s" % (indent, check % ("%s_v%d[it]" % (tag, i))) for i in range(len(pairs))]
    out.append("%s  }" % indent)
    out.append("%s}" % indent)
    return "\n".join(decl + issue + out)
