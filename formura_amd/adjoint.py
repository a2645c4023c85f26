"""Differentiating a Formura run: torch.autograd over the generated forward and
adjoint steps (``include/formura_adjoint.h``), with binomial checkpointing.

``Stepper`` binds one adjoint library.  ``run(state0, params, nsteps,
snapshots)`` is a ``torch.autograd.Function``.  Its forward advances
``nsteps`` steps and keeps at most ``snapshots`` states in HBM: the
checkpoints of the schedule's first path.  Its backward
reverses the run with the binomial checkpointing schedule of Griewank and
Walther ("Algorithm 799: Revolve", ACM TOMS 26(1), 2000).  A segment of l
steps with s free snapshots (c = s+1 checkpoints counting the segment's first
state) and repetition number r (the least r with C(c+r, c) >= l) is split after
j steps, j chosen in the interval where both parts stay within r repetitions;
the tail is reversed with s-1 snapshots and the head with s.  Forward steps
of the whole schedule: r*l - C(c+r, r-1), the binomial optimum; the forward
sweep does the first path's share of them.

The schedule runs over units of ``Formura_AD_Levels()`` steps (4 for the 2-D
tiled kernels): checkpoints sit at unit boundaries, and one fused call
(``Formura_AD_Adjoint_Steps``) reverses a unit from the state at its start,
recomputing the unit's inner states into a scratch of levels-1 states.  For
2000 steps and 32 snapshots this turns 2000 steps at repetition number 3 into
500 units at repetition number 2.

A step that is linear in the state and has no parameters (the heat equation)
has an adjoint that reads no state: its reversal runs the adjoint steps alone,
with no checkpoints and no recomputation.

The result is dL/d(initial state) and dL/d(parameters).  The parameters are
the program's named literals listed by the library (for Gray-Scott: F, k, Du,
Dv as f, k, ru, rv), reduced in FP64.
"""
from __future__ import annotations

import ctypes
import json
import math
import os
from typing import List

import torch

from . import build as _build


def _binom(n, k):
    return math.comb(n, k) if 0 <= k <= n else 0


def repetitions(l: int, s: int) -> int:
    """Least r with C(c + r, c) >= l, c = s + 1 checkpoints (s free snapshots plus
    the one holding the segment's first state): the repetition number."""
    c = s + 1
    r = 0
    while _binom(c + r, c) < l:
        r += 1
    return r


def optimal_cost(l: int, s: int) -> int:
    """Minimal forward steps to reverse l steps with s free snapshots, the state
    at the start given: r*l - C(c+r, r-1) with c = s + 1 (Griewank and Walther 2000)."""
    if l <= 1:
        return 0
    c = s + 1
    r = repetitions(l, s)
    return r * l - _binom(c + r, r - 1)


_SPLIT = {}


def split(l: int, s: int) -> int:
    """Steps to advance before the next snapshot: the j in the admissible interval
    [l - C(c-1+r, c-1), C(c+r-1, c)] that minimises j + cost(l-j, s-1) + cost(j, s)."""
    key = (l, s)
    if key not in _SPLIT:
        c = s + 1
        r = repetitions(l, s)
        lo = max(1, l - _binom(c - 1 + r, c - 1))
        hi = max(lo, min(l - 1, _binom(c + r - 1, c)))
        best = None
        for j in range(lo, hi + 1):
            cost = j + optimal_cost(l - j, s - 1) + optimal_cost(j, s)
            if best is None or cost < best[0]:
                best = (cost, j)
        _SPLIT[key] = best[1]
    return _SPLIT[key]


def first_descent(l: int, s: int) -> List[int]:
    """Step counts between the checkpoints that the forward sweep keeps: the
    schedule's first path (advance split(l, s), keep the state, continue with
    s-1 snapshots) until one step or no snapshot is left."""
    js = []
    while l > 1 and s > 0:
        j = split(l, s)
        js.append(j)
        l -= j
        s -= 1
    return js


def forward_steps_needed(l: int, s: int, stored: bool = True) -> int:
    """Forward steps the reversal recomputes, following exactly the schedule of
    ``_Run.backward``.  With ``stored`` the forward sweep has kept the
    checkpoints of ``first_descent``, so those advances are not repeated."""
    total = 0
    stack = [(l, s, stored)]
    while stack:
        l, s, pre = stack.pop()
        while l > 1:
            if s == 0:
                total += l * (l - 1) // 2
                break
            j = split(l, s)
            if not pre:
                total += j
            stack.append((l - j, s - 1, pre))
            pre = False
            l = j
    return total


def unit_steps(nsteps: int, unit: int, a: int, b: int) -> int:
    """Steps from unit a to unit b when ``nsteps`` steps are cut into units of
    ``unit`` steps (the last unit may be shorter)."""
    return min(nsteps, b * unit) - min(nsteps, a * unit)


def recompute_steps(nsteps: int, s: int, unit: int = 1) -> int:
    """Forward steps the reversal of ``_Run.backward`` recomputes (the forward
    sweep's kept checkpoints excluded) when the schedule runs over units of
    ``unit`` steps; the inner states of a unit are recomputed inside the fused
    adjoint call and are not counted here."""
    l = -(-nsteps // unit) if nsteps > 0 else 0
    total = 0
    stack = [(0, l, s, True)]
    while stack:
        pos, l, s, pre = stack.pop()
        while l > 1:
            if s == 0:
                total += sum(unit_steps(nsteps, unit, pos, pos + j) for j in range(1, l))
                break
            j = split(l, s)
            if not pre:
                total += unit_steps(nsteps, unit, pos, pos + j)
            stack.append((pos + j, l - j, s - 1, pre))
            pre = False
            l = j
    return total


def reversal_ops(nsteps: int, unit: int, snapshots: int):
    """The reversal as a list of operations in issue order, over numbered state
    buffers: 0 holds the initial state, 1..len(first_descent) the states the
    forward sweep kept, the others are work buffers.  (0, a, b, k): buffer b =
    k forward steps from buffer a; (1, a, -1, k): the adjoint of the k steps
    that start at buffer a.  Returns (ops, number of buffers).  A checkpoint's
    buffer (a kept snapshot's too) is freed once the units after it are
    reversed, and freed buffers are reused oldest-freed first, so forward work
    can run ahead of the adjoint; buffer 0 (the caller's initial state) is never
    written."""
    import collections
    n = int(nsteps)
    nunits = -(-n // unit) if n > 0 else 0
    path = first_descent(nunits, snapshots)
    kept_ids = list(range(1, 1 + len(path)))
    ops = []
    first_work = 1 + len(kept_ids)
    # three work buffers to start with: with the buffers freed one run before,
    # a recomputation chain can run one launch ahead of the adjoints it feeds
    free = collections.deque([first_work, first_work + 1, first_work + 2])
    count = [first_work + 3]

    def alloc():
        if free:
            return free.popleft()
        count[0] += 1
        return count[0] - 1

    def advance(src, pos, j):
        dst = alloc()
        ops.append((0, src, dst, unit_steps(n, unit, pos, pos + j)))
        return dst

    def adjoint(state, pos):
        ops.append((1, state, -1, unit_steps(n, unit, pos, pos + 1)))

    def reverse(base, pos, l, s, kept):
        while True:
            if l == 1:
                adjoint(base, pos)
                return
            if s == 0:
                for j in range(l - 1, -1, -1):
                    if j:
                        t = advance(base, pos, j)
                        adjoint(t, pos + j)
                        free.append(t)
                    else:
                        adjoint(base, pos)
                return
            j = split(l, s)
            mid = kept.pop(0) if kept else advance(base, pos, j)
            reverse(mid, pos + j, l - j, s - 1, kept)
            kept = []
            free.append(mid)  # no later operation reads this checkpoint
            l = j

    if nunits > 0:
        reverse(0, 0, nunits, snapshots, kept_ids)
    return ops, count[0]


class Stepper:
    """One adjoint library: forward/adjoint steps over torch device tensors."""

    def __init__(self, name: str):
        path = _build.lib_path(name)
        if not os.path.exists(path):
            raise RuntimeError("adjoint library %s is missing at %s; build it with "
                               "`python -m formura_amd.build %s`" % (name, path, name))
        L = self.lib = ctypes.CDLL(path, mode=os.RTLD_LOCAL)
        L.Formura_AD_Describe.restype = ctypes.c_char_p
        self.desc = json.loads(L.Formura_AD_Describe().decode())
        L.Formura_AD_Param_Slots.restype = ctypes.c_int
        L.Formura_AD_Levels.restype = ctypes.c_int
        L.Formura_AD_Forward.argtypes = [ctypes.c_void_p] * 3 + [ctypes.c_int, ctypes.c_void_p, ctypes.c_void_p]
        L.Formura_AD_Adjoint.argtypes = [ctypes.c_void_p] * 6
        L.Formura_AD_Adjoint_Steps.argtypes = [ctypes.c_void_p] * 4 + [ctypes.c_int] + [ctypes.c_void_p] * 3
        L.Formura_AD_Inner_States.argtypes = [ctypes.c_void_p] * 2 + [ctypes.c_int] + [ctypes.c_void_p] * 2
        L.Formura_AD_Adjoint_Levels.argtypes = [ctypes.c_void_p] * 4 + [ctypes.c_int] + [ctypes.c_void_p] * 3
        L.Formura_AD_Reverse.restype = ctypes.c_int
        L.Formura_AD_Reverse.argtypes = ([ctypes.c_void_p, ctypes.c_int, ctypes.c_void_p, ctypes.c_int]
                                         + [ctypes.c_void_p] * 8)
        self.dtype = torch.float64 if self.desc["real"] == "double" else torch.float32
        self.shape = [len(self.desc["fields"])] + list(self.desc["grid"])
        self.nparams = len(self.desc["params"])
        self.slots = int(L.Formura_AD_Param_Slots())
        # most steps one fused adjoint call reverses
        self.levels = int(L.Formura_AD_Levels())
        # a step linear in the state with no parameters: its adjoint reads no state
        self.state_free = bool(self.desc.get("state_free"))

    def _stream(self):
        return ctypes.c_void_p(torch.cuda.current_stream().cuda_stream)

    def _check(self, t):
        if not (t.is_cuda and t.dtype == self.dtype and list(t.shape) == self.shape and t.is_contiguous()):
            raise ValueError("expected a contiguous %s CUDA tensor of shape %s" % (self.dtype, self.shape))

    def scratch(self, device):
        """Inner-state buffer of the fused adjoint (levels - 1 states)."""
        return torch.empty([max(1, self.levels - 1)] + self.shape, dtype=self.dtype, device=device)

    def forward(self, x, out, tmp, nsteps, params_dev):
        self._check(x)
        self._check(out)
        self._check(tmp)
        self.lib.Formura_AD_Forward(x.data_ptr(), out.data_ptr(), tmp.data_ptr(), int(nsteps),
                                    params_dev.data_ptr(), self._stream())

    def adjoint(self, state_prev, lam_next, lam_prev, params_dev, pgrad):
        for t in (state_prev, lam_next, lam_prev):
            self._check(t)
        self.lib.Formura_AD_Adjoint(state_prev.data_ptr(), lam_next.data_ptr(), lam_prev.data_ptr(),
                                    params_dev.data_ptr(), pgrad.data_ptr(), self._stream())

    def adjoint_steps(self, state0, scratch, lam_next, lam_prev, nsteps, params_dev, pgrad):
        for t in (state0, lam_next, lam_prev):
            self._check(t)
        if not 1 <= nsteps <= self.levels:
            raise ValueError("adjoint_steps reverses 1..%d steps, got %d" % (self.levels, nsteps))
        if scratch.dtype != self.dtype or scratch.numel() < (nsteps - 1) * state0.numel() or not scratch.is_contiguous():
            raise ValueError("scratch holds fewer than %d states" % (nsteps - 1))
        self.lib.Formura_AD_Adjoint_Steps(state0.data_ptr(), scratch.data_ptr(), lam_next.data_ptr(),
                                          lam_prev.data_ptr(), int(nsteps), params_dev.data_ptr(),
                                          pgrad.data_ptr(), self._stream())


    def _check_scratch(self, scratch, nlevels, like):
        if scratch.dtype != self.dtype or scratch.numel() < nlevels * like.numel() or not scratch.is_contiguous():
            raise ValueError("scratch holds fewer than %d states" % nlevels)

    def inner_states(self, state0, scratch, nlevels, params_dev):
        """The ``nlevels`` states after ``state0`` into ``scratch`` (first half of adjoint_steps)."""
        self._check(state0)
        if not 0 <= nlevels <= self.levels - 1:
            raise ValueError("inner_states computes 0..%d levels, got %d" % (self.levels - 1, nlevels))
        self._check_scratch(scratch, nlevels, state0)
        self.lib.Formura_AD_Inner_States(state0.data_ptr(), scratch.data_ptr(), int(nlevels), params_dev.data_ptr(),
                                         self._stream())

    def adjoint_levels(self, state0, scratch, lam_next, lam_prev, nsteps, params_dev, pgrad):
        """Second half of adjoint_steps: the adjoint over inner states already in ``scratch``."""
        for t in (state0, lam_next, lam_prev):
            self._check(t)
        if not 1 <= nsteps <= self.levels:
            raise ValueError("adjoint_levels reverses 1..%d steps, got %d" % (self.levels, nsteps))
        self._check_scratch(scratch, nsteps - 1, state0)
        self.lib.Formura_AD_Adjoint_Levels(state0.data_ptr(), scratch.data_ptr(), lam_next.data_ptr(),
                                           lam_prev.data_ptr(), int(nsteps), params_dev.data_ptr(),
                                           pgrad.data_ptr(), self._stream())


    def reverse(self, ops, bufs, tmp, scratch, lam, lam2, params_dev, pgrad, stream_f=None):
        """Run a ``reversal_ops`` schedule natively: forward work on ``stream_f``
        (None: torch's current stream, i.e. one stream), adjoint work on torch's
        current stream.  Returns the buffer (lam or lam2) holding the result."""
        for t in list(bufs) + [tmp, lam, lam2]:
            self._check(t)
        if self.levels > 1:
            self._check_scratch(scratch, 2 * (self.levels - 1), lam)
        flat = (ctypes.c_int * (4 * len(ops)))(*[int(v) for op in ops for v in op])
        ptrs = (ctypes.c_void_p * len(bufs))(*[t.data_ptr() for t in bufs])
        a = self._stream()
        f = ctypes.c_void_p(stream_f.cuda_stream) if stream_f is not None else a
        which = self.lib.Formura_AD_Reverse(flat, len(ops), ptrs, len(bufs), tmp.data_ptr(), scratch.data_ptr(),
                                            lam.data_ptr(), lam2.data_ptr(), params_dev.data_ptr(),
                                            pgrad.data_ptr(), a, f)
        return lam2 if which else lam


class _Run(torch.autograd.Function):
    """The schedule runs over units of ``stepper.levels`` steps (the fused
    adjoint reverses one unit per call): checkpoints sit at unit boundaries and
    ``snapshots`` counts them."""

    @staticmethod
    def forward(ctx, state0, params, stepper, nsteps, snapshots, overlap=True):
        st = stepper
        n, unit = int(nsteps), max(1, getattr(st, "levels", 1))
        p = params.detach().to(device=state0.device, dtype=torch.float64).contiguous()
        x0 = state0.detach().contiguous()
        tmp = torch.empty_like(x0)
        # the sweep keeps the checkpoints of the schedule's first path, so the
        # reversal does not recompute them
        kept, base, pos = [], x0, 0
        nunits = -(-n // unit)
        path = [] if getattr(st, "state_free", False) else first_descent(nunits, int(snapshots))
        for j in path:
            nxt = torch.empty_like(x0)
            st.forward(base, nxt, tmp, unit_steps(n, unit, pos, pos + j), p)
            kept.append(nxt)
            base, pos = nxt, pos + j
        out = torch.empty_like(x0)
        st.forward(base, out, tmp, n - unit_steps(n, unit, 0, pos), p)
        ctx.save_for_backward(x0, p)
        ctx.kept = kept
        ctx.stepper, ctx.nsteps, ctx.snapshots = st, n, int(snapshots)
        ctx.overlap = overlap if overlap in ("one", "two", "python") else ("one" if overlap else "python")
        ctx.param_device, ctx.param_dtype = params.device, params.dtype
        return out

    @staticmethod
    def backward(ctx, grad_out):
        x0, p = ctx.saved_tensors
        st = ctx.stepper
        n, unit = ctx.nsteps, max(1, getattr(st, "levels", 1))
        lam = grad_out.detach().to(st.dtype).contiguous().clone()
        lam2 = torch.empty_like(lam)
        pgrad = torch.zeros(max(1, st.slots), max(1, st.nparams), dtype=torch.float64, device=x0.device)
        stats = {"forward_steps": 0, "adjoint_steps": 0, "inner_forward_steps": 0, "adjoint_calls": 0}
        if getattr(st, "state_free", False):
            # the adjoint of a state-independent linear step needs no forward state
            for _ in range(n):
                st.adjoint(x0, lam, lam2, p, pgrad)
                lam, lam2 = lam2, lam
                stats["adjoint_steps"] += 1
                stats["adjoint_calls"] += 1
        elif n > 0:
            ops, nbufs = reversal_ops(n, unit, ctx.snapshots)
            kept = list(ctx.kept)
            ctx.kept = None
            work = [torch.empty_like(x0) for _ in range(nbufs - 1 - len(kept))]
            bufs = [x0] + kept + work
            tmp = torch.empty_like(x0)
            for kind, a, b, k in ops:
                stats["forward_steps" if kind == 0 else "adjoint_steps"] += k
                if kind == 1:
                    stats["inner_forward_steps"] += k - 1
                    stats["adjoint_calls"] += 1
            native = ctx.overlap in ("one", "two") and hasattr(st, "reverse") and x0.is_cuda
            if native:
                # the library's executor; "two": forward work on a second stream
                scratch = torch.empty([max(1, 2 * (unit - 1))] + list(x0.shape), dtype=x0.dtype, device=x0.device)
                sf = torch.cuda.Stream(device=x0.device) if ctx.overlap == "two" else None
                lam = st.reverse(ops, bufs, tmp, scratch, lam, lam2, p, pgrad, stream_f=sf)
            else:
                lam = _execute(st, ops, bufs, tmp, unit, lam, lam2, p, pgrad)
        ctx.stepper.last_stats = stats
        gp = pgrad.sum(0)[: st.nparams].to(device=ctx.param_device, dtype=ctx.param_dtype) if st.nparams else None
        return lam.to(grad_out.dtype), (gp if gp is not None else None), None, None, None, None


def _execute(st, ops, bufs, tmp, unit, lam, lam2, p, pgrad):
    """Reference executor of a ``reversal_ops`` schedule, one call per operation
    on the current stream (any stepper)."""
    scratch = st.scratch(bufs[0].device) if unit > 1 else None
    for kind, a, b, k in ops:
        if kind == 0:
            st.forward(bufs[a], bufs[b], tmp, k, p)
        else:
            if unit == 1:
                st.adjoint(bufs[a], lam, lam2, p, pgrad)
            else:
                st.adjoint_steps(bufs[a], scratch, lam, lam2, k, p, pgrad)
            lam, lam2 = lam2, lam
    return lam


def run(stepper: Stepper, state0: torch.Tensor, params: torch.Tensor, nsteps: int, snapshots: int = 32,
        overlap="one"):
    """State after ``nsteps`` steps, differentiable in ``state0`` and ``params``.
    ``overlap``: "one" runs the reversal in the library's executor on one
    stream, "two" with its forward work on a second stream, "python" (or False)
    with one library call per operation."""
    return _Run.apply(state0, params, stepper, nsteps, snapshots, overlap)
