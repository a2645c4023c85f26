"""Differentiating a Formura run: torch.autograd over the generated forward and
adjoint steps (``include/formura_adjoint.h``), with binomial checkpointing.

``Stepper`` binds one adjoint library.  ``run(state0, params, nsteps,
snapshots)`` is a ``torch.autograd.Function``.  Its forward advances
``nsteps`` steps and keeps only ``snapshots`` states in HBM.  Its backward
reverses the run with the binomial checkpointing schedule of Griewank and
Walther ("Algorithm 799: Revolve", ACM TOMS 26(1), 2000).  A segment of l
steps with s free snapshots (c = s+1 checkpoints counting the segment's first
state) and repetition number r (the least r with C(c+r, c) >= l) is split after
j steps, j chosen in the interval where both parts stay within r repetitions;
the tail is reversed with s-1 snapshots and the head with s.  Forward steps
recomputed: r*l - C(c+r, r-1), the binomial optimum.  Each reversed step calls the adjoint kernel at the
recomputed input state.

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


def forward_steps_needed(l: int, s: int) -> int:
    """Forward steps the reversal recomputes (the first forward sweep excluded),
    following exactly the schedule of ``_Run.backward``."""
    total = 0
    stack = [(l, s)]
    while stack:
        l, s = stack.pop()
        while l > 1:
            if s == 0:
                total += l * (l - 1) // 2
                break
            j = split(l, s)
            total += j
            stack.append((l - j, s - 1))
            l = j
    return total


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
        L.Formura_AD_Forward.argtypes = [ctypes.c_void_p] * 3 + [ctypes.c_int, ctypes.c_void_p, ctypes.c_void_p]
        L.Formura_AD_Adjoint.argtypes = [ctypes.c_void_p] * 6
        self.dtype = torch.float64 if self.desc["real"] == "double" else torch.float32
        self.shape = [len(self.desc["fields"])] + list(self.desc["grid"])
        self.nparams = len(self.desc["params"])
        self.slots = int(L.Formura_AD_Param_Slots())

    def _stream(self):
        return ctypes.c_void_p(torch.cuda.current_stream().cuda_stream)

    def _check(self, t):
        if not (t.is_cuda and t.dtype == self.dtype and list(t.shape) == self.shape and t.is_contiguous()):
            raise ValueError("expected a contiguous %s CUDA tensor of shape %s" % (self.dtype, self.shape))

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


class _Run(torch.autograd.Function):
    @staticmethod
    def forward(ctx, state0, params, stepper, nsteps, snapshots):
        st = stepper
        p = params.detach().to(device=state0.device, dtype=torch.float64).contiguous()
        x0 = state0.detach().contiguous()
        out = torch.empty_like(x0)
        tmp = torch.empty_like(x0)
        st.forward(x0, out, tmp, nsteps, p)
        ctx.save_for_backward(x0, p)
        ctx.stepper, ctx.nsteps, ctx.snapshots = st, int(nsteps), int(snapshots)
        ctx.param_device, ctx.param_dtype = params.device, params.dtype
        return out

    @staticmethod
    def backward(ctx, grad_out):
        x0, p = ctx.saved_tensors
        st = ctx.stepper
        lam = grad_out.detach().to(st.dtype).contiguous().clone()
        lam2 = torch.empty_like(lam)
        pgrad = torch.zeros(max(1, st.slots), max(1, st.nparams), dtype=torch.float64, device=x0.device)
        tmp = torch.empty_like(x0)
        stats = {"forward_steps": 0, "adjoint_steps": 0}

        def advance(src, nsteps):
            dst = torch.empty_like(src)
            st.forward(src, dst, tmp, nsteps, p)
            stats["forward_steps"] += nsteps
            return dst

        def adjoint_at(state):
            nonlocal lam, lam2
            st.adjoint(state, lam, lam2, p, pgrad)
            lam, lam2 = lam2, lam
            stats["adjoint_steps"] += 1

        def reverse(base, l, s):
            # base: state at the start of the segment; reverse its l steps with s free
            # snapshots.  The tail recursion over the head of the segment is a loop,
            # so the depth is at most s.
            while True:
                if l == 1:
                    adjoint_at(base)
                    return
                if s == 0:
                    for j in range(l - 1, -1, -1):
                        adjoint_at(advance(base, j) if j else base)
                    return
                j = split(l, s)
                mid = advance(base, j)
                reverse(mid, l - j, s - 1)
                del mid
                l = j

        if ctx.nsteps > 0:
            reverse(x0, ctx.nsteps, ctx.snapshots)
        ctx.stepper.last_stats = stats
        gp = pgrad.sum(0)[: st.nparams].to(device=ctx.param_device, dtype=ctx.param_dtype) if st.nparams else None
        return lam.to(grad_out.dtype), (gp if gp is not None else None), None, None, None


def run(stepper: Stepper, state0: torch.Tensor, params: torch.Tensor, nsteps: int, snapshots: int = 32):
    """State after ``nsteps`` steps, differentiable in ``state0`` and ``params``."""
    return _Run.apply(state0, params, stepper, nsteps, snapshots)
