"""Host binding of a batched-ensemble library (``include/formura_batch.h``) and
the seeded Gray-Scott ensemble.

``Ensemble`` wraps ``Formura_Batch_*`` with numpy.  ``grayscott_ensemble``
draws the instances: square grids whose side is log-uniform in [64, 1024]
rounded to a multiple of 32, F ~ U[0.01, 0.06], k ~ U[0.045, 0.07], and the
single-instance initial state (the program's init graph: u = 1, v = 0, a
centred square of side 2w-1, w = N/20 + 1/2, with u = 1/2, v = 1/4) plus the
seeded 1 % noise of ``inputs.grayscott_perturb``.
"""
from __future__ import annotations

import ctypes
import json
import math
import os
from typing import List, Sequence

import numpy as np

from . import build as _build
from . import inputs


class Ensemble:
    def __init__(self, name: str, sizes: Sequence[tuple], params: np.ndarray, path: str = None):
        path = path or _build.lib_path(name)
        if not os.path.exists(path):
            raise RuntimeError("batch library %s is missing at %s; build it with "
                               "`python -m formura_amd.build %s`" % (name, path, name))
        self.lib = L = ctypes.CDLL(path, mode=os.RTLD_LOCAL)
        L.Formura_Batch_Describe.restype = ctypes.c_char_p
        self.desc = json.loads(L.Formura_Batch_Describe().decode())
        L.Formura_Batch_Create.restype = ctypes.c_void_p
        L.Formura_Batch_Create.argtypes = [ctypes.c_int, ctypes.POINTER(ctypes.c_int), ctypes.POINTER(ctypes.c_int),
                                           ctypes.POINTER(ctypes.c_double)]
        for fn in ("Set_State", "Get_State"):
            getattr(L, "Formura_Batch_" + fn).argtypes = [ctypes.c_void_p, ctypes.c_int, ctypes.POINTER(ctypes.c_double)]
        L.Formura_Batch_Offsets.argtypes = [ctypes.c_void_p, ctypes.c_int, ctypes.POINTER(ctypes.c_int)]
        L.Formura_Batch_Step.argtypes = [ctypes.c_void_p, ctypes.c_int]
        L.Formura_Batch_Step.restype = ctypes.c_double
        L.Formura_Batch_Last_Launches.argtypes = [ctypes.c_void_p]
        L.Formura_Batch_Last_Launches.restype = ctypes.c_longlong
        L.Formura_Batch_Cells.argtypes = [ctypes.c_void_p]
        L.Formura_Batch_Cells.restype = ctypes.c_longlong
        L.Formura_Batch_Destroy.argtypes = [ctypes.c_void_p]
        self.sizes = [tuple(int(x) for x in s) for s in sizes]
        self.nf = len(self.desc["fields"])
        self.np = max(1, len(self.desc["params"]))
        n1 = (ctypes.c_int * len(sizes))(*[s[0] for s in self.sizes])
        n2 = (ctypes.c_int * len(sizes))(*[s[1] for s in self.sizes])
        p = np.ascontiguousarray(np.asarray(params, dtype=np.float64).reshape(len(sizes), self.np))
        self.params = p
        self.h = L.Formura_Batch_Create(len(sizes), n1, n2, p.ctypes.data_as(ctypes.POINTER(ctypes.c_double)))

    def set_state(self, i: int, fields: dict):
        a = np.ascontiguousarray(np.stack([fields[f] for f in self.desc["fields"]]), dtype=np.float64)
        assert a.shape == (self.nf,) + self.sizes[i]
        self.lib.Formura_Batch_Set_State(self.h, i, a.ctypes.data_as(ctypes.POINTER(ctypes.c_double)))

    def get_state(self, i: int) -> dict:
        a = np.empty((self.nf,) + self.sizes[i], dtype=np.float64)
        self.lib.Formura_Batch_Get_State(self.h, i, a.ctypes.data_as(ctypes.POINTER(ctypes.c_double)))
        return {f: a[k] for k, f in enumerate(self.desc["fields"])}

    def offsets(self, i: int):
        o = (ctypes.c_int * 2)()
        self.lib.Formura_Batch_Offsets(self.h, i, o)
        return list(o)

    def step(self, n: int) -> float:
        return float(self.lib.Formura_Batch_Step(self.h, int(n)))

    def last_launches(self) -> int:
        return int(self.lib.Formura_Batch_Last_Launches(self.h))

    def cells(self) -> int:
        return int(self.lib.Formura_Batch_Cells(self.h))

    def close(self):
        if self.h:
            self.lib.Formura_Batch_Destroy(self.h)
            self.h = None


def grayscott_init(n: int, seed: int) -> dict:
    """The program's init graph on an n x n grid plus the seeded perturbation."""
    i = np.arange(n, dtype=np.float64)
    w = n / 20.0 + 0.5
    inside = (np.abs(n / 2.0 - i)[:, None] < w) & (np.abs(n / 2.0 - i)[None, :] < w)
    u = np.where(inside, 0.5, 1.0)
    v = np.where(inside, 0.25, 0.0)
    return inputs.grayscott_perturb(u, v, seed)


def grayscott_ensemble(count: int = 1024, seed: int = 2024, smin: int = 64, smax: int = 1024, mult: int = 32):
    """Sizes, (F, k) per instance, and per-instance seeds."""
    z = inputs.uniform(seed, (count, 3))
    sides = [int(max(smin, min(smax, mult * round(math.exp(math.log(smin) + z[i, 0] * (math.log(smax) - math.log(smin))) / mult))))
             for i in range(count)]
    F = 0.01 + 0.05 * z[:, 1]
    k = 0.045 + 0.025 * z[:, 2]
    return [(s, s) for s in sides], np.stack([F, k], axis=1), [seed + 1 + i for i in range(count)]
