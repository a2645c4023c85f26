"""Host-side binding of a generated Formura library (ctypes).

``FormuraLib`` loads any library that exports the API of the reference's
generated C -- ``Formura_Init``, ``Formura_Forward``, ``Formura_Finalize``,
``to_pos_<axis>`` and the global ``formura_data`` (Templates.hs:23-51,
221-250, 295-334, 383-386) -- and exposes the state as numpy arrays that alias
``formura_data``, the way a user's ``main.c`` reads and writes it directly
(examples/pearson2/main.c:13-22).  The HIP libraries also export the
device-resident extension of ``include/formura_hip.h``.

``open_program`` opens the HIP library of a bundled program and fails loudly
when it has not been built: there is no CPU fallback.
"""
from __future__ import annotations

import ctypes
import json
import os
from typing import Dict, List, Optional

import numpy as np

from . import build as _build

_CT = {"double": ctypes.c_double, "float": ctypes.c_float, "int": ctypes.c_int}
_NP = {"double": np.float64, "float": np.float32, "int": np.int32}


class LibraryMissing(RuntimeError):
    pass


def make_layout(om) -> dict:
    """Layout of a program from the front-end (the same JSON the HIP library describes)."""
    from .emit import lib_description
    return lib_description(om, "", "")


class FormuraLib:
    def __init__(self, path: str, layout: Optional[dict] = None):
        if not os.path.exists(path):
            raise LibraryMissing("library %s is not built (run __graft_entry__.build())" % path)
        self.path = path
        self.lib = ctypes.CDLL(path, mode=os.RTLD_LOCAL | getattr(os, "RTLD_NOW", 2))
        self.is_hip = hasattr(self.lib, "Formura_HIP_Describe")
        if layout is None:
            if not self.is_hip:
                raise ValueError("a layout is required for a library without Formura_HIP_Describe")
            self.lib.Formura_HIP_Describe.restype = ctypes.c_char_p
            layout = json.loads(self.lib.Formura_HIP_Describe().decode())
        self.layout = layout
        self.grid = list(layout["grid"])
        self.axes = list(layout["axes"])
        self.fields = [f["name"] for f in layout["fields"]]
        shape = 1
        for n in self.grid:
            shape *= n

        def arr(ct):
            t = _CT[ct]
            for n in reversed(self.grid):
                t = t * n
            return t

        class Grid(ctypes.Structure):
            _fields_ = [(f["name"], arr(f["ctype"])) for f in layout["fields"]]

        class Navi(ctypes.Structure):
            _fields_ = [(f["name"], _CT[f["ctype"]]) for f in layout["navi"]]

        self.Grid, self.Navi = Grid, Navi
        self.data = Grid.in_dll(self.lib, "formura_data")
        self._views: Dict[str, np.ndarray] = {}
        base = ctypes.addressof(self.data)
        for f in layout["fields"]:
            off = getattr(Grid, f["name"]).offset
            dt = _NP[f["ctype"]]
            buf = (ctypes.c_char * (shape * np.dtype(dt).itemsize)).from_address(base + off)
            self._views[f["name"]] = np.frombuffer(buf, dtype=dt).reshape(self.grid)
        L = self.lib
        L.Formura_Init.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.POINTER(Navi)]
        L.Formura_Init.restype = None
        L.Formura_Forward.argtypes = [ctypes.POINTER(Navi)]
        L.Formura_Forward.restype = None
        L.Formura_Finalize.argtypes = []
        L.Formura_Finalize.restype = None
        for a in self.axes:
            fn = getattr(L, "to_pos_" + a)
            fn.argtypes = [ctypes.c_int, Navi]
            fn.restype = ctypes.c_double
        if self.is_hip:
            L.Formura_Forward_Steps.argtypes = [ctypes.POINTER(Navi), ctypes.c_int]
            L.Formura_Forward_Steps.restype = None
            L.Formura_HIP_Set_Resident.argtypes = [ctypes.c_int]
            L.Formura_HIP_Last_Elapsed_Ms.restype = ctypes.c_double
            L.Formura_HIP_State_Device.restype = ctypes.c_void_p
            L.Formura_HIP_Stream.restype = ctypes.c_void_p
            L.Formura_HIP_Last_Launches.restype = ctypes.c_longlong
        self.navi = Navi()

    # -- the reference API ------------------------------------------------------
    def init(self):
        argc = ctypes.c_int(0)
        argv = ctypes.POINTER(ctypes.c_char_p)()
        self.lib.Formura_Init(ctypes.byref(argc), ctypes.byref(argv), ctypes.byref(self.navi))
        return self.navi

    def forward(self, n: int = 1):
        for _ in range(n):
            self.lib.Formura_Forward(ctypes.byref(self.navi))
        return self.navi

    def finalize(self):
        self.lib.Formura_Finalize()

    def to_pos(self, axis: str, i: int) -> float:
        return getattr(self.lib, "to_pos_" + axis)(i, self.navi)

    def field(self, name: str) -> np.ndarray:
        return self._views[name]

    def state(self) -> Dict[str, np.ndarray]:
        return {k: v.copy() for k, v in self._views.items()}

    def set_state(self, st: Dict[str, np.ndarray]):
        for k, v in st.items():
            self._views[k][...] = v

    def navi_dict(self) -> dict:
        return {f: getattr(self.navi, f) for f, _ in self.Navi._fields_}

    # -- HIP extension ------------------------------------------------------------
    def _need_hip(self):
        if not self.is_hip:
            raise RuntimeError("%s is not a HIP library" % self.path)

    def set_resident(self, on: bool):
        self._need_hip()
        self.lib.Formura_HIP_Set_Resident(1 if on else 0)

    def upload(self):
        self._need_hip()
        self.lib.Formura_HIP_Upload()

    def download(self):
        self._need_hip()
        self.lib.Formura_HIP_Download()

    def forward_steps(self, k: int):
        """``k`` Forward calls on the device; returns their device time in ms."""
        self._need_hip()
        self.lib.Formura_Forward_Steps(ctypes.byref(self.navi), int(k))
        return self.lib.Formura_HIP_Last_Elapsed_Ms()

    def last_launches(self) -> int:
        """Kernel launches issued by the last ``forward_steps`` call."""
        self._need_hip()
        return int(self.lib.Formura_HIP_Last_Launches())

    def totals(self) -> dict:
        """FP64 totals of the device state per field (fixed-order device sums)."""
        self._need_hip()
        out = (ctypes.c_double * len(self.fields))()
        self.lib.Formura_HIP_Totals(out)
        return {f: float(out[i]) for i, f in enumerate(self.fields)}

    def positivity_violations(self) -> int:
        """FP64 square-root arguments found not > 0 (mixed-precision libraries)."""
        self._need_hip()
        self.lib.Formura_HIP_Positivity_Violations.restype = ctypes.c_longlong
        return int(self.lib.Formura_HIP_Positivity_Violations())


def open_program(name: str) -> FormuraLib:
    """Open the HIP library of a bundled program (see ``formura_amd.build.PROGRAMS``)."""
    path = _build.lib_path(name)
    if not os.path.exists(path):
        raise LibraryMissing("HIP library for %s is missing at %s; build it with "
                             "`python -m formura_amd.build %s`" % (name, path, name))
    return FormuraLib(path)
