"""The allocation-free totals entry point, Formura_HIP_Totals_Static.

It sums the fields of a caller-provided device state image into caller-provided
device memory, with its partial sums in a file-scope device array instead of
memory the library allocates (Formura_HIP_Totals allocates its partial sums on
first use).  GPU: two calls in a row through one library -- first on that
library's own state, then on a smaller state of another program -- give
exactly what Formura_HIP_Totals of the owning library gives; an FP32 state
too.  CPU: sizes outside the stated limits stop with a message before any
device work.
"""
import ctypes
import os
import subprocess
import sys

import numpy as np
import pytest

from conftest import ROOT, has_gpu, hip_lib
from formura_amd import build as B, inputs


def _bind(lib):
    f = lib.lib.Formura_HIP_Totals_Static
    f.argtypes = [ctypes.c_void_p, ctypes.c_int, ctypes.c_longlong, ctypes.c_int, ctypes.c_void_p, ctypes.c_void_p]
    f.restype = None
    return f


def _resident(name, seed):
    h = hip_lib(name)
    h.init()
    h.set_state({f: inputs.uniform(seed + i, h.grid).astype(h.field(f).dtype) for i, f in enumerate(h.fields)})
    h.set_resident(True)
    h.upload()
    return h


@pytest.mark.gpu
def test_static_totals_match_totals_on_two_calls():
    import torch
    a = _resident("heat3d_32", 3)     # 1 FP64 field of 32^3 cells
    b = _resident("gs_64", 5)         # 2 FP64 fields of 64^2 cells: the smaller input
    c = _resident("heat3d_kahan_16", 7)  # 2 FP32 fields of 16^3 cells
    static = _bind(a)
    out = torch.full((8,), float("nan"), dtype=torch.float64, device="cuda")
    stream = a.lib.Formura_HIP_Stream()
    for lib, nbytes in ((a, 8), (b, 8), (c, 4)):
        cells = int(np.prod(lib.grid))
        print("Totals_Static of %s (%d fields x %d cells, %d-byte values)" % (lib.path, len(lib.fields), cells, nbytes),
              flush=True)
        out.fill_(float("nan"))
        torch.cuda.synchronize()
        static(lib.lib.Formura_HIP_State_Device(), len(lib.fields), cells, nbytes, out.data_ptr(), stream)
        torch.cuda.synchronize()
        got = out.cpu().numpy()
        want = lib.totals()
        for i, f in enumerate(lib.fields):
            assert got[i] == want[f], (lib.path, f, got[i], want[f])
        assert np.all(np.isnan(got[len(lib.fields):]))  # nothing written past nfields
    for lib in (a, b, c):
        lib.set_resident(False)
        lib.finalize()


@pytest.mark.parametrize("args,msg", [
    ("0, 8, 8", "nfields must be 1..64"),
    ("65, 8, 8", "nfields must be 1..64"),
    ("1, -1, 8", "cells must be 0..2^40"),
    ("1, (1 << 40) + 1, 8", "cells must be 0..2^40"),
    ("1, 8, 2", "elem_bytes must be 4 or 8"),
])
def test_static_totals_rejects_sizes_beyond_limits(args, msg):
    """Host-side checks, before any device call (runs without a GPU)."""
    p = B.lib_path("heat3d_16")
    if not os.path.exists(p):
        B.build_one("heat3d_16")
    code = ("import ctypes; L = ctypes.CDLL(%r); f = L.Formura_HIP_Totals_Static; "
            "f.argtypes = [ctypes.c_void_p, ctypes.c_int, ctypes.c_longlong, ctypes.c_int, ctypes.c_void_p, "
            "ctypes.c_void_p]; f(16, %s, 16, None)" % (p, args))
    r = subprocess.run([sys.executable, "-c", code], capture_output=True, text=True, timeout=120)
    assert r.returncode == 1, (r.returncode, r.stderr)
    assert "Formura_HIP_Totals_Static: " + msg in r.stderr
