"""Build recipe for the oracle (test infrastructure only).

Compiles the hand transcriptions of the reference's generated C (``*_ref.c``)
into shared libraries under ``oracle/_build/``.  Each library exports exactly
the API the reference's generated C exports (``Formura_Init``,
``Formura_Forward``, ``Formura_Finalize``, ``to_pos_*``, ``formura_data``).

Two builds per configuration:

* ``strict`` (parity): ``-O2 -ffp-contract=off`` -- no FMA contraction, no
  reassociation, so every operation is rounded as written;
* ``fast`` (CPU baseline timing): the reference Makefiles' flags
  ``-O3 -ffast-math -mavx2 -mfma -fopenmp`` (examples/diffusion2/Makefile:2,
  examples/pearson2/Makefile:2); ``-march=native`` (sample/3d_hllc/Makefile:2)
  is not used because the library is built here and run on another host.

Usage: ``python oracle/build_oracle.py [name ...]``.
"""
from __future__ import annotations

import hashlib
import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
OUT = os.path.join(HERE, "_build")

STRICT = ["-O2", "-ffp-contract=off", "-fno-fast-math", "-fopenmp", "-fPIC", "-shared", "-std=gnu11"]
FAST = ["-O3", "-ffast-math", "-mavx2", "-mfma", "-fopenmp", "-fPIC", "-shared", "-std=gnu11"]

# name -> (source, defines)
CONFIGS = {
    "heat3d_16": ("heat3d_ref.c", {"NX": 16}),
    "heat3d_16_tb": ("heat3d_ref.c", {"NX": 16, "TB": 1, "NT": 2, "GB1": 4, "GB2": 4, "GB3": 4}),
    "heat3d_32": ("heat3d_ref.c", {"NX": 32}),
    "heat3d_64": ("heat3d_ref.c", {"NX": 64}),
    "heat3d_64_tb": ("heat3d_ref.c", {"NX": 64, "TB": 1, "NT": 4, "GB1": 8, "GB2": 8, "GB3": 8}),
    "heat3d_256": ("heat3d_ref.c", {"NX": 256}),
    "gs_64": ("grayscott2d_ref.c", {"NX": 64}),
    "gs_72x120": ("grayscott2d_ref.c", {"NX": 72, "NY": 120}),
    "heat3d_24x40x56": ("heat3d_ref.c", {"NX": 24, "NY": 40, "NZ": 56}),
    "heat3d_24x40x56_tb": ("heat3d_ref.c", {"NX": 24, "NY": 40, "NZ": 56, "TB": 1, "NT": 2, "GB1": 7, "GB2": 11, "GB3": 6}),
    "gs_2048": ("grayscott2d_ref.c", {"NX": 2048}),
    "euler3d_16": ("euler3d_ref.c", {"NX": 16}),
    "euler3d_32": ("euler3d_ref.c", {"NX": 32}),
    "euler3d_256": ("euler3d_ref.c", {"NX": 256}),
    # the reference's sample/1d_hllc/hllc.fmr (MUSCL-Hancock HLLC) on one rank
    "hllc1d_64": ("hllc1d_ref.c", {"NX": 64, "LX": 4.0}),
    "hllc1d_400": ("hllc1d_ref.c", {"NX": 400, "LX": 4.0}),
    "hllc1d_128": ("hllc1d_ref.c", {"NX": 128, "LX": 4.0}),
    "hllc1d_256": ("hllc1d_ref.c", {"NX": 256, "LX": 4.0}),
    # the reference's shock tube (sample/3d_hllc: 400 cells over 4.0 along the
    # tube, dx = 0.01, dt = C_cfl*dx/max_signal_speed = (0.8*0.01)/4, :19-21)
    # along each axis in turn, 4 cells across
    "euler3d_tube_x": ("euler3d_ref.c", dict(NX=400, NY=4, NZ=4, LX=4.0, LY=0.04, LZ=0.04,
                                               DX=0.01, DY=0.01, DZ=0.01, DT="((0.8*0.01)/4.0)")),
    "euler3d_tube_y": ("euler3d_ref.c", dict(NX=4, NY=400, NZ=4, LX=0.04, LY=4.0, LZ=0.04,
                                               DX=0.01, DY=0.01, DZ=0.01, DT="((0.8*0.01)/4.0)")),
    "euler3d_tube_z": ("euler3d_ref.c", dict(NX=4, NY=4, NZ=400, LX=0.04, LY=0.04, LZ=4.0,
                                               DX=0.01, DY=0.01, DZ=0.01, DT="((0.8*0.01)/4.0)")),
}

# static arrays beyond 2 GiB need the medium code model
LARGE = {"euler3d_256"}


def lib_path(name: str, kind: str = "strict") -> str:
    return os.path.join(OUT, "liboracle_%s_%s.so" % (name, kind))


def build(name: str, kind: str = "strict", force: bool = False, source_text: str = None,
          out: str = None) -> str:
    """Compile configuration ``name``.  ``source_text`` replaces the source
    file's text (tests plant mistakes in a copy this way) and ``out`` names
    the library it goes to."""
    src, defs = CONFIGS[name]
    flags = (STRICT if kind == "strict" else FAST) + ["-D%s=%s" % kv for kv in sorted(defs.items())]
    if name in LARGE:
        flags.append("-mcmodel=medium")
    srcp = os.path.join(HERE, src)
    text = open(srcp).read() if source_text is None else source_text
    out = out or lib_path(name, kind)
    if source_text is not None:
        srcp = out + ".c"
        os.makedirs(os.path.dirname(srcp), exist_ok=True)
        with open(srcp, "w") as f:
            f.write(text)
    stamp = hashlib.sha1((" ".join(flags) + text).encode()).hexdigest()
    if not force and os.path.exists(out) and os.path.exists(out + ".stamp") and open(out + ".stamp").read() == stamp:
        return out
    os.makedirs(os.path.dirname(out), exist_ok=True)
    cmd = ["gcc"] + flags + ["-o", out + ".tmp", srcp, "-lm"]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError("oracle build failed: %s\n%s" % (" ".join(cmd), r.stderr))
    os.replace(out + ".tmp", out)
    with open(out + ".stamp", "w") as f:
        f.write(stamp)
    return out


def build_all(names=None, kinds=("strict", "fast")):
    out = []
    for n in names or CONFIGS:
        for k in kinds:
            out.append(build(n, k))
    return out


if __name__ == "__main__":
    for p in build_all(sys.argv[1:] or None):
        print(p)
