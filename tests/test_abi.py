"""The built libraries load and export every symbol the headers declare."""
import ctypes
import os
import re

import pytest

from conftest import ROOT

HEADER = os.path.join(ROOT, "include", "formura_hip.h")
REFERENCE_API = ["Formura_Init", "Formura_Forward", "Formura_Finalize", "formura_data"]


def declared(path):
    text = open(path).read()
    return sorted(set(re.findall(r"\b(Formura_\w+)\s*\(", text)))


@pytest.mark.parametrize("name", ["heat3d_16", "gs_64"])
def test_library_exports_declared_symbols(name):
    from formura_amd import build
    p = build.lib_path(name)
    if not os.path.exists(p):
        build.build_one(name)
    lib = ctypes.CDLL(p)
    syms = declared(HEADER) + REFERENCE_API
    hdr = open(build.header_path(name)).read()
    syms += re.findall(r"\b(to_pos_\w+)\(", hdr)
    missing = [s for s in syms if not hasattr(lib, s)]
    assert not missing, missing


def test_dropin_driver_compiles_against_generated_header():
    """The example user driver compiles (C11) against the generated header alone."""
    import subprocess
    from formura_amd import build
    d = build.generate_program("heat3d_64", os.path.join(build.PROG_DIR, "heat3d.fmr"),
                               os.path.join(build.PROG_DIR, "heat3d_64.yaml"), "auto")
    src = os.path.join(ROOT, "examples", "heat3d_dropin", "main.c")
    r = subprocess.run(["gcc", "-std=c11", "-Wall", "-fsyntax-only", "-I", d, "-I", os.path.join(ROOT, "include"), src],
                       capture_output=True, text=True)
    assert r.returncode == 0, r.stderr
