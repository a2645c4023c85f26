"""The built libraries load and export every symbol the headers declare."""
import ctypes
import os
import re

import pytest

from conftest import ROOT

HEADER = os.path.join(ROOT, "include", "formura_hip.h")
REFERENCE_API = ["Formura_Init", "Formura_Forward", "Formura_Finalize", "formura_data"]


def declared(path):
    text = re.sub(r"/\*.*?\*/", "", open(path).read(), flags=re.S)  # declarations, not comments
    text = re.sub(r"//[^\n]*", "", text)
    return sorted(set(re.findall(r"\b(Formura_\w+)\s*\(", text)))


@pytest.mark.parametrize("name", ["heat3d_16", "gs_64", "euler3d_16_mixed"])
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


@pytest.mark.parametrize("kind,names", [
    ("batch", ["gs_batch_f64", "gs_batch_f32k", "gs_batch_f32c"]),
    ("adjoint", ["gs_64_adj", "gs_64_adj_f32", "heat3d_16_adj"]),
])
def test_form_libraries_export_declared_symbols(kind, names):
    """Batch and adjoint libraries export every entry point of their headers."""
    from formura_amd import build
    header = os.path.join(ROOT, "include", "formura_%s.h" % kind)
    syms = declared(header)
    assert len(syms) >= 5
    for name in names:
        p = build.lib_path(name)
        if not os.path.exists(p):
            (build.build_batch if kind == "batch" else build.build_adjoint)(name)
        lib = ctypes.CDLL(p)
        missing = [x for x in syms if not hasattr(lib, x)]
        assert not missing, (name, missing)


def test_headers_compile_as_c():
    """Every public header is valid C (declarations only, no C++)."""
    import subprocess
    for h in ("formura_hip.h", "formura_batch.h", "formura_adjoint.h"):
        src = '#include <stddef.h>\ntypedef struct Formura_Navi Formura_Navi;\ntypedef struct FormuraBatch FormuraBatch;\n#include "%s"\n' % h
        r = subprocess.run(["gcc", "-std=c11", "-Wall", "-fsyntax-only", "-I", os.path.join(ROOT, "include"), "-x", "c", "-"],
                           input=src, capture_output=True, text=True)
        assert r.returncode == 0, (h, r.stderr)
