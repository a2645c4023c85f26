import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, ROOT)
sys.path.insert(0, os.path.join(ROOT, "oracle"))

try:  # torch before any HIP library: its HIP runtime is the one the libraries bind to
    import torch  # noqa: F401
except ImportError:
    pass


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs the HIP libraries)")


_LIBS = {}


def load_lib(path, layout=None):
    """One FormuraLib per shared object: a library's globals are shared by every handle."""
    from formura_amd.runtime import FormuraLib
    if path not in _LIBS:
        _LIBS[path] = FormuraLib(path, layout)
    return _LIBS[path]


def current_lib(name):
    """Path of the built library ``name`` (any kind).  Where a GPU is visible
    the library must have been built from the sources the current generator
    writes (a stale one fails the test: build before running on the GPU);
    without a GPU it is built when missing or stale."""
    from formura_amd import build
    if has_gpu():
        return build.check_current(name)
    return build.builder_of(name)(name)


def hip_lib(name):
    """The HIP library of a bundled program (checked as ``current_lib``)."""
    return load_lib(current_lib(name))


def oracle_lib(name, layout, kind="strict"):
    import build_oracle
    p = build_oracle.lib_path(name, kind)
    if not os.path.exists(p):
        build_oracle.build(name, kind)
    return load_lib(p, layout)


def layout_of(name):
    """Layout of a bundled program from the front-end (no library needed)."""
    from formura_amd import build
    from formura_amd.emit_hip import generate
    from formura_amd.emit import lib_description
    fmr, yml, sched = build.PROGRAMS[name]
    _, _, mm = generate(os.path.join(build.PROG_DIR, fmr), os.path.join(build.PROG_DIR, yml), name)
    return lib_description(mm.om, name, "")


def has_gpu():
    try:
        import torch
        return torch.cuda.is_available()
    except Exception:
        return False
