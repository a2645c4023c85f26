"""A C driver written against the reference's generated API links unchanged
against the HIP library and prints the same numbers as when linked against the
oracle (the transcription of the reference's generated C)."""
import os
import shutil
import subprocess

import pytest

from conftest import ROOT

pytestmark = pytest.mark.gpu


def _build_and_run(tmp_path, lib_dir, lib_name, header_dir):
    exe = tmp_path / ("main_" + lib_name)
    src = os.path.join(ROOT, "examples", "heat3d_dropin", "main.c")
    subprocess.run(["gcc", "-O2", "-std=c11", "-I", header_dir, "-I", os.path.join(ROOT, "include"), src,
                    "-L", lib_dir, "-l" + lib_name, "-Wl,-rpath," + lib_dir, "-lm", "-o", str(exe)], check=True)
    r = subprocess.run([str(exe)], capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stderr
    return r.stdout.strip()


def test_user_main_links_unchanged(tmp_path):
    from formura_amd import build
    import build_oracle
    from conftest import current_lib
    hip_so = current_lib("heat3d_64")
    hdr_dir = os.path.dirname(build.header_path("heat3d_64"))
    out_hip = _build_and_run(tmp_path, hdr_dir, "heat3d_64", hdr_dir)
    # the oracle as a library with the same soname pattern
    orc = build_oracle.lib_path("heat3d_64", "strict")
    if not os.path.exists(orc):
        build_oracle.build("heat3d_64", "strict")
    d = tmp_path / "oracle"
    d.mkdir()
    shutil.copy(orc, d / "liboracle_heat.so")
    out_ref = _build_and_run(tmp_path, str(d), "oracle_heat", hdr_dir)
    assert out_hip == out_ref
    assert "time_step 40 offset 24 24 24" in out_hip
