"""The reference's own sample programs through this front-end and HIP emitter.

sample/1d_hllc (HLLC, first order), sample/mhd_1d_hlld and sample/mhd_3d_hlld
(HLLD MHD, second order, GLM cleaning in 3-D) and sample/3d_hllc (HLLC with the
MUSCL-Hancock reconstruction, the program's default) of formura v2.3.2 are
compiled with one-rank configs of this package (``build.SAMPLES``).  The
fixtures in tests/golden/sample_*.npz are the numpy interpreter's runs of the
same graphs, checked when written for conservation, uniform-state invariance
and, for the 400-cell shock tube from the program's own init graph, against
the reference's exact Riemann solution (tests/golden/make_samples.py).

CPU: the interpreter reproduces every fixture bit for bit (needs the formura
checkout; skipped without it).  GPU: each built library, seeded the same way,
gives the fixture bit for bit after the same number of Formura_Forward calls,
with the same Navi offsets; the shock tube also sits within the fixture's
distance of the exact solution.
"""
import json
import os

import numpy as np
import pytest

from conftest import current_lib, load_lib
from formura_amd import build as B, inputs

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")
CASES = ["hllc1d_64", "mhd1d_64", "hllc3d_8", "mhd3d_8", "hllc1d_400"]


def fixture(name):
    with open(os.path.join(GOLDEN, "sample_%s.json" % name)) as f:
        meta = json.load(f)
    with np.load(os.path.join(GOLDEN, "sample_%s.npz" % name)) as z:
        out = {k: z[k] for k in z.files}
    return meta, out


def exact_interior_l1(rho_frame, steps):
    ex = np.loadtxt(os.path.join(GOLDEN, "hllc1d_exact.txt"))
    n = rho_frame.shape[0]
    k = (np.arange(n) + (-2 * steps) % n) % n
    rho = np.empty(n)
    rho[k] = rho_frame
    return float(np.mean(np.abs(rho - ex[:, 1])[150:350]))


@pytest.mark.skipif(not B.reference_present(), reason="formura checkout not present")
@pytest.mark.parametrize("name", CASES)
def test_interpreter_reproduces_fixture(name):
    import sys
    sys.path.insert(0, GOLDEN)
    import make_samples as M
    meta, want = fixture(name)
    mm, desc = M.front_end(name)
    state = M.initial_state(name, mm, desc)
    got = M.run(mm, desc, state, meta["steps"])
    assert sorted(got) == meta["fields"]
    for f in got:
        np.testing.assert_array_equal(got[f], want[f], err_msg=f)


def test_shock_tube_fixture_near_exact_solution():
    meta, out = fixture("hllc1d_400")
    assert exact_interior_l1(out["rho"], meta["steps"]) < 0.01


@pytest.mark.gpu
@pytest.mark.parametrize("name", CASES)
def test_sample_library_matches_fixture_bit_exact(name):
    path = B.lib_path(name)
    if not os.path.exists(path):
        pytest.skip("%s is built only where the formura checkout is present" % name)
    path = current_lib(name)
    meta, want = fixture(name)
    print("sample %s: %d steps, grid %s" % (name, meta["steps"], meta["grid"]), flush=True)
    lib = load_lib(path)
    lib.init()
    if meta["seed"] is not None:
        lib.set_state(inputs.conserved_state(meta["seed"], tuple(meta["grid"]), lib.fields, meta["gamma"]))
    lib.forward(meta["steps"])
    nv = lib.navi_dict()
    assert nv["time_step"] == meta["steps"]
    for ax, off in zip(lib.axes, meta["offset"]):
        assert nv["offset_" + ax] == off
    for f in lib.fields:
        got = lib.field(f)
        if not np.array_equal(got, want[f]):
            raise AssertionError("%s field %s differs from the fixture: max |diff| %g"
                                 % (name, f, float(np.max(np.abs(got - want[f])))))
    if name == "hllc1d_400":
        assert exact_interior_l1(lib.field("rho"), meta["steps"]) < 0.01


@pytest.mark.gpu
def test_reference_driver_links_unchanged(tmp_path):
    """sample/1d_hllc/main.c of the reference, compiled unchanged against the
    generated header and the HIP library (build.SAMPLE_DRIVERS): it runs
    Formura_Init, writes data/0_0.dat, advances to time_step 100 with
    Formura_Forward and writes data/100_0.dat.  Loaded with the drift-aware
    loader, step 0 is the program's shock tube and step 100 is the fixture's
    state, value for value (the files carry %.17e)."""
    import subprocess
    from formura_amd import output
    exe = B.driver_path("hllc1d_400")
    if not os.path.exists(exe):
        pytest.skip("the reference's driver is built only where the formura checkout is present")
    print("reference driver sample/1d_hllc/main.c on the HIP library", flush=True)
    (tmp_path / "data").mkdir()
    r = subprocess.run([exe], cwd=str(tmp_path), capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stderr[-2000:]
    d0 = output.load_data(str(tmp_path / "data"), 0, [400])[0, 0]
    x = d0[:, 0]
    np.testing.assert_array_equal(x, 0.01 * np.arange(400))
    np.testing.assert_array_equal(d0[:, 3], np.where(np.arange(400) < 200, 1.0, 0.125))
    meta, want = fixture("hllc1d_400")
    d1 = output.load_data(str(tmp_path / "data"), 100, [400])[0, 0]
    np.testing.assert_array_equal(d1[:, 0], 0.01 * np.arange(400))
    rho = np.empty(400)
    rho[(np.arange(400) + meta["offset"][0]) % 400] = want["rho"]
    np.testing.assert_array_equal(d1[:, 3], rho)


@pytest.mark.gpu
@pytest.mark.parametrize("name", ["mhd3d_8", "hllc3d_8"])
def test_mixed_sample_against_fp64(name):
    """The mixed form of the 3-D samples (FP32 fluxes; FP64 wave speeds,
    reconstructions, square-root arguments and the update chain): after the
    fixture's steps every field is within 1e-5 (relative max-norm) of the FP64
    run, every total but psi is conserved to 1e-13 of its scale, and no FP64
    square-root argument was found non-positive."""
    path = B.lib_path(name + "_mixed")
    if not os.path.exists(path):
        pytest.skip("%s_mixed is built only where the formura checkout is present" % name)
    path = current_lib(name + "_mixed")
    meta, want = fixture(name)
    print("mixed sample %s: %d steps" % (name, meta["steps"]), flush=True)
    lib = load_lib(path)
    lib.init()
    st = inputs.conserved_state(meta["seed"], tuple(meta["grid"]), lib.fields, meta["gamma"])
    lib.set_state(st)
    lib.set_resident(True)
    lib.upload()
    t0 = lib.totals()
    lib.forward_steps(meta["steps"])
    t1 = lib.totals()
    lib.download()
    lib.set_resident(False)
    assert lib.positivity_violations() == 0
    for f in lib.fields:
        err = float(np.max(np.abs(lib.field(f) - want[f])) / np.max(np.abs(want[f])))
        assert err < 1e-5, (f, err)
    cons = [f for f in t0 if f != "psi"]
    scale = max(abs(t0[f]) for f in cons)
    assert max(abs(t1[f] - t0[f]) for f in cons) <= 1e-13 * scale


EXACT = ["mhd1d_3200", "hllc3d_shock"]


def _exact_l1(meta, name, fields_frame):
    """Mean |f - f_exact| over the compared cells (the fixture's exact_<name>.txt
    holds x and the compared columns; output index i is cell (i + offset) mod N)."""
    ex = np.loadtxt(os.path.join(GOLDEN, "exact_%s.txt" % name))
    first, count = meta["exact_cells"]
    row0 = first if ex.shape[0] == count else 0
    out = {}
    for j, (f, _) in enumerate(meta["exact_columns"]):
        axis = 0 if fields_frame[f].ndim == 1 else 1  # the 3-D tube runs along y
        prof = fields_frame[f] if axis == 0 else fields_frame[f][0, :, 0]
        n = prof.shape[0]
        v = np.empty(n)
        v[(np.arange(n) + meta["offset"][axis]) % n] = prof
        out[f] = float(np.mean(np.abs(v[first:first + count] - ex[first - row0:first - row0 + count, j + 1])))
    return out


@pytest.mark.skipif(not B.reference_present(), reason="formura checkout not present")
@pytest.mark.parametrize("name", EXACT)
def test_interpreter_reproduces_reference_configuration(name):
    import sys
    sys.path.insert(0, GOLDEN)
    import make_samples as M
    import mm_interp
    with open(os.path.join(GOLDEN, "sample_%s.json" % name)) as f:
        meta = json.load(f)
    mm, desc = M.front_end(name)
    st = M.run(mm, desc, mm_interp.init(mm.init, tuple(desc["grid"])), meta["steps"])
    assert M.field_hashes(st) == meta["sha256"]


def test_exact_solution_fixtures_pinned():
    """The stored exact solutions and the bounds agree with the fixtures' own record."""
    for name in EXACT:
        with open(os.path.join(GOLDEN, "sample_%s.json" % name)) as f:
            meta = json.load(f)
        ex = np.loadtxt(os.path.join(GOLDEN, "exact_%s.txt" % name))
        assert ex.shape[1] == 1 + len(meta["exact_columns"]) and meta["l1_bound"] <= 0.01


@pytest.mark.gpu
@pytest.mark.parametrize("name", EXACT)
def test_reference_configuration_against_exact_solution(name):
    """The HLLD MHD (Ryu-Jones 2a) and MUSCL-Hancock HLLC shock tubes at the
    reference's configurations, from the programs' own init graphs: the GPU state
    equals the interpreter's bit for bit (SHA-256 per field), Navi included, the
    3-D tube stays uniform across x and z, and the profile is within the
    fixture's bound of the reference's exact solution."""
    import hashlib
    path = B.lib_path(name)
    if not os.path.exists(path):
        pytest.skip("%s is built only where the formura checkout is present" % name)
    path = current_lib(name)
    with open(os.path.join(GOLDEN, "sample_%s.json" % name)) as f:
        meta = json.load(f)
    print("reference configuration %s: %d steps" % (name, meta["steps"]), flush=True)
    lib = load_lib(path)
    lib.init()
    lib.set_resident(True)
    lib.upload()
    lib.forward_steps(meta["steps"])
    lib.download()
    lib.set_resident(False)
    nv = lib.navi_dict()
    assert [nv["offset_" + a] for a in lib.axes] == meta["offset"]
    st = {f: lib.field(f).copy() for f in lib.fields}
    got = {f: hashlib.sha256(np.ascontiguousarray(st[f], dtype=np.float64).tobytes()).hexdigest() for f in sorted(st)}
    assert got == meta["sha256"]
    if len(meta["grid"]) == 3:
        for f, v in st.items():
            assert (v == v[:1, :, :1]).all(), f
    for f, e in _exact_l1(meta, name, st).items():
        assert e < meta["l1_bound"], (f, e)


@pytest.mark.gpu
@pytest.mark.parametrize("name", ["hllc1d_64", "hllc1d_400"])
def test_sample_library_matches_hand_transcription(name):
    """The HIP library of sample/1d_hllc/hllc.fmr against oracle/hllc1d_ref.c,
    the transcription of the C formura emits for it (placement derived by hand,
    independent of this front-end): state and Navi equal bit for bit."""
    import test_oracle_hllc1d as T
    path = B.lib_path(name)
    if not os.path.exists(path):
        pytest.skip("%s is built only where the formura checkout is present" % name)
    lib = load_lib(current_lib(name))
    lib.init()
    o = T.oracle(name)
    meta, _ = fixture(name)
    print("sample %s against the transcription: %d steps" % (name, meta["steps"]), flush=True)
    if meta["seed"] is not None:
        st = inputs.conserved_state(meta["seed"], tuple(meta["grid"]), lib.fields, meta["gamma"])
        lib.set_state(st)
        o.set_state(st)
    lib.forward(meta["steps"])
    o.forward(meta["steps"])
    assert lib.navi_dict()["offset_x"] == o.navi_dict()["offset_x"]
    for f in lib.fields:
        np.testing.assert_array_equal(lib.field(f), o.field(f), err_msg=f)
