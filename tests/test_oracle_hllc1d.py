"""Pins the reference's own sample program sample/1d_hllc/hllc.fmr (MUSCL-Hancock
reconstruction with the minmod limiter, HLLC flux, half-offset face values,
three levels of manifest temporaries) independently of this package's
front-end: oracle/hllc1d_ref.c transcribes the C that formura emits for it,
with every nest's range, extent and cursor offset derived by hand from
Translate.hs (half offsets) and Templates.hs (calcRange, mkKernel).

* the transcription equals the fixtures of tests/golden (the front-end's graph
  evaluated by the numpy interpreter) bit for bit: 64 cells from a seeded
  state for 20 steps, and the 400-cell shock tube from the program's own init
  graph for 100 steps, Navi offsets included;
* the transcription is pinned by what the physics fixes: the shock tube within
  0.006 mean |rho - rho_exact| of the exact Riemann solution the reference
  ships, with the drifting frame's cell map closer than +-1, +-2 cells; the
  five totals conserved to 1e-13; a uniform state fixed bit for bit;
* a smooth density wave carried by a uniform flow (rho = 1 + 0.2 sin(2 pi x/4),
  u = 0.5, p = 1, exact solution its translation) converges at second order:
  the mean error falls by 3.7x per halving of dx (64, 128, 256 cells, t = 1;
  bound 3.4x), as MUSCL-Hancock does with the minmod limiter;
* planted mistakes in a copy of the transcription (a face read one cell off, a
  limiter branch swapped, the half-step predictor's sign) fail: each differs
  from the fixture, and each fails a physics pin (the shock tube, or the
  convergence order: the predictor's sign gives 2.2x, first order, while its
  shock tube stays close to the exact solution).
"""
import os

import numpy as np
import pytest

from conftest import ROOT
from formura_amd import inputs
from formura_amd.runtime import FormuraLib
import build_oracle

GOLDEN = os.path.join(ROOT, "tests", "golden")
SRC = os.path.join(ROOT, "oracle", "hllc1d_ref.c")
FIELDS = ["e", "rho", "rhou", "rhov", "rhow"]


def layout(n):
    navi = [("time_step", "int"), ("lower_x", "int"), ("upper_x", "int"), ("space_interval_x", "double"),
            ("my_rank", "int"), ("offset_x", "int"), ("length_x", "double"), ("total_grid_x", "int")]
    return {"name": "hllc1d", "dimension": 1, "axes": ["x"], "grid": [n], "sleeve": 2, "steps_per_forward": 1,
            "fields": [{"name": f, "ctype": "double"} for f in FIELDS],
            "navi": [{"name": a, "ctype": b} for a, b in navi], "filter_interval": None, "schedule": ""}


def oracle(name, src=None, tag=""):
    out = None
    if src is not None:
        out = os.path.join(build_oracle.OUT, "planted", "liboracle_%s_%s.so" % (name, tag))
    o = FormuraLib(build_oracle.build(name, "strict", source_text=src, out=out), layout(int(name.split("_")[1])))
    o.init()
    return o


def fixture(name):
    import json
    with open(os.path.join(GOLDEN, "sample_%s.json" % name)) as f:
        meta = json.load(f)
    with np.load(os.path.join(GOLDEN, "sample_%s.npz" % name)) as z:
        return meta, {k: z[k] for k in z.files}


def run(o, name):
    meta, _ = fixture(name)
    if meta["seed"] is not None:
        o.set_state(inputs.conserved_state(meta["seed"], tuple(meta["grid"]), FIELDS, meta["gamma"]))
    o.forward(meta["steps"])
    return meta


def tube_l1(rho, off):
    ex = np.loadtxt(os.path.join(GOLDEN, "hllc1d_exact.txt"))
    n = rho.shape[0]
    by_pos = np.empty(n)
    by_pos[(np.arange(n) + off) % n] = rho
    return float(np.mean(np.abs(by_pos - ex[:, 1])[150:350])), by_pos, ex


@pytest.mark.parametrize("name", ["hllc1d_64", "hllc1d_400"])
def test_transcription_equals_front_end_fixture(name):
    o = oracle(name)
    meta = run(o, name)
    _, want = fixture(name)
    assert [o.navi_dict()["offset_x"]] == meta["offset"]
    for f in FIELDS:
        np.testing.assert_array_equal(o.field(f), want[f], err_msg=f)


def test_shock_tube_near_exact_solution():
    o = oracle("hllc1d_400")
    meta = run(o, "hllc1d_400")
    off = o.navi_dict()["offset_x"]
    l1, by_pos, ex = tube_l1(o.field("rho"), off)
    assert l1 < 0.006, l1
    for s in (-2, -1, 1, 2):
        assert float(np.mean(np.abs(np.roll(by_pos, s) - ex[:, 1])[150:350])) > l1


def test_conservation_and_uniform_state():
    o = oracle("hllc1d_64")
    st = inputs.conserved_state(9, (64,), FIELDS, 1.4)
    o.set_state(st)
    o.forward(15)
    for f in FIELDS:
        assert abs(o.field(f).sum() - st[f].sum()) <= 1e-13 * np.abs(st[f]).sum(), f
    o.init()
    u = {"rho": np.full(64, 1.2), "rhou": np.full(64, 0.3), "rhov": np.full(64, -0.1), "rhow": np.full(64, 0.2),
         "e": np.full(64, 2.7)}
    o.set_state(u)
    o.forward(7)
    for f in FIELDS:
        assert np.all(o.field(f) == u[f][0]), f


def wave_errors(src=None, tag=""):
    """Mean |rho - exact| of the smooth wave after t = 1 at 64, 128, 256 cells."""
    errs = []
    for n in (64, 128, 256):
        o = oracle("hllc1d_%d" % n, src, tag)
        x = 4.0 * np.arange(n) / n
        rho = 1 + 0.2 * np.sin(2 * np.pi * x / 4.0)
        u, p = 0.5, 1.0
        o.set_state({"rho": rho, "rhou": rho * u, "rhov": 0 * rho, "rhow": 0 * rho, "e": p / 0.4 + 0.5 * rho * u * u})
        dt = (0.8 * (4.0 / n)) / 4.0
        steps = int(round(1.0 / dt))
        o.forward(steps)
        by_pos = np.empty(n)
        by_pos[(np.arange(n) + o.navi_dict()["offset_x"]) % n] = o.field("rho")
        errs.append(float(np.mean(np.abs(by_pos - (1 + 0.2 * np.sin(2 * np.pi * (x - u * steps * dt) / 4.0))))))
    return errs


def test_second_order_on_a_smooth_wave():
    e = wave_errors()
    assert e[0] / e[1] > 3.4 and e[1] / e[2] > 3.4, e


PLANTS = {
    # the right state of a face read at the left cell (a half offset placed wrong)
    "face": ("UR[c] = ul_x[c][i + 1] - ((0.5 * DT) * divf[c][i + 1]);",
             "UR[c] = ul_x[c][i] - ((0.5 * DT) * divf[c][i]);"),
    # the limiter's branch selection swapped
    "limiter": ("const double Delta = f > 0.0 ? DP : DN;", "const double Delta = f > 0.0 ? DN : DP;"),
    # the half-step predictor with the wrong sign
    "predictor": ("UL[c] = ur_x[c][i] - ((0.5 * DT) * divf[c][i]);",
                  "UL[c] = ur_x[c][i] + ((0.5 * DT) * divf[c][i]);"),
}


@pytest.mark.parametrize("plant", sorted(PLANTS))
def test_planted_mistake_fails(plant):
    src = open(SRC).read()
    a, b = PLANTS[plant]
    assert src.count(a) == 1
    o = oracle("hllc1d_400", src.replace(a, b), plant)
    run(o, "hllc1d_400")
    _, want = fixture("hllc1d_400")
    assert not np.array_equal(o.field("rho"), want["rho"])
    l1, _, _ = tube_l1(o.field("rho"), o.navi_dict()["offset_x"])
    if np.isfinite(l1) and l1 <= 0.006:  # the shock tube does not see it: the order must
        e = wave_errors(src.replace(a, b), plant)
        assert min(e[0] / e[1], e[1] / e[2]) < 3.4, (l1, e)
