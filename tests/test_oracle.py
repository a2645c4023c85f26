"""Pins the oracle (the hand transcription of the reference's generated C).

Nothing here re-types the oracle's formula: the checks are closed forms of the
discrete operators (Fourier modes of the FTCS heat step and of the linearised
Gray-Scott step), invariants (constant fields, conserved sum, the (1,0) fixed
point), the textbook Gray-Scott reaction terms with the configuration's
parameters, the equivalence of the no-blocking and temporal-blocking Forward
paths, and the committed golden fixtures.
"""
import json
import math
import os

import numpy as np
import pytest

from conftest import ROOT, layout_of, oracle_lib
from formura_amd import inputs

GOLDEN = os.path.join(ROOT, "tests", "golden")


def _heat(name="heat3d_16"):
    o = oracle_lib(name, layout_of(name))
    o.init()
    return o


def _mode(n, m):
    i = np.arange(n)
    ph = 2 * np.pi * (m[0] * i[:, None, None] + m[1] * i[None, :, None] + m[2] * i[None, None, :]) / n
    return np.cos(ph)


@pytest.mark.parametrize("m", [(1, 0, 0), (0, 2, 0), (1, 1, 3), (8, 8, 8)])
def test_heat_fourier_mode_decay_and_drift(m):
    """q = cos(2 pi m.i/N) -> after k steps lambda^k * q at index i-k per axis,
    lambda = 1 - 4 r sum_a sin^2(pi m_a/N), r = dt/dx^2 = 1/8."""
    o = _heat()
    n = o.grid[0]
    q0 = _mode(n, m)
    o.set_state({"q": q0})
    lam = 1 - 4 * 0.125 * sum(math.sin(math.pi * ma / n) ** 2 for ma in m)
    for k in range(1, 4):
        o.forward(1)
        expect = lam ** k * np.roll(q0, shift=(k, k, k), axis=(0, 1, 2))
        np.testing.assert_allclose(o.field("q"), expect, rtol=0, atol=1e-13)
        assert o.navi.offset_x == (-k) % n and o.navi.time_step == k


def test_heat_constant_field_invariant_bit_exact():
    o = _heat()
    o.set_state({"q": np.full(o.grid, 0.3)})
    o.forward(5)
    assert np.all(o.field("q") == 0.3)


def test_heat_sum_conserved():
    o = _heat()
    q0 = inputs.uniform(9, o.grid)
    o.set_state({"q": q0})
    o.forward(10)
    assert abs(o.field("q").sum() - q0.sum()) < 1e-12 * q0.sum()


def test_heat_to_pos_follows_offset():
    o = _heat()
    o.forward(3)
    n = o.grid[0]
    for i in (0, 1, 7, 15):
        assert o.to_pos("x", i) == o.navi.space_interval_x * ((i + (-3) % n) % n)


@pytest.mark.parametrize("nb,tb,forwards", [("heat3d_16", "heat3d_16_tb", 3), ("heat3d_64", "heat3d_64_tb", 2),
                                              ("heat3d_24x40x56", "heat3d_24x40x56_tb", 3)])
def test_no_blocking_and_temporal_blocking_bit_identical(nb, tb, forwards):
    a = oracle_lib(nb, layout_of(nb))
    b = oracle_lib(tb, layout_of(tb))
    a.init()
    b.init()
    q0 = inputs.uniform(2, a.grid)
    a.set_state({"q": q0})
    b.set_state({"q": q0})
    nt = b.layout["steps_per_forward"]
    a.forward(forwards * nt)
    b.forward(forwards)
    assert a.navi_dict() == b.navi_dict()
    np.testing.assert_array_equal(a.field("q"), b.field("q"))


@pytest.mark.parametrize("m", [(1, 0, 0), (0, 3, 0), (2, 5, 7)])
def test_heat_fourier_mode_uneven_grid(m):
    """The same closed form on a 24 x 40 x 56 grid (length 1 per axis, so
    d_a = 1/N_a and dt = dx^2/8): lambda = 1 - 4 sum_a (dt/d_a^2) sin^2(pi m_a/N_a)
    with dt/d_a^2 = (N_a/N_x)^2/8; an axis or extent mix-up fails it."""
    o = _heat("heat3d_24x40x56")
    n = o.grid
    assert list(n) == [24, 40, 56]
    i = [np.arange(x) for x in n]
    q0 = np.cos(2 * np.pi * (m[0] * i[0][:, None, None] / n[0] + m[1] * i[1][None, :, None] / n[1]
                             + m[2] * i[2][None, None, :] / n[2]))
    o.set_state({"q": q0})
    lam = 1 - 4 * sum((na / n[0]) ** 2 / 8 * math.sin(math.pi * ma / na) ** 2 for ma, na in zip(m, n))
    for k in range(1, 4):
        o.forward(1)
        expect = lam ** k * np.roll(q0, shift=(k, k, k), axis=(0, 1, 2))
        # per-axis factors up to 0.68 (q/dz^2 ~ 3e3 before the dt product): rounding ~1e-13
        np.testing.assert_allclose(o.field("q"), expect, rtol=0, atol=1e-12)
        assert (o.navi.offset_x, o.navi.offset_y, o.navi.offset_z) == ((-k) % 24, (-k) % 40, (-k) % 56)


def _gs():
    o = oracle_lib("gs_64", layout_of("gs_64"))
    o.init()
    return o


F, K, DU, DV, DT, DX = 0.04, 0.06, 2e-5, 1e-5, 1.0, 0.01


def test_grayscott_trivial_fixed_point():
    o = _gs()
    o.set_state({"u": np.ones(o.grid), "v": np.zeros(o.grid)})
    o.forward(7)
    assert np.all(o.field("u") == 1.0) and np.all(o.field("v") == 0.0)


def test_grayscott_uniform_state_is_reaction_ode():
    """A uniform state has no diffusion: one step is the Gray-Scott reaction
    u' = u + dt(-u v^2 + F(1-u)), v' = v + dt(u v^2 - (F+k) v) (Pearson 1993)."""
    o = _gs()
    u, v = 0.6, 0.2
    o.set_state({"u": np.full(o.grid, u), "v": np.full(o.grid, v)})
    o.forward(1)
    np.testing.assert_allclose(o.field("u"), u + DT * (-u * v * v + F * (1 - u)), rtol=1e-15)
    np.testing.assert_allclose(o.field("v"), v + DT * (u * v * v - (F + K) * v), rtol=1e-15)


@pytest.mark.parametrize("m", [(1, 0), (3, 5)])
def test_grayscott_linear_u_mode(m):
    """v = 0, u = 1 + e cos: v stays 0 and u-1 decays by g = 1 - dt(Du 4/dx^2 sum sin^2 + F), drifting."""
    o = _gs()
    n = o.grid[0]
    i = np.arange(n)
    c = np.cos(2 * np.pi * (m[0] * i[:, None] + m[1] * i[None, :]) / n)
    e = 1e-3
    o.set_state({"u": 1 + e * c, "v": np.zeros(o.grid)})
    g = 1 - DT * (DU * 4 / DX ** 2 * sum(math.sin(math.pi * a / n) ** 2 for a in m) + F)
    o.forward(4)
    assert np.all(o.field("v") == 0)
    np.testing.assert_allclose(o.field("u") - 1, e * g ** 4 * np.roll(c, (4, 4), (0, 1)), rtol=0, atol=1e-15)


@pytest.mark.parametrize("m", [(2, 1), (0, 7)])
def test_grayscott_linear_v_mode(m):
    """u = 0, v = e cos: one step gives v' = e(1 - dt(Dv 4/dx^2 sum sin^2 + F + k)) cos, u' = dt F."""
    o = _gs()
    n = o.grid[0]
    i = np.arange(n)
    c = np.cos(2 * np.pi * (m[0] * i[:, None] + m[1] * i[None, :]) / n)
    e = 1e-2
    o.set_state({"u": np.zeros(o.grid), "v": e * c})
    g = 1 - DT * (DV * 4 / DX ** 2 * sum(math.sin(math.pi * a / n) ** 2 for a in m) + F + K)
    o.forward(1)
    np.testing.assert_allclose(o.field("v"), e * g * np.roll(c, (1, 1), (0, 1)), rtol=0, atol=1e-16)
    np.testing.assert_allclose(o.field("u"), DT * F, rtol=1e-15)


@pytest.mark.parametrize("m", [(1, 0), (0, 1), (5, 7)])
def test_grayscott_linear_u_mode_uneven_grid(m):
    """As above on a 72 x 120 grid (dx = 0.01 on both axes): the decay factor
    uses each axis's own extent."""
    o = oracle_lib("gs_72x120", layout_of("gs_72x120"))
    o.init()
    n = o.grid
    assert list(n) == [72, 120]
    i, j = np.arange(n[0]), np.arange(n[1])
    c = np.cos(2 * np.pi * (m[0] * i[:, None] / n[0] + m[1] * j[None, :] / n[1]))
    e = 1e-3
    o.set_state({"u": 1 + e * c, "v": np.zeros(n)})
    g = 1 - DT * (DU * 4 / DX ** 2 * sum(math.sin(math.pi * a / na) ** 2 for a, na in zip(m, n)) + F)
    o.forward(4)
    assert np.all(o.field("v") == 0)
    np.testing.assert_allclose(o.field("u") - 1, e * g ** 4 * np.roll(c, (4, 4), (0, 1)), rtol=0, atol=1e-15)


def test_grayscott_init_square():
    """init: u=1/2, v=1/4 on a centred square of side 2w-1, w = N/20 + 1/2 (7 cells at N=64)."""
    o = _gs()
    u = o.field("u")
    inside = np.argwhere(u == 0.5)
    assert len(inside) == 49
    assert inside.min(0).tolist() == [29, 29] and inside.max(0).tolist() == [35, 35]
    assert np.all(o.field("v")[u == 0.5] == 0.25) and np.all(o.field("v")[u == 1.0] == 0.0)


@pytest.mark.parametrize("case", ["heat3d_16_seed11_steps9", "gs_64_seed7_steps40"])
def test_golden_fixtures(case):
    g = json.load(open(os.path.join(GOLDEN, case + ".json")))
    name = g["program"]
    o = oracle_lib(name, layout_of(name))
    o.init()
    if name.startswith("gs"):
        o.set_state(inputs.grayscott_perturb(o.field("u").copy(), o.field("v").copy(), g["seed"]))
    else:
        o.set_state(inputs.heat_field(g["seed"], o.grid))
    o.forward(g["steps"])
    assert o.navi_dict()["offset_x"] == g["offset_x"]
    for f, probes in g["probes"].items():
        for idx, val in probes:
            assert o.field(f)[tuple(idx)] == float.fromhex(val), (f, idx)
        assert abs(float(o.field(f).sum()) - g["sums"][f]) <= 1e-9 * abs(g["sums"][f])
