"""Pins the Euler HLLC oracle (oracle/euler3d_ref.c, the transcription of the
C that formura emits for the first-order form of sample/3d_hllc/3d_hllc.fmr).

None of these re-types the oracle's arithmetic:

* the shock tube of sample/3d_hllc (rho 1 / 0.125, normal velocity 0.75 / 0,
  p 1 / 0.1, diaphragm at the middle of 400 cells over 4.0, dt = 0.002,
  100 steps, 3d_hllc.fmr:62-71 and :19-21) run along each axis in turn
  against the exact Riemann solution the reference ships
  (sample/3d_hllc/ref_data/exact.dat, subsampled in tests/golden/hllc1d_exact.txt),
  with a first-order L1 bound, the drifting frame's cell map, and the state
  uniform across the tube;
* every face flux of a random state against the HLLC flux of Toro
  (Riemann Solvers and Numerical Methods for Fluid Dynamics, 3rd ed.,
  sec. 10.4) evaluated in exact rational arithmetic (square roots to 50
  digits): the contact speed S* from the condition p*_L(S*) = p*_R(S*)
  (10.36), p* = p_L + rho_L (S_L - u_L)(S* - u_L), the star states (10.39)
  with E* written through p_K, and the four-way flux selection (10.38); the
  wave-speed bounds are the program's choice S_L = min(u_L,u_R) - max(c_L,c_R),
  S_R = max(u_L,u_R) + max(c_L,c_R) (3d_hllc.fmr:107-113);
* invariants: a uniform flow is fixed bit for bit, the five totals are
  conserved, a stationary contact is preserved (the HLLC property, Toro
  sec. 10.4), and a cyclic rotation of the axes rotates the result;
* planted mistakes: copies of the oracle with a sign flipped in S_M, in the
  star energy, in the star density or in p*, or with another wave-speed bound
  (Davis), each fail at least one of the checks above.
"""
import ctypes
import decimal
import os
from fractions import Fraction

import numpy as np
import pytest

from conftest import ROOT, layout_of, oracle_lib
from formura_amd import inputs
from formura_amd.runtime import FormuraLib

import build_oracle

GOLDEN = os.path.join(ROOT, "tests", "golden")
GAMMA = Fraction(7, 5)
ORACLE_SRC = os.path.join(ROOT, "oracle", "euler3d_ref.c")

# first-order L1 distance of the density from the exact solution over the
# tube's interior (positions 1.5 to 3.5): measured 0.00752 on every axis
TUBE_L1_BOUND = 0.0080


def _euler(name="euler3d_16"):
    o = oracle_lib(name, layout_of(name))
    o.init()
    return o


def _lib(name, src=None, tag=""):
    """The strict oracle build of configuration ``name``; with ``src`` a copy
    of the oracle with that source text (a planted mistake), built apart."""
    grid = {"euler3d_tube_x": [400, 4, 4], "euler3d_tube_y": [4, 400, 4], "euler3d_tube_z": [4, 4, 400]}
    lay = dict(layout_of("euler3d_16"))
    if name in grid:
        lay["grid"] = grid[name]
    else:
        lay = layout_of(name)
    out = None
    if src is not None:
        out = os.path.join(build_oracle.OUT, "planted", "liboracle_%s_%s.so" % (name, tag))
    o = FormuraLib(build_oracle.build(name, "strict", source_text=src, out=out), lay)
    o.init()
    return o


# ---------------------------------------------------------------------------
# the shock tube against the reference's exact solution
# ---------------------------------------------------------------------------

def tube_run(o, ax, steps=100):
    """The sample's tube along axis ``ax``: returns (density by position, the
    density by position under the frame shifted by -2..2 cells, whether the
    state is uniform across the tube)."""
    g = o.grid
    idx = np.indices(g)[ax]
    left = idx < g[ax] // 2
    rho = np.where(left, 1.0, 0.125)
    p = np.where(left, 1.0, 0.1)
    un = np.where(left, 0.75, 0.0)
    st = {"rho": rho, "e": p / 0.4 + 0.5 * rho * un * un}
    for k, f in enumerate(("rhou", "rhov", "rhow")):
        st[f] = rho * un if k == ax else np.zeros(g)
    o.set_state(st)
    o.forward(steps)
    off = o.navi_dict()["offset_" + "xyz"[ax]]
    assert off == (-steps) % g[ax]
    r = np.moveaxis(o.field("rho"), ax, 0)
    uniform = bool(np.all(r == r[:, :1, :1]))
    n = g[ax]
    pos = (np.arange(n) + off) % n
    by_pos = np.empty(n)
    by_pos[pos] = r[:, 0, 0]
    return by_pos, uniform


def tube_l1(by_pos):
    ex = np.loadtxt(os.path.join(GOLDEN, "hllc1d_exact.txt"))
    return float(np.mean(np.abs(by_pos - ex[:, 1])[150:350]))


@pytest.mark.parametrize("ax", [0, 1, 2])
def test_shock_tube_against_exact_solution(ax):
    o = _lib("euler3d_tube_" + "xyz"[ax])
    by_pos, uniform = tube_run(o, ax)
    assert uniform
    l1 = tube_l1(by_pos)
    assert l1 <= TUBE_L1_BOUND, l1
    # the drifting frame's map (index i at position (i + offset) mod N) is the
    # one that fits: the same run read one or two cells off fits worse
    for s in (-2, -1, 1, 2):
        assert tube_l1(np.roll(by_pos, s)) > l1, s


# ---------------------------------------------------------------------------
# single-face fluxes in exact rational arithmetic
# ---------------------------------------------------------------------------

def _sqrt(x: Fraction) -> Fraction:
    with decimal.localcontext() as ctx:
        ctx.prec = 50
        r = (decimal.Decimal(x.numerator) / decimal.Decimal(x.denominator)).sqrt()
    return Fraction(r)


def toro_hllc(left, right):
    """HLLC flux (Toro sec. 10.4) of conserved states (rho, m_n, m_t1, m_t2, E)
    in exact rationals; returns (flux, region) with region 0: S_L >= 0,
    1: S_L < 0 <= S*, 2: S* < 0 <= S_R, 3: S_R < 0."""
    def prim(U):
        rho, mn, mt1, mt2, E = U
        u, v, w = mn / rho, mt1 / rho, mt2 / rho
        p = (GAMMA - 1) * (E - rho * (u * u + v * v + w * w) / 2)
        return rho, u, v, w, p, E

    rL, uL, vL, wL, pL, EL = prim(left)
    rR, uR, vR, wR, pR, ER = prim(right)
    cL, cR = _sqrt(GAMMA * pL / rL), _sqrt(GAMMA * pR / rR)
    c = max(cL, cR)
    SL, SR = min(uL, uR) - c, max(uL, uR) + c

    def p_star_of(S, rK, uK, pK, SK):
        # the pressure behind the K wave for a contact moving at S (Toro 10.36)
        return pK + rK * (SK - uK) * (S - uK)

    # p*_L(S) = p*_R(S) is linear in S: a S + b = 0
    a = p_star_of(1, rL, uL, pL, SL) - p_star_of(0, rL, uL, pL, SL) \
        - (p_star_of(1, rR, uR, pR, SR) - p_star_of(0, rR, uR, pR, SR))
    b = p_star_of(0, rL, uL, pL, SL) - p_star_of(0, rR, uR, pR, SR)
    Ss = -b / a
    assert p_star_of(Ss, rL, uL, pL, SL) == p_star_of(Ss, rR, uR, pR, SR)

    def phys(rho, u, v, w, p, E):
        return [rho * u, rho * u * u + p, rho * v * u, rho * w * u, (E + p) * u]

    def star(rho, u, v, w, p, E, SK):
        f = rho * (SK - u) / (SK - Ss)
        return [f, f * Ss, f * v, f * w, f * (E / rho + (Ss - u) * (Ss + p / (rho * (SK - u))))]

    FL, FR = phys(rL, uL, vL, wL, pL, EL), phys(rR, uR, vR, wR, pR, ER)
    if SL >= 0:
        return FL, 0
    if Ss >= 0:
        UsL = star(rL, uL, vL, wL, pL, EL, SL)
        return [FL[k] + SL * (UsL[k] - left[k]) for k in range(5)], 1
    if SR >= 0:
        UsR = star(rR, uR, vR, wR, pR, ER, SR)
        return [FR[k] + SR * (UsR[k] - right[k]) for k in range(5)], 2
    return FR, 3


def _random_state(seed, g):
    """Conserved fields with strong jumps between neighbours, so that the faces
    cover all four regions of the flux (supersonic both ways included)."""
    rng = np.random.default_rng(seed)
    rho = rng.uniform(0.4, 2.0, g)
    p = rng.uniform(0.3, 2.0, g)
    vel = [rng.uniform(-2.5, 2.5, g) for _ in range(3)]
    st = {"rho": rho, "e": p / 0.4 + 0.5 * rho * (vel[0] ** 2 + vel[1] ** 2 + vel[2] ** 2)}
    for f, v in zip(("rhou", "rhov", "rhow"), vel):
        st[f] = rho * v
    return st


ROT = {0: ("rhou", "rhov", "rhow"), 1: ("rhov", "rhow", "rhou"), 2: ("rhow", "rhou", "rhov")}


def face_flux_errors(o, seed=11, faces=120):
    """Largest |oracle - exact| / (1 + |exact|) over ``faces`` random faces of
    each direction, and the count of faces per flux region."""
    g = o.grid
    st = _random_state(seed, g)
    o.set_state(st)
    o.forward(1)  # the temporaries now hold the fluxes of ``st``
    T = [n + 1 for n in g]
    size = T[0] * T[1] * T[2]
    rng = np.random.default_rng(seed + 1)
    worst, regions = 0.0, [0, 0, 0, 0]
    for ax in range(3):
        names = ("rho",) + ROT[ax] + ("e",)
        tmp = {f: np.ctypeslib.as_array((ctypes.c_double * size).in_dll(o.lib, "flux_%s_%s" % ("xyz"[ax], f)))
               .reshape(T) for f in ("rho", "rhou", "rhov", "rhow", "e")}
        for _ in range(faces):
            j = [int(rng.integers(0, t)) for t in T]
            # buff[j] holds data[(j - 2s) mod N] (s = 1): the face between
            # buff[j] and buff[j + e_ax]
            a = tuple((j[k] - 2) % g[k] for k in range(3))
            b = tuple((j[k] - 2 + (k == ax)) % g[k] for k in range(3))
            U_L = [Fraction(float(st[f][a])) for f in names]
            U_R = [Fraction(float(st[f][b])) for f in names]
            F, reg = toro_hllc(U_L, U_R)
            regions[reg] += 1
            for k, f in enumerate(names):
                got = float(tmp[f][tuple(j)])
                err = abs(Fraction(got) - F[k]) / (1 + abs(F[k]))
                worst = max(worst, float(err))
    return worst, regions


# the oracle's double rounding against the exact flux: measured below 1e-14
FACE_TOL = 1e-12


def test_face_fluxes_match_exact_hllc():
    o = _lib("euler3d_16")
    worst, regions = face_flux_errors(o)
    assert all(r > 0 for r in regions), regions
    assert worst <= FACE_TOL, worst


# ---------------------------------------------------------------------------
# invariants
# ---------------------------------------------------------------------------

def test_uniform_flow_invariant_bit_exact():
    o = _euler()
    g = o.grid
    st = {"rho": np.full(g, 1.3), "rhou": np.full(g, 0.2), "rhov": np.full(g, -0.1), "rhow": np.full(g, 0.05),
          "e": np.full(g, 2.9)}
    o.set_state(st)
    o.forward(4)
    for f in st:
        assert np.all(o.field(f) == st[f][0, 0, 0]), f


def test_conserves_totals():
    o = _euler()
    st = inputs.euler_state(3, o.grid)
    o.set_state(st)
    o.forward(5)
    for f in st:
        scale = np.abs(st[f]).sum()
        assert abs(o.field(f).sum() - st[f].sum()) < 1e-12 * scale, f


def test_stationary_contact_preserved():
    """HLLC resolves a stationary contact (density jumps, p uniform, u = 0)
    exactly up to rounding (Toro sec. 10.4): the state does not change."""
    o = _euler()
    g = o.grid
    rho = 1.0 + 0.5 * inputs.uniform(4, g)
    p = 1.0
    st = {"rho": rho, "rhou": np.zeros(g), "rhov": np.zeros(g), "rhow": np.zeros(g), "e": np.full(g, p / 0.4)}
    o.set_state(st)
    o.forward(3)
    # identity up to rounding, then the drift of 3 cells per axis
    np.testing.assert_allclose(o.field("rho"), np.roll(rho, (3, 3, 3), (0, 1, 2)), rtol=1e-13)
    for f in ("rhou", "rhov", "rhow"):
        assert np.abs(o.field(f)).max() < 1e-13
    np.testing.assert_allclose(o.field("e"), p / 0.4, rtol=1e-13)


def test_cyclic_axis_rotation_symmetry():
    """Rotating the state x->y->z (fields and momentum components) rotates the
    result: the y and z fluxes are the x flux on rotated components."""
    a = _euler()
    st = inputs.euler_state(5, a.grid)
    a.set_state(st)
    a.forward(2)
    out_a = a.state()
    # rotation new[a,b,c] = old[c,a,b]: old axis x becomes new axis z, y -> x, z -> y
    rot = lambda f: np.transpose(f, (1, 2, 0))
    st_b = {"rho": rot(st["rho"]), "e": rot(st["e"]), "rhow": rot(st["rhou"]), "rhou": rot(st["rhov"]),
            "rhov": rot(st["rhow"])}
    a.init()
    a.set_state(st_b)
    a.forward(2)
    out_b = a.state()
    pairs = [("rho", "rho"), ("e", "e"), ("rhou", "rhow"), ("rhov", "rhou"), ("rhow", "rhov")]
    for fa, fb in pairs:
        np.testing.assert_allclose(out_b[fb], rot(out_a[fa]), rtol=1e-12, atol=1e-14)


# ---------------------------------------------------------------------------
# planted mistakes: each must fail a pin
# ---------------------------------------------------------------------------

PLANTS = {
    # sign of the pressure jump in the contact speed S_M
    "sm_sign": ("return (((R->p - L->p) +", "return (((L->p - R->p) +"),
    # sign of the p u term of the star energy
    "es_sign": ("(((S - K->u) * K->e) - (K->p * K->u))", "(((S - K->u) * K->e) + (K->p * K->u))"),
    # sign of the velocity jump in p*
    "pst_sign": ("* (L->u - R->u))) /", "* (R->u - L->u))) /"),
    # sign in the star density
    "rs_sign": ("const double rs = (K->rho * (S - K->u)) / (S - S_M);",
                "const double rs = (K->rho * (S + K->u)) / (S - S_M);"),
    # Davis's bound S_L = min(u_L - c_L, u_R - c_R) in place of the program's
    "davis": ("return (L->u > R->u ? R->u : L->u) - max_c(L, R);",
              "return (L->u - sound_speed(L)) < (R->u - sound_speed(R)) ? (L->u - sound_speed(L)) "
              ": (R->u - sound_speed(R));"),
}


@pytest.mark.parametrize("plant", sorted(PLANTS))
def test_planted_mistake_fails_a_pin(plant):
    src = open(ORACLE_SRC).read()
    old, new = PLANTS[plant]
    assert src.count(old) == 1
    bad = src.replace(old, new)
    worst, _ = face_flux_errors(_lib("euler3d_16", bad, plant))
    o = _lib("euler3d_tube_y", bad, plant)
    by_pos, _ = tube_run(o, 1)
    l1 = tube_l1(by_pos)
    tube_fails = not np.isfinite(l1) or l1 > TUBE_L1_BOUND
    assert worst > 1e3 * FACE_TOL or tube_fails, (worst, l1)
