"""The ``staged3d`` schedule's plan, checked on the CPU.

For a 3-D program whose manifest temporaries read only the state (the Euler
HLLC program's face fluxes, once the wave speeds they read are recomputed
inside them: ``"inline": "inner"``), ``Staged3DSchedule.analyse`` groups the
temporaries by the offsets at which the update reads them and derives the LDS
ring depths, halos and warm-up planes.  The kernel then addresses its rings
with ``slot = (plane - xs + 64*D) mod D``.  The test replays that bookkeeping
iteration by iteration -- which plane each ring slot holds after the loads and
group evaluations of iteration x -- and checks that every read the generated
code makes (group evaluations reading state planes, the update reading state
and temporaries) finds exactly the plane the flat schedule would read, for
every output plane of a chunk.  The GPU parity tests compare the kernels with
the oracle bit for bit (``tests/test_gpu_parity.py``).
"""
import os

import pytest

from formura_amd import build as B
from formura_amd.emit_hip import LibEmitter, Staged3DSchedule, candidate_specs, generate, EmitError


INNER = {"kind": "flat", "inline": "inner"}


def emitter(name="euler3d_32", spec=None):
    fmr, yml, _ = B.PROGRAMS[name]
    _, _, mm = generate(os.path.join(B.PROG_DIR, fmr), os.path.join(B.PROG_DIR, yml), name, "flat")
    return LibEmitter(mm, name, spec or (INNER if name.startswith("euler") else "flat")), mm


def test_wave_speeds_are_inner_temporaries():
    """S_L, S_R, S_M of each direction are read only by the fluxes: inlined
    into them, the fluxes read only the state, with their ranges unchanged."""
    from formura_amd.manifest import inner_temporaries
    em, mm = emitter()
    inner = inner_temporaries(mm.step)
    assert sorted(mm.step[k].name for k in inner) == sorted(["s_l", "s_r", "s_m"] * 3)
    assert len(em.step.temps) == 15
    assert all(m.inst[0] != "LoadCursor" for t in em.step.temps for m in em.step.mm[t].micro)
    with pytest.raises(EmitError):  # without the inlining the fluxes read other temporaries
        Staged3DSchedule.analyse(LibEmitter(mm, "euler3d_32", "flat"), 16, 16)


def test_euler_groups_and_rings():
    em, _ = emitter()
    G, st = Staged3DSchedule.analyse(em, 16, 16)
    # x-, y- and z-fluxes: five components each, read at two faces along their axis
    assert [len(g["temps"]) for g in G] == [5, 5, 5]
    assert [g["DG"] for g in G] == [2, 1, 1]
    assert [(g["RY"], g["RZ"]) for g in G] == [(16, 16), (17, 16), (16, 17)]
    assert (st["DS"], st["SRY"], st["SRZ"], st["W"]) == (2, 18, 18, 2)
    # 5 FP64 fields x 2 planes x 18 x 18 + 5 x (2 x 256 + 272 + 272) fluxes
    assert Staged3DSchedule.lds_bytes(em, 16, 16) == 8 * (5 * 2 * 324 + 5 * (2 * 256 + 272 + 272))


def _replay(em, ty, tz, cx):
    """Every ring read of one chunk resolves to the plane the flat schedule reads."""
    G, st = Staged3DSchedule.analyse(em, ty, tz)
    plan = em.step
    s = plan.sleeve
    DS = st["DS"]
    x0 = 3 * cx  # any chunk: the bookkeeping is relative to x0
    xs = x0 - st["W"]
    state_slot, group_slot = {}, [dict() for _ in G]
    checked = 0
    for x in range(xs, x0 + cx):
        P = x + st["Smax"]
        state_slot[(P - xs + 64 * DS) % DS] = P
        for gi, g in enumerate(G):
            if x < x0 - (g["DG"] - 1):
                continue
            Pg = x + g["oxmax"]
            for d in g["D"]:
                assert state_slot.get((Pg + d[0] - xs + 64 * DS) % DS) == Pg + d[0]
                checked += 1
            group_slot[gi][(Pg - xs + 64 * g["DG"]) % g["DG"]] = Pg
        if x < x0:
            continue
        gof = {t: gi for gi, g in enumerate(G) for t in g["temps"]}
        for k in plan.stores:
            base = plan.rng[k].offset - 2 * s
            for m in plan.mm[k].micro:
                if m.inst[0] == "LoadCursorStatic":
                    o = m.inst[1][0] + base
                    assert state_slot.get((x + o - xs + 64 * DS) % DS) == x + o
                    checked += 1
                elif m.inst[0] == "LoadCursor":
                    gi = gof[m.inst[2]]
                    dg = G[gi]["DG"]
                    o = m.inst[1][0] + base
                    assert group_slot[gi].get((x + o - xs + 64 * dg) % dg) == x + o
                    checked += 1
    return checked


@pytest.mark.parametrize("ty,tz,cx", [(16, 16, 8), (4, 64, 32), (8, 16, 4)])
def test_ring_bookkeeping(ty, tz, cx):
    em, _ = emitter()
    assert _replay(em, ty, tz, cx) > 0


def test_ring_bookkeeping_catches_a_short_warm_up():
    """The same replay with one warm-up plane fewer must find a stale slot."""
    em, _ = emitter()
    orig = Staged3DSchedule.analyse

    def short(em_, ty, tz):
        G, st = orig(em_, ty, tz)
        st = dict(st, W=st["W"] - 1)
        return G, st
    Staged3DSchedule.analyse = staticmethod(short)
    try:
        with pytest.raises(AssertionError):
            _replay(em, 16, 16, 8)
    finally:
        Staged3DSchedule.analyse = staticmethod(orig)


def test_applicability():
    em, _ = emitter()
    specs = [c for c in candidate_specs(emitter("euler3d_256", "flat")[0]) if c["kind"] == "staged3d"]
    assert specs and all(c.get("inline") == "inner" for c in specs)
    assert all(Staged3DSchedule.lds_bytes(emitter("euler3d_256")[0], c["ty"], c["tz"]) <= 160 * 1024
               for c in specs)
    heat, _ = emitter("heat3d_32")
    with pytest.raises(EmitError):
        LibEmitter(heat.mm, "heat3d_32", {"kind": "staged3d", "ty": 8, "tz": 8, "cx": 8})
    with pytest.raises(EmitError):  # temporaries recomputed in place leave nothing to stage
        LibEmitter(em.mm, "euler3d_32", {"kind": "staged3d", "ty": 8, "tz": 8, "cx": 8, "inline": 1})
