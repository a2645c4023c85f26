"""Index reads in a temporally blocked step (``temporal_blocking_interval`` > 1).

Reference semantics (Templates.hs:311-381, 481): inside one Formura_Forward
every one of the nt sub-steps calls Formura_Step with the Navi of the
Forward's start and the block's floor offset, so an index reads
``i + offset + block_offset`` with i the block-local output index -- the same
offset for all sub-steps, unlike nt separate steps.  ``mm_interp.forward_tb``
restates the blocked update literally (floor, blocks from high to low
positions, walls between neighbouring blocks).  The restatement is pinned on
the CPU by (1) equality with nt plain steps for a step without index reads,
(2) a closed form for the step with its diffusion switched off (r = 0): the
final cell G holds the start value of cell G - s nt plus c times the indices
G + s (nt - 1 - it) + offset of each sub-step it, and (3) equality with an
independent statement of the GPU scheme: level j of a Forward over unwrapped
coordinates g in [-2s(nt - j), N), an index reading g + 2s(nt - j) + offset.
On the GPU the tile2d kernel (2-D, two tile shapes) and the stream3d kernel
(3-D) with k = nt (one Forward per launch) give the restatement bit for bit,
state and Navi.
"""
import os

import numpy as np
import pytest

import mm_interp
from conftest import hip_lib
from formura_amd import build as B
from formura_amd.emit_hip import generate, EmitError
from formura_amd.manifest import calc_range

N, S, NT, GPB = 32, 1, 2, 12
C = 1.0 / 4096


def program(yml="index2d_32.yaml"):
    return generate(os.path.join(B.PROG_DIR, "index2d.fmr"), os.path.join(B.PROG_DIR, yml), "index2d", "flat")[2]


def test_restatement_without_index_equals_plain_steps():
    mm = generate(os.path.join(B.PROG_DIR, "hooks2d.fmr"), os.path.join(B.PROG_DIR, "hooks2d_32.yaml"),
                  "h", "flat")[2]
    q = {"q": np.random.default_rng(1).random((N, N))}
    nb, off = q, [0, 0]
    for t in range(NT):
        nb = mm_interp.step(mm.step, S, nb, offsets=off)
        off = [(o - S) % N for o in off]
    tb = mm_interp.forward_tb(mm.step, S, q, NT, GPB, [0, 0])
    np.testing.assert_array_equal(tb["q"], nb["q"])


@pytest.mark.parametrize("off", [[0, 0], [30, 7], [5, 31]])
def test_restatement_closed_form_without_diffusion(off):
    mm = program()
    q0 = np.random.default_rng(3).random((N, N))
    got = mm_interp.forward_tb(mm.step, S, {"q": q0}, NT, GPB, off, params={"r": 0.0})["q"]
    G = np.arange(N)
    want = q0[np.ix_((G - S * NT) % N, (G - S * NT) % N)].copy()
    for it in range(NT):
        ix = (G + S * (NT - 1 - it) + off[0]).astype(np.float64)
        iy = (G + S * (NT - 1 - it) + off[1]).astype(np.float64)
        want = (want + 0.0) + C * (ix[:, None] + 3.0 * iy[None, :])
    np.testing.assert_array_equal(got, want)


def _trapezoid(g, q, offT):
    rng = calc_range(g)
    n = next(iter(q.values())).shape[0]
    dim = len(offT)
    L = {f: v[np.ix_(*[np.arange(-2 * S * NT, n) % n] * dim)] for f, v in q.items()}
    for j in range(1, NT + 1):
        size = n + 2 * S * (NT - j)
        # local index l is g = l - 2s(nt - j); the index reads g + 2s(nt - j) + offset = l + offset
        index = [np.arange(size, dtype=np.float64) + float(offT[a]) for a in range(dim)]
        L = {g[k].micro[-1].inst[1]: np.array(np.broadcast_to(np.asarray(
            mm_interp._eval_block(g[k], L, [rng[k].offset] * dim, [size] * dim, None, index), dtype=np.float64),
            [size] * dim)) for k in sorted(g)}
    return L


def test_restatement_equals_the_unwrapped_level_scheme():
    g = program().step
    q = {"q": np.random.default_rng(2).random((N, N))}
    a, b, off = dict(q), dict(q), [0, 0]
    for fw in range(5):
        a = mm_interp.forward_tb(g, S, a, NT, GPB, off)
        b = _trapezoid(g, b, off)
        np.testing.assert_array_equal(a["q"], b["q"])
        off = [(o - S * NT) % N for o in off]


def program3d():
    return generate(os.path.join(B.PROG_DIR, "index3d.fmr"), os.path.join(B.PROG_DIR, "index3d_16.yaml"),
                    "index3d", "flat")[2]


def test_restatement_equals_the_unwrapped_level_scheme_3d():
    g = program3d().step
    q = {"q": np.random.default_rng(5).random((16, 16, 16))}
    a, b, off = dict(q), dict(q), [0, 0, 0]
    for fw in range(3):
        a = mm_interp.forward_tb(g, S, a, NT, 10, off)
        b = _trapezoid(g, b, off)
        np.testing.assert_array_equal(a["q"], b["q"])
        off = [(o - S * NT) % 16 for o in off]


def test_schedules_that_cannot_reproduce_it_are_refused():
    for spec in ("flat", {"kind": "tile2d", "tx": 16, "ty": 16, "k": 1},
                 {"kind": "tile2d", "tx": 16, "ty": 16, "k": 2, "inplace": 1}):
        with pytest.raises(EmitError):
            generate(os.path.join(B.PROG_DIR, "index2d.fmr"), os.path.join(B.PROG_DIR, "index2d_32_tb.yaml"),
                     "index2d", spec)


@pytest.mark.gpu
@pytest.mark.parametrize("name,forwards,resident", [("index2d_32_tb", 1, False), ("index2d_32_tb", 7, True),
                                                     ("index2d_32_tb8", 5, False)])
def test_index_tb_bit_exact(name, forwards, resident):
    print("%s: %d forwards, resident=%s" % (name, forwards, resident), flush=True)
    g = program().step
    h = hip_lib(name)
    h.init()
    q = {"q": h.field("q").copy()}
    off = [h.navi_dict()["offset_x"], h.navi_dict()["offset_y"]]
    if resident:
        h.set_resident(True)
        h.upload()
        h.forward_steps(forwards)
        h.download()
        h.set_resident(False)
    else:
        h.forward(forwards)
    for _ in range(forwards):
        q = mm_interp.forward_tb(g, S, q, NT, GPB, off)
        off = [(o - S * NT) % N for o in off]
    nv = h.navi_dict()
    assert (nv["time_step"], nv["offset_x"], nv["offset_y"]) == (NT * forwards, off[0], off[1])
    got = h.field("q")
    if not np.array_equal(got, q["q"]):
        raise AssertionError("%s differs after %d forwards: max |diff| %g"
                             % (name, forwards, float(np.max(np.abs(got - q["q"])))))


@pytest.mark.gpu
@pytest.mark.parametrize("forwards,resident", [(1, False), (4, True)])
def test_index_tb_bit_exact_3d(forwards, resident):
    print("index3d_16_tb: %d forwards, resident=%s" % (forwards, resident), flush=True)
    g = program3d().step
    h = hip_lib("index3d_16_tb")
    h.init()
    q = {"q": h.field("q").copy()}
    off = [h.navi_dict()["offset_" + a] for a in "xyz"]
    if resident:
        h.set_resident(True)
        h.upload()
        h.forward_steps(forwards)
        h.download()
        h.set_resident(False)
    else:
        h.forward(forwards)
    for _ in range(forwards):
        q = mm_interp.forward_tb(g, S, q, NT, 10, off)
        off = [(o - S * NT) % 16 for o in off]
    nv = h.navi_dict()
    assert [nv["offset_" + a] for a in "xyz"] == off and nv["time_step"] == NT * forwards
    got = h.field("q")
    if not np.array_equal(got, q["q"]):
        raise AssertionError("index3d_16_tb differs after %d forwards: max |diff| %g"
                             % (forwards, float(np.max(np.abs(got - q["q"])))))
