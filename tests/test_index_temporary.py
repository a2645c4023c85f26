"""Index reads inside a manifest temporary (programs/indextmp2d.fmr).

The reference evaluates a temporary in its own loop nest over its own extent
N + 2s - size(range), with its own loop index (Templates.hs:442-457, 481), and
a consumer at index i reads it at i + d + off(consumer) - off(temporary)
(:474): a temporary read at a neighbour carries that nest's index.  The
emitter recomputes such a temporary where it is read, its index reads shifted
by the same constant (``manifest.inline_temporaries``).

``mm_interp.step_nests`` restates the generated C literally (buff with the
low halo, every node over its own extent).  It is pinned by: the
interpreter's periodic evaluation for programs whose temporaries read no
index (the Euler HLLC program with 24 temporaries; the reference's 1-D
MUSCL-Hancock sample, whose fixture a hand transcription of the emitted C
reproduces); and for the index program by the closed form with the diffusion
switched off (r = 0): the new value at i is q[i - s] + c ((i + 1 + off_x) +
3 (j + 1 + off_y)), the temporary's nest index being i + 1.  The inlined
graph equals it on the CPU; the GPU library equals it bit for bit.
"""
import os

import numpy as np
import pytest

import mm_interp
from conftest import hip_lib
from formura_amd import build as B, inputs
from formura_amd.emit_hip import EmitError, LibEmitter, generate
from formura_amd.manifest import inline_temporaries

C = 1.0 / 4096


def program(name="indextmp2d.fmr", yml="indextmp2d_32.yaml"):
    return generate(os.path.join(B.PROG_DIR, name), os.path.join(B.PROG_DIR, yml), "x", "flat")[2]


def test_nest_restatement_equals_periodic_interpreter_without_index_temporaries():
    mm = program("euler3d.fmr", "euler3d_16.yaml")
    st = inputs.euler_state(4, (16, 16, 16))
    a = mm_interp.step_nests(mm.step, 1, st)
    b = mm_interp.step(mm.step, 1, st)
    for f in a:
        np.testing.assert_array_equal(a[f], b[f])


@pytest.mark.skipif(not B.reference_present(), reason="formura checkout not present")
def test_nest_restatement_reproduces_the_transcribed_sample():
    import json
    rel, yml, _ = B.SAMPLES["hllc1d_64"]
    mm = generate(os.path.join(B.REFERENCE_ROOT, rel), os.path.join(B.PROG_DIR, yml), "s", "flat")[2]
    gold = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")
    meta = json.load(open(os.path.join(gold, "sample_hllc1d_64.json")))
    with np.load(os.path.join(gold, "sample_hllc1d_64.npz")) as z:
        want = {k: z[k] for k in z.files}
    st = inputs.conserved_state(meta["seed"], tuple(meta["grid"]), meta["fields"], meta["gamma"])
    for _ in range(meta["steps"]):
        st = mm_interp.step_nests(mm.step, 2, st)
    for f in want:
        np.testing.assert_array_equal(st[f], want[f])


@pytest.mark.parametrize("off", [[0, 0], [7, 30], [31, 2]])
def test_closed_form_without_diffusion(off):
    mm = program()
    q = np.random.default_rng(5).random((32, 32))
    got = mm_interp.step_nests(mm.step, 1, {"q": q}, offsets=off, params={"r": 0.0})["q"]
    i = np.arange(32)[:, None]
    j = np.arange(32)[None, :]
    want = np.roll(q, (1, 1), (0, 1)) + C * ((i + 1 + off[0]) + 3 * (j + 1 + off[1]))
    np.testing.assert_allclose(got, want, rtol=0, atol=1e-15)
    # the periodic reading of the temporary (its position's index) differs
    per = mm_interp.step(mm.step, 1, {"q": q}, offsets=off, params={"r": 0.0})["q"]
    assert np.abs(per - want).max() > 1e-4


@pytest.mark.parametrize("off", [[0, 0], [5, 29]])
def test_inlined_graph_equals_nest_restatement(off):
    mm = program()
    q = np.random.default_rng(3).random((32, 32))
    a = mm_interp.step_nests(mm.step, 1, {"q": q}, offsets=off)["q"]
    b = mm_interp.step(inline_temporaries(mm.step), 1, {"q": q}, offsets=off)["q"]
    np.testing.assert_array_equal(a, b)


def test_refused_under_temporal_blocking(tmp_path):
    y = tmp_path / "tb.yaml"
    y.write_text("length_per_node: [1.0, 1.0]\ngrid_per_node: [32, 32]\n"
                 "temporal_blocking_interval: 2\ngrid_per_block: [12, 12]\n")
    with pytest.raises(EmitError):
        generate(os.path.join(B.PROG_DIR, "indextmp2d.fmr"), str(y), "x", "auto")


@pytest.mark.gpu
@pytest.mark.parametrize("steps", [1, 9])
def test_gpu_bit_exact(steps):
    mm = program()
    h = hip_lib("indextmp2d_32")
    h.init()
    q = {"q": np.random.default_rng(11).random((32, 32))}
    h.set_state(q)
    print("indextmp2d_32: %d steps" % steps, flush=True)
    off = [0, 0]
    for _ in range(steps):
        q = mm_interp.step_nests(mm.step, 1, q, offsets=off)
        off = [(o - 1) % 32 for o in off]
    h.forward(steps)
    assert [h.navi.offset_x, h.navi.offset_y] == off
    np.testing.assert_array_equal(h.field("q"), q["q"])
