"""The optional ``first_step`` and ``filter`` hooks of the generated C API.

Reference semantics (Templates.hs:227-250, 252-270, 295-310, 336-346):
``Formura_Init`` runs the init graph (Setup) and then, if the program defines
``first_step``, one application of that graph through ``noBlocking`` (halo,
kernel with the current Navi, offsets -= sleeve0).  Each ``Formura_Forward``
runs the step (offsets -= sleeve), then, when ``time_step %
filter_interval == 0`` (before the increment), the filter graph through
``noBlocking`` (offsets -= filter sleeve), then ``time_step += 1``.  An index
``i`` inside a graph reads ``i + offset`` of the Navi the kernel is called
with (Templates.hs:481).

The program ``hooks2d.fmr`` uses both hooks.  The CPU restatement below drives
the front-end's graphs with the numpy interpreter in that order.  It is pinned
by closed forms: the hooks' averages and the diffusion step keep the total, so
the total after Init is sum(init) + d N^2, and every filter adds
c N sum_i (i + offset_x + 1) with the offset of that call, which fixes which
Forward calls run the filter and with which offsets.  The GPU test compares the
HIP library with the restatement bit for bit, state and Navi.
"""
import os

import numpy as np
import pytest

import mm_interp
from conftest import hip_lib
from formura_amd import build as B
from formura_amd.emit_hip import generate

NAME = "hooks2d_32"
N = 32
D, C, INTERVAL = 0.25, 1.0 / 1024, 3


def program():
    fmr, yml, _ = B.PROGRAMS[NAME]
    return generate(os.path.join(B.PROG_DIR, fmr), os.path.join(B.PROG_DIR, yml), NAME, "flat")[2]


class Restatement:
    """Init and Forward of the generated C, over the interpreter's graphs."""

    def __init__(self, mm):
        self.mm = mm
        cfg = mm.om.config
        self.s, self.s0, self.sf = cfg.sleeve, cfg.sleeve0, cfg.filter_sleeve
        self.time_step, self.off = 0, [0, 0]
        self.q = mm_interp.init(mm.init, (N, N))
        self.filter_calls = []
        self._apply(mm.first_step, self.s0)

    def _apply(self, g, s):
        self.q = mm_interp.step(g, s, self.q, offsets=list(self.off))
        self.off = [(o - s) % N for o in self.off]

    def forward(self):
        self._apply(self.mm.step, self.s)
        if self.time_step % INTERVAL == 0:
            self.filter_calls.append(self.off[0])
            self._apply(self.mm.filter, self.sf)
        self.time_step += 1


def test_restatement_closed_forms():
    mm = program()
    q0 = mm_interp.init(mm.init, (N, N))
    i, j = np.meshgrid(np.arange(N), np.arange(N), indexing="ij")
    np.testing.assert_array_equal(q0["q"], (i + 2 * j) / 64.0)
    r = Restatement(mm)
    total = float(q0["q"].sum()) + D * N * N
    assert abs(float(r.q["q"].sum()) - total) <= 1e-12 * total
    for k in range(10):
        r.forward()
    # filter calls at time_step 0, 3, 6, 9; before the k-th, Init's first step,
    # 3k + 1 steps and k filters have each moved the offset by one cell
    assert r.filter_calls == [(-1 - (3 * k + 1) - k) % N for k in range(4)]
    for off in r.filter_calls:
        total += C * N * (N * (N + 1) / 2 + N * off)
    assert abs(float(r.q["q"].sum()) - total) <= 1e-12 * total
    assert r.off == [(-1 - 10 - 4) % N] * 2 and r.time_step == 10


def test_filter_interval_must_be_a_multiple_of_the_blocking_interval():
    from formura_amd.config import ConfigException, read_config, convert_config
    import tempfile
    with tempfile.NamedTemporaryFile("w", suffix=".yaml", delete=False) as f:
        f.write("length_per_node: [1.0, 1.0]\ngrid_per_node: [32, 32]\nfilter_interval: 3\n"
                "temporal_blocking_interval: 2\ngrid_per_block: [6, 6]\n")
    try:
        with pytest.raises(ConfigException):
            convert_config(1, 1, 1, read_config(f.name))
    finally:
        os.unlink(f.name)


@pytest.mark.gpu
@pytest.mark.parametrize("forwards,resident", [(1, False), (10, False), (10, True)])
def test_hooks_bit_exact(forwards, resident):
    print("hooks2d_32: %d forwards, resident=%s" % (forwards, resident), flush=True)
    mm = program()
    h = hip_lib(NAME)
    h.init()
    r = Restatement(mm)
    np.testing.assert_array_equal(h.field("q"), r.q["q"])
    assert h.navi_dict()["offset_x"] == r.off[0]
    if resident:
        h.set_resident(True)
        h.upload()
        h.forward_steps(forwards)
        h.download()
        h.set_resident(False)
    else:
        h.forward(forwards)
    for _ in range(forwards):
        r.forward()
    nv = h.navi_dict()
    assert (nv["time_step"], nv["offset_x"], nv["offset_y"]) == (r.time_step, r.off[0], r.off[1])
    got = h.field("q")
    if not np.array_equal(got, r.q["q"]):
        raise AssertionError("hooks2d_32 differs after %d forwards: max |diff| %g"
                             % (forwards, float(np.max(np.abs(got - r.q["q"])))))


@pytest.mark.gpu
@pytest.mark.parametrize("forwards", [1, 5])
def test_hooks_under_temporal_blocking_bit_exact(forwards):
    """The same program with temporal_blocking_interval 3 and filter_interval 6
    (hooks2d_32_tb.yaml): a Forward advances 3 steps (offsets -= 3 s at its
    end, Templates.hs:325-328), then runs the filter when time_step, before
    the += 3, is a multiple of 6 (:329-333).  The step has no index reads, so
    the blocked update equals 3 plain steps (checked for the restatement in
    tests/test_index_tb.py); the filter reads the offsets after the shift."""
    print("hooks2d_32_tb: %d forwards" % forwards, flush=True)
    fmr, yml, _ = B.PROGRAMS["hooks2d_32_tb"]
    mm = generate(os.path.join(B.PROG_DIR, fmr), os.path.join(B.PROG_DIR, "hooks2d_32.yaml"), NAME, "flat")[2]
    h = hip_lib("hooks2d_32_tb")
    h.init()
    r = Restatement(mm)
    np.testing.assert_array_equal(h.field("q"), r.q["q"])
    filters = []
    for _ in range(forwards):
        for _ in range(3):
            r._apply(mm.step, r.s)
        if r.time_step % 6 == 0:
            filters.append(r.off[0])
            r._apply(mm.filter, r.sf)
        r.time_step += 3
    h.forward(forwards)
    nv = h.navi_dict()
    assert (nv["time_step"], nv["offset_x"], nv["offset_y"]) == (r.time_step, r.off[0], r.off[1])
    assert len(filters) == (forwards + 1) // 2
    np.testing.assert_array_equal(h.field("q"), r.q["q"])
