"""FP64 totals of the device state, and the mixed-precision Euler form.

* Formura_HIP_Totals equals the host's exactly rounded sum of the state
  (math.fsum) to 1e-13;
* the heat step conserves the total of q and the Euler step the five
  conserved totals (flux differences telescope on a periodic grid), to
  rounding -- also in the mixed form, whose update chain is FP64;
* mixed Euler (FP32 flux arithmetic; FP64 wave speeds, positivity checks and
  update) stays within 1e-5 relative max-norm of the FP64 oracle per field, with
  no positivity violation, and its flat and LDS-streamed schedules agree bit
  for bit.
"""
import math

import numpy as np
import pytest

from conftest import hip_lib, layout_of, oracle_lib
from formura_amd import inputs

pytestmark = pytest.mark.gpu


def _seeded(name, state):
    h = hip_lib(name)
    h.init()
    h.set_state(state(h))
    h.set_resident(True)
    h.upload()
    return h


def _gs(h):
    return inputs.grayscott_perturb(h.field("u").copy(), h.field("v").copy(), 3)


def _heat(h):
    return {"q": inputs.uniform(4, h.grid)}


def _euler(h):
    return inputs.euler_state(1, h.grid)


@pytest.mark.parametrize("name,state", [("gs_64", _gs), ("heat3d_32", _heat), ("euler3d_16", _euler)])
def test_totals_equal_host_sum(name, state):
    h = _seeded(name, state)
    tot = h.totals()
    h.download()
    for f in h.fields:
        want = math.fsum(h.field(f).ravel().tolist())
        assert abs(tot[f] - want) <= 1e-13 * max(abs(want), 1.0), (f, tot[f], want)
    h.set_resident(False)
    h.finalize()


@pytest.mark.parametrize("name,state,steps,tol", [
    ("heat3d_64", _heat, 100, 1e-13),
    ("euler3d_32_inline_stream", _euler, 20, 1e-13),
    ("euler3d_32_mixed_stream", _euler, 20, 1e-13),
    ("euler3d_32_staged", _euler, 20, 1e-13),
    ("euler3d_32_mixed_staged", _euler, 20, 1e-13),
])
def test_totals_conserved(name, state, steps, tol):
    h = _seeded(name, state)
    t0 = h.totals()
    h.forward_steps(steps)
    t1 = h.totals()
    scale = max(abs(v) for v in t0.values())
    for f in h.fields:
        assert abs(t1[f] - t0[f]) <= tol * scale, (f, t0[f], t1[f])
    h.set_resident(False)
    h.finalize()


def test_mixed_euler_close_to_fp64_oracle():
    steps = 40
    o = oracle_lib("euler3d_32", layout_of("euler3d_32"))
    o.init()
    st = inputs.euler_state(1, o.grid)
    o.set_state(st)
    o.forward(steps)
    h = _seeded("euler3d_32_mixed_stream", lambda h: st)
    h.forward_steps(steps)
    h.download()
    assert h.positivity_violations() == 0
    errs = {}
    for f in o.fields:
        a, b = o.field(f), h.field(f)
        errs[f] = float(np.max(np.abs(a - b)) / np.max(np.abs(a)))
    print("mixed Euler 32^3, %d steps: relative max-norm error vs FP64 per field %s" % (steps, errs))
    assert max(errs.values()) <= 1e-5, errs
    assert max(errs.values()) > 0  # the FP32 flux arithmetic is really in use
    h.set_resident(False)
    h.finalize()


def test_mixed_flat_and_streamed_agree():
    st = None
    out = []
    for name in ("euler3d_32_mixed", "euler3d_32_mixed_stream"):
        h = _seeded(name, lambda h: inputs.euler_state(2, h.grid))
        h.forward_steps(9)
        h.download()
        out.append({f: h.field(f).copy() for f in h.fields})
        h.set_resident(False)
        h.finalize()
    for f in out[0]:
        np.testing.assert_array_equal(out[0][f], out[1][f])


def test_mixed_staged_matches_materialized_flat():
    """The mixed form with materialized flux temporaries (the wave speeds
    recomputed inside them, precision decided over that graph, FP32 flux
    storage): the staged schedule, which keeps the fluxes in LDS rings, gives
    the flat schedule's state bit for bit.  Both, and the form with every
    temporary materialized (S_L, S_R, S_M in their own FP64 buffers), are
    within 1e-5 of the FP64 oracle, with no positivity violation."""
    steps = 12
    o = oracle_lib("euler3d_32", layout_of("euler3d_32"))
    o.init()
    st = inputs.euler_state(3, o.grid)
    o.set_state(st)
    o.forward(steps)
    out = []
    for name in ("euler3d_32_mixed_inner", "euler3d_32_mixed_staged", "euler3d_32_mixed_mat"):
        print("mixed Euler 32^3 %s, %d steps" % (name, steps), flush=True)
        h = _seeded(name, lambda h: st)
        h.forward_steps(steps)
        h.download()
        assert h.positivity_violations() == 0
        out.append({f: h.field(f).copy() for f in h.fields})
        h.set_resident(False)
        h.finalize()
    for f in out[0]:
        np.testing.assert_array_equal(out[0][f], out[1][f])
        for res in out:
            err = float(np.max(np.abs(res[f] - o.field(f))) / np.max(np.abs(o.field(f))))
            assert 0 < err <= 1e-5, (f, err)


def test_mixed_configuration_256_200_steps():
    """The configuration's mixed variant: Euler HLLC 256^3, 200 steps, the
    seeded state, FP32 fluxes (recomputed in place, and staged in LDS) against
    the FP64 library (itself bit-exact with the oracle): relative max-norm
    error per field within 1e-5, the five totals conserved to 1e-13 of their
    scale, no positivity violation."""
    steps = 200
    st = None
    out = {}
    for name in ("euler3d_256", "euler3d_256_mixed", "euler3d_256_mixed_staged"):
        print("Euler 256^3 %s, %d steps" % (name, steps), flush=True)
        h = _seeded(name, lambda h: inputs.euler_state(1, h.grid))
        t0 = h.totals()
        h.forward_steps(steps)
        t1 = h.totals()
        h.download()
        scale = max(abs(v) for v in t0.values())
        assert max(abs(t1[f] - t0[f]) for f in t0) <= 1e-13 * scale
        assert h.positivity_violations() == 0
        out[name] = {f: h.field(f).copy() for f in h.fields}
        h.set_resident(False)
        h.finalize()
    a = out["euler3d_256"]
    for name in ("euler3d_256_mixed", "euler3d_256_mixed_staged"):
        for f in a:
            err = float(np.max(np.abs(out[name][f] - a[f])) / np.max(np.abs(a[f])))
            print("  %s %s: %.3g" % (name, f, err), flush=True)
            assert 0 < err <= 1e-5, (name, f, err)
