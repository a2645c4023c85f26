"""The flat schedule's grouping of manifest temporaries (``emit_hip.temp_groups``):
one kernel per dependency level.  Every temporary is in exactly one group, a
group reads only temporaries of lower levels, and every temporary a group
reads is among the inputs its kernel is given (``temp_group_inputs``) -- on the
Euler program (one level of 15 fluxes) and, where the formura checkout is
present, on the reference's MUSCL-HLLC and HLLD samples (5 levels each)."""
import os

import pytest

from formura_amd import build as B
from formura_amd.emit_hip import LibEmitter, generate, temp_groups, temp_group_inputs


def _plan(fmr, yml, name):
    _, _, mm = generate(fmr, yml, name, "flat")
    return LibEmitter(mm, name, "flat").step


def _check(plan):
    groups = temp_groups(plan)
    seen = [k for g in groups for k in g]
    assert sorted(seen) == sorted(plan.temps) and len(seen) == len(set(seen))
    done = set()
    for g in groups:
        inputs = set(temp_group_inputs(plan, g))
        for k in g:
            reads = {m.inst[2] for m in plan.mm[k].micro if m.inst[0] == "LoadCursor"}
            assert reads <= done, (k, reads - done)      # computed by an earlier launch
            assert reads <= inputs, (k, reads - inputs)  # and handed to this kernel
        done |= set(g)
    return groups


def test_euler_temporaries_in_three_levels():
    """S_L, S_R (read the state), S_M (reads S_L, S_R), then the 15 flux
    components (read all three): one kernel per level."""
    plan = _plan(os.path.join(B.PROG_DIR, "euler3d.fmr"), os.path.join(B.PROG_DIR, "euler3d_16.yaml"), "e")
    assert [len(g) for g in _check(plan)] == [6, 3, 15]


@pytest.mark.skipif(not B.reference_present(), reason="formura checkout not present")
@pytest.mark.parametrize("name,levels", [("hllc3d_8", 5), ("mhd3d_8", 5)])
def test_sample_levels(name, levels):
    rel, yml, _ = B.SAMPLES[name]
    plan = _plan(os.path.join(B.REFERENCE_ROOT, rel), os.path.join(B.PROG_DIR, yml), name)
    assert len(_check(plan)) == levels
