"""Writes the golden fixtures of the oracle (run: python tests/golden/make_golden.py).

Each fixture is the oracle's state after a seeded run: a few probe values
(as exact hex floats), the per-field sums and the Navi offset.  The values
were cross-checked against the closed-form and restatement tests of
tests/test_oracle.py when generated; the fixtures then pin the oracle against
accidental change.
"""
import json
import os
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.dirname(HERE))
from conftest import layout_of, oracle_lib  # noqa: E402
from formura_amd import inputs  # noqa: E402

CASES = [("heat3d_16", 11, 9), ("gs_64", 7, 40)]

for name, seed, steps in CASES:
    o = oracle_lib(name, layout_of(name))
    o.init()
    if name.startswith("gs"):
        o.set_state(inputs.grayscott_perturb(o.field("u").copy(), o.field("v").copy(), seed))
    else:
        o.set_state(inputs.heat_field(seed, o.grid))
    o.forward(steps)
    probes = {}
    for f in o.fields:
        a = o.field(f)
        idxs = [tuple(0 for _ in a.shape), tuple(n // 2 for n in a.shape), tuple(n - 1 for n in a.shape),
                tuple((3 * k + 1) % n for k, n in enumerate(a.shape))]
        probes[f] = [[list(i), float(a[i]).hex()] for i in idxs]
    out = {"program": name, "seed": seed, "steps": steps, "offset_x": o.navi.offset_x,
           "probes": probes, "sums": {f: float(o.field(f).sum()) for f in o.fields}}
    with open(os.path.join(HERE, "%s_seed%d_steps%d.json" % (name, seed, steps)), "w") as fh:
        json.dump(out, fh, indent=1)
    print("wrote", name)
