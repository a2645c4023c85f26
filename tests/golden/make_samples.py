"""Writes the fixtures of the reference's sample programs (run from the repo root:
python tests/golden/make_samples.py; needs the formura checkout, FORMURA_REFERENCE).

The sample programs (sample/1d_hllc, sample/mhd_1d_hlld, sample/mhd_3d_hlld,
sample/3d_hllc of formura v2.3.2) are compiled by this package's front-end and
their step graphs are evaluated by the numpy interpreter (tests/mm_interp.py:
IEEE double, one rounding per operation, the reference's placement).  Before a
fixture is written the run is checked against what the mathematics fixes:

* conservation: a finite-volume update U' = U - dt (F_i - F_{i-1})/dx on a
  periodic grid keeps every conserved total (all fields but the GLM cleaning
  potential psi, which is damped) to rounding;
* a uniform state is a fixed point, bit for bit (equal fluxes cancel exactly);
* the shock tube of sample/1d_hllc, run from the program's own init graph at
  the reference's configuration, stays within a bounded L1 distance of the
  reference's exact Riemann solution (sample/1d_hllc/ref_data/exact.dat) in
  the tube's interior, and the drifting frame's cell map gives a smaller
  distance than the same map shifted by one or two cells.

Each fixture holds the final state (the drifting frame of the reference:
index i holds position (i + offset) mod N) in ``sample_<name>.npz`` and the
run's parameters in ``sample_<name>.json``.  The exact solution is kept,
subsampled to the 400 cell positions, in ``hllc1d_exact.txt``.
"""
import json
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)
sys.path.insert(0, os.path.dirname(HERE))
import mm_interp  # noqa: E402
from formura_amd import build as B, inputs  # noqa: E402
from formura_amd.emit import lib_description  # noqa: E402
from formura_amd.emit_hip import generate  # noqa: E402

GAMMA = {"hllc": 7.0 / 5.0, "mhd": 5.0 / 3.0}

# name -> (seed or None for the program's init graph, steps)
CASES = {
    "hllc1d_64": (3, 20),
    "mhd1d_64": (3, 20),
    "hllc3d_8": (3, 6),
    "mhd3d_8": (3, 6),
    "hllc1d_400": (None, 100),
}


def front_end(name):
    rel, yml, _ = B.SAMPLES[name]
    _, _, mm = generate(os.path.join(B.REFERENCE_ROOT, rel), os.path.join(B.PROG_DIR, yml), name, "flat")
    return mm, lib_description(mm.om, name, "")


def gamma_of(name):
    return GAMMA["mhd" if name.startswith("mhd") else "hllc"]


def initial_state(name, mm, desc):
    seed, _ = CASES[name]
    shape = tuple(desc["grid"])
    fields = [f["name"] for f in desc["fields"]]
    if seed is None:
        return mm_interp.init(mm.init, shape)
    return inputs.conserved_state(seed, shape, fields, gamma_of(name))


def run(mm, desc, state, steps):
    s = desc["sleeve"]
    st = {k: v.copy() for k, v in state.items()}
    for _ in range(steps):
        st = mm_interp.step(mm.step, s, st)
    return st


def check_invariants(name, mm, desc, state, final):
    for f in final:
        if f == "psi":
            continue
        t0, t1 = float(np.sum(state[f])), float(np.sum(final[f]))
        scale = float(np.sum(np.abs(state[f]))) + 1e-300
        assert abs(t1 - t0) <= 1e-13 * scale, (name, f, t0, t1)
    uni = {f: np.full_like(v, float(v.flat[0])) for f, v in state.items()}
    out = run(mm, desc, uni, 2)
    for f in out:
        if f == "psi":
            continue
        assert np.array_equal(out[f], uni[f]), (name, "uniform state moved", f)


def exact_hllc1d():
    """The reference's exact solution (x, rho, u, p, e per line), sampled at the
    cell positions x = 0.01 k of the 400-cell run (its lines sit at 0.0005 + 0.001 j)."""
    path = os.path.join(B.REFERENCE_ROOT, "sample/1d_hllc/ref_data/exact.dat")
    ex = np.loadtxt(path)
    return ex[::10]


def l1_vs_exact(final, steps, exact, shift=0):
    """Mean |rho - rho_exact| over the cells 150..349 (x in [1.5, 3.5)), the tube's
    interior: on the periodic grid the wrap at x = 0 = 4.0 is a second Riemann
    problem, which the reference's infinite-tube solution does not contain.
    Output index i holds cell (i + offset) mod 400, offset = -2*steps (sleeve 2);
    ``shift`` moves that map, so the true placement must give the least error."""
    n = 400
    k = (np.arange(n) + (-2 * steps) % n + shift) % n
    rho = np.empty(n)
    rho[k] = final["rho"]
    return float(np.mean(np.abs(rho - exact[:, 1])[150:350]))


def main():
    if not B.reference_present():
        sys.exit("formura checkout not found at %s (set FORMURA_REFERENCE)" % B.REFERENCE_ROOT)
    exact = exact_hllc1d()
    np.savetxt(os.path.join(HERE, "hllc1d_exact.txt"), exact, fmt="%.6f",
               header="sample/1d_hllc/ref_data/exact.dat of formura v2.3.2, every 10th line "
                      "(x rho u p e at the 400 cell positions of the 4.0-long grid)")
    for name, (seed, steps) in CASES.items():
        mm, desc = front_end(name)
        state = initial_state(name, mm, desc)
        final = run(mm, desc, state, steps)
        meta = {"program": name, "source": B.SAMPLES[name][0], "seed": seed, "steps": steps,
                "gamma": gamma_of(name), "fields": sorted(final), "grid": desc["grid"],
                "offset": [(-desc["sleeve"] * steps) % n for n in desc["grid"]]}
        if seed is not None:
            check_invariants(name, mm, desc, state, final)
            meta["checked"] = "conservation of every total but psi to 1e-13; uniform state fixed bit for bit"
        else:
            err = l1_vs_exact(final, steps, exact)
            others = [l1_vs_exact(final, steps, exact, sh) for sh in (-2, -1, 1, 2)]
            assert err < 0.01 and err < min(others), (err, others)
            meta["checked"] = ("mean |rho - rho_exact| = %.4f on x in [1.5, 3.5) against the reference's exact "
                               "solution; frame shifts of -2..2 cells give %s" % (err, ", ".join("%.4f" % o for o in others)))
        np.savez_compressed(os.path.join(HERE, "sample_%s.npz" % name), **{f: final[f] for f in final})
        with open(os.path.join(HERE, "sample_%s.json" % name), "w") as fh:
            json.dump(meta, fh, indent=1)
        print("wrote", name, meta["checked"], flush=True)


if __name__ == "__main__":
    main()
