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
* the shock tubes of sample/1d_hllc, sample/3d_hllc (MUSCL-Hancock, along y)
  and sample/mhd_1d_hlld (Ryu-Jones 2a, HLLD), each run from the program's own
  init graph at the reference's configuration, stays within a bounded L1 distance of the
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


# The reference's own configurations (one rank over the whole grid) run from the
# program's init graph and compared with the exact solutions the reference ships:
# name -> (steps, exact-solution file, [(field, column of the file)], cells the
# cells compared (first, count), L1 bound per field, cell of the file's first row)
EXACT_CASES = {
    "mhd1d_3200": (1600, "sample/mhd_1d_hlld/ref_data/Bx2.dat", [("rho", 3), ("B_y", 9)], (1200, 800), 0.002, 1200),
    "hllc3d_shock": (100, "sample/1d_hllc/ref_data/exact.dat", [("rho", 1)], (150, 200), 0.01, 0),
}


def field_hashes(st):
    import hashlib
    return {f: hashlib.sha256(np.ascontiguousarray(st[f], dtype=np.float64).tobytes()).hexdigest() for f in sorted(st)}


def profile(name, st):
    """The first axis' profile of a 1-D run, the y profile (x = z = 0) of the 3-D tube."""
    return {f: (v if v.ndim == 1 else v[0, :, 0]) for f, v in st.items()}


def exact_l1(name, prof, steps, exact_rows, shift=0):
    """Mean |f - f_exact| over the cells the exact solution covers, per compared field.
    Output index i holds cell (i + offset) mod N with offset = -2*steps (sleeve 2)."""
    _, _, cols, (first, count), _, row0 = EXACT_CASES[name]
    n = len(next(iter(prof.values())))
    k = (np.arange(n) + (-2 * steps) % n + shift) % n
    out = {}
    for f, c in cols:
        v = np.empty(n)
        v[k] = prof[f]
        out[f] = float(np.mean(np.abs(v[first:first + count] - exact_rows[first - row0:first - row0 + count, c])))
    return out


def exact_rows_of(name):
    path = os.path.join(B.REFERENCE_ROOT, EXACT_CASES[name][1])
    ex = np.loadtxt(path)
    return ex[::10] if name == "hllc3d_shock" else ex


def main_exact():
    for name, (steps, rel, cols, (first, count), bound, row0) in EXACT_CASES.items():
        mm, desc = front_end(name)
        st = run(mm, desc, mm_interp.init(mm.init, tuple(desc["grid"])), steps)
        if len(desc["grid"]) == 3:
            for f, v in st.items():
                assert (v == v[:1, :, :1]).all(), (name, f, "not uniform across x and z")
        ex = exact_rows_of(name)
        keep = [0] + [c for _, c in cols]
        np.savetxt(os.path.join(HERE, "exact_%s.txt" % name), ex[:, keep], fmt="%.9e",
                   header="%s of formura v2.3.2, columns x %s%s" % (
                       rel, " ".join(f for f, _ in cols), " (every 10th line)" if name == "hllc3d_shock" else ""))
        prof = profile(name, st)
        ref = np.loadtxt(os.path.join(HERE, "exact_%s.txt" % name))
        ref_full = np.zeros((ref.shape[0], max(c for _, c in cols) + 1))
        for j, (_, c) in enumerate(cols):
            ref_full[:, c] = ref[:, j + 1]
        err = exact_l1(name, prof, steps, ref_full)
        shifted = {sh: exact_l1(name, prof, steps, ref_full, sh) for sh in (-2, -1, 1, 2)}
        for f in err:
            assert err[f] < bound and all(err[f] < e[f] for e in shifted.values()), (name, f, err, shifted)
        meta = {"program": name, "source": B.SAMPLES[name][0], "seed": None, "steps": steps, "grid": desc["grid"],
                "offset": [(-desc["sleeve"] * steps) % n for n in desc["grid"]], "sha256": field_hashes(st),
                "exact": rel, "exact_columns": cols, "exact_cells": [first, count], "l1_bound": bound,
                "checked": "mean |f - f_exact| %s over cells %d..%d; frame shifts -2..2 give %s" % (
                    ", ".join("%s %.5f" % kv for kv in err.items()), first, first + count - 1,
                    "; ".join("%+d: %s" % (sh, ", ".join("%.5f" % v for v in e.values())) for sh, e in shifted.items()))}
        with open(os.path.join(HERE, "sample_%s.json" % name), "w") as fh:
            json.dump(meta, fh, indent=1)
        print("wrote", name, meta["checked"], flush=True)


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
    main_exact()
