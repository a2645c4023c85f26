"""Autotuner for the kernel schedule of a program (tile shape, steps per launch).

Two phases, so that compilation runs on a host without a GPU and only the
timing needs one:

``python -m formura_amd.autotune build NAME...``
    generates and compiles one library per candidate schedule
    (``emit_hip.candidate_specs``) under ``_build/NAME/variants/``;
``python -m formura_amd.autotune run NAME... --out DIR``
    on the GPU: runs every variant on the same seeded state for the same
    number of time steps (state resident in HBM), checks that every variant's
    final state is bit-identical to the first one's, and writes
    ``DIR/NAME.json`` with the GLUPS of every variant and the fastest spec.

Copying that JSON to ``formura_amd/tuning/NAME.json`` makes the build use the
winner (``emit_hip.choose_schedule``).

Batched ensembles (``build.BATCHES``) are tuned the same way over tile shape,
steps per launch and persistent grid size (``build_batch_variants`` /
``run_batch_variants``), on a seeded ensemble of the same size distribution;
every variant's final states must be bit-identical to the first variant's.
"""
from __future__ import annotations

import argparse
import hashlib
import json
import os
import sys
import time
import concurrent.futures as cf

from . import build as B


def _tag(spec: dict) -> str:
    return "_".join("%s%s" % (k, spec[k]) for k in sorted(spec)) if spec else "default"


def variant_dir(name: str) -> str:
    return os.path.join(B.BUILD_DIR, name, "variants")


def build_variants(name: str, jobs: int = 8):
    from .emit_hip import LibEmitter, candidate_specs, generate
    from .emit import emit_header
    fmr, yml, _ = B.PROGRAMS[name]
    _, _, mm = generate(os.path.join(B.PROG_DIR, fmr), os.path.join(B.PROG_DIR, yml), name)
    base = LibEmitter(mm, name, "flat")
    specs = candidate_specs(base)
    todo = []
    for spec in specs:
        d = os.path.join(variant_dir(name), _tag(spec))
        os.makedirs(d, exist_ok=True)
        src = LibEmitter(mm, name, spec).emit()
        for fn, text in (("%s.h" % name, emit_header(mm.om)), ("%s.hip" % name, src)):
            with open(os.path.join(d, fn), "w") as f:
                f.write(text)
        with open(os.path.join(d, "spec.json"), "w") as f:
            json.dump(spec, f)
        todo.append(d)
    with cf.ThreadPoolExecutor(max_workers=jobs) as ex:
        list(ex.map(lambda d: B.compile_library(name, d), todo))
    return todo


_STATE_CACHE = {}


def _seed_state(lib, seed):
    """Seeded input of the program (Gray-Scott, Euler, or i.i.d. uniform fields), cached per shape."""
    from . import inputs
    key = (tuple(lib.fields), tuple(lib.grid), seed)
    if key not in _STATE_CACHE:
        if "u" in lib.fields and "v" in lib.fields:
            st = inputs.grayscott_perturb(lib.field("u").copy(), lib.field("v").copy(), seed)
        elif set(lib.fields) == {"rho", "rhou", "rhov", "rhow", "e"}:
            st = inputs.euler_state(1, lib.grid)
        else:
            st = {f: inputs.uniform(seed + i, lib.grid) for i, f in enumerate(lib.fields)}
        _STATE_CACHE[key] = st
    return _STATE_CACHE[key]


def run_variants(name: str, out_dir: str, seconds: float = 0.3, seed: int = 5, settle: float = 0.8,
                 job_steps: int = 0):
    """Time every built variant of ``name`` on the same seeded state.

    The GPU's clocks ramp up from idle over about a second of load and fall
    back within a fraction of a second of idling, so each variant first runs
    ``settle`` seconds of untimed steps.  With ``job_steps``, the state is then
    reset to the seeded one and exactly ``job_steps`` steps are timed (a whole
    job from its initial state: for Gray-Scott, the early steps whose
    diffusion front carries dividends outside the fused division's range are
    included); otherwise the timed run continues from the settled state and
    lasts about ``seconds``.  Every variant's final state must be bit-identical
    to the first variant's."""
    import numpy as np
    from .runtime import FormuraLib
    root = variant_dir(name)
    dirs = sorted(d for d in os.listdir(root) if os.path.exists(os.path.join(root, d, "lib%s.so" % name)))
    results, ref_state, ref_tag = [], None, None
    steps = job_steps or None
    settle_steps = 0
    for d in dirs:
        spec = json.load(open(os.path.join(root, d, "spec.json")))
        lib = FormuraLib(os.path.join(root, d, "lib%s.so" % name))
        lib.init()
        lib.set_state(_seed_state(lib, seed))
        lib.set_resident(True)
        lib.upload()
        cells = int(np.prod(lib.grid))
        if job_steps:
            t0 = time.time()
            while time.time() - t0 < settle:
                lib.forward_steps(48)
            lib.set_state(_seed_state(lib, seed))
            lib.upload()
        else:
            if steps is None:  # size the run from the first variant: ~seconds of work
                ms = lib.forward_steps(4)
                per = max(ms / 4, 1e-3)
                steps = int(max(24, min(10 ** 6, seconds * 1e3 / per))) // 24 * 24
                # the same number of settling steps for every variant: final states stay comparable
                settle_steps = int(max(24, min(10 ** 6, settle * 1e3 / per))) // 24 * 24
                lib.set_state(_seed_state(lib, seed))
                lib.upload()
            lib.forward_steps(settle_steps)
        ms = lib.forward_steps(steps)
        lib.download()
        st = lib.state()
        if ref_state is None:
            ref_state, ref_tag = st, d
            same = True
        else:
            same = all(np.array_equal(st[f], ref_state[f]) for f in st)
        glups = cells * steps / (ms * 1e-3) / 1e9
        results.append({"spec": spec, "tag": d, "glups": glups, "ms": ms, "steps": steps, "settle_s": settle,
                        "from_initial_state": bool(job_steps), "launches": lib.last_launches(),
                        "identical_to_first": bool(same)})
        print("%-40s %8.2f GLUPS  %s" % (d, glups, "ok" if same else "MISMATCH"), flush=True)
        lib.set_resident(False)
        lib.finalize()
    good = [r for r in results if r["identical_to_first"]]
    best = max(good, key=lambda r: r["glups"]) if good else None
    out = {"name": name, "best": best["spec"] if best else None, "results": results,
           "reference_variant": ref_tag, "steps_per_variant": steps}
    os.makedirs(out_dir, exist_ok=True)
    with open(os.path.join(out_dir, "%s.json" % name), "w") as f:
        json.dump(out, f, indent=1)
    return out


def batch_candidates(precision: str):
    """Tile (tx, ty), steps per launch k and persistent grid size for a batch
    library; ty and tx divide every side of the ensemble (multiples of 32) and
    the LDS of a workgroup stays within 64 KiB."""
    arrays, size = (4, 4) if precision in ("f32k", "f32c", "f32x") else (2, 8)
    out = []
    for tx, ty in ((16, 32), (32, 32), (8, 32), (16, 16), (32, 16)):
        for k in (2, 3, 4, 5, 6, 8):
            lds = (tx + 2 * k) * (ty + 2 * k) * arrays * 2 * size
            if lds > 64 * 1024:
                continue
            for grid in (1024, 2048, 4096):
                out.append({"tx": tx, "ty": ty, "k": k, "grid": grid})
            if tx * ty >= 512 and k in (3, 4, 5):
                for grid in (1024, 2048):
                    out.append({"tx": tx, "ty": ty, "k": k, "grid": grid, "threads": 512})
    return out


def build_batch_variants(name: str, jobs: int = 8):
    prec = B.BATCHES[name][2]
    todo = []
    for spec in batch_candidates(prec):
        d = os.path.join(variant_dir(name), _tag(spec))
        os.makedirs(d, exist_ok=True)
        with open(os.path.join(d, "spec.json"), "w") as f:
            json.dump(spec, f)
        todo.append((spec, d))
    with cf.ThreadPoolExecutor(max_workers=jobs) as ex:
        list(ex.map(lambda a: B.build_batch(name, spec=a[0], out_dir=a[1]), todo))
    return [d for _, d in todo]


def run_batch_variants(name: str, out_dir: str, count: int = 256, steps: int = 480, seed: int = 77):
    import numpy as np
    from . import batch as BT
    sizes, params, seeds = BT.grayscott_ensemble(count, seed=seed)
    init = [BT.grayscott_init(sz[0], sd) for sz, sd in zip(sizes, seeds)]
    root = variant_dir(name)
    dirs = sorted(d for d in os.listdir(root) if os.path.exists(os.path.join(root, d, "lib%s.so" % name)))
    probe = [0, count // 3, count - 1]
    results, ref, ref_tag = [], None, None
    for d in dirs:
        spec = json.load(open(os.path.join(root, d, "spec.json")))
        e = BT.Ensemble(name, sizes, params, path=os.path.join(root, d, "lib%s.so" % name))
        for i, st in enumerate(init):
            e.set_state(i, st)
        e.step(48)  # warm-up: code objects, graph capture
        for i, st in enumerate(init):  # every variant then runs the same steps from the same state
            e.set_state(i, st)
        ms = e.step(steps)
        got = [e.get_state(i) for i in probe]
        if ref is None:
            ref, ref_tag, same = got, d, True
        else:
            same = all(np.array_equal(a[f], b[f]) for a, b in zip(got, ref) for f in a)
        glups = e.cells() * steps / (ms * 1e-3) / 1e9
        e.close()
        results.append({"spec": spec, "tag": d, "glups": glups, "ms": ms, "steps": steps, "identical_to_first": bool(same)})
        print("%-28s %8.2f GLUPS  %s" % (d, glups, "ok" if same else "MISMATCH"), flush=True)
    good = [r for r in results if r["identical_to_first"]]
    best = max(good, key=lambda r: r["glups"]) if good else None
    out = {"name": name, "best": best["spec"] if best else None, "results": results,
           "reference_variant": ref_tag, "ensemble": {"count": count, "seed": seed, "steps": steps}}
    os.makedirs(out_dir, exist_ok=True)
    with open(os.path.join(out_dir, "%s.json" % name), "w") as f:
        json.dump(out, f, indent=1)
    return out


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("phase", choices=["build", "run"])
    ap.add_argument("names", nargs="+")
    ap.add_argument("--out", default="tuning_out")
    ap.add_argument("--seconds", type=float, default=0.3)
    ap.add_argument("--settle", type=float, default=0.8,
                    help="seconds of untimed steps before each variant's timed run (clock ramp)")
    ap.add_argument("--job-steps", type=int, default=0,
                    help="time exactly this many steps from the seeded state (after settling)")
    a = ap.parse_args()
    for n in a.names:
        batch = n in B.BATCHES
        if a.phase == "build":
            print(n, len((build_batch_variants if batch else build_variants)(n)), "variants", flush=True)
        elif batch:
            run_batch_variants(n, a.out)
        else:
            run_variants(n, a.out, a.seconds, settle=a.settle, job_steps=a.job_steps)


if __name__ == "__main__":
    main()
