#!/usr/bin/env python3
"""Benchmark of the generated Formura_Forward on one MI355X.

Workload (one "step" = one Formura_Forward = one time step of the whole grid):
2-D Gray-Scott reaction-diffusion, 2048 x 2048 periodic grid, FP64, fields u and
v, dx = 0.01, dt = 1, Du = 2e-5, Dv = 1e-5, F = 0.04, k = 0.06; initial state
from the program's init graph (u = 1, v = 0, central square of 205 cells with
u = 0.5, v = 0.25) plus seeded 1 % noise on the square; 10000 steps by default.

The state is resident in HBM for the timed region (Formura_HIP_Set_Resident).
The GPU's clocks ramp up over about a second of load from idle and fall back
within a fraction of a second of idling, so every timed region follows
``--settle`` seconds (default 1) of untimed steps of the same kernel, after
which the state is reset to the seeded one: the timed steps are the job's own
steps from its initial state (then the --warmup steps), at settled clocks.
``value`` is GLUPS = cells x steps / s / 1e9.  ``roofline`` is measured on the
dominant kernel: algorithmic bytes per launch (32 B per cell update: read u, v,
write u', v' in FP64) over its mean launch duration from HIP events recorded on
the library's stream around the timed launches.  ``cpu_baseline`` times the
oracle (the transcription of the reference's generated C, compiled with the
reference Makefile flags and OpenMP) on the same workload for a bounded number
of steps.
"""
from __future__ import annotations

import argparse
import json
import os
import re
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0

# seconds of untimed work of the kernel about to be timed (set from --settle)
SETTLE_S = 1.0


def settle(step, seconds=None):
    """Call ``step()`` (untimed steps of the kernel about to be timed) until
    ``seconds`` (default SETTLE_S) of wall time have passed; returns the calls."""
    t0, n = time.perf_counter(), 0
    while time.perf_counter() - t0 < (SETTLE_S if seconds is None else seconds):
        step()
        n += 1
    return n


def _issue_from_profile(kernel: str):
    """VALU issue fraction of a kernel from the committed PMC summary
    (profiles/valu_fp64_kernels.md: SQ_INSTS_VALU over the wave-instruction issue
    peak during the kernel), or None."""
    p = os.path.join(ROOT, "profiles", "valu_fp64_kernels.md")
    if not os.path.exists(p):
        return None
    for line in open(p):
        cells = [c.strip() for c in line.split("|")]
        if len(cells) > 7 and cells[1] == "`%s`" % kernel:
            m = re.match(r"([0-9.]+)\s*%", cells[7])
            if m:
                return float(m.group(1)) / 100.0
    return None


def _traffic_from_profile(kernel_tag: str):
    """HBM bytes per launch from the committed rocprofv3 PMC summary, if present."""
    p = os.path.join(ROOT, "profiles", "hbm_traffic.txt")
    if not os.path.exists(p):
        return None
    for line in open(p):
        m = re.match(r"\s*traffic_bytes_per_launch\[%s\]\s*=\s*([0-9.eE+]+)" % re.escape(kernel_tag), line)
        if m:
            return float(m.group(1))
    return None


def run_gpu(name: str, steps: int, warmup: int, bytes_per_lup: float, seed: int = 7):
    import numpy as np
    from formura_amd import inputs
    from formura_amd.runtime import open_program
    lib = open_program(name)
    lib.init()
    if "u" in lib.fields:
        st = inputs.grayscott_perturb(lib.field("u").copy(), lib.field("v").copy(), seed)
    else:
        st = inputs.heat_field(seed, lib.grid)
    lib.set_state(st)
    lib.set_resident(True)
    lib.upload()
    spf = lib.layout["steps_per_forward"]
    t_settle = time.perf_counter()
    settle(lambda: lib.forward_steps(max(1, 48 // spf)))
    t_settle = time.perf_counter() - t_settle
    lib.set_state(st)  # the job's steps from its initial state
    lib.upload()
    if warmup:
        lib.forward_steps(max(1, warmup // spf))
    try:
        import torch
        if torch.cuda.is_available():
            torch.cuda.synchronize()
    except Exception:
        pass
    t0 = time.perf_counter()
    ms = lib.forward_steps(max(1, steps // spf))
    wall = time.perf_counter() - t0
    nsteps = max(1, steps // spf) * spf
    cells = int(np.prod(lib.grid))
    launches = lib.last_launches()
    lib.download()
    finite = all(bool(np.isfinite(lib.field(f)).all()) for f in lib.fields)
    lib.finalize()
    glups = cells * nsteps / (ms * 1e-3) / 1e9
    return {
        "glups": glups, "ms": ms, "wall_s": wall, "steps": nsteps, "cells": cells,
        "launches": launches, "mean_launch_us": ms * 1e3 / launches,
        "bytes_per_launch": bytes_per_lup * cells * nsteps / launches,
        # a launch must read the state once and write it once, however many
        # steps it fuses: the minimum HBM bytes of one launch
        "min_bytes_per_launch": bytes_per_lup * cells,
        "steps_per_launch": nsteps / launches,
        "schedule": lib.layout.get("schedule"), "finite": finite, "settle_s": round(t_settle, 3),
    }


# the FP32 forms of the batched ensemble: f32k (Kahan-compensated accumulation),
# f32c (and the increment in float-float FP32 from the FP32 state), f32x (the
# increment in float-float FP32 from the compensated state)
FP32_FORMS = ("f32k", "f32c", "f32x")


def run_ensemble(count: int, steps: int):
    """The batched FP32 Gray-Scott ensemble in its two forms (f32k: Kahan-compensated
    accumulation; f32c: and the increment in float-float FP32) and its FP64 twin on
    the GPU.

    Returns GLUPS of each (all instances, whole ensemble, state resident) and the
    relative L2 error of every instance of each FP32 run against its own FP64 run
    (maximum, median, instances within 1e-5)."""
    import numpy as np
    from formura_amd import batch as B
    sizes, params, seeds = B.grayscott_ensemble(count)
    inits = [B.grayscott_init(s[0], sd) for s, sd in zip(sizes, seeds)]
    cells = sum(a * b for a, b in sizes)
    out = {"instances": count, "cells": cells, "steps": steps,
           "sides": [min(s[0] for s in sizes), max(s[0] for s in sizes)]}
    ens = {}
    for name in ("gs_batch_f32k", "gs_batch_f32c", "gs_batch_f32x", "gs_batch_f64"):
        e = B.Ensemble(name, sizes, params)
        for i, st in enumerate(inits):
            e.set_state(i, st)
        e.step(e.desc["steps_per_launch"] * e.desc["graph_launches"])  # warm-up, captures the graph
        settle(lambda: e.step(e.desc["steps_per_launch"] * e.desc["graph_launches"]))
        for i, st in enumerate(inits):
            e.set_state(i, st)
        ms = e.step(steps)
        out[name] = {"glups": cells * steps / (ms * 1e-3) / 1e9, "ms": ms,
                     "tile": e.desc["tile"], "steps_per_launch": e.desc["steps_per_launch"]}
        ens[name] = e
    all_errs = {}
    for prec in FP32_FORMS:
        errs, rms = [], {"u": 0.0, "v": 0.0}
        for i in range(count):
            x, y = ens["gs_batch_f64"].get_state(i), ens["gs_batch_" + prec].get_state(i)
            num = sum(float(np.sum((y[f] - x[f]) ** 2)) for f in ("u", "v"))
            den = sum(float(np.sum(x[f] ** 2)) for f in ("u", "v"))
            errs.append((num / den) ** 0.5)
            for f in ("u", "v"):
                rms[f] = max(rms[f], float(np.sqrt(np.mean((y[f] - x[f]) ** 2))))
        errs = np.array(errs)
        all_errs[prec] = errs
        out[prec + "_rel_l2_vs_f64"] = {"max": float(errs.max()), "median": float(np.median(errs)),
                                        "instances_le_1e-5": int((errs <= 1e-5).sum())}
        out[prec + "_max_rms_error_per_field"] = rms
    errs = all_errs["f32k"]
    # Control: the FP64 ensemble again, from initial states perturbed at the scale
    # of FP32 rounding (u scaled by 1 + 2^-24 (2r - 1), r ~ U[0,1) seeded).  An
    # instance whose FP64 run moves by more than 1e-5 under that perturbation
    # amplifies rounding-size differences; FP32 arithmetic cannot follow it.
    from formura_amd import inputs
    ref = {i: ens["gs_batch_f64"].get_state(i) for i in range(count)}
    e = ens["gs_batch_f64"]
    for i, st in enumerate(inits):
        r = inputs.uniform(977 + i, st["u"].shape)
        e.set_state(i, {"u": st["u"] * (1.0 + 2.0 ** -24 * (2.0 * r - 1.0)), "v": st["v"]})
    e.step(steps)
    sens = []
    for i in range(count):
        x, y = ref[i], e.get_state(i)
        num = sum(float(np.sum((y[f] - x[f]) ** 2)) for f in ("u", "v"))
        den = sum(float(np.sum(x[f] ** 2)) for f in ("u", "v"))
        sens.append((num / den) ** 0.5)
    sens = np.array(sens)
    out["f64_perturbed_2^-24_rel_l2"] = {
        "median": float(np.median(sens)), "instances_le_1e-5": int((sens <= 1e-5).sum())}
    for prec in FP32_FORMS:
        e = all_errs[prec]
        out["f64_perturbed_2^-24_rel_l2"].update({
            prec + "_above_1e-5_where_f64_control_le_1e-5": int(((e > 1e-5) & (sens <= 1e-5)).sum()),
            prec + "_le_1e-5_where_f64_control_above_1e-5": int(((e <= 1e-5) & (sens > 1e-5)).sum())})
    out["per_instance"] = {"side": [int(a) for a, _ in sizes], "control": [float(x) for x in sens],
                           **{prec: [float(x) for x in all_errs[prec]] for prec in FP32_FORMS}}
    for e in ens.values():
        e.close()
    return out


def run_fp32_single(steps: int = 10000, seed: int = 7):
    """The headline workload (Gray-Scott 2048^2, F = 0.04, k = 0.06, the same
    seeded state) for 10^4 steps in the two FP32 forms -- f32k (Kahan-compensated
    accumulation) and f32c (and the increment in float-float FP32) -- and in FP64,
    each as a one-instance batch (the FP64 batch is bit-identical to the plain
    path at the program's own parameters): GLUPS and the error of each FP32 run
    against FP64 per field (relative L2, maximum absolute, relative maximum), and
    the same for FP64 from the state with u perturbed at FP32 rounding scale
    (how far this run amplifies a rounding-size change of its initial state)."""
    import numpy as np
    from formura_amd import batch as B
    st = B.grayscott_init(2048, seed)
    params = np.array([[0.04, 0.06]])
    out = {"grid": [2048, 2048], "steps": steps}
    res = {}
    for name in ("gs_batch_f64",) + tuple("gs_batch_" + p for p in FP32_FORMS):
        e = B.Ensemble(name, [(2048, 2048)], params)
        e.set_state(0, st)
        e.step(e.desc["steps_per_launch"] * e.desc["graph_launches"])  # warm-up, captures the graph
        settle(lambda: e.step(e.desc["steps_per_launch"] * e.desc["graph_launches"]))
        e.set_state(0, st)
        ms = e.step(steps)
        res[name] = e.get_state(0)
        out[name] = {"glups": 2048 * 2048 * steps / (ms * 1e-3) / 1e9, "ms": ms}
        e.close()
    # control: FP64 again from the state with u perturbed at FP32 rounding scale
    from formura_amd import inputs
    e = B.Ensemble("gs_batch_f64", [(2048, 2048)], params)
    r = inputs.uniform(977, st["u"].shape)
    e.set_state(0, {"u": st["u"] * (1.0 + 2.0 ** -24 * (2.0 * r - 1.0)), "v": st["v"]})
    e.step(steps)
    res["control"] = e.get_state(0)
    e.close()
    x = res["gs_batch_f64"]
    for prec in FP32_FORMS + ("control",):
        y = res["gs_batch_" + prec] if prec != "control" else res["control"]
        out[prec + "_error_vs_fp64" if prec != "control" else "f64_perturbed_2^-24_vs_fp64"] = {
            f: {"rel_l2": float(np.sqrt(np.sum((y[f] - x[f]) ** 2) / np.sum(x[f] ** 2))),
                "max_abs": float(np.max(np.abs(y[f] - x[f]))),
                "rel_max": float(np.max(np.abs(y[f] - x[f])) / np.max(np.abs(x[f])))} for f in ("u", "v")}
    return out


def run_fp32_diffusion(steps: int = 10000, seed: int = 7):
    """2-D diffusion (diffusion2d.fmr: 5-point FTCS, dt = dx^2/10) on a
    2048^2 periodic grid for 10^4 steps, as a one-instance batch in FP64 and in
    the two FP32 forms (f32k: FP32 state and arithmetic, Kahan-compensated
    accumulation; f32c: and the increment in float-float FP32): GLUPS, and the
    error of each FP32 run against FP64 (relative L2, maximum absolute,
    relative maximum).  Initial state: a Gaussian bump (sigma = 256 cells) plus
    1 % seeded uniform noise."""
    import numpy as np
    from formura_amd import batch as B, inputs
    n = 2048
    i = np.arange(n)[:, None] - n / 2
    j = np.arange(n)[None, :] - n / 2
    q0 = np.exp(-(i * i + j * j) / (2 * (n / 8) ** 2)) + 0.01 * inputs.uniform(seed, (n, n))
    out = {"grid": [n, n], "steps": steps, "program": "diffusion2d.fmr (dt = dx^2/10)"}
    res = {}
    for name in ("diff_batch_f64", "diff_batch_f32k", "diff_batch_f32c"):
        e = B.Ensemble(name, [(n, n)], np.zeros((1, 1)))
        e.set_state(0, {"q": q0})
        e.step(e.desc["steps_per_launch"] * e.desc["graph_launches"])  # warm-up, captures the graph
        settle(lambda: e.step(e.desc["steps_per_launch"] * e.desc["graph_launches"]))
        e.set_state(0, {"q": q0})
        ms = e.step(steps)
        res[name] = e.get_state(0)["q"]
        out[name] = {"glups": n * n * steps / (ms * 1e-3) / 1e9, "ms": ms}
        e.close()
    x = res["diff_batch_f64"]
    for prec in ("f32k", "f32c"):
        y = res["diff_batch_" + prec]
        out[prec + "_error_vs_fp64"] = {"q": {
            "rel_l2": float(np.sqrt(np.sum((y - x) ** 2) / np.sum(x ** 2))),
            "max_abs": float(np.max(np.abs(y - x))),
            "rel_max": float(np.max(np.abs(y - x)) / np.max(np.abs(x)))}}
    return out


def run_per_instance_fp64(sizes, steps: int, probe_steps: int = 600):
    """The plain per-instance FP64 GPU path for the same ensemble: every instance
    advanced by its own plain Formura_Forward library (state resident, one
    library per grid side, ``build.ENSEMBLE_SIDES``).  Each distinct side is
    timed on one instance over ``probe_steps`` steps; the ensemble time is the
    sum over instances of (time per step of its side) x ``steps``.  The time per
    step of a plain Gray-Scott run does not depend on F and k."""
    from collections import Counter
    from formura_amd import build as B, inputs
    from formura_amd.runtime import open_program
    counts = Counter(int(s[0]) for s in sizes)
    per, total_ms = {}, 0.0
    for n, c in sorted(counts.items()):
        lib = open_program(B.side_name(n))
        lib.init()
        st = inputs.grayscott_perturb(lib.field("u").copy(), lib.field("v").copy(), 7)
        lib.set_state(st)
        lib.set_resident(True)
        lib.upload()
        spf = lib.layout["steps_per_forward"]
        # the sides are timed back to back: the first settles the clocks, the rest keep them
        settle(lambda: lib.forward_steps(max(1, 48 // spf)), None if not per else 0.2)
        lib.set_state(st)
        lib.upload()
        lib.forward_steps(max(1, 30 // spf))
        k = max(1, probe_steps // spf)
        ms = lib.forward_steps(k)
        lib.finalize()
        per[n] = ms / (k * spf)
        total_ms += c * per[n] * steps
    cells = sum(int(a) * int(b) for a, b in sizes)
    return {"glups": cells * steps / (total_ms * 1e-3) / 1e9, "ms": total_ms, "steps": steps,
            "sides_timed": len(per), "probe_steps": probe_steps,
            "us_per_step": {str(n): round(v * 1e3, 3) for n, v in sorted(per.items())}}


def _roof(r, bytes_per_lup):
    """GLUPS, the HBM fraction of the minimum bytes per launch (the state read
    and written once), and the effective rate (bytes per update x updates)."""
    min_gbs = r["min_bytes_per_launch"] / (r["mean_launch_us"] * 1e-6) / 1e9
    return {"glups": r["glups"], "schedule": r["schedule"], "bytes_per_lup": bytes_per_lup,
            "steps_per_launch": r["steps_per_launch"], "mean_launch_us": r["mean_launch_us"], "steps": r["steps"],
            "min_traffic_gbs": min_gbs, "min_traffic_frac": min_gbs / HBM_PEAK_GBS,
            "effective_gbs": r["glups"] * bytes_per_lup, "effective_frac": r["glups"] * bytes_per_lup / HBM_PEAK_GBS}


def run_heat_roofline(name: str = "heat3d_512", steps: int = 200):
    """Plain FP64 diffusion at full size: GLUPS, the HBM fraction of the minimum
    traffic per launch, and the effective HBM fraction (16 B per cell update:
    read q, write q')."""
    return _roof(run_gpu(name, steps, 20, 16.0), 16.0)


def run_sizes():
    """The other plain-path configurations: Gray-Scott 8192^2 (1 GiB of state,
    beyond the 256 MiB Infinity Cache) and the 3-D heat 64^3 of the CPU-runnable
    configuration, without and with temporal blocking (4 steps per Forward)."""
    return {"gs_8192": _roof(run_gpu("gs_8192", 400, 20, 32.0), 32.0),
            "heat3d_64": _roof(run_gpu("heat3d_64", 100, 20, 16.0), 16.0),
            "heat3d_64_tb": _roof(run_gpu("heat3d_64_tb", 100, 20, 16.0), 16.0)}


def run_contract():
    """The opt-in contract form (FMA contraction, reciprocal division by
    constants) against the plain FP64 path from the same seeded state: GLUPS and
    the relative max-norm error per field at the configurations' steps."""
    import numpy as np
    from formura_amd import inputs
    from formura_amd.runtime import open_program
    out = {}
    for plain, fast, steps in (("gs_2048", "gs_2048_contract", 10000), ("euler3d_256", "euler3d_256_contract", 200)):
        res = {}
        for name in (plain, fast):
            lib = open_program(name)
            lib.init()
            if "u" in lib.fields:
                lib.set_state(inputs.grayscott_perturb(lib.field("u").copy(), lib.field("v").copy(), 7))
            else:
                lib.set_state(inputs.euler_state(1, lib.grid))
            lib.set_resident(True)
            lib.upload()
            st = lib.state()
            settle(lambda: lib.forward_steps(8))
            lib.set_state(st)
            lib.upload()
            ms = lib.forward_steps(steps)
            lib.download()
            res[name] = ({f: lib.field(f).copy() for f in lib.fields}, int(np.prod(lib.grid)) * steps / (ms * 1e-3) / 1e9)
            lib.set_resident(False)
            lib.finalize()
        a, b = res[plain][0], res[fast][0]
        out[fast] = {"steps": steps, "glups": res[fast][1], "plain_glups": res[plain][1],
                     "speedup": res[fast][1] / res[plain][1],
                     "rel_maxnorm_err_vs_fp64": {f: float(np.max(np.abs(a[f] - b[f])) / np.max(np.abs(a[f])))
                                                 for f in a}}
    return out


def run_fp32_heat(steps: int = 200, seed: int = 23):
    """3-D heat on 512^3 from one seeded state (i.i.d. U[0,1)): the plain FP64
    library (heat3d_512) against heat3d_kahan.fmr, the same 7-point step on FP32
    state with Kahan-compensated accumulation written in the language (fields
    q and c, FP32 arithmetic), both on the stream3d schedule, state resident:
    GLUPS and the FP32 error against FP64 (relative L2, maximum absolute), for
    q and for the compensated value q - c."""
    import numpy as np
    from formura_amd import inputs
    from formura_amd.runtime import open_program
    out = {"steps": steps, "grid": [512, 512, 512]}
    res = {}
    for name in ("heat3d_512", "heat3d_kahan_512"):
        lib = open_program(name)
        lib.init()
        q0 = inputs.uniform(seed, lib.grid)
        st = {"q": q0} if name == "heat3d_512" else {"q": q0.astype(np.float32),
                                                      "c": np.zeros(lib.grid, np.float32)}
        lib.set_state(st)
        lib.set_resident(True)
        lib.upload()
        settle(lambda: lib.forward_steps(8))
        lib.set_state(st)
        lib.upload()
        ms = lib.forward_steps(steps)
        lib.download()
        res[name] = {f: lib.field(f).copy() for f in lib.fields}
        # 16 B per cell update in both: one double, or the two floats q and c, read and written
        out[name] = {"glups": float(np.prod(lib.grid)) * steps / (ms * 1e-3) / 1e9, "ms": ms, "bytes_per_lup": 16}
        lib.finalize()
    x = res["heat3d_512"]["q"]
    k = res["heat3d_kahan_512"]
    for key, y in (("q", k["q"].astype(np.float64)), ("q_minus_c", k["q"].astype(np.float64) - k["c"])):
        out["error_vs_fp64_" + key] = {"rel_l2": float(np.sqrt(np.sum((y - x) ** 2) / np.sum(x ** 2))),
                                       "max_abs": float(np.max(np.abs(y - x)))}
    return out


def run_euler(steps: int = 200):
    """3-D Euler HLLC at 256^3: the plain FP64 path and two mixed forms (FP32 flux
    arithmetic; FP64 wave speeds, positivity checks, update and totals) from the
    same seeded state -- fluxes recomputed in place (``euler3d_256_mixed``) and
    fluxes evaluated once per face into LDS in FP32 (``euler3d_256_mixed_staged``);
    error of each mixed run against FP64 per field."""
    import numpy as np
    from formura_amd import inputs
    from formura_amd.runtime import open_program
    st, out = None, {"steps": steps}
    res = {}
    for name in ("euler3d_256", "euler3d_256_mixed", "euler3d_256_mixed_staged"):
        lib = open_program(name)
        lib.init()
        st = st or inputs.euler_state(1, lib.grid)
        lib.set_state(st)
        lib.set_resident(True)
        lib.upload()
        settle(lambda: lib.forward_steps(4))
        lib.set_state(st)
        lib.upload()
        t0 = lib.totals()
        ms = lib.forward_steps(steps)
        t1 = lib.totals()
        lib.download()
        res[name] = {f: lib.field(f).copy() for f in lib.fields}
        cells = int(np.prod(lib.grid))
        scale = max(abs(v) for v in t0.values())
        out[name] = {"glups": cells * steps / (ms * 1e-3) / 1e9, "ms": ms, "schedule": lib.layout.get("schedule"),
                     "max_rel_total_drift": max(abs(t1[f] - t0[f]) for f in t0) / scale,
                     "positivity_violations": lib.positivity_violations()}
        lib.set_resident(False)
        lib.finalize()
    a, b = res["euler3d_256"], res["euler3d_256_mixed"]
    out["mixed_rel_maxnorm_err_vs_fp64"] = {f: float(np.max(np.abs(a[f] - b[f])) / np.max(np.abs(a[f]))) for f in a}
    out["mixed_speedup"] = out["euler3d_256_mixed"]["glups"] / out["euler3d_256"]["glups"]
    c = res["euler3d_256_mixed_staged"]
    out["mixed_staged_rel_maxnorm_err_vs_fp64"] = {f: float(np.max(np.abs(a[f] - c[f])) / np.max(np.abs(a[f])))
                                                   for f in a}
    out["mixed_staged_speedup"] = out["euler3d_256_mixed_staged"]["glups"] / out["euler3d_256"]["glups"]
    return out


def run_adjoint(nsteps: int = 2000, snapshots: int = 32, directions: int = 16, eps: float = 1e-6):
    """Gray-Scott 1024^2 adjoint: loss 0.5*||v_T - v_target||^2 with v_target from a
    seeded forward run at (F, k) = (0.035, 0.065); gradients w.r.t. the initial
    u, v and F, k, Du, Dv by the generated adjoint with binomial checkpointing;
    checked by FP64 central differences along seeded random directions.  Also
    the FP32 form (gradient error against FP64) and the 3-D heat adjoint at 256^3."""
    import numpy as np
    import torch
    from formura_amd import adjoint as A, batch as B, inputs
    dev = "cuda"

    def ev():
        e = torch.cuda.Event(enable_timing=True)
        e.record()
        return e

    out = {"config": {"grid": [1024, 1024], "steps": nsteps, "snapshots": snapshots, "directions": directions,
                      "eps": eps}}
    st = B.grayscott_init(1024, 11)
    x0 = np.stack([st["u"], st["v"]])
    theta = [0.04, 0.06, 2e-5, 1e-5]
    grads = {}
    for name, dt in (("gs_1024_adj", torch.float64), ("gs_1024_adj_f32", torch.float32)):
        s = A.Stepper(name)
        xd = torch.tensor(x0, device=dev, dtype=dt)
        with torch.no_grad():
            vt = A.run(s, xd, torch.tensor([0.035, 0.065, 2e-5, 1e-5], dtype=torch.float64), nsteps, 0)[1].clone()
        fwd, bwd = [], []

        def fwd_only():
            with torch.no_grad():
                A.run(s, xd, torch.tensor(theta, dtype=torch.float64), 100, 0)
            torch.cuda.synchronize()
        settle(fwd_only)
        for rep in range(4):  # one warm-up, then the best of three timed repetitions
            x = xd.clone().requires_grad_(True)
            p = torch.tensor(theta, dtype=torch.float64, requires_grad=True)
            torch.cuda.synchronize()
            e0 = ev()
            y = A.run(s, x, p, nsteps, snapshots)
            e1 = ev()
            loss = 0.5 * ((y[1] - vt) ** 2).sum()
            e2 = ev()
            loss.backward()
            e3 = ev()
            torch.cuda.synchronize()
            if rep > 0:
                fwd.append(e0.elapsed_time(e1))
                bwd.append(e2.elapsed_time(e3))
            del y, loss
        rec = {"forward_ms": min(fwd), "backward_ms": min(bwd), "ratio": min(bwd) / min(fwd),
               "forward_ms_all": fwd, "backward_ms_all": bwd, "stats": dict(s.last_stats), "unit_steps": s.levels}
        grads[name] = (x.grad.double().cpu().numpy(), p.grad.cpu().numpy(), s, vt)
        out[name] = rec
    gx, gp, s, vt = grads["gs_1024_adj"]
    rng = np.random.default_rng(2024)

    def L(x, th):
        with torch.no_grad():
            y = A.run(s, torch.tensor(x, device=dev), torch.tensor(th, dtype=torch.float64), nsteps, 0)
            return 0.5 * float(((y[1] - vt) ** 2).sum())

    errs = []
    for _ in range(directions):
        dx = rng.standard_normal(x0.shape)
        dp = rng.standard_normal(4) * np.array(theta)
        fd = (L(x0 + eps * dx, np.array(theta) + eps * dp) - L(x0 - eps * dx, np.array(theta) - eps * dp)) / (2 * eps)
        an = float((gx * dx).sum() + (gp * dp).sum())
        errs.append(abs(fd - an) / max(abs(an), 1e-300))
    out["gs_1024_adj"]["fd_rel_err"] = {"max": max(errs), "median": float(np.median(errs))}
    g32x, g32p = grads["gs_1024_adj_f32"][:2]
    out["gs_1024_adj_f32"]["grad_rel_l2_vs_fp64"] = {
        "state": float(np.linalg.norm(g32x - gx) / np.linalg.norm(gx)),
        "params": [float(abs(a - b) / abs(b)) for a, b in zip(g32p, gp)]}
    # 3-D heat: linear step without parameters, no checkpoint needed
    s = A.Stepper("heat3d_256_adj")
    q0 = inputs.uniform(5, (256, 256, 256))[None]
    tgt = torch.tensor(inputs.uniform(6, (256, 256, 256)), device=dev)
    hsteps = 100
    x = torch.tensor(q0, device=dev, requires_grad=True)

    def heat_fwd():
        with torch.no_grad():
            A.run(s, x.detach(), torch.zeros(0, dtype=torch.float64), 10, 0)
        torch.cuda.synchronize()
    settle(heat_fwd)
    torch.cuda.synchronize()
    e0 = ev()
    y = A.run(s, x, torch.zeros(0, dtype=torch.float64), hsteps, snapshots)
    e1 = ev()
    loss = 0.5 * ((y[0] - tgt) ** 2).sum()
    e2 = ev()
    loss.backward()
    e3 = ev()
    torch.cuda.synchronize()
    g = x.grad.cpu().numpy()
    herr = []
    # the heat map is linear and the loss quadratic, so central differences have
    # no truncation error at any step; 1e-2 lifts them above the roundoff of a
    # loss of ~7e5 whose gradient is ~4e-3 per cell after 100 smoothing steps
    heps = 1e-2
    for _ in range(4):
        d = rng.standard_normal(q0.shape)
        with torch.no_grad():
            lp = 0.5 * float(((A.run(s, torch.tensor(q0 + heps * d, device=dev), torch.zeros(0, dtype=torch.float64), hsteps, 0)[0] - tgt) ** 2).sum())
            lm = 0.5 * float(((A.run(s, torch.tensor(q0 - heps * d, device=dev), torch.zeros(0, dtype=torch.float64), hsteps, 0)[0] - tgt) ** 2).sum())
        fd = (lp - lm) / (2 * heps)
        an = float((g * d).sum())
        herr.append(abs(fd - an) / abs(an))
    out["heat3d_256_adj"] = {"steps": hsteps, "forward_ms": e0.elapsed_time(e1), "backward_ms": e2.elapsed_time(e3),
                             "ratio": e2.elapsed_time(e3) / e0.elapsed_time(e1), "stats": dict(s.last_stats),
                             "fd_eps": heps, "fd_rel_err_max": max(herr)}
    return out


def host_cores():
    """(physical cores of the host per lscpu, CPUs this process may run on, the
    OpenMP threads a parallel region of the oracle runs with)."""
    import ctypes
    import subprocess
    phys = None
    try:
        out = subprocess.run(["lscpu", "-p=CORE,SOCKET"], capture_output=True, text=True, timeout=30).stdout
        phys = len({ln for ln in out.splitlines() if ln and not ln.startswith("#")})
    except Exception:
        pass
    try:
        avail = len(os.sched_getaffinity(0))
    except Exception:
        avail = os.cpu_count() or 1
    try:  # the OpenMP runtime the oracle libraries link (already initialised by this process's imports)
        threads = int(ctypes.CDLL("libgomp.so.1").omp_get_max_threads())
    except Exception:
        threads = int(os.environ.get("OMP_NUM_THREADS", "0")) or avail
    return phys, avail, threads


# CPU baseline configurations: name -> (oracle library, seed function name, steps per timed sample)
CPU_CONFIGS = {
    "gs_2048": ("gs", 200),
    "heat3d_64": ("heat", 100),
    "heat3d_64_tb": ("heat", 25),       # 25 Forwards = 100 steps (4 per Forward)
    "heat3d_256": ("heat", 4),
    "euler3d_256": ("euler", 1),
}


def cpu_time(name: str, samples: int = 5, min_s: float = 1.0, threads: int = 1):
    """The oracle's fast build (the reference Makefile flags, OpenMP) on the host
    cores: ``samples`` timed samples of whole Formura_Forward calls (each at least
    ``min_s`` seconds) after one warm-up call; returns the median GLUPS."""
    import statistics
    import numpy as np
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import build_oracle
    from formura_amd import inputs
    from formura_amd.runtime import FormuraLib
    from formura_amd.emit_hip import generate
    from formura_amd.emit import lib_description
    from formura_amd import build as B
    fmr, yml, _ = B.PROGRAMS[name]
    _, _, mm = generate(os.path.join(B.PROG_DIR, fmr), os.path.join(B.PROG_DIR, yml), name)
    path = build_oracle.lib_path(name, "fast")
    if not os.path.exists(path):
        build_oracle.build(name, "fast")
    lib = FormuraLib(path, lib_description(mm.om, name, ""))
    lib.init()
    kind, chunk = CPU_CONFIGS[name]
    if kind == "gs":
        lib.set_state(inputs.grayscott_perturb(lib.field("u").copy(), lib.field("v").copy(), 7))
    elif kind == "euler":
        lib.set_state(inputs.euler_state(1, lib.grid))
    else:
        lib.set_state(inputs.heat_field(7, lib.grid))
    spf = lib.layout["steps_per_forward"]
    lib.forward(1)  # warm-up
    cells = int(np.prod(lib.grid))
    rates, total = [], 0.0
    for _ in range(samples):
        t0 = time.perf_counter()
        n = 0
        while True:
            lib.forward(chunk)
            n += chunk
            el = time.perf_counter() - t0
            if el >= min_s:
                break
        rates.append(cells * n * spf / el / 1e9)
        total += el
    return {"glups": statistics.median(rates), "samples": [round(r, 4) for r in rates], "seconds": round(total, 2),
            "steps_per_forward": spf, "blocking": "temporal" if spf > 1 else "none", "threads": threads}


def cpu_baseline(name: str, threads: int, samples: int = 5, min_s: float = 2.0):
    """The headline workload on the host cores (see ``cpu_time``)."""
    phys, avail, _ = host_cores()
    r = cpu_time(name, samples, min_s, threads)
    return {"value": r["glups"], "unit": "GLUPS", "cores": threads, "kind": "port",
            "sample": "median of %d samples of >= %.0f s of whole Formura_Forward calls (%.1f s in all) of the %s "
                      "workload; oracle C (transcription of the reference's generated C) built with -O3 "
                      "-ffast-math -mavx2 -mfma -fopenmp, %d OpenMP threads (OMP_PROC_BIND=%s); host: %s "
                      "physical cores (lscpu), %d CPUs available to this process"
                      % (samples, min_s, r["seconds"], name, threads, os.environ.get("OMP_PROC_BIND", "unset"),
                         phys, avail),
            "samples": r["samples"]}


def cpu_baselines(threads: int):
    """CPU rates of the other configurations (C1 64^3 without and with temporal
    blocking, heat 256^3, C3 Euler 256^3), each the median of 5 samples."""
    out = {}
    for name in ("heat3d_64", "heat3d_64_tb", "heat3d_256", "euler3d_256"):
        try:
            out[name] = cpu_time(name, 5, 1.0, threads)
        except Exception as e:
            out[name] = {"error": str(e)[:200]}
    return out


def run_samples(steps: int = 50):
    """The reference's sample programs at 128^3 (3-D HLLD MHD with GLM cleaning,
    3-D HLLC with MUSCL-Hancock), compiled from their .fmr by this front-end
    with the default schedule (cheap temporaries recomputed where read), with
    every temporary materialized (``_flat``), and in the mixed form: GLUPS with
    the state resident in HBM, and the FP64 totals' relative drift (every field
    but the damped psi)."""
    import numpy as np
    from formura_amd import build as B, inputs
    from formura_amd.runtime import open_program
    out = {"steps": steps}
    finals = {}
    for name, gamma in (("mhd3d_128", 5.0 / 3.0), ("mhd3d_128_mixed", 5.0 / 3.0), ("mhd3d_128_flat", 5.0 / 3.0),
                        ("hllc3d_128", 7.0 / 5.0), ("hllc3d_128_mixed", 7.0 / 5.0), ("hllc3d_128_flat", 7.0 / 5.0)):
        if not os.path.exists(B.lib_path(name)):
            out[name] = {"skipped": "built only where the formura checkout is present"}
            continue
        lib = open_program(name)
        lib.init()
        st = inputs.conserved_state(5, tuple(lib.grid), lib.fields, gamma)
        lib.set_state(st)
        lib.set_resident(True)
        lib.upload()
        settle(lambda: lib.forward_steps(2))
        lib.set_state(st)
        lib.upload()
        t0 = lib.totals()
        ms = lib.forward_steps(steps)
        t1 = lib.totals()
        cells = int(np.prod(lib.grid))
        cons = [f for f in t0 if f != "psi"]
        out[name] = {"glups": cells * steps / (ms * 1e-3) / 1e9, "ms": ms, "schedule": lib.layout.get("schedule"),
                     "fields": len(lib.fields),
                     "max_rel_total_drift": max(abs(t1[f] - t0[f]) for f in cons) / max(abs(t0[f]) for f in cons),
                     "positivity_violations": lib.positivity_violations()}
        lib.download()
        finals[name] = {f: lib.field(f).copy() for f in lib.fields}
        lib.set_resident(False)
        lib.finalize()
    for base in ("mhd3d_128", "hllc3d_128"):
        a, c = finals.get(base), finals.get(base + "_flat")
        if a and c:
            out[base]["identical_to_materialized_flat"] = all(bool(np.array_equal(a[f], c[f])) for f in a)
            out[base]["speedup_vs_materialized_flat"] = out[base]["glups"] / out[base + "_flat"]["glups"]
        a, b = finals.get(base), finals.get(base + "_mixed")
        if a and b:
            out[base + "_mixed"]["rel_maxnorm_err_vs_fp64"] = {
                f: float(np.max(np.abs(a[f] - b[f])) / np.max(np.abs(a[f]))) for f in a}
            out[base + "_mixed"]["speedup_vs_fp64"] = out[base + "_mixed"]["glups"] / out[base]["glups"]
    return out


def run_adjoint_probe(nsteps: int = 200, snapshots: int = 8):
    """A short FP64 Gray-Scott 1024^2 forward + backward (for profiling the adjoint
    kernels without the finite-difference checks of the adjoint leg)."""
    import numpy as np
    import torch
    from formura_amd import adjoint as A, batch as B
    st = B.grayscott_init(1024, 11)
    s = A.Stepper("gs_1024_adj")
    x = torch.tensor(np.stack([st["u"], st["v"]]), device="cuda").requires_grad_(True)
    p = torch.tensor([0.04, 0.06, 2e-5, 1e-5], dtype=torch.float64, requires_grad=True)
    y = A.run(s, x, p, nsteps, snapshots)
    (0.5 * (y[1] ** 2).sum()).backward()
    torch.cuda.synchronize()
    return {"steps": nsteps, "snapshots": snapshots, "dL_dF": float(p.grad[0]), "stats": dict(s.last_stats)}


# the legs after the headline, in order: (key in the JSON line, leg name, function)
LEGS = (("grayscott2048_fp32", "fp32", run_fp32_single),
        ("diffusion2d_2048_fp32", "fp32", run_fp32_diffusion),
        ("heat3d_512_fp32", "fp32", run_fp32_heat),
        ("heat3d_512", "heat", run_heat_roofline), ("sizes", "sizes", run_sizes),
        ("euler3d_256", "euler", run_euler), ("contract", "contract", run_contract),
        ("adjoint", "adjoint", run_adjoint), ("samples", "samples", run_samples))


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10000)
    ap.add_argument("--warmup", type=int, default=200)
    ap.add_argument("--settle", type=float, default=1.0,
                    help="seconds of untimed steps before each timed region (GPU clock ramp); the state is "
                         "then reset to the seeded one")
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline leg")
    ap.add_argument("--cpu-seconds", type=float, default=10.0,
                    help="CPU time of the headline's CPU baseline (5 samples of a fifth of it)")
    ap.add_argument("--ensemble", type=int, default=1024,
                    help="instances of the batched FP32 Gray-Scott ensemble (0: skip)")
    ap.add_argument("--ensemble-steps", type=int, default=5000)
    ap.add_argument("--no-extras", action="store_true",
                    help="skip the heat 512^3, Euler 256^3 and adjoint legs")
    ap.add_argument("--legs", default="plain,fp32,ensemble,heat,sizes,euler,contract,adjoint,samples,cpu",
                    help="comma-separated legs to run (for profiling one kernel family): plain (Gray-Scott "
                         "2048^2, the headline value), fp32 (the same in the FP32 forms, error vs FP64), "
                         "ensemble, heat, euler, adjoint, samples, cpu; "
                         "adjoint_probe (a short adjoint run for profiling, not in the default)")
    ap.add_argument("--calibrate-copy", action="store_true",
                    help="only run a plain device-to-device copy of 256 MiB, 20 times (the reference "
                         "kernel for scaling rocprofv3 FETCH_SIZE/WRITE_SIZE counters)")
    args = ap.parse_args()
    global SETTLE_S
    SETTLE_S = args.settle
    if args.calibrate_copy:
        import torch
        a = torch.empty(256 << 20, dtype=torch.uint8, device="cuda")
        b = torch.empty_like(a)
        a.fill_(1)
        for _ in range(20):
            b.copy_(a)
        torch.cuda.synchronize()
        print(json.dumps({"calibrate_copy_bytes": a.numel(), "launches": 20}))
        return
    # torch first: its HIP runtime is then the one every library of this process
    # binds to (the adjoint leg runs through torch)
    try:
        import torch
        torch.cuda.is_available()
    except ImportError:
        pass
    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) > 1:
        print("bench.py: this build runs on one GPU only (the periodic halo is wrapped in HBM; "
              "no inter-GPU exchange is built). Use --gpus 1.", file=sys.stderr)
        sys.exit(2)
    legs = set(args.legs.split(","))
    if args.no_extras:
        legs -= {"heat", "sizes", "euler", "contract", "adjoint", "samples"}
    if args.no_cpu:
        legs.discard("cpu")
    if args.ensemble <= 0:
        legs.discard("ensemble")
    if "plain" not in legs:
        line = {"legs": sorted(legs)}
        for key, leg, fn in LEGS + (("adjoint_probe", "adjoint_probe", run_adjoint_probe),):
            if leg in legs:
                line[key] = fn()
        if "ensemble" in legs:
            line["ensemble"] = run_ensemble(args.ensemble, args.ensemble_steps)
        if "cpu" in legs:
            line["cpu_baselines"] = cpu_baselines(host_cores()[2])
        print(json.dumps(line), flush=True)
        return
    name = "gs_2048"
    r = run_gpu(name, args.steps, args.warmup, 32.0)
    # roofline of the dominant kernel: the minimum HBM bytes of one launch (the
    # state read once and written once, 32 B per cell, whatever the number of
    # fused steps) over its mean duration
    achieved = r["min_bytes_per_launch"] / (r["mean_launch_us"] * 1e-6) / 1e9
    traffic = _traffic_from_profile(name)
    kernel = "k_step_tile<%d>" % round(r["steps_per_launch"])
    issue = _issue_from_profile(kernel)
    line = {
        "metric": "GLUPS & % HBM roofline, FP32/batched vs plain FP64; adjoint/forward ratio",
        "value": round(r["glups"], 3),
        "unit": "GLUPS",
        "n_gpus": 1,
        "steps": r["steps"],
        "warmup": args.warmup,
        "settle_s": r["settle_s"],
        "ms_per_step": r["ms"] / r["steps"],
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic",
        "config": {"workload": "grayscott2d 2048x2048 periodic, FP64 u,v, dx=0.01 dt=1 Du=2e-5 Dv=1e-5 "
                               "F=0.04 k=0.06, seeded init, plain (bit-exact) path",
                   "grid": [2048, 2048], "fields": 2, "schedule": r["schedule"],
                   "bytes_per_lup": 32, "resident": True},
        "roofline": {"bound": "hbm", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": round(achieved / HBM_PEAK_GBS, 4), "traffic": traffic,
                     "kernel": kernel,
                     "bytes_model": "state read once and written once per launch: 32 B x cells",
                     "effective_gbs": round(r["glups"] * 32.0, 1),
                     "valu_issue_frac": issue},
    }
    if not r["finite"]:
        line["error"] = "non-finite state after the run"
    if "ensemble" in legs:
        try:
            ens = run_ensemble(args.ensemble, args.ensemble_steps)
            from formura_amd import batch as B
            sizes, _, _ = B.grayscott_ensemble(args.ensemble)
            pi = run_per_instance_fp64(sizes, args.ensemble_steps)
            ens["per_instance_fp64"] = pi
            for prec in FP32_FORMS:
                ens[prec + "_batched_vs_per_instance_fp64"] = ens["gs_batch_" + prec]["glups"] / pi["glups"]
                ens[prec + "_batched_vs_plain_fp64_2048"] = ens["gs_batch_" + prec]["glups"] / r["glups"]
            line["ensemble"] = ens
        except Exception as e:
            line["ensemble"] = {"error": str(e)[:300]}
    for key, leg, fn in LEGS:
        if leg in legs:
            try:
                line[key] = fn()
            except Exception as e:
                line[key] = {"error": str(e)[:300]}
    if "cpu" in legs and int(os.environ.get("RANK", "0")) == 0:
        threads = host_cores()[2]
        try:
            line["cpu_baseline"] = cpu_baseline(name, threads, min_s=args.cpu_seconds / 5)
            line["cpu_baseline"]["gpu_vs_cpu"] = r["glups"] / line["cpu_baseline"]["value"]
        except Exception as e:  # the GPU number stands without it
            line["cpu_baseline"] = {"error": str(e)[:300]}
        line["cpu_baselines"] = cpu_baselines(threads)
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
