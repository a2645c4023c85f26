"""Build the HIP libraries of the bundled programs (and of user programs).

``python -m formura_amd.build`` generates ``<name>.h`` / ``<name>.hip`` for every
entry of ``PROGRAMS`` into ``formura_amd/_build/<name>/`` and compiles
``lib<name>.so`` for gfx950 with hipcc.  Built libraries are loaded in place
(they are not installed anywhere else).
"""
from __future__ import annotations

import concurrent.futures as cf
import hashlib
import os
import subprocess
import sys
from typing import Dict, List, Optional

PKG = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(PKG)
PROG_DIR = os.path.join(PKG, "programs")
BUILD_DIR = os.path.join(PKG, "_build")
CSRC = os.path.join(PKG, "csrc")
INCLUDE = os.path.join(ROOT, "include")

HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
ARCH = os.environ.get("FORMURA_HIP_ARCH", "gfx950")

# -ffp-contract=off: no FMA contraction, so every + - * / is rounded as in the
# reference's C.  The packed-fp32 feature is switched off for all HIP code.
HIP_FLAGS = ["--offload-arch=" + ARCH, "-O3", "-std=c++17", "-fPIC", "-shared",
             "-ffp-contract=off", "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
             "-Wno-unused-variable", "-Wno-unused-but-set-variable", "-Wno-unused-parameter"]

# name -> (program, config, schedule)
PROGRAMS: Dict[str, tuple] = {
    "heat3d_16": ("heat3d.fmr", "heat3d_16.yaml", "auto"),
    "heat3d_16_tb": ("heat3d.fmr", "heat3d_16_tb.yaml", "auto"),
    "heat3d_32": ("heat3d.fmr", "heat3d_32.yaml", "auto"),
    "heat3d_64": ("heat3d.fmr", "heat3d_64.yaml", "auto"),
    "heat3d_64_tb": ("heat3d.fmr", "heat3d_64_tb.yaml", "auto"),
    "heat3d_256": ("heat3d.fmr", "heat3d_256.yaml", "auto"),
    "heat3d_512": ("heat3d.fmr", "heat3d_512.yaml", "auto"),
    "gs_64": ("grayscott2d.fmr", "gs_64.yaml", "auto"),
    "hooks2d_32": ("hooks2d.fmr", "hooks2d_32.yaml", "auto"),
    "float2d_32": ("float2d.fmr", "float2d_32.yaml", "flat"),
    "float2d_32_tile": ("float2d.fmr", "float2d_32.yaml", {"kind": "tile2d", "tx": 16, "ty": 16, "k": 3}),
    # FP32 state divided by a literal (double division, rounded to float) and by a float constant
    "floatdiv2d_32": ("floatdiv2d.fmr", "float2d_32.yaml", "flat"),
    "floatdiv2d_32_tile": ("floatdiv2d.fmr", "float2d_32.yaml", {"kind": "tile2d", "tx": 16, "ty": 16, "k": 3}),
    "hooks2d_32_tb": ("hooks2d.fmr", "hooks2d_32_tb.yaml", "auto"),
    "gs_1024": ("grayscott2d.fmr", "gs_1024.yaml", "auto"),
    "gs_2048": ("grayscott2d.fmr", "gs_2048.yaml", "auto"),
    "gs_8192": ("grayscott2d.fmr", "gs_8192.yaml", "auto"),
    "euler3d_16": ("euler3d.fmr", "euler3d_16.yaml", "auto"),
    "euler3d_32": ("euler3d.fmr", "euler3d_32.yaml", "auto"),
    "euler3d_256": ("euler3d.fmr", "euler3d_256.yaml", "auto"),
    # fixed schedules, so that every schedule is parity-tested at small sizes
    "heat3d_32_stream": ("heat3d.fmr", "heat3d_32.yaml", {"kind": "stream3d", "ty": 8, "tz": 16, "cx": 8, "k": 2}),
    "heat3d_64_tb_stream": ("heat3d.fmr", "heat3d_64_tb.yaml", {"kind": "stream3d", "ty": 4, "tz": 16, "cx": 16, "k": 3}),
    "heat3d_32_stream512": ("heat3d.fmr", "heat3d_32.yaml", {"kind": "stream3d", "ty": 16, "tz": 32, "cx": 8, "k": 2, "threads": 512}),
    # skewed levels: one barrier per plane (rings of 2s+2 planes)
    "heat3d_32_skew": ("heat3d.fmr", "heat3d_32.yaml", {"kind": "stream3d", "ty": 8, "tz": 16, "cx": 8, "k": 2, "skew": 1}),
    "heat3d_64_tb_skew": ("heat3d.fmr", "heat3d_64_tb.yaml", {"kind": "stream3d", "ty": 4, "tz": 16, "cx": 16, "k": 3,
                                                             "skew": 1}),
    "heat3d_24x40x56_skew": ("heat3d.fmr", "heat3d_24x40x56.yaml", {"kind": "stream3d", "ty": 8, "tz": 8, "cx": 8,
                                                                   "k": 2, "skew": 1}),
    "heat3d_kahan_16_skew": ("heat3d_kahan.fmr", "heat3d_kahan_16.yaml", {"kind": "stream3d", "ty": 8, "tz": 8, "cx": 8,
                                                                         "k": 2, "skew": 1}),
    "gs_64_flat": ("grayscott2d.fmr", "gs_64.yaml", "flat"),
    # grids no power of two divides: 72 x 120 and 24 x 40 x 56
    "gs_72x120": ("grayscott2d.fmr", "gs_72x120.yaml", "auto"),
    "gs_72x120_tile": ("grayscott2d.fmr", "gs_72x120.yaml", {"kind": "tile2d", "tx": 24, "ty": 40, "k": 3}),
    "gs_72x120_wave": ("grayscott2d.fmr", "gs_72x120.yaml", {"kind": "wave2d", "rows": 8, "k": 2}),
    "heat3d_24x40x56": ("heat3d.fmr", "heat3d_24x40x56.yaml", "auto"),
    "heat3d_24x40x56_stream": ("heat3d.fmr", "heat3d_24x40x56.yaml", {"kind": "stream3d", "ty": 8, "tz": 8, "cx": 8, "k": 2}),
    "heat3d_24x40x56_tb": ("heat3d.fmr", "heat3d_24x40x56_tb.yaml", "auto"),
    "diffusion2d_64": ("diffusion2d.fmr", "diffusion2d_64.yaml", "auto"),
    # FP32 state with Kahan-compensated accumulation written in the language (fields q, c)
    "heat3d_kahan_16": ("heat3d_kahan.fmr", "heat3d_kahan_16.yaml", "flat"),
    "heat3d_kahan_16_stream": ("heat3d_kahan.fmr", "heat3d_kahan_16.yaml", {"kind": "stream3d", "ty": 8, "tz": 8, "cx": 8, "k": 2}),
    "heat3d_kahan_512": ("heat3d_kahan.fmr", "heat3d_kahan_512.yaml", {"kind": "stream3d", "ty": 32, "tz": 64, "cx": 256, "k": 2,
                                                                        "skew": 1, "threads": 1024}),
    "diffusion2d_64_wave": ("diffusion2d.fmr", "diffusion2d_64.yaml", {"kind": "wave2d", "rows": 8, "k": 3}),
    "heat3d_24x40x56_tb_stream": ("heat3d.fmr", "heat3d_24x40x56_tb.yaml", {"kind": "stream3d", "ty": 8, "tz": 8, "cx": 8, "k": 2}),
    "euler3d_16_inline": ("euler3d.fmr", "euler3d_16.yaml", {"kind": "flat", "inline": 1}),
    "euler3d_32_inline_stream": ("euler3d.fmr", "euler3d_32.yaml",
                                 {"kind": "stream3d", "ty": 8, "tz": 16, "cx": 8, "k": 1, "inline": 1}),
    "euler3d_16_staged": ("euler3d.fmr", "euler3d_16.yaml",
                          {"kind": "staged3d", "ty": 8, "tz": 16, "cx": 4, "inline": "inner"}),
    "euler3d_32_staged": ("euler3d.fmr", "euler3d_32.yaml",
                          {"kind": "staged3d", "ty": 16, "tz": 16, "cx": 8, "threads": 512, "inline": "inner"}),
    "index2d_32_tb": ("index2d.fmr", "index2d_32_tb.yaml", {"kind": "tile2d", "tx": 16, "ty": 16, "k": 2}),
    "index2d_32_tb8": ("index2d.fmr", "index2d_32_tb.yaml", {"kind": "tile2d", "tx": 8, "ty": 32, "k": 2, "threads": 128}),
    # index reads inside a manifest temporary (recomputed where read, with its nest's index)
    "indextmp2d_32": ("indextmp2d.fmr", "indextmp2d_32.yaml", "auto"),
    "index3d_16_tb": ("index3d.fmr", "index3d_16_tb.yaml", {"kind": "stream3d", "ty": 8, "tz": 16, "cx": 8, "k": 2}),
    "gs_64_tile_k8": ("grayscott2d.fmr", "gs_64.yaml", {"kind": "tile2d", "tx": 16, "ty": 32, "k": 8}),
    "gs_64_tile_k5_inplace": ("grayscott2d.fmr", "gs_64.yaml", {"kind": "tile2d", "tx": 16, "ty": 32, "k": 5, "inplace": 1}),
    "gs_64_wave_k3": ("grayscott2d.fmr", "gs_64.yaml", {"kind": "wave2d", "rows": 16, "k": 3}),
    "gs_64_tile_t128": ("grayscott2d.fmr", "gs_64.yaml", {"kind": "tile2d", "tx": 16, "ty": 32, "k": 3, "threads": 128}),
    "gs_64_tile_92k": ("grayscott2d.fmr", "gs_64.yaml", {"kind": "tile2d", "tx": 64, "ty": 64, "k": 6, "inplace": 1, "threads": 1024}),
    "gs_2048_wave": ("grayscott2d.fmr", "gs_2048.yaml", {"kind": "wave2d", "rows": 32, "k": 3}),
    # how a division by a constant meets dividends outside the fused sequence's
    # range: the whole point again for such a wave (spec), or a per-lane branch
    # in each division (guard); the default redoes the division for such a wave
    "gs_64_tile_spec": ("grayscott2d.fmr", "gs_64.yaml", {"kind": "tile2d", "tx": 16, "ty": 32, "k": 4, "div": "spec"}),
    "gs_64_tile_guard": ("grayscott2d.fmr", "gs_64.yaml", {"kind": "tile2d", "tx": 16, "ty": 32, "k": 4, "div": "guard"}),
    # a level's work-items on a fixed column each; two rows per work-item
    "gs_64_tile_colmap": ("grayscott2d.fmr", "gs_64.yaml", {"kind": "tile2d", "tx": 16, "ty": 32, "k": 4, "colmap": 1}),
    # divisions without tests where a level's values passed the per-level range test
    "gs_64_tile_level": ("grayscott2d.fmr", "gs_64.yaml",
                         {"kind": "tile2d", "tx": 16, "ty": 32, "k": 4, "colmap": 1, "div": "level"}),
    "gs_64_tile_level_flat": ("grayscott2d.fmr", "gs_64.yaml", {"kind": "tile2d", "tx": 16, "ty": 16, "k": 3,
                                                               "threads": 128, "div": "level"}),
    "gs_2048_level": ("grayscott2d.fmr", "gs_2048.yaml", {"kind": "tile2d", "tx": 32, "ty": 32, "k": 4, "threads": 512,
                                                         "colmap": 1, "div": "level"}),
    "gs_64_tile_k5_inplace_colmap": ("grayscott2d.fmr", "gs_64.yaml",
                                     {"kind": "tile2d", "tx": 16, "ty": 32, "k": 5, "inplace": 1, "colmap": 1}),
    "gs_64_tile_vpair": ("grayscott2d.fmr", "gs_64.yaml", {"kind": "tile2d", "tx": 16, "ty": 32, "k": 4, "vpair": 1}),
    "gs_72x120_tile_colmap": ("grayscott2d.fmr", "gs_72x120.yaml",
                              {"kind": "tile2d", "tx": 24, "ty": 40, "k": 3, "colmap": 1}),
    # the opt-in contract form (FMA contraction, reciprocal division; not bit-exact)
    "gs_2048_contract": ("grayscott2d.fmr", "gs_2048.yaml", {"kind": "tile2d", "tx": 32, "ty": 64, "k": 5,
                                                              "threads": 1024, "inplace": 1, "colmap": 1,
                                                              "contract": 1}),
    "euler3d_256_contract": ("euler3d.fmr", "euler3d_256.yaml", {"kind": "staged3d", "ty": 4, "tz": 64, "cx": 32,
                                                                  "threads": 512, "inline": "inner", "contract": 1}),
    "gs_64_contract": ("grayscott2d.fmr", "gs_64.yaml", {"kind": "tile2d", "tx": 16, "ty": 32, "k": 4, "contract": 1}),
    # two adjacent points per work-item
    "gs_64_tile_pair": ("grayscott2d.fmr", "gs_64.yaml", {"kind": "tile2d", "tx": 16, "ty": 32, "k": 4,
                                                           "threads": 128, "pair": 1}),
    "gs_72x120_tile_pair": ("grayscott2d.fmr", "gs_72x120.yaml", {"kind": "tile2d", "tx": 24, "ty": 40, "k": 3,
                                                                   "pair": 1}),
    "gs_2048_tile_pair": ("grayscott2d.fmr", "gs_2048.yaml", {"kind": "tile2d", "tx": 32, "ty": 32, "k": 4,
                                                               "threads": 512, "pair": 1}),
    # mixed precision: FP32 flux arithmetic, FP64 wave speeds, positivity checks and update
    "euler3d_16_mixed": ("euler3d.fmr", "euler3d_16.yaml", {"kind": "flat", "inline": 1, "mixed": 1}),
    "euler3d_32_mixed": ("euler3d.fmr", "euler3d_32.yaml", {"kind": "flat", "inline": 1, "mixed": 1}),
    "euler3d_32_mixed_mat": ("euler3d.fmr", "euler3d_32.yaml", {"kind": "flat", "mixed": 1}),
    "euler3d_32_mixed_inner": ("euler3d.fmr", "euler3d_32.yaml", {"kind": "flat", "mixed": 1, "inline": "inner"}),
    "euler3d_32_mixed_staged": ("euler3d.fmr", "euler3d_32.yaml",
                                {"kind": "staged3d", "ty": 8, "tz": 16, "cx": 8, "mixed": 1, "inline": "inner"}),
    "euler3d_32_mixed_stream": ("euler3d.fmr", "euler3d_32.yaml",
                                {"kind": "stream3d", "ty": 8, "tz": 16, "cx": 8, "k": 1, "inline": 1, "mixed": 1}),
    "euler3d_256_mixed_staged": ("euler3d.fmr", "euler3d_256.yaml",
                                 {"kind": "staged3d", "ty": 4, "tz": 64, "cx": 32, "threads": 512, "mixed": 1,
                                  "inline": "inner"}),
    "euler3d_256_mixed": ("euler3d.fmr", "euler3d_256.yaml",
                          {"kind": "stream3d", "ty": 8, "tz": 32, "cx": 32, "k": 1, "inline": 1, "mixed": 1}),
}


# The reference's own sample programs (the .fmr files under sample/ of the
# formura checkout), compiled by this front-end and emitter with one-rank
# configs of this package: name -> (program path relative to the checkout,
# config, schedule).  They are built only where the checkout is present
# (FORMURA_REFERENCE, default /root/reference); the built libraries are then
# used like any other.  Their .fmr text is not part of this package.
REFERENCE_ROOT = os.environ.get("FORMURA_REFERENCE", "/root/reference")
SAMPLES: Dict[str, tuple] = {
    "hllc1d_64": ("sample/1d_hllc/hllc.fmr", "hllc1d_64.yaml", "auto"),
    "hllc1d_400": ("sample/1d_hllc/hllc.fmr", "hllc1d_400.yaml", "auto"),
    # the reference's own configurations, on one rank (for their exact solutions)
    "mhd1d_3200": ("sample/mhd_1d_hlld/mhd_hlld.fmr", "mhd1d_3200.yaml", "auto"),
    "hllc3d_shock": ("sample/3d_hllc/3d_hllc.fmr", "hllc3d_shock.yaml", "auto"),
    "mhd1d_64": ("sample/mhd_1d_hlld/mhd_hlld.fmr", "mhd1d_64.yaml", "auto"),
    "mhd3d_8": ("sample/mhd_3d_hlld/mhd_3d_hlld.fmr", "mhd3d_8.yaml", "auto"),
    "hllc3d_8": ("sample/3d_hllc/3d_hllc.fmr", "hllc3d_8.yaml", "auto"),
    "mhd3d_128": ("sample/mhd_3d_hlld/mhd_3d_hlld.fmr", "mhd3d_128.yaml", "auto"),
    "hllc3d_128": ("sample/3d_hllc/3d_hllc.fmr", "hllc3d_128.yaml", "auto"),
    # mixed form: FP32 fluxes, FP64 wave speeds, reconstructions, positivity checks and update
    "mhd3d_8_mixed": ("sample/mhd_3d_hlld/mhd_3d_hlld.fmr", "mhd3d_8.yaml", {"kind": "flat", "mixed": 1, "inline": "cheap"}),
    "hllc3d_8_mixed": ("sample/3d_hllc/3d_hllc.fmr", "hllc3d_8.yaml", {"kind": "flat", "mixed": 1, "inline": "cheap"}),
    "mhd3d_128_mixed": ("sample/mhd_3d_hlld/mhd_3d_hlld.fmr", "mhd3d_128.yaml", {"kind": "flat", "mixed": 1, "inline": "cheap"}),
    "hllc3d_128_mixed": ("sample/3d_hllc/3d_hllc.fmr", "hllc3d_128.yaml", {"kind": "flat", "mixed": 1, "inline": "cheap"}),
    # every temporary materialized (the default recomputes the cheap ones where read)
    "hllc3d_8_flat": ("sample/3d_hllc/3d_hllc.fmr", "hllc3d_8.yaml", {"kind": "flat"}),
    "mhd3d_8_flat": ("sample/mhd_3d_hlld/mhd_3d_hlld.fmr", "mhd3d_8.yaml", {"kind": "flat"}),
    "hllc3d_128_flat": ("sample/3d_hllc/3d_hllc.fmr", "hllc3d_128.yaml", {"kind": "flat"}),
    "mhd3d_128_flat": ("sample/mhd_3d_hlld/mhd_3d_hlld.fmr", "mhd3d_128.yaml", {"kind": "flat"}),
}


def reference_present() -> bool:
    return os.path.isdir(os.path.join(REFERENCE_ROOT, "sample"))


# The reference's own driver of a sample, built unchanged against the generated
# header (under the name the driver includes) and the HIP library:
# name -> (driver path relative to the checkout, header name it includes).
SAMPLE_DRIVERS: Dict[str, tuple] = {
    "hllc1d_400": ("sample/1d_hllc/main.c", "hllc.h"),
}


def driver_path(name: str) -> str:
    return os.path.join(BUILD_DIR, name, "driver", "main")


def build_sample(name: str, force: bool = False) -> str:
    rel, yml, schedule = SAMPLES[name]
    d = generate_program(name, os.path.join(REFERENCE_ROOT, rel), os.path.join(PROG_DIR, yml), schedule)
    lib = compile_library(name, d, force=force)
    if name in SAMPLE_DRIVERS:
        src, hdr = SAMPLE_DRIVERS[name]
        dd = os.path.join(d, "driver")
        os.makedirs(dd, exist_ok=True)
        with open(os.path.join(d, "%s.h" % name)) as f:
            text = f.read()
        with open(os.path.join(dd, hdr), "w") as f:
            f.write(text)
        cmd = ["gcc", "-O2", "-I", dd, "-I", INCLUDE, os.path.join(REFERENCE_ROOT, src), "-L", d, "-l" + name,
               "-Wl,-rpath,$ORIGIN/..", "-lm", "-o", driver_path(name)]
        r = subprocess.run(cmd, capture_output=True, text=True)
        if r.returncode != 0:
            raise RuntimeError("driver build failed for %s:\n%s\n%s" % (name, " ".join(cmd), r.stderr[-4000:]))
    return lib


# batched ensembles: name -> (program, config, precision, lifted literals, tile x, tile y, steps per launch)
BATCHES: Dict[str, tuple] = {
    "gs_batch_f64": ("grayscott2d.fmr", "gs_64.yaml", "f64", ["f", "k"], 16, 32, 3),
    "gs_batch_f32k": ("grayscott2d.fmr", "gs_64.yaml", "f32k", ["f", "k"], 16, 32, 4),
    "gs_batch_f32c": ("grayscott2d.fmr", "gs_64.yaml", "f32c", ["f", "k"], 16, 32, 4),
    "gs_batch_f32x": ("grayscott2d.fmr", "gs_64.yaml", "f32x", ["f", "k"], 32, 32, 4),
    # 2-D diffusion (dx = 1/64 from the template config, dt = dx^2/10), no lifted parameters
    "diff_batch_f64": ("diffusion2d.fmr", "diffusion2d_64.yaml", "f64", [], 32, 32, 4),
    "diff_batch_f32k": ("diffusion2d.fmr", "diffusion2d_64.yaml", "f32k", [], 32, 32, 4),
    "diff_batch_f32c": ("diffusion2d.fmr", "diffusion2d_64.yaml", "f32c", [], 32, 32, 4),
}


def build_batch(name: str, force: bool = False, spec: Optional[dict] = None, out_dir: Optional[str] = None) -> str:
    from .emit_batch import BatchEmitter
    from .emit_hip import generate
    fmr, yml, prec, params, tx, ty, k = BATCHES[name]
    if spec is None:
        spec = _tuned_batch(name)
    grid, threads = 2048, 256
    if spec:
        tx, ty, k = int(spec["tx"]), int(spec["ty"]), int(spec["k"])
        grid = int(spec.get("grid", grid))
        threads = int(spec.get("threads", threads))
    _, _, mm = generate(os.path.join(PROG_DIR, fmr), os.path.join(PROG_DIR, yml), name, "flat")
    src = BatchEmitter(mm, name, prec, params, tx, ty, k, grid_wgs=grid, threads=threads).emit()
    d = out_dir or os.path.join(BUILD_DIR, name)
    os.makedirs(d, exist_ok=True)
    p = os.path.join(d, "%s.hip" % name)
    if not os.path.exists(p) or open(p).read() != src:
        with open(p, "w") as f:
            f.write(src)
    hdr = os.path.join(d, "%s.h" % name)
    if not os.path.exists(hdr):
        open(hdr, "w").write("/* batch library: see include/formura_batch.h */\n")
    return compile_library(name, d, force=force)


# adjoint libraries: name -> (program, config, lifted parameters, real type)
ADJOINTS: Dict[str, tuple] = {
    "gs_64_adj": ("grayscott2d.fmr", "gs_64.yaml", ["f", "k", "ru", "rv"], "double"),
    "gs_64_adj_f32": ("grayscott2d.fmr", "gs_64.yaml", ["f", "k", "ru", "rv"], "float"),
    "gs_1024_adj": ("grayscott2d.fmr", "gs_1024.yaml", ["f", "k", "ru", "rv"], "double"),
    "gs_1024_adj_f32": ("grayscott2d.fmr", "gs_1024.yaml", ["f", "k", "ru", "rv"], "float"),
    "heat3d_16_adj": ("heat3d.fmr", "heat3d_16.yaml", [], "double"),
    "heat3d_256_adj": ("heat3d.fmr", "heat3d_256.yaml", [], "double"),
    # the Euler HLLC step: the adjoint of the flux Jacobian (wave speeds, star
    # states and the flux selection differentiated through their branches)
    "euler3d_16_adj": ("euler3d.fmr", "euler3d_16.yaml", [], "double"),
}


def build_adjoint(name: str, force: bool = False) -> str:
    from .emit_adjoint import AdjointEmitter
    from .emit_hip import generate
    fmr, yml, params, real = ADJOINTS[name]
    _, _, mm = generate(os.path.join(PROG_DIR, fmr), os.path.join(PROG_DIR, yml), name, "flat")
    src = AdjointEmitter(mm, name, params, real).emit()
    d = os.path.join(BUILD_DIR, name)
    os.makedirs(d, exist_ok=True)
    p = os.path.join(d, "%s.hip" % name)
    if not os.path.exists(p) or open(p).read() != src:
        with open(p, "w") as f:
            f.write(src)
    hdr = os.path.join(d, "%s.h" % name)
    if not os.path.exists(hdr):
        open(hdr, "w").write("/* adjoint library: see include/formura_adjoint.h */\n")
    return compile_library(name, d, force=force)


def _tuned_batch(name: str):
    p = os.path.join(PKG, "tuning", "%s.json" % name)
    if os.path.exists(p):
        import json
        with open(p) as f:
            return json.load(f).get("best")
    return None


# Plain single-instance Gray-Scott programs for every side of the batched
# ensemble (64..1024 in steps of 32): the per-instance FP64 GPU path that the
# batched ensemble is compared with.  Schedule: the one tuned for 2048^2.
ENSEMBLE_SIDES = tuple(range(64, 1025, 32))
SIDE_SPEC = {"kind": "tile2d", "tx": 32, "ty": 32, "k": 4, "threads": 512}


def side_name(n: int) -> str:
    return "gs_side_%d" % n


def build_side(name: str, force: bool = False) -> str:
    n = int(name.rsplit("_", 1)[1])
    d = os.path.join(BUILD_DIR, name)
    os.makedirs(d, exist_ok=True)
    y = os.path.join(d, "%s.yaml" % name)
    text = "length_per_node: [%d.%02d,%d.%02d]\ngrid_per_node: [%d,%d]\n" % (n // 100, n % 100, n // 100, n % 100, n, n)
    if not os.path.exists(y) or open(y).read() != text:
        with open(y, "w") as f:
            f.write(text)
    generate_program(name, os.path.join(PROG_DIR, "grayscott2d.fmr"), y, dict(SIDE_SPEC), out_dir=d)
    return compile_library(name, d, force=force)


def lib_path(name: str) -> str:
    return os.path.join(BUILD_DIR, name, "lib%s.so" % name)


def header_path(name: str) -> str:
    return os.path.join(BUILD_DIR, name, "%s.h" % name)


def generate_program(name: str, fmr: str, yaml: str, schedule: str = "auto", out_dir: Optional[str] = None):
    from .emit_hip import generate
    h, src, mm = generate(fmr, yaml, name, schedule)
    d = out_dir or os.path.join(BUILD_DIR, name)
    os.makedirs(d, exist_ok=True)
    for fn, text in (("%s.h" % name, h), ("%s.hip" % name, src)):
        p = os.path.join(d, fn)
        if not os.path.exists(p) or open(p).read() != text:
            with open(p, "w") as f:
                f.write(text)
    return d


def source_stamp(name: str, src_dir: str) -> str:
    """Hash of everything a library is compiled from (its generated source and
    header, the runtime and public headers, the compiler flags)."""
    src = os.path.join(src_dir, "%s.hip" % name)
    deps = [src, os.path.join(src_dir, "%s.h" % name), os.path.join(CSRC, "formura_hip_rt.h"),
            os.path.join(INCLUDE, "formura_hip.h"), os.path.join(INCLUDE, "formura_batch.h"),
            os.path.join(INCLUDE, "formura_adjoint.h")]
    return hashlib.sha1(("\0".join(HIP_FLAGS) + "".join(open(p).read() for p in deps)).encode()).hexdigest()


def compile_library(name: str, src_dir: str, force: bool = False) -> str:
    src = os.path.join(src_dir, "%s.hip" % name)
    out = os.path.join(src_dir, "lib%s.so" % name)
    stamp = source_stamp(name, src_dir)
    stamp_file = out + ".stamp"
    if not force and os.path.exists(out) and os.path.exists(stamp_file) and open(stamp_file).read() == stamp:
        return out
    cmd = [HIPCC] + HIP_FLAGS + ["-I", CSRC, "-I", INCLUDE, "-I", src_dir, "-o", out + ".tmp", src]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError("hipcc failed for %s:\n%s\n%s" % (name, " ".join(cmd), r.stderr[-6000:]))
    os.replace(out + ".tmp", out)
    with open(stamp_file, "w") as f:
        f.write(stamp)
    return out


def build_one(name: str, force: bool = False) -> str:
    fmr, yaml, schedule = PROGRAMS[name]
    d = generate_program(name, os.path.join(PROG_DIR, fmr), os.path.join(PROG_DIR, yaml), schedule)
    return compile_library(name, d, force=force)


def regenerate(name: str) -> str:
    """Write the current generated source of library ``name`` (whatever its
    kind) into its build folder, compiling nothing; returns the folder."""
    if name in BATCHES or name in ADJOINTS:
        from .emit_hip import generate
        if name in BATCHES:
            from .emit_batch import BatchEmitter
            fmr, yml, prec, params, tx, ty, k = BATCHES[name]
            spec = _tuned_batch(name)
            grid, threads = 2048, 256
            if spec:
                tx, ty, k = int(spec["tx"]), int(spec["ty"]), int(spec["k"])
                grid = int(spec.get("grid", grid))
                threads = int(spec.get("threads", threads))
            _, _, mm = generate(os.path.join(PROG_DIR, fmr), os.path.join(PROG_DIR, yml), name, "flat")
            src = BatchEmitter(mm, name, prec, params, tx, ty, k, grid_wgs=grid, threads=threads).emit()
        else:
            from .emit_adjoint import AdjointEmitter
            fmr, yml, params, real = ADJOINTS[name]
            _, _, mm = generate(os.path.join(PROG_DIR, fmr), os.path.join(PROG_DIR, yml), name, "flat")
            src = AdjointEmitter(mm, name, params, real).emit()
        d = os.path.join(BUILD_DIR, name)
        p = os.path.join(d, "%s.hip" % name)
        if not os.path.exists(p) or open(p).read() != src:
            os.makedirs(d, exist_ok=True)
            with open(p, "w") as f:
                f.write(src)
        return d
    if name.startswith("gs_side_"):
        d = os.path.join(BUILD_DIR, name)
        return generate_program(name, os.path.join(PROG_DIR, "grayscott2d.fmr"), os.path.join(d, "%s.yaml" % name),
                                dict(SIDE_SPEC), out_dir=d)
    if name in SAMPLES:
        rel, yml, schedule = SAMPLES[name]
        return generate_program(name, os.path.join(REFERENCE_ROOT, rel), os.path.join(PROG_DIR, yml), schedule)
    fmr, yaml, schedule = PROGRAMS[name]
    return generate_program(name, os.path.join(PROG_DIR, fmr), os.path.join(PROG_DIR, yaml), schedule)


def check_current(name: str) -> str:
    """The built library of ``name``, after checking that it was compiled from
    the sources the current generator writes (the reference's sample programs,
    where the checkout is absent: from the generated source in the folder).
    Raises LibraryStale otherwise; compiles nothing."""
    if name in SAMPLES and not reference_present():
        d = os.path.join(BUILD_DIR, name)
    else:
        d = regenerate(name)
    out = lib_path(name)
    if not os.path.exists(out):
        raise LibraryStale("%s is not built; run __graft_entry__.build()" % out)
    stamp_file = out + ".stamp"
    if not os.path.exists(stamp_file) or open(stamp_file).read() != source_stamp(name, d):
        raise LibraryStale("%s was built from other sources than the current generator writes; "
                           "run __graft_entry__.build()" % out)
    return out


class LibraryStale(RuntimeError):
    pass


def builder_of(name: str):
    """The build function of library ``name`` (program, batch, adjoint, side, sample)."""
    if name in BATCHES:
        return build_batch
    if name in ADJOINTS:
        return build_adjoint
    if name.startswith("gs_side_"):
        return build_side
    if name in SAMPLES:
        return build_sample
    return build_one


def build_all(names: Optional[List[str]] = None, force: bool = False, jobs: int = 0) -> List[str]:
    names = names or (list(PROGRAMS) + list(BATCHES) + list(ADJOINTS) + [side_name(n) for n in ENSEMBLE_SIDES]
                      + (list(SAMPLES) if reference_present() else []))
    jobs = jobs or min(8, os.cpu_count() or 4)
    outs = []

    with cf.ThreadPoolExecutor(max_workers=jobs) as ex:
        futs = {ex.submit(builder_of(n), n, force): n for n in names}
        for f in cf.as_completed(futs):
            outs.append(f.result())
    return outs


if __name__ == "__main__":
    args = [a for a in sys.argv[1:] if not a.startswith("-")]
    force = "--force" in sys.argv
    for p in build_all(args or None, force=force):
        print(p)
