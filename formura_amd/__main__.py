"""Command line: ``python -m formura_amd PROG.fmr [--nc CONFIG.yaml] [-o OUT] [--compile]``.

Mirrors the reference's ``formura FILES [-o out.c] [--nc cfg.yaml]``
(src/Formura/CommandLineOption.hs:33-51): the config defaults to the program
name with ``.yaml``; the outputs are ``NAME.h`` (the drop-in header, same
layout as the reference's) and ``NAME.hip`` (the HIP library source).  With
``--compile`` it also builds ``libNAME.so`` for gfx950.
"""
import argparse
import os
import sys

from .build import compile_library, generate_program


def main(argv=None):
    ap = argparse.ArgumentParser(prog="formura_amd")
    ap.add_argument("fmr")
    ap.add_argument("--nc", help="numerical config (default: PROG.yaml)")
    ap.add_argument("-o", "--output", help="output .hip path (default: PROG.hip next to PROG.fmr)")
    ap.add_argument("--schedule", default="auto", help="auto | flat | tile2d | stream3d")
    ap.add_argument("--compile", action="store_true", help="also build libNAME.so with hipcc")
    a = ap.parse_args(argv)
    base = os.path.splitext(a.fmr)[0]
    nc = a.nc or base + ".yaml"
    out = a.output or base + ".hip"
    name = os.path.splitext(os.path.basename(out))[0]
    d = generate_program(name, a.fmr, nc, a.schedule, out_dir=os.path.dirname(os.path.abspath(out)))
    print("Generate: %s.h %s.hip" % (os.path.join(d, name), os.path.join(d, name)))
    if a.compile:
        print(compile_library(name, d))


if __name__ == "__main__":
    sys.exit(main())
